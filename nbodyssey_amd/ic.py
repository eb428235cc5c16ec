"""Initial-condition generators (host side, NumPy).

Two families:

* ``generate_ic_disks`` reproduces scripts/generate_ic.py:16-128 draw for draw
  (same PCG64 stream, same call order), so its ``ic.bin`` bytes are the ones
  the reference's published energy traces were run on
  (benchmarks/energy_gpu_*.csv used ``--particles 12000``, seed 0,
  scripts/gpu_bench.sh:65).  The order of the random draws has to match the
  reference's for the bytes to match; the code is otherwise this package's own.
* ``two_disk_encounter`` is the benchmark generator: two exponential disks,
  each of total mass 1, surface density ~ exp(-R/Rd) (radius ~ Gamma(2, Rd),
  resampled above ``r_max``), z ~ N(0, z_sigma), circular speed
  sqrt(M(<R) / sqrt(R^2 + eps^2)) plus N(0, 0.05 v) per component; disk A at
  (-10, -2, 0) moving (+0.35, 0, 0), disk B at (+10, +2, 0) moving
  (-0.35, 0, 0) with opposite spin, tilted 30 degrees about x.

Both return (pos4, vel4) float32 arrays of shape (n, 4) in the reference
layout (docs/FORMATS.md:11-16).
"""
import math

import numpy as np


def _rotate_x(v, angle):
    c, s = math.cos(angle), math.sin(angle)
    out = v.copy()
    out[:, 1] = c * v[:, 1] - s * v[:, 2]
    out[:, 2] = s * v[:, 1] + c * v[:, 2]
    return out


def _disk_reference(n, m_total, rd, center, bulk, incl, spin, eps, rng, r_max_factor=4.0):
    """One disk with the draw sequence of generate_ic.py:build_disk."""
    mp = m_total / n
    r_max = r_max_factor * rd
    r = rng.gamma(shape=2.0, scale=rd, size=n)
    bad = r > r_max
    while np.any(bad):
        r[bad] = rng.gamma(shape=2.0, scale=rd, size=int(bad.sum()))
        bad = r > r_max
    r = np.sort(r)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
    z = rng.normal(0.0, 0.05 * rd, size=n)
    m_enc = mp * (np.arange(n) + 0.5)
    accel = m_enc * r / np.power(r * r + eps * eps, 1.5)
    vc = np.sqrt(np.maximum(r * accel, 0.0))
    pos = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    vel = np.stack([-spin * vc * np.sin(phi), spin * vc * np.cos(phi), np.zeros(n)], axis=1)
    if incl != 0.0:
        c, s = np.cos(incl), np.sin(incl)
        rot = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
        pos = pos @ rot.T
        vel = vel @ rot.T
    pos += np.asarray(center)
    vel += np.asarray(bulk)
    pos4 = np.column_stack([pos, np.full(n, mp)]).astype("<f4")
    vel4 = np.column_stack([vel, np.zeros(n)]).astype("<f4")
    return pos4, vel4


def generate_ic_disks(particles=40000, separation=16.0, impact=3.0, approach=0.55, mass=1.0,
                      scale_radius=1.0, eps=0.05, inclination=30.0, seed=0):
    """The IC of scripts/generate_ic.py with its CLI defaults (generate_ic.py:89-128)."""
    rng = np.random.default_rng(seed)
    n_each = particles // 2
    hs, hi = 0.5 * separation, 0.5 * impact
    a = _disk_reference(n_each, mass, scale_radius, [-hs, -hi, 0.0], [approach, 0.0, 0.0],
                        0.0, +1.0, eps, rng)
    b = _disk_reference(n_each, mass, scale_radius, [hs, hi, 0.0], [-approach, 0.0, 0.0],
                        np.radians(inclination), -1.0, eps, rng)
    return np.concatenate([a[0], b[0]]), np.concatenate([a[1], b[1]])


def exponential_disk(n, rng, m_total=1.0, rd=1.0, r_max=5.0, z_sigma=0.1, eps=0.05,
                     dispersion=0.05, spin=1.0):
    """Disk in its own frame: (pos (n,3), vel (n,3), mass (n,)) in float64."""
    mp = m_total / n
    r = rng.gamma(2.0, rd, size=n)
    bad = r > r_max
    while np.any(bad):
        r[bad] = rng.gamma(2.0, rd, size=int(bad.sum()))
        bad = r > r_max
    r.sort()
    phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
    z = rng.normal(0.0, z_sigma, size=n)
    m_enc = mp * (np.arange(n) + 0.5)
    vc = np.sqrt(m_enc / np.sqrt(r * r + eps * eps))
    pos = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    vel = np.stack([-spin * vc * np.sin(phi), spin * vc * np.cos(phi), np.zeros(n)], axis=1)
    vel += rng.normal(0.0, 1.0, size=(n, 3)) * (dispersion * vc)[:, None]
    return pos, vel, np.full(n, mp)


def two_disk_encounter(n_total=12000, seed=0, eps=0.05, r_max=5.0, z_sigma=0.1,
                       inclination_deg=30.0, center=(10.0, 2.0, 0.0), speed=0.35,
                       mass_b=1.0):
    """The benchmark two-disk encounter; n_total is split evenly (A gets the odd one)."""
    rng = np.random.default_rng(seed)
    na = n_total - n_total // 2
    nb = n_total // 2
    pa, va, ma = exponential_disk(na, rng, 1.0, 1.0, r_max, z_sigma, eps, spin=+1.0)
    pb, vb, mb = exponential_disk(nb, rng, mass_b, 1.0, r_max, z_sigma, eps, spin=-1.0)
    incl = math.radians(inclination_deg)
    pb, vb = _rotate_x(pb, incl), _rotate_x(vb, incl)
    c = np.asarray(center, dtype=np.float64)
    pa += -c
    pb += c
    va += np.array([speed, 0.0, 0.0])
    vb += np.array([-speed, 0.0, 0.0])
    pos = np.concatenate([pa, pb])
    vel = np.concatenate([va, vb])
    m = np.concatenate([ma, mb])
    pos4 = np.column_stack([pos, m]).astype(np.float32)
    vel4 = np.column_stack([vel, np.zeros(len(m))]).astype(np.float32)
    return pos4, vel4


def parabolic_encounter(n_total, seed, mass_ratio, inclination_deg, separation, pericentre,
                        eps=0.05, r_max=5.0, z_sigma=0.1):
    """Two disks on a parabolic orbit (ensemble member).

    Disk A has mass 1, disk B mass ``mass_ratio``; particles have equal mass,
    so B gets round(n q / (1 + q)) of them.  B is tilted by the inclination
    about x and spins opposite to A.  The centres start ``separation`` apart
    on the x axis with the parabolic relative velocity of two point masses
    M = 1 + q whose pericentre is ``pericentre``: |v| = sqrt(2M/d), tangential
    part sqrt(2 M r_p) / d, the rest radial and closing; both are placed in
    the centre-of-mass frame.
    """
    rng = np.random.default_rng(seed)
    q = float(mass_ratio)
    nb = int(round(n_total * q / (1.0 + q)))
    nb = min(max(nb, 1), n_total - 1)
    na = n_total - nb
    pa, va, ma = exponential_disk(na, rng, 1.0, 1.0, r_max, z_sigma, eps, spin=+1.0)
    pb, vb, mb = exponential_disk(nb, rng, q * 1.0, 1.0, r_max, z_sigma, eps, spin=-1.0)
    # equal particle masses: (1 + q) / n_total each
    ma[:] = mb[:] = (1.0 + q) / n_total
    incl = math.radians(inclination_deg)
    pb, vb = _rotate_x(pb, incl), _rotate_x(vb, incl)
    m_tot = 1.0 + q
    d, rp = float(separation), float(pericentre)
    v2 = 2.0 * m_tot / d
    vt = math.sqrt(2.0 * m_tot * rp) / d
    vr = -math.sqrt(max(v2 - vt * vt, 0.0))
    r_rel = np.array([d, 0.0, 0.0])
    v_rel = np.array([vr, vt, 0.0])
    pa += -(q / m_tot) * r_rel
    va += -(q / m_tot) * v_rel
    pb += (1.0 / m_tot) * r_rel
    vb += (1.0 / m_tot) * v_rel
    pos = np.concatenate([pa, pb])
    vel = np.concatenate([va, vb])
    m = np.concatenate([ma, mb])
    pos4 = np.column_stack([pos, m]).astype(np.float32)
    vel4 = np.column_stack([vel, np.zeros(len(m))]).astype(np.float32)
    return pos4, vel4


def ensemble_encounters(nsys=256, seed=0, n_min=2048, n_max=65536, eps=0.05):
    """Batched-ensemble workload: nsys independent parabolic encounters.

    System k draws its parameters from default_rng(seed + k): particle count
    log-uniform on [n_min, n_max], mass ratio U[0.25, 1], inclination
    U[0, 90] degrees, separation U[15, 25], pericentre U[1, 5] (disk scale
    lengths), then builds ``parabolic_encounter`` with seed
    (seed + k) * 7919 + 1.  Returns (list of (pos4, vel4), list of dicts).
    """
    systems, meta = [], []
    for k in range(nsys):
        r = np.random.default_rng(seed + k)
        n = int(round(math.exp(r.uniform(math.log(n_min), math.log(n_max)))))
        n = min(max(n, n_min), n_max)
        prm = dict(n=n, mass_ratio=float(r.uniform(0.25, 1.0)),
                   inclination_deg=float(r.uniform(0.0, 90.0)),
                   separation=float(r.uniform(15.0, 25.0)),
                   pericentre=float(r.uniform(1.0, 5.0)))
        systems.append(parabolic_encounter(n, (seed + k) * 7919 + 1, prm["mass_ratio"],
                                           prm["inclination_deg"], prm["separation"],
                                           prm["pericentre"], eps=eps))
        meta.append(prm)
    return systems, meta


def _jeans_sigma2(r_grid, rho, m_total_enc):
    """Isotropic Jeans equation, sigma_r^2(r) = (1/rho) int_r^inf rho G M(<r') / r'^2 dr'
    (G = 1), by the trapezoid rule on the grid from the outside in."""
    integrand = rho * m_total_enc / (r_grid * r_grid)
    seg = 0.5 * (integrand[1:] + integrand[:-1]) * np.diff(r_grid)
    tail = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
    return tail / np.maximum(rho, 1e-300)


def composite_galaxy(n, rng, m_total=1.0, f_disk=0.6, f_bulge=0.1, f_halo=0.3, rd=1.0,
                     z0=0.1, a_bulge=0.2, a_halo=5.0, eps=0.02, m_cut=0.99):
    """Disk + Hernquist bulge + Plummer halo in equilibrium (own frame, float64).

    * disk: exponential surface density (R ~ Gamma(2, rd), resampled above
      6 rd), z ~ N(0, z0), circular speed sqrt(M(<R) / sqrt(R^2 + eps^2))
      from the spherical enclosed mass of all three components, plus
      N(0, 0.05 v_c) per component;
    * bulge: Hernquist, M(<r) = M r^2 / (r + a)^2, r by inverse CDF;
    * halo: Plummer, M(<r) = M r^3 / (r^2 + a^2)^(3/2), r by inverse CDF;
      both radially truncated at enclosed-mass fraction m_cut, isotropic
      directions, velocities Gaussian with the isotropic Jeans dispersion
      of each component in the total potential (disk mass taken as
      spherically distributed, M_d(<r) = M_d [1 - (1 + r/rd) e^(-r/rd)]).
    Particle counts are split by mass fraction (equal particle masses).
    Returns (pos (n,3), vel (n,3), mass (n,)).
    """
    nb_ = int(round(n * f_bulge))
    nh = int(round(n * f_halo))
    nd = n - nb_ - nh
    mp = m_total / n
    md, mb, mh = m_total * f_disk, m_total * f_bulge, m_total * f_halo

    def m_disk(r):
        x = r / rd
        return md * (1.0 - (1.0 + x) * np.exp(-x))

    def m_bulge(r):
        return mb * r * r / (r + a_bulge) ** 2

    def m_halo(r):
        return mh * r ** 3 / (r * r + a_halo * a_halo) ** 1.5

    def m_enc(r):
        return m_disk(r) + m_bulge(r) + m_halo(r)

    # disk
    R = rng.gamma(2.0, rd, size=nd)
    bad = R > 6.0 * rd
    while np.any(bad):
        R[bad] = rng.gamma(2.0, rd, size=int(bad.sum()))
        bad = R > 6.0 * rd
    phi = rng.uniform(0.0, 2.0 * np.pi, size=nd)
    z = rng.normal(0.0, z0, size=nd)
    vc = np.sqrt(m_enc(R) * R * R / (R * R + eps * eps) ** 1.5)
    pd = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    vd = np.stack([-vc * np.sin(phi), vc * np.cos(phi), np.zeros(nd)], axis=1)
    vd += rng.normal(0.0, 1.0, size=(nd, 3)) * (0.05 * vc)[:, None]

    def iso_dirs(k):
        u = rng.uniform(-1.0, 1.0, size=k)
        t = rng.uniform(0.0, 2.0 * np.pi, size=k)
        s = np.sqrt(1.0 - u * u)
        return np.stack([s * np.cos(t), s * np.sin(t), u], axis=1)

    # spheroids: inverse-CDF radii, truncated at m_cut of their mass
    q = rng.uniform(0.0, m_cut, size=nb_)
    sq = np.sqrt(q)
    rb = a_bulge * sq / (1.0 - sq)
    q = rng.uniform(0.0, m_cut, size=nh)
    rh = a_halo / np.sqrt(q ** (-2.0 / 3.0) - 1.0)
    r_grid = np.logspace(-4, np.log10(max(rh.max(), rb.max(), 10.0) * 4.0), 4000)
    menc = m_enc(r_grid)
    rho_b = mb * a_bulge / (2 * np.pi * r_grid * (r_grid + a_bulge) ** 3)
    rho_h = 3 * mh / (4 * np.pi * a_halo ** 3) * (1 + r_grid ** 2 / a_halo ** 2) ** -2.5
    sig_b = np.sqrt(np.maximum(_jeans_sigma2(r_grid, rho_b, menc), 0.0))
    sig_h = np.sqrt(np.maximum(_jeans_sigma2(r_grid, rho_h, menc), 0.0))
    pb = iso_dirs(nb_) * rb[:, None]
    vb = rng.normal(0.0, 1.0, size=(nb_, 3)) * np.interp(rb, r_grid, sig_b)[:, None]
    ph = iso_dirs(nh) * rh[:, None]
    vh = rng.normal(0.0, 1.0, size=(nh, 3)) * np.interp(rh, r_grid, sig_h)[:, None]
    pos = np.concatenate([pd, pb, ph])
    vel = np.concatenate([vd, vb, vh])
    return pos, vel, np.full(n, mp)


def composite_encounter(n_total=1 << 25, seed=5, pericentre=3.0, separation=30.0,
                        inclination_deg=45.0, eps=0.02):
    """Two equal composite galaxies (``composite_galaxy``, mass 1 each) on a
    parabolic orbit of the given pericentre, starting ``separation`` apart;
    the second is tilted by the inclination about x and counter-rotates."""
    rng = np.random.default_rng(seed)
    na = n_total - n_total // 2
    nb_ = n_total // 2
    pa, va, ma = composite_galaxy(na, rng, eps=eps)
    pb, vb, mb = composite_galaxy(nb_, rng, eps=eps)
    ma[:] = mb[:] = 2.0 / n_total
    vb = -vb   # counter-rotating (the spheroids' isotropic velocities stay isotropic)
    incl = math.radians(inclination_deg)
    pb, vb = _rotate_x(pb, incl), _rotate_x(vb, incl)
    m_tot, d, rp = 2.0, float(separation), float(pericentre)
    v2 = 2.0 * m_tot / d
    vt = math.sqrt(2.0 * m_tot * rp) / d
    vr = -math.sqrt(max(v2 - vt * vt, 0.0))
    r_rel = np.array([d, 0.0, 0.0])
    v_rel = np.array([vr, vt, 0.0])
    pa += -0.5 * r_rel
    va += -0.5 * v_rel
    pb += 0.5 * r_rel
    vb += 0.5 * v_rel
    pos4 = np.empty((n_total, 4), np.float32)
    vel4 = np.zeros((n_total, 4), np.float32)
    pos4[:na, :3] = pa
    pos4[na:, :3] = pb
    pos4[:, 3] = 2.0 / n_total
    vel4[:na, :3] = va
    vel4[na:, :3] = vb
    return pos4, vel4


# Workloads the benchmark and the tests name (sizes, softening, step, theta).
WORKLOADS = {
    "two-disk-12k-naive": dict(n=12000, seed=0, eps=0.05, dt=0.01, steps=200, theta=0.5,
                               force="naive"),
    "two-disk-1M-bh": dict(n=1048576, seed=1, eps=0.02, dt=0.005, steps=100, theta=0.5,
                           force="bh", key_bits=30),
    "two-disk-4M-precision": dict(n=4194304, seed=2, eps=0.02, dt=0.005, steps=500,
                                  theta=0.5, force="bh", key_bits=30),
    "ensemble-256-parabolic": dict(nsys=256, seed=0, eps=0.05, dt=0.01, steps=200, theta=0.5,
                                   force="bh"),
    "composite-33M-63bit": dict(n=1 << 25, seed=5, eps=0.02, dt=0.002, steps=50, theta=0.6,
                                force="bh", key_bits=63),
}


def workload_ic(name, n=None):
    w = WORKLOADS[name]
    return two_disk_encounter(n_total=n or w["n"], seed=w["seed"], eps=w["eps"])
