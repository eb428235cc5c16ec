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
}


def workload_ic(name, n=None):
    w = WORKLOADS[name]
    return two_disk_encounter(n_total=n or w["n"], seed=w["seed"], eps=w["eps"])
