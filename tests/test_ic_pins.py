"""Pins of the initial-condition generators of the ensemble and composite
workloads (nbodyssey_amd/ic.py) -- CPU only.

These generators extend the reference's disk-only scripts/generate_ic.py:16-75
(spheroids, Jeans velocities, parabolic orbits), so nothing in the reference
fixes their output; they are checked against the mathematics instead:
  * sampled bulge / halo radii against the Hernquist and Plummer enclosed-mass
    closed forms (Kolmogorov-Smirnov distance at the 99.9% level);
  * the isotropic Jeans solver against the self-consistent Plummer closed
    form, and the sampled spheroid velocities against a quadrature of the
    Jeans integral in the total (disk + bulge + halo) potential;
  * the orbit: zero two-body energy and the requested pericentre, from the
    particles' own centres of mass.
Each check is chosen so that a plausible slip (a swapped profile, a missing
component in the potential, sqrt(2) in a speed) fails it.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from nbodyssey_amd import ic
from oracle import numpy_ref, orc

KS99 = 1.95  # Kolmogorov-Smirnov critical value at the 99.9% level, times sqrt(N)


def hernquist_m(r, m, a):
    return m * r * r / (r + a) ** 2


def plummer_m(r, m, a):
    return m * r ** 3 / (r * r + a * a) ** 1.5


def hernquist_rho(r, m, a):
    return m * a / (2 * np.pi * r * (r + a) ** 3)


def plummer_rho(r, m, a):
    return 3 * m / (4 * np.pi * a ** 3) * (1 + r * r / (a * a)) ** -2.5


def disk_m_spherical(r, m, rd):
    x = r / rd
    return m * (1 - (1 + x) * np.exp(-x))


def ks_distance(samples, cdf):
    x = np.sort(samples)
    f = cdf(x)
    k = np.arange(1, len(x) + 1) / len(x)
    return max(np.abs(k - f).max(), np.abs(k - 1.0 / len(x) - f).max())


@pytest.fixture(scope="module")
def galaxy():
    n = 200000
    pos, vel, m = ic.composite_galaxy(n, np.random.default_rng(11))
    nb = int(round(n * 0.1))
    nh = int(round(n * 0.3))
    nd = n - nb - nh
    return dict(pos=pos, vel=vel, m=m, disk=slice(0, nd), bulge=slice(nd, nd + nb),
                halo=slice(nd + nb, n))


def test_component_masses(galaxy):
    m = galaxy["m"]
    assert np.allclose(m.sum(), 1.0)
    assert abs(m[galaxy["disk"]].sum() - 0.6) < 1e-4
    assert abs(m[galaxy["bulge"]].sum() - 0.1) < 1e-4
    assert abs(m[galaxy["halo"]].sum() - 0.3) < 1e-4


def test_bulge_radii_follow_hernquist(galaxy):
    """M(<r) = M r^2 / (r + a)^2 with a = 0.2, truncated at 99% of the mass."""
    r = np.linalg.norm(galaxy["pos"][galaxy["bulge"]], axis=1)
    d = ks_distance(r, lambda x: hernquist_m(x, 1.0, 0.2) / 0.99)
    assert d < KS99 / math.sqrt(len(r)), d
    # a plausible slip -- the Plummer profile with the same scale -- is far off
    assert ks_distance(r, lambda x: plummer_m(x, 1.0, 0.2) / 0.99) > 10 * KS99 / math.sqrt(len(r))


def test_halo_radii_follow_plummer(galaxy):
    """M(<r) = M r^3 / (r^2 + a^2)^(3/2) with a = 5, truncated at 99% of the mass."""
    r = np.linalg.norm(galaxy["pos"][galaxy["halo"]], axis=1)
    d = ks_distance(r, lambda x: plummer_m(x, 1.0, 5.0) / 0.99)
    assert d < KS99 / math.sqrt(len(r)), d
    assert ks_distance(r, lambda x: hernquist_m(x, 1.0, 5.0) / 0.99) > 10 * KS99 / math.sqrt(len(r))


def test_spheroid_directions_isotropic(galaxy):
    for part in ("bulge", "halo"):
        p = galaxy["pos"][galaxy[part]]
        u = p / np.linalg.norm(p, axis=1, keepdims=True)
        # mean direction ~ 0 and <u_z^2> = 1/3 for isotropy
        assert np.abs(u.mean(axis=0)).max() < 5 / math.sqrt(len(u))
        assert abs((u[:, 2] ** 2).mean() - 1 / 3) < 5 * math.sqrt(4 / 45 / len(u))


@pytest.mark.parametrize("m,a", [(1.0, 1.0), (0.3, 5.0)])
def test_densities_integrate_to_enclosed_mass(m, a):
    """The densities the Jeans solve uses integrate to the closed-form M(<r)."""
    for r in (0.1 * a, a, 7.0 * a):
        mp = integrate.quad(lambda x: 4 * np.pi * x * x * plummer_rho(x, m, a), 0, r)[0]
        mh = integrate.quad(lambda x: 4 * np.pi * x * x * hernquist_rho(x, m, a), 0, r)[0]
        assert abs(mp / plummer_m(r, m, a) - 1) < 1e-8
        assert abs(mh / hernquist_m(r, m, a) - 1) < 1e-8


def test_jeans_solver_plummer_closed_form():
    """A self-consistent Plummer sphere (G = 1) has the isotropic dispersion
    sigma^2(r) = M / (6 sqrt(r^2 + a^2))."""
    m, a = 1.0, 1.0
    r = np.logspace(-4, np.log10(400.0), 4000)
    s2 = ic._jeans_sigma2(r, plummer_rho(r, m, a), plummer_m(r, m, a))
    exact = m / (6 * np.sqrt(r * r + a * a))
    sel = (r > 0.01) & (r < 5.0)
    assert np.abs(s2[sel] / exact[sel] - 1).max() < 2e-3


def jeans_quad(r, rho, m_tot):
    """sigma^2(r) = (1/rho) int_r^inf rho(x) M_tot(<x) / x^2 dx by quadrature."""
    f = lambda x: rho(x) * m_tot(x) / (x * x)
    v = integrate.quad(f, r, 50 * r + 200.0, limit=400)[0] + integrate.quad(f, 50 * r + 200.0,
                                                                           np.inf)[0]
    return v / rho(r)


@pytest.mark.parametrize("part", ["bulge", "halo"])
def test_spheroid_velocities_match_jeans_quadrature(galaxy, part):
    """In radial bins, the sampled 1-D velocity variance of each spheroid
    equals the Jeans sigma^2 in the total potential of the three components
    (G = 1) to within its sampling error; leaving the disk out of the
    potential (a plausible slip) fails the inner bins."""
    md, mb, mh = 0.6, 0.1, 0.3

    def m_tot(x):
        return disk_m_spherical(x, md, 1.0) + hernquist_m(x, mb, 0.2) + plummer_m(x, mh, 5.0)

    def m_no_disk(x):
        return hernquist_m(x, mb, 0.2) + plummer_m(x, mh, 5.0)

    rho = (lambda x: hernquist_rho(x, mb, 0.2)) if part == "bulge" else \
          (lambda x: plummer_rho(x, mh, 5.0))
    p = galaxy["pos"][galaxy[part]]
    v = galaxy["vel"][galaxy[part]]
    r = np.linalg.norm(p, axis=1)
    edges = np.quantile(r, np.linspace(0.05, 0.85, 9))
    bad_fit = 0
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (r >= lo) & (r < hi)
        k = int(sel.sum())
        var = float((v[sel] ** 2).mean())  # mean of vx^2, vy^2, vz^2 (zero-mean Gaussians)
        rs = np.quantile(r[sel], np.linspace(0.05, 0.95, 7))
        expect = float(np.mean([jeans_quad(x, rho, m_tot) for x in rs]))
        wrong = float(np.mean([jeans_quad(x, rho, m_no_disk) for x in rs]))
        tol = 5 * math.sqrt(2.0 / (3 * k)) + 0.02  # sampling error + bin averaging
        assert abs(var / expect - 1) < tol, (part, lo, hi, var, expect)
        bad_fit += abs(var / wrong - 1) > tol
    assert bad_fit >= 1


def orbit_elements(pos4, vel4, split):
    """Relative separation / velocity of the two halves' centres of mass."""
    m = pos4[:, 3].astype(np.float64)
    p = pos4[:, :3].astype(np.float64)
    v = vel4[:, :3].astype(np.float64)
    ca = (m[:split, None] * p[:split]).sum(0) / m[:split].sum()
    cb = (m[split:, None] * p[split:]).sum(0) / m[split:].sum()
    va = (m[:split, None] * v[:split]).sum(0) / m[:split].sum()
    vb = (m[split:, None] * v[split:]).sum(0) / m[split:].sum()
    return cb - ca, vb - va, m[:split].sum() + m[split:].sum()


@pytest.mark.parametrize("gen", ["composite", "parabolic"])
def test_parabolic_orbit_energy_and_pericentre(gen):
    """The galaxies' centres start on a parabola (two-body energy 0) closing
    towards the requested pericentre r_p = h^2 / (2 M), h = |r x v|.  The
    tolerances are about three times the sampling noise of the centres'
    velocities (each galaxy's own mean velocity is zero only to
    ~ v_c / sqrt(n)); a slip such as |v|^2 = M / d instead of 2 M / d moves
    the energy ratio by 0.5."""
    if gen == "composite":
        n, rp, d = 200000, 3.0, 30.0
        pos, vel = ic.composite_encounter(n, seed=3, pericentre=rp, separation=d)
        split = n - n // 2
    else:
        n, rp, d, q = 400000, 2.5, 20.0, 0.5
        pos, vel = ic.parabolic_encounter(n, 5, q, 40.0, d, rp)
        split = n - int(round(n * q / (1 + q)))
    r, v, m = orbit_elements(pos, vel, split)
    dist = np.linalg.norm(r)
    energy = 0.5 * v @ v - m / dist
    h = np.linalg.norm(np.cross(r, v))
    assert abs(dist / d - 1) < 5e-3
    assert abs(energy) / (m / dist) < 2e-2
    assert abs(h * h / (2 * m) / rp - 1) < 4e-2
    assert r @ v < 0  # approaching


def test_ensemble_parameters_in_their_ranges():
    systems, meta = ic.ensemble_encounters(64, seed=7)
    for (p, v), prm in zip(systems, meta):
        assert 2048 <= prm["n"] <= 65536 and len(p) == prm["n"]
        assert 0.25 <= prm["mass_ratio"] <= 1.0 and 0.0 <= prm["inclination_deg"] <= 90.0
        assert 15.0 <= prm["separation"] <= 25.0 and 1.0 <= prm["pericentre"] <= 5.0
        q = prm["mass_ratio"]
        np.testing.assert_allclose(p[:, 3].sum(dtype=np.float64), 1 + q, rtol=1e-5)
    n = np.array([m["n"] for m in meta])
    # log-uniform on [2048, 65536]: the mean of log n sits mid-way
    assert abs(np.log(n).mean() - 0.5 * (np.log(2048) + np.log(65536))) < 0.3


def test_numpy_kdk_run_matches_oracle_steps():
    """numpy_ref.kdk_run (the timed CPU baseline, reference_nbody.py:129-136)
    against the oracle's KDK with the FP32 force of the same law."""
    pos, vel = ic.two_disk_encounter(1500, seed=12, eps=0.05)
    p1, v1, _ = numpy_ref.kdk_run(pos[:, :3], vel[:, :3], pos[:, 3], 0.05, 0.01, 5, chunk=512)
    p2, v2 = pos.copy(), vel.copy()
    acc = orc.allpairs_f32(p2, 0.05)
    for _ in range(5):
        orc.kdk_step_naive(p2, v2, acc, 0.01, 0.05)
    assert np.abs(p1 - p2[:, :3]).max() < 1e-5
    dv = np.linalg.norm(v1 - v2[:, :3], axis=1) / np.linalg.norm(v2[:, :3], axis=1)
    assert np.median(dv) < 1e-5


def test_numpy_pool_matches_serial():
    """The pooled CPU baseline (one target block per worker) computes what
    the serial restatement computes."""
    from multiprocessing import get_context
    pos, _ = ic.two_disk_encounter(3000, seed=13, eps=0.05)
    p, m = pos[:, :3].copy(), pos[:, 3].copy()
    eps2, g = np.float32(0.0025), np.float32(1.0)
    a1 = numpy_ref.accelerations(p, m, eps2, g, chunk=512)
    with get_context("fork").Pool(3) as pool:
        a2 = numpy_ref.accelerations(p, m, eps2, g, chunk=512, pool=pool)
    np.testing.assert_allclose(a2, a1, rtol=1e-6, atol=0)
