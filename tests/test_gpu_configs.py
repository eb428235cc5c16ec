"""GPU tests of the precision study and the large composite system.

* Precision forms (FP32, compensated, FP16 far field) on the two-disk
  encounter at theta 0.3 / 0.5 / 0.7: node decisions (interaction and opening
  counts) equal the oracle's wave64 group walk; the force error against FP64
  all-pairs is stated and bounded per theta; and the arithmetic error alone
  -- against the same interaction lists summed in FP64 by the oracle -- shows
  the compensated sums at or below the plain FP32 ones.
* The 4,194,304-particle precision-study instance, one force evaluation per
  form, through size-independent properties (sorted keys, permutation, root
  mass, finite forces, error against the GPU's FP64 direct sum).
* The composite-galaxy generator with 63-bit keys at theta 0.6: the whole
  tree bit for bit against the oracle at 1,048,576 particles, and the full
  33,554,432-particle system through the properties.
"""
import numpy as np
import pytest

from nbodyssey_amd import engine, ic
from oracle import orc

pytestmark = pytest.mark.gpu

FORMS = {"fp32": {}, "kahan": {"compensated": True}, "fp16far": {"far16": True}}


def rel_err(a, b):
    a = np.asarray(a, np.float64)[:, :3]
    b = np.asarray(b, np.float64)[:, :3]
    return np.linalg.norm(a - b, axis=1) / (np.linalg.norm(b, axis=1) + 1e-300)


@pytest.fixture(scope="module")
def disk65k():
    pos, _ = ic.two_disk_encounter(65536, seed=1, eps=0.02)
    idx = np.random.default_rng(0).choice(len(pos), 512, replace=False).astype(np.int32)
    return pos, idx, orc.allpairs_f64(pos, 0.02, targets=idx), orc.Tree(pos, key_bits=30)


# median force error bound vs FP64 all-pairs, per theta (the multipole error;
# the reference quotes ~0.3% / ~1% / ~3% at 0.3 / 0.5 / 0.8, PERFORMANCE.md:84-88)
MEDIAN_BOUND = {0.3: 0.004, 0.5: 0.01, 0.7: 0.025}


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_precision_forms_vs_fp64(disk65k, theta):
    pos, idx, ref, tree = disk65k
    eps = 0.02
    _, _, gi, go, _ = tree.traverse_group(theta, eps, 64)
    lists64, _ = tree.traverse_group_f64sum(theta, eps, 64)
    kw = dict(eps=eps, theta=theta, force="bh", traverse="warp", group_width=64, key_bits=30)
    stats = {}
    for form, extra in FORMS.items():
        print(f"precision theta={theta} form={form}", flush=True)
        with engine.Simulation(pos, **kw, **extra) as sim:
            a = sim.compute_forces()
            ci, co = sim.traverse_counts("warp", 64)
        assert np.all(np.isfinite(a)), form
        assert np.array_equal(ci, gi) and np.array_equal(co, go), form
        e = rel_err(a[idx], ref)
        arith = rel_err(a, lists64)
        stats[form] = (np.median(e), np.quantile(e, 0.99), np.median(arith),
                       np.quantile(arith, 0.99))
        print(f"  total err median {stats[form][0]:.4e} p99 {stats[form][1]:.4e}; "
              f"arithmetic median {stats[form][2]:.3e} p99 {stats[form][3]:.3e}", flush=True)
    f32, kah, f16 = stats["fp32"], stats["kahan"], stats["fp16far"]
    assert f32[0] < MEDIAN_BOUND[theta] and f32[1] < 3 * MEDIAN_BOUND[theta]
    # compensated sums: at most half the arithmetic error of plain FP32
    # (measured: 4-8x lower), and the total error is the multipole error
    # either way
    assert kah[2] <= 0.5 * f32[2] and kah[3] <= 0.5 * f32[3]
    assert abs(kah[0] - f32[0]) < 1e-4 * f32[0]
    # FP16 far field: rounding well under the multipole error
    assert f16[0] < f32[0] + 5e-4 and f16[1] < f32[1] + 2e-3


def test_allpairs_compensated_arithmetic_65536():
    """All-pairs at 65,536 (no multipole error: the whole error is
    arithmetic): the compensated tile sums are no less accurate than plain
    FP32 against the FP64 direct sum, in median and p99 (measured: about 16x
    lower; at least 4x is required)."""
    pos, _ = ic.two_disk_encounter(65536, seed=1, eps=0.02)
    idx = np.random.default_rng(4).choice(len(pos), 512, replace=False).astype(np.int32)
    ref = orc.allpairs_f64(pos, 0.02, targets=idx)
    with engine.Simulation(pos, eps=0.02, force="naive") as sim:
        a = sim.compute_forces()
        sim.update(compensated=True)
        ac = sim.compute_forces()
    e, ec = rel_err(a[idx], ref), rel_err(ac[idx], ref)
    print(f"allpairs 65536: fp32 median {np.median(e):.3e} p99 {np.quantile(e, .99):.3e}; "
          f"compensated median {np.median(ec):.3e} p99 {np.quantile(ec, .99):.3e}", flush=True)
    assert np.median(e) < 1e-5
    assert np.median(ec) <= 0.25 * np.median(e)
    assert np.quantile(ec, 0.99) <= 0.25 * np.quantile(e, 0.99)


def check_tree_properties(sim, pos, key_bits):
    t = sim.tree()
    k, perm = t["keys"], t["perm"].astype(np.int64)
    n = len(pos)
    assert np.all(k[1:] >= k[:-1]), "keys not sorted"
    assert np.all(np.bincount(perm, minlength=n) == 1), "perm is not a permutation"
    ties = k[1:] == k[:-1]
    assert np.all(perm[1:][ties] > perm[:-1][ties]), "ties not in index order"
    assert t["parent"][0] == -1 and np.all(t["parent"][1:] >= 0)
    m = pos[:, 3].astype(np.float64).sum()
    assert abs(t["com4"][0, 3] / m - 1) < 1e-6, "root mass"
    assert np.all(np.isfinite(t["com4"])) and np.all(t["size2"] >= 0)
    return t


@pytest.mark.parametrize("form", list(FORMS))
def test_precision_4m_properties(form):
    """The 4,194,304-particle instance of the precision study, theta 0.5."""
    pos, _ = ic.workload_ic("two-disk-4M-precision")
    print(f"4M form={form} n={len(pos)}", flush=True)
    idx = np.random.default_rng(2).choice(len(pos), 256, replace=False).astype(np.int32)
    with engine.Simulation(pos, eps=0.02, theta=0.5, force="bh", traverse="warp",
                           group_width=64, key_bits=30, **FORMS[form]) as sim:
        a = sim.compute_forces()
        assert np.all(np.isfinite(a))
        check_tree_properties(sim, pos, 30)
        e = sim.force_error(idx, a)
    print(f"  median {np.median(e):.4e} p99 {np.quantile(e, 0.99):.4e}", flush=True)
    assert np.median(e) < 0.01


def test_composite_1m_tree_bit_exact():
    """Composite galaxies (disk + Hernquist bulge + Plummer halo, parabolic
    orbit) at 1,048,576 particles, 63-bit keys, theta 0.6: keys, permutation,
    topology, node COM and size^2 equal the oracle's bit for bit; the median
    force error vs FP64 all-pairs is within 1%."""
    pos, _ = ic.composite_encounter(1 << 20, seed=5)
    idx = np.random.default_rng(3).choice(len(pos), 256, replace=False).astype(np.int32)
    with engine.Simulation(pos, eps=0.02, theta=0.6, force="bh", traverse="warp",
                           group_width=64, key_bits=63) as sim:
        a = sim.compute_forces()
        t = sim.tree()
    ref = orc.Tree(pos, key_bits=63)
    assert np.array_equal(t["keys"], ref.keys_sorted)
    assert np.array_equal(t["perm"], ref.perm)
    assert np.array_equal(t["children"], ref.children)
    assert np.array_equal(t["parent"], ref.parent)
    assert np.array_equal(t["com4"], ref.com4)
    assert np.array_equal(t["size2"], ref.size2)
    e = rel_err(a[idx], orc.allpairs_f64(pos, 0.02, targets=idx))
    print(f"composite 1M: median {np.median(e):.4e} p99 {np.quantile(e, 0.99):.4e}", flush=True)
    assert np.median(e) < 0.01 and np.all(np.isfinite(a))


def test_composite_33m_properties():
    """The 33,554,432-particle composite system (63-bit keys, theta 0.6,
    dt 0.002): one force evaluation and one leapfrog step through the
    properties, and the median force error vs the GPU's FP64 direct sum
    within 1%."""
    pos, vel = ic.composite_encounter(1 << 25, seed=5)
    print(f"composite 33M generated n={len(pos)}", flush=True)
    idx = np.random.default_rng(5).choice(len(pos), 256, replace=False).astype(np.int32)
    with engine.Simulation(pos, vel, dt=0.002, eps=0.02, theta=0.6, force="bh",
                           traverse="warp", group_width=64, key_bits=63) as sim:
        a = sim.compute_forces()
        assert np.all(np.isfinite(a))
        check_tree_properties(sim, pos, 63)
        e = sim.force_error(idx, a)
        print(f"  median {np.median(e):.4e} p99 {np.quantile(e, 0.99):.4e}", flush=True)
        assert np.median(e) < 0.01
        sim.step(1)
        p1, v1 = sim.download()
    assert np.all(np.isfinite(p1)) and np.all(np.isfinite(v1))
    # one step moves particles by about v dt
    d = np.linalg.norm(p1[:, :3] - pos[:, :3], axis=1)
    assert np.median(d) < 0.01 and np.median(d) > 0
