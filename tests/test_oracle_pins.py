"""Pins of the CPU oracle (oracle/nbody_oracle.c) -- CPU only.

Each oracle function is checked against something other than its own formula:
brute force, closed forms, an independent construction, the reference's own
test assertions (scripts/lbvh_check.py) and its committed data
(benchmarks/energy_gpu_naive.csv).  The tests also plant mistakes in the oracle
(a mis-wired interleave, an off-by-one split) to show the pins catch them.
"""
import numpy as np
import pytest

from nbodyssey_amd import ic
from oracle import numpy_ref, orc


def blobs_with_dupes():
    """scripts/lbvh_check.py:220-229 input (FP64 there; FP32-rounded here)."""
    rng = np.random.default_rng(1)
    a = rng.normal([-2.0, 0.0, 0.0], 0.7, size=(1500, 3))
    b = rng.normal([2.0, 0.5, 0.1], 0.7, size=(1500, 3))
    d = np.tile([[0.123, -0.456, 0.789]], (24, 1))
    p = np.concatenate([a, b, d])
    return np.column_stack([p, np.full(len(p), 1.0 / len(p))]).astype(np.float32)


# ------------------------------------------------------------------ Morton

def spread_bruteforce(v):
    """bit k of v -> bit 3k, by a plain loop."""
    out = np.zeros_like(v, dtype=np.uint64)
    for k in range(21):
        out |= ((v >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k)
    return out


def test_spread21_all_inputs():
    v = np.arange(1 << 21, dtype=np.uint64)
    ref = spread_bruteforce(v)
    lib = orc.lib()
    # the C function on a strided subset plus both ends (all 2^21 via numpy loop
    # would be slow through ctypes); the closed forms below cover the rest
    idx = np.concatenate([np.arange(0, 1 << 21, 4099), np.arange((1 << 21) - 64, 1 << 21)])
    got = np.array([lib.orc_spread21(int(x)) for x in idx], dtype=np.uint64)
    assert np.array_equal(got, ref[idx])
    assert orc.spread21(0x1FFFFF) == 0x1249249249249249
    assert orc.spread21(1) == 1 and orc.spread21(2) == 8 and orc.spread21(1 << 20) == 1 << 60


def test_morton_all_quantised_values_one_axis():
    """Every 21-bit grid value on the x axis maps to spread(q) << 2."""
    q = np.arange(0, 1 << 21, 997, dtype=np.int64)
    # positions exactly on grid points of a [0, 2^21) cube -> q recovered exactly
    n = len(q) + 1
    pos = np.zeros((n, 4), np.float32)
    pos[:-1, 0] = q.astype(np.float32)
    pos[-1, :3] = float(1 << 21)  # sets the cube extent to 2^21
    keys = orc.morton(pos, 63)
    qq = np.minimum(q, (1 << 21) - 1).astype(np.uint64)
    assert np.array_equal(keys[:-1], spread_bruteforce(qq) << np.uint64(2))


def test_morton_axis_order_closed_forms():
    """q = (1,0,0) -> 4, (0,1,0) -> 2, (0,0,1) -> 1 (forces_bh.cu:218)."""
    s = float(1 << 21)
    pos = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [s, s, s, 1]],
                   np.float32)
    k = orc.morton(pos, 63)
    assert list(k[:4]) == [0, 4, 2, 1]
    assert k[4] == (1 << 63) - 1  # clamped to 2^21-1 on every axis


def test_key30_is_key63_shifted():
    for pos in [blobs_with_dupes(), ic.two_disk_encounter(20000, seed=4)[0]]:
        assert np.array_equal(orc.morton(pos, 30), orc.morton(pos, 63) >> np.uint64(33))


def test_morton_quantisation_fp64_closed_form():
    """q = trunc(((p - lo) * (1/ext)) * 2^21) in FP64 (lbvh_check.py:45-52)."""
    pos = ic.two_disk_encounter(5000, seed=9)[0]
    p = pos[:, :3].astype(np.float64)
    lo = p.min(axis=0)
    inv = 1.0 / (p.max(axis=0) - lo).max()
    q = np.clip(np.trunc((p - lo) * inv * 2097152.0), 0, 2097151).astype(np.uint64)
    ref = (spread_bruteforce(q[:, 0]) << np.uint64(2)) | (spread_bruteforce(q[:, 1]) << np.uint64(1)) \
        | spread_bruteforce(q[:, 2])
    assert np.array_equal(orc.morton(pos, 63), ref)


def test_lbvh_check_duplicate_keys():
    """lbvh_check.py:237-239: the clump of 24 coincident points collides."""
    keys = orc.morton(blobs_with_dupes(), 63)
    assert len(keys) - len(np.unique(keys)) >= 23


def test_degenerate_extent_gives_zero_keys():
    pos = np.tile(np.array([[1.5, -2.0, 3.0, 1.0]], np.float32), (10, 1))
    assert np.all(orc.morton(pos, 63) == 0)


# -------------------------------------------------------------------- sort

@pytest.mark.parametrize("n", [1, 2, 3, 1000, 50001])
def test_sort_matches_numpy_stable(n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, max(n // 5, 1), n).astype(np.uint64) << np.uint64(40)
    ks, perm = orc.sort_pairs(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, ref)
    assert np.array_equal(ks, keys[ref])


# ------------------------------------------------------------------ Karras

def topology_ok(children, parent, n):
    """lbvh_check.py:103-118: parents consistent, each leaf reached once."""
    assert parent[0] == -1
    assert np.all(parent[1:] >= 0)
    seen, stack, visited = 0, [0], set()
    while stack:
        i = stack.pop()
        assert i not in visited
        visited.add(i)
        if i >= n - 1:
            seen += 1
        else:
            stack.extend(int(c) for c in children[i])
    assert seen == n and len(visited) == 2 * n - 1


@pytest.mark.parametrize("bits", [63, 30])
@pytest.mark.parametrize("case", ["blobs", "disk", "same", "n2", "n3", "pairs"])
def test_karras_equals_topdown(case, bits):
    if case == "blobs":
        pos = blobs_with_dupes()
    elif case == "disk":
        pos = ic.two_disk_encounter(8000, seed=2)[0]
    elif case == "same":
        pos = np.tile(np.array([[0.1, 0.2, 0.3, 1.0]], np.float32), (300, 1))
    elif case == "n2":
        pos = np.array([[0, 0, 0, 1], [1, 1, 1, 1]], np.float32)
    elif case == "n3":
        pos = np.array([[0, 0, 0, 1], [1, 1, 1, 1], [0.5, 0.2, 0.9, 1]], np.float32)
    else:
        pos = np.repeat(ic.two_disk_encounter(400, seed=3)[0], 3, axis=0)
    ks, _ = orc.sort_pairs(orc.morton(pos, bits))
    ch, par = orc.karras(ks, bits)
    tch, tpar = orc.radix_topdown(ks, bits)
    assert np.array_equal(ch, tch)
    assert np.array_equal(par, tpar)
    topology_ok(ch, par, len(pos))


def test_key30_topology_equals_64bit_word_topology():
    """30-bit keys built on a 32-bit word give the same tree as the same keys
    held in 64-bit words (delta shifted by 32 on both branches)."""
    pos = ic.two_disk_encounter(6000, seed=8)[0]
    ks, _ = orc.sort_pairs(orc.morton(pos, 30))
    a = orc.karras(ks, 30)
    b = orc.karras(ks, 63)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_planted_split_error_is_caught():
    """An off-by-one in the split would break the topology pin."""
    pos = ic.two_disk_encounter(3000, seed=2)[0]
    ks, _ = orc.sort_pairs(orc.morton(pos, 63))
    ch, par = orc.karras(ks, 63)
    bad = ch.copy()
    i = int(np.flatnonzero(bad[:, 0] < len(pos) - 1)[0])
    bad[i, 0] += 1
    with pytest.raises(AssertionError):
        tch, _ = orc.radix_topdown(ks, 63)
        assert np.array_equal(bad, tch)
        topology_ok(bad, par, len(pos))


# --------------------------------------------------------------------- COM

def test_com_matches_fp64_and_root_mass():
    pos = blobs_with_dupes()
    t = orc.Tree(pos, 63)
    com, size = numpy_ref.com_f64(t.children, t.pos_sorted[:, :3], t.pos_sorted[:, 3])
    np.testing.assert_allclose(t.com4[:, :3], com[:, :3], rtol=0, atol=2e-6)
    np.testing.assert_allclose(t.com4[:, 3], com[:, 3], rtol=1e-6)
    np.testing.assert_allclose(np.sqrt(t.size2), size, rtol=1e-6, atol=1e-7)
    # root mass = total mass (lbvh_check.py:245-248), root COM = weighted mean
    m = pos[:, 3].astype(np.float64)
    assert abs(t.com4[0, 3] - m.sum()) < 1e-6
    np.testing.assert_allclose(t.com4[0, :3], (pos[:, :3] * m[:, None]).sum(0) / m.sum(),
                               atol=1e-6)


# --------------------------------------------------------------- traversal

def fp64_tree(t):
    com, size = numpy_ref.com_f64(t.children, t.pos_sorted[:, :3].astype(np.float64),
                                  t.pos_sorted[:, 3].astype(np.float64))
    return com, size * size


def test_theta0_walk_is_exact_fp64():
    """lbvh_check.py:250-264: theta=0 touches n leaves and equals the direct
    sum to 1e-12 in FP64."""
    pos = blobs_with_dupes()
    t = orc.Tree(pos, 63)
    com, s2 = fp64_tree(t)
    ps = t.pos_sorted.astype(np.float64)
    acc, inter, depth = orc.traverse_thread_f64(ps, t.children, com, s2, 0.0, 0.05, 0, len(pos))
    assert np.all(inter == len(pos))
    d = ps[None, :, :3] - ps[:, None, :3]
    inv = (np.sum(d * d, axis=2) + 0.05 ** 2) ** -1.5
    np.fill_diagonal(inv, 0.0)
    ref = np.einsum("tn,tnk->tk", ps[None, :, 3] * inv, d)
    err = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 1e-12
    assert depth < 128


def test_theta05_band_and_pruning():
    """lbvh_check.py:266-283: theta=0.5 median error < 0.05, far fewer than n
    interactions, stack depth < 128; FP32 walk vs FP64 all-pairs."""
    pos = blobs_with_dupes()
    t = orc.Tree(pos, 63)
    acc, _, inter, opens, depth = t.traverse_thread(0.5, 0.05)
    ref = orc.allpairs_f64(pos, 0.05)
    err = np.linalg.norm(acc[:, :3] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(err) < 0.05
    assert inter.mean() < 0.5 * len(pos)
    assert depth < 128


def test_group_walk_width1_is_thread_walk():
    pos = ic.two_disk_encounter(3000, seed=5)[0]
    t = orc.Tree(pos, 63)
    a1, _, i1, o1, _ = t.traverse_thread(0.6, 0.05)
    a2, _, i2, o2, _ = t.traverse_group(0.6, 0.05, 1)
    assert np.array_equal(i1, i2) and np.array_equal(o1, o2)
    assert np.array_equal(a1, a2)


@pytest.mark.parametrize("width", [16, 32, 64])
def test_group_walk_dominates_thread_walk(width):
    """A group opens every node any member opens, so each group does at least
    as many interactions as its busiest member, and theta=0 touches n leaves."""
    pos = ic.two_disk_encounter(4000, seed=6)[0]
    t = orc.Tree(pos, 63)
    _, _, ti, _, _ = t.traverse_thread(0.5, 0.05)
    _, _, gi, _, _ = t.traverse_group(0.5, 0.05, width)
    n = len(pos)
    busiest = np.array([ti[g * width:(g + 1) * width].max() for g in range(len(gi))])
    assert np.all(gi >= busiest)
    _, _, g0, _, _ = t.traverse_group(0.0, 0.05, width)
    assert np.all(g0 == n)


def test_warp_walk_lbvh_check_bands():
    """lbvh_check.py:285-316: warp walk theta=0 exact, theta=0.5 median < 0.05."""
    pos = blobs_with_dupes()
    t = orc.Tree(pos, 63)
    a0, _, g0, _, _ = t.traverse_group(0.0, 0.05, 32)
    a5, _, _, _, _ = t.traverse_group(0.5, 0.05, 32)
    ref = orc.allpairs_f64(pos, 0.05)
    e0 = np.linalg.norm(a0[:, :3] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    e5 = np.linalg.norm(a5[:, :3] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.all(g0 == len(pos))
    assert np.median(e0) < 1e-5       # FP32 walk; FP64 exactness is pinned above
    assert np.median(e5) < 0.05


def test_group_walk_fp64_sum():
    """orc_traverse_group_f64sum: the group walk's lists summed in FP64.  At
    theta = 0 (every leaf; eps = 1/16, so the walk's FP32 eps^2 is exact) it
    is the FP64 direct sum to summation-order round-off; at theta 0.5 it
    keeps the FP32 walk's decisions (same counts) and the FP32 walk sits
    within FP32 round-off of it."""
    pos = ic.two_disk_encounter(3000, seed=8)[0]
    t = orc.Tree(pos, 30)
    a0, g0 = t.traverse_group_f64sum(0.0, 0.0625, 32)
    ref = orc.allpairs_f64(pos, 0.0625)
    e0 = np.linalg.norm(a0 - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert e0.max() < 1e-12 and np.all(g0 == len(pos))
    for width in (16, 64):
        a5, g5 = t.traverse_group_f64sum(0.5, 0.05, width)
        a32, _, gi, _, _ = t.traverse_group(0.5, 0.05, width)
        assert np.array_equal(g5, gi)
        e = np.linalg.norm(a32[:, :3] - a5, axis=1) / np.linalg.norm(a5, axis=1)
        assert np.median(e) < 2e-6 and e.max() < 1e-4
        # and it is not the direct sum: the multipole error is still in it
        e_ref = np.linalg.norm(a5 - orc.allpairs_f64(pos, 0.05), axis=1)
        assert np.median(e_ref / np.linalg.norm(a5, axis=1)) > 1e-4


# ---------------------------------------------------------- force & energy

def test_two_body_closed_form():
    """a = G m r / (r^2 + eps^2)^1.5 along the separation (forces.cu:12-15)."""
    eps, r, m1, m2 = 0.05, 1.7, 0.3, 0.9
    pos = np.array([[0, 0, 0, m1], [r, 0, 0, m2]], np.float32)
    acc = orc.allpairs_f64(pos, eps, G=2.0)
    r, m1, m2 = (float(np.float32(x)) for x in (r, m1, m2))  # the stored FP32 values
    ref = 2.0 * r / (r * r + eps * eps) ** 1.5
    assert abs(acc[0, 0] - m2 * ref) < 1e-14 and abs(acc[1, 0] + m1 * ref) < 1e-14
    assert np.all(acc[:, 1:] == 0)
    a32 = orc.allpairs_f32(pos, eps, G=2.0)
    assert abs(a32[0, 0] / (m2 * ref) - 1) < 1e-6


def test_newton_third_law():
    pos = ic.two_disk_encounter(2000, seed=1)[0]
    a = orc.allpairs_f64(pos, 0.05)
    f = a * pos[:, 3:4].astype(np.float64)
    assert np.abs(f.sum(axis=0)).max() < 1e-15 * len(pos)


def test_allpairs_fp32_vs_fp64():
    pos = ic.two_disk_encounter(3000, seed=2)[0]
    a32 = orc.allpairs_f32(pos, 0.05)
    a64 = orc.allpairs_f64(pos, 0.05)
    e = np.linalg.norm(a32[:, :3] - a64, axis=1) / np.linalg.norm(a64, axis=1)
    assert np.median(e) < 1e-6


def test_numpy_baseline_matches_oracle():
    """The CPU baseline (reference_nbody.py:45-61 restated) agrees with FP64."""
    pos = ic.two_disk_encounter(3000, seed=3)[0]
    a = numpy_ref.accelerations(pos[:, :3], pos[:, 3], np.float32(0.05 ** 2), np.float32(1.0),
                                chunk=512)
    a64 = orc.allpairs_f64(pos, 0.05)
    e = np.linalg.norm(a - a64, axis=1) / np.linalg.norm(a64, axis=1)
    assert np.median(e) < 1e-5


def test_energy_two_body_closed_form():
    eps, r = 0.05, 2.0
    pos = np.array([[0, 0, 0, 0.5], [r, 0, 0, 0.25]], np.float32)
    vel = np.array([[0, 1, 0, 0], [0, -2, 0, 0]], np.float32)
    e = orc.energy_f64(pos, vel, eps)
    ke = 0.5 * 0.5 * 1 + 0.5 * 0.25 * 4
    pe = -0.5 * 0.25 / np.sqrt(r * r + eps * eps)
    np.testing.assert_allclose(e, [ke + pe, ke, pe], rtol=1e-14)
    np.testing.assert_allclose(
        numpy_ref.total_energy(pos[:, :3], vel[:, :3], pos[:, 3], eps * eps, 1.0), e, rtol=1e-12)


def test_reference_trace_initial_energy():
    """benchmarks/energy_gpu_naive.csv (tests/golden) row 0: E = -1.17839549e-01
    with rel_error 5.740e-08 gives E0 = -0.11783955576; the restated
    generate_ic.py IC (12,000 particles, seed 0) has the same energy to FP32
    round-off."""
    from nbodyssey_amd import formats
    import os
    log = formats.read_energy_csv(os.path.join(os.path.dirname(__file__), "golden",
                                               "energy_gpu_naive.csv"))
    pos, vel = ic.generate_ic_disks(particles=12000, seed=0)
    e = orc.energy_f64(pos, vel, 0.05)
    e0_ref = log["energy"][0] / (1 - log["rel_error"][0])
    assert abs(e0_ref / -0.11783955576 - 1) < 1e-9
    assert abs(e[0] / e0_ref - 1) < 1e-6
    # a wrong draw order (e.g. disks swapped) changes the energy by far more
    pos_b, vel_b = ic.generate_ic_disks(particles=12000, seed=1)
    assert abs(orc.energy_f64(pos_b, vel_b, 0.05)[0] / e0_ref - 1) > 1e-4


def test_kick_drift_closed_form():
    pos = np.array([[1, 2, 3, 1]], np.float32)
    vel = np.array([[0.5, 0, -1, 0]], np.float32)
    acc = np.array([[2, -4, 0, 0]], np.float32)
    orc.lib().orc_kick(vel, acc, 1, np.float32(0.05))
    np.testing.assert_allclose(vel[0, :3], [0.6, -0.2, -1.0], rtol=1e-6)
    orc.lib().orc_drift(pos, vel, 1, np.float32(0.1))
    np.testing.assert_allclose(pos[0, :3], [1.06, 1.98, 2.9], rtol=1e-6)
    assert pos[0, 3] == 1.0


@pytest.mark.slow
def test_oracle_energy_drift_two_body_orbit():
    """A circular two-body orbit integrated by the oracle's KDK conserves
    energy to O(dt^2) over many periods (symplectic)."""
    m, r = 0.5, 1.0
    eps = 1e-3
    v = np.sqrt(2 * m * r * r / (r * r + eps * eps) ** 1.5 / 2) * 0.5 ** 0.5 * 2 ** 0.5 / 2 ** 0.5
    pos = np.array([[-r / 2, 0, 0, m], [r / 2, 0, 0, m]], np.float32)
    vel = np.array([[0, -v, 0, 0], [0, v, 0, 0]], np.float32)
    acc = orc.allpairs_f32(pos, eps)
    e0 = orc.energy_f64(pos, vel, eps)[0]
    worst = 0.0
    for _ in range(2000):
        orc.kdk_step_naive(pos, vel, acc, 0.01, eps)
        worst = max(worst, abs(orc.energy_f64(pos, vel, eps)[0] / e0 - 1))
    assert worst < 1e-3
