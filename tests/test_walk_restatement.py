"""Independent pin of the oracle's walk counts at theta > 0 -- CPU only.

scripts/lbvh_check.py:151-209 defines the two walks the reference's kernels
follow: `traverse` (per target, forces_bh.cu:335-377) and `traverse_warp`
(32 lanes share one stack, a node opens if any live lane opens it,
forces_bh.cu:395-447).  Below they are restated in plain Python/NumPy (FP64
distances from the FP32 tree data, group width W instead of 32) and run on
the oracle's own tree; their interaction and opening-test counts must equal
orc_traverse_thread / orc_traverse_group exactly.

The two sides evaluate d2 differently (FP64 here, the canonical FP32 order in
the oracle), so a walk that meets an opening test within 1e-6 relative of its
threshold could legitimately decide it differently; such walks are skipped
(and counted, so the test fails if too many are skipped).
"""
import numpy as np
import pytest

from nbodyssey_amd import ic
from oracle import orc

NEAR = 1e-6


def blobs_with_dupes():
    """scripts/lbvh_check.py:220-229 input (FP32-rounded)."""
    rng = np.random.default_rng(1)
    a = rng.normal([-2.0, 0.0, 0.0], 0.7, size=(1500, 3))
    b = rng.normal([2.0, 0.5, 0.1], 0.7, size=(1500, 3))
    d = np.tile([[0.123, -0.456, 0.789]], (24, 1))
    p = np.concatenate([a, b, d])
    return np.column_stack([p, np.full(len(p), 1.0 / len(p))]).astype(np.float32)


def inputs():
    return {
        "lbvh_check3024": blobs_with_dupes(),
        "two_disk2000": ic.two_disk_encounter(2000, seed=9)[0],
    }


class TreeData:
    """The oracle's tree as plain Python/NumPy data for the restated walks."""

    def __init__(self, pos4, key_bits=63):
        t = orc.Tree(pos4, key_bits)
        self.tree = t
        self.n = t.n
        self.pos = t.pos_sorted[:, :3].astype(np.float64)
        self.com = t.com4[:, :3].astype(np.float64)
        self.size2 = t.size2.astype(np.float64)
        self.children = [(int(a), int(b)) for a, b in t.children]


def walk_target(td, t, theta2):
    """lbvh_check.traverse: one target's DFS; (interactions, opens, near_tie)."""
    n1 = td.n - 1
    b = td.pos[t]
    stack = [0]
    inter = opens = 0
    near = False
    while stack:
        node = stack.pop()
        if node >= n1:
            inter += 1
            continue
        opens += 1
        d = td.com[node] - b
        rhs = theta2 * float(d @ d)
        s2 = td.size2[node]
        if abs(s2 - rhs) <= NEAR * rhs:
            near = True
        if s2 < rhs:
            inter += 1
        else:
            stack.extend(td.children[node])  # left, then right (popped first)
    return inter, opens, near


def walk_group(td, base, width, theta2):
    """lbvh_check.traverse_warp over `width` lanes: (interactions, opens, near)."""
    n1 = td.n - 1
    lanes = td.pos[base:base + width]
    stack = [0]
    inter = opens = 0
    near = False
    while stack:
        node = stack.pop()
        if node >= n1:
            inter += 1
            continue
        opens += 1
        d = td.com[node] - lanes
        rhs = theta2 * np.einsum("ij,ij->i", d, d)
        s2 = td.size2[node]
        if np.any(np.abs(s2 - rhs) <= NEAR * rhs):
            near = True
        if np.any(s2 >= rhs):  # any live lane votes to open
            stack.extend(td.children[node])
        else:
            inter += 1
    return inter, opens, near


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.8])
@pytest.mark.parametrize("case", ["lbvh_check3024", "two_disk2000"])
def test_thread_walk_counts_match_restatement(case, theta):
    td = TreeData(inputs()[case])
    theta2 = float(np.float32(theta) * np.float32(theta))  # FP32 theta^2, as the walks use
    _, _, inter, opens, _ = td.tree.traverse_thread(theta, 0.05)
    targets = np.random.default_rng(17).choice(td.n, 240, replace=False)
    checked = 0
    for t in targets:
        ni, no, near = walk_target(td, int(t), theta2)
        if near:
            continue
        assert (ni, no) == (inter[t], opens[t]), (case, theta, int(t))
        checked += 1
    assert checked >= 0.9 * len(targets), checked
    assert 0 < inter[targets].mean() < td.n  # theta > 0 prunes


@pytest.mark.parametrize("width", [1, 16, 32, 64])
@pytest.mark.parametrize("theta", [0.3, 0.5, 0.8])
def test_group_walk_counts_match_restatement(width, theta):
    td = TreeData(inputs()["lbvh_check3024"])
    theta2 = float(np.float32(theta) * np.float32(theta))
    _, _, gi, go, _ = td.tree.traverse_group(theta, 0.05, width)
    ngroups = len(gi)
    groups = np.random.default_rng(width * 10 + 3).choice(ngroups, min(ngroups, 16),
                                                        replace=False)
    groups = np.union1d(groups, [ngroups - 1])  # the ragged last group too
    checked = 0
    for g in groups:
        ni, no, near = walk_group(td, int(g) * width, width, theta2)
        if near:
            continue
        assert (ni, no) == (gi[g], go[g]), (theta, width, int(g))
        checked += 1
    assert checked >= 0.8 * len(groups), checked


def test_restatement_catches_a_planted_mistake():
    """Walking the children in the wrong order changes nothing in the counts
    (they are order-free), but a wrong opening test -- the cell edge size
    instead of size^2, as a plausible slip -- must change them."""
    td = TreeData(inputs()["lbvh_check3024"])
    theta2 = float(np.float32(0.5) ** 2)
    _, _, inter, opens, _ = td.tree.traverse_thread(0.5, 0.05)
    saved = td.size2
    td.size2 = np.sqrt(saved)
    diffs = sum(walk_target(td, t, theta2)[:2] != (inter[t], opens[t]) for t in range(0, 3000, 97))
    td.size2 = saved
    assert diffs > 20
