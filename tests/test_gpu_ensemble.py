"""GPU parity of the batched ensemble (nb_create_ensemble).

The forest build must give every system exactly its standalone tree (keys,
order, topology, node values bit-exact against the oracle), and the
persistent walk over (system, 64-target block) items must reproduce the
standalone wave64 walk: interaction/opening counts equal to the oracle's
64-wide group walk, accelerations and steps bitwise equal to running each
system alone on the plain path.
"""
import numpy as np
import pytest

from nbodyssey_amd import NbError, engine, ic
from oracle import orc

pytestmark = pytest.mark.gpu

EPS = 0.05


def mixed_systems():
    """Systems of awkward sizes and shapes: 2, 3, 64, 65 particles, coincident
    points, duplicate-heavy blobs, two-disk encounters."""
    rng = np.random.default_rng(3)

    def cloud(n, seed):
        r = np.random.default_rng(seed)
        p = r.normal(0.0, 1.0, size=(n, 3))
        return np.column_stack([p, np.full(n, 1.0 / n)]).astype(np.float32)

    same = np.tile(np.array([[0.3, 0.3, 0.3, 0.01]], np.float32), (777, 1))
    blobs = np.concatenate([cloud(1500, 1), cloud(1500, 2) + np.float32([4, 0.5, 0.1, 0]),
                            np.tile(np.float32([[0.123, -0.456, 0.789, 1.0 / 3024]]), (24, 1))])
    out = [cloud(2, 5), cloud(3, 6), cloud(64, 7), cloud(65, 8), same, blobs,
           ic.two_disk_encounter(5000, seed=7)[0], cloud(130, 9),
           ic.parabolic_encounter(4096, 11, 0.5, 45.0, 18.0, 2.0)[0]]
    order = rng.permutation(len(out))
    systems = []
    for k in order:
        p = out[k]
        v = rng.normal(0.0, 0.1, size=(len(p), 4)).astype(np.float32)
        v[:, 3] = 0
        systems.append((p, v))
    return systems


def local_to_forest(c, m, o, n, root):
    """Standalone node id c (m particles) -> forest id (system at offset o,
    n particles in the ensemble, subtree root `root`)."""
    c = np.asarray(c, np.int64)
    return np.where(c >= m - 1, (n - 1) + o + (c - (m - 1)), np.where(c == 0, root, o + c))


def subtree_matches(t, off, n, roots, s, m, ref):
    """System s (m particles) of forest tree t equals the standalone tree ref
    (dict with keys, perm, pos_sorted, children, com4, size2, parent)."""
    o = int(off[s])
    k = t["keys"][o:o + m]
    assert np.all((k >> np.uint64(30)) == s), s
    assert np.array_equal(k & np.uint64((1 << 30) - 1), ref["keys"]), s
    assert np.array_equal(t["perm"][o:o + m].astype(np.int64) - o, ref["perm"]), s
    assert np.array_equal(t["pos_sorted"][o:o + m], ref["pos_sorted"]), s
    g = local_to_forest(np.arange(m - 1), m, o, n, roots[s])
    assert np.array_equal(t["children"][g], local_to_forest(ref["children"], m, o, n, roots[s])), s
    assert np.array_equal(t["com4"][g], ref["com4"]), s
    assert np.array_equal(t["size2"][g], ref["size2"]), s
    # every node of the subtree but its root has its standalone parent
    nodes = np.arange(1, 2 * m - 1)
    assert np.array_equal(t["parent"][local_to_forest(nodes, m, o, n, roots[s])],
                          local_to_forest(ref["parent"][1:], m, o, n, roots[s])), s


def oracle_tree(pos):
    r = orc.Tree(pos, key_bits=30)
    return dict(keys=r.keys_sorted, perm=r.perm, pos_sorted=r.pos_sorted, children=r.children,
                com4=r.com4, size2=r.size2, parent=r.parent)


def test_forest_top_pass_shares():
    """Forests above one 32,768-leaf superblock run the top COM pass as up to
    16 workgroups, each owning whole systems.  24 systems of 3,000-12,000
    particles (about 180K, six superblocks; each workgroup stages its share)
    against the oracle's trees, then two systems of 2.4M particles, whose
    shares each list more nodes than one workgroup stages (the unstaged
    climb), against the standalone build of each system."""
    rng = np.random.default_rng(17)
    sizes = rng.integers(3000, 12001, size=24)
    systems = [(ic.two_disk_encounter(int(m), seed=100 + k)[0], None) for k, m in enumerate(sizes)]
    systems = [(p, np.zeros_like(p)) for p, _ in systems]
    with engine.Ensemble(systems, eps=EPS, theta=0.5) as ens:
        ens.compute_forces()
        t, roots, off = ens.tree(), ens.roots(), ens.offsets
    n = int(off[-1])
    print(f"forest n={n} systems={len(systems)}", flush=True)
    for s, (pos, _) in enumerate(systems):
        subtree_matches(t, off, n, roots, s, len(pos), oracle_tree(pos))

    big = [ic.two_disk_encounter(2_400_000, seed=21 + k)[0] for k in range(2)]
    with engine.Ensemble([(p, np.zeros_like(p)) for p in big], eps=EPS, theta=0.5) as ens:
        ens.compute_forces()
        t, roots, off = ens.tree(), ens.roots(), ens.offsets
    n = int(off[-1])
    print(f"forest n={n} systems=2", flush=True)
    for s, pos in enumerate(big):
        with engine.Simulation(pos, eps=EPS, theta=0.5, force="bh", traverse="warp",
                               group_width=64, key_bits=30) as sim:
            sim.compute_forces()
            ref = sim.tree()
        subtree_matches(t, off, n, roots, s, len(pos), ref)


def test_forest_matches_standalone_trees():
    systems = mixed_systems()
    with engine.Ensemble(systems, eps=EPS, theta=0.5) as ens:
        ens.compute_forces()
        t = ens.tree()
        roots = ens.roots()
        off = ens.offsets
    n = int(off[-1])
    print(f"forest n={n} systems={len(systems)} roots={roots.tolist()}", flush=True)
    for s, (pos, _) in enumerate(systems):
        o, m = int(off[s]), len(pos)
        ref = orc.Tree(pos, key_bits=30)
        k = t["keys"][o:o + m]
        assert np.all((k >> np.uint64(30)) == s)
        assert np.array_equal(k & np.uint64((1 << 30) - 1), ref.keys_sorted)
        assert np.array_equal(t["perm"][o:o + m].astype(np.int64) - o, ref.perm)
        assert np.array_equal(t["pos_sorted"][o:o + m], ref.pos_sorted)
        g = local_to_forest(np.arange(m - 1), m, o, n, roots[s])
        assert np.array_equal(t["children"][g], local_to_forest(ref.children, m, o, n, roots[s]))
        assert np.array_equal(t["com4"][g], ref.com4)
        assert np.array_equal(t["size2"][g], ref.size2)
        # every node of the subtree but its root has its standalone parent
        nodes = np.arange(1, 2 * m - 1)
        assert np.array_equal(t["parent"][local_to_forest(nodes, m, o, n, roots[s])],
                              local_to_forest(ref.parent[1:], m, o, n, roots[s]))


@pytest.mark.parametrize("theta", [0.0, 0.5, 0.8])
def test_item_walk_counts_and_forces(theta):
    """Per-item counts equal the oracle's 64-wide group walk of each system
    alone; accelerations are bitwise those of the standalone wave64 walk."""
    systems = mixed_systems()
    with engine.Ensemble(systems, eps=EPS, theta=theta) as ens:
        acc = ens.compute_forces()
        inter, opens = ens.traverse_counts()
        items = ens.items()
        off = ens.offsets
    print(f"ensemble walk theta={theta} items={len(items)}", flush=True)
    assert np.array_equal(np.sort(items[:, 1]), np.sort(np.concatenate(
        [np.arange(off[s], off[s + 1], 64) for s in range(len(systems))])))
    for s, (pos, _) in enumerate(systems):
        o = int(off[s])
        ref = orc.Tree(pos, key_bits=30)
        _, _, ri, ro, _ = ref.traverse_group(theta, EPS, 64)
        mine = items[:, 0] == s
        grp = (items[mine, 1] - o) // 64
        assert np.array_equal(inter[mine], ri[grp])
        assert np.array_equal(opens[mine], ro[grp])
        with engine.Simulation(pos, eps=EPS, theta=theta, force="bh", traverse="warp",
                               group_width=64, key_bits=30) as sim:
            a1 = sim.compute_forces()
        assert np.array_equal(acc[o:o + len(pos)], a1), f"system {s} forces differ"


def test_steps_and_energy_match_standalone():
    """Three KDK steps of a 24-system ensemble (sizes 2,048..16,384) equal the
    systems stepped one at a time, bitwise; per-system FP64 energies match
    the oracle (rtol 2e-6) and nb_total_energy is their sum."""
    systems, _ = ic.ensemble_encounters(24, seed=100, n_min=2048, n_max=16384)
    kw = dict(dt=0.01, eps=EPS, theta=0.5)
    with engine.Ensemble(systems, **kw) as ens:
        e_sys = ens.system_energy()
        e_tot = ens.total_energy()
        ens.step(3)
        p_e, v_e = ens.download()
        off = ens.offsets
    np.testing.assert_allclose(e_tot, e_sys.sum(axis=0), rtol=1e-12)
    for s, (pos, vel) in enumerate(systems):
        r = orc.energy_f64(pos, vel, EPS)
        np.testing.assert_allclose(e_sys[s], r, rtol=2e-6, atol=1e-12)
        with engine.Simulation(pos, vel, force="bh", traverse="warp", group_width=64,
                               key_bits=30, **kw) as sim:
            sim.step(3)
            p1, v1 = sim.download()
        o, m = int(off[s]), len(pos)
        assert np.array_equal(p_e[o:o + m], p1) and np.array_equal(v_e[o:o + m], v1), s


def test_ensemble_graph_matches_eager():
    """Ten graph replays of the ensemble step, with eager work (downloads,
    per-system energies) between them, equal the eager steps bitwise."""
    systems, _ = ic.ensemble_encounters(16, seed=200, n_min=2048, n_max=8192)
    kw = dict(dt=0.01, eps=EPS, theta=0.5)
    with engine.Ensemble(systems, **kw) as eager, engine.Ensemble(systems, **kw) as graph:
        graph.set_graph(True)
        for r in range(10):
            print(f"ensemble graph replay {r}", flush=True)
            eager.step(1)
            graph.step(1)
            pe, ve = eager.download()
            pg, vg = graph.download()
            assert np.array_equal(pe, pg) and np.array_equal(ve, vg), f"replay {r} differs"
            assert np.array_equal(eager.system_energy(), graph.system_energy())


def test_ensemble_parameter_checks():
    systems = [(np.zeros((4, 4), np.float32) + np.float32([[0], [1], [2], [3]]), None)] * 2
    with engine.Ensemble(systems) as ens:
        with pytest.raises(NbError):
            ens.update(force="naive")
        with pytest.raises(NbError):
            ens.update(group_width=32)
        ens.update(theta=0.7, compensated=True)
        ens.compute_forces()
    with pytest.raises(NbError):
        engine.Ensemble([(np.zeros((1, 4), np.float32), None)])


def test_direct_f64_per_system():
    """In an ensemble the FP64 reference sums over the target's own system
    only: equal to the oracle on that system alone."""
    systems = mixed_systems()
    eps = float(np.float32(EPS))
    with engine.Ensemble(systems, eps=eps) as ens:
        off = ens.offsets
        idx = np.concatenate([np.arange(off[s], min(off[s + 1], off[s] + 40))
                              for s in range(len(systems))])[::-1].astype(np.int32)
        got = ens.direct_f64(idx)
    for s, (pos, _) in enumerate(systems):
        mine = (idx >= off[s]) & (idx < off[s + 1])
        ref = orc.allpairs_f64(pos, eps, targets=(idx[mine] - off[s]).astype(np.int32))
        np.testing.assert_allclose(got[mine], ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_ensemble_far16():
    """FP16 far field in the persistent ensemble walk: same per-system result
    as the standalone FP16 walk, bitwise."""
    systems, _ = ic.ensemble_encounters(8, seed=400, n_min=2048, n_max=8192)
    with engine.Ensemble(systems, eps=EPS, theta=0.5, far16=True) as ens:
        acc = ens.compute_forces()
        off = ens.offsets
    for s, (pos, _) in enumerate(systems):
        with engine.Simulation(pos, eps=EPS, theta=0.5, force="bh", traverse="warp",
                               group_width=64, key_bits=30, far16=True) as sim:
            assert np.array_equal(acc[off[s]:off[s + 1]], sim.compute_forces()), s


def test_ensemble_256_systems_queue():
    """The benchmark ensemble at full size (256 parabolic encounters, 2,048 to
    65,536 particles each, 5.4M in all): the persistent walk drains the whole
    item queue (every item records a non-empty walk), six graph-replayed
    steps equal six eager steps bitwise, and the forces of the smallest, the
    largest and six random systems equal their standalone walks bitwise."""
    w = ic.WORKLOADS["ensemble-256-parabolic"]
    systems, _ = ic.ensemble_encounters(256, seed=w["seed"])
    print(f"ensemble 256: {sum(len(p) for p, _ in systems)} particles", flush=True)
    kw = dict(dt=w["dt"], eps=w["eps"], theta=w["theta"])
    sizes = np.array([len(p) for p, _ in systems])
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    pick |= set(np.random.default_rng(9).choice(256, 6, replace=False).tolist())
    with engine.Ensemble(systems, **kw) as ens:
        acc = ens.compute_forces()
        inter, opens = ens.traverse_counts()
        assert len(inter) == len(ens.items()) and np.all(inter > 0) and np.all(opens > 0)
        off = ens.offsets
        for s in sorted(pick):
            pos, _ = systems[s]
            with engine.Simulation(pos, eps=w["eps"], theta=w["theta"], force="bh",
                                   traverse="warp", group_width=64, key_bits=30) as sim:
                a1 = sim.compute_forces()
            o = int(off[s])
            assert np.array_equal(acc[o:o + len(pos)], a1), f"system {s} forces differ"
    with engine.Ensemble(systems, **kw) as eager, engine.Ensemble(systems, **kw) as graph:
        graph.set_graph(True)
        for r in range(6):
            print(f"ensemble 256 step {r}", flush=True)
            eager.step(1)
            graph.step(1)
            pe, ve = eager.download()
            pg, vg = graph.download()
            assert np.array_equal(pe, pg) and np.array_equal(ve, vg), f"step {r} differs"
