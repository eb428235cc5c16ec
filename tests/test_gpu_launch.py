"""GPU tests of the launch configuration (nb_set_launch / nb_get_launch) and
of the autotuner's launch candidates: all-pairs LDS tile width and the wave64
walk's LDS stack depth change speed, not the physics."""
import numpy as np
import pytest

from nbodyssey_amd import _lib, engine, ic
from oracle import orc

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = np.asarray(a, np.float64)[:, :3]
    b = np.asarray(b, np.float64)[:, :3]
    return np.linalg.norm(a - b, axis=1) / (np.linalg.norm(b, axis=1) + 1e-300)


def test_allpairs_tile_widths():
    """One source partition: the plain FP32 sums are the same bits at tile
    widths 256, 512 and 1024 (the sources are added in the same order; only
    the LDS staging differs), and agree with the oracle's FP32 all-pairs.
    With 4 partitions the partial sums are cut at tile multiples, so widths
    agree to FP32 order effects; compensated sums at every width stay at
    least twice as accurate as plain FP32 against FP64."""
    pos, _ = ic.two_disk_encounter(20011, seed=71, eps=0.05)
    ref = orc.allpairs_f32(pos, 0.05)
    out = {}
    with engine.Simulation(pos, eps=0.05, force="naive") as sim:
        for split in (1, 4):
            for tile in (256, 512, 1024):
                print(f"allpairs split={split} tile={tile}", flush=True)
                sim.set_launch(allpairs_split=split, allpairs_tile=tile)
                assert sim.launch()["allpairs_tile"] == tile
                out[split, tile] = sim.compute_forces()
        sim.update(compensated=True)
        comp = {}
        for tile in (256, 1024):
            sim.set_launch(allpairs_split=1, allpairs_tile=tile)
            comp[tile] = sim.compute_forces()
    for tile in (512, 1024):
        assert np.array_equal(out[1, tile], out[1, 256]), tile
    for tile in (256, 512, 1024):
        # partition sums against one running sum: FP32 order effects (the
        # bound of the other all-pairs order tests)
        e = rel_err(out[4, tile], out[1, 256])
        assert np.median(e) < 5e-6 and e.max() < 1e-4, (tile, np.median(e), e.max())
    e = rel_err(out[1, 256], ref)
    assert np.median(e) < 5e-6 and e.max() < 1e-4
    # compensated: each tile's FP32 partial is Kahan-added, so the widths
    # differ by the partials' rounding; against FP64 both stay at least twice
    # as accurate as the plain sum
    idx = np.random.default_rng(7).choice(len(pos), 512, replace=False).astype(np.int32)
    r64 = orc.allpairs_f64(pos, 0.05, targets=idx)
    plain = np.median(rel_err(out[1, 256][idx], r64))
    for tile, a in comp.items():
        ec = np.median(rel_err(a[idx], r64))
        print(f"fp64 error: plain {plain:.3e}, compensated tile {tile} {ec:.3e}", flush=True)
        assert ec <= 0.5 * plain, (tile, ec, plain)


@pytest.mark.parametrize("theta", [0.5, 0.2])
def test_walk_stack_depths(theta):
    """Stack depths 512, 768, 1024 and 2048 entries per wave: the group
    walk's interaction and opening counts equal the oracle's wave64 vote at
    every depth, and the accelerations agree with the default depth to FP32
    order effects.  At theta 0.2 the 512-entry stack reaches its reserve and
    pops single nodes (measured: 993 of the 1024 groups stack more than the
    192 entries below the reserve), the same node sets in another order."""
    pos, _ = ic.two_disk_encounter(65536, seed=72, eps=0.02)
    tree = orc.Tree(pos, key_bits=30)
    _, _, gi, go, _ = tree.traverse_group(theta, 0.02, 64)
    acc = {}
    with engine.Simulation(pos, eps=0.02, theta=theta, force="bh", traverse="warp",
                           group_width=64, key_bits=30) as sim:
        for stack in (1024, 512, 768, 2048):
            print(f"walk theta={theta} stack={stack}", flush=True)
            sim.set_launch(walk_stack=stack)
            acc[stack] = sim.compute_forces()
            ci, co = sim.traverse_counts("warp", 64)
            assert np.array_equal(ci, gi) and np.array_equal(co, go), stack
    for stack in (512, 768, 2048):
        e = rel_err(acc[stack], acc[1024])
        assert np.median(e) < 1e-6 and e.max() < 1e-4, (stack, np.median(e), e.max())


def test_walk_stack_step_graph():
    """A leapfrog step captured in a HIP graph with a non-default stack depth
    replays to the eager steps' bits, and changing the depth re-captures."""
    pos, vel = ic.two_disk_encounter(30000, seed=73, eps=0.02)
    kw = dict(dt=0.005, eps=0.02, theta=0.5, force="bh", traverse="warp", group_width=64,
              key_bits=30)
    with engine.Simulation(pos, vel, **kw) as eager, engine.Simulation(pos, vel, **kw) as gr:
        eager.set_launch(walk_stack=768)
        gr.set_launch(walk_stack=768)
        gr.set_graph(True)
        for k in range(6):
            eager.step(1)
            gr.step(1)
            pe, ve = eager.download()
            pg, vg = gr.download()
            assert np.array_equal(pe, pg) and np.array_equal(ve, vg), k
        gr.set_launch(walk_stack=2048)
        eager.set_launch(walk_stack=2048)
        gr.step(2)
        eager.step(2)
        assert np.array_equal(gr.download()[0], eager.download()[0])


def test_set_launch_validation():
    pos, _ = ic.two_disk_encounter(4096, seed=74, eps=0.05)
    with engine.Simulation(pos, eps=0.05, force="naive") as sim:
        assert sim.launch() == {"allpairs_split": 0, "allpairs_tile": 0, "walk_stack": 0,
                                "walk_grid": 0}
        for bad in (dict(allpairs_tile=128), dict(allpairs_tile=300), dict(walk_stack=256),
                    dict(walk_stack=1000), dict(allpairs_split=65), dict(allpairs_split=-1),
                    dict(walk_grid=-2)):
            with pytest.raises(_lib.NbError):
                sim.set_launch(**bad)
        # a refused setting leaves the configuration as it was
        assert sim.launch()["allpairs_tile"] == 0
        with pytest.raises(ValueError):
            sim.set_launch(tile=256)
        sim.set_launch(allpairs_tile=1024, walk_stack=768)
        assert sim.launch()["allpairs_tile"] == 1024 and sim.launch()["walk_stack"] == 768


@pytest.mark.parametrize("force", ["bh", "naive"])
def test_autotune_launch_candidates(force):
    """The autotuner's log covers the tile widths (all-pairs) and the stack
    depths (group walk); the choice is set on the engine (nb_get_launch)."""
    pos, vel = ic.two_disk_encounter(8192, seed=75, eps=0.05)
    with engine.Simulation(pos, vel, dt=0.01, eps=0.05, theta=0.5, force=force,
                           traverse="warp", group_width=64, key_bits=30) as sim:
        choice, log = sim.autotune(steps=1, max_median_err=0.01)
        print(choice, *log, sep="\n", flush=True)
        got = sim.launch()
    if force == "naive":
        assert {f"tile={t}" for t in (256, 512, 1024)} <= {w for l in log for w in l.split()}
        assert choice["allpairs_tile"] in (256, 512, 1024)
        assert got["allpairs_tile"] == choice["allpairs_tile"]
        assert got["allpairs_split"] == choice["allpairs_split"]
    else:
        assert {f"stack={s}" for s in (512, 768, 1024, 2048)} <= {w for l in log for w in l.split()}
        assert choice["walk_stack"] in (512, 768, 1024, 2048)
        assert got["walk_stack"] == choice["walk_stack"]
    assert choice["median_err"] <= 0.01
