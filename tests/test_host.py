"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared entry point, the file formats, the galaxy_sim command line's
error behaviour (main.cu:28-70,114-128), and the IC generators."""
import os
import re
import subprocess

import numpy as np
import pytest

from nbodyssey_amd import GALAXY_SIM, LIB_PATH, formats, ic
from nbodyssey_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nbody.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["nb_create", "nb_destroy", "nb_set_params", "nb_upload", "nb_download",
                 "nb_compute_forces", "nb_step", "nb_total_energy", "nb_force_ms_total",
                 "nb_force_calls", "nb_reset_force_timing", "nb_bh_phase_ms", "nb_bh_release",
                 "nb_morton", "nb_sort_pairs", "nb_karras_build", "nb_get_tree",
                 "nb_traverse_counts"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "build libnbody.so first (make -C nbodyssey_amd/csrc)"
    L = _lib.load()
    for name in declared_functions():
        assert hasattr(L, name), f"{name} declared in nbody.h but not exported"
    # and the ctypes binding covers exactly the declared set
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_exports_have_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for name in declared_functions():
        assert name in exported, f"{name} is not an unmangled dynamic symbol"


def test_last_error_is_empty_string_initially():
    assert isinstance(_lib.load().nb_last_error(), bytes)


def test_default_params_match_reference():
    """main.cu:19-25, simulation.hpp:11-14."""
    p = _lib.NbParams()
    _lib.load().nb_default_params(p)
    assert (p.dt, p.epsilon, p.G, p.theta) == (np.float32(0.01), np.float32(0.05), 1.0, 0.5)
    assert (p.force, p.traverse, p.key_bits, p.group_width) == (0, 0, 63, 32)


# ------------------------------------------------------------------ formats

def test_ic_roundtrip(tmp_path):
    pos, vel = ic.two_disk_encounter(1000, seed=1)
    path = tmp_path / "ic.bin"
    formats.write_ic(path, pos, vel)
    assert os.path.getsize(path) == 4 + 32 * 1000  # FORMATS.md:11-16
    p2, v2 = formats.read_ic(path)
    assert np.array_equal(p2, pos) and np.array_equal(v2, vel)


def test_frame_roundtrip(tmp_path):
    pos, _ = ic.two_disk_encounter(100, seed=1)
    path = formats.frame_path(str(tmp_path), 7)
    assert path.endswith("frame_00007.bin")
    formats.write_frame(path, pos)
    assert os.path.getsize(path) == 4 + 16 * 100
    assert np.array_equal(formats.read_frame(path), pos)


@pytest.mark.parametrize("payload,msg", [
    (b"", "bad particle count"),
    (np.int32(0).tobytes(), "bad particle count"),
    (np.int32(-3).tobytes(), "bad particle count"),
    (np.int32(5).tobytes() + b"\0" * 40, "truncated IC file"),
])
def test_read_ic_errors(tmp_path, payload, msg):
    path = tmp_path / "bad.bin"
    path.write_bytes(payload)
    with pytest.raises(formats.FormatError, match=msg):
        formats.read_ic(path)


def test_energy_row_format():
    """main.cu:174-177: %d,%.5f,%.8e,%.8e,%.8e,%.3e; rel = (E - E0)/|E0|."""
    row = formats.energy_row(100, 0.01, (-0.117839655, 0.601599732, -0.719439386),
                             (-0.1178395557, 0, 0))
    assert row == "100,1.00000,-1.17839655e-01,6.01599732e-01,-7.19439386e-01,-8.427e-07\n"


# ---------------------------------------------- the reference's committed data

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ENERGY_LOGS = ["energy_gpu_naive.csv", "energy_gpu_bh_thread.csv", "energy_gpu_bh_warp.csv"]


def golden_e0(log):
    """E0 from row 0 of an energy log: rel = (E - E0)/|E0| with E0 < 0."""
    return log["energy"][0] / (1.0 - log["rel_error"][0])


@pytest.mark.parametrize("name", ENERGY_LOGS)
def test_reference_energy_logs_reformat(name):
    """The reference's 12k logs (benchmarks/, gpu_bench.sh:64-71): rows after
    steps 0, 100, ..., 1500 (main.cu:186-192), and our row writer reproduces
    every row from its parsed values; rel_error agrees to the rounding of the
    printed energies and rel_error's last printed digit."""
    path = os.path.join(GOLDEN, name)
    log = formats.read_energy_csv(path)
    assert list(log["step"]) == list(range(0, 1501, 100))
    e0 = golden_e0(log)
    lines = open(path).read().splitlines()[1:]
    for k, line in enumerate(lines):
        mine = formats.energy_row(int(log["step"][k]), 0.01,
                                  (log["energy"][k], log["kinetic"][k], log["potential"][k]),
                                  (e0, 0.0, 0.0))
        assert mine.rsplit(",", 1)[0] == line.rsplit(",", 1)[0]
        # %.3e: one unit of the last digit, or the energies' own rounding
        rel = log["rel_error"][k]
        assert abs(float(mine.rsplit(",", 1)[1]) - rel) <= max(5e-8, 1.01e-3 * abs(rel))


def test_reference_energy_logs_share_e0():
    """The three logs start from one IC: their inverted E0 agree to 1e-8."""
    e0 = [golden_e0(formats.read_energy_csv(os.path.join(GOLDEN, n))) for n in ENERGY_LOGS]
    assert max(e0) - min(e0) < 1e-8 * abs(e0[0])


def test_gpu_bench_covers_reference_sweep():
    """python -m nbodyssey_amd.gpu_bench measures exactly the (n, method, theta)
    points of the reference's gpu_results.csv, under the same header."""
    from nbodyssey_amd import gpu_bench
    lines = open(os.path.join(GOLDEN, "gpu_results.csv")).read().splitlines()
    assert lines[0] == "n,method,theta,ms_per_call"
    ref = {(int(n), m, float(t)) for n, m, t, _ in (l.split(",") for l in lines[1:])}
    ours = {(n, m, 0.5) for n in gpu_bench.SIZES for m in ("naive", "bh-thread", "bh-warp")}
    ours |= {(100000, m, t) for t in (0.3, 0.8) for m in ("bh-thread", "bh-warp")}
    assert ours == ref


def test_gpu_bench_scrapes_like_the_reference():
    """ms/call is the next-to-last token of the force-pipeline line
    (gpu_bench.sh:31, awk '{print $(NF-1)}') and the gate reads the mean of
    the --compare-forces report (gpu_bench.sh:41-44)."""
    from nbodyssey_amd import gpu_bench
    out = ("n=4096 dt=0.01 eps=0.05 steps=10 force=naive\n"
           "force pipeline: 11 calls, 0.5 ms total, 0.043 ms/call\n")
    assert gpu_bench.ms_per_call(out) == 0.043
    gate = ("naive vs barnes-hut, theta=0, traverse=warp, n=12000\n"
            "relative force error: max 2.801e-04, mean 1.694e-05\n")
    assert gpu_bench.gate_mean(gate) == 1.694e-05


# ------------------------------------------------------------- galaxy_sim CLI

def run_cli(*args):
    return subprocess.run([GALAXY_SIM, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args,msg", [
    ([], "usage: galaxy_sim --ic ic.bin"),
    (["--bogus"], "unknown argument: --bogus"),
    (["--ic"], "missing value for --ic"),
    (["--ic", "x", "--force", "tree"], "--force must be naive or bh, got: tree"),
    (["--ic", "x", "--traverse", "block"], "--traverse must be thread or warp, got: block"),
    (["--ic", "/nonexistent/ic.bin"], "cannot open IC file: /nonexistent/ic.bin"),
    (["--ic", "x", "--steps", "ten"], "bad value for --steps: ten"),
    (["--ic", "x", "--keys", "32"], "--keys must be 30 or 63, got: 32"),
    (["--ic", "x", "--group-width", "48"], "--group-width must be 16, 32 or 64, got: 48"),
    (["--ic", "x", "--steps", "10abc"], "bad value for --steps: 10abc"),
    (["--ic", "x", "--dt", " 0.1"], "bad value for --dt:  0.1"),
    (["--ic", "x", "--dump-every", "0"], "--dump-every must be positive"),
    (["--ic", "x", "--energy-theta", "0.6"], "--energy-theta must be in [0, 0.577)"),
    (["--ic", "x", "--theta"], "missing value for --theta"),
])
def test_cli_errors(args, msg):
    assert os.path.exists(GALAXY_SIM)
    r = run_cli(*args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_cli_bad_ic_files(tmp_path):
    bad = tmp_path / "neg.bin"
    bad.write_bytes(np.int32(-1).tobytes())
    r = run_cli("--ic", str(bad))
    assert r.returncode == 1 and f"bad particle count in {bad}" in r.stderr
    trunc = tmp_path / "trunc.bin"
    trunc.write_bytes(np.int32(10).tobytes() + b"\0" * 100)
    r = run_cli("--ic", str(trunc))
    assert r.returncode == 1 and f"truncated IC file: {trunc}" in r.stderr


# -------------------------------------------------------------- IC generators

def test_two_disk_encounter_layout():
    pos, vel = ic.two_disk_encounter(12000, seed=0)
    assert pos.shape == (12000, 4) and vel.shape == (12000, 4)
    assert pos.dtype == np.float32 and np.all(vel[:, 3] == 0)
    np.testing.assert_allclose(pos[:, 3].sum(dtype=np.float64), 2.0, rtol=1e-6)
    a, b = pos[:6000], pos[6000:]
    # disk A around (-10,-2,0), B around (+10,+2,0), within the 5.0 truncation
    assert np.linalg.norm(a[:, :3] - [-10, -2, 0], axis=1).max() < 5.0 + 0.6
    assert np.linalg.norm(b[:, :3] - [10, 2, 0], axis=1).max() < 5.0 + 0.6
    np.testing.assert_allclose(vel[:6000, 0].mean(), 0.35, atol=0.02)
    np.testing.assert_allclose(vel[6000:, 0].mean(), -0.35, atol=0.02)
    p2, _ = ic.two_disk_encounter(12000, seed=0)
    assert np.array_equal(p2, pos)


def test_generate_ic_defaults_shape():
    pos, vel = ic.generate_ic_disks(particles=2000, seed=3)
    assert pos.shape == (2000, 4)
    np.testing.assert_allclose(pos[:, 3], 1.0 / 1000, rtol=1e-6)


def test_bench_refuses_several_gpus():
    """bench.py --gpus N > 1 exits with a message before touching a GPU."""
    import sys
    for argv in (["--gpus", "2"], ["--gpus", "8", "--steps", "1"]):
        r = subprocess.run([sys.executable, "bench.py", *argv], capture_output=True, text=True,
                           timeout=300, cwd=ROOT)
        assert r.returncode == 2, argv
        assert "one GPU" in r.stderr and r.stdout == ""
