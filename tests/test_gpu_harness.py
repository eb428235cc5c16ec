"""The gpu_bench.sh-equivalent session driver (nbodyssey_amd/gpu_bench.py) run
end to end on the GPU in its quick mode: the theta = 0 --compare-forces gate
for both walks (scripts/gpu_bench.sh:35-48, mean < 1e-4), then the timing
sweep written with the reference's CSV header (gpu_bench.sh:20-33)."""
import os
import subprocess
import sys

import pytest

from nbodyssey_amd import gpu_bench

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_gpu_bench_quick_session(tmp_path):
    out = tmp_path / "bench"
    r = subprocess.run([sys.executable, "-m", "nbodyssey_amd.gpu_bench", "--quick", "--out",
                        str(out)], capture_output=True, text=True, timeout=600, cwd=ROOT)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0
    for walk in ("thread", "warp"):
        assert f"gate passed ({walk})" in r.stdout
        log = (out / "logs" / f"gate_theta0_{walk}.txt").read_text()
        assert f"theta=0, traverse={walk}, n=12000" in log
        assert gpu_bench.gate_mean(log) < 1e-4
    lines = (out / "gpu_results.csv").read_text().splitlines()
    assert lines[0] == "n,method,theta,ms_per_call"
    rows = [l.split(",") for l in lines[1:]]
    expect = [(n, m) for n in gpu_bench.QUICK_SIZES for m in ("naive", "bh-thread", "bh-warp")]
    assert [(int(n), m) for n, m, _, _ in rows] == expect
    for n, m, theta, ms in rows:
        assert float(theta) == 0.5 and float(ms) > 0
        assert (out / "logs" / f"{m}_n{n}_theta0.5.txt").exists()


def test_bench_default_path_small():
    """bench.py's default path (headline steps, CPU baseline, energy drift,
    all-pairs launch, batched ensemble against serial) end to end at small
    sizes: one JSON line with the fields of the contract."""
    import json
    cmd = [sys.executable, "bench.py", "--n", "65536", "--steps", "3", "--warmup", "1",
           "--cpu-sample", "2048", "--cpu-c1-budget", "0.5", "--systems", "6", "--ens-steps", "3",
           "--ens-serial-steps", "1", "--accuracy-targets", "16", "--phase-steps", "2"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline", "energy", "allpairs", "ensemble", "accuracy"):
        assert k in line, k
    assert line["n_gpus"] == 1 and line["steps"] == 3 and line["warmup"] == 1
    assert line["value"] > 0 and line["ms_per_step"] > 0
    assert line["config"]["n"] == 65536 and line["config"]["walk_stack"] == 1024
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in line["roofline"], k
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in line["cpu_baseline"], k
    assert line["ensemble"]["batched_forces_bitwise_equal_serial"] is True
