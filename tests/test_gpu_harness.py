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
