"""Copy the reference's committed benchmark data into tests/golden/.

The reference holds four small CSV files under ``benchmarks/`` that its
documentation and ``scripts/gpu_bench.sh`` are built on (data, not source):

* ``energy_gpu_naive.csv``, ``energy_gpu_bh_thread.csv``,
  ``energy_gpu_bh_warp.csv``: the 12,000-particle ``generate_ic.py`` run (seed 0,
  dt 0.01, eps 0.05, 1,500 steps, a row every 100 steps), one per force module
  (``scripts/gpu_bench.sh:64-71``; ``benchmarks/README.md``);
* ``gpu_results.csv``: the T4 timing sweep, header ``n,method,theta,ms_per_call``
  (``scripts/gpu_bench.sh:20-62``).

They are copied byte for byte; the tests read them as expected outputs.  Run
from the repository root in a checkout that has the reference next to it:

    python tests/golden/make_golden.py /root/reference
"""
import shutil
import sys
from pathlib import Path

FILES = ["energy_gpu_naive.csv", "energy_gpu_bh_thread.csv", "energy_gpu_bh_warp.csv",
         "gpu_results.csv"]


def main(ref_root):
    src = Path(ref_root) / "benchmarks"
    dst = Path(__file__).resolve().parent
    for name in FILES:
        shutil.copyfile(src / name, dst / name)
        print(f"{src / name} -> {dst / name}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
