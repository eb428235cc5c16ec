"""GPU session driver with the reference's gate and sweep (scripts/gpu_bench.sh:20-80).

  python -m nbodyssey_amd.gpu_bench [--quick] [--out benchmarks]

1. theta = 0 correctness gate: ``galaxy_sim --compare-forces --theta 0`` for the
   thread and warp walks on a 20,000-particle IC; the mean relative force
   error must be < 1e-4 or the session stops with exit status 1
   (gpu_bench.sh:35-48).
2. Timing sweep: naive, bh-thread and bh-warp at theta 0.5 over the
   reference's particle counts, 10 steps each, ``ms/call`` scraped from the
   ``force pipeline`` line as gpu_bench.sh:31 does, written to
   ``gpu_results.csv`` with the header ``n,method,theta,ms_per_call``.
3. The theta dial (0.3, 0.8) at 100,000 particles for both walks.
4. Energy logs: 12,000 particles, 1,500 steps, dump every 100, naive and warp.

ICs come from the restated generate_ic.py (nbodyssey_amd.ic.generate_ic_disks,
same defaults), so the inputs are the reference's.  Additive: ``--group-width``
for the warp walk (default 32 = the reference's warp).
"""
import argparse
import os
import re
import shutil
import subprocess
import sys
import tempfile

from . import GALAXY_SIM, formats, ic

SIZES = [4096, 12000, 50000, 100000, 200000, 400000, 1000000]
QUICK_SIZES = [4096, 12000, 50000]


def run_sim(args, log_path):
    proc = subprocess.run([GALAXY_SIM, *args], capture_output=True, text=True)
    with open(log_path, "w") as f:
        f.write(proc.stdout)
        f.write(proc.stderr)
    if proc.returncode != 0:
        raise RuntimeError(f"galaxy_sim {' '.join(args)} failed ({proc.returncode}): "
                           f"{proc.stderr.strip()}")
    return proc.stdout


def ms_per_call(stdout):
    line = [l for l in stdout.splitlines() if "ms/call" in l][0]
    return float(line.split()[-2])   # awk '{print $(NF-1)}'


def gate_mean(stdout):
    line = [l for l in stdout.splitlines() if "relative force error" in l][0]
    return float(re.sub(r".*mean ", "", line))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default="benchmarks")
    ap.add_argument("--quick", action="store_true", help="small sweep, no energy runs")
    ap.add_argument("--group-width", default="32")
    a = ap.parse_args(argv)
    os.makedirs(os.path.join(a.out, "logs"), exist_ok=True)
    scratch = tempfile.mkdtemp()
    try:
        return session(a, scratch)
    finally:
        shutil.rmtree(scratch, ignore_errors=True)


def session(a, scratch):
    warp = ["--traverse", "warp", "--group-width", a.group_width]

    def make_ic(n, name):
        path = os.path.join(scratch, name)
        pos, vel = ic.generate_ic_disks(particles=n, seed=0)
        formats.write_ic(path, pos, vel)
        return path

    gate_ic = make_ic(20000 if not a.quick else 12000, "ic_gate.bin")
    for tv, extra in (("thread", ["--traverse", "thread"]), ("warp", warp)):
        out = run_sim(["--ic", gate_ic, "--compare-forces", "--theta", "0", *extra],
                      os.path.join(a.out, "logs", f"gate_theta0_{tv}.txt"))
        mean = gate_mean(out)
        print(out.strip(), flush=True)
        if not mean < 1e-4:
            print(f"GATE FAILED ({tv}): theta=0 mean error {mean}", file=sys.stderr)
            return 1
        print(f"gate passed ({tv}): theta=0 mean error {mean}", flush=True)

    csv_path = os.path.join(a.out, "gpu_results.csv")
    rows = ["n,method,theta,ms_per_call"]

    def case(n, label, theta, extra):
        path = os.path.join(scratch, f"ic_{n}.bin")
        if not os.path.exists(path):
            make_ic(n, f"ic_{n}.bin")
        out = run_sim(["--ic", path, "--steps", "10", "--dump-every", "1000000", "--out",
                       os.path.join(scratch, "frames"), "--energy-log",
                       os.path.join(scratch, "e.csv"), "--theta", str(theta), *extra],
                      os.path.join(a.out, "logs", f"{label}_n{n}_theta{theta}.txt"))
        ms = ms_per_call(out)
        rows.append(f"{n},{label},{theta},{ms:.3f}")
        print(rows[-1], flush=True)

    for n in (QUICK_SIZES if a.quick else SIZES):
        case(n, "naive", 0.5, ["--force", "naive"])
        case(n, "bh-thread", 0.5, ["--force", "bh", "--traverse", "thread"])
        case(n, "bh-warp", 0.5, ["--force", "bh", *warp])
    if not a.quick:
        for th in (0.3, 0.8):
            case(100000, "bh-thread", th, ["--force", "bh", "--traverse", "thread"])
            case(100000, "bh-warp", th, ["--force", "bh", *warp])
    with open(csv_path, "w") as f:
        f.write("\n".join(rows) + "\n")

    if not a.quick:
        e_ic = make_ic(12000, "ic_e.bin")
        for label, extra in (("naive", []), ("bh_warp", ["--force", "bh", "--theta", "0.5",
                                                         *warp])):
            run_sim(["--ic", e_ic, "--steps", "1500", "--dump-every", "100", "--out",
                     os.path.join(scratch, "frames_e"), "--energy-log",
                     os.path.join(a.out, f"energy_gpu_{label}.csv"), *extra],
                    os.path.join(a.out, "logs", f"energy_{label}.txt"))
            e = formats.read_energy_csv(os.path.join(a.out, f"energy_gpu_{label}.csv"))
            print(f"final energy error {label}: {e['rel_error'][-1]:.3e}", flush=True)
    print(open(csv_path).read())
    return 0


if __name__ == "__main__":
    sys.exit(main())
