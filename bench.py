"""Benchmark: Barnes-Hut leapfrog steps on one MI355X.

Workload (default): two-disk encounter, 1,048,576 particles, FP32 state,
Barnes-Hut LBVH rebuilt every step with 30-bit Morton keys, theta = 0.5,
eps = 0.02, dt = 0.005, wave64 group walk.  A step = fused kick+drift+bbox,
Morton, onesweep sort, Karras build, COM, tree walk + closing kick; inputs are
resident in HBM when the timed region starts.

Prints one JSON line (the metric is particle-steps/s) with:
  roofline      the tree walk kernel (the dominant one) against the FP32 VALU
                peak: algorithmic flops = 20 per interaction + 10 per opening
                test of the per-target walk on the same tree (counted on the
                device), divided by the walk kernel's mean duration measured
                with HIP events on the engine stream during the timed steps;
  accuracy      median / p99 relative force error of the final state's
                accelerations against FP64 direct summation on the GPU
                (nb_direct_f64) for sampled targets;
  energy        relative FP64 total-energy drift over the timed steps;
  allpairs      one tiled all-pairs launch on the initial 1M state:
                Ginteractions/s and its fraction of the FP32 peak;
  ensemble      256 parabolic two-disk encounters batched in one engine
                against the same systems one at a time: both rates, their
                ratio, whether the forces are bitwise equal, force error;
  cpu_baseline  the reference CPU algorithm (NumPy chunked float32 all-pairs
                + KDK, scripts/reference_nbody.py:45-61,130-136 restated)
                timed on this host's cores: the 12,000-particle run and one
                65,536-particle step extrapolated to 1M.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak
HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_target, sample_n=65536, c1_budget_s=15.0, workers=None):
    """The reference's CPU algorithm (scripts/reference_nbody.py:45-61,129-136:
    NumPy float32 chunked (b, n, 3) all-pairs + KDK, restated in
    oracle/numpy_ref.py) on this host, over a pool of one worker process per
    host core (up to 16), each evaluating one contiguous block of targets:

      * the 12,000-particle two-disk run (eps 0.05, dt 0.01, 200 naive KDK
        steps) timed wall-clock step by step, energy logging off, until
        c1_budget_s has passed or all 200 steps have run; the full-run time is
        extrapolated from the steps done;
      * one KDK step of the two-disk encounter at sample_n particles (eps
        0.02, dt 0.005), extrapolated to n_target as O(n^2) all-pairs (the
        reference's CPU path has no tree code at scale).

    Must run before this process touches the GPU (the pool forks)."""
    from multiprocessing import get_context
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    from nbodyssey_amd import ic
    from oracle import numpy_ref
    workers = workers or max(1, min(16, os.cpu_count() or 1))
    with get_context("fork").Pool(workers) as pool:
        p4, v4 = ic.two_disk_encounter(12000, seed=0, eps=0.05)
        t0 = time.perf_counter()
        numpy_ref.kdk_run(p4[:, :3], v4[:, :3], p4[:, 3], 0.05, 0.01, 200, pool=pool,
                          budget_s=c1_budget_s)
        c1_s = time.perf_counter() - t0
        c1_steps = numpy_ref.kdk_run.steps_done
        p4, v4 = ic.two_disk_encounter(sample_n, seed=1, eps=0.02)
        pos, mass, vel = p4[:, :3].copy(), p4[:, 3].copy(), v4[:, :3].copy()
        eps2, dt = np.float32(0.02 * 0.02), np.float32(0.005)
        acc = numpy_ref.accelerations(pos, mass, eps2, np.float32(1.0), pool=pool)
        t0 = time.perf_counter()
        vel += np.float32(0.5) * dt * acc        # one KDK step
        pos += dt * vel
        acc = numpy_ref.accelerations(pos, mass, eps2, np.float32(1.0), pool=pool)
        vel += np.float32(0.5) * dt * acc
        step_s = time.perf_counter() - t0
    scale = (n_target / sample_n) ** 2
    per_step = step_s * scale
    # 200 steps are 201 force evaluations (a(x0) first); c1_steps steps took
    # c1_steps + 1 of them
    c1_full = c1_s * 201 / (c1_steps + 1)
    return {"value": n_target / per_step, "unit": "particle-steps/s", "cores": workers,
            "kind": "port",
            "sample": f"NumPy float32 chunked all-pairs KDK (reference_nbody.py algorithm) on "
                      f"{workers} worker processes of {os.cpu_count()} host cores "
                      f"({cpu_model()}): one {sample_n}-particle step in {step_s:.2f} s, "
                      f"extrapolated x{scale:.0f} to {n_target} particles (O(n^2))",
            "extrapolated": True,
            "config1": {"workload": "two-disk 12000 particles, naive all-pairs KDK, eps 0.05, "
                                    "dt 0.01, 200 steps, energy logging off",
                        "steps_timed": c1_steps, "wall_s": c1_s,
                        "full_run_s": c1_full,
                        "full_run_extrapolated": c1_steps < 200,
                        "particle_steps_per_s": 12000 * 200 / c1_full},
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}


def profiled_walk_traffic(args):
    """HBM bytes per walk launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against plain streaming kernels), for the
    default workload only; None otherwise."""
    default = (args.workload == "two-disk" and args.n == 1 << 20 and args.theta == 0.5
               and args.traverse == "warp" and args.group_width == 64 and args.key_bits == 30)
    path = os.path.join(ROOT, "profiles", "bh_step_1m_theta05_traffic.md")
    if not default or not os.path.exists(path):
        return None
    import re
    m = re.search(r"traffic_bytes_per_launch = (\d+)", open(path).read())
    return float(m.group(1)) if m else None


def build_roofline(phase, calls, n, key_bits, passes=None):
    """Achieved HBM rate of the build phases (algorithmic bytes per particle
    over the phase's event time) against the 8 TB/s peak.  Bytes per
    particle: kick+drift+bbox 80 (read pos/vel/acc, write pos/vel); Morton 16
    + key; sort: per pass read and write key + index (the first pass reads no
    index) + 32 for the last pass's gather of pos_sorted; Karras: key 1x +
    node 12 + two upward records 32 + arrival counter 4; COM: pos_sorted 16 +
    two upward records 32 + node com/size 20."""
    kb = 4 if key_bits == 30 else 8
    passes = passes or (4 if key_bits == 30 else 8)
    per = {"bbox": 80.0, "morton": 16.0 + kb,
           "sort": passes * (2 * kb + 8) - 4 + 32,
           "build": kb + 12 + 32 + 4, "com": 16 + 32 + 20}
    out = {}
    for k, b in per.items():
        ms = phase[["bbox", "morton", "sort", "build", "com"].index(k)] / max(calls, 1)
        gbs = b * n / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[k] = {"bytes_per_particle": b, "ms": ms, "achieved_GBps": gbs,
                  "frac_of_8TBps": None if gbs is None else gbs / HBM_PEAK_GBS}
    return out


def naive_main(args, sim, eps):
    """All-pairs leapfrog steps; the step time is dominated by the all-pairs
    kernel (the fused kick/drift add ~48 B/particle of HBM traffic)."""
    sim.compute_forces()
    sim.step(args.warmup)
    sim.synchronize()
    sim.reset_force_timing()
    ms = sim.time_steps(args.steps)
    # the force kernel's own time from further timed calls (outside the bracket)
    sim.reset_force_timing()
    sim.step(2)
    sim.synchronize()
    force_ms = sim.force_ms_total() / max(sim.force_calls(), 1)
    n = args.n
    flops = 20.0 * n * n
    achieved = flops / (force_ms * 1e-3) / 1e12
    out = {
        "metric": "particle-steps/s (all-pairs)", "value": n * args.steps / (ms * 1e-3),
        "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic two-disk encounter",
        "config": {"workload": f"two-disk {n} particles, tiled all-pairs, eps={eps}", "n": n},
        "ginteractions_per_s": n * n / (force_ms * 1e-3) / 1e9,
        "roofline": {"bound": "valu", "kernel": "allpairs", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": flops, "force_ms": force_ms},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    sim.close()


def ensemble_run(args, steps, serial_steps, serial_warmup):
    """Batched ensemble of independent parabolic two-disk encounters against
    the same systems run one at a time on the plain path.

    Batched: one engine (nb_create_ensemble), launch configuration chosen by
    the autotuner (persistent-walk workgroups x graph replay, within 1% median
    force error).  Serial: one plain engine per system (Barnes-Hut, wave64
    walk, 30-bit keys, graph mode: the fastest plain configuration), started
    from the batched run's final state, timed per system and summed.  Both
    compute bitwise the same forces for a given state (the forest holds each
    system's standalone tree, checked for every system), so the accuracy is
    the same by construction.  Returns a dict of the measurements."""
    from nbodyssey_amd import engine, ic
    w = ic.WORKLOADS["ensemble-256-parabolic"]
    eps, dt, theta = w["eps"], w["dt"], args.theta
    t_ic = time.perf_counter()
    systems, _ = ic.ensemble_encounters(args.systems, seed=w["seed"])
    counts = np.array([len(p) for p, _ in systems])
    n = int(counts.sum())
    print(f"bench: ensemble {args.systems} systems, {n} particles, ic "
          f"{time.perf_counter() - t_ic:.1f}s", file=sys.stderr, flush=True)

    ens = engine.Ensemble(systems, dt=dt, eps=eps, theta=theta)
    ens.set_graph(not args.no_graph)
    ens.compute_forces()
    tuned = None
    if not args.no_graph:
        choice, log = ens.autotune(steps=3, max_median_err=0.01)
        tuned = {"choice": choice, "log": log}
        print("bench: autotune " + log[0], file=sys.stderr, flush=True)
    ens.step(args.warmup)
    ens.synchronize()
    ms_batched = ens.time_steps(steps)
    print(f"bench: batched {steps} steps {ms_batched:.2f} ms", file=sys.stderr, flush=True)
    # phase times (walk kernel duration) from eager steps of the same engine
    ens.set_graph(False)
    ens.reset_force_timing()
    ens.step(args.phase_steps)
    ens.synchronize()
    phase, calls = ens.bh_phase_ms()
    walk_ms = phase[5] / max(calls, 1)
    acc_b = ens.acc()
    p_b, v_b = ens.download()
    off = ens.offsets
    ens.close()

    # serial: each system alone from the batched run's final state
    ms_serial = 0.0
    mismatched = []
    inter_sum = opens_sum = 0.0
    err = []
    rng = np.random.default_rng(0)
    per_sys = max(1, args.accuracy_targets // args.systems)
    for s in range(args.systems):
        o, m = int(off[s]), int(counts[s])
        with engine.Simulation(p_b[o:o + m], v_b[o:o + m], dt=dt, eps=eps, theta=theta,
                               force="bh", traverse="warp", group_width=64,
                               key_bits=30) as sim:
            a_s = sim.compute_forces()
            if not np.array_equal(a_s, acc_b[o:o + m]):
                mismatched.append(s)
            ti, to = sim.traverse_counts("thread")
            inter_sum += ti.astype(np.float64).sum()
            opens_sum += to.astype(np.float64).sum()
            idx = rng.choice(m, per_sys, replace=False).astype(np.int32)
            err.append(sim.force_error(idx, a_s))
            sim.set_graph(not args.no_graph)
            sim.step(serial_warmup)
            sim.synchronize()
            ms_serial += sim.time_steps(serial_steps) / serial_steps
    value = n * steps / (ms_batched * 1e-3)
    serial_value = n / (ms_serial * 1e-3)
    print(f"bench: serial {ms_serial:.3f} ms per ensemble step (sum over systems)",
          file=sys.stderr, flush=True)
    flops = 20.0 * inter_sum + 10.0 * opens_sum
    e = np.concatenate(err)
    return dict(
        systems=args.systems, n=n, theta=theta, eps=eps, dt=dt, steps=steps,
        ms_batched=ms_batched, value=value, serial_value=serial_value, ms_serial=ms_serial,
        serial_steps=serial_steps, tuned=tuned, bitwise=not mismatched, phase=phase,
        calls=calls, walk_ms=walk_ms, flops=flops,
        accuracy={"median_rel_force_err_vs_fp64": float(np.median(e)),
                  "p99_rel_force_err_vs_fp64": float(np.quantile(e, 0.99)),
                  "targets": int(len(e)),
                  "reference": "FP64 direct summation on the GPU (nb_direct_f64)"})


def ensemble_summary(r, graph=True):
    """The ensemble measurements as one JSON object."""
    return {
        "workload": f"{r['systems']} parabolic two-disk encounters, {r['n']} particles "
                    f"(n_s log-uniform 2048..65536), Barnes-Hut rebuilt every step, (system, "
                    f"30-bit) keys, theta={r['theta']}, eps={r['eps']}, dt={r['dt']}, wave64 "
                    f"persistent walk, {'graph-captured' if graph else 'eager'} step",
        "particle_steps_per_s": r["value"], "ms_per_step": r["ms_batched"] / r["steps"],
        "steps": r["steps"],
        "serial_particle_steps_per_s": r["serial_value"],
        "serial_ms_per_ensemble_step": r["ms_serial"],
        "serial_how": f"each system alone on the plain path (bh, wave64, 30-bit keys, "
                      f"{'graph mode' if graph else 'eager'}), {r['serial_steps']} timed steps "
                      f"each, per-step times summed",
        "batched_over_serial": r["value"] / r["serial_value"],
        "batched_forces_bitwise_equal_serial": r["bitwise"],
        "accuracy": r["accuracy"],
        "autotune_choice": None if r["tuned"] is None else r["tuned"]["choice"],
    }


def ensemble_main(args):
    r = ensemble_run(args, args.steps, args.serial_steps, args.serial_warmup)
    achieved = r["flops"] / (r["walk_ms"] * 1e-3) / 1e12
    names = ["bbox", "morton", "sort", "build", "com", "traverse"]
    summary = ensemble_summary(r, not args.no_graph)
    out = {
        "metric": "ensemble particle-steps/s (batched vs serial)",
        "value": r["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_batched"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic parabolic two-disk encounters (seeded generator)",
        "config": {"workload": summary["workload"], "systems": args.systems, "n": r["n"],
                   "theta": r["theta"]},
        "serial": {"value": r["serial_value"], "ms_per_ensemble_step": r["ms_serial"],
                   "how": summary["serial_how"]},
        "batched_over_serial": summary["batched_over_serial"],
        "autotune": r["tuned"],
        "batched_forces_bitwise_equal_serial": r["bitwise"],
        "phase_ms": {k: r["phase"][i] / max(r["calls"], 1) for i, k in enumerate(names)},
        "build_roofline_hbm": build_roofline(r["phase"], r["calls"], r["n"], 63, passes=5),
        "roofline": {"bound": "valu", "kernel": "walk_items",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": r["flops"], "walk_ms": r["walk_ms"],
                     "note": f"walk duration from {args.phase_steps} eager steps after the "
                             f"timed graph steps"},
        "accuracy": r["accuracy"],
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def allpairs_probe(pos, eps, launches=1):
    """Tiled all-pairs on the same positions: one warm-up launch, then
    `launches` timed force evaluations (HIP events around the kernels);
    20 flop per interaction, n^2 interactions per launch (main.cu:203-207)."""
    from nbodyssey_amd import engine
    n = len(pos)
    with engine.Simulation(pos, eps=eps, force="naive") as sim:
        sim.compute_forces()
        sim.reset_force_timing()
        for _ in range(launches):
            sim.compute_forces()
        ms = sim.force_ms_total() / max(sim.force_calls(), 1)
    gint = n * n / (ms * 1e-3) / 1e9
    return {"n": n, "launches": launches, "ms_per_launch": ms, "ginteractions_per_s": gint,
            "achieved_tflops": gint * 20.0 / 1e3, "peak_tflops": FP32_PEAK_TFLOPS,
            "frac_of_fp32_peak": gint * 20.0 / 1e3 / FP32_PEAK_TFLOPS,
            "convention": "20 flop per interaction, n^2 interactions (main.cu:203-207)"}


def precision_main(args):
    """Precision study on one two-disk system (default 4,194,304 particles,
    eps 0.02, dt 0.005): for theta in {0.3, 0.5, 0.7} and the FP32, compensated
    (Kahan) and FP16-far-field walks, the step time, the force error on a seeded
    subsample against FP64 direct summation on the GPU, and the relative
    total-energy drift over `drift_steps` steps (FP64 energy on the GPU)."""
    from nbodyssey_amd import engine, ic
    w = ic.WORKLOADS["two-disk-4M-precision"]
    n = args.n if "--n" in sys.argv else w["n"]
    eps, dt = w["eps"], w["dt"]
    pos, vel = ic.two_disk_encounter(n, seed=w["seed"], eps=eps)
    idx = np.random.default_rng(0).choice(n, min(args.error_targets, n),
                                          replace=False).astype(np.int32)
    rows = []
    headline = None
    for theta in (0.3, 0.5, 0.7):
        for form in ("fp32", "kahan", "fp16far"):
            comp = form == "kahan"
            with engine.Simulation(pos, vel, dt=dt, eps=eps, theta=theta, force="bh",
                                   traverse="warp", group_width=64, key_bits=30,
                                   compensated=comp, far16=form == "fp16far") as sim:
                a = sim.compute_forces()
                err = sim.force_error(idx, a)
                sim.step(args.warmup)
                ms = sim.time_steps(args.steps)
                e0 = None
                drift = None
                if args.drift_steps > 0:
                    sim.upload(pos, vel)
                    sim.compute_forces()
                    e0 = sim.total_energy()
                    sim.step(args.drift_steps)
                    e1 = sim.total_energy()
                    drift = float((e1[0] - e0[0]) / abs(e0[0]))
            row = {"theta": theta, "precision": form,
                   "ms_per_step": ms / args.steps,
                   "median_rel_force_err_vs_fp64": float(np.median(err)),
                   "p99_rel_force_err_vs_fp64": float(np.quantile(err, 0.99)),
                   "energy_drift": drift, "drift_steps": args.drift_steps}
            rows.append(row)
            print(f"bench: {row}", file=sys.stderr, flush=True)
            if theta == 0.5 and form == "fp32":
                headline = row
    value = n / (headline["ms_per_step"] * 1e-3)
    out = {
        "metric": f"particle-steps/s @{n} BH theta=0.5 incl. rebuild (precision study)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": headline["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic two-disk encounter (seeded generator)",
        "config": {"workload": f"two-disk {n} particles, Barnes-Hut rebuilt every step, 30-bit "
                               f"keys, wave64 walk, eps={eps}, dt={dt}; theta x precision sweep",
                   "n": n, "error_targets": int(len(idx))},
        "precision_study": rows,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--force", default="bh", choices=["bh", "naive"],
                    help="naive: time the tiled all-pairs step instead (roofline = 20 flop "
                         "per interaction, n^2 interactions per launch)")
    ap.add_argument("--traverse", default="warp", choices=["warp", "thread"])
    ap.add_argument("--group-width", type=int, default=64)
    ap.add_argument("--walk-stack", type=int, default=0,
                    help="LDS stack entries per wave of the group walk (0: automatic)")
    ap.add_argument("--key-bits", type=int, default=30)
    ap.add_argument("--cpu-sample", type=int, default=65536,
                    help="particles of the one timed CPU step extrapolated to n")
    ap.add_argument("--cpu-c1-budget", type=float, default=15.0,
                    help="seconds of the 12,000-particle CPU run (up to its 200 steps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="two-disk default: skip the all-pairs and ensemble measurements "
                         "that the default line carries beside the headline")
    ap.add_argument("--ens-steps", type=int, default=50,
                    help="default line: batched ensemble steps timed")
    ap.add_argument("--ens-serial-steps", type=int, default=5,
                    help="default line: timed steps per system of the serial ensemble run")
    ap.add_argument("--accuracy-targets", type=int, default=256)
    ap.add_argument("--far16", action="store_true",
                    help="FP16 far field for accepted cells (wave64 walk precision mode 2); "
                         "reported with dtype f32+f16")
    ap.add_argument("--graph", action="store_true",
                    help="replay each step as a captured HIP graph (no per-phase timing)")
    ap.add_argument("--error-targets", type=int, default=65536)
    ap.add_argument("--drift-steps", type=int, default=500)
    ap.add_argument("--workload", default="two-disk",
                    choices=["two-disk", "ensemble", "precision", "composite"],
                    help="ensemble: batched independent systems vs the same systems one at a "
                         "time (defaults: 256 systems, 200 steps)")
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--serial-steps", type=int, default=20)
    ap.add_argument("--serial-warmup", type=int, default=2)
    ap.add_argument("--phase-steps", type=int, default=10)
    ap.add_argument("--energy-theta", type=float, default=0.25,
                    help="composite: theta of the tree-potential energy taken before and after "
                         "the timed steps (0 = none)")
    ap.add_argument("--no-graph", action="store_true",
                    help="ensemble: time eager steps instead of graph replays, batched and "
                         "serial (for kernel tracing)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the hot path (one N-body system per GPU) has "
              "no multi-GPU mode in this build; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ensemble":
        if "--steps" not in sys.argv:
            args.steps = 200
        return ensemble_main(args)
    if args.workload == "precision":
        if "--steps" not in sys.argv:
            args.steps = 20
        return precision_main(args)

    from nbodyssey_amd import ic
    if args.workload == "composite":
        w = ic.WORKLOADS["composite-33M-63bit"]
        n = args.n if "--n" in sys.argv else w["n"]
        for k in ("theta", "key_bits"):
            if "--" + k.replace("_", "-") not in sys.argv:
                setattr(args, k, w[k])
        t_ic = time.perf_counter()
        pos, vel = ic.composite_encounter(n, seed=w["seed"], eps=w["eps"])
        print(f"bench: composite n={n} ic {time.perf_counter() - t_ic:.1f}s", file=sys.stderr,
              flush=True)
        args.n = n
        return bh_main(args, pos, vel, w["eps"], w["dt"],
                       f"two composite galaxies (disk 60% + Hernquist bulge 10% + Plummer halo "
                       f"30%) on a parabolic orbit, {n} particles",
                       f"particle-steps/s @{n} composite BH theta={args.theta} incl. rebuild",
                       "synthetic composite-galaxy encounter (seeded generator)",
                       energy_theta=args.energy_theta)
    eps, dt = 0.02, 0.005
    headline = args.force == "bh" and not args.headline_only
    # the CPU baseline forks its worker pool, so it runs before the GPU is touched
    cpu_line = None
    if args.force == "bh" and not args.no_cpu_baseline:
        t_cpu = time.perf_counter()
        cpu_line = cpu_baseline(args.n, args.cpu_sample, args.cpu_c1_budget)
        print(f"bench: cpu baseline {time.perf_counter() - t_cpu:.1f}s", file=sys.stderr,
              flush=True)
    t_ic = time.perf_counter()
    pos, vel = ic.two_disk_encounter(args.n, seed=1, eps=eps)
    print(f"bench: n={args.n} ic {time.perf_counter() - t_ic:.1f}s", file=sys.stderr, flush=True)
    if args.force == "naive":
        from nbodyssey_amd import engine
        sim = engine.Simulation(pos, vel, dt=dt, eps=eps, force="naive")
        return naive_main(args, sim, eps)
    return bh_main(args, pos, vel, eps, dt, f"two-disk {args.n} particles",
                   "particle-steps/s @1M BH θ=0.5 incl. rebuild; ensemble vs serial at ≤1% "
                   "force err",
                   "synthetic two-disk encounter (seeded generator)",
                   energy_theta=None, cpu_line=cpu_line, headline=headline)


def bh_main(args, pos, vel, eps, dt, what, metric, data, energy_theta=None, cpu_line=None,
            headline=False):
    """Barnes-Hut leapfrog steps of one system: timed steps, phase times,
    walk roofline from the per-target walk counts, accuracy vs FP64 direct
    summation, optional tree-energy drift over the timed steps and the CPU
    baseline.  headline: also the FP64 energy drift over the timed steps, one
    all-pairs launch on the initial state and the batched ensemble against
    its serial runs (the three targets the metric names)."""
    from nbodyssey_amd import engine
    n = len(pos)
    sim = engine.Simulation(pos, vel, dt=dt, eps=eps, theta=args.theta, force="bh",
                            traverse=args.traverse, group_width=args.group_width,
                            key_bits=args.key_bits, far16=args.far16)
    if args.walk_stack:
        sim.set_launch(walk_stack=args.walk_stack)
    walk_stack = sim.launch()["walk_stack"] or 1024
    if args.graph:
        sim.set_graph(True)
    sim.compute_forces()              # a(x0)
    sim.step(args.warmup)
    sim.synchronize()
    print("bench: warmup done", file=sys.stderr, flush=True)
    energy = None
    if energy_theta:
        e0 = sim.total_energy_tree(energy_theta)
    elif headline:
        e0 = sim.total_energy()   # exact FP64 sums, untimed

    sim.synchronize()
    t0 = time.perf_counter()
    ms_events = sim.time_steps(args.steps)   # no timing events inside the bracket
    sim.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ms_events / args.steps
    value = n * args.steps / (ms_events * 1e-3)
    print(f"bench: {args.steps} steps {ms_events:.2f} ms (wall {wall * 1e3:.2f} ms)",
          file=sys.stderr, flush=True)
    if energy_theta or headline:
        e1 = sim.total_energy_tree(energy_theta) if energy_theta else sim.total_energy()
        how = (f"tree potential at theta={energy_theta} (nb_total_energy_tree)" if energy_theta
               else "exact O(n^2) FP32 pair potentials with FP64 sums (nb_total_energy)")
        energy = {"method": how, "E_start": float(e0[0]), "E_end": float(e1[0]),
                  "rel_drift": float((e1[0] - e0[0]) / abs(e0[0])), "steps": args.steps}
    # per-phase times (and the walk kernel's duration) from further eager steps
    # with phase events, outside the timed region
    if args.graph:
        sim.set_graph(False)
    sim.reset_force_timing()
    sim.step(args.phase_steps)
    sim.synchronize()
    phase, calls = sim.bh_phase_ms()

    # algorithmic flops of the walk on the final tree: per-target walk counts
    inter, opens = sim.traverse_counts("thread")
    flops = 20.0 * inter.astype(np.float64).sum() + 10.0 * opens.astype(np.float64).sum()
    walk_ms = phase[5] / max(calls, 1)
    achieved = flops / (walk_ms * 1e-3) / 1e12
    ginter, gopen = sim.traverse_counts(args.traverse, args.group_width)
    width = 1 if args.traverse == "thread" else args.group_width
    lane_inter = ginter.astype(np.float64).sum() * width if width > 1 else ginter.sum()

    # accuracy of the final state's forces vs FP64 direct summation on the GPU
    idx = np.random.default_rng(0).choice(n, args.accuracy_targets,
                                          replace=False).astype(np.int32)
    err = sim.force_error(idx)
    accuracy = {"median_rel_force_err_vs_fp64": float(np.median(err)),
                "p99_rel_force_err_vs_fp64": float(np.quantile(err, 0.99)),
                "targets": int(args.accuracy_targets),
                "reference": "FP64 direct summation on the GPU (nb_direct_f64)"}
    extras = {}
    if headline:
        sim.close()
        print("bench: all-pairs launch", file=sys.stderr, flush=True)
        extras["allpairs"] = allpairs_probe(pos, eps)
        extras["ensemble"] = ensemble_summary(
            ensemble_run(args, args.ens_steps, args.ens_serial_steps, 1), not args.no_graph)

    names = ["bbox", "morton", "sort", "build", "com", "traverse"]
    build = build_roofline(phase, calls, n, args.key_bits)
    # whole-step roofline: each phase at the larger of its HBM time (algorithmic
    # bytes at 8 TB/s) and its FP32 time (algorithmic flops at 157.3 TFLOP/s);
    # the walk streams pos_sorted, perm and acc (36 B per particle)
    t_roof_ms = 1e3 * (sum(v["bytes_per_particle"] for v in build.values()) * n
                       / (HBM_PEAK_GBS * 1e9)
                       + max(36.0 * n / (HBM_PEAK_GBS * 1e9), flops / (FP32_PEAK_TFLOPS * 1e12)))
    out = {
        "metric": metric,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f16 far field" if args.far16 else "f32",
        "data": data,
        "config": {"workload": f"{what}, Barnes-Hut LBVH rebuilt every step, "
                               f"{args.key_bits}-bit keys, theta={args.theta}, eps={eps}, dt={dt}, "
                               f"walk={args.traverse}{'' if width == 1 else width}",
                   "n": n, "theta": args.theta, "key_bits": args.key_bits,
                   "traverse": args.traverse, "group_width": width,
                   "walk_stack": walk_stack},
        "phase_ms": {k: phase[i] / max(calls, 1) for i, k in enumerate(names)},
        "build_roofline_hbm": build,
        "step_roofline": {"t_roof_ms": t_roof_ms, "t_meas_ms": ms_per_step,
                          "frac": t_roof_ms / ms_per_step,
                          "convention": "sum over phases of max(algorithmic bytes / 8 TB/s, "
                                        "per-target-walk flops / 157.3 TFLOP/s)"},
        "roofline": {"bound": "valu", "kernel": "walk",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": profiled_walk_traffic(args),
                     "traffic_unit": "bytes/launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, calibrated)",
                     "flops_per_launch": flops, "walk_ms": walk_ms,
                     "per_target_interactions": float(inter.mean()),
                     "per_target_opens": float(opens.mean()),
                     "group_lane_interactions_per_target": float(lane_inter / n)},
        "accuracy": accuracy,
        "cpu_baseline": cpu_line,
    }
    if energy is not None:
        out["energy"] = energy
    out.update(extras)
    print(json.dumps(out))
    sim.close()


if __name__ == "__main__":
    main()
