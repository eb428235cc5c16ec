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
  cpu_baseline  the reference CPU algorithm (NumPy chunked float32 all-pairs
                + KDK, scripts/reference_nbody.py:45-61,130-136 restated)
                timed on this host on a bounded sample and extrapolated.
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


def cpu_baseline(n_target, sample_n, eps, dt):
    """reference_nbody.py's algorithm (NumPy, float32, chunk 2048) on one step
    of a sample_n-particle two-disk system; extrapolated to n_target as
    O(n^2) all-pairs (the reference's CPU path has no tree code at scale)."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from nbodyssey_amd import ic
    from oracle import numpy_ref
    pos4, vel4 = ic.two_disk_encounter(sample_n, seed=1, eps=eps)
    pos, mass, vel = pos4[:, :3].copy(), pos4[:, 3].copy(), vel4[:, :3].copy()
    t0 = time.perf_counter()
    numpy_ref.kdk_run(pos, vel, mass, eps, dt, 1)
    sec = time.perf_counter() - t0   # a(x0) + one step = 2 force evaluations
    per_step = sec / 2.0 * (n_target / sample_n) ** 2
    return {"value": n_target / per_step, "unit": "particle-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"NumPy float32 chunked all-pairs KDK (reference_nbody.py algorithm), "
                      f"{sample_n} particles, 2 force evaluations in {sec:.1f} s on 1 of "
                      f"{os.cpu_count()} host cores ({cpu_model()}), extrapolated x"
                      f"{(n_target / sample_n) ** 2:.0f} to {n_target} particles (O(n^2))"}


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


def ensemble_main(args):
    """Batched ensemble of independent parabolic two-disk encounters against
    the same systems run one at a time on the plain path.

    Batched: one engine (nb_create_ensemble), launch configuration chosen by
    the autotuner (persistent-walk workgroups, graph replay).  Serial: one plain engine per system (Barnes-Hut, wave64 walk,
    30-bit keys, graph mode: the fastest plain configuration), started from
    the batched run's final state, timed per system and summed.  Both compute
    bitwise the same forces for a given state (the forest holds each system's
    standalone tree), so the accuracy is the same by construction.
    """
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
        # autotuner: persistent-walk workgroups x graph replay, within 1% median error
        choice, log = ens.autotune(steps=3, max_median_err=0.01)
        tuned = {"choice": choice, "log": log}
        print("bench: autotune " + log[0], file=sys.stderr, flush=True)
    ens.step(args.warmup)
    ens.synchronize()
    ms_batched = ens.time_steps(args.steps)
    print(f"bench: batched {args.steps} steps {ms_batched:.2f} ms", file=sys.stderr, flush=True)
    # phase times (walk kernel duration) from eager steps of the same engine
    ens.set_graph(False)
    ens.reset_force_timing()
    ens.time_steps(args.phase_steps)
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
            sim.step(args.serial_warmup)
            sim.synchronize()
            ms_serial += sim.time_steps(args.serial_steps) / args.serial_steps
    ms_serial_total = ms_serial * args.steps   # same step count as the batched run
    value = n * args.steps / (ms_batched * 1e-3)
    serial_value = n * args.steps / (ms_serial_total * 1e-3)
    print(f"bench: serial {ms_serial:.3f} ms per ensemble step (sum over systems)",
          file=sys.stderr, flush=True)
    flops = 20.0 * inter_sum + 10.0 * opens_sum
    achieved = flops / (walk_ms * 1e-3) / 1e12
    names = ["bbox", "morton", "sort", "build", "com", "traverse"]
    accuracy = None
    if err:
        e = np.concatenate(err)
        accuracy = {"median_rel_force_err_vs_fp64": float(np.median(e)),
                    "p99_rel_force_err_vs_fp64": float(np.quantile(e, 0.99)),
                    "targets": int(len(e))}
    out = {
        "metric": "ensemble particle-steps/s (batched vs serial)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_batched / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic parabolic two-disk encounters (seeded generator)",
        "config": {"workload": f"ensemble of {args.systems} two-disk encounters, {n} particles "
                               f"(n_s log-uniform 2048..65536), Barnes-Hut rebuilt every step, "
                               f"(system, 30-bit) keys, theta={theta}, eps={eps}, dt={dt}, "
                               f"wave64 persistent walk, "
                               f"{'eager' if args.no_graph else 'graph-captured'} step",
                   "systems": args.systems, "n": n, "theta": theta},
        "serial": {"value": serial_value, "ms_per_ensemble_step": ms_serial,
                   "how": "each system alone on the plain path (bh, wave64, 30-bit keys, "
                          + ("eager" if args.no_graph else "graph mode") + "), summed"},
        "batched_over_serial": value / serial_value,
        "autotune": tuned,
        "batched_forces_bitwise_equal_serial": not mismatched,
        "phase_ms": {k: phase[i] / max(calls, 1) for i, k in enumerate(names)},
        "build_roofline_hbm": build_roofline(phase, calls, n, 63, passes=5),
        "roofline": {"bound": "valu", "kernel": "walk_items",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": flops, "walk_ms": walk_ms,
                     "note": f"walk duration from {args.phase_steps} eager steps after the "
                             f"timed graph steps"},
        "accuracy": accuracy,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


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
    ap.add_argument("--key-bits", type=int, default=30)
    ap.add_argument("--cpu-sample", type=int, default=16384)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
                       energy_theta=args.energy_theta, cpu=False)
    eps, dt = 0.02, 0.005
    t_ic = time.perf_counter()
    pos, vel = ic.two_disk_encounter(args.n, seed=1, eps=eps)
    print(f"bench: n={args.n} ic {time.perf_counter() - t_ic:.1f}s", file=sys.stderr, flush=True)
    if args.force == "naive":
        from nbodyssey_amd import engine
        sim = engine.Simulation(pos, vel, dt=dt, eps=eps, force="naive")
        return naive_main(args, sim, eps)
    return bh_main(args, pos, vel, eps, dt, f"two-disk {args.n} particles",
                   "particle-steps/s @1M BH θ=0.5 incl. rebuild",
                   "synthetic two-disk encounter (seeded generator)",
                   energy_theta=None, cpu=not args.no_cpu_baseline)


def bh_main(args, pos, vel, eps, dt, what, metric, data, energy_theta=None, cpu=True):
    """Barnes-Hut leapfrog steps of one system: timed steps, phase times,
    walk roofline from the per-target walk counts, accuracy vs FP64 direct
    summation, optional tree-energy drift over the timed steps and the CPU
    baseline."""
    from nbodyssey_amd import engine
    n = len(pos)
    sim = engine.Simulation(pos, vel, dt=dt, eps=eps, theta=args.theta, force="bh",
                            traverse=args.traverse, group_width=args.group_width,
                            key_bits=args.key_bits, far16=args.far16)
    if args.graph:
        sim.set_graph(True)
    sim.compute_forces()              # a(x0)
    sim.step(args.warmup)
    sim.synchronize()
    print("bench: warmup done", file=sys.stderr, flush=True)
    energy = None
    if energy_theta:
        e0 = sim.total_energy_tree(energy_theta)

    sim.reset_force_timing()
    sim.synchronize()
    t0 = time.perf_counter()
    ms_events = sim.time_steps(args.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    phase, calls = sim.bh_phase_ms()
    ms_per_step = ms_events / args.steps
    value = n * args.steps / (ms_events * 1e-3)
    print(f"bench: {args.steps} steps {ms_events:.2f} ms (wall {wall * 1e3:.2f} ms)",
          file=sys.stderr, flush=True)
    if energy_theta:
        e1 = sim.total_energy_tree(energy_theta)
        energy = {"method": f"tree potential at theta={energy_theta} (nb_total_energy_tree)",
                  "E_start": float(e0[0]), "E_end": float(e1[0]),
                  "rel_drift": float((e1[0] - e0[0]) / abs(e0[0])), "steps": args.steps}

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
    cpu_line = cpu_baseline(n, args.cpu_sample, eps, dt) if cpu else None

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
                   "traverse": args.traverse, "group_width": width},
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
    print(json.dumps(out))
    sim.close()


if __name__ == "__main__":
    main()
