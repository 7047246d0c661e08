"""Throughput of the MI355X MD-step solver on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched]
    python bench.py --calib-copy      # plain 512 MiB device copies, for HBM counter scaling
    python bench.py --bilayer-only    # the flat bilayer of 999,698 lipids (2,999,094 beads)

Workload (`value`): the 80,000-lipid vesicle (240,000 beads) from the liposome generator
(seed 1234, areal density 3.45, box 400^3, dt 0.02, T 3, gamma 1, tension 0 so no box
moves), built in memory with the product's generator (byte-identical to the reference
generator).  A step is one full MD step: Verlet first half + cell key, radix sort,
reorder + cell table, pair force, CHAIN forces + Langevin + second half-kick.  Inputs are
resident in HBM before timing starts; no measurement or file output inside the timed loop.

`batched`: the 64-replica tension x temperature sweep (softmold_amd.sweep) stepped as one
batch after the autotuner picked its launch shape, with box moves every 8 steps, next to
the same replicas run one after another through the single-system solver.

Prints one JSON line: particle-steps/s, the pair kernel's roofline record (algorithmic
bytes per launch over its HIP-event-timed average duration), the batched sweep, and the
CPU baseline (the reference MD binary built from source in oracle/_ref, or the FP64
oracle port when it is absent), timed on this host on a bounded sample of the plain
workload.  The same CPU leg reports the FP32-vs-FP64 error of the workload at step 0
(`fp64_check`), with the FP64 oracle as the checker.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1 GPU, plain vs batched FP32, and % of HBM roofline"
HBM_PEAK_GBS = 8000.0
VALU_FP32_PEAK_TFLOPS = 157.3          # published vector FP32 peak (counts packed FP32 FMA)
VALU_FP32_UNPACKED_TFLOPS = 78.6       # the same without packed FP32, which this build disables
                                       # (and the FP64 vector rate of the kernel's exact test)
SEED, NLIPIDS, DENSITY = 1234, 80000, 3.45
# HBM bytes per pair_force launch on this workload, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs scaled by the same counters on a copy of known size (bench.py --calib-copy)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pair_force_vesicle240k_traffic.json")
COUNTERS_FILE = os.path.join(ROOT, "profiles", "pair_force_vesicle240k_counters.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stencil_stats(keys: np.ndarray, nc) -> tuple[int, float]:
    """Occupied cells and mean 27-stencil candidates per bead."""
    nx, ny, nz = (int(x) for x in nc)
    uk, cnt = np.unique(keys, return_counts=True)
    occ = dict(zip(uk.tolist(), cnt.tolist()))
    x, y, z = uk % nx, (uk // nx) % ny, uk // (nx * ny)
    cand = np.zeros(len(uk))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                kk = ((x + dx) % nx) + ((y + dy) % ny) * nx + ((z + dz) % nz) * nx * ny
                cand += np.array([occ.get(int(k), 0) for k in kk])
    return len(uk), float((cand * cnt).sum() / cnt.sum())


def cpu_baseline(n_beads: int) -> dict:
    """Reference MD binary (oracle/_ref/MD, compiled from the reference sources) on this
    host: (wall of S steps - wall of 0 steps) / S with store and measure disabled."""
    import softmold_amd as sm

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = int(os.environ.get("OMP_NUM_THREADS", cores))
    md = os.path.join(ROOT, "oracle", "_ref", "MD")
    work = tempfile.mkdtemp(prefix="smcpu_")
    try:
        m = sm.Mpd.liposome(SEED, NLIPIDS, DENSITY)
        m["storeInterval"] = 1.0e6
        m["measureInterval"] = 1.0e6
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        if os.access(md, os.X_OK):
            def run(steps):
                m["finalTime"] = steps * 0.02
                m["initialTime"] = 0.0
                m.write(os.path.join(work, "cpu"))
                t0 = time.perf_counter()
                subprocess.run([md, "cpu"], cwd=work, env=env, check=True, capture_output=True)
                return time.perf_counter() - t0
            t_zero = run(0)
            t_probe = run(10) - t_zero
            steps = int(max(20, min(2000, 15.0 / max(t_probe / 10, 1e-4))))
            dt = run(steps) - t_zero
            return {"value": n_beads * steps / dt, "unit": "particle-steps/s", "cores": cores,
                    "kind": "reference",
                    "sample": f"MD.cpp built from source, {steps} steps of the same vesicle "
                              f"(wall of {steps} steps minus wall of 0 steps), OMP_NUM_THREADS={cores}"}
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.set_threads(cores)
        d = m.system()
        s = O.System(d["pos"], d["type"], d["vel"], d["box"], d["fconst"], d["uconst"], d["n_types"],
                     d["chain"], d["chain_const"], cutoff=d["cutoff"], dt=d["dt"], gamma=d["gamma"],
                     temp=d["temperature"], seed=d["seed"])
        mdo = O.MD(s, rng_mode=1)
        mdo.init_forces()
        t0 = time.perf_counter()
        mdo.run(0, 9)
        probe = (time.perf_counter() - t0) / 10
        steps = int(max(10, min(1000, 15.0 / max(probe, 1e-4))))
        t0 = time.perf_counter()
        mdo.run(10, 10 + steps - 1)
        dt = time.perf_counter() - t0
        return {"value": n_beads * steps / dt, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                "sample": f"FP64 oracle restatement of MD.cpp, {steps} steps of the same vesicle"}
    finally:
        shutil.rmtree(work, ignore_errors=True)


def fp64_check(d: dict, cores: int) -> dict:
    """The FP32-vs-FP64 error of this workload at step 0, with the FP64 oracle as the
    checker: conservative forces (pair + CHAIN) per particle, relative to
    max(|F_i|, 0.05 rms|F|), and in L2; |dU|/|U| and |dK|/K of the device's FP64 sums."""
    import softmold_amd as sm
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.set_threads(cores)
    s = sm.Solver.from_system(d)
    fp, fc = s.compute_forces()
    upair, uchain, kin = s.measure()
    s.close()
    o = O.System(d["pos"], d["type"], d["vel"], d["box"], d["fconst"], d["uconst"], d["n_types"], d["chain"],
                 d["chain_const"], cutoff=d["cutoff"])
    want = o.pair_force() + o.chain_force()
    got = fp + fc
    err = np.linalg.norm(got - want, axis=1)
    scale = np.maximum(np.linalg.norm(want, axis=1), 0.05 * np.sqrt((want ** 2).sum(1).mean()))
    u_ref = o.pair_potential() + o.chain_potential()
    k_ref = o.kinetic()
    return {"force_max_rel": float((err / scale).max()),
            "force_l2_rel": float(np.linalg.norm(got - want) / np.linalg.norm(want)),
            "u_rel": float(abs(upair + uchain - u_ref) / abs(u_ref)),
            "k_rel": float(abs(kin - k_ref) / k_ref),
            "checker": "FP64 oracle restatement of MD.cpp, step 0 of the same vesicle"}


def bench_batched(steps: int, warmup: int, seq_steps: int, fixed: dict | None = None) -> dict:
    """The 64-replica tension x temperature sweep stepped as one batch, against the same
    replicas run one after another through the single-system solver (seq_steps 0: no
    sequential run).  fixed: knob values to step with instead of the autotuner's choice."""
    import torch

    import softmold_amd as sm
    from softmold_amd.sweep import replica_sweep

    systems = replica_sweep()
    nb = sum(len(d["type"]) for d in systems)
    b = sm.Solver.batch(systems)
    b.init_forces()
    tuned_ms = None
    if fixed:
        for k, v in fixed.items():
            b.tune(k, v)
    else:
        tuned_ms = b.autotune(32)
    knobs = {k: b.tune_get(k) for k in ("pair_chunk", "pair_variant", "graph", "sort_mode")}
    b.run(1, warmup)
    b.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(1 + warmup, steps)
    b.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    trials, accepted = b.mc_stats()
    b.close()
    batched = nb * steps / wall
    log(f"batched: {len(systems)} replicas, {nb} beads, {1e3 * wall / steps:.3f} ms/step, "
        f"{batched:.4g} particle-steps/s, knobs {knobs}, MC acceptance {accepted.sum() / max(trials.sum(), 1):.3f}")

    # the plain solver on each replica alone, at that replica's own best launch shape: the
    # autotuner picks its pair chunk, sort mode and the faster of captured graph and eager
    # stepping (where a step without a box move ends in the fused kick + next-first-half
    # kernel); the same steps and warmup as the batch, summed walls
    seq_wall, seq_work, modes = 0.0, 0, {"graph": 0, "eager": 0}
    for d in (systems if seq_steps > 0 else []):
        s = sm.Solver.from_system(d)
        s.init_forces()
        s.autotune(32)
        modes["graph" if s.tune_get("graph") else "eager"] += 1
        s.run(1, warmup)
        s.sync()
        t0 = time.perf_counter()
        s.run(1 + warmup, seq_steps)
        s.sync()
        seq_wall += time.perf_counter() - t0
        seq_work += len(d["type"]) * seq_steps
        s.close()
    seq = seq_work / seq_wall if seq_steps > 0 else None
    if seq:
        log(f"sequential plain solver: {seq:.4g} particle-steps/s ({seq_steps} steps per replica, "
            f"autotuned: {modes})")
    return {"value": batched, "unit": "particle-steps/s", "replicas": len(systems), "beads": nb,
            "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * wall / steps,
            "sequential_value": seq, "sequential_steps_per_replica": seq_steps,
            "sequential_modes": modes,
            "speedup_vs_sequential": batched / seq if seq else None, ("knobs" if fixed else "autotuned"): knobs,
            "autotune_best_ms_per_step": tuned_ms,
            "workload": "64 liposome replicas (seeds 1000..1063), lipid counts log-uniform in "
                        "[2000, 40000] (MT19937 seed 20261202), tension linspace(0,1,8) x "
                        "temperature 3*linspace(0.9,1.1,8), deltaLXY 0.01, MC box move every 8 steps"}


def bench_bilayer(steps: int, warmup: int) -> dict:
    """The reference generator's flat bilayer `bilayer 1234 1000000 3.45` (999,698 lipids,
    2,999,094 beads, box 538.38^2 x 40, deltaLXY 0.01: a box move every 8 steps)."""
    import torch

    import softmold_amd as sm

    d = sm.Mpd.bilayer(1234, 1000000, 3.45).system()
    n = len(d["type"])
    s = sm.Solver.from_system(d)
    s.init_forces()
    s.run(0, 1 + warmup)
    s.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(1 + warmup, steps)
    s.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    trials, accepted = s.mc_stats()
    s.close()
    return {"value": n * steps / wall, "unit": "particle-steps/s", "beads": n, "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * wall / steps, "mc_trials": trials, "mc_accepted": accepted,
            "workload": "flat bilayer, reference generator bilayer 1234 1000000 3.45 (999,698 lipids), "
                        "deltaLXY 0.01, box move every 8 steps"}


def calib_copy():
    """Five copies of a 512 MiB buffer: the reference dispatch for scaling HBM counters."""
    import torch

    torch.cuda.set_device(0)
    nbytes = 512 << 20
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "copies": 5}), flush=True)


def valu_issue() -> float | None:
    """VALU issue fraction of the pair kernel on this workload, from its SQ counters
    (profiles/pair_force_vesicle240k_counters.json): VALU wave-instructions over what the
    chip can issue in the kernel's active cycles -- the bound this kernel is held to."""
    try:
        with open(COUNTERS_FILE) as f:
            return float(json.load(f)["valu_issue_frac"])
    except (OSError, KeyError, ValueError):
        return None


def traffic() -> float | None:
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return float(t["bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=1000)
    ap.add_argument("--calib-copy", action="store_true")
    ap.add_argument("--batched-only", action="store_true", help="only the replica sweep (for profiling it)")
    ap.add_argument("--bilayer-only", action="store_true", help="only the 1M-lipid flat bilayer (for profiling it)")
    ap.add_argument("--batched-knobs", default=None, metavar="CHUNK,SORT_MODE,GRAPH",
                    help="with --batched-only: step the batch with these knobs, without the autotuner "
                         "and without the sequential baseline")
    ap.add_argument("--sort-mode", type=int, choices=(0, 1, 2), default=None,
                    help="the plain workload's cell sort: 0 every key, 1 incremental, 2 every key in a window "
                         "(default: the solver's)")
    args = ap.parse_args()
    if args.calib_copy:
        calib_copy()
        return
    if args.bilayer_only:
        print(json.dumps({"bilayer": bench_bilayer(args.steps, args.warmup)}), flush=True)
        return
    if args.batched_only:
        fixed, seq_steps = None, args.batched_steps
        if args.batched_knobs:
            chunk, sort_mode, graph = (int(v) for v in args.batched_knobs.split(","))
            fixed, seq_steps = {"pair_chunk": chunk, "sort_mode": sort_mode, "graph": graph}, 0
        print(json.dumps({"batched": bench_batched(args.batched_steps, 100, seq_steps, fixed)}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the solver has no multi-GPU path "
              "(a replica sweep would shard one replica subset per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch

    import softmold_amd as sm

    assert torch.cuda.is_available(), "bench.py needs the MI355X"
    torch.cuda.set_device(0)
    d = sm.Mpd.liposome(SEED, NLIPIDS, DENSITY).system()
    n = len(d["type"])
    s = sm.Solver.from_system(d)
    if args.sort_mode is not None:
        s.tune("sort_mode", args.sort_mode)
    keys = s.cell_keys()
    nc = (d["box"] / d["cutoff"]).astype(int)
    occ, cand = stencil_stats(keys, nc)
    _, npairs = s.pair_count()
    s.init_forces()

    # warmup, then the timed loop bracketed by device synchronisation
    s.run(0, args.warmup)
    s.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(args.warmup, args.steps)
    s.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = n * args.steps / wall

    # per-kernel averages from HIP events on the solver's stream, same number of steps
    s.profile(True)
    s.run(args.warmup + args.steps, args.steps)
    prof = s.profile_read()
    s.profile(False)
    sort_mode = s.tune_get("sort_mode")
    window_path = s.tune_get("sort_window_path") == 1
    pair_ms = prof["pair_force"][0]
    for k, (avg, cnt) in prof.items():
        if cnt:
            log(f"  {k:26s} {avg * 1e3:9.2f} us x {cnt}")
    algo_bytes = 28.0 * n + 8.0 * occ      # p16 read + f12 write per bead, 8 B per occupied cell
    achieved = algo_bytes / (pair_ms * 1e-3) / 1e9
    accepted = 2.0 * npairs / n
    flops = n * (9.0 * cand + 14.0 * accepted)
    tflops = flops / (pair_ms * 1e-3) / 1e12
    log(f"N={n} occupied cells={occ} candidates/bead={cand:.1f} neighbours/bead={accepted:.1f}")
    # the whole step against HBM: the per-bead algorithmic bytes of an unfused FP32 step
    # (integrate + key 76, radix sort 16 per pass, reorder 36, pair 28, bonded 40,
    # scatter 16, second half + Langevin 36) and 8 B per occupied cell; the sort takes
    # ceil(bits / 8) passes over the cell indices, or two over the keys of a window of at
    # most 2^18 cells
    bits = max(1, int(np.ceil(np.log2(float(np.prod(nc))))))
    passes = 2 if window_path else (bits + 7) // 8
    step_bytes = (232.0 + 16.0 * passes) * n + 8.0 * occ
    step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": "f32 pair arithmetic; f64 positions, bonded forces, energies and dU",
        "data": "synthetic (liposome generator output, seed 1234)",
        "config": {"workload": "plain solver, vesicle 80000 lipids = 240000 beads, box 400^3, "
                               "dt 0.02, T 3, gamma 1, tension 0",
                   "n_beads": n, "parallelism": "single GPU"},
        # the dominant kernel (the pair force, two thirds of the step) is bound by VALU
        # issue: its algorithmic flops (9 per stencil candidate + 14 per neighbour, per bead)
        # against the published FP32 vector peak, with its compulsory HBM bytes beside them
        "roofline": {"bound": "valu", "achieved": tflops, "peak": VALU_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / VALU_FP32_PEAK_TFLOPS, "traffic": traffic(),
                     "kernel": "pair_force", "kernel_avg_us": pair_ms * 1e3,
                     "algorithmic_flops_per_launch": flops,
                     "frac_unpacked_fp32": tflops / VALU_FP32_UNPACKED_TFLOPS,
                     "valu_issue_frac": valu_issue(),
                     "hbm": {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": algo_bytes}},
        "step_roofline": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_step": step_bytes,
                          "sort_passes": passes, "sort_mode": sort_mode},
    }
    s.close()
    if not args.no_batched:
        out["batched"] = bench_batched(args.batched_steps, 100, args.batched_steps)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(n)
        except Exception as e:  # the GPU number stands on its own
            log("cpu baseline failed:", e)
            out["cpu_baseline"] = None
        try:
            cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
            out["fp64_check"] = fp64_check(d, int(os.environ.get("OMP_NUM_THREADS", cores)))
        except Exception as e:
            log("fp64 check failed:", e)
            out["fp64_check"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
