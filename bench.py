"""Benchmark of the XPBD Neo-Hookean substep on one MI355X.

Workload (one "step" = one frame of the single-large-mesh configuration):
  64^3 hex box jittered +-0.2 cell (seed 0), 6 tets/hex -> 274 625 vertices,
  1 572 864 tets; E=1e6, nu=0.48, density 1000; plane y=0 + fixed sphere r=0.3;
  dt = 1/60, 10 substeps x 100 Chebyshev-accelerated block Jacobi iterations.

Prints one JSON line: tet-iterations/s over the timed frames (inputs resident
in HBM), the roofline of the dominant kernel (algorithmic bytes / HIP-event
duration against 8 TB/s), the per-frame latency percentiles, and the CPU
oracle timed on the host cores over a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "rdc-xpbd-taichi_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBPS = 8000.0


def b_alg_iteration(T: int, V: int) -> int:
    """Algorithmic bytes of one Jacobi iteration (block + Chebyshev, f32 state,
    i32 indices, each vertex gathered once): 136 T + 140 V."""
    return 136 * T + 140 * V


def b_alg_project(T: int, V: int) -> int:
    """Algorithmic bytes of one projection launch: per tet 16 (ids) + 36
    (inv_Dm) + 4 (V0) + 8 + 8 (lambda read/write); per vertex 16 (x, inv_m
    gather) + 12 (delta_x write)."""
    return 72 * T + 28 * V


def cpu_baseline(scene, substeps: int, threads: int) -> dict:
    from oracle import Oracle  # the CPU restatement, used here only as the baseline
    p = dict(scene.params)
    o = Oracle(scene.mesh.verts, scene.mesh.tets, scene.sdfs, precision=32, threads=threads,
               gravity=scene.gravity, **p)
    o.compute_lumped_mass(scene.total_mass)
    o.set_log(False)
    sub_dt = p["dt"] / p["substeps_num"]
    t0 = time.perf_counter()
    for _ in range(substeps):
        o.substep(sub_dt)
    el = time.perf_counter() - t0
    tit = scene.mesh.n_tets * p["constraints_iter"] * substeps
    return {"value": tit / el, "unit": "tet-iterations/s", "cores": threads, "kind": "port",
            "sample": f"{substeps} substeps x {p['constraints_iter']} iterations of the same "
                      f"mesh ({tit / 1e6:.0f} M tet-iterations, {el:.1f} s), f32 oracle, "
                      f"OpenMP over tets"}


def batched_measurement(n_scenes: int, frames: int, tune: bool) -> dict:
    """256 ragged scenes in one batched launch sequence vs the same scenes
    solved one after another with plain per-scene solves."""
    from xpbd_gpu.batch import SceneBatch, autotune, default_candidates
    from xpbd_gpu.scenes import batched_scenes, make_solver

    scenes = batched_scenes(n_scenes=n_scenes)
    t0 = time.perf_counter()
    solvers = [make_solver(s) for s in scenes]
    setup = time.perf_counter() - t0
    T = sum(s.n_tets for s in solvers)
    iters = scenes[0].params["constraints_iter"]
    subs = scenes[0].params["substeps_num"]
    print(f"[bench] batched: {len(solvers)} scenes, {T} tets, setup {setup:.1f} s",
          file=sys.stderr, flush=True)
    opts, trials = {}, []
    if tune:
        cands = [c for c in default_candidates(subs) if c["grid_pct"] == 100]
        opts, trials = autotune(solvers, candidates=cands)
    with SceneBatch(solvers, **opts) as b:
        b.frame(1)  # warm-up (graph capture)
        b.synchronize()
        lat = []
        t0 = time.perf_counter()
        for _ in range(frames):
            s0 = time.perf_counter()
            b.frame(1)
            b.synchronize()
            lat.append(time.perf_counter() - s0)
        el = time.perf_counter() - t0
    batched = T * iters * subs * frames / el
    # sequential plain solves of the same scenes, one frame each
    for s in solvers[:4]:
        s.solve()  # warm-up of per-scene graphs
    seq_lat = []
    t0 = time.perf_counter()
    for s in solvers:
        s0 = time.perf_counter()
        s.solve()
        seq_lat.append(time.perf_counter() - s0)
    seq_el = time.perf_counter() - t0
    sequential = T * iters * subs / seq_el
    return {
        "workload": f"{len(solvers)} box scenes, per-axis resolution uniform in 6..40 (seed 0), "
                    f"{T} tets in total, {subs} substeps x {iters} iterations per frame",
        "tet_iterations_per_s": batched,
        "sequential_plain_tet_iterations_per_s": sequential,
        "speedup_vs_sequential": batched / sequential,
        "frame_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                     "p99": float(np.percentile(lat, 99) * 1e3)},
        "sequential_frame_ms_per_scene": {"p50": float(np.percentile(seq_lat, 50) * 1e3),
                                          "p99": float(np.percentile(seq_lat, 99) * 1e3)},
        "sequential_all_scenes_ms": seq_el * 1e3,
        "tuned_opts": opts,
        "tuning_trials": len(trials),
    }


def served_measurement(n_scenes: int, rate: float, seconds: float) -> dict:
    """Poisson request stream over resident scenes: continuous batching vs the
    same trace served one request at a time with plain solves."""
    from xpbd_gpu.serve import Server, poisson_trace, run_sequential, served_scenes, summarize

    scenes = served_scenes(n_scenes=n_scenes)
    reqs = poisson_trace(n_scenes, rate=rate, duration=seconds, seed=1)
    srv = Server(scenes, hbm_budget=200e9)
    srv.warm(range(n_scenes))
    warm = [r for r in poisson_trace(n_scenes, rate=rate, duration=0.2, seed=7)]
    srv.run(warm)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    wall = time.perf_counter() - t0
    served = summarize(done, wall)
    print(f"[bench] served: {len(reqs)} requests in {wall:.1f} s, p99 {served['p99_ms']:.1f} ms",
          file=sys.stderr, flush=True)
    seq_solvers = {}
    for k in range(n_scenes):
        from xpbd_gpu.scenes import make_solver
        seq_solvers[k] = make_solver(scenes[k])
        seq_solvers[k].solve()  # warm-up of per-scene graphs
    t0 = time.perf_counter()
    sdone = run_sequential(scenes, reqs, solvers=seq_solvers)
    swall = time.perf_counter() - t0
    seq = summarize(sdone, swall)
    tets = np.array([s.mesh.n_tets for s in scenes])
    return {
        "workload": f"{n_scenes} resident scenes, tet counts log-normal median 2e4 in [1e3, 5e5] "
                    f"(seed 0; min {tets.min()}, max {tets.max()}), Poisson {rate:.0f} req/s over "
                    f"{seconds:.0f} s (seed 1), 1-4 frames per request, HBM budget 200 GB",
        "requests": served["requests"],
        "p50_step_latency_ms": served["p50_ms"],
        "p99_step_latency_ms": served["p99_ms"],
        "tet_iterations_per_s": served["tet_iterations_per_s"],
        "sequential_plain": {"p50_step_latency_ms": seq["p50_ms"],
                             "p99_step_latency_ms": seq["p99_ms"],
                             "tet_iterations_per_s": seq["tet_iterations_per_s"],
                             "makespan_s": seq["makespan_s"]},
        "speedup_vs_sequential": served["tet_iterations_per_s"] / seq["tet_iterations_per_s"],
        "batch_rebuilds": srv.rebuilds,
        "substeps": srv.substeps_run,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="hexes per axis")
    ap.add_argument("--accumulate", default="csr", choices=["atomic", "csr"])
    ap.add_argument("--layout", default="patch", choices=["patch", "flat"])
    ap.add_argument("--exec-path", default="graph", choices=["graph", "kernels"])
    ap.add_argument("--cpu-substeps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-frames", type=int, default=1)
    ap.add_argument("--batched-scenes", type=int, default=256)
    ap.add_argument("--batched-frames", type=int, default=3)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--serve-scenes", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=100.0)
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--plain-only", action="store_true",
                    help="skip the batched and served measurements")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this size (counter calibration)")
    args = ap.parse_args()

    if args.copy_calibration_mib:
        # A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE readings
        # give the scale between those counters and bytes moved.
        import torch
        n = args.copy_calibration_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    from xpbd_gpu.scenes import cfg2_plain_large, make_solver

    scene = cfg2_plain_large(n=args.n)
    T, V = scene.mesh.n_tets, scene.mesh.n_verts
    iters = scene.params["constraints_iter"]
    subs = scene.params["substeps_num"]
    print(f"[bench] workload cfg2: {V} verts, {T} tets, {subs} substeps x {iters} iterations "
          f"per frame; accumulate={args.accumulate} exec={args.exec_path}", file=sys.stderr,
          flush=True)
    solver = make_solver(scene, accumulate_mode=args.accumulate, exec_path=args.exec_path,
                         layout=args.layout)

    for _ in range(args.warmup):
        solver.solve()
    solver.synchronize()

    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        solver.solve()  # ends with a stream synchronization
        lat.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - t0
    solver.check_finite()
    tet_iters = T * iters * subs * args.steps
    value = tet_iters / elapsed
    gbps = b_alg_iteration(T, V) * iters * subs * args.steps / elapsed / 1e9

    # dominant kernel: HIP events around every projection launch (its stream)
    solver.set_kernel_timing(True)
    for _ in range(args.roofline_frames):
        solver.solve()
    ms, launches = solver.kernel_time()
    solver.set_kernel_timing(False)
    avg_s = ms / 1e3 / max(1, launches)
    ach = b_alg_project(T, V) / avg_s / 1e9 if avg_s > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": "k_project", "avg_us": round(avg_s * 1e6, 2), "launches": launches,
                "bytes_per_launch": b_alg_project(T, V)}

    out = {
        "metric": "tet-iterations/s & HBM GB/s at 1.57M tets; p99 step latency batched vs plain",
        "value": value,
        "unit": "tet-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"cfg2 single large mesh: {args.n}^3 jittered box, {T} tets, {V} verts, "
                        f"{subs} substeps x {iters} iterations per step (one frame)",
            "accumulate_mode": args.accumulate,
            "layout": args.layout,
            "exec_path": args.exec_path,
        },
        "hbm_gbps_algorithmic": gbps,
        "b_alg_per_iteration": b_alg_iteration(T, V),
        "step_latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                            "p99": float(np.percentile(lat, 99) * 1e3)},
        "roofline": roofline,
    }
    del solver
    if not args.plain_only:
        out["batched"] = batched_measurement(args.batched_scenes, args.batched_frames,
                                             not args.no_tune)
        out["served"] = served_measurement(args.serve_scenes, args.serve_rate,
                                           args.serve_seconds)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_substeps, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
