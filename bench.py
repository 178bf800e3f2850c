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
import gc
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
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_patch_project_jittered_box64_pmc.txt")
PMC_COMMANDS = """\
cd <repo root>; export TMPDIR=/tmp; D=<output folder>
rocprofv3 --pmc FETCH_SIZE --output-format csv -d $D/fetch -o run -- python3 bench.py --steps 1 --warmup 0 --no-cpu-baseline --roofline-frames 0 --plain-only
rocprofv3 --pmc WRITE_SIZE --output-format csv -d $D/write -o run -- python3 bench.py --steps 1 --warmup 0 --no-cpu-baseline --roofline-frames 0 --plain-only
rocprofv3 --pmc FETCH_SIZE --output-format csv -d $D/cfetch -o run -- python3 bench.py --copy-calibration-mib 512
rocprofv3 --pmc WRITE_SIZE --output-format csv -d $D/cwrite -o run -- python3 bench.py --copy-calibration-mib 512
python3 bench.py --summarize-pmc $D"""


def b_alg_iteration(T: int, V: int) -> int:
    """Algorithmic bytes of one Jacobi iteration (block + Chebyshev, f32 state,
    i32 indices, each vertex gathered once): 136 T + 140 V."""
    return 136 * T + 140 * V


def b_alg_project(T: int, V: int, iters: int) -> int:
    """Algorithmic bytes of one projection launch, averaged over the `iters`
    launches of a substep: per tet 16 (ids) + 36 (inv_Dm) + 4 (V0) + 8 + 8
    (lambda read/write); per vertex 16 (x, inv_m gather) + 12 (delta_x write);
    and per vertex 60 for the Chebyshev blend of the previous iteration (x_J,
    y^k, prev read 36; x, prev write 24), which every launch but a substep's
    first performs while staging its vertices.  With the average's 40 V and
    the residual's 64 T + 12 V this partitions 136 T + 140 V exactly."""
    return 72 * T + 28 * V + (60 * V * (iters - 1)) // iters


def _pmc_per_kernel(path: str) -> dict:
    """{kernel short name: [counter value per dispatch, ...]} from a rocprofv3
    counter_collection.csv (values in KiB for FETCH_SIZE / WRITE_SIZE)."""
    import csv
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
            out.setdefault(name, []).append(float(r["Counter_Value"]))
    return out


def summarize_pmc(folder: str, copy_mib: int = 512) -> str:
    """Per-dispatch HBM-side traffic of every kernel of one frame of the 64^3 jittered box, from the
    FETCH_SIZE and WRITE_SIZE passes, each scaled by what the same counter
    reads for a plain device copy of known size (the copy's own kernel)."""
    f = _pmc_per_kernel(os.path.join(folder, "fetch", "run_counter_collection.csv"))
    w = _pmc_per_kernel(os.path.join(folder, "write", "run_counter_collection.csv"))
    cf = _pmc_per_kernel(os.path.join(folder, "cfetch", "run_counter_collection.csv"))
    cw = _pmc_per_kernel(os.path.join(folder, "cwrite", "run_counter_collection.csv"))
    copy_kib = copy_mib * 1024.0
    ck = next(k for k in cf if "copyBuffer" in k)
    f_read = float(np.mean(cf[ck]))
    w_read = float(np.mean(cw[ck]))
    f_scale, w_scale = copy_kib / f_read, copy_kib / w_read
    lines = ["# HBM-side traffic per dispatch of a 64^3 jittered-box frame (1 572 864 tets,",
             "# 274 625 vertices, 10 substeps x 100 iterations), rocprofv3 FETCH_SIZE / WRITE_SIZE.",
             "# Made by:"]
    lines += ["#   " + c for c in PMC_COMMANDS.splitlines()]
    lines += [f"copy calibration: {ck}, {copy_mib} MiB each way per launch;",
              f"  FETCH_SIZE reads {f_read:.1f} KiB -> scale {f_scale:.4f};"
              f" WRITE_SIZE reads {w_read:.1f} KiB -> scale {w_scale:.4f}",
              "",
              f"{'kernel':28s} {'dispatches':>10s} {'FETCH KiB':>12s} {'WRITE KiB':>12s}"
              f" {'bytes/dispatch (scaled)':>24s}"]
    rows = []
    for k in sorted(set(f) | set(w)):
        if not k.startswith("xpbd::"):
            continue
        fv = float(np.mean(f.get(k, [0.0])))
        wv = float(np.mean(w.get(k, [0.0])))
        b = (fv * f_scale + wv * w_scale) * 1024.0
        rows.append((b * len(f.get(k, [])), k, len(f.get(k, [])), fv, wv, b))
    for _, k, n, fv, wv, b in sorted(rows, reverse=True):
        lines.append(f"{k:28s} {n:10d} {fv:12.1f} {wv:12.1f} {b:24.0f}")
    proj = next(r for r in rows if r[1].startswith("xpbd::k_patch_project"))
    lines += ["", f"traffic_bytes_per_launch {proj[1]}: {proj[5]:.0f}"]
    return "\n".join(lines) + "\n"


COUNTER_SUMMARY = os.path.join(ROOT, "profiles", "iteration_kernels_jittered_box64_counters.txt")
COUNTER_PASSES = {
    "valu": "SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_INSTS_VALU_FMA_F32 SQ_INSTS_VALU_TRANS_F32 "
            "SQ_INSTS_SALU SQ_WAVES SQ_BUSY_CYCLES SQ_WAVE_CYCLES GRBM_GUI_ACTIVE",
    "lds": "SQ_INSTS_LDS SQ_ACTIVE_INST_LDS SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE "
           "SQ_WAIT_INST_LDS SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_INST_CYCLES_VMEM_RD GRBM_GUI_ACTIVE",
}
COUNTER_COMMANDS = "\n".join(
    [f"rocprofv3 --pmc {c} --output-format csv -d $D/{k} -o run -- python3 bench.py --steps 1 "
     "--warmup 0 --no-cpu-baseline --roofline-frames 0 --plain-only"
     for k, c in COUNTER_PASSES.items()] +
    [PMC_COMMANDS.splitlines()[i] for i in (1, 2, 3, 4)] +
    ["rocprofv3 --kernel-trace --stats --output-format csv -d $D/stats -o run -- python3 bench.py "
     "--steps 3 --warmup 1 --no-cpu-baseline --plain-only",
     "python3 bench.py --summarize-counters $D"])


def summarize_counters(folder: str, copy_mib: int = 512) -> str:
    """The three kernels of a Jacobi iteration of the 64^3 jittered box, side
    by side: average duration (kernel trace), algorithmic and measured HBM
    bytes per launch with the fraction of 8 TB/s each gives, VALU issue
    (wave instructions per launch against one per CU per cycle over the
    launch's busy cycles), and LDS bank-conflict stalls."""
    import csv

    def per_kernel(path):
        out = {}
        with open(path) as fh:
            for r in csv.DictReader(fh):
                name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
                out.setdefault(name, {}).setdefault(r["Counter_Name"], []).append(
                    float(r["Counter_Value"]))
        return {k: {c: float(np.mean(v)) for c, v in d.items()} for k, d in out.items()}

    cnt = {}
    for k in COUNTER_PASSES:
        for name, d in per_kernel(os.path.join(folder, k, "run_counter_collection.csv")).items():
            cnt.setdefault(name, {}).update(d)
    with open(os.path.join(folder, "valu", "run_agent_info.csv")) as fh:
        agent = next(r for r in csv.DictReader(fh) if r.get("Name", "").startswith("gfx"))
    n_cu, n_se = int(agent["Cu_Count"]), int(agent["Array_Count"])
    f = _pmc_per_kernel(os.path.join(folder, "fetch", "run_counter_collection.csv"))
    w = _pmc_per_kernel(os.path.join(folder, "write", "run_counter_collection.csv"))
    cf = _pmc_per_kernel(os.path.join(folder, "cfetch", "run_counter_collection.csv"))
    cw = _pmc_per_kernel(os.path.join(folder, "cwrite", "run_counter_collection.csv"))
    ck = next(k for k in cf if "copyBuffer" in k)
    f_scale = copy_mib * 1024.0 / float(np.mean(cf[ck]))
    w_scale = copy_mib * 1024.0 / float(np.mean(cw[ck]))
    with open(os.path.join(folder, "stats", "run_kernel_stats.csv")) as fh:
        dur = {r["Name"].split("(")[0].replace("void ", "").strip(): float(r["AverageNs"]) / 1e3
               for r in csv.DictReader(fh)}
    T, V, iters = 1572864, 274625, 100
    alg = {"xpbd::k_patch_project<false>": b_alg_project(T, V, iters),
           "xpbd::k_patch_average": 40 * V, "xpbd::k_patch_residual": 64 * T + 12 * V}
    lines = ["# The kernels of one Jacobi iteration of the 64^3 jittered box (1 572 864 tets,",
             "# 274 625 vertices), from rocprofv3 counter passes over one frame (1000 launches",
             f"# each) on {agent['Name']} ({n_cu} CUs, {n_se} shader engines).  Made by:",
             "#   cd <repo root>; export TMPDIR=/tmp; D=<output folder>"]
    lines += ["#   " + c for c in COUNTER_COMMANDS.splitlines()]
    lines += ["#",
              "# alg MB: the algorithmic bytes per launch (136 T + 140 V per iteration, split",
              "#   72 T + 28 V + 60 V (it-1)/it | 40 V | 64 T + 12 V); meas MB: FETCH_SIZE +",
              "#   WRITE_SIZE scaled by a plain copy of known size; frac: MB / us against 8 TB/s.",
              "# VALU/tet: SQ_INSTS_VALU x 64 / T (lane instructions per tet); VALU issue: the",
              "#   wave instructions at one per CU per cycle over the launch's busy cycles",
              "#   (SQ_BUSY_CYCLES per shader engine); clock: those cycles over the duration.",
              "# LDS conflict: SQ_LDS_BANK_CONFLICT per CU over the busy cycles, and conflict",
              "#   cycles per access cycle (SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - it)).",
              "",
              f"{'kernel':28s} {'us':>6s} {'alg MB':>7s} {'frac':>5s} {'meas MB':>7s} {'frac':>5s}"
              f" {'VALU/tet':>8s} {'FMA%':>5s} {'VALU issue':>10s} {'clock':>6s}"
              f" {'LDS conflict':>12s} {'per acc':>7s}"]
    for k in ("xpbd::k_patch_project<false>", "xpbd::k_patch_residual", "xpbd::k_patch_average"):
        c = cnt.get(k, {})
        us = dur.get(k, float("nan"))
        meas = (float(np.mean(f[k])) * f_scale + float(np.mean(w[k])) * w_scale) * 1024.0
        busy = c.get("SQ_BUSY_CYCLES", 0.0) / n_se
        valu = c.get("SQ_INSTS_VALU", 0.0)
        conf = c.get("SQ_LDS_BANK_CONFLICT", 0.0)
        idx = c.get("SQ_LDS_IDX_ACTIVE", 0.0)
        lines.append(
            f"{k:28s} {us:6.2f} {alg[k] / 1e6:7.1f} {alg[k] / us / 1e6 / 8:5.2f}"
            f" {meas / 1e6:7.1f} {meas / us / 1e6 / 8:5.2f} {valu * 64 / T:8.0f}"
            f" {100 * c.get('SQ_INSTS_VALU_FMA_F32', 0.0) / max(valu, 1):5.1f}"
            f" {valu / n_cu / max(busy, 1):10.2f} {busy / us / 1e3:5.2f}G"
            f" {conf / n_cu / max(busy, 1):12.2f} {conf / max(idx - conf, 1):7.2f}")
    c = cnt.get("xpbd::k_patch_project<false>", {})
    busy = c.get("SQ_BUSY_CYCLES", 0.0) / n_se
    lines += ["", "valu_issue_frac xpbd::k_patch_project<false>: "
                  f"{c.get('SQ_INSTS_VALU', 0.0) / n_cu / max(busy, 1):.3f}",
              "lds_conflict_stall_frac xpbd::k_patch_project<false>: "
              f"{c.get('SQ_LDS_BANK_CONFLICT', 0.0) / n_cu / max(busy, 1):.3f}"]
    lines += ["", "raw counters per launch:"]
    for k in ("xpbd::k_patch_project<false>", "xpbd::k_patch_residual", "xpbd::k_patch_average"):
        lines.append(f"  {k}: " + ", ".join(f"{n} {v:.0f}" for n, v in sorted(cnt.get(k, {}).items())))
    return "\n".join(lines) + "\n"


STATS_COMMAND = ("cd <repo root>; export TMPDIR=/tmp; rocprofv3 --kernel-trace --stats "
                 "--output-format csv -d <D> -o run -- python3 bench.py --steps 3 --warmup 1 "
                 "--no-cpu-baseline --plain-only; python3 bench.py --summarize-stats <D>")


def summarize_stats(folder: str) -> str:
    """Per-kernel launch counts and durations from rocprofv3 --kernel-trace
    --stats, plus the projection kernel's average over the graph-replayed
    frames alone (warmup + timed steps, before the event-timed frames)."""
    import csv
    with open(os.path.join(folder, "run_kernel_stats.csv")) as fh:
        stats = list(csv.DictReader(fh))
    with open(os.path.join(folder, "run_kernel_trace.csv")) as fh:
        proj = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
                      for r in csv.DictReader(fh) if "k_patch_project" in r["Kernel_Name"])
    lines = ["# rocprofv3 kernel statistics of the 64^3 jittered-box bench (1 572 864 tets,",
             "# 274 625 vertices, 10 substeps x 100 iterations per frame): 1 warmup + 3 timed",
             "# graph frames, then the event-timed frames (8 projection replays per iteration,",
             "# then one event pair per projection).",
             "# Made by: " + STATS_COMMAND, "",
             f"{'kernel':30s} {'calls':>7s} {'avg us':>9s} {'min us':>9s} {'max us':>9s} {'%':>7s}"]
    for r in stats:
        name = r["Name"].split("(")[0].replace("void ", "").strip()
        lines.append(f"{name:30s} {int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:9.2f} "
                     f"{float(r['MinNs']) / 1e3:9.2f} {float(r['MaxNs']) / 1e3:9.2f} "
                     f"{float(r['Percentage']):7.2f}")
    graph = proj[:4000]
    if graph:
        avg = sum(e - b for b, e in graph) / len(graph) / 1e3
        lines += ["", f"xpbd::k_patch_project<false> over the {len(graph)} graph-replayed "
                      f"launches: avg {avg:.2f} us"]
    return "\n".join(lines) + "\n"


BATCHED_STATS_COMMAND = ("cd <repo root>; export TMPDIR=/tmp; rocprofv3 --kernel-trace --stats "
                         "--output-format csv -d <D> -o run -- python3 bench.py --batched-only "
                         "--no-tune; python3 bench.py --summarize-batched-stats <D>")


SMALL_STATS_COMMAND = ("cd <repo root>; export TMPDIR=/tmp; rocprofv3 --kernel-trace --stats "
                       "--output-format csv -d <D> -o run -- python3 bench.py --small-only; "
                       "python3 bench.py --summarize-small-stats <D>")


def _kernel_stats_table(folder: str, header: list) -> str:
    import csv
    with open(os.path.join(folder, "run_kernel_stats.csv")) as fh:
        stats = list(csv.DictReader(fh))
    lines = header + ["",
                      f"{'kernel':30s} {'calls':>7s} {'avg us':>9s} {'min us':>9s} {'max us':>9s}"
                      f" {'%':>7s}"]
    for r in stats:
        name = r["Name"].split("(")[0].replace("void ", "").strip()
        lines.append(f"{name:30s} {int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:9.2f} "
                     f"{float(r['MinNs']) / 1e3:9.2f} {float(r['MaxNs']) / 1e3:9.2f} "
                     f"{float(r['Percentage']):7.2f}")
    return "\n".join(lines) + "\n"


def summarize_small_stats(folder: str) -> str:
    """Per-kernel launch counts and durations of the small-configuration
    measurement (the 16^3 cube falling and resting, the 32^3 twisted grid)."""
    return _kernel_stats_table(folder, [
        "# rocprofv3 kernel statistics of the small single-scene configurations: the 16^3 cube",
        "# (24 576 tets, 5 substeps x 30 iterations per frame) falling and resting on the plane,",
        "# and the 32^3 twisted grid (196 608 tets, 200 iterations per frame); each 1 untimed +",
        "# 10 timed frames on the default path (auto: the cube's substeps run as one",
        "# k_substep_persistent launch each, the grid as phased kernels) and again on the graph",
        "# path (phased kernels for both); then batches of 2, 4, 8 and 14 copies of the 14^3",
        "# cube (5 x 30 iterations per frame, 1 untimed + 5 timed frames) on the batched",
        "# persistent substep (k_substep_persistent<true>) and on the phased kernels",
        "# (k_patch_project<true>).  Kernels of all runs are pooled per name.",
        "# Made by: " + SMALL_STATS_COMMAND])


def summarize_batched_stats(folder: str) -> str:
    """Per-kernel launch counts and durations of the batched measurement (256
    ragged scenes, default launch options) from rocprofv3 --kernel-trace --stats."""
    return _kernel_stats_table(folder, [
        "# rocprofv3 kernel statistics of the batched measurement: 256 box scenes, per-axis",
        "# resolution uniform in 6..40 (seed 0), 19 999 224 tets, 5 substeps x 50 iterations",
        "# per frame, default launch options (one graph per substep, bisection order),",
        "# 1 warmup + 3 timed batched frames.",
        "# Made by: " + BATCHED_STATS_COMMAND])


def committed_traffic(path: str = PMC_SUMMARY):
    """The projection kernel's measured bytes per launch from the committed
    counter summary, or None when it is absent."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch xpbd::k_patch_project"):
                    return int(float(line.rsplit(":", 1)[1]))
    except OSError:
        pass
    return None


def committed_counters(path: str = COUNTER_SUMMARY):
    """The projection kernel's VALU issue and LDS bank-conflict stall, as
    fractions of its busy cycles, from the committed counter summary (None
    when it is absent)."""
    out = {}
    try:
        with open(path) as fh:
            for line in fh:
                for key in ("valu_issue_frac", "lds_conflict_stall_frac"):
                    if line.startswith(key + " xpbd::k_patch_project"):
                        out[key] = float(line.rsplit(":", 1)[1])
    except OSError:
        return None
    if not out:
        return None
    out["source"] = os.path.relpath(path, ROOT)
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _cpu_leg(scene, substeps: int, threads: int) -> dict:
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
    return {"tet_iterations_per_s": tit / el, "seconds": el, "tet_iterations": tit,
            "substeps": substeps}


def cpu_baseline(scene, substeps: int, threads: int) -> dict:
    """The CPU restatement of the reference kernel sequence (the f32 oracle,
    OpenMP over tets, `threads` threads) on the host of the benchmarked GPU: a bounded
    sample of the headline mesh (`substeps` substeps of the 64^3 box) and, as
    a second leg, the 16^3 cube's 10 frames in full."""
    from xpbd_gpu.scenes import resting_cube
    big = _cpu_leg(scene, substeps, threads)
    cube = resting_cube()
    small = _cpu_leg(cube, 10 * cube.params["substeps_num"], threads)
    p = scene.params
    return {"value": big["tet_iterations_per_s"], "unit": "tet-iterations/s", "cores": threads,
            "kind": "port", "host_nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": f"{substeps} substeps x {p['constraints_iter']} iterations of the same "
                      f"64^3 mesh ({big['tet_iterations'] / 1e6:.0f} M tet-iterations, "
                      f"{big['seconds']:.1f} s), f32 oracle (CPU restatement of the reference "
                      f"kernel sequence), OpenMP over tets on {threads} threads",
            "small_leg": {"workload": "16^3 cube resting on the plane, 10 frames of 5 x 30 "
                                      "iterations (24 576 tets)",
                          "tet_iterations_per_s": small["tet_iterations_per_s"],
                          "seconds": small["seconds"]}}


def batched_measurement(n_scenes: int, frames: int, tune: bool, sequential: bool = True) -> dict:
    """256 ragged scenes in one batched launch sequence vs the same scenes
    solved one after another with plain per-scene solves."""
    from xpbd_gpu.batch import SceneBatch, autotune, autotune_ordering, default_candidates
    from xpbd_gpu.scenes import batched_scenes, make_solver

    scenes = batched_scenes(n_scenes=n_scenes)
    order, order_trials = "bisect", []
    if tune:  # the tet ordering is fixed at creation: pick it first
        order, order_trials = autotune_ordering(
            lambda o: [make_solver(s, reorder=o) for s in scenes])
        print(f"[bench] batched: ordering trials "
              f"{[(o, round(t * 1e3, 2)) for o, t in order_trials]} ms/substep",
              file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    solvers = [make_solver(s, reorder=order) for s in scenes]
    setup = time.perf_counter() - t0
    T = sum(s.n_tets for s in solvers)
    iters = scenes[0].params["constraints_iter"]
    subs = scenes[0].params["substeps_num"]
    print(f"[bench] batched: {len(solvers)} scenes, {T} tets, setup {setup:.1f} s",
          file=sys.stderr, flush=True)
    opts, trials = {}, []
    if tune:
        cands = [c for c in default_candidates(subs) if c.get("grid_pct", 100) == 100]
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
    workload = (f"{len(solvers)} box scenes, per-axis resolution uniform in 6..40 (seed 0), "
                f"{T} tets in total, {subs} substeps x {iters} iterations per frame")
    if not sequential:
        for s in solvers:
            s.close()
        return {"workload": workload, "tet_iterations_per_s": batched,
                "frame_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                             "p99": float(np.percentile(lat, 99) * 1e3)},
                "tuned_opts": dict(opts, tet_order=order)}
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
    for s in solvers:
        s.close()
    return {
        "workload": workload,
        "tet_iterations_per_s": batched,
        "sequential_plain_tet_iterations_per_s": sequential,
        "speedup_vs_sequential": batched / sequential,
        "frame_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                     "p99": float(np.percentile(lat, 99) * 1e3)},
        "sequential_frame_ms_per_scene": {"p50": float(np.percentile(seq_lat, 50) * 1e3),
                                          "p99": float(np.percentile(seq_lat, 99) * 1e3)},
        "sequential_all_scenes_ms": seq_el * 1e3,
        "tuned_opts": dict(opts, tet_order=order),
        "tuning_trials": len(trials) + len(order_trials),
        "order_trials_ms_per_substep": {o: round(t * 1e3, 3) for o, t in order_trials},
    }


def _served_run(srv, reqs) -> dict:
    """One trace through the server: latency and service percentiles, throughput,
    and where the host's time went (re-targeting, enqueueing, waiting)."""
    from xpbd_gpu.serve import summarize

    # scenes of earlier measurements that are only reachable through reference
    # cycles are freed (device frees synchronize) before the clock starts
    gc.collect()
    srv.reset_counters()
    t0 = time.perf_counter()
    done = srv.run(reqs)
    wall = time.perf_counter() - t0
    st = summarize(done, wall)
    busy = srv.build_s + srv.launch_s
    return {
        "requests": st["requests"],
        "p50_step_latency_ms": st["p50_ms"],
        "p99_step_latency_ms": st["p99_ms"],
        "service_p50_ms": st["service_p50_ms"],
        "service_p99_ms": st["service_p99_ms"],
        "tet_iterations_per_s": st["tet_iterations_per_s"],
        "makespan_s": st["makespan_s"],
        "batch_rebuilds": srv.rebuilds,
        "substeps": srv.substeps_run,
        "device_waits": srv.syncs,
        "host_s": {"wall": wall, "build": srv.build_s, "launch": srv.launch_s,
                   "wait": srv.wait_s},
        "host_busy_frac": busy / wall if wall > 0 else 0.0,
    }


def _sequential_run(scenes, reqs, solvers) -> dict:
    from xpbd_gpu.serve import run_sequential, summarize

    gc.collect()
    t0 = time.perf_counter()
    done = run_sequential(scenes, reqs, solvers=solvers)
    st = summarize(done, time.perf_counter() - t0)
    return {"requests": st["requests"],
            "p50_step_latency_ms": st["p50_ms"], "p99_step_latency_ms": st["p99_ms"],
            "service_p50_ms": st["service_p50_ms"], "service_p99_ms": st["service_p99_ms"],
            "tet_iterations_per_s": st["tet_iterations_per_s"], "makespan_s": st["makespan_s"]}


def served_measurement(n_scenes: int, rate: float, seconds: float, low_rate: float,
                       low_seconds: float, seq_seconds: float) -> dict:
    """Poisson request streams over resident scenes, continuous batching against
    the same requests served one at a time with plain solves.

    The stated trace (`rate` req/s over `seconds`) runs through the server.
    The sequential baseline cannot keep up with that rate, so it serves only
    the trace's first `seq_seconds` (its queue grows without bound; only its
    throughput is comparable), and the same prefix, queued at time 0, gives the
    server's capacity.  A second trace at `low_rate`, which the sequential
    baseline sustains, runs through both, so their latencies compare two
    queues that stay bounded."""
    from xpbd_gpu.scenes import make_solver
    from xpbd_gpu.serve import Request, Server, poisson_trace, served_scenes

    scenes = served_scenes(n_scenes=n_scenes)
    reqs = poisson_trace(n_scenes, rate=rate, duration=seconds, seed=1)
    srv = Server(scenes, hbm_budget=200e9)
    srv.warm(range(n_scenes))
    srv.run(poisson_trace(n_scenes, rate=rate, duration=0.3, seed=7))  # graphs, caches
    stated = _served_run(srv, reqs)
    print(f"[bench] served: {stated['requests']} requests at {rate:.0f}/s in "
          f"{stated['host_s']['wall']:.1f} s, p99 {stated['p99_step_latency_ms']:.1f} ms, "
          f"host busy {stated['host_busy_frac']:.2f}", file=sys.stderr, flush=True)
    prefix = [r for r in reqs if r.arrival < seq_seconds]
    cap = _served_run(srv, [Request(r.rid, r.scene, r.frames, 0.0) for r in prefix])
    low = poisson_trace(n_scenes, rate=low_rate, duration=low_seconds, seed=2)
    low_served = _served_run(srv, low)
    srv.close()
    seq_solvers = {}
    for k in range(n_scenes):
        seq_solvers[k] = make_solver(scenes[k])
        seq_solvers[k].solve()  # warm-up of per-scene graphs
    seq = _sequential_run(scenes, prefix, seq_solvers)
    print(f"[bench] sequential: {len(prefix)} requests in {seq['makespan_s']:.1f} s",
          file=sys.stderr, flush=True)
    low_seq = _sequential_run(scenes, low, seq_solvers)
    for s in seq_solvers.values():
        s.close()
    tets = np.array([s.mesh.n_tets for s in scenes])
    per_frame = [s.mesh.n_tets * s.params["substeps_num"] * s.params["constraints_iter"]
                 for s in scenes]
    offered = sum(per_frame[r.scene] * r.frames for r in reqs) / seconds
    return {
        "workload": f"{n_scenes} resident scenes, tet counts log-normal median 2e4 in [1e3, 5e5] "
                    f"(seed 0; min {tets.min()}, max {tets.max()}), Poisson {rate:.0f} req/s over "
                    f"{seconds:.0f} s (seed 1), 1-4 frames per request, HBM budget 200 GB",
        **stated,
        "offered_load_tet_iterations_per_s": offered,
        "speedup_vs_sequential": stated["tet_iterations_per_s"] / seq["tet_iterations_per_s"],
        "speedup_note": "the server keeps up with the stated rate, so its throughput here is the "
                        "trace's offered load and this ratio is the load's, not the server's; "
                        "capacity.speedup_vs_sequential compares the two at saturation",
        "sequential_plain": {
            "note": f"the first {seq_seconds:.0f} s of the same trace ({len(prefix)} requests); "
                    "the sequential queue grows without bound at this rate, so its latencies "
                    "measure queueing and only its throughput compares",
            **seq},
        "capacity": {
            "note": f"the first {seq_seconds:.0f} s of the trace queued at time 0",
            **{k: cap[k] for k in ("requests", "tet_iterations_per_s", "makespan_s",
                                   "host_busy_frac")},
            "speedup_vs_sequential": cap["tet_iterations_per_s"] / seq["tet_iterations_per_s"]},
        "sustained_rate": {
            "workload": f"Poisson {low_rate:.0f} req/s over {low_seconds:.0f} s (seed 2), the "
                        "same scenes; a rate the sequential baseline keeps up with",
            "batched": low_served,
            "sequential_plain": low_seq,
        },
    }


def small_configs_measurement(frames: int = 10) -> dict:
    """Throughput of the two smaller single-scene configurations: the 16^3
    cube (24 576 tets, 5 substeps x 30 iterations per frame; as stated it
    falls from 0.5, and the same cube resting on the plane is the contact
    case) and the 32^3 twisted grid (196 608 tets, one substep of 200
    iterations per frame, relaxing from its stretched and twisted start).
    Each runs one untimed frame, then `frames` timed frames, on the default
    execution path ("auto": the persistent substep kernel for the cube, the
    graph of phased kernels for the grid) and, for comparison, on the graph
    path."""
    from xpbd_gpu.scenes import dropped_cube, make_solver, resting_cube, twisted_grid

    out = {}
    runs = []
    for name, scene in (("dropped_cube16", dropped_cube()), ("resting_cube16", resting_cube()),
                        ("twisted_grid32", twisted_grid())):
        runs.append((name, scene, "auto"))
        runs.append((name + "_graph", scene, "graph"))
    for name, scene, path in runs:
        solver = make_solver(scene, exec_path=path)
        T, V = scene.mesh.n_tets, scene.mesh.n_verts
        it_per_frame = scene.params["substeps_num"] * scene.params["constraints_iter"]
        solver.solve()
        solver.synchronize()
        lat = []
        t0 = time.perf_counter()
        for _ in range(frames):
            s0 = time.perf_counter()
            solver.solve()
            lat.append(time.perf_counter() - s0)
        el = time.perf_counter() - t0
        solver.check_finite()
        out[name] = {"tets": T, "verts": V, "exec_path": path,
                     "iterations_per_frame": it_per_frame,
                     "frames": frames, "ms_per_frame": el / frames * 1e3,
                     "frame_p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "tet_iterations_per_s": T * it_per_frame * frames / el,
                     "us_per_iteration": el / (frames * it_per_frame) * 1e6,
                     "hbm_gbps_algorithmic":
                         b_alg_iteration(T, V) * it_per_frame * frames / el / 1e9}
        solver.close()
    return out


def small_batch_measurement(sizes=(2, 4, 8, 14), frames: int = 5) -> dict:
    """Batches of small scenes on the persistent substep against the phased
    kernels (graph): k copies of the 14^3 cube resting on the plane (16 464
    tets, 33 patches, 17 workgroups each), 5 substeps x 30 iterations per
    frame, one untimed frame, then `frames` timed frames per path."""
    from xpbd_gpu.batch import SceneBatch
    from xpbd_gpu.scenes import make_solver, resting_cube

    out = {}
    for k in sizes:
        row = {"scenes": k}
        for per in ("always", "never"):
            sols = [make_solver(resting_cube(n=14, frames=frames + 1)) for _ in range(k)]
            with SceneBatch(sols, persistent=per) as b:
                b.frame(1)
                b.synchronize()
                t0 = time.perf_counter()
                b.frame(frames)
                b.synchronize()
                el = time.perf_counter() - t0
                it = frames * sols[0].substeps_num * sols[0].constraints_iter
                row["patches"] = b.n_patches
                row[b.path + "_us_per_iteration"] = el / it * 1e6
            for s in sols:
                s.close()
        out[f"{k}_scenes"] = row
        print(f"[bench] small batch {k} x cube14: {row}", file=sys.stderr, flush=True)
    return out


def presets_measurement(frames: int = 301) -> dict:
    """The reference's two scripted presets, each through the command-line
    driver's solver set-up (run_neohookean.make_solver, the same flags) on its
    own TetGen mesh (taken from the golden fixtures, which hold the reference's
    meshes), on the default execution path:
    scripts/demo.bat: the cow at scale 1.5, E 5e8, nu 0.4999, dt 0.01 at
      60 fps (two time chunks per frame), one substep of 200 iterations,
      block + Chebyshev, 301 frames;
    scripts/compare_convergence_armadillo.bat: the armadillo at scale 0.01 with
      the same parameters and the convergence monitor on, 301 frames for each
      of its four methods (split and block, each without and with Chebyshev).
    Positions are read back after every frame, as the driver's USD writer
    does; USD, plots and reports are not written."""
    import tempfile

    import run_neohookean
    from xpbd_gpu.mesh import TetMesh, write_tetgen

    golden = os.path.join(ROOT, "tests", "golden")
    common = ["--young", "500000000", "--dt", "0.01", "--fps", "60", "--frames", str(frames),
              "--substeps", "1", "--iterations", "200", "--no-usd"]
    out = {}
    with tempfile.TemporaryDirectory() as d:
        models = {}
        for name, fixture in (("cow", "cow_demo_contact"), ("armadillo",
                                                           "armadillo_compare_block_dc")):
            with np.load(os.path.join(golden, fixture + ".npz"), allow_pickle=False) as z:
                write_tetgen(os.path.join(d, name), TetMesh(z["verts"], z["tets"]))
            models[name] = os.path.join(d, name + ".node")
        runs = [("demo_cow", ["--model", models["cow"], "--scale", "1.5", "--block-neohookean",
                              "--cheb"])]
        for tag, flags in (("NH", ["--split-neohookean", "--no-cheb"]),
                           ("NH-DC", ["--split-neohookean", "--cheb"]),
                           ("BNH", ["--block-neohookean", "--no-cheb"]),
                           ("BNH-DC", ["--block-neohookean", "--cheb"])):
            runs.append(("armadillo_" + tag, ["--model", models["armadillo"], "--scale", "0.01",
                                              "--monitor-convergence", *flags]))
        for name, argv in runs:
            args = run_neohookean.build_parser().parse_args(argv + common)
            solver, _ = run_neohookean.make_solver(args)
            t0 = time.perf_counter()
            for _ in range(frames):
                solver.solve()
                solver.mesh.verts.x.to_numpy()
            el = time.perf_counter() - t0
            solver.check_finite()
            n_tets = int(solver.n_tets)
            # dt 0.01 at 60 fps: each frame is a 0.01 s and a 1/150 s chunk
            it_per_frame = 2 * args.substeps * args.iterations
            out[name] = {"tets": n_tets, "frames": frames, "seconds": el,
                         "ms_per_frame": el / frames * 1e3,
                         "iterations_per_frame": it_per_frame,
                         "tet_iterations_per_s": n_tets * it_per_frame * frames / el}
            print(f"[bench] preset {name}: {n_tets} tets, {frames} frames in {el:.2f} s",
                  file=sys.stderr, flush=True)
            solver.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="hexes per axis")
    ap.add_argument("--accumulate", default="csr", choices=["atomic", "csr"])
    ap.add_argument("--layout", default="patch", choices=["patch", "flat"])
    ap.add_argument("--exec-path", default="auto",
                    choices=["auto", "graph", "persistent", "kernels"])
    ap.add_argument("--cpu-substeps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-frames", type=int, default=1)
    ap.add_argument("--batched-scenes", type=int, default=256)
    ap.add_argument("--batched-frames", type=int, default=3)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--serve-scenes", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=400.0)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-low-rate", type=float, default=60.0)
    ap.add_argument("--serve-low-seconds", type=float, default=20.0)
    ap.add_argument("--serve-seq-seconds", type=float, default=10.0,
                    help="prefix of the stated trace that the sequential baseline serves")
    ap.add_argument("--plain-only", action="store_true",
                    help="skip the batched and served measurements")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this size (counter calibration)")
    ap.add_argument("--summarize-pmc", default="",
                    help="only summarize the FETCH_SIZE/WRITE_SIZE passes in this folder")
    ap.add_argument("--summarize-counters", default="",
                    help="only summarize the counter passes of the iteration kernels in this folder")
    ap.add_argument("--summarize-stats", default="",
                    help="only summarize a rocprofv3 --kernel-trace --stats folder")
    ap.add_argument("--timing-replays", type=int, default=8,
                    help="back-to-back projection replays per event pair for the roofline")
    ap.add_argument("--presets-only", action="store_true",
                    help="time only the reference's scripted presets (demo, armadillo comparison)")
    ap.add_argument("--small-only", action="store_true",
                    help="only measure the small single-scene configurations, print JSON")
    ap.add_argument("--summarize-small-stats", default="",
                    help="only summarize a --kernel-trace --stats folder of --small-only")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the batched frames (no sequential comparison), print JSON")
    ap.add_argument("--summarize-batched-stats", default="",
                    help="only summarize a rocprofv3 --stats folder of a --batched-only run")
    args = ap.parse_args()

    if args.summarize_batched_stats:
        sys.stdout.write(summarize_batched_stats(args.summarize_batched_stats))
        return
    if args.summarize_small_stats:
        sys.stdout.write(summarize_small_stats(args.summarize_small_stats))
        return
    if args.summarize_counters:
        sys.stdout.write(summarize_counters(args.summarize_counters))
        return
    if args.summarize_pmc:
        sys.stdout.write(summarize_pmc(args.summarize_pmc))
        return
    if args.summarize_stats:
        sys.stdout.write(summarize_stats(args.summarize_stats))
        return

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
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    if args.small_only:
        print(json.dumps({"small_configs": small_configs_measurement(),
                          "small_batches": small_batch_measurement()}), flush=True)
        return

    if args.presets_only:
        print(json.dumps(presets_measurement()), flush=True)
        return

    if args.batched_only:
        print(json.dumps(batched_measurement(args.batched_scenes, args.batched_frames,
                                             not args.no_tune, sequential=False)), flush=True)
        return

    from xpbd_gpu.scenes import jittered_box, make_solver

    scene = jittered_box(n=args.n)
    T, V = scene.mesh.n_tets, scene.mesh.n_verts
    iters = scene.params["constraints_iter"]
    subs = scene.params["substeps_num"]
    print(f"[bench] workload: {V} verts, {T} tets, {subs} substeps x {iters} iterations "
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

    # dominant kernel: HIP events on the solver's stream around runs of
    # back-to-back projection replays (one event pair per run, so the pair's
    # own overhead is spread over the run), then around single launches
    def timed_avg(replays):
        solver.set_kernel_timing(True, replays=replays)
        for _ in range(args.roofline_frames):
            solver.solve()
        ms, n = solver.kernel_time()
        solver.set_kernel_timing(False)
        return ms / 1e3 / max(1, n), n

    avg_s, launches = timed_avg(max(2, args.timing_replays))
    avg_single_s, _ = timed_avg(1)
    b_proj = b_alg_project(T, V, iters)
    ach = b_proj / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic = committed_traffic() if (args.n == 64 and args.layout == "patch") else None
    counters = committed_counters() if (args.n == 64 and args.layout == "patch") else None
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "kernel": "xpbd::k_patch_project<false>" if args.layout == "patch"
                          else "xpbd::k_project",
                "avg_us": round(avg_s * 1e6, 2), "launches": launches,
                "avg_us_one_launch_per_event_pair": round(avg_single_s * 1e6, 2),
                "bytes_per_launch": b_proj,
                # the same kernel by rocprofv3 counters: the fraction of its busy
                # cycles that issue a VALU instruction, and LDS conflict stalls
                "counters": counters}

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
            "workload": f"single large mesh: {args.n}^3 jittered box, {T} tets, {V} verts, "
                        f"{subs} substeps x {iters} iterations per step (one frame)",
            "accumulate_mode": args.accumulate,
            "layout": args.layout,
            "exec_path": args.exec_path,
        },
        "hbm_gbps_algorithmic": gbps,
        "hbm_frac_frame": round(gbps / HBM_PEAK_GBPS, 4),  # all kernels of a frame together
        "b_alg_per_iteration": b_alg_iteration(T, V),
        "step_latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                            "p99": float(np.percentile(lat, 99) * 1e3)},
        "roofline": roofline,
    }
    solver.close()
    del solver
    if not args.plain_only:
        out["small_configs"] = small_configs_measurement()
        out["small_batches"] = small_batch_measurement()
        out["presets"] = presets_measurement()
        out["batched"] = batched_measurement(args.batched_scenes, args.batched_frames,
                                             not args.no_tune)
        out["served"] = served_measurement(args.serve_scenes, args.serve_rate,
                                           args.serve_seconds, args.serve_low_rate,
                                           args.serve_low_seconds, args.serve_seq_seconds)
    if not args.no_cpu_baseline:
        # the box's CPU share is 16 threads per GPU, whatever nproc shows
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_substeps, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
