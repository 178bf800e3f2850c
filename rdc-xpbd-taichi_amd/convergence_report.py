"""Convergence reporting for XPBDNeoHookeanSolver runs.

Same functions and output formats as the reference's `convergence_report.py`
(lines 12-30 timing stats, 32-96 per-substep JSON export and plot, 164-176
callback, 179-195 frame print, 198-223 statistics, 226-293 analysis and CSV):
the JSON keys, the CSV columns and the file names are what the reference's
tools (`tools/summarize_errors.py`, `tools/compare_convergence.py`) read, so
runs of this solver can be compared with runs of the reference.

The records come from the solver's `convergence_history` (with
`monitor_convergence=True`): dicts with frame, substep, iter,
iter_compute_time_sec, hydro_error (mean |det F - 1|) and dev_error
(mean |‖F‖² - 3|) per Jacobi iteration, evaluated on the device.
"""
from __future__ import annotations

import csv
import json
import os
from collections import defaultdict
from pathlib import Path

import numpy as np

RECORD_FIELDS = ["frame", "substep", "iter", "iter_compute_time_sec", "hydro_error", "dev_error"]


def get_compute_timing_stats(solver) -> dict:
    calls = solver.solve_compute_calls
    iters = solver.solve_compute_iterations
    total = solver.solve_compute_time_total
    return {
        "solve_step_compute_count": calls,
        "constraint_iteration_count": iters,
        "total_compute_time_sec": total,
        "avg_compute_time_per_solve_step_sec": total / calls if calls > 0 else 0.0,
        "avg_compute_time_per_iteration_sec": total / iters if iters > 0 else 0.0,
    }


def substep_export(solver, frame, substep, convergence_data) -> dict:
    """The JSON document written for one substep."""
    return {
        "frame": frame,
        "substep": substep,
        "iterations": [d["iter"] for d in convergence_data],
        "iter_compute_times_sec": [d["iter_compute_time_sec"] for d in convergence_data],
        "hydro_errors": [d["hydro_error"] for d in convergence_data],
        "dev_errors": [d["dev_error"] for d in convergence_data],
        "records": list(convergence_data),
        "solver_type": solver.__class__.__name__,
        "compute_timing": get_compute_timing_stats(solver),
        "config": {
            "constraints_iter": solver.constraints_iter,
            "hydrostatic_relaxation": solver.hydrostatic_relaxation,
            "deviatoric_relaxation": solver.deviatoric_relaxation,
            "cheb_enable": solver.cheb_enable,
            "cheb_gamma": solver.cheb_gamma,
            "cheb_warmup": solver.cheb_warmup,
        },
    }


def save_substep_convergence(solver, frame, substep, convergence_data, save_path,
                             log_scale=False, export_data=True, plot=True):
    """Writes <save_path stem>.json (export_data) and the two-panel error plot at
    save_path (plot; needs matplotlib)."""
    if not convergence_data:
        return
    if export_data:
        with open(Path(save_path).with_suffix(".json"), "w", encoding="utf-8") as f:
            json.dump(substep_export(solver, frame, substep, convergence_data), f, indent=2)
    if plot:
        _plot_substep(frame, substep, convergence_data, save_path, log_scale)


def _plot_substep(frame, substep, data, save_path, log_scale):
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    its = [d["iter"] for d in data]
    fig, axes = plt.subplots(1, 2, figsize=(12, 4))
    panels = (("hydro_error", "Hydrostatic Error |det(F) - 1|", "Hydrostatic", "#2E86AB", "o"),
              ("dev_error", "Deviatoric Error ||F||^2 - 3|", "Deviatoric", "#A23B72", "s"))
    for ax, (key, ylabel, name, color, marker) in zip(axes, panels):
        err = [d[key] for d in data]
        ax.plot(its, err, marker + "-", color=color, linewidth=2, markersize=5, alpha=0.8)
        ax.set_xlabel("Iteration")
        ax.set_ylabel(ylabel)
        ax.set_title(f"Frame {frame}, Substep {substep} - {name}", fontweight="bold")
        ax.grid(True, alpha=0.3, linestyle="--")
        if log_scale and err and max(err) > 0:
            ax.set_yscale("log")
        if len(err) > 1:
            red = err[0] / err[-1] if err[-1] > 0 else float("inf")
            ax.text(0.95, 0.95, f"Initial: {err[0]:.2e}\nFinal: {err[-1]:.2e}\nReduction: {red:.1f}x",
                    transform=ax.transAxes, fontsize=9, va="top", ha="right",
                    bbox=dict(boxstyle="round", facecolor="wheat", alpha=0.7))
    fig.tight_layout()
    fig.savefig(save_path, dpi=150, bbox_inches="tight")
    plt.close(fig)


def setup_convergence_callback(solver, run_output_dir, frame_ids=None, verbose=False, plot=True):
    """Writes f{frame:04d}_s{substep:02d}.json (and .png) under
    <run_output_dir>/convergence_plots for the selected frames (all if None)."""
    plot_dir = os.path.join(run_output_dir, "convergence_plots")
    os.makedirs(plot_dir, exist_ok=True)
    wanted = set(frame_ids) if frame_ids else None

    def callback(frame, substep, convergence_data):
        if wanted is None or frame in wanted:
            path = os.path.join(plot_dir, f"f{frame:04d}_s{substep:02d}.png")
            save_substep_convergence(solver, frame, substep, convergence_data, path, plot=plot)
            if verbose:
                print(f"  Saved convergence plot: {path}")

    solver.convergence_callback = callback
    return plot_dir


def print_frame_convergence(solver, frame, substeps):
    rows = [d for d in solver.convergence_history
            if d["frame"] == frame and d["substep"] == substeps - 1]
    if not rows:
        return
    last = rows[-1]
    print(f"  Final substep {substeps - 1}, Iter {last['iter']}: "
          f"t_iter={last['iter_compute_time_sec'] * 1e3:.3f} ms, "
          f"H_err={last['hydro_error']:.3e}, D_err={last['dev_error']:.3e}")


def get_convergence_stats(solver):
    if not solver.monitor_convergence or not solver.convergence_history:
        return None
    by_sub = defaultdict(lambda: ([], []))
    for d in solver.convergence_history:
        h, v = by_sub[(d["frame"], d["substep"])]
        h.append(d["hydro_error"])
        v.append(d["dev_error"])
    fh = np.array([h[-1] for h, _ in by_sub.values()])
    fd = np.array([v[-1] for _, v in by_sub.values()])
    return {
        "total_records": len(solver.convergence_history),
        "num_substeps": len(by_sub),
        "hydro": {"mean_final": float(fh.mean()), "max_final": float(fh.max()),
                  "min_final": float(fh.min())},
        "dev": {"mean_final": float(fd.mean()), "max_final": float(fd.max()),
                "min_final": float(fd.min())},
    }


def write_convergence_csv(solver, path):
    with open(path, "w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=RECORD_FIELDS)
        w.writeheader()
        for d in solver.convergence_history:
            w.writerow({k: d[k] for k in RECORD_FIELDS})


def print_convergence_analysis(solver, iterations, run_output_dir, export_csv=False):
    if not solver.monitor_convergence:
        return
    print("\n" + "=" * 70)
    print("CONSTRAINT CONVERGENCE ANALYSIS")
    print("=" * 70)
    stats = get_convergence_stats(solver)
    if stats:
        print("\nData Collection:")
        print(f"  Total records: {stats['total_records']:,}")
        print(f"  Substeps monitored: {stats['num_substeps']}")
        for key, title in (("hydro", "Hydrostatic Constraint: C_H = det(F) - 1"),
                           ("dev", "Deviatoric Constraint: C_D = ||F||^2 - 3")):
            print(f"\n--- {title} ---")
            print(f"  Mean final error: {stats[key]['mean_final']:.6e}")
            print(f"  Max final error:  {stats[key]['max_final']:.6e}")
            print(f"  Min final error:  {stats[key]['min_final']:.6e}")
        first = [d for d in solver.convergence_history if d["iter"] == 0]
        last = [d for d in solver.convergence_history if d["iter"] == iterations - 1]
        if first and last:
            print("\n--- Convergence Rate Analysis ---")
            for key, name in (("hydro_error", "Hydrostatic"), ("dev_error", "Deviatoric: ")):
                a = float(np.mean([d[key] for d in first]))
                b = float(np.mean([d[key] for d in last]))
                red = a / b if b > 0 else float("inf")
                print(f"  {name}: {a:.3e} -> {b:.3e} ({red:.2f}x reduction)")
    if export_csv:
        path = os.path.join(run_output_dir, "convergence_data.csv")
        write_convergence_csv(solver, path)
        print(f"\nConvergence CSV: {path}")
    print("=" * 70)


__all__ = ["get_compute_timing_stats", "substep_export", "save_substep_convergence",
           "setup_convergence_callback", "print_frame_convergence", "get_convergence_stats",
           "write_convergence_csv", "print_convergence_analysis", "RECORD_FIELDS"]
