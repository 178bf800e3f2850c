"""Command-line driver: the reference's run_neohookean_usd.py on the MI355X solver.

Same flags and defaults (run_neohookean_usd.py:13-75), the same scene (a fixed
plane at y=-2 and a fixed sphere of radius 0.3 at (0,-1,0), the mesh offset by
(0,1,0), :82-85), the same output folders (output/<model>/xpbd_<method>_young_
<E>/, usd_export.py:10-34) and the same console report and convergence files.
The USD animation is written by this package's `usd_export` (text .usda, since
the `pxr` library is not available; a binary .usdc request is written as
.usda).  `--arch cpu` is refused: this solver runs on the GPU only.

Example (from rdc-xpbd-taichi_amd/):
  python run_neohookean.py --model assets/tetmesh/cow.node --frames 30 \\
      --monitor-convergence --export-csv
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="Run Neo-Hookean XPBD.")
    p.add_argument("--model", default="assets/tetmesh/cow.node")
    p.add_argument("--scale", default=1.0, type=float)
    p.add_argument("--arch", default="gpu", choices=["cpu", "gpu"])
    p.add_argument("--frames", default=300, type=int)
    p.add_argument("--fps", default=60, type=int)
    p.add_argument("--dt", default=0.01, type=float)
    p.add_argument("--substeps", default=1, type=int)
    p.add_argument("--iterations", default=200, type=int)
    p.add_argument("--young", default=5e8, type=float)
    p.add_argument("--poisson", default=0.4999, type=float)
    p.add_argument("--hydro-relax", default=1.2, type=float)
    p.add_argument("--dev-relax", default=1.2, type=float)
    m = p.add_mutually_exclusive_group()
    m.add_argument("--block-neohookean", dest="block_neohookean", action="store_true",
                   default=True, help="Use the block Neo-Hookean constraint solve (default)")
    m.add_argument("--split-neohookean", dest="block_neohookean", action="store_false",
                   help="Use the decoupled split Neo-Hookean baseline")
    c = p.add_mutually_exclusive_group()
    c.add_argument("--cheb", dest="cheb", action="store_true", default=True,
                   help="Use residual-driven dynamic Chebyshev acceleration (default)")
    c.add_argument("--no-cheb", dest="cheb", action="store_false",
                   help="Disable dynamic Chebyshev acceleration")
    p.add_argument("--cheb-warmup", default=5, type=int)
    p.add_argument("--cheb-gamma", default=0.666, type=float)
    p.add_argument("--optimize-usd", action="store_true")
    p.add_argument("--use-binary", action="store_true", default=True)
    p.add_argument("--streaming", action="store_true")
    p.add_argument("--usd-root-prim", default="/SimAsset")
    p.add_argument("--no-usd", action="store_true")
    p.add_argument("--monitor-convergence", action="store_true")
    p.add_argument("--print-convergence", action="store_true")
    p.add_argument("--plot-frames", action="store_true")
    p.add_argument("--plot-frame-ids", type=int, nargs="+", default=None)
    p.add_argument("--export-csv", action="store_true")
    p.add_argument("--output-root", default=None,
                   help="folder that holds output/ (default: the current directory)")
    return p


from usd_export import method_tag, young_tag  # noqa: E402  (usd_export.py:10-17)


def build_output_paths(model_path, block_neohookean, dynamic_cheby, young, root=None):
    """(output_dir, run_output_dir) of usd_export.build_output_paths."""
    import usd_export
    out, run_dir, _ = usd_export.build_output_paths(model_path, block_neohookean, dynamic_cheby,
                                                    young, root=root)
    return out, run_dir


def make_solver(args):
    import xpbd_gpu.neohookean_solver as xpbd
    from xpbd_gpu.sdf import PlaneSdf, SphereSdf
    sdf = [PlaneSdf(point=[0.0, -2.0, 0.0], dir=[0.0, 1.0, 0.0], fixed=True),
           SphereSdf(center=[0.0, -1.0, 0.0], radius=0.3, fixed=True)]
    solver = xpbd.XPBDNeoHookeanSolver(
        rest_pose=args.model, sdf=sdf, scale=args.scale, offset=(0.0, 1.0, 0.0),
        hydrostatic_relaxation=args.hydro_relax, deviatoric_relaxation=args.dev_relax,
        frame_dt=1.0 / args.fps, dt=args.dt, substeps_num=args.substeps,
        constraints_iter=args.iterations, young_modulus=args.young,
        poisson_ratio=args.poisson, cheb_enable=args.cheb, cheb_warmup=args.cheb_warmup,
        cheb_gamma=args.cheb_gamma, block_neohookean_enable=args.block_neohookean,
        monitor_convergence=args.monitor_convergence)
    return solver, sdf


def main(argv=None) -> int:
    args = build_parser().parse_args(argv)
    if args.arch != "gpu":
        print("run_neohookean.py: this solver runs on the GPU only (--arch gpu)", file=sys.stderr)
        return 2
    import convergence_report
    import usd_export

    solver, sdf = make_solver(args)
    _, run_dir, usd_path = usd_export.build_output_paths(
        args.model, args.block_neohookean, args.cheb, args.young, use_binary=args.use_binary,
        root=args.output_root)
    renderer = usd_export.create_renderer(args, usd_path, args.frames, args.fps)
    if renderer is not None:
        usd_path = renderer.filepath
    usd_export.add_solver_mesh(renderer, solver)
    usd_export.add_sdfs(renderer, sdf)
    plot_dir = None
    if args.monitor_convergence and args.plot_frames:
        plot_dir = convergence_report.setup_convergence_callback(
            solver, run_dir, frame_ids=args.plot_frame_ids, verbose=args.print_convergence)

    print("=" * 70)
    print("NEO-HOOKEAN XPBD SIMULATION")
    print("=" * 70)
    print(f"Model: {args.model}")
    print(f"Output directory: {run_dir}")
    print(f"Architecture: {args.arch.upper()}")
    print(f"Method: {'Block Neo-Hookean' if args.block_neohookean else 'Split Neo-Hookean'}")
    print(f"Dynamic Chebyshev: {'ON' if args.cheb else 'OFF'}")
    print(f"Frames: {args.frames}, substeps: {args.substeps}, iterations: {args.iterations}")
    print(f"Young's modulus: {args.young:.2e}, Poisson ratio: {args.poisson}")
    print(f"USD export: {'OFF' if renderer is None else usd_path}")
    print("=" * 70)

    times, usd_times = [], []
    t0 = time.time()
    for frame in range(args.frames):
        if frame % 10 == 0:
            print(f"\n[Frame {frame}/{args.frames}] ({frame / args.frames * 100:.1f}%)")
        s = time.time()
        solver.solve()
        times.append(time.time() - s)
        if args.monitor_convergence and args.print_convergence and not args.plot_frames:
            convergence_report.print_frame_convergence(solver, frame, args.substeps)
        if renderer is not None:
            s = time.time()
            renderer.render()
            usd_times.append(time.time() - s)
    save_time = 0.0
    if renderer is not None:
        print("\n" + "=" * 70)
        print("Saving USD file...")
        s = time.time()
        renderer.save()
        save_time = time.time() - s
    total = time.time() - t0
    solver.check_finite()

    print("\n" + "=" * 70)
    print("\nPERFORMANCE")
    print("=" * 70)
    print(f"Total runtime: {total:.3f} s")
    print(f"Solver total: {np.sum(times):.3f} s")
    print(f"Solver per frame: {np.mean(times):.4f} s")
    print(f"Solver FPS: {1.0 / np.mean(times):.2f}")
    print(f"USD save time: {save_time:.3f} s")
    if renderer is not None and usd_times:
        print(f"USD export total: {np.sum(usd_times):.3f} s")
        print(f"USD file size: {renderer.get_file_size_mb():.1f} MB")
    n_iter = args.frames * args.substeps * args.iterations
    print(f"Total constraint iterations: {n_iter:,}")
    print(f"Time per iteration: {np.sum(times) / n_iter * 1e6:.2f} us")
    convergence_report.print_convergence_analysis(solver, iterations=args.iterations,
                                                  run_output_dir=run_dir,
                                                  export_csv=args.export_csv)
    print("\nOUTPUTS")
    print("=" * 70)
    if renderer is not None:
        print(f"USD animation: {usd_path}")
    if plot_dir is not None:
        print(f"Convergence plots: {plot_dir}")
    print("=" * 70)
    return 0


if __name__ == "__main__":
    sys.exit(main())
