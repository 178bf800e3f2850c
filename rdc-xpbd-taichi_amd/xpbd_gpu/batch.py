"""Batched form: many independent scenes advanced by shared kernel launches.

A `SceneBatch` concatenates the patches and vertex blocks of its member
`XPBDNeoHookeanSolver` scenes into one work list, so each kernel of a substep
(and one replayed hipGraph per substep or per frame) covers every scene.
Batches of small scenes (each at most 288 patches, 512 in all) take the
persistent substep instead: one launch per substep, in which each member's
workgroups meet at that member's own barriers (option `persistent`).
Scenes keep their own buffers, materials, SDFs and rho/omega state; the
results of a scene are bit-identical alone and in any batch.

`autotune` picks the batch's launch knobs (vertex workgroup size, persistent
grid size, XCD-contiguous work split, substeps per graph, and whether a
substep's iterations run in one persistent launch) by timing them on the
actual scenes, restoring the scenes' state after each trial;
`autotune_exec_path` does the same for one scene's execution path.
"""
from __future__ import annotations

import ctypes
import itertools
import time

from . import _abi

DEFAULT_OPTS = dict(graph=True, graph_substeps=1, vertex_block=256, grid_pct=100,
                    xcd_split=True, persistent="auto")

_PERSISTENT = {"never": 0, False: 0, "auto": 1, "always": 2, True: 2}


def _opts_struct(opts: dict) -> _abi.XpbdBatchOpts:
    o = _abi.XpbdBatchOpts()
    _abi.check(_abi.load().xpbd_batch_opts_init(ctypes.byref(o)))
    o.graph = int(bool(opts.get("graph", True)))
    o.graph_substeps = int(opts.get("graph_substeps", 1))
    o.vertex_block = int(opts.get("vertex_block", 256))
    o.grid_pct = int(opts.get("grid_pct", 100))
    o.xcd_split = int(bool(opts.get("xcd_split", True)))
    per = opts.get("persistent", "auto")
    if per not in _PERSISTENT:
        raise ValueError(f"persistent must be one of 'never', 'auto', 'always' (got {per!r})")
    o.persistent = _PERSISTENT[per]
    return o


class SceneBatch:
    """Member scenes stepped together (one substep or frame of every member)."""

    def __init__(self, solvers, **opts):
        self._lib = _abi.load()
        self.solvers = list(solvers)
        if not self.solvers:
            raise ValueError("a batch needs at least one scene")
        self.opts = dict(DEFAULT_OPTS)
        self.opts.update(opts)
        for s in self.solvers:
            s._sync_params()
            s.synchronize()
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h for s in self.solvers])
        h = ctypes.c_void_p()
        o = _opts_struct(self.opts)
        _abi.check(self._lib.xpbd_batch_create(arr, len(self.solvers), ctypes.byref(o),
                                               ctypes.byref(h)))
        self._h = h.value
        self._refresh_info()

    def _refresh_info(self):
        ns = ctypes.c_int32()
        nv = ctypes.c_int64()
        nt = ctypes.c_int64()
        npch = ctypes.c_int32()
        _abi.check(self._lib.xpbd_batch_info(self._h, ctypes.byref(ns), ctypes.byref(nv),
                                             ctypes.byref(nt), ctypes.byref(npch)))
        self.n_scenes, self.n_verts, self.n_tets, self.n_patches = (ns.value, nv.value, nt.value,
                                                                     npch.value)
        s0 = self.solvers[0]
        self.iters = int(s0.constraints_iter)
        self.substeps_num = int(s0.substeps_num)

    def grid(self) -> dict:
        """Persistent grids of the patch kernels and the most scenes one
        projection workgroup's patch range touches (xpbd_batch_grid)."""
        gp = ctypes.c_int32()
        gr = ctypes.c_int32()
        ms = ctypes.c_int32()
        _abi.check(self._lib.xpbd_batch_grid(self._h, ctypes.byref(gp), ctypes.byref(gr),
                                             ctypes.byref(ms)))
        return {"project": gp.value, "residual": gr.value, "max_scenes_per_range": ms.value}

    @property
    def path(self) -> str:
        """The path the next substeps take: "persistent" (one launch per
        substep), "graph" or "kernels" (xpbd_batch_path)."""
        v = ctypes.c_int32()
        _abi.check(self._lib.xpbd_batch_path(self._h, ctypes.byref(v)))
        return {_abi.PATH_PERSISTENT: "persistent", _abi.PATH_GRAPH: "graph",
                _abi.PATH_KERNELS: "kernels"}[v.value]

    def set_barrier_timeout(self, microseconds: int):
        """xpbd_batch_set_barrier_timeout: the persistent substep's barrier wait
        (default 2 s)."""
        _abi.check(self._lib.xpbd_batch_set_barrier_timeout(self._h, int(microseconds)))

    def set_members(self, solvers):
        """Re-targets the batch at another list of scenes (xpbd_batch_set_members):
        the stream and large-enough device buffers are kept; queued work is
        waited for."""
        solvers = list(solvers)
        if not solvers:
            raise ValueError("a batch needs at least one scene")
        for s in solvers:
            s._sync_params()
        arr = (ctypes.c_void_p * len(solvers))(*[s._h for s in solvers])
        _abi.check(self._lib.xpbd_batch_set_members(self._h, arr, len(solvers)))
        self.solvers = solvers
        self._refresh_info()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.xpbd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def substep(self, n: int = 1):
        """n substeps of every member (asynchronous)."""
        for _ in range(n):
            _abi.check(self._lib.xpbd_batch_substep(self._h))

    def frame(self, n: int = 1):
        """n frames of every member (asynchronous); member counters follow."""
        _abi.check(self._lib.xpbd_batch_frame(self._h, int(n)))

    def synchronize(self):
        _abi.check(self._lib.xpbd_batch_synchronize(self._h))
        for s in self.solvers:
            s._refresh_counters()

    @property
    def tet_iterations_per_substep(self) -> int:
        return self.n_tets * self.iters


def _snapshot(solvers):
    return [(s.mesh.verts.x.to_numpy(), s.mesh.verts.v.to_numpy(), s.rho_hat, s.current_frame,
             s.time) for s in solvers]


def _restore(solvers, snap):
    for s, (x, v, rho, frame, t) in zip(solvers, snap):
        s.mesh.verts.x.from_numpy(x)
        s.mesh.verts.v.from_numpy(v)
        s.current_frame = frame
        s.time = t
        _abi.check(s._lib.xpbd_set_state(s._h, int(frame), float(t), float(rho)))


def default_candidates(substeps_num: int):
    """The phased kernels' launch knobs, and the persistent substep (all of a
    substep's iterations in one launch; skipped by `autotune` where it does
    not apply)."""
    grid = itertools.product([128, 256, 512], [50, 100], [True, False],
                             sorted({1, int(substeps_num)}))
    # graph_substeps > 1 holds whole frames (dt == frame_dt)
    return [dict(graph=True, vertex_block=vb, grid_pct=gp, xcd_split=xs, graph_substeps=gs,
                 persistent="never")
            for vb, gp, xs, gs in grid] + [dict(persistent="always")]


def autotune(solvers, candidates=None, substeps: int | None = None, warmup: int = 1):
    """Times each candidate option set on the scenes and returns
    (best_opts, [(opts, seconds_per_substep), ...]).  Scene state (positions,
    velocities, rho_hat, counters) is restored after every trial.  A candidate
    that does not apply to these scenes (persistent="always" for scenes too
    large to be resident together) is listed with an infinite time."""
    solvers = list(solvers)
    s0 = solvers[0]
    if candidates is None:
        candidates = default_candidates(s0.substeps_num)
    snap = _snapshot(solvers)
    results = []
    try:
        for opts in candidates:
            per = int(opts.get("graph_substeps", 1))
            n = substeps if substeps is not None else max(2, per)
            n = max(per, (n // per) * per)
            with SceneBatch(solvers, **opts) as b:
                if opts.get("persistent") == "always" and b.path != "persistent":
                    results.append((dict(opts), float("inf")))
                    continue
                # whole-frame graphs are replayed by frame(); single substeps by substep()
                step = (lambda k: b.frame(k // per)) if per > 1 else b.substep
                step(max(warmup, 0) * per)
                b.synchronize()
                t0 = time.perf_counter()
                step(n)
                b.synchronize()
                el = (time.perf_counter() - t0) / n
            results.append((dict(opts), el))
            _restore(solvers, snap)
    finally:
        _restore(solvers, snap)
    best = min(results, key=lambda r: r[1])[0]
    return best, results


def autotune_exec_path(solver, paths=("persistent", "graph", "kernels"), frames: int = 1,
                       warmup: int = 1):
    """Times one scene's execution paths, `frames` frames of solve() each after
    `warmup`, and leaves the solver on the fastest.  The paths differ in
    iterations per launch: "persistent" runs all of a substep's iterations in
    one launch (where the scene is small enough, else it takes the graph),
    "graph" and "kernels" one launch per kernel, replayed from a graph or not;
    every path gives the same bits.  The scene's state, frame counters,
    convergence history and solve timers are restored after each trial.
    Returns (best_path, [(path, seconds_per_frame), ...])."""
    snap = _snapshot([solver])
    hist = len(solver.convergence_history)
    timers = (solver.solve_compute_time_total, solver.solve_compute_calls,
              solver.solve_compute_iterations)

    def restore():
        _restore([solver], snap)
        del solver.convergence_history[hist:]
        (solver.solve_compute_time_total, solver.solve_compute_calls,
         solver.solve_compute_iterations) = timers

    results = []
    try:
        for path in paths:
            solver.exec_path = path
            for _ in range(max(warmup, 0)):
                solver.solve()
            solver.synchronize()
            t0 = time.perf_counter()
            for _ in range(frames):
                solver.solve()
            solver.synchronize()
            results.append((path, (time.perf_counter() - t0) / max(1, frames)))
            restore()
    finally:
        restore()
    best = min(results, key=lambda r: r[1])[0]
    solver.exec_path = best
    return best, results


def autotune_ordering(build, orders=("bisect", "morton", "none"), opts=None,
                      substeps: int = 2, warmup: int = 1):
    """Picks the internal tet/vertex ordering of the scenes: for each order,
    build(order) creates the scenes (a list of solvers, with reorder=order),
    a batch with launch options `opts` (default DEFAULT_OPTS) runs `warmup`
    then `substeps` timed substeps, and the scenes are freed.  Returns
    (best_order, [(order, seconds_per_substep), ...])."""
    opts = dict(DEFAULT_OPTS if opts is None else opts)
    opts["graph_substeps"] = 1
    results = []
    for order in orders:
        solvers = build(order)
        try:
            with SceneBatch(solvers, **opts) as b:
                b.substep(max(warmup, 0))
                b.synchronize()
                t0 = time.perf_counter()
                b.substep(substeps)
                b.synchronize()
                results.append((order, (time.perf_counter() - t0) / max(1, substeps)))
        finally:
            for s in solvers:
                s.close()
            del solvers
    best = min(results, key=lambda r: r[1])[0]
    return best, results


def run_sequential(solvers, frames: int = 1):
    """The same scenes solved one after another with plain per-scene solves."""
    for s in solvers:
        for _ in range(frames):
            s.solve()


def scene_bytes(solver) -> int:
    b = ctypes.c_int64()
    _abi.check(_abi.load().xpbd_get_device_bytes(solver._h, ctypes.byref(b)))
    return int(b.value)


__all__ = ["SceneBatch", "autotune", "autotune_ordering", "default_candidates",
           "run_sequential", "scene_bytes", "DEFAULT_OPTS"]
