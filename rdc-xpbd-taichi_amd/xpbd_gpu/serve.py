"""Served form: a continuous-batching request loop over device-resident scenes.

Each request asks to advance one scene by a number of frames.  Scenes are kept
resident in HBM: created on first use and admitted only while the sum of their
device footprints fits the budget; idle scenes are evicted least recently used
first, their state parked in host memory and restored when they return.  The
loop admits the requests that have arrived and whose scene is free and
resident, then enqueues substeps of all active scenes (`SceneBatch`: for few
small scenes one persistent launch per substep, otherwise one batched launch
sequence per substep replayed from a captured graph while the membership
holds) up to the next request completion -- at most one frame, so
that new arrivals wait at most that long for admission -- and only then waits
for the device and retires the completed requests.  Requests that arrive while
others run join the batch at the next such boundary, so their work overlaps on
the device.

`run_sequential` serves the same trace one request at a time with plain
per-scene solves, the baseline the batched loop is compared with.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass

import numpy as np

from .batch import SceneBatch, scene_bytes
from .mesh import box_grid
from .scenes import Scene, base_params, make_solver, plane


@dataclass
class Request:
    rid: int
    scene: int
    frames: int
    arrival: float  # seconds after the start of the run


@dataclass
class Done:
    rid: int
    scene: int
    frames: int
    arrival: float
    start: float
    finish: float
    tet_iterations: int

    @property
    def latency(self) -> float:
        """Arrival to completion."""
        return self.finish - self.arrival

    @property
    def service(self) -> float:
        """Admission (start of its first substep) to completion."""
        return self.finish - self.start


def served_scenes(n_scenes: int = 64, seed: int = 0, median_tets: float = 2e4,
                  sigma: float = 1.0, tet_range=(1e3, 5e5), substeps: int = 5,
                  iters: int = 50, cell: float = 1.0 / 32.0) -> list:
    """Scenes of the served configuration: tet counts log-normal with the given
    median (clipped to tet_range), built as seeded jittered (+-0.2 cell) box
    grids with 6 tets per hex; E log-uniform in [1e4, 1e7], nu uniform in
    [0.3, 0.49], density 1000, plane y=0; dt = 1/60, 5 substeps x 50 iterations."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_scenes):
        t = float(np.clip(np.exp(rng.normal(np.log(median_tets), sigma)), *tet_range))
        hexes = t / 6.0
        # a box of aspect close to 1 with about `hexes` cells
        a = max(1, int(round(hexes ** (1.0 / 3.0))))
        nx = a
        ny = max(1, int(round((hexes / nx) ** 0.5)))
        nz = max(1, int(round(hexes / (nx * ny))))
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        nu = float(rng.uniform(0.3, 0.49))
        size = (nx * cell, ny * cell, nz * cell)
        m = box_grid(nx, ny, nz, size=size, origin=(0.0, 0.05, 0.0), jitter=0.2,
                     seed=seed * 100003 + i)
        vol = size[0] * size[1] * size[2]
        out.append(Scene(f"served{i}", m, [plane(0.0)],
                         base_params(young_modulus=E, poisson_ratio=nu, substeps_num=substeps,
                                     constraints_iter=iters),
                         total_mass=1000.0 * vol, frames=1, extra=dict(dims=(nx, ny, nz))))
    return out


def poisson_trace(n_scenes: int, rate: float, duration: float, seed: int = 0,
                  frames=(1, 4)) -> list:
    """Poisson arrivals at `rate` per second over `duration` seconds, each for a
    uniformly chosen scene and 1-4 frames."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        out.append(Request(len(out), int(rng.integers(n_scenes)),
                           int(rng.integers(frames[0], frames[1] + 1)), t))
    return out


class Server:
    """Resident scenes, admission by device-memory budget, continuous batching."""

    def __init__(self, scenes: list, hbm_budget: float = 200e9, batch_opts=None,
                 solver_opts=None, max_run: int | None = None):
        self.specs = list(scenes)
        self.budget = float(hbm_budget)
        self.batch_opts = dict(graph=True)
        self.batch_opts.update(batch_opts or {})
        # substeps enqueued between two waits for the device (default: a frame)
        self.max_run = max_run
        self.solver_opts = dict(solver_opts or {})
        self.resident = collections.OrderedDict()  # scene -> solver, least recently used first
        self.bytes = {}
        self.footprint = {}  # scene -> device bytes, measured when first built
        self.parked = {}  # evicted scenes' state, restored when they return
        self.rebuilds = 0
        self.substeps_run = 0
        self.evictions = 0
        # host seconds: re-targeting batches, enqueueing substeps (launches or
        # graph replays, graph capture included), waiting for the device
        self.build_s = 0.0
        self.launch_s = 0.0
        self.wait_s = 0.0
        self.syncs = 0

    @property
    def step_s(self) -> float:
        """Host seconds spent stepping batches (enqueue + wait)."""
        return self.launch_s + self.wait_s

    def reset_counters(self):
        self.rebuilds = self.substeps_run = self.syncs = 0
        self.build_s = self.launch_s = self.wait_s = 0.0

    def close(self):
        """Frees every resident scene (parked states stay in host memory)."""
        for s in self.resident.values():
            s.close()
        self.resident.clear()
        self.bytes.clear()

    @property
    def used_bytes(self) -> int:
        return sum(self.bytes.values())

    def make_resident(self, k: int, busy=()) -> bool:
        """Makes scene k resident if it fits the budget once idle scenes are
        evicted (least recently used first); otherwise changes nothing and
        returns False.  A scene's device footprint is measured when it is
        first built and remembered, so a scene that has to wait is not built
        again at every check."""
        if k in self.resident:
            self.resident.move_to_end(k)
            return True
        s = None
        if k not in self.footprint:
            s = make_solver(self.specs[k], **self.solver_opts)
            self.footprint[k] = scene_bytes(s)
        need = self.footprint[k]
        idle = sum(b for v, b in self.bytes.items() if v not in busy)
        if self.used_bytes - idle + need > self.budget:
            if s is not None:
                s.close()
            return False
        while self.used_bytes + need > self.budget:
            victim = next(v for v in self.resident if v not in busy)
            self.parked[victim] = self.resident[victim].save_state()
            self.resident.pop(victim).close()
            del self.bytes[victim]
            self.evictions += 1
        if s is None:
            s = make_solver(self.specs[k], **self.solver_opts)
        if k in self.parked:
            s.load_state(self.parked.pop(k))
        self.resident[k] = s
        self.bytes[k] = need
        return True

    def warm(self, ks):
        """Makes the given scenes resident before the clock starts."""
        for k in ks:
            if not self.make_resident(k):
                break

    def run(self, requests: list, idle_sleep: float = 2e-4) -> list:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = {}  # scene -> its arrived requests, oldest first
        active = {}  # scene -> [request, substeps left, start]
        done = []
        batch = None
        key = None
        t0 = time.perf_counter()
        while pending or waiting or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                req = pending.popleft()
                waiting.setdefault(req.scene, collections.deque()).append(req)
            # admit the oldest waiting request of every free scene, oldest first
            for j in sorted((j for j in waiting if j not in active),
                            key=lambda j: (waiting[j][0].arrival, waiting[j][0].rid)):
                if not self.make_resident(j, busy=active):
                    continue
                req = waiting[j].popleft()
                if not waiting[j]:
                    del waiting[j]
                subs = req.frames * self.specs[j].params["substeps_num"]
                active[j] = [req, subs, now]
            if not active:
                if pending:
                    wait = pending[0].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, idle_sleep * 10))
                    continue
                raise RuntimeError("queued requests cannot be admitted within the HBM budget")
            k = tuple(sorted(active))
            if k != key:
                tb = time.perf_counter()
                members = [self.resident[j] for j in k]
                if batch is None:
                    batch = SceneBatch(members, **self.batch_opts)
                else:  # same stream and buffers, new member list
                    batch.set_members(members)
                key = k
                self.rebuilds += 1
                self.build_s += time.perf_counter() - tb
            # every active request runs at least `run` more substeps
            run = min(e[1] for e in active.values())
            run = min(run, self.max_run or min(self.specs[j].params["substeps_num"] for j in k))
            ts = time.perf_counter()
            batch.substep(run)
            tw = time.perf_counter()
            batch.synchronize()
            self.launch_s += tw - ts
            self.wait_s += time.perf_counter() - tw
            self.substeps_run += run
            self.syncs += 1
            now = time.perf_counter() - t0
            for j in k:
                entry = active[j]
                entry[1] -= run
                if entry[1] == 0:
                    req = entry[0]
                    s = self.resident[j]
                    tit = s.n_tets * s.constraints_iter * req.frames * s.substeps_num
                    done.append(Done(req.rid, j, req.frames, req.arrival, entry[2], now, tit))
                    del active[j]
        if batch is not None:
            batch.close()
        return done


def run_sequential(scenes: list, requests: list, solver_opts=None, solvers=None) -> list:
    """Serves the trace first-come first-served, one request at a time, each
    with plain per-scene solve() calls."""
    solvers = solvers if solvers is not None else {}
    done = []
    t0 = time.perf_counter()
    for req in sorted(requests, key=lambda r: r.arrival):
        now = time.perf_counter() - t0
        if now < req.arrival:
            time.sleep(req.arrival - now)
        if req.scene not in solvers:
            solvers[req.scene] = make_solver(scenes[req.scene], **(solver_opts or {}))
        s = solvers[req.scene]
        start = time.perf_counter() - t0
        for _ in range(req.frames):
            s.solve()
        finish = time.perf_counter() - t0
        done.append(Done(req.rid, req.scene, req.frames, req.arrival, start, finish,
                         s.n_tets * s.constraints_iter * req.frames * s.substeps_num))
    return done


def summarize(done: list, wall: float | None = None) -> dict:
    lat = np.array([d.latency for d in done]) if done else np.zeros(1)
    svc = np.array([d.service for d in done]) if done else np.zeros(1)
    work = sum(d.tet_iterations for d in done)
    span = wall if wall is not None else (max(d.finish for d in done) if done else 0.0)
    return {
        "requests": len(done),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "service_p50_ms": float(np.percentile(svc, 50) * 1e3),
        "service_p99_ms": float(np.percentile(svc, 99) * 1e3),
        "tet_iterations": int(work),
        "makespan_s": float(span),
        "tet_iterations_per_s": float(work / span) if span > 0 else 0.0,
    }


__all__ = ["Request", "Done", "Server", "served_scenes", "poisson_trace", "run_sequential",
           "summarize"]
