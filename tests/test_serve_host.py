"""Host-side pieces of the served form (no GPU)."""
import numpy as np

from xpbd_gpu.serve import Done, poisson_trace, served_scenes, summarize


def test_poisson_trace_rate_and_fields():
    reqs = poisson_trace(10, rate=50.0, duration=20.0, seed=1)
    assert abs(len(reqs) / 20.0 - 50.0) < 5.0
    assert all(0 <= r.scene < 10 and 1 <= r.frames <= 4 for r in reqs)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0) and t[-1] < 20.0
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 50.0) < 0.003


def test_served_scene_sizes_follow_lognormal():
    sc = served_scenes(n_scenes=200, seed=0)
    t = np.array([s.mesh.n_tets for s in sc])
    assert 1e3 * 0.5 <= t.min() and t.max() <= 5e5 * 1.2
    assert 1.0e4 < np.median(t) < 4.0e4


def test_summary_percentiles():
    done = [Done(i, 0, 1, 0.0, 0.0, 0.001 * (i + 1), 10) for i in range(100)]
    st = summarize(done)
    assert abs(st["p50_ms"] - 50.5) < 1.0 and abs(st["p99_ms"] - 99.0) < 1.1
    assert st["tet_iterations"] == 1000


class _FakeSolver:
    """Stands in for a device solver in the admission logic."""
    built = 0

    def __init__(self, spec):
        _FakeSolver.built += 1
        self.spec = spec
        self.closed = False
        self.state = None

    def save_state(self):
        return {"of": self.spec}

    def load_state(self, st):
        self.state = st

    def close(self):
        self.closed = True


def _server(monkeypatch, sizes, budget):
    import xpbd_gpu.serve as serve
    monkeypatch.setattr(serve, "make_solver", lambda spec, **kw: _FakeSolver(spec))
    monkeypatch.setattr(serve, "scene_bytes", lambda s: sizes[s.spec])
    _FakeSolver.built = 0
    return serve.Server(list(range(len(sizes))), hbm_budget=budget)


def test_admission_evicts_least_recently_used_idle_scenes(monkeypatch):
    srv = _server(monkeypatch, [40, 40, 40, 50], budget=100)
    assert srv.make_resident(0) and srv.make_resident(1)
    assert srv.make_resident(0)  # touch: 1 is now the least recently used
    first = srv.resident[1]
    assert srv.make_resident(2)  # 120 > 100: evicts scene 1
    assert list(srv.resident) == [0, 2] and srv.evictions == 1
    assert first.closed and srv.parked[1] == {"of": 1}
    assert srv.used_bytes == 80
    assert srv.make_resident(1, busy={2})  # evicts 0 (idle), not 2 (busy)
    assert list(srv.resident) == [2, 1] and srv.resident[1].state == {"of": 1}
    assert 1 not in srv.parked


def test_admission_that_cannot_fit_changes_nothing(monkeypatch):
    srv = _server(monkeypatch, [60, 30, 50], budget=100)
    assert srv.make_resident(0) and srv.make_resident(1)
    built = _FakeSolver.built
    # scene 2 needs 50: only scene 1 (30) is idle, so even evicting it leaves 110
    assert not srv.make_resident(2, busy={0})
    assert list(srv.resident) == [0, 1] and srv.evictions == 0 and not srv.parked
    assert _FakeSolver.built == built + 1  # built once to measure it
    # while it waits it is not built again
    assert not srv.make_resident(2, busy={0})
    assert _FakeSolver.built == built + 1
    # once scene 0 is idle, admission evicts 0 (least recently used) and builds it
    assert srv.make_resident(2)
    assert list(srv.resident) == [1, 2] and srv.evictions == 1


class _Spec:
    def __init__(self, substeps):
        self.params = {"substeps_num": substeps}


class _StepSolver(_FakeSolver):
    n_tets, constraints_iter = 100, 2

    def __init__(self, spec):
        super().__init__(spec)
        self.substeps_num = spec.params["substeps_num"]


class _FakeBatch:
    """Records what the serving loop enqueues and when it waits."""
    log = []

    def __init__(self, members, **opts):
        _FakeBatch.log.append(("create", len(members)))

    def set_members(self, members):
        _FakeBatch.log.append(("members", len(members)))

    def substep(self, n):
        _FakeBatch.log.append(("substep", n))

    def synchronize(self):
        _FakeBatch.log.append(("wait",))

    def close(self):
        pass


def _stepping_server(monkeypatch, max_run=None):
    import xpbd_gpu.serve as serve
    monkeypatch.setattr(serve, "make_solver", lambda spec, **kw: _StepSolver(spec))
    monkeypatch.setattr(serve, "scene_bytes", lambda s: 1)
    monkeypatch.setattr(serve, "SceneBatch", _FakeBatch)
    _FakeBatch.log = []
    return serve.Server([_Spec(3), _Spec(3)], max_run=max_run)


def test_loop_enqueues_up_to_the_next_completion_then_waits_once(monkeypatch):
    from xpbd_gpu.serve import Request
    srv = _stepping_server(monkeypatch)
    reqs = [Request(0, 0, 2, 0.0), Request(1, 1, 1, 0.0), Request(2, 0, 1, 0.0)]
    done = {d.rid: d for d in srv.run(reqs)}
    # request 1 (one frame of 3 substeps) ends the first run; request 0 runs its
    # second frame alone; request 2 follows on the same scene (same membership)
    assert _FakeBatch.log == [("create", 2), ("substep", 3), ("wait",), ("members", 1),
                              ("substep", 3), ("wait",), ("substep", 3), ("wait",)]
    assert srv.substeps_run == 9 and srv.syncs == 3 and srv.rebuilds == 2
    assert done[1].finish <= done[0].finish <= done[2].start <= done[2].finish
    assert all(d.latency >= d.service >= 0.0 for d in done.values())
    assert done[0].tet_iterations == 100 * 2 * 2 * 3
    assert srv.step_s == srv.launch_s + srv.wait_s


def test_loop_run_length_is_capped(monkeypatch):
    from xpbd_gpu.serve import Request
    srv = _stepping_server(monkeypatch, max_run=1)
    srv.run([Request(0, 0, 2, 0.0)])
    assert [e for e in _FakeBatch.log if e[0] == "substep"] == [("substep", 1)] * 6
    assert srv.syncs == 6
