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
