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
