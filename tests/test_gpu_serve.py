"""Served form: continuous batching over resident scenes; every result is the
same (bit-exact, deterministic mode) alone and in the mix."""
import numpy as np
import pytest

from xpbd_gpu.scenes import make_solver
from xpbd_gpu.serve import Request, Server, poisson_trace, run_sequential, served_scenes, summarize

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _alone_after(scene, frames):
    s = make_solver(scene)
    for _ in range(frames):
        s.solve()
    return s.mesh.verts.x.to_numpy()


@pytest.mark.parametrize("budget_scenes,graph,max_run", [(None, True, None), (2, True, None),
                                                         (None, False, None), (2, True, 1)])
def test_served_results_match_alone(budget_scenes, graph, max_run):
    _print_case("served_alone_vs_mix", budget_scenes=budget_scenes, graph=graph, max_run=max_run)
    scenes = served_scenes(n_scenes=5, seed=1, median_tets=2e3, tet_range=(3e2, 6e3), iters=10)
    reqs = poisson_trace(len(scenes), rate=60.0, duration=0.5, seed=2)
    assert len(reqs) >= 10
    budget = 200e9
    if budget_scenes is not None:
        srv0 = Server(scenes)
        for k in range(len(scenes)):
            srv0.make_resident(k)
        sizes = sorted(srv0.bytes.values())
        budget = sizes[-1] + sizes[-2] + 1  # any two scenes fit, never all five
        del srv0
    srv = Server(scenes, hbm_budget=budget, batch_opts=dict(graph=graph), max_run=max_run)
    done = srv.run(reqs)
    assert sorted(d.rid for d in done) == sorted(r.rid for r in reqs)
    assert srv.used_bytes <= budget
    # one wait for the device per run of substeps, never more than one per substep
    assert srv.syncs <= srv.substeps_run
    if max_run == 1:
        assert srv.syncs == srv.substeps_run
    if budget_scenes is not None:
        assert srv.evictions > 0
    frames = {}
    for r in reqs:
        frames[r.scene] = frames.get(r.scene, 0) + r.frames
    for k, f in frames.items():
        s = srv.resident.get(k)
        x = s.mesh.verts.x.to_numpy() if s is not None else srv.parked[k]["x"]
        assert np.array_equal(x, _alone_after(scenes[k], f)), k
    st = summarize(done)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_requests_overlap_and_latency_accounting():
    _print_case("served_overlap")
    scenes = served_scenes(n_scenes=3, seed=4, median_tets=1e3, tet_range=(3e2, 3e3), iters=5)
    reqs = [Request(0, 0, 2, 0.0), Request(1, 1, 1, 0.0), Request(2, 0, 1, 0.0)]
    srv = Server(scenes)
    done = {d.rid: d for d in srv.run(reqs)}
    # requests 0 and 1 (different scenes) ran in the same batch; 2 waited for 0
    assert done[1].start == done[0].start
    assert done[2].start >= done[0].finish
    assert all(d.finish >= d.start >= d.arrival for d in done.values())
    assert all(d.latency >= d.service > 0 for d in done.values())
    # request 1 (one frame) ends the first run: 5 substeps, then 0 runs 5 more
    # alone, then 2; three waits for the device in all
    assert srv.substeps_run == 15 and srv.syncs == 3
    srv.close()


def test_sequential_baseline_serves_trace():
    scenes = served_scenes(n_scenes=3, seed=6, median_tets=1e3, tet_range=(3e2, 3e3), iters=5)
    reqs = poisson_trace(3, rate=40.0, duration=0.2, seed=3)
    done = run_sequential(scenes, reqs)
    assert len(done) == len(reqs)
    assert summarize(done)["tet_iterations"] > 0
