"""Batched form: one set of launches for many scenes, checked against the same
scenes alone (bit-exact, deterministic mode) and against the oracle.  The
launch knobs (graphs, grid size, XCD split, vertex workgroup) belong to the
phased kernels, so those cases turn the persistent substep off; the default
options take it for these small members."""
import numpy as np
import pytest

from scenes import batched_scenes, resting_cube, make_oracle, make_solver
from xpbd_gpu.batch import SceneBatch, autotune, autotune_ordering

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


NEVER = dict(persistent="never")


@pytest.mark.parametrize("opts", [dict(), dict(persistent="always"), NEVER,
                                  dict(graph=False, **NEVER), dict(graph_substeps=5, **NEVER),
                                  dict(xcd_split=False, vertex_block=128, grid_pct=50, **NEVER),
                                  dict(grid_pct=1, **NEVER), dict(grid_pct=1, graph=False, **NEVER)])
def test_batch_members_match_alone_bitwise(opts):
    _print_case("batch_vs_alone", **opts)
    scenes = batched_scenes(n_scenes=6, seed=3, n_range=(3, 9), iters=20)
    alone = [make_solver(s) for s in scenes]
    mixed = [make_solver(s) for s in scenes]
    for s in alone:
        s.solve()
        s.solve()
    with SceneBatch(mixed, **opts) as b:
        assert b.path == ("graph" if opts.get("graph", True) else "kernels") \
            if opts.get("persistent") == "never" else "persistent"
        if opts.get("grid_pct") == 1:  # workgroups walk patches of several scenes
            assert b.grid()["max_scenes_per_range"] >= 2, b.grid()
        b.frame(2)
        b.synchronize()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.mesh.verts.x.to_numpy(), m.mesh.verts.x.to_numpy())
        assert np.array_equal(a.mesh.verts.v.to_numpy(), m.mesh.verts.v.to_numpy())
        assert a.rho_hat == m.rho_hat
        assert m.current_frame == 2


@pytest.mark.parametrize("persistent", ["auto", "never"])
def test_batch_member_matches_oracle(persistent):
    _print_case("batch_vs_oracle", persistent=persistent)
    scenes = batched_scenes(n_scenes=4, seed=5, n_range=(3, 8), iters=30)
    solvers = [make_solver(s) for s in scenes]
    with SceneBatch(solvers, persistent=persistent) as b:
        b.frame(3)
        b.synchronize()
    for sc, s in zip(scenes, solvers):
        o = make_oracle(sc)
        o.frame(3)
        err = np.abs(s.mesh.verts.x.to_numpy() - o.x).max()
        assert err <= 1e-4 * sc.diag, (sc.name, err)


@pytest.mark.parametrize("persistent", ["auto", "never"])
def test_batch_substeps_and_mixed_sizes(persistent):
    _print_case("batch_substeps", persistent=persistent)
    scenes = batched_scenes(n_scenes=3, seed=7, n_range=(2, 12), iters=10) + [resting_cube(n=4)]
    scenes[-1].params.update(substeps_num=5, constraints_iter=10)
    solvers = [make_solver(s) for s in scenes]
    ref = [make_solver(s) for s in scenes]
    with SceneBatch(solvers, persistent=persistent) as b:
        b.substep(3)
        b.synchronize()
        assert all(s.current_substep == 2 for s in solvers)
        b.substep(2)
        b.frame(1)
        b.synchronize()
    for s in ref:
        s.solve()
        s.solve()
    for a, m in zip(ref, solvers):
        assert np.array_equal(a.mesh.verts.x.to_numpy(), m.mesh.verts.x.to_numpy())


@pytest.mark.parametrize("graph,grid_pct,persistent", [(False, 100, "never"), (True, 100, "never"),
                                                       (True, 1, "never"), (True, 100, "auto")])
def test_set_members_matches_fresh_batches(graph, grid_pct, persistent):
    """A batch re-targeted at other members (growing, shrinking and reordering
    its list) steps each scene exactly as plain solves do."""
    _print_case("set_members", graph=graph, grid_pct=grid_pct, persistent=persistent)
    scenes = batched_scenes(n_scenes=5, seed=11, n_range=(3, 10), iters=15)
    alone = [make_solver(s) for s in scenes]
    mixed = [make_solver(s) for s in scenes]
    plan = [[0, 1], [1, 2, 3, 4], [4], [3, 0, 2]]  # members per substep round
    done = [0] * len(scenes)
    with SceneBatch([mixed[0]], graph=graph, grid_pct=grid_pct, persistent=persistent) as b:
        for members in plan:
            b.set_members([mixed[k] for k in members])
            assert b.n_scenes == len(members)
            b.substep(2)
            b.synchronize()
            for k in members:
                done[k] += 2
    for k, n in enumerate(done):
        sub_dt = alone[k].dt / alone[k].substeps_num
        for _ in range(n):
            alone[k].apply_external_forces(sub_dt)
            alone[k].solve_step(sub_dt)
            alone[k].advance(sub_dt, alone[k].damping)
        assert np.array_equal(alone[k].mesh.verts.x.to_numpy(), mixed[k].mesh.verts.x.to_numpy()), k
    with pytest.raises(ValueError):
        b2 = SceneBatch([mixed[0]])
        b2.set_members([mixed[1], mixed[1]])


def test_batch_rejects_mismatched_members():
    a = make_solver(resting_cube(n=2))
    sc = resting_cube(n=2)
    sc.params["constraints_iter"] = 7
    b = make_solver(sc)
    with pytest.raises(ValueError):
        SceneBatch([a, b])
    with pytest.raises(ValueError):
        SceneBatch([a, a])


def test_autotune_ordering_frees_and_picks():
    _print_case("autotune_ordering")
    scenes = batched_scenes(n_scenes=4, seed=5, n_range=(3, 8), iters=10)
    built = []

    def build(order):
        solvers = [make_solver(s, reorder=order) for s in scenes]
        built.append(solvers)
        return solvers

    best, res = autotune_ordering(build, substeps=1)
    assert best in ("bisect", "morton", "none") and [o for o, _ in res] == ["bisect", "morton",
                                                                           "none"]
    assert all(t > 0 for _, t in res)
    assert all(s._h is None for solvers in built for s in solvers)  # freed after timing


def test_autotune_restores_state():
    _print_case("autotune")
    scenes = batched_scenes(n_scenes=4, seed=9, n_range=(3, 8), iters=10)
    solvers = [make_solver(s) for s in scenes]
    x0 = [s.mesh.verts.x.to_numpy() for s in solvers]
    best, res = autotune(solvers, candidates=[dict(vertex_block=128), dict(vertex_block=256),
                                              dict(graph_substeps=5)])
    assert best in [r[0] for r in res] and len(res) == 3
    for s, x in zip(solvers, x0):
        assert np.array_equal(s.mesh.verts.x.to_numpy(), x)
        assert s.current_frame == 0


def test_autotune_skips_persistent_where_it_does_not_apply():
    """The persistent candidate is timed on small members and listed with an
    infinite time where the members cannot all be resident at once."""
    _print_case("autotune_persistent")
    small = [make_solver(s) for s in batched_scenes(n_scenes=3, seed=9, n_range=(3, 8), iters=5)]
    best, res = autotune(small, candidates=[dict(persistent="never"), dict(persistent="always")])
    assert all(np.isfinite(t) for _, t in res)
    big = [make_solver(s) for s in batched_scenes(n_scenes=6, seed=5, n_range=(22, 22), iters=5)]
    best, res = autotune(big, candidates=[dict(persistent="never"), dict(persistent="always")])
    assert best == dict(persistent="never")
    assert np.isfinite(res[0][1]) and res[1][1] == float("inf")


def test_autotune_exec_path_restores_state_and_keeps_the_bits():
    _print_case("autotune_exec_path")
    from xpbd_gpu.batch import autotune_exec_path
    scene = resting_cube(n=6, frames=2)
    s = make_solver(scene, monitor_convergence=True)
    ref = make_solver(scene, monitor_convergence=True, exec_path="graph")
    x0 = s.mesh.verts.x.to_numpy()
    best, res = autotune_exec_path(s, frames=1)
    assert best in ("persistent", "graph", "kernels") and s.exec_path == best
    assert [p for p, _ in res] == ["persistent", "graph", "kernels"]
    assert np.array_equal(s.mesh.verts.x.to_numpy(), x0)
    assert s.current_frame == 0 and s.convergence_history == [] and s.solve_compute_calls == 0
    for _ in range(2):
        s.solve()
        ref.solve()
    assert np.array_equal(s.mesh.verts.x.to_numpy(), ref.mesh.verts.x.to_numpy())
