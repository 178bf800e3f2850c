"""The persistent substep kernel (exec_path="persistent": one launch per
substep for small scenes, every iteration inside it with scene-wide barriers)
against the phased kernels replayed from a graph: bit-identical positions,
velocities, multipliers and per-iteration logs in deterministic mode, over
contact, split mode with the convergence monitor, the iteration counts at the
edges of the sequence, and the reference's demo (cow) case; in atomic mode
against the oracle.  Its batched form (the members of a batch side by side in
one launch, each meeting at its own barriers) against the phased batch and
the scenes alone, and its path selection."""
import numpy as np
import pytest

import golden_cases
from scenes import batched_scenes, dropped_cube, make_oracle, make_solver, resting_cube
from test_gpu_parity import check_positions, check_residual_curve, residual_rows, run_pair
from xpbd_gpu.batch import SceneBatch
from xpbd_gpu.mesh import box_grid

pytestmark = pytest.mark.gpu


def _pair(scene, frames, setup=None, **opts):
    out = []
    for path in ("graph", "persistent"):
        s = make_solver(scene, log_residuals=True, exec_path=path, **opts)
        if setup is not None:
            setup(s)
        for _ in range(frames):
            s.solve()
        out.append(s)
    return out


def _same(a, b):
    assert np.array_equal(a.mesh.verts.x.to_numpy(), b.mesh.verts.x.to_numpy())
    assert np.array_equal(a.mesh.verts.v.to_numpy(), b.mesh.verts.v.to_numpy())
    assert np.array_equal(a.mesh.cells.lambda_D.to_numpy(), b.mesh.cells.lambda_D.to_numpy())
    assert np.array_equal(a.mesh.cells.lambda_H.to_numpy(), b.mesh.cells.lambda_H.to_numpy())
    assert np.array_equal(residual_rows(a), residual_rows(b))
    assert a.rho_hat == b.rho_hat


@pytest.mark.parametrize("name,scene", [
    ("dropped_cube16", dropped_cube()),
    ("resting_cube16", resting_cube()),
    ("resting_cube5", resting_cube(n=5, frames=4)),
])
def test_persistent_matches_graph_bitwise(name, scene):
    print(f"[case] persistent {name}", flush=True)
    a, b = _pair(scene, scene.frames)
    _same(a, b)


@pytest.mark.parametrize("iters,cheb", [(1, True), (2, True), (5, True), (6, True), (40, True),
                                        (1, False), (2, False)])
def test_persistent_iteration_counts(iters, cheb):
    print(f"[case] persistent iters {iters} cheb {cheb}", flush=True)
    scene = resting_cube(n=6, frames=2)
    scene.params["constraints_iter"] = iters
    scene.params["cheb_enable"] = cheb
    scene.params["monitor_convergence"] = not cheb
    a, b = _pair(scene, 2)
    _same(a, b)


def test_persistent_split_mode_and_monitor():
    print("[case] persistent split+monitor", flush=True)
    scene = resting_cube(n=8, frames=3)
    scene.params["block_neohookean_enable"] = False
    scene.params["monitor_convergence"] = True
    a, b = _pair(scene, 3)
    _same(a, b)
    ha = [(r["hydro_error"], r["dev_error"]) for r in a.convergence_history]
    hb = [(r["hydro_error"], r["dev_error"]) for r in b.convergence_history]
    assert ha == hb


@pytest.mark.parametrize("block,monitor", [(True, False), (True, True), (False, True),
                                           (False, False)])
def test_persistent_without_chebyshev(block, monitor):
    """Chebyshev off (the reference's NH and BNH methods): the persistent
    substep takes x_J as the next iterate and runs the residual pass only for
    the convergence monitor; the same bits as the graph.  A wait forced to
    give up is reported, so the persistent path is the one that ran."""
    from xpbd_gpu._abi import XpbdError
    print(f"[case] persistent no-cheb block={block} monitor={monitor}", flush=True)
    scene = resting_cube(frames=2)  # 48 patches: 24 workgroups
    scene.params.update(cheb_enable=False, block_neohookean_enable=block,
                        monitor_convergence=monitor)
    a, b = _pair(scene, 2)
    _same(a, b)
    if monitor:
        ha = [(r["hydro_error"], r["dev_error"]) for r in a.convergence_history]
        hb = [(r["hydro_error"], r["dev_error"]) for r in b.convergence_history]
        assert len(ha) == 2 * 5 * 30 and ha == hb
    b.set_barrier_timeout(0)
    with pytest.raises(XpbdError, match="timed out"):
        b.solve()


def test_batch_persistent_without_chebyshev():
    print("[case] batch persistent no-cheb", flush=True)
    scenes = _mixed_batch_scenes()
    for sc in scenes:
        sc.params.update(cheb_enable=False, monitor_convergence=True)
    alone = [make_solver(s, exec_path="graph") for s in scenes]
    per = [make_solver(s) for s in scenes]
    for s in alone:
        s.solve()
    with SceneBatch(per) as b:
        assert b.path == "persistent"
        b.frame(1)
        b.synchronize()
    for a, p in zip(alone, per):
        assert np.array_equal(a.mesh.verts.x.to_numpy(), p.mesh.verts.x.to_numpy())
        assert np.array_equal(a.mesh.verts.v.to_numpy(), p.mesh.verts.v.to_numpy())
        assert np.array_equal(a.mesh.cells.lambda_D.to_numpy(), p.mesh.cells.lambda_D.to_numpy())


def test_persistent_reference_demo():
    print("[case] persistent cow demo", flush=True)
    case = golden_cases.load("cow_demo_contact")
    sc = golden_cases.scene(case)

    def setup(s):
        s.mesh.verts.v.from_numpy(case["v0"])
        s.rho_hat = case["rho0"]
    a, b = _pair(sc, case["frames"], setup=setup)
    _same(a, b)


def test_persistent_hang_and_cube_colliders():
    from xpbd_gpu.scenes import Scene, base_params, cube, hang, plane
    print("[case] persistent colliders", flush=True)
    m = box_grid(4, origin=(0.0, 0.3, 0.0))
    sdfs = [cube((0.5, 0.1, 0.5), (2.0, 0.4, 2.0), fixed=False),
            hang([[0.0, 1.3, 0.0], [1.0, 1.3, 1.0]]), plane(-1.0)]
    sc = Scene("cube_hang", m, sdfs, base_params(young_modulus=2e5, poisson_ratio=0.4,
                                                 substeps_num=4, constraints_iter=20),
               total_mass=1000.0, frames=5)
    a, b = _pair(sc, 5)
    _same(a, b)


@pytest.mark.parametrize("cheb", [True, False])
def test_persistent_atomic_mode_matches_oracle(cheb):
    print(f"[case] persistent atomic cheb {cheb}", flush=True)
    scene = resting_cube(frames=5)
    scene.params["cheb_enable"] = cheb
    solver, orc = run_pair(scene, accumulate_mode="atomic", exec_path="persistent")
    check_positions(scene, solver, orc)
    if cheb:
        check_residual_curve(solver, orc)


def test_large_scene_falls_back_to_graph():
    """Above the persistent kernel's size limit the solver takes the graph path
    (same bits)."""
    from scenes import jittered_box
    print("[case] persistent fallback", flush=True)
    scene = jittered_box(n=38, frames=1, iters=10)  # 643 patches
    a, b = _pair(scene, 1)
    _same(a, b)


def _mixed_batch_scenes():
    """Four batched-configuration boxes of 3-12 cells per axis and the 16^3
    cube resting on the plane (48 patches), 5 substeps x 20 iterations."""
    scenes = batched_scenes(n_scenes=4, seed=13, n_range=(3, 12), iters=20)
    cube = resting_cube(frames=2)
    cube.params.update(substeps_num=5, constraints_iter=20)
    return scenes + [cube]


@pytest.mark.parametrize("persistent", ["auto", "always"])
def test_batch_persistent_matches_phased_and_alone(persistent):
    print(f"[case] batch persistent {persistent}", flush=True)
    scenes = _mixed_batch_scenes()
    alone = [make_solver(s, exec_path="graph") for s in scenes]
    per = [make_solver(s) for s in scenes]
    phased = [make_solver(s) for s in scenes]
    for s in alone:
        s.solve()
        s.solve()
    with SceneBatch(per, persistent=persistent) as b:
        assert b.path == "persistent"
        b.frame(1)
        b.substep(3)
        b.synchronize()
        b.substep(2)
        b.synchronize()
    with SceneBatch(phased, persistent="never") as b:
        assert b.path == "graph"
        b.frame(2)
        b.synchronize()
    for a, p, q in zip(alone, per, phased):
        for s in (p, q):
            assert np.array_equal(a.mesh.verts.x.to_numpy(), s.mesh.verts.x.to_numpy())
            assert np.array_equal(a.mesh.verts.v.to_numpy(), s.mesh.verts.v.to_numpy())
            assert np.array_equal(a.mesh.cells.lambda_D.to_numpy(),
                                  s.mesh.cells.lambda_D.to_numpy())
            assert a.rho_hat == s.rho_hat
            assert s.current_frame == 2


def test_persistent_placement_keeps_the_bits(monkeypatch):
    """The single-scene persistent substep places its workgroups on one XCD
    when they fit there (24 for the cube); spread over all XCDs (one XCD
    declared) the results are the same bits."""
    print("[case] persistent placement", flush=True)
    scene = resting_cube(frames=2)
    one = make_solver(scene, log_residuals=True, exec_path="persistent")
    monkeypatch.setenv("XPBD_XCDS", "1")
    spread = make_solver(scene, log_residuals=True, exec_path="persistent")
    for s in (one, spread):
        for _ in range(scene.frames):
            s.solve()
    _same(one, spread)


@pytest.mark.parametrize("n_members,xcd_split", [(4, True), (4, False), (10, True)])
def test_batch_persistent_layouts_keep_the_bits(n_members, xcd_split):
    """Members of 17 groups each: four fit one XCD apiece (per-XCD layout with
    holes), and without the XCD split they run contiguously; ten do not fit
    (two on one XCD would need 34 of its 32 CUs), so the layout stays
    contiguous.  Every member matches the scene run alone, bit for bit."""
    print(f"[case] batch persistent layout {n_members} {xcd_split}", flush=True)
    scene = resting_cube(n=14, frames=2)
    scene.params.update(substeps_num=2, constraints_iter=12)
    alone = make_solver(scene, exec_path="graph")
    for _ in range(2):
        alone.solve()
    members = [make_solver(scene) for _ in range(n_members)]
    with SceneBatch(members, persistent="always", xcd_split=xcd_split) as b:
        assert b.path == "persistent"
        b.frame(2)
        b.synchronize()
    for s in members:
        assert np.array_equal(alone.mesh.verts.x.to_numpy(), s.mesh.verts.x.to_numpy())
        assert np.array_equal(alone.mesh.verts.v.to_numpy(), s.mesh.verts.v.to_numpy())
        assert np.array_equal(alone.mesh.cells.lambda_H.to_numpy(),
                              s.mesh.cells.lambda_H.to_numpy())
        assert alone.rho_hat == s.rho_hat


def test_batch_persistent_path_selection():
    print("[case] batch persistent path selection", flush=True)
    small = [make_solver(s) for s in batched_scenes(n_scenes=3, seed=5, n_range=(3, 6), iters=5)]
    with SceneBatch(small) as b:
        assert b.path == "persistent"
    with SceneBatch(small, persistent="never") as b:
        assert b.path == "graph"
    with SceneBatch(small, persistent="never", graph=False) as b:
        assert b.path == "kernels"
    # a member of more than 512 patches: the persistent substep does not apply
    # (automatic selection keeps the phased kernels, "always" refuses)
    huge = make_solver(batched_scenes(n_scenes=1, seed=5, n_range=(38, 38), iters=5)[0])
    with SceneBatch([huge], persistent="never") as b:
        assert b.n_patches > 512, b.n_patches
    with SceneBatch([huge] + small) as b:
        assert b.path == "graph"
    with SceneBatch([huge] + small, persistent="always") as b:
        with pytest.raises(ValueError):
            b.substep(1)
    # a member of 289-512 patches: automatic selection keeps the phased
    # kernels, "always" takes the persistent substep
    mid = make_solver(batched_scenes(n_scenes=1, seed=5, n_range=(30, 30), iters=5)[0])
    with SceneBatch([mid], persistent="never") as b:
        assert 288 < b.n_patches <= 512, b.n_patches
    with SceneBatch([mid] + small) as b:
        assert b.path == "graph"
    with SceneBatch([mid] + small, persistent="always") as b:
        assert b.path == "persistent"
    # a member of 97-288 patches takes it automatically, with the same bits
    big_scene = batched_scenes(n_scenes=1, seed=5, n_range=(22, 22), iters=5)[0]
    big = make_solver(big_scene)
    with SceneBatch([big], persistent="never") as b:
        groups = (b.n_patches + 1) // 2
    assert 48 < groups <= 128, groups
    ref = make_solver(big_scene, exec_path="graph")
    with SceneBatch([big] + small) as b:
        assert b.path == "persistent"
        b.substep(2)
        b.synchronize()
    for _ in range(2):
        sub_dt = ref.dt / ref.substeps_num
        ref.apply_external_forces(sub_dt)
        ref.solve_step(sub_dt)
        ref.advance(sub_dt, ref.damping)
    assert np.array_equal(ref.mesh.verts.x.to_numpy(), big.mesh.verts.x.to_numpy())
    # more workgroups than the device holds at once: "always" refuses before
    # launching anything, the automatic choice falls back
    bigs = [make_solver(big_scene) for _ in range(6)]
    with SceneBatch(bigs) as b:
        assert b.path == "graph"
    with SceneBatch(bigs, persistent="always") as b:
        with pytest.raises(ValueError):
            b.substep(1)


def test_barrier_timeout_is_reported_and_the_solver_recovers():
    """A scene barrier that gives up (forced here by a zero wait, so every
    workgroup still leaves its loops) is reported by the next synchronizing
    call; the error word and the barrier counters are reset, so the solver
    continues normally from a saved state."""
    from xpbd_gpu._abi import XpbdError
    print("[case] persistent barrier timeout", flush=True)
    scene = resting_cube(frames=2)  # 48 patches: 24 workgroups
    s = make_solver(scene, exec_path="persistent")
    ref = make_solver(scene, exec_path="persistent")
    st = s.save_state()
    s.set_barrier_timeout(0)
    with pytest.raises(XpbdError, match="timed out"):
        s.solve()
    s.set_barrier_timeout(2_000_000)
    s.load_state(st)
    for _ in range(2):
        s.solve()
        ref.solve()
    assert np.array_equal(s.mesh.verts.x.to_numpy(), ref.mesh.verts.x.to_numpy())


def test_batch_barrier_timeout_is_reported_and_the_batch_recovers():
    from xpbd_gpu._abi import XpbdError
    print("[case] batch persistent barrier timeout", flush=True)
    scenes = _mixed_batch_scenes()
    per = [make_solver(s) for s in scenes]
    ref = [make_solver(s) for s in scenes]
    states = [s.save_state() for s in per]
    with SceneBatch(per) as b:
        assert b.path == "persistent"
        b.set_barrier_timeout(0)
        b.substep(1)
        with pytest.raises(XpbdError, match="timed out"):
            b.synchronize()
        b.set_barrier_timeout(2_000_000)
        for s, st in zip(per, states):
            s.load_state(st)
        b.frame(1)
        b.synchronize()
    with SceneBatch(ref, persistent="never") as b:
        b.frame(1)
        b.synchronize()
    for a, r in zip(per, ref):
        assert np.array_equal(a.mesh.verts.x.to_numpy(), r.mesh.verts.x.to_numpy())
