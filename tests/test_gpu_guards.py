"""GPU cases for the reference's guard paths and for state that crosses scene
boundaries inside one launch: the 2x2 solve's diagonal fallback
(neohookean_solver.py:425-432), tets of non-positive volume
(xpbd_base.py:104-106), the obstacle velocity of a collider (sdf.py:13-27),
atomic-mode batches whose workgroups walk patches of several scenes, and a
batch whose member changes its parameters.  All through libxpbd_hip.so."""
import numpy as np
import pytest

import oracle
from scenes import batched_scenes, make_oracle, make_solver, resting_cube, Scene, base_params
from test_gpu_parity import check_residual_curve, run_pair
from xpbd_gpu.batch import SceneBatch
from xpbd_gpu.mesh import TetMesh, box_grid

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


# ---- the 2x2 solve's diagonal fallback ---------------------------------------------

GIANT_EDGE = 1.34e4  # V0 = edge^3 / 6 = 4.0e11


def fallback_scene():
    """A 3^3 box (E 1e6, nu 0.1) resting on the plane y = 0, and a separate
    giant tet pinned inside a fixed sphere.  With nu = 0.1, lambda = mu / 4, so
    for the giant tet alpha~_D = 1/(mu V0 dt^2) ~ 0.5e-12 and alpha~_H = 4
    alpha~_D ~ 2e-12.  Its vertices are pinned (w = 0), so A = diag(alpha~_D,
    alpha~_H): the first pivot is below EPSILON, the LU reports failure and the
    fallback gives dlambda_D = 0 (|A11| <= EPSILON) and dlambda_H = b2 / A22."""
    box = box_grid(3, origin=(0.0, 0.0, 0.0))
    o = np.array([-3.0e4, 0.0, 0.0])
    giant = o + GIANT_EDGE * np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    verts = np.vstack([box.verts, giant.astype(np.float32)])
    nb = box.n_verts
    tets = np.vstack([box.tets, np.array([[nb, nb + 1, nb + 2, nb + 3]], np.int32)])
    mesh = TetMesh(verts.astype(np.float32), tets.astype(np.int32))
    cen = giant.mean(0)
    sdfs = [dict(kind="plane", point=(0.0, 0.0, 0.0), dir=(0.0, 1.0, 0.0), fixed=False),
            dict(kind="sphere", center=tuple(cen), radius=1.5e4, fixed=True)]
    vol = 1.0 + GIANT_EDGE ** 3 / 6.0
    sc = Scene("fallback", mesh, sdfs,
               base_params(young_modulus=1e6, poisson_ratio=0.1, substeps_num=5,
                           constraints_iter=20),
               total_mass=1000.0 * vol, frames=2)
    return sc, nb, mesh.n_tets - 1


@pytest.mark.parametrize("layout", ["patch", "flat"])
def test_diagonal_fallback_of_the_2x2_solve(layout):
    _print_case("diagonal_fallback", layout=layout)
    sc, nb, gt = fallback_scene()
    solver, orc = run_pair(sc, layout=layout)
    # the oracle's own per-tet function on the giant tet at rest confirms the branch
    X = sc.mesh.verts[sc.mesh.tets[gt]].astype(np.float64)
    Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], 1)
    lame1, lame2 = oracle.lame(1e6, 0.1)
    dt = sc.params["dt"] / sc.params["substeps_num"]
    r = oracle.block_tet(X, np.zeros(4), np.linalg.inv(Dm), np.linalg.det(Dm) / 6, lame1, lame2,
                         dt, precision=32)
    assert r["ok"] == 0 and abs(r["A"][0, 0]) < 1e-12 < abs(r["A"][1, 1])
    assert r["dlam"][0] == 0.0
    # device and oracle both took it: lambda_D stays exactly 0, lambda_H moved
    ld, lh = solver.mesh.cells.lambda_D.to_numpy(), solver.mesh.cells.lambda_H.to_numpy()
    _, _, old, olh = orc.tet_state()
    assert ld[gt] == 0.0 and old[gt] == 0.0
    assert lh[gt] != 0.0 and olh[gt] != 0.0
    # the giant tet stays pinned; the box matches the oracle
    x = solver.mesh.verts.x.to_numpy().astype(np.float64)
    np.testing.assert_array_equal(x[nb:], sc.mesh.verts[nb:].astype(np.float64))
    diag = float(np.linalg.norm(x[:nb].max(0) - x[:nb].min(0)))
    err = np.abs(x[:nb] - orc.x[:nb]).max()
    print(f"  box max|dx| = {err / diag:.3e} x diag", flush=True)
    assert err <= 1e-4 * diag
    check_residual_curve(solver, orc)


# ---- tets of non-positive volume ----------------------------------------------------

def test_degenerate_tets_are_reported():
    _print_case("degenerate_tets")
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0]], dtype=np.float32)
    t = np.array([[0, 1, 2, 3],      # positive
                  [0, 2, 1, 3],      # inverted
                  [0, 1, 4, 2]],     # flat (z = 0)
                 dtype=np.int32)
    from xpbd_gpu.neohookean_solver import XPBDNeoHookeanSolver
    from xpbd_gpu.sdf import PlaneSdf
    with pytest.warns(RuntimeWarning, match="Cell 1 has negative or zero volume"):
        s = XPBDNeoHookeanSolver((v, t), [PlaneSdf([0.0, -1.0, 0.0], [0.0, 1.0, 0.0])])
    assert s.degenerate_tets() == (2, 1)
    good = make_solver(resting_cube(n=2))
    assert good.degenerate_tets() == (0, -1)


# ---- the obstacle's velocity ----------------------------------------------------------

def _rising_cube(vel):
    """A 4^3 cube starting 0.05 below a non-fixed plane y = 0 and rising at 0.2:
    at rest, the plane lets its submerged vertices through (n . (0 - v) < 0);
    moving up at 0.4 it catches them (n . (vel - v) > 0); moving down it
    releases even those that fall back."""
    sc = resting_cube(n=4, frames=3)
    sc.mesh = box_grid(4, origin=(0.0, -0.05, 0.0))
    sc.sdfs = [dict(sc.sdfs[0], fixed=False, vel=vel)]
    return sc


@pytest.mark.parametrize("vel", [(0.0, 0.4, 0.0), (0.0, -0.4, 0.0)])
def test_moving_obstacle_matches_oracle(vel):
    _print_case("moving_obstacle", vel=vel)
    sc = _rising_cube(vel)
    v0 = np.tile([0.0, 0.2, 0.0], (sc.mesh.n_verts, 1))
    solver = make_solver(sc)
    solver.mesh.verts.v.from_numpy(v0.astype(np.float32))
    orc = make_oracle(sc)
    orc.v = v0
    for _ in range(sc.frames):
        solver.solve()
        orc.frame(1)
    x = solver.mesh.verts.x.to_numpy().astype(np.float64)
    err = np.abs(x - orc.x).max()
    print(f"  max|dx| = {err / sc.diag:.3e} x diag", flush=True)
    assert err <= 1e-5 * sc.diag, err
    still = make_oracle(_rising_cube((0.0, 0.0, 0.0)))
    still.v = v0
    still.frame(sc.frames)
    assert np.abs(still.x - orc.x).max() > 1e-3  # the obstacle's velocity changed the result


# ---- batches whose workgroups cross scenes ----------------------------------------------

def _grid_crossing(b):
    g = b.grid()
    assert g["max_scenes_per_range"] >= 2, g
    return g


def test_atomic_batch_with_ranges_crossing_scenes():
    """Atomic (LDS float atomic) accumulation in a batch on 1 % of the
    persistent grid, so each projection workgroup walks patches of several
    scenes and re-evaluates omega at each scene change: every member matches
    the same scene stepped alone (to float-atomic order) and the oracle."""
    _print_case("atomic_batch_cross_scenes")
    scenes = batched_scenes(n_scenes=6, seed=3, n_range=(3, 9), iters=20)
    alone = [make_solver(s, accumulate_mode="atomic") for s in scenes]
    mixed = [make_solver(s, accumulate_mode="atomic") for s in scenes]
    for s in alone:
        s.solve()
        s.solve()
    with SceneBatch(mixed, grid_pct=1) as b:
        print(f"  grid {_grid_crossing(b)}", flush=True)
        b.frame(2)
        b.synchronize()
    for sc, a, m in zip(scenes, alone, mixed):
        xa, xm = a.mesh.verts.x.to_numpy(), m.mesh.verts.x.to_numpy()
        assert np.abs(xa - xm).max() <= 1e-5 * sc.diag, sc.name
        o = make_oracle(sc)
        o.frame(2)
        assert np.abs(xm - o.x).max() <= 1e-4 * sc.diag, sc.name


# ---- a batch member whose parameters change ---------------------------------------------

def test_batch_follows_member_parameter_changes():
    """A member's iteration count changed while it sits in a batch: the batch
    refuses to run while the members disagree, and once both are changed it
    runs the new count (bit-identical to the scenes stepped alone)."""
    _print_case("batch_member_params")
    scenes = batched_scenes(n_scenes=2, seed=5, n_range=(3, 6), iters=10)
    mixed = [make_solver(s) for s in scenes]
    alone = [make_solver(s) for s in scenes]
    with SceneBatch(mixed) as b:
        b.substep(1)
        b.synchronize()
        mixed[0].constraints_iter = 7
        mixed[0]._sync_params()
        with pytest.raises(Exception, match="no longer share"):
            b.substep(1)
        mixed[1].constraints_iter = 7
        mixed[1]._sync_params()
        b.substep(2)
        b.synchronize()
    for a in alone:
        sub_dt = a.dt / a.substeps_num
        for k in range(3):
            if k == 1:
                a.constraints_iter = 7
            a.apply_external_forces(sub_dt)
            a.solve_step(sub_dt)
            a.advance(sub_dt, a.damping)
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.mesh.verts.x.to_numpy(), m.mesh.verts.x.to_numpy())


def test_dropped_solver_is_freed_at_once():
    """The mesh views and the gravity callback hold the solver weakly, so a
    dropped solver frees its device state when its last reference goes, not
    whenever the cycle collector runs (a device free waits for the device)."""
    import gc
    import weakref
    print("[case] dropped solver", flush=True)
    gc.disable()
    try:
        s = make_solver(resting_cube(n=4, frames=1))
        s.gravity[None] = (0.0, -5.0, 0.0)
        s.solve()
        x = s.mesh.verts.x
        ref = weakref.ref(s)
        del s
        assert ref() is None
        with pytest.raises(ReferenceError):
            x.to_numpy()
    finally:
        gc.enable()
