"""Pins of the CPU oracle against closed forms, invariants, library routines
and brute force (no GPU needed).

The reference has no tests and no golden vectors (its only fixtures are the
TetGen meshes under assets/tetmesh), so every oracle function is pinned here
by mathematics that does not restate the oracle's own formula.
"""
import math

import numpy as np
import pytest

import oracle
from scenes import resting_cube, dropped_cube, twisted_grid, make_oracle
from xpbd_gpu.mesh import box_grid

RNG = np.random.default_rng(1234)


def rand_rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def rest_tet(rng, jitter=0.2):
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    return base + rng.uniform(-jitter, jitter, size=base.shape)


def inv_dm_of(X):
    Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    return np.linalg.inv(Dm), np.linalg.det(Dm) / 6.0


# ---- a1: Lame constants ----------------------------------------------------------

@pytest.mark.parametrize("E,nu", [(1e5, 0.45), (1e6, 0.48), (2e5, 0.495), (5e8, 0.4999)])
def test_lame_textbook(E, nu):
    lam, mu = oracle.lame(E, nu)
    # textbook forms: lambda = E nu / ((1+nu)(1-2nu)), mu = E / (2(1+nu))
    assert lam == pytest.approx(E * nu / ((1 + nu) * (1 - 2 * nu)), rel=1e-12)
    assert mu == pytest.approx(E / 2 / (1 + nu), rel=1e-12)


def test_lame_values():
    lam, mu = oracle.lame(1e5, 0.45)
    assert lam == pytest.approx(310344.83, abs=0.01)
    assert mu == pytest.approx(34482.76, abs=0.01)
    lam, mu = oracle.lame(1e6, 0.48)
    assert lam == pytest.approx(8108108.1, abs=0.1)
    assert mu == pytest.approx(337837.8, abs=0.1)


# ---- a3: inverse and determinant ----------------------------------------------------

def test_inv3_det3_vs_numpy():
    for _ in range(200):
        a = RNG.normal(size=(3, 3))
        inv, det = oracle.inv3(a)
        np.testing.assert_allclose(inv, np.linalg.inv(a), rtol=1e-9, atol=1e-9)
        assert det == pytest.approx(np.linalg.det(a), rel=1e-10, abs=1e-12)


def test_setup_rest_state_of_grid():
    """Unjittered grid (h a power of two, so exact in f32): V0 = h^3/6 for every
    tet; inv_Dm maps the rest edges to the identity."""
    n = 4
    m = box_grid(n)
    o = make_oracle(dropped_cube(n=n), precision=64)
    v0, idm, _, _ = o.tet_state()
    np.testing.assert_allclose(v0, (1.0 / n) ** 3 / 6.0, rtol=1e-12)
    X = m.verts.astype(np.float64) + np.array([0.0, 0.5, 0.0])
    t = m.tets
    Dm = np.stack([X[t[:, 1]] - X[t[:, 0]], X[t[:, 2]] - X[t[:, 0]], X[t[:, 3]] - X[t[:, 0]]],
                  axis=2)
    np.testing.assert_allclose(Dm @ idm, np.broadcast_to(np.eye(3), Dm.shape), atol=1e-12)


# ---- a4: lumped mass -------------------------------------------------------------

def test_lumped_mass_closed_form():
    """Sum m = total_mass; on a regular grid an interior vertex touches 24 tets
    (mass = 24 V/4 scaled), a corner on the Kuhn diagonal touches 6."""
    n = 4
    s = dropped_cube(n=n, drop=0.5)
    o = make_oracle(s, precision=64)
    m, im = o.mass()
    assert m.sum() == pytest.approx(1000.0, rel=1e-12)
    h = 1.0 / n
    vt = h ** 3 / 6
    scale = 1000.0 / (n ** 3 * 6 * vt)
    val = np.bincount(s.mesh.tets.ravel(), minlength=s.mesh.n_verts)
    np.testing.assert_allclose(m, val * vt / 4 * scale, rtol=1e-12)
    assert val.max() == 24
    np.testing.assert_allclose(im, 1.0 / m, rtol=1e-12)


def test_fixed_sdf_pins_initially_inside():
    from scenes import jittered_box
    s = jittered_box(n=8, frames=1, iters=2)
    o = make_oracle(s, precision=64)
    _, im = o.mass()
    X = s.mesh.verts.astype(np.float64)
    inside = np.linalg.norm(X - np.array([0.5, 0.0, 0.5]), axis=1) < 0.3
    inside |= X[:, 1] < 0.0
    assert np.array_equal(im == 0, inside)


# ---- a9/a10: block projection on single tets ----------------------------------

def C_of(P, X):
    idm, _ = inv_dm_of(X)
    Ds = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
    F = Ds @ idm
    return (F * F).sum() - 3.0, np.linalg.det(F) - 1.0


def test_gradients_match_finite_differences():
    for _ in range(20):
        X = rest_tet(RNG)
        P = X @ (np.eye(3) + 0.3 * RNG.normal(size=(3, 3))).T + RNG.normal(size=3)
        idm, v0 = inv_dm_of(X)
        r = oracle.block_tet(P, np.ones(4), idm, v0, 1e5, 1e4, 1e-3)
        h = 1e-6
        for k in range(4):
            for c in range(3):
                Pp, Pm = P.copy(), P.copy()
                Pp[k, c] += h
                Pm[k, c] -= h
                dD = (C_of(Pp, X)[0] - C_of(Pm, X)[0]) / (2 * h)
                dH = (C_of(Pp, X)[1] - C_of(Pm, X)[1]) / (2 * h)
                assert r["gD"][k, c] == pytest.approx(dD, rel=1e-6, abs=1e-6)
                assert r["gH"][k, c] == pytest.approx(dH, rel=1e-6, abs=1e-6)
        cd, ch = C_of(P, X)
        assert r["C_D"] == pytest.approx(cd, rel=1e-10, abs=1e-12)
        assert r["C_H"] == pytest.approx(ch, rel=1e-10, abs=1e-12)


def test_gradients_sum_to_zero():
    for _ in range(20):
        X = rest_tet(RNG)
        P = X + 0.2 * RNG.normal(size=X.shape)
        idm, v0 = inv_dm_of(X)
        r = oracle.block_tet(P, RNG.uniform(0.5, 2, 4), idm, v0, 1e5, 1e4, 1e-3)
        np.testing.assert_allclose(r["gD"].sum(axis=0), 0, atol=1e-12)
        np.testing.assert_allclose(r["gH"].sum(axis=0), 0, atol=1e-12)


def test_rest_and_rigid_motion_give_no_correction():
    for _ in range(10):
        X = rest_tet(RNG)
        idm, v0 = inv_dm_of(X)
        Rm = rand_rot(RNG)
        P = X @ Rm.T + RNG.normal(size=3)
        r = oracle.block_tet(P, np.ones(4), idm, v0, 1e5, 1e4, 1e-3)
        assert abs(r["C_D"]) < 1e-12 and abs(r["C_H"]) < 1e-12
        np.testing.assert_allclose(r["corner"], 0, atol=1e-12)
        np.testing.assert_allclose(r["dlam"], 0, atol=1e-10)


def test_uniform_stretch_closed_form():
    """F = s I: C_D = 3 s^2 - 3, C_H = s^3 - 1."""
    X = rest_tet(RNG)
    idm, v0 = inv_dm_of(X)
    for s in (0.5, 0.9, 1.1, 2.0):
        r = oracle.block_tet(s * X, np.ones(4), idm, v0, 1e5, 1e4, 1e-3)
        assert r["C_D"] == pytest.approx(3 * s * s - 3, rel=1e-10)
        assert r["C_H"] == pytest.approx(s ** 3 - 1, rel=1e-10)


def test_block_2x2_system_is_solved():
    """A dlam = b with A = [[sum w gD.gD + aD, sum w gD.gH], [., sum w gH.gH + aH]],
    b = -(C + alpha_tilde lambda)."""
    for _ in range(50):
        X = rest_tet(RNG)
        idm, v0 = inv_dm_of(X)
        P = X + 0.1 * RNG.normal(size=X.shape)
        w = RNG.uniform(0.1, 3.0, 4)
        lam0 = RNG.normal(size=2) * 1e-3
        lame1, lame2, dt = 8e6, 3e5, 1 / 600
        r = oracle.block_tet(P, w, idm, v0, lame1, lame2, dt, lam=lam0)
        gD, gH = r["gD"], r["gH"]
        aD = 1 / (lame2 * v0) / dt ** 2
        aH = 1 / (lame1 * v0) / dt ** 2
        A = np.array([[np.sum(w * np.sum(gD * gD, 1)) + aD, np.sum(w * np.sum(gD * gH, 1))],
                      [np.sum(w * np.sum(gD * gH, 1)), np.sum(w * np.sum(gH * gH, 1)) + aH]])
        b = -np.array([r["C_D"] + aD * lam0[0], r["C_H"] + aH * lam0[1]])
        np.testing.assert_allclose(A @ r["dlam"], b, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(r["lam"], lam0 + r["dlam"], rtol=1e-12)
        np.testing.assert_allclose(
            r["corner"], w[:, None] * (gD * r["dlam"][0] + gH * r["dlam"][1]), rtol=1e-12,
            atol=1e-15)


def test_solve2x2_vs_numpy_and_fallback():
    for _ in range(200):
        A = RNG.normal(size=(2, 2))
        A = A @ A.T + 0.1 * np.eye(2)
        b = RNG.normal(size=2)
        ok, x = oracle.solve2x2(A[0, 0], A[0, 1], A[1, 0], A[1, 1], b[0], b[1])
        assert ok == 1
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9)
    # singular: rank one
    ok, x = oracle.solve2x2(1.0, 2.0, 2.0, 4.0, 1.0, 1.0)
    assert ok == 0
    # zero pivot after swap
    ok, x = oracle.solve2x2(0.0, 0.0, 0.0, 0.0, 1.0, 1.0)
    assert ok == 0


def test_fully_pinned_tet_leaves_diag_alpha():
    """w = 0: A = diag(aD, aH), so dlam = -(C + a lambda) / a and no motion."""
    X = rest_tet(RNG)
    idm, v0 = inv_dm_of(X)
    P = 1.2 * X
    r = oracle.block_tet(P, np.zeros(4), idm, v0, 1e5, 1e4, 1e-2)
    aD = 1 / (1e4 * v0) / 1e-4
    aH = 1 / (1e5 * v0) / 1e-4
    np.testing.assert_allclose(r["dlam"], [-r["C_D"] / aD, -r["C_H"] / aH], rtol=1e-12)
    np.testing.assert_allclose(r["corner"], 0, atol=0)


# ---- a13/a14: rho and omega ----------------------------------------------------

def test_omega_recurrence_fixed_point():
    """For fixed rho the Chebyshev recurrence w <- 4/(4 - rho^2 w), started at
    2/(2 - rho^2), converges to the optimal 2/(1 + sqrt(1 - rho^2))."""
    for rho in (0.3, 0.9, 0.99):
        w = 2.0 / (2.0 - rho * rho)
        for _ in range(2000):
            w = 4.0 / (4.0 - rho * rho * w)
        assert w == pytest.approx(2.0 / (1.0 + math.sqrt(1.0 - rho * rho)), rel=1e-6)


def test_oracle_omega_schedule_in_log():
    s = resting_cube(n=4, frames=1)
    o = make_oracle(s, precision=64)
    o.frame(1)
    L = o.log()
    warm = s.params["cheb_warmup"]
    iters = s.params["constraints_iter"]
    for sub in range(s.params["substeps_num"]):
        rows = L[L[:, 1] == sub]
        assert np.all(rows[:warm, 5] == 1.0)
        rho = rows[warm, 4]
        assert rows[warm, 5] == pytest.approx(2 / (2 - rho * rho), rel=1e-15)
        for k in range(warm + 1, iters):
            rho = rows[k, 4]
            assert rows[k, 5] == pytest.approx(4 / (4 - rho * rho * rows[k - 1, 5]), rel=1e-15)
        # rho estimator: EMA of clamped r ratio
        for k in range(1, iters):
            if rows[k - 1, 3] > 1e-12:
                inst = min(max(rows[k, 3] / (rows[k - 1, 3] + 1e-12), 1e-5), 0.99999)
                prev_rho = rows[k - 1, 4]
                assert rows[k, 4] == pytest.approx(0.8 * prev_rho + 0.2 * inst, rel=1e-12)


# ---- a5/a6/a16: free fall, collision, advance ------------------------------------

def test_free_fall_is_ballistic_with_damping():
    """No strain: v_{k+1} = (v_k + g dt) e^{-5 dt}, x_{k+1} = x_k + (v_k + g dt) dt."""
    s = dropped_cube(n=2, drop=5.0, frames=3)
    o = make_oracle(s, precision=64)
    o.frame(3)
    dt = s.params["dt"] / s.params["substeps_num"]
    y = s.mesh.verts[:, 1].astype(np.float64)
    v = 0.0
    for _ in range(3 * s.params["substeps_num"]):
        v = v - 9.8 * dt
        y = y + v * dt
        v = v * math.exp(-5.0 * dt)
    np.testing.assert_allclose(o.x[:, 1], y, atol=1e-12)
    np.testing.assert_allclose(o.x[:, [0, 2]], s.mesh.verts[:, [0, 2]], atol=1e-12)


def test_plane_and_sphere_projection_land_on_surface():
    from scenes import plane, sphere
    from oracle import Oracle
    v = np.array([[0, 0.05, 0], [1, 0.05, 0], [0, 1.05, 0], [0, 0.05, 1]], dtype=np.float32)
    t = np.array([[0, 1, 2, 3]], dtype=np.int32)
    o = Oracle(v, t, [plane(0.0, fixed=False)], precision=64, young_modulus=1e3,
               constraints_iter=0, frame_dt=0.1, dt=0.1, substeps_num=1, cheb_enable=False)
    o.v = np.tile([0.0, -2.0, 0.0], (4, 1))
    o.set_gravity([0, 0, 0])
    o.substep(0.1)  # pred = x + v dt = 0.05 - 0.2 < 0 for the bottom three
    x = o.x
    np.testing.assert_allclose(x[[0, 1, 3], 1], 0.0, atol=1e-15)
    assert x[2, 1] == pytest.approx(float(np.float32(1.05)) - 0.2, abs=1e-12)
    # sphere: a point pushed in lands on radius
    o2 = Oracle(v, t, [sphere((0.0, -1.0, 0.0), 1.2, fixed=True)], precision=64,
                young_modulus=1e3, constraints_iter=0, frame_dt=0.1, dt=0.1, substeps_num=1,
                cheb_enable=False)
    _, im = o2.mass()
    assert im[0] == 0  # starts inside a fixed sphere -> pinned
    o2.v = np.tile([0.0, -1.0, 0.0], (4, 1))
    o2.set_gravity([0, 0, 0])
    o2.substep(0.1)
    r = np.linalg.norm(o2.x - np.array([0, -1.0, 0]), axis=1)
    free_inside = im > 0
    assert np.all(r[free_inside] >= 1.2 - 1e-9)


def test_outward_moving_vertex_untouched_by_nonfixed_sdf():
    from scenes import plane
    from oracle import Oracle
    v = np.array([[0, -0.5, 0], [1, -0.5, 0], [0, 0.5, 0], [0, -0.5, 1]], dtype=np.float32)
    t = np.array([[0, 1, 2, 3]], dtype=np.int32)
    o = Oracle(v, t, [plane(0.0, fixed=False)], precision=64, constraints_iter=0, frame_dt=0.1,
               dt=0.1, substeps_num=1, cheb_enable=False)
    o.set_gravity([0, 0, 0])
    o.v = np.tile([0.0, 1.0, 0.0], (4, 1))  # moving out of the plane
    o.substep(0.1)
    np.testing.assert_allclose(o.x[:, 1], v[:, 1] + 0.1, atol=1e-12)


# ---- a11/a17: brute force Jacobi iteration on 2 tets ------------------------------

def test_two_tet_jacobi_average_brute_force():
    """Oracle pred after one non-Chebyshev iteration equals y + (sum of the
    per-tet corrections) / (number of tets at the vertex)."""
    from scenes import plane
    from oracle import Oracle
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=np.float64)
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], dtype=np.int32)
    # orient second tet positively
    if np.linalg.det(np.stack([X[2] - X[1], X[3] - X[1], X[4] - X[1]], 1)) < 0:
        tets[1] = [1, 3, 2, 4]
    o = Oracle(X.astype(np.float32), tets, [plane(-10.0)], precision=64, constraints_iter=1,
               frame_dt=0.01, dt=0.01, substeps_num=1, cheb_enable=False,
               block_neohookean_enable=True, young_modulus=1e4, poisson_ratio=0.3)
    o.set_gravity([0, 0, 0])
    v0 = RNG.normal(size=X.shape)
    o.v = v0
    m, im = o.mass()
    o.substep(0.01)
    y = X + v0 * 0.01
    lam1, lam2 = oracle.lame(1e4, 0.3)
    acc = np.zeros_like(X)
    cnt = np.zeros(len(X))
    for t in tets:
        idm, vol = inv_dm_of(X[t])
        r = oracle.block_tet(y[t], im[t], idm, vol, lam1, lam2, 0.01)
        acc[t] += r["corner"]
        cnt[t] += 1
    pred = y + acc / cnt[:, None]
    np.testing.assert_allclose(o.x, pred, rtol=1e-9, atol=1e-12)


# ---- properties of full runs ------------------------------------------------------

def test_threads_do_not_change_bits():
    s = resting_cube(n=6, frames=2)
    a = make_oracle(s, threads=1)
    b = make_oracle(s, threads=4)
    a.frame(2)
    b.frame(2)
    assert np.array_equal(a.x, b.x)
    assert np.array_equal(a.log(), b.log())


def test_f32_tracks_f64():
    s = resting_cube(n=6, frames=3)
    a = make_oracle(s, precision=32)
    b = make_oracle(s, precision=64)
    a.frame(3)
    b.frame(3)
    assert np.abs(a.x - b.x).max() <= 1e-4 * s.diag


def test_residual_zero_at_rest():
    s = dropped_cube(n=3, drop=5.0, frames=1)
    s.gravity = (0.0, 0.0, 0.0)
    o = make_oracle(s, precision=64)
    o.set_gravity([0, 0, 0])
    o.frame(1)
    assert np.abs(o.log()[:, 3]).max() < 1e-12
    np.testing.assert_allclose(o.x, s.mesh.verts, atol=1e-12)


def test_stress_run_stays_finite_and_relaxes():
    s = twisted_grid(n=6, iters=200)
    o = make_oracle(s, precision=64)
    o.frame(1)
    L = o.log()
    assert np.all(np.isfinite(o.x))
    assert L[-1, 3] < L[0, 3]


def _sdf_oracle(sdfs, precision=64):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float32) + 10.0
    t = np.array([[0, 1, 2, 3]], dtype=np.int32)
    from oracle import Oracle
    return Oracle(v, t, sdfs, precision=precision, threads=1)


def test_cube_sdf_closed_forms():
    """CubeSdf (sdf.py:98-124): distance to a box of half size h around c."""
    o = _sdf_oracle([dict(kind="cube", center=(1.0, 2.0, 3.0), size=(2.0, 4.0, 6.0), fixed=False)])
    c = np.array([1.0, 2.0, 3.0])
    d, n, _ = o.sdf_eval(0, c + [1.5, 0, 0])        # 0.5 outside the +x face
    assert d == pytest.approx(0.5) and np.array_equal(n, [1, 0, 0])
    d, n, _ = o.sdf_eval(0, c + [0, -2.25, 0])       # 0.25 outside the -y face
    assert d == pytest.approx(0.25) and np.array_equal(n, [0, -1, 0])
    d, _, _ = o.sdf_eval(0, c + [2.0, 3.0, 4.0])     # outside a corner: Euclidean
    assert d == pytest.approx(np.sqrt(1.0 + 1.0 + 1.0))
    d, n, _ = o.sdf_eval(0, c + [0.2, 0.5, -2.6])    # inside: minus the nearest face distance
    assert d == pytest.approx(-0.4)
    assert np.array_equal(n, [0, 0, -1])             # axis of the largest |x - c|, not nearest face
    d, n, _ = o.sdf_eval(0, c + [0.9, 0.1, 0.1])
    assert d == pytest.approx(-0.1) and np.array_equal(n, [1, 0, 0])


def test_cube_normal_uses_largest_offset_not_nearest_face():
    """The reference's normal picks the axis of the largest |x - c| (sdf.py:114-123),
    which for a non-cubic box is not always the nearest face."""
    o = _sdf_oracle([dict(kind="cube", center=(0.0, 0.0, 0.0), size=(10.0, 1.0, 1.0), fixed=True)])
    d, n, inside = o.sdf_eval(0, (1.0, 0.45, 0.0))
    assert d == pytest.approx(-0.05)                  # nearest face is +y
    assert np.array_equal(n, [1, 0, 0]) and inside    # but the normal is +x; fixed: inside


def test_hang_sdf_union_and_sphere0_normal():
    """HangSdfModel (sdf.py:29-56): min over spheres of sqrt(|q|^2 + eps) - r, and
    the normal is taken towards sphere 0 whichever sphere is nearest."""
    pts = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]
    o = _sdf_oracle([dict(kind="hang", points=pts, radius=0.013)])
    d, n, inside = o.sdf_eval(0, (1.0, 0.005, 0.0))   # inside sphere 1
    assert d == pytest.approx(0.005 - 0.013, abs=1e-9) and inside  # always fixed
    assert n == pytest.approx([1.0, 0.005, 0.0] / np.hypot(1.0, 0.005))  # towards sphere 0
    d, _, inside = o.sdf_eval(0, (0.5, 0.0, 0.0))
    assert d == pytest.approx(0.5 - 0.013) and not inside


def test_nonfixed_cube_leaves_outward_moving_vertex():
    o = _sdf_oracle([dict(kind="cube", center=(0.0, 0.0, 0.0), size=(2.0, 2.0, 2.0), fixed=False)])
    _, n, inside = o.sdf_eval(0, (0.9, 0.0, 0.0), v=(1.0, 0.0, 0.0))   # moving out along +x
    assert not inside
    _, n, inside = o.sdf_eval(0, (0.9, 0.0, 0.0), v=(-1.0, 0.0, 0.0))  # moving in
    assert inside and np.array_equal(n, [1, 0, 0])
