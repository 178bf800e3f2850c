"""Pins of the oracle's per-iteration functions that the single-tet projection
pins leave open: the residual (a12), the Chebyshev blend and its prev/curr
rotation (a15), the split-mode scalar projections with their relaxation and
cons_num count (f1), the convergence monitor's means (f1), and the collider's
obstacle velocity (a6).  No GPU is needed.

Each pin checks the oracle's own code path (the per-tet functions the solver
loop calls, or the solver loop itself through its trace and log) against
quantities computed here independently: deformation gradients, determinants
and norms by numpy, gradients by finite differences, sums in float64, and the
reference's kernel sequence (neohookean_solver.py:131-142) for the blend.
"""
import numpy as np
import pytest

import oracle
from oracle import Oracle
from xpbd_gpu.mesh import box_grid

RNG = np.random.default_rng(4321)


def rest_tet(rng, jitter=0.2):
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    return base + rng.uniform(-jitter, jitter, size=base.shape)


def inv_dm_of(X):
    Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    return np.linalg.inv(Dm), np.linalg.det(Dm) / 6.0


def F_of(P, X):
    Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    Ds = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
    return Ds @ np.linalg.inv(Dm)


def plane(y):
    return dict(kind="plane", point=[0.0, y, 0.0], dir=[0.0, 1.0, 0.0], fixed=False)


# ---- a12: residual ------------------------------------------------------------------

@pytest.mark.parametrize("s", [0.7, 0.95, 1.0, 1.05, 1.6])
def test_residual_closed_form_uniform_stretch(s):
    """F = s I: C_D = 3 s^2 - 3, C_H = s^3 - 1, and with alpha~_D = 1/(mu V0 dt^2),
    alpha~_H = 1/(lambda V0 dt^2), r = sqrt((C_D + alpha~_D l_D)^2 + (C_H + alpha~_H l_H)^2)
    (neohookean_solver.py:298-329)."""
    X = rest_tet(RNG)
    idm, v0 = inv_dm_of(X)
    lame1, lame2, dt = 8.1e6, 3.4e5, 1.0 / 600.0
    for _ in range(5):
        lam_d, lam_h = RNG.normal(size=2) * 1e-4
        t = RNG.normal(size=3)
        r, cd, ch = oracle.residual_tet(s * X + t, idm, v0, lame1, lame2, dt, lam_d, lam_h)
        aD = 1.0 / (lame2 * v0 * dt * dt)
        aH = 1.0 / (lame1 * v0 * dt * dt)
        assert cd == pytest.approx(3 * s * s - 3, abs=1e-12)
        assert ch == pytest.approx(s ** 3 - 1, abs=1e-12)
        want = np.hypot(3 * s * s - 3 + aD * lam_d, s ** 3 - 1 + aH * lam_h)
        assert r == pytest.approx(want, rel=1e-9, abs=1e-12)


def test_residual_general_affine_and_lame_roles():
    """x = A X + t gives F = A exactly: C_D = |A|_F^2 - 3, C_H = det A - 1.  Swapping
    the two Lame constants changes r (each multiplier is paired with its own
    compliance)."""
    lame1, lame2, dt = 5.0e5, 2.0e5, 1e-3
    for _ in range(20):
        X = rest_tet(RNG)
        idm, v0 = inv_dm_of(X)
        A = np.eye(3) + 0.3 * RNG.normal(size=(3, 3))
        P = X @ A.T + RNG.normal(size=3)
        lam_d, lam_h = RNG.normal(size=2) * 1e-3
        r, cd, ch = oracle.residual_tet(P, idm, v0, lame1, lame2, dt, lam_d, lam_h)
        CD = (A * A).sum() - 3.0
        CH = np.linalg.det(A) - 1.0
        assert cd == pytest.approx(CD, rel=1e-10, abs=1e-12)
        assert ch == pytest.approx(CH, rel=1e-10, abs=1e-12)
        aD = 1.0 / (lame2 * v0 * dt * dt)
        aH = 1.0 / (lame1 * v0 * dt * dt)
        assert r == pytest.approx(np.hypot(CD + aD * lam_d, CH + aH * lam_h), rel=1e-9)
        r_sw, _, _ = oracle.residual_tet(P, idm, v0, lame2, lame1, dt, lam_d, lam_h)
        assert abs(r_sw - r) > 1e-6 * r


def _box_oracle(n=3, jitter=0.15, **kw):
    m = box_grid(n, jitter=jitter, seed=7)
    p = dict(precision=64, threads=2, frame_dt=0.01, dt=0.01, substeps_num=1,
             constraints_iter=1, block_neohookean_enable=True, young_modulus=2e5,
             poisson_ratio=0.4)
    p.update(kw)
    return m, Oracle(m.verts, m.tets, [plane(-10.0)], **p)


def _numpy_residuals(m, x, lam_d, lam_h, E, nu, dt):
    lame1, lame2 = oracle.lame(E, nu)
    X = m.verts.astype(np.float32).astype(np.float64)
    r = np.empty(m.n_tets)
    ch = np.empty(m.n_tets)
    cd = np.empty(m.n_tets)
    for i, t in enumerate(m.tets):
        F = F_of(x[t], X[t])
        v0 = np.linalg.det(np.stack([X[t[1]] - X[t[0]], X[t[2]] - X[t[0]],
                                     X[t[3]] - X[t[0]]], 1)) / 6.0
        cd[i] = (F * F).sum() - 3.0
        ch[i] = np.linalg.det(F) - 1.0
        r[i] = np.hypot(cd[i] + lam_d[i] / (lame2 * v0 * dt * dt),
                        ch[i] + lam_h[i] / (lame1 * v0 * dt * dt))
    return r, ch, cd


def test_residual_sums_through_the_solver():
    """The solver's residual pass on a jittered box with random positions and
    multipliers: sum r, sum |C_H| and sum |C_D| equal the per-tet values
    computed here, summed over all tets."""
    m, o = _box_oracle()
    x = m.verts.astype(np.float64) + 0.05 * RNG.normal(size=(m.n_verts, 3))
    lam_d = RNG.normal(size=m.n_tets) * 1e-4
    lam_h = RNG.normal(size=m.n_tets) * 1e-4
    o.pred_x = x
    o.set_lambda(lam_d, lam_h)
    dt = 1.0 / 300.0
    got = o.residual_now(dt)
    r, ch, cd = _numpy_residuals(m, x, lam_d, lam_h, 2e5, 0.4, dt)
    assert got[0] == pytest.approx(r.sum(), rel=1e-10)
    assert got[1] == pytest.approx(np.abs(ch).sum(), rel=1e-10)
    assert got[2] == pytest.approx(np.abs(cd).sum(), rel=1e-10)


def test_logged_residual_and_monitor_means_in_a_substep():
    """One substep of one iteration with the monitor on: the logged r_hat is the
    mean residual over the tets at the Jacobi iterate with the updated
    multipliers, and hydro_error / dev_error are the means of |C_H| and |C_D|
    there (neohookean_solver.py:256-329, 494-532)."""
    m, o = _box_oracle(monitor_convergence=True, cheb_enable=True)
    o.v = 0.3 * RNG.normal(size=(m.n_verts, 3))
    o.set_log(True)
    o.set_trace(True)
    o.substep(0.01)
    tr = o.trace()
    assert len(tr) == 1
    _, _, lam_d, lam_h = o.tet_state()
    r, ch, cd = _numpy_residuals(m, tr[0]["xj"], lam_d, lam_h, 2e5, 0.4, 0.01)
    row = o.log()[0]
    assert row[3] == pytest.approx(r.mean(), rel=1e-10)
    assert row[6] == pytest.approx(np.abs(ch).mean(), rel=1e-10)
    assert row[7] == pytest.approx(np.abs(cd).mean(), rel=1e-10)
    assert row[3] > 1e-6 and row[6] > 1e-6 and row[7] > 1e-6


# ---- a15: Chebyshev blend and rotation ------------------------------------------

def reference_chebyshev_sequence(y0, xjs, omegas, gamma):
    """The reference's buffers over a substep (neohookean_solver.py:121-142,
    466-549): prev <- pred once; per iteration curr <- pred, pred <- Jacobi
    iterate, _chebyshev_update(omega), then prev <- curr.  Here the Jacobi
    iterate of iteration k is given as an offset from curr, so the table is a
    function of y0, the offsets, omega and gamma alone."""
    pred = y0.copy()
    prev = pred.copy()
    out = []
    for dx, omega in zip(xjs, omegas):
        curr = pred.copy()
        pred = curr + dx
        x_new = gamma * (pred - curr) + curr
        pred = prev + omega * (x_new - prev)
        prev = curr.copy()
        out.append(pred.copy())
    return out


def test_blend_table_on_random_vectors():
    """Three iterations of the oracle's blend on random vectors, with omega = 1
    for the first (a warmup iteration, where gamma still under-relaxes) and
    then two Chebyshev weights, against the reference's buffer sequence."""
    n = 37
    y0 = RNG.normal(size=n)
    dxs = [RNG.normal(size=n) for _ in range(3)]
    omegas = [1.0, 1.7, 1.85]
    gamma = 0.6666
    want = reference_chebyshev_sequence(y0, dxs, omegas, gamma)
    x = y0.copy()
    prev = y0.copy()
    lib = oracle._load()
    for k in range(3):
        curr = x.copy()
        x = curr + dxs[k]
        lib.oracle_cheb_blend64(oracle._dptr(x), oracle._dptr(curr), oracle._dptr(prev), n,
                                omegas[k], gamma)
        prev = curr
        np.testing.assert_allclose(x, want[k], rtol=1e-14, atol=1e-14)
    # warmup iteration: omega = 1 leaves the gamma blend, not the Jacobi iterate
    first = y0 + gamma * dxs[0]
    np.testing.assert_allclose(want[0], first, rtol=1e-14)
    assert np.abs(want[0] - (y0 + dxs[0])).max() > 0.1


def test_blend_sequence_inside_a_substep():
    """The solver loop's own iterates (trace of one substep with contact):
    every blended iterate is prev + omega (gamma (x_J - curr) + curr - prev),
    with prev the iterate of the previous iteration (the substep's start for
    the first), omega the logged one (1 during the warmup), and the next
    iteration starts from it."""
    m, o = _box_oracle(n=3, constraints_iter=9, cheb_enable=True, cheb_warmup=5,
                       cheb_gamma=0.6666)
    o.v = 0.5 * RNG.normal(size=(m.n_verts, 3))
    o.set_log(True)
    o.set_trace(True)
    o.substep(0.01)
    tr = o.trace()
    log = o.log()
    assert len(tr) == 9
    prev = tr[0]["curr"]
    for k, it in enumerate(tr):
        assert it["omega"] == log[k, 5]
        if k < 5:
            assert it["omega"] == 1.0
        else:
            assert it["omega"] > 1.0
        want = prev + it["omega"] * (0.6666 * (it["xj"] - it["curr"]) + it["curr"] - prev)
        np.testing.assert_allclose(it["next"], want, rtol=1e-13, atol=1e-15)
        assert np.abs(it["next"] - it["xj"]).max() > 0  # gamma applies in every iteration
        if k + 1 < len(tr):
            assert np.array_equal(tr[k + 1]["curr"], it["next"])
        prev = it["curr"]


# ---- f1: split-mode projections ------------------------------------------------------

def fd_grad(C, P, h=1e-6):
    g = np.zeros_like(P)
    for k in range(4):
        for c in range(3):
            Pp, Pm = P.copy(), P.copy()
            Pp[k, c] += h
            Pm[k, c] -= h
            g[k, c] = (C(Pp) - C(Pm)) / (2 * h)
    return g


@pytest.mark.parametrize("hydro", [True, False])
def test_split_projection_closed_form(hydro):
    """dlambda = -relax (C + alpha~ lambda) / (sum_i w_i |g_i|^2 + alpha~) with the
    constraint's own Lame constant (lambda for C_H, mu for C_D), corner i =
    dlambda w_i g_i (neohookean_solver.py:144-254); gradients by finite
    differences."""
    lame = 5.0e5 if hydro else 2.0e5
    dt, relax = 2e-3, 1.2
    for _ in range(10):
        X = rest_tet(RNG)
        idm, v0 = inv_dm_of(X)
        P = X @ (np.eye(3) + 0.2 * RNG.normal(size=(3, 3))).T
        w = RNG.uniform(0.2, 2.0, 4)
        lam0 = RNG.normal() * 1e-3

        def C(Q):
            F = F_of(Q, X)
            return np.linalg.det(F) - 1.0 if hydro else (F * F).sum() - 3.0

        g = fd_grad(C, P)
        at = 1.0 / (lame * v0 * dt * dt)
        ws = float(np.sum(w * np.sum(g * g, 1)))
        dl = -relax * (C(P) + at * lam0) / (ws + at)
        r = oracle.split_tet(P, w, idm, v0, lame, dt, relax, hydro, lam=lam0)
        assert r["ok"] == 1
        assert r["C"] == pytest.approx(C(P), rel=1e-10, abs=1e-12)
        assert r["w_sum"] == pytest.approx(ws, rel=1e-6)
        assert r["lam"] == pytest.approx(lam0 + dl, rel=1e-6, abs=1e-12)
        np.testing.assert_allclose(r["corner"], dl * w[:, None] * g, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("hydro", [True, False])
def test_split_projection_skips_when_w_sum_is_at_most_eps(hydro):
    """w_sum <= EPSILON (all four vertices pinned, or masses so large that
    sum w |g|^2 <= 1e-12): no multiplier change, no correction, no cons_num."""
    X = rest_tet(RNG)
    idm, v0 = inv_dm_of(X)
    P = 1.3 * X
    for w in (np.zeros(4), np.full(4, 1e-16)):
        r = oracle.split_tet(P, w, idm, v0, 3e5, 1e-3, 1.2, hydro, lam=0.25)
        assert r["ok"] == 0 and r["w_sum"] <= 1e-12
        assert r["lam"] == 0.25
        assert np.all(r["corner"] == 0)


def test_split_cons_num_counts_only_projected_constraints():
    """Two tets sharing vertex 3; the second is collapsed to a point, so both of
    its constraints have zero gradients and are skipped.  After one split-mode
    iteration (no Chebyshev) each vertex moves by the sum of the corrections
    of its projected constraints divided by their number (cons_num +1 per
    projected constraint, neohookean_solver.py:186-194, 246-254, 115-119): 2
    for vertex 3, not 4."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                  [1, 0, 1], [0, 1, 1], [0, 0, 2]], dtype=np.float64)
    tets = np.array([[0, 1, 2, 3], [3, 4, 5, 6]], dtype=np.int32)
    D = np.stack([X[4] - X[3], X[5] - X[3], X[6] - X[3]], 1)
    assert np.linalg.det(D) > 0
    E, nu, dt = 1e4, 0.3, 0.01
    o = Oracle(X.astype(np.float32), tets, [plane(-10.0)], precision=64, threads=1,
               constraints_iter=1, frame_dt=dt, dt=dt, substeps_num=1, cheb_enable=False,
               block_neohookean_enable=False, young_modulus=E, poisson_ratio=nu,
               hydrostatic_relaxation=1.2, deviatoric_relaxation=1.1)
    o.set_gravity([0.0, 0.0, 0.0])
    x0 = X.copy()
    x0[0] += [0.1, -0.05, 0.02]
    x0[4:] = x0[3]  # tet 1 collapsed onto vertex 3
    o.x = x0
    o.set_trace(True)
    o.substep(dt)
    y = o.trace()[0]["curr"]
    np.testing.assert_array_equal(y, x0)
    _, im = o.mass()
    lame1, lame2 = oracle.lame(E, nu)
    acc = np.zeros_like(X)
    cnt = np.zeros(len(X))
    for t in tets:
        idm, v0 = inv_dm_of(X[t])
        for hydro, lame, relax in ((True, lame1, 1.2), (False, lame2, 1.1)):
            r = oracle.split_tet(y[t], im[t], idm, v0, lame, dt, relax, hydro)
            if r["ok"]:
                acc[t] += r["corner"]
                cnt[t] += 1
    assert cnt[3] == 2 and cnt[4] == 0
    want = y.copy()
    moved = cnt > 0
    want[moved] += acc[moved] / cnt[moved, None]
    np.testing.assert_allclose(o.trace()[0]["xj"], want, rtol=1e-12, atol=1e-15)
    assert np.abs(want[3] - y[3]).max() > 1e-6


# ---- a6: the obstacle's velocity ------------------------------------------------------

def test_obstacle_velocity_enters_the_inside_test():
    """SdfModel.check (sdf.py:13-27): a vertex below a non-fixed plane counts as
    inside when n . (vel_sdf - v) > 0.  A vertex at rest is outside for an
    obstacle moving away (vel = -n) and inside for one moving towards it."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float32) + 10.0
    t = np.array([[0, 1, 2, 3]], dtype=np.int32)
    for vel, inside in (((0.0, -1.0, 0.0), False), ((0.0, 1.0, 0.0), True),
                        ((0.0, 0.0, 0.0), False)):
        sd = dict(plane(0.0), vel=vel)
        o = Oracle(v, t, [sd], precision=64, threads=1)
        d, n, ins = o.sdf_eval(0, (0.3, -0.2, 0.1), v=(0.0, 0.0, 0.0))
        assert d == pytest.approx(-0.2) and ins is inside
    # a vertex moving up at 2 through an obstacle moving up at 1: n . (1 - 2) < 0
    o = Oracle(v, t, [dict(plane(0.0), vel=(0.0, 1.0, 0.0))], precision=64, threads=1)
    assert o.sdf_eval(0, (0.0, -0.1, 0.0), v=(0.0, 2.0, 0.0))[2] is False
    assert o.sdf_eval(0, (0.0, -0.1, 0.0), v=(0.0, 0.5, 0.0))[2] is True
