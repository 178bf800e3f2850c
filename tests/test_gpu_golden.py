"""The gfx950 solver against the committed golden fixtures.

Each fixture is run through the C ABI from its stored inputs (and starting
state) and compared with the fixture's f32 oracle positions and residual log.
The f32 oracle is also run three more times from the same inputs with its
rounding changed (scenes.ROUNDING_VARIANTS: scatter and sum order permuted,
then also 1/x and sqrt moved by up to one ulp, then also fused multiply-adds
and the compliance reassociated; seeded), which measures the case's own f32
noise.
Gates, as in test_gpu_parity:
  positions  max |x_gpu - x_f32| <= max(1e-4 x bbox diagonal, 2 x the
             largest distance between two of the fixture's f32 run and its
             rounding variants).
  residual   r_hat within 1 % of the fixture's above the case's f32 floor (the
             level above which the variants agree within 0.5 %).  Where that
             floor leaves less than half of the iterations gated (the
             reference demo in contact, E 5e8, nu 0.4999: its curve is
             round-off dominated, the f32 rounding variants differ from each
             other by tens of percent), the curve is also compared with the
             f64 one: per substep the device's median relative deviation from
             f64 must be at most 3x the f32 oracle's plus 1 %, and its largest
             at most 10x the oracle's largest plus 1 %.
  monitor    with the convergence monitor on, each of its per-iteration means
             (|C_H|, |C_D|) within 1 % of the fixture's above that column's
             f32 floor (found from the variants as for r_hat); where the floor
             leaves less than half of the iterations gated (the stiff
             comparison-preset cases, whose nearly incompressible |C_H| is
             round-off), also the FP64-relative check above on that column.
"""
import numpy as np
import pytest

import golden_cases
from scenes import make_solver, rounding_variants, position_spread, rhat_floor
from test_gpu_parity import POS_TOL, RES_TOL, VARIANT_AGREE, check_residual_curve

# gating of the FP64-relative curve comparison: r_hat at least 1 % of the
# substep's first and at least 1e-5
RES_FLOOR = 1e-2
RES_ABS = 1e-5

pytestmark = pytest.mark.gpu


class _Log:
    """check_residual_curve's view of the stored oracle log and its
    rounding variants."""

    def __init__(self, rows, variants):
        self._rows = rows
        self.variants = variants

    def log(self):
        return self._rows


def golden_variants(case):
    """Rounding variants of the f32 oracle from the fixture's inputs."""
    sc = golden_cases.scene(case)

    def setup(o):
        if case["v0"] is not None:
            o.v = case["v0"]
        if case.get("rho0") is not None:
            o.rho_hat = case["rho0"]
    return rounding_variants(sc, frames=case["frames"], setup=setup)


def gated_rows(log, iters, col=3):
    """Rows of the residual gate of test_gpu_parity: r_hat at least 1 % of
    the substep's first and at least RES_ABS.  For a monitor column: at least
    1e-4 of the column's largest value."""
    r = log[:, col]
    if col != 3:
        return r >= 1e-4 * r.max()
    m = np.zeros(len(r), bool)
    for s0 in range(0, len(r), iters):
        ro = r[s0:s0 + iters]
        if ro[0] > 0:
            m[s0:s0 + iters] = (ro >= RES_FLOOR * ro[0]) & (ro >= RES_ABS)
    return m


def check_curve_vs_fp64(solver, case, col=3, values=None):
    """Round-off dominated curves: per substep, the median relative deviation of
    the device's r_hat (or, with `col` and `values`, another logged column)
    from f64 over the gated iterations must be at most 3x the f32 oracle's
    plus 1 %, and no gated iteration may deviate by more than 10x the f32
    oracle's worst plus 1 % (no divergence)."""
    it = int(case["params"]["constraints_iter"])
    g = np.array([r["r_hat"] for r in solver.residual_history]) if values is None else values
    assert len(g) == len(case["log64"])
    m = gated_rows(case["log64"], it, col)
    worst_med, worst_max = 0.0, 0.0
    for s0 in range(0, len(g), it):
        mm = m[s0:s0 + it]
        if not mm.any():
            continue
        r64 = case["log64"][s0:s0 + it, col][mm]
        d32 = np.abs(case["log32"][s0:s0 + it, col][mm] - r64) / r64
        dg = np.abs(g[s0:s0 + it][mm] - r64) / r64
        worst_med = max(worst_med, float(np.median(dg) / (3.0 * np.median(d32) + RES_TOL)))
        worst_max = max(worst_max, float(dg.max() / (10.0 * d32.max() + RES_TOL)))
    what = "residual curve" if col == 3 else f"log column {col}"
    print(f"  {what} (round-off dominated): median ratio {worst_med:.3f}, "
          f"max ratio {worst_max:.3f} over {int(m.sum())} gated iterations", flush=True)
    assert m.sum() > 0
    assert worst_med <= 1.0 and worst_max <= 1.0


@pytest.mark.parametrize("name", golden_cases.names())
def test_gpu_matches_golden(name):
    print(f"[case] golden {name}", flush=True)
    case = golden_cases.load(name)
    sc = golden_cases.scene(case)
    solver = make_solver(sc, log_residuals=True)
    if case["v0"] is not None:
        solver.mesh.verts.v.from_numpy(case["v0"])
    if case.get("rho0") is not None:
        solver.rho_hat = case["rho0"]
    for _ in range(case["frames"]):
        solver.solve()
    x = solver.mesh.verts.x.to_numpy().astype(np.float64)
    d = golden_cases.diag(case["x32"])
    variants = golden_variants(case)
    x32 = case["x32"].astype(np.float64)
    spread = position_spread(x32, variants) / d
    e32 = np.abs(x - x32).max() / d
    e64 = np.abs(x - case["x64"]).max() / d
    eo = np.abs(x32 - case["x64"]).max() / d
    print(f"  vs f32 oracle {e32:.2e} x diag; vs f64 {e64:.2e} (f32 oracle vs f64 {eo:.2e}); "
          f"f32 rounding spread {spread:.2e}", flush=True)
    assert np.all(np.isfinite(x))
    assert e32 <= max(POS_TOL, 2.0 * spread)
    if case["params"].get("cheb_enable", True):
        log = case["log32"]
        floor = rhat_floor(log, variants, VARIANT_AGREE)
        gated = 0 if floor is None else int((log[:, 3] > floor).sum())
        if gated:
            check_residual_curve(solver, _Log(log, variants))
        if 2 * gated < len(log):
            check_curve_vs_fp64(solver, case)
    if case["params"].get("monitor_convergence", False):
        hist = solver.convergence_history
        log = case["log32"]
        assert len(hist) == len(log)
        h = np.array([[r["hydro_error"], r["dev_error"]] for r in hist])
        for k, col in enumerate((6, 7)):
            ref = log[:, col]
            floor = rhat_floor(log, variants, VARIANT_AGREE, col=col)
            above = ref > (np.inf if floor is None else floor)
            above &= ref > 1e-4 * ref.max()
            worst = float((np.abs(h[:, k] - ref)[above] / ref[above]).max()) if above.any() else 0.0
            print(f"  monitor column {col}: f32 floor {floor}, worst rel {worst:.3e} over "
                  f"{int(above.sum())} of {len(ref)} iterations", flush=True)
            assert worst <= RES_TOL
            if 2 * above.sum() < len(ref):
                check_curve_vs_fp64(solver, case, col=col, values=h[:, k])
