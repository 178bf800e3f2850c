"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a C++ restatement of the reference solver
(`xpbd_gpu/neohookean_solver.py`, `xpbd_gpu/xpbd_base.py`, `xpbd_gpu/sdf.py`)
in float32 (the reference's field type) or float64 (the FP64 error run), with a
fixed ascending-tet scatter order.  Only `tests/`, the smoke check and the
benchmark's CPU-baseline leg use it, as the checker; the product path never
imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_xpbd.so")
# The same source built with floating-point contraction (fused multiply-add)
# on: an arithmetic variant, used only to measure f32 rounding sensitivity.
_FMA_LIB_PATH = os.path.join(_HERE, "liboracle_xpbd_fma.so")
_libs = {}

# Run-time arithmetic variants (bit flags of the `variant` keyword; 0 is the
# restatement itself): VAR_ORDER takes every scatter and sum in a seeded
# random order, VAR_ULP moves 1/x and sqrt by a seeded -1/0/+1 ulp where a GPU
# uses its native instructions, VAR_FOLD forms the compliance as
# (6 / (lame dt^2)) / det Dm.
VAR_ORDER = 1
VAR_ULP = 2
VAR_FOLD = 4

# Order of the 22 doubles in the oracle's parameter block.
PARAM_KEYS = (
    "young_modulus", "poisson_ratio", "scale", "off_x", "off_y", "off_z",
    "frame_dt", "dt", "substeps_num", "constraints_iter",
    "hydrostatic_relaxation", "deviatoric_relaxation",
    "cheb_enable", "cheb_warmup", "cheb_gamma", "block_neohookean_enable",
    "monitor_convergence", "damping", "gx", "gy", "gz", "total_mass",
    "variant", "variant_seed",
)

# Constructor defaults of the reference (neohookean_solver.py:18-38,
# xpbd_base.py:27-36).
DEFAULTS = dict(
    young_modulus=200000.0, poisson_ratio=0.495, scale=1.0, offset=(0.0, 0.0, 0.0),
    frame_dt=1e-2, dt=1e-2, substeps_num=20, constraints_iter=5,
    hydrostatic_relaxation=1.2, deviatoric_relaxation=1.2,
    cheb_enable=True, cheb_warmup=5, cheb_gamma=0.6666,
    block_neohookean_enable=False, monitor_convergence=False,
    damping=5.0, gravity=(0.0, -9.8, 0.0), total_mass=1.0,
    variant=0, variant_seed=0,
)


def build(force: bool = False) -> str:
    src = os.path.getmtime(os.path.join(_HERE, "xpbd_oracle.cpp"))
    stale = any(not os.path.exists(p) or os.path.getmtime(p) < src
                for p in (_LIB_PATH, _FMA_LIB_PATH))
    if force or stale:
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def _load(flavor: str = "plain"):
    if flavor in _libs:
        return _libs[flavor]
    path = {"plain": _LIB_PATH, "fma": _FMA_LIB_PATH}[flavor]
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    vp, dp, ip, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), \
        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
    lib.oracle_create.restype = vp
    lib.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, ip, dp,
                                  ctypes.c_int, ip, ip, dp, dp, dp, ctypes.c_int]
    for name in ("oracle_destroy", "oracle_clear_log"):
        getattr(lib, name).argtypes = [vp]
    lib.oracle_frame.argtypes = [vp, ctypes.c_int]
    lib.oracle_substep.argtypes = [vp, ctypes.c_double]
    for name in ("oracle_get_x", "oracle_set_x", "oracle_get_v", "oracle_set_v",
                 "oracle_get_pred", "oracle_set_gravity", "oracle_get_log"):
        getattr(lib, name).argtypes = [vp, dp]
    lib.oracle_get_mass.argtypes = [vp, dp, dp]
    lib.oracle_get_tet_state.argtypes = [vp, dp, dp, dp, dp]
    lib.oracle_compute_lumped_mass.argtypes = [vp, ctypes.c_double]
    lib.oracle_rho_hat.argtypes = [vp]
    lib.oracle_rho_hat.restype = ctypes.c_double
    lib.oracle_set_rho_hat.argtypes = [vp, ctypes.c_double]
    lib.oracle_set_rho_hat.restype = None
    lib.oracle_set_sdf_points.argtypes = [vp, ctypes.c_int, dp, ctypes.c_int]
    lib.oracle_set_sdf_points.restype = None
    lib.oracle_sdf_eval.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    lib.oracle_sdf_eval.restype = ctypes.c_int
    lib.oracle_set_log.argtypes = [vp, ctypes.c_int]
    lib.oracle_log_size.argtypes = [vp]
    lib.oracle_log_size.restype = ctypes.c_int
    lib.oracle_block_tet64.argtypes = [dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
    lib.oracle_block_tet.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
    lib.oracle_split_tet64.argtypes = [dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp, dp]
    lib.oracle_split_tet64.restype = ctypes.c_int
    lib.oracle_residual_tet64.argtypes = [dp, dp, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, dp]
    lib.oracle_cheb_blend64.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double]
    lib.oracle_solve2x2_64.argtypes = [ctypes.c_double] * 7 + [dp]
    lib.oracle_solve2x2_64.restype = ctypes.c_int
    lib.oracle_inv3_64.argtypes = [dp, dp, dp]
    lib.oracle_set_pred.argtypes = [vp, dp]
    lib.oracle_set_lambda.argtypes = [vp, dp, dp]
    lib.oracle_residual_now.argtypes = [vp, ctypes.c_double, dp]
    lib.oracle_set_trace.argtypes = [vp, ctypes.c_int]
    lib.oracle_trace_size.argtypes = [vp]
    lib.oracle_trace_size.restype = ctypes.c_int
    lib.oracle_get_trace.argtypes = [vp, dp]
    _libs[flavor] = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def sdf_arrays(sdfs):
    """sdfs: list of dicts {kind: 'plane'|'sphere'|'cube'|'hang', point/dir |
    center/radius | center/size | points/radius, fixed}."""
    n = len(sdfs)
    kind = np.zeros(n, dtype=np.int32)
    fixed = np.zeros(n, dtype=np.int32)
    p = np.zeros((n, 4), dtype=np.float64)
    d = np.zeros((n, 3), dtype=np.float64)
    vel = np.zeros((n, 3), dtype=np.float64)
    for i, s in enumerate(sdfs):
        fixed[i] = int(bool(s.get("fixed", False)))
        vel[i] = s.get("vel", (0.0, 0.0, 0.0))
        if s["kind"] == "plane":
            kind[i] = 0
            p[i, :3] = s["point"]
            nrm = np.asarray(s["dir"], dtype=np.float64)
            d[i] = nrm / np.linalg.norm(nrm)  # sdf.py:66 normalizes at construction
        elif s["kind"] == "sphere":
            kind[i] = 1
            p[i, :3] = s["center"]
            p[i, 3] = s["radius"]
        elif s["kind"] == "cube":
            kind[i] = 2
            p[i, :3] = s["center"]
            d[i] = s["size"]
        elif s["kind"] == "hang":
            kind[i] = 3
            fixed[i] = 1
            p[i, 3] = s.get("radius", 0.013)
        else:
            raise ValueError(s["kind"])
    return kind, fixed, p, d, vel


class Oracle:
    """One reference-equivalent solver instance on the CPU."""

    def __init__(self, verts, tets, sdfs, precision: int = 32, threads: int | None = None,
                 flavor: str = "plain", **kw):
        lib = _load(flavor)
        prm = dict(DEFAULTS)
        prm.update(kw)
        off = prm.pop("offset")
        g = prm.pop("gravity")
        prm.update(off_x=off[0], off_y=off[1], off_z=off[2], gx=g[0], gy=g[1], gz=g[2])
        vals = np.array([float(prm[k]) for k in PARAM_KEYS], dtype=np.float64)
        self.params = prm
        verts = np.ascontiguousarray(verts, dtype=np.float32)
        tets = np.ascontiguousarray(tets, dtype=np.int32)
        self.n_verts, self.n_tets = verts.shape[0], tets.shape[0]
        kind, fixed, p, d, vel = sdf_arrays(sdfs)
        if threads is None:
            threads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
        self.threads = threads
        self._keep = (verts, tets, kind, fixed, p, d, vel)
        self._h = lib.oracle_create(
            precision, self.n_verts, self.n_tets,
            verts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            tets.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dptr(vals), len(sdfs),
            kind.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            fixed.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dptr(p), _dptr(d), _dptr(vel),
            threads)
        self._lib = lib
        for i, s in enumerate(sdfs):
            if s["kind"] == "hang":
                pts = np.ascontiguousarray(s["points"], dtype=np.float64).reshape(-1, 3)
                lib.oracle_set_sdf_points(self._h, i, _dptr(pts), pts.shape[0])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.oracle_destroy(h)
            self._h = None

    def frame(self, n: int = 1):
        self._lib.oracle_frame(self._h, n)

    def substep(self, sub_dt: float):
        self._lib.oracle_substep(self._h, float(sub_dt))

    def _get3(self, fn):
        out = np.empty((self.n_verts, 3), dtype=np.float64)
        fn(self._h, _dptr(out))
        return out

    @property
    def x(self):
        return self._get3(self._lib.oracle_get_x)

    @x.setter
    def x(self, val):
        a = np.ascontiguousarray(val, dtype=np.float64).reshape(self.n_verts, 3)
        self._lib.oracle_set_x(self._h, _dptr(a))

    @property
    def v(self):
        return self._get3(self._lib.oracle_get_v)

    @v.setter
    def v(self, val):
        a = np.ascontiguousarray(val, dtype=np.float64).reshape(self.n_verts, 3)
        self._lib.oracle_set_v(self._h, _dptr(a))

    @property
    def pred_x(self):
        return self._get3(self._lib.oracle_get_pred)

    def mass(self):
        m = np.empty(self.n_verts)
        im = np.empty(self.n_verts)
        self._lib.oracle_get_mass(self._h, _dptr(m), _dptr(im))
        return m, im

    def tet_state(self):
        v0 = np.empty(self.n_tets)
        idm = np.empty((self.n_tets, 3, 3))
        ld = np.empty(self.n_tets)
        lh = np.empty(self.n_tets)
        self._lib.oracle_get_tet_state(self._h, _dptr(v0), _dptr(idm), _dptr(ld), _dptr(lh))
        return v0, idm, ld, lh

    def set_gravity(self, g):
        a = np.asarray(g, dtype=np.float64)
        self._lib.oracle_set_gravity(self._h, _dptr(a))

    def compute_lumped_mass(self, total_mass: float):
        self._lib.oracle_compute_lumped_mass(self._h, float(total_mass))

    def sdf_eval(self, i: int, x, v=(0.0, 0.0, 0.0)):
        """(dist, normal, inside) of collider i at x for a vertex moving with v."""
        xa = np.ascontiguousarray(x, dtype=np.float64).reshape(3)
        va = np.ascontiguousarray(v, dtype=np.float64).reshape(3)
        out = np.zeros(4, dtype=np.float64)
        inside = self._lib.oracle_sdf_eval(self._h, int(i), _dptr(xa), _dptr(va), _dptr(out))
        return out[0], out[1:].copy(), bool(inside)

    @property
    def rho_hat(self) -> float:
        return self._lib.oracle_rho_hat(self._h)

    @rho_hat.setter
    def rho_hat(self, value: float):
        self._lib.oracle_set_rho_hat(self._h, float(value))

    def set_log(self, enabled: bool):
        self._lib.oracle_set_log(self._h, int(enabled))

    def log(self) -> np.ndarray:
        """Rows: frame, substep, iter, r_hat, rho_hat, omega, hydro_error, dev_error."""
        n = self._lib.oracle_log_size(self._h)
        out = np.empty((n, 8), dtype=np.float64)
        if n:
            self._lib.oracle_get_log(self._h, _dptr(out))
        return out

    def clear_log(self):
        self._lib.oracle_clear_log(self._h)

    @pred_x.setter
    def pred_x(self, val):
        a = np.ascontiguousarray(val, dtype=np.float64).reshape(self.n_verts, 3)
        self._lib.oracle_set_pred(self._h, _dptr(a))

    def set_lambda(self, lam_d, lam_h):
        ld = np.ascontiguousarray(lam_d, dtype=np.float64).reshape(self.n_tets)
        lh = np.ascontiguousarray(lam_h, dtype=np.float64).reshape(self.n_tets)
        self._lib.oracle_set_lambda(self._h, _dptr(ld), _dptr(lh))

    def residual_now(self, dt: float):
        """(sum r, sum |C_H|, sum |C_D|) over the tets at the current pred_x and
        multipliers (compute_xpbd_residual and the monitor kernels)."""
        out = np.zeros(3)
        self._lib.oracle_residual_now(self._h, float(dt), _dptr(out))
        return out

    def set_trace(self, on: bool = True):
        """Records the next substep's Chebyshev sequence (see trace())."""
        self._lib.oracle_set_trace(self._h, int(on))

    def trace(self):
        """List of per-iteration dicts: curr (the iterate the iteration started
        from), xj (after the Jacobi step), next (after the blend), omega."""
        n = self._lib.oracle_trace_size(self._h)
        buf = np.empty(n)
        if n:
            self._lib.oracle_get_trace(self._h, _dptr(buf))
        per = 9 * self.n_verts + 1
        out = []
        for k in range(n // per):
            r = buf[k * per:(k + 1) * per]
            v3 = 3 * self.n_verts
            out.append(dict(curr=r[:v3].reshape(-1, 3), xj=r[v3:2 * v3].reshape(-1, 3),
                            next=r[2 * v3:3 * v3].reshape(-1, 3), omega=float(r[-1])))
        return out


def block_tet(p, w, inv_dm, v0, lame1, lame2, dt, lam=(0.0, 0.0), precision: int = 64):
    """One block projection of an isolated tet (the solver's own per-tet
    function) in float64 or float32.

    Returns dict(corner (4,3), lam (2,), gD (4,3), gH (4,3), C_D, C_H, A (2,2),
    dlam (2,), ok).
    """
    lib = _load()
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(4, 3)
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(4)
    idm = np.ascontiguousarray(inv_dm, dtype=np.float64).reshape(3, 3)
    lam = np.array(lam, dtype=np.float64)
    corner = np.empty((4, 3))
    grads = np.empty(24)
    aux = np.empty(8)
    lib.oracle_block_tet(int(precision), _dptr(p), _dptr(w), _dptr(idm), float(v0),
                         float(lame1), float(lame2), float(dt), _dptr(lam), _dptr(corner),
                         _dptr(grads), _dptr(aux))
    return dict(corner=corner, lam=lam, gD=grads[:12].reshape(4, 3),
                gH=grads[12:].reshape(4, 3), C_D=aux[0], C_H=aux[1],
                A=np.array([[aux[2], aux[3]], [aux[3], aux[4]]]),
                dlam=aux[5:7].copy(), ok=int(aux[7]))


def split_tet(p, w, inv_dm, v0, lame, dt, relax, hydro, lam=0.0):
    """One split-mode constraint of an isolated tet in float64 (hydro: C_H =
    det F - 1 with lame = lambda; else C_D = |F|^2 - 3 with lame = mu).
    Returns dict(ok, lam, corner (4,3), C, w_sum)."""
    lib = _load()
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(4, 3)
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(4)
    idm = np.ascontiguousarray(inv_dm, dtype=np.float64).reshape(3, 3)
    lam_a = np.array([lam], dtype=np.float64)
    corner = np.zeros((4, 3))
    info = np.zeros(2)
    ok = lib.oracle_split_tet64(_dptr(p), _dptr(w), _dptr(idm), float(v0), float(lame),
                                float(dt), float(relax), int(bool(hydro)), _dptr(lam_a),
                                _dptr(corner), _dptr(info))
    return dict(ok=int(ok), lam=float(lam_a[0]), corner=corner, C=float(info[0]),
                w_sum=float(info[1]))


def residual_tet(p, inv_dm, v0, lame1, lame2, dt, lam_d, lam_h):
    """compute_xpbd_residual of one tet in float64: (r, C_D, C_H)."""
    lib = _load()
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(4, 3)
    idm = np.ascontiguousarray(inv_dm, dtype=np.float64).reshape(3, 3)
    out = np.zeros(3)
    lib.oracle_residual_tet64(_dptr(p), _dptr(idm), float(v0), float(lame1), float(lame2),
                              float(dt), float(lam_d), float(lam_h), _dptr(out))
    return float(out[0]), float(out[1]), float(out[2])


def solve2x2(a00, a01, a10, a11, b0, b1, eps=1e-12):
    lib = _load()
    x = np.zeros(2)
    ok = lib.oracle_solve2x2_64(a00, a01, a10, a11, b0, b1, eps, _dptr(x))
    return ok, x


def inv3(a):
    lib = _load()
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(3, 3)
    out = np.empty((3, 3))
    det = np.empty(1)
    lib.oracle_inv3_64(_dptr(a), _dptr(out), _dptr(det))
    return out, float(det[0])


def lame(E, nu):
    """calc_first_lame, calc_second_lame (neohookean_solver.py:7-12) in float64."""
    mu = E / (2.0 * (1.0 + nu))
    return 2.0 * nu / (1.0 - 2.0 * nu) * mu, mu
