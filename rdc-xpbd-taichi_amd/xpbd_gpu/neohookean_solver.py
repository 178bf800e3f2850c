"""Residual-driven Chebyshev Jacobi block Neo-Hookean XPBD solver on MI355X.

Drop-in for `xpbd_gpu/neohookean_solver.py` of the reference: same module
functions (`calc_first_lame`, `calc_second_lame`, :7-12), same class name,
constructor arguments and defaults (:18-38), same methods (`solve` :562-587,
`solve_step` :453-560) and attributes read by the reference's callers
(`convergence_report.py`, `usd_export.py`).  The substep runs as hand-written
gfx950 kernels in libxpbd_hip.so; the per-iteration rho/omega control that the
reference evaluates in host Python (:492-546) runs on the device, so a frame
needs no host round trip.

Execution options with no reference counterpart are keyword-only:
  accumulate_mode  "csr" (deterministic fixed-order sums, default) or "atomic"
                   (float atomics, the reference's unordered accumulation)
  reorder          internal tet/vertex order: "bisect" (compact patches by
                   coordinate bisection; True, the default), "morton" or
                   "none" (the caller's order; False)
  device           HIP device ordinal
  exec_path        "auto" (default: "persistent" for scenes of at most 288
                   patches, "graph" otherwise), "graph" (hipGraph per substep),
                   "persistent" (one launch per substep, scenes of at most 512
                   patches) or "kernels"
  layout           "patch" (workgroup per LDS-staged patch, default) or "flat"
  log_residuals    keep r_hat / rho_hat / omega per iteration in
                   `residual_history` (one device sync per substep)
"""
from __future__ import annotations

import time

import numpy as np

from . import _abi
from .xpbd_base import XPBDSolverBase


def calc_first_lame(E, nu):
    """Lame's first parameter lambda (neohookean_solver.py:7-8)."""
    return 2.0 * nu / (1.0 - 2.0 * nu) * calc_second_lame(E, nu)


def calc_second_lame(E, nu):
    """Shear modulus mu (neohookean_solver.py:11-12)."""
    return E / (2.0 * (1.0 + nu))


_ACCUM = {"atomic": _abi.ACCUM_ATOMIC, "csr": _abi.ACCUM_CSR}
_PATHS = {"graph": _abi.PATH_GRAPH, "kernels": _abi.PATH_KERNELS,
          "persistent": _abi.PATH_PERSISTENT, "auto": _abi.PATH_AUTO}
_LAYOUTS = {"patch": _abi.LAYOUT_PATCH, "flat": _abi.LAYOUT_FLAT}
_ORDERS = {"bisect": _abi.ORDER_BISECT, "morton": _abi.ORDER_MORTON, "none": _abi.ORDER_NONE,
           True: _abi.ORDER_BISECT, False: _abi.ORDER_NONE}

# Python attributes that feed the device parameter block.
_PARAM_ATTRS = ("young_modulus", "poisson_ratio", "frame_dt", "dt", "substeps_num",
                "constraints_iter", "hydrostatic_relaxation", "deviatoric_relaxation",
                "cheb_enable", "cheb_warmup", "cheb_gamma", "block_neohookean_enable",
                "monitor_convergence", "damping", "total_mass")


class XPBDNeoHookeanSolver(XPBDSolverBase):

    def __init__(self,
                 rest_pose,
                 sdf,
                 young_modulus=200000.0,
                 poisson_ratio=0.495,
                 scale=1.0,
                 offset=(0.0, 0.0, 0.0),
                 frame_dt=1e-2,
                 dt=1e-2,
                 substeps_num=20,
                 constraints_iter=5,
                 reorder_all=False,
                 block_size=256,
                 hydrostatic_relaxation=1.2,
                 deviatoric_relaxation=1.2,
                 cheb_enable=True,
                 cheb_warmup=5,
                 cheb_gamma=0.6666,
                 block_neohookean_enable=False,
                 monitor_convergence=False,
                 *,
                 accumulate_mode="csr",
                 reorder=True,
                 device=0,
                 exec_path="auto",
                 layout="patch",
                 log_residuals=False):
        self.constraints_iter = constraints_iter
        self.hydrostatic_relaxation = hydrostatic_relaxation
        self.deviatoric_relaxation = deviatoric_relaxation
        self.young_modulus = young_modulus
        self.poisson_ratio = poisson_ratio
        self.cheb_enable = cheb_enable
        self.cheb_gamma = cheb_gamma
        self.cheb_warmup = cheb_warmup
        self.block_neohookean_enable = block_neohookean_enable
        self.monitor_convergence = monitor_convergence
        self.convergence_history = []
        self.current_substep_convergence = []
        self.residual_history = []
        self.log_residuals = log_residuals
        self.current_frame = 0
        self.current_substep = 0
        self.convergence_callback = None
        # rho estimator constants (neohookean_solver.py:65-68); rho_hat itself
        # is device state, read through the rho_hat property
        self.rho_min = 0.00001
        self.rho_max = 0.99999
        self.rho_beta = 0.2
        self.prev_r_hat = None
        self.solve_compute_time_total = 0.0
        self.solve_compute_calls = 0
        self.solve_compute_iterations = 0
        if accumulate_mode not in _ACCUM:
            raise ValueError(f"accumulate_mode must be one of {sorted(_ACCUM)}")
        if exec_path not in _PATHS:
            raise ValueError(f"exec_path must be one of {sorted(_PATHS)}")
        if layout not in _LAYOUTS:
            raise ValueError(f"layout must be one of {sorted(_LAYOUTS)}")
        self._layout = layout
        self._accumulate_mode = accumulate_mode
        if reorder not in _ORDERS:
            raise ValueError(f"reorder must be True, False or one of "
                             f"{sorted(k for k in _ORDERS if isinstance(k, str))}")
        self._reorder = _ORDERS[reorder]
        self._device = int(device)
        self._exec_path = exec_path
        self._scale = scale
        self._offset = tuple(float(o) for o in offset)
        self._pushed = None
        self.block_size = block_size
        super().__init__(rest_pose, sdf, scale, offset, frame_dt, dt, substeps_num,
                         reorder_all, block_size)

    # ---- Lame constants (neohookean_solver.py:43-44) ---------------------------
    @property
    def lame1(self):
        return calc_first_lame(self.young_modulus, self.poisson_ratio)

    @property
    def lame2(self):
        return calc_second_lame(self.young_modulus, self.poisson_ratio)

    @property
    def rho_hat(self) -> float:
        rho = ctypes_double()
        _abi.check(self._lib.xpbd_get_state(self._h, None, None, None, _ref(rho)))
        return rho.value

    @rho_hat.setter
    def rho_hat(self, value):
        if self._h:
            _abi.check(self._lib.xpbd_set_state(self._h, int(self.current_frame),
                                                float(self.time), float(value)))

    # ---- construction -----------------------------------------------------
    def _desc(self) -> _abi.XpbdDesc:
        d = _abi.XpbdDesc()
        _abi.check(self._lib.xpbd_desc_init(_ref(d)))
        d.young_modulus = float(self.young_modulus)
        d.poisson_ratio = float(self.poisson_ratio)
        d.scale = float(self._scale)
        for k in range(3):
            d.offset[k] = self._offset[k]
            d.gravity[k] = float(np.asarray(self.gravity[None], dtype=np.float64)[k])
        d.frame_dt = float(self.frame_dt)
        d.dt = float(self.dt)
        d.substeps_num = int(self.substeps_num)
        d.constraints_iter = int(self.constraints_iter)
        d.hydrostatic_relaxation = float(self.hydrostatic_relaxation)
        d.deviatoric_relaxation = float(self.deviatoric_relaxation)
        d.cheb_enable = int(bool(self.cheb_enable))
        d.cheb_warmup = int(self.cheb_warmup)
        d.cheb_gamma = float(self.cheb_gamma)
        d.block_neohookean_enable = int(bool(self.block_neohookean_enable))
        d.monitor_convergence = int(bool(self.monitor_convergence))
        d.damping = float(self.damping)
        d.total_mass = float(self.total_mass)
        d.accumulate_mode = _ACCUM[self._accumulate_mode]
        d.reorder = self._reorder
        d.block_size = int(self.block_size)
        d.device = self._device
        d.layout = _LAYOUTS[self._layout]
        return d

    def initialize_specific(self, scale, offset):
        """Uploads the mesh and runs initialize + compute_lumped_mass on the
        device side (neohookean_solver.py:79-107)."""
        d = self._desc()
        verts = np.ascontiguousarray(self._rest.verts, dtype=np.float32)
        tets = np.ascontiguousarray(self._rest.tets, dtype=np.int32)
        d.n_verts = verts.shape[0]
        d.n_tets = tets.shape[0]
        d.rest_positions = _abi.fptr(verts)
        d.tets = tets.ctypes.data_as(_abi.ctypes.POINTER(_abi.ctypes.c_int32))
        sdfs = self._sdf_array()
        d.n_sdf = len(self.sdfs)
        d.sdfs = sdfs
        h = _abi.ctypes.c_void_p()
        _abi.check(self._lib.xpbd_create(_ref(d), _ref(h)))
        self._h = h.value
        self._warn_degenerate()
        _abi.check(self._lib.xpbd_set_exec_path(self._h, _PATHS[self._exec_path]))
        self._pushed = self._param_key()
        self._logging = None

    @property
    def exec_path(self) -> str:
        """How a substep reaches the GPU: "auto", "persistent", "graph" or
        "kernels" (xpbd_set_exec_path); every path gives the same bits."""
        return self._exec_path

    @exec_path.setter
    def exec_path(self, path: str):
        if path not in _PATHS:
            raise ValueError(f"exec_path must be one of {sorted(_PATHS)}")
        _abi.check(self._lib.xpbd_set_exec_path(self._h, _PATHS[path]))
        self._exec_path = path

    def set_barrier_timeout(self, microseconds: int):
        """How long a scene barrier of the persistent substep waits before it
        gives up and the next synchronizing call raises (xpbd_set_barrier_timeout;
        default 2 s)."""
        _abi.check(self._lib.xpbd_set_barrier_timeout(self._h, int(microseconds)))

    def _param_key(self):
        g = tuple(float(x) for x in np.asarray(self.gravity[None], dtype=np.float64))
        sd = tuple(s.key() for s in self.sdfs)
        return tuple(getattr(self, a) for a in _PARAM_ATTRS) + (g, sd)

    def _sync_params(self):
        """Pushes changed Python attributes to the device parameter block."""
        key = self._param_key()
        if key != self._pushed:
            if key[-1] != self._pushed[-1]:
                arr = self._sdf_array()
                _abi.check(self._lib.xpbd_set_sdfs(self._h, arr, len(self.sdfs)))
            _abi.check(self._lib.xpbd_update_params(self._h, _ref(self._desc())))
            self._pushed = key
        want_log = bool(self.monitor_convergence or self.log_residuals)
        if want_log != self._logging:
            _abi.check(self._lib.xpbd_set_logging(self._h, int(want_log)))
            self._logging = want_log

    # ---- the hot path -------------------------------------------------------
    def _drain_log(self):
        n = _abi.ctypes.c_int64()
        _abi.check(self._lib.xpbd_log_size(self._h, _ref(n)))
        if n.value == 0:
            return np.empty((0, 8))
        rows = np.empty((n.value, 8), dtype=np.float64)
        _abi.check(self._lib.xpbd_get_log(self._h, _abi.dptr(rows), n.value))
        _abi.check(self._lib.xpbd_clear_log(self._h))
        return rows

    def _records(self, rows, elapsed_per_iter):
        recs = []
        for r in rows:
            recs.append({
                "frame": int(r[0]),
                "substep": int(r[1]),
                "iter": int(r[2]),
                "iter_compute_time_sec": elapsed_per_iter,
                "hydro_error": float(r[6]),
                "dev_error": float(r[7]),
            })
            self.residual_history.append({
                "frame": int(r[0]), "substep": int(r[1]), "iter": int(r[2]),
                "r_hat": float(r[3]), "rho_hat": float(r[4]), "omega": float(r[5]),
            })
        return recs

    def solve_step(self, dt):
        """One XPBD substep's constraint solve (neohookean_solver.py:453-560):
        lambda reset, SDF collision, then constraints_iter Jacobi iterations of
        block (or split) projection, averaging, residual, rho/omega update and
        Chebyshev blend."""
        self._sync_params()
        t0 = time.perf_counter()
        _abi.check(self._lib.xpbd_solve_step(self._h, float(dt)))
        _abi.check(self._lib.xpbd_synchronize(self._h))
        elapsed = time.perf_counter() - t0
        self.solve_compute_time_total += elapsed
        self.solve_compute_calls += 1
        self.solve_compute_iterations += self.constraints_iter
        if self._logging:
            rows = self._drain_log()
            rows[:, 0] = self.current_frame
            rows[:, 1] = self.current_substep
            recs = self._records(rows, elapsed / max(1, self.constraints_iter))
            if self.monitor_convergence:
                self.convergence_history.extend(recs)
                self.current_substep_convergence.extend(recs)

    def solve(self):
        """Advance one frame (neohookean_solver.py:562-587)."""
        self._sync_params()
        _abi.check(self._lib.xpbd_set_state(self._h, int(self.current_frame), float(self.time),
                                            float("nan")))
        t0 = time.perf_counter()
        _abi.check(self._lib.xpbd_frame(self._h, 1))
        _abi.check(self._lib.xpbd_synchronize(self._h))
        elapsed = time.perf_counter() - t0
        frame = _abi.ctypes.c_int32()
        substep = _abi.ctypes.c_int32()
        t = ctypes_double()
        _abi.check(self._lib.xpbd_get_state(self._h, _ref(frame), _ref(substep), _ref(t), None))
        n_sub = self._substeps_per_frame()
        self.solve_compute_time_total += elapsed
        self.solve_compute_calls += n_sub
        self.solve_compute_iterations += n_sub * self.constraints_iter
        rows = self._drain_log() if self._logging else np.empty((0, 8))
        per_iter = elapsed / max(1, n_sub * self.constraints_iter)
        recs = self._records(rows, per_iter)
        if self.monitor_convergence:
            self.convergence_history.extend(recs)
            # one callback per substep of every time chunk, in order, as the
            # reference calls it after each substep (neohookean_solver.py:576-578)
            runs = []
            for rec in recs:
                if rec["iter"] == 0 or not runs:
                    runs.append([])
                runs[-1].append(rec)
            for run in runs:
                sub = run[0]["substep"]
                self.current_substep = sub
                self.current_substep_convergence = run
                if self.convergence_callback:
                    self.convergence_callback(self.current_frame, sub, run)
        self.current_substep = substep.value
        self.time = t.value
        self.current_frame = frame.value

    def save_state(self) -> dict:
        """Host checkpoint of everything a later substep reads: x, v, rho_hat
        (persists across substeps, neohookean_solver.py:65), frame counter and
        time.  Multipliers and Chebyshev buffers are rebuilt every substep."""
        self.synchronize()
        self._refresh_counters()
        return {"x": self.mesh.verts.x.to_numpy(), "v": self.mesh.verts.v.to_numpy(),
                "rho_hat": self.rho_hat, "current_frame": self.current_frame, "time": self.time}

    def load_state(self, state: dict):
        self.mesh.verts.x.from_numpy(state["x"])
        self.mesh.verts.v.from_numpy(state["v"])
        self.current_frame = int(state["current_frame"])
        self.time = float(state["time"])
        _abi.check(self._lib.xpbd_set_state(self._h, self.current_frame, self.time,
                                            float(state["rho_hat"])))

    def _refresh_counters(self):
        """Reads current_frame / current_substep / time back from the device
        solver (after frames stepped by a batch)."""
        frame = _abi.ctypes.c_int32()
        substep = _abi.ctypes.c_int32()
        t = ctypes_double()
        _abi.check(self._lib.xpbd_get_state(self._h, _ref(frame), _ref(substep), _ref(t), None))
        self.current_frame = frame.value
        self.current_substep = substep.value
        self.time = t.value

    def _substeps_per_frame(self) -> int:
        left = self.frame_dt
        chunks = 0
        while left > 0.0:
            left -= min(self.dt, left)
            chunks += 1
        return chunks * self.substeps_num

    # ---- execution options ----------------------------------------------------
    def set_kernel_timing(self, enabled, replays: int = 1):
        """Times the tet projection kernel with HIP events on the solver's
        stream.  replays == 1 brackets every projection launch; replays > 1
        times that many back-to-back replays of each projection (one event
        pair per run, lambda restored afterwards, so the solve is unchanged)."""
        mode = 0 if not enabled else max(1, int(replays))
        _abi.check(self._lib.xpbd_set_kernel_timing(self._h, mode))

    def kernel_time(self):
        ms = ctypes_double()
        n = _abi.ctypes.c_int64()
        _abi.check(self._lib.xpbd_get_kernel_time(self._h, _ref(ms), _ref(n)))
        return ms.value, n.value

    def stream_handle(self) -> int:
        p = _abi.ctypes.c_void_p()
        _abi.check(self._lib.xpbd_get_stream(self._h, _ref(p)))
        return p.value or 0

    def positions_to_device(self, dev_ptr: int):
        """Writes x (caller order, float32 (V,3)) into device memory at dev_ptr."""
        _abi.check(self._lib.xpbd_get_positions_device(self._h, dev_ptr))

    def positions_from_device(self, dev_ptr: int):
        _abi.check(self._lib.xpbd_set_positions_device(self._h, dev_ptr))


def ctypes_double():
    return _abi.ctypes.c_double()


def _ref(x):
    return _abi.ctypes.byref(x)
