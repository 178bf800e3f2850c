"""Solver base with the reference's Python surface (xpbd_gpu/xpbd_base.py).

`XPBDSolverBase` keeps the reference's constructor arguments, attributes and
methods (`xpbd_base.py:11-204`): mesh loading, lumped masses, gravity
prediction, SDF collision, damped advance and the base frame loop.  All state
lives on the GPU inside libxpbd_hip.so; this class only forwards calls through
the C ABI and exposes the device buffers as field-like objects
(`mesh.verts.x.to_numpy()` / `.from_numpy()`, `indices.to_numpy()`).
"""
from __future__ import annotations

import ctypes
import warnings
import weakref

import numpy as np

from . import _abi
from .mesh import TetMesh, load_mesh
from .sdf import _Scalar, _vec, as_sdf_list


def calc_tet_volume(p0, p1, p2, p3):
    """Signed tet volume (xpbd_base.py:6-8), on host arrays."""
    p0, p1, p2, p3 = (np.asarray(p, dtype=np.float64) for p in (p0, p1, p2, p3))
    return np.einsum("...i,...i->...", np.cross(p1 - p0, p2 - p0), p3 - p0) / 6.0


class _SolverRef:
    """A weak reference to the owning solver: the mesh views must not keep
    the solver alive, so that a dropped solver frees its device state at once
    rather than whenever the cycle collector runs (a device free waits for the
    device)."""

    def __init__(self, solver):
        self._ref = weakref.ref(solver)

    @property
    def _solver(self):
        s = self._ref()
        if s is None:
            raise ReferenceError("the solver of this field has been freed")
        return s


def _weak_method(method):
    """Calls `method` while its object is alive (for change callbacks held by
    the object's own fields)."""
    ref = weakref.WeakMethod(method)

    def call(*args):
        m = ref()
        return None if m is None else m(*args)
    return call


class _VertexAttr(_SolverRef):
    """A per-vertex attribute of the device mesh, read and written as numpy;
    `getter(solver)` and `setter(solver, arr)` do the transfers."""

    def __init__(self, solver, getter, setter=None, width=3):
        super().__init__(solver)
        self._getter = getter
        self._setter = setter
        self._width = width

    @property
    def shape(self):
        return (self._solver.n_verts,)

    def to_numpy(self) -> np.ndarray:
        return self._getter(self._solver)

    def from_numpy(self, arr):
        if self._setter is None:
            raise AttributeError("this attribute is read-only")
        self._setter(self._solver, arr)


class _CellAttr(_SolverRef):
    def __init__(self, solver, getter):
        super().__init__(solver)
        self._getter = getter

    @property
    def shape(self):
        return (self._solver.n_tets,)

    def to_numpy(self) -> np.ndarray:
        return self._getter(self._solver)


class _Elements:
    def __init__(self, n, **attrs):
        self._n = n
        for k, v in attrs.items():
            setattr(self, k, v)

    def __len__(self):
        return self._n


class _Mesh:
    def __init__(self, rest: TetMesh, verts: _Elements, cells: _Elements):
        self._rest = rest
        self.verts = verts
        self.cells = cells

    def get_position_as_numpy(self) -> np.ndarray:
        return self._rest.verts.copy()


class _IndexField(_SolverRef):
    """The `indices` field (xpbd_base.py:57-58): 4 faces x 3 ids per tet, uint32."""

    def __init__(self, solver):
        super().__init__(solver)
        self._cache = None

    @property
    def shape(self):
        return (self._solver.n_tets * 12,)

    def to_numpy(self) -> np.ndarray:
        if self._cache is None:
            out = np.empty(self._solver.n_tets * 12, dtype=np.uint32)
            _abi.check(_abi.load().xpbd_get_visual_indices(
                self._solver._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
            self._cache = out
        return self._cache.copy()


class XPBDSolverBase:
    """Base class for deformable object solvers (xpbd_base.py:11-204)."""

    def __init__(self, rest_pose, sdf, scale=1.0, offset=(0.0, 0.0, 0.0), frame_dt=1e-2,
                 dt=1e-2, substeps_num=20, reorder_all=False, block_size=128):
        self._h = None
        self._lib = _abi.load()
        # Time stepping parameters
        self.time = 0.0
        self.frame_dt = frame_dt
        self.dt = dt
        self.substeps_num = substeps_num
        # Physics constants (xpbd_base.py:33-36)
        self.gravity = _Scalar(_vec((0.0, -9.8, 0.0)), on_change=_weak_method(self._push_gravity))
        self.damping = 5.0
        self.total_mass = 1.0
        self.block_size = block_size
        self.reorder_all = reorder_all
        # Mesh (replaces Patcher.load_mesh, xpbd_base.py:42)
        self._rest = load_mesh(rest_pose)
        self.n_verts = self._rest.n_verts
        self.n_tets = self._rest.n_tets
        # Collision (xpbd_base.py:61-70)
        self.sdfs = as_sdf_list(sdf)
        self.sdf = self.sdfs[0]
        self.initialize_specific(scale, offset)
        self.mesh = _Mesh(
            self._rest,
            _Elements(
                self.n_verts,
                x=_VertexAttr(self, lambda s: s._get_positions(),
                              lambda s, a: s._set_positions(a)),
                v=_VertexAttr(self, lambda s: s._get_velocities(),
                              lambda s, a: s._set_velocities(a)),
                pred_x=_VertexAttr(self, lambda s: s._get_predicted()),
                m=_VertexAttr(self, lambda s: s._get_masses()[0], width=1),
                inv_m=_VertexAttr(self, lambda s: s._get_masses()[1], width=1),
            ),
            _Elements(
                self.n_tets,
                lambda_D=_CellAttr(self, lambda s: s._get_multipliers()[0]),
                lambda_H=_CellAttr(self, lambda s: s._get_multipliers()[1]),
            ),
        )
        self.indices = _IndexField(self)

    # ---- handle plumbing ----------------------------------------------------
    def close(self):
        """Frees the scene's device state now; the solver is unusable after."""
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            self._lib.xpbd_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _push_gravity(self):
        if self._h:
            g = np.asarray(self.gravity[None], dtype=np.float64)
            _abi.check(self._lib.xpbd_set_gravity(self._h, _abi.dptr(g)))

    def _sdf_array(self):
        arr = (_abi.XpbdSdf * len(self.sdfs))()
        for i, s in enumerate(self.sdfs):
            arr[i] = s.to_abi()
        return arr

    def _get3(self, fn) -> np.ndarray:
        out = np.empty((self.n_verts, 3), dtype=np.float32)
        _abi.check(fn(self._h, _abi.fptr(out)))
        return out

    def _get_positions(self):
        return self._get3(self._lib.xpbd_get_positions)

    def _set_positions(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(self.n_verts, 3)
        _abi.check(self._lib.xpbd_set_positions(self._h, _abi.fptr(a)))

    def _get_velocities(self):
        return self._get3(self._lib.xpbd_get_velocities)

    def _set_velocities(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(self.n_verts, 3)
        _abi.check(self._lib.xpbd_set_velocities(self._h, _abi.fptr(a)))

    def _get_predicted(self):
        return self._get3(self._lib.xpbd_get_predicted)

    def _get_masses(self):
        m = np.empty(self.n_verts, dtype=np.float32)
        im = np.empty(self.n_verts, dtype=np.float32)
        _abi.check(self._lib.xpbd_get_masses(self._h, _abi.fptr(m), _abi.fptr(im)))
        return m, im

    def degenerate_tets(self):
        """(count, first): the tets whose volume was at most EPSILON when the
        masses were last computed, and the first of them (-1 when none)."""
        n = ctypes.c_int32()
        first = ctypes.c_int32()
        _abi.check(self._lib.xpbd_get_degenerate_tets(self._h, ctypes.byref(n),
                                                       ctypes.byref(first)))
        return n.value, first.value

    def _warn_degenerate(self):
        """The reference prints a warning per cell of volume <= EPSILON from its
        mass kernel and goes on (xpbd_base.py:104-106); one warning names the
        first such cell and their number."""
        n, first = self.degenerate_tets()
        if n:
            warnings.warn(f"Warning: Cell {first} has negative or zero volume "
                          f"({n} cell{'s' if n > 1 else ''} in total)", RuntimeWarning,
                          stacklevel=3)

    def _get_multipliers(self):
        ld = np.empty(self.n_tets, dtype=np.float32)
        lh = np.empty(self.n_tets, dtype=np.float32)
        _abi.check(self._lib.xpbd_get_multipliers(self._h, _abi.fptr(ld), _abi.fptr(lh)))
        return ld, lh

    # ---- reference methods ----------------------------------------------------
    def compute_lumped_mass(self):
        """Masses from tet volumes scaled to total_mass, then fixed-SDF pinning
        (xpbd_base.py:133-137)."""
        _abi.check(self._lib.xpbd_compute_lumped_mass(self._h, float(self.total_mass)))
        self._warn_degenerate()

    def apply_external_forces(self, dt):
        """v += g dt for free vertices; pred_x = x + v dt (xpbd_base.py:139-146)."""
        self._sync_params()
        _abi.check(self._lib.xpbd_apply_external_forces(self._h, float(dt)))

    def advance(self, dt, damping):
        """v = (pred_x - x)/dt * exp(-dt damping); x = pred_x (xpbd_base.py:148-160)."""
        self._sync_params()
        _abi.check(self._lib.xpbd_advance(self._h, float(dt), float(damping)))

    def solve_sdf_collision(self, dt):
        """Project pred_x out of each SDF in list order (xpbd_base.py:176-179)."""
        self._sync_params()
        _abi.check(self._lib.xpbd_solve_sdf_collision(self._h, float(dt)))

    def solve(self):
        """Advance one frame (xpbd_base.py:181-194)."""
        frame_time_left = self.frame_dt
        while frame_time_left > 0.0:
            dt0 = min(self.dt, frame_time_left)
            frame_time_left -= dt0
            sub_dt = dt0 / self.substeps_num
            for _ in range(self.substeps_num):
                self.apply_external_forces(sub_dt)
                self.solve_step(sub_dt)
                self.advance(sub_dt, self.damping)
            self.time += dt0

    def synchronize(self):
        """Waits for all queued device work of this solver."""
        _abi.check(self._lib.xpbd_synchronize(self._h))

    def check_finite(self):
        """Raises when a vertex position is NaN or infinite."""
        _abi.check(self._lib.xpbd_check_finite(self._h))

    # ---- to be provided by the derived solver --------------------------------
    def initialize_specific(self, scale, offset):
        raise NotImplementedError("Derived classes must implement initialize_specific")

    def solve_step(self, dt):
        raise NotImplementedError("Derived classes must implement solve_step")

    def _sync_params(self):
        pass
