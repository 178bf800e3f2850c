"""Signed-distance colliders with the reference's Python surface.

Mirrors `xpbd_gpu/sdf.py` of the reference: `SdfModel(fixed)` (:6-27),
`HangSdfModel(pos)` (:29-56), `PlaneSdf(point, dir, fixed)` (:59-74),
`SphereSdf(center, radius, fixed)` (:77-95) and `CubeSdf(center, size, fixed)`
(:98-124).  The distance/normal/check functions run inside the gfx950 kernels;
these objects only carry the parameters, exposed as `field[None]` like the
reference's Taichi scalar fields (`usd_export.py:76-87` reads them that way).
"""
from __future__ import annotations

import numpy as np

from . import _abi


class _Vec(np.ndarray):
    """A small float32 vector that also answers `.to_numpy()` like a ti.Vector."""

    def to_numpy(self):
        return np.array(self, dtype=np.float32)


def _vec(v) -> _Vec:
    return np.asarray(v, dtype=np.float32).reshape(3).copy().view(_Vec)


class _Scalar:
    """Stand-in for a 0-d Taichi field: `f[None]` reads, `f[None] = v` writes."""

    def __init__(self, value, on_change=None):
        self._value = value
        self._on_change = on_change

    def __getitem__(self, key):
        if key is not None and key != ():
            raise KeyError(key)
        return self._value

    def __setitem__(self, key, value):
        if key is not None and key != ():
            raise KeyError(key)
        if isinstance(self._value, np.ndarray):
            value = _vec(value)
        self._value = value
        if self._on_change is not None:
            self._on_change()


class SdfModel:
    """Base collider (sdf.py:6-27).  `vel[None]` is the obstacle's velocity
    (zero unless set): check() counts a vertex moving with v as inside when
    n . (vel - v) > 0 (sdf.py:20-24)."""

    kind = None

    def __init__(self, fixed):
        self.vel = _Scalar(_vec((0.0, 0.0, 0.0)))
        self.fixed = _Scalar(int(bool(fixed)))

    def _abi_struct(self) -> _abi.XpbdSdf:
        s = _abi.XpbdSdf()
        s.kind = self.kind
        s.fixed = int(bool(self.fixed[None]))
        v = np.asarray(self.vel[None], dtype=np.float64)
        for k in range(3):
            s.vel[k] = float(v[k])
        return s

    def to_abi(self) -> _abi.XpbdSdf:
        raise NotImplementedError

    def key(self):
        """What the device copy depends on (a change is pushed before the next step)."""
        a = self.to_abi()
        return (a.kind, a.fixed, tuple(a.p), tuple(a.dir), tuple(a.vel))


class PlaneSdf(SdfModel):
    """phi(p) = (p - point) . dir with dir normalized at construction (sdf.py:59-74)."""

    kind = _abi.SDF_PLANE

    def __init__(self, point, dir, fixed=False):
        super().__init__(fixed=fixed)
        d = np.asarray(dir, dtype=np.float64).reshape(3)
        n = np.linalg.norm(d)
        if n == 0:
            raise ValueError("PlaneSdf dir must be non-zero")
        self.point = _Scalar(_vec(point))
        self.dir = _Scalar(_vec(d / n))
        self._dir64 = d / n

    def to_abi(self) -> _abi.XpbdSdf:
        s = self._abi_struct()
        pt = np.asarray(self.point[None], dtype=np.float64)
        for k in range(3):
            s.p[k] = float(pt[k])
            s.dir[k] = float(np.asarray(self.dir[None], dtype=np.float64)[k])
        return s


class SphereSdf(SdfModel):
    """phi(p) = |p - center| - radius (sdf.py:77-95)."""

    kind = _abi.SDF_SPHERE

    def __init__(self, center, radius, fixed=False):
        super().__init__(fixed=fixed)
        self.center = _Scalar(_vec(center))
        self.radius = _Scalar(float(radius))

    def to_abi(self) -> _abi.XpbdSdf:
        s = self._abi_struct()
        c = np.asarray(self.center[None], dtype=np.float64)
        for k in range(3):
            s.p[k] = float(c[k])
        s.p[3] = float(self.radius[None])
        return s


class CubeSdf(SdfModel):
    """Axis-aligned box of edge lengths `size` around `center` (sdf.py:98-124):
    phi = |max(d, 0)| + min(max(d.x, d.y, d.z), 0) with d = |p - center| - size/2;
    the normal is the sign of p - center along its largest component."""

    kind = _abi.SDF_CUBE

    def __init__(self, center, size, fixed=False):
        super().__init__(fixed=fixed)
        self.center = _Scalar(_vec(center))
        self.half_size = _Scalar(_vec(np.asarray(size, dtype=np.float64).reshape(3) * 0.5))

    def to_abi(self) -> _abi.XpbdSdf:
        s = self._abi_struct()
        c = np.asarray(self.center[None], dtype=np.float64)
        h = np.asarray(self.half_size[None], dtype=np.float64)
        for k in range(3):
            s.p[k] = float(c[k])
            s.dir[k] = float(2.0 * h[k])  # the library takes the edge lengths
        return s


class _Points:
    """A (n, 3) float32 array read through `.to_numpy()` like a Taichi field."""

    def __init__(self, arr):
        self._a = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1, 3)

    @property
    def shape(self):
        return (self._a.shape[0],)

    def to_numpy(self):
        return self._a.copy()

    def from_numpy(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1, 3)
        if a.shape != self._a.shape:
            raise ValueError("the number of hang points is fixed")
        self._a = a


class HangSdfModel(SdfModel):
    """Union of spheres of radius 0.013 at `pos` (sdf.py:29-56), always fixed.
    As in the reference, the normal is taken towards sphere 0 once any sphere
    is nearer than 1e5 (sdf.py:45-53)."""

    kind = _abi.SDF_HANG

    def __init__(self, pos):
        super().__init__(fixed=True)
        pos = np.asarray(pos)
        if pos.ndim != 2 or pos.shape[1] != 3 or not (1 <= pos.shape[0] <= _abi.MAX_HANG_POINTS):
            raise ValueError(f"HangSdfModel needs 1..{_abi.MAX_HANG_POINTS} points of shape (n, 3)")
        self.sphere_pos = _Points(pos)
        self.sphere_radius = 0.013

    def to_abi(self) -> _abi.XpbdSdf:
        s = self._abi_struct()
        s.fixed = 1
        s.p[3] = float(self.sphere_radius)
        a = self.sphere_pos._a  # held by this object for the duration of the call
        s.points = a.ctypes.data_as(_abi.ctypes.POINTER(_abi.ctypes.c_float))
        s.n_points = a.shape[0]
        return s

    def key(self):
        return (self.kind, float(self.sphere_radius), self.sphere_pos._a.tobytes(),
                tuple(np.asarray(self.vel[None], dtype=np.float64)))


def as_sdf_list(sdf):
    """The reference's SDF argument handling (xpbd_base.py:61-67)."""
    if isinstance(sdf, (list, tuple)):
        sdfs = [s for s in sdf if s is not None]
    else:
        sdfs = [sdf] if sdf is not None else []
    if len(sdfs) == 0:
        raise ValueError("At least one SDF object must be provided.")
    for s in sdfs:
        if not isinstance(s, SdfModel):
            raise TypeError(f"unsupported SDF type {type(s).__name__}")
    return sdfs
