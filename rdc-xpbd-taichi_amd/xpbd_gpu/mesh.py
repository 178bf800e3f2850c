"""Tetrahedral mesh sources for the solver: seeded box grids and TetGen files.

The reference loads meshes through the (un-vendored) MeshTaichi patcher
(`xpbd_gpu/xpbd_base.py:42`, `Patcher.load_mesh(rest_pose, relations=["VV", "CV"])`),
which reads TetGen `.node/.ele` files such as `assets/tetmesh/cow.node`.
This module replaces that loader with a plain reader/writer for the same text
format plus a generator of jittered box hex grids cut into 6 tets per hex
(Kuhn decomposition along the hex main diagonal).  All tets come out positively
oriented, which the reference needs (`xpbd_base.py:104-113` warns on
non-positive volume and then produces negative masses).
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from itertools import permutations

import numpy as np


@dataclass
class TetMesh:
    """Rest positions (V,3) float32 and tet-vertex table (T,4) int32."""

    verts: np.ndarray
    tets: np.ndarray

    @property
    def n_verts(self) -> int:
        return int(self.verts.shape[0])

    @property
    def n_tets(self) -> int:
        return int(self.tets.shape[0])

    def bbox_diagonal(self) -> float:
        v = self.verts.astype(np.float64)
        return float(np.linalg.norm(v.max(axis=0) - v.min(axis=0)))


def _perm_parity(p) -> int:
    p = list(p)
    parity = 0
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            if p[i] > p[j]:
                parity ^= 1
    return parity


def _kuhn_local_tets() -> np.ndarray:
    """The 6 Kuhn tets of a unit hex as corner ids (corner bit0=x, bit1=y, bit2=z).

    Each tet walks from corner 0 to corner 7 along the axes in the order of one
    permutation of (x, y, z).  Odd permutations give negatively oriented tets;
    their last two path corners are swapped so every tet has positive volume.
    """
    out = []
    for perm in permutations(range(3)):
        c0 = 0
        c1 = c0 | (1 << perm[0])
        c2 = c1 | (1 << perm[1])
        c3 = 7
        tet = [c0, c1, c2, c3]
        if _perm_parity(perm) == 1:
            tet = [c0, c2, c1, c3]
        out.append(tet)
    return np.asarray(out, dtype=np.int64)


_KUHN = _kuhn_local_tets()


def box_grid(nx: int, ny: int | None = None, nz: int | None = None, *,
             size=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
             jitter: float = 0.0, seed: int = 0) -> TetMesh:
    """Box of nx*ny*nz hexes, 6 tets per hex; vertex id is x-fastest.

    `jitter` is the half-width, in cells, of the uniform per-coordinate
    perturbation of every vertex (drawn with numpy.random.default_rng(seed)).
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    size = np.asarray(size, dtype=np.float64)
    origin = np.asarray(origin, dtype=np.float64)
    h = size / np.array([nx, ny, nz], dtype=np.float64)
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    k = np.arange(nz + 1)
    kk, jj, ii = np.meshgrid(k, j, i, indexing="ij")
    verts = np.stack([ii.ravel() * h[0], jj.ravel() * h[1], kk.ravel() * h[2]], axis=1)
    verts += origin
    if jitter:
        rng = np.random.default_rng(seed)
        verts += rng.uniform(-jitter, jitter, size=verts.shape) * h
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = (ci + cj * sx + ck * sy).transpose(2, 1, 0).ravel()  # hex order: x fastest
    corner_off = np.array([(c & 1) + ((c >> 1) & 1) * sx + ((c >> 2) & 1) * sy
                           for c in range(8)], dtype=np.int64)
    tets = base[:, None, None] + corner_off[_KUHN][None, :, :]
    tets = tets.reshape(-1, 4).astype(np.int32)
    return TetMesh(verts.astype(np.float32), tets)


def signed_volumes(verts: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """(x1-x0)x(x2-x0).(x3-x0)/6 per tet, in float64."""
    x = verts.astype(np.float64)
    a = x[tets[:, 1]] - x[tets[:, 0]]
    b = x[tets[:, 2]] - x[tets[:, 0]]
    c = x[tets[:, 3]] - x[tets[:, 0]]
    return np.einsum("ij,ij->i", np.cross(a, b), c) / 6.0


def _node_ele_paths(path: str):
    root, ext = os.path.splitext(path)
    if ext in (".node", ".ele"):
        return root + ".node", root + ".ele"
    return path + ".node", path + ".ele"


def read_tetgen(path: str) -> TetMesh:
    """Read a TetGen `.node`/`.ele` pair (0- or 1-based ids, extra columns ignored)."""
    node_path, ele_path = _node_ele_paths(path)

    def rows(p):
        with open(p, "r", encoding="utf-8") as f:
            for line in f:
                line = line.split("#", 1)[0].strip()
                if line:
                    yield line.split()

    it = rows(node_path)
    header = next(it)
    nv, dim = int(header[0]), int(header[1])
    if dim != 3:
        raise ValueError(f"{node_path}: expected 3D nodes, got dim={dim}")
    ids = np.empty(nv, dtype=np.int64)
    verts = np.empty((nv, 3), dtype=np.float64)
    for n in range(nv):
        r = next(it)
        ids[n] = int(r[0])
        verts[n] = [float(r[1]), float(r[2]), float(r[3])]
    first = int(ids.min()) if nv else 0

    it = rows(ele_path)
    header = next(it)
    nt, npe = int(header[0]), int(header[1])
    if npe < 4:
        raise ValueError(f"{ele_path}: expected 4 nodes per tet, got {npe}")
    tets = np.empty((nt, 4), dtype=np.int64)
    for n in range(nt):
        r = next(it)
        tets[n] = [int(r[1]), int(r[2]), int(r[3]), int(r[4])]
    tets -= first
    if nt and (tets.min() < 0 or tets.max() >= nv):
        raise ValueError(f"{ele_path}: vertex id out of range")
    return TetMesh(verts.astype(np.float32), tets.astype(np.int32))


def write_tetgen(prefix: str, mesh: TetMesh) -> tuple[str, str]:
    """Write `prefix.node` / `prefix.ele` in the layout of `assets/tetmesh/*` (0-based)."""
    node_path, ele_path = prefix + ".node", prefix + ".ele"
    with open(node_path, "w", encoding="utf-8") as f:
        f.write(f"{mesh.n_verts} 3 0 0\n")
        for n, (x, y, z) in enumerate(mesh.verts.astype(np.float64)):
            f.write(f"{n} {x:.9g} {y:.9g} {z:.9g}\n")
    with open(ele_path, "w", encoding="utf-8") as f:
        f.write(f"{mesh.n_tets} 4 0\n")
        for n, t in enumerate(mesh.tets):
            f.write(f"{n} {t[0]} {t[1]} {t[2]} {t[3]}\n")
    return node_path, ele_path


def load_mesh(rest_pose) -> TetMesh:
    """Accept a TetGen path, a TetMesh, or a (verts, tets) pair."""
    if isinstance(rest_pose, TetMesh):
        return TetMesh(np.ascontiguousarray(rest_pose.verts, dtype=np.float32),
                       np.ascontiguousarray(rest_pose.tets, dtype=np.int32))
    if isinstance(rest_pose, (tuple, list)) and len(rest_pose) == 2:
        v, t = rest_pose
        return TetMesh(np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3),
                       np.ascontiguousarray(t, dtype=np.int32).reshape(-1, 4))
    if isinstance(rest_pose, (str, os.PathLike)):
        return read_tetgen(os.fspath(rest_pose))
    raise TypeError(f"unsupported rest_pose type {type(rest_pose).__name__}")
