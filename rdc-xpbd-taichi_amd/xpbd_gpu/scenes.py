"""Synthetic scenes of the benchmark configurations (seeded box grids).

Each scene is a mesh plus the reference constructor arguments, the SDF list,
and the harness steps the reference's driver would run
(`run_neohookean_usd.py:81-106,146-161`).  Density is expressed, as the
reference allows, by `total_mass = density * volume` followed by
`compute_lumped_mass()` (xpbd_base.py:36,118).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import TetMesh, box_grid


@dataclass
class Scene:
    name: str
    mesh: TetMesh
    sdfs: list                     # dicts: kind plane/sphere
    params: dict                   # reference constructor keywords
    total_mass: float = 1.0
    gravity: tuple = (0.0, -9.8, 0.0)
    init_x: np.ndarray | None = None   # positions set after construction (from_numpy)
    frames: int = 1
    substeps_only: int | None = None   # run this many substeps instead of frames
    notes: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def diag(self) -> float:
        return self.mesh.bbox_diagonal()


def plane(y=0.0, fixed=True):
    return dict(kind="plane", point=(0.0, y, 0.0), dir=(0.0, 1.0, 0.0), fixed=fixed)


def sphere(c, r, fixed=True):
    return dict(kind="sphere", center=tuple(c), radius=r, fixed=fixed)


def cube(c, size, fixed=False):
    return dict(kind="cube", center=tuple(c), size=tuple(size), fixed=fixed)


def hang(points, radius=0.013):
    return dict(kind="hang", points=[list(map(float, p)) for p in points], radius=radius,
                fixed=True)


def base_params(**kw):
    p = dict(frame_dt=1.0 / 60.0, dt=1.0 / 60.0, block_neohookean_enable=True, cheb_enable=True,
             cheb_warmup=5, cheb_gamma=0.6666)
    p.update(kw)
    return p


def dropped_cube(n=16, drop=0.5, frames=10) -> Scene:
    """Unit cube, n^3 hexes x 6 tets, no jitter, dropped from `drop` onto the
    plane y=0; E=1e5, nu=0.45, density 1000; 5 substeps x 30 iterations."""
    m = box_grid(n, origin=(0.0, drop, 0.0))
    return Scene("dropped_cube", m, [plane(0.0)],
                 base_params(young_modulus=1e5, poisson_ratio=0.45, substeps_num=5,
                             constraints_iter=30),
                 total_mass=1000.0, frames=frames)


def resting_cube(n=16, frames=10) -> Scene:
    """The dropped cube resting on the plane from the start, so that gravity and
    contact deform it within the 10 frames (dropped from 0.5 it free-falls and never
    reaches the plane)."""
    s = dropped_cube(n=n, drop=0.0, frames=frames)
    s.name = "resting_cube"
    return s


def jittered_box(n=64, seed=0, frames=10, substeps_only=None, iters=100) -> Scene:
    """[0,1]^3 box, n^3 hexes x 6 tets, jitter +-0.2 cell (seed), E=1e6,
    nu=0.48, density 1000, resting on the plane y=0 with a fixed sphere r=0.3
    centred on the bottom face; 10 substeps x 100 iterations."""
    m = box_grid(n, jitter=0.2, seed=seed)
    return Scene("jittered_box", m, [plane(0.0), sphere((0.5, 0.0, 0.5), 0.3)],
                 base_params(young_modulus=1e6, poisson_ratio=0.48, substeps_num=10,
                             constraints_iter=iters),
                 total_mass=1000.0, frames=frames, substeps_only=substeps_only)


def twist_stretch(verts: np.ndarray, seed: int) -> np.ndarray:
    """x = twist_z(pi * z / L) (A X): A has principal stretches log-uniform in
    [0.3, 3] and a random rotation; the twist turns 180 degrees over the height."""
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(0.3), np.log(3.0), size=3))
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    A = q @ np.diag(s) @ q.T
    X = verts.astype(np.float64)
    c = X.mean(axis=0)
    Y = (X - c) @ A.T
    z = X[:, 2] - X[:, 2].min()
    L = max(X[:, 2].max() - X[:, 2].min(), 1e-12)
    th = math.pi * z / L
    cs, sn = np.cos(th), np.sin(th)
    out = Y.copy()
    out[:, 0] = cs * Y[:, 0] - sn * Y[:, 1]
    out[:, 1] = sn * Y[:, 0] + cs * Y[:, 1]
    return (out + c).astype(np.float32)


def twisted_grid(n=32, seed=0, iters=200) -> Scene:
    """Unjittered n^3 grid, initial x from a seeded affine stretch plus a
    180-degree twist about z, zero velocity, no gravity; one substep of 200
    iterations (frame_dt = dt, substeps 1).  E=1e6, nu=0.48; a plane far below
    keeps the reference's required SDF inactive."""
    m = box_grid(n)
    s = Scene("twisted_grid", m, [plane(-10.0)],
              base_params(young_modulus=1e6, poisson_ratio=0.48, substeps_num=1,
                          constraints_iter=iters),
              total_mass=1000.0, gravity=(0.0, 0.0, 0.0), frames=1)
    s.init_x = twist_stretch(m.verts, seed)
    return s


def batched_scenes(n_scenes: int = 256, seed: int = 0, n_range=(6, 40), cell: float = 1.0 / 32.0,
                   substeps: int = 5, iters: int = 50, drop: float = 0.05) -> list:
    """The batched configuration: n_scenes boxes whose resolution per axis is
    drawn uniformly from n_range (independently per axis), 6 tets per hex, cell
    size `cell`, each dropped from `drop` above the shared plane y=0; E
    log-uniform in [1e4, 1e7], nu uniform in [0.3, 0.49], density 1000;
    dt = 1/60, 5 substeps x 50 iterations."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_scenes):
        nx, ny, nz = (int(v) for v in rng.integers(n_range[0], n_range[1] + 1, size=3))
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        nu = float(rng.uniform(0.3, 0.49))
        size = (nx * cell, ny * cell, nz * cell)
        m = box_grid(nx, ny, nz, size=size, origin=(0.0, drop, 0.0))
        vol = size[0] * size[1] * size[2]
        out.append(Scene(f"batch{i}", m, [plane(0.0)],
                         base_params(young_modulus=E, poisson_ratio=nu, substeps_num=substeps,
                                     constraints_iter=iters),
                         total_mass=1000.0 * vol, frames=1,
                         extra=dict(dims=(nx, ny, nz))))
    return out


def make_solver(scene: Scene, **opts):
    from .neohookean_solver import XPBDNeoHookeanSolver
    from .sdf import CubeSdf, HangSdfModel, PlaneSdf, SphereSdf
    sd = []
    for s in scene.sdfs:
        if s["kind"] == "plane":
            sd.append(PlaneSdf(point=s["point"], dir=s["dir"], fixed=s["fixed"]))
        elif s["kind"] == "sphere":
            sd.append(SphereSdf(center=s["center"], radius=s["radius"], fixed=s["fixed"]))
        elif s["kind"] == "cube":
            sd.append(CubeSdf(center=s["center"], size=s["size"], fixed=s["fixed"]))
        else:
            h = HangSdfModel(np.asarray(s["points"], dtype=np.float32))
            h.sphere_radius = s.get("radius", 0.013)
            sd.append(h)
        if "vel" in s:
            sd[-1].vel[None] = s["vel"]
    p = dict(scene.params)
    solver = XPBDNeoHookeanSolver((scene.mesh.verts, scene.mesh.tets), sd, **p, **opts)
    solver.gravity[None] = scene.gravity
    solver.total_mass = scene.total_mass
    solver.compute_lumped_mass()
    if scene.init_x is not None:
        solver.mesh.verts.x.from_numpy(scene.init_x)
    return solver
