"""Oracle side of the shared scenes (test infrastructure)."""
from __future__ import annotations

from xpbd_gpu.scenes import *  # noqa: F401,F403
from xpbd_gpu.scenes import Scene


def make_oracle(scene: Scene, precision=32, threads=None, **over):
    """The oracle set up exactly like make_solver sets up the device solver."""
    from oracle import Oracle
    p = dict(scene.params)
    p.update(over)
    o = Oracle(scene.mesh.verts, scene.mesh.tets, scene.sdfs, precision=precision,
               threads=threads, gravity=scene.gravity, **p)
    o.compute_lumped_mass(scene.total_mass)
    if scene.init_x is not None:
        o.x = scene.init_x
    return o
