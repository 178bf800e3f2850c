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


def fan_mesh(n_tets=600, pool=30, seed=0):
    """n_tets tets that all share vertex 0 (at the origin), their other three
    vertices drawn from `pool` points on the unit sphere; each tet is
    positively oriented and not flat.  Vertex 0 has more corners than a patch
    may rank in a byte (255), so every patch holding its tets is split."""
    import itertools
    from xpbd_gpu.mesh import TetMesh
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(pool, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    verts = np.vstack([np.zeros((1, 3)), p]).astype(np.float32)
    combos = list(itertools.combinations(range(1, pool + 1), 3))
    order = rng.permutation(len(combos))
    tets = []
    for i in order:
        a, b, c = combos[i]
        vol = np.dot(np.cross(verts[b] - verts[0], verts[c] - verts[0]), verts[a] - verts[0])
        if abs(vol) < 0.05:
            continue
        tets.append((0, a, b, c) if vol > 0 else (0, a, c, b))
        if len(tets) == n_tets:
            break
    return TetMesh(verts, np.array(tets, dtype=np.int32))
