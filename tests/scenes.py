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


# The rounding variants of the f32 oracle: (library flavour, variant flags, seed).
#  1. every scatter and sum in a seeded random order (oracle.VAR_ORDER): the
#     reference's float atomics fix no order;
#  2. the same, with 1/x and sqrt also moved by a seeded -1/0/+1 ulp
#     (VAR_ULP), which perturbs even a mesh too small for order to matter;
#  3. the same, in the oracle built with floating-point contraction (fused
#     multiply-add) and with the compliance reassociated as (6 / (lame dt^2))
#     / det Dm (VAR_FOLD): the reference's driver initialises Taichi with its
#     defaults (ti.init(arch=...), run_neohookean_usd.py:81), and Taichi's
#     default fast-math mode permits contraction, reassociation and
#     approximate 1/x and sqrt.
ROUNDING_VARIANTS = (("plain", 1, 1), ("plain", 3, 2), ("fma", 7, 3))


def rounding_variants(scene: Scene, frames=None, substeps=None, threads=None, setup=None,
                      kinds=ROUNDING_VARIANTS, repeats=1, **over):
    """The f32 oracle on the same scene and schedule once per rounding variant
    (ROUNDING_VARIANTS), each kind `repeats` times with other seeds; their
    spread around the default run is the f32 noise of the case itself (more
    repeats sample it better where it is wide).  `setup(o)` prepares each
    oracle like the caller's own.  Returns a list of (positions, log, flavour)."""
    out = []
    kinds = [(f, v, seed + 100 * r) for r in range(repeats) for f, v, seed in kinds]
    for flavor, var, seed in kinds:
        o = make_oracle(scene, threads=threads, variant=var, variant_seed=seed, flavor=flavor,
                        **over)
        if setup is not None:
            setup(o)
        o.set_log(True)
        if substeps is not None:
            sub_dt = scene.params["dt"] / scene.params["substeps_num"]
            for _ in range(substeps):
                o.substep(sub_dt)
        else:
            o.frame(frames if frames is not None else scene.frames)
        out.append((o.x, o.log(), flavor))
    return out


def position_spread(x0, variants) -> float:
    """The f32 noise of a case's final positions: the largest distance between
    two of the runs that differ only in rounding (the default run x0 and the
    rounding variants)."""
    xs = [np.asarray(x0, np.float64)] + [np.asarray(v[0], np.float64) for v in variants]
    return max((float(np.abs(a - b).max()) for i, a in enumerate(xs) for b in xs[i + 1:]),
               default=0.0)


def rhat_floor(log0, variants, agree=0.005, col=3):
    """The f32 floor of a logged curve (r_hat, column 3 of the oracle log, by
    default; 6 and 7 are the monitor's mean |C_H| and |C_D|): the level above
    which every rounding variant agrees with the default run within `agree`
    (relative); iterations below it are round-off.  0.0 when they agree
    everywhere, None when they disagree even at the largest value."""
    r0 = log0[:, col]
    spread = np.zeros_like(r0)
    for v in variants:
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.where(r0 > 0, np.abs(v[1][:, col] - r0) / r0, np.inf)
        spread = np.maximum(spread, rel)
    order = np.argsort(-r0, kind="stable")
    worst = np.maximum.accumulate(spread[order])
    bad = np.nonzero(worst > agree)[0]
    if bad.size == 0:
        return 0.0
    if bad[0] == 0:
        return None
    return float(r0[order[bad[0]]])


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
