"""Loader of the golden fixtures in tests/golden (inputs + oracle outputs)."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        case = {k: z[k] for k in z.files}
    meta = json.loads(str(case.pop("meta")))
    case.update(meta)
    case["name"] = name
    if case.get("base"):  # mesh and starting state kept in another fixture
        base = load(case["base"])
        for k in ("verts", "tets", "x0", "v0", "rho0"):
            case[k] = base[k]
        return case
    if case["x0"].size == 0:
        case["x0"] = None
    if case["v0"].size == 0:
        case["v0"] = None
    return case


def scene(case):
    """The case as a Scene of xpbd_gpu.scenes (what make_solver takes)."""
    from xpbd_gpu.mesh import TetMesh
    from xpbd_gpu.scenes import Scene
    return Scene(case["name"], TetMesh(case["verts"], case["tets"]), case["sdfs"],
                 dict(case["params"]), total_mass=case["total_mass"],
                 gravity=tuple(case["gravity"]), init_x=case["x0"], frames=case["frames"])


def diag(x) -> float:
    x = np.asarray(x, np.float64)
    return float(np.linalg.norm(x.max(0) - x.min(0)))
