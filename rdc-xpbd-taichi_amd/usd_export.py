"""USD export of a run: the reference's `usd_export.py` and `utils/renderer.py`
surface, written as USD text (.usda) without the `pxr` library.

Same helpers (`method_tag`, `young_tag`, `build_output_paths`,
`create_renderer`, `add_solver_mesh`, `add_sdfs`; usd_export.py:10-92) and the
same renderer classes and methods (`USDMeshRenderer`,
`OptimizedUSDMeshRenderer`: `add_dynamic_mesh`, `add_static_sphere`,
`add_static_cube`, `add_static_plane`, `render`, `save`; renderer.py:7-198).
The stage has the reference's layout: a root Xform (the default prim), up axis
Y, time codes 1..frames at `fps` per second, the deformable mesh with its
points sampled at time code frame + 1, triangle faces from `solver.indices`
and no subdivision, and the colliders as a Sphere, a scaled Cube and a
two-triangle plane Mesh.  `pxr` is not available here, so the crate (.usdc)
format the reference writes when `use_binary` is set is replaced by the text
format of the same stage: a `.usdc` path is written as `.usda`.
"""
from __future__ import annotations

import os
from pathlib import Path

import numpy as np


def method_tag(block_neohookean: bool, dynamic_cheby: bool) -> str:
    model = "block_neohookean" if block_neohookean else "split_neohookean"
    accel = "dynamic_cheby" if dynamic_cheby else "no_cheby"
    return f"{model}_{accel}"


def young_tag(young: float) -> str:
    return f"{young:.3e}".replace("+", "").replace(".", "p")


def build_output_paths(model_path, block_neohookean, dynamic_cheby, young, use_binary=True,
                       root=None):
    """(output_dir, run_output_dir, usd_output_path) as usd_export.py:20-34, under
    `root` (default: the current directory)."""
    name = Path(model_path).name
    prefix = (name.split(".")[0] if "." in name else Path(name).stem) or "model"
    output_dir = os.path.join(root or os.getcwd(), "output", prefix)
    os.makedirs(output_dir, exist_ok=True)
    run_name = f"xpbd_{method_tag(block_neohookean, dynamic_cheby)}_young_{young_tag(young)}"
    usd_path = os.path.join(output_dir, f"{run_name}.{'usdc' if use_binary else 'usda'}")
    run_dir = os.path.join(output_dir, run_name)
    os.makedirs(run_dir, exist_ok=True)
    return output_dir, run_dir, usd_path


def _f(x) -> str:
    return "%.9g" % float(x)


def _vec3(v) -> str:
    v = np.asarray(v, dtype=np.float64).reshape(3)
    return f"({_f(v[0])}, {_f(v[1])}, {_f(v[2])})"


def _points(p: np.ndarray) -> str:
    p = np.asarray(p, dtype=np.float32).reshape(-1, 3)
    return "[" + ", ".join("(%.9g, %.9g, %.9g)" % (a, b, c) for a, b, c in p.tolist()) + "]"


def _ints(a) -> str:
    return "[" + ", ".join(str(int(x)) for x in np.asarray(a).ravel()) + "]"


class USDMeshRenderer:
    """Stage writer with the reference renderer's interface (renderer.py:7-129)."""

    def __init__(self, filepath, totalframes, fps, use_binary=True, enable_streaming=False,
                 stream_chunk_size=100, root_prim_path="/root") -> None:
        filepath = str(filepath)
        if filepath.endswith(".usdc"):
            filepath = filepath[:-5] + ".usda"
        elif not filepath.endswith(".usda"):
            filepath += ".usda"
        self.filepath = filepath
        self.totalframes = int(totalframes)
        self.fps = fps
        self.frame = 0
        self.time = 0.0
        self.enable_streaming = enable_streaming
        self.stream_chunk_size = stream_chunk_size
        self.root_prim_path = (root_prim_path if root_prim_path.startswith("/")
                               else f"/{root_prim_path}")
        self._meshes = []   # (name, vertex field, faces, rest points, [(time code, points)])
        self._statics = []  # prim text blocks

    # ---- scene content -------------------------------------------------------
    def add_dynamic_mesh(self, vert, face, meshname="mesh"):
        faces = np.asarray(face.to_numpy() if hasattr(face, "to_numpy") else face,
                           dtype=np.int64).ravel()
        rest = np.asarray(vert.to_numpy() if hasattr(vert, "to_numpy") else vert,
                          dtype=np.float32).copy()
        self._meshes.append((meshname, vert, faces, rest, []))

    def add_static_sphere(self, center, radius, name="sdf_sphere"):
        self._statics.append(
            f'    def Sphere "{name}"\n    {{\n'
            f"        double radius = {_f(radius)}\n"
            f"        double3 xformOp:translate = {_vec3(np.asarray(center, np.float32))}\n"
            f'        uniform token[] xformOpOrder = ["xformOp:translate"]\n    }}\n')

    def add_static_cube(self, center, size, name="sdf_cube"):
        self._statics.append(
            f'    def Cube "{name}"\n    {{\n'
            f"        float3 xformOp:scale = {_vec3(np.asarray(size, np.float32))}\n"
            f"        double3 xformOp:translate = {_vec3(np.asarray(center, np.float32))}\n"
            f'        uniform token[] xformOpOrder = ["xformOp:scale", "xformOp:translate"]\n'
            f"    }}\n")

    def add_static_plane(self, point, normal, extent=2.0, name="sdf_plane"):
        # a square of side `extent` around `point` spanned by two tangents
        n = np.asarray(normal, dtype=np.float32)
        nn = np.linalg.norm(n)
        n = np.array([0.0, 1.0, 0.0], np.float32) if nn < 1e-8 else n / nn
        ref = np.array([1.0, 0.0, 0.0], np.float32)
        if abs(float(np.dot(ref, n))) > 0.9:
            ref = np.array([0.0, 0.0, 1.0], np.float32)
        u = np.cross(n, ref)
        u = u / max(float(np.linalg.norm(u)), 1e-8)
        v = np.cross(n, u)
        p = np.asarray(point, dtype=np.float32)
        h = float(extent) * 0.5
        pts = np.array([p - h * u - h * v, p - h * u + h * v, p + h * u + h * v,
                        p + h * u - h * v], dtype=np.float32)
        self._statics.append(
            f'    def Mesh "{name}"\n    {{\n'
            f"        int[] faceVertexCounts = [3, 3]\n"
            f"        int[] faceVertexIndices = [0, 1, 2, 0, 2, 3]\n"
            f"        point3f[] points = {_points(pts)}\n"
            f'        uniform token subdivisionScheme = "none"\n    }}\n')

    # ---- time samples ----------------------------------------------------------
    def render(self):
        """Samples every dynamic mesh's points at time code frame + 1."""
        if self.frame >= self.totalframes:
            return
        for _, vert, _, _, samples in self._meshes:
            x = vert.to_numpy() if hasattr(vert, "to_numpy") else vert
            samples.append((self.frame + 1, np.asarray(x, dtype=np.float32).copy()))
        if self.enable_streaming and (self.frame + 1) % self.stream_chunk_size == 0:
            print(f"Streaming save at frame {self.frame + 1}/{self.totalframes}")
            self.save()
        self.frame += 1
        self.time = self.frame / self.fps

    def _text(self) -> str:
        root = self.root_prim_path.strip("/").split("/")
        out = ["#usda 1.0\n(\n", f'    defaultPrim = "{root[0]}"\n',
               f"    endTimeCode = {self.totalframes}\n", "    startTimeCode = 1\n",
               f"    timeCodesPerSecond = {_f(self.fps)}\n", '    upAxis = "Y"\n)\n\n']
        depth = 0
        for i, part in enumerate(root):
            pad = "    " * depth
            out.append(f'{pad}def Xform "{part}"\n{pad}{{\n')
            depth += 1
        pad = "    " * (depth - 1)
        for name, _, faces, rest, samples in self._meshes:
            out.append(f'{pad}    def Mesh "{name}"\n{pad}    {{\n')
            out.append(f"{pad}        int[] faceVertexCounts = {_ints(np.full(len(faces) // 3, 3))}\n")
            out.append(f"{pad}        int[] faceVertexIndices = {_ints(faces)}\n")
            out.append(f"{pad}        point3f[] points = {_points(rest)}\n")
            if samples:
                out.append(f"{pad}        point3f[] points.timeSamples = {{\n")
                for t, x in samples:
                    out.append(f"{pad}            {t}: {_points(x)},\n")
                out.append(f"{pad}        }}\n")
            out.append(f'{pad}        uniform token subdivisionScheme = "none"\n{pad}    }}\n')
        for block in self._statics:
            out.append("".join(pad + line + "\n" for line in block.rstrip("\n").split("\n")))
        for d in range(depth - 1, -1, -1):
            out.append("    " * d + "}\n")
        return "".join(out)

    def save(self):
        tmp = self.filepath + ".tmp"
        with open(tmp, "w", encoding="utf-8") as fh:
            fh.write(self._text())
        os.replace(tmp, self.filepath)

    def get_root_prim_path(self):
        return self.root_prim_path

    def get_file_size_mb(self):
        if os.path.exists(self.filepath):
            return os.path.getsize(self.filepath) / (1024 * 1024)
        return 0.0


class OptimizedUSDMeshRenderer(USDMeshRenderer):
    """The reference's large-file variant (renderer.py:132-198): streaming saves
    every 50 frames, tightened to 25 past 1 GB."""

    def __init__(self, filepath, totalframes, fps, max_memory_frames=500,
                 compression_level="high", root_prim_path="/root"):
        super().__init__(filepath, totalframes, fps, use_binary=True, enable_streaming=True,
                         stream_chunk_size=50, root_prim_path=root_prim_path)
        self.max_memory_frames = max_memory_frames
        self.compression_level = compression_level

    def render(self):
        if self.frame >= self.totalframes:
            return
        if self.frame % 100 == 0:
            size = self.get_file_size_mb()
            if size > 1000:
                self.stream_chunk_size = 25
            elif size > 500:
                self.stream_chunk_size = 50
        super().render()

    def save(self):
        super().save()
        print(f"USD file size: {self.get_file_size_mb():.1f}MB")


def create_renderer(args, usd_output_path, total_frames, fps):
    """usd_export.py:37-56: None with --no-usd, else the renderer the flags ask for."""
    if args.no_usd:
        print("USD export disabled (--no-usd)")
        return None
    if args.optimize_usd:
        print("Using OptimizedUSDMeshRenderer for large file handling")
        return OptimizedUSDMeshRenderer(usd_output_path, total_frames, fps,
                                        root_prim_path=args.usd_root_prim)
    return USDMeshRenderer(usd_output_path, total_frames, fps, use_binary=args.use_binary,
                           enable_streaming=args.streaming, root_prim_path=args.usd_root_prim)


def add_solver_mesh(usd_renderer, solver):
    if usd_renderer is not None:
        usd_renderer.add_dynamic_mesh(solver.mesh.verts.x, solver.indices, "deformable_mesh")


def add_sdfs(usd_renderer, sdf_config):
    """usd_export.py:64-92: each collider as a static prim (hang colliders have none)."""
    if usd_renderer is None:
        return
    from xpbd_gpu.sdf import CubeSdf, PlaneSdf, SphereSdf
    items = sdf_config if isinstance(sdf_config, (list, tuple)) else [sdf_config]
    for i, s in enumerate(items):
        sfx = f"_{i}" if isinstance(sdf_config, (list, tuple)) else ""
        if isinstance(s, SphereSdf):
            usd_renderer.add_static_sphere(center=s.center[None].to_numpy(),
                                           radius=float(s.radius[None]), name=f"sdf_sphere{sfx}")
        elif isinstance(s, CubeSdf):
            size = (np.asarray(s.half_size[None], np.float32) * 2.0).astype(np.float32)
            usd_renderer.add_static_cube(center=s.center[None].to_numpy(), size=size,
                                         name=f"sdf_cube{sfx}")
        elif isinstance(s, PlaneSdf):
            usd_renderer.add_static_plane(point=s.point[None].to_numpy(),
                                          normal=s.dir[None].to_numpy(), extent=4.0,
                                          name=f"sdf_plane{sfx}")
        else:
            print(f"SDF type {type(s).__name__} has no USD export hook yet.")


__all__ = ["method_tag", "young_tag", "build_output_paths", "create_renderer", "add_solver_mesh",
           "add_sdfs", "USDMeshRenderer", "OptimizedUSDMeshRenderer"]
