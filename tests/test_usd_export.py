"""The USD text writer (usd_export.py): stage metadata, the dynamic mesh's time
samples, faces and the collider prims, read back from the written file.  The
`pxr` library the reference's renderer uses is absent, so the file is checked
by parsing it here (the text format is USD's own .usda syntax)."""
import os
import re
from types import SimpleNamespace

import numpy as np
import pytest

import usd_export


class _Field:
    """A vertex field whose positions move every frame, like mesh.verts.x."""

    def __init__(self, x):
        self.x = np.asarray(x, dtype=np.float32)

    def to_numpy(self):
        return self.x.copy()


def _block(text, header):
    i = text.index(header)
    depth, j = 0, text.index("{", i)
    for k in range(j, len(text)):
        depth += text[k] == "{"
        depth -= text[k] == "}"
        if depth == 0:
            return text[j:k + 1]
    raise AssertionError("unbalanced block")


def _tuples(s):
    return np.array([[float(v) for v in t.split(",")] for t in re.findall(r"\(([^()]*)\)", s)],
                    dtype=np.float32)


def _array_after(block, key):
    i = block.index(key) + len(key)
    j = block.index("[", i)
    return block[j:block.index("]", j) + 1]


def test_stage_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    x0 = rng.standard_normal((7, 3)).astype(np.float32)
    faces = np.array([0, 1, 2, 2, 3, 4, 4, 5, 6], dtype=np.uint32)
    f = _Field(x0)
    r = usd_export.USDMeshRenderer(str(tmp_path / "run.usdc"), 3, 60, root_prim_path="/SimAsset")
    assert r.filepath.endswith(".usda")
    r.add_dynamic_mesh(f, faces, "deformable_mesh")
    r.add_static_sphere(center=[0.0, -1.0, 0.0], radius=0.3, name="sdf_sphere_1")
    r.add_static_cube(center=[1.0, 2.0, 3.0], size=[0.5, 1.0, 2.0], name="sdf_cube_2")
    r.add_static_plane(point=[0.0, -2.0, 0.0], normal=[0.0, 2.0, 0.0], extent=4.0,
                       name="sdf_plane_0")
    frames = []
    for k in range(4):  # the fourth render is past totalframes and records nothing
        f.x = (x0 + 0.1 * k * rng.standard_normal((7, 3))).astype(np.float32)
        frames.append(f.x.copy())
        r.render()
    r.save()
    text = open(r.filepath, encoding="utf-8").read()
    assert text.startswith("#usda 1.0")
    for key in ('defaultPrim = "SimAsset"', "startTimeCode = 1", "endTimeCode = 3",
                "timeCodesPerSecond = 60", 'upAxis = "Y"'):
        assert key in text
    assert text.count("{") == text.count("}")
    mesh = _block(text, 'def Mesh "deformable_mesh"')
    assert np.array_equal(_tuples(_array_after(mesh, "point3f[] points =")), x0)
    counts = _array_after(mesh, "faceVertexCounts")
    assert [int(v) for v in counts.strip("[]").split(",")] == [3, 3, 3]
    idx = _array_after(mesh, "faceVertexIndices")
    assert [int(v) for v in idx.strip("[]").split(",")] == faces.tolist()
    samples = _block(mesh, "points.timeSamples")
    got = re.findall(r"(\d+): (\[[^\]]*\])", samples)
    assert [int(t) for t, _ in got] == [1, 2, 3]
    for (t, arr), ref in zip(got, frames):
        assert np.array_equal(_tuples(arr), ref)  # %.9g keeps every float32 exactly
    assert 'subdivisionScheme = "none"' in mesh
    sphere = _block(text, 'def Sphere "sdf_sphere_1"')
    assert "double radius = 0.3" in sphere and "xformOp:translate = (0, -1, 0)" in sphere
    cube = _block(text, 'def Cube "sdf_cube_2"')
    assert "float3 xformOp:scale = (0.5, 1, 2)" in cube
    assert '["xformOp:scale", "xformOp:translate"]' in cube
    plane = _block(text, 'def Mesh "sdf_plane_0"')
    p = _tuples(_array_after(plane, "point3f[] points ="))
    # a square of side 4 in the plane y = -2, centred on the given point
    assert np.allclose(p[:, 1], -2.0) and np.allclose(p.mean(axis=0), [0.0, -2.0, 0.0],
                                                      atol=1e-6)
    sides = np.linalg.norm(np.roll(p, -1, axis=0) - p, axis=1)
    assert np.allclose(sides, 4.0)
    assert np.isclose(abs(np.dot(p[1] - p[0], p[3] - p[0])), 0.0, atol=1e-5)


def test_streaming_saves_and_nested_root(tmp_path):
    f = _Field(np.zeros((3, 3)))
    r = usd_export.USDMeshRenderer(str(tmp_path / "s.usda"), 4, 30, enable_streaming=True,
                                   stream_chunk_size=2, root_prim_path="a/b")
    r.add_dynamic_mesh(f, np.array([0, 1, 2]), "m")
    r.render()
    assert not os.path.exists(r.filepath)
    r.render()
    assert os.path.exists(r.filepath)  # saved at frame 2
    text = open(r.filepath).read()
    assert 'defaultPrim = "a"' in text and 'def Xform "b"' in text
    assert text.count("{") == text.count("}")
    assert r.get_root_prim_path() == "/a/b"


def test_renderer_selection_follows_the_flags(tmp_path):
    base = dict(no_usd=False, optimize_usd=False, use_binary=True, streaming=False,
                usd_root_prim="/SimAsset")
    assert usd_export.create_renderer(SimpleNamespace(**dict(base, no_usd=True)), "x.usdc",
                                      10, 60) is None
    r = usd_export.create_renderer(SimpleNamespace(**base), str(tmp_path / "x.usdc"), 10, 60)
    assert type(r) is usd_export.USDMeshRenderer and r.filepath.endswith("x.usda")
    o = usd_export.create_renderer(SimpleNamespace(**dict(base, optimize_usd=True)),
                                   str(tmp_path / "y.usdc"), 10, 60)
    assert isinstance(o, usd_export.OptimizedUSDMeshRenderer) and o.stream_chunk_size == 50
    out, run, usd = usd_export.build_output_paths("assets/tetmesh/cow.node", True, True, 5e8,
                                                  use_binary=True, root=str(tmp_path))
    assert usd == os.path.join(out, "xpbd_block_neohookean_dynamic_cheby_young_5p000e08.usdc")


def test_pxr_reads_the_stage_when_available(tmp_path):
    pxr = pytest.importorskip("pxr")
    f = _Field(np.eye(3))
    r = usd_export.USDMeshRenderer(str(tmp_path / "p.usda"), 2, 24)
    r.add_dynamic_mesh(f, np.array([0, 1, 2]), "m")
    r.render()
    r.render()
    r.save()
    stage = pxr.Usd.Stage.Open(r.filepath)
    assert stage.GetEndTimeCode() == 2
