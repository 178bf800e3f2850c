"""Mesh sources: box grids (Kuhn 6-tet split) and the TetGen reader/writer."""
import os

import numpy as np
import pytest

from xpbd_gpu.mesh import box_grid, load_mesh, read_tetgen, signed_volumes, write_tetgen

ASSETS = "/root/reference/assets/tetmesh"


@pytest.mark.parametrize("n", [1, 2, 5, 16])
def test_box_grid_counts_and_orientation(n):
    m = box_grid(n)
    assert m.n_verts == (n + 1) ** 3
    assert m.n_tets == 6 * n ** 3
    vol = signed_volumes(m.verts, m.tets)
    np.testing.assert_allclose(vol, (1.0 / n) ** 3 / 6.0, rtol=1e-6)
    assert vol.sum() == pytest.approx(1.0, rel=1e-6)


def test_box_grid_is_conforming():
    """Every interior triangle is shared by exactly two tets; boundary faces
    cover the 6 box faces with 2 triangles per cell face."""
    n = 3
    m = box_grid(n)
    faces = {}
    for t in m.tets:
        for f in ([0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]):
            key = tuple(sorted(t[f]))
            faces[key] = faces.get(key, 0) + 1
    counts = np.array(list(faces.values()))
    assert counts.max() == 2
    assert (counts == 1).sum() == 6 * n * n * 2


def test_anisotropic_grid_and_jitter_seed():
    a = box_grid(3, 5, 7, jitter=0.2, seed=3)
    b = box_grid(3, 5, 7, jitter=0.2, seed=3)
    c = box_grid(3, 5, 7, jitter=0.2, seed=4)
    assert a.n_tets == 6 * 3 * 5 * 7
    assert np.array_equal(a.verts, b.verts) and not np.array_equal(a.verts, c.verts)
    assert (signed_volumes(a.verts, a.tets) > 0).all()


def test_jittered_box_mesh_has_no_inverted_tets():
    m = box_grid(64, jitter=0.2, seed=0)
    assert m.n_verts == 274625 and m.n_tets == 1572864
    vol = signed_volumes(m.verts, m.tets)
    assert vol.min() > 0
    ref = (1 / 64) ** 3 / 6
    assert vol.min() / ref == pytest.approx(0.0593, abs=0.002)


def test_tetgen_roundtrip(tmp_path):
    m = box_grid(3, jitter=0.1, seed=1)
    write_tetgen(str(tmp_path / "box"), m)
    r = read_tetgen(str(tmp_path / "box.node"))
    assert np.array_equal(r.tets, m.tets)
    np.testing.assert_allclose(r.verts, m.verts, rtol=1e-6, atol=1e-7)
    r2 = load_mesh(str(tmp_path / "box"))
    assert np.array_equal(r2.tets, m.tets)


@pytest.mark.skipif(not os.path.isdir(ASSETS), reason="reference assets not mounted")
def test_reads_reference_assets():
    m = read_tetgen(os.path.join(ASSETS, "cow.node"))
    assert (m.n_verts, m.n_tets) == (4784, 17360)
    assert (signed_volumes(m.verts, m.tets) > 0).all()


def test_patch_builder_invariants():
    """Patches hold <= 512 tets and <= 256 vertices, every tet's local ids map
    back to its vertices (checked inside the library), and slots cover each
    vertex once per patch that touches it."""
    from xpbd_gpu import _abi
    if not _abi.library_available():
        pytest.skip("libxpbd_hip.so not built")
    for m in (box_grid(1), box_grid(5, 3, 7), box_grid(16, jitter=0.2, seed=2)):
        for pos in (None, m.verts):
            st = _abi.patch_stats(m.tets, m.n_verts, pos)
            assert st["max_patch_tets"] <= 512 and st["max_patch_verts"] <= 256
            assert st["n_patches"] >= (m.n_tets + 511) // 512
            assert st["n_slots"] >= m.n_verts
    with pytest.raises(ValueError):
        _abi.patch_stats(np.array([[0, 1, 2, 9]]), 4)


def test_patch_builder_splits_at_255_corners():
    """A vertex in 600 tets: corner ranks are bytes, so no patch may hold more
    than 255 of its tets, greedy or bisected."""
    from scenes import fan_mesh
    from xpbd_gpu import _abi
    if not _abi.library_available():
        pytest.skip("libxpbd_hip.so not built")
    m = fan_mesh()
    assert (m.tets == 0).any(axis=1).all()
    for pos in (None, m.verts):
        st = _abi.patch_stats(m.tets, m.n_verts, pos)
        assert st["max_patch_tets"] <= 255 and st["n_patches"] >= 3


def test_bisection_patches_are_compact():
    """Coordinate bisection stages far fewer vertices per tet than greedy cuts
    of the given order."""
    from xpbd_gpu import _abi
    if not _abi.library_available():
        pytest.skip("libxpbd_hip.so not built")
    m = box_grid(24, jitter=0.2, seed=0)
    greedy = _abi.patch_stats(m.tets, m.n_verts)
    rcb = _abi.patch_stats(m.tets, m.n_verts, m.verts)
    assert rcb["n_slots"] < 0.8 * greedy["n_slots"]
