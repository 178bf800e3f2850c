"""The C-ABI library loads and exports every symbol include/xpbd_abi.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from xpbd_gpu import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xpbd_abi.h")


def declared_symbols():
    text = open(HEADER, encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(xpbd_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_abi.EXPORTED)


@pytest.mark.skipif(not _abi.library_available(), reason="libxpbd_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not _abi.library_available(), reason="libxpbd_hip.so not built")
def test_struct_layout_matches():
    lib = _abi.load()  # raises on a layout mismatch
    d = _abi.XpbdDesc()
    assert lib.xpbd_desc_init(ctypes.byref(d)) == 0
    assert d.young_modulus == 200000.0 and d.poisson_ratio == 0.495
    assert d.substeps_num == 20 and d.constraints_iter == 5
    assert d.cheb_gamma == 0.6666 and d.cheb_warmup == 5 and d.cheb_enable == 1
    assert d.block_neohookean_enable == 0 and d.gravity[1] == -9.8 and d.damping == 5.0


@pytest.mark.skipif(not _abi.library_available(), reason="libxpbd_hip.so not built")
def test_errors_are_reported_without_gpu_work():
    lib = _abi.load()
    d = _abi.XpbdDesc()
    lib.xpbd_desc_init(ctypes.byref(d))
    h = ctypes.c_void_p()
    assert lib.xpbd_create(ctypes.byref(d), ctypes.byref(h)) == _abi.XPBD_EINVAL
    assert "mesh" in _abi.last_error()


@pytest.mark.skipif(not _abi.library_available(), reason="libxpbd_hip.so not built")
def test_unknown_order_is_rejected():
    import numpy as np
    from xpbd_gpu.neohookean_solver import XPBDNeoHookeanSolver
    from xpbd_gpu.sdf import PlaneSdf
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float32)
    t = np.array([[0, 1, 2, 3]], dtype=np.int32)
    with pytest.raises(ValueError, match="reorder"):
        XPBDNeoHookeanSolver((v, t), [PlaneSdf([0.0, -1.0, 0.0], [0.0, 1.0, 0.0])],
                             reorder="hilbert")
    lib = _abi.load()
    d = _abi.XpbdDesc()
    lib.xpbd_desc_init(ctypes.byref(d))
    assert d.reorder == _abi.ORDER_BISECT
    vf, tf = v.ravel(), t.ravel()
    d.rest_positions = vf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    d.tets = tf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    d.n_verts, d.n_tets = 4, 1
    d.reorder = 7
    h = ctypes.c_void_p()
    assert lib.xpbd_create(ctypes.byref(d), ctypes.byref(h)) == _abi.XPBD_EINVAL
    assert "order" in _abi.last_error()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_abi, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_abi, "_lib", None)
    with pytest.raises(ImportError):
        _abi.load()


@pytest.mark.skipif(not _abi.library_available(), reason="libxpbd_hip.so not built")
@pytest.mark.parametrize("bad", [4, -1])
def test_out_of_range_vertex_index_is_rejected_before_gpu_work(bad):
    """A tet that names a vertex outside [0, n_verts) is refused on the host,
    before any device call, instead of reaching the kernels' gathers."""
    import numpy as np
    lib = _abi.load()
    d = _abi.XpbdDesc()
    lib.xpbd_desc_init(ctypes.byref(d))
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float32)
    t = np.array([[0, 1, 2, 3], [0, 1, 2, bad]], dtype=np.int32)
    sdf = _abi.XpbdSdf()
    sdf.kind = _abi.SDF_PLANE
    sdf.dir[1] = 1.0
    d.n_verts, d.n_tets = 4, 2
    d.rest_positions = v.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    d.tets = t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    d.n_sdf = 1
    d.sdfs = ctypes.pointer(sdf)
    h = ctypes.c_void_p()
    assert lib.xpbd_create(ctypes.byref(d), ctypes.byref(h)) == _abi.XPBD_EINVAL
    assert "tet 1 references vertex" in _abi.last_error()
    assert not h.value


def test_batch_opts_defaults_and_tuning_candidates():
    from xpbd_gpu.batch import _opts_struct, default_candidates
    o = _abi.XpbdBatchOpts()
    assert _abi.load().xpbd_batch_opts_init(ctypes.byref(o)) == 0
    assert (o.graph, o.graph_substeps, o.vertex_block, o.grid_pct, o.xcd_split,
            o.persistent) == (1, 1, 256, 100, 1, 1)
    assert _opts_struct(dict(persistent="never")).persistent == 0
    assert _opts_struct(dict(persistent="always")).persistent == 2
    with pytest.raises(ValueError):
        _opts_struct(dict(persistent="sometimes"))
    cands = default_candidates(5)
    # the phased knobs are timed on the phased kernels; one candidate runs every
    # iteration of a substep in one persistent launch
    assert [c for c in cands if c.get("persistent") == "always"] == [dict(persistent="always")]
    assert all(c["persistent"] == "never" for c in cands if "grid_pct" in c)
    assert {c["graph_substeps"] for c in cands if "grid_pct" in c} == {1, 5}
