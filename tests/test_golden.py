"""The oracle against the committed golden fixtures (CPU).

The fixtures were written by tests/golden/make_golden.py from this oracle; the
oracle is deterministic for any thread count, so rerunning it must give the
stored positions and logs bit for bit.  This pins the oracle against silent
changes; the device solver is checked against the same fixtures in
test_gpu_golden.py.
"""
import os

import numpy as np
import pytest

import golden_cases
from oracle import Oracle


def run(case, precision):
    o = Oracle(case["verts"], case["tets"], case["sdfs"], precision=precision, threads=4,
               gravity=tuple(case["gravity"]), **case["params"])
    o.compute_lumped_mass(case["total_mass"])
    if case["x0"] is not None:
        o.x = case["x0"]
    if case["v0"] is not None:
        o.v = case["v0"]
    if case.get("rho0") is not None:
        o.rho_hat = case["rho0"]
    o.set_log(True)
    o.frame(case["frames"])
    return o.x, o.log()


def test_fixtures_present():
    assert set(golden_cases.names()) >= {"tet1_stretch", "tet2_fall", "box2_contact",
                                         "box4_split_monitor", "box4_split_cheb",
                                         "cow_demo_contact", "cow_compare_nh",
                                         "cow_compare_nh_dc", "cow_compare_block",
                                         "cow_compare_block_dc"}


@pytest.mark.parametrize("name", golden_cases.names())
@pytest.mark.parametrize("precision", [32, 64])
def test_oracle_reproduces_fixture(name, precision):
    case = golden_cases.load(name)
    x, log = run(case, precision)
    xs, ls = (case["x32"], case["log32"]) if precision == 32 else (case["x64"], case["log64"])
    if precision == 32:
        assert np.array_equal(x.astype(np.float32), xs)
    else:
        assert np.array_equal(x, xs)
    assert np.array_equal(log, ls)


def test_fixtures_move_and_converge():
    """Each case does something: positions change, residuals are positive and
    (with Chebyshev) omega leaves 1 after the warm-up."""
    for name in golden_cases.names():
        c = golden_cases.load(name)
        start = c["x0"] if c["x0"] is not None else c["verts"]
        assert np.abs(c["x32"] - start).max() > 0, name
        if c["params"].get("cheb_enable", True):
            assert c["log32"][:, 3].max() > 0, name
            assert (c["log32"][:, 5] > 1.0).any(), name


@pytest.mark.skipif(not os.path.isdir("/root/reference/assets/tetmesh"),
                    reason="reference checkout not present")
def test_cow_fixture_is_the_reference_mesh():
    from xpbd_gpu.mesh import read_tetgen
    m = read_tetgen("/root/reference/assets/tetmesh/cow.node")
    c = golden_cases.load("cow_demo_contact")
    assert np.array_equal(m.verts, c["verts"]) and np.array_equal(m.tets, c["tets"])
