"""The batch re-target that allocates nothing (xpbd_batch_set_members_noalloc),
through the C ABI: the same scenes re-targeted by it and by
xpbd_batch_set_members step bit for bit alike (deterministic accumulation),
on two calls in a row, the second with a smaller membership; oversized
memberships, graph-replaying batches and a batch whose array another batch
took over are refused."""
import ctypes

import numpy as np
import pytest

from scenes import batched_scenes, make_solver, resting_cube
from xpbd_gpu import _abi
from xpbd_gpu.batch import SceneBatch

pytestmark = pytest.mark.gpu


def _handles(solvers):
    return (ctypes.c_void_p * len(solvers))(*[s._h for s in solvers])


def _noalloc(batch, solvers):
    return _abi.load().xpbd_batch_set_members_noalloc(batch._h, _handles(solvers), len(solvers))


@pytest.mark.parametrize("persistent", ["auto", "never"])
def test_noalloc_retarget_matches_allocating_retarget(persistent):
    """With "auto" these small members take the persistent substep, whose
    workgroup list then lives in the array too; with "never" the phased
    kernels."""
    print(f"[case] noalloc_vs_set_members {persistent}", flush=True)
    scenes = batched_scenes(n_scenes=6, seed=21, n_range=(3, 10), iters=15)
    a = [make_solver(s) for s in scenes]
    b = [make_solver(s) for s in scenes]
    rounds = [[0, 1, 2, 3, 4, 5], [4, 1]]  # the second call with a smaller membership
    opts = dict(graph=False, persistent=persistent)
    with SceneBatch([a[0]], **opts) as ba, SceneBatch([b[0]], **opts) as bb:
        for members in rounds:
            ba.set_members([a[k] for k in members])
            _abi.check(_noalloc(bb, [b[k] for k in members]))
            assert bb.path == ("persistent" if persistent == "auto" else "kernels")
            ba.substep(3)
            _abi.check(_abi.load().xpbd_batch_substep(bb._h))
            _abi.check(_abi.load().xpbd_batch_substep(bb._h))
            _abi.check(_abi.load().xpbd_batch_substep(bb._h))
            ba.synchronize()
            _abi.check(_abi.load().xpbd_batch_synchronize(bb._h))
            for k in members:
                xa = a[k].mesh.verts.x.to_numpy()
                xb = b[k].mesh.verts.x.to_numpy()
                assert np.array_equal(xa, xb), (members, k)
                assert np.array_equal(a[k].mesh.verts.v.to_numpy(), b[k].mesh.verts.v.to_numpy())
    for k in range(6):  # scenes outside the last membership were not touched by it
        assert np.array_equal(a[k].mesh.verts.x.to_numpy(), b[k].mesh.verts.x.to_numpy())


def test_noalloc_refusals_and_takeover():
    print("[case] noalloc_refusals", flush=True)
    lib = _abi.load()
    tiny = [make_solver(resting_cube(n=1, frames=1)) for _ in range(_abi.NOALLOC_MAX_SCENES + 1)]
    with SceneBatch(tiny[:1], graph=False) as b1, SceneBatch(tiny[1:2], graph=False) as b2:
        # more members than the array's scene table holds: refused, batch unchanged
        assert _noalloc(b1, tiny) == _abi.XPBD_EINVAL
        assert "at most" in _abi.last_error()
        _abi.check(lib.xpbd_batch_substep(b1._h))
        # b1 holds the array until b2 takes it over
        _abi.check(_noalloc(b1, tiny[:3]))
        _abi.check(lib.xpbd_batch_substep(b1._h))
        _abi.check(_noalloc(b2, tiny[3:5]))
        assert lib.xpbd_batch_substep(b1._h) == _abi.XPBD_ESTATE
        assert "re-target" in _abi.last_error()
        _abi.check(lib.xpbd_batch_substep(b2._h))
        b1.set_members(tiny[:3])  # back to its own buffers
        _abi.check(lib.xpbd_batch_substep(b1._h))
        _abi.check(lib.xpbd_batch_synchronize(b1._h))
        _abi.check(lib.xpbd_batch_synchronize(b2._h))
    with SceneBatch(tiny[:1]) as bg:  # replays graphs: refused
        assert _noalloc(bg, tiny[:2]) == _abi.XPBD_EINVAL
        assert "graph" in _abi.last_error()
