"""rustyfem_solve_device_noalloc against rustyfem_solve_device through the C ABI.

The allocation-free solve runs the same kernels on work vectors carved from a
fixed per-device scratch array, eagerly instead of graph-replayed, so its
result must equal the existing entry point's bit for bit.  Two calls in a
row share the scratch array: the second has a smaller input (fewer
right-hand sides, then a smaller problem), so stale contents of the first
call would show.  A call beyond the scratch limit must be rejected with an
error and leave the handle usable.
"""
import numpy as np
import pytest

from rustyfem_amd import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


def _both(rf, h, b):
    """(solve_device result, solve_device_noalloc result, noalloc return) for b (n x nrhs)."""
    from rustyfem_amd import hip
    nrhs = b.shape[1]
    flat = np.asfortranarray(b).reshape(-1, order="F")
    d0 = hip.DeviceArray.from_host(flat)
    d1 = hip.DeviceArray.from_host(flat)
    h.solve_device(d0.ptr, nrhs)
    met = h.solve_device_noalloc(d1.ptr, nrhs)
    x0 = d0.download().reshape(b.shape, order="F")
    x1 = d1.download().reshape(b.shape, order="F")
    d0.free()
    d1.free()
    return x0, x1, met


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_noalloc_matches_solve_device(rf, ora, precision):
    big = meshes.hex_block(8, nrhs=6)
    small = meshes.hex_block(5, nrhs=2)
    print(f"case noalloc precision={precision} n={big.n}/{small.n}", flush=True)
    hb, eb, _ = rf.Handle.factor(big, precision=precision)
    hs, es, _ = rf.Handle.factor(small, precision=precision)
    assert hb is not None and hs is not None, rf.last_error()
    xo = ora.Handle.factor(big)[0].solve(big.rhs)
    # first call: 6 right-hand sides; second: 2 of them on the same handle;
    # third: a smaller problem on another handle
    for h, p, b in ((hb, big, big.rhs), (hb, big, big.rhs[:, :2].copy()), (hs, small, small.rhs)):
        x0, x1, met = _both(rf, h, b)
        assert met
        np.testing.assert_array_equal(x1, x0)
        a = p.to_scipy()
        rel = np.linalg.norm(b - a @ x1, axis=0) / np.linalg.norm(b, axis=0)
        assert np.all(rel <= 1e-10)
        if p is big:
            ref = xo[:, :b.shape[1]]
            assert np.max(np.abs(x1 - ref)) / np.max(np.abs(ref)) <= 1e-9
    hb.free()
    hs.free()


def test_noalloc_rejects_oversized_call(rf):
    from rustyfem_amd import hip
    p = meshes.hex_block(4, nrhs=1)
    h, e, _ = rf.Handle.factor(p)
    assert h is not None, rf.last_error()
    b = np.tile(p.rhs, (1, 65))
    d = hip.DeviceArray.from_host(np.asfortranarray(b).reshape(-1, order="F"))
    with pytest.raises(RuntimeError, match="at most 64"):
        h.solve_device_noalloc(d.ptr, 65)
    d.free()
    # the handle still solves, with and without allocation
    x0, x1, met = _both(rf, h, p.rhs)
    assert met
    np.testing.assert_array_equal(x1, x0)
    h.free()
