"""fvc_locate_noalloc (no allocation; staging in the library's fixed device
array) against fvc_locate on the same inputs: identical element indices, on
two calls in a row (the second with fewer points, so stale results of the
first would show), in every walk mode, and a call above the size limit
refused."""
import ctypes

import numpy as np
import pytest

from fvcpy import fvc, host

pytestmark = pytest.mark.gpu

BOUNDS = [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0]


def _noalloc(ctx, xyz, guess, mode):
    L = fvc.lib()
    L.fvc_locate_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                     ctypes.c_int32, ctypes.c_void_p]
    xyz = np.ascontiguousarray(xyz, np.float64)
    out = np.full(xyz.shape[0], -7, np.int32)
    g = None if guess is None else np.ascontiguousarray(guess, np.int32)
    rc = L.fvc_locate_noalloc(ctx.h, xyz.ctypes.data_as(ctypes.c_void_p),
                              None if g is None else g.ctypes.data_as(ctypes.c_void_p), xyz.shape[0], mode,
                              out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_locate_noalloc_equals_locate(precision):
    t = host.pipe(12, 12, 40, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    ctx = fvc.Context(3, BOUNDS, precision, 2, coords=xyz, conn=conn, adj=adj)
    rng = np.random.default_rng(8)
    for n, seed in ((60000, 1), (7000, 2)):  # second call smaller than the first
        q = np.concatenate([host.tube_seeds(t, n - n // 10, 0.0, 10.0, seed),
                            rng.uniform(-1.5, 1.5, size=(n // 10, 3)) * [1, 1, 8]])
        guess = rng.integers(0, conn.shape[0], q.shape[0]).astype(np.int32)
        for mode in (fvc.LOCATE_WALK, fvc.LOCATE_GLOBAL, fvc.LOCATE_WALK_GLOBAL):
            print("locate_noalloc %s n=%d mode=%d" % (precision, q.shape[0], mode), flush=True)
            want = ctx.locate(q, None if mode == fvc.LOCATE_GLOBAL else guess, mode)
            rc, got = _noalloc(ctx, q, None if mode == fvc.LOCATE_GLOBAL else guess, mode)
            assert rc == fvc.OK
            np.testing.assert_array_equal(got, want)
        assert (want >= 0).sum() > n // 2
    big = np.zeros((fvc.NOALLOC_MAX_POINTS + 1, 3))
    rc, _ = _noalloc(ctx, big, None, fvc.LOCATE_GLOBAL)
    assert rc == fvc.E_INVALID
    ctx.close()
