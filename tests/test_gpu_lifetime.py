"""Object lifetimes in the C ABI.

fvc_destroy releases the device memory of every object created on the
context (particle sets, FTLE slides, batches, servers) and detaches them: a
detached handle accepts only its *_destroy call, and every other call on it
returns FVC_E_INVALID.  Destroying a particle set first releases the batches
that advance it.  The calls below go through the C ABI directly, in the
orders a C caller may use."""
import ctypes

import numpy as np
import pytest

from fvcpy import fvc, host

pytestmark = pytest.mark.gpu

BOUNDS = [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0]
FDT = 0.02


@pytest.fixture(scope="module")
def pipe():
    t = host.pipe(10, 10, 30, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    slices = [host.tube_velocity(t, xyz, s * FDT, 0.01, 5, s) for s in range(2)]
    return t, xyz, conn, adj, slices


def _ctx(pipe, precision="f64"):
    t, xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, BOUNDS, precision, 2, coords=xyz, conn=conn, adj=adj)
    for s in range(2):
        ctx.load_slice(s, s * FDT, slices[s])
    return ctx


def _advect_once(ctx, recs):
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    P.advect(0.0, FDT, 0, 1, 0.0, FDT, 2e-3)
    out = P.get()
    P.close()
    return out


def _take_handle(obj):
    """The raw handle; the wrapper forgets it, so only this test calls the ABI on it."""
    h = obj.h
    obj.h = None
    return h


def test_destroy_context_before_its_objects(pipe):
    t = pipe[0]
    L = fvc.lib()
    ctx = _ctx(pipe)
    seeds = host.tube_seeds(t, 300, 0.5, 9.0, 3)
    recs = ctx.seed_list(seeds, 1)
    ref = _advect_once(ctx, recs)
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    F = fvc.FtleSlide(ctx, [-0.5, -0.5, 1.0], [0.5, 0.5, 8.0], [4, 4, 9], 1, 0.0, FDT)
    B = fvc.Batch(ctx, P, [(0, recs.shape[0], 0.0, 0.01, fvc.INT_RK4)], 0.0, FDT, 2, 2e-3)
    S = fvc.Server(ctx, 0.0, FDT, 2, 2e-3, pool_particles=2048, steps_per_launch=8)
    rid = S.submit_tracer(seeds[:50], 0.0, 0.01)
    hp, hf, hb, hs, hc = (_take_handle(o) for o in (P, F, B, S, ctx))
    print("destroying the context first", flush=True)
    L.fvc_destroy(hc)
    buf = np.zeros(recs.shape[0], host.POINT_DTYPE)
    assert L.fvc_particles_get(hp, 0, buf.ctypes.data_as(ctypes.c_void_p), recs.shape[0]) == fvc.E_INVALID
    assert L.fvc_particles_set(hp, 0, recs.ctypes.data_as(ctypes.c_void_p), recs.shape[0]) == fvc.E_INVALID
    assert L.fvc_ftle_reset(hf) == fvc.E_INVALID
    n = ctypes.c_int64()
    assert L.fvc_ftle_points(hf, ctypes.byref(n), None) == fvc.E_INVALID
    assert L.fvc_batch_reset(hb) == fvc.E_INVALID
    bp = fvc.BatchParams(256, 0, 0, 0)
    assert L.fvc_batch_run(hb, ctypes.byref(bp), None, None) == fvc.E_INVALID
    assert L.fvc_server_poll(hs, None, None) == fvc.E_INVALID
    info = fvc.RequestInfo()
    assert L.fvc_server_info(hs, rid, ctypes.byref(info)) == fvc.E_INVALID
    # the destroy calls free what is left of each object, in any order
    L.fvc_server_destroy(hs)
    L.fvc_particles_destroy(hp)
    L.fvc_batch_destroy(hb)
    L.fvc_ftle_destroy(hf)
    # the device is usable afterwards: a new context gives the same result
    ctx2 = _ctx(pipe)
    got = _advect_once(ctx2, recs)
    np.testing.assert_array_equal(got["X"], ref["X"])
    np.testing.assert_array_equal(got["ElementIndex"], ref["ElementIndex"])
    ctx2.close()


def test_destroy_particles_before_their_batch(pipe):
    t = pipe[0]
    L = fvc.lib()
    ctx = _ctx(pipe, "f32")
    recs = ctx.seed_list(host.tube_seeds(t, 200, 0.5, 9.0, 4), 1)
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    B = fvc.Batch(ctx, P, [(0, recs.shape[0], 0.0, 0.01, fvc.INT_RK4)], 0.0, FDT, 2, 2e-3)
    st, _ = B.run(256, 0, 0)
    assert st > 0
    hp, hb = _take_handle(P), _take_handle(B)
    L.fvc_particles_destroy(hp)  # releases the batch that advances it
    bp = fvc.BatchParams(256, 0, 0, 0)
    assert L.fvc_batch_run(hb, ctypes.byref(bp), None, None) == fvc.E_INVALID
    L.fvc_batch_destroy(hb)
    # the context and its other objects are unaffected
    got = _advect_once(ctx, recs)
    assert (got["LeftDomain"] == 0).sum() > 0
    ctx.close()
