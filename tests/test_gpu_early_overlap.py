"""The early FTLE of an RK4 slide frame runs on the context's second stream,
beside what the host enqueues next, and reads the frame's two slots.  A
slice loaded into one of those slots right after the frame must wait for it.
Here the same slide runs over five frames on a context holding all six slices
and on a two-slot ring that reloads, before each frame, the slot the previous
frame read as its lower slice; the FTLE fields must agree bit for bit.  The
grid sits at the pipe's outlet, so points leave in every frame and their
stencils take the early FTLE."""
import numpy as np
import pytest

from fvcpy import fvc, host

pytestmark = pytest.mark.gpu

PIPE_BOUNDS = [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0]
FDT = 0.02
NF = 5


@pytest.fixture(scope="module")
def pipe():
    t = host.pipe(12, 12, 40, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    slices = [host.tube_velocity(t, xyz, s * FDT, 0.01, 77, s) for s in range(NF + 1)]
    return xyz, conn, adj, slices


def _run(pipe, precision, ring):
    xyz, conn, adj, slices = pipe
    nslots = 2 if ring else NF + 1
    ctx = fvc.Context(3, PIPE_BOUNDS, precision, nslots, coords=xyz, conn=conn, adj=adj)
    if not ring:
        for s in range(NF + 1):
            ctx.load_slice(s, s * FDT, slices[s])
    F = fvc.FtleSlide(ctx, [-0.9, -0.9, 9.8], [0.9, 0.9, 10.0], [19, 19, 21], 1, 0.0, NF * FDT)
    outside0 = F.records()["LeftDomain"] != 0  # grid corners beyond the wall
    for f in range(NF):
        if ring:
            if f == 0:
                ctx.load_slice(0, 0.0, slices[0])
            ctx.load_slice((f + 1) % 2, (f + 1) * FDT, slices[f + 1])
            lo, hi = f % 2, (f + 1) % 2
        else:
            lo, hi = f, f + 1
        F.frame(1, lo, hi, f * FDT, (f + 1) * FDT, 1e-3)
    wt, wot = F.output(NF * FDT)
    left = int(((F.records()["LeftDomain"] != 0) & ~outside0).sum())  # left during the frames
    F.close()
    ctx.close()
    return wt, wot, left


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_slot_reload_waits_for_the_early_ftle(pipe, precision):
    print("outlet FTLE, %s, resident slices" % precision, flush=True)
    wt_a, wot_a, left_a = _run(pipe, precision, ring=False)
    print("outlet FTLE, %s, two-slot ring" % precision, flush=True)
    wt_b, wot_b, left_b = _run(pipe, precision, ring=True)
    assert left_a == left_b and left_a > 1000, (left_a, left_b)
    np.testing.assert_array_equal(wt_a, wt_b)
    np.testing.assert_array_equal(wot_a, wot_b)


def test_frames_beyond_the_snapshot_cap_match_the_server(pipe):
    """A frame that triggers more than 2^18 points takes the live launch on
    the main stream instead of the snapshot.  A thin grid at the outlet loses
    most of its 500k inside points in the first frame; its FTLE must equal
    the same grid served as a request (the server's early FTLE always reads
    the live state)."""
    xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f32", NF + 1, coords=xyz, conn=conn, adj=adj)
    for s in range(NF + 1):
        ctx.load_slice(s, s * FDT, slices[s])
    mn, mx, res = [-0.99, -0.99, 9.99], [0.99, 0.99, 10.0], [128, 128, 40]
    ts, te = 0.0, 3 * FDT
    print("outlet FTLE slide, %d points" % (128 * 128 * 40), flush=True)
    F = fvc.FtleSlide(ctx, mn, mx, res, 1, ts, te)
    outside0 = F.records()["LeftDomain"] != 0
    F.frame(1, 0, 1, 0.0, FDT, 1e-3)
    left0 = int(((F.records()["LeftDomain"] != 0) & ~outside0).sum())
    assert left0 > 300_000, left0  # each of them triggers its own early FTLE
    for f in range(1, 3):
        F.frame(1, f, f + 1, f * FDT, (f + 1) * FDT, 1e-3)
    wt, wot = F.output(te - ts)
    F.close()
    print("the same grid as a served request", flush=True)
    S = fvc.Server(ctx, 0.0, FDT, NF + 1, 1e-3, pool_particles=128 * 128 * 40, steps_per_launch=16)
    rid = S.submit_ftle(mn, mx, res, ts, te)
    S.run_until_idle()
    assert S.info(rid)["state"] == fvc.RS_DONE
    swt, swot = S.take(rid)
    S.close()
    ctx.close()
    np.testing.assert_array_equal(wt, swt)
    np.testing.assert_array_equal(wot, swot)
