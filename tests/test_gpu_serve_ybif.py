"""Serving on the Y-bifurcation with 16-bit velocity slices held resident.

The served workload of the benchmark (parent vessel R=1 along -z, daughters
R=0.7 at +/-30 degrees, per-branch pulsatile Poiseuille profiles, RK4 at
dt=1e-3, tracer and FTLE requests) at a reduced mesh size: requests of mixed
kinds and sizes are submitted while others run, into a pool too small to hold
them all, and each served result must equal, bit for bit, the same request
run alone through the plain path (fvc_seed_list + fvc_batch_run, or
fvc_ftle_create + fvc_ftle_frame + fvc_ftle_output).  With FP64 slices a
served tracer request must also equal the oracle."""
import math
import os
import sys

import numpy as np
import pytest

import cases
from fvcpy import fvc, host

sys.path.insert(0, os.path.join(cases.REPO, "oracle"))
import oracle  # noqa: E402

pytestmark = pytest.mark.gpu

FDT = 0.01
NSLOTS = 6
DT = 1e-3


@pytest.fixture(scope="module")
def ybif():
    t = host.ybif(24, 9, 32)
    xyz, conn, adj = host.tube_mesh(t)
    slices = [host.tube_velocity(t, xyz, s * FDT, 0.01, 4321, s) for s in range(NSLOTS)]
    lo, hi = xyz.min(0), xyz.max(0)
    bounds = [float(lo[0]), float(hi[0]), float(lo[1]), float(hi[1]), float(lo[2]), float(hi[2])]
    return t, xyz, conn, adj, slices, bounds


def _ctx(y, precision):
    t, xyz, conn, adj, slices, bounds = y
    ctx = fvc.Context(3, bounds, precision, NSLOTS, coords=xyz, conn=conn, adj=adj)
    for s in range(NSLOTS):
        ctx.load_slice(s, s * FDT, slices[s])
    return ctx


def frame_of(ts):
    f = int(math.floor(ts / FDT))
    while f * FDT > ts:
        f -= 1
    while (f + 1) * FDT <= ts:
        f += 1
    return f


def make_requests(t, seed):
    """Tracer (RK4 or AB3) and FTLE requests, seed counts spread over two
    decades, start times and horizons drawn from a seed."""
    rng = np.random.default_rng(seed)
    lo, hi = np.array(t.lo[:3]), np.array(t.hi[:3])
    reqs = []
    for r in range(9):
        ts = float(rng.uniform(0.0, 0.02))
        te = ts + float(rng.uniform(0.005, 0.03))
        if r % 3 == 2:
            n = int(np.exp(rng.uniform(np.log(200), np.log(2000))))
            h = (np.prod(hi - lo) / n) ** (1.0 / 3.0)
            res = [max(2, int(round(e / h))) for e in hi - lo]
            reqs.append(("ftle", (lo.tolist(), hi.tolist(), res), ts, te, fvc.INT_RK4))
        else:
            n = int(np.exp(rng.uniform(np.log(30), np.log(2500))))
            xyz = host.tube_seeds(t, n, lo[2], hi[2], seed * 100 + r)
            reqs.append(("tracer", xyz, ts, te, fvc.INT_AB3 if r % 3 == 1 else fvc.INT_RK4))
    return reqs


def size_of(rq):
    return rq[1].shape[0] if rq[0] == "tracer" else int(np.prod(rq[1][2]))


def alone_plain(ctx, rq):
    kind, spec, ts, te, it = rq
    if kind == "tracer":
        recs = ctx.seed_list(spec, 1)
        P = fvc.Particles(ctx, recs.shape[0])
        P.set(recs)
        B = fvc.Batch(ctx, P, [(0, recs.shape[0], ts, te, it)], 0.0, FDT, NSLOTS, DT)
        B.run(256, 0, 0)
        out = P.get()
        B.close()
        P.close()
        return out
    mn, mx, res = spec
    F = fvc.FtleSlide(ctx, mn, mx, res, 1, ts, te)
    f = frame_of(ts)
    while F.frame(1, f % NSLOTS, (f + 1) % NSLOTS, f * FDT, f * FDT + FDT, DT) != 2:
        f += 1
    out = F.output(te - ts)
    F.close()
    return out


def submit(S, rq):
    kind, spec, ts, te, it = rq
    if kind == "tracer":
        return S.submit_tracer(spec, ts, te, it)
    mn, mx, res = spec
    return S.submit_ftle(mn, mx, res, ts, te)


def same(a, b, kind):
    if kind == "tracer":
        for k in ("X", "V", "ElementIndex", "LeftDomain", "LeftDomainTime"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    else:
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("precision", ["bf16", "f16"])
def test_served_equals_alone_on_the_bifurcation(ybif, precision):
    t = ybif[0]
    ctx = _ctx(ybif, precision)
    reqs = make_requests(t, 12)
    sizes = [size_of(rq) for rq in reqs]
    print("Y-bifurcation serving %s: %d requests, sizes %s" % (precision, len(reqs), sizes), flush=True)
    ref = [alone_plain(ctx, rq) for rq in reqs]
    pool = max(sizes) + 64
    assert pool < sum(sizes) // 2
    S = fvc.Server(ctx, 0.0, FDT, NSLOTS, DT, pool_particles=pool, steps_per_launch=16)
    # first half submitted together; the rest arrive while those run
    half = len(reqs) // 2
    ids = [submit(S, rq) for rq in reqs[:half]]
    saw_queue, peak = False, 0
    for _ in range(4):
        r, q = S.poll()
        saw_queue, peak = saw_queue or q > 0, max(peak, r)
    ids += [submit(S, rq) for rq in reqs[half:]]
    for _ in range(200000):
        r, q = S.poll()
        saw_queue, peak = saw_queue or q > 0, max(peak, r)
        if r == 0 and q == 0:
            break
    assert saw_queue and peak >= 2
    for rq, r, rid in zip(reqs, ref, ids):
        inf = S.info(rid)
        assert inf["state"] == fvc.RS_DONE
        got = S.take(rid)
        same(got, r, rq[0])
    S.close()
    ctx.close()


def test_served_fp64_tracer_equals_oracle(ybif):
    t, xyz, conn, adj, slices, bounds = ybif
    ctx = _ctx(ybif, "f64")
    om = oracle.Mesh(3, bounds, coords=xyz, conn=conn, adj=adj)
    lo, hi = t.lo[:3], t.hi[:3]
    seeds = host.tube_seeds(t, 800, lo[2], hi[2], 99)
    ts, te = 0.0043, 0.0337
    S = fvc.Server(ctx, 0.0, FDT, NSLOTS, DT, pool_particles=4096, steps_per_launch=16)
    other = S.submit_ftle(list(lo), list(hi), [9, 5, 13], 0.002, 0.03)  # runs beside it
    rid = S.submit_tracer(seeds, ts, te, fvc.INT_RK4)
    S.run_until_idle()
    got = S.take(rid)
    assert S.info(other)["state"] == fvc.RS_DONE
    pts, rc = om.seed_list(seeds, 1)
    assert rc == 0
    f = frame_of(ts)
    om.set_slot(0, slices[f])
    om.set_slot(1, slices[f + 1])
    om.set_interval(f * FDT, f * FDT + FDT)
    live = np.nonzero(pts["LeftDomain"] == 0)[0]
    sub = pts[live].copy()
    v, rc = om.velocity(sub, ts)
    assert rc == 0
    sub["V"] = v
    pts[live] = sub
    while True:
        tmin, tmax = f * FDT, f * FDT + FDT
        om.set_slot(0, slices[f])
        om.set_slot(1, slices[f + 1])
        om.set_interval(tmin, tmax)
        assert om.advect(pts, max(ts, tmin), min(te, tmax), DT) == 0
        if te - tmax < 1e-10:
            break
        f += 1
    assert (pts["LeftDomain"] == 0).sum() > 100
    for k in ("X", "V", "ElementIndex", "LeftDomain", "LeftDomainTime"):
        np.testing.assert_array_equal(got[k], pts[k], err_msg=k)
    S.close()
    ctx.close()
