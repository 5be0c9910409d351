"""Batched releases (fvc_batch_*) against the same releases run one at a time.

A release run alone is GetVelocity at its start time followed by one Advect
per data frame (tracers.c:202-229 through main.c:170-227).  The fused
persistent kernel must give every release the same result bit for bit,
whatever the block size, step budget per launch or graph capture, and after
a reset; in FP64 that result must also equal the oracle's."""
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

PIPE_BOUNDS = [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0]
FDT = 0.02
NSLOTS = 4
DT = 2e-3


@pytest.fixture(scope="module")
def pipe():
    t = host.pipe(12, 12, 40, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    slices = [host.tube_velocity(t, xyz, s * FDT, 0.01, 77, s) for s in range(NSLOTS)]
    om = oracle.Mesh(3, PIPE_BOUNDS, coords=xyz, conn=conn, adj=adj)
    return t, xyz, conn, adj, slices, om


def make_releases(t, om, seed, n_rel=7):
    """Uneven releases: sizes, start times, horizons and integrators from a seed."""
    rng = np.random.default_rng(seed)
    types = [fvc.INT_RK4, fvc.INT_AB3, fvc.INT_AB2, fvc.INT_EULER]
    rel, recs, begin = [], [], 0
    for r in range(n_rel):
        cnt = int(rng.integers(1, 900)) if r else 0  # one empty release
        t0 = float(rng.uniform(0.0, 0.05))
        if r == 1:
            t0 = 2 * FDT  # starts exactly on a frame boundary
        horizon = float(rng.uniform(0.0, 0.06)) if r != 2 else 0.0  # one zero-length release
        seeds = host.tube_seeds(t, max(cnt, 1), 0.0, 10.0, seed * 100 + r)[:cnt]
        # a few seeds outside the pipe (left at seeding) and right at the outlet
        if cnt > 10:
            seeds[0] = [0.99, 0.99, 5.0]
            seeds[1] = [0.0, 0.0, 9.9999]
        pts, rc = om.seed_list(seeds, 1)
        assert rc == 0
        recs.append(pts)
        rel.append((begin, cnt, t0, t0 + horizon, types[r % 4]))
        begin += cnt
    return rel, np.concatenate(recs)


def frame_of(ts):
    f = int(math.floor(ts / FDT))
    while f * FDT > ts:
        f -= 1
    while (f + 1) * FDT <= ts:
        f += 1
    return f


def run_one_at_a_time(ctx, recs, rel):
    """The releases through fvc_launch_velocity + fvc_advect, frame by frame."""
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    steps = 0
    for (b, c, ts, te, it) in rel:
        if c == 0:
            continue
        f = frame_of(ts)
        P.launch_velocity(ts, f % NSLOTS, (f + 1) % NSLOTS, f * FDT, f * FDT + FDT, begin=b, count=c)
        while True:
            tmin = f * FDT
            tmax = tmin + FDT
            steps += P.advect(max(ts, tmin), min(te, tmax), f % NSLOTS, (f + 1) % NSLOTS, tmin, tmax, DT,
                              int_type=it, begin=b, count=c)
            if te - tmax < 1e-10:
                break
            f += 1
    return P.get(), steps


def run_oracle(om, slices, recs, rel):
    out = recs.copy()
    for (b, c, ts, te, it) in rel:
        if c == 0:
            continue
        pts = out[b:b + c].copy()
        hist = np.zeros(c, oracle.HIST_DTYPE)
        f = frame_of(ts)
        om.set_slot(0, slices[f % NSLOTS])
        om.set_slot(1, slices[(f + 1) % NSLOTS])
        om.set_interval(f * FDT, f * FDT + FDT)
        live = np.nonzero(pts["LeftDomain"] == 0)[0]
        sub = pts[live].copy()
        v, rc = om.velocity(sub, ts)
        assert rc == 0
        sub["V"] = v
        pts[live] = sub
        while True:
            tmin = f * FDT
            tmax = tmin + FDT
            om.set_slot(0, slices[f % NSLOTS])
            om.set_slot(1, slices[(f + 1) % NSLOTS])
            om.set_interval(tmin, tmax)
            a, bb = max(ts, tmin), min(te, tmax)
            if it in (fvc.INT_AB2, fvc.INT_AB3):
                rc, _ = oracle.ab_advect(om, pts, hist, a, bb, DT, 2 if it == fvc.INT_AB2 else 3)
            else:
                rc = om.advect(pts, a, bb, DT, int_type=it)
            assert rc == 0
            if te - tmax < 1e-10:
                break
            f += 1
        out[b:b + c] = pts
    return out


def _ctx(pipe, precision):
    t, xyz, conn, adj, slices, om = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, precision, NSLOTS, coords=xyz, conn=conn, adj=adj)
    for s in range(NSLOTS):
        ctx.load_slice(s, s * FDT, slices[s])
    return ctx


def _same(a, b):
    for k in ("X", "V", "ElementIndex", "LeftDomain", "LeftDomainTime"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


CONFIGS = {
    "f64": [(256, 64, 1), (256, 1, 0), (256, 0, 1)],
    "f32": [(64, 1, 0), (128, 5, 1), (256, 64, 1), (512, 256, 1), (1024, 0, 1)],
    "bf16": [(256, 3, 1)],
}


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16"])
def test_batch_matches_releases_one_at_a_time(pipe, precision):
    t, xyz, conn, adj, slices, om = pipe
    rel, recs = make_releases(t, om, 5)
    ctx = _ctx(pipe, precision)
    print("batch %s n=%d releases=%d" % (precision, recs.shape[0], len(rel)), flush=True)
    ref, ref_steps = run_one_at_a_time(ctx, recs, rel)
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    B = fvc.Batch(ctx, P, rel, 0.0, FDT, NSLOTS, DT)
    for (bs, budget, graph) in CONFIGS[precision]:
        B.reset()
        print("  block=%d budget=%d graph=%d" % (bs, budget, graph), flush=True)
        steps, launches = B.run(bs, budget, graph)
        _same(P.get(), ref)
        assert steps == ref_steps
        assert launches >= 1
    B.close()
    assert ref_steps > 0
    moved = (ref["LeftDomain"] == 0) & (np.abs(ref["X"] - recs["X"]).max(1) > 0)
    assert moved.sum() > recs.shape[0] // 3


def test_batch_f64_matches_oracle(pipe):
    t, xyz, conn, adj, slices, om = pipe
    rel, recs = make_releases(t, om, 9)
    ctx = _ctx(pipe, "f64")
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    B = fvc.Batch(ctx, P, rel, 0.0, FDT, NSLOTS, DT)
    print("batch f64 vs oracle n=%d" % recs.shape[0], flush=True)
    B.run(256, 16, 1)
    got = P.get()
    B.close()
    ref = run_oracle(om, slices, recs, rel)
    _same(got, ref)
    assert (ref["LeftDomain"] == 1).sum() > 0


def test_batch_tuner_returns_a_swept_config(pipe):
    t, xyz, conn, adj, slices, om = pipe
    rel, recs = make_releases(t, om, 3, n_rel=5)
    ctx = _ctx(pipe, "f32")
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    B = fvc.Batch(ctx, P, rel, 0.0, FDT, NSLOTS, DT)
    best, ms = B.tune(blocks=(64, 256), budgets=(4, 0))
    assert best["block"] in (64, 256) and best["steps_per_launch"] in (4, 0) and ms > 0
    B.run(**best)
    got = P.get()
    ref, _ = run_one_at_a_time(ctx, recs, rel)
    _same(got, ref)
    B.close()


def test_batch_rejects_bad_input(pipe):
    t, xyz, conn, adj, slices, om = pipe
    ctx = _ctx(pipe, "f32")
    P = fvc.Particles(ctx, 100)
    with pytest.raises(fvc.FvcError):
        fvc.Batch(ctx, P, [(0, 60, 0.0, 0.01, fvc.INT_RK4), (50, 20, 0.0, 0.01, fvc.INT_RK4)], 0.0, FDT, NSLOTS, DT)
    with pytest.raises(fvc.FvcError):
        fvc.Batch(ctx, P, [(0, 60, 0.01, 0.0, fvc.INT_RK4)], 0.0, FDT, NSLOTS, DT)
    with pytest.raises(fvc.FvcError):
        fvc.Batch(ctx, P, [(0, 200, 0.0, 0.01, fvc.INT_RK4)], 0.0, FDT, NSLOTS, DT)
    B = fvc.Batch(ctx, P, [(0, 100, 0.0, 0.0, fvc.INT_RK4)], 0.0, FDT, NSLOTS, DT)
    with pytest.raises(fvc.FvcError):
        B.run(block=96)
    B.close()
