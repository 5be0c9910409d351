"""Request serving with continuous batching (fvc_server_*).

Every request must give, bit for bit, what it gives when run alone through
the plain path: a tracer request what fvc_seed_list + one release of
fvc_batch_run give (itself equal to per-frame fvc_advect calls and, in FP64,
to the oracle), an FTLE request what fvc_ftle_create + fvc_ftle_frame over
its frames + fvc_ftle_output give.  It must do so alone on the server and in
a mix admitted while others run, with a pool small enough that requests wait
for memory."""
import math

import numpy as np
import pytest

from fvcpy import fvc, host

pytestmark = pytest.mark.gpu

PIPE_BOUNDS = [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0]
FDT = 0.02
NSLOTS = 4
DT = 2e-3


@pytest.fixture(scope="module")
def pipe():
    t = host.pipe(12, 12, 40, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    slices = [host.tube_velocity(t, xyz, s * FDT, 0.01, 91, s) for s in range(NSLOTS)]
    return t, xyz, conn, adj, slices


def _ctx(pipe, precision):
    t, xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, precision, NSLOTS, coords=xyz, conn=conn, adj=adj)
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
    rng = np.random.default_rng(seed)
    reqs = []
    types = [fvc.INT_RK4, fvc.INT_AB3, fvc.INT_RK4, fvc.INT_EULER, fvc.INT_AB2, fvc.INT_RK4]
    for r, it in enumerate(types):
        n = int(rng.integers(50, 1500))
        ts = float(rng.uniform(0.0, 0.05))
        te = ts + float(rng.uniform(0.005, 0.07))
        xyz = host.tube_seeds(t, n, 0.0, 10.0, seed * 10 + r)
        xyz[0] = [0.99, 0.99, 5.0]  # outside: left at seeding
        reqs.append(("tracer", xyz, ts, te, it))
    reqs.append(("ftle", ([-0.8, -0.8, 0.2], [0.8, 0.8, 9.9], [7, 6, 31]), 0.013, 0.061, fvc.INT_RK4))
    reqs.append(("ftle", ([-1.0, -1.0, 8.0], [1.0, 1.0, 10.0], [9, 9, 12]), 0.02, 0.075, fvc.INT_RK4))
    reqs.append(("tracer", np.zeros((0, 3)), 0.01, 0.03, fvc.INT_RK4))  # empty request
    return reqs


def alone_plain(ctx, rq, lsc=1):
    kind, spec, ts, te, it = rq
    if kind == "tracer":
        xyz = spec
        if xyz.shape[0] == 0:
            return np.zeros(0, host.POINT_DTYPE)
        recs = ctx.seed_list(xyz, lsc)
        P = fvc.Particles(ctx, recs.shape[0])
        P.set(recs)
        B = fvc.Batch(ctx, P, [(0, recs.shape[0], ts, te, it)], 0.0, FDT, NSLOTS, DT)
        B.run(256, 0, 0)
        out = P.get()
        B.close()
        return out
    mn, mx, res = spec
    F = fvc.FtleSlide(ctx, mn, mx, res, lsc, ts, te)
    f = frame_of(ts)
    while True:
        tmin = f * FDT
        done = F.frame(1, f % NSLOTS, (f + 1) % NSLOTS, tmin, tmin + FDT, DT)
        if done == 2:
            break
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
        for k in ("X", "V", "ElementIndex", "AuxElementIndex", "LeftDomain", "LeftDomainTime"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    else:
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("precision,graph", [("f64", True), ("f32", True), ("f32", False)])
def test_served_results_equal_alone_and_in_mix(pipe, precision, graph):
    t = pipe[0]
    ctx = _ctx(pipe, precision)
    reqs = make_requests(t, 4)
    print("serve %s: %d requests" % (precision, len(reqs)), flush=True)
    ref = [alone_plain(ctx, rq) for rq in reqs]
    # alone on the server
    sizes = [rq[1].shape[0] if rq[0] == "tracer" else int(np.prod(rq[1][2])) for rq in reqs]
    pool = 3000  # holds the largest request, not the first five together
    assert max(sizes) <= pool < sum(sizes[:5])
    S = fvc.Server(ctx, 0.0, FDT, NSLOTS, DT, pool_particles=pool, steps_per_launch=7, use_graph=graph)
    for rq, r in zip(reqs, ref):
        rid = submit(S, rq)
        S.run_until_idle()
        assert S.info(rid)["state"] == fvc.RS_DONE
        print("  alone on the server: %s n=%d ok?" % (rq[0], S.info(rid)["n"]), flush=True)
        same(S.take(rid), r, rq[0])
    # the mix: all submitted at once into a pool that cannot hold them all,
    # then more arriving while the first ones run
    ids = [submit(S, rq) for rq in reqs[:5]]
    saw_queue = False
    for k in range(3):
        _, q = S.poll()
        saw_queue = saw_queue or q > 0
    ids += [submit(S, rq) for rq in reqs[5:]]
    for _ in range(100000):
        r, q = S.poll()
        saw_queue = saw_queue or q > 0
        if r == 0 and q == 0:
            break
    assert saw_queue
    st = S.stats()
    assert st["ticks"] > 0 and st["wait_s"] >= 0.0
    for rq, r, rid in zip(reqs, ref, ids):
        inf = S.info(rid)
        assert inf["state"] == fvc.RS_DONE
        assert inf["t_done"] >= inf["t_admit"] >= inf["t_submit"]
        same(S.take(rid), r, rq[0])
    S.close()


def test_server_steps_counted_and_admission_limits(pipe):
    t = pipe[0]
    ctx = _ctx(pipe, "f32")
    S = fvc.Server(ctx, 0.0, FDT, NSLOTS, DT, pool_particles=4096, steps_per_launch=64)
    cap, nbytes = S.capacity()
    assert cap == 4096 and nbytes > cap * 128
    with pytest.raises(fvc.FvcError):
        S.submit_tracer(host.tube_seeds(t, 5000, 0.0, 10.0, 1), 0.0, 0.01)
    with pytest.raises(fvc.FvcError):
        S.submit_ftle([-1, -1, 0], [1, 1, 10], [4, 4, 1], 0.0, 0.02)
    xyz = host.tube_seeds(t, 1000, 0.0, 3.0, 2)
    rid = S.submit_tracer(xyz, 0.0, 0.04, fvc.INT_RK4)
    S.run_until_idle()
    inf = S.info(rid)
    recs = S.take(rid)
    inside = int(((recs["LeftDomain"] == 0) | (recs["LeftDomainTime"] > 0)).sum())
    assert inf["particle_steps"] >= 10 * (inside - 5)
    S.close()


def test_failed_admission_leaves_the_queue(pipe):
    """A request whose admission fails is retired as failed and leaves the
    queue: later polls do not admit it again, and requests after it run.
    The failure is forced from the host: a context whose walks may take at
    most 8 hops, and a point list without LocalSearchChecking whose second
    seed lies 9 units down the pipe from the first (its chain walk crosses
    dozens of elements)."""
    t, xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f32", NSLOTS, coords=xyz, conn=conn, adj=adj, walk_limit=8)
    for s in range(NSLOTS):
        ctx.load_slice(s, s * FDT, slices[s])
    S = fvc.Server(ctx, 0.0, FDT, NSLOTS, DT, pool_particles=4096, steps_per_launch=16, lsc=0)
    bad = S.submit_tracer(np.array([[0.0, 0.0, 0.4], [0.0, 0.0, 9.6]]), 0.0, 0.01)
    with pytest.raises(fvc.FvcError) as e:
        S.poll()
    assert e.value.code == fvc.E_WALK_LIMIT
    assert S.info(bad)["state"] == fvc.RS_FAILED
    for _ in range(3):
        r, q = S.poll()
        assert r == 0 and q == 0
        assert S.info(bad)["state"] == fvc.RS_FAILED
    good = S.submit_tracer(np.array([[0.1, 0.05, 5.0]]), 0.0, 0.01)
    S.run_until_idle()
    assert S.info(good)["state"] == fvc.RS_DONE
    rec = S.take(good)
    assert rec.shape[0] == 1 and rec["LeftDomain"][0] == 0 and rec["X"][0, 2] > 5.0
    S.close()
    ctx.close()
