"""GPU kernels through the C ABI against the oracle (CPU FP64 restatement) on
the same seeded inputs: element indices bit-exact for identical query points
(walk from a guess, lowest-index global search) in every precision mode; FP64
trajectories bit-exact; FP32/BF16/FP16 trajectories within stated tolerances;
Z-order permutation ordered by the reference comparator."""
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


@pytest.fixture(scope="module")
def pipe():
    t = host.pipe(12, 12, 40, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    slices = [host.tube_velocity(t, xyz, s * 0.02, 0.01, 1234, s) for s in range(3)]
    return t, xyz, conn, adj, slices


def query_points(t, n, seed):
    rng = np.random.default_rng(seed)
    pts = host.tube_seeds(t, n, 0.0, 10.0, seed)
    # add points on element faces/edges/vertices (exact grid coordinates) and outside
    g = rng.integers(0, 13, size=(n // 4, 3)).astype(np.float64)
    grid = np.stack([-1 + g[:, 0] * 2 / 12, -1 + g[:, 1] * 2 / 12, g[:, 2] * 10 / 40], 1)
    out = rng.uniform(-1.5, 1.5, size=(n // 8, 3)) * [1, 1, 8]
    return np.concatenate([pts, grid, out])


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16"])
def test_locate_walk_and_global_bit_exact(pipe, precision):
    t, xyz, conn, adj, _ = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, precision, 2, coords=xyz, conn=conn, adj=adj)
    om = oracle.Mesh(3, PIPE_BOUNDS, coords=xyz, conn=conn, adj=adj)
    q = query_points(t, 4000, 11)
    rng = np.random.default_rng(5)
    guess = rng.integers(0, conn.shape[0], q.shape[0]).astype(np.int32)
    print("locate %s n=%d" % (precision, q.shape[0]), flush=True)
    g_gpu = ctx.locate(q, guess, fvc.LOCATE_WALK)
    g_ref, rc = om.locate_local(q, guess)
    assert rc == 0
    np.testing.assert_array_equal(g_gpu, g_ref)
    glob_gpu = ctx.locate(q, None, fvc.LOCATE_GLOBAL)
    glob_ref = om.locate_global(q)
    np.testing.assert_array_equal(glob_gpu, glob_ref)
    both = ctx.locate(q, guess, fvc.LOCATE_WALK_GLOBAL)
    np.testing.assert_array_equal(both, np.where(g_ref >= 0, g_ref, glob_ref))
    assert (glob_ref >= 0).sum() > q.shape[0] // 2


def _stretch(p):
    """z squeezed 50 times: tets 0.17 wide and 0.005 high (condition ~30 times the pipe's)"""
    return p * [1.0, 1.0, 0.02]


def _grade(p):
    """z -> 10 (z/10)^3: layers from 2e-5 to 0.7 high along the pipe"""
    q = p.copy()
    q[:, 2] = 10.0 * (p[:, 2] / 10.0) ** 3
    return q


@pytest.mark.parametrize("shape", ["stretched", "graded"])
@pytest.mark.parametrize("precision", ["f32", "bf16"])
def test_locate_bit_exact_on_badly_shaped_meshes(pipe, shape, precision):
    """The FP32 frames carry each element's own guard band, set by its
    condition number: on flat and strongly graded tets the reduced-precision
    walks must still return the FP64 walk's element for every probe (faces,
    edges and vertices included: the probes are mapped like the nodes)."""
    t, xyz, conn, adj, _ = pipe
    f = _stretch if shape == "stretched" else _grade
    X = f(xyz)
    b = [X[:, 0].min(), X[:, 0].max(), X[:, 1].min(), X[:, 1].max(), X[:, 2].min(), X[:, 2].max()]
    ctx = fvc.Context(3, b, precision, 2, coords=X, conn=conn, adj=adj)
    om = oracle.Mesh(3, b, coords=X, conn=conn, adj=adj)
    q = f(query_points(t, 4000, 17))
    rng = np.random.default_rng(6)
    guess = rng.integers(0, conn.shape[0], q.shape[0]).astype(np.int32)
    print("locate %s %s n=%d" % (shape, precision, q.shape[0]), flush=True)
    g_ref, rc = om.locate_local(q, guess)
    assert rc == 0
    np.testing.assert_array_equal(ctx.locate(q, guess, fvc.LOCATE_WALK), g_ref)
    np.testing.assert_array_equal(ctx.locate(q, None, fvc.LOCATE_GLOBAL), om.locate_global(q))
    assert (g_ref >= 0).sum() > q.shape[0] // 4
    ctx.close()


def test_seed_list_chain_matches_oracle(pipe):
    t, xyz, conn, adj, _ = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f64", 2, coords=xyz, conn=conn, adj=adj)
    om = oracle.Mesh(3, PIPE_BOUNDS, coords=xyz, conn=conn, adj=adj)
    q = query_points(t, 2000, 3)
    for lsc in (1, 0):
        a = ctx.seed_list(q, lsc)
        b, rc = om.seed_list(q, lsc)
        assert rc == 0
        np.testing.assert_array_equal(a["LeftDomain"], b["LeftDomain"])
        np.testing.assert_array_equal(a["ElementIndex"][a["LeftDomain"] == 0], b["ElementIndex"][b["LeftDomain"] == 0])
        np.testing.assert_array_equal(a["LeftDomainTime"], b["LeftDomainTime"])


def _advect_both(pipe, precision, n=3000, t1=0.0, t2=0.02, dt=1e-3):
    t, xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, precision, 4, coords=xyz, conn=conn, adj=adj)
    om = oracle.Mesh(3, PIPE_BOUNDS, coords=xyz, conn=conn, adj=adj)
    for s in range(2):
        ctx.load_slice(s, s * 0.02, slices[s])
        om.set_slot(s, slices[s])
    om.set_interval(0.0, 0.02)
    seeds = host.tube_seeds(t, n, 0.0, 10.0, 99)
    recs, rc = om.seed_list(seeds, 1)
    assert rc == 0
    ref = recs.copy()
    P = fvc.Particles(ctx, n)
    P.set(recs)
    print("advect %s n=%d" % (precision, n), flush=True)
    steps = P.advect(t1, t2, 0, 1, 0.0, 0.02, dt)
    got = P.get()
    assert om.advect(ref, t1, t2, dt) == 0
    return got, ref, steps


def test_advect_f64_bit_exact(pipe):
    got, ref, steps = _advect_both(pipe, "f64")
    np.testing.assert_array_equal(got["X"], ref["X"])
    np.testing.assert_array_equal(got["ElementIndex"], ref["ElementIndex"])
    np.testing.assert_array_equal(got["LeftDomain"], ref["LeftDomain"])
    np.testing.assert_array_equal(got["LeftDomainTime"], ref["LeftDomainTime"])
    inside = (ref["LeftDomainTime"] > -1e-10) & True
    assert steps >= 20 * (inside.sum() - ref["LeftDomain"].sum())


@pytest.mark.parametrize("precision,tol", [("f32", 1e-7), ("bf16", 3e-4), ("f16", 5e-5)])
def test_advect_reduced_precision_within_tolerance(pipe, precision, tol):
    """Position error after one frame (20 RK4 steps) against the oracle's FP64;
    tol is absolute (domain 10), about three times the measured maxima
    (FP32 2.2e-9, BF16 6.5e-5, FP16 8.6e-6 on these 3,000 seeds)."""
    got, ref, _ = _advect_both(pipe, precision)
    same = (got["LeftDomain"] == ref["LeftDomain"])
    assert same.mean() > 0.99
    live = same & (ref["LeftDomain"] == 0) & (ref["LeftDomainTime"] > -1e-10)
    err = np.abs(got["X"][live] - ref["X"][live]).max()
    assert err <= tol, err


def test_advect_cartesian_f64_bit_exact():
    nx, ny = 64, 32
    s0 = host.double_gyre_slice(nx, ny, 0.0)
    s1 = host.double_gyre_slice(nx, ny, 0.5)
    b = [0.0, 2.0, 0.0, 1.0, 0.0, 0.0]
    ctx = fvc.Context(2, b, "f64", 2, cart_min=[0, 0, 0], cart_max=[2, 1, 0], cart_res=[nx, ny, 1])
    om = oracle.Mesh(2, b, cart_min=[0, 0, 0], cart_max=[2, 1, 0], cart_res=[nx, ny, 1])
    ctx.load_slice(0, 0.0, s0)
    ctx.load_slice(1, 0.5, s1)
    om.set_slot(0, s0)
    om.set_slot(1, s1)
    om.set_interval(0.0, 0.5)
    rng = np.random.default_rng(1)
    n = 5000
    recs = np.zeros(n, oracle.POINT_DTYPE)
    recs["X"][:, 0] = rng.uniform(0, 2, n)
    recs["X"][:, 1] = rng.uniform(0, 1, n)
    recs["X"][:10, 0] = 2.0  # boundary points (floor/ceil exits)
    ref = recs.copy()
    P = fvc.Particles(ctx, n)
    P.set(recs)
    P.advect(0.0, 0.5, 0, 1, 0.0, 0.5, 0.01)
    got = P.get()
    assert om.advect(ref, 0.0, 0.5, 0.01) == 0
    np.testing.assert_array_equal(got["X"], ref["X"])
    np.testing.assert_array_equal(got["LeftDomain"], ref["LeftDomain"])
    np.testing.assert_array_equal(got["LeftDomainTime"], ref["LeftDomainTime"])


def _zorder_less(a, b):
    """cudaVC/ZOrderSort/main.cpp:86-101 comparator on integer coordinates."""
    j, x = 0, 0
    for k in range(3):
        y = int(a[k]) ^ int(b[k])
        if x < y and x < (x ^ y):
            j, x = k, y
    return a[j] < b[j]


def test_zsort_order_matches_reference_comparator(pipe):
    t, xyz, conn, adj, _ = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f32", 2, coords=xyz, conn=conn, adj=adj)
    rng = np.random.default_rng(2)
    n = 5000
    pts = rng.uniform([-1, -1, 0], [1, 1, 10], size=(n, 3))
    pts[:50] = pts[50:100]  # duplicates: stable order by index
    lo, hi = np.array([-1.0, -1.0, 0.0]), np.array([1.0, 1.0, 10.0])
    perm = ctx.zsort(pts, lo, hi)
    assert sorted(perm.tolist()) == list(range(n))
    q = np.clip(np.floor((pts - lo) * (2097152.0 / (hi - lo))), 0, 2097151).astype(np.int64)
    qs = q[perm]
    for i in range(n - 1):
        assert not _zorder_less(qs[i + 1], qs[i]), i
        if (qs[i] == qs[i + 1]).all():
            assert perm[i] < perm[i + 1]


@pytest.mark.parametrize("order", [2, 3])
def test_adams_bashforth_f64_bit_exact(pipe, order):
    """AB2/AB3 over two frames (history carried across calls) against the oracle."""
    t, xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f64", 4, coords=xyz, conn=conn, adj=adj)
    om = oracle.Mesh(3, PIPE_BOUNDS, coords=xyz, conn=conn, adj=adj)
    for s in range(3):
        ctx.load_slice(s, s * 0.02, slices[s])
    seeds = host.tube_seeds(t, 2000, 0.0, 10.0, 17)
    recs, rc = om.seed_list(seeds, 1)
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    ref = recs.copy()
    hist = np.zeros(ref.shape[0], oracle.HIST_DTYPE)
    it = fvc.INT_AB2 if order == 2 else fvc.INT_AB3
    for f in range(2):
        tmin, tmax = f * 0.02, (f + 1) * 0.02
        P.advect(tmin, tmax, f, f + 1, tmin, tmax, 1e-3, int_type=it)
        om.set_slot(0, slices[f])
        om.set_slot(1, slices[f + 1])
        om.set_interval(tmin, tmax)
        assert oracle.ab_advect(om, ref, hist, tmin, tmax, 1e-3, order)[0] == 0
    got = P.get()
    np.testing.assert_array_equal(got["X"], ref["X"])
    np.testing.assert_array_equal(got["LeftDomain"], ref["LeftDomain"])
    np.testing.assert_array_equal(got["ElementIndex"], ref["ElementIndex"])


def _face_grid(res):
    """A grid whose points are the pipe mesh's voxel corners: every point lies on
    faces of several elements, so each one's element depends on the chain."""
    return [-1.0, -1.0, 0.0], [1.0, 1.0, 10.0], res


@pytest.mark.parametrize("lsc", [1, 0])
def test_grid_chains_on_faces_match_oracle(pipe, lsc):
    """Tracer-grid (GenerateTracerMesh) and FTLE-grid (InitializeFTLEArray)
    location chains where every point is on element faces, so the chain's
    guesses decide: element indices equal the oracle's."""
    t, xyz, conn, adj, _ = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f32", 2, coords=xyz, conn=conn, adj=adj)
    om = oracle.Mesh(3, PIPE_BOUNDS, coords=xyz, conn=conn, adj=adj)
    mn, mx, res = _face_grid([13, 13, 41])
    dl = [(mx[a] - mn[a]) / (res[a] - 1) for a in range(3)]
    print("tracer grid chain lsc=%d" % lsc, flush=True)
    a = ctx.seed_grid(mn, dl, res, lsc)
    b, rc = om.seed_grid(mn, dl, res, lsc)
    assert rc == 0
    inside = b["LeftDomain"] == 0
    np.testing.assert_array_equal(a["LeftDomain"], b["LeftDomain"])
    np.testing.assert_array_equal(a["ElementIndex"][inside], b["ElementIndex"][inside])
    np.testing.assert_array_equal(a["LeftDomainTime"], b["LeftDomainTime"])
    print("FTLE grid chain lsc=%d" % lsc, flush=True)
    F = fvc.FtleSlide(ctx, mn, mx, res, lsc, 0.0, 0.02)
    g = F.records()
    o, rc = om.ftle_grid(mn, mx, res, lsc)
    assert rc == 0
    inside = o["Pt"]["LeftDomain"] == 0
    np.testing.assert_array_equal(g["LeftDomain"], o["Pt"]["LeftDomain"])
    np.testing.assert_array_equal(g["ElementIndex"][inside], o["Pt"]["ElementIndex"][inside])
    np.testing.assert_array_equal(g["LeftDomainTime"], o["Pt"]["LeftDomainTime"])
    assert (g["LeftDomain"] == 0).sum() > 1000


def _fan(nt):
    """2-D fan: nt triangles around a hub node, so the hub is held by all of
    them and points on a spoke by two."""
    ang = 2 * np.pi * np.arange(nt) / nt
    xyz = np.zeros((nt + 1, 3))
    xyz[1:, 0], xyz[1:, 1] = np.cos(ang), np.sin(ang)
    k = np.arange(nt)
    conn = np.stack([np.zeros(nt, np.int64), 1 + k, 1 + (k + 1) % nt, np.zeros(nt, np.int64)], 1).astype(np.int32)
    # slot 0 crosses the edge opposite node 1, slot 1 opposite node 2, slot 2 opposite node 0
    adj = np.stack([(k + 1) % nt, (k - 1) % nt, -np.ones(nt, np.int64), -np.ones(nt, np.int64)], 1).astype(np.int32)
    return xyz, conn, adj


@pytest.mark.parametrize("lsc", [1, 0])
def test_fan_chains_match_oracle(lsc):
    """List and grid chains on a 2-D fan: grids through the hub (held by 40
    triangles, more than a chain table keeps, so the chain is replayed point by
    point) and grids along spokes only (two triangles each, tabulated)."""
    xyz, conn, adj = _fan(40)
    bounds = [-1.0, 1.0, -1.0, 1.0, 0.0, 0.0]
    ctx = fvc.Context(2, bounds, "f64", 2, coords=xyz, conn=conn, adj=adj)
    om = oracle.Mesh(2, bounds, coords=xyz, conn=conn, adj=adj)
    for mn, res in (([-1.0, -1.0, 0.0], [21, 21, 1]), ([-0.95, -1.0, 0.0], [20, 21, 1])):
        dl = [0.1, 0.1, 0.0]
        print("fan grid chain lsc=%d mn=%s" % (lsc, mn), flush=True)
        a = ctx.seed_grid(mn, dl, res, lsc)
        b, rc = om.seed_grid(mn, dl, res, lsc)
        assert rc == 0
        inside = b["LeftDomain"] == 0
        assert inside.sum() > 100
        np.testing.assert_array_equal(a["LeftDomain"], b["LeftDomain"])
        np.testing.assert_array_equal(a["ElementIndex"][inside], b["ElementIndex"][inside])
        np.testing.assert_array_equal(a["LeftDomainTime"], b["LeftDomainTime"])
        q = np.stack(np.meshgrid(mn[0] + 0.1 * np.arange(res[0]), mn[1] + 0.1 * np.arange(res[1]), indexing="ij"), -1)
        q = np.concatenate([q.reshape(-1, 2), np.zeros((res[0] * res[1], 1))], 1)
        q = q[np.random.default_rng(5).permutation(q.shape[0])]
        a = ctx.seed_list(q, lsc)
        b, rc = om.seed_list(q, lsc)
        assert rc == 0
        inside = b["LeftDomain"] == 0
        np.testing.assert_array_equal(a["LeftDomain"], b["LeftDomain"])
        np.testing.assert_array_equal(a["ElementIndex"][inside], b["ElementIndex"][inside])
        np.testing.assert_array_equal(a["LeftDomainTime"], b["LeftDomainTime"])


def test_ftle_chain_on_voxel_edges_matches_oracle():
    """FTLE grid over a Y-bifurcation's bounding box with a spacing of two
    voxels in x and z, so every grid point inside lies on a voxel edge and its
    element depends on the chain's guesses."""
    t = host.ybif(24, 9, 32)
    xyz, conn, adj = host.tube_mesh(t)
    lo, hi = xyz.min(0), xyz.max(0)
    bounds = [float(lo[0]), float(hi[0]), float(lo[1]), float(hi[1]), float(lo[2]), float(hi[2])]
    ctx = fvc.Context(3, bounds, "f32", 2, coords=xyz, conn=conn, adj=adj)
    om = oracle.Mesh(3, bounds, coords=xyz, conn=conn, adj=adj)
    mn, mx = [t.lo[0], t.lo[1], t.lo[2]], [t.hi[0], t.hi[1], t.hi[2]]
    for res in ([13, 7, 17], [25, 5, 33]):
        print("FTLE grid chain on voxel edges res=%s" % res, flush=True)
        F = fvc.FtleSlide(ctx, mn, mx, res, 1, 0.0, 0.02)
        g = F.records()
        o, rc = om.ftle_grid(mn, mx, res, 1)
        assert rc == 0
        inside = o["Pt"]["LeftDomain"] == 0
        assert inside.sum() > 100
        np.testing.assert_array_equal(g["LeftDomain"], o["Pt"]["LeftDomain"])
        np.testing.assert_array_equal(g["ElementIndex"][inside], o["Pt"]["ElementIndex"][inside])
        np.testing.assert_array_equal(g["LeftDomainTime"], o["Pt"]["LeftDomainTime"])


def _chain_both(ctx, fn):
    ctx.set_chain_mode(1)
    a = fn()
    ctx.set_chain_mode(0)
    b = fn()
    for k in ("ElementIndex", "LeftDomain", "LeftDomainTime"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    return b


def test_tabulated_chains_equal_replay_at_full_size():
    """Seed chains at the sizes the benchmarks use: the tabulated chain and the
    point-by-point replay give the same records (FTLE grid of the serving
    benchmark over the Y-bifurcation, where every inside point lies on a voxel
    edge; the pipe benchmark's 128x128x640 FTLE grid; a tracer grid and a
    4M-point list)."""
    t = host.ybif(184, 65, 256)
    xyz, conn, adj = host.tube_mesh(t)
    lo, hi = xyz.min(0), xyz.max(0)
    ctx = fvc.Context(3, [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]], "bf16", 2, coords=xyz, conn=conn, adj=adj)
    mn, mx = [t.lo[0], t.lo[1], t.lo[2]], [t.hi[0], t.hi[1], t.hi[2]]
    for res in ([93, 33, 129], [70, 25, 97]):
        print("Y-bifurcation FTLE grid chain res=%s" % res, flush=True)
        r = _chain_both(ctx, lambda: fvc.FtleSlide(ctx, mn, mx, res, 1, 0.0, 0.01).records())
        assert (r["LeftDomain"] == 0).sum() > 10000
    del ctx
    p = host.pipe(60, 60, 330, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(p)
    ctx = fvc.Context(3, PIPE_BOUNDS, "f32", 2, coords=xyz, conn=conn, adj=adj)
    print("pipe FTLE grid chain 128x128x640", flush=True)
    _chain_both(ctx, lambda: fvc.FtleSlide(ctx, [-1, -1, 0], [1, 1, 10], [128, 128, 640], 1, 0.0, 0.02).records())
    print("pipe tracer grid chain on voxel corners", flush=True)
    _chain_both(ctx, lambda: ctx.seed_grid([-1.0, -1.0, 0.0], [2.0 / 60, 2.0 / 60, 10.0 / 330], [61, 61, 331], 1))
    seeds = np.concatenate([host.tube_seeds(p, 4000000, 0.0, 10.0, 5), query_points(p, 40000, 6)])
    print("pipe list chain n=%d" % seeds.shape[0], flush=True)
    _chain_both(ctx, lambda: ctx.seed_list(seeds, 1))


def test_particle_records_with_foreign_element_ids(pipe):
    """Records from files carry element ids the kernels walk from: an id
    outside [-1, elements) on a point not flagged as left is refused before
    anything reaches the device; points flagged as left are stored without
    an element when their id lies outside the mesh (ompVC leaves those ids
    unset in its stored grids) and keep it otherwise."""
    t, xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f64", 2, coords=xyz, conn=conn, adj=adj)
    recs = ctx.seed_list(host.tube_seeds(t, 64, 0.0, 10.0, 3), 1)
    assert (recs["LeftDomain"] == 0).all()
    P = fvc.Particles(ctx, recs.shape[0])
    for bad in (conn.shape[0], -2, 2 ** 31 - 1):
        r = recs.copy()
        r["ElementIndex"][5] = bad
        with pytest.raises(fvc.FvcError):
            P.set(r)
    r = recs.copy()
    r["LeftDomain"][7:9] = 1
    r["ElementIndex"][7] = -2072131723
    P.set(r)
    back = P.get()
    assert back["ElementIndex"][7] == -1
    np.testing.assert_array_equal(np.delete(back["ElementIndex"], 7), np.delete(recs["ElementIndex"], 7))


@pytest.mark.parametrize("precision,pos_tol,pos_p99,ftle_p99", [("f32", 2e-5, 2e-7, 1e-5), ("f16", 0.03, 6e-4, 1e-2),
                                                                ("bf16", 0.25, 5e-3, 0.08)])
def test_double_gyre_ftle_reduced_precision(precision, pos_tol, pos_p99, ftle_p99):
    """The 2-D double-gyre FTLE (256x128 nodes, 21 slices, RK4 dt=0.01 over
    T=10) on a 256x128 grid against the FP64 path (bit-identical to the
    oracle).  FP32 meets 1e-5 of the domain size (2) in final position and
    1e-3 relative FTLE.  The bounds are about three times the measured values
    (FP32: position max 1.2e-6, p99 3.6e-8, FTLE p99 8.0e-7; FP16: 8.6e-3,
    2.2e-4, 4.3e-3; BF16: 0.081, 1.7e-3, 0.034): the flow is chaotic
    (separations grow up to ~100x over T), so a velocity rounding of 2^-12
    (FP16) or 2^-9 (BF16) ends as these position errors."""
    nx, ny = 256, 128
    slices = [host.double_gyre_slice(nx, ny, 0.5 * s) for s in range(21)]
    out = {}
    for prec in ("f64", precision):
        ctx = fvc.Context(2, [0.0, 2.0, 0.0, 1.0, 0.0, 0.0], prec, 21, cart_min=[0.0, 0.0, 0.0],
                          cart_max=[2.0, 1.0, 0.0], cart_res=[nx, ny, 1])
        for s, sl in enumerate(slices):
            ctx.load_slice(s, 0.5 * s, sl)
        print("double gyre FTLE %s" % prec, flush=True)
        F = fvc.FtleSlide(ctx, [0.0, 0.0, 0.0], [2.0, 1.0, 0.0], [256, 128, 1], 1, 0.0, 10.0)
        for f in range(20):
            F.frame(1, f, f + 1, 0.5 * f, 0.5 * (f + 1), 0.01)
        wt, _ = F.output(10.0)
        out[prec] = (wt, F.records())
    (w64, r64), (w, r) = out["f64"], out[precision]
    np.testing.assert_array_equal(r["LeftDomain"], r64["LeftDomain"])
    both = r64["LeftDomain"] == 0
    assert both.sum() > 30000
    d = np.linalg.norm(r["X"][both] - r64["X"][both], axis=1)
    assert d.max() <= pos_tol and np.percentile(d, 99) <= pos_p99, (d.max(), np.percentile(d, 99))
    big = np.abs(w64) > 1e-3
    assert np.percentile(np.abs(w[big] - w64[big]) / np.abs(w64[big]), 99) <= ftle_p99


def test_ftle_slide_from_stored_initial_state(pipe):
    """A slide rebuilt from another slide's initial state (the FTLE_ICFile
    round trip: fvc_ftle_initial, fvc_ftle_create_from) integrates to the
    same FTLE, bit for bit, and the rebuilt slide reports the same initial
    state; grids whose size disagrees with the arrays are refused."""
    t, xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f64", 3, coords=xyz, conn=conn, adj=adj)
    for s in range(3):
        ctx.load_slice(s, s * 0.02, slices[s])
    mn, mx, res = [-1.0, -1.0, 0.0], [1.0, 1.0, 10.0], [13, 13, 41]
    dl = [(mx[a] - mn[a]) / (res[a] - 1) for a in range(3)]
    A = fvc.FtleSlide(ctx, mn, mx, res, 1, 0.0, 0.04)
    init = A.initial()
    assert (init[0]["LeftDomain"] == 0).sum() > 3000 and init[1].sum() > 0
    B = fvc.FtleSlide.from_records(ctx, mn, dl, res, *init, 0.0, 0.04)
    for k, v in zip(("recs", "have", "wt", "wot"), B.initial()):
        if k == "recs":
            for f in ("X", "ElementIndex", "LeftDomain", "LeftDomainTime"):
                np.testing.assert_array_equal(v[f], init[0][f], err_msg=f)
        else:
            np.testing.assert_array_equal(v, init[("recs", "have", "wt", "wot").index(k)], err_msg=k)
    out = []
    for F in (A, B):
        for f in range(2):
            F.frame(1, f, f + 1, 0.02 * f, 0.02 * (f + 1), 1e-3)
        out.append(F.output(0.04))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    with pytest.raises(ValueError):
        fvc.FtleSlide.from_records(ctx, mn, dl, [13, 13, 40], *init, 0.0, 0.04)


@pytest.mark.parametrize("extrapolate", [0, 1])
def test_tracer_list_of_live_particles(pipe, extrapolate):
    """Seeds near the outlet mostly leave within a few frames: fvc_advect
    then launches over a list of the particles still to integrate, and every
    record equals the run that launches the whole range each frame (its
    state rewritten from the host between frames, which drops any list).
    With extrapolation nothing is idle, so no list is built."""
    t, xyz, conn, adj, slices = pipe
    ctx = fvc.Context(3, PIPE_BOUNDS, "f64", 3, coords=xyz, conn=conn, adj=adj)
    for s in range(3):
        ctx.load_slice(s, s * 0.02, slices[s])
    seeds = host.tube_seeds(t, 20000, 9.97, 10.0, 5)
    recs = ctx.seed_list(seeds, 1)
    A, B = fvc.Particles(ctx, recs.shape[0]), fvc.Particles(ctx, recs.shape[0])
    A.set(recs)
    B.set(recs)
    listed = []
    for P in (A, B):
        P.launch_velocity(0.0, 0, 1, 0.0, 0.02)
    sa = sb = 0
    for k in range(8):
        t1, t2 = 0.0025 * k, 0.0025 * (k + 1)
        sa += A.advect(t1, t2, 0, 1, 0.0, 0.02, 1e-3, extrapolate=extrapolate)
        listed.append(A.active())
        sb += B.advect(t1, t2, 0, 1, 0.0, 0.02, 1e-3, extrapolate=extrapolate)
        B.set(B.get())
        assert B.active() == -1
    print("active-list lengths:", listed, flush=True)
    a, b = A.get(), B.get()
    assert sa == sb
    for f in ("X", "V", "ElementIndex", "LeftDomain", "LeftDomainTime"):
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)
    if extrapolate:
        assert listed == [-1] * 8
    else:
        gone = int((a["LeftDomain"] != 0).sum())
        assert gone > recs.shape[0] // 4
        assert any(0 <= n < recs.shape[0] for n in listed)
        assert listed[-1] >= int((a["LeftDomain"] == 0).sum())
