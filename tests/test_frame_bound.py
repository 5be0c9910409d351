"""The FP32 element frames of the reduced-precision walk (csrc/fvc_device.h,
ld_frame / frame_nat / decide_fast; built by fvc_api.hip's build_frames):
natural coordinates (r, s, t) = lo + M (q - o) with M the FP32 inverse edge
matrix, o node 0 rounded to FP32 and lo the natural coordinates of o.  The
device decides in FP32 only outside a band of 4 (18 cond + 12) 2^-24 L +
12 2^-24 |M| |d| around each threshold, and otherwise asks the reference's
FP64 formula, so the element a walk returns depends on the FP32 error staying
below the band.

These tests emulate that FP32 arithmetic with numpy (float32 rounding after
every operation, FMAs rounded once from the exact double product) on the
synthetic pipe and Y-bifurcation meshes and on the reference's Stenosis2D
triangles, for query points inside, within 1e-4 of a face or an edge, and up
to two elements away, with probe offsets d of a stage of an RK4 step, and
check the measured error against the first-order bound (without the factor
4).  A frame built wrongly (a transposed M, a dropped lo, an origin one ulp
off) exceeds the same bound, which the last test shows on a long pipe."""
import numpy as np
import pytest

import cases
from fvcpy import host

F32 = np.float32
U = 2.0 ** -24


def frames(P, conn, dims):
    """per element: M (float32), o (float32), lo (float32), cond, row norm"""
    P4 = P[conn[:, :dims + 1]]
    E = np.stack([P4[:, q + 1, :dims] - P4[:, 0, :dims] for q in range(dims)], axis=2)  # E[:, a, q]
    Minv = np.linalg.inv(E)
    rown = np.maximum(np.abs(Minv).sum(2).max(1), np.abs(Minv.sum(1)).sum(1))
    cond = rown * np.abs(E).sum(2).max(1)
    o = P4[:, 0, :dims].astype(F32)
    lo = np.einsum("nij,nj->ni", Minv, o.astype(np.float64) - P4[:, 0, :dims]).astype(F32)
    return Minv, Minv.astype(F32), o, lo, cond, rown


def fma(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def fp32_nat(M, o, lo, A, d, dims):
    q = ((A - o).astype(F32) + d).astype(F32)
    out = []
    for i in range(dims):
        v = lo[:, i]
        for j in range(dims):
            v = fma(M[:, i, j], q[:, j], v)
        out.append(v)
    l0 = F32(1) - out[0]
    for v in out[1:]:
        l0 = (l0 - v).astype(F32)
    return np.stack(out + [l0], 1).astype(np.float64)


def queries(rng, n, dims):
    """natural coordinates (lam[:, 1:] = r, s[, t]; lam[:, 0] = 1 - sum):
    inside, near a face, near an edge (3-D) or vertex (2-D), outside"""
    k = dims + 1
    lam = rng.dirichlet(np.ones(k), n)
    m = n // 4
    f = rng.integers(0, k, n)
    lam[m:2 * m][np.arange(m), f[m:2 * m]] = rng.uniform(-1e-4, 1e-4, m)
    lam[2 * m:3 * m][np.arange(m), f[2 * m:3 * m]] = rng.uniform(-1e-4, 1e-4, m)
    lam[2 * m:3 * m][np.arange(m), (f[2 * m:3 * m] + 1) % k] = rng.uniform(-1e-4, 1e-4, m)
    lam[3 * m:] = rng.uniform(-2.0, 3.0, (n - 3 * m, k))
    lam[:, 0] = 1.0 - lam[:, 1:].sum(1)
    return lam


def check_mesh(P, conn, dims, n=200_000, seed=0, stage=2e-3, corrupt=None):
    rng = np.random.default_rng(seed)
    Minv, M, o, lo, cond, rown = frames(P, conn, dims)
    if corrupt is not None:
        M, o, lo = corrupt(M, o, lo)
    e = rng.integers(0, conn.shape[0], n)
    lam = queries(rng, n, dims)
    P0 = P[conn[e, 0], :dims]
    E = np.stack([P[conn[e, q + 1], :dims] - P0 for q in range(dims)], axis=2)
    X = P0 + np.einsum("nij,nj->ni", E, lam[:, 1:])  # the query minus the mesh centre
    A = X.astype(F32)
    D = (X - A.astype(np.float64)).astype(F32)
    d = (D + rng.uniform(-stage, stage, (n, dims)).astype(F32)).astype(F32)  # a probe offset
    Q = A.astype(np.float64) + d.astype(np.float64)  # the point (A, d) stands for
    ex = np.einsum("nij,nj->ni", Minv[e], Q - P0)
    ex = np.concatenate([ex, (1.0 - ex.sum(1))[:, None]], 1)
    got = fp32_nat(M[e], o[e], lo[e], A, d, dims)
    err = np.abs(got - ex).max(1)
    L = np.maximum(1.0, np.abs(ex).max(1))
    bound = (18.0 * cond[e] + 12.0) * U * L + 3.0 * U * rown[e] * np.abs(d).max(1)
    return err, bound


def _pipe():
    xyz, conn, _ = host.tube_mesh(host.pipe(12, 12, 40, 1.0, 10.0))
    return xyz - (xyz.max(0) + xyz.min(0)) / 2, conn


def test_pipe_frames_within_first_order_bound():
    P, conn = _pipe()
    err, bound = check_mesh(P, conn, 3)
    assert (err <= bound).all(), float((err / bound).max())
    assert (err / bound).max() < 0.25  # the band (4x the bound) keeps a margin of 16


def test_ybif_frames_within_first_order_bound():
    xyz, conn, _ = host.tube_mesh(host.ybif(24, 9, 32))
    P = xyz - (xyz.max(0) + xyz.min(0)) / 2
    err, bound = check_mesh(P, conn, 3, seed=1)
    assert (err <= bound).all(), float((err / bound).max())


def test_stenosis2d_frames_within_first_order_bound():
    xyz, conn = cases.sten_mesh()
    P = xyz - (xyz.max(0) + xyz.min(0)) / 2
    err, bound = check_mesh(P, conn, 2, seed=2, stage=2e-6)  # a stage of dt 1e-5 at the mesh's speeds
    assert (err <= bound).all(), float((err / bound).max())


def _long_pipe():
    """a pipe 100 long with cells about 0.25 wide: node rounding at the domain's
    extent (2^-24 x 50) is large beside the elements' own rounding (397
    layers, so the node coordinates are not binary fractions)"""
    xyz, conn, _ = host.tube_mesh(host.pipe(8, 8, 397, 1.0, 100.0))
    return xyz - (xyz.max(0) + xyz.min(0)) / 2, conn


def test_long_pipe_frames_within_first_order_bound():
    P, conn = _long_pipe()
    err, bound = check_mesh(P, conn, 3, seed=3)
    assert (err <= bound).all(), float((err / bound).max())


@pytest.mark.parametrize("mistake", ["transposed_M", "no_lo", "origin_off_by_one_ulp"])
def test_planted_frame_mistakes_exceed_the_bound(mistake):
    """the check above is sharp enough to see a frame built wrongly"""
    P, conn = _long_pipe()

    def corrupt(M, o, lo):
        if mistake == "transposed_M":
            return np.ascontiguousarray(M.transpose(0, 2, 1)), o, lo
        if mistake == "no_lo":
            return M, o, np.zeros_like(lo)
        return M, np.nextafter(o, F32(np.inf)), lo

    err, bound = check_mesh(P, conn, 3, n=50_000, corrupt=corrupt)
    assert (err > bound).any()
