"""Large device -> host copies (fvc_particles_get, fvc_ftle_output) travel in
8 MB chunks through a pinned staging buffer of the context (at most 64 MB, so
larger arrays go in rounds), host threads copying each chunk out in stripes
while the next chunks' DMA runs; blocks under 4 MB are copied directly.
Whatever the path, the host must receive the device's bytes exactly."""
import numpy as np
import pytest

from fvcpy import fvc, host

pytestmark = pytest.mark.gpu


def _records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, host.POINT_DTYPE)
    r["X"] = rng.standard_normal((n, 3))
    r["V"] = rng.standard_normal((n, 3)) * 1e-300  # subnormals and tiny values survive bit for bit
    r["ElementIndex"] = rng.integers(-1, 1 << 30, n)
    r["AuxElementIndex"] = rng.integers(-1, 1 << 30, n)
    r["LeftDomain"] = rng.integers(0, 2, n)
    r["LeftDomainTime"] = rng.uniform(-1.0, 1.0, n)
    r["Scalar"] = rng.standard_normal(n)
    r["X"][:5, 0] = [0.0, -0.0, np.inf, -np.inf, 5e-324]
    return r


def _same_bytes(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_particle_records_round_trip_through_staged_copies():
    """9M records: each FP64 array is 72 MB (two rounds of the staging
    buffer, nine chunks, the last one partial), each int32 array 36 MB; then
    an unaligned sub-range, and a small range that takes the direct copy."""
    n = 9_000_000
    ctx = fvc.Context(2, [0.0, 2.0, 0.0, 1.0, 0.0, 0.0], "f64", 2, cart_min=[0, 0, 0], cart_max=[2, 1, 0],
                      cart_res=[16, 8, 1])
    recs = _records(n, 11)
    P = fvc.Particles(ctx, n)
    print("round trip of %d records" % n, flush=True)
    P.set(recs)
    assert _same_bytes(P.get(), recs)
    b = 12_345
    assert _same_bytes(P.get(n - 2 * b - 7, begin=b), recs[b:n - b - 7])
    assert _same_bytes(P.get(1000, begin=777), recs[777:1777])
    P.close()
    ctx.close()


def test_large_ftle_output_equals_a_narrow_grid_of_the_same_spacing():
    """Double-gyre FTLE on a 1101 x 1001 grid (8.8 MB per field: staged) and
    on its first 101 columns as a grid of their own (0.8 MB: direct copy),
    both with spacing 2^-9 so that every grid point is the same double.
    Away from the narrow grid's last column, each point has the same stencil
    in both grids, so both fields agree bit for bit there; wot = wt * T with
    T = 2 holds on both."""
    nx, ny = 256, 128
    ctx = fvc.Context(2, [0.0, 2.0, 0.0, 1.0, 0.0, 0.0], "f64", 5, cart_min=[0, 0, 0], cart_max=[2, 1, 0],
                      cart_res=[nx, ny, 1])
    for s in range(5):
        ctx.load_slice(s, 0.5 * s, host.double_gyre_slice(nx, ny, 0.5 * s))
    h = 2.0 ** -9
    out = {}
    for name, cols in (("large", 1101), ("narrow", 101)):
        print("double gyre FTLE, %d x 1001 grid" % cols, flush=True)
        F = fvc.FtleSlide(ctx, [0.0, 0.0, 0.0], [(cols - 1) * h, 1000 * h, 0.0], [cols, 1001, 1], 1, 0.0, 2.0)
        for f in range(4):
            F.frame(1, f, f + 1, 0.5 * f, 0.5 * (f + 1), 0.01)
        wt, wot = F.output(2.0)
        out[name] = (wt.reshape(1001, cols), wot.reshape(1001, cols))
        F.close()
    (Lt, Lo), (St, So) = out["large"], out["narrow"]
    np.testing.assert_array_equal(Lt[:, :100], St[:, :100])
    np.testing.assert_array_equal(Lo[:, :100], So[:, :100])
    live = Lt > 0
    assert live.sum() > 500_000
    np.testing.assert_array_equal(Lt[live] * 2.0 == Lo[live], np.ones(live.sum(), bool))
    ctx.close()
