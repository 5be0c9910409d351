"""Device mesh Z-order renumbering (fvc_mesh_zorder, fvc_gather_rows and the
fvc_zorder tool) against the reference tool's files and the oracle: node and
element orders, renumbered coordinates, connectivity and adjacency
bit-identical; trajectories unchanged by the renumbering."""
import os
import subprocess
import sys

import numpy as np
import pytest

import cases
import make_zorder_goldens as zg
from fvcpy import fvc, host

sys.path.insert(0, os.path.join(cases.REPO, "oracle"))
import oracle  # noqa: E402

pytestmark = pytest.mark.gpu


def read_new(path, kind):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    if kind == "coordinates":
        return np.frombuffer(raw[4:], np.float64).reshape(n, 3)
    return np.frombuffer(raw[4:], np.int32).reshape(n, 4)


@pytest.mark.parametrize("case", sorted(zg.ZCASES))
def test_device_renumbering_matches_reference_files(case):
    xyz, conn, adj = zg.make_mesh(case)
    print("zorder %s: %d nodes %d elements" % (case, xyz.shape[0], conn.shape[0]), flush=True)
    no, eo, xn, cn, an = fvc.mesh_zorder(xyz, conn, adj)
    fx = os.path.join(cases.HERE, case)
    np.testing.assert_array_equal(xn, read_new(os.path.join(fx, "coordinates.bin.new"), "coordinates"))
    np.testing.assert_array_equal(cn, read_new(os.path.join(fx, "connectivity.bin.new"), "connectivity"))
    np.testing.assert_array_equal(an, read_new(os.path.join(fx, "adjacency.bin.new"), "adjacency"))


@pytest.mark.parametrize("dims", [(12, 12, 40), (60, 60, 330)])
def test_device_renumbering_matches_oracle_on_pipes(dims):
    t = host.pipe(*dims, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    print("zorder pipe %s: %d nodes %d elements" % (dims, xyz.shape[0], conn.shape[0]), flush=True)
    got = fvc.mesh_zorder(xyz, conn, adj)
    ref = oracle.mesh_zorder(xyz, conn, adj)
    assert ref[-1] == 0
    for g, r in zip(got, ref[:5]):
        np.testing.assert_array_equal(g, r)


def test_device_point_order_with_wide_ranks():
    """More than 2^21 distinct values on an axis: the 96-bit key path."""
    rng = np.random.default_rng(8)
    n = (1 << 21) + 5000
    pts = rng.uniform(-1.0, 1.0, size=(n, 3))
    pts[::7, 2] = np.round(pts[::7, 2], 2)  # ties on z
    empty = np.zeros((0, 4), np.int32)
    print("zorder points n=%d" % n, flush=True)
    no = fvc.mesh_zorder(pts, empty, empty)[0]
    ref = oracle.mesh_zorder(pts, empty, empty)
    assert ref[-1] == 0
    np.testing.assert_array_equal(no, ref[0])


def test_gather_rows_and_rejects_bad_input():
    rng = np.random.default_rng(1)
    data = rng.normal(size=(1000, 3))
    order = rng.permutation(1000).astype(np.int32)
    np.testing.assert_array_equal(fvc.gather_rows(order, data), data[order])
    order[5] = 1000
    with pytest.raises(fvc.FvcError):
        fvc.gather_rows(order, data)
    xyz, conn, adj = zg.make_mesh("zorder_pipe")
    bad = conn.copy()
    bad[0, 0] = -1
    with pytest.raises(fvc.FvcError):
        fvc.mesh_zorder(xyz, bad, adj)


def test_tool_writes_reference_files_and_velocity(tmp_path):
    xyz, conn, adj = zg.make_mesh("zorder_ybif")
    d = str(tmp_path)
    host.write_unstruct_mesh(d, "m", xyz, conn, adj)
    rng = np.random.default_rng(3)
    vel = [rng.normal(size=(xyz.shape[0], 3)) for _ in range(3)]
    for k, v in enumerate(vel):
        host.write_vel_slice(os.path.join(d, "m_vel.%d.bin" % (10 * k)), 0.5 * k, v)
    files = [os.path.join(d, "m_%s.bin" % k) for k in zg.MESH_FILES]
    tool = os.path.join(cases.REPO, "flowvc-1_amd", "fvc_zorder")
    print("fvc_zorder tool", flush=True)
    r = subprocess.run([tool] + files + [os.path.join(d, "m_vel"), "0", "20", "10"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    fx = os.path.join(cases.HERE, "zorder_ybif")
    for k, f in zip(zg.MESH_FILES, files):
        assert open(f + ".new", "rb").read() == open(os.path.join(fx, "%s.bin.new" % k), "rb").read(), k
    no = oracle.mesh_zorder(xyz, conn, adj)[0]
    for k, v in enumerate(vel):
        raw = open(os.path.join(d, "m_vel.%d.bin.new" % (10 * k)), "rb").read()
        assert np.frombuffer(raw[:8], np.float64)[0] == 0.5 * k
        np.testing.assert_array_equal(np.frombuffer(raw[8:], np.float64).reshape(-1, 3), v[no])


def test_trajectories_unchanged_by_renumbering():
    """FP64 tracers on the pipe before and after renumbering: identical
    positions and exits; element ids map through the element order."""
    t = host.pipe(12, 12, 40, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    slices = [host.tube_velocity(t, xyz, s * 0.02, 0.01, 5, s) for s in range(2)]
    no, eo, xn, cn, an = fvc.mesh_zorder(xyz, conn, adj)
    bounds = [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0]
    seeds = host.tube_seeds(t, 3000, 0.0, 10.0, 12)
    out = []
    for (X, C, A, S) in ((xyz, conn, adj, slices), (xn, cn, an, [s[no] for s in slices])):
        ctx = fvc.Context(3, bounds, "f64", 2, coords=X, conn=C, adj=A)
        for s in range(2):
            ctx.load_slice(s, s * 0.02, S[s])
        recs = ctx.seed_list(seeds, 1)
        P = fvc.Particles(ctx, recs.shape[0])
        P.set(recs)
        P.launch_velocity(0.0, 0, 1, 0.0, 0.02)
        P.advect(0.0, 0.02, 0, 1, 0.0, 0.02, 1e-3)
        out.append(P.get())
    a, b = out
    for k in ("X", "V", "LeftDomain", "LeftDomainTime"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    inv = np.empty_like(eo)
    inv[eo] = np.arange(eo.size)
    live = a["LeftDomain"] == 0
    np.testing.assert_array_equal(inv[a["ElementIndex"][live]], b["ElementIndex"][live])
