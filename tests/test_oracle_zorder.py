"""Pin the oracle's mesh Z-order renumbering (restatement of
cudaVC/ZOrderSort/main.cpp:103-259) against the reference tool's own output
files, against a brute-force sort with the reference comparator, and against
the invariants a renumbering must keep (same geometry per element, symmetric
adjacency, permutations)."""
import functools
import json
import os
import sys

import numpy as np
import pytest

import cases
import make_zorder_goldens as zg

sys.path.insert(0, os.path.join(cases.REPO, "oracle"))
import oracle  # noqa: E402


def read_new(fx, kind):
    raw = open(os.path.join(fx, "%s.bin.new" % kind), "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    if kind == "coordinates":
        return np.frombuffer(raw[4:], np.float64).reshape(n, 3)
    return np.frombuffer(raw[4:], np.int32).reshape(n, 4)


@pytest.mark.parametrize("case", sorted(zg.ZCASES))
def test_oracle_matches_reference_files(case, oracle_built, host_built, tmp_path):
    fx = os.path.join(cases.HERE, case)
    man = json.load(open(os.path.join(fx, "manifest.json")))
    xyz, conn, adj = zg.make_mesh(case)
    from fvcpy import host
    host.write_unstruct_mesh(str(tmp_path), "m", xyz, conn, adj)
    for k in zg.MESH_FILES:
        assert zg.digest(os.path.join(str(tmp_path), "m_%s.bin" % k)) == man["inputs"][k], k
    no, eo, xn, cn, an, rc = oracle.mesh_zorder(xyz, conn, adj)
    assert rc == 0
    np.testing.assert_array_equal(xn, read_new(fx, "coordinates"))
    np.testing.assert_array_equal(cn, read_new(fx, "connectivity"))
    np.testing.assert_array_equal(an, read_new(fx, "adjacency"))


def _zless(a, b):
    """main.cpp:86-101, on integer triples."""
    j, x = 0, 0
    for k in range(3):
        y = int(a[k]) ^ int(b[k])
        if x < y and x < (x ^ y):
            j, x = k, y
    return a[j] < b[j]


def test_point_order_equals_brute_force_comparator_sort(oracle_built):
    """Points on a small integer lattice (ties in every axis, negative and
    -0.0 coordinates): the oracle's order equals a comparison sort with the
    reference comparator on lower_bound ranks computed directly."""
    rng = np.random.default_rng(4)
    pts = rng.integers(-6, 7, size=(600, 3)).astype(np.float64) * 0.5
    pts[:5, 0] = -0.0
    pts = np.unique(pts, axis=0)[rng.permutation(np.unique(pts, axis=0).shape[0])]
    conn = np.zeros((0, 4), np.int32)
    no, eo, xn, cn, an, rc = oracle.mesh_zorder(pts, conn, conn)
    assert rc == 0
    ranks = np.stack([np.searchsorted(np.unique(pts[:, a]), pts[:, a]) for a in range(3)], 1)

    def cmp(i, j):
        return -1 if _zless(ranks[i], ranks[j]) else (1 if _zless(ranks[j], ranks[i]) else 0)

    brute = sorted(range(pts.shape[0]), key=functools.cmp_to_key(cmp))
    np.testing.assert_array_equal(no, brute)
    np.testing.assert_array_equal(xn, pts[no])


def test_renumbering_keeps_geometry_and_adjacency(oracle_built, host_built):
    xyz, conn, adj = zg.make_mesh("zorder_ybif")
    no, eo, xn, cn, an, rc = oracle.mesh_zorder(xyz, conn, adj)
    assert rc == 0
    assert sorted(no.tolist()) == list(range(xyz.shape[0]))
    assert sorted(eo.tolist()) == list(range(conn.shape[0]))
    # element i of the new mesh is old element eo[i], with the same corner coordinates in the same order
    np.testing.assert_array_equal(xn[cn], xyz[conn[eo]])
    # adjacency stays symmetric and boundary faces stay boundary faces
    assert ((an == -1) == (adj[eo] == -1)).all()
    e, q = np.nonzero(an >= 0)
    assert (an[an[e, q]] == e[:, None]).any(1).all()
    # neighbours are the renumbered old neighbours
    inv = np.empty_like(eo)
    inv[eo] = np.arange(eo.size)
    np.testing.assert_array_equal(an[an >= 0], inv[adj[eo][an >= 0]])


def test_rejects_out_of_range_ids(oracle_built, host_built):
    xyz, conn, adj = zg.make_mesh("zorder_pipe")
    bad = conn.copy()
    bad[3, 2] = xyz.shape[0]
    assert oracle.mesh_zorder(xyz, bad, adj)[-1] != 0
    bad = adj.copy()
    bad[0, 0] = conn.shape[0]
    assert oracle.mesh_zorder(xyz, conn, bad)[-1] != 0
