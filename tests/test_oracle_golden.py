"""Pin the oracle (CPU FP64 restatement) against golden outputs of the
reference ompVC.  Tracer positions and element indices must be bit-identical;
FTLE values may differ only by the eigen-solver's rounding (the oracle uses
Jacobi rotations, the reference Householder+QL), so they are compared at
rtol 1e-12 with an absolute floor of 1e-12 on values divided by a short
exit time (early-exit points)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import cases
from fvcpy import host

sys.path.insert(0, os.path.join(cases.REPO, "oracle"))
import oracle  # noqa: E402


def run_oracle(case, tmp_path, oracle_dir):
    deck_path, data, out = cases.prepare(case, str(tmp_path))
    man = cases.load_manifest(case)
    got = cases.input_digests(data)
    for f, h in man["inputs"].items():
        assert got.get(f) == h, "regenerated input %s differs from the fixture's" % f
    r = subprocess.run([os.path.join(oracle_dir, "fvc_cpu"), deck_path, str(man["nproc"])],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return man, data, out


@pytest.mark.parametrize("case", ["dg_tracers", "pipe_tracers", "pipe_locate", "dg_tracers_euler",
                                  "pipe_tracers_euler", "pipe_tracers_extrap", "box_tracers", "chan_tracers",
                                  "pipe_tracers_grid", "dg_tracers_back", "pipe_tracers_back",
                                  "pipe_tracers_periods", "dg_tracers_rkf", "pipe_tracers_rkf", "sten_locate",
                                  "sten_tracers"])
def test_trajectories_bit_identical(case, tmp_path, oracle_built, host_built):
    man, data, out = run_oracle(case, tmp_path, oracle_built)
    fx = cases.fixture_dir(case)
    assert sorted(f for f in os.listdir(out) if not f.endswith(".info")) == sorted(man["outputs"])
    for f in man["outputs"]:
        a = np.fromfile(os.path.join(fx, f), np.uint8)
        b = np.fromfile(os.path.join(out, f), np.uint8)
        assert a.size == b.size and np.array_equal(a, b), f
    if man.get("ic"):
        ra = host.read_ic(os.path.join(fx, man["ic"]))
        rb = host.read_ic(os.path.join(data, man["ic"]))
        np.testing.assert_array_equal(ra["ElementIndex"][ra["LeftDomain"] == 0], rb["ElementIndex"][rb["LeftDomain"] == 0])
        np.testing.assert_array_equal(ra["LeftDomain"], rb["LeftDomain"])
        np.testing.assert_array_equal(ra["X"], rb["X"])


@pytest.mark.parametrize("case", ["dg_ftle", "dg_ftle_back", "pipe_ftle", "dg_ftle_euler", "pipe_ftle_extrap",
                                  "box_ftle", "chan_ftle", "pipe_ftle_vertices",
                                  "dg_ftle_rkf"])
def test_ftle_matches_reference(case, tmp_path, oracle_built, host_built):
    man, data, out = run_oracle(case, tmp_path, oracle_built)
    fx = cases.fixture_dir(case)
    assert sorted(f for f in os.listdir(out) if not f.endswith(".info")) == sorted(man["outputs"])
    for f in man["outputs"]:
        a = np.fromfile(os.path.join(fx, f))
        b = np.fromfile(os.path.join(out, f))
        if f.endswith("_Cartesian.bin"):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
            continue
        assert a.shape == b.shape
        # masks (-1 outside, 0 for neighbours/non-stretching) are exact
        np.testing.assert_array_equal(a == -1.0, b == -1.0)
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12)


def test_locate_probes_cover_special_points(tmp_path, oracle_built, host_built):
    """The location fixture holds points on faces/edges/vertices and outside."""
    man = cases.load_manifest("pipe_locate")
    ic = host.read_ic(os.path.join(cases.fixture_dir("pipe_locate"), man["ic"]))
    assert ic.size > 60
    assert ic["LeftDomain"].sum() >= 3          # outside probes
    assert (ic["ElementIndex"][ic["LeftDomain"] == 0] >= 0).all()


FTLE_IC = np.dtype({"names": ["X", "ElementIndex", "LeftDomain", "HaveFTLE", "FTLEwT"],
                    "formats": [("<f8", 3), "<i4", "<i4", "<i4", "<f8"], "offsets": [0, 48, 56, 80, 88],
                    "itemsize": 104})


def _reference_ftle_grid(case):
    raw = open(os.path.join(cases.fixture_dir(case), cases.CASES[case]["keep"][0]), "rb").read()
    return np.frombuffer(raw[:12], "<i4"), np.frombuffer(raw[12:], FTLE_IC)


def _pipe_oracle():
    t = host.pipe(*cases.PIPE_DIMS, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(t)
    return oracle.Mesh(3, [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0], coords=xyz, conn=conn, adj=adj)


@pytest.mark.parametrize("case", ["pipe_ftle", "pipe_ftle_vertices"])
def test_ftle_grid_chain_matches_reference_records(case, oracle_built, host_built):
    """The FTLE grid's location chain (ftle.c:456-601) against the grid
    records the reference stores in FTLE_ICFile: masks, HaveFTLE and the
    element id of every inside point, bit-identical."""
    res, ref = _reference_ftle_grid(case)
    d = cases.CASES[case]["deck"]
    mn = [d["FTLE_MeshBounds_XMin"], d["FTLE_MeshBounds_YMin"], d["FTLE_MeshBounds_ZMin"]]
    mx = [d["FTLE_MeshBounds_XMax"], d["FTLE_MeshBounds_YMax"], d["FTLE_MeshBounds_ZMax"]]
    got, rc = _pipe_oracle().ftle_grid(mn, mx, list(res), d["LocalSearchChecking"])
    assert rc == 0
    np.testing.assert_array_equal(got["Pt"]["X"], ref["X"])
    np.testing.assert_array_equal(got["Pt"]["LeftDomain"], ref["LeftDomain"])
    np.testing.assert_array_equal(got["HaveFTLE"], ref["HaveFTLE"])
    inside = ref["LeftDomain"] == 0
    assert inside.sum() > 500
    np.testing.assert_array_equal(got["Pt"]["ElementIndex"][inside], ref["ElementIndex"][inside])


def test_vertex_grid_fixture_separates_guess_reset_rules(oracle_built, host_built):
    """The chain replayed point by point with the oracle's walk and global
    search under two rules: the guess reset after each plane to the first
    found element of its last row (ftle.c:597-598), and a reset after every
    row as the velocity output grid does (velout.c:158-162).  The reference's
    records follow the first and differ from the second, so the fixture
    catches a chain that resets per row."""
    res, ref = _reference_ftle_grid("pipe_ftle_vertices")
    om = _pipe_oracle()
    X = ref["X"].reshape(tuple(res) + (3,))
    inside = (ref["LeftDomain"] == 0).reshape(tuple(res))

    def walk(x, g):
        e, rc = om.locate_local(x[None, :], [g])
        assert rc == 0
        return int(e[0])

    def chain(per_row):
        seed = int(om.locate_global(X[res[0] // 2, res[1] // 2, res[2] // 2][None, :])[0])
        assert seed >= 0
        guess, out = seed, np.full(tuple(res), -1, np.int64)
        for i in range(res[0]):
            for j in range(res[1]):
                jguess = -1
                for k in range(res[2]):
                    if not inside[i, j, k]:
                        continue
                    e = walk(X[i, j, k], guess)
                    if e < 0:
                        e = walk(X[i, j, k], seed)
                    if e < 0:
                        e = int(om.locate_global(X[i, j, k][None, :])[0])
                    assert e >= 0
                    out[i, j, k] = guess = e
                    jguess = e if jguess < 0 else jguess
                if per_row and jguess >= 0:
                    guess = jguess
            if jguess >= 0:
                guess = jguess
        return out.reshape(-1)

    refe = ref["ElementIndex"].astype(np.int64)
    ins = inside.reshape(-1)
    np.testing.assert_array_equal(chain(False)[ins], refe[ins])
    assert (chain(True)[ins] != refe[ins]).sum() > 0
