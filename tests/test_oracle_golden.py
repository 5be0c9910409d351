"""Pin the oracle (CPU FP64 restatement) against golden outputs of the
reference ompVC.  Tracer positions and element indices must be bit-identical;
FTLE values may differ only by the eigen-solver's rounding (the oracle uses
Jacobi rotations, the reference Householder+QL), so they are compared at
rtol 1e-12 with an absolute floor of 1e-12 on values divided by a short
exit time (early-exit points)."""
import os
import subprocess

import numpy as np
import pytest

import cases
from fvcpy import host


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
                                  "pipe_tracers_euler", "pipe_tracers_extrap", "box_tracers", "chan_tracers"])
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
                                  "box_ftle", "chan_ftle"])
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
