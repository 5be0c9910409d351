"""Pin the oracle where no reference output exists, through what the
mathematics fixes (no oracle formula is retyped here):

* largest eigenvalue of symmetric 3x3 matrices against numpy.linalg.eigvalsh;
* P1 interpolation reproduces fields linear in space and time exactly;
* bilinear Cartesian interpolation reproduces bilinear fields exactly;
* RK4 on a linear field equals the 4th-order Taylor polynomial of exp(hA)
  applied N times (the exact discrete map);
* FTLE of a shear flow equals the closed form ln(sigma_max)/T.
A plausible slip (wrong node order, dropped stage, transposed gradient, wrong
Taylor coefficient) fails at least one of these.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import cases
from fvcpy import deck, host

sys.path.insert(0, os.path.join(cases.REPO, "oracle"))
import oracle  # noqa: E402

PIPE_B = [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0]


@pytest.fixture(scope="module")
def pipe_mesh(oracle_built, host_built):
    t = host.pipe(10, 10, 30, 1.0, 10.0)
    return t, host.tube_mesh(t)


def test_max_eigenvalue_matches_eigvalsh(oracle_built):
    rng = np.random.default_rng(0)
    mats = []
    for k in range(400):
        B = rng.normal(size=(3, 3)) * 10 ** rng.uniform(-3, 3, size=(3, 1))
        if k % 4 == 0:
            B[2, :] = 0.0  # 2-D gradients: zero third row and column
            B[:, 2] = 0.0
        mats.append(B.T @ B)
    mats.append(np.diag([1.0, 1.0, 1.0]))
    mats.append(np.diag([1e12, 1e-12, 1.0]))
    mats = np.array(mats)
    got = oracle.max_eig(mats)
    want = np.linalg.eigvalsh(mats)[:, -1]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


def test_p1_reproduces_linear_fields(pipe_mesh):
    t, (xyz, conn, adj) = pipe_mesh
    A = np.array([[0.1, -0.3, 0.2], [0.05, 0.0, -0.1], [0.2, 0.1, 0.3]])
    b = np.array([0.5, -0.2, 1.0])
    f0 = xyz @ A.T + b
    f1 = xyz @ (2 * A).T - b
    om = oracle.Mesh(3, PIPE_B, coords=xyz, conn=conn, adj=adj)
    om.set_slot(0, f0)
    om.set_slot(1, f1)
    om.set_interval(1.0, 3.0)
    q = host.tube_seeds(t, 500, 0.5, 9.5, 5) * [0.9, 0.9, 1.0]
    pts, rc = om.seed_list(q, 1)
    assert rc == 0 and (pts["LeftDomain"] == 0).all()
    for tq in (1.0, 1.5, 3.0):
        tl = (tq - 1.0) / 2.0
        want = (1 - tl) * (q @ A.T + b) + tl * (q @ (2 * A).T - b)
        got, rc = om.velocity(pts, tq)
        assert rc == 0
        np.testing.assert_allclose(got, want, atol=1e-13, rtol=0)


def test_bilinear_cartesian_reproduction(oracle_built):
    nx, ny = 21, 11
    xs, ys = np.meshgrid(np.linspace(0, 2, nx), np.linspace(0, 1, ny))  # rows j, cols i
    f = lambda x, y: np.stack([0.3 + 0.2 * x - 0.7 * y + 0.4 * x * y, -0.1 * x + 0.5 * y - 0.2 * x * y,
                               np.zeros_like(x)], -1)
    uvw = f(xs, ys).reshape(-1, 3)  # k, j, i order with i fastest
    om = oracle.Mesh(2, [0, 2, 0, 1, 0, 0], cart_min=[0, 0, 0], cart_max=[2, 1, 0], cart_res=[nx, ny, 1])
    om.set_slot(0, uvw)
    om.set_slot(1, uvw)
    om.set_interval(0.0, 1.0)
    rng = np.random.default_rng(1)
    pts = np.zeros(300, oracle.POINT_DTYPE)
    pts["X"][:, 0] = rng.uniform(0, 2, 300)
    pts["X"][:, 1] = rng.uniform(0, 1, 300)
    got, rc = om.velocity(pts, 0.25)
    assert rc == 0
    np.testing.assert_allclose(got, f(pts["X"][:, 0], pts["X"][:, 1]), atol=1e-14)


def test_rk4_equals_exact_discrete_map(pipe_mesh):
    """Rotation about the pipe axis plus axial drift: u = A x + b stays inside."""
    t, (xyz, conn, adj) = pipe_mesh
    w = 2.0
    A = np.array([[0.0, -w, 0.0], [w, 0.0, 0.0], [0.0, 0.0, 0.0]])
    b = np.array([0.0, 0.0, 0.5])
    field = xyz @ A.T + b
    om = oracle.Mesh(3, PIPE_B, coords=xyz, conn=conn, adj=adj)
    om.set_slot(0, field)
    om.set_slot(1, field)
    om.set_interval(0.0, 1.0)
    q = host.tube_seeds(t, 200, 1.0, 4.0, 9) * [0.8, 0.8, 1.0]
    pts, rc = om.seed_list(q, 1)
    assert rc == 0
    h, n = 0.01, 100
    assert om.advect(pts, 0.0, 1.0, h) == 0
    hA = h * A
    T = np.eye(3) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    # affine part: x_{n+1} = T x_n + h (I + hA/2 + (hA)^2/6 + (hA)^3/24) b
    c = h * (np.eye(3) + hA / 2 + hA @ hA / 6 + hA @ hA @ hA / 24) @ b
    x = q.copy()
    for _ in range(n):
        x = x @ T.T + c
    np.testing.assert_allclose(pts["X"], x, atol=1e-11)
    assert (pts["LeftDomain"] == 0).all()


def test_ftle_of_shear_flow_closed_form(tmp_path, oracle_built):
    """u = g*y, v = 0: flow map gradient [[1, gT],[0,1]] (RK4 exact), FTLE =
    ln(sigma_max)/T with sigma_max^2 = 1 + s^2/2 + s*sqrt(1 + s^2/4), s = gT."""
    g, T = 0.5, 1.0
    d = str(tmp_path)
    nx, ny = 41, 21
    for s in range(3):
        xs, ys = np.meshgrid(np.linspace(0, 2, nx), np.linspace(0, 1, ny))
        uvw = np.stack([g * ys, np.zeros_like(ys), np.zeros_like(ys)], -1).reshape(-1, 3)
        host.write_vel_slice(os.path.join(d, "sh_vel.%d.bin" % s), s * 0.5, uvw)
    with open(os.path.join(d, "sh_Cartesian.bin"), "wb") as f:
        np.array([0.0, 2.0], "<f8").tofile(f); np.array([nx], "<i4").tofile(f)
        np.array([0.0, 1.0], "<f8").tofile(f); np.array([ny], "<i4").tofile(f)
        np.array([0.0, 0.0], "<f8").tofile(f); np.array([1], "<i4").tofile(f)
    deck.write_deck(os.path.join(d, "sh.in"), Path_Data=d + "/", Path_Output=d + "/", Dimensions=2,
                    Data_MeshType=0, Data_InFilePrefix="sh", Data_TRes=3, Data_TDelta=0.5, Data_TPeriodic=0,
                    Data_MeshBounds_XMin=0.0, Data_MeshBounds_XMax=2.0, Data_MeshBounds_YMin=0.0,
                    Data_MeshBounds_YMax=1.0, Data_MeshBounds_ZMin=0.0, Data_MeshBounds_ZMax=0.0, Output_TStart=0.0,
                    Output_TRes=1, Int_Type=1, Int_TimeStep=0.01, FTLE_Compute=1, FTLE_MeshBounds_XMin=0.2,
                    FTLE_MeshBounds_XMax=0.6, FTLE_MeshBounds_YMin=0.2, FTLE_MeshBounds_YMax=0.8,
                    FTLE_MeshBounds_XRes=9, FTLE_MeshBounds_YRes=13, FTLE_MeshBounds_ZRes=1, FTLE_IntTLength=T,
                    FTLE_OutFilePrefix="shf", FTLE_ICFile="shf_ic.bin")
    r = subprocess.run([os.path.join(oracle_built, "fvc_cpu"), os.path.join(d, "sh.in")], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    _, ftle = host.read_ftle_output(os.path.join(d, "shf.0.bin"), (9, 13, 1))
    s = g * T
    want = 0.5 * np.log(1 + s * s / 2 + s * np.sqrt(1 + s * s / 4)) / T
    np.testing.assert_allclose(ftle, want, rtol=1e-10)


def _rotation_case(pipe_mesh):
    t, (xyz, conn, adj) = pipe_mesh
    w = 2.0
    A = np.array([[0.0, -w, 0.0], [w, 0.0, 0.0], [0.0, 0.0, 0.0]])
    b = np.array([0.0, 0.0, 0.5])
    om = oracle.Mesh(3, PIPE_B, coords=xyz, conn=conn, adj=adj)
    om.set_slot(0, xyz @ A.T + b)
    om.set_slot(1, xyz @ A.T + b)
    om.set_interval(0.0, 1.0)
    q = host.tube_seeds(t, 100, 1.0, 4.0, 9) * [0.8, 0.8, 1.0]
    return om, A, b, q


def test_ab3_equals_exact_recurrence(pipe_mesh):
    """AB3 with two RK4 starting steps on u = A x + b, split over two calls
    (history carried): equals the recurrence applied in numpy."""
    om, A, b, q = _rotation_case(pipe_mesh)
    pts, rc = om.seed_list(q, 1)
    hist = np.zeros(q.shape[0], oracle.HIST_DTYPE)
    h = 0.01
    rc1, n1 = oracle.ab_advect(om, pts, hist, 0.0, 0.25, h, 3)
    rc2, n2 = oracle.ab_advect(om, pts, hist, 0.25, 0.5, h, 3)
    assert rc1 == 0 and rc2 == 0
    f = lambda x: x @ A.T + b
    hA = h * A
    T = np.eye(3) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    c = h * (np.eye(3) + hA / 2 + hA @ hA / 6 + hA @ hA @ hA / 24) @ b
    xs = [q.copy()]
    xs.append(xs[0] @ T.T + c)
    xs.append(xs[1] @ T.T + c)
    while len(xs) < 51:
        x2, x1, x0 = xs[-1], xs[-2], xs[-3]
        xs.append(x2 + h * (23 * f(x2) - 16 * f(x1) + 5 * f(x0)) / 12)
    np.testing.assert_allclose(pts["X"], xs[50], atol=1e-11)
    assert n1 + n2 == 50 * q.shape[0]


def test_ab3_is_third_order(pipe_mesh):
    om, A, b, q = _rotation_case(pipe_mesh)
    q = q[:20]
    errs = []
    for h in (0.02, 0.01, 0.005):
        pts, rc = om.seed_list(q, 1)
        hist = np.zeros(q.shape[0], oracle.HIST_DTYPE)
        assert oracle.ab_advect(om, pts, hist, 0.0, 1.0, h, 3)[0] == 0
        ang = 2.0 * 1.0
        R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
        exact = q @ R.T + [0, 0, 0.5]
        errs.append(np.abs(pts["X"] - exact).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 6.0 < r1 < 10.0 and 6.0 < r2 < 10.0, errs


def test_ab2_equals_exact_recurrence(pipe_mesh):
    """AB2 with one RK4 starting step on u = A x + b, split over two calls
    (history carried): equals x+ = x + h (3 f(x) - f(x-)) / 2 applied in numpy
    after the RK4 map of the first step."""
    om, A, b, q = _rotation_case(pipe_mesh)
    pts, rc = om.seed_list(q, 1)
    hist = np.zeros(q.shape[0], oracle.HIST_DTYPE)
    h = 0.01
    rc1, n1 = oracle.ab_advect(om, pts, hist, 0.0, 0.25, h, 2)
    rc2, n2 = oracle.ab_advect(om, pts, hist, 0.25, 0.5, h, 2)
    assert rc1 == 0 and rc2 == 0
    f = lambda x: x @ A.T + b
    hA = h * A
    T = np.eye(3) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    c = h * (np.eye(3) + hA / 2 + hA @ hA / 6 + hA @ hA @ hA / 24) @ b
    xs = [q.copy(), q @ T.T + c]
    while len(xs) < 51:
        x1, x0 = xs[-1], xs[-2]
        xs.append(x1 + h * (3 * f(x1) - f(x0)) / 2)
    np.testing.assert_allclose(pts["X"], xs[50], atol=1e-11)
    assert n1 + n2 == 50 * q.shape[0]
    # the recurrence is not AB3's, nor Euler's: both land far from it
    xe = q.copy()
    for _ in range(50):
        xe = xe + h * f(xe)
    assert np.abs(xe - xs[50]).max() > 1e-4


def test_ab2_is_second_order(pipe_mesh):
    """Errors against the exact rotation fall by about 4 when h halves."""
    om, A, b, q = _rotation_case(pipe_mesh)
    q = q[:20]
    errs = []
    for h in (0.02, 0.01, 0.005):
        pts, rc = om.seed_list(q, 1)
        hist = np.zeros(q.shape[0], oracle.HIST_DTYPE)
        assert oracle.ab_advect(om, pts, hist, 0.0, 1.0, h, 2)[0] == 0
        ang = 2.0 * 1.0
        R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
        exact = q @ R.T + [0, 0, 0.5]
        errs.append(np.abs(pts["X"] - exact).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.0 < r1 < 5.0 and 3.0 < r2 < 5.0, errs
