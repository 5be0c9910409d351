"""ctypes bindings of libfvc_host.so: formats and deterministic synthetic inputs.

No GPU code is loaded here.  Arrays are numpy; the C side fills caller buffers.
"""
import ctypes
import os

import numpy as np

from . import PKG_DIR

_LIB = None


class Tube(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32), ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
                ("R", ctypes.c_double), ("L", ctypes.c_double), ("rd", ctypes.c_double),
                ("ang_deg", ctypes.c_double), ("Ld", ctypes.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(PKG_DIR, "libfvc_host.so")
        if not os.path.exists(path):
            raise RuntimeError("libfvc_host.so is not built; run `make -C flowvc-1_amd`")
        L = ctypes.CDLL(path)
        d, i32, i64, u64, vp = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p
        L.fvc_tube_pipe.argtypes = [ctypes.POINTER(Tube), ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d]
        L.fvc_tube_ybif.argtypes = [ctypes.POINTER(Tube), ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, d, d, d]
        L.fvc_tube_count.argtypes = [ctypes.POINTER(Tube), ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.fvc_tube_fill.argtypes = [ctypes.POINTER(Tube), vp, vp, vp]
        L.fvc_tube_velocity.argtypes = [ctypes.POINTER(Tube), i64, vp, d, d, u64, i32, vp]
        L.fvc_tube_seeds.argtypes = [ctypes.POINTER(Tube), i64, d, d, u64, vp]
        L.fvc_tube_inside.argtypes = [ctypes.POINTER(Tube), d, d, d]
        L.fvc_gen_double_gyre_slice.argtypes = [ctypes.c_int, ctypes.c_int, d, d, d, d, vp]
        L.fvc_gen_double_gyre_files.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, d, d, d, d, ctypes.c_int, ctypes.c_int]
        L.fvc_write_unstruct_mesh.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i64, vp, i64, vp, vp]
        L.fvc_write_vel_slice.argtypes = [ctypes.c_char_p, d, vp, i64]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def pipe(nx=60, ny=60, nz=330, R=1.0, L=10.0):
    t = Tube()
    lib().fvc_tube_pipe(ctypes.byref(t), nx, ny, nz, R, L)
    return t


def ybif(nx, ny, nz, R=1.0, rd=0.7, ang_deg=30.0, L=4.0, Ld=4.0):
    t = Tube()
    lib().fvc_tube_ybif(ctypes.byref(t), nx, ny, nz, R, rd, ang_deg, L, Ld)
    return t


def tube_mesh(t):
    """Return (xyz[nn,3] f64, conn[ne,4] i32, adj[ne,4] i32)."""
    nn, ne = ctypes.c_int64(), ctypes.c_int64()
    rc = lib().fvc_tube_count(ctypes.byref(t), ctypes.byref(nn), ctypes.byref(ne))
    if rc:
        raise RuntimeError("fvc_tube_count failed (%d)" % rc)
    xyz = np.zeros((nn.value, 3), np.float64)
    conn = np.zeros((ne.value, 4), np.int32)
    adj = np.zeros((ne.value, 4), np.int32)
    rc = lib().fvc_tube_fill(ctypes.byref(t), _p(xyz), _p(conn), _p(adj))
    if rc:
        raise RuntimeError("fvc_tube_fill failed (%d)" % rc)
    return xyz, conn, adj


def tube_velocity(t, xyz, time, sigma=0.01, seed=1234, slice_index=0):
    xyz = np.ascontiguousarray(xyz, np.float64)
    uvw = np.zeros_like(xyz)
    lib().fvc_tube_velocity(ctypes.byref(t), xyz.shape[0], _p(xyz), time, sigma, seed, slice_index, _p(uvw))
    return uvw


def tube_seeds(t, n, z0, z1, seed=42):
    xyz = np.zeros((n, 3), np.float64)
    lib().fvc_tube_seeds(ctypes.byref(t), n, z0, z1, seed, _p(xyz))
    return xyz


def double_gyre_slice(nx, ny, t, A=0.1, eps=0.25, omega=2 * np.pi / 10):
    uvw = np.zeros((nx * ny, 3), np.float64)
    lib().fvc_gen_double_gyre_slice(nx, ny, t, A, eps, omega, _p(uvw))
    return uvw


def write_double_gyre(dirpath, prefix, nx, ny, nt, tdelta, A=0.1, eps=0.25, omega=2 * np.pi / 10,
                      suffix_min=0, suffix_delta=1):
    rc = lib().fvc_gen_double_gyre_files(dirpath.encode(), prefix.encode(), nx, ny, nt, tdelta, A, eps, omega,
                                         suffix_min, suffix_delta)
    if rc:
        raise RuntimeError("double gyre generation failed (%d)" % rc)


def write_unstruct_mesh(dirpath, prefix, xyz, conn, adj):
    xyz = np.ascontiguousarray(xyz, np.float64)
    conn = np.ascontiguousarray(conn, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    rc = lib().fvc_write_unstruct_mesh(dirpath.encode(), prefix.encode(), xyz.shape[0], _p(xyz), conn.shape[0],
                                       _p(conn), _p(adj))
    if rc:
        raise RuntimeError("mesh write failed")


def write_vel_slice(path, t, uvw):
    uvw = np.ascontiguousarray(uvw, np.float64)
    rc = lib().fvc_write_vel_slice(path.encode(), t, _p(uvw), uvw.shape[0])
    if rc:
        raise RuntimeError("slice write failed: %s" % path)


def write_binary_list(path, xyz, dims=3):
    xyz = np.ascontiguousarray(xyz, np.float64)
    with open(path, "wb") as f:
        np.array([xyz.shape[0]], "<i4").tofile(f)
        np.ascontiguousarray(xyz[:, :dims]).tofile(f)


POINT_DTYPE = np.dtype([("X", "<f8", 3), ("V", "<f8", 3), ("ElementIndex", "<i4"), ("AuxElementIndex", "<i4"),
                        ("LeftDomain", "<i4"), ("pad", "<i4"), ("LeftDomainTime", "<f8"), ("Scalar", "<f8")])
assert POINT_DTYPE.itemsize == 80

FTLE_DTYPE = np.dtype([("Pt", POINT_DTYPE), ("HaveFTLE", "<i4"), ("pad", "<i4"), ("FTLEwT", "<f8"),
                       ("FTLEwoT", "<f8")])
assert FTLE_DTYPE.itemsize == 104


def read_ic(path):
    """Tracer .IC file: int32 N then N LagrangianPoint records."""
    with open(path, "rb") as f:
        n = int(np.fromfile(f, "<i4", 1)[0])
        return np.fromfile(f, POINT_DTYPE, n)


def read_tracer_output(path, scalar=False):
    """<prefix>.<k>.bin: float64 time, then records of 3 (+1) float64."""
    raw = np.fromfile(path, "<f8")
    w = 4 if scalar else 3
    return raw[0], raw[1:].reshape(-1, w)


def read_ftle_output(path, res):
    """<prefix>.<n>.bin: float64 start time, then values in k,j,i order (i fastest)."""
    raw = np.fromfile(path, "<f8")
    return raw[0], raw[1:].reshape(res[2], res[1], res[0])
