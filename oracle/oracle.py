"""ctypes binding of liboracle.so - TEST INFRASTRUCTURE ONLY.

The oracle is the checker and the timed CPU baseline; only tests/, the smoke
check and bench.py's cpu_baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

POINT_DTYPE = np.dtype([("X", "<f8", 3), ("V", "<f8", 3), ("ElementIndex", "<i4"), ("AuxElementIndex", "<i4"),
                        ("LeftDomain", "<i4"), ("pad", "<i4"), ("LeftDomainTime", "<f8"), ("Scalar", "<f8")])

_LIB = None


class OcInt(ctypes.Structure):
    _fields_ = [("int_type", ctypes.c_int32), ("time_step", ctypes.c_double), ("extrapolate", ctypes.c_int32),
                ("min_step", ctypes.c_double), ("max_step", ctypes.c_double), ("accuracy", ctypes.c_double)]


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.oc_mesh_cartesian.restype = vp
        L.oc_mesh_cartesian.argtypes = [ctypes.c_int, vp, vp, vp, vp]
        L.oc_mesh_unstructured.restype = vp
        L.oc_mesh_unstructured.argtypes = [ctypes.c_int, vp, i64, vp, i64, vp, vp]
        L.oc_mesh_free.argtypes = [vp]
        L.oc_mesh_set_slot.argtypes = [vp, ctypes.c_int, vp]
        L.oc_mesh_set_interval.argtypes = [vp, d, d]
        L.oc_locate_local_batch.argtypes = [vp, i64, vp, vp, vp]
        L.oc_locate_global_batch.argtypes = [vp, i64, vp, vp, ctypes.c_int]
        L.oc_advect_batch.argtypes = [vp, ctypes.POINTER(OcInt), i64, vp, d, d, ctypes.c_int, i64, ctypes.POINTER(i64)]
        L.oc_velocity_batch.argtypes = [vp, d, i64, vp, vp]
        L.oc_eig_batch.argtypes = [i64, vp, vp]
        L.oc_seed_list.argtypes = [vp, i64, vp, ctypes.c_int, vp]
        L.oc_seed_grid.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp]
        L.oc_mesh_zorder.argtypes = [i64, vp, i64, vp, vp, vp, vp, vp, vp, vp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def mesh_zorder(xyz, conn, adj):
    """Z-order renumbering (cudaVC/ZOrderSort/main.cpp:103-259): returns
    (node_order, elem_order, xyz_new, conn_new, adj_new, rc)."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    conn = np.ascontiguousarray(conn, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    nn, ne = xyz.shape[0], conn.shape[0]
    no, eo = np.zeros(nn, np.int32), np.zeros(ne, np.int32)
    xn, cn, an = np.zeros_like(xyz), np.zeros_like(conn), np.zeros_like(adj)
    rc = lib().oc_mesh_zorder(nn, _p(xyz), ne, _p(conn), _p(adj), _p(no), _p(eo), _p(xn), _p(cn), _p(an))
    return no, eo, xn, cn, an, rc


class Mesh:
    def __init__(self, dims, bounds, coords=None, conn=None, adj=None, cart_min=None, cart_max=None, cart_res=None):
        L = lib()
        b = np.asarray(bounds, np.float64)
        if coords is not None:
            c = np.ascontiguousarray(coords, np.float64)
            cn = np.ascontiguousarray(conn, np.int32)
            ad = np.ascontiguousarray(adj, np.int32)
            self.h = L.oc_mesh_unstructured(dims, _p(b), c.shape[0], _p(c), cn.shape[0], _p(cn), _p(ad))
            self.nvel = c.shape[0]
        else:
            mn, mx = np.asarray(cart_min, np.float64), np.asarray(cart_max, np.float64)
            rs = np.asarray(cart_res, np.int32)
            self.h = L.oc_mesh_cartesian(dims, _p(b), _p(mn), _p(mx), _p(rs))
            self.nvel = int(rs[0] * rs[1] * (rs[2] if dims == 3 else 1))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_mesh_free(self.h)
            self.h = None

    def set_slot(self, slot, uvw):
        uvw = np.ascontiguousarray(uvw, np.float64)
        assert uvw.shape[0] == self.nvel
        lib().oc_mesh_set_slot(self.h, slot, _p(uvw))

    def set_interval(self, tmin, tmax):
        lib().oc_mesh_set_interval(self.h, tmin, tmax)

    def locate_local(self, xyz, guess):
        xyz = np.ascontiguousarray(xyz, np.float64)
        g = np.ascontiguousarray(guess, np.int32)
        out = np.empty(xyz.shape[0], np.int32)
        rc = lib().oc_locate_local_batch(self.h, xyz.shape[0], _p(xyz), _p(g), _p(out))
        return out, rc

    def locate_global(self, xyz, nthreads=0):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.empty(xyz.shape[0], np.int32)
        lib().oc_locate_global_batch(self.h, xyz.shape[0], _p(xyz), _p(out), nthreads)
        return out

    def advect(self, pts, t1, t2, dt, int_type=1, extrapolate=0, nthreads=1, chunk=64):
        """Advect records in place (the LOWERLOOP filter applies); returns the error code."""
        assert pts.dtype == POINT_DTYPE and pts.flags.c_contiguous
        ip = OcInt(int_type, dt, extrapolate)
        self.last_steps = ctypes.c_int64(0)
        return lib().oc_advect_batch(self.h, ctypes.byref(ip), pts.shape[0], _p(pts), t1, t2, nthreads, chunk,
                                     ctypes.byref(self.last_steps))

    def velocity(self, pts, tq):
        out = np.zeros((pts.shape[0], 3))
        rc = lib().oc_velocity_batch(self.h, tq, pts.shape[0], _p(pts), _p(out))
        return out, rc

    def seed_grid(self, mn, dl, res, lsc=1):
        """GenerateTracerMesh chain (tracers.c:1298-1532)."""
        mn, dl = np.asarray(mn, np.float64), np.asarray(dl, np.float64)
        res = np.asarray(res, np.int32)
        out = np.zeros(int(np.prod(res)), POINT_DTYPE)
        rc = lib().oc_seed_grid(self.h, _p(mn), _p(dl), _p(res), lsc, _p(out))
        return out, rc

    def ftle_grid(self, mn, mx, res, lsc=1):
        """InitializeFTLEArray's grid records (ftle.c:383-609), grid index (i*res1 + j)*res2 + k."""
        mn, mx = np.asarray(mn, np.float64), np.asarray(mx, np.float64)
        res = np.asarray(res, np.int32)
        err = ctypes.c_int(0)
        L = lib()
        L.oc_ftle_create.restype = ctypes.c_void_p
        L.oc_ftle_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int)]
        L.oc_ftle_free.argtypes = [ctypes.c_void_p]
        L.oc_ftle_points.restype = ctypes.c_void_p
        L.oc_ftle_points.argtypes = [ctypes.c_void_p]
        f = L.oc_ftle_create(self.h, _p(mn), _p(mx), _p(res), lsc, ctypes.byref(err))
        if not f:
            raise RuntimeError("oc_ftle_create failed")
        n = int(np.prod(res))
        buf = (ctypes.c_char * (104 * n)).from_address(L.oc_ftle_points(f))
        rec = np.frombuffer(bytes(buf), dtype=FTLE_PT_DTYPE).copy()
        L.oc_ftle_free(f)
        return rec, err.value

    def seed_list(self, xyz, lsc=1):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.zeros(xyz.shape[0], POINT_DTYPE)
        rc = lib().oc_seed_list(self.h, xyz.shape[0], _p(xyz), lsc, _p(out))
        return out, rc


def max_eig(mats):
    mats = np.ascontiguousarray(mats, np.float64).reshape(-1, 9)
    out = np.empty(mats.shape[0])
    lib().oc_eig_batch(mats.shape[0], _p(mats), _p(out))
    return out


FTLE_PT_DTYPE = np.dtype([("Pt", POINT_DTYPE), ("HaveFTLE", "<i4"), ("pad", "<i4"), ("FTLEwT", "<f8"),
                          ("FTLEwoT", "<f8")])

HIST_DTYPE = np.dtype([("F1", "<f8", 3), ("F2", "<f8", 3), ("nhist", "<i4"), ("pad", "<i4")])


def _ab_lib():
    L = lib()
    if not getattr(L, "_ab_ready", False):
        L.oc_ab_batch.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int64)]
        L._ab_ready = True
    return L


def ab_advect(mesh, pts, hist, t1, t2, dt, order=3, nthreads=1):
    """Adams-Bashforth (order 2/3) over [t1, t2]; pts and hist updated in place."""
    assert pts.dtype == POINT_DTYPE and hist.dtype == HIST_DTYPE
    steps = ctypes.c_int64(0)
    rc = _ab_lib().oc_ab_batch(mesh.h, pts.shape[0], _p(pts), _p(hist), t1, t2, dt, order, nthreads,
                               ctypes.byref(steps))
    return rc, steps.value
