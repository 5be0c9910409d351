"""ctypes bindings of libfvc.so, the C ABI of include/fvc.h (HIP path on gfx950).

Loading fails loudly when the library is not built; there is no CPU fallback.
"""
import ctypes
import os
import weakref

import numpy as np

from . import PKG_DIR
from .host import POINT_DTYPE

OK = 0
E_INVALID, E_NOMEM, E_HIP, E_INDETERMINATE_NEIGHBOR, E_TLOC_RANGE, E_ELEM_NEG1, E_WALK_LIMIT = -1, -2, -3, -4, -5, -6, -7
E_ADMISSION, E_STATE = -9, -10
NOALLOC_MAX_POINTS = 1 << 18  # FVC_NOALLOC_MAX_POINTS (include/fvc.h)
PREC = {"f64": 0, "f32": 1, "bf16": 2, "f16": 3}
LOCATE_WALK, LOCATE_GLOBAL, LOCATE_WALK_GLOBAL = 0, 1, 2
INT_EULER, INT_RK4, INT_RKF, INT_AB2, INT_AB3 = 0, 1, 2, 3, 4

_LIB = None


class MeshDesc(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32), ("mesh_type", ctypes.c_int32), ("bounds", ctypes.c_double * 6),
                ("cart_min", ctypes.c_double * 3), ("cart_max", ctypes.c_double * 3), ("cart_res", ctypes.c_int32 * 3),
                ("num_nodes", ctypes.c_int64), ("coords", ctypes.c_void_p), ("num_elems", ctypes.c_int64),
                ("conn", ctypes.c_void_p), ("adj", ctypes.c_void_p)]


class Params(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("max_slices", ctypes.c_int32), ("device", ctypes.c_int32),
                ("hash_cells", ctypes.c_int32), ("walk_limit", ctypes.c_int32)]


class IntParams(ctypes.Structure):
    _fields_ = [("int_type", ctypes.c_int32), ("extrapolate", ctypes.c_int32), ("time_step", ctypes.c_double),
                ("block", ctypes.c_int32), ("steps_per_launch", ctypes.c_int32),
                ("activation_potential", ctypes.c_int32), ("exposure_substeps", ctypes.c_int32),
                ("rkf_min_step", ctypes.c_double), ("rkf_max_step", ctypes.c_double), ("rkf_accuracy", ctypes.c_double)]


class Release(ctypes.Structure):
    _fields_ = [("begin", ctypes.c_int64), ("count", ctypes.c_int64), ("t_start", ctypes.c_double),
                ("t_stop", ctypes.c_double), ("int_type", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class BatchParams(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("steps_per_launch", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32)]


class ServerParams(ctypes.Structure):
    _fields_ = [("frame_t0", ctypes.c_double), ("frame_dt", ctypes.c_double), ("nslots", ctypes.c_int32),
                ("time_step", ctypes.c_double), ("pool_particles", ctypes.c_int64), ("hbm_fraction", ctypes.c_double),
                ("max_requests", ctypes.c_int32), ("block", ctypes.c_int32), ("steps_per_launch", ctypes.c_int32),
                ("lsc", ctypes.c_int32), ("no_graph", ctypes.c_int32)]


class RequestDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("int_type", ctypes.c_int32), ("t_start", ctypes.c_double),
                ("t_stop", ctypes.c_double), ("n", ctypes.c_int64), ("xyz", ctypes.c_void_p),
                ("ftle_min", ctypes.c_double * 3), ("ftle_max", ctypes.c_double * 3), ("ftle_res", ctypes.c_int32 * 3),
                ("pad_", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("kind", ctypes.c_int32), ("n", ctypes.c_int64),
                ("particle_steps", ctypes.c_int64), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]


REQ_TRACER, REQ_FTLE = 0, 1
RS_QUEUED, RS_RUNNING, RS_DONE, RS_FAILED = 0, 1, 2, 3


class FvcError(RuntimeError):
    def __init__(self, code, what, detail=""):
        self.code = code
        super().__init__("%s failed: %s (%d) %s" % (what, lib().fvc_strerror(code).decode(), code, detail))


def lib_path():
    return os.path.join(PKG_DIR, "libfvc.so")


def lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("libfvc.so is not built (run `make -C flowvc-1_amd`); the HIP path has no fallback")
        L = ctypes.CDLL(path)
        vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.fvc_strerror.restype = ctypes.c_char_p
        L.fvc_strerror.argtypes = [ctypes.c_int]
        L.fvc_last_error.restype = ctypes.c_char_p
        L.fvc_last_error.argtypes = [vp]
        L.fvc_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.fvc_destroy.argtypes = [vp]
        L.fvc_sync.argtypes = [vp]
        L.fvc_memory.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.fvc_load_slice.argtypes = [vp, i32, d, vp, i64]
        L.fvc_particles_create.argtypes = [vp, i64, ctypes.POINTER(vp)]
        L.fvc_particles_destroy.argtypes = [vp]
        L.fvc_particles_set.argtypes = [vp, i64, vp, i64]
        L.fvc_particles_get.argtypes = [vp, i64, vp, i64]
        L.fvc_particles_capacity.restype = i64
        L.fvc_particles_capacity.argtypes = [vp]
        L.fvc_particles_active.argtypes = [vp, ctypes.POINTER(i64)]
        L.fvc_launch_velocity.argtypes = [vp, vp, i64, i64, d, i32, i32, d, d]
        L.fvc_advect.argtypes = [vp, vp, i64, i64, d, d, i32, i32, d, d, ctypes.POINTER(IntParams)]
        L.fvc_last_step_count.argtypes = [vp, ctypes.POINTER(i64)]
        L.fvc_locate.argtypes = [vp, vp, vp, i64, i32, vp]
        L.fvc_seed_list.argtypes = [vp, vp, i64, i32, vp]
        L.fvc_seed_grid.argtypes = [vp, vp, vp, vp, i32, vp]
        L.fvc_ftle_create.argtypes = [vp, vp, vp, vp, i32, d, d, ctypes.POINTER(vp)]
        L.fvc_ftle_destroy.argtypes = [vp]
        L.fvc_ftle_frame.argtypes = [vp, i32, i32, i32, d, d, ctypes.POINTER(IntParams), ctypes.POINTER(i32)]
        L.fvc_ftle_output.argtypes = [vp, d, vp, vp]
        L.fvc_ftle_points.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.fvc_ftle_records.argtypes = [vp, vp]
        L.fvc_ftle_initial.argtypes = [vp, vp, vp, vp, vp]
        L.fvc_ftle_create_from.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, d, d, ctypes.POINTER(vp)]
        L.fvc_zsort.argtypes = [vp, vp, i64, vp, vp, vp]
        L.fvc_particles_copy.argtypes = [vp, i64, vp, i64, i64]
        L.fvc_particles_get_xyz.argtypes = [vp, i64, vp, vp, i64]
        L.fvc_ftle_reset.argtypes = [vp]
        L.fvc_set_timing.argtypes = [vp, i32]
        L.fvc_kernel_time.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(i64), ctypes.POINTER(i64), i32]
        L.fvc_copy_probe.argtypes = [vp, i64]
        L.fvc_batch_create.argtypes = [vp, vp, ctypes.POINTER(Release), i32, d, d, i32, d, ctypes.POINTER(vp)]
        L.fvc_batch_destroy.argtypes = [vp]
        L.fvc_batch_reset.argtypes = [vp]
        L.fvc_batch_run.argtypes = [vp, ctypes.POINTER(BatchParams), ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.fvc_batch_tune.argtypes = [vp, vp, i32, vp, i32, i32, ctypes.POINTER(BatchParams), ctypes.POINTER(d)]
        L.fvc_server_create.argtypes = [vp, ctypes.POINTER(ServerParams), ctypes.POINTER(vp)]
        L.fvc_server_destroy.argtypes = [vp]
        L.fvc_server_capacity.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.fvc_server_submit.argtypes = [vp, ctypes.POINTER(RequestDesc), ctypes.POINTER(i64)]
        L.fvc_server_poll.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.fvc_server_clock.restype = d
        L.fvc_server_clock.argtypes = [vp]
        L.fvc_server_info.argtypes = [vp, i64, ctypes.POINTER(RequestInfo)]
        L.fvc_server_stats.argtypes = [vp, vp]
        L.fvc_server_take.argtypes = [vp, i64, vp, vp, vp]
        L.fvc_mesh_zorder.argtypes = [i64, vp, i64, vp, vp, vp, vp, vp, vp, vp]
        L.fvc_set_chain_mode.argtypes = [vp, ctypes.c_int32]
        L.fvc_gather_rows.argtypes = [i64, ctypes.c_int, vp, vp, vp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _chk(rc, what, ctx=None):
    if rc != OK:
        detail = lib().fvc_last_error(ctx).decode() if ctx else ""
        raise FvcError(rc, what, detail)


def mesh_zorder(xyz, conn, adj=None):
    """Z-order renumbering of a tetrahedral mesh on the device
    (fvc_mesh_zorder; cudaVC/ZOrderSort/main.cpp:103-259).  Returns
    (node_order, elem_order, xyz_new, conn_new, adj_new)."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    conn = np.ascontiguousarray(conn, np.int32)
    nn, ne = xyz.shape[0], conn.shape[0]
    no, eo = np.zeros(nn, np.int32), np.zeros(ne, np.int32)
    xn, cn = np.zeros_like(xyz), np.zeros_like(conn)
    an = None
    if adj is not None:
        adj = np.ascontiguousarray(adj, np.int32)
        an = np.zeros_like(adj)
    _chk(lib().fvc_mesh_zorder(nn, _p(xyz), ne, _p(conn), _p(adj) if adj is not None else None, _p(no), _p(eo),
                               _p(xn), _p(cn), _p(an) if an is not None else None), "fvc_mesh_zorder")
    return no, eo, xn, cn, an


def gather_rows(order, data):
    """Rows of `data` (n x width doubles) in the order `order` (fvc_gather_rows)."""
    order = np.ascontiguousarray(order, np.int32)
    data = np.ascontiguousarray(data, np.float64)
    out = np.empty_like(data)
    width = 1 if data.ndim == 1 else data.shape[1]
    _chk(lib().fvc_gather_rows(data.shape[0], width, _p(order), _p(data), _p(out)), "fvc_gather_rows")
    return out


class Context:
    """A device context: mesh, spatial hash and resident velocity slices."""

    def __init__(self, dims, bounds, precision="f64", max_slices=2, coords=None, conn=None, adj=None,
                 cart_min=None, cart_max=None, cart_res=None, walk_limit=0, device=0):
        L = lib()
        md = MeshDesc()
        md.dims = dims
        md.bounds[:] = list(bounds)
        self._keep = []
        if coords is not None:
            md.mesh_type = 1
            coords = np.ascontiguousarray(coords, np.float64)
            conn = np.ascontiguousarray(conn, np.int32)
            adj = np.ascontiguousarray(adj, np.int32)
            self._keep += [coords, conn, adj]
            md.num_nodes, md.num_elems = coords.shape[0], conn.shape[0]
            md.coords, md.conn, md.adj = _p(coords), _p(conn), _p(adj)
            self.nvel = coords.shape[0]
        else:
            md.mesh_type = 0
            md.cart_min[:] = list(cart_min)
            md.cart_max[:] = list(cart_max)
            md.cart_res[:] = list(cart_res)
            self.nvel = int(cart_res[0] * cart_res[1] * (cart_res[2] if dims == 3 else 1))
        pp = Params(PREC[precision], max_slices, device, 0, walk_limit)
        h = ctypes.c_void_p()
        _chk(L.fvc_create(ctypes.byref(md), ctypes.byref(pp), ctypes.byref(h)), "fvc_create")
        self.h = h
        self.dims = dims
        self.precision = precision
        self.mesh_type = md.mesh_type
        self._children = weakref.WeakSet()

    def _adopt(self, child):
        """Objects on this context are destroyed before it (also when the
        garbage collector finalises a reference cycle in any order)."""
        self._children.add(child)

    def close(self):
        if self.h:
            for child in list(self._children):
                child.close()
            lib().fvc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_slice(self, slot, t, uvw):
        uvw = np.ascontiguousarray(uvw, np.float64)
        _chk(lib().fvc_load_slice(self.h, slot, t, _p(uvw), uvw.shape[0]), "fvc_load_slice", self.h)

    def memory(self):
        u, f = ctypes.c_int64(), ctypes.c_int64()
        _chk(lib().fvc_memory(self.h, ctypes.byref(u), ctypes.byref(f)), "fvc_memory", self.h)
        return u.value, f.value

    def locate(self, xyz, guess=None, mode=LOCATE_WALK):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.empty(xyz.shape[0], np.int32)
        g = None if guess is None else np.ascontiguousarray(guess, np.int32)
        _chk(lib().fvc_locate(self.h, _p(xyz), _p(g), xyz.shape[0], mode, _p(out)), "fvc_locate", self.h)
        return out

    def seed_list(self, xyz, lsc=1):
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.zeros(xyz.shape[0], POINT_DTYPE)
        _chk(lib().fvc_seed_list(self.h, _p(xyz), xyz.shape[0], lsc, _p(out)), "fvc_seed_list", self.h)
        return out

    def set_chain_mode(self, mode):
        """0: tabulated seed chains where possible (default); 1: point-by-point replay."""
        _chk(lib().fvc_set_chain_mode(self.h, mode), "fvc_set_chain_mode", self.h)

    def seed_grid(self, mn, dl, res, lsc=1):
        mn, dl = np.asarray(mn, np.float64), np.asarray(dl, np.float64)
        res = np.asarray(res, np.int32)
        out = np.zeros(int(np.prod(res)), POINT_DTYPE)
        _chk(lib().fvc_seed_grid(self.h, _p(mn), _p(dl), _p(res), lsc, _p(out)), "fvc_seed_grid", self.h)
        return out

    def zsort(self, xyz, lo, hi):
        xyz = np.ascontiguousarray(xyz, np.float64)
        lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
        perm = np.empty(xyz.shape[0], np.int32)
        _chk(lib().fvc_zsort(self.h, _p(xyz), xyz.shape[0], _p(lo), _p(hi), _p(perm)), "fvc_zsort", self.h)
        return perm

    def last_steps(self):
        n = ctypes.c_int64()
        lib().fvc_last_step_count(self.h, ctypes.byref(n))
        return n.value

    def set_timing(self, enable=True):
        _chk(lib().fvc_set_timing(self.h, 1 if enable else 0), "fvc_set_timing", self.h)

    def kernel_time(self, reset=False):
        """(milliseconds, launches, particle-steps) of timed integration kernels."""
        ms, nl, ns = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        _chk(lib().fvc_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(ns), 1 if reset else 0),
             "fvc_kernel_time", self.h)
        return ms.value, nl.value, ns.value

    def copy_probe(self, nbytes):
        _chk(lib().fvc_copy_probe(self.h, nbytes), "fvc_copy_probe", self.h)


class Particles:
    """Device-resident LagrangianPoint state."""

    def __init__(self, ctx, capacity):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _chk(lib().fvc_particles_create(ctx.h, capacity, ctypes.byref(h)), "fvc_particles_create", ctx.h)
        self.h = h
        ctx._adopt(self)
        self.capacity = capacity

    def close(self):
        if self.h:  # also safe after the context's destroy, which detached this object
            lib().fvc_particles_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, recs, begin=0):
        recs = np.ascontiguousarray(recs, POINT_DTYPE)
        _chk(lib().fvc_particles_set(self.h, begin, _p(recs), recs.shape[0]), "fvc_particles_set", self.ctx.h)

    def get(self, n=None, begin=0):
        n = self.capacity - begin if n is None else n
        out = np.zeros(n, POINT_DTYPE)
        _chk(lib().fvc_particles_get(self.h, begin, _p(out), n), "fvc_particles_get", self.ctx.h)
        return out

    def copy_from(self, src, n=None, begin=0, src_begin=0):
        n = src.capacity - src_begin if n is None else n
        _chk(lib().fvc_particles_copy(self.h, begin, src.h, src_begin, n), "fvc_particles_copy", self.ctx.h)

    def get_xyz(self, n=None, begin=0):
        n = self.capacity - begin if n is None else n
        xyz = np.empty((n, 3), np.float64)
        left = np.empty(n, np.int32)
        _chk(lib().fvc_particles_get_xyz(self.h, begin, _p(xyz), _p(left), n), "fvc_particles_get_xyz", self.ctx.h)
        return xyz, left

    def launch_velocity(self, t, lo, hi, tmin, tmax, begin=0, count=None):
        count = self.capacity - begin if count is None else count
        _chk(lib().fvc_launch_velocity(self.ctx.h, self.h, begin, count, t, lo, hi, tmin, tmax),
             "fvc_launch_velocity", self.ctx.h)

    def active(self):
        """Length of the list of particles still to integrate that the next
        advect over the last range launches, or -1 (whole range)."""
        n = ctypes.c_int64()
        _chk(lib().fvc_particles_active(self.h, ctypes.byref(n)), "fvc_particles_active", self.ctx.h)
        return n.value

    def advect(self, t1, t2, lo, hi, tmin, tmax, dt, int_type=INT_RK4, extrapolate=0, begin=0, count=None,
               rkf=(0.0, 0.0, 0.0)):
        """rkf: (Int_MinTimeStep, Int_MaxTimeStep, Int_Accuracy) for INT_RKF."""
        count = self.capacity - begin if count is None else count
        ip = IntParams(int_type, extrapolate, dt, 0, 0, 0, 0, *rkf)
        _chk(lib().fvc_advect(self.ctx.h, self.h, begin, count, t1, t2, lo, hi, tmin, tmax, ctypes.byref(ip)),
             "fvc_advect", self.ctx.h)
        return self.ctx.last_steps()


class FtleSlide:
    def __init__(self, ctx, mn, mx, res, lsc, start, stop, _handle=None):
        self.ctx = ctx
        if _handle is None:
            mn, mx = np.asarray(mn, np.float64), np.asarray(mx, np.float64)
            res = np.asarray(res, np.int32)
            h = ctypes.c_void_p()
            _chk(lib().fvc_ftle_create(ctx.h, _p(mn), _p(mx), _p(res), lsc, start, stop, ctypes.byref(h)),
                 "fvc_ftle_create", ctx.h)
        else:
            h = _handle
        self.h = h
        ctx._adopt(self)
        n, ins = ctypes.c_int64(), ctypes.c_int64()
        lib().fvc_ftle_points(h, ctypes.byref(n), ctypes.byref(ins))
        self.n, self.inside = n.value, ins.value

    @classmethod
    def from_records(cls, ctx, mn, dl, res, recs, have, ftle_wt, ftle_wot, start, stop):
        """A slide started from stored initial state (FTLE_GenerateMesh = 0)."""
        mn, dl = np.asarray(mn, np.float64), np.asarray(dl, np.float64)
        res = np.asarray(res, np.int32)
        recs = np.ascontiguousarray(recs, POINT_DTYPE)
        have = np.ascontiguousarray(have, np.int32)
        ftle_wt = np.ascontiguousarray(ftle_wt, np.float64)
        ftle_wot = np.ascontiguousarray(ftle_wot, np.float64)
        n = int(np.prod(res))
        if not (recs.size == have.size == ftle_wt.size == ftle_wot.size == n):
            raise ValueError("initial state arrays must hold res[0]*res[1]*res[2] entries")
        h = ctypes.c_void_p()
        _chk(lib().fvc_ftle_create_from(ctx.h, _p(mn), _p(dl), _p(res), _p(recs), _p(have), _p(ftle_wt),
                                        _p(ftle_wot), start, stop, ctypes.byref(h)), "fvc_ftle_create_from", ctx.h)
        return cls(ctx, None, None, None, None, None, None, _handle=h)

    def initial(self):
        """The slide's state right after creation: (records, HaveFTLE, FTLEwT, FTLEwoT)."""
        recs = np.zeros(self.n, POINT_DTYPE)
        have = np.zeros(self.n, np.int32)
        wt, wot = np.zeros(self.n), np.zeros(self.n)
        _chk(lib().fvc_ftle_initial(self.h, _p(recs), _p(have), _p(wt), _p(wot)), "fvc_ftle_initial", self.ctx.h)
        return recs, have, wt, wot

    def frame(self, direction, lo, hi, tmin, tmax, dt, int_type=INT_RK4, extrapolate=0, rkf=(0.0, 0.0, 0.0)):
        ip = IntParams(int_type, extrapolate, dt, 0, 0, 0, 0, *rkf)
        done = ctypes.c_int32()
        _chk(lib().fvc_ftle_frame(self.h, direction, lo, hi, tmin, tmax, ctypes.byref(ip), ctypes.byref(done)),
             "fvc_ftle_frame", self.ctx.h)
        return done.value

    def reset(self):
        _chk(lib().fvc_ftle_reset(self.h), "fvc_ftle_reset", self.ctx.h)

    def records(self):
        out = np.zeros(self.n, POINT_DTYPE)
        _chk(lib().fvc_ftle_records(self.h, _p(out)), "fvc_ftle_records", self.ctx.h)
        return out

    def output(self, int_t):
        wt = np.empty(self.n, np.float64)
        wot = np.empty(self.n, np.float64)
        _chk(lib().fvc_ftle_output(self.h, int_t, _p(wt), _p(wot)), "fvc_ftle_output", self.ctx.h)
        return wt, wot

    def close(self):
        if self.h:  # also safe after the context's destroy, which detached this object
            lib().fvc_ftle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Releases fused into one persistent kernel (fvc_batch_*).

    releases: sequence of (begin, count, t_start, t_stop, int_type).  Data
    frame f spans [t0 + f*frame_dt, t0 + (f+1)*frame_dt] and interpolates
    slots f % nslots and (f+1) % nslots."""

    def __init__(self, ctx, particles, releases, t0, frame_dt, nslots, dt):
        self.ctx = ctx
        arr = (Release * len(releases))()
        for k, (b, c, ts, te, it) in enumerate(releases):
            arr[k] = Release(int(b), int(c), float(ts), float(te), int(it), 0)
        h = ctypes.c_void_p()
        _chk(lib().fvc_batch_create(ctx.h, particles.h, arr, len(releases), t0, frame_dt, nslots, dt,
                                    ctypes.byref(h)), "fvc_batch_create", ctx.h)
        self.h = h
        ctx._adopt(self)

    def reset(self):
        _chk(lib().fvc_batch_reset(self.h), "fvc_batch_reset", self.ctx.h)

    def run(self, block=256, steps_per_launch=64, use_graph=1, waves_per_cu=0):
        bp = BatchParams(block, steps_per_launch, use_graph, waves_per_cu)
        st, nl = ctypes.c_int64(), ctypes.c_int64()
        _chk(lib().fvc_batch_run(self.h, ctypes.byref(bp), ctypes.byref(st), ctypes.byref(nl)), "fvc_batch_run",
             self.ctx.h)
        return st.value, nl.value

    def tune(self, blocks=(64, 128, 256, 512, 1024), budgets=(1, 4, 16, 64, 256), use_graph=1):
        bl = np.asarray(blocks, np.int32)
        bu = np.asarray(budgets, np.int32)
        best = BatchParams()
        ms = ctypes.c_double()
        _chk(lib().fvc_batch_tune(self.h, _p(bl), len(bl), _p(bu), len(bu), use_graph, ctypes.byref(best),
                                  ctypes.byref(ms)), "fvc_batch_tune", self.ctx.h)
        return {"block": best.block, "steps_per_launch": best.steps_per_launch, "use_graph": best.use_graph,
                "waves_per_cu": best.waves_per_cu}, ms.value

    def close(self):
        if self.h:  # also safe after the context's destroy, which detached this object
            lib().fvc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server over a context's resident slices
    (fvc_server_*)."""

    def __init__(self, ctx, frame_t0, frame_dt, nslots, dt, pool_particles=0, hbm_fraction=0.5, max_requests=256,
                 block=256, steps_per_launch=32, lsc=1, use_graph=True):
        self.ctx = ctx
        sp = ServerParams(frame_t0, frame_dt, nslots, dt, pool_particles, hbm_fraction, max_requests, block,
                          steps_per_launch, lsc, 0 if use_graph else 1)
        h = ctypes.c_void_p()
        _chk(lib().fvc_server_create(ctx.h, ctypes.byref(sp), ctypes.byref(h)), "fvc_server_create", ctx.h)
        self.h = h
        ctx._adopt(self)
        self._keep = {}

    def capacity(self):
        n, b = ctypes.c_int64(), ctypes.c_int64()
        _chk(lib().fvc_server_capacity(self.h, ctypes.byref(n), ctypes.byref(b)), "fvc_server_capacity", self.ctx.h)
        return n.value, b.value

    def submit_tracer(self, xyz, t_start, t_stop, int_type=INT_RK4):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        rq = RequestDesc(REQ_TRACER, int_type, t_start, t_stop, xyz.shape[0], _p(xyz).value)
        rid = ctypes.c_int64()
        _chk(lib().fvc_server_submit(self.h, ctypes.byref(rq), ctypes.byref(rid)), "fvc_server_submit", self.ctx.h)
        return rid.value

    def submit_ftle(self, mn, mx, res, t_start, t_stop):
        rq = RequestDesc(REQ_FTLE, INT_RK4, t_start, t_stop, 0, None)
        for a in range(3):
            rq.ftle_min[a], rq.ftle_max[a], rq.ftle_res[a] = float(mn[a]), float(mx[a]), int(res[a])
        rid = ctypes.c_int64()
        _chk(lib().fvc_server_submit(self.h, ctypes.byref(rq), ctypes.byref(rid)), "fvc_server_submit", self.ctx.h)
        return rid.value

    def poll(self):
        r, q = ctypes.c_int32(), ctypes.c_int32()
        _chk(lib().fvc_server_poll(self.h, ctypes.byref(r), ctypes.byref(q)), "fvc_server_poll", self.ctx.h)
        return r.value, q.value

    def clock(self):
        return lib().fvc_server_clock(self.h)

    def stats(self):
        out = np.zeros(5, np.float64)
        _chk(lib().fvc_server_stats(self.h, _p(out)), "fvc_server_stats", self.ctx.h)
        return {"wait_s": out[0], "admit_s": out[1], "ftle_frame_s": out[2], "retire_s": out[3],
                "ticks": int(out[4])}

    def info(self, rid):
        inf = RequestInfo()
        _chk(lib().fvc_server_info(self.h, rid, ctypes.byref(inf)), "fvc_server_info", self.ctx.h)
        return {"state": inf.state, "kind": inf.kind, "n": inf.n, "particle_steps": inf.particle_steps,
                "t_submit": inf.t_submit, "t_admit": inf.t_admit, "t_done": inf.t_done}

    def take(self, rid):
        inf = self.info(rid)
        if inf["kind"] == REQ_TRACER:
            out = np.zeros(inf["n"], POINT_DTYPE)
            _chk(lib().fvc_server_take(self.h, rid, _p(out), None, None), "fvc_server_take", self.ctx.h)
            return out
        wt = np.empty(inf["n"], np.float64)
        wot = np.empty(inf["n"], np.float64)
        _chk(lib().fvc_server_take(self.h, rid, None, _p(wt), _p(wot)), "fvc_server_take", self.ctx.h)
        return wt, wot

    def run_until_idle(self, max_ticks=1000000):
        for _ in range(max_ticks):
            r, q = self.poll()
            if r == 0 and q == 0:
                return
        raise RuntimeError("server still busy after %d ticks" % max_ticks)

    def close(self):
        if self.h:  # also safe after the context's destroy, which detached this object
            lib().fvc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
