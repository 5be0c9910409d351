/*
 * fvc_device.h - gfx950 device code of the advection path: point location,
 * interpolation and the per-particle integrators.
 *
 * One lane owns one particle for a whole time interval; its state (FP64
 * position, walk elements of the four RK4 stage probes) lives in registers.
 * Mesh geometry and velocity are random gathers (no MFMA: this is a sparse
 * stencil, not a contraction).
 *
 * Precision modes (template MODE):
 *   P_F64  : FP64 everything, expressions in the reference's order with FP
 *            contraction off, so results are bit-identical to ompVC's.
 *   P_F32  : FP32 element frames (an inverse edge matrix and origin per
 *            element) and FP32 velocity; FP64 position/time state.  Every
 *            walk decision (accept, exit face) is taken in FP32 only when it
 *            is farther than the element's guard band from its threshold;
 *            otherwise it is recomputed with the FP64 formula, so the
 *            element sequence of a walk equals the FP64 walk's for the same
 *            query point.
 *   P_BF16/P_F16 : as P_F32 with velocity stored in 16 bits times a
 *            per-slice scale.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include <type_traits>

#include "fvc_root4.h"

/* waves per SIMD the hot persistent and per-interval RK4 kernels are
 * compiled for (their VGPR budget is 512 / waves, rounded to the
 * allocation granule) */
#ifndef FVC_HOT_WAVES
#define FVC_HOT_WAVES 3
#endif

namespace fvc {

constexpr double TINY = 1.0e-10; /* ompVC/macros.h:17 */

enum Mode { P_F64 = 0, P_F32 = 1, P_BF16 = 2, P_F16 = 3 };
enum Accept { ACC_SEARCH = 0, ACC_INTERP = 1 }; /* d > -TINY (mesh.c:549) / d+TINY >= 0 (velocity.c:545) */

/* device error codes (match include/fvc.h) */
enum { E_OK = 0, E_INDET = -4, E_TLOC = -5, E_ELEM = -6, E_WALK = -7, E_EIG = -8, E_LEFT = -11, E_CET = -13 };

struct DevMesh {
    int dims, mesh_type;
    double bounds[6];
    /* Cartesian velocity grid */
    double cmin[3], cdel[3];
    int cres[3];
    /* unstructured */
    const double4 *__restrict__ X64;
    const float4 *__restrict__ ER; /* FP32 affine frame of each element (4 x float4) */
    const int4 *__restrict__ conn;
    const int4 *__restrict__ adj;
    double centre[3];
    int64_t nn, ne;
    float gq;       /* FP32 guard coefficient of the query offset: 12 2^-24 max |M| */
    int walk_limit; /* hop limit */
    /* spatial hash */
    const int *__restrict__ cell_start;
    const int *__restrict__ cell_elems;
    double hlo[3], hinv[3];
    int hres[3];
};

/* Velocity slices bound for one interval: slot lo at tmin, hi at tmax. */
struct DevField {
    const void *lo, *hi;
    float scale_lo, scale_hi;
    double tmin, tmax;
    /* strain-rate slices at tmin, tmax when the activation potential is integrated, else null */
    const double *sr_lo = nullptr, *sr_hi = nullptr;
    /* exposure time accumulators (per element: 128-bit fixed-point time, encounters), else null */
    unsigned long long *cet_lo = nullptr, *cet_hi = nullptr;
    int *cet_enc = nullptr;
    int cet_sub = 0; /* Trace_CETSubsteps */
    /* Runge-Kutta-Fehlberg: Int_MinTimeStep, Int_MaxTimeStep, Int_Accuracy */
    double rkf_min = 0.0, rkf_max = 0.0, rkf_tol = 0.0;
};

struct ErrBuf {
    int *code;     /* first error */
    int64_t *where;
};

__device__ inline void raise_err(ErrBuf e, int code, int64_t where) {
    if (atomicCAS(e.code, 0, code) == 0) *e.where = where;
}

/* Gather of element idx of an unstructured-mesh array with a 32-bit byte
 * offset: node, connectivity, adjacency and node-velocity arrays are below
 * 4 GiB (fvc_create checks it), so the address is the array base plus a
 * zero-extended 32-bit offset, which gfx950's global loads take as a scalar
 * base and one vector register, without 64-bit address arithmetic. */
template <typename T>
__device__ __forceinline__ T ld32(const T *__restrict__ base, uint32_t idx) {
    const uint32_t off = idx * (uint32_t)sizeof(T);
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (uint64_t)off);
}

/* A wave claims the next 64 entries of a persistent kernel's work queue
 * (a counter per queue, reset before the launch) and returns their offset;
 * an offset >= len means the queue is exhausted.  Lane 0 reads the counter
 * first and skips the atomic once it has passed len, so the waves that
 * arrive after the queue ran out (most of a launch with little work) do not
 * all contend on one address; the counter only grows, so a stale read only
 * costs an atomic. */
__device__ __forceinline__ int claim64(int *queue, int len) {
    int base = 0;
    if ((threadIdx.x & 63) == 0) {
        base = __hip_atomic_load(queue, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (base < len) base = atomicAdd(queue, 64);
    }
    return __shfl(base, 0, 64);
}

/* ------------------------------------------------------------------------ */
/* FP64 natural coordinates in the reference's expression order               */
/* (ompVC/mesh.c:527-547, velocity.c:497-517)                                */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ void nat3_exact_p(const double4 *__restrict__ X64, const int4 *__restrict__ conn, int e,
                                             double x, double y, double z, double &r, double &s, double &t) {
#pragma clang fp contract(off)
    int4 nd = ld32(conn, e);
    double4 p0 = ld32(X64, nd.x), p1 = ld32(X64, nd.y), p2 = ld32(X64, nd.z), p3 = ld32(X64, nd.w);
    double x0 = p0.x, y0 = p0.y, z0 = p0.z, x1 = p1.x, y1 = p1.y, z1 = p1.z;
    double x2 = p2.x, y2 = p2.y, z2 = p2.z, x3 = p3.x, y3 = p3.y, z3 = p3.z;
    double c11 = (z3 - z0) * (y2 - y3) - (z2 - z3) * (y3 - y0);
    double c21 = (z3 - z0) * (y0 - y1) - (z0 - z1) * (y3 - y0);
    double c31 = (z1 - z2) * (y0 - y1) - (z0 - z1) * (y1 - y2);
    double c12 = (x3 - x0) * (z2 - z3) - (x2 - x3) * (z3 - z0);
    double c22 = (x3 - x0) * (z0 - z1) - (x0 - x1) * (z3 - z0);
    double c32 = (x1 - x2) * (z0 - z1) - (x0 - x1) * (z1 - z2);
    double c13 = (y3 - y0) * (x2 - x3) - (y2 - y3) * (x3 - x0);
    double c23 = (y3 - y0) * (x0 - x1) - (y0 - y1) * (x3 - x0);
    double c33 = (y1 - y2) * (x0 - x1) - (y0 - y1) * (x1 - x2);
    double det = (x1 - x0) * ((y2 - y0) * (z3 - z0) - (z2 - z0) * (y3 - y0)) +
                 (x2 - x0) * ((y0 - y1) * (z3 - z0) - (z0 - z1) * (y3 - y0)) +
                 (x3 - x0) * ((y1 - y0) * (z2 - z0) - (z1 - z0) * (y2 - y0));
    double dx = x - x0, dy = y - y0, dz = z - z0;
    r = (c11 * dx + c12 * dy + c13 * dz) / det;
    s = (c21 * dx + c22 * dy + c23 * dz) / det;
    t = (c31 * dx + c32 * dy + c33 * dz) / det;
}

__device__ __forceinline__ void nat2_exact_p(const double4 *__restrict__ X64, const int4 *__restrict__ conn, int e,
                                             double x, double y, double &r, double &s) {
#pragma clang fp contract(off)
    int4 nd = ld32(conn, e);
    double4 p0 = ld32(X64, nd.x), p1 = ld32(X64, nd.y), p2 = ld32(X64, nd.z);
    double x0 = p0.x, y0 = p0.y, x1 = p1.x, y1 = p1.y, x2 = p2.x, y2 = p2.y;
    double c11 = y2 - y0, c12 = x0 - x2, c21 = y0 - y1, c22 = x1 - x0;
    double det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    r = (c11 * (x - x0) + c12 * (y - y0)) / det;
    s = (c21 * (x - x0) + c22 * (y - y0)) / det;
}

__device__ __forceinline__ void nat3_exact(const DevMesh &m, int e, double x, double y, double z, double &r, double &s,
                                           double &t) {
    nat3_exact_p(m.X64, m.conn, e, x, y, z, r, s, t);
}

__device__ __forceinline__ void nat2_exact(const DevMesh &m, int e, double x, double y, double &r, double &s) {
    nat2_exact_p(m.X64, m.conn, e, x, y, r, s);
}

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

/* exact decision on natural coordinates; returns 1 accept, 0 move (slot set), -1 indeterminate */
template <int DIM>
__device__ __forceinline__ int decide_exact(double r, double s, double t, int acc, int &slot, int *tie = nullptr) {
#pragma clang fp contract(off)
    double d, l0;
    if (DIM == 3) {
        l0 = 1 - r - s - t;
        d = dmin(r, dmin(s, dmin(t, l0)));
    } else {
        l0 = 1 - r - s;
        d = dmin(r, dmin(s, l0));
    }
    if (tie && d == -TINY) *tie = 1; /* the one value on which the two acceptance rules differ */
    bool in = acc == ACC_SEARCH ? (d > -TINY) : ((d + TINY) >= 0);
    if (in) return 1;
    if (fabs(r - d) < TINY) slot = 0;
    else if (fabs(s - d) < TINY) slot = 1;
    else if (DIM == 3 && fabs(t - d) < TINY) slot = 2;
    else if (fabs(d - l0) < TINY) slot = DIM == 3 ? 3 : 2;
    else return -1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP32 natural coordinates from per-element affine frames                  */
/* ------------------------------------------------------------------------ */
/* Element e's frame is four float4 at ER[4e .. 4e+3], built at fvc_create
 * from the FP64 nodes (relative to the mesh centre):
 *   ER[4e+a] = (M[a][0], M[a][1], M[a][2], o[a])   a = 0, 1, 2
 *   ER[4e+3] = (lo[0], lo[1], lo[2], g)
 * o = node 0 rounded to FP32 (an FP32 origin, exact in FP32), M = the inverse
 * of the edge matrix [n1-n0, n2-n0, n3-n0] rounded to FP32, lo the natural
 * coordinates of o itself (o - node0 is the rounding of node 0), and g the
 * element's guard coefficient 4 (18 cond + 12) 2^-24, cond = |M| |E| in the
 * row-sum norm with the 1-r-s-t row included.  2-D frames use the (x, y)
 * parts of rows 0, 1 and 3.  Natural coordinates are then
 *   (r, s, t) = lo + M (q - o),
 * q - o formed as (A - o) + d from the anchor A and offset d of the query
 * (or rounded once from the FP64 difference), three FMAs per coordinate.
 * Per coordinate, the rounding of M, of (A - o), of the sum with d and of
 * the three FMAs err by at most 6 2^-24 cond L + 2^-24 |M| |d| to first
 * order, L = max(1, |r|, |s|, |t|, |1-r-s-t|); 1-r-s-t sums three such
 * errors and rounds three times more: (18 cond + 12) 2^-24 L +
 * 3 2^-24 |M| |d| in all, and the band is four times that.  Node rounding
 * enters only through lo, so the band scales with the element's own shape,
 * not with the mesh's extent.  An element whose edge matrix is singular or
 * whose frame is not finite gets M = 0, lo = 0, g = inf: every test on it
 * is decided in FP64. */
struct Frame {
    float4 a0, a1, a2, a3;
};

template <int DIM>
__device__ __forceinline__ Frame ld_frame(const DevMesh &m, int e) {
    Frame F;
    const uint32_t b = 4u * (uint32_t)e;
    F.a0 = ld32(m.ER, b);
    F.a1 = ld32(m.ER, b + 1);
    if (DIM == 3) F.a2 = ld32(m.ER, b + 2);
    else F.a2 = make_float4(0.f, 0.f, 0.f, 0.f);
    F.a3 = ld32(m.ER, b + 3);
    return F;
}

/* (r, s, t) = lo + M q with q = query - o, each an explicit FMA chain */
template <int DIM>
__device__ __forceinline__ void frame_nat(const Frame &F, float qx, float qy, float qz, float &r, float &s, float &t) {
    if (DIM == 3) {
        r = __builtin_fmaf(F.a0.z, qz, __builtin_fmaf(F.a0.y, qy, __builtin_fmaf(F.a0.x, qx, F.a3.x)));
        s = __builtin_fmaf(F.a1.z, qz, __builtin_fmaf(F.a1.y, qy, __builtin_fmaf(F.a1.x, qx, F.a3.y)));
        t = __builtin_fmaf(F.a2.z, qz, __builtin_fmaf(F.a2.y, qy, __builtin_fmaf(F.a2.x, qx, F.a3.z)));
    } else {
        r = __builtin_fmaf(F.a0.y, qy, __builtin_fmaf(F.a0.x, qx, F.a3.x));
        s = __builtin_fmaf(F.a1.y, qy, __builtin_fmaf(F.a1.x, qx, F.a3.y));
        t = 0.0f;
    }
}

/* query given in FP64 (x, y, z): q - o rounded once from the FP64 difference */
template <int DIM>
__device__ __forceinline__ void nat_fast(const DevMesh &m, int e, double x, double y, double z, float &r, float &s,
                                         float &t, float &ge) {
#pragma clang fp contract(off)
    const Frame F = ld_frame<DIM>(m, e);
    const float qx = (float)((x - m.centre[0]) - (double)F.a0.w);
    const float qy = (float)((y - m.centre[1]) - (double)F.a1.w);
    const float qz = DIM == 3 ? (float)((z - m.centre[2]) - (double)F.a2.w) : 0.0f;
    frame_nat<DIM>(F, qx, qy, qz, r, s, t);
    ge = F.a3.w;
}

/* query anchored in FP32: the point (centre + A) + d, q - o = (A - o) + d */
template <int DIM>
__device__ __forceinline__ void nat_fast_c(const DevMesh &m, int e, const float A[3], const float d[3], float &r,
                                           float &s, float &t, float &ge) {
#pragma clang fp contract(off)
    const Frame F = ld_frame<DIM>(m, e);
    const float qx = (A[0] - F.a0.w) + d[0];
    const float qy = (A[1] - F.a1.w) + d[1];
    const float qz = DIM == 3 ? (A[2] - F.a2.w) + d[2] : 0.0f;
    frame_nat<DIM>(F, qx, qy, qz, r, s, t);
    ge = F.a3.w;
}

/* guard term of a query offset d (its rounding in (A - o) + d) */
__device__ __forceinline__ float guard_offset(const DevMesh &m, const float d[3]) {
    return m.gq * fmaxf(fabsf(d[0]), fmaxf(fabsf(d[1]), fabsf(d[2])));
}

/* FP32 decision: 1 accept, 0 move (slot set), 2 undecided (escalate); the
 * band is ge L + gq around each threshold */
template <int DIM>
__device__ __forceinline__ int decide_fast(float r, float s, float t, float ge, float gq, int &slot) {
#pragma clang fp contract(off)
    float l0 = DIM == 3 ? ((1.0f - r) - s) - t : (1.0f - r) - s;
    float v[4] = {r, s, DIM == 3 ? t : l0, l0};
    const int n = DIM == 3 ? 4 : 3;
    int a = 0;
    float d = v[0], L = fabsf(v[0]);
    for (int q = 1; q < n; q++) {
        L = fmaxf(L, fabsf(v[q]));
        if (v[q] < d) {
            d = v[q];
            a = q;
        }
    }
    const float g = __builtin_fmaf(ge, fmaxf(L, 1.0f), gq);
    if (d > g) return 1;
    if (d >= -g) return 2;
    float second = 3.0e38f;
    for (int q = 0; q < n; q++)
        if (q != a && v[q] < second) second = v[q];
    if (second - d <= 2.0f * g) return 2;
    slot = a;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* walk (ompVC/mesh.c:496-611 and velocity.c:523-600)                       */
/* ------------------------------------------------------------------------ */
/* Walks from `e` towards (x,y,z).  Returns the accepting element, -1 when an
 * adjacency of -1 is crossed, or -2 on error (err set).  Natural coordinates
 * of the accepting element are returned in r,s,t (for interpolation). */
template <int MODE, int DIM>
__device__ __forceinline__ int walk(const DevMesh &m, double x, double y, double z, int e, int acc, double &r, double &s,
                                    double &t, int &err, int *tie = nullptr) {
    for (int hop = 0; hop < m.walk_limit; hop++) {
        int slot = -1, dec;
        if constexpr (MODE == P_F64) {
            if (DIM == 3) nat3_exact(m, e, x, y, z, r, s, t);
            else {
                nat2_exact(m, e, x, y, r, s);
                t = 0.0;
            }
            dec = decide_exact<DIM>(r, s, t, acc, slot, tie);
        } else {
            float fr, fs, ft, ge;
            nat_fast<DIM>(m, e, x, y, z, fr, fs, ft, ge);
            dec = decide_fast<DIM>(fr, fs, ft, ge, 0.0f, slot);
            if (dec == 2) {
                if (DIM == 3) nat3_exact(m, e, x, y, z, r, s, t);
                else {
                    nat2_exact(m, e, x, y, r, s);
                    t = 0.0;
                }
                dec = decide_exact<DIM>(r, s, t, acc, slot, tie);
            } else {
                r = fr;
                s = fs;
                t = ft;
            }
        }
        if (dec == 1) return e;
        if (dec < 0) {
            err = E_INDET;
            return -2;
        }
        e = ld32(reinterpret_cast<const int *>(m.adj), 4 * (uint32_t)e + slot);
        if (e == -1) return -1;
    }
    err = E_WALK;
    return -2;
}

/* ------------------------------------------------------------------------ */
/* walk on a query point given in FP32 relative to the mesh centre           */
/* ------------------------------------------------------------------------ */
/* The RK4 stage probes of the reduced-precision modes are formed in FP32 as
 * an anchor A = (float)(X - centre) and a small offset d = D + c h k, D the
 * FP32 residual of X - centre - A; the probe stands for the point
 * (centre + A) + d.  The offset keeps FP32 resolution at its own small
 * magnitude, so probes sit within ~1e-10 of the FP64 probe positions.  An
 * FP32 decision forms q - node0 as (A - node0) + d, and an undecided one is
 * recomputed with the FP64 formula at (centre + A) + d, so the element
 * returned is the one the reference's walk returns for that point. */
/* the FP64 decision at the double point (x, y, z), for the few tests that
 * fall inside the guard band (node coordinates and connectivity by pointer,
 * so the call can stay out of line without the mesh descriptor); tie reports
 * d == -TINY, where the two acceptance rules differ */
struct ExactDec {
    float r, s, t;
    int slot, dec, tie;
};

/* inlined: an out-of-line call costs the hot kernels registers around the
 * call (build with FVC_ESC_NOINLINE to compare) */
#ifdef FVC_ESC_NOINLINE
#define FVC_ESC_ATTR __noinline__
#else
#define FVC_ESC_ATTR __forceinline__
#endif

template <int DIM>
__device__ FVC_ESC_ATTR ExactDec decide_exact_at(const double4 *__restrict__ X64, const int4 *__restrict__ conn, int e,
                                                 double x, double y, double z, int acc) {
#pragma clang fp contract(off)
    double r, s, t = 0.0;
    if (DIM == 3) nat3_exact_p(X64, conn, e, x, y, z, r, s, t);
    else nat2_exact_p(X64, conn, e, x, y, r, s);
    ExactDec o;
    o.slot = -1;
    o.tie = 0;
    o.dec = decide_exact<DIM>(r, s, t, acc, o.slot, &o.tie);
    o.r = (float)r;
    o.s = (float)s;
    o.t = (float)t;
    return o;
}

/* walk<MODE != P_F64> with the query point (centre + A) + d: returns the
 * accepting element, -1 when an adjacency of -1 is crossed, -2 on error */
template <int DIM>
__device__ __forceinline__ int walk_c(const DevMesh &m, const float A[3], const float d[3], int e, int acc, float &r,
                                      float &s, float &t, int4 &nodes, int &err) {
    const float gd = guard_offset(m, d);
    for (int hop = 0; hop < m.walk_limit; hop++) {
        int slot = -1;
        nodes = ld32(m.conn, e);
        float fr, fs, ft, ge;
        nat_fast_c<DIM>(m, e, A, d, fr, fs, ft, ge);
        int dec = decide_fast<DIM>(fr, fs, ft, ge, gd, slot);
        if (dec == 2) {
            const ExactDec x = decide_exact_at<DIM>(m.X64, m.conn, e, (m.centre[0] + (double)A[0]) + (double)d[0],
                                                    (m.centre[1] + (double)A[1]) + (double)d[1],
                                                    (m.centre[2] + (double)A[2]) + (double)d[2], acc);
            fr = x.r;
            fs = x.s;
            ft = x.t;
            slot = x.slot;
            dec = x.dec;
        }
        if (dec == 1) {
            r = fr;
            s = fs;
            t = ft;
            return e;
        }
        if (dec < 0) {
            err = E_INDET;
            return -2;
        }
        e = ld32(reinterpret_cast<const int *>(m.adj), 4 * (uint32_t)e + slot);
        if (e == -1) return -1;
    }
    err = E_WALK;
    return -2;
}

/* ------------------------------------------------------------------------ */
/* searches (ompVC/mesh.c:304-398, 496-611)                                  */
/* ------------------------------------------------------------------------ */
/* lowest-index element with d > -TINY (mesh.c:321-363) and whether a second
 * element also passes; cell lists are ascending, so the first pass is the
 * lowest index. */
template <int DIM>
__device__ int hash_global(const DevMesh &m, double x, double y, double z, int &second) {
    second = -1;
    double fx = (x - m.hlo[0]) * m.hinv[0], fy = (y - m.hlo[1]) * m.hinv[1];
    double fz = DIM == 3 ? (z - m.hlo[2]) * m.hinv[2] : 0.0;
    if (!(fx >= 0.0 && fy >= 0.0 && fz >= 0.0)) return -1;
    if (!(fx < (double)m.hres[0] && fy < (double)m.hres[1] && fz < (double)m.hres[2])) return -1;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    int64_t cell = ((int64_t)iz * m.hres[1] + iy) * m.hres[0] + ix;
    int first = -1;
    for (int q = m.cell_start[cell]; q < m.cell_start[cell + 1]; q++) {
        int e = m.cell_elems[q];
        double r, s, t = 0.0;
        if (DIM == 3) nat3_exact(m, e, x, y, z, r, s, t);
        else nat2_exact(m, e, x, y, r, s);
        int slot;
        if (decide_exact<DIM>(r, s, t, ACC_SEARCH, slot) == 1) {
            if (first < 0) first = e;
            else {
                second = e;
                break;
            }
        }
    }
    return first;
}

/* every element with d > -TINY, ascending: the first `cap` are written to
 * `list` (when given), the lowest is returned in `first`, the count is
 * returned whole */
template <int DIM>
__device__ int hash_accepting(const DevMesh &m, double x, double y, double z, int *list, int cap, int &first) {
    first = -1;
    double fx = (x - m.hlo[0]) * m.hinv[0], fy = (y - m.hlo[1]) * m.hinv[1];
    double fz = DIM == 3 ? (z - m.hlo[2]) * m.hinv[2] : 0.0;
    if (!(fx >= 0.0 && fy >= 0.0 && fz >= 0.0)) return 0;
    if (!(fx < (double)m.hres[0] && fy < (double)m.hres[1] && fz < (double)m.hres[2])) return 0;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    int64_t cell = ((int64_t)iz * m.hres[1] + iy) * m.hres[0] + ix;
    int cnt = 0;
    for (int q = m.cell_start[cell]; q < m.cell_start[cell + 1]; q++) {
        int e = m.cell_elems[q];
        double r, s, t = 0.0;
        if (DIM == 3) nat3_exact(m, e, x, y, z, r, s, t);
        else nat2_exact(m, e, x, y, r, s);
        int slot;
        if (decide_exact<DIM>(r, s, t, ACC_SEARCH, slot) == 1) {
            if (cnt == 0) first = e;
            if (list && cnt < cap) list[cnt] = e;
            cnt++;
        }
    }
    return cnt;
}

template <int DIM>
__device__ int local_search(const DevMesh &m, double x, double y, double z, int guess, int &err) {
    double r, s, t;
    int e = walk<P_F64, DIM>(m, x, y, z, guess, ACC_SEARCH, r, s, t, err);
    return e == -2 ? -1 : e;
}

/* ------------------------------------------------------------------------ */
/* velocity loads                                                            */
/* ------------------------------------------------------------------------ */
template <int MODE>
struct VelT {
    using T = double;
};
template <>
struct VelT<P_F32> {
    using T = float;
};
template <>
struct VelT<P_BF16> {
    using T = float;
};
template <>
struct VelT<P_F16> {
    using T = float;
};

/* velocity of node n of an unstructured mesh (32-bit offsets, see ld32) */
template <int MODE>
__device__ __forceinline__ void load_vel_node(const void *base, float scale, uint32_t n, typename VelT<MODE>::T out[3]) {
    if constexpr (MODE == P_F64) {
        const double4 v = ld32(reinterpret_cast<const double4 *>(base), n);
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
    } else if constexpr (MODE == P_F32) {
        const float4 v = ld32(reinterpret_cast<const float4 *>(base), n);
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
    } else {
        const uint2 w = ld32(reinterpret_cast<const uint2 *>(base), n);
        unsigned short h0 = (unsigned short)(w.x & 0xffffu), h1 = (unsigned short)(w.x >> 16);
        unsigned short h2 = (unsigned short)(w.y & 0xffffu);
        if constexpr (MODE == P_BF16) {
            out[0] = __uint_as_float((unsigned)h0 << 16) * scale;
            out[1] = __uint_as_float((unsigned)h1 << 16) * scale;
            out[2] = __uint_as_float((unsigned)h2 << 16) * scale;
        } else {
            out[0] = __half2float(__ushort_as_half(h0)) * scale;
            out[1] = __half2float(__ushort_as_half(h1)) * scale;
            out[2] = __half2float(__ushort_as_half(h2)) * scale;
        }
    }
}

template <int MODE>
__device__ __forceinline__ void load_vel(const void *base, float scale, int64_t n, typename VelT<MODE>::T out[3]) {
    if constexpr (MODE == P_F64) {
        double4 v = reinterpret_cast<const double4 *>(base)[n];
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
    } else if constexpr (MODE == P_F32) {
        float4 v = reinterpret_cast<const float4 *>(base)[n];
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
    } else {
        uint2 w = reinterpret_cast<const uint2 *>(base)[n];
        unsigned short h0 = (unsigned short)(w.x & 0xffffu), h1 = (unsigned short)(w.x >> 16);
        unsigned short h2 = (unsigned short)(w.y & 0xffffu);
        if constexpr (MODE == P_BF16) {
            out[0] = __uint_as_float((unsigned)h0 << 16) * scale;
            out[1] = __uint_as_float((unsigned)h1 << 16) * scale;
            out[2] = __uint_as_float((unsigned)h2 << 16) * scale;
        } else {
            out[0] = __half2float(__ushort_as_half(h0)) * scale;
            out[1] = __half2float(__ushort_as_half(h1)) * scale;
            out[2] = __half2float(__ushort_as_half(h2)) * scale;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* unstructured interpolation (ompVC/velocity.c:459-621)                     */
/* ------------------------------------------------------------------------ */
/* Probe state: element (walk start, updated in place) and a left flag. */
struct Probe {
    int elem;
    int left;
};

template <int MODE, int DIM>
__device__ __forceinline__ void interp_unstruct(const DevMesh &m, const DevField &f, double tloc, int e, double r,
                                                double s, double t, double out[3]);

/* reduced-precision modes: velocities of an element's nodes in the two
 * bound slices as loaded (FP32 values, or 16-bit words before scaling), so
 * the gather can be issued before the containment test that decides whether
 * it is used */
struct F3 {
    float x, y, z;
};

template <int MODE, int DIM>
struct NodeVel {
    static constexpr int NN = DIM == 3 ? 4 : 3;
    typename std::conditional<MODE == P_F32, F3, uint2>::type lo[NN], hi[NN];
};

template <int MODE>
__device__ __forceinline__ void ld_vel_word(const void *base, uint32_t n, F3 &w) {
    const float4 v = ld32(reinterpret_cast<const float4 *>(base), n);
    w.x = v.x;
    w.y = v.y;
    w.z = v.z;
}

template <int MODE>
__device__ __forceinline__ void ld_vel_word(const void *base, uint32_t n, uint2 &w) {
    w = ld32(reinterpret_cast<const uint2 *>(base), n);
}

template <int MODE, int DIM>
__device__ __forceinline__ void gather_vel(const DevField &f, int4 nd, NodeVel<MODE, DIM> &g) {
    const uint32_t ids[4] = {(uint32_t)nd.x, (uint32_t)nd.y, (uint32_t)nd.z, (uint32_t)nd.w};
    for (int q = 0; q < NodeVel<MODE, DIM>::NN; q++) {
        ld_vel_word<MODE>(f.lo, ids[q], g.lo[q]);
        ld_vel_word<MODE>(f.hi, ids[q], g.hi[q]);
    }
}

template <int MODE>
__device__ __forceinline__ void node_comp(const F3 &w, float v[3]) {
    v[0] = w.x;
    v[1] = w.y;
    v[2] = w.z;
}

template <int MODE>
__device__ __forceinline__ void node_comp(const uint2 &w, float v[3]) {
    const unsigned short h0 = (unsigned short)(w.x & 0xffffu), h1 = (unsigned short)(w.x >> 16);
    const unsigned short h2 = (unsigned short)(w.y & 0xffffu);
    if constexpr (MODE == P_BF16) {
        v[0] = __uint_as_float((unsigned)h0 << 16);
        v[1] = __uint_as_float((unsigned)h1 << 16);
        v[2] = __uint_as_float((unsigned)h2 << 16);
    } else {
        v[0] = __half2float(__ushort_as_half(h0));
        v[1] = __half2float(__ushort_as_half(h1));
        v[2] = __half2float(__ushort_as_half(h2));
    }
}

/* the reference's weights fabs(r), fabs(s), fabs(t) and 1 minus their sum
 * (velocity.c:574-581, n0 + (n1-n0)|r| + (n2-n0)|s| + (n3-n0)|t|) applied to
 * each slice, then the time weights (1 - tl) and tl, which carry the 16-bit
 * slices' scales; FP32 throughout */
template <int MODE, int DIM>
__device__ __forceinline__ void interp_gathered(const DevField &f, const NodeVel<MODE, DIM> &g, float tl, float r,
                                                float s, float t, float out[3]) {
#pragma clang fp contract(off)
    constexpr int NN = NodeVel<MODE, DIM>::NN;
    float w[4];
    w[1] = fabsf(r);
    w[2] = fabsf(s);
    w[3] = DIM == 3 ? fabsf(t) : 0.0f;
    w[0] = DIM == 3 ? ((1.0f - w[1]) - w[2]) - w[3] : (1.0f - w[1]) - w[2];
    float lo[3], hi[3];
    for (int q = 0; q < NN; q++) {
        float a[3], b[3];
        node_comp<MODE>(g.lo[q], a);
        node_comp<MODE>(g.hi[q], b);
        for (int c = 0; c < 3; c++) {
            lo[c] = q == 0 ? w[0] * a[c] : __builtin_fmaf(w[q], a[c], lo[c]);
            hi[c] = q == 0 ? w[0] * b[c] : __builtin_fmaf(w[q], b[c], hi[c]);
        }
    }
    const float tlo = MODE == P_F32 ? 1.0f - tl : (1.0f - tl) * f.scale_lo;
    const float thi = MODE == P_F32 ? tl : tl * f.scale_hi;
    for (int c = 0; c < 3; c++) out[c] = (DIM == 2 && c == 2) ? 0.0f : __builtin_fmaf(thi, hi[c], tlo * lo[c]);
}

template <int MODE, int DIM>
__device__ __forceinline__ void interp_nodes_f(const DevField &f, float tl, int4 nd, float r, float s, float t,
                                               float out[3]) {
    NodeVel<MODE, DIM> g;
    gather_vel<MODE, DIM>(f, nd, g);
    interp_gathered<MODE, DIM>(f, g, tl, r, s, t, out);
}

template <int MODE, int DIM>
__device__ __forceinline__ void interp_unstruct_f(const DevMesh &m, const DevField &f, float tl, int e, float r,
                                                  float s, float t, float out[3]) {
    interp_nodes_f<MODE, DIM>(f, tl, ld32(m.conn, e), r, s, t, out);
}

template <int MODE, int DIM>
__device__ __forceinline__ void vel_unstruct(const DevMesh &m, const DevField &f, double tloc, double x, double y,
                                             double z, Probe &p, double out[3], int &err) {
    if (p.elem < 0 || p.elem >= m.ne) {
        err = E_ELEM;
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    double r, s, t;
    int e = walk<MODE, DIM>(m, x, y, z, p.elem, ACC_INTERP, r, s, t, err);
    if (e == -2) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    if (e == -1) {
        p.elem = -1;
        p.left = 1;
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    p.elem = e;
    interp_unstruct<MODE, DIM>(m, f, tloc, e, r, s, t, out);
}

/* velocity at natural coordinates (r,s,t) of element e (velocity.c:603-619) */
template <int MODE, int DIM>
__device__ __forceinline__ void interp_unstruct(const DevMesh &m, const DevField &f, double tloc, int e, double r,
                                                double s, double t, double out[3]) {
    int4 nd = ld32(m.conn, e);
    using VT = typename VelT<MODE>::T;
    if constexpr (MODE == P_F64) {
#pragma clang fp contract(off)
        double a[4][3], b[4][3], n[4][3];
        const int nn = DIM == 3 ? 4 : 3;
        int ids[4] = {nd.x, nd.y, nd.z, nd.w};
        for (int q = 0; q < nn; q++) {
            load_vel_node<MODE>(f.lo, 1.0f, ids[q], a[q]);
            load_vel_node<MODE>(f.hi, 1.0f, ids[q], b[q]);
            for (int c = 0; c < 3; c++) n[q][c] = (1 - tloc) * a[q][c] + tloc * b[q][c];
        }
        double ar = fabs(r), as = fabs(s), at = fabs(t);
        if (DIM == 3) {
            for (int c = 0; c < 3; c++)
                out[c] = n[0][c] + (n[1][c] - n[0][c]) * ar + (n[2][c] - n[0][c]) * as + (n[3][c] - n[0][c]) * at;
        } else {
            for (int c = 0; c < 2; c++) out[c] = n[0][c] + (n[1][c] - n[0][c]) * ar + (n[2][c] - n[0][c]) * as;
            out[2] = 0.0;
        }
    } else {
        float v[3];
        interp_unstruct_f<MODE, DIM>(m, f, (float)tloc, e, (float)r, (float)s, (float)t, v);
        for (int c = 0; c < 3; c++) out[c] = v[c];
    }
}

/* ------------------------------------------------------------------------ */
/* Cartesian domain tests and interpolation (ompVC/velocity.c:257-457,624-669) */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ int outside_domain(const DevMesh &m, double x, double y, double z) {
#pragma clang fp contract(off)
    if (x < (m.bounds[0] - TINY) || x > (m.bounds[1] + TINY)) return 1;
    if (y < (m.bounds[2] - TINY) || y > (m.bounds[3] + TINY)) return 1;
    if (z < (m.bounds[4] - TINY) || z > (m.bounds[5] + TINY)) return 1;
    return 0;
}

template <int DIM>
__device__ __forceinline__ int outside_cart_vel(const DevMesh &m, const double X[3]) {
#pragma clang fp contract(off)
    for (int a = 0; a < DIM; a++) {
        double q = (X[a] - m.cmin[a]) / m.cdel[a];
        int lo = (int)floor(q), hi = (int)ceil(q);
        if (lo < 0 || hi > m.cres[a] - 1) return 1;
    }
    return 0;
}

template <int MODE, int DIM>
__device__ __forceinline__ void vel_cart(const DevMesh &m, const DevField &f, double tloc, const double X[3], int &left,
                                         double out[3]) {
#pragma clang fp contract(off)
    if (outside_cart_vel<DIM>(m, X)) {
        out[0] = out[1] = out[2] = 0.0;
        left = 1;
        return;
    }
    int cell[3] = {0, 0, 0};
    double loc[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < DIM; a++) {
        cell[a] = (int)floor((X[a] - m.cmin[a]) / m.cdel[a]);
        if (cell[a] == m.cres[a] - 1) {
            cell[a] = m.cres[a] - 2;
            loc[a] = 1.0;
        } else {
            loc[a] = (X[a] - m.cmin[a] - cell[a] * m.cdel[a]) / m.cdel[a];
        }
    }
    const int64_t nx = m.cres[0], ny = m.cres[1];
    const int i = cell[0], j = cell[1], k = cell[2], k1 = DIM == 3 ? k + 1 : k;
    int64_t c[8] = {((int64_t)k * ny + j) * nx + i,           ((int64_t)k1 * ny + j) * nx + i,
                    ((int64_t)k * ny + j + 1) * nx + i,       ((int64_t)k1 * ny + j + 1) * nx + i,
                    ((int64_t)k * ny + j) * nx + i + 1,       ((int64_t)k1 * ny + j) * nx + i + 1,
                    ((int64_t)k * ny + j + 1) * nx + i + 1,   ((int64_t)k1 * ny + j + 1) * nx + i + 1};
    if constexpr (MODE == P_F64) {
        double tl = tloc, xl = loc[0], yl = loc[1], zl = loc[2];
        /* weights w[slot][corner], corner bits (x,y,z) = (4,2,1); products left to right */
        double w0[8], w1[8];
        for (int q = 0; q < 8; q++) {
            double wx = (q & 4) ? xl : (1 - xl), wy = (q & 2) ? yl : (1 - yl), wz = (q & 1) ? zl : (1 - zl);
            w1[q] = (tl) * wx * wy * wz;
            w0[q] = (1 - tl) * wx * wy * wz;
        }
        double A[8][3], B[8][3];
        for (int q = 0; q < 8; q++) {
            load_vel<MODE>(f.lo, 1.0f, c[q], A[q]);
            load_vel<MODE>(f.hi, 1.0f, c[q], B[q]);
        }
        for (int cc = 0; cc < 3; cc++) {
            double acc = A[0][cc] * w0[0];
            for (int q = 1; q < 8; q++) acc = acc + A[q][cc] * w0[q];
            for (int q = 0; q < 8; q++) acc = acc + B[q][cc] * w1[q];
            out[cc] = acc;
        }
    } else {
        using VT = typename VelT<MODE>::T;
        float tl = (float)tloc, xl = (float)loc[0], yl = (float)loc[1], zl = (float)loc[2];
        float acc[3] = {0.f, 0.f, 0.f};
        const int nq = DIM == 3 ? 8 : 8;
        for (int q = 0; q < nq; q++) {
            if (DIM == 2 && (q & 1)) continue; /* 2-D: the k+1 plane repeats k with zero weight */
            float wx = (q & 4) ? xl : (1.f - xl), wy = (q & 2) ? yl : (1.f - yl);
            float wz = DIM == 3 ? ((q & 1) ? zl : (1.f - zl)) : 1.f;
            float w = wx * wy * wz;
            VT a[3], b[3];
            load_vel<MODE>(f.lo, f.scale_lo, c[q], a);
            load_vel<MODE>(f.hi, f.scale_hi, c[q], b);
            for (int cc = 0; cc < 3; cc++) acc[cc] += w * (a[cc] + tl * (b[cc] - a[cc]));
        }
        out[0] = acc[0];
        out[1] = acc[1];
        out[2] = DIM == 3 ? (double)acc[2] : 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* particle state and integrators (ompVC/integration.c)                     */
/* ------------------------------------------------------------------------ */
struct PState {
    double X[3];
    double V[3];
    int elem;
    int left;
    double left_time;
    /* multistep history: velocities at the two previous nodes, count held */
    double F1[3], F2[3];
    int nhist;
    double scalar; /* activation potential */
};

__device__ __forceinline__ double sign_of(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

/* GetVelocity on the particle itself (side effects on elem / left). */
template <int MODE, int DIM>
__device__ __forceinline__ void get_velocity(const DevMesh &m, const DevField &f, double tq, PState &p, double out[3],
                                             int &err) {
#pragma clang fp contract(off)
    double tloc = (tq - f.tmin) / (f.tmax - f.tmin);
    if (m.mesh_type == 1) {
        if (p.elem < 0 || p.elem >= m.ne) {
            err = E_ELEM;
            return;
        }
        if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) {
            err = E_TLOC;
            return;
        }
        Probe pr{p.elem, p.left};
        vel_unstruct<MODE, DIM>(m, f, tloc, p.X[0], p.X[1], p.X[2], pr, out, err);
        p.elem = pr.elem;
        p.left = pr.left;
    } else {
        int lft = 0;
        if (outside_cart_vel<DIM>(m, p.X)) {
            out[0] = out[1] = out[2] = 0.0;
            p.left = 1;
            return;
        }
        if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) {
            err = E_TLOC;
            return;
        }
        vel_cart<MODE, DIM>(m, f, tloc, p.X, lft, out);
        if (lft) p.left = 1;
    }
}

/* velocity at a probe position for a stage (probe owns its element / flag) */
template <int MODE, int DIM, int UNSTRUCT>
__device__ __forceinline__ void stage_velocity(const DevMesh &m, const DevField &f, double tq, const double X[3],
                                               Probe &pr, double out[3], int &err) {
#pragma clang fp contract(off)
    double tloc = (tq - f.tmin) / (f.tmax - f.tmin);
    if (UNSTRUCT) {
        if (pr.elem == -1) {
            err = E_ELEM;
            return;
        }
        if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) {
            err = E_TLOC;
            return;
        }
        vel_unstruct<MODE, DIM>(m, f, tloc, X[0], X[1], X[2], pr, out, err);
    } else {
        if (outside_cart_vel<DIM>(m, X)) {
            out[0] = out[1] = out[2] = 0.0;
            pr.left = 1;
            return;
        }
        if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) {
            err = E_TLOC;
            return;
        }
        int lft = 0;
        vel_cart<MODE, DIM>(m, f, tloc, X, lft, out);
        if (lft) pr.left = 1;
    }
}

/* Reduced-precision modes on unstructured meshes: the stage velocity at an
 * FP32 probe (anchor A, offset d; the point (centre + A) + d) and at the
 * FP64 position X, with FP32 results. */
template <int MODE, int DIM>
__device__ __forceinline__ void stage_velocity_c(const DevMesh &m, const DevField &f, double tloc, const float A[3],
                                                 const float d[3], Probe &pr, float out[3], int &err) {
#pragma clang fp contract(off)
    out[0] = out[1] = out[2] = 0.0f;
    if (pr.elem == -1) {
        err = E_ELEM;
        return;
    }
    if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) {
        err = E_TLOC;
        return;
    }
    if (pr.elem < 0 || pr.elem >= m.ne) {
        err = E_ELEM;
        return;
    }
    float r, s, t;
    int4 nd;
    const int e = walk_c<DIM>(m, A, d, pr.elem, ACC_INTERP, r, s, t, nd, err);
    if (e == -2) return;
    if (e == -1) {
        pr.elem = -1;
        pr.left = 1;
        return;
    }
    pr.elem = e;
    interp_nodes_f<MODE, DIM>(f, (float)tloc, nd, r, s, t, out);
}

template <int MODE, int DIM>
__device__ __forceinline__ void stage_velocity_x(const DevMesh &m, const DevField &f, double tloc, const double X[3],
                                                 Probe &pr, float out[3], int &err) {
#pragma clang fp contract(off)
    out[0] = out[1] = out[2] = 0.0f;
    if (pr.elem == -1) {
        err = E_ELEM;
        return;
    }
    if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) {
        err = E_TLOC;
        return;
    }
    if (pr.elem < 0 || pr.elem >= m.ne) {
        err = E_ELEM;
        return;
    }
    double r, s, t;
    const int e = walk<MODE, DIM>(m, X[0], X[1], X[2], pr.elem, ACC_INTERP, r, s, t, err);
    if (e == -2) return;
    if (e == -1) {
        pr.elem = -1;
        pr.left = 1;
        return;
    }
    pr.elem = e;
    interp_unstruct_f<MODE, DIM>(m, f, (float)tloc, e, (float)r, (float)s, (float)t, out);
}

/* ------------------------------------------------------------------------ */
/* Resumable integration of one interval [a, b]: one Advect call of the      */
/* reference.  IState keeps what the reference keeps in locals across its    */
/* step loop: the current time, the step size (which stays clipped once      */
/* clipped, integration.c:346-352), and the four RK4 stage probes, each with */
/* its own walk element and left flag (integration.c:337-341).  With         */
/* max_steps unlimited this is the reference's loop; with a budget, a        */
/* persistent kernel can stop and resume it with identical arithmetic.       */
/* ------------------------------------------------------------------------ */
struct IState {
    double tc, h;
    Probe mp[4];
    int started;
};

enum { INT_EULER = 0, INT_RK4 = 1, INT_RKF = 2, INT_AB2 = 3, INT_AB3 = 4 };

/* Strain rate at the point (GetStrainRate_Unstructured, ompVC/strainrate.c:
 * 232-246; GetStrainRate_Cartesian, :196-230), linear in time between the two
 * loaded strain-rate slices.  Unstructured: the reference first walks from
 * the point's element (search acceptance, d > -TINY); the element held here
 * was located by that same acceptance at this same position (the step's
 * final locate, the seed chain or a launch walk that keeps it), so that walk
 * returns it unchanged and is not repeated. */
/* Exposure-time sums are kept exactly: a time t >= 0 adds floor(t 2^CET_FRAC)
 * to a 128-bit fixed-point accumulator (two 64-bit words, the carry of the low
 * word added to the high one), so the per-element totals do not depend on the
 * order in which the parallel steps arrive; the reference adds in its
 * threads' arrival order (exposuretime.c:135-164). */
constexpr int CET_FRAC = 100;

__device__ inline void cet_add(unsigned long long *lo, unsigned long long *hi, double v) {
    if (!(v > 0.0)) return;
    int e;
    const double fr = frexp(v, &e);                            /* v = fr 2^e, fr in [0.5, 1) */
    const unsigned long long mnt = (unsigned long long)ldexp(fr, 53); /* v = mnt 2^(e-53) */
    const int sh = e - 53 + CET_FRAC;
    unsigned long long l = 0, h = 0;
    if (sh <= -64) return;
    if (sh < 0) l = mnt >> (-sh);
    else if (sh < 64) {
        l = mnt << sh;
        h = sh ? (mnt >> (64 - sh)) : 0ull;
    } else if (sh < 128) {
        h = mnt << (sh - 64);
    }
    const unsigned long long old = atomicAdd(lo, l);
    const unsigned long long carry = (old + l < old) ? 1ull : 0ull;
    if (h + carry) atomicAdd(hi, h + carry);
}

/* ComputeExposureTime (exposuretime.c:120-178) for a step from X0 to X1 of
 * length h, X1 in element e1: the whole step to the element of X0 if that is
 * e1, else h/(S-1) to the element of each of the S-1 points X0 + a (X1 - X0),
 * a = s/(S-1), s = 1..S-1, walking from point to point, with an encounter
 * where the element changes; points in no element are skipped. */
template <int DIM>
__device__ inline void cet_step(const DevMesh &m, const DevField &f, const double X0[3], const double X1[3], int e1,
                                double h, int &err) {
#pragma clang fp contract(off)
    int e0 = local_search<DIM>(m, X0[0], X0[1], X0[2], e1, err);
    if (err) return;
    if (e0 < 0) {
        int second;
        e0 = hash_global<DIM>(m, X0[0], X0[1], X0[2], second);
        if (e0 < 0) {
            err = E_CET;
            return;
        }
    }
    if (e0 == e1) {
        cet_add(f.cet_lo + e0, f.cet_hi + e0, h);
        return;
    }
    const int S = f.cet_sub;
    int last = e0;
    for (int s = 1; s < S; s++) {
        const double a = ((double)s) / ((double)(S - 1));
        double Xt[3];
        for (int i = 0; i < 3; i++) Xt[i] = (1 - a) * X0[i] + a * X1[i];
        int et = local_search<DIM>(m, Xt[0], Xt[1], Xt[2], last, err);
        if (err) return;
        if (et < 0) {
            int second;
            et = hash_global<DIM>(m, Xt[0], Xt[1], Xt[2], second);
            if (et < 0) continue; /* path crosses the outside: no contribution (exposuretime.c:160-165) */
        }
        cet_add(f.cet_lo + et, f.cet_hi + et, h / (S - 1));
        if (et != last) {
            atomicAdd(f.cet_enc + et, 1);
            last = et;
        }
    }
}

template <int DIM, int UNSTRUCT>
__device__ inline double strain_rate(const DevMesh &m, const DevField &f, double t, const PState &p, int &err) {
#pragma clang fp contract(off)
    const double tloc = (t - f.tmin) / (f.tmax - f.tmin);
    if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) {
        err = E_TLOC;
        return 0.0;
    }
    int64_t idx;
    if (UNSTRUCT) {
        if (p.elem < 0 || p.elem >= m.ne) {
            err = E_ELEM;
            return 0.0;
        }
        idx = p.elem;
    } else {
        if (outside_domain(m, p.X[0], p.X[1], p.X[2])) {
            err = E_LEFT;
            return 0.0;
        }
        int cell[3] = {0, 0, 0};
        for (int a = 0; a < DIM; a++) {
            cell[a] = (int)floor((p.X[a] - m.cmin[a]) / m.cdel[a]);
            if (cell[a] == m.cres[a] - 1) cell[a] = m.cres[a] - 2;
            if (cell[a] < 0 || cell[a] > m.cres[a] - 2) { /* the reference asserts here */
                err = E_ELEM;
                return 0.0;
            }
        }
        const int64_t imax = m.cres[0] - 1, jmax = m.cres[1] - 1;
        idx = ((int64_t)cell[2] * jmax + cell[1]) * imax + cell[0];
    }
    return (1 - tloc) * f.sr_lo[idx] + tloc * f.sr_hi[idx];
}


/* ------------------------------------------------------------------------ */
/* RK4 of the reduced-precision modes on unstructured meshes, one           */
/* containment test per lane per loop iteration                              */
/* ------------------------------------------------------------------------ */
/* The same integrator as interval_run's RK4 (integration.c:317-462): four
 * stage probes that keep their own walk elements, the increment
 * h (f0 + 2 k2 + 2 k3 + k4) / 6, a locate walk at the new position, the exit
 * revert, the first stage taking the locate's element and coordinates.  It is
 * written as a state machine: a lane is in one of five phases (first-stage
 * walk at X, the k2, k3, k4 probe walks, the locate walk), and each pass of
 * the loop makes one containment test for every lane, followed by the
 * interpolation of the lanes whose test accepted and the bookkeeping of their
 * next phase.  Walks of different lengths then overlap instead of making the
 * whole wavefront repeat the longest one.  Every lane performs the same
 * floating-point operations, in the same order, as a straight-line walk would.
 * Query points are anchored FP32 offsets (A, d), see nat_fast_c; an
 * undecided test of the first-stage or locate walk is decided in FP64 at the
 * double position X, a probe's at (centre + A) + d. */
enum { PH_F0 = 0, PH_K2 = 1, PH_K3 = 2, PH_K4 = 3, PH_LOC = 4, PH_DONE = 5 };

template <int MODE, int DIM>
__device__ int rk4_sm(const DevMesh &m, const DevField &f, PState &p, IState &is, double a, double b, double dt,
                      int max_steps, int &err, uint32_t &nsteps, double &stop) {
#pragma clang fp contract(off)
    if (!is.started) {
        if (p.left) {
            err = E_LEFT;
            stop = b;
            return 1;
        }
        if (p.elem < 0 || p.elem >= m.ne) {
            err = E_ELEM;
            stop = b;
            return 1;
        }
        is.tc = a;
        is.h = sign_of(dt, b - a);
        for (int q = 0; q < 4; q++) {
            is.mp[q].elem = p.elem;
            is.mp[q].left = 0;
        }
        is.started = 1;
    }
    const double hmin = 1e-12 * fabs(dt);
    const double inv_span = 1.0 / (f.tmax - f.tmin);
    int taken = 0, ret = 1;
    int phase = PH_F0, e = 0, hops = 0, start = 0;
    /* S: f0, then f0 + 2 k2, f0 + 2 k2 + 2 k3, and from the k4 stage to the
     * locate's end the step's f0 + 2 k2 + 2 k3 + k4 (increment h/6 S) */
    float A[3], D[3], dq[3], S[3] = {0.f, 0.f, 0.f};
    float gD = 0.f, gdq = 0.f; /* guard terms of the offsets D and dq */
    /* anchors of X: A = (float)(X - centre), D its FP32 residual */
    auto anchor = [&]() {
        for (int c = 0; c < 3; c++) {
            const double xc = p.X[c] - m.centre[c];
            A[c] = (float)xc;
            D[c] = (float)(xc - (double)A[c]);
        }
        gD = guard_offset(m, D);
    };
    /* time location of a phase: tc, tc + h/2, tc + h */
    auto tloc_of = [&](int ph) {
        const double tq = ph == PH_F0 ? is.tc : (ph == PH_K4 ? is.tc + is.h : is.tc + is.h / 2);
        return (tq - f.tmin) * inv_span;
    };
    /* enter a probe phase: its walk starts at the probe's own element */
    auto enter_probe = [&](int ph) {
        phase = ph;
        hops = 0;
        const double tl = tloc_of(ph);
        if (is.mp[ph].elem == -1) err = E_ELEM;
        else if ((tl > 1 + TINY) || (tl < 0 - TINY)) err = E_TLOC;
        else if (is.mp[ph].elem < 0 || is.mp[ph].elem >= m.ne) err = E_ELEM;
        e = is.mp[ph].elem;
    };
    /* start of a step (the loop head of integration.c:344); false when the
     * interval is done or the step budget spent */
    auto begin_step = [&]() -> bool {
        if (!((is.tc - b) * (b - a) < 0.0)) {
            stop = b;
            return false;
        }
        if (taken == max_steps) {
            ret = 0;
            return false;
        }
        if ((is.tc + is.h - b) * (is.tc + is.h - a) > 0.0) is.h = sign_of(b - is.tc, b - a);
        return true;
    };
    anchor();
    float r = 0.f, s = 0.f, t = 0.f;
    int ltie = 0; /* a locate-walk test hit d == -TINY */
    if (!begin_step()) phase = PH_DONE;
    else enter_probe(PH_F0);
    if (err) phase = PH_DONE;
    while (phase != PH_DONE) {
        /* ---- one containment test of this lane's walk (the element's
         * connectivity is loaded beside its frame, for the velocity gather
         * when the walk ends here) */
        const int4 nd = ld32(m.conn, e);
        const bool at_x = phase == PH_F0 || phase == PH_LOC;
        int slot = -1, dec;
        float ge;
        nat_fast_c<DIM>(m, e, A, at_x ? D : dq, r, s, t, ge);
        dec = decide_fast<DIM>(r, s, t, ge, at_x ? gD : gdq, slot);
        if (dec == 2) {
            double x, y, z;
            if (at_x) {
                x = p.X[0];
                y = p.X[1];
                z = p.X[2];
            } else {
                x = (m.centre[0] + (double)A[0]) + (double)dq[0];
                y = (m.centre[1] + (double)A[1]) + (double)dq[1];
                z = (m.centre[2] + (double)A[2]) + (double)dq[2];
            }
            const ExactDec xd =
                decide_exact_at<DIM>(m.X64, m.conn, e, x, y, z, phase == PH_LOC ? ACC_SEARCH : ACC_INTERP);
            r = xd.r;
            s = xd.s;
            t = xd.t;
            slot = xd.slot;
            dec = xd.dec;
            ltie |= xd.tie;
        }
        if (dec < 0) {
            err = E_INDET;
            break;
        }
        if (dec == 0) { /* across the face */
            e = ld32(reinterpret_cast<const int *>(m.adj), 4 * (uint32_t)e + slot);
            if (e != -1) {
                if (++hops >= m.walk_limit) {
                    err = E_WALK;
                    break;
                }
                continue; /* the next test of this walk */
            }
        }
        /* ---- the walk ended: accepted at e, or left the mesh (e == -1) */
        bool want = e >= 0;
        double tl_i = 0.0;
        if (phase == PH_LOC) {
            p.elem = want ? e : -1;
            /* a probe that left holds element -1 (and its left flag) */
            const bool out = !want || is.mp[0].elem < 0 || is.mp[1].elem < 0 || is.mp[2].elem < 0 ||
                             is.mp[3].elem < 0;
            if (out) { /* exit: back to the step's start (integration.c:430-440) */
                const double h6 = is.h * (1.0 / 6.0);
                p.left = 1;
                p.left_time = is.tc - is.h;
                for (int c = 0; c < 3; c++) p.X[c] = p.X[c] - h6 * (double)S[c];
                stop = p.left_time;
                break;
            }
            /* the next step; its first stage takes this element and these
             * coordinates when the locate started where the first stage
             * started and no test met the tie */
            const bool reuse = !ltie && start == is.mp[0].elem;
            if (!begin_step()) break;
            if (!reuse) {
                enter_probe(PH_F0);
                if (err) break;
                continue;
            }
            tl_i = tloc_of(PH_F0);
            if ((tl_i > 1 + TINY) || (tl_i < 0 - TINY)) {
                err = E_TLOC;
                break;
            }
            is.mp[0].elem = e;
            phase = PH_F0;
        } else {
            if (want) is.mp[phase].elem = e;
            else { /* the probe left: zero velocity for this stage (velocity.c:588-592) */
                is.mp[phase].elem = -1;
                is.mp[phase].left = 1;
            }
            tl_i = tloc_of(phase);
        }
        /* ---- this stage's velocity and the next phase */
        float k[3] = {0.f, 0.f, 0.f};
        if (want) interp_nodes_f<MODE, DIM>(f, (float)tl_i, nd, r, s, t, k);
        const float hf = (float)is.h, hh = 0.5f * hf;
        if (phase == PH_F0) {
            for (int c = 0; c < 3; c++) {
                S[c] = k[c];
                dq[c] = D[c] + hh * k[c];
            }
            gdq = guard_offset(m, dq);
            enter_probe(PH_K2);
        } else if (phase == PH_K2 || phase == PH_K3) {
            const float w = phase == PH_K2 ? hh : hf;
            for (int c = 0; c < 3; c++) {
                S[c] = S[c] + 2.0f * k[c];
                dq[c] = D[c] + w * k[c];
            }
            gdq = guard_offset(m, dq);
            enter_probe(phase + 1);
        } else { /* PH_K4: the step's increment, position and time */
            const double h = is.h, h6 = h * (1.0 / 6.0);
            for (int c = 0; c < 3; c++) {
                S[c] = S[c] + k[c];
                p.X[c] = p.X[c] + h6 * (double)S[c];
            }
            is.tc = is.tc + h;
            if (!(fabs(h) <= hmin)) nsteps++;
            taken++;
            anchor();
            phase = PH_LOC;
            hops = 0;
            ltie = 0;
            start = p.elem;
            e = p.elem; /* a located element: the walks above keep p.elem >= 0 */
        }
        if (err) break;
    }
    if (err) stop = b;
    return ret;
}

/* Returns 1 when the interval is finished (end reached, point exited or an
 * error), with *stop the stop time; 0 when max_steps were taken first.
 *   RK4   ompVC/integration.c:317-462
 *   Euler ompVC/integration.c:222-311 (GetVelocity on the point itself)
 *   AB2/AB3 (no ompVC counterpart; cudaVC/src/multistep.cu:249-304 is not
 *   followed): x+ = x + h (3 f0 - f1)/2, x+ = x + h (23 f0 - 16 f1 + 5 f2)/12
 *   with f0 = u(x, t); RK4 steps while the history is short or the step is
 *   clipped by more than 1e-9 h (history restarts); steps below 1e-12 dt
 *   (accumulation remainders) are Euler steps that keep the history. */
template <int MODE, int DIM, int UNSTRUCT, int GEN>
__device__ int interval_run(const DevMesh &m, const DevField &f, PState &p, IState &is, double a, double b, double dt,
                            int itype_rt, int extrap, int max_steps, int &err, uint32_t &nsteps, double &stop) {
#pragma clang fp contract(off)
    /* GEN = 0 compiles RK4 only (no multistep state in registers) */
    const int itype = GEN ? itype_rt : INT_RK4;
    if constexpr (MODE != P_F64 && UNSTRUCT && !GEN) {
        /* GEN = 0 builds integrate RK4 only, without strain rate or exposure
         * time; the GEN = 1 builds (other integrators, RK4 with extrapolation
         * or those integrals) keep the straight-line walks below and leave
         * the state machine's registers out */
        if (!extrap) return rk4_sm<MODE, DIM>(m, f, p, is, a, b, dt, max_steps, err, nsteps, stop);
    }
    if (!is.started) {
        if (p.left) {
            err = E_LEFT;
            stop = b;
            return 1;
        }
        if (UNSTRUCT && (p.elem < 0 || p.elem >= m.ne)) {
            err = E_ELEM;
            stop = b;
            return 1;
        }
        is.tc = a;
        is.h = sign_of(dt, b - a);
        for (int q = 0; q < 4; q++) {
            is.mp[q].elem = p.elem;
            is.mp[q].left = 0;
        }
        is.started = 1;
    }
    const double hfull = sign_of(dt, b - a);
    const double hmin = 1e-12 * fabs(dt);
    const int order = itype == INT_AB2 ? 2 : 3;
    int taken = 0;
    /* The locate walk at the end of a step starts from the point's element
     * and seeks the new position; the next step's first-stage walk seeks the
     * same position from the first probe's element.  When the two start
     * elements are equal (the usual case: both hold the previous position)
     * the walks visit the same elements and take the same decisions (the
     * acceptance rules d > -TINY and d + TINY >= 0 differ only at
     * d == -TINY, which voids the reuse), so the first stage takes the
     * locate's element and natural coordinates without walking again. */
    int c_elem = -1;
    /* activation potential (integration.c:355-360, 449-455): Scalar gains
     * h (S0 + S1) / 2 per step that stays in the domain, S0 and S1 the strain
     * rate at the step's start and end; only the generic integrator build */
    const bool ap = GEN && f.sr_lo != nullptr;
    const bool cet = GEN && UNSTRUCT && f.cet_lo != nullptr; /* exposure time (integration.c:427-448) */
    using CT = typename VelT<MODE>::T; /* reduced-precision interpolation reads (float) r, s, t */
    CT c_r = 0, c_s = 0, c_t = 0;
    /* Reduced-precision modes on unstructured meshes integrate in FP32: stage
     * velocities, probes (relative to the mesh centre) and the weighted sum
     * of the four slopes; the position and time stay FP64, the increment is
     * h times the FP32 slope sum over 6 in FP64. */
    constexpr bool FAST = MODE != P_F64 && UNSTRUCT;
    using KT = typename std::conditional<FAST, float, double>::type;
    /* FAST: time locations as (t - tmin) times one reciprocal of the frame span */
    const double inv_span = FAST ? 1.0 / (f.tmax - f.tmin) : 0.0;
    while ((is.tc - b) * (b - a) < 0.0) {
        if (taken == max_steps) return 0;
        if ((is.tc + is.h - b) * (is.tc + is.h - a) > 0.0) is.h = sign_of(b - is.tc, b - a);
        const double h = is.h, tc = is.tc;
        const bool micro = fabs(h) <= hmin;
        double inc[3], sr0 = 0.0;
        KT f0[3];
        if (ap) {
            sr0 = strain_rate<DIM, UNSTRUCT>(m, f, tc, p, err);
            if (err) break;
        }
        bool ab_step = false;
        if (itype == INT_EULER) {
            double ve[3];
            get_velocity<MODE, DIM>(m, f, tc, p, ve, err);
            if (err) break;
            for (int c = 0; c < 3; c++) inc[c] = h * ve[c];
        } else {
            if (UNSTRUCT && c_elem >= 0) {
                const double tloc = FAST ? (tc - f.tmin) * inv_span : (tc - f.tmin) / (f.tmax - f.tmin);
                if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) {
                    err = E_TLOC;
                    break;
                }
                is.mp[0].elem = c_elem;
                if constexpr (FAST) interp_unstruct_f<MODE, DIM>(m, f, (float)tloc, c_elem, c_r, c_s, c_t, f0);
                else interp_unstruct<MODE, DIM>(m, f, tloc, c_elem, c_r, c_s, c_t, f0);
            } else {
                if constexpr (FAST) stage_velocity_x<MODE, DIM>(m, f, (tc - f.tmin) * inv_span, p.X, is.mp[0], f0, err);
                else stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc, p.X, is.mp[0], f0, err);
            }
            c_elem = -1;
            if (err) break;
            const bool full = fabs(h - hfull) <= 1e-9 * fabs(hfull);
            if (itype != INT_RK4 && micro) {
                for (int c = 0; c < 3; c++) inc[c] = h * (double)f0[c];
            } else if (itype != INT_RK4 && full && p.nhist >= order - 1) {
                ab_step = true;
                if (order == 2)
                    for (int c = 0; c < 3; c++) inc[c] = h * (3 * (double)f0[c] - p.F1[c]) / 2;
                else
                    for (int c = 0; c < 3; c++)
                        inc[c] = h * (23 * (double)f0[c] - 16 * p.F1[c] + 5 * p.F2[c]) / 12;
            } else if constexpr (FAST) {
                /* k2, k3, k4 at probes X + h/2 f0, X + h/2 k2, X + h k3 (anchor A,
                 * offset D + c h k); S = f0 + 2 k2 + 2 k3 + k4 */
                const float hf = (float)h, hh = 0.5f * hf;
                const double tl_half = (tc + h / 2 - f.tmin) * inv_span, tl_end = (tc + h - f.tmin) * inv_span;
                float A[3], D[3], d[3], k[3], S[3];
                for (int c = 0; c < 3; c++) {
                    const double xc = p.X[c] - m.centre[c];
                    A[c] = (float)xc;
                    D[c] = (float)(xc - (double)A[c]);
                    S[c] = f0[c];
                    d[c] = D[c] + hh * f0[c];
                }
                stage_velocity_c<MODE, DIM>(m, f, tl_half, A, d, is.mp[1], k, err);
                if (err) break;
                for (int c = 0; c < 3; c++) {
                    S[c] = S[c] + 2.0f * k[c];
                    d[c] = D[c] + hh * k[c];
                }
                stage_velocity_c<MODE, DIM>(m, f, tl_half, A, d, is.mp[2], k, err);
                if (err) break;
                for (int c = 0; c < 3; c++) {
                    S[c] = S[c] + 2.0f * k[c];
                    d[c] = D[c] + hf * k[c];
                }
                stage_velocity_c<MODE, DIM>(m, f, tl_end, A, d, is.mp[3], k, err);
                if (err) break;
                const double h6 = h * (1.0 / 6.0);
                for (int c = 0; c < 3; c++) inc[c] = h6 * (double)(S[c] + k[c]);
            } else {
                double k2[3], k3[3], k4[3], Y[3];
                for (int c = 0; c < 3; c++) Y[c] = p.X[c] + 0.5 * f0[c] * h;
                stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc + h / 2, Y, is.mp[1], k2, err);
                if (err) break;
                for (int c = 0; c < 3; c++) Y[c] = p.X[c] + 0.5 * k2[c] * h;
                stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc + h / 2, Y, is.mp[2], k3, err);
                if (err) break;
                for (int c = 0; c < 3; c++) Y[c] = p.X[c] + k3[c] * h;
                stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc + h, Y, is.mp[3], k4, err);
                if (err) break;
                for (int c = 0; c < 3; c++) inc[c] = h * (f0[c] + 2 * k2[c] + 2 * k3[c] + k4[c]) / 6;
            }
            (void)ab_step;
        }
        for (int c = 0; c < 3; c++) p.X[c] = p.X[c] + inc[c];
        is.tc = tc + h;
        if (!micro) nsteps++;
        taken++;
        bool out;
        if (UNSTRUCT) {
            double r, s, t;
            const int start = p.elem;
            int tie = 0;
            int e = p.elem < 0 ? -1
                               : walk<MODE, DIM>(m, p.X[0], p.X[1], p.X[2], p.elem, ACC_SEARCH, r, s, t, err, &tie);
            if (e == -2) break;
            p.elem = e;
            out = e < 0;
            /* RK4: the next first-stage walk starts at mp[0]; AB: at the located element */
            if (e >= 0 && !tie && itype != INT_EULER && (itype != INT_RK4 || start == is.mp[0].elem)) {
                c_elem = e;
                c_r = (CT)r;
                c_s = (CT)s;
                c_t = (CT)t;
            }
        } else {
            out = outside_domain(m, p.X[0], p.X[1], p.X[2]);
        }
        if (itype != INT_EULER)
            out = out || is.mp[0].left || is.mp[1].left || is.mp[2].left || is.mp[3].left;
        if (out) {
            p.left = 1;
            p.left_time = is.tc - h;
            for (int c = 0; c < 3; c++) p.X[c] = p.X[c] - inc[c];
            stop = p.left_time;
            return 1;
        }
        if (cet) { /* the step's start as the reference recomputes it: X1 minus the increment */
            double X0[3];
            for (int c = 0; c < 3; c++) X0[c] = p.X[c] - inc[c];
            cet_step<DIM>(m, f, X0, p.X, p.elem, h, err);
            if (err) break;
        }
        if (ap) {
            const double sr1 = strain_rate<DIM, UNSTRUCT>(m, f, is.tc, p, err);
            if (err) break;
            p.scalar += (h * ((sr0 + sr1) / 2));
        }
        if (itype == INT_AB2 || itype == INT_AB3) {
            if (!micro) {
                if (fabs(h - hfull) <= 1e-9 * fabs(hfull)) {
                    for (int c = 0; c < 3; c++) {
                        p.F2[c] = p.F1[c];
                        p.F1[c] = f0[c];
                    }
                    p.nhist = p.nhist < 2 ? p.nhist + 1 : 2;
                } else {
                    p.nhist = 0;
                }
            }
            if (UNSTRUCT) is.mp[0].elem = p.elem; /* next f0 walks from the located element */
        }
        if (extrap) {
            get_velocity<MODE, DIM>(m, f, is.tc, p, p.V, err);
            if (err) break;
        }
    }
    stop = b;
    return 1;
}

/* Runge-Kutta-Fehlberg 4(5) over [a, b] (ompVC/integration.c:464-699): the
 * fourth-order result is kept and the step adapted from its difference to
 * the fifth-order one; the step starts at Int_MaxTimeStep every call and six
 * probes keep their own elements within the call.  A probe outside stops the
 * point at the step's start time, without revert; a position outside after
 * an accepted step is reverted as in RK4.  Accepted steps are counted. */
template <int MODE, int DIM, int UNSTRUCT>
__device__ double rkf_run(const DevMesh &m, const DevField &f, PState &p, double a, double b, int extrap, int &err,
                          uint32_t &nsteps) {
#pragma clang fp contract(off)
    if (p.left) {
        err = E_LEFT;
        return b;
    }
    if (UNSTRUCT && (p.elem < 0 || p.elem >= m.ne)) {
        err = E_ELEM;
        return b;
    }
    Probe mp[6];
    for (int q = 0; q < 6; q++) {
        mp[q].elem = p.elem;
        mp[q].left = 0;
    }
    const double a3 = 3.0 / 32.0, b3 = 9.0 / 32.0;
    const double a4 = 1932.0 / 2197.0, b4 = -7200.0 / 2197.0, c4 = 7296.0 / 2197.0;
    const double a5 = 439.0 / 216.0, b5 = -8.0, c5 = 3680.0 / 513.0, d5 = -845.0 / 4104.0;
    const double a6 = -8.0 / 27.0, b6 = 2.0, c6 = -3544.0 / 2565.0, d6 = 1859.0 / 4104.0, e6 = -11.0 / 40.0;
    const double a7 = 1.0 / 360.0, b7 = -128.0 / 4275.0, c7 = -2197.0 / 75240.0, d7 = 1.0 / 50.0, e7 = 2.0 / 55.0;
    const double a8 = 25.0 / 216.0, b8 = 1408.0 / 2565.0, c8 = 2197.0 / 4104.0, d8 = -1.0 / 5.0;
    const double hmax = f.rkf_max, hmin = f.rkf_min, tol = f.rkf_tol;
    const bool ap = f.sr_lo != nullptr;                 /* activation potential (integration.c:534-539, 671-677) */
    const bool cet = UNSTRUCT && f.cet_lo != nullptr;   /* exposure time (integration.c:650-670) */
    double h = sign_of(hmax, b - a), tc = a;
    double k1[3], k2[3], k3[3], k4[3], k5[3], k6[3], Y[3];
    while ((tc - b) * (b - a) < 0.0) {
        if ((tc + h - b) * (tc + h - a) > 0.0) h = sign_of(b - tc, b - a);
        double sr0 = 0.0;
        if (ap) {
            sr0 = strain_rate<DIM, UNSTRUCT>(m, f, tc, p, err);
            if (err) return b;
        }
        stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc, p.X, mp[0], k1, err);
        if (err) return b;
        for (int c = 0; c < 3; c++) Y[c] = p.X[c] + 0.25 * k1[c] * h;
        stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc + 0.25 * h, Y, mp[1], k2, err);
        if (err) return b;
        for (int c = 0; c < 3; c++) Y[c] = p.X[c] + (a3 * k1[c] + b3 * k2[c]) * h;
        stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc + 0.375 * h, Y, mp[2], k3, err);
        if (err) return b;
        for (int c = 0; c < 3; c++) Y[c] = p.X[c] + (a4 * k1[c] + b4 * k2[c] + c4 * k3[c]) * h;
        stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc + (12.0 * h) / 13.0, Y, mp[3], k4, err);
        if (err) return b;
        for (int c = 0; c < 3; c++) Y[c] = p.X[c] + (a5 * k1[c] + b5 * k2[c] + c5 * k3[c] + d5 * k4[c]) * h;
        stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc + h, Y, mp[4], k5, err);
        if (err) return b;
        for (int c = 0; c < 3; c++)
            Y[c] = p.X[c] + (a6 * k1[c] + b6 * k2[c] + c6 * k3[c] + d6 * k4[c] + e6 * k5[c]) * h;
        stage_velocity<MODE, DIM, UNSTRUCT>(m, f, tc + 0.5 * h, Y, mp[5], k6, err);
        if (err) return b;
        if (mp[0].left || mp[1].left || mp[2].left || mp[3].left || mp[4].left || mp[5].left) {
            p.left = 1;
            p.left_time = tc;
            return tc;
        }
        double er = 0.0;
        for (int c = 0; c < 3; c++) {
            const double ec = fabs(a7 * k1[c] + b7 * k3[c] + c7 * k4[c] + d7 * k5[c] + e7 * k6[c]);
            er = c == 0 ? ec : fmax(er, ec);
        }
        double ss = er == 0.0 ? 0.0 : fvc_root4_rn(fabs((tol * h) / (2 * er)));
        if (er < tol || fabs(h) <= hmin) {
            double inc[3];
            for (int c = 0; c < 3; c++) {
                inc[c] = (a8 * k1[c] + b8 * k3[c] + c8 * k4[c] + d8 * k5[c]) * h;
                p.X[c] = p.X[c] + inc[c];
            }
            tc = tc + h;
            nsteps++;
            bool out;
            if (UNSTRUCT) {
                double r, s, t;
                const int e = walk<MODE, DIM>(m, p.X[0], p.X[1], p.X[2], p.elem, ACC_SEARCH, r, s, t, err);
                if (e == -2) return b;
                p.elem = e;
                out = e < 0;
            } else {
                out = outside_domain(m, p.X[0], p.X[1], p.X[2]);
            }
            if (out) {
                p.left = 1;
                p.left_time = tc - h;
                for (int c = 0; c < 3; c++) p.X[c] = p.X[c] - inc[c];
                return p.left_time;
            }
            if (extrap) {
                get_velocity<MODE, DIM>(m, f, tc, p, p.V, err);
                if (err) return b;
            }
            if (cet) {
                double X0[3];
                for (int c = 0; c < 3; c++) X0[c] = p.X[c] - inc[c];
                cet_step<DIM>(m, f, X0, p.X, p.elem, h, err);
                if (err) return b;
            }
            if (ap) {
                const double sr1 = strain_rate<DIM, UNSTRUCT>(m, f, tc, p, err);
                if (err) return b;
                p.scalar += (h * ((sr0 + sr1) / 2));
            }
            if (fabs(h) > hmin && ss > 1.0) {
                ss = ss > 4.0 ? 4.0 : ss;
                h = sign_of(fabs(ss * h) > hmax ? hmax : ss * h, b - a);
            }
        } else if (ss < 1.0) {
            ss = ss < 0.1 ? 0.1 : ss;
            h = sign_of(fabs(ss * h) < hmin ? hmin : ss * h, b - a);
        } else {
            h = sign_of(fabs(0.5 * h) < hmin ? hmin : 0.5 * h, b - a);
        }
    }
    return b;
}

/* Advect (ompVC/integration.c:22-45) over [t1, t2] in one go */
template <int MODE, int DIM, int UNSTRUCT, int GEN = 1>
__device__ __forceinline__ double advect(const DevMesh &m, const DevField &f, PState &p, double t1, double t2, int itype,
                                         double dt, int extrap, int &err, uint32_t &nsteps) {
#pragma clang fp contract(off)
    if (p.left && extrap) {
        for (int c = 0; c < 3; c++) p.X[c] = p.X[c] + p.V[c] * (t2 - t1);
        return t2;
    }
    double stop = t2;
    if (GEN && itype == INT_RKF) {
        stop = rkf_run<MODE, DIM, UNSTRUCT>(m, f, p, t1, t2, extrap, err, nsteps);
    } else {
        IState is;
        is.started = 0;
        interval_run<MODE, DIM, UNSTRUCT, GEN>(m, f, p, is, t1, t2, dt, itype, extrap, 0x7fffffff, err, nsteps, stop);
    }
    if ((fabs(stop - t2) > TINY) && extrap)
        for (int c = 0; c < 3; c++) p.X[c] = p.X[c] + p.V[c] * (t2 - stop);
    return stop;
}


/* ------------------------------------------------------------------------ */
/* largest eigenvalue of a symmetric 3x3, cyclic Jacobi rotations (FP64)     */
/* (Golub & Van Loan, Matrix Computations, sec. 8.5)                          */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ double max_eig_sym3(double a[3][3], int &err) {
#pragma clang fp contract(off)
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++) a[i][j] = a[j][i] = 0.5 * (a[i][j] + a[j][i]);
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        double diag = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
        if (off == 0.0 || off <= 1e-300 || off < 1e-18 * diag) break;
        if (sweep == 63) err = E_EIG;
#pragma unroll
        for (int pq = 0; pq < 3; pq++) {
            int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            double apq = a[p][q];
            if (apq == 0.0) continue;
            double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            if (!isfinite(theta) || fabs(theta) > 1e150) t = 0.5 / theta;
            double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
            double app = a[p][p], aqq = a[q][q];
            a[p][p] = app - t * apq;
            a[q][q] = aqq + t * apq;
            a[p][q] = a[q][p] = 0.0;
            int r = 3 - p - q;
            double arp = a[r][p], arq = a[r][q];
            a[r][p] = a[p][r] = c * arp - s * arq;
            a[r][q] = a[q][r] = s * arp + c * arq;
        }
    }
    double m = a[0][0] > a[1][1] ? a[0][0] : a[1][1];
    return m > a[2][2] ? m : a[2][2];
}

} // namespace fvc
