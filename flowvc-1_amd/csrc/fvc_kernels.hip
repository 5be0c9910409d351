/*
 * fvc_kernels.hip - gfx950 kernels of the advection path and their launchers.
 *
 *  k_advect        Advect over an interval, one lane per particle, in place
 *                  (the loop of ompVC/tracers.c:224-229).
 *  k_launch_vel    GetVelocity at release time with its side effects.
 *  k_ftle_*        one frame of ComputeFTLE for a slide: staged advection,
 *                  early FTLE for exits, update, gradient + largest eigenvalue.
 *  k_locate_*      walk / spatial-hash global search (lowest passing index).
 *  k_seed_*        exact seed-location chains: a parallel pass finds, per
 *                  point, the lowest containing element and whether more than
 *                  one element contains it; one wave then replays the
 *                  reference's sequential chain, walking only where the
 *                  result can depend on the chain's guess.
 *  k_morton/k_rs_* Z-order keys and a stable LSD radix sort (wave64 ballot
 *                  ranking).
 *  k_cmp_*         stable stream compaction (ballot + block prefix).
 * Determinism: no output depends on atomic ordering; counters are integer sums.
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fvc_internal.h"

namespace fvc {

constexpr int BLOCK = 256;

/* ------------------------------------------------------------------------ */
/* helpers                                                                   */
/* ------------------------------------------------------------------------ */
/* The dispatcher deals workgroups to the eight XCDs round-robin (block b runs
 * on XCD b % 8).  Renumbering so that XCD x takes the x-th contiguous eighth
 * of the grid keeps the blocks in flight on one XCD on neighbouring,
 * Z-ordered particles, so one L2 holds their mesh and velocity gathers
 * instead of all eight L2s holding the whole in-flight region.  Bijective on
 * [0, gridDim.x). */
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t G = gridDim.x, b = blockIdx.x;
    const int64_t per = G >> 3, rem = G & 7, x = b & 7, k = b >> 3;
    return x * per + (x < rem ? x : rem) + k;
}

__device__ inline void add_steps(unsigned long long *ctr, uint32_t ns) {
    __shared__ unsigned int red[BLOCK / 64];
    unsigned long long v = ns;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = (unsigned int)v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[q];
        if (s) atomicAdd(ctr, s);
    }
}

/* add_steps, plus the number of the block's lanes whose particle was not
 * integrated into ctr[1] */
__device__ inline void add_steps_idle(unsigned long long *ctr, uint32_t ns, bool idle) {
    __shared__ unsigned int red[BLOCK / 64], ired[BLOCK / 64];
    unsigned long long v = ns;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const unsigned int ni = (unsigned int)__popcll(__ballot(idle));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[w] = (unsigned int)v;
        ired[w] = ni;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0, si = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
            s += red[q];
            si += ired[q];
        }
        if (s) atomicAdd(ctr, s);
        if (si) atomicAdd(ctr + 1, si);
    }
}

/* a particle is integrated when it started inside the domain and has not
 * left it, or left and is extrapolated (ompVC/tracers.c:109, 225) */
__device__ __forceinline__ bool particle_live(const PtrSet &p, int64_t i, int extrap) {
    return p.left_time[i] > -TINY && (!p.left[i] || extrap);
}

__device__ inline void load_state(const PtrSet &p, int64_t i, PState &s, bool hist = false, bool vel = true) {
    for (int c = 0; c < 3; c++) {
        s.X[c] = p.X[c][i];
        s.V[c] = vel ? p.V[c][i] : 0.0;
    }
    s.elem = p.elem[i];
    s.left = p.left[i];
    s.left_time = p.left_time[i];
    s.nhist = 0;
    s.scalar = p.scalar ? p.scalar[i] : 0.0;
    if (hist) {
        for (int c = 0; c < 3; c++) {
            s.F1[c] = p.H[c][i];
            s.F2[c] = p.H[3 + c][i];
        }
        s.nhist = p.nhist[i];
    }
}

__device__ inline void store_state(const PtrSet &p, int64_t i, const PState &s, bool hist = false, bool vel = true) {
    for (int c = 0; c < 3; c++) {
        p.X[c][i] = s.X[c];
        if (vel) p.V[c][i] = s.V[c];
    }
    p.elem[i] = s.elem;
    p.left[i] = s.left;
    p.left_time[i] = s.left_time;
    if (p.scalar) p.scalar[i] = s.scalar;
    if (hist) {
        for (int c = 0; c < 3; c++) {
            p.H[c][i] = s.F1[c];
            p.H[3 + c][i] = s.F2[c];
        }
        p.nhist[i] = s.nhist;
    }
}

/* ------------------------------------------------------------------------ */
/* advect / launch velocity                                                   */
/* ------------------------------------------------------------------------ */
template <int MODE, int DIM, int UNS, int GEN>
__device__ __forceinline__ void advect_body(const DevMesh &m, const DevField &f, const PtrSet &p,
                                            const int *__restrict__ index, int64_t begin, int64_t count, double t1,
                                            double t2, int itype, double dt, int extrap, ErrBuf eb,
                                            unsigned long long *steps) {
    int64_t q = xcd_block() * blockDim.x + threadIdx.x;
    uint32_t ns = 0;
    bool idle = false;
    if (q < count) {
        int64_t i = begin + (index ? (int64_t)index[q] : q);
        if (!particle_live(p, i, extrap)) {
            idle = true;
        } else {
            const bool hist = GEN && itype >= 3;
            PState s;
            load_state(p, i, s, hist, extrap);
            int err = 0;
            advect<MODE, DIM, UNS, GEN>(m, f, s, t1, t2, itype, dt, extrap, err, ns);
            if (err)
                raise_err(eb, err, i);
            else
                store_state(p, i, s, hist, extrap);
        }
    }
    add_steps_idle(steps, ns, idle);
}

/* RK4 without extrapolation: no multistep state or particle velocity in
 * registers (the launchers send RK4 with extrapolation to k_advect_gen).
 * The reduced-precision state machine on unstructured meshes fits 4
 * waves/SIMD (at most 128 VGPRs) without spills; the FP64 and Cartesian
 * builds are held to 3 (168), which hides more of the gather latency than
 * their natural allocation.  The 3-D FP64 Cartesian build spills about 80
 * VGPRs at 3 waves and still runs 16% faster than at 2 waves without spills
 * (2M particles on a 128^3 grid: 0.283 s against 0.336 s). */
#define FVC_RK4_WAVES(MODE, UNS) ((MODE) != P_F64 && (UNS) ? 4 : FVC_HOT_WAVES)
template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(FVC_RK4_WAVES(MODE, UNS)))) void k_advect(
    DevMesh m, DevField f, PtrSet p, const int *__restrict__ index, int64_t begin, int64_t count, double t1, double t2,
    int itype, double dt, int extrap, ErrBuf eb, unsigned long long *steps) {
    (void)extrap;
    advect_body<MODE, DIM, UNS, 0>(m, f, p, index, begin, count, t1, t2, itype, dt, 0, eb, steps);
}

/* Euler, AB2, AB3 */
template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_advect_gen(DevMesh m, DevField f, PtrSet p, const int *__restrict__ index,
                                                     int64_t begin, int64_t count, double t1, double t2, int itype,
                                                     double dt, int extrap, ErrBuf eb, unsigned long long *steps) {
    advect_body<MODE, DIM, UNS, 1>(m, f, p, index, begin, count, t1, t2, itype, dt, extrap, eb, steps);
}

/* Each particle over its own segment: tt[2q], tt[2q+1] = t1, t2 of particle
 * begin+q, skipped where t1 is NaN (staggered release, tracers.c:336-397). */
template <int MODE, int DIM, int UNS, int GEN>
__device__ __forceinline__ void k_advect_times_body(const DevMesh &m, const DevField &f, const PtrSet &p, int64_t begin,
                                                    int64_t count, const double *__restrict__ tt, int itype, double dt,
                                                    int extrap, ErrBuf eb, unsigned long long *steps) {
    const int64_t q = xcd_block() * blockDim.x + threadIdx.x;
    uint32_t ns = 0;
    if (q < count) {
        const double t1 = tt[2 * q], t2 = tt[2 * q + 1];
        const int64_t i = begin + q;
        if (t1 == t1 && p.left_time[i] > -TINY && (!p.left[i] || extrap)) {
            const bool hist = GEN && itype >= 3;
            PState s;
            load_state(p, i, s, hist, extrap);
            int err = 0;
            advect<MODE, DIM, UNS, GEN>(m, f, s, t1, t2, itype, dt, extrap, err, ns);
            if (err)
                raise_err(eb, err, i);
            else
                store_state(p, i, s, hist, extrap);
        }
    }
    add_steps(steps, ns);
}

template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_launch_vel(DevMesh m, DevField f, PtrSet p, int64_t begin, int64_t count,
                                                     double t, ErrBuf eb) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    int64_t i = begin + q;
    if (p.left[i]) return;
    PState s;
    load_state(p, i, s);
    int err = 0;
    double v[3];
    get_velocity<MODE, DIM>(m, f, t, s, v, err);
    if (err) {
        raise_err(eb, err, i);
        return;
    }
    for (int c = 0; c < 3; c++) s.V[c] = v[c];
    store_state(p, i, s);
}

#define FVC_DISPATCH(MODE_, DIM_, UNS_, KERNEL, ...)                                                                   \
    do {                                                                                                               \
        if (UNS_) {                                                                                                    \
            if (DIM_ == 3) {                                                                                           \
                switch (MODE_) {                                                                                       \
                case P_F64: KERNEL<P_F64, 3, 1> __VA_ARGS__; break;                                                     \
                case P_F32: KERNEL<P_F32, 3, 1> __VA_ARGS__; break;                                                     \
                case P_BF16: KERNEL<P_BF16, 3, 1> __VA_ARGS__; break;                                                   \
                default: KERNEL<P_F16, 3, 1> __VA_ARGS__; break;                                                        \
                }                                                                                                      \
            } else {                                                                                                   \
                switch (MODE_) {                                                                                       \
                case P_F64: KERNEL<P_F64, 2, 1> __VA_ARGS__; break;                                                     \
                case P_F32: KERNEL<P_F32, 2, 1> __VA_ARGS__; break;                                                     \
                case P_BF16: KERNEL<P_BF16, 2, 1> __VA_ARGS__; break;                                                   \
                default: KERNEL<P_F16, 2, 1> __VA_ARGS__; break;                                                        \
                }                                                                                                      \
            }                                                                                                          \
        } else {                                                                                                       \
            if (DIM_ == 3) {                                                                                           \
                switch (MODE_) {                                                                                       \
                case P_F64: KERNEL<P_F64, 3, 0> __VA_ARGS__; break;                                                     \
                case P_F32: KERNEL<P_F32, 3, 0> __VA_ARGS__; break;                                                     \
                case P_BF16: KERNEL<P_BF16, 3, 0> __VA_ARGS__; break;                                                   \
                default: KERNEL<P_F16, 3, 0> __VA_ARGS__; break;                                                        \
                }                                                                                                      \
            } else {                                                                                                   \
                switch (MODE_) {                                                                                       \
                case P_F64: KERNEL<P_F64, 2, 0> __VA_ARGS__; break;                                                     \
                case P_F32: KERNEL<P_F32, 2, 0> __VA_ARGS__; break;                                                     \
                case P_BF16: KERNEL<P_BF16, 2, 0> __VA_ARGS__; break;                                                   \
                default: KERNEL<P_F16, 2, 0> __VA_ARGS__; break;                                                        \
                }                                                                                                      \
            }                                                                                                          \
        }                                                                                                              \
    } while (0)

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

static ErrBuf errbuf(fvc_ctx *c) { return ErrBuf{c->d_err, c->d_where}; }

__global__ __launch_bounds__(BLOCK) void k_cet_encounters(PtrSet p, int64_t begin, int64_t count, int *enc) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int64_t i = begin + q;
    if (!p.left[i] && p.elem[i] >= 0) atomicAdd(enc + p.elem[i], 1);
}

int launch_cet_encounters(fvc_ctx *c, PtrSet p, int64_t begin, int64_t count, int *enc) {
    if (count <= 0) return FVC_OK;
    k_cet_encounters<<<grid_for(count, BLOCK), BLOCK, 0, c->stream>>>(p, begin, count, enc);
    FVC_HIP(hipGetLastError());
    return FVC_OK;
}

int launch_velocity(fvc_ctx *c, PtrSet p, int64_t begin, int64_t count, double t, const DevField &f) {
    if (count <= 0) return FVC_OK;
    FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_launch_vel,
                 <<<grid_for(count, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, f, p, begin, count, t, errbuf(c)));
    FVC_HIP(hipGetLastError());
    return check_err(c, "fvc_launch_velocity");
}

template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_advect_times_rk4(DevMesh m, DevField f, PtrSet p, int64_t begin,
                                                           int64_t count, const double *__restrict__ tt, int itype,
                                                           double dt, int extrap, ErrBuf eb, unsigned long long *steps);
template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_advect_times_gen(DevMesh m, DevField f, PtrSet p, int64_t begin,
                                                           int64_t count, const double *__restrict__ tt, int itype,
                                                           double dt, int extrap, ErrBuf eb, unsigned long long *steps);

int launch_advect_times(fvc_ctx *c, PtrSet p, int64_t begin, int64_t count, const double *dtt, const DevField &f,
                        const fvc_int_params &ip) {
    FVC_HIP(hipMemsetAsync(c->d_steps, 0, sizeof(unsigned long long), c->stream));
    if (count > 0) {
        if (ip.int_type == FVC_INT_RK4 && !f.sr_lo && !f.cet_lo)
            FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_advect_times_rk4,
                         <<<grid_for(count, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, f, p, begin, count, dtt,
                                                                           ip.int_type, ip.time_step, ip.extrapolate,
                                                                           errbuf(c), c->d_steps));
        else
            FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_advect_times_gen,
                         <<<grid_for(count, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, f, p, begin, count, dtt,
                                                                           ip.int_type, ip.time_step, ip.extrapolate,
                                                                           errbuf(c), c->d_steps));
        FVC_HIP(hipGetLastError());
    }
    int rc = check_err(c, "fvc_advect_times");
    c->last_steps = (int64_t)c->h_status->steps[0];
    return rc;
}

template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_advect_times_rk4(DevMesh m, DevField f, PtrSet p, int64_t begin,
                                                           int64_t count, const double *__restrict__ tt, int itype,
                                                           double dt, int extrap, ErrBuf eb, unsigned long long *steps) {
    k_advect_times_body<MODE, DIM, UNS, 0>(m, f, p, begin, count, tt, itype, dt, extrap, eb, steps);
}
template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_advect_times_gen(DevMesh m, DevField f, PtrSet p, int64_t begin,
                                                           int64_t count, const double *__restrict__ tt, int itype,
                                                           double dt, int extrap, ErrBuf eb, unsigned long long *steps) {
    k_advect_times_body<MODE, DIM, UNS, 1>(m, f, p, begin, count, tt, itype, dt, extrap, eb, steps);
}

__global__ __launch_bounds__(BLOCK) void k_live_flags(PtrSet p, int64_t begin, int64_t count, int extrap,
                                                      int *__restrict__ flags) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < count) flags[q] = particle_live(p, begin + q, extrap) ? 1 : 0;
}

/* indices (relative to begin) of the particles of [begin, begin + count)
 * still to integrate, in order */
int launch_active_list(fvc_ctx *c, PtrSet p, int64_t begin, int64_t count, int extrap, int *list, int64_t *n) {
    *n = 0;
    if (count <= 0) return FVC_OK;
    int *flags = (int *)scratch(c, sizeof(int) * (size_t)count);
    if (!flags) return FVC_E_NOMEM;
    k_live_flags<<<grid_for(count, BLOCK), BLOCK, 0, c->stream>>>(p, begin, count, extrap, flags);
    FVC_HIP(hipGetLastError());
    return launch_compact(c, flags, count, list, n, nullptr);
}

/* index: null for every particle of [begin, begin + count), else `count`
 * indices relative to begin */
int launch_advect(fvc_ctx *c, PtrSet p, const int *index, int64_t begin, int64_t count, double t1, double t2,
                  const DevField &f, const fvc_int_params &ip) {
    FVC_HIP(hipMemsetAsync(c->d_steps, 0, 2 * sizeof(unsigned long long), c->stream));
    if (count > 0) {
        int block = ip.block > 0 ? ip.block : BLOCK;
        if (block != BLOCK) block = BLOCK; /* kernels are compiled for 256-lane blocks */
        timing_begin(c);
        /* strain rate, exposure, extrapolation: generic build only */
        if (ip.int_type == FVC_INT_RK4 && !ip.extrapolate && !f.sr_lo && !f.cet_lo)
            FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_advect,
                         <<<grid_for(count, block), block, 0, c->stream>>>(c->mesh, f, p, index, begin, count, t1, t2,
                                                                           ip.int_type, ip.time_step, ip.extrapolate,
                                                                           errbuf(c), c->d_steps));
        else
            FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_advect_gen,
                         <<<grid_for(count, block), block, 0, c->stream>>>(c->mesh, f, p, index, begin, count, t1, t2,
                                                                           ip.int_type, ip.time_step, ip.extrapolate,
                                                                           errbuf(c), c->d_steps));
        timing_end(c);
        FVC_HIP(hipGetLastError());
    }
    int rc = check_err(c, "fvc_advect");
    c->last_steps = (int64_t)c->h_status->steps[0];
    c->last_idle = (int64_t)c->h_status->steps[1];
    if (count > 0) timing_collect(c, c->last_steps);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* location                                                                   */
/* ------------------------------------------------------------------------ */
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_locate(DevMesh m, const double *__restrict__ xyz, const int *__restrict__ guess,
                                                 int64_t n, int mode, int *__restrict__ out, ErrBuf eb) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = DIM == 3 ? xyz[3 * i + 2] : 0.0;
    int res = -1;
    if (mode != FVC_LOCATE_GLOBAL) {
        int g = guess[i];
        if (g >= 0 && g < m.ne) {
            int err = 0;
            res = local_search<DIM>(m, x, y, z, g, err);
            if (err) {
                raise_err(eb, err, i);
                res = -1;
            }
        }
    }
    if (mode == FVC_LOCATE_GLOBAL || (mode == FVC_LOCATE_WALK_GLOBAL && res < 0)) {
        int second;
        res = hash_global<DIM>(m, x, y, z, second);
    }
    out[i] = res;
}

int launch_locate(fvc_ctx *c, const double *dxyz, const int *dguess, int64_t n, int mode, int *dout) {
    if (n <= 0) return FVC_OK;
    if (c->mesh.dims == 3)
        k_locate<3><<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, dxyz, dguess, n, mode, dout, errbuf(c));
    else
        k_locate<2><<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, dxyz, dguess, n, mode, dout, errbuf(c));
    FVC_HIP(hipGetLastError());
    return check_err(c, "fvc_locate");
}

/* ------------------------------------------------------------------------ */
/* seed chains                                                                */
/* ------------------------------------------------------------------------ */
/* per-point chain flags: outside the bounding box, inside more than one
 * element, and (grid chains) row and plane starts and the row / plane ends
 * where the guess is reset */
enum { SF_OUTB = 1, SF_MULTI = 2, SF_ROWSTART = 4, SF_PLANESTART = 8, SF_ROWEND = 16, SF_PLANEEND = 32 };

struct ChainGeom {
    int chain;
    const double *xyz; /* list chain */
    double mn[3], del[3];
    int res[3];
};

/* Chains 2 and 3 walk rows and planes and reset the guess at row ends.
 * Chain 3, the velocity output grid (x fastest, velout.c:58-163), sets it to
 * the row's first found element after every row, then to the plane's after
 * every plane.  Chain 2, the FTLE grid (z fastest, ftle.c:456-601), sets it
 * only after each plane, to the first found element of the plane's last row
 * (its plane guess is applied after the whole grid, when no point follows). */
__host__ __device__ inline int64_t chain_rowlen(const ChainGeom &g) { return g.chain == 2 ? g.res[2] : g.res[0]; }

/* chain-order point i -> coordinates (ftle.c:412-420, tracers.c:1344-1346, velout.c:67-69) */
__device__ inline void chain_point(const ChainGeom &g, int64_t i, double X[3]) {
#pragma clang fp contract(off)
    if (g.chain == 0) {
        X[0] = g.xyz[3 * i];
        X[1] = g.xyz[3 * i + 1];
        X[2] = g.xyz[3 * i + 2];
        return;
    }
    int ii, jj, kk;
    if (g.chain == 1 || g.chain == 3) { /* k, j, i order: x fastest */
        ii = (int)(i % g.res[0]);
        jj = (int)((i / g.res[0]) % g.res[1]);
        kk = (int)(i / ((int64_t)g.res[0] * g.res[1]));
    } else { /* FTLE: i, j, k order: z fastest */
        kk = (int)(i % g.res[2]);
        jj = (int)((i / g.res[2]) % g.res[1]);
        ii = (int)(i / ((int64_t)g.res[2] * g.res[1]));
    }
    X[0] = g.mn[0] + ii * g.del[0];
    X[1] = g.mn[1] + jj * g.del[1];
    X[2] = g.mn[2] + kk * g.del[2];
}

/* bits 8-15 of the flags: number of elements holding the point (capped at 255) */
constexpr int SF_CNT_SHIFT = 8;
__device__ __forceinline__ int sf_count(int flags) { return (flags >> SF_CNT_SHIFT) & 255; }

/* bookkeeping for the tabulated chain (all null for the replayed chain) */
struct TabPrep {
    int *mflag;      /* 1 where a found point lies in more than one element */
    int *misc;       /* [0] first found point, [1] largest element count */
    int *rowfirst;   /* grid chains: first found point of each row */
    int *planefirst; /* grid chains: first found point of each plane */
};

template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_seed_prepare(DevMesh m, ChainGeom g, int64_t n, int2 *__restrict__ info,
                                                       TabPrep tp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n; /* no early return: the bookkeeping below uses whole-wave ballots */
    int flags = 0, emin = -1, cnt = 0;
    if (active) {
        double X[3];
        chain_point(g, i, X);
        if (g.chain >= 2) {
            const int64_t rl = chain_rowlen(g);
            const int kk = (int)(i % rl), jj = (int)((i / rl) % g.res[1]);
            if (kk == 0) flags |= SF_ROWSTART | (jj == 0 ? SF_PLANESTART : 0);
            if (kk == rl - 1) {
                if (g.chain == 3)
                    flags |= SF_ROWEND | (jj == g.res[1] - 1 ? SF_PLANEEND : 0);
                else if (jj == g.res[1] - 1) /* FTLE: one reset per plane, from its last row */
                    flags |= SF_ROWEND;
            }
        }
        if (outside_domain(m, X[0], X[1], X[2]))
            flags |= SF_OUTB;
        else if (m.mesh_type == 1) {
            cnt = hash_accepting<DIM>(m, X[0], X[1], X[2], nullptr, 0, emin);
            if (cnt > 1) flags |= SF_MULTI;
        }
        info[i] = make_int2(emin, flags | (min(cnt, 255) << SF_CNT_SHIFT));
    }
    if (!tp.mflag) return;
    const bool found = active && !(flags & SF_OUTB) && emin >= 0;
    if (active) tp.mflag[i] = (found && cnt > 1) ? 1 : 0;
    int wmax = cnt;
    for (int off = 32; off; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    const unsigned long long F = __ballot(found);
    const int lane = threadIdx.x & 63;
    if (F && lane == __ffsll((long long)F) - 1) {
        atomicMin(tp.misc, (int)i);
        atomicMax(tp.misc + 1, wmax);
    }
    if (g.chain >= 2 && found) {
        /* one atomic per row (plane) and wave: the wave's first found point in it */
        const unsigned long long below = F & ((1ull << lane) - 1ull);
        const int64_t prev = below ? i - lane + (63 - __clzll(below)) : -1;
        const int64_t rowlen = chain_rowlen(g), planelen = rowlen * g.res[1];
        if (prev < 0 || prev / rowlen != i / rowlen) atomicMin(tp.rowfirst + i / rowlen, (int)i);
        if (prev < 0 || prev / planelen != i / planelen) atomicMin(tp.planefirst + i / planelen, (int)i);
    }
}

/* One wave replays the chain over chunks of 64 consecutive points.  A chunk
 * whose results cannot depend on the chain state (every point outside, or
 * inside exactly one element with LocalSearchChecking, after the first found
 * point, and no row boundary of a grid chain) is resolved by all lanes at once and the
 * state (last found element, first found element of the row and plane) is
 * advanced with ballots; any other chunk is replayed point by point. */
constexpr int CHAIN_STAGE = 1024; /* points staged through LDS per round */

template <int DIM>
__global__ __launch_bounds__(CHAIN_STAGE) void k_seed_chain(DevMesh m, ChainGeom g, int64_t n,
                                                           const int2 *__restrict__ info, int lsc, int seed0,
                                                           int *__restrict__ out_elem, int *__restrict__ out_left,
                                                           double *__restrict__ out_lt, ErrBuf eb) {
    /* The block stages CHAIN_STAGE points through LDS.  Every wave resolves
     * its own chunk of 64 as far as the state allows (results that cannot
     * depend on the chain, and ballot masks of the events that move the
     * state); wave 0 then walks the chunks in order, advancing the state from
     * the masks alone where the chunk allowed it and replaying the chunk
     * point by point where it did not. */
    __shared__ int2 s_info[CHAIN_STAGE];
    __shared__ int s_elem[CHAIN_STAGE], s_left[CHAIN_STAGE];
    __shared__ double s_lt[CHAIN_STAGE];
    __shared__ unsigned long long s_mask[CHAIN_STAGE / 64][5]; /* hits, row end, row start, plane start, plane end */
    __shared__ int s_flag[CHAIN_STAGE / 64][2];                /* chunk independent of the state, all outside */
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool uns = m.mesh_type == 1;
    int found = 0;                 /* list chain */
    int seed = seed0, guess = seed0 >= 0 ? seed0 : -1;
    int iguess = -1, jguess = -1;  /* grid chains */
    for (int64_t stage = 0; stage < n; stage += CHAIN_STAGE) {
        {
            const int64_t i = stage + threadIdx.x;
            const bool active = i < n;
            const int2 my = active ? info[i] : make_int2(-1, SF_OUTB);
            s_info[threadIdx.x] = my;
            const bool outb = (my.y & SF_OUTB) != 0;
            const bool indep = outb || !uns || (lsc && !(my.y & SF_MULTI));
            int left = outb ? 1 : 0, elem = -1;
            double lt = 0.0;
            bool hit = false;
            if (active && !left && uns) {
                if (my.x < 0) {
                    left = 1;
                    lt = -1.0;
                } else {
                    elem = my.x;
                    hit = true;
                }
            }
            if (g.chain == 0) lt = left ? -1.0 : 0.0;
            if (g.chain >= 2 && left) lt = -1.0;
            s_elem[threadIdx.x] = elem;
            s_left[threadIdx.x] = left;
            s_lt[threadIdx.x] = lt;
            const unsigned long long hits = __ballot(hit);
            const unsigned long long rowend = __ballot(active && (my.y & SF_ROWEND));
            const unsigned long long rowstart = __ballot(active && (my.y & SF_ROWSTART));
            const unsigned long long planestart = __ballot(active && (my.y & SF_PLANESTART));
            const unsigned long long planeend = __ballot(active && (my.y & SF_PLANEEND));
            const bool chunk_indep = __ballot(active && !indep) == 0ull;
            const bool all_outb = __ballot(active && !(outb || !uns)) == 0ull;
            if (lane == 0) {
                s_mask[wv][0] = hits;
                s_mask[wv][1] = rowend;
                s_mask[wv][2] = rowstart;
                s_mask[wv][3] = planestart;
                s_mask[wv][4] = planeend;
                s_flag[wv][0] = chunk_indep;
                s_flag[wv][1] = all_outb;
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            for (int64_t base = stage; base < n && base < stage + CHAIN_STAGE; base += 64) {
                const int ck = (int)((base - stage) >> 6);
                const int2 my = s_info[base - stage + lane];
                const bool state_ok = g.chain == 0 ? found > 0 : seed >= 0;
                if (s_flag[ck][0] && (state_ok || s_flag[ck][1])) {
                    /* advance the state segment by segment: a segment ends at a
                     * row-end event (grid chains) or just before a row start, so a
                     * row start can only open one */
                    const unsigned long long hits = s_mask[ck][0], rowend = s_mask[ck][1],
                                             rowstart = s_mask[ck][2], planestart = s_mask[ck][3],
                                             planeend = s_mask[ck][4];
                    const int *ce = s_elem + (base - stage);
                    int a = 0;
                    while (a < 64) {
                        const unsigned long long from_a = ~0ull << a;
                        const unsigned long long re = rowend & from_a;
                        const unsigned long long rs = rowstart & (from_a << 1); /* row starts after a */
                        int b = re ? __ffsll((long long)re) - 1 : 63;
                        if (rs) b = min(b, __ffsll((long long)rs) - 2);
                        const unsigned long long seg = (b == 63 ? ~0ull : ((1ull << (b + 1)) - 1ull)) & from_a;
                        if ((rowstart >> a) & 1ull) {
                            jguess = -1;
                            if ((planestart >> a) & 1ull) iguess = -1;
                        }
                        const unsigned long long sh = hits & seg;
                        if (sh) {
                            const int last = 63 - __clzll(sh), first = __ffsll((long long)sh) - 1;
                            guess = ce[last];
                            if (g.chain == 0) found += __popcll(sh);
                            if (g.chain >= 2) {
                                const int e1 = ce[first];
                                if (jguess < 0) jguess = e1;
                                if (iguess < 0) iguess = e1;
                            }
                        }
                        if ((rowend >> b) & 1ull) {
                            if (jguess >= 0) guess = jguess;
                            if (((planeend >> b) & 1ull) && iguess >= 0) guess = iguess;
                        }
                        a = b + 1;
                    }
                } else {
                    /* Points whose result depends on the guess: every lane walks at
                     * once from the guess the chain would hand it given the other
                     * lanes' current results (a 64-step scan of the chain state,
                     * row and plane resets included), and walks again while that
                     * guess changes.  Lane l is final once lanes < l are, so this
                     * ends within 64 rounds, usually 2-3, with the sequential
                     * results.  Before the chain is seeded, or where a walk
                     * fails, the chunk is replayed point by point. */
                    const bool state_ok2 = g.chain == 0 ? found > 0 : seed >= 0;
                    bool done = false;
                    if (uns && state_ok2) {
                        const bool active = base + lane < n;
                        const bool outb = !active || (my.y & SF_OUTB);
                        const bool walker = !outb && !(lsc && !(my.y & SF_MULTI));
                        const int lim = (int)min((int64_t)64, n - base);
                        double X[3] = {0.0, 0.0, 0.0};
                        if (walker) chain_point(g, base + lane, X);
                        int r = outb ? -1 : my.x; /* direct lanes: final; walkers: first estimate */
                        int used = -3;
                        bool bad = false;
                        int gs = guess, jg = jguess, ig = iguess, fc = found;
                        for (int round = 0;; round++) {
                            /* the chain state before each lane, from the current results */
                            gs = guess;
                            jg = jguess;
                            ig = iguess;
                            fc = found;
                            int my_g = -1;
                            for (int l = 0; l < lim; l++) {
                                const int fl = __shfl(my.y, l, 64);
                                const int rl = __shfl(r, l, 64);
                                if (g.chain >= 2 && (fl & SF_ROWSTART)) {
                                    jg = -1;
                                    if (fl & SF_PLANESTART) ig = -1;
                                }
                                if (lane == l) my_g = gs;
                                if (!(fl & SF_OUTB) && rl >= 0) {
                                    gs = rl;
                                    fc++;
                                    if (g.chain >= 2) {
                                        if (jg < 0) jg = rl;
                                        if (ig < 0) ig = rl;
                                    }
                                }
                                if (g.chain >= 2 && (fl & SF_ROWEND)) {
                                    if (jg >= 0) gs = jg;
                                    if ((fl & SF_PLANEEND) && ig >= 0) gs = ig;
                                }
                            }
                            const bool need = walker && my_g != used;
                            if (__ballot(need) == 0ull) break;
                            if (round == 65) {
                                bad = true;
                                break;
                            }
                            if (need) {
                                int err = 0;
                                int idx = local_search<DIM>(m, X[0], X[1], X[2], my_g, err);
                                if (g.chain != 0 && idx < 0 && !err)
                                    idx = local_search<DIM>(m, X[0], X[1], X[2], seed, err);
                                if (idx < 0 && !err && lsc) idx = my.x;
                                if (err) bad = true;
                                r = idx;
                                used = my_g;
                            }
                            if (__ballot(bad)) {
                                bad = true;
                                break;
                            }
                        }
                        if (!bad) {
                            const bool left = outb || r < 0;
                            double lt = (!outb && r < 0) ? -1.0 : 0.0;
                            if (g.chain == 0) lt = left ? -1.0 : 0.0;
                            if (g.chain >= 2 && left) lt = -1.0;
                            s_elem[base - stage + lane] = left ? -1 : r;
                            s_left[base - stage + lane] = left ? 1 : 0;
                            s_lt[base - stage + lane] = lt;
                            guess = gs;
                            jguess = jg;
                            iguess = ig;
                            found = fc;
                            done = true;
                        }
                    }
                    if (!done) {
                    int r_elem = -1, r_left = 0;
                    double r_lt = 0.0;
                    const int lim = (int)min((int64_t)64, n - base);
                    for (int l = 0; l < lim; l++) {
                        const int64_t i = base + l;
                        const int emin = __shfl(my.x, l, 64), flags = __shfl(my.y, l, 64);
                        int elem = -1, left = (flags & SF_OUTB) ? 1 : 0;
                        double lt = 0.0;
                        if (g.chain >= 2 && (flags & SF_ROWSTART)) {
                            jguess = -1;
                            if (flags & SF_PLANESTART) iguess = -1;
                        }
                        if (!left && uns) {
                            const bool direct = lsc && !(flags & SF_MULTI); /* result cannot depend on the guess */
                            double X[3] = {0.0, 0.0, 0.0};
                            if (!direct) chain_point(g, i, X);
                            int err = 0, index = -1;
                            if (g.chain == 0) {
                                if (!found) {
                                    index = emin;
                                } else {
                                    index = direct ? emin : local_search<DIM>(m, X[0], X[1], X[2], guess, err);
                                    if (!direct && index < 0 && lsc) index = emin;
                                }
                                if (index < 0) {
                                    left = 1;
                                    elem = -1;
                                } else {
                                    elem = index;
                                    guess = index;
                                    found++;
                                }
                            } else {
                                if (seed < 0) {
                                    seed = emin;
                                    if (seed < 0) {
                                        left = 1;
                                        elem = -1;
                                        lt = -1.0;
                                    } else {
                                        elem = seed;
                                        guess = seed;
                                        if (g.chain >= 2) jguess = iguess = seed;
                                    }
                                } else {
                                    if (direct)
                                        index = emin;
                                    else {
                                        index = local_search<DIM>(m, X[0], X[1], X[2], guess, err);
                                        if (index < 0 && !err) index = local_search<DIM>(m, X[0], X[1], X[2], seed, err);
                                        if (index < 0 && !err && lsc) index = emin;
                                    }
                                    if (index < 0) {
                                        left = 1;
                                        elem = -1;
                                        lt = -1.0;
                                    } else {
                                        elem = index;
                                        guess = index;
                                        if (g.chain >= 2) {
                                            if (jguess < 0) jguess = index;
                                            if (iguess < 0) iguess = index;
                                        }
                                    }
                                }
                            }
                            if (err && lane == 0) raise_err(eb, err, i);
                        }
                        if (g.chain == 0) lt = left ? -1.0 : 0.0;
                        if (g.chain >= 2 && left) lt = -1.0;
                        if (g.chain >= 2 && (flags & SF_ROWEND)) {
                            if (jguess >= 0) guess = jguess;
                            if ((flags & SF_PLANEEND) && iguess >= 0) guess = iguess;
                        }
                        if (lane == l) {
                            r_elem = elem;
                            r_left = left;
                            r_lt = lt;
                        }
                    }
                    s_elem[base - stage + lane] = r_elem;
                    s_left[base - stage + lane] = r_left;
                    s_lt[base - stage + lane] = r_lt;
                    }
                }
            }
        }
        __syncthreads();
        {
            const int64_t i = stage + threadIdx.x;
            if (i < n) {
                out_elem[i] = s_elem[threadIdx.x];
                out_left[i] = s_left[threadIdx.x];
                out_lt[i] = s_lt[threadIdx.x];
            }
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------------ */
/* tabulated chains (LocalSearchChecking on, unstructured mesh)                */
/* ------------------------------------------------------------------------ */
/* With LocalSearchChecking a point inside the mesh always ends up found (its
 * walks fall back to the lowest holding element), so which points are found,
 * and hence which earlier point's element each guess comes from, is fixed by
 * the flags alone.  A point held by one element gets that element whatever
 * the guess.  A point held by several (a face, edge or vertex point) gets
 * whichever of them its walk reaches, a function of the guess; the guess is
 * the element of its source point, one of that point's few holding elements.
 * So every such point walks once from each holding element of its source
 * (all points at once), giving a small table, and the tables are composed
 * along the source links by pointer jumping.  The results equal the
 * point-by-point replay (tracers.c:1318-1346, ftle.c:456-516, mesh.c:496-611). */
constexpr int TAB_K = 32;     /* holding elements kept per point; more -> replay */
constexpr int TAB_SEED = -2;  /* source: the centre seed element */
constexpr int TAB_NONE = -3;  /* source: none (the first found point) */

/* Source of every point: the last event before it in chain order that set
 * the guess, as a point index.  A found point sets it to itself; a row-end
 * event of a grid chain sets it to the row's first found point, a plane-end
 * event to the plane's.
 * Within each 64-point chunk from ballots; -1 where no event precedes the
 * point inside its chunk.  chunk_last[c]: the chunk's last event, or -1. */
__global__ __launch_bounds__(BLOCK) void k_tab_src(ChainGeom g, int64_t n, const int2 *__restrict__ info,
                                                   const int *__restrict__ rowfirst,
                                                   const int *__restrict__ planefirst, int *__restrict__ src,
                                                   int *__restrict__ chunk_last) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool active = i < n;
    const int2 my = active ? info[i] : make_int2(-1, SF_OUTB);
    const bool found = active && !(my.y & SF_OUTB) && my.x >= 0;
    int rv = -1; /* row-end event value */
    if (g.chain >= 2 && active && (my.y & SF_ROWEND)) {
        const int64_t rowlen = chain_rowlen(g), planelen = rowlen * g.res[1];
        const int pf = planefirst[i / planelen], rf = rowfirst[i / rowlen];
        if ((my.y & SF_PLANEEND) && pf < n) rv = pf;
        else if (rf < n) rv = rf;
    }
    const unsigned long long F = __ballot(found), E = __ballot(rv >= 0);
    const int64_t base = i - lane;
    /* a found point's event comes before its own row-end event */
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const unsigned long long fb = F & below, eb = E & below;
    const int jf = fb ? 63 - __clzll(fb) : -1, je = eb ? 63 - __clzll(eb) : -1;
    const int rvj = __shfl(rv, je < 0 ? 0 : je, 64);
    if (active) src[i] = (je >= 0 && je >= jf) ? rvj : (jf >= 0 ? (int)(base + jf) : -1);
    const int cf = F ? 63 - __clzll(F) : -1, ce = E ? 63 - __clzll(E) : -1;
    const int rvc = __shfl(rv, ce < 0 ? 0 : ce, 64);
    if (lane == 0 && base < n) chunk_last[base >> 6] = (ce >= 0 && ce >= cf) ? rvc : (cf >= 0 ? (int)(base + cf) : -1);
}

/* carry[c]: the last event before chunk c (gs0 where there is none) */
__global__ __launch_bounds__(1024) void k_tab_carry(int64_t nch, const int *__restrict__ chunk_last, int gs0,
                                                    int *__restrict__ carry) {
    __shared__ int s[1024];
    __shared__ int s_run;
    const int tid = threadIdx.x;
    if (tid == 0) s_run = gs0;
    __syncthreads();
    for (int64_t b = 0; b < nch; b += 1024) {
        const int64_t c = b + tid;
        const int v = c < nch ? chunk_last[c] : -1;
        s[tid] = v >= 0 ? tid : -1;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) { /* inclusive prefix max of positions with an event */
            const int o = tid >= off ? s[tid - off] : -1;
            __syncthreads();
            s[tid] = max(s[tid], o);
            __syncthreads();
        }
        const int ex = tid ? s[tid - 1] : -1;
        if (c < nch) carry[c] = ex >= 0 ? chunk_last[b + ex] : s_run;
        __syncthreads();
        if (tid == 0 && s[1023] >= 0) s_run = chunk_last[b + s[1023]];
        __syncthreads();
    }
}

__global__ __launch_bounds__(BLOCK) void k_tab_slots(int64_t nm, const int *__restrict__ mlist, int *__restrict__ mslot) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nm) mslot[mlist[k]] = (int)k;
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_tab_accept(DevMesh m, ChainGeom g, int64_t nm, const int *__restrict__ mlist,
                                                      int *__restrict__ A) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    double X[3];
    chain_point(g, mlist[k], X);
    int first;
    hash_accepting<DIM>(m, X[0], X[1], X[2], A + k * TAB_K, TAB_K, first);
}

/* Walks of every multi-element point from each holding element of its
 * source: T[k][q] = index (into the point's holding list) of the element
 * reached from the source's q-th holding element, or an error code.  Points
 * whose source has one possible element are resolved here (par = -1). */
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_tab_walk(DevMesh m, ChainGeom g, int64_t n, int64_t nm,
                                                    const int2 *__restrict__ info, const int *__restrict__ mlist,
                                                    const int *__restrict__ mslot, const int *__restrict__ A,
                                                    const int *__restrict__ src, const int *__restrict__ carry,
                                                    int seed0, const int *__restrict__ misc, int *__restrict__ par,
                                                    int *__restrict__ val, signed char *__restrict__ T) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const int i = mlist[k];
    int s = src[i];
    if (s < 0) s = carry[i >> 6];
    double X[3];
    chain_point(g, i, X);
    const int *Ai = A + k * TAB_K;
    const int ci = sf_count(info[i].y);
    int seedv = seed0; /* walk fallback of the grid chains: the centre seed, else the first found element */
    if (seedv < 0) {
        const int f = misc[0];
        seedv = f < n ? info[f].x : -1;
    }
    auto tab = [&](int guess) -> int {
        int err = 0;
        int e = local_search<DIM>(m, X[0], X[1], X[2], guess, err);
        if (e < 0 && !err && g.chain != 0) e = local_search<DIM>(m, X[0], X[1], X[2], seedv, err);
        if (err) return err;
        if (e < 0) return 0; /* LocalSearchChecking: the lowest holding element */
        for (int q = 0; q < ci; q++)
            if (Ai[q] == e) return q;
        return E_INDET;
    };
    if (s == TAB_NONE) {
        par[k] = -1;
        val[k] = 0;
        return;
    }
    if (s == TAB_SEED) {
        par[k] = -1;
        val[k] = tab(seed0);
        return;
    }
    const int2 ps = info[s];
    const int cs = sf_count(ps.y);
    if (cs <= 1) {
        par[k] = -1;
        val[k] = tab(ps.x);
        return;
    }
    const int *As = A + (int64_t)mslot[s] * TAB_K;
    signed char *Tk = T + k * TAB_K;
    for (int q = 0; q < cs; q++) Tk[q] = (signed char)tab(As[q]);
    par[k] = s;
    val[k] = 0;
}

/* one pointer-jumping round: compose each unresolved table with its parent's */
__global__ __launch_bounds__(BLOCK) void k_tab_jump(int64_t nm, const int2 *__restrict__ info,
                                                    const int *__restrict__ mslot, const int *__restrict__ par_in,
                                                    const int *__restrict__ val_in, const signed char *__restrict__ T_in,
                                                    int *__restrict__ par_out, int *__restrict__ val_out,
                                                    signed char *__restrict__ T_out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const int p = par_in[k];
    if (p < 0) {
        par_out[k] = -1;
        val_out[k] = val_in[k];
        return;
    }
    const int64_t q = mslot[p];
    const signed char *Tk = T_in + k * TAB_K;
    const int pq = par_in[q];
    if (pq < 0) {
        const int vq = val_in[q];
        par_out[k] = -1;
        val_out[k] = vq < 0 ? vq : (int)Tk[vq];
        return;
    }
    const signed char *Tq = T_in + q * TAB_K;
    signed char *To = T_out + k * TAB_K;
    const int dom = sf_count(info[pq].y);
    for (int j = 0; j < dom; j++) {
        const int t = Tq[j];
        To[j] = t < 0 ? (signed char)t : Tk[t];
    }
    par_out[k] = pq;
    val_out[k] = 0;
}

/* final records (the replay's rules for LeftDomain and LeftDomainTime) */
__global__ __launch_bounds__(BLOCK) void k_tab_out(ChainGeom g, int64_t n, const int2 *__restrict__ info,
                                                   const int *__restrict__ mslot, const int *__restrict__ A,
                                                   const int *__restrict__ par, const int *__restrict__ val,
                                                   int *__restrict__ out_elem, int *__restrict__ out_left,
                                                   double *__restrict__ out_lt, ErrBuf eb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 my = info[i];
    int elem = -1, left = 1;
    double lt = -1.0;
    if (my.y & SF_OUTB) {
        if (g.chain == 1) lt = 0.0;
    } else if (my.x >= 0) {
        if (sf_count(my.y) <= 1) {
            elem = my.x;
            left = 0;
            lt = 0.0;
        } else {
            const int64_t k = mslot[i];
            const int v = par[k] >= 0 ? E_INDET : val[k];
            if (v < 0) {
                raise_err(eb, v, i);
            } else {
                elem = A[k * TAB_K + v];
                left = 0;
                lt = 0.0;
            }
        }
    }
    out_elem[i] = elem;
    out_left[i] = left;
    out_lt[i] = lt;
}

template <int DIM>
static int seed_chain_tab(fvc_ctx *c, const ChainGeom &g, int64_t n, const int2 *info, const int *mlist, int64_t nm,
                          int *mslot, int *src, int *chunk_last, int *carry, const TabPrep &tp, int seed0,
                          int *d_elem, int *d_left, double *d_left_time) {
    const int64_t nch = (n + 63) / 64;
    k_tab_src<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(g, n, info, tp.rowfirst, tp.planefirst, src, chunk_last);
    const int gs0 = (g.chain != 0 && seed0 >= 0) ? TAB_SEED : TAB_NONE;
    k_tab_carry<<<1, 1024, 0, c->stream>>>(nch, chunk_last, gs0, carry);
    int *pv = nullptr;
    int *A = nullptr;
    signed char *T = nullptr;
    int fin = 0;
    if (nm > 0) {
        if (hipMallocAsync((void **)&A, sizeof(int) * TAB_K * (size_t)nm, c->stream) != hipSuccess ||
            hipMallocAsync((void **)&T, 2 * TAB_K * (size_t)nm, c->stream) != hipSuccess ||
            hipMallocAsync((void **)&pv, sizeof(int) * 4 * (size_t)nm, c->stream) != hipSuccess) {
            if (A) hipFreeAsync(A, c->stream);
            if (T) hipFreeAsync(T, c->stream);
            return FVC_E_NOMEM;
        }
        int *par[2] = {pv, pv + nm}, *val[2] = {pv + 2 * nm, pv + 3 * nm};
        signed char *TT[2] = {T, T + TAB_K * nm};
        k_tab_slots<<<grid_for(nm, BLOCK), BLOCK, 0, c->stream>>>(nm, mlist, mslot);
        k_tab_accept<DIM><<<grid_for(nm, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, g, nm, mlist, A);
        k_tab_walk<DIM><<<grid_for(nm, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, g, n, nm, info, mlist, mslot, A, src,
                                                                      carry, seed0, tp.misc, par[0], val[0], TT[0]);
        /* a link chain of length L resolves in ceil(log2 L) + 1 rounds */
        int rounds = 1;
        while ((int64_t)1 << (rounds - 1) < nm) rounds++;
        for (int r = 0; r < rounds; r++) {
            const int a = r & 1;
            k_tab_jump<<<grid_for(nm, BLOCK), BLOCK, 0, c->stream>>>(nm, info, mslot, par[a], val[a], TT[a],
                                                                     par[a ^ 1], val[a ^ 1], TT[a ^ 1]);
        }
        fin = rounds & 1;
        k_tab_out<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(g, n, info, mslot, A, par[fin], val[fin], d_elem,
                                                               d_left, d_left_time, errbuf(c));
    } else {
        k_tab_out<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(g, n, info, mslot, A, nullptr, nullptr, d_elem, d_left,
                                                               d_left_time, errbuf(c));
    }
    FVC_HIP(hipGetLastError());
    if (A) FVC_HIP(hipFreeAsync(A, c->stream));
    if (T) FVC_HIP(hipFreeAsync(T, c->stream));
    if (pv) FVC_HIP(hipFreeAsync(pv, c->stream));
    return FVC_OK;
}

int launch_seed_chain(fvc_ctx *c, int chain, int64_t n, const double *dxyz, const double mn[3], const double del[3],
                      const int32_t res[3], int lsc, int *d_elem, int *d_left, double *d_left_time) {
    if (n <= 0) return FVC_OK;
    ChainGeom g;
    g.chain = chain;
    g.xyz = dxyz;
    for (int a = 0; a < 3; a++) {
        g.mn[a] = mn ? mn[a] : 0.0;
        g.del[a] = del ? del[a] : 0.0;
        g.res[a] = res ? res[a] : 1;
    }
    const bool uns = c->mesh.mesh_type == 1;
    int seed0 = -1;
    if (chain != 0 && uns) {
        /* centre seed: global search at the middle grid point (ftle.c:456-465, tracers.c:1318-1322) */
        double xs[3] = {g.mn[0] + ((int)(g.res[0] / 2)) * g.del[0], g.mn[1] + ((int)(g.res[1] / 2)) * g.del[1],
                        g.mn[2] + ((int)(g.res[2] / 2)) * g.del[2]};
        double *dxs = (double *)scratch(c, 64);
        if (!dxs) return FVC_E_NOMEM;
        FVC_HIP(hipMemcpyAsync(dxs, xs, sizeof xs, hipMemcpyHostToDevice, c->stream));
        int *dres = (int *)(dxs + 4);
        int rc = launch_locate(c, dxs, nullptr, 1, FVC_LOCATE_GLOBAL, dres);
        if (rc) return rc;
        FVC_HIP(hipMemcpyAsync(&seed0, dres, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipStreamSynchronize(c->stream));
    }
    /* point indices of the tabulated chain are ints; 0x7f7f7f7f marks "none" */
    const bool tab = uns && lsc && n < 0x7f000000LL && !c->chain_replay;
    const int64_t nch = (n + 63) / 64;
    const int64_t nrows = chain >= 2 ? n / chain_rowlen(g) : 1,
                  nplanes = chain >= 2 ? n / (chain_rowlen(g) * g.res[1]) : 1;
    size_t bytes = sizeof(int2) * (size_t)n;
    if (tab) bytes += sizeof(int) * (4 * (size_t)n + 2 * (size_t)nch + (size_t)nrows + (size_t)nplanes + 2);
    char *base = (char *)scratch(c, bytes);
    if (!base) return FVC_E_NOMEM;
    int2 *info = (int2 *)base;
    int *ip = (int *)(info + n);
    TabPrep tp{nullptr, nullptr, nullptr, nullptr};
    int *mlist = nullptr, *mslot = nullptr, *src = nullptr, *chunk_last = nullptr, *carry = nullptr;
    if (tab) {
        tp.mflag = ip;
        mlist = ip + n;
        mslot = ip + 2 * n;
        src = ip + 3 * n;
        chunk_last = ip + 4 * n;
        carry = chunk_last + nch;
        tp.rowfirst = carry + nch;
        tp.planefirst = tp.rowfirst + nrows;
        tp.misc = tp.planefirst + nplanes;
        FVC_HIP(hipMemsetAsync(tp.rowfirst, 0x7f, sizeof(int) * (size_t)(nrows + nplanes + 1), c->stream));
        FVC_HIP(hipMemsetAsync(tp.misc + 1, 0, sizeof(int), c->stream));
    }
    if (c->mesh.dims == 3)
        k_seed_prepare<3><<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, g, n, info, tp);
    else
        k_seed_prepare<2><<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, g, n, info, tp);
    FVC_HIP(hipGetLastError());
    bool replay = !tab;
    if (tab) {
        int misc[2] = {0, 0};
        FVC_HIP(hipMemcpyAsync(misc, tp.misc, sizeof misc, hipMemcpyDeviceToHost, c->stream));
        int64_t nm = 0;
        int rc = launch_compact(c, tp.mflag, n, mlist, &nm, nullptr); /* synchronizes the stream */
        if (rc) return rc;
        if (misc[1] > TAB_K) {
            replay = true; /* a point held by more elements than a table keeps */
        } else {
            rc = c->mesh.dims == 3 ? seed_chain_tab<3>(c, g, n, info, mlist, nm, mslot, src, chunk_last, carry, tp,
                                                       seed0, d_elem, d_left, d_left_time)
                                   : seed_chain_tab<2>(c, g, n, info, mlist, nm, mslot, src, chunk_last, carry, tp,
                                                       seed0, d_elem, d_left, d_left_time);
            if (rc) return rc;
        }
    }
    if (replay) {
        if (c->mesh.dims == 3)
            k_seed_chain<3><<<1, CHAIN_STAGE, 0, c->stream>>>(c->mesh, g, n, info, lsc, seed0, d_elem, d_left,
                                                              d_left_time, errbuf(c));
        else
            k_seed_chain<2><<<1, CHAIN_STAGE, 0, c->stream>>>(c->mesh, g, n, info, lsc, seed0, d_elem, d_left,
                                                              d_left_time, errbuf(c));
    }
    FVC_HIP(hipGetLastError());
    return check_err(c, "seed chain");
}

/* ------------------------------------------------------------------------ */
/* FTLE                                                                       */
/* ------------------------------------------------------------------------ */
/* ------------------------------------------------------------------------ */
/* stable compaction: indices of nonzero flags, in order                      */
/* ------------------------------------------------------------------------ */
constexpr int CMP_TILE = 4096; /* items per block */

__global__ __launch_bounds__(BLOCK) void k_cmp_count(const int *__restrict__ flags, int64_t n, int *__restrict__ counts) {
    int64_t base = (int64_t)blockIdx.x * CMP_TILE;
    int cnt = 0;
    for (int q = threadIdx.x; q < CMP_TILE; q += BLOCK) {
        int64_t i = base + q;
        if (i < n && flags[i]) cnt++;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    __shared__ int red[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

/* exclusive scan of a small int array in one block (in place); total at a[n] */
__global__ __launch_bounds__(1024) void k_scan_small(int *__restrict__ a, int n) {
    __shared__ int carry;
    __shared__ int tmp[1024];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < n ? a[i] : 0;
        tmp[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int add = threadIdx.x >= off ? tmp[threadIdx.x - off] : 0;
            __syncthreads();
            tmp[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < n) a[i] = carry + tmp[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += tmp[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) a[n] = carry;
}

__global__ __launch_bounds__(BLOCK) void k_cmp_scatter(const int *__restrict__ flags, int64_t n,
                                                      const int *__restrict__ offs, const int *__restrict__ map,
                                                      int *__restrict__ out) {
    __shared__ int wsum[BLOCK / 64];
    __shared__ int run;
    if (threadIdx.x == 0) run = offs[blockIdx.x];
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * CMP_TILE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int r = 0; r < CMP_TILE; r += BLOCK) {
        int64_t i = base + r + threadIdx.x;
        int fl = (i < n && flags[i]) ? 1 : 0;
        unsigned long long b = __ballot(fl);
        int before = __popcll(b & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(b);
        __syncthreads();
        int pre = 0;
        for (int q = 0; q < w; q++) pre += wsum[q];
        if (fl) out[run + pre + before] = map ? map[i] : (int)i;
        __syncthreads();
        if (threadIdx.x == 0) run += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

FtleDev ftle_dev(fvc_ftle *f) {
    FtleDev d;
    for (int a = 0; a < 3; a++) {
        d.res[a] = f->res[a];
        d.del[a] = f->del[a];
        d.Xn[a] = f->Xn[a];
        d.Vn[a] = f->Vn[a];
    }
    d.start_time = f->start_time;
    d.p = ptrs(f->pts);
    d.have = f->have;
    d.wt = f->wt;
    d.wot = f->wot;
    d.elem_new = f->elem_new;
    d.exited = f->exited;
    d.advected = f->advected;
    d.exit_time = f->exit_time;
    return d;
}

/* staged advection of the active points (ftle.c:109-169 without the early
 * part).  The particle velocity V only changes with extrapolation
 * (integration.c:419-425), so it moves only then. */
template <int MODE, int DIM, int UNS, int GEN>
__device__ __forceinline__ void ftle_advect_body(const DevMesh &m, const DevField &f, const FtleDev &d,
                                                 const int *__restrict__ active, int64_t nactive, double t1, double t2,
                                                 int itype, double dt, int extrap, ErrBuf eb,
                                                 unsigned long long *steps) {
    int64_t q = xcd_block() * blockDim.x + threadIdx.x;
    uint32_t ns = 0;
    if (q < nactive) {
        int64_t i = active[q];
        PState s;
        load_state(d.p, i, s, false, extrap);
        int err = 0;
        advect<MODE, DIM, UNS, GEN>(m, f, s, t1, t2, itype, dt, extrap, err, ns);
        if (err) raise_err(eb, err, i);
        d.advected[i] = 1;
        int ex = s.left && !extrap;
        d.exited[i] = ex;
        d.exit_time[i] = s.left_time;
        for (int c = 0; c < 3; c++) {
            d.Xn[c][i] = s.X[c];
            if (extrap) d.Vn[c][i] = s.V[c];
        }
        d.elem_new[i] = s.elem;
        if (extrap) d.p.left[i] = s.left;
    }
    add_steps(steps, ns);
}

template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(FVC_RK4_WAVES(MODE, UNS)))) void k_ftle_advect(
    DevMesh m, DevField f, FtleDev d, const int *__restrict__ active, int64_t nactive, double t1, double t2, int itype,
    double dt, int extrap, ErrBuf eb, unsigned long long *steps) {
    (void)extrap; /* RK4 without extrapolation; the launcher sends the rest to k_ftle_advect_gen */
    ftle_advect_body<MODE, DIM, UNS, 0>(m, f, d, active, nactive, t1, t2, itype, dt, 0, eb, steps);
}

template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_ftle_advect_gen(DevMesh m, DevField f, FtleDev d,
                                                          const int *__restrict__ active, int64_t nactive, double t1,
                                                          double t2, int itype, double dt, int extrap, ErrBuf eb,
                                                          unsigned long long *steps) {
    ftle_advect_body<MODE, DIM, UNS, 1>(m, f, d, active, nactive, t1, t2, itype, dt, extrap, eb, steps);
}

/* Advect_FTLEPoint (ftle.c:1205-1227) from the frame-start state; the element
 * is the one the reference's array holds when the trigger is processed. */
template <int MODE, int DIM, int UNS, int GEN = 1>
__device__ void advect_copy(const DevMesh &m, const DevField &f, const FtleDev &d, int64_t idx, int64_t trig, double t1,
                            double t2, int itype, double dt, PState &s, int &err) {
    for (int c = 0; c < 3; c++) {
        s.X[c] = d.p.X[c][idx];
        s.V[c] = 0.0;
    }
    s.left = d.p.left[idx];
    bool updated = idx < trig && d.advected[idx] && !d.exited[idx];
    s.elem = updated ? d.elem_new[idx] : d.p.elem[idx];
    s.left_time = 0.0;
    uint32_t ns = 0;
    advect<MODE, DIM, UNS, GEN>(m, f, s, t1, t2, itype, dt, 0, err, ns);
}

/* Frame-start state of the stencils of one frame's triggered points, taken
 * before the frame's position update so that their early FTLE can run on a
 * second stream beside the next frame's advection.  Slot q = 0-5 is
 * neighbour nb[q] (grid_nbrs), q = 6 the point itself; elem is the element
 * advect_copy would start from (the loop-order rule already applied). */
struct EarlySnap {
    int *nt;          /* triggered points of the frame (also when above cap) */
    int *P;
    double *te;
    double *X;        /* [(q * 3 + c) * cap + t] */
    int *left, *elem; /* [q * cap + t] */
    int64_t cap;
};

static size_t early_snap_bytes(int64_t cap) { return 256 + (size_t)cap * (4 + 8 + 7 * (24 + 4 + 4)); }

static EarlySnap early_snap_view(void *buf, int64_t cap) {
    EarlySnap s;
    char *b = (char *)buf;
    s.nt = (int *)b;
    b += 256;
    s.X = (double *)b;
    b += sizeof(double) * 21 * (size_t)cap;
    s.te = (double *)b;
    b += sizeof(double) * (size_t)cap;
    s.P = (int *)b;
    b += sizeof(int) * (size_t)cap;
    s.left = (int *)b;
    b += sizeof(int) * 7 * (size_t)cap;
    s.elem = (int *)b;
    s.cap = cap;
    return s;
}

template <int MODE, int DIM, int UNS, int GEN>
__device__ void advect_snap(const DevMesh &m, const DevField &f, const EarlySnap &sn, int q, int64_t t, double t1,
                            double t2, int itype, double dt, PState &s, int &err) {
    for (int c = 0; c < 3; c++) {
        s.X[c] = sn.X[(q * 3 + c) * sn.cap + t];
        s.V[c] = 0.0;
    }
    s.left = sn.left[q * sn.cap + t];
    s.elem = sn.elem[q * sn.cap + t];
    s.left_time = 0.0;
    uint32_t ns = 0;
    advect<MODE, DIM, UNS, GEN>(m, f, s, t1, t2, itype, dt, 0, err, ns);
}

__device__ inline void ftle_value(double D[3][3], double dt_int, double &wt, double &wot, int &err, bool zero_time) {
#pragma clang fp contract(off)
    double C[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            C[a][b] = 0;
            for (int c = 0; c < 3; c++) C[a][b] += D[c][a] * D[c][b];
        }
    double lam = max_eig_sym3(C, err);
    if (lam > 1.0 - TINY) {
        if (zero_time) {
            wt = 0.0;
            wot = 0.0;
        } else {
            wt = (0.5 * log(lam)) / dt_int;
            wot = 0.5 * log(lam);
        }
    } else {
        wt = 0.0;
        wot = 0.0;
    }
}

/* gradient entries from lo/hi neighbour positions per axis (ftle.c:994-1020) */
__device__ inline void gradient(const double lo[3][3], const double hi[3][3], const double fac[3], const double del[3],
                                int dims, double D[3][3]) {
#pragma clang fp contract(off)
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) D[a][b] = 0.0;
    if (dims == 3) {
        D[0][2] = (hi[2][0] - lo[2][0]) / (fac[2] * del[2]);
        D[1][2] = (hi[2][1] - lo[2][1]) / (fac[2] * del[2]);
        D[2][0] = (hi[0][2] - lo[0][2]) / (fac[0] * del[0]);
        D[2][1] = (hi[1][2] - lo[1][2]) / (fac[1] * del[1]);
        D[2][2] = (hi[2][2] - lo[2][2]) / (fac[2] * del[2]);
    }
    D[0][0] = (hi[0][0] - lo[0][0]) / (fac[0] * del[0]);
    D[0][1] = (hi[1][0] - lo[1][0]) / (fac[1] * del[1]);
    D[1][0] = (hi[0][1] - lo[0][1]) / (fac[0] * del[0]);
    D[1][1] = (hi[1][1] - lo[1][1]) / (fac[1] * del[1]);
}

/* GetFTLEForPointEarly (ftle.c:788-1048).  The trigger of point P is the
 * first exiting point, in the reference's loop order, among P and its six
 * neighbours; te is the trigger's exit time.  The re-advections of a
 * triggered point are independent, so they run as separate lanes: pass A
 * advects the (up to) six neighbours to te, pass B the point itself and the
 * gradient neighbours to MinExitTime, pass C forms the gradient and the FTLE.
 * Each lane does exactly the reference's computation for its advection. */
__device__ inline void grid_nbrs(const FtleDev &d, int64_t P, int cc[3], int64_t nb[6], bool has[6]) {
    const int64_t sj = d.res[2], si = (int64_t)d.res[2] * d.res[1];
    cc[0] = (int)(P / si);
    cc[1] = (int)((P / d.res[2]) % d.res[1]);
    cc[2] = (int)(P % d.res[2]);
    nb[0] = P - si;
    nb[1] = P + si;
    nb[2] = P - sj;
    nb[3] = P + sj;
    nb[4] = P - 1;
    nb[5] = P + 1;
    for (int ax = 0; ax < 3; ax++) {
        has[2 * ax] = cc[ax] > 0;
        has[2 * ax + 1] = cc[ax] < d.res[ax] - 1;
    }
}

__device__ inline int64_t early_trigger(const FtleDev &d, int64_t P, const int64_t nb[6], const bool has[6]) {
    int64_t trig = d.exited[P] ? P : INT64_MAX;
    for (int q = 0; q < 6; q++)
        if (has[q] && d.exited[nb[q]] && nb[q] < trig) trig = nb[q];
    return trig;
}

__global__ __launch_bounds__(BLOCK) void k_early_flag(FtleDev d, int64_t n, int *__restrict__ flags) {
    const int64_t P = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= n) return;
    int fl = 0;
    if (!d.have[P]) {
        int cc[3];
        int64_t nb[6];
        bool has[6];
        grid_nbrs(d, P, cc, nb, has);
        fl = early_trigger(d, P, nb, has) != INT64_MAX;
    }
    flags[P] = fl;
}

/* the stencil snapshot of the first cap triggered points (EarlySnap) */
__global__ __launch_bounds__(BLOCK) void k_early_snap(FtleDev d, const int *__restrict__ list,
                                                     const int *__restrict__ d_nt, EarlySnap s) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = *d_nt;
    if (t == 0) *s.nt = nt;
    if (t >= nt || t >= s.cap) return;
    const int64_t P = list[t];
    int cc[3];
    int64_t nb[6];
    bool has[6];
    grid_nbrs(d, P, cc, nb, has);
    const int64_t trig = early_trigger(d, P, nb, has);
    s.P[t] = (int)P;
    s.te[t] = d.exit_time[trig];
    for (int q = 0; q < 7; q++) {
        const int64_t idx = q < 6 ? nb[q] : P;
        const bool ok = q == 6 || has[q];
        for (int c = 0; c < 3; c++) s.X[(q * 3 + c) * s.cap + t] = ok ? d.p.X[c][idx] : 0.0;
        int left = 0, elem = -1;
        if (ok) {
            left = d.p.left[idx];
            const bool updated = idx < trig && d.advected[idx] && !d.exited[idx];
            elem = updated ? d.elem_new[idx] : d.p.elem[idx];
        }
        s.left[q * s.cap + t] = left;
        s.elem[q * s.cap + t] = elem;
    }
}

/* One triggered point per group of eight lanes (grid-stride over the list,
 * whose length is read from device memory):
 *   lanes 0-5 advect the six neighbours to te -> their exit times; every lane
 *   of the group then forms MinExitTime from the same six values;
 *   lane 0 advects the point itself and lanes 1-6 the lower (odd) or upper
 *   (even) neighbour along each axis that the difference formula uses, to
 *   MinExitTime; lane 0 gathers the positions, forms the gradient (one-sided
 *   at the grid edges, ftle.c:994-1020) and the FTLE. */
template <int MODE, int DIM, int UNS, int GEN, int SNAP>
__device__ __forceinline__ void early_body(const DevMesh &m, const DevField &f, const FtleDev &d,
                                           const int *__restrict__ list, const int *__restrict__ d_nt,
                                           const EarlySnap &sn, int64_t above, double t1, int itype, double dt,
                                           ErrBuf eb) {
#pragma clang fp contract(off)
    /* live input (list, d_nt, the slide's state): runs when the frame has
     * more than `above` triggered points; snapshot input: when it has at
     * most sn.cap (the two launches of a frame split the cases) */
    int nt;
    if constexpr (SNAP) {
        nt = *sn.nt;
        if (nt > sn.cap) return;
    } else {
        nt = *d_nt;
        if (nt <= above) return;
    }
    const int sub = threadIdx.x & 7;
    const int64_t ngroups = ((int64_t)gridDim.x * blockDim.x) >> 3;
    for (int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3; t < nt; t += ngroups) {
        const int64_t P = SNAP ? (int64_t)sn.P[t] : (int64_t)list[t];
        int cc[3];
        int64_t nb[6];
        bool has[6];
        grid_nbrs(d, P, cc, nb, has);
        int64_t trig = 0;
        double te;
        if constexpr (SNAP) {
            te = sn.te[t];
        } else {
            trig = early_trigger(d, P, nb, has);
            te = d.exit_time[trig];
        }
        int err = 0;
        double mt = te;
        double X[3] = {0.0, 0.0, 0.0};
        /* one advection call site for both phases keeps a single inlined copy */
        for (int phase = 0; phase < 2; phase++) {
            bool need;
            int64_t who;
            int q; /* snapshot slot of `who` */
            if (phase == 0) {
                need = sub < 6 && has[sub];
                who = sub < 6 ? nb[sub] : P;
                q = sub < 6 ? sub : 6;
            } else {
                need = sub == 0;
                who = P;
                q = 6;
                if (sub >= 1 && sub <= 6) {
                    const int ax = (sub - 1) >> 1, side = (sub - 1) & 1;
                    if (ax < DIM) {
                        const int last = d.res[ax] - 1;
                        need = !(side == 0 && cc[ax] == 0) && !(side == 1 && cc[ax] != 0 && cc[ax] == last);
                        who = nb[2 * ax + side];
                        q = 2 * ax + side;
                    }
                }
            }
            PState s;
            s.left = 0;
            s.left_time = 0.0;
            if (need) {
                if constexpr (SNAP)
                    advect_snap<MODE, DIM, UNS, GEN>(m, f, sn, q, t, t1, phase == 0 ? te : mt, itype, dt, s, err);
                else
                    advect_copy<MODE, DIM, UNS, GEN>(m, f, d, who, trig, t1, phase == 0 ? te : mt, itype, dt, s, err);
                for (int c = 0; c < 3; c++) X[c] = s.X[c];
            }
            if (phase == 0) {
                const double lt = (need && s.left) ? s.left_time : te;
                double L[6];
                for (int q = 0; q < 6; q++) L[q] = __shfl(lt, q, 8);
                mt = dmin(te, dmin(L[5], dmin(L[4], dmin(L[3], dmin(L[2], dmin(L[1], L[0]))))));
            }
        }
        if (err) raise_err(eb, err, P);
        double pos[7][3];
        for (int r = 0; r < 7; r++)
            for (int c = 0; c < 3; c++) pos[r][c] = __shfl(X[c], r, 8);
        if (sub == 0) {
            double lo[3][3], hi[3][3], fac[3] = {1.0, 1.0, 1.0};
            for (int ax = 0; ax < DIM; ax++) {
                const int last = d.res[ax] - 1;
                for (int c = 0; c < 3; c++) {
                    if (cc[ax] == 0) {
                        lo[ax][c] = pos[0][c];
                        hi[ax][c] = pos[2 + 2 * ax][c];
                    } else if (cc[ax] == last) {
                        lo[ax][c] = pos[1 + 2 * ax][c];
                        hi[ax][c] = pos[0][c];
                    } else {
                        lo[ax][c] = pos[1 + 2 * ax][c];
                        hi[ax][c] = pos[2 + 2 * ax][c];
                    }
                }
                if (cc[ax] != 0 && cc[ax] != last) fac[ax] = 2.0;
            }
            double D[3][3];
            gradient(lo, hi, fac, d.del, DIM, D);
            int e2 = 0;
            double wt, wot;
            ftle_value(D, fabs(mt - d.start_time), wt, wot, e2, fabs(mt - d.start_time) < TINY);
            if (e2) raise_err(eb, e2, P);
            d.wt[P] = wt;
            d.wot[P] = wot;
            d.have[P] = 1;
        }
    }
}

/* RK4 frames, live input: the frame's points when it has more than `above` */
template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_early(DevMesh m, DevField f, FtleDev d, const int *__restrict__ list,
                                                const int *__restrict__ d_nt, int64_t above, double t1, int itype,
                                                double dt, ErrBuf eb) {
    early_body<MODE, DIM, UNS, 0, 0>(m, f, d, list, d_nt, EarlySnap{}, above, t1, itype, dt, eb);
}

/* RK4 frames, from the stencil snapshot (at most cap triggered points) */
template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_early_snapped(DevMesh m, DevField f, FtleDev d, EarlySnap sn, double t1,
                                                        int itype, double dt, ErrBuf eb) {
    early_body<MODE, DIM, UNS, 0, 1>(m, f, d, nullptr, nullptr, sn, 0, t1, itype, dt, eb);
}

/* Euler frames */
template <int MODE, int DIM, int UNS>
__global__ __launch_bounds__(BLOCK) void k_early_gen(DevMesh m, DevField f, FtleDev d, const int *__restrict__ list,
                                                    const int *__restrict__ d_nt, double t1, int itype, double dt,
                                                    ErrBuf eb) {
    early_body<MODE, DIM, UNS, 1, 0>(m, f, d, list, d_nt, EarlySnap{}, -1, t1, itype, dt, eb);
}

/* GetFTLEForPoint (ftle.c:1050-1203) */
__global__ __launch_bounds__(BLOCK) void k_ftle_fill(FtleDev d, int64_t n, int dims, double int_t, ErrBuf eb) {
#pragma clang fp contract(off)
    int64_t P = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= n || d.have[P]) return;
    if (int_t < TINY) {
        d.wt[P] = 0.0;
        d.wot[P] = 0.0;
        return;
    }
    const int64_t st[3] = {(int64_t)d.res[2] * d.res[1], d.res[2], 1};
    int cc[3] = {(int)(P / st[0]), (int)((P / d.res[2]) % d.res[1]), (int)(P % d.res[2])};
    double lo[3][3], hi[3][3], fac[3] = {1.0, 1.0, 1.0};
    for (int ax = 0; ax < dims; ax++) {
        int64_t l = P, h = P;
        if (cc[ax] == 0)
            h = P + st[ax];
        else if (cc[ax] == d.res[ax] - 1)
            l = P - st[ax];
        else {
            l = P - st[ax];
            h = P + st[ax];
            fac[ax] = 2.0;
        }
        for (int c = 0; c < 3; c++) {
            lo[ax][c] = d.p.X[c][l];
            hi[ax][c] = d.p.X[c][h];
        }
    }
    double D[3][3];
    gradient(lo, hi, fac, d.del, dims, D);
    int err = 0;
    double wt, wot;
    ftle_value(D, int_t, wt, wot, err, false);
    if (err) raise_err(eb, err, P);
    d.wt[P] = wt;
    d.wot[P] = wot;
}

/* outside masks after location (ftle.c:425-452, 484-501, 521-572) */
__global__ __launch_bounds__(BLOCK) void k_ftle_mask(FtleDev d, int64_t n) {
    int64_t P = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= n) return;
    const int64_t si = (int64_t)d.res[2] * d.res[1], sj = d.res[2];
    int k = (int)(P % d.res[2]), j = (int)((P / d.res[2]) % d.res[1]), i = (int)(P / si);
    int own = d.p.left[P];
    int have = own;
    if (i > 0 && d.p.left[P - si]) have = 1;
    if (i < d.res[0] - 1 && d.p.left[P + si]) have = 1;
    if (j > 0 && d.p.left[P - sj]) have = 1;
    if (j < d.res[1] - 1 && d.p.left[P + sj]) have = 1;
    if (k > 0 && d.p.left[P - 1]) have = 1;
    if (k < d.res[2] - 1 && d.p.left[P + 1]) have = 1;
    d.have[P] = have;
    d.wt[P] = own ? -1.0 : 0.0;
    d.wot[P] = own ? -1.0 : 0.0;
    d.exited[P] = 0;
    d.advected[P] = 0;
}

/* active list: points integrated this frame (ftle.c:113), flagged in the
 * grid's Z-order so that the compacted list is Z-ordered too */
__global__ __launch_bounds__(BLOCK) void k_ftle_flags(FtleDev d, int64_t n, int extrap, const int *__restrict__ zperm,
                                                     int *__restrict__ flags) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t P = zperm[q];
    flags[q] = d.p.left_time[P] > -TINY && (!d.p.left[P] || extrap);
}

int launch_ftle_mask_dev(fvc_ctx *c, const FtleDev &d, int64_t n) {
    if (n <= 0) return FVC_OK;
    k_ftle_mask<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(d, n);
    FVC_HIP(hipGetLastError());
    return FVC_OK;
}

int launch_ftle_mask(fvc_ftle *f) { return launch_ftle_mask_dev(f->ctx, ftle_dev(f), f->n); }

/* early FTLE of one frame for a grid of n points (after its advection, before
 * the position update) */
int launch_ftle_early(fvc_ctx *c, const FtleDev &d, int64_t n, const DevField &fld, double t1, int itype, double dt) {
    /* scratch: flags [n], triggered list [n], per-tile offsets and the count
     * [nb + 1]; the count stays on the device, so nothing here waits */
    const int nb = (int)((n + CMP_TILE - 1) / CMP_TILE);
    int *eflags = (int *)scratch(c, sizeof(int) * (2 * (size_t)n + nb + 1));
    if (!eflags) return FVC_E_NOMEM;
    int *elist = eflags + n, *offs = elist + n;
    k_early_flag<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(d, n, eflags);
    k_cmp_count<<<nb, BLOCK, 0, c->stream>>>(eflags, n, offs);
    k_scan_small<<<1, 1024, 0, c->stream>>>(offs, nb);
    k_cmp_scatter<<<nb, BLOCK, 0, c->stream>>>(eflags, n, offs, nullptr, elist);
    const unsigned groups = (unsigned)std::min<int64_t>(std::max<int64_t>(n / 8, 1), 16384);
    const unsigned grid = (unsigned)(((int64_t)groups * 8 + BLOCK - 1) / BLOCK);
    if (itype == INT_RK4)
        FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_early,
                     <<<grid, BLOCK, 0, c->stream>>>(c->mesh, fld, d, elist, offs + nb, (int64_t)-1, t1, itype, dt,
                                                     errbuf(c)));
    else
        FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_early_gen,
                     <<<grid, BLOCK, 0, c->stream>>>(c->mesh, fld, d, elist, offs + nb, t1, itype, dt, errbuf(c)));
    FVC_HIP(hipGetLastError());
    return FVC_OK;
}

/* Early FTLE of an RK4 slide frame, overlapped with what follows it.  On the
 * context's stream: the triggered list, then the frame-start state of the
 * first EARLY_CAP triggered points' stencils into one of the slide's two
 * snapshot buffers (they alternate by frame), then the live kernel, which
 * runs only when the frame triggered more points than that.  On the side
 * stream, after the snapshot: the early FTLE from the snapshot, which reads
 * nothing the position update or the next frame's advection writes, and
 * writes only have / wt / wot of the triggered points.  Everything that
 * reads those or changes the slices joins the side stream first (side_join);
 * so does the next frame, before its triggered list, which also keeps a
 * snapshot buffer from being rewritten while its early FTLE still reads it. */
int launch_ftle_early_async(fvc_ftle *fo, const FtleDev &d, const DevField &fld, double t1, int itype, double dt) {
    fvc_ctx *c = fo->ctx;
    const int64_t n = fo->n;
    constexpr int64_t EARLY_CAP = 1 << 18;
    if (itype != INT_RK4) {
        int rc = side_join(c);
        return rc ? rc : launch_ftle_early(c, d, n, fld, t1, itype, dt);
    }
    if (!fo->snap[0]) {
        const int64_t cap = std::min<int64_t>(n, EARLY_CAP);
        const size_t bytes = early_snap_bytes(cap);
        if (hipMalloc(&fo->snap[0], bytes) != hipSuccess) fo->snap[0] = nullptr;
        if (fo->snap[0] && hipMalloc(&fo->snap[1], bytes) != hipSuccess) {
            hipFree(fo->snap[0]);
            fo->snap[0] = fo->snap[1] = nullptr;
        }
        if (!fo->snap[0]) {
            (void)hipGetLastError();
            c->last_error = "early FTLE snapshot alloc";
            return FVC_E_NOMEM;
        }
        fo->snap_cap = cap;
        c->bytes_used += 2 * (int64_t)bytes;
    }
    int rc = side_join(c);
    if (rc) return rc;
    const int nb = (int)((n + CMP_TILE - 1) / CMP_TILE);
    int *eflags = (int *)scratch(c, sizeof(int) * (2 * (size_t)n + nb + 1));
    if (!eflags) return FVC_E_NOMEM;
    int *elist = eflags + n, *offs = elist + n;
    k_early_flag<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(d, n, eflags);
    k_cmp_count<<<nb, BLOCK, 0, c->stream>>>(eflags, n, offs);
    k_scan_small<<<1, 1024, 0, c->stream>>>(offs, nb);
    k_cmp_scatter<<<nb, BLOCK, 0, c->stream>>>(eflags, n, offs, nullptr, elist);
    const EarlySnap sn = early_snap_view(fo->snap[fo->snap_sel], fo->snap_cap);
    k_early_snap<<<grid_for(fo->snap_cap, BLOCK), BLOCK, 0, c->stream>>>(d, elist, offs + nb, sn);
    const unsigned groups = (unsigned)std::min<int64_t>(std::max<int64_t>(n / 8, 1), 16384);
    const unsigned grid = (unsigned)(((int64_t)groups * 8 + BLOCK - 1) / BLOCK);
    FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_early,
                 <<<grid, BLOCK, 0, c->stream>>>(c->mesh, fld, d, elist, offs + nb, fo->snap_cap, t1, itype, dt,
                                                 errbuf(c)));
    FVC_HIP(hipGetLastError());
    FVC_HIP(hipEventRecord(c->side_ready, c->stream));
    FVC_HIP(hipStreamWaitEvent(c->side, c->side_ready, 0));
    FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_early_snapped,
                 <<<grid, BLOCK, 0, c->side>>>(c->mesh, fld, d, sn, t1, itype, dt, errbuf(c)));
    FVC_HIP(hipGetLastError());
    FVC_HIP(hipEventRecord(c->side_done, c->side));
    c->side_pending = 1;
    fo->snap_sel ^= 1;
    return FVC_OK;
}

/* position update (ftle.c:160-169, UpdateFTLELocations ftle.c:1229-1242)
 * of the points advected this frame, over the whole grid in grid order:
 * coalesced, where the active list's Z-order scatters the copies */
__global__ __launch_bounds__(BLOCK) void k_ftle_update_all(FtleDev d, int64_t n, int uns, int extrap) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !d.advected[i]) return;
    if (d.exited[i]) {
        d.p.left[i] = 1;
    } else {
        for (int c = 0; c < 3; c++) {
            d.p.X[c][i] = d.Xn[c][i];
            if (extrap) d.p.V[c][i] = d.Vn[c][i];
        }
        if (uns) d.p.elem[i] = d.elem_new[i];
    }
    d.exited[i] = 0;
    d.advected[i] = 0;
}

int launch_ftle_update_all(fvc_ctx *c, const FtleDev &d, int64_t n, int extrap) {
    if (n <= 0) return FVC_OK;
    k_ftle_update_all<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(d, n, c->mesh.mesh_type, extrap);
    FVC_HIP(hipGetLastError());
    return FVC_OK;
}

int launch_ftle_frame(fvc_ftle *f, int lo, int hi, double tmin, double tmax, double t1, double t2,
                      const fvc_int_params &ip) {
    fvc_ctx *c = f->ctx;
    DevField fld = bind_field(c, lo, hi, tmin, tmax);
    fld.rkf_min = ip.rkf_min_step;
    fld.rkf_max = ip.rkf_max_step;
    fld.rkf_tol = ip.rkf_accuracy;
    FtleDev d = ftle_dev(f);
    const int itype = ip.int_type == FVC_INT_RK4 ? INT_RK4 : ip.int_type == FVC_INT_RKF ? INT_RKF : INT_EULER;
    int *flags = (int *)scratch(c, sizeof(int) * (size_t)f->n);
    if (!flags) return FVC_E_NOMEM;
    k_ftle_flags<<<grid_for(f->n, BLOCK), BLOCK, 0, c->stream>>>(d, f->n, ip.extrapolate, f->zperm, flags);
    FVC_HIP(hipGetLastError());
    /* the counter is cleared before the compaction's count read-back, so the
     * advection launches right after that synchronisation */
    FVC_HIP(hipMemsetAsync(c->d_steps, 0, sizeof(unsigned long long), c->stream));
    int rc = launch_compact(c, flags, f->n, f->active, &f->nactive, f->zperm);
    if (rc) return rc;
    if (f->nactive > 0) {
        timing_begin(c);
        if (itype == 1 && !ip.extrapolate)
            FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_ftle_advect,
                         <<<grid_for(f->nactive, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, fld, d, f->active,
                                                                                f->nactive, t1, t2, itype,
                                                                                ip.time_step, ip.extrapolate,
                                                                                errbuf(c), c->d_steps));
        else
            FVC_DISPATCH(c->precision, c->mesh.dims, c->mesh.mesh_type, k_ftle_advect_gen,
                         <<<grid_for(f->nactive, BLOCK), BLOCK, 0, c->stream>>>(c->mesh, fld, d, f->active,
                                                                                f->nactive, t1, t2, itype,
                                                                                ip.time_step, ip.extrapolate,
                                                                                errbuf(c), c->d_steps));
        timing_end(c);
        FVC_HIP(hipGetLastError());
        if (!ip.extrapolate && (rc = launch_ftle_early_async(f, d, fld, t1, itype, ip.time_step))) return rc;
        if ((rc = launch_ftle_update_all(c, d, f->n, ip.extrapolate))) return rc;
    }
    rc = check_err(c, "fvc_ftle_frame");
    c->last_steps = (int64_t)c->h_status->steps[0];
    if (f->nactive > 0) timing_collect(c, c->last_steps);
    return rc;
}

/* plain streaming copy, 16 B per lane (HBM byte-counter calibration) */
__global__ __launch_bounds__(BLOCK) void k_copy_probe(uint4 *__restrict__ dst, const uint4 *__restrict__ src,
                                                     int64_t n16) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

int launch_copy_probe(fvc_ctx *c, void *dst, const void *src, int64_t bytes) {
    int64_t n16 = bytes / 16;
    unsigned grid = (unsigned)std::min<int64_t>((n16 + BLOCK - 1) / BLOCK, 65536);
    if (grid == 0) return FVC_OK;
    k_copy_probe<<<grid, BLOCK, 0, c->stream>>>((uint4 *)dst, (const uint4 *)src, n16);
    FVC_HIP(hipGetLastError());
    FVC_HIP(hipStreamSynchronize(c->stream));
    return FVC_OK;
}

int launch_ftle_fill_dev(fvc_ctx *c, const FtleDev &d, int64_t n, double int_t) {
    if (n <= 0) return FVC_OK;
    k_ftle_fill<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(d, n, c->mesh.dims, int_t, errbuf(c));
    FVC_HIP(hipGetLastError());
    return FVC_OK;
}

int launch_ftle_fill(fvc_ftle *f, double int_t) {
    fvc_ctx *c = f->ctx;
    int rc = launch_ftle_fill_dev(c, ftle_dev(f), f->n, int_t);
    return rc ? rc : check_err(c, "fvc_ftle_output");
}

int launch_compact(fvc_ctx *c, const int *flags, int64_t n, int *out_idx, int64_t *count, const int *map) {
    int nb = (int)((n + CMP_TILE - 1) / CMP_TILE);
    if (nb == 0) {
        *count = 0;
        return FVC_OK;
    }
    int *offs = nullptr;
    FVC_HIP(hipMallocAsync((void **)&offs, sizeof(int) * (nb + 1), c->stream));
    k_cmp_count<<<nb, BLOCK, 0, c->stream>>>(flags, n, offs);
    k_scan_small<<<1, 1024, 0, c->stream>>>(offs, nb);
    k_cmp_scatter<<<nb, BLOCK, 0, c->stream>>>(flags, n, offs, map, out_idx);
    /* the count lands in pinned host memory (the status block's slot) */
    FVC_HIP(hipMemcpyAsync(&c->h_status->count, offs + nb, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FVC_HIP(hipFreeAsync(offs, c->stream));
    FVC_HIP(hipStreamSynchronize(c->stream));
    FVC_HIP(hipGetLastError());
    *count = c->h_status->count;
    return FVC_OK;
}

/* ------------------------------------------------------------------------ */
/* Z-order keys + stable LSD radix sort                                       */
/* ------------------------------------------------------------------------ */
__device__ inline uint64_t spread3(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ __launch_bounds__(BLOCK) void k_morton(const double *__restrict__ xyz, int64_t n, double lo0, double lo1,
                                                 double lo2, double s0, double s1, double s2, uint64_t *__restrict__ keys,
                                                 int *__restrict__ vals) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double lo[3] = {lo0, lo1, lo2}, sc[3] = {s0, s1, s2};
    uint64_t q[3];
    for (int a = 0; a < 3; a++) {
        double v = (xyz[3 * i + a] - lo[a]) * sc[a];
        v = v < 0.0 ? 0.0 : (v > 2097151.0 ? 2097151.0 : v);
        q[a] = (uint64_t)v;
    }
    /* x bit most significant within each level, then y, then z */
    keys[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    vals[i] = (int)i;
}

/* grid point P = (i*r1 + j)*r2 + k (k fastest) keyed by its integer (i,j,k) */
__global__ __launch_bounds__(BLOCK) void k_grid_morton(int r1, int r2, int64_t n, uint64_t *__restrict__ keys,
                                                      int *__restrict__ vals) {
    int64_t P = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= n) return;
    const uint64_t k = (uint64_t)(P % r2), j = (uint64_t)((P / r2) % r1), i = (uint64_t)(P / ((int64_t)r2 * r1));
    keys[P] = (spread3(i) << 2) | (spread3(j) << 1) | spread3(k);
    vals[P] = (int)P;
}

constexpr int RS_TILE = 4096;

__global__ __launch_bounds__(BLOCK) void k_rs_hist(const uint64_t *__restrict__ keys, int64_t n, int shift, int nblocks,
                                                  int *__restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int q = threadIdx.x; q < RS_TILE; q += BLOCK) {
        int64_t i = base + q;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(BLOCK) void k_rs_scatter(const uint64_t *__restrict__ kin, const int *__restrict__ vin,
                                                     int64_t n, int shift, int nblocks, const int *__restrict__ offs,
                                                     uint64_t *__restrict__ kout, int *__restrict__ vout) {
    __shared__ int run[256];
    __shared__ int wcnt[BLOCK / 64][256];
    run[threadIdx.x] = offs[threadIdx.x * nblocks + blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_TILE; r += BLOCK) {
        for (int q = 0; q < BLOCK / 64; q++) wcnt[q][threadIdx.x] = 0;
        __syncthreads();
        int64_t i = base + r + threadIdx.x;
        bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0;
        int val = valid ? vin[i] : 0;
        int dig = valid ? (int)((key >> shift) & 255) : 256;
        /* lanes of this wave with the same digit */
        unsigned long long same = __ballot(valid);
        for (int b = 0; b < 9; b++) {
            unsigned long long bb = __ballot((dig >> b) & 1);
            same &= ((dig >> b) & 1) ? bb : ~bb;
        }
        int before = __popcll(same & ((1ull << lane) - 1ull));
        if (valid && before == 0) wcnt[w][dig] = __popcll(same);
        __syncthreads();
        if (valid) {
            int pos = run[dig] + before;
            for (int q = 0; q < w; q++) pos += wcnt[q][dig];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int tot = 0;
        for (int q = 0; q < BLOCK / 64; q++) tot += wcnt[q][threadIdx.x];
        run[threadIdx.x] += tot;
        __syncthreads();
    }
}

/* exclusive scan over the digit-major histogram (hist[d*nb + b]) */
__global__ __launch_bounds__(1024) void k_rs_scan(int *__restrict__ a, int total) { /* single block */
    __shared__ int carry;
    __shared__ int tmp[1024];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < total; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < total ? a[i] : 0;
        tmp[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int add = threadIdx.x >= off ? tmp[threadIdx.x - off] : 0;
            __syncthreads();
            tmp[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < total) a[i] = carry + tmp[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += tmp[1023];
        __syncthreads();
    }
}

/* stable LSD radix sort of (key, value) pairs on the low `bits` bits of the
 * keys, ping-ponging between (k0, v0) and (k1, v1); *which = 0 when the sorted
 * pairs end in (k0, v0), 1 when in (k1, v1) */
int radix_sort_pairs(hipStream_t s, uint64_t *k0, int *v0, uint64_t *k1, int *v1, int64_t n, int bits, int *which) {
    *which = 0;
    if (n <= 0) return FVC_OK;
    const int nb = (int)((n + RS_TILE - 1) / RS_TILE);
    int *hist = nullptr;
    if (hipMallocAsync((void **)&hist, sizeof(int) * 256 * (size_t)nb, s) != hipSuccess) return FVC_E_NOMEM;
    for (int shift = 0; shift < bits && shift < 64; shift += 8) {
        k_rs_hist<<<nb, BLOCK, 0, s>>>(k0, n, shift, nb, hist);
        k_rs_scan<<<1, 1024, 0, s>>>(hist, 256 * nb);
        k_rs_scatter<<<nb, BLOCK, 0, s>>>(k0, v0, n, shift, nb, hist, k1, v1);
        std::swap(k0, k1);
        std::swap(v0, v1);
        *which ^= 1;
    }
    (void)hipFreeAsync(hist, s);
    return hipGetLastError() == hipSuccess ? FVC_OK : FVC_E_HIP;
}

/* exclusive prefix sum of a[0..n) in place (one block) */
int scan_exclusive_small(hipStream_t s, int *a, int n) {
    if (n > 0) k_rs_scan<<<1, 1024, 0, s>>>(a, n);
    return hipGetLastError() == hipSuccess ? FVC_OK : FVC_E_HIP;
}

/* stable sort of (key, value) pairs by key; dperm receives the values */
static int sort_pairs(fvc_ctx *c, uint64_t *k0, int *v0, int64_t n, int *dperm, uint64_t *k1, int *v1) {
    int which = 0;
    int rc = radix_sort_pairs(c->stream, k0, v0, k1, v1, n, 64, &which);
    if (rc) return rc;
    FVC_HIP(hipMemcpyAsync(dperm, which ? v1 : v0, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
    return FVC_OK;
}

static int zsort_keys(fvc_ctx *c, int64_t n, int *dperm, const double *dxyz, const double lo[3], const double hi[3],
                      const int32_t *res) {
    uint64_t *k[2] = {nullptr, nullptr};
    int *v[2] = {nullptr, nullptr};
    FVC_HIP(hipMallocAsync((void **)&k[0], sizeof(uint64_t) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&k[1], sizeof(uint64_t) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&v[0], sizeof(int) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&v[1], sizeof(int) * n, c->stream));
    if (dxyz) {
        double sc[3];
        for (int a = 0; a < 3; a++) sc[a] = hi[a] > lo[a] ? 2097152.0 / (hi[a] - lo[a]) : 0.0;
        k_morton<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(dxyz, n, lo[0], lo[1], lo[2], sc[0], sc[1], sc[2], k[0],
                                                              v[0]);
    } else {
        k_grid_morton<<<grid_for(n, BLOCK), BLOCK, 0, c->stream>>>(res[1], res[2], n, k[0], v[0]);
    }
    int rc = sort_pairs(c, k[0], v[0], n, dperm, k[1], v[1]);
    for (int a = 0; a < 2; a++) {
        (void)hipFreeAsync(k[a], c->stream);
        (void)hipFreeAsync(v[a], c->stream);
    }
    if (rc) return rc;
    FVC_HIP(hipStreamSynchronize(c->stream));
    return FVC_OK;
}

int launch_zsort(fvc_ctx *c, const double *dxyz, int64_t n, const double lo[3], const double hi[3], int *dperm) {
    if (n <= 0) return FVC_OK;
    if (n > INT32_MAX / 2) return FVC_E_INVALID;
    return zsort_keys(c, n, dperm, dxyz, lo, hi, nullptr);
}

int launch_grid_zorder(fvc_ctx *c, const int32_t res[3], int64_t n, int *dperm) {
    if (n <= 0) return FVC_OK;
    if (n > INT32_MAX / 2 || res[0] > (1 << 21) || res[1] > (1 << 21) || res[2] > (1 << 21)) return FVC_E_INVALID;
    return zsort_keys(c, n, dperm, nullptr, nullptr, nullptr, res);
}

} // namespace fvc
