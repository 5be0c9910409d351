/*
 * fvc_batch.hip - many tracer releases fused into one persistent kernel.
 *
 * A release is a contiguous range of particles with its own start and stop
 * time and integrator.  The work per release is what the tracer driver does
 * for it (tracers.c:202-229 through main.c:170-227): GetVelocity at the
 * release time on the data frame holding it, then one Advect per data frame
 * over [max(t_start, tmin), min(t_stop, tmax)] until the frame whose end
 * is within TINY of t_stop or beyond it, skipping points that left (LeftDomain) or carry LeftDomainTime <=
 * -TINY.  Frame f interpolates between slots f mod nslots and (f+1) mod
 * nslots over [t0 + f*dt_frame, t0 + (f+1)*dt_frame].
 *
 * Each particle carries its frame and the resumable state of its current
 * interval (time, step size, stage-probe elements), so a launch gives every
 * active particle a budget of S steps and the next launch continues with the
 * same arithmetic (interval_run in fvc_device.h, the code fvc_advect runs):
 * results do not depend on S, on the block size or on the order of work.
 * After each launch the active list is compacted on the device (stable);
 * every count lives in device memory, so a group of launches is captured once
 * in a hipGraph and replayed until the list is empty.
 *
 * Work distribution: a persistent grid sized to the occupancy limit.  Each
 * wave takes 64 consecutive list entries at a time from one of eight queues;
 * queue x covers the x-th eighth of the (index-ordered) active list and is
 * served first by the blocks placed on XCD x (the hardware deals blocks to
 * the eight XCDs round-robin), so each XCD's L2 holds one compact region of
 * the mesh; a wave whose queue is empty then takes from the others.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "fvc_internal.h"

using namespace fvc;

namespace {

constexpr int NQ = 8;       /* XCDs per MI355X */
constexpr int TILE = 4096;  /* list entries per compaction block */
constexpr int GROUP = 4;    /* launches per graph replay (even: parity returns) */

struct DevRelease {
    double t_start, t_stop;
    int itype;
    int f0; /* data frame holding t_start */
};

struct DevSlot {
    const void *ptr;
    float scale;
    float pad;
};

enum { FL_DONE = 2, FL_STARTED = 4 };

struct BatchDev {
    PtrSet p;
    const DevRelease *rel;
    const DevSlot *slots;
    int nslots;
    double t0, fdt, dt;
    const int *rel_id;
    int *frame, *flags, *mpl;
    double *tc, *h;
    int4 *mp;
    const int *list_in;
    int *list_out;
    const int *n_in;
    int *n_out;
    int *queue;
    unsigned long long *steps;
    int *err;
    int64_t *where;
};

__device__ inline DevField frame_field(const BatchDev &b, int f) {
#pragma clang fp contract(off)
    DevField fld;
    int lo = f % b.nslots, hi = (f + 1) % b.nslots;
    fld.lo = b.slots[lo].ptr;
    fld.hi = b.slots[hi].ptr;
    fld.scale_lo = b.slots[lo].scale;
    fld.scale_hi = b.slots[hi].scale;
    fld.tmin = b.t0 + f * b.fdt; /* main.c:188-190 */
    fld.tmax = fld.tmin + b.fdt;
    return fld;
}

/* one particle, at most `budget` steps; returns with its state stored */
template <int MODE, int GEN>
__device__ void batch_particle(const DevMesh &m, const BatchDev &b, int i, int budget, uint32_t &ns) {
#pragma clang fp contract(off)
    int fl = b.flags[i];
    if (fl & FL_DONE) return;
    const DevRelease R = b.rel[b.rel_id[i]];
    const bool hist = GEN && (R.itype == INT_AB2 || R.itype == INT_AB3);
    PState s;
    /* V is set at release time and not used again (no extrapolation) */
    for (int c = 0; c < 3; c++) {
        s.X[c] = b.p.X[c][i];
        s.V[c] = 0.0;
    }
    s.elem = b.p.elem[i];
    s.left = b.p.left[i];
    s.left_time = b.p.left_time[i];
    s.nhist = 0;
    if (hist) {
        for (int c = 0; c < 3; c++) {
            s.F1[c] = b.p.H[c][i];
            s.F2[c] = b.p.H[3 + c][i];
        }
        s.nhist = b.p.nhist[i];
    }
    int frame = b.frame[i];
    IState is;
    is.started = (fl & FL_STARTED) ? 1 : 0;
    is.tc = 0.0;
    is.h = 0.0;
    for (int q = 0; q < 4; q++) is.mp[q] = Probe{0, 0};
    if (is.started) {
        is.tc = b.tc[i];
        is.h = b.h[i];
        const int4 me = b.mp[i];
        const int ml = b.mpl[i];
        is.mp[0].elem = me.x;
        is.mp[1].elem = me.y;
        is.mp[2].elem = me.z;
        is.mp[3].elem = me.w;
        for (int q = 0; q < 4; q++) is.mp[q].left = (ml >> q) & 1;
    }
    int err = 0;
    if (!(s.left_time > -TINY && !s.left)) fl |= FL_DONE; /* integration filter, tracers.c:224 */
    int left_budget = budget;
    while (!err && !(fl & FL_DONE) && left_budget > 0) {
        const DevField fld = frame_field(b, frame);
        const double ia = R.t_start > fld.tmin ? R.t_start : fld.tmin;
        const double ib = R.t_stop < fld.tmax ? R.t_stop : fld.tmax;
        const uint32_t before = ns;
        double stop;
        const int finished =
            interval_run<MODE, 3, 1, GEN>(m, fld, s, is, ia, ib, b.dt, R.itype, 0, left_budget, err, ns, stop);
        if (err) break;
        if (!finished) { /* budget spent inside the interval */
            fl |= FL_STARTED;
            break;
        }
        fl &= ~FL_STARTED;
        is.started = 0;
        /* every pass consumes at least one unit, so a launch always ends */
        left_budget -= ns - before > 0 ? (int)(ns - before) : 1;
        if (s.left || R.t_stop - fld.tmax < TINY) { /* status 2 rule, tracers.c:283-284 */
            fl |= FL_DONE;
            break;
        }
        frame++;
    }
    if (err) {
        if (atomicCAS(b.err, 0, err) == 0) *b.where = i;
        fl |= FL_DONE;
    }
    for (int c = 0; c < 3; c++) b.p.X[c][i] = s.X[c];
    b.p.elem[i] = s.elem;
    b.p.left[i] = s.left;
    b.p.left_time[i] = s.left_time;
    if (hist) {
        for (int c = 0; c < 3; c++) {
            b.p.H[c][i] = s.F1[c];
            b.p.H[3 + c][i] = s.F2[c];
        }
        b.p.nhist[i] = s.nhist;
    }
    b.frame[i] = frame;
    b.flags[i] = fl;
    if (fl & FL_STARTED) {
        b.tc[i] = is.tc;
        b.h[i] = is.h;
        b.mp[i] = make_int4(is.mp[0].elem, is.mp[1].elem, is.mp[2].elem, is.mp[3].elem);
        b.mpl[i] = is.mp[0].left | (is.mp[1].left << 1) | (is.mp[2].left << 2) | (is.mp[3].left << 3);
    }
}

/* GEN = 0: RK4 releases only (no multistep state in registers), held to
 * 3 waves/SIMD (at most 168 VGPRs); 1: the other integrators, at 2 waves/SIMD
 * (about 235 VGPRs, where 3 waves spill 70-87 of them). */
template <int MODE, int BS, int GEN>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(GEN ? 2 : FVC_HOT_WAVES))) void k_batch(DevMesh m, BatchDev b,
                                                                                      int budget) {
    const int lane = threadIdx.x & 63;
    const int home = blockIdx.x % NQ;
    const int n = *b.n_in;
    uint32_t ns = 0;
    for (int k = 0; k < NQ; k++) {
        const int q = (home + k) % NQ;
        const int seg_lo = (int)((int64_t)n * q / NQ), seg_hi = (int)((int64_t)n * (q + 1) / NQ);
        for (;;) {
            const int base = claim64(&b.queue[q], seg_hi - seg_lo);
            if (seg_lo + base >= seg_hi) break;
            const int idx = seg_lo + base + lane;
            if (idx < seg_hi) batch_particle<MODE, GEN>(m, b, b.list_in[idx], budget, ns);
        }
    }
    unsigned long long v = ns;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0 && v) atomicAdd(b.steps, v);
}

/* release: GetVelocity at the launch time, on the frame holding it, for
 * points still inside (tracers.c:202-215) */
template <int MODE>
__global__ __launch_bounds__(256) void k_batch_launch(DevMesh m, BatchDev b, int64_t n) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = b.rel_id[i];
    if (r < 0 || b.p.left[i]) return;
    const DevRelease R = b.rel[r];
    PState s;
    for (int c = 0; c < 3; c++) {
        s.X[c] = b.p.X[c][i];
        s.V[c] = 0.0;
    }
    s.elem = b.p.elem[i];
    s.left = 0;
    s.left_time = b.p.left_time[i];
    int err = 0;
    double v[3];
    get_velocity<MODE, 3>(m, frame_field(b, R.f0), R.t_start, s, v, err);
    if (err) {
        if (atomicCAS(b.err, 0, err) == 0) *b.where = i;
        b.flags[i] = FL_DONE;
        return;
    }
    for (int c = 0; c < 3; c++) b.p.V[c][i] = v[c];
    b.p.elem[i] = s.elem;
    b.p.left[i] = s.left;
}

__global__ void k_batch_prep(int *queue) {
    if (threadIdx.x < NQ) queue[threadIdx.x] = 0;
}

/* reset: frame and flags from the release table */
__global__ void k_batch_init(const DevRelease *rel, const int *rel_id, int64_t n, int *frame, int *flags) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r = rel_id[i];
    frame[i] = r >= 0 ? rel[r].f0 : 0;
    flags[i] = r >= 0 ? 0 : FL_DONE;
}

__global__ void k_set_count(int *dst, int v) {
    if (threadIdx.x == 0) *dst = v;
}

/* stable compaction of list_in -> list_out keeping entries not done */
__global__ __launch_bounds__(256) void k_bc_count(const int *list_in, const int *n_in, const int *flags,
                                                  int *tile_counts) {
    const int n = *n_in;
    const int64_t base = (int64_t)blockIdx.x * TILE;
    int cnt = 0;
    if (base < n)
        for (int q = threadIdx.x; q < TILE; q += 256) {
            const int64_t i = base + q;
            if (i < n && !(flags[list_in[i]] & FL_DONE)) cnt++;
        }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    __shared__ int red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(1024) void k_bc_scan(int *a, int ntiles, int *n_out) {
    __shared__ int tmp[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < ntiles; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < ntiles ? a[i] : 0;
        tmp[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int add = threadIdx.x >= off ? tmp[threadIdx.x - off] : 0;
            __syncthreads();
            tmp[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < ntiles) a[i] = carry + tmp[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += tmp[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *n_out = carry;
}

__global__ __launch_bounds__(256) void k_bc_scatter(const int *list_in, const int *n_in, const int *flags,
                                                    const int *offs, int *list_out) {
    const int n = *n_in;
    const int64_t base = (int64_t)blockIdx.x * TILE;
    if (base >= n) return;
    __shared__ int wsum[4];
    __shared__ int run;
    if (threadIdx.x == 0) run = offs[blockIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int r = 0; r < TILE; r += 256) {
        const int64_t i = base + r + threadIdx.x;
        int keep = 0, v = 0;
        if (i < n) {
            v = list_in[i];
            keep = !(flags[v] & FL_DONE);
        }
        const unsigned long long bal = __ballot(keep);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(bal);
        __syncthreads();
        int pre = 0;
        for (int q = 0; q < w; q++) pre += wsum[q];
        if (keep) list_out[run + pre + before] = v;
        __syncthreads();
        if (threadIdx.x == 0) run += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

/* frame holding t: t0 + f*dt <= t < t0 + (f+1)*dt, evaluated as the kernels do */
static int frame_of(double t, double t0, double fdt) {
#pragma clang fp contract(off)
    double q = std::floor((t - t0) / fdt);
    if (!(std::fabs(q) < 1.0e9)) return INT32_MIN;
    int f = (int)q;
    while (t0 + f * fdt > t) f--;
    while (t0 + (f + 1) * fdt <= t) f++;
    return f;
}

} // namespace

struct fvc_batch {
    fvc_ctx *ctx;
    fvc_particles *p;
    fvc_particles *snap;
    int64_t n;
    int64_t n0;         /* particles in releases */
    int64_t cls_n0[2];  /* class 0: RK4 releases, class 1: the others */
    int nrel;
    double t0, fdt, dt;
    int nslots;
    int64_t max_budget; /* per-particle bound on budget units for the whole run */
    DevRelease *d_rel;
    DevSlot *d_slots;
    void *buf;
    int *rel_id, *frame, *flags, *mpl, *list0, *lists[2], *counts, *queue, *tiles;
    double *tc, *h;
    int4 *mp;
    int ntiles[2];
    int launched; /* release-time velocities done */
    /* captured launch groups, one per (block, budget, grid) configuration
     * run so far; kept until the batch is released, so a sweep captures each
     * configuration once and never destroys a graph between runs */
    struct Group {
        int block, budget, grid0, grid1;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Group> groups;
};

/* class k's lists live at offset cls_off(k) of each list buffer */
static int64_t cls_off(const fvc_batch *b, int k) { return k ? b->cls_n0[0] : 0; }

static BatchDev batch_dev(fvc_batch *b, int k, int parity) {
    BatchDev d;
    d.p = ptrs(b->p);
    d.rel = b->d_rel;
    d.slots = b->d_slots;
    d.nslots = b->nslots;
    d.t0 = b->t0;
    d.fdt = b->fdt;
    d.dt = b->dt;
    d.rel_id = b->rel_id;
    d.frame = b->frame;
    d.flags = b->flags;
    d.mpl = b->mpl;
    d.tc = b->tc;
    d.h = b->h;
    d.mp = b->mp;
    d.list_in = b->lists[parity] + cls_off(b, k);
    d.list_out = b->lists[parity ^ 1] + cls_off(b, k);
    d.n_in = b->counts + 2 * k + parity;
    d.n_out = b->counts + 2 * k + (parity ^ 1);
    d.queue = b->queue + NQ * k;
    d.steps = b->ctx->d_steps;
    d.err = b->ctx->d_err;
    d.where = b->ctx->d_where;
    return d;
}

/* block sizes with a compiled kernel; FP64 / BF16 / FP16 use 256 only */
static bool block_ok(int prec, int bs) {
    if (prec == FVC_PREC_F32) return bs == 64 || bs == 128 || bs == 256 || bs == 512 || bs == 1024;
    return bs == 256;
}

template <int MODE, int BS, int GEN>
static const void *kfun() {
    return reinterpret_cast<const void *>(&k_batch<MODE, BS, GEN>);
}

template <int GEN>
static const void *batch_kernel(int prec, int bs) {
    switch (prec) {
    case FVC_PREC_F64: return kfun<P_F64, 256, GEN>();
    case FVC_PREC_BF16: return kfun<P_BF16, 256, GEN>();
    case FVC_PREC_F16: return kfun<P_F16, 256, GEN>();
    default:
        switch (bs) {
        case 64: return kfun<P_F32, 64, GEN>();
        case 128: return kfun<P_F32, 128, GEN>();
        case 512: return kfun<P_F32, 512, GEN>();
        case 1024: return kfun<P_F32, 1024, GEN>();
        default: return kfun<P_F32, 256, GEN>();
        }
    }
}

static int grid_for_batch(fvc_batch *b, int k, int bs, int waves_per_cu) {
    fvc_ctx *c = b->ctx;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0)
        cus = 256;
    int blocks_per_cu = 0;
    const void *fn = k ? batch_kernel<1>(c->precision, bs) : batch_kernel<0>(c->precision, bs);
    if (waves_per_cu > 0) {
        blocks_per_cu = std::max(1, waves_per_cu / (bs / 64));
    } else if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, fn, bs, 0) != hipSuccess ||
               blocks_per_cu <= 0) {
        blocks_per_cu = std::max(1, 8 / (bs / 64));
    }
    const int grid = cus * blocks_per_cu;
    return (grid + NQ - 1) / NQ * NQ;
}

template <int MODE, int BS>
static void launch_k(fvc_batch *b, int k, int grid, int budget, const BatchDev &d) {
    if (k)
        k_batch<MODE, BS, 1><<<grid, BS, 0, b->ctx->stream>>>(b->ctx->mesh, d, budget);
    else
        k_batch<MODE, BS, 0><<<grid, BS, 0, b->ctx->stream>>>(b->ctx->mesh, d, budget);
}

/* one launch of the sequence for class k: queue reset, batch kernel, compaction */
static void enqueue_class(fvc_batch *b, int k, int bs, int grid, int budget, int parity) {
    fvc_ctx *c = b->ctx;
    const BatchDev d = batch_dev(b, k, parity);
    k_batch_prep<<<1, 64, 0, c->stream>>>(d.queue);
    switch (c->precision) {
    case FVC_PREC_F64: launch_k<P_F64, 256>(b, k, grid, budget, d); break;
    case FVC_PREC_BF16: launch_k<P_BF16, 256>(b, k, grid, budget, d); break;
    case FVC_PREC_F16: launch_k<P_F16, 256>(b, k, grid, budget, d); break;
    default:
        switch (bs) {
        case 64: launch_k<P_F32, 64>(b, k, grid, budget, d); break;
        case 128: launch_k<P_F32, 128>(b, k, grid, budget, d); break;
        case 512: launch_k<P_F32, 512>(b, k, grid, budget, d); break;
        case 1024: launch_k<P_F32, 1024>(b, k, grid, budget, d); break;
        default: launch_k<P_F32, 256>(b, k, grid, budget, d); break;
        }
    }
    int *tiles = b->tiles + (k ? b->ntiles[0] + 1 : 0);
    k_bc_count<<<b->ntiles[k], 256, 0, c->stream>>>(d.list_in, d.n_in, b->flags, tiles);
    k_bc_scan<<<1, 1024, 0, c->stream>>>(tiles, b->ntiles[k], d.n_out);
    k_bc_scatter<<<b->ntiles[k], 256, 0, c->stream>>>(d.list_in, d.n_in, b->flags, tiles, d.list_out);
}

static void enqueue_group(fvc_batch *b, int bs, const int grid[2], int budget) {
    for (int r = 0; r < GROUP; r++)
        for (int k = 0; k < 2; k++)
            if (b->cls_n0[k]) enqueue_class(b, k, bs, grid[k], budget, r & 1);
}

static int set_err(fvc_ctx *c, int code, const char *msg) {
    c->last_error = msg;
    return code;
}

extern "C" int fvc_batch_create(fvc_ctx *c, fvc_particles *p, const fvc_release *rel, int32_t nrel, double t0,
                                double fdt, int32_t nslots, double dt, fvc_batch **out) {
    if (!c || !p || !rel || nrel <= 0 || !out || p->ctx != c) return FVC_E_INVALID;
    *out = nullptr;
    if (!(fdt > 0) || !(dt > 1.0e-10) || nslots < 1 || nslots > c->max_slices)
        return set_err(c, FVC_E_INVALID, "batch: frame spacing, time step or slot count out of range");
    if (c->mesh.mesh_type != 1 || c->mesh.dims != 3)
        return set_err(c, FVC_E_INVALID, "batched releases need a 3-D unstructured mesh");
    for (int s = 0; s < nslots; s++)
        if (!c->slices[s]) return set_err(c, FVC_E_STATE, "batched releases need slots 0..nslots-1 loaded");
    std::vector<DevRelease> dr(nrel);
    int64_t n = 0, n0 = 0, max_budget = 1;
    for (int r = 0; r < nrel; r++) {
        const fvc_release &R = rel[r];
        if (R.begin < 0 || R.count < 0 || R.begin + R.count > p->cap || !(R.t_stop >= R.t_start) ||
            (R.int_type != FVC_INT_RK4 && R.int_type != FVC_INT_EULER && R.int_type != FVC_INT_AB2 &&
             R.int_type != FVC_INT_AB3))
            return set_err(c, FVC_E_INVALID, "batch: invalid release (range, times or integrator)");
        const int f0 = frame_of(R.t_start, t0, fdt);
        if (f0 == INT32_MIN || f0 < 0) return set_err(c, FVC_E_INVALID, "batch: release starts before frame 0");
        const double frames = std::ceil((R.t_stop - R.t_start) / fdt) + 2;
        /* budget units: one per counted step, one per frame pass */
        const double units = std::ceil((R.t_stop - R.t_start) / dt) + 3 * frames;
        if (frames + f0 > 1.0e9 || units > 1.0e12) return set_err(c, FVC_E_INVALID, "batch: release too long");
        dr[r] = DevRelease{R.t_start, R.t_stop, R.int_type, f0};
        max_budget = std::max(max_budget, (int64_t)units);
        n = std::max(n, R.begin + R.count);
        n0 += R.count;
    }
    if (n > INT32_MAX) return set_err(c, FVC_E_INVALID, "batch: more than 2^31-1 particles");
    std::vector<int> rid((size_t)n, -1);
    for (int r = 0; r < nrel; r++)
        for (int64_t i = rel[r].begin; i < rel[r].begin + rel[r].count; i++) {
            if (rid[i] >= 0) return set_err(c, FVC_E_INVALID, "batch: releases overlap");
            rid[i] = r;
        }
    /* initial active lists: RK4 particles, then the others, each in index order */
    std::vector<int> list;
    list.reserve((size_t)n0);
    int64_t cls_n0[2] = {0, 0};
    for (int k = 0; k < 2; k++)
        for (int64_t i = 0; i < n; i++)
            if (rid[i] >= 0 && (rel[rid[i]].int_type != FVC_INT_RK4) == (k == 1)) {
                list.push_back((int)i);
                cls_n0[k]++;
            }

    fvc_batch *b = new fvc_batch();
    b->ctx = c;
    c->live_batches.insert(b);
    b->p = p;
    b->n = n;
    b->n0 = n0;
    b->cls_n0[0] = cls_n0[0];
    b->cls_n0[1] = cls_n0[1];
    b->nrel = nrel;
    b->t0 = t0;
    b->fdt = fdt;
    b->dt = dt;
    b->nslots = nslots;
    b->max_budget = max_budget;
    for (int k = 0; k < 2; k++) b->ntiles[k] = (int)std::max<int64_t>(1, (cls_n0[k] + TILE - 1) / TILE);
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    const size_t sizes[] = {8 * nn, 8 * nn, 16 * nn, 4 * nn, 4 * nn, 4 * nn, 4 * nn, 4 * nn, 4 * nn, 4 * nn,
                            64,     8 * NQ, 4 * ((size_t)b->ntiles[0] + b->ntiles[1] + 2)};
    size_t bytes = 0;
    for (size_t s : sizes) bytes += (s + 255) / 256 * 256;
    if (hipMalloc(&b->buf, bytes) != hipSuccess) {
        fvc_batch_destroy(b);
        return set_err(c, FVC_E_NOMEM, "batch: device memory");
    }
    char *q = (char *)b->buf;
    int k = 0;
    auto take = [&]() {
        void *r = q;
        q += (sizes[k++] + 255) / 256 * 256;
        return r;
    };
    b->tc = (double *)take();
    b->h = (double *)take();
    b->mp = (int4 *)take();
    b->rel_id = (int *)take();
    b->frame = (int *)take();
    b->flags = (int *)take();
    b->mpl = (int *)take();
    b->list0 = (int *)take();
    b->lists[0] = (int *)take();
    b->lists[1] = (int *)take();
    b->counts = (int *)take();
    b->queue = (int *)take();
    b->tiles = (int *)take();
    int rc = FVC_OK;
    if (hipMalloc((void **)&b->d_rel, sizeof(DevRelease) * nrel) != hipSuccess ||
        hipMalloc((void **)&b->d_slots, sizeof(DevSlot) * nslots) != hipSuccess)
        rc = set_err(c, FVC_E_NOMEM, "batch: device memory");
    std::vector<DevSlot> sl(nslots);
    for (int s = 0; s < nslots; s++) sl[s] = DevSlot{c->slices[s], c->slice_scale[s], 0.0f};
    if (!rc && (hipMemcpyAsync(b->d_rel, dr.data(), sizeof(DevRelease) * nrel, hipMemcpyHostToDevice, c->stream) ||
                hipMemcpyAsync(b->d_slots, sl.data(), sizeof(DevSlot) * nslots, hipMemcpyHostToDevice, c->stream) ||
                hipMemcpyAsync(b->rel_id, rid.data(), 4 * (size_t)n, hipMemcpyHostToDevice, c->stream) ||
                (n0 && hipMemcpyAsync(b->list0, list.data(), 4 * (size_t)n0, hipMemcpyHostToDevice, c->stream)) ||
                hipStreamSynchronize(c->stream)))
        rc = set_err(c, FVC_E_HIP, "batch: upload failed");
    if (!rc) rc = fvc_particles_create(c, n, &b->snap);
    if (!rc) rc = fvc_particles_copy(b->snap, 0, p, 0, n);
    if (!rc) rc = fvc_batch_reset(b);
    if (rc) {
        fvc_batch_destroy(b);
        return rc;
    }
    c->bytes_used += (int64_t)bytes;
    *out = b;
    return FVC_OK;
}

void fvc::batch_release(fvc_batch *b) {
    fvc_ctx *c = b->ctx;
    if (!c) return;
    hipStreamSynchronize(c->stream);
    for (auto &g : b->groups) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        if (g.graph) hipGraphDestroy(g.graph);
    }
    b->groups.clear();
    if (b->buf) hipFree(b->buf);
    if (b->d_rel) hipFree(b->d_rel);
    if (b->d_slots) hipFree(b->d_slots);
    b->buf = nullptr;
    b->d_rel = nullptr;
    b->d_slots = nullptr;
    fvc_particles_destroy(b->snap);
    b->snap = nullptr;
    c->live_batches.erase(b);
    b->ctx = nullptr;
    b->p = nullptr;
}

void fvc::batches_release_on(fvc_ctx *c, fvc_particles *p) {
    std::vector<fvc_batch *> on;
    for (fvc_batch *b : c->live_batches)
        if (b->p == p) on.push_back(b);
    for (fvc_batch *b : on) batch_release(b);
}

extern "C" void fvc_batch_destroy(fvc_batch *b) {
    if (!b) return;
    batch_release(b);
    delete b;
}

extern "C" int fvc_batch_reset(fvc_batch *b) {
    if (!b || !b->ctx) return FVC_E_INVALID;
    fvc_ctx *c = b->ctx;
    int rc = fvc_particles_copy(b->p, 0, b->snap, 0, b->n);
    if (rc) return rc;
    if (b->n > 0)
        k_batch_init<<<(unsigned)((b->n + 255) / 256), 256, 0, c->stream>>>(b->d_rel, b->rel_id, b->n, b->frame,
                                                                            b->flags);
    if (b->n0) FVC_HIP(hipMemcpyAsync(b->lists[0], b->list0, 4 * (size_t)b->n0, hipMemcpyDeviceToDevice, c->stream));
    for (int k = 0; k < 2; k++) {
        k_set_count<<<1, 64, 0, c->stream>>>(b->counts + 2 * k, (int)b->cls_n0[k]);
        k_set_count<<<1, 64, 0, c->stream>>>(b->counts + 2 * k + 1, 0);
    }
    FVC_HIP(hipGetLastError());
    FVC_HIP(hipStreamSynchronize(c->stream));
    b->launched = 0;
    return FVC_OK;
}

static int launch_releases(fvc_batch *b) {
    fvc_ctx *c = b->ctx;
    if (b->launched || b->n == 0) return FVC_OK;
    const BatchDev d = batch_dev(b, 0, 0);
    const unsigned g = (unsigned)((b->n + 255) / 256);
    switch (c->precision) {
    case FVC_PREC_F64: k_batch_launch<P_F64><<<g, 256, 0, c->stream>>>(c->mesh, d, b->n); break;
    case FVC_PREC_BF16: k_batch_launch<P_BF16><<<g, 256, 0, c->stream>>>(c->mesh, d, b->n); break;
    case FVC_PREC_F16: k_batch_launch<P_F16><<<g, 256, 0, c->stream>>>(c->mesh, d, b->n); break;
    default: k_batch_launch<P_F32><<<g, 256, 0, c->stream>>>(c->mesh, d, b->n); break;
    }
    FVC_HIP(hipGetLastError());
    b->launched = 1;
    return FVC_OK;
}

extern "C" int fvc_batch_run(fvc_batch *b, const fvc_batch_params *bp, int64_t *particle_steps, int64_t *launches) {
    if (!b || !b->ctx || !bp) return FVC_E_INVALID;
    fvc_ctx *c = b->ctx;
    const int bs = bp->block;
    if (!block_ok(c->precision, bs))
        return set_err(c, FVC_E_INVALID, "batch: block size must be 64..1024 (a power of two) in FP32, 256 otherwise");
    if (bp->steps_per_launch < 0 || bp->waves_per_cu < 0 || bp->waves_per_cu > 64)
        return set_err(c, FVC_E_INVALID, "batch: steps_per_launch or waves_per_cu out of range");
    const int budget = bp->steps_per_launch == 0 ? 0x7fffffff : bp->steps_per_launch;
    const int grid[2] = {grid_for_batch(b, 0, bs, bp->waves_per_cu), grid_for_batch(b, 1, bs, bp->waves_per_cu)};
    /* every launch takes at least one budget unit from each active particle */
    const int64_t max_rounds = b->max_budget / std::min<int64_t>(budget, b->max_budget) + 2;
    FVC_HIP(hipMemsetAsync(c->d_steps, 0, sizeof(unsigned long long), c->stream));
    timing_begin(c); /* with timing on: device time of the whole run */
    int rc0 = launch_releases(b);
    if (rc0) return rc0;
    hipGraphExec_t gexec = nullptr;
    if (bp->use_graph) {
        for (const auto &g : b->groups)
            if (g.block == bs && g.budget == budget && g.grid0 == grid[0] && g.grid1 == grid[1]) gexec = g.exec;
        if (!gexec) {
            fvc_batch::Group g{bs, budget, grid[0], grid[1], nullptr, nullptr};
            FVC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            enqueue_group(b, bs, grid, budget);
            const hipError_t e = hipStreamEndCapture(c->stream, &g.graph);
            if (e != hipSuccess) {
                c->last_error = std::string("batch: graph capture: ") + hipGetErrorString(e);
                return FVC_E_HIP;
            }
            const hipError_t ei = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                hipGraphDestroy(g.graph);
                c->last_error = std::string("batch: graph instantiate: ") + hipGetErrorString(ei);
                return FVC_E_HIP;
            }
            b->groups.push_back(g);
            gexec = g.exec;
        }
    }
    int64_t rounds = 0;
    int active[4] = {(int)b->cls_n0[0], 0, (int)b->cls_n0[1], 0};
    while (active[0] > 0 || active[2] > 0) {
        if (rounds >= max_rounds + GROUP)
            return set_err(c, FVC_E_STATE, "batch: particles still active after the step bound");
        if (bp->use_graph) {
            FVC_HIP(hipGraphLaunch(gexec, c->stream));
        } else {
            enqueue_group(b, bs, grid, budget);
            FVC_HIP(hipGetLastError());
        }
        rounds += GROUP;
        int code = 0;
        FVC_HIP(hipMemcpyAsync(active, b->counts, sizeof active, hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipMemcpyAsync(&code, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipStreamSynchronize(c->stream));
        if (code) break;
    }
    timing_end(c);
    int rc = check_err(c, "fvc_batch_run");
    unsigned long long st = 0;
    if (hipMemcpy(&st, c->d_steps, sizeof st, hipMemcpyDeviceToHost) != hipSuccess && !rc) rc = FVC_E_HIP;
    c->last_steps = (int64_t)st;
    timing_collect(c, (int64_t)st);
    if (particle_steps) *particle_steps = (int64_t)st;
    if (launches) *launches = rounds;
    return rc;
}

extern "C" int fvc_batch_tune(fvc_batch *b, const int32_t *blocks, int32_t nblocks, const int32_t *budgets,
                              int32_t nbudgets, int32_t use_graph, fvc_batch_params *best, double *best_ms) {
    if (!b || !b->ctx || !blocks || !budgets || nblocks <= 0 || nbudgets <= 0 || !best) return FVC_E_INVALID;
    fvc_ctx *c = b->ctx;
    double best_t = 1e300;
    fvc_batch_params bp_best{};
    bool warmed = false;
    for (int i = 0; i < nblocks; i++) {
        if (!block_ok(c->precision, blocks[i])) continue;
        for (int j = 0; j < nbudgets; j++) {
            if (budgets[j] < 0) return set_err(c, FVC_E_INVALID, "batch: negative step budget in the sweep");
            fvc_batch_params bp{blocks[i], budgets[j], use_graph, 0};
            int rc;
            if (!warmed) { /* the first run pays code-object loading */
                if ((rc = fvc_batch_reset(b)) || (rc = fvc_batch_run(b, &bp, nullptr, nullptr))) return rc;
                warmed = true;
            }
            if ((rc = fvc_batch_reset(b))) return rc;
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = fvc_batch_run(b, &bp, nullptr, nullptr))) return rc;
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (ms < best_t) {
                best_t = ms;
                bp_best = bp;
            }
        }
    }
    if (!warmed) return set_err(c, FVC_E_INVALID, "batch: no block size in the sweep is available");
    *best = bp_best;
    if (best_ms) *best_ms = best_t;
    return fvc_batch_reset(b);
}
