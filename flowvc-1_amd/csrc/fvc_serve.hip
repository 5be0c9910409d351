/*
 * fvc_serve.hip - tracer and FTLE requests served with continuous batching.
 *
 * Device side: one particle pool (sized from free HBM) holds the particles of
 * every running request in contiguous ranges.  Each tick runs one launch of a
 * persistent kernel over the active list of all requests (XCD-aware queues as
 * in fvc_batch.hip), with the per-particle resumable interval state, so a
 * request admitted between two ticks simply adds its range to the list.
 *
 *   tracer particles run frame after frame until their stop time or exit,
 *   exactly as one release of fvc_batch_run;
 *   FTLE points run one data frame and then wait at a per-request barrier:
 *   their frame-start state stays in place (X, elem) and the frame's result
 *   is staged (Xn, elem_new, exited, exit_time) as in ComputeFTLE
 *   (ftle.c:109-169).  When every point of the request has reached the
 *   barrier the host runs the frame's early-FTLE passes and position update
 *   on that request's grid and either releases the next frame or finishes
 *   the request (GetFTLEForPoint fill, ftle.c:1304-1315).
 *
 * A per-request counter of particles still in the current phase (whole run
 * for tracers, frame for FTLE) is decremented by the kernel; the host reads
 * the counters after each tick.  Launches of one tick are replayed from a
 * captured hipGraph (one per list parity).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

#include "fvc_internal.h"

using namespace fvc;

namespace {

constexpr int NQ = 8;      /* XCDs per MI355X */
constexpr int TILE = 4096; /* list entries per compaction block */

enum { SF_DONE = 2, SF_STARTED = 4, SF_WAIT = 8 };

struct SrvReq {
    double t_start, t_stop;
    int kind, itype;
};

struct SrvSlot {
    const void *ptr;
    float scale;
    float pad;
};

struct SrvDev {
    PtrSet p;
    double *Xn[3];
    int *elem_new, *exited, *advected;
    double *exit_time;
    const SrvReq *req;
    const int *req_id;
    int *frame, *flags, *mpl;
    double *tc, *h;
    int4 *mp;
    const int *list_in;
    int *list_out;
    const int *n_in;
    int *n_out;
    int *queue;
    int *remaining;
    unsigned long long *req_steps;
    const SrvSlot *slots;
    int nslots;
    double t0, fdt, dt;
    unsigned long long *steps;
    int *err;
    int64_t *where;
};

__device__ inline DevField srv_field(const SrvDev &S, int f) {
#pragma clang fp contract(off)
    DevField fld;
    const int lo = f % S.nslots, hi = (f + 1) % S.nslots;
    fld.lo = S.slots[lo].ptr;
    fld.hi = S.slots[hi].ptr;
    fld.scale_lo = S.slots[lo].scale;
    fld.scale_hi = S.slots[hi].scale;
    fld.tmin = S.t0 + f * S.fdt; /* main.c:188-190 */
    fld.tmax = fld.tmin + S.fdt;
    return fld;
}

/* one particle for at most `budget` steps; returns the steps it took */
template <int MODE, int GEN>
__device__ uint32_t serve_particle(const DevMesh &m, const SrvDev &S, int i, int budget, int &r_out, int &fin) {
#pragma clang fp contract(off)
    int fl = S.flags[i];
    if (fl & (SF_DONE | SF_WAIT)) return 0;
    const int r = S.req_id[i];
    r_out = r;
    const SrvReq R = S.req[r];
    const bool ftle = R.kind == FVC_REQ_FTLE;
    const bool hist = GEN && !ftle && (R.itype == INT_AB2 || R.itype == INT_AB3);
    const bool started = (fl & SF_STARTED) != 0;
    PState s;
    for (int c = 0; c < 3; c++) {
        s.X[c] = (ftle && started) ? S.Xn[c][i] : S.p.X[c][i];
        s.V[c] = 0.0;
    }
    s.elem = (ftle && started) ? S.elem_new[i] : S.p.elem[i];
    s.left = S.p.left[i];
    s.left_time = S.p.left_time[i];
    s.nhist = 0;
    if (hist) {
        for (int c = 0; c < 3; c++) {
            s.F1[c] = S.p.H[c][i];
            s.F2[c] = S.p.H[3 + c][i];
        }
        s.nhist = S.p.nhist[i];
    }
    int frame = S.frame[i];
    IState is;
    is.started = started ? 1 : 0;
    is.tc = 0.0;
    is.h = 0.0;
    for (int q = 0; q < 4; q++) is.mp[q] = Probe{0, 0};
    if (started) {
        is.tc = S.tc[i];
        is.h = S.h[i];
        const int4 me = S.mp[i];
        const int ml = S.mpl[i];
        is.mp[0].elem = me.x;
        is.mp[1].elem = me.y;
        is.mp[2].elem = me.z;
        is.mp[3].elem = me.w;
        for (int q = 0; q < 4; q++) is.mp[q].left = (ml >> q) & 1;
    }
    if (!(s.left_time > -TINY && !s.left)) { /* not integrated (tracers.c:224, ftle.c:113) */
        S.flags[i] = ftle ? SF_WAIT : SF_DONE;
        fin = 1;
        return 0;
    }
    int err = 0;
    uint32_t ns = 0;
    int left_budget = budget;
    while (!err && left_budget > 0) {
        const DevField fld = srv_field(S, frame);
        const double ia = R.t_start > fld.tmin ? R.t_start : fld.tmin;
        const double ib = R.t_stop < fld.tmax ? R.t_stop : fld.tmax;
        const uint32_t before = ns;
        double stop;
        const int finished =
            interval_run<MODE, 3, 1, GEN>(m, fld, s, is, ia, ib, S.dt, R.itype, 0, left_budget, err, ns, stop);
        if (err) break;
        if (!finished) {
            fl |= SF_STARTED;
            break;
        }
        fl &= ~SF_STARTED;
        is.started = 0;
        left_budget -= ns - before > 0 ? (int)(ns - before) : 1;
        if (ftle) { /* frame barrier */
            fl |= SF_WAIT;
            break;
        }
        if (s.left || R.t_stop - fld.tmax < TINY) { /* status 2 rule, tracers.c:283-284 */
            fl |= SF_DONE;
            break;
        }
        frame++;
    }
    if (err) {
        if (atomicCAS(S.err, 0, err) == 0) *S.where = i;
        fl = ftle ? SF_WAIT : SF_DONE;
    }
    if (ftle) {
        for (int c = 0; c < 3; c++) S.Xn[c][i] = s.X[c];
        S.elem_new[i] = s.elem;
        if (fl & SF_WAIT) {
            S.exited[i] = s.left;
            S.exit_time[i] = s.left_time;
            S.advected[i] = 1;
        }
    } else {
        for (int c = 0; c < 3; c++) S.p.X[c][i] = s.X[c];
        S.p.elem[i] = s.elem;
        S.p.left[i] = s.left;
        S.p.left_time[i] = s.left_time;
        if (hist) {
            for (int c = 0; c < 3; c++) {
                S.p.H[c][i] = s.F1[c];
                S.p.H[3 + c][i] = s.F2[c];
            }
            S.p.nhist[i] = s.nhist;
        }
    }
    S.frame[i] = frame;
    S.flags[i] = fl;
    if (fl & SF_STARTED) {
        S.tc[i] = is.tc;
        S.h[i] = is.h;
        S.mp[i] = make_int4(is.mp[0].elem, is.mp[1].elem, is.mp[2].elem, is.mp[3].elem);
        S.mpl[i] = is.mp[0].left | (is.mp[1].left << 1) | (is.mp[2].left << 2) | (is.mp[3].left << 3);
    }
    fin = (fl & (SF_DONE | SF_WAIT)) ? 1 : 0;
    return ns;
}

/* GEN = 0: RK4 requests (tracer or FTLE), 3 waves/SIMD; 1: the other
 * integrators, 2 waves/SIMD (no spills; 3 waves spill 70 VGPRs) */
template <int MODE, int BS, int GEN>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(GEN ? 2 : FVC_HOT_WAVES))) void k_serve(DevMesh m, SrvDev S,
                                                                                      int budget) {
    const int lane = threadIdx.x & 63;
    const int home = blockIdx.x % NQ;
    const int n = *S.n_in;
    unsigned long long total = 0;
    for (int k = 0; k < NQ; k++) {
        const int q = (home + k) % NQ;
        const int seg_lo = (int)((int64_t)n * q / NQ), seg_hi = (int)((int64_t)n * (q + 1) / NQ);
        for (;;) {
            const int base = claim64(&S.queue[q], seg_hi - seg_lo);
            if (seg_lo + base >= seg_hi) break;
            const int idx = seg_lo + base + lane;
            int r = -1, fin = 0;
            uint32_t ns = 0;
            if (idx < seg_hi) ns = serve_particle<MODE, GEN>(m, S, S.list_in[idx], budget, r, fin);
            total += ns;
            /* per-request step counts and phase counters: one atomic per wave
             * when the wave serves one request (the usual case: requests own
             * contiguous list ranges), per lane otherwise */
            const int r0 = __shfl(r, 0, 64);
            const bool uniform = __ballot(r != r0 && r >= 0) == 0ull;
            const unsigned long long fb = __ballot(fin);
            unsigned long long v = ns;
            if (uniform) {
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane == 0 && r0 >= 0 && v) atomicAdd(&S.req_steps[r0], v);
                if (lane == 0 && r0 >= 0 && fb) atomicSub(&S.remaining[r0], __popcll(fb));
            } else {
                if (r >= 0 && ns) atomicAdd(&S.req_steps[r], (unsigned long long)ns);
                if (fin) atomicSub(&S.remaining[r], 1);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off, 64);
    if (lane == 0 && total) atomicAdd(S.steps, total);
}

__global__ void k_srv_prep(int *queue) {
    if (threadIdx.x < 2 * NQ) queue[threadIdx.x] = 0;
}

/* append pool range [base, base+n) to a list whose length is *count */
__global__ void k_srv_append(int *list, const int *count, int64_t base, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) list[*count + k] = (int)(base + k);
}

__global__ void k_srv_add(int *count, int64_t n) {
    if (threadIdx.x == 0) *count += (int)n;
}

/* The list compaction works on the live tiles only (the list length is read
 * on the device): a fixed grid strides over them, so a tick costs the same
 * whatever the pool's capacity. */
constexpr int CMP_GRID = 1024;

__global__ __launch_bounds__(256) void k_srv_count(const int *list_in, const int *n_in, const int *flags,
                                                   int *tile_counts) {
    const int n = *n_in;
    const int nt = (n + TILE - 1) / TILE;
    __shared__ int red[4];
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const int64_t base = (int64_t)t * TILE;
        int cnt = 0;
        for (int q = threadIdx.x; q < TILE; q += 256) {
            const int64_t i = base + q;
            if (i < n && !(flags[list_in[i]] & (SF_DONE | SF_WAIT))) cnt++;
        }
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) tile_counts[t] = red[0] + red[1] + red[2] + red[3];
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_srv_scan(int *a, const int *n_in, int *n_out) {
    __shared__ int tmp[1024];
    __shared__ int carry;
    const int ntiles = (*n_in + TILE - 1) / TILE;
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

__global__ __launch_bounds__(256) void k_srv_scatter(const int *list_in, const int *n_in, const int *flags,
                                                     const int *offs, int *list_out) {
    const int n = *n_in;
    const int nt = (n + TILE - 1) / TILE;
    __shared__ int wsum[4];
    __shared__ int run;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const int64_t base = (int64_t)t * TILE;
        if (threadIdx.x == 0) run = offs[t];
        __syncthreads();
        for (int r = 0; r < TILE; r += 256) {
            const int64_t i = base + r + threadIdx.x;
            int keep = 0, v = 0;
            if (i < n) {
                v = list_in[i];
                keep = !(flags[v] & (SF_DONE | SF_WAIT));
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
}

/* tracer admission: located seeds (input order) gathered into the pool in Z-order */
__global__ void k_srv_tracer_init(PtrSet p, int64_t base, int64_t n, const double *xyz, const int *perm,
                                  const int *elem, const int *left, const double *lt, int *aux, double *scalar,
                                  int *req_id, int *frame, int *flags, int r, int f0) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = base + k, src = perm[k];
    for (int c = 0; c < 3; c++) {
        p.X[c][i] = xyz[3 * src + c];
        p.V[c][i] = 0.0;
    }
    for (int c = 0; c < 6; c++) p.H[c][i] = 0.0;
    p.nhist[i] = 0;
    p.elem[i] = elem[src];
    p.left[i] = left[src];
    p.left_time[i] = lt[src];
    aux[i] = left[src] ? 0 : elem[src]; /* AuxElementIndex as fvc_seed_list sets it */
    scalar[i] = 0.0;
    req_id[i] = r;
    frame[i] = f0;
    flags[i] = 0;
}

/* FTLE admission: grid records (fvc_ftle_create's k_ftle_grid) */
__global__ void k_srv_ftle_init(PtrSet p, int64_t base, int64_t n, double mn0, double mn1, double mn2, double d0,
                                double d1, double d2, int r1, int r2, const int *celem, const int *cleft,
                                const double *clt, int *aux, double *scalar, int *req_id, int *frame, int *flags,
                                int r, int f0) {
#pragma clang fp contract(off)
    const int64_t P = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= n) return;
    const int64_t i = base + P;
    const int k = (int)(P % r2), j = (int)((P / r2) % r1), ii = (int)(P / ((int64_t)r2 * r1));
    p.X[0][i] = mn0 + ii * d0;
    p.X[1][i] = mn1 + j * d1;
    p.X[2][i] = mn2 + k * d2;
    for (int c = 0; c < 3; c++) p.V[c][i] = 0.0;
    p.elem[i] = celem[P];
    p.left[i] = cleft[P];
    p.left_time[i] = clt[P];
    aux[i] = 0;
    scalar[i] = 0.0;
    req_id[i] = r;
    frame[i] = f0;
    flags[i] = 0;
}

/* release-time GetVelocity over a pool range (tracers.c:202-215, ftle.c:80-93) */
template <int MODE>
__global__ __launch_bounds__(256) void k_srv_launch(DevMesh m, DevField fld, PtrSet p, int64_t base, int64_t n,
                                                    double t, int *errc, int64_t *where) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = base + k;
    if (p.left[i]) return;
    PState s;
    for (int c = 0; c < 3; c++) {
        s.X[c] = p.X[c][i];
        s.V[c] = 0.0;
    }
    s.elem = p.elem[i];
    s.left = 0;
    s.left_time = p.left_time[i];
    int err = 0;
    double v[3];
    get_velocity<MODE, 3>(m, fld, t, s, v, err);
    if (err) {
        if (atomicCAS(errc, 0, err) == 0) *where = i;
        return;
    }
    for (int c = 0; c < 3; c++) p.V[c][i] = v[c];
    p.elem[i] = s.elem;
    p.left[i] = s.left;
}

/* next FTLE frame: clear the barrier of a request's points */
__global__ void k_srv_release(int *flags, int *frame, int64_t base, int64_t n, int f) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    flags[base + k] = 0;
    frame[base + k] = f;
}

__global__ void k_srv_zero(int *a, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = 0;
}

int frame_of(double t, double t0, double fdt) {
#pragma clang fp contract(off)
    const double q = std::floor((t - t0) / fdt);
    if (!(std::fabs(q) < 1.0e9)) return INT32_MIN;
    int f = (int)q;
    while (t0 + f * fdt > t) f--;
    while (t0 + (f + 1) * fdt <= t) f++;
    return f;
}

unsigned grid256(int64_t n) { return (unsigned)((n + 255) / 256); }

PtrSet offset(PtrSet s, int64_t b) {
    for (int c = 0; c < 3; c++) {
        s.X[c] += b;
        s.V[c] += b;
    }
    for (int c = 0; c < 6; c++) s.H[c] += b;
    s.elem += b;
    s.left += b;
    s.left_time += b;
    s.nhist += b;
    return s;
}

struct HostReq {
    int64_t id = 0;
    fvc_request_desc d{};
    std::vector<double> xyz;
    int64_t n = 0; /* pool particles */
    int state = FVC_RS_QUEUED;
    int slot = -1;
    int64_t base = -1;
    int f0 = 0, frame = 0;
    double del[3] = {0, 0, 0};
    double t_submit = 0, t_admit = 0, t_done = 0;
    int64_t steps = 0;
    std::vector<fvc_point> recs;
    std::vector<double> wt, wot;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

struct fvc_server {
    fvc_ctx *ctx;
    fvc_server_params sp;
    int64_t cap;
    int64_t pool_bytes;
    fvc_particles *pool;
    void *buf;
    /* pool-wide per-slot arrays */
    double *Xn[3], *exit_time, *tc, *h, *wt, *wot;
    int *elem_new, *exited, *advected, *have, *req_id, *frame, *flags, *mpl, *perm, *lists[2][2];
    int4 *mp;
    /* small device tables */
    SrvReq *d_req;
    SrvSlot *d_slots;
    int *counts; /* [class][parity] */
    int *queue;  /* [class][NQ] */
    int *remaining;
    unsigned long long *req_steps;
    int *tiles;
    int ntiles;
    int cur; /* parity of the list the next launch reads */
    int grid[2];
    hipGraphExec_t gexec[2];
    hipGraph_t graph[2];
    /* host bookkeeping */
    std::map<int64_t, int64_t> free_ranges; /* base -> length */
    std::vector<int> free_slots;
    std::deque<int64_t> queued;
    std::unordered_map<int64_t, HostReq> reqs;
    std::vector<int64_t> slot_owner;
    int64_t next_id;
    bool pending;
    double t_origin;
    double st_wait, st_admit, st_frame, st_retire, st_ticks;
};

static int srv_err(fvc_ctx *c, int code, const char *msg) {
    c->last_error = msg;
    return code;
}

static SrvDev srv_dev(fvc_server *S, int k, int parity) {
    SrvDev d;
    d.p = ptrs(S->pool);
    for (int c = 0; c < 3; c++) d.Xn[c] = S->Xn[c];
    d.elem_new = S->elem_new;
    d.exited = S->exited;
    d.advected = S->advected;
    d.exit_time = S->exit_time;
    d.req = S->d_req;
    d.req_id = S->req_id;
    d.frame = S->frame;
    d.flags = S->flags;
    d.mpl = S->mpl;
    d.tc = S->tc;
    d.h = S->h;
    d.mp = S->mp;
    d.list_in = S->lists[k][parity];
    d.list_out = S->lists[k][parity ^ 1];
    d.n_in = S->counts + 2 * k + parity;
    d.n_out = S->counts + 2 * k + (parity ^ 1);
    d.queue = S->queue + NQ * k;
    d.remaining = S->remaining;
    d.req_steps = S->req_steps;
    d.slots = S->d_slots;
    d.nslots = S->sp.nslots;
    d.t0 = S->sp.frame_t0;
    d.fdt = S->sp.frame_dt;
    d.dt = S->sp.time_step;
    d.steps = S->ctx->d_steps;
    d.err = S->ctx->d_err;
    d.where = S->ctx->d_where;
    return d;
}

template <int MODE, int BS>
static void launch_serve(fvc_server *S, int k, int grid, const SrvDev &d, int budget) {
    if (k)
        k_serve<MODE, BS, 1><<<grid, BS, 0, S->ctx->stream>>>(S->ctx->mesh, d, budget);
    else
        k_serve<MODE, BS, 0><<<grid, BS, 0, S->ctx->stream>>>(S->ctx->mesh, d, budget);
}

template <int MODE, int BS, int GEN>
static const void *serve_fn() {
    return reinterpret_cast<const void *>(&k_serve<MODE, BS, GEN>);
}

static const void *serve_kernel(int prec, int bs, int k) {
    switch (prec) {
    case FVC_PREC_F64: return k ? serve_fn<P_F64, 256, 1>() : serve_fn<P_F64, 256, 0>();
    case FVC_PREC_BF16: return k ? serve_fn<P_BF16, 256, 1>() : serve_fn<P_BF16, 256, 0>();
    case FVC_PREC_F16: return k ? serve_fn<P_F16, 256, 1>() : serve_fn<P_F16, 256, 0>();
    default:
        if (bs == 64) return k ? serve_fn<P_F32, 64, 1>() : serve_fn<P_F32, 64, 0>();
        if (bs == 512) return k ? serve_fn<P_F32, 512, 1>() : serve_fn<P_F32, 512, 0>();
        return k ? serve_fn<P_F32, 256, 1>() : serve_fn<P_F32, 256, 0>();
    }
}

/* one tick's launches for list parity `par` */
static void enqueue_tick(fvc_server *S, int par) {
    fvc_ctx *c = S->ctx;
    const int budget = S->sp.steps_per_launch;
    k_srv_prep<<<1, 64, 0, c->stream>>>(S->queue);
    for (int k = 0; k < 2; k++) {
        const SrvDev d = srv_dev(S, k, par);
        const int bs = S->sp.block;
        switch (c->precision) {
        case FVC_PREC_F64: launch_serve<P_F64, 256>(S, k, S->grid[k], d, budget); break;
        case FVC_PREC_BF16: launch_serve<P_BF16, 256>(S, k, S->grid[k], d, budget); break;
        case FVC_PREC_F16: launch_serve<P_F16, 256>(S, k, S->grid[k], d, budget); break;
        default:
            if (bs == 64)
                launch_serve<P_F32, 64>(S, k, S->grid[k], d, budget);
            else if (bs == 512)
                launch_serve<P_F32, 512>(S, k, S->grid[k], d, budget);
            else
                launch_serve<P_F32, 256>(S, k, S->grid[k], d, budget);
        }
        int *tiles = S->tiles + k * (S->ntiles + 1);
        const int cg = std::min(S->ntiles, CMP_GRID);
        k_srv_count<<<cg, 256, 0, c->stream>>>(d.list_in, d.n_in, S->flags, tiles);
        k_srv_scan<<<1, 1024, 0, c->stream>>>(tiles, d.n_in, d.n_out);
        k_srv_scatter<<<cg, 256, 0, c->stream>>>(d.list_in, d.n_in, S->flags, tiles, d.list_out);
    }
}

static int pool_alloc(fvc_server *S, int64_t n, int64_t *base) {
    for (auto it = S->free_ranges.begin(); it != S->free_ranges.end(); ++it)
        if (it->second >= n) {
            *base = it->first;
            const int64_t rest = it->second - n, nb = it->first + n;
            S->free_ranges.erase(it);
            if (rest > 0) S->free_ranges[nb] = rest;
            return 1;
        }
    return 0;
}

static void pool_free(fvc_server *S, int64_t base, int64_t n) {
    if (n <= 0) return;
    auto it = S->free_ranges.emplace(base, n).first;
    auto nx = std::next(it);
    if (nx != S->free_ranges.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        S->free_ranges.erase(nx);
    }
    if (it != S->free_ranges.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            S->free_ranges.erase(it);
        }
    }
}

static DevField srv_field_host(fvc_server *S, int f) {
#pragma clang fp contract(off)
    const int lo = f % S->sp.nslots, hi = (f + 1) % S->sp.nslots;
    const double tmin = S->sp.frame_t0 + f * S->sp.frame_dt;
    return bind_field(S->ctx, lo, hi, tmin, tmin + S->sp.frame_dt);
}

static FtleDev ftle_view(fvc_server *S, const HostReq &R) {
    FtleDev d;
    for (int a = 0; a < 3; a++) {
        d.res[a] = R.d.ftle_res[a];
        d.del[a] = R.del[a];
        d.Xn[a] = S->Xn[a] + R.base;
        d.Vn[a] = S->Xn[a] + R.base; /* no extrapolation: velocities are not staged */
    }
    d.start_time = R.d.t_start;
    d.p = offset(ptrs(S->pool), R.base);
    d.have = S->have + R.base;
    d.wt = S->wt + R.base;
    d.wot = S->wot + R.base;
    d.elem_new = S->elem_new + R.base;
    d.exited = S->exited + R.base;
    d.advected = S->advected + R.base;
    d.exit_time = S->exit_time + R.base;
    return d;
}

static int launch_release_velocity(fvc_server *S, const HostReq &R) {
    fvc_ctx *c = S->ctx;
    const DevField fld = srv_field_host(S, R.f0);
    const PtrSet p = ptrs(S->pool);
    const unsigned g = grid256(R.n);
    switch (c->precision) {
    case FVC_PREC_F64: k_srv_launch<P_F64><<<g, 256, 0, c->stream>>>(c->mesh, fld, p, R.base, R.n, R.d.t_start, c->d_err, c->d_where); break;
    case FVC_PREC_BF16: k_srv_launch<P_BF16><<<g, 256, 0, c->stream>>>(c->mesh, fld, p, R.base, R.n, R.d.t_start, c->d_err, c->d_where); break;
    case FVC_PREC_F16: k_srv_launch<P_F16><<<g, 256, 0, c->stream>>>(c->mesh, fld, p, R.base, R.n, R.d.t_start, c->d_err, c->d_where); break;
    default: k_srv_launch<P_F32><<<g, 256, 0, c->stream>>>(c->mesh, fld, p, R.base, R.n, R.d.t_start, c->d_err, c->d_where); break;
    }
    FVC_HIP(hipGetLastError());
    return FVC_OK;
}

static int append_range(fvc_server *S, const HostReq &R) {
    fvc_ctx *c = S->ctx;
    const int k = R.d.int_type == FVC_INT_RK4 ? 0 : 1;
    int *cnt = S->counts + 2 * k + S->cur;
    k_srv_append<<<grid256(R.n), 256, 0, c->stream>>>(S->lists[k][S->cur], cnt, R.base, R.n);
    k_srv_add<<<1, 64, 0, c->stream>>>(cnt, R.n);
    FVC_HIP(hipGetLastError());
    return FVC_OK;
}

__global__ void k_srv_set(int *dst, int v) {
    if (threadIdx.x == 0) *dst = v;
}

static int set_remaining(fvc_server *S, int slot, int64_t n) {
    k_srv_set<<<1, 64, 0, S->ctx->stream>>>(S->remaining + slot, (int)n); /* ordered, no host wait */
    fvc_ctx *c = S->ctx;
    FVC_HIP(hipGetLastError());
    return FVC_OK;
}

/* sort keys of the pool slots: the element holding each slot's seed, seeds
 * outside the mesh last */
__global__ void k_srv_elem_keys(const int *__restrict__ elem, const int *__restrict__ perm, int64_t n,
                                uint64_t *__restrict__ key, int *__restrict__ val) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = perm[k];
    const int e = elem[p];
    key[k] = e < 0 ? (uint64_t)1 << 31 : (uint64_t)(uint32_t)e;
    val[k] = p;
}

/* A request's pool order: Z-order of its seeds, then (stable) the element
 * holding each, so that the lanes of a wave gather from neighbouring
 * elements (the mesh is numbered along its own curve when renumbered);
 * results return in input order through perm either way. */
static int order_by_element(fvc_ctx *c, const int *elem, int *perm, int64_t n) {
    if (n <= 1) return FVC_OK;
    uint64_t *k[2] = {nullptr, nullptr};
    int *v[2] = {nullptr, nullptr};
    FVC_HIP(hipMallocAsync((void **)&k[0], sizeof(uint64_t) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&k[1], sizeof(uint64_t) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&v[0], sizeof(int) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&v[1], sizeof(int) * n, c->stream));
    k_srv_elem_keys<<<grid256(n), 256, 0, c->stream>>>(elem, perm, n, k[0], v[0]);
    int which = 0;
    int rc = hipGetLastError() == hipSuccess ? FVC_OK : FVC_E_HIP;
    if (!rc) rc = radix_sort_pairs(c->stream, k[0], v[0], k[1], v[1], n, 32, &which);
    if (!rc && hipMemcpyAsync(perm, v[which], sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        rc = FVC_E_HIP;
    for (int a = 0; a < 2; a++) {
        (void)hipFreeAsync(k[a], c->stream);
        (void)hipFreeAsync(v[a], c->stream);
    }
    return rc;
}

static int admit(fvc_server *S, HostReq &R) {
    fvc_ctx *c = S->ctx;
    int rc;
    const int slot = R.slot;
    SrvReq dr{R.d.t_start, R.d.t_stop, R.d.kind, R.d.int_type};
    FVC_HIP(hipMemcpyAsync(S->d_req + slot, &dr, sizeof dr, hipMemcpyHostToDevice, c->stream));
    unsigned long long zero = 0;
    FVC_HIP(hipMemcpyAsync(S->req_steps + slot, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    if (R.d.kind == FVC_REQ_TRACER) {
        const int64_t n = R.n;
        double *dx = nullptr, *dlt = nullptr;
        int *de = nullptr, *dl = nullptr;
        int *dperm = S->perm + R.base; /* pool slot k holds seed perm[k] */
        FVC_HIP(hipMallocAsync((void **)&dx, sizeof(double) * 3 * n, c->stream));
        FVC_HIP(hipMallocAsync((void **)&dlt, sizeof(double) * n, c->stream));
        FVC_HIP(hipMallocAsync((void **)&de, sizeof(int) * n, c->stream));
        FVC_HIP(hipMallocAsync((void **)&dl, sizeof(int) * n, c->stream));
        FVC_HIP(hipMemcpyAsync(dx, R.xyz.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
        /* location chain in input order (ReadInTraceMesh), then Z-order for the pool */
        rc = launch_seed_chain(c, 0, n, dx, nullptr, nullptr, nullptr, S->sp.lsc, de, dl, dlt);
        double lo[3] = {c->mesh.bounds[0], c->mesh.bounds[2], c->mesh.bounds[4]};
        double hi[3] = {c->mesh.bounds[1], c->mesh.bounds[3], c->mesh.bounds[5]};
        if (!rc) rc = launch_zsort(c, dx, n, lo, hi, dperm);
        if (!rc && c->mesh.mesh_type == 1) rc = order_by_element(c, de, dperm, n);
        if (!rc) {
            k_srv_tracer_init<<<grid256(n), 256, 0, c->stream>>>(ptrs(S->pool), R.base, n, dx, dperm, de, dl, dlt,
                                                                S->pool->aux, S->pool->scalar, S->req_id, S->frame,
                                                                S->flags, slot, R.f0);
        }
        (void)hipFreeAsync(dx, c->stream);
        (void)hipFreeAsync(dlt, c->stream);
        (void)hipFreeAsync(de, c->stream);
        (void)hipFreeAsync(dl, c->stream);
        if (rc) return rc;
        std::vector<double>().swap(R.xyz);
    } else {
        const int32_t *res = R.d.ftle_res;
        const int64_t n = R.n;
        /* the chain writes into the staging arrays, as fvc_ftle_create does */
        int *ce = S->elem_new + R.base, *cl = S->exited + R.base;
        double *clt = S->exit_time + R.base;
        rc = launch_seed_chain(c, 2, n, nullptr, R.d.ftle_min, R.del, res, S->sp.lsc, ce, cl, clt);
        if (rc) return rc;
        k_srv_ftle_init<<<grid256(n), 256, 0, c->stream>>>(
            ptrs(S->pool), R.base, n, R.d.ftle_min[0], R.d.ftle_min[1], R.d.ftle_min[2], R.del[0], R.del[1],
            R.del[2], res[1], res[2], ce, cl, clt, S->pool->aux, S->pool->scalar, S->req_id, S->frame, S->flags, slot,
            R.f0);
        FVC_HIP(hipGetLastError());
        if ((rc = launch_ftle_mask_dev(c, ftle_view(S, R), n))) return rc;
    }
    FVC_HIP(hipGetLastError());
    if ((rc = launch_release_velocity(S, R))) return rc;
    if ((rc = set_remaining(S, slot, R.n))) return rc;
    return append_range(S, R);
}

/* finished tracer request -> records in input order (slot k holds seed perm[k]) */
__global__ void k_srv_records(PtrSet p, const int *aux, const double *scalar, const int *perm, int64_t base, int64_t n,
                              fvc_point *out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = base + k;
    fvc_point r;
    for (int c = 0; c < 3; c++) {
        r.X[c] = p.X[c][i];
        r.V[c] = p.V[c][i];
    }
    r.ElementIndex = p.elem[i];
    r.AuxElementIndex = aux[i];
    r.LeftDomain = p.left[i];
    r.pad_ = 0;
    r.LeftDomainTime = p.left_time[i];
    r.Scalar = scalar[i];
    out[perm[i]] = r;
}

static int copy_records(fvc_server *S, HostReq &R) {
    fvc_ctx *c = S->ctx;
    R.recs.resize(R.n);
    if (R.n == 0) return FVC_OK;
    fvc_point *d = nullptr;
    FVC_HIP(hipMallocAsync((void **)&d, sizeof(fvc_point) * R.n, c->stream));
    k_srv_records<<<grid256(R.n), 256, 0, c->stream>>>(ptrs(S->pool), S->pool->aux, S->pool->scalar, S->perm, R.base,
                                                        R.n, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(R.recs.data(), d, sizeof(fvc_point) * R.n, hipMemcpyDeviceToHost, c->stream);
    (void)hipFreeAsync(d, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->last_error = std::string("served tracer records: ") + hipGetErrorString(e);
        return FVC_E_HIP;
    }
    return FVC_OK;
}

static int finish_ftle_frame(fvc_server *S, HostReq &R, bool *complete) {
#pragma clang fp contract(off)
    fvc_ctx *c = S->ctx;
    const FtleDev v = ftle_view(S, R);
    const DevField fld = srv_field_host(S, R.frame);
    const double t1 = R.d.t_start > fld.tmin ? R.d.t_start : fld.tmin;
    int rc = launch_ftle_early(c, v, R.n, fld, t1, INT_RK4, S->sp.time_step);
    if (!rc) rc = launch_ftle_update_all(c, v, R.n, 0);
    if (rc) return rc;
    *complete = R.d.t_stop - fld.tmax < TINY; /* StopTime reached (ftle.c:352-377) */
    if (*complete) {
        if ((rc = launch_ftle_fill_dev(c, v, R.n, R.d.t_stop - R.d.t_start))) return rc;
        R.wt.resize(R.n);
        R.wot.resize(R.n);
        return copy_ftle_kji(c, S->wt + R.base, S->wot + R.base, R.d.ftle_res, R.wt.data(), R.wot.data());
    }
    R.frame++;
    k_srv_release<<<grid256(R.n), 256, 0, c->stream>>>(S->flags, S->frame, R.base, R.n, R.frame);
    FVC_HIP(hipGetLastError());
    if ((rc = set_remaining(S, R.slot, R.n))) return rc;
    return append_range(S, R);
}

static void retire(fvc_server *S, HostReq &R, int state) {
    pool_free(S, R.base, R.n);
    if (R.slot >= 0) {
        S->slot_owner[R.slot] = -1;
        S->free_slots.push_back(R.slot);
    }
    R.slot = -1;
    R.state = state;
    R.t_done = now_s() - S->t_origin;
}

extern "C" int fvc_server_create(fvc_ctx *c, const fvc_server_params *sp_in, fvc_server **out) {
    if (!c || !sp_in || !out) return FVC_E_INVALID;
    *out = nullptr;
    fvc_server_params sp = *sp_in;
    if (sp.hbm_fraction <= 0) sp.hbm_fraction = 0.5;
    if (sp.max_requests <= 0) sp.max_requests = 256;
    if (sp.steps_per_launch <= 0) sp.steps_per_launch = 32;
    if (sp.block <= 0) sp.block = 256;
    if (c->mesh.mesh_type != 1 || c->mesh.dims != 3)
        return srv_err(c, FVC_E_INVALID, "serving needs a 3-D unstructured mesh");
    if (!(sp.frame_dt > 0) || !(sp.time_step > 1.0e-10) || sp.nslots < 1 || sp.nslots > c->max_slices ||
        sp.hbm_fraction > 0.95 || sp.max_requests > 65536 || sp.steps_per_launch > (1 << 30) ||
        (c->precision == FVC_PREC_F32 ? (sp.block != 64 && sp.block != 256 && sp.block != 512) : sp.block != 256))
        return srv_err(c, FVC_E_INVALID, "server parameters out of range");
    for (int s = 0; s < sp.nslots; s++)
        if (!c->slices[s]) return srv_err(c, FVC_E_STATE, "serving needs slots 0..nslots-1 loaded");
    /* per pool slot: particle record SoA 128 B, serving state 52 B, FTLE staging 64 B, lists 16 B */
    const int64_t per = 128 + 52 + 64 + 16;
    int64_t cap = sp.pool_particles;
    if (cap <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return srv_err(c, FVC_E_HIP, "hipMemGetInfo failed");
        cap = (int64_t)((double)fr * sp.hbm_fraction) / per;
    }
    cap = std::min<int64_t>(cap, INT32_MAX / 2);
    if (cap < 1024) return srv_err(c, FVC_E_NOMEM, "no device memory for a request pool");
    fvc_server *S = new fvc_server();
    S->ctx = c;
    c->live_servers.insert(S);
    S->sp = sp;
    S->cap = cap;
    S->ntiles = (int)((cap + TILE - 1) / TILE);
    int rc = fvc_particles_create(c, cap, &S->pool);
    if (rc) {
        fvc_server_destroy(S);
        return rc;
    }
    const size_t n = (size_t)cap;
    const size_t sizes[] = {8 * n, 8 * n, 8 * n, 8 * n, 8 * n, 8 * n, 8 * n, 8 * n, /* Xn[3] exit_time tc h wt wot */
                            /* elem_new exited advected have req_id frame flags mpl perm */
                            4 * n, 4 * n, 4 * n, 4 * n, 4 * n, 4 * n, 4 * n, 4 * n, 4 * n,
                            16 * n, 4 * n, 4 * n, 4 * n, 4 * n,                     /* mp lists */
                            sizeof(SrvReq) * sp.max_requests, sizeof(SrvSlot) * sp.nslots, 64, 4 * 2 * NQ,
                            4 * (size_t)sp.max_requests, 8 * (size_t)sp.max_requests, 4 * 2 * ((size_t)S->ntiles + 1)};
    size_t bytes = 0;
    for (size_t s : sizes) bytes += (s + 255) / 256 * 256;
    if (hipMalloc(&S->buf, bytes) != hipSuccess) {
        fvc_server_destroy(S);
        return srv_err(c, FVC_E_NOMEM, "server pool");
    }
    char *q = (char *)S->buf;
    int k = 0;
    auto take = [&]() {
        void *r = q;
        q += (sizes[k++] + 255) / 256 * 256;
        return r;
    };
    for (int a = 0; a < 3; a++) S->Xn[a] = (double *)take();
    S->exit_time = (double *)take();
    S->tc = (double *)take();
    S->h = (double *)take();
    S->wt = (double *)take();
    S->wot = (double *)take();
    S->elem_new = (int *)take();
    S->exited = (int *)take();
    S->advected = (int *)take();
    S->have = (int *)take();
    S->req_id = (int *)take();
    S->frame = (int *)take();
    S->flags = (int *)take();
    S->mpl = (int *)take();
    S->perm = (int *)take();
    S->mp = (int4 *)take();
    S->lists[0][0] = (int *)take();
    S->lists[0][1] = (int *)take();
    S->lists[1][0] = (int *)take();
    S->lists[1][1] = (int *)take();
    S->d_req = (SrvReq *)take();
    S->d_slots = (SrvSlot *)take();
    S->counts = (int *)take();
    S->queue = (int *)take();
    S->remaining = (int *)take();
    S->req_steps = (unsigned long long *)take();
    S->tiles = (int *)take();
    S->pool_bytes = (int64_t)bytes + cap * 128;
    c->bytes_used += (int64_t)bytes;
    std::vector<SrvSlot> sl(sp.nslots);
    for (int s = 0; s < sp.nslots; s++) sl[s] = SrvSlot{c->slices[s], c->slice_scale[s], 0.0f};
    if (hipMemcpyAsync(S->d_slots, sl.data(), sizeof(SrvSlot) * sp.nslots, hipMemcpyHostToDevice, c->stream) ||
        hipMemsetAsync(S->flags, 0, 4 * n, c->stream) || hipMemsetAsync(S->counts, 0, 64, c->stream) ||
        hipMemsetAsync(S->exited, 0, 4 * n, c->stream) || hipMemsetAsync(S->advected, 0, 4 * n, c->stream) ||
        hipStreamSynchronize(c->stream)) {
        fvc_server_destroy(S);
        return srv_err(c, FVC_E_HIP, "server init");
    }
    for (int kk = 0; kk < 2; kk++) {
        int bpc = 0;
        const void *fn = serve_kernel(c->precision, sp.block, kk);
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, sp.block, 0) != hipSuccess || bpc <= 0)
            bpc = std::max(1, 8 / (sp.block / 64));
        S->grid[kk] = (cus * bpc + NQ - 1) / NQ * NQ;
    }
    S->free_ranges[0] = cap;
    S->slot_owner.assign(sp.max_requests, -1);
    for (int s = sp.max_requests - 1; s >= 0; s--) S->free_slots.push_back(s);
    S->next_id = 1;
    S->cur = 0;
    S->pending = false;
    S->t_origin = now_s();
    /* launch groups for both list parities */
    for (int par = 0; par < 2 && !sp.no_graph; par++) {
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            fvc_server_destroy(S);
            return srv_err(c, FVC_E_HIP, "server graph capture");
        }
        enqueue_tick(S, par);
        if (hipStreamEndCapture(c->stream, &S->graph[par]) != hipSuccess ||
            hipGraphInstantiate(&S->gexec[par], S->graph[par], nullptr, nullptr, 0) != hipSuccess) {
            fvc_server_destroy(S);
            return srv_err(c, FVC_E_HIP, "server graph instantiate");
        }
    }
    *out = S;
    return FVC_OK;
}

void fvc::server_release(fvc_server *S) {
    fvc_ctx *c = S->ctx;
    if (!c) return;
    hipStreamSynchronize(c->stream);
    for (int par = 0; par < 2; par++) {
        if (S->gexec[par]) hipGraphExecDestroy(S->gexec[par]);
        if (S->graph[par]) hipGraphDestroy(S->graph[par]);
        S->gexec[par] = nullptr;
        S->graph[par] = nullptr;
    }
    if (S->buf) hipFree(S->buf);
    S->buf = nullptr;
    fvc_particles_destroy(S->pool);
    S->pool = nullptr;
    c->live_servers.erase(S);
    S->ctx = nullptr;
}

extern "C" void fvc_server_destroy(fvc_server *S) {
    if (!S) return;
    server_release(S);
    delete S;
}

extern "C" int fvc_server_capacity(fvc_server *S, int64_t *particles, int64_t *bytes) {
    if (!S) return FVC_E_INVALID;
    if (particles) *particles = S->cap;
    if (bytes) *bytes = S->pool_bytes;
    return FVC_OK;
}

extern "C" double fvc_server_clock(fvc_server *S) { return S ? now_s() - S->t_origin : 0.0; }

extern "C" int fvc_server_submit(fvc_server *S, const fvc_request_desc *rq, int64_t *id) {
    if (!S || !S->ctx || !rq || !id) return FVC_E_INVALID;
    fvc_ctx *c = S->ctx;
    HostReq R;
    R.d = *rq;
    R.d.xyz = nullptr;
    if (!(rq->t_stop >= rq->t_start)) return srv_err(c, FVC_E_INVALID, "request: t_stop < t_start");
    R.f0 = frame_of(rq->t_start, S->sp.frame_t0, S->sp.frame_dt);
    if (R.f0 == INT32_MIN || R.f0 < 0) return srv_err(c, FVC_E_INVALID, "request starts before frame 0");
    if (rq->kind == FVC_REQ_TRACER) {
        if (rq->n < 0 || (rq->n > 0 && !rq->xyz) ||
            (rq->int_type != FVC_INT_RK4 && rq->int_type != FVC_INT_EULER && rq->int_type != FVC_INT_AB2 &&
             rq->int_type != FVC_INT_AB3))
            return srv_err(c, FVC_E_INVALID, "tracer request: seeds or integrator");
        R.n = rq->n;
        R.xyz.assign(rq->xyz, rq->xyz + 3 * rq->n);
    } else if (rq->kind == FVC_REQ_FTLE) {
        if (rq->int_type != FVC_INT_RK4) return srv_err(c, FVC_E_INVALID, "FTLE requests integrate with RK4");
        for (int a = 0; a < 3; a++)
            if (rq->ftle_res[a] < 2 || !(rq->ftle_max[a] > rq->ftle_min[a]))
                return srv_err(c, FVC_E_INVALID, "FTLE request: grid");
        R.n = (int64_t)rq->ftle_res[0] * rq->ftle_res[1] * rq->ftle_res[2];
        for (int a = 0; a < 3; a++) /* parameters.c:310-320 */
            R.del[a] = (rq->ftle_max[a] - rq->ftle_min[a]) / (rq->ftle_res[a] - 1);
    } else {
        return srv_err(c, FVC_E_INVALID, "request kind");
    }
    if (R.n > S->cap) return srv_err(c, FVC_E_ADMISSION, "request larger than the server's pool");
    R.id = S->next_id++;
    R.t_submit = now_s() - S->t_origin;
    R.frame = R.f0;
    *id = R.id;
    S->queued.push_back(R.id);
    S->reqs.emplace(R.id, std::move(R));
    return FVC_OK;
}

extern "C" int fvc_server_poll(fvc_server *S, int32_t *running, int32_t *queued) {
    if (!S || !S->ctx) return FVC_E_INVALID;
    fvc_ctx *c = S->ctx;
    int rc = FVC_OK;
    const int M = S->sp.max_requests;
    if (S->pending) {
        std::vector<int> rem(M);
        std::vector<unsigned long long> st(M);
        double t0 = now_s();
        FVC_HIP(hipMemcpyAsync(rem.data(), S->remaining, sizeof(int) * M, hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipMemcpyAsync(st.data(), S->req_steps, sizeof(unsigned long long) * M, hipMemcpyDeviceToHost,
                               c->stream));
        if ((rc = check_err(c, "fvc_server_poll"))) return rc;
        S->st_wait += now_s() - t0;
        S->pending = false;
        for (int s = 0; s < M; s++) {
            const int64_t id = S->slot_owner[s];
            if (id < 0) continue;
            HostReq &R = S->reqs[id];
            R.steps = (int64_t)st[s];
            if (rem[s] != 0) continue;
            t0 = now_s();
            if (R.d.kind == FVC_REQ_TRACER) {
                if ((rc = copy_records(S, R))) return rc;
                retire(S, R, FVC_RS_DONE);
                S->st_retire += now_s() - t0;
            } else {
                bool complete = false;
                if ((rc = finish_ftle_frame(S, R, &complete))) return rc;
                if (complete) retire(S, R, FVC_RS_DONE);
                S->st_frame += now_s() - t0;
            }
        }
    }
    const double ta = now_s();
    /* admission: arrival order, first fit (a request that does not fit waits;
     * smaller later ones may pass it unless it has waited more than 2 s) */
    const double tnow = now_s() - S->t_origin;
    for (auto it = S->queued.begin(); it != S->queued.end();) {
        HostReq &R = S->reqs[*it];
        if (S->free_slots.empty()) break;
        int64_t base = 0;
        if (!pool_alloc(S, R.n, &base)) {
            if (tnow - R.t_submit > 2.0) break;
            ++it;
            continue;
        }
        R.base = base;
        R.slot = S->free_slots.back();
        S->free_slots.pop_back();
        S->slot_owner[R.slot] = R.id;
        R.state = FVC_RS_RUNNING;
        R.t_admit = now_s() - S->t_origin;
        if (R.n == 0) {
            retire(S, R, FVC_RS_DONE);
        } else if ((rc = admit(S, R))) {
            retire(S, R, FVC_RS_FAILED);
            S->queued.erase(it); /* retired: a later poll must not admit it again */
            return rc;
        }
        it = S->queued.erase(it);
    }
    S->st_admit += now_s() - ta;
    int nrun = 0;
    for (int s = 0; s < M; s++) nrun += S->slot_owner[s] >= 0;
    if (nrun > 0) {
        if (S->sp.no_graph) {
            enqueue_tick(S, S->cur);
            FVC_HIP(hipGetLastError());
        } else {
            FVC_HIP(hipGraphLaunch(S->gexec[S->cur], c->stream));
        }
        S->cur ^= 1;
        S->pending = true;
        S->st_ticks += 1;
    }
    if (running) *running = nrun;
    if (queued) *queued = (int32_t)S->queued.size();
    return rc;
}

extern "C" int fvc_server_stats(fvc_server *S, double out[5]) {
    if (!S || !out) return FVC_E_INVALID;
    out[0] = S->st_wait;
    out[1] = S->st_admit;
    out[2] = S->st_frame;
    out[3] = S->st_retire;
    out[4] = S->st_ticks;
    return FVC_OK;
}

extern "C" int fvc_server_info(fvc_server *S, int64_t id, fvc_request_info *info) {
    if (!S || !S->ctx || !info) return FVC_E_INVALID;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return srv_err(S->ctx, FVC_E_INVALID, "unknown request");
    const HostReq &R = it->second;
    info->state = R.state;
    info->kind = R.d.kind;
    info->n = R.n;
    info->particle_steps = R.steps;
    info->t_submit = R.t_submit;
    info->t_admit = R.t_admit;
    info->t_done = R.t_done;
    return FVC_OK;
}

extern "C" int fvc_server_take(fvc_server *S, int64_t id, fvc_point *records, double *ftle_wt, double *ftle_wot) {
    if (!S || !S->ctx) return FVC_E_INVALID;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return srv_err(S->ctx, FVC_E_INVALID, "unknown request");
    HostReq &R = it->second;
    if (R.state != FVC_RS_DONE) return srv_err(S->ctx, FVC_E_STATE, "request not finished");
    if (R.d.kind == FVC_REQ_TRACER) {
        if (records && R.n) memcpy(records, R.recs.data(), sizeof(fvc_point) * R.n);
    } else {
        if (ftle_wt) memcpy(ftle_wt, R.wt.data(), sizeof(double) * R.n);
        if (ftle_wot) memcpy(ftle_wot, R.wot.data(), sizeof(double) * R.n);
    }
    S->reqs.erase(it);
    return FVC_OK;
}
