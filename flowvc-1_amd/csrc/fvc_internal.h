/*
 * fvc_internal.h - shared between the C ABI (fvc_api.hip) and the kernel
 * launchers (fvc_kernels.hip).  Not part of the public interface.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <set>
#include <string>
#include <vector>

#include "fvc.h"
#include "fvc_device.h"

/* what the kernels report to the host after a launch */
struct fvc_status {
    unsigned long long steps[2]; /* particle-steps; idle lanes (fvc_advect) */
    int64_t where;               /* point of the first error */
    int err;                     /* first error code, 0 if none */
    int count;                   /* length of the last compacted list */
};

struct fvc_ctx {
    int device;
    int precision;
    hipStream_t stream;
    fvc::DevMesh mesh;
    /* device allocations */
    double4 *X64;
    float4 *ER; /* element frames, 4 per element */
    int4 *conn, *adj;
    int *cell_start, *cell_elems;
    int64_t ncells, nentries;
    /* resident velocity slices */
    int max_slices;
    std::vector<void *> slices;
    std::vector<float> slice_scale;
    std::vector<double> slice_time;
    int64_t nvel;       /* nodes per slice */
    size_t vel_stride;  /* bytes per node */
    /* resident strain-rate slices (activation potential), FP64, one per slot */
    std::vector<double *> sr_slices;
    int64_t nsr; /* values per strain-rate slice: elements, or Cartesian cells */
    /* exposure time: per element 128-bit fixed-point sums (low, high words) and encounters */
    unsigned long long *cet_lo, *cet_hi;
    int *cet_enc;
    /* errors and counters: d_err, d_where and d_steps point into d_status,
     * which check_err copies back whole into the pinned h_status with one
     * copy and one stream synchronisation */
    fvc_status *d_status, *h_status;
    int *d_err;
    int64_t *d_where;
    unsigned long long *d_steps;
    int64_t last_steps;
    int64_t last_idle; /* lanes of the last fvc_advect launch whose particle was not integrated */
    std::string last_error;
    int64_t bytes_used;
    /* scratch (grown on demand) */
    void *scratch;
    size_t scratch_bytes;
    /* pinned staging of large device -> host copies (copy_to_host), grown on
     * demand, with one event per chunk */
    char *stage;
    size_t stage_bytes;
    std::vector<hipEvent_t> stage_ev;
    /* side stream: an FTLE slide frame's early FTLE (from a snapshot of its
     * stencils) runs there beside the next frame's advection; side_done
     * follows the last launch on it, side_pending until the main stream or
     * the host has waited for it */
    hipStream_t side;
    hipEvent_t side_ready, side_done;
    int side_pending;
    /* kernel timing (HIP events on `stream`) */
    int timing;
    int chain_replay; /* seed chains: 1 = always the point-by-point replay */
    hipEvent_t ev0, ev1;
    double t_ms;
    int64_t t_launches, t_steps;
    /* objects created on this context: fvc_destroy releases their device
     * memory and detaches them (their handles then accept only *_destroy) */
    std::set<fvc_particles *> live_particles;
    std::set<fvc_ftle *> live_ftle;
    std::set<fvc_batch *> live_batches;
    std::set<fvc_server *> live_servers;
};

struct fvc_particles {
    fvc_ctx *ctx; /* null once detached by fvc_destroy */
    int64_t cap;
    void *buf;
    double *X[3];
    double *V[3];
    int *elem, *aux, *left;
    double *left_time, *scalar;
    double *H[6]; /* multistep history F1 (x,y,z), F2 (x,y,z) */
    int *nhist;
    /* particles still to integrate in [act_begin, act_begin + act_count), as
     * indices relative to act_begin (built by fvc_advect, dropped by any
     * write of particle state from outside the integrators) */
    int *act = nullptr;
    int64_t act_begin = -1, act_count = 0, act_n = 0;
    int act_extrap = 0;
};

struct fvc_ftle {
    fvc_ctx *ctx; /* null once detached by fvc_destroy */
    double mn[3], mx[3], del[3];
    int32_t res[3];
    int64_t n;
    double start_time, stop_time;
    int status; /* 0 unlaunched, 1 launched, 2 complete */
    fvc_particles *pts;
    /* per point */
    int *have;
    double *wt, *wot;
    /* staging of one frame */
    double *Xn[3];
    int *elem_new;
    int *exited;   /* 1 when the point left during this frame */
    int *advected; /* 1 when the point was integrated this frame */
    double *exit_time;
    double *Vn[3];
    int *active;    /* compacted indices of points to integrate (Z-ordered) */
    int *zperm;     /* grid points in Z-order */
    int64_t nactive;
    /* stencil snapshots of the early FTLE, used by alternate frames */
    void *snap[2];
    int64_t snap_cap;
    int snap_sel;
    void *buf;
    int64_t inside;
    /* state right after creation, for fvc_ftle_reset */
    fvc_particles *init_pts;
    int *init_have;
    double *init_w;
};

/* launchers (fvc_kernels.hip); all return FVC_OK or a negative code */
namespace fvc {
struct PtrSet {
    double *X[3];
    double *V[3];
    int *elem;
    int *left;
    double *left_time;
    double *H[6];
    int *nhist;
    double *scalar = nullptr; /* LagrangianPoint.Scalar (activation potential) */
};

inline PtrSet ptrs(fvc_particles *p) {
    PtrSet s;
    for (int c = 0; c < 3; c++) {
        s.X[c] = p->X[c];
        s.V[c] = p->V[c];
    }
    for (int c = 0; c < 6; c++) s.H[c] = p->H[c];
    s.elem = p->elem;
    s.left = p->left;
    s.left_time = p->left_time;
    s.nhist = p->nhist;
    return s; /* scalar stays null: only activation-potential runs carry it */
}

/* one FTLE grid's device arrays (grid index P = (i*res1 + j)*res2 + k) */
struct FtleDev {
    int res[3];
    double del[3];
    double start_time;
    PtrSet p;
    int *have;
    double *wt, *wot;
    double *Xn[3], *Vn[3];
    int *elem_new, *exited, *advected;
    double *exit_time;
};
FtleDev ftle_dev(fvc_ftle *f);

DevField bind_field(fvc_ctx *c, int lo, int hi, double tmin, double tmax);

int launch_velocity(fvc_ctx *c, PtrSet p, int64_t begin, int64_t count, double t, const DevField &f);
int launch_cet_encounters(fvc_ctx *c, PtrSet p, int64_t begin, int64_t count, int *enc);
int launch_advect_times(fvc_ctx *c, PtrSet p, int64_t begin, int64_t count, const double *dtt, const DevField &f,
                        const fvc_int_params &ip);
int launch_active_list(fvc_ctx *c, PtrSet p, int64_t begin, int64_t count, int extrap, int *list, int64_t *n);
int launch_advect(fvc_ctx *c, PtrSet p, const int *index, int64_t begin, int64_t count, double t1, double t2,
                  const DevField &f, const fvc_int_params &ip);
int launch_locate(fvc_ctx *c, const double *dxyz, const int *dguess, int64_t n, int mode, int *dout);
/* chain: 0 = point list (ReadInTraceMesh), 1 = tracer grid (GenerateTracerMesh), 2 = FTLE grid */
/* device sorting helpers (any stream) */
int radix_sort_pairs(hipStream_t s, uint64_t *k0, int *v0, uint64_t *k1, int *v1, int64_t n, int bits, int *which);
int scan_exclusive_small(hipStream_t s, int *a, int n);
int launch_seed_chain(fvc_ctx *c, int chain, int64_t n, const double *dxyz, const double mn[3], const double del[3],
                      const int32_t res[3], int lsc, int *d_elem, int *d_left, double *d_left_time);
int launch_ftle_frame(fvc_ftle *f, int lo, int hi, double tmin, double tmax, double t1, double t2,
                      const fvc_int_params &ip);
int launch_ftle_fill(fvc_ftle *f, double int_t);
int launch_ftle_mask(fvc_ftle *f);
/* the same on any FTLE grid view (enqueue only, no synchronisation) */
int launch_ftle_early(fvc_ctx *c, const FtleDev &d, int64_t n, const DevField &fld, double t1, int itype, double dt);
int launch_ftle_update_all(fvc_ctx *c, const FtleDev &d, int64_t n, int extrap);
int launch_ftle_fill_dev(fvc_ctx *c, const FtleDev &d, int64_t n, double int_t);
int launch_ftle_mask_dev(fvc_ctx *c, const FtleDev &d, int64_t n);
/* indices i (or map[i]) of the nonzero flags, in order */
int launch_compact(fvc_ctx *c, const int *flags, int64_t n, int *out_idx, int64_t *count, const int *map = nullptr);
/* Z-order (Morton, i most significant) permutation of the points of a grid */
int launch_grid_zorder(fvc_ctx *c, const int32_t res[3], int64_t n, int *dperm);
int launch_zsort(fvc_ctx *c, const double *dxyz, int64_t n, const double lo[3], const double hi[3], int *dperm);

int launch_copy_probe(fvc_ctx *c, void *dst, const void *src, int64_t bytes);
/* a grid's FTLE values (i-major on the device) to host arrays in k, j, i order */
int copy_ftle_kji(fvc_ctx *c, const double *d_wt, const double *d_wot, const int32_t res[3], double *wt, double *wot);

int check_err(fvc_ctx *c, const char *what);
/* order the main stream after the side stream's last launch / wait for it on the host */
int side_join(fvc_ctx *c);
int side_sync(fvc_ctx *c);
int copy_to_host(fvc_ctx *c, void *dst, const void *src, size_t bytes);
/* release an object's device memory and detach it from its context (the
 * handle stays valid for its *_destroy call only); no-ops when detached */
void particles_release(fvc_particles *p);
void ftle_release(fvc_ftle *f);
void batch_release(fvc_batch *b);
void server_release(fvc_server *s);
/* batches running over particle set p (released before p) */
void batches_release_on(fvc_ctx *c, fvc_particles *p);
void *scratch(fvc_ctx *c, size_t bytes);
/* bracket the hot kernel launch with events when timing is on */
inline void timing_begin(fvc_ctx *c) {
    if (c->timing) (void)hipEventRecord(c->ev0, c->stream);
}
inline void timing_end(fvc_ctx *c) {
    if (c->timing) (void)hipEventRecord(c->ev1, c->stream);
}
inline void timing_collect(fvc_ctx *c, int64_t steps) {
    if (!c->timing) return;
    float ms = 0.0f;
    if (hipEventSynchronize(c->ev1) == hipSuccess && hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) {
        c->t_ms += ms;
        c->t_launches++;
        c->t_steps += steps;
    }
}
} // namespace fvc

#define FVC_HIP(call)                                                                                                  \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) {                                                                                        \
            c->last_error = std::string(#call) + ": " + hipGetErrorString(e_);                                         \
            return e_ == hipErrorOutOfMemory ? FVC_E_NOMEM : FVC_E_HIP;                                                \
        }                                                                                                              \
    } while (0)
