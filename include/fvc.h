/*
 * fvc.h - C ABI of the MI355X (gfx950) flowVC advection path.
 *
 * ompVC has no plugin layer: its hot path is the per-point loop that calls
 * Advect() (ompVC/integration.h:13) from ComputeFTLE (ompVC/ftle.c:109-173) and
 * Compute_Tracer (ompVC/tracers.c:217-229), with the mesh and the two loaded
 * velocity slices in globals (ompVC/globals.h:34-67) and errors reported by
 * FatalError -> exit(1) (ompVC/io.c:15-52).  This ABI replaces those loops one
 * level up, with batched calls over device-resident state.  Plain pointers and
 * sizes only; host arrays are lent for the duration of a call.
 *
 * Errors: every call returns FVC_OK (0) or a negative code; fvc_strerror()
 * names it.  A caller keeping ompVC's behaviour turns non-zero into exit(1).
 * Threading: one host thread per context; calls on one context are ordered on
 * the context's stream, and every call returns when its device work is done
 * (fvc_server_poll returns with one tick of the server enqueued), except the
 * early FTLE of an RK4 fvc_ftle_frame: it runs on the context's second
 * stream beside what follows, and every call that depends on it (the next
 * frame, fvc_ftle_output / _reset / _destroy, fvc_load_slice,
 * fvc_load_strain_slice, fvc_sync, fvc_destroy) waits for it first; an error
 * it raises is returned by a later call on the context.
 *
 * Lifetimes: particle sets, FTLE slides, batches and servers belong to the
 * context they were created on.  They may be destroyed in any order, before
 * or after the context: fvc_destroy releases the device memory of every
 * object still alive on the context and detaches it, and a detached handle
 * accepts only its *_destroy call (any other call returns FVC_E_INVALID).
 * Destroying a particle set first releases the batches created over it in
 * the same way.
 */
#ifndef FVC_H
#define FVC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FVC_OK 0
#define FVC_E_INVALID (-1)                /* bad argument */
#define FVC_E_NOMEM (-2)                  /* device or host allocation failed */
#define FVC_E_HIP (-3)                    /* HIP runtime error */
#define FVC_E_INDETERMINATE_NEIGHBOR (-4) /* ompVC/mesh.c:561, velocity.c:597 */
#define FVC_E_TLOC_RANGE (-5)             /* ompVC/velocity.c:277, :478 */
#define FVC_E_ELEM_NEG1 (-6)              /* ompVC/velocity.c:472, integration.c:335 */
#define FVC_E_WALK_LIMIT (-7)             /* walk exceeded the hop limit (no reference counterpart) */
#define FVC_E_EIG (-8)                    /* ompVC/mymath.c:107 "Too many iterations" */
#define FVC_E_ADMISSION (-9)              /* serving: request does not fit free device memory */
#define FVC_E_STATE (-10)                 /* call out of sequence */
#define FVC_E_LEFT (-11)                  /* ompVC/integration.c:326 integrate a point that left */
#define FVC_E_NODEVICE (-12)              /* no usable gfx950 device */
#define FVC_E_CET_OUTSIDE (-13)           /* ompVC/exposuretime.c:129 step start outside the mesh */

/* ---- enums -------------------------------------------------------------- */
enum fvc_mesh_type { FVC_CARTESIAN = 0, FVC_UNSTRUCTURED = 1 };
/* storage/arithmetic of the velocity field and element geometry */
enum fvc_precision {
    FVC_PREC_F64 = 0,  /* FP64 throughout, bit-identical to ompVC's arithmetic */
    FVC_PREC_F32 = 1,  /* FP32 geometry+velocity, FP64 state, exact FP64 walk decisions */
    FVC_PREC_BF16 = 2, /* as F32 with velocity stored bf16 x per-slice scale */
    FVC_PREC_F16 = 3   /* as F32 with velocity stored fp16 x per-slice scale */
};
/* Int_Type (ompVC/integration.c:32-37) plus multistep extensions */
enum fvc_int_type { FVC_INT_EULER = 0, FVC_INT_RK4 = 1, FVC_INT_RKF = 2, FVC_INT_AB2 = 3, FVC_INT_AB3 = 4 };
/* point location modes (ompVC/mesh.c:304-398, :496-611) */
enum fvc_locate_mode {
    FVC_LOCATE_WALK = 0,        /* Get_Element_Local_Search(X, guess): -1 when the walk leaves */
    FVC_LOCATE_GLOBAL = 1,      /* Get_Element_Global_Search(X): lowest-index containing element */
    FVC_LOCATE_WALK_GLOBAL = 2  /* walk, then global search when the walk leaves */
};

/* ---- records (byte layouts of ompVC/structs.h:16-59) -------------------- */
typedef struct fvc_point { /* LagrangianPoint: 80 bytes */
    double X[3];
    double V[3];
    int32_t ElementIndex;
    int32_t AuxElementIndex;
    int32_t LeftDomain;
    int32_t pad_;
    double LeftDomainTime;
    double Scalar;
} fvc_point;

typedef struct fvc_mesh_desc {
    int32_t dims;      /* Dimensions: 2 or 3 */
    int32_t mesh_type; /* Data_MeshType */
    double bounds[6];  /* Data_MeshBounds X/Y/Z Min,Max (ompVC/velocity.c:654-669) */
    /* Cartesian velocity grid, <prefix>_Cartesian.bin (ompVC/mesh.c:67-122) */
    double cart_min[3], cart_max[3];
    int32_t cart_res[3];
    /* unstructured: <prefix>_coordinates / _connectivity / _adjacency.bin
     * (ompVC/mesh.c:124-213); 2-D meshes carry 3 node ids and -1 per element */
    int64_t num_nodes;
    const double *coords;  /* num_nodes x 3 */
    int64_t num_elems;
    const int32_t *conn;   /* num_elems x 4 */
    const int32_t *adj;    /* num_elems x 4 */
} fvc_mesh_desc;

typedef struct fvc_params {
    int32_t precision;       /* enum fvc_precision */
    int32_t max_slices;      /* resident velocity slices (>= 2) */
    int32_t device;          /* HIP device ordinal */
    int32_t hash_cells;      /* spatial-hash cells along the longest axis; 0 = auto */
    int32_t walk_limit;      /* hop limit of one walk; 0 = 1<<20 */
} fvc_params;

typedef struct fvc_int_params {
    int32_t int_type;    /* enum fvc_int_type */
    int32_t extrapolate; /* Int_Extrapolate (ompVC/integration.c:27-43) */
    double time_step;    /* Int_TimeStep */
    int32_t block;       /* threads per block; 0 = default */
    int32_t steps_per_launch; /* 0 = whole interval in one launch */
    /* Trace_APCompute: integrate the strain rate of the loaded strain-rate
     * slices into each point's Scalar (ompVC/integration.c:355-360, 449-455);
     * fvc_advect and fvc_advect_times only */
    int32_t activation_potential;
    /* Trace_CETSubsteps when exposure times are accumulated (ompVC/
     * integration.c:427-448, exposuretime.c:120-178), 0 otherwise; needs
     * fvc_cet_reset first; unstructured meshes, fvc_advect and
     * fvc_advect_times only */
    int32_t exposure_substeps;
    /* FVC_INT_RKF (ompVC/integration.c:464-699): Int_MinTimeStep,
     * Int_MaxTimeStep, Int_Accuracy, each positive; fvc_advect,
     * fvc_advect_times and fvc_ftle_frame */
    double rkf_min_step, rkf_max_step, rkf_accuracy;
} fvc_int_params;

typedef struct fvc_ctx fvc_ctx;
typedef struct fvc_particles fvc_particles;
typedef struct fvc_ftle fvc_ftle;

/* ---- context ------------------------------------------------------------- */
/* Upload the mesh (LoadMeshData, ompVC/mesh.c:21-43) and build the device
 * spatial hash.  Replaces the mesh globals of ompVC/globals.h:38-47.
 * Device memory of an unstructured mesh: 32 B per node, 96 B per element
 * (connectivity, adjacency and the 64-byte FP32 frame the reduced-precision
 * walks test against), plus the hash; at most 2^26 elements and 2^27 nodes.
 * Host memory: device -> host copies of 4 MB or more (fvc_particles_get,
 * fvc_ftle_output) stage through up to 64 MB of pinned memory that the
 * context allocates on first use and holds until fvc_destroy. */
int fvc_create(const fvc_mesh_desc *mesh, const fvc_params *params, fvc_ctx **out);
/* Frees the context; objects still alive on it are released and detached
 * (see Lifetimes above). */
void fvc_destroy(fvc_ctx *ctx);
const char *fvc_strerror(int code);
/* Message of the last failure on this context (HIP error text, offending index). */
const char *fvc_last_error(const fvc_ctx *ctx);
int fvc_sync(fvc_ctx *ctx);
/* Device memory in use by the context, and free device memory (bytes). */
int fvc_memory(fvc_ctx *ctx, int64_t *used, int64_t *free_bytes);

/* Store one velocity slice, n x (u,v,w) FP64 as read from <prefix>_vel.<n>.bin
 * (ompVC/velocity.c:126-141, 184-211), into resident slot `slot`.  Replaces the
 * two-slot ring of LoadCartVelDataFrame / LoadUnstructVelDataFrame
 * (ompVC/velocity.c:61-234): callers name the two slots to interpolate between.
 * BF16/F16 precision stores the slice scaled by its max-abs value. */
int fvc_load_slice(fvc_ctx *ctx, int32_t slot, double timestamp, const double *uvw, int64_t n);
/* Strain-rate data of a frame for the activation potential
 * (LoadUnstructStrainRateDataFrame / LoadCartStrainRateDataFrame, ompVC/
 * strainrate.c:53-173): one value per element, or per Cartesian cell in k,j,i
 * order (i fastest; (XRes-1)(YRes-1)(ZRes-1 or 1) values), kept in FP64 in
 * `slot` beside that slot's velocity. */
int fvc_load_strain_slice(fvc_ctx *ctx, int32_t slot, double timestamp, const double *sr, int64_t n);

/* ---- particles (device-resident LagrangianPoint state) ------------------- */
int fvc_particles_create(fvc_ctx *ctx, int64_t capacity, fvc_particles **out);
void fvc_particles_destroy(fvc_particles *p);
/* copy n host records into positions [begin, begin+n) / back out.  On an
 * unstructured mesh a record not flagged LeftDomain whose element id lies
 * outside [-1, number of elements) is rejected (FVC_E_INVALID); on records
 * flagged LeftDomain such an id is stored as -1. */
int fvc_particles_set(fvc_particles *p, int64_t begin, const fvc_point *recs, int64_t n);
int fvc_particles_get(fvc_particles *p, int64_t begin, fvc_point *recs, int64_t n);
int64_t fvc_particles_capacity(const fvc_particles *p);
/* Particles that leave the domain keep their slots.  Once more than an
 * eighth of the lanes of an fvc_advect call over a range (4096 particles or
 * more) found their particle gone, fvc_advect lists that range's particles
 * still to integrate and later calls on the same range launch over the list
 * only.  *listed receives the list's length, or -1 when the next call
 * launches its whole range.  Writing particle state (fvc_particles_set,
 * fvc_particles_copy) drops the list.  No counterpart in the reference, whose
 * loops test every point (ompVC/tracers.c:109, 225). */
int fvc_particles_active(const fvc_particles *p, int64_t *listed);
/* device-to-device copy of n records (a release re-launched from resident
 * initial conditions, as CreateNewTraceLaunch copies the .IC file,
 * ompVC/tracers.c:1579-1632) */
int fvc_particles_copy(fvc_particles *dst, int64_t dst_begin, const fvc_particles *src, int64_t src_begin, int64_t n);
/* positions only, n x 3 doubles (output frames, ompVC/tracers.c:243-252) */
int fvc_particles_get_xyz(fvc_particles *p, int64_t begin, double *xyz, int32_t *left, int64_t n);

/* GetVelocity at launch for points with !LeftDomain (ompVC/tracers.c:202-215,
 * ftle.c:80-93): sets V and, as the reference's side effects, LeftDomain and
 * ElementIndex.  Slices `lo`,`hi` hold the data at tmin,tmax. */
int fvc_launch_velocity(fvc_ctx *ctx, fvc_particles *p, int64_t begin, int64_t count, double t,
                        int32_t lo, int32_t hi, double tmin, double tmax);

/* Advect(&pt, t1, t2) (ompVC/integration.c:22-45) for every point of
 * [begin, begin+count) with LeftDomainTime > -TINY && (!LeftDomain ||
 * extrapolate), the filter of ompVC/tracers.c:224 and ftle.c:113.  Slices lo,hi
 * hold the loaded interval [tmin, tmax] (main.c:188-195). */
int fvc_advect(fvc_ctx *ctx, fvc_particles *p, int64_t begin, int64_t count, double t1, double t2,
               int32_t lo, int32_t hi, double tmin, double tmax, const fvc_int_params *ip);

/* Advect(&pt, t1_i, t2_i) with a segment of its own for every point i of
 * [begin, begin+count): t12[2q], t12[2q+1] for point begin+q; a NaN t1 skips
 * the point.  The staggered release integrates each released point this way
 * between output times (ompVC/tracers.c:336-397).  Same filter as fvc_advect;
 * RK4 or Euler; segments must lie in [tmin, tmax]. */
int fvc_advect_times(fvc_ctx *ctx, fvc_particles *p, int64_t begin, int64_t count, const double *t12,
                     int32_t lo, int32_t hi, double tmin, double tmax, const fvc_int_params *ip);
/* Points advanced by the last fvc_advect (particle-steps, the step in which a
 * point exits included, floating-point micro-steps below 1e-12*h excluded). */
int fvc_last_step_count(fvc_ctx *ctx, int64_t *particle_steps);

/* ---- point location -------------------------------------------------------- */
/* n query points xyz (n x 3, host), guesses (walk modes), results elem (host).
 * Element indices are those ompVC returns for the same (X, guess). */
int fvc_locate(fvc_ctx *ctx, const double *xyz, const int32_t *guess, int64_t n, int32_t mode, int32_t *elem);
/* fvc_locate without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the query points,
 * guesses and results are staged in a fixed device array of the library
 * (one per GPU, held until the process ends) instead of memory allocated per
 * call.  Same arguments and results as fvc_locate; at most
 * FVC_NOALLOC_MAX_POINTS points per call (more: FVC_E_INVALID).  Calls on one
 * GPU share the array: calls that overlap on one GPU (several host threads,
 * contexts on the same device) race. */
#define FVC_NOALLOC_MAX_POINTS (1 << 18)
int fvc_locate_noalloc(fvc_ctx *ctx, const double *xyz, const int32_t *guess, int64_t n, int32_t mode,
                       int32_t *elem);
/* Seed location chains of ReadInTraceMesh (ompVC/tracers.c:1111-1251) for a
 * point list, and GenerateTracerMesh (tracers.c:1298-1532) for a grid: writes
 * X, LeftDomain, ElementIndex and LeftDomainTime of n records (grid: k,j,i
 * order).  Exact chain semantics; see fvc_seed_grid for the grid form. */
int fvc_seed_list(fvc_ctx *ctx, const double *xyz, int64_t n, int32_t local_search_checking, fvc_point *out);
int fvc_seed_grid(fvc_ctx *ctx, const double mn[3], const double del[3], const int32_t res[3],
                  int32_t local_search_checking, fvc_point *out);
/* Location chain of GenerateVelOutMesh (ompVC/velout.c:21-186) for the
 * velocity output grid: k,j,i order (x fastest), centre seed, guess reset to
 * the first found element of each x-row and each z-plane at their ends.
 * Writes X, LeftDomain and ElementIndex of res[0]*res[1]*res[2] records. */
int fvc_seed_velout_grid(fvc_ctx *ctx, const double mn[3], const double del[3], const int32_t res[3],
                         int32_t local_search_checking, fvc_point *out);

/* ---- FTLE (ompVC/ftle.c) ---------------------------------------------------- */
/* Grid generation, outside masks and the location chain of InitializeFTLEArray
 * (ompVC/ftle.c:383-609) for one output slide. */
int fvc_ftle_create(fvc_ctx *ctx, const double mn[3], const double mx[3], const int32_t res[3],
                    int32_t local_search_checking, double start_time, double stop_time, fvc_ftle **out);
void fvc_ftle_destroy(fvc_ftle *f);
/* One frame of ComputeFTLE for this slide (ompVC/ftle.c:68-214): launch when
 * not yet launched, advect, early FTLE for exits (ftle.c:788-1048) with the
 * reference's loop-order semantics, position update.  *done receives 0 when
 * the slide was not integrated in this interval (complete, or not launched
 * yet), 1 when it was, 2 when it was and reached its stop time.  With RK4 the
 * early FTLE reads a snapshot of the triggered points' stencils (240 B per
 * point in two buffers of up to 2^18 points, allocated at the first frame)
 * and overlaps the next frame's advection (see Threading). */
int fvc_ftle_frame(fvc_ftle *f, int32_t direction, int32_t lo, int32_t hi, double tmin, double tmax,
                   const fvc_int_params *ip, int32_t *done);
/* Missing values via GetFTLEForPoint (ftle.c:1050-1203) with integration time
 * int_t, then FTLEwT / FTLEwoT in k,j,i order (i fastest) as OutputFTLE writes. */
int fvc_ftle_output(fvc_ftle *f, double int_t, double *ftle_wt, double *ftle_wot);
int fvc_ftle_points(fvc_ftle *f, int64_t *n, int64_t *inside);
/* The slide's point records (n of them, grid index (i*res1 + j)*res2 + k), as
 * the FTLE temp files hold them (ftle.c:616-636). */
int fvc_ftle_records(fvc_ftle *f, fvc_point *recs);
/* Restore the slide to its state right after fvc_ftle_create (unlaunched,
 * located grid, masks), from a device-resident copy. */
int fvc_ftle_reset(fvc_ftle *f);
/* The slide's state right after creation, as InitializeFTLEArray stores it in
 * FTLE_ICFile (ompVC/ftle.c:611-625): per grid point, index
 * (i*res1 + j)*res2 + k, the point record, HaveFTLE and the FTLE values the
 * outside masks set.  Any output pointer may be NULL. */
int fvc_ftle_initial(fvc_ftle *f, fvc_point *recs, int32_t *have, double *ftle_wt, double *ftle_wot);
/* A slide from stored initial state instead of generating and locating the
 * grid: the FTLE_GenerateMesh = 0 branch of InitializeFTLEArray
 * (ompVC/ftle.c:626-660), which takes the resolution from the file and the
 * origin and spacing from the parameters.  Records are checked as
 * fvc_particles_set checks them. */
int fvc_ftle_create_from(fvc_ctx *ctx, const double mn[3], const double del[3], const int32_t res[3],
                         const fvc_point *recs, const int32_t *have, const double *ftle_wt, const double *ftle_wot,
                         double start_time, double stop_time, fvc_ftle **out);

/* ---- measurement ------------------------------------------------------------ */
/* When enabled, every launch of the integration kernels (fvc_advect's and the
 * FTLE frame's advection kernel) is bracketed by HIP events on the context's
 * stream, and so is each fvc_batch_run as a whole (all of its kernels);
 * fvc_kernel_time returns the summed milliseconds, launch (run) count and
 * particle-steps since the last reset. */
int fvc_set_timing(fvc_ctx *ctx, int32_t enable);
int fvc_kernel_time(fvc_ctx *ctx, double *ms, int64_t *launches, int64_t *particle_steps, int32_t reset);
/* Seed location chains: 0 (default) tabulates the chain where it can (walk
 * tables per multi-element point composed by pointer jumping, all points in
 * parallel), 1 always replays it point by point.  Both give the reference's
 * result; the replay is the cross-check of the tabulated form. */
int fvc_set_chain_mode(fvc_ctx *ctx, int32_t mode);
/* A plain device copy of `bytes` bytes (HBM traffic calibration for counters). */
int fvc_copy_probe(fvc_ctx *ctx, int64_t bytes);

/* ---- batched releases (no reference counterpart: Compute_Tracer,
 *      tracers.c:25-289, advances releases one after another) ---------------- */
/* Many releases of different sizes, start times, horizons and integrators in
 * one persistent kernel.  Release r covers particles [begin, begin+count) of
 * the particle set and integrates them from t_start to t_stop across data
 * frames f = [t0 + f dt, t0 + (f+1) dt] interpolating slots (f mod nslots) and
 * ((f+1) mod nslots) - per frame exactly as fvc_advect would (same stepping,
 * same arithmetic), so a release's result does not depend on the batch it
 * runs in.  Releases begin with the launch-time velocity lookup; points that
 * have left, or carry LeftDomainTime <= -1e-10, are not integrated.  Needs a
 * 3-D unstructured mesh, t_stop >= t_start, t_start >= t0; releases must not
 * overlap.  FP32 contexts accept every block size, others 256 only. */
typedef struct fvc_release {
    int64_t begin, count;
    double t_start, t_stop;
    int32_t int_type; /* enum fvc_int_type */
    int32_t pad_;
} fvc_release;

typedef struct fvc_batch_params {
    int32_t block;            /* threads per block: 64, 128, 256, 512 or 1024 */
    int32_t steps_per_launch; /* step budget of one particle per launch; 0 = unlimited */
    int32_t use_graph;        /* capture the launch sequence in a hipGraph and replay it */
    int32_t waves_per_cu;     /* persistent grid size: 0 = occupancy limit */
} fvc_batch_params;

typedef struct fvc_batch fvc_batch;
int fvc_batch_create(fvc_ctx *ctx, fvc_particles *p, const fvc_release *rel, int32_t nrel, double frame_t0,
                     double frame_dt, int32_t nslots, double time_step, fvc_batch **out);
void fvc_batch_destroy(fvc_batch *b);
/* Restore the particles of every release to their state at fvc_batch_create. */
int fvc_batch_reset(fvc_batch *b);
/* Run every release to its stop time (or exit).  Returns particle-steps and
 * the number of kernel launches. */
int fvc_batch_run(fvc_batch *b, const fvc_batch_params *bp, int64_t *particle_steps, int64_t *launches);
/* Autotuner: times fvc_batch_run over block sizes and step budgets (resetting
 * the batch between trials) and returns the fastest parameters. */
int fvc_batch_tune(fvc_batch *b, const int32_t *blocks, int32_t nblocks, const int32_t *budgets, int32_t nbudgets,
                   int32_t use_graph, fvc_batch_params *best, double *best_ms);

/* ---- request serving (no reference counterpart: ompVC runs one FTLE slide
 *      or tracer job per process, ftle.c:24-377 / tracers.c:25-289) -------- */
/* A server keeps the context's velocity slices resident and runs tracer and
 * FTLE requests with continuous batching: requests are admitted while others
 * run, as soon as their particles fit in a device pool sized from free HBM,
 * and the particles of every running request advance together in one
 * persistent kernel (the fvc_batch_* kernel extended with FTLE frames).  Each
 * request gives the result it gives when run alone:
 *   tracer: seeds located as fvc_seed_list does (chain in input order), then
 *           what fvc_batch_run does for one release; result = the records;
 *   FTLE:   the grid of fvc_ftle_create (start = t_start, stop = t_stop), the
 *           frames fvc_ftle_frame runs from the frame holding t_start, then
 *           fvc_ftle_output(t_stop - t_start); RK4 only.
 * Data frame f spans [t0 + f dt, t0 + (f+1) dt] over slots f mod nslots and
 * (f+1) mod nslots, as for fvc_batch_create.  One fvc_server_poll call runs
 * one scheduling tick: finish FTLE frames and retire completed requests,
 * admit queued ones that fit, run one launch of the persistent kernel. */
enum fvc_request_kind { FVC_REQ_TRACER = 0, FVC_REQ_FTLE = 1 };
enum fvc_request_state { FVC_RS_QUEUED = 0, FVC_RS_RUNNING = 1, FVC_RS_DONE = 2, FVC_RS_FAILED = 3 };

typedef struct fvc_server_params {
    double frame_t0, frame_dt; /* data frames as for fvc_batch_create */
    int32_t nslots;
    double time_step;
    int64_t pool_particles;    /* 0: from free HBM x hbm_fraction */
    double hbm_fraction;       /* share of free HBM the pool may take (0 = 0.5) */
    int32_t max_requests;      /* running at once (0 = 256) */
    int32_t block;             /* as fvc_batch_params */
    int32_t steps_per_launch;  /* step budget per particle per tick (0 = 32) */
    int32_t lsc;               /* LocalSearchChecking for seed location */
    int32_t no_graph;          /* 1: launch each tick directly instead of replaying a captured hipGraph */
} fvc_server_params;

typedef struct fvc_request_desc {
    int32_t kind;     /* enum fvc_request_kind */
    int32_t int_type; /* tracer: RK4, Euler, AB2, AB3; FTLE: RK4 */
    double t_start, t_stop;
    int64_t n;        /* tracer: seed count */
    const double *xyz;/* tracer seeds (n x 3), copied by fvc_server_submit */
    double ftle_min[3], ftle_max[3];
    int32_t ftle_res[3];
    int32_t pad_;
} fvc_request_desc;

typedef struct fvc_request_info {
    int32_t state;       /* enum fvc_request_state */
    int32_t kind;
    int64_t n;           /* tracer: records; FTLE: grid points */
    int64_t particle_steps;
    double t_submit, t_admit, t_done; /* seconds on the server's steady clock */
} fvc_request_info;

typedef struct fvc_server fvc_server;
int fvc_server_create(fvc_ctx *ctx, const fvc_server_params *sp, fvc_server **out);
void fvc_server_destroy(fvc_server *s);
/* Pool capacity in particles and its bytes. */
int fvc_server_capacity(fvc_server *s, int64_t *particles, int64_t *bytes);
/* Queue a request (inputs are copied); *id identifies it from then on. */
int fvc_server_submit(fvc_server *s, const fvc_request_desc *rq, int64_t *id);
/* One scheduling tick; returns counts of running and queued requests. */
int fvc_server_poll(fvc_server *s, int32_t *running, int32_t *queued);
/* Server clock (seconds), the clock of fvc_request_info's times. */
double fvc_server_clock(fvc_server *s);
int fvc_server_info(fvc_server *s, int64_t id, fvc_request_info *info);
/* Cumulative host seconds: [0] waiting for ticks to finish, [1] admitting,
 * [2] FTLE frame ends, [3] retiring tracer requests; [4] ticks launched. */
int fvc_server_stats(fvc_server *s, double out[5]);
/* Results of a finished request (then forgotten): tracer records (info.n of
 * them, input order) or FTLE values with and without the time factor (k, j,
 * i order, as fvc_ftle_output). */
int fvc_server_take(fvc_server *s, int64_t id, fvc_point *records, double *ftle_wt, double *ftle_wot);

/* ---- exposure time (ompVC/exposuretime.c) ---------------------------------- */
/* Zero the per-element exposure-time sums and encounter counts (main.c:84-90). */
int fvc_cet_reset(fvc_ctx *ctx);
/* One encounter in the element of every point of the range still in the
 * domain, at a release (tracers.c:180-188). */
int fvc_cet_launch(fvc_ctx *ctx, fvc_particles *p, int64_t begin, int64_t count);
/* Per-element exposure-time sums (CETsum) and encounter counts; either may be
 * NULL.  The sums are accumulated exactly in fixed point (2^-100 s) and
 * rounded once here, so they do not depend on the order of the steps. */
int fvc_cet_get(fvc_ctx *ctx, double *cetsum, int32_t *encounters);

/* ---- Z-order seed sort (cudaVC/ZOrderSort/main.cpp:86-101 ordering) ------- */
/* Sort n points by the interleaved-bit (x most significant) order of their
 * coordinates quantised to 21 bits on [lo, hi]; perm receives the stable order. */
int fvc_zsort(fvc_ctx *ctx, const double *xyz, int64_t n, const double lo[3], const double hi[3], int32_t *perm);

/* ---- Z-order mesh renumbering (cudaVC/ZOrderSort/main.cpp:103-299) -------- */
/* Renumbers a tetrahedral mesh (host arrays in the ompVC file layouts:
 * coords[3*num_nodes], conn and adj [4*num_elems], adj -1 at boundaries) so
 * that nodes and elements follow the Z-order of their coordinate ranks and
 * centroid ranks; replaces the reference program's Work() (main.cpp:103-259).
 * node_order[i] / elem_order[i] = old id of new node / element i (the
 * reference's sortedPointIDs / sortedCellIDs).  Outputs may be NULL; adj may
 * be NULL (then adj_out is not written).  Runs on the current device with no
 * context.  Returns FVC_E_INVALID for out-of-range node or neighbour ids. */
int fvc_mesh_zorder(int64_t num_nodes, const double *coords, int64_t num_elems, const int32_t *conn,
                    const int32_t *adj, int32_t *node_order, int32_t *elem_order, double *coords_out,
                    int32_t *conn_out, int32_t *adj_out);
/* out row i = in row order[i], rows of `width` doubles (node data such as a
 * velocity slice renumbered with node_order; main.cpp:262-291).  order must
 * be a permutation of 0..n-1 (checked). */
int fvc_gather_rows(int64_t n, int width, const int32_t *order, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
