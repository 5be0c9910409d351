/*
 * fvc_oracle.h - CPU FP64 restatement of ompVC's advection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU path is compared
 * against, and the timed CPU baseline ("port").  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * It restates, in its own code, the arithmetic of:
 *   ompVC/mesh.c:304-398 (global search), mesh.c:496-611 (local walk),
 *   ompVC/velocity.c:257-457 (Cartesian interpolation), velocity.c:459-621
 *   (unstructured interpolation), velocity.c:624-669 (domain tests),
 *   ompVC/integration.c:22-45 (Advect), :222-311 (Euler), :317-462 (RK4), :464-699 (RKF),
 *   ompVC/ftle.c:24-377 (per-frame FTLE), :383-744 (grid init + location chain),
 *   :788-1048 (early FTLE), :1050-1203 (flow-map gradient), :1244-1329 (output),
 *   ompVC/tracers.c:25-289 (tracer frames), :998-1577 (seed location chains),
 *   :1700-1816 (outputs), ompVC/main.c:170-227 (frame loop),
 *   ompVC/velocity.c:61-234 (two-slot slice ring).
 * Expression order follows the reference so FP64 results are bit-identical
 * (gcc, no FMA contraction on x86-64).  The largest eigenvalue uses a cyclic
 * Jacobi method instead of the reference's Householder+QL; it agrees to
 * rounding.
 */
#ifndef FVC_ORACLE_H
#define FVC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OC_TINY 1.0e-10

enum {
    OC_OK = 0,
    OC_E_INDETERMINATE = -4, /* mesh.c:561, velocity.c:597 */
    OC_E_TLOC = -5,          /* velocity.c:277, :478 */
    OC_E_ELEM_NEG1 = -6,     /* velocity.c:472, integration.c:335 */
    OC_E_LEFT = -11,         /* integration.c:326 "already left domain" */
    OC_E_EIG = -8,
    OC_E_IO = -12,
    OC_E_PARAM = -13,
    OC_E_NOMEM = -2,
    OC_E_BADMESH = -3, /* connectivity or adjacency index out of range */
};

typedef struct oc_point { /* ompVC LagrangianPoint, 80 bytes */
    double X[3];
    double V[3];
    int32_t ElementIndex;
    int32_t AuxElementIndex;
    int32_t LeftDomain;
    int32_t pad_;
    double LeftDomainTime;
    double Scalar;
} oc_point;

typedef struct oc_mesh {
    int32_t dims;      /* 2 or 3 */
    int32_t mesh_type; /* 0 Cartesian, 1 unstructured */
    double bounds[6];  /* Data_MeshBounds */
    /* Cartesian velocity grid */
    double cmin[3], cmax[3], cdel[3];
    int32_t cres[3];
    /* unstructured */
    int64_t nn, ne;
    double *xyz;   /* nn x 3 */
    int32_t *conn; /* ne x 4 */
    int32_t *adj;  /* ne x 4 */
    /* two loaded slices: slot 0 at tmin, slot 1 at tmax; per node (u,v,w) */
    double *vel[2]; /* n x 3 each; Cartesian nodes in k,j,i order (i fastest) */
    double tmin, tmax;
    int64_t nvel;
} oc_mesh;

typedef struct oc_int {
    int32_t int_type;    /* 0 Euler, 1 RK4, 2 RKF */
    double time_step;    /* Int_TimeStep */
    int32_t extrapolate; /* Int_Extrapolate */
    double min_step, max_step, accuracy; /* RKF: Int_MinTimeStep, Int_MaxTimeStep, Int_Accuracy */
} oc_int;

/* construction (arrays are copied) */
oc_mesh *oc_mesh_cartesian(int dims, const double bounds[6], const double cmin[3], const double cmax[3],
                           const int32_t cres[3]);
oc_mesh *oc_mesh_unstructured(int dims, const double bounds[6], int64_t nn, const double *xyz, int64_t ne,
                              const int32_t *conn, const int32_t *adj);
void oc_mesh_free(oc_mesh *m);
/* copy a slice into slot (0 = earlier time, 1 = later time) */
int oc_mesh_set_slot(oc_mesh *m, int slot, const double *uvw);
/* ring semantics of velocity.c:61-234: swap then fill the incoming slot */
int oc_mesh_push_frame(oc_mesh *m, int direction, const double *uvw);
void oc_mesh_set_interval(oc_mesh *m, double tmin, double tmax);

/* primitives */
int oc_test_outside_domain(const oc_mesh *m, const double X[3]);
int oc_test_outside_cart_vel(const oc_mesh *m, const double X[3]);
int oc_local_search(const oc_mesh *m, const double X[3], int guess, int *err);
int oc_global_search(const oc_mesh *m, const double X[3]);
int oc_velocity(const oc_mesh *m, double tq, oc_point *p, double out[3]); /* returns error code */
double oc_advect(const oc_mesh *m, const oc_int *ip, oc_point *p, double t1, double t2, int *err);
double oc_advect_counted(const oc_mesh *m, const oc_int *ip, oc_point *p, double t1, double t2, int *err,
                         int64_t *steps);
double oc_max_eig_sym3(const double a[3][3], int *err);

/* Adams-Bashforth 2/3 with history carried across calls (no reference counterpart). */
typedef struct oc_hist {
    double F1[3], F2[3];
    int32_t nhist;
    int32_t pad_;
} oc_hist;
double oc_ab_counted(const oc_mesh *m, oc_point *p, oc_hist *h, double t0, double t1, double dt, int order, int *err,
                     int64_t *ns);
int oc_ab_batch(const oc_mesh *m, int64_t n, oc_point *pts, oc_hist *hist, double t1, double t2, double dt, int order,
                int nthreads, int64_t *steps);

/* batches (OpenMP over points; nthreads <= 0 leaves the runtime default) */
int oc_locate_local_batch(const oc_mesh *m, int64_t n, const double *xyz, const int32_t *guess, int32_t *out);
int oc_locate_global_batch(const oc_mesh *m, int64_t n, const double *xyz, int32_t *out, int nthreads);
int oc_advect_batch(const oc_mesh *m, const oc_int *ip, int64_t n, oc_point *pts, double t1, double t2,
                    int nthreads, int64_t chunk, int64_t *steps);
int oc_velocity_batch(const oc_mesh *m, double tq, int64_t n, oc_point *pts, double *uvw);
int oc_eig_batch(int64_t n, const double *mats, double *lam);

/* FTLE grid state for one output slide (ftle.c).  Points are stored i-major,
 * then j, then k (index (i*YRes + j)*ZRes + k), as the reference's IC file. */
typedef struct oc_ftle_pt { /* FTLEPoint, 104 bytes */
    oc_point Pt;
    int32_t HaveFTLE;
    int32_t pad_;
    double FTLEwT;
    double FTLEwoT;
} oc_ftle_pt;

typedef struct oc_ftle {
    double min[3], max[3], del[3];
    int32_t res[3];
    int64_t n;
    oc_ftle_pt *pts;
    double start_time, stop_time;
    int32_t status; /* 0 unlaunched, 1 launched, 2 complete */
    int32_t local_search_checking;
    int64_t n_integrated; /* particle-advect calls of the last frame */
} oc_ftle;

oc_ftle *oc_ftle_create(const oc_mesh *m, const double min[3], const double max[3], const int32_t res[3],
                        int local_search_checking, int *err);
void oc_ftle_free(oc_ftle *f);
const oc_ftle_pt *oc_ftle_points(const oc_ftle *f); /* n records, index (i*res1 + j)*res2 + k */
/* One frame of ComputeFTLE for this slide (ftle.c:68-214) */
int oc_ftle_frame(const oc_mesh *m, const oc_int *ip, oc_ftle *f, int direction, double loaded_tmin,
                  double loaded_tmax);
/* Fill missing values (ftle.c:1304-1315) with integration time int_t */
int oc_ftle_fill(oc_ftle *f, int dims, double int_t);
double oc_ftle_time(const oc_ftle *f); /* seconds spent in oc_ftle_frame */

/* Tracer seeds: location chain of ReadInTraceMesh (tracers.c:1111-1251), points in
 * file order; pts receive X, LeftDomain, ElementIndex, LeftDomainTime. */
int oc_seed_list(const oc_mesh *m, int64_t n, const double *xyz, int local_search_checking, oc_point *pts);
/* GenerateTracerMesh chain (tracers.c:1298-1532), k,j,i order, x fastest. */
int oc_seed_grid(const oc_mesh *m, const double min[3], const double del[3], const int32_t res[3],
                 int local_search_checking, oc_point *pts);

/* Z-order renumbering of a tetrahedral mesh (cudaVC/ZOrderSort/main.cpp:103-259):
 * node_order[i] / elem_order[i] = old id of new node / element i; any of the
 * renumbered outputs may be NULL. */
int oc_mesh_zorder(int64_t nn, const double *xyz, int64_t ne, const int32_t *conn, const int32_t *adj,
                   int32_t *node_order, int32_t *elem_order, double *xyz_out, int32_t *conn_out, int32_t *adj_out);

/* Deck-driven run writing ompVC-format outputs.  nproc emulates ompVC's loop
 * choice (parameters.c:398-421); <= 0 uses the host count.  Returns 0 or an error. */
int oc_run_deck(const char *deck_path, int nproc, char *msg, int msglen);

#ifdef __cplusplus
}
#endif
#endif
