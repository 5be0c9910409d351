/*
 * fvc_host.h - host-side C support for the flowVC advection path on MI355X:
 * ompVC-compatible parameter decks, binary file formats, and deterministic
 * synthetic inputs (double gyre, Kuhn-split pipe, Y-bifurcation).
 *
 * Nothing here touches the GPU.  The record layouts mirror ompVC's structs
 * (ompVC/structs.h:16-59) byte for byte so that .IC / slide / output files
 * are interchangeable with the reference program.
 */
#ifndef FVC_HOST_H
#define FVC_HOST_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVC_TINY 1.0e-10 /* ompVC/macros.h:17 */

/* ---- record layouts (ompVC/structs.h) ---------------------------------- */
typedef struct fvc_point_rec { /* LagrangianPoint, 80 bytes */
    double X[3];
    double V[3];
    int32_t ElementIndex;
    int32_t AuxElementIndex;
    int32_t LeftDomain;
    int32_t pad_;
    double LeftDomainTime;
    double Scalar;
} fvc_point_rec;

typedef struct fvc_ftle_rec { /* FTLEPoint, 104 bytes */
    fvc_point_rec Pt;
    int32_t HaveFTLE;
    int32_t pad_;
    double FTLEwT;
    double FTLEwoT;
} fvc_ftle_rec;

typedef struct fvc_cart { /* CartMesh, 96 bytes */
    double XMin, XMax; int32_t XRes; int32_t pad0_; double XDelta;
    double YMin, YMax; int32_t YRes; int32_t pad1_; double YDelta;
    double ZMin, ZMax; int32_t ZRes; int32_t pad2_; double ZDelta;
} fvc_cart;

/* ---- parameter deck (ompVC/parameters.c:20-138, 294-568) --------------- */
#define FVC_STR 200
typedef struct fvc_deck {
    char Path_Data[FVC_STR], Path_Output[FVC_STR];
    int32_t Dimensions, Data_MeshType;
    char Data_InFilePrefix[FVC_STR];
    int32_t Data_SuffixTMin, Data_SuffixTDelta, Data_TRes;
    double Data_TDelta, Data_TMin;
    int32_t Data_TPeriodic;
    double Bounds[6]; /* Data_MeshBounds X/Y/Z Min,Max */
    double Fluid_Density, Fluid_Viscosity;
    double Output_TStart; int32_t Output_TRes; double Output_TDelta;
    int32_t Int_Type; double Int_TimeStep, Int_Accuracy, Int_MinTimeStep, Int_MaxTimeStep;
    int32_t Int_TimeDirection, Int_NormalFlow; double Int_NormalFlowScaling; int32_t Int_Extrapolate;
    double Particle_Radius, Particle_Density; int32_t Particle_ICType;
    double Gravity[3];
    int32_t LocalSearchChecking, Num_threads;
    int32_t FTLE_Compute, FTLE_GenerateMesh; char FTLE_ICFile[FVC_STR];
    double FTLE_Bounds[6]; int32_t FTLE_Res[3];
    double FTLE_IntTLength; int32_t FTLE_ComputeVariation, FTLE_VariationOutFreq;
    char FTLE_OutFilePrefix[FVC_STR];
    int32_t Trace_Compute, Trace_ReleaseStrategy; double Trace_ReleaseTMax;
    int32_t Trace_GenerateMesh; char Trace_InFile[FVC_STR];
    int32_t Trace_MultipleInFiles, Trace_InFileFormat; char Trace_OutFilePrefix[FVC_STR];
    int32_t Trace_NumLaunchTimes; double Trace_LaunchTimeSpacing, Trace_IntTLength;
    int32_t Trace_AlwaysOutput;
    double Trace_Bounds[6]; int32_t Trace_Res[3];
    int32_t Trace_APCompute, Trace_CETCompute, Trace_CETAuxillaryMesh; char Trace_CETMeshPrefix[FVC_STR];
    int32_t Trace_CETSubsteps, Trace_RTCompute; char Trace_RTOutFilePrefix[FVC_STR];
    int32_t VelOut_Compute, VelOut_GenerateMesh; char VelOut_InFile[FVC_STR];
    int32_t VelOut_InFileFormat; char VelOut_FilePrefix[FVC_STR];
    double VelOut_Bounds[6]; int32_t VelOut_Res[3];
    /* derived (parameters.c:294-430) */
    double Data_TMax, Output_TEnd;
    int32_t Data_FirstFrame, Data_LastFrame;
    fvc_cart FTLE_Cart, Trace_Cart;
    int32_t legacy_96; /* deck lacked Num_threads (96 keys) */
} fvc_deck;

/* Parse a deck.  Returns 0, or a negative value and a message in err. */
int fvc_deck_read(const char *path, fvc_deck *d, char *err, int errlen);
/* Checks of parameters.c:432-568 that the advection path relies on. */
int fvc_deck_check(const fvc_deck *d, char *err, int errlen);
/* Derived values of parameters.c:294-385 (no thread-count logic). */
int fvc_deck_derive(fvc_deck *d, char *err, int errlen);
/* Write a deck in the 97-key order. */
int fvc_deck_write(const char *path, const fvc_deck *d);
void fvc_deck_defaults(fvc_deck *d);

/* ---- binary formats (Appendix of the reference formats) --------------- */
int fvc_write_cart_header(const char *path, const double mn[3], const double mx[3], const int32_t res[3]);
int fvc_read_cart_header(const char *path, double mn[3], double mx[3], int32_t res[3]);
int fvc_write_vel_slice(const char *path, double t, const double *uvw, int64_t n);
int fvc_read_vel_slice(const char *path, double *t, double *uvw, int64_t n);
int fvc_write_unstruct_mesh(const char *dir, const char *prefix, int64_t nn, const double *xyz,
                            int64_t ne, const int32_t *conn, const int32_t *adj);
int fvc_read_unstruct_counts(const char *dir, const char *prefix, int64_t *nn, int64_t *ne);
int fvc_read_unstruct_mesh(const char *dir, const char *prefix, double *xyz, int32_t *conn, int32_t *adj);

/* ---- PRNG: SplitMix64 seeding xoshiro256** ----------------------------- */
typedef struct fvc_rng { uint64_t s[4]; int has_spare; double spare; } fvc_rng;
void fvc_rng_seed(fvc_rng *r, uint64_t seed);
uint64_t fvc_rng_next(fvc_rng *r);
double fvc_rng_uniform(fvc_rng *r); /* [0,1) with 53 random bits */
double fvc_rng_normal(fvc_rng *r);  /* Box-Muller */

/* ---- synthetic inputs -------------------------------------------------- */
/* Double gyre velocity at (x,y,t) (u,v). */
void fvc_double_gyre(double A, double eps, double omega, double x, double y, double t, double *u, double *v);
/* Write <dir>/<prefix>_Cartesian.bin and <prefix>_vel.<suffix_min+s*suffix_delta>.bin,
 * s = 0..nt-1 at t = s*tdelta, on [0,2]x[0,1] with nx x ny nodes. */
int fvc_gen_double_gyre_files(const char *dir, const char *prefix, int nx, int ny, int nt, double tdelta,
                              double A, double eps, double omega, int suffix_min, int suffix_delta);
/* Fill a (nx*ny) x 3 slice in k,j,i order (i fastest). */
void fvc_gen_double_gyre_slice(int nx, int ny, double t, double A, double eps, double omega, double *uvw);

/* Voxelised tube meshes, Kuhn 6-tet split, face adjacency in the reference's
 * slot convention ([0] opposite node 1, [1] opposite 2, [2] opposite 3, [3] opposite 0).
 * shape 0 = straight pipe of radius R along z in [0,L] (voxels on [-R,R]^2 x [0,L]);
 * shape 1 = Y-bifurcation (parent radius R along -z, daughters of radius rd at +/- ang deg). */
typedef struct fvc_tube {
    int32_t shape;
    int32_t nx, ny, nz;       /* voxel counts */
    double lo[3], hi[3];      /* voxel box */
    double R, L;              /* parent radius, length */
    double rd, ang_deg, Ld;   /* daughters (shape 1) */
} fvc_tube;
void fvc_tube_pipe(fvc_tube *t, int nx, int ny, int nz, double R, double L);
void fvc_tube_ybif(fvc_tube *t, int nx, int ny, int nz, double R, double rd, double ang_deg, double L, double Ld);
int fvc_tube_count(const fvc_tube *t, int64_t *nn, int64_t *ne);
int fvc_tube_fill(const fvc_tube *t, double *xyz, int32_t *conn, int32_t *adj);
int fvc_tube_inside(const fvc_tube *t, double x, double y, double z); /* analytic body (voxel-centre test) */

/* Pulsatile field on the tube nodes: pipe = u_z 2(1-r^2)(1+0.5 sin 2 pi t) plus swirl
 * u_theta = 0.1 r (1-r^2) sin(pi z / L); Y-bifurcation = per-branch Poiseuille blended
 * at the junction, scaled by 1+0.6 sin 2 pi t.  Gaussian noise of sd sigma per
 * component per node, seeded by (seed, slice). */
void fvc_tube_velocity(const fvc_tube *t, int64_t nn, const double *xyz, double time, double sigma,
                       uint64_t seed, int32_t slice, double *uvw);
/* n seeds uniform in {r<R, z in [z0,z1]} (pipe) or inside the body (Y), seeded. */
int fvc_tube_seeds(const fvc_tube *t, int64_t n, double z0, double z1, uint64_t seed, double *xyz);

#ifdef __cplusplus
}
#endif
#endif
