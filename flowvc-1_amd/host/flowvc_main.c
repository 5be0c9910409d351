/*
 * flowvc_main.c - ompVC-compatible driver over the MI355X C ABI (include/fvc.h).
 *
 *   flowvc <deck.in> [--precision f64|f32|bf16|f16] [--nproc N] [--quiet]
 *
 * Reads the same 97-entry deck and binary mesh/velocity files as ompflowVC and
 * writes the same FTLE (<prefix>.<n>.bin, _noT, _Cartesian.bin, .info) and
 * tracer (<prefix>.<k>.bin, <prefix>.IC) files.  The frame loop follows
 * ompVC/main.c:170-227; FTLE slides follow ftle.c:24-377 and tracer releases
 * tracers.c:25-289.  Velocity slices stay resident on the device: each data
 * file is uploaded once into its own slot (the reference re-reads two slices
 * per frame, velocity.c:61-234).  --nproc emulates the host processor count
 * that selects ompVC's tracer output loop (parameters.c:411-421).
 * Staggered tracer release follows tracers.c:293-1000 and velocity output
 * velout.c:21-435; the activation potential (Trace_APCompute) integrates
 * the strain-rate data along tracer paths (strainrate.c, integration.c);
 * exposure times (exposuretime.c) and residence times (residencetime.c) are
 * written per element and per seed.  RKF (Int_Type 2) and normal flow
 * (Int_NormalFlow) run as the reference does; inertial particles
 * (Particle_Radius > 0) stop with an error.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "fvc.h"
#include "fvc_host.h"

#define TINY FVC_TINY

static void fatal(const char *fmt, const char *a) {
    fprintf(stderr, "Error: ");
    fprintf(stderr, fmt, a);
    fprintf(stderr, "\n");
    exit(1);
}

static void check(int rc, fvc_ctx *c, const char *what) {
    if (rc) {
        fprintf(stderr, "Error: %s: %s (%s)\n", what, fvc_strerror(rc), c ? fvc_last_error(c) : "");
        exit(1);
    }
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

typedef struct {
    fvc_ctx *c;
    const fvc_deck *d;
    int64_t n;
    int nslots;
    int *slot_file; /* file index held by each slot, -1 empty */
    double *buf;
    int next_victim;
    double io_s;
    int64_t nsr;   /* strain-rate values per frame when the activation potential is integrated, else 0 */
    double *srbuf;
    int32_t nwall;  /* Int_NormalFlow: wall nodes whose velocity every slice replaces */
    int32_t *wall_ids;
    double *wall_v; /* their velocity, 3 per node */
} slices_t;

static int file_of(const fvc_deck *d, int frame) {
    int mod = (int)fmod(frame, d->Data_TRes - 1); /* velocity.c:77-80 */
    if (mod < 0) mod += d->Data_TRes - 1;
    return mod;
}

/* slot holding `frame`, loading the file when absent; `keep` is not evicted */
static int slot_for(slices_t *s, int frame, int keep) {
    int f = file_of(s->d, frame);
    for (int q = 0; q < s->nslots; q++)
        if (s->slot_file[q] == f) return q;
    int q = -1;
    for (int k = 0; k < s->nslots; k++)
        if (s->slot_file[k] < 0) {
            q = k;
            break;
        }
    if (q < 0) {
        q = s->next_victim;
        if (q == keep) q = (q + 1) % s->nslots;
        s->next_victim = (q + 1) % s->nslots;
    }
    char path[3 * FVC_STR];
    snprintf(path, sizeof path, "%s%s_vel.%d.bin", s->d->Path_Data, s->d->Data_InFilePrefix,
             s->d->Data_SuffixTMin + f * s->d->Data_SuffixTDelta);
    double t0 = now_s(), ts;
    if (fvc_read_vel_slice(path, &ts, s->buf, s->n)) fatal("Could not load velocity data from %s", path);
    for (int32_t i = 0; i < s->nwall; i++) /* inward flow on the walls (velocity.c:213-228) */
        for (int a = 0; a < 3; a++) s->buf[3 * (int64_t)s->wall_ids[i] + a] = s->wall_v[3 * i + a];
    check(fvc_load_slice(s->c, q, ts, s->buf, s->n), s->c, "fvc_load_slice");
    if (s->nsr) { /* LoadUnstructStrainRateDataFrame / LoadCartStrainRateDataFrame (strainrate.c:53-173) */
        snprintf(path, sizeof path, "%s%s_strain-rate.%d.bin", s->d->Path_Data, s->d->Data_InFilePrefix,
                 s->d->Data_SuffixTMin + f * s->d->Data_SuffixTDelta);
        FILE *fs = fopen(path, "rb");
        double st;
        if (!fs) fatal("Could not open file %s", path);
        if (fread(&st, 8, 1, fs) != 1) fatal("Could not read time stamp from file %s", path);
        if (fread(s->srbuf, 8, (size_t)s->nsr, fs) != (size_t)s->nsr) fatal("Could not read strain rate data from %s", path);
        fclose(fs);
        check(fvc_load_strain_slice(s->c, q, st, s->srbuf, s->nsr), s->c, "fvc_load_strain_slice");
    }
    s->io_s += now_s() - t0;
    s->slot_file[q] = f;
    return q;
}

static void write_info(const fvc_deck *d) {
    char path[3 * FVC_STR];
    snprintf(path, sizeof path, "%s%s.info", d->Path_Output, d->FTLE_OutFilePrefix);
    FILE *o = fopen(path, "w");
    if (!o) fatal("Could not open file %s", path);
    /* ftle.c:1253-1270 */
    fprintf(o, "Int_Type\t\t=\t%10d\n", d->Int_Type);
    fprintf(o, "Int_TimeStep\t\t=\t%10g\n", d->Int_TimeStep);
    fprintf(o, "Int_Accuracy\t\t=\t%10g\n", d->Int_Accuracy);
    fprintf(o, "Int_MinTimeStep\t\t=\t%10g\n", d->Int_MinTimeStep);
    fprintf(o, "Int_MaxTimeStep\t\t=\t%10g\n", d->Int_MaxTimeStep);
    fprintf(o, "Int_TimeDirection\t=\t%10d\n", d->Int_TimeDirection);
    fprintf(o, "Int_NormalFlow\t\t=\t%10d\n", d->Int_NormalFlow);
    fprintf(o, "Int_NormalFlowScaling\t=\t%10g\n\n", d->Int_NormalFlowScaling);
    fprintf(o, "LocalSearchChecking\t=\t%10d\n\n", d->LocalSearchChecking);
    fprintf(o, "FTLE_GenerateMesh\t=\t%10d\n", d->FTLE_GenerateMesh);
    fprintf(o, "FTLE_ICFile\t\t=\t%s\n", d->FTLE_ICFile);
    fprintf(o, "FTLE_IntTLength\t\t=\t%10g\n", d->FTLE_IntTLength);
    fclose(o);
    snprintf(path, sizeof path, "%s%s_Cartesian.bin", d->Path_Output, d->FTLE_OutFilePrefix);
    double mn[3] = {d->FTLE_Cart.XMin, d->FTLE_Cart.YMin, d->FTLE_Cart.ZMin};
    double mx[3] = {d->FTLE_Cart.XMax, d->FTLE_Cart.YMax, d->FTLE_Cart.ZMax};
    int32_t res[3] = {d->FTLE_Cart.XRes, d->FTLE_Cart.YRes, d->FTLE_Cart.ZRes};
    if (fvc_write_cart_header(path, mn, mx, res)) fatal("Could not open file %s", path);
}

static void write_ftle(const fvc_deck *d, int outnum, double start, const double *wt, const double *wot, int64_t n) {
    char path[3 * FVC_STR];
    for (int which = 0; which < 2; which++) {
        snprintf(path, sizeof path, "%s%s%s.%d.bin", d->Path_Output, d->FTLE_OutFilePrefix, which ? "_noT" : "",
                 outnum);
        FILE *o = fopen(path, "wb");
        if (!o) fatal("Could not open file %s", path);
        fwrite(&start, 8, 1, o);
        fwrite(which ? wot : wt, 8, (size_t)n, o);
        fclose(o);
    }
}

/* One seed file of ReadInTraceMesh (tracers.c:1010-1104): binary list, ASCII
 * list, or legacy VTK (four header lines, "POINTS n type", coordinates). */
static double *read_seed_file(const fvc_deck *d, const char *path, int64_t *n) {
    double *xyz = NULL;
    int cnt = 0;
    if (d->Trace_InFileFormat == 4) { /* tracers.c:1012-1036 */
        FILE *f = fopen(path, "rb");
        if (!f) fatal("Could not open file %s", path);
        int32_t c32;
        if (fread(&c32, 4, 1, f) != 1 || c32 < 0) fatal("Could not read Trace_NumTracers from %s", path);
        cnt = c32;
        xyz = (double *)calloc((size_t)3 * (cnt > 0 ? cnt : 1), sizeof(double));
        for (int q = 0; q < cnt; q++)
            if (fread(xyz + 3 * q, 8, (size_t)d->Dimensions, f) != (size_t)d->Dimensions)
                fatal("Could not read tracer coordinates from %s", path);
        fclose(f);
    } else if (d->Trace_InFileFormat == 1) { /* tracers.c:1039-1066 */
        FILE *f = fopen(path, "r");
        if (!f) fatal("Could not open file %s", path);
        if (fscanf(f, "%d", &cnt) != 1 || cnt < 0) fatal("Could not read Trace_NumTracers from %s", path);
        xyz = (double *)calloc((size_t)3 * (cnt > 0 ? cnt : 1), sizeof(double));
        for (int q = 0; q < cnt; q++)
            for (int a = 0; a < d->Dimensions; a++)
                if (fscanf(f, "%lf", xyz + 3 * q + a) != 1) fatal("Could not read tracer coordinates from %s", path);
        fclose(f);
    } else if (d->Trace_InFileFormat == 2 || d->Trace_InFileFormat == 3) { /* tracers.c:1067-1100 */
        if (d->Dimensions == 2) fatal("%s", "The specified Trace_InFileFormat is not currently supported for 2D analysis");
        FILE *f = fopen(path, "r");
        if (!f) fatal("Could not open file %s", path);
        char buf[4096];
        for (int q = 0; q < 4; q++)
            if (!fgets(buf, sizeof buf, f)) fatal("%s", "Skipping of first four header lines failed.");
        if (fscanf(f, "%*s %d %*s\n", &cnt) != 1 || cnt < 0) fatal("Could not read Trace_NumTracers from %s", path);
        xyz = (double *)calloc((size_t)3 * (cnt > 0 ? cnt : 1), sizeof(double));
        for (int q = 0; q < cnt; q++)
            if (fscanf(f, "%lf %lf %lf", xyz + 3 * q, xyz + 3 * q + 1, xyz + 3 * q + 2) != 3)
                fatal("Could not read tracer coordinates from %s", path);
        fclose(f);
    } else {
        fatal("%s", "Unsupported value for Trace_InFileFormat in ReadInTraceMesh()");
    }
    *n = cnt;
    return xyz;
}

/* A release's stored records (the .IC files, tracers.c:1258-1276, read back
 * by CreateNewTraceLaunch, tracers.c:1579-1631): int32 count, then
 * LagrangianPoint records. */
static fvc_point *read_ic_file(const char *path, int64_t *n) {
    FILE *f = fopen(path, "rb");
    if (!f) fatal("Could not open file %s", path);
    int32_t cnt;
    if (fread(&cnt, 4, 1, f) != 1 || cnt < 0) fatal("Could not read Trace_NumTracers from %s", path);
    fvc_point *r = (fvc_point *)calloc((size_t)(cnt > 0 ? cnt : 1), sizeof(fvc_point));
    if (fread(r, sizeof(fvc_point), (size_t)cnt, f) != (size_t)cnt)
        fatal("Could not completely read TraceIC_Array data from %s", path);
    fclose(f);
    *n = cnt;
    return r;
}

static void write_ic_file(const char *path, const fvc_point *ic, int64_t n) {
    FILE *o = fopen(path, "wb");
    if (!o) fatal("Could not open file %s", path);
    int32_t n32 = (int32_t)n;
    if (fwrite(&n32, 4, 1, o) != 1 || fwrite(ic, sizeof(fvc_point), (size_t)n, o) != (size_t)n)
        fatal("Could not completely write TraceIC_Array to %s", path);
    fclose(o);
}

/* Description of the seed set beside per-seed outputs (OutputRT, OutputAP:
 * residencetime.c:29-66, strainrate.c:250-290): the generated grid's header,
 * or the seeds' initial coordinates. */
static void write_trace_mesh(const fvc_deck *d, const char *prefix, const fvc_point *ic, int64_t n) {
    char path[3 * FVC_STR];
    if (d->Trace_GenerateMesh) {
        snprintf(path, sizeof path, "%s%s_Cartesian.bin", d->Path_Output, prefix);
        double cmn[3] = {d->Trace_Cart.XMin, d->Trace_Cart.YMin, d->Trace_Cart.ZMin};
        double cmx[3] = {d->Trace_Cart.XMax, d->Trace_Cart.YMax, d->Trace_Cart.ZMax};
        int32_t cres[3] = {d->Trace_Cart.XRes, d->Trace_Cart.YRes, d->Trace_Cart.ZRes};
        if (fvc_write_cart_header(path, cmn, cmx, cres)) fatal("Could not open file %s", path);
    } else {
        snprintf(path, sizeof path, "%s%s_coordinates.bin", d->Path_Output, prefix);
        FILE *o = fopen(path, "wb");
        if (!o) fatal("Could not open file %s", path);
        int32_t n32 = (int32_t)n;
        fwrite(&n32, 4, 1, o);
        for (int64_t q = 0; q < n; q++) fwrite(ic[q].X, 8, 3, o);
        fclose(o);
    }
}

/* ------------------------------------------------------------------------ */
/* velocity output (velout.c:21-435)                                        */
/* ------------------------------------------------------------------------ */
/* Interpolated velocity at every output point and output time: a generated
 * grid (GenerateVelOutMesh) or a point list (ReadInVelOutMesh) located once,
 * then GetVelocity at each Output time inside the loaded interval (main.c:
 * 205-213), each point walking on from its element of the previous output. */
static void run_velout(const fvc_deck *d, fvc_ctx *c, slices_t *sl, int quiet) {
    const int dir = d->Int_TimeDirection;
    int64_t n = 0;
    fvc_point *pts = NULL;
    char path[3 * FVC_STR];
    double mn[3] = {d->VelOut_Bounds[0], d->VelOut_Bounds[2], d->VelOut_Bounds[4]};
    double mx[3] = {d->VelOut_Bounds[1], d->VelOut_Bounds[3], d->VelOut_Bounds[5]};
    int32_t res[3] = {d->VelOut_Res[0], d->VelOut_Res[1], d->VelOut_Res[2]};
    double del[3];
    if (d->VelOut_GenerateMesh) { /* velout.c:28-48 */
        for (int a = 0; a < 2; a++) del[a] = res[a] == 1 ? 0 : (mx[a] - mn[a]) / (res[a] - 1);
        if (res[2] == 1)
            del[2] = 0;
        else if (d->Dimensions == 3)
            del[2] = (mx[2] - mn[2]) / (res[2] - 1);
        else {
            fprintf(stderr, "Warning: VelOut_CartMesh.ZRes being set to 1 since velocity data is 2D\n");
            res[2] = 1;
            del[2] = 0;
        }
        n = (int64_t)res[0] * res[1] * res[2];
        pts = (fvc_point *)calloc((size_t)(n > 0 ? n : 1), sizeof(fvc_point));
        check(fvc_seed_velout_grid(c, mn, del, res, d->LocalSearchChecking, pts), c, "fvc_seed_velout_grid");
    } else { /* velout.c:188-293 */
        snprintf(path, sizeof path, "%s%s", d->Path_Data, d->VelOut_InFile);
        FILE *f = fopen(path, "r");
        if (!f) fatal("Could not open file %s", path);
        int cnt = 0;
        double *xyz = NULL;
        if (d->VelOut_InFileFormat == 1) {
            if (fscanf(f, "%d\n", &cnt) != 1) fatal("Could not read the number of points from %s", path);
            xyz = (double *)calloc((size_t)3 * (cnt > 0 ? cnt : 1), sizeof(double));
            for (int q = 0; q < cnt; q++) {
                int ok = d->Dimensions == 3 ? fscanf(f, "%lf %lf %lf\n", xyz + 3 * q, xyz + 3 * q + 1, xyz + 3 * q + 2) == 3
                                            : fscanf(f, "%lf %lf\n", xyz + 3 * q, xyz + 3 * q + 1) == 2;
                if (!ok) fatal("Could not read coordinates from %s", path);
            }
        } else if (d->VelOut_InFileFormat == 2) { /* VTK polydata: 4 header lines, then POINTS n type */
            char buf[1024];
            for (int q = 0; q < 4; q++)
                if (!fgets(buf, sizeof buf, f)) fatal("Reached end of file in %s", path);
            if (fscanf(f, "%*s %d %*s\n", &cnt) != 1) fatal("Could not read the number of points from %s", path);
            xyz = (double *)calloc((size_t)3 * (cnt > 0 ? cnt : 1), sizeof(double));
            for (int q = 0; q < cnt; q++)
                if (fscanf(f, "%lf %lf %lf", xyz + 3 * q, xyz + 3 * q + 1, xyz + 3 * q + 2) < 3)
                    fatal("Failed to read coordinates from %s", path);
        } else {
            fatal("%s", "VelOut_InFileFormat 3 (VTK unstructured connectivity output) is not supported by this driver");
        }
        fclose(f);
        n = cnt;
        pts = (fvc_point *)calloc((size_t)(n > 0 ? n : 1), sizeof(fvc_point));
        check(fvc_seed_list(c, xyz, n, d->LocalSearchChecking, pts), c, "fvc_seed_list");
        free(xyz);
    }
    if (!quiet) {
        int64_t found = 0;
        for (int64_t q = 0; q < n; q++) found += !pts[q].LeftDomain;
        printf("  %lld of %lld located in domain\n", (long long)found, (long long)n);
    }
    fvc_particles *P = NULL;
    check(fvc_particles_create(c, n > 0 ? n : 1, &P), c, "fvc_particles_create");
    if (n) check(fvc_particles_set(P, 0, pts, n), c, "fvc_particles_set");
    const int TR = d->Output_TRes;
    int *done = (int *)calloc((size_t)TR, sizeof(int));
    fvc_point *cur = (fvc_point *)calloc((size_t)(n > 0 ? n : 1), sizeof(fvc_point));
    for (int df = d->Data_FirstFrame; df != d->Data_LastFrame; df += dir) {
        int flo = dir > 0 ? df : df - 1;
        int lo = slot_for(sl, flo, -1);
        int hi = slot_for(sl, flo + 1, lo);
        const double tmin = d->Data_TMin + flo * d->Data_TDelta, tmax = tmin + d->Data_TDelta;
        for (int ss = 0; ss < TR; ss++) {
            const double t = d->Output_TStart + dir * ss * d->Output_TDelta; /* velout.c:183-184 */
            if (!(t - tmin > -TINY && t - tmax < TINY) || done[ss]) continue;
            if (ss == 0) { /* mesh description with the first output (velout.c:318-350) */
                if (d->VelOut_GenerateMesh) {
                    snprintf(path, sizeof path, "%s%s_Cartesian.bin", d->Path_Output, d->VelOut_FilePrefix);
                    if (fvc_write_cart_header(path, mn, mx, res)) fatal("Could not open file %s", path);
                } else {
                    snprintf(path, sizeof path, "%s%s_coordinates.bin", d->Path_Output, d->VelOut_FilePrefix);
                    FILE *o = fopen(path, "wb");
                    if (!o) fatal("Could not open file %s", path);
                    int32_t n32 = (int32_t)n;
                    fwrite(&n32, 4, 1, o);
                    for (int64_t q = 0; q < n; q++) fwrite(pts[q].X, 8, 3, o);
                    fclose(o);
                }
            }
            if (n) {
                check(fvc_launch_velocity(c, P, 0, n, t, lo, hi, tmin, tmax), c, "fvc_launch_velocity");
                check(fvc_particles_get(P, 0, cur, n), c, "fvc_particles_get");
            }
            snprintf(path, sizeof path, "%s%s.%d.bin", d->Path_Output, d->VelOut_FilePrefix, ss);
            FILE *o = fopen(path, "wb");
            if (!o) fatal("Could not open file %s", path);
            fwrite(&t, 8, 1, o);
            const double zero[3] = {0.0, 0.0, 0.0};
            for (int64_t q = 0; q < n; q++) fwrite(pts[q].LeftDomain ? zero : cur[q].V, 8, 3, o);
            fclose(o);
            for (int64_t q = 0; q < n; q++) pts[q].LeftDomain = cur[q].LeftDomain;
            done[ss] = 1;
            if (!quiet) printf("Outputting velocity field at time %f...OK!\n", t);
        }
    }
    free(done);
    free(cur);
    free(pts);
    fvc_particles_destroy(P);
}

/* ------------------------------------------------------------------------ */
/* staggered release (tracers.c:293-1000)                                    */
/* ------------------------------------------------------------------------ */
/* Release points from Trace_InFile (binary list), located by the list chain
 * with the global fallback always on; the spacing voxdim is the smallest
 * distance between consecutive points (LoadReleasePoints, tracers.c:798-990). */
static fvc_point *stagger_points(const fvc_deck *d, fvc_ctx *c, int64_t *nloc, double *voxdim) {
    char path[3 * FVC_STR];
    snprintf(path, sizeof path, "%s%s", d->Path_Data, d->Trace_InFile);
    FILE *f = fopen(path, "rb");
    int32_t cnt = 0;
    if (!f) fatal("Could not open %s", path);
    if (fread(&cnt, 4, 1, f) != 1) fatal("Could not read Trace_NumReleaseLocations from %s", path);
    double *xyz = (double *)calloc((size_t)3 * (cnt > 0 ? cnt : 1), sizeof(double));
    for (int q = 0; q < cnt; q++)
        if (fread(xyz + 3 * q, 8, 3, f) != 3) fatal("Could not read complete coordinate list from %s", path);
    fclose(f);
    if (cnt > 1) { /* dist() of mymath.c:29-35 */
        double m = -1.0;
        for (int q = 0; q + 1 < cnt; q++) {
            double dv[3], s2 = 0;
            for (int a = 0; a < 3; a++) dv[a] = xyz[3 * q + a] - xyz[3 * (q + 1) + a];
            for (int a = 0; a < 3; a++) s2 = s2 + dv[a] * dv[a];
            double dd = sqrt(s2);
            m = q == 0 ? dd : fmin(m, dd);
        }
        *voxdim = m;
        if (*voxdim < TINY) fatal("%s", "Spacing between release points becomes nonpositive");
    } else if (cnt > 0) {
        fprintf(stderr, "\nWarning: Only 1 release location\n");
        fprintf(stderr, "Enter desired spacing:");
        if (fscanf(stdin, "%lf", voxdim) != 1) fatal("%s", "no spacing given for a single release location");
    } else {
        fatal("%s", "Number of release locations cannot be zero");
    }
    fvc_point *rp = (fvc_point *)calloc((size_t)cnt, sizeof(fvc_point));
    check(fvc_seed_list(c, xyz, cnt, 1, rp), c, "fvc_seed_list");
    int32_t found = 0;
    for (int q = 0; q < cnt; q++) found += !rp[q].LeftDomain;
    if (found < cnt) { /* the reference rewrites the file with the interior points (tracers.c:970-990) */
        FILE *o = fopen(path, "wb");
        if (!o) fatal("Could not open file %s", path);
        fwrite(&found, 4, 1, o);
        for (int q = 0; q < cnt; q++)
            if (!rp[q].LeftDomain) fwrite(xyz + 3 * q, 8, 3, o);
        fclose(o);
    }
    free(xyz);
    *nloc = cnt;
    return rp;
}

typedef struct {
    double start, stop;
    int32_t loc, elem; /* release location, element copied at creation */
} snode_t;

typedef struct {
    snode_t *v;
    int64_t n, cap;
} snodes_t;

static void snode_add(snodes_t *L, const fvc_deck *d, int32_t loc, int32_t elem, double start) {
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 1024;
        L->v = (snode_t *)realloc(L->v, sizeof(snode_t) * (size_t)L->cap);
        if (!L->v) fatal("%s", "Malloc failed for nnode");
    }
    snode_t *s = &L->v[L->n++];
    s->start = start;
    s->stop = fmin(d->Output_TEnd, start + d->Trace_IntTLength);
    s->loc = loc;
    s->elem = elem;
}

static double vdot3(const double *a, const double *b) {
    double sum = 0;
    for (int i = 0; i < 3; i++) sum = sum + a[i] * b[i];
    return sum;
}

static double sign_of(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); } /* macros.h:32 */

/* Release times per location from the flow through it (GenerateStaggeredRelease,
 * tracers.c:418-796): a new point is released whenever the previous one has
 * travelled voxdim, with the signed speed linear in time over each data
 * interval (linear or quadratic solve).  Velocities at the release points are
 * GetVelocity at the interval bounds, on the device.  Returns nodes grouped
 * by location in release order (the reference's linked lists). */
static snodes_t stagger_schedule(const fvc_deck *d, fvc_ctx *c, slices_t *sl, const fvc_point *rp, int64_t nloc,
                                 double voxdim) {
    int64_t nn = 0;
    snodes_t *per = (snodes_t *)calloc((size_t)nloc, sizeof(snodes_t));
    double *voxdim0 = (double *)calloc((size_t)nloc, sizeof(double)), *u0 = (double *)calloc((size_t)nloc * 3, 8);
    int *ptdone = (int *)calloc((size_t)nloc, sizeof(int));
    for (int64_t ss = 0; ss < nloc; ss++)
        if (!rp[ss].LeftDomain) {
            snode_add(&per[ss], d, (int32_t)ss, rp[ss].ElementIndex, d->Output_TStart);
            nn++;
        }
    fvc_particles *P = NULL;
    check(fvc_particles_create(c, nloc, &P), c, "fvc_particles_create");
    check(fvc_particles_set(P, 0, rp, nloc), c, "fvc_particles_set");
    fvc_point *us = (fvc_point *)malloc(sizeof(fvc_point) * (size_t)nloc);
    fvc_point *ue = (fvc_point *)malloc(sizeof(fvc_point) * (size_t)nloc);
    for (int df = d->Data_FirstFrame; df < d->Data_LastFrame; df++) {
        int lo = slot_for(sl, df, -1);
        int hi = slot_for(sl, df + 1, lo);
        const double tmin = d->Data_TMin + df * d->Data_TDelta, tmax = tmin + d->Data_TDelta;
        /* GetVelocity_* at both bounds; it walks from, and updates, each point's element */
        check(fvc_launch_velocity(c, P, 0, nloc, tmin, lo, hi, tmin, tmax), c, "fvc_launch_velocity");
        check(fvc_particles_get(P, 0, us, nloc), c, "fvc_particles_get");
        if (df == d->Data_FirstFrame)
            for (int64_t ss = 0; ss < nloc; ss++)
                if (!rp[ss].LeftDomain) memcpy(u0 + 3 * ss, us[ss].V, 3 * sizeof(double));
        if (tmin >= d->Trace_ReleaseTMax) break;
        check(fvc_launch_velocity(c, P, 0, nloc, tmax, lo, hi, tmin, tmax), c, "fvc_launch_velocity");
        check(fvc_particles_get(P, 0, ue, nloc), c, "fvc_particles_get");
        for (int64_t ss = 0; ss < nloc; ss++) {
            if (rp[ss].LeftDomain || ptdone[ss]) continue;
            snodes_t *L = &per[ss];
            const int32_t elem = ue[ss].ElementIndex;
            const double usbar = sign_of(sqrt(vdot3(us[ss].V, us[ss].V)), vdot3(us[ss].V, u0 + 3 * ss));
            const double uebar = sign_of(sqrt(vdot3(ue[ss].V, ue[ss].V)), vdot3(ue[ss].V, u0 + 3 * ss));
            double ti, tii, dd;
            if (usbar < TINY && uebar < TINY) continue; /* reverse or stagnant flow */
            if (fabs(usbar - uebar) < 1.0e-6) {          /* steady: linear solve (tracers.c:557-620) */
                if (voxdim0[ss] > TINY)
                    ti = tmin;
                else if (voxdim0[ss] < -TINY) {
                    ti = tmin;
                    voxdim0[ss] = 0;
                } else
                    ti = L->v[L->n - 1].start;
                for (;;) {
                    dd = voxdim - voxdim0[ss];
                    if (dd < 0) fatal("%s", "d < 0");
                    tii = ti + dd / usbar;
                    if (tii <= ti) fatal("%s", "Next output time not after previous");
                    if (tii >= d->Trace_ReleaseTMax) {
                        ptdone[ss] = 1;
                        break;
                    } else if (tii > tmax) {
                        voxdim0[ss] = voxdim0[ss] + uebar * (tmax - ti);
                        if (voxdim0[ss] < 0) fatal("%s", "voxdim0 < 0");
                        break;
                    }
                    snode_add(L, d, (int32_t)ss, elem, tii);
                    nn++;
                    voxdim0[ss] = 0;
                    ti = tii;
                }
            } else { /* speed linear in time: quadratic solve (tracers.c:621-765) */
                const double M = (usbar - uebar) / (tmin - tmax);
                const double N = usbar - M * tmin;
                if (usbar < 0 && uebar > 0) { /* release when the flow turns positive */
                    snode_add(L, d, (int32_t)ss, elem, -N / M);
                    nn++;
                    voxdim0[ss] = 0;
                    ti = -N / M;
                } else {
                    if (voxdim0[ss] > TINY)
                        ti = tmin;
                    else if (voxdim0[ss] < -TINY) {
                        ti = tmin;
                        voxdim0[ss] = 0;
                    } else
                        ti = L->v[L->n - 1].start;
                }
                for (;;) {
                    const double uibar = M * ti + N;
                    dd = voxdim - voxdim0[ss];
                    if (dd < 0) fatal("%s", "d < 0");
                    const double a = M, b = N + uibar - M * ti, cc = -N * ti - uibar * ti - 2 * dd;
                    if (b * b - 4 * a * cc < 0) {
                        if (usbar < uebar) fatal("%s", "Next output time has imaginary roots even with increasing flow");
                        voxdim0[ss] = -1; /* next interval starts from its own TMin */
                        break;
                    }
                    tii = (-b + sqrt(b * b - 4 * a * cc)) / (2 * a);
                    if (tii <= ti) fatal("%s", "Next output time not after previous");
                    if (tii >= d->Trace_ReleaseTMax) {
                        ptdone[ss] = 1;
                        break;
                    } else if (tii > tmax) {
                        voxdim0[ss] = voxdim0[ss] + 0.5 * (uebar + uibar) * (tmax - ti);
                        if (voxdim0[ss] < 0) fatal("%s", "voxdim0 < 0");
                        break;
                    }
                    snode_add(L, d, (int32_t)ss, elem, tii);
                    nn++;
                    voxdim0[ss] = 0;
                    ti = tii;
                }
            }
        }
    }
    fvc_particles_destroy(P);
    snodes_t all = {NULL, 0, 0};
    all.v = (snode_t *)malloc(sizeof(snode_t) * (size_t)(nn > 0 ? nn : 1));
    all.cap = nn;
    for (int64_t ss = 0; ss < nloc; ss++) {
        if (per[ss].n) memcpy(all.v + all.n, per[ss].v, sizeof(snode_t) * (size_t)per[ss].n);
        all.n += per[ss].n;
        free(per[ss].v);
    }
    free(per);
    free(voxdim0);
    free(u0);
    free(ptdone);
    free(us);
    free(ue);
    return all;
}

typedef struct {
    int64_t node;
    int32_t file;
    double X[3];
} srec_t;

static int srec_cmp(const void *a, const void *b) {
    const srec_t *x = (const srec_t *)a, *y = (const srec_t *)b;
    if (x->file != y->file) return x->file < y->file ? -1 : 1;
    return x->node < y->node ? -1 : (x->node > y->node);
}

/* Tracer_StaggeredRelease (tracers.c:293-416) over every data interval: each
 * launched point integrates to each output time inside the interval, then to
 * the interval's end or its stop time; positions are written per output time
 * in location and release order.  Every round of segments is one device call
 * with a segment per point. */
static int64_t run_staggered(const fvc_deck *d, fvc_ctx *c, slices_t *sl, const fvc_int_params *ip, int quiet,
                             double *t_compute) {
    int64_t nloc = 0, steps_total = 0;
    double voxdim = 0.0;
    fvc_point *rp = stagger_points(d, c, &nloc, &voxdim);
    if (!quiet) printf("Determining release times...");
    /* release times come from the velocity as read: ompVC determines them
     * before it loads the wall normals (main.c:55-67, tracers.c:488-495) */
    const int32_t nwall = sl->nwall;
    sl->nwall = 0;
    snodes_t S = stagger_schedule(d, c, sl, rp, nloc, voxdim);
    if (nwall) { /* integration reads every slice with the wall velocities */
        sl->nwall = nwall;
        for (int q = 0; q < sl->nslots; q++) sl->slot_file[q] = -1;
    }
    if (!quiet) printf("OK!\nTotal number of points to be released is %lld\n", (long long)S.n);
    const int64_t n = S.n;
    const int TR = d->Output_TRes;
    const double TS = d->Output_TStart, DT = d->Output_TDelta;
    FILE **of = (FILE **)calloc((size_t)TR, sizeof(FILE *));
    char path[3 * FVC_STR];
    for (int ss = 0; ss < TR; ss++) { /* main.c:92-106 */
        snprintf(path, sizeof path, "%s%s.%d.bin", d->Path_Output, d->Trace_OutFilePrefix, ss);
        if (!(of[ss] = fopen(path, "wb"))) fatal("Could not open file %s", path);
        double tt = TS + ss * DT;
        fwrite(&tt, 8, 1, of[ss]);
    }
    fvc_particles *P = NULL;
    check(fvc_particles_create(c, n > 0 ? n : 1, &P), c, "fvc_particles_create");
    fvc_point *recs = (fvc_point *)calloc((size_t)(n > 0 ? n : 1), sizeof(fvc_point));
    for (int64_t k = 0; k < n; k++) {
        memcpy(recs[k].X, rp[S.v[k].loc].X, sizeof recs[k].X);
        recs[k].ElementIndex = S.v[k].elem;
    }
    if (n) check(fvc_particles_set(P, 0, recs, n), c, "fvc_particles_set");
    double *X = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    int32_t *left = (int32_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    for (int64_t k = 0; k < n; k++) memcpy(X + 3 * k, recs[k].X, sizeof recs[k].X);
    int *status = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    int *nout = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int)), *first_ii = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    double *tt = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1)), *t2last = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    unsigned char *act = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
    srec_t *rec = NULL;
    int64_t nrec = 0, caprec = 0;
#define SREC(k_, file_)                                                                                                \
    do {                                                                                                               \
        if (nrec == caprec) {                                                                                          \
            caprec = caprec ? 2 * caprec : 4096;                                                                       \
            rec = (srec_t *)realloc(rec, sizeof(srec_t) * (size_t)caprec);                                             \
        }                                                                                                              \
        rec[nrec].node = (k_);                                                                                         \
        rec[nrec].file = (file_);                                                                                      \
        memcpy(rec[nrec].X, X + 3 * (k_), 3 * sizeof(double));                                                        \
        nrec++;                                                                                                        \
    } while (0)
    const double NaN = nan("");
    for (int df = d->Data_FirstFrame; df != d->Data_LastFrame; df++) {
        int lo = slot_for(sl, df, -1);
        int hi = slot_for(sl, df + 1, lo);
        const double tmin = d->Data_TMin + df * d->Data_TDelta, tmax = tmin + d->Data_TDelta;
        nrec = 0;
        int maxout = 0, any = 0;
        for (int64_t k = 0; k < n; k++) {
            const snode_t *r = &S.v[k];
            act[k] = 0;
            tt[2 * k] = NaN;
            tt[2 * k + 1] = NaN;
            if (status[k] == 2 || !(status[k] == 1 || (r->start >= tmin && r->start <= tmax))) continue;
            if (left[k] && !ip->extrapolate) continue;
            if (status[k] == 0) {
                status[k] = 1;
                if (fabs(r->start - TS) < TINY) SREC(k, 0);
            }
            double t1 = fmax(r->start, tmin), t2 = fmin(r->stop, tmax);
            int oc = 0;
            for (int ii = 0; ii < TR; ii++)
                if ((TS + ii * DT > t1 + TINY) && (TS + ii * DT <= t2 + TINY)) oc++;
            nout[k] = oc;
            if (oc) {
                int ii;
                double next = 0.0;
                for (ii = 0; ii < TR; ii++) {
                    next = TS + ii * DT;
                    if (next > t1 + TINY) break;
                }
                first_ii[k] = ii;
                t2 = next;
            }
            tt[2 * k] = t1;
            tt[2 * k + 1] = t2;
            t2last[k] = t2;
            act[k] = 1;
            any = 1;
            if (oc > maxout) maxout = oc;
        }
        if (any) {
            double t0 = now_s();
            /* round 0: to the first output time or the segment end */
            check(fvc_advect_times(c, P, 0, n, tt, lo, hi, tmin, tmax, ip), c, "fvc_advect_times");
            int64_t st = 0;
            fvc_last_step_count(c, &st);
            steps_total += st;
            check(fvc_particles_get_xyz(P, 0, X, left, n), c, "fvc_particles_get_xyz");
            for (int64_t k = 0; k < n; k++)
                if (act[k] && nout[k] && (!left[k] || ip->extrapolate)) {
                    if (first_ii[k] >= TR) fatal("%s", "output index out of range");
                    SREC(k, first_ii[k]);
                }
            /* later output times inside the interval (t2 += Output_TDelta, as tracers.c:366-368) */
            for (int r = 1; r < maxout; r++) {
                int more = 0;
                for (int64_t k = 0; k < n; k++) {
                    tt[2 * k] = NaN;
                    if (act[k] && nout[k] > r && (!left[k] || ip->extrapolate)) {
                        tt[2 * k] = t2last[k];
                        tt[2 * k + 1] = t2last[k] + DT;
                        t2last[k] = tt[2 * k + 1];
                        more = 1;
                    }
                }
                if (!more) break;
                check(fvc_advect_times(c, P, 0, n, tt, lo, hi, tmin, tmax, ip), c, "fvc_advect_times");
                fvc_last_step_count(c, &st);
                steps_total += st;
                check(fvc_particles_get_xyz(P, 0, X, left, n), c, "fvc_particles_get_xyz");
                for (int64_t k = 0; k < n; k++)
                    if (tt[2 * k] == tt[2 * k] && (!left[k] || ip->extrapolate)) {
                        const int file = (int)((tt[2 * k + 1] + TINY) / DT);
                        if (file < 0 || file >= TR) fatal("%s", "output index out of range");
                        SREC(k, file);
                    }
            }
            /* remainder of the interval, up to the stop time */
            int more = 0;
            for (int64_t k = 0; k < n; k++) {
                tt[2 * k] = NaN;
                if (act[k] && nout[k] && (!left[k] || ip->extrapolate)) {
                    const double t1 = t2last[k], t2 = fmin(S.v[k].stop, tmax);
                    if (t2 - t1 > 0) {
                        tt[2 * k] = t1;
                        tt[2 * k + 1] = t2;
                        more = 1;
                    } else if (t2 - t1 < -TINY) {
                        fatal("%s", "Integration of point beyond requested stop time or range of loaded data");
                    }
                }
            }
            if (more) {
                check(fvc_advect_times(c, P, 0, n, tt, lo, hi, tmin, tmax, ip), c, "fvc_advect_times");
                fvc_last_step_count(c, &st);
                steps_total += st;
                check(fvc_particles_get_xyz(P, 0, X, left, n), c, "fvc_particles_get_xyz");
            }
            *t_compute += now_s() - t0;
            for (int64_t k = 0; k < n; k++)
                if (act[k] && S.v[k].stop <= tmax) status[k] = 2;
        }
        /* this interval's records: per output file, in location and release order */
        if (nrec) {
            qsort(rec, (size_t)nrec, sizeof(srec_t), srec_cmp);
            for (int64_t q = 0; q < nrec; q++) fwrite(rec[q].X, 8, 3, of[rec[q].file]);
        }
        if (!quiet) printf(" %lld tracers integrated.\n", (long long)0);
    }
#undef SREC
    for (int ss = 0; ss < TR; ss++) fclose(of[ss]);
    free(of);
    free(rec);
    free(act);
    free(tt);
    free(t2last);
    free(nout);
    free(first_ii);
    free(status);
    free(left);
    free(X);
    free(recs);
    free(S.v);
    free(rp);
    fvc_particles_destroy(P);
    return steps_total;
}

int main(int argc, char **argv) {
    const char *deck_path = NULL;
    int precision = FVC_PREC_F64, nproc = (int)sysconf(_SC_NPROCESSORS_ONLN), quiet = 0;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--precision") && a + 1 < argc) {
            const char *p = argv[++a];
            precision = !strcmp(p, "f32") ? FVC_PREC_F32 : !strcmp(p, "bf16") ? FVC_PREC_BF16
                      : !strcmp(p, "f16") ? FVC_PREC_F16 : FVC_PREC_F64;
        } else if (!strcmp(argv[a], "--nproc") && a + 1 < argc) {
            nproc = atoi(argv[++a]);
        } else if (!strcmp(argv[a], "--quiet")) {
            quiet = 1;
        } else {
            deck_path = argv[a];
        }
    }
    if (!deck_path) {
        fprintf(stderr, "usage: %s <deck.in> [--precision f64|f32|bf16|f16] [--nproc N] [--quiet]\n", argv[0]);
        return 2;
    }
    double t_start = now_s();
    fvc_deck d;
    char err[512];
    if (fvc_deck_read(deck_path, &d, err, sizeof err) || fvc_deck_check(&d, err, sizeof err) ||
        fvc_deck_derive(&d, err, sizeof err))
        fatal("%s", err);
    if (d.legacy_96 && !quiet) fprintf(stderr, "Note: deck has 96 entries (no Num_threads); read as Num_threads = 0\n");
    if (d.Particle_Radius > TINY) fatal("%s", "deck selects a mode outside the advection path (inertial particles)");
    const int ap = d.Trace_Compute && d.Trace_APCompute;
    const int cet = d.Trace_Compute && d.Trace_CETCompute, rt = d.Trace_Compute && d.Trace_RTCompute;
    if (cet && (d.Trace_CETAuxillaryMesh || d.Data_MeshType != 1))
        fatal("%s", "exposure times are supported on the velocity mesh itself (unstructured, no auxiliary mesh)");
    if ((cet || rt) && d.Trace_ReleaseStrategy == 1)
        fatal("%s", "exposure or residence times with the staggered release are not supported");
    if (ap && d.Trace_ReleaseStrategy == 1)
        fatal("%s", "Trace_APCompute with the staggered release is not supported (the reference's OutputAP reads "
                    "release files that the staggered release never writes)");
    const int dir = d.Int_TimeDirection;

    /* mesh (mesh.c:21-213) */
    fvc_mesh_desc md;
    memset(&md, 0, sizeof md);
    md.dims = d.Dimensions;
    md.mesh_type = d.Data_MeshType;
    memcpy(md.bounds, d.Bounds, sizeof md.bounds);
    double *xyz = NULL;
    int32_t *conn = NULL, *adj = NULL;
    char path[3 * FVC_STR];
    if (d.Data_MeshType == 0) {
        snprintf(path, sizeof path, "%s%s_Cartesian.bin", d.Path_Data, d.Data_InFilePrefix);
        if (fvc_read_cart_header(path, md.cart_min, md.cart_max, md.cart_res)) fatal("Could not open file %s", path);
    } else {
        int64_t nn, ne;
        if (fvc_read_unstruct_counts(d.Path_Data, d.Data_InFilePrefix, &nn, &ne))
            fatal("Could not open mesh files with prefix %s", d.Data_InFilePrefix);
        xyz = (double *)malloc(sizeof(double) * 3 * (size_t)nn);
        conn = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)ne);
        adj = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)ne);
        int rc = fvc_read_unstruct_mesh(d.Path_Data, d.Data_InFilePrefix, xyz, conn, adj);
        if (rc == -3) fatal("%s", "Incompatible number of elements listed in connectivity and adjacency files");
        if (rc) fatal("Could not read mesh files with prefix %s", d.Data_InFilePrefix);
        md.num_nodes = nn;
        md.num_elems = ne;
        md.coords = xyz;
        md.conn = conn;
        md.adj = adj;
    }
    fvc_params pp = {precision, d.Data_TRes - 1 >= 2 ? d.Data_TRes - 1 : 2, 0, 0, 0};
    fvc_ctx *c = NULL;
    check(fvc_create(&md, &pp, &c), NULL, "fvc_create");
    slices_t sl;
    memset(&sl, 0, sizeof sl);
    sl.c = c;
    sl.d = &d;
    sl.n = d.Data_MeshType == 0 ? (int64_t)md.cart_res[0] * md.cart_res[1] * (d.Dimensions == 3 ? md.cart_res[2] : 1)
                                : md.num_nodes;
    sl.nslots = pp.max_slices;
    sl.slot_file = (int *)malloc(sizeof(int) * sl.nslots);
    for (int q = 0; q < sl.nslots; q++) sl.slot_file[q] = -1;
    sl.buf = (double *)malloc(sizeof(double) * 3 * (size_t)sl.n);
    if (d.Int_NormalFlow && d.Data_MeshType == 1) { /* LoadMeshSurfaceNormals (mesh.c:255-300) */
        snprintf(path, sizeof path, "%s%s_normals.bin", d.Path_Data, d.Data_InFilePrefix);
        FILE *fn = fopen(path, "rb");
        if (!fn) fatal("Could not open file %s", path);
        if (fread(&sl.nwall, 4, 1, fn) != 1 || sl.nwall < 0)
            fatal("Could not read value for Vel_SurfaceMeshNumNodes from %s", path);
        sl.wall_ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)(sl.nwall > 0 ? sl.nwall : 1));
        sl.wall_v = (double *)malloc(sizeof(double) * 3 * (size_t)(sl.nwall > 0 ? sl.nwall : 1));
        if (fread(sl.wall_ids, 4, (size_t)sl.nwall, fn) != (size_t)sl.nwall)
            fatal("Could not completely read Vel_SurfaceMeshNodeIDs array from %s", path);
        if (fread(sl.wall_v, 8, 3 * (size_t)sl.nwall, fn) != 3 * (size_t)sl.nwall)
            fatal("Could not completely read inward normal vector data from %s", path);
        fclose(fn);
        for (int32_t i = 0; i < sl.nwall; i++) {
            if (sl.wall_ids[i] < 0 || sl.wall_ids[i] >= md.num_nodes) fatal("surface node id out of range in %s", path);
            for (int a = 0; a < 3; a++) /* Int_TimeDirection * Int_NormalFlowScaling * normal, in that order */
                sl.wall_v[3 * i + a] = d.Int_TimeDirection * d.Int_NormalFlowScaling * sl.wall_v[3 * i + a];
        }
    }
    if (ap) {
        sl.nsr = d.Data_MeshType == 1 ? md.num_elems
                                      : (int64_t)(md.cart_res[0] - 1) * (md.cart_res[1] - 1) *
                                            ((d.Dimensions == 2 || md.cart_res[2] == 1) ? 1 : md.cart_res[2] - 1);
        sl.srbuf = (double *)malloc(sizeof(double) * (size_t)(sl.nsr > 0 ? sl.nsr : 1));
    }
    /* Advect's dispatch (integration.c:32-37): 1 RK4, 2 RKF, anything else Euler */
    fvc_int_params ip = {d.Int_Type == 1 ? FVC_INT_RK4 : d.Int_Type == 2 ? FVC_INT_RKF : FVC_INT_EULER,
                         d.Int_Extrapolate, d.Int_TimeStep, 0, 0, ap, cet ? d.Trace_CETSubsteps : 0,
                         d.Int_MinTimeStep, d.Int_MaxTimeStep, d.Int_Accuracy};
    if (cet) check(fvc_cet_reset(c), c, "fvc_cet_reset");
    double t_compute = 0.0;
    int64_t steps_total = 0;

    if (d.VelOut_Compute) run_velout(&d, c, &sl, quiet);
    if (d.FTLE_Compute) {
        int ns = d.Output_TRes;
        fvc_ftle **f = (fvc_ftle **)calloc((size_t)ns, sizeof(fvc_ftle *));
        double mn[3] = {d.FTLE_Bounds[0], d.FTLE_Bounds[2], d.FTLE_Bounds[4]};
        double mx[3] = {d.FTLE_Bounds[1], d.FTLE_Bounds[3], d.FTLE_Bounds[5]};
        int32_t res[3] = {d.FTLE_Res[0], d.FTLE_Res[1], d.FTLE_Res[2]};
        char icpath[3 * FVC_STR];
        snprintf(icpath, sizeof icpath, "%s%s", d.Path_Data, d.FTLE_ICFile);
        /* FTLE_GenerateMesh = 0: every slide starts from the records of FTLE_ICFile, whose resolution replaces
         * the parameters' while origin and spacing stay (ftle.c:626-660) */
        fvc_point *icp = NULL;
        int32_t *ich = NULL;
        double *icw = NULL, *icwo = NULL;
        if (!d.FTLE_GenerateMesh) {
            FILE *fi = fopen(icpath, "rb");
            if (!fi) {
                fprintf(stderr, "Error: Could not open file %s (file must be generated by flowVC with FTLE_GenerateMesh = 0)\n",
                        icpath);
                exit(1);
            }
            int32_t r3[3];
            if (fread(r3, 4, 3, fi) != 3) fatal("Could not completely read FTLE grid resolution from %s", icpath);
            if (r3[0] < 1 || r3[1] < 1 || r3[2] < 1 || (int64_t)r3[0] * r3[1] * r3[2] > 0x7fffffffLL)
                fatal("FTLE grid resolution in %s out of range", icpath);
            const int64_t ni = (int64_t)r3[0] * r3[1] * r3[2];
            fvc_ftle_rec *rec = (fvc_ftle_rec *)malloc(sizeof(fvc_ftle_rec) * (size_t)ni);
            icp = (fvc_point *)malloc(sizeof(fvc_point) * (size_t)ni);
            ich = (int32_t *)malloc(sizeof(int32_t) * (size_t)ni);
            icw = (double *)malloc(sizeof(double) * (size_t)ni);
            icwo = (double *)malloc(sizeof(double) * (size_t)ni);
            if (!rec || !icp || !ich || !icw || !icwo) fatal("%s", "Malloc failed for FTLE_MeshPt_temp");
            if (fread(rec, sizeof(fvc_ftle_rec), (size_t)ni, fi) != (size_t)ni)
                fatal("Could not completely read FTLE_MeshPt_temp data from %s", icpath);
            fclose(fi);
            for (int64_t q = 0; q < ni; q++) {
                memcpy(&icp[q], &rec[q].Pt, sizeof(fvc_point));
                ich[q] = rec[q].HaveFTLE;
                icw[q] = rec[q].FTLEwT;
                icwo[q] = rec[q].FTLEwoT;
            }
            free(rec);
            d.FTLE_Cart.XRes = r3[0];
            d.FTLE_Cart.YRes = r3[1];
            d.FTLE_Cart.ZRes = r3[2];
        }
        const double gmn[3] = {d.FTLE_Cart.XMin, d.FTLE_Cart.YMin, d.FTLE_Cart.ZMin};
        const double gdel[3] = {d.FTLE_Cart.XDelta, d.FTLE_Cart.YDelta, d.FTLE_Cart.ZDelta};
        const int32_t gres[3] = {d.FTLE_Cart.XRes, d.FTLE_Cart.YRes, d.FTLE_Cart.ZRes};
        for (int ss = 0; ss < ns; ss++) {
            /* ftle.c:672-685 */
            double start = d.Output_TStart + ss * dir * d.Output_TDelta, stop;
            if (!d.Data_TPeriodic) {
                stop = fmin(d.Output_TStart + ss * dir * d.Output_TDelta + (double)dir * d.FTLE_IntTLength, d.Data_TMax);
                stop = fmax(stop, d.Data_TMin);
                if (fabs(stop - start) < d.FTLE_IntTLength - TINY && !quiet)
                    fprintf(stderr, "Warning: FTLE at time %f will have a maximum integration length of %f\n", start,
                            fabs(stop - start));
            } else {
                stop = d.Output_TStart + ss * dir * d.Output_TDelta + (double)dir * d.FTLE_IntTLength;
            }
            if (d.FTLE_GenerateMesh)
                check(fvc_ftle_create(c, mn, mx, res, d.LocalSearchChecking, start, stop, &f[ss]), c, "fvc_ftle_create");
            else
                check(fvc_ftle_create_from(c, gmn, gdel, gres, icp, ich, icw, icwo, start, stop, &f[ss]), c,
                      "fvc_ftle_create_from");
        }
        if (d.FTLE_GenerateMesh && ns > 0) { /* FTLE_ICFile: resolution, then the grid's records (ftle.c:611-625) */
            int64_t ni = 0, nin = 0;
            fvc_ftle_points(f[0], &ni, &nin);
            fvc_point *pr = (fvc_point *)malloc(sizeof(fvc_point) * (size_t)(ni > 0 ? ni : 1));
            int32_t *hv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ni > 0 ? ni : 1));
            double *w1 = (double *)malloc(sizeof(double) * (size_t)(ni > 0 ? ni : 1)),
                   *w2 = (double *)malloc(sizeof(double) * (size_t)(ni > 0 ? ni : 1));
            fvc_ftle_rec *rec = (fvc_ftle_rec *)calloc((size_t)(ni > 0 ? ni : 1), sizeof(fvc_ftle_rec));
            if (!pr || !hv || !w1 || !w2 || !rec) fatal("%s", "Malloc failed for FTLE_MeshPt_temp");
            check(fvc_ftle_initial(f[0], pr, hv, w1, w2), c, "fvc_ftle_initial");
            for (int64_t q = 0; q < ni; q++) {
                memcpy(&rec[q].Pt, &pr[q], sizeof(fvc_point));
                rec[q].HaveFTLE = hv[q];
                rec[q].FTLEwT = w1[q];
                rec[q].FTLEwoT = w2[q];
            }
            FILE *fo = fopen(icpath, "wb");
            if (!fo) fatal("Could not open file %s", icpath);
            const int32_t r3[3] = {gres[0], gres[1], gres[2]};
            if (fwrite(r3, 4, 3, fo) != 3 || fwrite(rec, sizeof(fvc_ftle_rec), (size_t)ni, fo) != (size_t)ni)
                fatal("Could not completely write FTLE_MeshPt_temp data to %s", icpath);
            fclose(fo);
            free(pr);
            free(hv);
            free(w1);
            free(w2);
            free(rec);
        }
        free(icp);
        free(ich);
        free(icw);
        free(icwo);
        int64_t npts = 0, inside = 0;
        fvc_ftle_points(f[0], &npts, &inside);
        if (!quiet) printf("  %lld of %lld located in domain\n", (long long)inside, (long long)npts);
        double *wt = (double *)malloc(sizeof(double) * (size_t)npts), *wot = (double *)malloc(sizeof(double) * (size_t)npts);
        int written_first = 0, numout = 0;
        for (int df = d.Data_FirstFrame; df != d.Data_LastFrame; df += dir) {
            int flo = dir > 0 ? df : df - 1;
            int lo = slot_for(&sl, flo, -1);
            int hi = slot_for(&sl, flo + 1, lo);
            double tmin = d.Data_TMin + flo * d.Data_TDelta, tmax = tmin + d.Data_TDelta;
            for (int ss = 0; ss < ns; ss++) {
                int32_t done = 0;
                double t0 = now_s();
                check(fvc_ftle_frame(f[ss], dir, lo, hi, tmin, tmax, &ip, &done), c, "fvc_ftle_frame");
                if (done == 0) continue; /* complete earlier, or not yet launched */
                int64_t st = 0;
                fvc_last_step_count(c, &st);
                steps_total += st;
                t_compute += now_s() - t0;
                if (done == 2) {
                    if ((ss == 0 && !d.FTLE_ComputeVariation) || (d.FTLE_ComputeVariation && !written_first)) {
                        write_info(&d);
                        written_first = 1;
                    }
                    check(fvc_ftle_output(f[ss], d.FTLE_IntTLength, wt, wot), c, "fvc_ftle_output");
                    int outnum = d.FTLE_ComputeVariation ? numout++ : (dir > 0 ? ss : d.Output_TRes - 1 - ss);
                    double start = d.Output_TStart + ss * dir * d.Output_TDelta;
                    write_ftle(&d, outnum, start, wt, wot, npts);
                    if (!quiet) printf("Stopping FTLE output slide %d and writing to output file.\n", ss);
                } else if (d.FTLE_ComputeVariation && !fmod(df - d.Data_FirstFrame, d.FTLE_VariationOutFreq)) {
                    double start = d.Output_TStart + ss * dir * d.Output_TDelta;
                    double t2 = dir > 0 ? tmax : tmin; /* slide not complete: t2 is the loaded bound */
                    if (!written_first) {
                        write_info(&d);
                        written_first = 1;
                    }
                    check(fvc_ftle_output(f[ss], fabs(t2 - start), wt, wot), c, "fvc_ftle_output");
                    write_ftle(&d, numout++, start, wt, wot, npts);
                }
            }
        }
        free(wt);
        free(wot);
        for (int ss = 0; ss < ns; ss++) fvc_ftle_destroy(f[ss]);
        free(f);
    } else if (d.Trace_Compute && d.Trace_ReleaseStrategy == 1) {
        steps_total += run_staggered(&d, c, &sl, &ip, quiet, &t_compute);
    } else if (d.Trace_Compute) {
        /* seeds: GenerateTracerMesh / ReadInTraceMesh location chains, or the
         * stored .IC records (Trace_InFileFormat 0); with Trace_MultipleInFiles
         * every release has its own seed file and .IC.<release> */
        const int nl = d.Trace_NumLaunchTimes;
        const int multi = !d.Trace_GenerateMesh && d.Trace_MultipleInFiles;
        /* OutputAP / OutputRT read the single .IC file (strainrate.c:293, residencetime.c:54), which separate
         * seed files never write; exposure times divide by one release's count (exposuretime.c:242) */
        if (multi && (ap || rt || cet))
            fatal("%s", "Trace_MultipleInFiles with activation potential, residence or exposure times is not "
                        "supported by this driver");
        fvc_point **ics = (fvc_point **)calloc((size_t)nl, sizeof(fvc_point *));
        int64_t *rn = (int64_t *)calloc((size_t)nl, sizeof(int64_t)), *roff = (int64_t *)calloc((size_t)nl + 1, 8);
        char icpath[3 * FVC_STR + 16];
        snprintf(icpath, sizeof icpath, "%s%s.IC", d.Path_Data, d.Trace_OutFilePrefix);
        if (d.Trace_GenerateMesh) {
            double mn[3] = {d.Trace_Cart.XMin, d.Trace_Cart.YMin, d.Trace_Cart.ZMin};
            double del[3] = {d.Trace_Cart.XDelta, d.Trace_Cart.YDelta, d.Trace_Cart.ZDelta};
            int32_t res[3] = {d.Trace_Cart.XRes, d.Trace_Cart.YRes, d.Trace_Cart.ZRes};
            rn[0] = (int64_t)res[0] * res[1] * res[2];
            ics[0] = (fvc_point *)calloc((size_t)(rn[0] > 0 ? rn[0] : 1), sizeof(fvc_point));
            check(fvc_seed_grid(c, mn, del, res, d.LocalSearchChecking, ics[0]), c, "fvc_seed_grid");
            write_ic_file(icpath, ics[0], rn[0]);
        } else if (d.Trace_InFileFormat != 0) {
            static const char *ext[5] = {"", "dat", "vtk", "vtk", "bin"};
            for (int ff = 0; ff < (multi ? nl : 1); ff++) {
                char sp[3 * FVC_STR + 32], ip2[3 * FVC_STR + 32];
                if (multi) {
                    snprintf(sp, sizeof sp, "%s%s.%d.%s", d.Path_Data, d.Trace_InFile, ff,
                             ext[d.Trace_InFileFormat >= 1 && d.Trace_InFileFormat <= 4 ? d.Trace_InFileFormat : 0]);
                    snprintf(ip2, sizeof ip2, "%s.%d", icpath, ff);
                } else {
                    snprintf(sp, sizeof sp, "%s%s", d.Path_Data, d.Trace_InFile);
                    snprintf(ip2, sizeof ip2, "%s", icpath);
                }
                double *sx = read_seed_file(&d, sp, &rn[ff]);
                ics[ff] = (fvc_point *)calloc((size_t)(rn[ff] > 0 ? rn[ff] : 1), sizeof(fvc_point));
                check(fvc_seed_list(c, sx, rn[ff], d.LocalSearchChecking, ics[ff]), c, "fvc_seed_list");
                free(sx);
                write_ic_file(ip2, ics[ff], rn[ff]);
            }
        } else { /* tracers.c:1006: no reading or location; the releases copy the stored records */
            for (int ff = 0; ff < (multi ? nl : 1); ff++) {
                char ip2[3 * FVC_STR + 32];
                if (multi)
                    snprintf(ip2, sizeof ip2, "%s.%d", icpath, ff);
                else
                    snprintf(ip2, sizeof ip2, "%s", icpath);
                ics[ff] = read_ic_file(ip2, &rn[ff]);
            }
        }
        for (int ss = 0; ss < nl; ss++) {
            if (!multi && ss > 0) {
                ics[ss] = ics[0];
                rn[ss] = rn[0];
            }
            roff[ss + 1] = roff[ss] + rn[ss];
        }
        const int64_t n = rn[0]; /* the count of every release unless each reads its own file */
        int64_t nmax = 0;
        for (int ss = 0; ss < nl; ss++) nmax = rn[ss] > nmax ? rn[ss] : nmax;
        fvc_point *ic = ics[0];
        fvc_particles *P = NULL;
        check(fvc_particles_create(c, roff[nl] > 0 ? roff[nl] : 1, &P), c, "fvc_particles_create");
        double *rstart = (double *)malloc(sizeof(double) * nl), *rstop = (double *)malloc(sizeof(double) * nl);
        int *status = (int *)calloc((size_t)nl, sizeof(int));
        for (int ss = 0; ss < nl; ss++) {
            check(fvc_particles_set(P, roff[ss], ics[ss], rn[ss]), c, "fvc_particles_set");
            rstart[ss] = d.Output_TStart + dir * ss * d.Trace_LaunchTimeSpacing; /* tracers.c:1287-1291 */
            rstop[ss] = dir > 0 ? fmin(rstart[ss] + d.Trace_IntTLength, d.Output_TEnd)
                                : fmax(rstart[ss] - d.Trace_IntTLength, d.Output_TEnd);
        }
        int upper = nl >= nproc / 2;
        int64_t *cnt = (int64_t *)calloc((size_t)d.Output_TRes, sizeof(int64_t));
        FILE **of = (FILE **)calloc((size_t)d.Output_TRes, sizeof(FILE *));
        fvc_point *tmp = (fvc_point *)malloc(sizeof(fvc_point) * (size_t)(nmax > 0 ? nmax : 1));
        int outframe = 0;
        double nextout = d.Output_TStart;
        for (int df = d.Data_FirstFrame; df != d.Data_LastFrame; df += dir) {
            int flo = dir > 0 ? df : df - 1;
            int lo = slot_for(&sl, flo, -1);
            int hi = slot_for(&sl, flo + 1, lo);
            double tmin = d.Data_TMin + flo * d.Data_TDelta, tmax = tmin + d.Data_TDelta;
            double t1 = dir > 0 ? tmin : tmax, t2;
            while (((dir > 0) ? (t1 - tmax) : (tmin - t1)) < 0) {
                t2 = dir > 0 ? fmin(nextout, tmax) : fmax(tmin, nextout);
                for (int ss = 0; ss < nl; ss++) {
                    if (status[ss] != 2 && (status[ss] == 1 || (rstart[ss] <= fmax(t1, t2) && rstart[ss] >= fmin(t1, t2)))) {
                        double t0 = now_s();
                        if (status[ss] == 0) {
                            status[ss] = 1;
                            if (cet) check(fvc_cet_launch(c, P, roff[ss], rn[ss]), c, "fvc_cet_launch");
                            t1 = rstart[ss];
                            check(fvc_launch_velocity(c, P, roff[ss], rn[ss], rstart[ss], lo, hi, tmin, tmax), c,
                                  "fvc_launch_velocity");
                        }
                        check(fvc_advect(c, P, roff[ss], rn[ss], t1, t2, lo, hi, tmin, tmax, &ip), c, "fvc_advect");
                        int64_t st = 0;
                        fvc_last_step_count(c, &st);
                        steps_total += st;
                        t_compute += now_s() - t0;
                        if (fabs(nextout - t2) < TINY) {
                            check(fvc_particles_get(P, roff[ss], tmp, rn[ss]), c, "fvc_particles_get");
                            for (int64_t q = 0; q < rn[ss]; q++) {
                                if (tmp[q].LeftDomainTime > -TINY && (!tmp[q].LeftDomain || d.Trace_AlwaysOutput)) {
                                    if (!of[outframe]) {
                                        snprintf(path, sizeof path, "%s%s.%d.bin", d.Path_Output, d.Trace_OutFilePrefix,
                                                 outframe);
                                        of[outframe] = fopen(path, "wb");
                                        if (!of[outframe]) fatal("Could not open file %s", path);
                                        double tt = d.Output_TStart + outframe * dir * d.Output_TDelta;
                                        fwrite(&tt, 8, 1, of[outframe]);
                                    }
                                    fwrite(tmp[q].X, 8, 3, of[outframe]);
                                    if (ap) fwrite(&tmp[q].Scalar, 8, 1, of[outframe]); /* tracers.c:136-138 */
                                    cnt[outframe]++;
                                }
                            }
                        }
                    }
                    if (((dir > 0) ? (rstop[ss] - t2) : (t2 - rstop[ss])) < TINY) status[ss] = 2;
                }
                if (fabs(t2 - nextout) < TINY) {
                    nextout += dir * d.Output_TDelta;
                    outframe++;
                    if (outframe >= d.Output_TRes) break;
                }
                t1 = t2;
            }
        }
        /* OutputTracers_OMP writes every frame file (tracers.c:1700-1761); OutputTracers only non-empty ones */
        for (int i = 0; i < d.Output_TRes; i++) {
            if (!of[i] && upper) {
                snprintf(path, sizeof path, "%s%s.%d.bin", d.Path_Output, d.Trace_OutFilePrefix, i);
                of[i] = fopen(path, "wb");
                if (!of[i]) fatal("Could not open file %s", path);
                double tt = d.Output_TStart + i * dir * d.Output_TDelta;
                fwrite(&tt, 8, 1, of[i]);
            }
            if (of[i]) fclose(of[i]);
        }
        if (rt) { /* OutputRT (residencetime.c:21-185): mesh description, then each release's residence times */
            write_trace_mesh(&d, d.Trace_RTOutFilePrefix, ic, n);
            for (int ss = 0; ss < nl; ss++) {
                check(fvc_particles_get(P, roff[ss], tmp, n), c, "fvc_particles_get");
                snprintf(path, sizeof path, "%s%s.%d.bin", d.Path_Output, d.Trace_RTOutFilePrefix, ss);
                FILE *o = fopen(path, "wb");
                if (!o) fatal("Could not open file %s", path);
                const double t0r = d.Output_TStart + ss * d.Trace_LaunchTimeSpacing;
                fwrite(&t0r, 8, 1, o);
                for (int64_t q = 0; q < n; q++) {
                    double rtv;
                    if (fabs(tmp[q].LeftDomainTime + 1) < TINY)
                        rtv = -1.0;
                    else
                        rtv = (tmp[q].LeftDomain ? tmp[q].LeftDomainTime : rstop[ss]) - t0r;
                    fwrite(&rtv, 8, 1, o);
                }
                fclose(o);
            }
        } else if (cet) { /* OutputCET (exposuretime.c:180-300) */
            const int64_t ne = md.num_elems;
            double *sum = (double *)malloc(sizeof(double) * (size_t)ne), *cv = (double *)calloc((size_t)ne, 8),
                   *mv = (double *)calloc((size_t)ne, 8);
            int32_t *enc = (int32_t *)malloc(sizeof(int32_t) * (size_t)ne);
            check(fvc_cet_get(c, sum, enc), c, "fvc_cet_get");
            for (int64_t e = 0; e < ne; e++) {
                if (!(sum[e] > TINY)) continue;
                double xc;
                if (d.Dimensions == 3) {
                    xc = 1.0; /* pow(vol, (1/3)): the exponent 1/3 is integer division, 0 (exposuretime.c:222) */
                } else {
                    const double *p1 = xyz + 3 * conn[4 * e], *p2 = xyz + 3 * conn[4 * e + 1],
                                 *p3 = xyz + 3 * conn[4 * e + 2];
                    const double A0 = p2[0] - p1[0], A1 = p2[1] - p1[1], B0 = p3[0] - p1[0], B1 = p3[1] - p1[1];
                    xc = sqrt(fabs(A0 * B1 - A1 * B0) / 2);
                }
                cv[e] = sum[e] / xc / (double)n;
                mv[e] = sum[e] / xc / enc[e];
            }
            const double dummy = -1.0;
            const char *suf[3] = {"CET", "MET", "ENC"};
            for (int k = 0; k < 3; k++) {
                snprintf(path, sizeof path, "%s%s_%s.0.bin", d.Path_Output, d.Trace_RTOutFilePrefix, suf[k]);
                FILE *o = fopen(path, "wb");
                if (!o) fatal("Could not open %s", path);
                fwrite(&dummy, 8, 1, o);
                if (k == 0) fwrite(cv, 8, (size_t)ne, o);
                if (k == 1) fwrite(mv, 8, (size_t)ne, o);
                if (k == 2) fwrite(enc, 4, (size_t)ne, o);
                fclose(o);
            }
            free(sum);
            free(cv);
            free(mv);
            free(enc);
        }
        if (ap) { /* OutputAP (strainrate.c:248-395): mesh description, then each release's Scalar */
            write_trace_mesh(&d, d.Trace_OutFilePrefix, ic, n);
            for (int ss = 0; ss < nl; ss++) {
                check(fvc_particles_get(P, roff[ss], tmp, n), c, "fvc_particles_get");
                snprintf(path, sizeof path, "%s%s_AP.%d.bin", d.Path_Output, d.Trace_OutFilePrefix, ss);
                FILE *o = fopen(path, "wb");
                if (!o) fatal("Could not open file %s", path);
                double tt = d.Output_TStart + ss * d.Trace_LaunchTimeSpacing;
                fwrite(&tt, 8, 1, o);
                for (int64_t q = 0; q < n; q++) fwrite(&tmp[q].Scalar, 8, 1, o);
                fclose(o);
            }
        }
        free(tmp);
        free(cnt);
        free(of);
        free(rstart);
        free(rstop);
        free(status);
        for (int ss = 0; ss < (multi ? nl : 1); ss++) free(ics[ss]);
        free(ics);
        free(rn);
        free(roff);
        fvc_particles_destroy(P);
    }
    if (!quiet)
        printf("Simulation complete: %.3f s total, %.3f s on the GPU path, %.3f s velocity I/O, %lld particle-steps\n",
               now_s() - t_start, t_compute, sl.io_s, (long long)steps_total);
    fvc_destroy(c);
    free(sl.slot_file);
    free(sl.buf);
    free(sl.srbuf);
    free(sl.wall_ids);
    free(sl.wall_v);
    free(xyz);
    free(conn);
    free(adj);
    return 0;
}
