/*
 * oc_driver.c - oracle drivers: FTLE slides, seed location chains, tracer
 * releases and the deck-driven frame loop.  TEST INFRASTRUCTURE ONLY.
 *
 * Semantics restated from ompVC/ftle.c, tracers.c, main.c and parameters.c
 * (line references at each function).  Within a slide the point loop runs in
 * the reference's sequential order (i, j, k with k fastest), which fixes the
 * order-dependent early-FTLE results (ftle.c:109-173).
 */
#include "fvc_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TINY OC_TINY
#define MIN2(a, b) ((a) < (b) ? (a) : (b))
#define MAX2(a, b) ((a) > (b) ? (a) : (b))

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* FTLE                                                                      */
/* ------------------------------------------------------------------------ */
#define FIDX(f, i, j, k) (((int64_t)(i) * (f)->res[1] + (j)) * (f)->res[2] + (k))

static void mask_outside(oc_ftle *f, int i, int j, int k) {
    oc_ftle_pt *q = &f->pts[FIDX(f, i, j, k)];
    q->Pt.LeftDomain = 1;
    q->Pt.LeftDomainTime = -1.0;
    q->FTLEwT = -1.0;
    q->FTLEwoT = -1.0;
    q->HaveFTLE = 1;
    if (i > 0) f->pts[FIDX(f, i - 1, j, k)].HaveFTLE = 1;
    if (i < f->res[0] - 1) f->pts[FIDX(f, i + 1, j, k)].HaveFTLE = 1;
    if (j > 0) f->pts[FIDX(f, i, j - 1, k)].HaveFTLE = 1;
    if (j < f->res[1] - 1) f->pts[FIDX(f, i, j + 1, k)].HaveFTLE = 1;
    if (k > 0) f->pts[FIDX(f, i, j, k - 1)].HaveFTLE = 1;
    if (k < f->res[2] - 1) f->pts[FIDX(f, i, j, k + 1)].HaveFTLE = 1;
}

/* ftle.c:383-609 (FTLE_GenerateMesh = 1) */
oc_ftle *oc_ftle_create(const oc_mesh *m, const double min[3], const double max[3], const int32_t res[3],
                        int local_search_checking, int *err) {
    oc_ftle *f = (oc_ftle *)calloc(1, sizeof *f);
    if (!f) return NULL;
    *err = OC_OK;
    for (int a = 0; a < 3; a++) {
        f->min[a] = min[a];
        f->max[a] = max[a];
        f->res[a] = res[a];
    }
    /* parameters.c:310-320 */
    f->del[0] = (f->max[0] - f->min[0]) / (f->res[0] - 1);
    f->del[1] = (f->max[1] - f->min[1]) / (f->res[1] - 1);
    if (m->dims == 3)
        f->del[2] = (f->max[2] - f->min[2]) / (f->res[2] - 1);
    else {
        f->min[2] = f->max[2] = 0;
        f->del[2] = 0;
        f->res[2] = 1;
    }
    f->n = (int64_t)f->res[0] * f->res[1] * f->res[2];
    f->pts = (oc_ftle_pt *)calloc((size_t)f->n, sizeof(oc_ftle_pt));
    f->local_search_checking = local_search_checking;
    for (int i = 0; i < f->res[0]; i++)
        for (int j = 0; j < f->res[1]; j++)
            for (int k = 0; k < f->res[2]; k++) {
                oc_ftle_pt *q = &f->pts[FIDX(f, i, j, k)];
                q->Pt.X[0] = f->min[0] + i * f->del[0];
                q->Pt.X[1] = f->min[1] + j * f->del[1];
                q->Pt.X[2] = f->min[2] + k * f->del[2];
            }
    for (int i = 0; i < f->res[0]; i++)
        for (int j = 0; j < f->res[1]; j++)
            for (int k = 0; k < f->res[2]; k++)
                if (oc_test_outside_domain(m, f->pts[FIDX(f, i, j, k)].Pt.X)) mask_outside(f, i, j, k);
    if (m->mesh_type == 1) {
        int seed = -1, guess = -1, jguess = -1;
        double xs[3] = {f->min[0] + ((int)(f->res[0] / 2)) * f->del[0], f->min[1] + ((int)(f->res[1] / 2)) * f->del[1],
                        f->min[2] + ((int)(f->res[2] / 2)) * f->del[2]};
        seed = oc_global_search(m, xs);
        if (seed >= 0) guess = seed;
        for (int i = 0; i < f->res[0]; i++) {
            for (int j = 0; j < f->res[1]; j++) {
                jguess = -1;
                for (int k = 0; k < f->res[2]; k++) {
                    oc_ftle_pt *q = &f->pts[FIDX(f, i, j, k)];
                    if (q->Pt.LeftDomain) continue;
                    if (seed < 0) {
                        seed = oc_global_search(m, q->Pt.X);
                        if (seed < 0)
                            mask_outside(f, i, j, k);
                        else {
                            q->Pt.ElementIndex = seed;
                            guess = jguess = seed;
                        }
                        continue;
                    }
                    int e2 = OC_OK;
                    int index = oc_local_search(m, q->Pt.X, guess, &e2);
                    if (index < 0 && !e2) index = oc_local_search(m, q->Pt.X, seed, &e2);
                    if (index < 0 && !e2 && local_search_checking) index = oc_global_search(m, q->Pt.X);
                    if (e2) {
                        *err = e2;
                        return f;
                    }
                    if (index < 0)
                        mask_outside(f, i, j, k);
                    else {
                        q->Pt.ElementIndex = index;
                        guess = index;
                        if (jguess < 0) jguess = index;
                    }
                }
            }
            /* after each plane, the first found element of its last row (ftle.c:597-598) */
            if (jguess >= 0) guess = jguess;
        }
        /* the plane guess of ftle.c:600-601 is applied after the whole grid, when no point follows */
    }
    return f;
}

const oc_ftle_pt *oc_ftle_points(const oc_ftle *f) { return f->pts; }

void oc_ftle_free(oc_ftle *f) {
    if (!f) return;
    free(f->pts);
    free(f);
}

/* Advect_FTLEPoint (ftle.c:1205-1227): copies X, LeftDomain, ElementIndex */
static oc_point advect_copy(const oc_mesh *m, const oc_int *ip, const oc_ftle *f, int64_t idx, double t1, double t2,
                            int *err) {
    oc_point tp;
    memset(&tp, 0, sizeof tp);
    const oc_point *src = &f->pts[idx].Pt;
    tp.X[0] = src->X[0];
    tp.X[1] = src->X[1];
    tp.X[2] = src->X[2];
    tp.LeftDomain = src->LeftDomain;
    tp.ElementIndex = src->ElementIndex;
    oc_advect(m, ip, &tp, t1, t2, err);
    return tp;
}

/* Cauchy-Green largest eigenvalue -> FTLE (ftle.c:1174-1202) from a gradient */
static void ftle_from_gradient(double D[3][3], double *lam_out, int *err) {
    double C[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            C[a][b] = 0;
            for (int c = 0; c < 3; c++) C[a][b] += D[c][a] * D[c][b];
        }
    *lam_out = oc_max_eig_sym3((const double(*)[3])C, err);
}

/* ftle.c:788-1048 */
static int ftle_early(const oc_mesh *m, const oc_int *ip, oc_ftle *f, int i, int j, int k, double t1, double te) {
    int err = OC_OK;
    double lt[6];
    int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
    int have[6] = {i > 0, i < f->res[0] - 1, j > 0, j < f->res[1] - 1, k > 0, k < f->res[2] - 1};
    for (int q = 0; q < 6; q++) {
        lt[q] = te;
        if (!have[q]) continue;
        oc_point tp = advect_copy(m, ip, f, FIDX(f, nb[q][0], nb[q][1], nb[q][2]), t1, te, &err);
        if (err) return err;
        if (tp.LeftDomain) lt[q] = tp.LeftDomainTime;
    }
    double mt = MIN2(te, MIN2(lt[5], MIN2(lt[4], MIN2(lt[3], MIN2(lt[2], MIN2(lt[1], lt[0]))))));
    oc_point self = advect_copy(m, ip, f, FIDX(f, i, j, k), t1, mt, &err);
    if (err) return err;
    double lo[3][3], hi[3][3], fac[3];
    int hasax[3] = {1, 1, m->dims == 3};
    for (int ax = 0; ax < 3; ax++) {
        if (!hasax[ax]) continue;
        int c = ax == 0 ? i : (ax == 1 ? j : k);
        int last = f->res[ax] - 1;
        int lo_n = 2 * ax, hi_n = 2 * ax + 1;
        if (c == 0) {
            memcpy(lo[ax], self.X, sizeof self.X);
            oc_point tp = advect_copy(m, ip, f, FIDX(f, nb[hi_n][0], nb[hi_n][1], nb[hi_n][2]), t1, mt, &err);
            memcpy(hi[ax], tp.X, sizeof tp.X);
            fac[ax] = 1.0;
        } else if (c == last) {
            oc_point tp = advect_copy(m, ip, f, FIDX(f, nb[lo_n][0], nb[lo_n][1], nb[lo_n][2]), t1, mt, &err);
            memcpy(lo[ax], tp.X, sizeof tp.X);
            memcpy(hi[ax], self.X, sizeof self.X);
            fac[ax] = 1.0;
        } else {
            oc_point tp = advect_copy(m, ip, f, FIDX(f, nb[lo_n][0], nb[lo_n][1], nb[lo_n][2]), t1, mt, &err);
            memcpy(lo[ax], tp.X, sizeof tp.X);
            if (err) return err;
            tp = advect_copy(m, ip, f, FIDX(f, nb[hi_n][0], nb[hi_n][1], nb[hi_n][2]), t1, mt, &err);
            memcpy(hi[ax], tp.X, sizeof tp.X);
            fac[ax] = 2.0;
        }
        if (err) return err;
    }
    double D[3][3] = {{0}};
    if (m->dims == 3) {
        D[0][2] = (hi[2][0] - lo[2][0]) / (fac[2] * f->del[2]);
        D[1][2] = (hi[2][1] - lo[2][1]) / (fac[2] * f->del[2]);
        D[2][0] = (hi[0][2] - lo[0][2]) / (fac[0] * f->del[0]);
        D[2][1] = (hi[1][2] - lo[1][2]) / (fac[1] * f->del[1]);
        D[2][2] = (hi[2][2] - lo[2][2]) / (fac[2] * f->del[2]);
    }
    D[0][0] = (hi[0][0] - lo[0][0]) / (fac[0] * f->del[0]);
    D[0][1] = (hi[1][0] - lo[1][0]) / (fac[1] * f->del[1]);
    D[1][0] = (hi[0][1] - lo[0][1]) / (fac[0] * f->del[0]);
    D[1][1] = (hi[1][1] - lo[1][1]) / (fac[1] * f->del[1]);
    double lam;
    ftle_from_gradient(D, &lam, &err);
    oc_ftle_pt *q = &f->pts[FIDX(f, i, j, k)];
    if (lam > 1.0 - TINY) {
        if (fabs(mt - f->start_time) < TINY) {
            q->FTLEwT = 0.0;
            q->FTLEwoT = 0.0;
        } else {
            q->FTLEwT = (0.5 * log(lam)) / fabs(mt - f->start_time);
            q->FTLEwoT = 0.5 * log(lam);
        }
    } else {
        q->FTLEwT = 0.0;
        q->FTLEwoT = 0.0;
    }
    q->HaveFTLE = 1;
    return err;
}

static double ftle_acc_time = 0.0;
double oc_ftle_time(const oc_ftle *f) {
    (void)f;
    return ftle_acc_time;
}

/* ftle.c:68-214 for one slide (UPPERLOOP order) */
int oc_ftle_frame(const oc_mesh *m, const oc_int *ip, oc_ftle *f, int dir, double lmin, double lmax) {
    double t0c = now_s();
    if (f->status == 2) return OC_OK;
    if (!(f->status == 1 || (f->start_time <= lmax + TINY && f->start_time >= lmin - TINY))) return OC_OK;
    double t1, t2;
    int rc = OC_OK;
    if (f->status == 0) {
        f->status = 1;
        t1 = f->start_time;
        for (int64_t q = 0; q < f->n; q++)
            if (!f->pts[q].Pt.LeftDomain) {
                rc = oc_velocity(m, f->start_time, &f->pts[q].Pt, f->pts[q].Pt.V);
                if (rc) return rc;
            }
    } else {
        t1 = dir > 0 ? lmin : lmax;
    }
    t2 = dir > 0 ? MIN2(lmax, f->stop_time) : MAX2(lmin, f->stop_time);
    double *newx = (double *)malloc(sizeof(double) * 3 * (size_t)f->n);
    f->n_integrated = 0;
    for (int i = 0; i < f->res[0]; i++)
        for (int j = 0; j < f->res[1]; j++)
            for (int k = 0; k < f->res[2]; k++) {
                int64_t idx = FIDX(f, i, j, k);
                oc_ftle_pt *q = &f->pts[idx];
                if (!(q->Pt.LeftDomainTime > -TINY && (!q->Pt.LeftDomain || ip->extrapolate))) continue;
                oc_point tp;
                memset(&tp, 0, sizeof tp);
                memcpy(tp.X, q->Pt.X, sizeof tp.X);
                tp.LeftDomain = q->Pt.LeftDomain;
                tp.LeftDomainTime = q->Pt.LeftDomainTime;
                memcpy(tp.V, q->Pt.V, sizeof tp.V);
                tp.ElementIndex = q->Pt.ElementIndex;
                int err = OC_OK;
                oc_advect(m, ip, &tp, t1, t2, &err);
                if (err) {
                    free(newx);
                    return err;
                }
                f->n_integrated++;
                if (tp.LeftDomain && !ip->extrapolate) {
                    double te = tp.LeftDomainTime;
                    if (!q->HaveFTLE && (rc = ftle_early(m, ip, f, i, j, k, t1, te))) goto out;
                    if (i > 0 && !f->pts[FIDX(f, i - 1, j, k)].HaveFTLE && (rc = ftle_early(m, ip, f, i - 1, j, k, t1, te))) goto out;
                    if (i < f->res[0] - 1 && !f->pts[FIDX(f, i + 1, j, k)].HaveFTLE && (rc = ftle_early(m, ip, f, i + 1, j, k, t1, te))) goto out;
                    if (j > 0 && !f->pts[FIDX(f, i, j - 1, k)].HaveFTLE && (rc = ftle_early(m, ip, f, i, j - 1, k, t1, te))) goto out;
                    if (j < f->res[1] - 1 && !f->pts[FIDX(f, i, j + 1, k)].HaveFTLE && (rc = ftle_early(m, ip, f, i, j + 1, k, t1, te))) goto out;
                    if (k > 0 && !f->pts[FIDX(f, i, j, k - 1)].HaveFTLE && (rc = ftle_early(m, ip, f, i, j, k - 1, t1, te))) goto out;
                    if (k < f->res[2] - 1 && !f->pts[FIDX(f, i, j, k + 1)].HaveFTLE && (rc = ftle_early(m, ip, f, i, j, k + 1, t1, te))) goto out;
                    q->Pt.LeftDomain = 1;
                } else {
                    memcpy(newx + 3 * idx, tp.X, sizeof tp.X);
                    memcpy(q->Pt.V, tp.V, sizeof tp.V);
                    if (ip->extrapolate) q->Pt.LeftDomain = tp.LeftDomain;
                    if (m->mesh_type == 1) q->Pt.ElementIndex = tp.ElementIndex;
                }
            }
    /* UpdateFTLELocations (ftle.c:1229-1242) */
    for (int64_t idx = 0; idx < f->n; idx++)
        if (!f->pts[idx].Pt.LeftDomain || ip->extrapolate) memcpy(f->pts[idx].Pt.X, newx + 3 * idx, 3 * sizeof(double));
    if (((dir > 0) ? (f->stop_time - lmax) : (lmin - f->stop_time)) < TINY) f->status = 2;
out:
    free(newx);
    ftle_acc_time += now_s() - t0c;
    return rc;
}

/* GetFTLEForPoint (ftle.c:1050-1203) for points without a value */
int oc_ftle_fill(oc_ftle *f, int dims, double int_t) {
    int rc = OC_OK;
    for (int k = 0; k < f->res[2]; k++)
        for (int j = 0; j < f->res[1]; j++)
            for (int i = 0; i < f->res[0]; i++) {
                oc_ftle_pt *q = &f->pts[FIDX(f, i, j, k)];
                if (q->HaveFTLE) continue;
                if (int_t < TINY) {
                    q->FTLEwT = 0.0;
                    q->FTLEwoT = 0.0;
                    continue;
                }
                int c[3] = {i, j, k};
                const double *lo[3], *hi[3];
                double fac[3] = {1, 1, 1};
                for (int ax = 0; ax < 3; ax++) {
                    if (ax == 2 && dims != 3) continue;
                    int cl[3] = {i, j, k}, ch[3] = {i, j, k};
                    if (c[ax] == 0) {
                        ch[ax] = c[ax] + 1;
                    } else if (c[ax] == f->res[ax] - 1) {
                        cl[ax] = c[ax] - 1;
                    } else {
                        cl[ax] = c[ax] - 1;
                        ch[ax] = c[ax] + 1;
                        fac[ax] = 2.0;
                    }
                    lo[ax] = f->pts[FIDX(f, cl[0], cl[1], cl[2])].Pt.X;
                    hi[ax] = f->pts[FIDX(f, ch[0], ch[1], ch[2])].Pt.X;
                }
                double D[3][3] = {{0}};
                if (dims == 3) {
                    D[0][2] = (hi[2][0] - lo[2][0]) / (fac[2] * f->del[2]);
                    D[1][2] = (hi[2][1] - lo[2][1]) / (fac[2] * f->del[2]);
                    D[2][0] = (hi[0][2] - lo[0][2]) / (fac[0] * f->del[0]);
                    D[2][1] = (hi[1][2] - lo[1][2]) / (fac[1] * f->del[1]);
                    D[2][2] = (hi[2][2] - lo[2][2]) / (fac[2] * f->del[2]);
                }
                D[0][0] = (hi[0][0] - lo[0][0]) / (fac[0] * f->del[0]);
                D[0][1] = (hi[1][0] - lo[1][0]) / (fac[1] * f->del[1]);
                D[1][0] = (hi[0][1] - lo[0][1]) / (fac[0] * f->del[0]);
                D[1][1] = (hi[1][1] - lo[1][1]) / (fac[1] * f->del[1]);
                double lam;
                ftle_from_gradient(D, &lam, &rc);
                if (lam > 1.0 - TINY) {
                    q->FTLEwT = (0.5 * log(lam)) / int_t;
                    q->FTLEwoT = 0.5 * log(lam);
                } else {
                    q->FTLEwT = 0.0;
                    q->FTLEwoT = 0.0;
                }
            }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* tracer seeds                                                              */
/* ------------------------------------------------------------------------ */
/* ReadInTraceMesh (tracers.c:1111-1251), no auxiliary mesh */
int oc_seed_list(const oc_mesh *m, int64_t n, const double *xyz, int lsc, oc_point *pts) {
    int found = 0, guess = -1;
    for (int64_t q = 0; q < n; q++) {
        oc_point *p = &pts[q];
        memset(p, 0, sizeof *p);
        memcpy(p->X, xyz + 3 * q, 3 * sizeof(double));
        if (m->dims == 2) p->X[2] = 0;
        p->LeftDomain = oc_test_outside_domain(m, p->X);
        if (m->mesh_type == 1 && !p->LeftDomain) {
            int index, err = OC_OK;
            if (!found) {
                index = oc_global_search(m, p->X);
                if (index < 0) {
                    p->LeftDomain = 1;
                    p->ElementIndex = -1;
                } else {
                    p->ElementIndex = index;
                    guess = index;
                    found++;
                }
            } else {
                index = oc_local_search(m, p->X, guess, &err);
                if (err) return err;
                if (index < 0) {
                    if (lsc) {
                        index = oc_global_search(m, p->X);
                        if (index < 0) {
                            p->LeftDomain = 1;
                            p->ElementIndex = -1;
                        } else {
                            p->ElementIndex = index;
                            p->LeftDomain = 0;
                            guess = index;
                            found++;
                        }
                    } else {
                        p->LeftDomain = 1;
                        p->ElementIndex = -1;
                    }
                } else {
                    p->ElementIndex = index;
                    p->AuxElementIndex = index;
                    p->LeftDomain = 0;
                    guess = index;
                    found++;
                }
            }
        }
        p->LeftDomainTime = p->LeftDomain ? -1.0 : 0.0;
    }
    return OC_OK;
}

/* GenerateTracerMesh (tracers.c:1298-1532), no auxiliary mesh.  For Cartesian
 * data the reference leaves LeftDomainTime uninitialised; fresh allocations
 * read as 0, which is what is written here. */
int oc_seed_grid(const oc_mesh *m, const double mn[3], const double del[3], const int32_t res[3], int lsc,
                 oc_point *pts) {
    int seed = -1, guess = -1;
    if (m->mesh_type == 1) {
        double xs[3] = {mn[0] + ((int)(res[0] / 2)) * del[0], mn[1] + ((int)(res[1] / 2)) * del[1],
                        mn[2] + ((int)(res[2] / 2)) * del[2]};
        seed = oc_global_search(m, xs);
        if (seed >= 0) guess = seed;
    }
    int64_t count = 0;
    for (int kk = 0; kk < res[2]; kk++)
        for (int jj = 0; jj < res[1]; jj++)
            for (int ii = 0; ii < res[0]; ii++) {
                oc_point *p = &pts[count++];
                memset(p, 0, sizeof *p);
                p->X[0] = mn[0] + ii * del[0];
                p->X[1] = mn[1] + jj * del[1];
                p->X[2] = mn[2] + kk * del[2];
                p->LeftDomain = oc_test_outside_domain(m, p->X);
                if (p->LeftDomain || m->mesh_type != 1) continue;
                if (seed < 0) {
                    seed = oc_global_search(m, p->X);
                    if (seed < 0) {
                        p->LeftDomain = 1;
                        p->LeftDomainTime = -1.0;
                        p->ElementIndex = -1;
                        p->AuxElementIndex = -1;
                    } else {
                        p->ElementIndex = seed;
                        guess = seed;
                    }
                    continue;
                }
                int err = OC_OK;
                int index = oc_local_search(m, p->X, guess, &err);
                if (index < 0 && !err) index = oc_local_search(m, p->X, seed, &err);
                if (index < 0 && !err && lsc) index = oc_global_search(m, p->X);
                if (err) return err;
                if (index < 0) {
                    p->LeftDomain = 1;
                    p->LeftDomainTime = -1.0;
                    p->ElementIndex = -1;
                } else {
                    p->ElementIndex = index;
                    p->LeftDomain = 0;
                    p->LeftDomainTime = 0.0;
                    guess = index;
                }
            }
    return OC_OK;
}

/* ------------------------------------------------------------------------ */
/* deck                                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    char path_data[256], path_out[256], prefix[256];
    int dims, mesh_type, suffix_min, suffix_delta, tres, periodic;
    double tdelta, tmin, bounds[6];
    double out_tstart, out_tdelta;
    int out_tres;
    int int_type, direction, extrapolate;
    double dt, accuracy, min_step, max_step;
    double particle_radius;
    int lsc;
    int ftle, ftle_gen, ftle_res[3], ftle_var, ftle_varfreq;
    double ftle_b[6], ftle_len;
    char ftle_prefix[256], ftle_ic[256];
    int trace, trace_strategy, trace_gen, trace_multi, trace_fmt, trace_nl, trace_always, trace_res[3], trace_ap;
    double trace_spacing, trace_len, trace_b[6];
    char trace_in[256], trace_prefix[256];
    int trace_cet, trace_rt;
    int velout;
} deck_t;

static int deck_tokens(FILE *f, char *buf, int len, int comments) {
    int c;
    for (;;) {
        c = fgetc(f);
        if (c == EOF) return 0;
        if (isspace(c)) continue;
        if (c == '#' && comments) {
            while (c != '\n' && c != EOF) c = fgetc(f);
            continue;
        }
        break;
    }
    int n = 0;
    while (c != EOF && !isspace(c)) {
        if (n < len - 1) buf[n++] = (char)c;
        c = fgetc(f);
    }
    buf[n] = 0;
    return 1;
}

static int read_deck(const char *path, deck_t *d, char *msg, int msglen) {
    FILE *f = fopen(path, "r");
    if (!f) {
        snprintf(msg, msglen, "Could not open input file %s", path);
        return OC_E_IO;
    }
    char vals[97][256], name[256], eq[256];
    int k = 0;
    while (k < 97 && deck_tokens(f, name, sizeof name, 1)) {
        if (k == 38 && strcmp(name, "FTLE_Compute") == 0) { /* 96-entry deck */
            strcpy(vals[k++], "0");
        }
        if (!deck_tokens(f, eq, sizeof eq, 0) || !deck_tokens(f, vals[k], sizeof vals[k], 0)) break;
        k++;
    }
    fclose(f);
    if (k < 97) {
        snprintf(msg, msglen, "deck %s has %d entries", path, k);
        return OC_E_PARAM;
    }
#define S(i) vals[i]
#define I(i) ((int)strtol(vals[i], NULL, 10))
#define D(i) strtod(vals[i], NULL)
    memset(d, 0, sizeof *d);
    strcpy(d->path_data, S(0));
    strcpy(d->path_out, S(1));
    d->dims = I(2);
    d->mesh_type = I(3);
    strcpy(d->prefix, S(4));
    d->suffix_min = I(5);
    d->suffix_delta = I(6);
    d->tres = I(7);
    d->tdelta = D(8);
    d->tmin = D(9);
    d->periodic = I(10);
    for (int a = 0; a < 6; a++) d->bounds[a] = D(11 + a);
    d->out_tstart = D(19);
    d->out_tres = I(20);
    d->out_tdelta = D(21);
    d->int_type = I(22);
    d->dt = D(23);
    d->accuracy = D(24);
    d->min_step = D(25);
    d->max_step = D(26);
    d->direction = I(27);
    d->extrapolate = I(30);
    d->particle_radius = D(31);
    d->lsc = I(37);
    d->ftle = I(39);
    d->ftle_gen = I(40);
    strcpy(d->ftle_ic, S(41));
    for (int a = 0; a < 6; a++) d->ftle_b[a] = D(42 + a);
    for (int a = 0; a < 3; a++) d->ftle_res[a] = I(48 + a);
    d->ftle_len = D(51);
    d->ftle_var = I(52);
    d->ftle_varfreq = I(53);
    strcpy(d->ftle_prefix, S(54));
    d->trace = I(55);
    d->trace_strategy = I(56);
    d->trace_gen = I(58);
    strcpy(d->trace_in, S(59));
    d->trace_multi = I(60);
    d->trace_fmt = I(61);
    strcpy(d->trace_prefix, S(62));
    d->trace_nl = I(63);
    d->trace_spacing = D(64);
    d->trace_len = D(65);
    d->trace_always = I(66);
    for (int a = 0; a < 6; a++) d->trace_b[a] = D(67 + a);
    for (int a = 0; a < 3; a++) d->trace_res[a] = I(73 + a);
    d->trace_ap = I(76);
    d->trace_cet = I(77);
    d->trace_rt = I(81);
    d->velout = I(83);
#undef S
#undef I
#undef D
    for (int p = 0; p < 2; p++) {
        char *s = p ? d->path_out : d->path_data;
        if (strcmp(s, "pwd") == 0)
            strcpy(s, "./");
        else if (s[strlen(s) - 1] != '/')
            strcat(s, "/");
    }
    return OC_OK;
}

static int load_slice(const deck_t *d, int frame, int64_t n, double *uvw) {
    int mod = (int)fmod(frame, d->tres - 1);
    if (mod < 0) mod += d->tres - 1;
    char path[600];
    snprintf(path, sizeof path, "%s%s_vel.%d.bin", d->path_data, d->prefix, d->suffix_min + mod * d->suffix_delta);
    FILE *f = fopen(path, "rb");
    if (!f) return OC_E_IO;
    double ts;
    int ok = fread(&ts, 8, 1, f) == 1 && fread(uvw, 8, (size_t)(3 * n), f) == (size_t)(3 * n);
    fclose(f);
    return ok ? OC_OK : OC_E_IO;
}

static oc_mesh *load_mesh(const deck_t *d) {
    char path[600];
    if (d->mesh_type == 0) {
        snprintf(path, sizeof path, "%s%s_Cartesian.bin", d->path_data, d->prefix);
        FILE *f = fopen(path, "rb");
        if (!f) return NULL;
        double mn[3], mx[3];
        int32_t res[3];
        int ok = 1;
        for (int a = 0; a < 3; a++)
            ok &= fread(&mn[a], 8, 1, f) == 1 && fread(&mx[a], 8, 1, f) == 1 && fread(&res[a], 4, 1, f) == 1;
        fclose(f);
        if (!ok) return NULL;
        return oc_mesh_cartesian(d->dims, d->bounds, mn, mx, res);
    }
    int32_t nn, ne, ne2;
    snprintf(path, sizeof path, "%s%s_coordinates.bin", d->path_data, d->prefix);
    FILE *f = fopen(path, "rb");
    if (!f || fread(&nn, 4, 1, f) != 1) return NULL;
    double *xyz = (double *)malloc(sizeof(double) * 3 * (size_t)nn);
    if (fread(xyz, 8, (size_t)3 * nn, f) != (size_t)3 * nn) return NULL;
    fclose(f);
    snprintf(path, sizeof path, "%s%s_connectivity.bin", d->path_data, d->prefix);
    f = fopen(path, "rb");
    if (!f || fread(&ne, 4, 1, f) != 1) return NULL;
    int32_t *conn = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)ne);
    int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)ne);
    if (fread(conn, 4, (size_t)4 * ne, f) != (size_t)4 * ne) return NULL;
    fclose(f);
    snprintf(path, sizeof path, "%s%s_adjacency.bin", d->path_data, d->prefix);
    f = fopen(path, "rb");
    if (!f || fread(&ne2, 4, 1, f) != 1 || ne2 != ne) return NULL;
    if (fread(adj, 4, (size_t)4 * ne, f) != (size_t)4 * ne) return NULL;
    fclose(f);
    oc_mesh *m = oc_mesh_unstructured(d->dims, d->bounds, nn, xyz, ne, conn, adj);
    free(xyz);
    free(conn);
    free(adj);
    return m;
}

static int write_ftle_outputs(const deck_t *d, oc_ftle *f, int ss, int outnum, int first) {
    char path[700];
    FILE *o;
    if (first) {
        snprintf(path, sizeof path, "%s%s_Cartesian.bin", d->path_out, d->ftle_prefix);
        if (!(o = fopen(path, "wb"))) return OC_E_IO;
        for (int a = 0; a < 3; a++) {
            fwrite(&f->min[a], 8, 1, o);
            fwrite(&f->max[a], 8, 1, o);
            fwrite(&f->res[a], 4, 1, o);
        }
        fclose(o);
    }
    (void)ss;
    for (int which = 0; which < 2; which++) {
        snprintf(path, sizeof path, "%s%s%s.%d.bin", d->path_out, d->ftle_prefix, which ? "_noT" : "", outnum);
        if (!(o = fopen(path, "wb"))) return OC_E_IO;
        fwrite(&f->start_time, 8, 1, o);
        for (int k = 0; k < f->res[2]; k++)
            for (int j = 0; j < f->res[1]; j++)
                for (int i = 0; i < f->res[0]; i++) {
                    const oc_ftle_pt *q = &f->pts[FIDX(f, i, j, k)];
                    fwrite(which ? &q->FTLEwoT : &q->FTLEwT, 8, 1, o);
                }
        fclose(o);
    }
    return OC_OK;
}

/* frame bookkeeping of parameters.c:301-341 */
static void frames(const deck_t *d, int *first, int *last) {
    int dir = d->direction;
    double tmax = d->tmin + (d->tres - 1) * d->tdelta;
    double tend = d->out_tstart + (d->out_tres - 1) * dir * d->out_tdelta;
    if (d->ftle) {
        double a = d->out_tstart, b = tend + dir * d->ftle_len;
        double lo = MIN2(a, b), hi = MAX2(a, b);
        if (lo < d->tmin && !d->periodic) lo = d->tmin;
        if (hi > tmax && !d->periodic) hi = tmax;
        *first = (int)(dir > 0 ? floor((lo - d->tmin) / d->tdelta) : ceil((hi - d->tmin) / d->tdelta));
        *last = (int)(dir > 0 ? ceil((hi - d->tmin) / d->tdelta) : floor((lo - d->tmin) / d->tdelta));
    } else {
        *first = (int)(dir > 0 ? floor((d->out_tstart - d->tmin) / d->tdelta) : ceil((d->out_tstart - d->tmin) / d->tdelta));
        *last = (int)(dir > 0 ? ceil((tend - d->tmin) / d->tdelta) : floor((tend - d->tmin) / d->tdelta));
    }
}

static int read_seed_file(const deck_t *d, int64_t *n, double **xyz) {
    char path[600];
    snprintf(path, sizeof path, "%s%s", d->path_data, d->trace_in);
    if (d->trace_fmt == 4) {
        FILE *f = fopen(path, "rb");
        int32_t cnt;
        if (!f || fread(&cnt, 4, 1, f) != 1) return OC_E_IO;
        *n = cnt;
        *xyz = (double *)calloc((size_t)3 * cnt, sizeof(double));
        for (int q = 0; q < cnt; q++)
            if (fread(*xyz + 3 * q, 8, (size_t)d->dims, f) != (size_t)d->dims) return OC_E_IO;
        fclose(f);
        return OC_OK;
    }
    if (d->trace_fmt == 1) {
        FILE *f = fopen(path, "r");
        int cnt;
        if (!f || fscanf(f, "%d", &cnt) != 1) return OC_E_IO;
        *n = cnt;
        *xyz = (double *)calloc((size_t)3 * cnt, sizeof(double));
        for (int q = 0; q < cnt; q++)
            for (int a = 0; a < d->dims; a++)
                if (fscanf(f, "%lf", *xyz + 3 * q + a) != 1) return OC_E_IO;
        fclose(f);
        return OC_OK;
    }
    return OC_E_PARAM;
}

typedef struct {
    double start, stop;
    int status;
    oc_point *pts;
} release_t;

int oc_run_deck(const char *deck_path, int nproc, char *msg, int msglen) {
    deck_t d;
    int rc = read_deck(deck_path, &d, msg, msglen);
    if (rc) return rc;
    if (d.particle_radius > TINY || d.trace_strategy == 1 || d.trace_cet || d.trace_rt ||
        d.trace_ap || d.velout || d.ftle_var) {
        snprintf(msg, msglen, "deck uses a mode outside the advection path");
        return OC_E_PARAM;
    }
    if (nproc <= 0) nproc = (int)sysconf(_SC_NPROCESSORS_ONLN);
    oc_mesh *m = load_mesh(&d);
    if (!m) {
        snprintf(msg, msglen, "could not load mesh");
        return OC_E_IO;
    }
    oc_int ip = {d.int_type == 1 || d.int_type == 2 ? d.int_type : 0, d.dt, d.extrapolate, d.min_step, d.max_step,
                 d.accuracy};
    int dir = d.direction;
    int first, last;
    frames(&d, &first, &last);
    double *uvw = (double *)malloc(sizeof(double) * 3 * (size_t)m->nvel);
    if ((rc = load_slice(&d, first, m->nvel, uvw))) goto done;
    oc_mesh_push_frame(m, dir, uvw);

    if (d.ftle) {
        oc_ftle **sl = (oc_ftle **)calloc((size_t)d.out_tres, sizeof(oc_ftle *));
        double tmax = d.tmin + (d.tres - 1) * d.tdelta;
        int32_t res[3] = {d.ftle_res[0], d.ftle_res[1], d.ftle_res[2]};
        double mn[3] = {d.ftle_b[0], d.ftle_b[2], d.ftle_b[4]}, mx[3] = {d.ftle_b[1], d.ftle_b[3], d.ftle_b[5]};
        for (int ss = 0; ss < d.out_tres; ss++) {
            sl[ss] = oc_ftle_create(m, mn, mx, res, d.lsc, &rc);
            if (rc) goto done;
            sl[ss]->start_time = d.out_tstart + ss * dir * d.out_tdelta;
            if (!d.periodic) {
                double st = MIN2(d.out_tstart + ss * dir * d.out_tdelta + (double)dir * d.ftle_len, tmax);
                sl[ss]->stop_time = MAX2(st, d.tmin);
            } else {
                sl[ss]->stop_time = d.out_tstart + ss * dir * d.out_tdelta + (double)dir * d.ftle_len;
            }
        }
        int written = 0;
        for (int df = first; df != last; df += dir) {
            if ((rc = load_slice(&d, df + dir, m->nvel, uvw))) goto done;
            oc_mesh_push_frame(m, dir, uvw);
            double lmin = dir > 0 ? d.tmin + df * d.tdelta : d.tmin + (df - 1) * d.tdelta;
            oc_mesh_set_interval(m, lmin, lmin + d.tdelta);
            for (int ss = 0; ss < d.out_tres; ss++) {
                int was = sl[ss]->status;
                if ((rc = oc_ftle_frame(m, &ip, sl[ss], dir, m->tmin, m->tmax))) goto done;
                if (was != 2 && sl[ss]->status == 2) {
                    oc_ftle_fill(sl[ss], d.dims, d.ftle_len);
                    int outnum = dir > 0 ? ss : d.out_tres - 1 - ss;
                    if ((rc = write_ftle_outputs(&d, sl[ss], ss, outnum, ss == 0 && !written))) goto done;
                    if (ss == 0) written = 1;
                }
            }
        }
        for (int ss = 0; ss < d.out_tres; ss++) oc_ftle_free(sl[ss]);
        free(sl);
    } else if (d.trace) {
        int64_t n = 0;
        oc_point *ic = NULL;
        if (d.trace_gen) {
            double mn[3] = {d.trace_b[0], d.trace_b[2], d.trace_b[4]}, del[3];
            int32_t res[3] = {d.trace_res[0], d.trace_res[1], d.trace_res[2]};
            del[0] = res[0] == 1 ? 0.0 : (d.trace_b[1] - d.trace_b[0]) / ((double)res[0] - 1.0);
            del[1] = res[1] == 1 ? 0.0 : (d.trace_b[3] - d.trace_b[2]) / ((double)res[1] - 1.0);
            if (res[2] == 1)
                del[2] = 0.0;
            else if (d.dims == 3)
                del[2] = (d.trace_b[5] - d.trace_b[4]) / ((double)res[2] - 1.0);
            else {
                mn[2] = 0;
                del[2] = 0;
                res[2] = 1;
            }
            n = (int64_t)res[0] * res[1] * res[2];
            ic = (oc_point *)calloc((size_t)n, sizeof(oc_point));
            if ((rc = oc_seed_grid(m, mn, del, res, d.lsc, ic))) goto done;
        } else {
            double *xyz = NULL;
            if ((rc = read_seed_file(&d, &n, &xyz))) goto done;
            ic = (oc_point *)calloc((size_t)n, sizeof(oc_point));
            rc = oc_seed_list(m, n, xyz, d.lsc, ic);
            free(xyz);
            if (rc) goto done;
        }
        {
            char path[700];
            snprintf(path, sizeof path, "%s%s.IC", d.path_data, d.trace_prefix);
            FILE *o = fopen(path, "wb");
            if (o) {
                int32_t nn = (int32_t)n;
                fwrite(&nn, 4, 1, o);
                fwrite(ic, sizeof(oc_point), (size_t)n, o);
                fclose(o);
            }
        }
        double tend = d.out_tstart + (d.out_tres - 1) * dir * d.out_tdelta;
        release_t *rl = (release_t *)calloc((size_t)d.trace_nl, sizeof(release_t));
        for (int ss = 0; ss < d.trace_nl; ss++) {
            rl[ss].start = d.out_tstart + dir * ss * d.trace_spacing;
            rl[ss].stop = dir > 0 ? MIN2(rl[ss].start + d.trace_len, tend) : MAX2(rl[ss].start - d.trace_len, tend);
            rl[ss].pts = (oc_point *)malloc(sizeof(oc_point) * (size_t)n);
            memcpy(rl[ss].pts, ic, sizeof(oc_point) * (size_t)n);
        }
        int upper = d.trace_nl >= nproc / 2; /* parameters.c:411-421 */
        /* per output frame, records in release order */
        double **outbuf = (double **)calloc((size_t)d.out_tres, sizeof(double *));
        int64_t *outcnt = (int64_t *)calloc((size_t)d.out_tres, sizeof(int64_t));
        int64_t *outcap = (int64_t *)calloc((size_t)d.out_tres, sizeof(int64_t));
        int outframe = 0;
        double nextout = d.out_tstart;
        for (int df = first; df != last; df += dir) {
            if ((rc = load_slice(&d, df + dir, m->nvel, uvw))) goto done;
            oc_mesh_push_frame(m, dir, uvw);
            double lmin = dir > 0 ? d.tmin + df * d.tdelta : d.tmin + (df - 1) * d.tdelta;
            oc_mesh_set_interval(m, lmin, lmin + d.tdelta);
            /* Compute_Tracer (tracers.c:25-289), sequential release order */
            double t1 = dir > 0 ? m->tmin : m->tmax, t2;
            while (((dir > 0) ? (t1 - m->tmax) : (m->tmin - t1)) < 0) {
                t2 = dir > 0 ? MIN2(nextout, m->tmax) : MAX2(m->tmin, nextout);
                for (int ss = 0; ss < d.trace_nl; ss++) {
                    release_t *r = &rl[ss];
                    if (r->status != 2 && (r->status == 1 || (r->start <= MAX2(t1, t2) && r->start >= MIN2(t1, t2)))) {
                        if (r->status == 0) {
                            r->status = 1;
                            t1 = r->start;
                            for (int64_t q = 0; q < n; q++)
                                if (!r->pts[q].LeftDomain) {
                                    if ((rc = oc_velocity(m, r->start, &r->pts[q], r->pts[q].V))) goto done;
                                }
                        }
                        /* LOWERLOOP's parallel loop over the tracers, schedule(dynamic, N/nproc) (tracers.c:218-226) */
                        if ((rc = oc_advect_batch(m, &ip, n, r->pts, t1, t2, nproc, n / nproc > 0 ? n / nproc : 1, NULL)))
                            goto done;
                        if (fabs(nextout - t2) < TINY) {
                            for (int64_t q = 0; q < n; q++) {
                                const oc_point *p = &r->pts[q];
                                if (p->LeftDomainTime > -TINY && (!p->LeftDomain || d.trace_always)) {
                                    if (outcnt[outframe] == outcap[outframe]) {
                                        outcap[outframe] = outcap[outframe] ? 2 * outcap[outframe] : 1024;
                                        outbuf[outframe] = (double *)realloc(outbuf[outframe],
                                                                              sizeof(double) * 3 * (size_t)outcap[outframe]);
                                    }
                                    memcpy(outbuf[outframe] + 3 * outcnt[outframe], p->X, 3 * sizeof(double));
                                    outcnt[outframe]++;
                                }
                            }
                        }
                    }
                    if (((dir > 0) ? (r->stop - t2) : (t2 - r->stop)) < TINY) r->status = 2;
                }
                if (fabs(t2 - nextout) < TINY) {
                    nextout += dir * d.out_tdelta;
                    outframe++;
                    if (outframe >= d.out_tres) break;
                }
                t1 = t2;
            }
        }
        /* OutputTracers(_OMP) (tracers.c:1700-1816) */
        for (int i = 0; i < d.out_tres; i++) {
            if (!upper && outcnt[i] == 0) continue;
            char path[700];
            snprintf(path, sizeof path, "%s%s.%d.bin", d.path_out, d.trace_prefix, i);
            FILE *o = fopen(path, "wb");
            if (!o) {
                rc = OC_E_IO;
                goto done;
            }
            double tt = d.out_tstart + i * dir * d.out_tdelta;
            fwrite(&tt, 8, 1, o);
            if (outcnt[i]) fwrite(outbuf[i], 8, (size_t)(3 * outcnt[i]), o);
            fclose(o);
        }
        for (int i = 0; i < d.out_tres; i++) free(outbuf[i]);
        free(outbuf);
        free(outcnt);
        free(outcap);
        for (int ss = 0; ss < d.trace_nl; ss++) free(rl[ss].pts);
        free(rl);
        free(ic);
    }
done:
    if (rc && msg && !msg[0]) snprintf(msg, msglen, "error %d", rc);
    free(uvw);
    oc_mesh_free(m);
    return rc;
}
