/*
 * oc_core.c - oracle primitives: point location, interpolation, integrators,
 * largest eigenvalue.  TEST INFRASTRUCTURE ONLY (see fvc_oracle.h).
 *
 * Arithmetic order follows the reference expressions cited at each function so
 * that FP64 results match ompVC bit for bit; the code itself is this
 * project's.  Minimum macro semantics: (a < b ? a : b) as ompVC/macros.h:33.
 */
#include "fvc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TINY OC_TINY
#define MIN2(a, b) ((a) < (b) ? (a) : (b))
#define MAX2(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------------------ */
/* mesh                                                                       */
/* ------------------------------------------------------------------------ */
oc_mesh *oc_mesh_cartesian(int dims, const double bounds[6], const double cmin[3], const double cmax[3],
                           const int32_t cres[3]) {
    oc_mesh *m = (oc_mesh *)calloc(1, sizeof *m);
    if (!m) return NULL;
    m->dims = dims;
    m->mesh_type = 0;
    memcpy(m->bounds, bounds, sizeof m->bounds);
    for (int a = 0; a < 3; a++) {
        m->cmin[a] = cmin[a];
        m->cmax[a] = cmax[a];
        m->cres[a] = cres[a];
    }
    /* mesh.c:103-121 */
    m->cdel[0] = (m->cmax[0] - m->cmin[0]) / (m->cres[0] - 1);
    m->cdel[1] = (m->cmax[1] - m->cmin[1]) / (m->cres[1] - 1);
    if (dims == 3)
        m->cdel[2] = (m->cmax[2] - m->cmin[2]) / (m->cres[2] - 1);
    else {
        m->cmin[2] = 0;
        m->cmax[2] = 0;
        m->cres[2] = 1;
        m->cdel[2] = 0;
    }
    m->nvel = (int64_t)m->cres[0] * m->cres[1] * m->cres[2];
    m->vel[0] = (double *)calloc((size_t)(3 * m->nvel), sizeof(double));
    m->vel[1] = (double *)calloc((size_t)(3 * m->nvel), sizeof(double));
    m->tmin = 0.0;
    m->tmax = 1.0;
    return m;
}

oc_mesh *oc_mesh_unstructured(int dims, const double bounds[6], int64_t nn, const double *xyz, int64_t ne,
                              const int32_t *conn, const int32_t *adj) {
    oc_mesh *m = (oc_mesh *)calloc(1, sizeof *m);
    if (!m) return NULL;
    m->dims = dims;
    m->mesh_type = 1;
    memcpy(m->bounds, bounds, sizeof m->bounds);
    m->nn = nn;
    m->ne = ne;
    m->xyz = (double *)malloc(sizeof(double) * 3 * (size_t)nn);
    m->conn = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)ne);
    m->adj = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)ne);
    memcpy(m->xyz, xyz, sizeof(double) * 3 * (size_t)nn);
    memcpy(m->conn, conn, sizeof(int32_t) * 4 * (size_t)ne);
    memcpy(m->adj, adj, sizeof(int32_t) * 4 * (size_t)ne);
    m->nvel = nn;
    m->vel[0] = (double *)calloc((size_t)(3 * nn), sizeof(double));
    m->vel[1] = (double *)calloc((size_t)(3 * nn), sizeof(double));
    m->tmin = 0.0;
    m->tmax = 1.0;
    return m;
}

void oc_mesh_free(oc_mesh *m) {
    if (!m) return;
    free(m->xyz);
    free(m->conn);
    free(m->adj);
    free(m->vel[0]);
    free(m->vel[1]);
    free(m);
}

int oc_mesh_set_slot(oc_mesh *m, int slot, const double *uvw) {
    if (slot < 0 || slot > 1) return OC_E_PARAM;
    memcpy(m->vel[slot], uvw, sizeof(double) * 3 * (size_t)m->nvel);
    return OC_OK;
}

/* velocity.c:67-75,196-204: forward fills slot 1 after moving slot 1 to 0;
 * backward fills slot 0 after moving slot 0 to 1. */
int oc_mesh_push_frame(oc_mesh *m, int direction, const double *uvw) {
    int keep = direction > 0 ? 0 : 1, fill = direction > 0 ? 1 : 0;
    double *tmp = m->vel[keep];
    m->vel[keep] = m->vel[fill];
    m->vel[fill] = tmp;
    memcpy(m->vel[fill], uvw, sizeof(double) * 3 * (size_t)m->nvel);
    return OC_OK;
}

void oc_mesh_set_interval(oc_mesh *m, double tmin, double tmax) {
    m->tmin = tmin;
    m->tmax = tmax;
}

/* ------------------------------------------------------------------------ */
/* domain tests (velocity.c:624-669)                                         */
/* ------------------------------------------------------------------------ */
int oc_test_outside_domain(const oc_mesh *m, const double X[3]) {
    const double *b = m->bounds;
    if (X[0] < (b[0] - TINY) || X[0] > (b[1] + TINY)) return 1;
    if (X[1] < (b[2] - TINY) || X[1] > (b[3] + TINY)) return 1;
    if (X[2] < (b[4] - TINY) || X[2] > (b[5] + TINY)) return 1;
    return 0;
}

int oc_test_outside_cart_vel(const oc_mesh *m, const double X[3]) {
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        if (a == 2 && m->dims != 3) {
            lo[a] = hi[a] = 0;
            continue;
        }
        double q = (X[a] - m->cmin[a]) / m->cdel[a];
        lo[a] = (int)floor(q);
        hi[a] = (int)ceil(q);
    }
    for (int a = 0; a < 3; a++)
        if (lo[a] < 0 || hi[a] > m->cres[a] - 1) return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* natural coordinates of X in element e (mesh.c:527-547, velocity.c:497-517) */
/* ------------------------------------------------------------------------ */
static inline void natural3(const oc_mesh *m, int e, double x, double y, double z, double *pr, double *ps,
                            double *pt) {
    const int32_t *nd = m->conn + 4 * (int64_t)e;
    const double *q0 = m->xyz + 3 * (int64_t)nd[0], *q1 = m->xyz + 3 * (int64_t)nd[1];
    const double *q2 = m->xyz + 3 * (int64_t)nd[2], *q3 = m->xyz + 3 * (int64_t)nd[3];
    double x0 = q0[0], y0 = q0[1], z0 = q0[2], x1 = q1[0], y1 = q1[1], z1 = q1[2];
    double x2 = q2[0], y2 = q2[1], z2 = q2[2], x3 = q3[0], y3 = q3[1], z3 = q3[2];
    /* cofactors of the physical-to-natural map */
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
    *pr = (c11 * dx + c12 * dy + c13 * dz) / det;
    *ps = (c21 * dx + c22 * dy + c23 * dz) / det;
    *pt = (c31 * dx + c32 * dy + c33 * dz) / det;
}

static inline void natural2(const oc_mesh *m, int e, double x, double y, double *pr, double *ps) {
    const int32_t *nd = m->conn + 4 * (int64_t)e;
    const double *q0 = m->xyz + 3 * (int64_t)nd[0], *q1 = m->xyz + 3 * (int64_t)nd[1];
    const double *q2 = m->xyz + 3 * (int64_t)nd[2];
    double x0 = q0[0], y0 = q0[1], x1 = q1[0], y1 = q1[1], x2 = q2[0], y2 = q2[1];
    double c11 = y2 - y0, c12 = x0 - x2, c21 = y0 - y1, c22 = x1 - x0;
    double det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    *pr = (c11 * (x - x0) + c12 * (y - y0)) / det;
    *ps = (c21 * (x - x0) + c22 * (y - y0)) / det;
}

/* exit face of a rejected element: first of r, s, t, 1-r-s-t within TINY of
 * the minimum (mesh.c:551-563); returns slot or -1 when indeterminate */
static inline int exit_slot3(double r, double s, double t, double d) {
    if (fabs(r - d) < TINY) return 0;
    if (fabs(s - d) < TINY) return 1;
    if (fabs(t - d) < TINY) return 2;
    if (fabs(d - (1 - r - s - t)) < TINY) return 3;
    return -1;
}

static inline int exit_slot2(double r, double s, double d) {
    if (fabs(r - d) < TINY) return 0;
    if (fabs(s - d) < TINY) return 1;
    if (fabs(d - (1 - r - s)) < TINY) return 2;
    return -1;
}

/* mesh.c:496-611 (no hop limit, as the reference) */
int oc_local_search(const oc_mesh *m, const double X[3], int guess, int *err) {
    double x = X[0], y = X[1], z = X[2];
    for (;;) {
        double r, s, t, d;
        int slot;
        if (m->dims == 3) {
            natural3(m, guess, x, y, z, &r, &s, &t);
            d = MIN2(r, MIN2(s, MIN2(t, 1 - r - s - t)));
            if (d > -TINY) return guess;
            slot = exit_slot3(r, s, t, d);
        } else {
            natural2(m, guess, x, y, &r, &s);
            d = MIN2(r, MIN2(s, 1 - r - s));
            if (d > -TINY) return guess;
            slot = exit_slot2(r, s, d);
        }
        if (slot < 0) {
            if (err) *err = OC_E_INDETERMINATE;
            return -1;
        }
        guess = m->adj[4 * (int64_t)guess + slot];
        if (guess == -1) return -1;
    }
}

/* mesh.c:304-398: lowest index with d > -TINY */
int oc_global_search(const oc_mesh *m, const double X[3]) {
    for (int64_t e = 0; e < m->ne; e++) {
        double r, s, t, d;
        if (m->dims == 3) {
            natural3(m, (int)e, X[0], X[1], X[2], &r, &s, &t);
            d = MIN2(r, MIN2(s, MIN2(t, 1 - r - s - t)));
        } else {
            natural2(m, (int)e, X[0], X[1], &r, &s);
            d = MIN2(r, MIN2(s, 1 - r - s));
        }
        if (d > -TINY) return (int)e;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* interpolation                                                             */
/* ------------------------------------------------------------------------ */
/* velocity.c:257-457 */
static int velocity_cart(const oc_mesh *m, double tq, oc_point *p, double out[3]) {
    if (oc_test_outside_cart_vel(m, p->X)) {
        out[0] = out[1] = out[2] = 0.0;
        p->LeftDomain = 1;
        return OC_OK;
    }
    double tloc = (tq - m->tmin) / (m->tmax - m->tmin);
    if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) return OC_E_TLOC;
    int cell[3];
    double loc[3];
    for (int a = 0; a < 3; a++) {
        if (a == 2 && m->dims != 3) {
            cell[a] = 0;
            loc[a] = 0.0;
            continue;
        }
        cell[a] = (int)floor((p->X[a] - m->cmin[a]) / m->cdel[a]);
        if (cell[a] == m->cres[a] - 1) {
            cell[a] = m->cres[a] - 2;
            loc[a] = 1.0;
        } else {
            loc[a] = (p->X[a] - m->cmin[a] - cell[a] * m->cdel[a]) / m->cdel[a];
        }
    }
    double tl = tloc, xl = loc[0], yl = loc[1], zl = loc[2];
    /* weights of the 16 space-time corners, named w<t><x><y><z> for the corner
     * (t = slot, x/y/z = +1 offsets) they multiply; products left to right */
    double w1111 = (tl) * (xl) * (yl) * (zl);
    double w1110 = (tl) * (xl) * (yl) * (1 - zl);
    double w1101 = (tl) * (xl) * (1 - yl) * (zl);
    double w1100 = (tl) * (xl) * (1 - yl) * (1 - zl);
    double w1011 = (tl) * (1 - xl) * (yl) * (zl);
    double w1010 = (tl) * (1 - xl) * (yl) * (1 - zl);
    double w1001 = (tl) * (1 - xl) * (1 - yl) * (zl);
    double w1000 = (tl) * (1 - xl) * (1 - yl) * (1 - zl);
    double w0111 = (1 - tl) * (xl) * (yl) * (zl);
    double w0110 = (1 - tl) * (xl) * (yl) * (1 - zl);
    double w0101 = (1 - tl) * (xl) * (1 - yl) * (zl);
    double w0100 = (1 - tl) * (xl) * (1 - yl) * (1 - zl);
    double w0011 = (1 - tl) * (1 - xl) * (yl) * (zl);
    double w0010 = (1 - tl) * (1 - xl) * (yl) * (1 - zl);
    double w0001 = (1 - tl) * (1 - xl) * (1 - yl) * (zl);
    double w0000 = (1 - tl) * (1 - xl) * (1 - yl) * (1 - zl);
    int64_t nx = m->cres[0], ny = m->cres[1];
    int i = cell[0], j = cell[1], k = cell[2];
    int k1 = m->dims == 3 ? k + 1 : k; /* 2-D repeats the k plane */
#define NODE(ii, jj, kk) ((((int64_t)(kk) * ny + (jj)) * nx + (ii)) * 3)
    int64_t c000 = NODE(i, j, k), c001 = NODE(i, j, k1), c010 = NODE(i, j + 1, k), c011 = NODE(i, j + 1, k1);
    int64_t c100 = NODE(i + 1, j, k), c101 = NODE(i + 1, j, k1), c110 = NODE(i + 1, j + 1, k),
            c111 = NODE(i + 1, j + 1, k1);
#undef NODE
    const double *A = m->vel[0], *B = m->vel[1];
    for (int c = 0; c < 3; c++) {
        /* sum order of velocity.c:345-346 */
        out[c] = A[c000 + c] * w0000 + A[c001 + c] * w0001 + A[c010 + c] * w0010 + A[c011 + c] * w0011 +
                 A[c100 + c] * w0100 + A[c101 + c] * w0101 + A[c110 + c] * w0110 + A[c111 + c] * w0111 +
                 B[c000 + c] * w1000 + B[c001 + c] * w1001 + B[c010 + c] * w1010 + B[c011 + c] * w1011 +
                 B[c100 + c] * w1100 + B[c101 + c] * w1101 + B[c110 + c] * w1110 + B[c111 + c] * w1111;
    }
    return OC_OK;
}

/* velocity.c:459-621: walks from p->ElementIndex (updated in place) */
static int velocity_unstruct(const oc_mesh *m, double tq, oc_point *p, double out[3]) {
    if (p->ElementIndex == -1) return OC_E_ELEM_NEG1;
    double x = p->X[0], y = p->X[1], z = p->X[2];
    double tloc = (tq - m->tmin) / (m->tmax - m->tmin);
    if ((tloc > 1 + TINY) || (tloc < 0 - TINY)) return OC_E_TLOC;
    for (;;) {
        double r, s, t = 0.0, d;
        int slot;
        if (m->dims == 3) {
            natural3(m, p->ElementIndex, x, y, z, &r, &s, &t);
            d = MIN2(r, MIN2(s, MIN2(t, 1 - r - s - t)));
        } else {
            natural2(m, p->ElementIndex, x, y, &r, &s);
            t = 0;
            d = MIN2(r, MIN2(s, 1 - r - s));
        }
        if ((d + TINY) >= 0) {
            const int32_t *nd = m->conn + 4 * (int64_t)p->ElementIndex;
            const double *A = m->vel[0], *B = m->vel[1];
            double nt[4][3];
            int nnod = m->dims == 3 ? 4 : 3;
            for (int q = 0; q < nnod; q++)
                for (int c = 0; c < 3; c++)
                    nt[q][c] = (1 - tloc) * A[3 * (int64_t)nd[q] + c] + tloc * B[3 * (int64_t)nd[q] + c];
            double ar = fabs(r), as = fabs(s), at = fabs(t);
            if (m->dims == 3) {
                for (int c = 0; c < 3; c++)
                    out[c] = nt[0][c] + (nt[1][c] - nt[0][c]) * ar + (nt[2][c] - nt[0][c]) * as +
                             (nt[3][c] - nt[0][c]) * at;
            } else {
                for (int c = 0; c < 2; c++)
                    out[c] = nt[0][c] + (nt[1][c] - nt[0][c]) * ar + (nt[2][c] - nt[0][c]) * as;
                out[2] = 0.0;
            }
            return OC_OK;
        }
        slot = m->dims == 3 ? exit_slot3(r, s, t, d) : exit_slot2(r, s, d);
        if (slot < 0) return OC_E_INDETERMINATE;
        p->ElementIndex = m->adj[4 * (int64_t)p->ElementIndex + slot];
        if (p->ElementIndex == -1) {
            p->LeftDomain = 1;
            out[0] = out[1] = out[2] = 0.0;
            return OC_OK;
        }
    }
}

int oc_velocity(const oc_mesh *m, double tq, oc_point *p, double out[3]) {
    return m->mesh_type == 0 ? velocity_cart(m, tq, p, out) : velocity_unstruct(m, tq, p, out);
}

/* ------------------------------------------------------------------------ */
/* integrators                                                               */
/* ------------------------------------------------------------------------ */
static inline double sign_of(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

/* integration.c:222-311 (tracer side computations omitted: not on the path) */
static double euler(const oc_mesh *m, oc_point *p, double t0, double t1, double dt, int extrap, int *err, int64_t *ns) {
    if (p->LeftDomain) {
        *err = OC_E_LEFT;
        return t1;
    }
    if (m->mesh_type == 1 && p->ElementIndex < 0) {
        *err = OC_E_ELEM_NEG1;
        return t1;
    }
    double tc = t0, h = sign_of(dt, t1 - t0), ve[3];
    while ((tc - t1) * (t1 - t0) < 0.0) {
        if ((tc + h - t1) * (tc + h - t0) > 0.0) h = sign_of(t1 - tc, t1 - t0);
        int rc = oc_velocity(m, tc, p, ve);
        if (rc) {
            *err = rc;
            return t1;
        }
        p->X[0] = p->X[0] + h * ve[0];
        p->X[1] = p->X[1] + h * ve[1];
        p->X[2] = p->X[2] + h * ve[2];
        tc = tc + h;
        if (fabs(h) > 1e-12 * fabs(dt)) (*ns)++;
        if (m->mesh_type == 1) {
            /* a walk that already left yields -1 without an out-of-range read */
            p->ElementIndex = p->ElementIndex < 0 ? -1 : oc_local_search(m, p->X, p->ElementIndex, err);
            if (*err) return t1;
        }
        if ((m->mesh_type == 0 && oc_test_outside_domain(m, p->X)) || (m->mesh_type == 1 && p->ElementIndex < 0)) {
            p->LeftDomain = 1;
            p->LeftDomainTime = tc - h;
            p->X[0] = p->X[0] - h * ve[0];
            p->X[1] = p->X[1] - h * ve[1];
            p->X[2] = p->X[2] - h * ve[2];
            return p->LeftDomainTime;
        }
        if (extrap) {
            rc = oc_velocity(m, tc, p, p->V);
            if (rc) {
                *err = rc;
                return t1;
            }
        }
    }
    return t1;
}

/* integration.c:317-462 */
static double rk4(const oc_mesh *m, oc_point *p, double t0, double t1, double dt, int extrap, int *err, int64_t *ns) {
    oc_point mp[4];
    if (p->LeftDomain) {
        *err = OC_E_LEFT;
        return t1;
    }
    memset(mp, 0, sizeof mp);
    if (m->mesh_type == 1) {
        if (p->ElementIndex < 0) {
            *err = OC_E_ELEM_NEG1;
            return t1;
        }
        for (int q = 0; q < 4; q++) mp[q].ElementIndex = p->ElementIndex;
    }
    double tc = t0, h = sign_of(dt, t1 - t0);
    double k1[3], k2[3], k3[3], k4[3];
    int rc;
    while ((tc - t1) * (t1 - t0) < 0.0) {
        if ((tc + h - t1) * (tc + h - t0) > 0.0) h = sign_of(t1 - tc, t1 - t0);
        mp[0].X[0] = p->X[0];
        mp[0].X[1] = p->X[1];
        mp[0].X[2] = p->X[2];
        if ((rc = oc_velocity(m, tc, &mp[0], k1))) goto fail;
        mp[1].X[0] = p->X[0] + 0.5 * k1[0] * h;
        mp[1].X[1] = p->X[1] + 0.5 * k1[1] * h;
        mp[1].X[2] = p->X[2] + 0.5 * k1[2] * h;
        if ((rc = oc_velocity(m, tc + h / 2, &mp[1], k2))) goto fail;
        mp[2].X[0] = p->X[0] + 0.5 * k2[0] * h;
        mp[2].X[1] = p->X[1] + 0.5 * k2[1] * h;
        mp[2].X[2] = p->X[2] + 0.5 * k2[2] * h;
        if ((rc = oc_velocity(m, tc + h / 2, &mp[2], k3))) goto fail;
        mp[3].X[0] = p->X[0] + k3[0] * h;
        mp[3].X[1] = p->X[1] + k3[1] * h;
        mp[3].X[2] = p->X[2] + k3[2] * h;
        if ((rc = oc_velocity(m, tc + h, &mp[3], k4))) goto fail;
        for (int c = 0; c < 3; c++) p->X[c] = p->X[c] + h * (k1[c] + 2 * k2[c] + 2 * k3[c] + k4[c]) / 6;
        tc = tc + h;
        if (fabs(h) > 1e-12 * fabs(dt)) (*ns)++;
        if (m->mesh_type == 1) {
            p->ElementIndex = oc_local_search(m, p->X, p->ElementIndex, err);
            if (*err) return t1;
        }
        if ((m->mesh_type == 0 && oc_test_outside_domain(m, p->X)) || (m->mesh_type == 1 && p->ElementIndex < 0) ||
            mp[0].LeftDomain || mp[1].LeftDomain || mp[2].LeftDomain || mp[3].LeftDomain) {
            p->LeftDomain = 1;
            p->LeftDomainTime = tc - h;
            for (int c = 0; c < 3; c++) p->X[c] = p->X[c] - h * (k1[c] + 2 * k2[c] + 2 * k3[c] + k4[c]) / 6;
            return p->LeftDomainTime;
        }
        if (extrap) {
            if ((rc = oc_velocity(m, tc, p, p->V))) goto fail;
        }
    }
    return t1;
fail:
    *err = rc;
    return t1;
}

/* integration.c:464-699: Runge-Kutta-Fehlberg 4(5), the fourth-order result
 * kept, step size from the difference of the two; six probes that keep their
 * own elements within the call; step starts at Int_MaxTimeStep each call.
 * Steps counted: accepted ones. */
static double rkf(const oc_mesh *m, oc_point *p, double t0, double t1, const oc_int *ip, int *err, int64_t *ns) {
    oc_point mp[6];
    if (p->LeftDomain) {
        *err = OC_E_LEFT;
        return t1;
    }
    memset(mp, 0, sizeof mp);
    if (m->mesh_type == 1) {
        if (p->ElementIndex == -1) {
            *err = OC_E_ELEM_NEG1;
            return t1;
        }
        for (int q = 0; q < 6; q++) mp[q].ElementIndex = p->ElementIndex;
    }
    const double a3 = 3.0 / 32.0, b3 = 9.0 / 32.0;
    const double a4 = 1932.0 / 2197.0, b4 = -7200.0 / 2197.0, c4 = 7296.0 / 2197.0;
    const double a5 = 439.0 / 216.0, b5 = -8.0, c5 = 3680.0 / 513.0, d5 = -845.0 / 4104.0;
    const double a6 = -8.0 / 27.0, b6 = 2.0, c6 = -3544.0 / 2565.0, d6 = 1859.0 / 4104.0, e6 = -11.0 / 40.0;
    const double a7 = 1.0 / 360.0, b7 = -128.0 / 4275.0, c7 = -2197.0 / 75240.0, d7 = 1.0 / 50.0, e7 = 2.0 / 55.0;
    const double a8 = 25.0 / 216.0, b8 = 1408.0 / 2565.0, c8 = 2197.0 / 4104.0, d8 = -1.0 / 5.0;
    double h = sign_of(ip->max_step, t1 - t0), hmax = ip->max_step, tol = ip->accuracy, tc = t0, ss;
    double k1[3], k2[3], k3[3], k4[3], k5[3], k6[3];
    int rc;
    while ((tc - t1) * (t1 - t0) < 0.0) {
        if ((tc + h - t1) * (tc + h - t0) > 0.0) h = sign_of(t1 - tc, t1 - t0);
        for (int c = 0; c < 3; c++) mp[0].X[c] = p->X[c];
        if ((rc = oc_velocity(m, tc, &mp[0], k1))) goto fail;
        for (int c = 0; c < 3; c++) mp[1].X[c] = p->X[c] + 0.25 * k1[c] * h;
        if ((rc = oc_velocity(m, tc + 0.25 * h, &mp[1], k2))) goto fail;
        for (int c = 0; c < 3; c++) mp[2].X[c] = p->X[c] + (a3 * k1[c] + b3 * k2[c]) * h;
        if ((rc = oc_velocity(m, tc + 0.375 * h, &mp[2], k3))) goto fail;
        for (int c = 0; c < 3; c++) mp[3].X[c] = p->X[c] + (a4 * k1[c] + b4 * k2[c] + c4 * k3[c]) * h;
        if ((rc = oc_velocity(m, tc + (12.0 * h) / 13.0, &mp[3], k4))) goto fail;
        for (int c = 0; c < 3; c++) mp[4].X[c] = p->X[c] + (a5 * k1[c] + b5 * k2[c] + c5 * k3[c] + d5 * k4[c]) * h;
        if ((rc = oc_velocity(m, tc + h, &mp[4], k5))) goto fail;
        for (int c = 0; c < 3; c++)
            mp[5].X[c] = p->X[c] + (a6 * k1[c] + b6 * k2[c] + c6 * k3[c] + d6 * k4[c] + e6 * k5[c]) * h;
        if ((rc = oc_velocity(m, tc + 0.5 * h, &mp[5], k6))) goto fail;
        if (mp[0].LeftDomain || mp[1].LeftDomain || mp[2].LeftDomain || mp[3].LeftDomain || mp[4].LeftDomain ||
            mp[5].LeftDomain) {
            p->LeftDomain = 1;
            p->LeftDomainTime = tc;
            return p->LeftDomainTime;
        }
        double e[3];
        for (int c = 0; c < 3; c++) e[c] = fabs(a7 * k1[c] + b7 * k3[c] + c7 * k4[c] + d7 * k5[c] + e7 * k6[c]);
        const double er = fmax(fmax(e[0], e[1]), e[2]);
        ss = er == 0.0 ? 0.0 : pow(fabs((tol * h) / (2 * er)), 0.25);
        if (er < tol || fabs(h) <= ip->min_step) {
            for (int c = 0; c < 3; c++) p->X[c] = p->X[c] + (a8 * k1[c] + b8 * k3[c] + c8 * k4[c] + d8 * k5[c]) * h;
            tc = tc + h;
            (*ns)++;
            if (m->mesh_type == 1) {
                p->ElementIndex = oc_local_search(m, p->X, p->ElementIndex, err);
                if (*err) return t1;
            }
            if ((m->mesh_type == 0 && oc_test_outside_domain(m, p->X)) || (m->mesh_type == 1 && p->ElementIndex < 0)) {
                p->LeftDomain = 1;
                p->LeftDomainTime = tc - h;
                for (int c = 0; c < 3; c++)
                    p->X[c] = p->X[c] - (a8 * k1[c] + b8 * k3[c] + c8 * k4[c] + d8 * k5[c]) * h;
                return p->LeftDomainTime;
            }
            if (ip->extrapolate && (rc = oc_velocity(m, tc, p, p->V))) goto fail;
            if (fabs(h) > ip->min_step && ss > 1.0) {
                ss = (ss > 4.0) ? 4.0 : ss;
                h = sign_of((fabs(ss * h) > hmax) ? hmax : ss * h, t1 - t0);
            }
        } else if (ss < 1.0) {
            ss = (ss < 0.1) ? 0.1 : ss;
            h = sign_of((fabs(ss * h) < ip->min_step) ? ip->min_step : ss * h, t1 - t0);
        } else {
            h = sign_of((fabs(0.5 * h) < ip->min_step) ? ip->min_step : 0.5 * h, t1 - t0);
        }
    }
    return t1;
fail:
    *err = rc;
    return t1;
}

/* Adams-Bashforth 2/3 (no ompVC counterpart; this project's definition, the
 * same as the GPU path's): fixed step x+ = x + h (3 f0 - f1)/2 (AB2) or
 * x+ = x + h (23 f0 - 16 f1 + 5 f2)/12 (AB3), RK4 steps while history is short
 * or when the step is clipped (history restarts), Euler for remainders below
 * 1e-12 dt (history kept); exits as RK4. */
double oc_ab_counted(const oc_mesh *m, oc_point *p, oc_hist *hs, double t0, double t1, double dt, int order, int *err,
                     int64_t *ns) {
    oc_point mp[4];
    *err = OC_OK;
    if (p->LeftDomain) {
        *err = OC_E_LEFT;
        return t1;
    }
    memset(mp, 0, sizeof mp);
    if (m->mesh_type == 1) {
        if (p->ElementIndex < 0) {
            *err = OC_E_ELEM_NEG1;
            return t1;
        }
        for (int q = 0; q < 4; q++) mp[q].ElementIndex = p->ElementIndex;
    }
    const double hfull = sign_of(dt, t1 - t0), hmin = 1e-12 * fabs(dt);
    double tc = t0, h = hfull;
    int rc;
    while ((tc - t1) * (t1 - t0) < 0.0) {
        if ((tc + h - t1) * (tc + h - t0) > 0.0) h = sign_of(t1 - tc, t1 - t0);
        int micro = fabs(h) <= hmin, full = fabs(h - hfull) <= 1e-9 * fabs(hfull);
        double f0[3], inc[3];
        memcpy(mp[0].X, p->X, sizeof p->X);
        if ((rc = oc_velocity(m, tc, &mp[0], f0))) goto fail;
        if (micro) {
            for (int c = 0; c < 3; c++) inc[c] = h * f0[c];
        } else if (full && hs->nhist >= order - 1) {
            if (order == 2)
                for (int c = 0; c < 3; c++) inc[c] = h * (3 * f0[c] - hs->F1[c]) / 2;
            else
                for (int c = 0; c < 3; c++) inc[c] = h * (23 * f0[c] - 16 * hs->F1[c] + 5 * hs->F2[c]) / 12;
        } else {
            double k2[3], k3[3], k4[3];
            for (int c = 0; c < 3; c++) mp[1].X[c] = p->X[c] + 0.5 * f0[c] * h;
            if ((rc = oc_velocity(m, tc + h / 2, &mp[1], k2))) goto fail;
            for (int c = 0; c < 3; c++) mp[2].X[c] = p->X[c] + 0.5 * k2[c] * h;
            if ((rc = oc_velocity(m, tc + h / 2, &mp[2], k3))) goto fail;
            for (int c = 0; c < 3; c++) mp[3].X[c] = p->X[c] + k3[c] * h;
            if ((rc = oc_velocity(m, tc + h, &mp[3], k4))) goto fail;
            for (int c = 0; c < 3; c++) inc[c] = h * (f0[c] + 2 * k2[c] + 2 * k3[c] + k4[c]) / 6;
        }
        for (int c = 0; c < 3; c++) p->X[c] = p->X[c] + inc[c];
        tc = tc + h;
        if (!micro) (*ns)++;
        if (m->mesh_type == 1) {
            p->ElementIndex = oc_local_search(m, p->X, p->ElementIndex, err);
            if (*err) return t1;
        }
        if ((m->mesh_type == 0 && oc_test_outside_domain(m, p->X)) || (m->mesh_type == 1 && p->ElementIndex < 0) ||
            mp[0].LeftDomain || mp[1].LeftDomain || mp[2].LeftDomain || mp[3].LeftDomain) {
            p->LeftDomain = 1;
            p->LeftDomainTime = tc - h;
            for (int c = 0; c < 3; c++) p->X[c] = p->X[c] - inc[c];
            return p->LeftDomainTime;
        }
        if (!micro) {
            if (full) {
                memcpy(hs->F2, hs->F1, sizeof hs->F1);
                memcpy(hs->F1, f0, sizeof f0);
                hs->nhist = hs->nhist < 2 ? hs->nhist + 1 : 2;
            } else {
                hs->nhist = 0;
            }
        }
        if (m->mesh_type == 1) mp[0].ElementIndex = p->ElementIndex;
    }
    return t1;
fail:
    *err = rc;
    return t1;
}

int oc_ab_batch(const oc_mesh *m, int64_t n, oc_point *pts, oc_hist *hist, double t1, double t2, double dt, int order,
                int nthreads, int64_t *steps) {
    int rc = OC_OK;
    int64_t total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
#endif
    for (int64_t q = 0; q < n; q++) {
        oc_point *p = &pts[q];
        if (p->LeftDomainTime > -TINY && !p->LeftDomain) {
            int err = OC_OK;
            int64_t ns = 0;
            oc_ab_counted(m, p, &hist[q], t1, t2, dt, order, &err, &ns);
            total += ns;
            if (err) {
#ifdef _OPENMP
#pragma omp critical
#endif
                if (!rc) rc = err;
            }
        }
    }
    (void)nthreads;
    if (steps) *steps = total;
    return rc;
}

/* integration.c:22-45 (inertial particles and RKF are outside this path) */
double oc_advect(const oc_mesh *m, const oc_int *ip, oc_point *p, double t1, double t2, int *err) {
    int64_t ns = 0;
    return oc_advect_counted(m, ip, p, t1, t2, err, &ns);
}

/* particle-steps counted as the GPU path counts them: every executed step,
 * the exit step included, floating-point micro-steps (|h| <= 1e-12 dt) not */
double oc_advect_counted(const oc_mesh *m, const oc_int *ip, oc_point *p, double t1, double t2, int *err,
                         int64_t *ns) {
    *err = OC_OK;
    if (p->LeftDomain && ip->extrapolate) {
        for (int c = 0; c < 3; c++) p->X[c] = p->X[c] + p->V[c] * (t2 - t1);
        return t2;
    }
    double stop = ip->int_type == 1   ? rk4(m, p, t1, t2, ip->time_step, ip->extrapolate, err, ns)
                  : ip->int_type == 2 ? rkf(m, p, t1, t2, ip, err, ns)
                                      : euler(m, p, t1, t2, ip->time_step, ip->extrapolate, err, ns);
    if ((fabs(stop - t2) > TINY) && ip->extrapolate)
        for (int c = 0; c < 3; c++) p->X[c] = p->X[c] + p->V[c] * (t2 - stop);
    return stop;
}

/* ------------------------------------------------------------------------ */
/* largest eigenvalue of a symmetric 3x3: cyclic Jacobi rotations            */
/* (Golub & Van Loan, Matrix Computations, sec. 8.5)                          */
/* ------------------------------------------------------------------------ */
double oc_max_eig_sym3(const double a_in[3][3], int *err) {
    double a[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = 0.5 * (a_in[i][j] + a_in[j][i]);
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        double diag = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
        if (off == 0.0 || off <= 1e-300 || off < 1e-18 * diag) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
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
        if (sweep == 63 && err) *err = OC_E_EIG;
    }
    return MAX2(a[0][0], MAX2(a[1][1], a[2][2]));
}

/* ------------------------------------------------------------------------ */
/* batches                                                                   */
/* ------------------------------------------------------------------------ */
int oc_locate_local_batch(const oc_mesh *m, int64_t n, const double *xyz, const int32_t *guess, int32_t *out) {
    int rc = OC_OK;
    for (int64_t q = 0; q < n; q++) {
        int err = OC_OK;
        out[q] = guess[q] < 0 ? -1 : oc_local_search(m, xyz + 3 * q, guess[q], &err);
        if (err && !rc) rc = err;
    }
    return rc;
}

int oc_locate_global_batch(const oc_mesh *m, int64_t n, const double *xyz, int32_t *out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t q = 0; q < n; q++) out[q] = oc_global_search(m, xyz + 3 * q);
    (void)nthreads;
    return OC_OK;
}

/* LOWERLOOP tracer integration (tracers.c:217-229): dynamic chunks over points */
int oc_advect_batch(const oc_mesh *m, const oc_int *ip, int64_t n, oc_point *pts, double t1, double t2,
                    int nthreads, int64_t chunk, int64_t *steps) {
    int rc = OC_OK;
    int64_t total = 0;
    if (chunk <= 0) chunk = 64;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, chunk) reduction(+ : total)
#endif
    for (int64_t q = 0; q < n; q++) {
        oc_point *p = &pts[q];
        if (p->LeftDomainTime > -TINY && (!p->LeftDomain || ip->extrapolate)) {
            int err = OC_OK;
            int64_t ns = 0;
            oc_advect_counted(m, ip, p, t1, t2, &err, &ns);
            total += ns;
            if (err) {
#ifdef _OPENMP
#pragma omp critical
#endif
                if (!rc) rc = err;
            }
        }
    }
    (void)nthreads;
    if (steps) *steps = total;
    return rc;
}

int oc_velocity_batch(const oc_mesh *m, double tq, int64_t n, oc_point *pts, double *uvw) {
    int rc = OC_OK;
    for (int64_t q = 0; q < n; q++) {
        int e = oc_velocity(m, tq, &pts[q], uvw + 3 * q);
        if (e && !rc) rc = e;
    }
    return rc;
}

int oc_eig_batch(int64_t n, const double *mats, double *lam) {
    int rc = OC_OK;
    for (int64_t q = 0; q < n; q++) {
        double a[3][3];
        memcpy(a, mats + 9 * q, sizeof a);
        lam[q] = oc_max_eig_sym3((const double(*)[3])a, &rc);
    }
    return rc;
}
