/*
 * fvc_gen.c - deterministic synthetic inputs for the advection path.
 *
 *  - double gyre on [0,2]x[0,1] (the field of cudaVC/data/DoubleGyre, and of the
 *    2-D Cartesian FTLE workload with A=0.1, eps=0.25, omega=2 pi/10);
 *  - voxelised tubes (straight pipe, Y-bifurcation) split into Kuhn tetrahedra
 *    with face adjacency in ompVC's neighbour-slot convention
 *    (ompVC/velocity.c:589-596: r -> slot 0, s -> 1, t -> 2, 1-r-s-t -> 3, where
 *    r,s,t are the natural coordinates attached to nodes 1,2,3);
 *  - pulsatile velocity slices with seeded Gaussian noise;
 *  - seeded tracer seed clouds.
 * PRNG: SplitMix64 expands the seed into xoshiro256** state (Blackman & Vigna).
 */
#include "fvc_host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* PRNG                                                                       */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void fvc_rng_seed(fvc_rng *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
    r->has_spare = 0;
    r->spare = 0.0;
}

static inline uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

uint64_t fvc_rng_next(fvc_rng *r) {
    uint64_t *s = r->s;
    uint64_t out = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return out;
}

double fvc_rng_uniform(fvc_rng *r) { return (double)(fvc_rng_next(r) >> 11) * 0x1.0p-53; }

double fvc_rng_normal(fvc_rng *r) {
    if (r->has_spare) {
        r->has_spare = 0;
        return r->spare;
    }
    double u1, u2;
    do {
        u1 = fvc_rng_uniform(r);
    } while (u1 <= 0.0);
    u2 = fvc_rng_uniform(r);
    double mag = sqrt(-2.0 * log(u1));
    r->spare = mag * sin(2.0 * M_PI * u2);
    r->has_spare = 1;
    return mag * cos(2.0 * M_PI * u2);
}

/* ------------------------------------------------------------------------ */
/* Double gyre                                                               */
/* ------------------------------------------------------------------------ */
void fvc_double_gyre(double A, double eps, double omega, double x, double y, double t, double *u, double *v) {
    double a = eps * sin(omega * t);
    double b = 1.0 - 2.0 * a;
    double f = a * x * x + b * x;
    double dfdx = 2.0 * a * x + b;
    *u = -M_PI * A * sin(M_PI * f) * cos(M_PI * y);
    *v = M_PI * A * cos(M_PI * f) * sin(M_PI * y) * dfdx;
}

void fvc_gen_double_gyre_slice(int nx, int ny, double t, double A, double eps, double omega, double *uvw) {
    double dx = 2.0 / (nx - 1), dy = 1.0 / (ny - 1);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            double u, v;
            fvc_double_gyre(A, eps, omega, i * dx, j * dy, t, &u, &v);
            int64_t n = (int64_t)j * nx + i;
            uvw[3 * n + 0] = u;
            uvw[3 * n + 1] = v;
            uvw[3 * n + 2] = 0.0;
        }
}

int fvc_gen_double_gyre_files(const char *dir, const char *prefix, int nx, int ny, int nt, double tdelta,
                              double A, double eps, double omega, int suffix_min, int suffix_delta) {
    char path[2 * FVC_STR];
    double mn[3] = {0.0, 0.0, 0.0}, mx[3] = {2.0, 1.0, 0.0};
    int32_t res[3] = {nx, ny, 1};
    snprintf(path, sizeof path, "%s/%s_Cartesian.bin", dir, prefix);
    if (fvc_write_cart_header(path, mn, mx, res)) return -1;
    double *uvw = (double *)malloc(sizeof(double) * 3 * (size_t)nx * ny);
    if (!uvw) return -2;
    for (int s = 0; s < nt; s++) {
        double t = s * tdelta;
        fvc_gen_double_gyre_slice(nx, ny, t, A, eps, omega, uvw);
        snprintf(path, sizeof path, "%s/%s_vel.%d.bin", dir, prefix, suffix_min + s * suffix_delta);
        if (fvc_write_vel_slice(path, t, uvw, (int64_t)nx * ny)) {
            free(uvw);
            return -1;
        }
    }
    free(uvw);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Tubes                                                                     */
/* ------------------------------------------------------------------------ */
void fvc_tube_pipe(fvc_tube *t, int nx, int ny, int nz, double R, double L) {
    memset(t, 0, sizeof *t);
    t->shape = 0;
    t->nx = nx; t->ny = ny; t->nz = nz;
    t->lo[0] = -R; t->hi[0] = R;
    t->lo[1] = -R; t->hi[1] = R;
    t->lo[2] = 0.0; t->hi[2] = L;
    t->R = R; t->L = L;
}

/* Parent: axis from (0,0,-L) to (0,0,0), radius R.  Daughters: axes from the
 * origin along (+-sin a, 0, cos a) for length Ld, radius rd. */
void fvc_tube_ybif(fvc_tube *t, int nx, int ny, int nz, double R, double rd, double ang_deg, double L, double Ld) {
    memset(t, 0, sizeof *t);
    t->shape = 1;
    t->nx = nx; t->ny = ny; t->nz = nz;
    double a = ang_deg * M_PI / 180.0;
    double xr = fmax(R, Ld * sin(a) + rd / cos(a));
    t->lo[0] = -xr; t->hi[0] = xr;
    t->lo[1] = -R;  t->hi[1] = R;
    t->lo[2] = -L;  t->hi[2] = Ld * cos(a) + rd * sin(a);
    t->R = R; t->L = L; t->rd = rd; t->ang_deg = ang_deg; t->Ld = Ld;
}

/* distance-to-segment test for a capsule-free finite cylinder */
static int in_cyl(double x, double y, double z, const double a[3], const double d[3], double len, double r) {
    double px = x - a[0], py = y - a[1], pz = z - a[2];
    double s = px * d[0] + py * d[1] + pz * d[2];
    if (s < 0.0 || s > len) return 0;
    double qx = px - s * d[0], qy = py - s * d[1], qz = pz - s * d[2];
    return qx * qx + qy * qy + qz * qz <= r * r;
}

int fvc_tube_inside(const fvc_tube *t, double x, double y, double z) {
    if (t->shape == 0) {
        return x * x + y * y <= t->R * t->R && z >= t->lo[2] && z <= t->hi[2];
    }
    double a = t->ang_deg * M_PI / 180.0;
    double o[3] = {0.0, 0.0, -t->L}, dz[3] = {0.0, 0.0, 1.0};
    if (in_cyl(x, y, z, o, dz, t->L, t->R)) return 1;
    double c[3] = {0.0, 0.0, 0.0};
    double d1[3] = {sin(a), 0.0, cos(a)}, d2[3] = {-sin(a), 0.0, cos(a)};
    if (in_cyl(x, y, z, c, d1, t->Ld, t->rd)) return 1;
    if (in_cyl(x, y, z, c, d2, t->Ld, t->rd)) return 1;
    /* junction sphere closes the gap between the parent end and daughter starts */
    return x * x + y * y + z * z <= t->R * t->R;
}

static unsigned char *tube_voxel_mask(const fvc_tube *t) {
    int64_t nv = (int64_t)t->nx * t->ny * t->nz;
    unsigned char *m = (unsigned char *)malloc((size_t)nv);
    if (!m) return NULL;
    double dx = (t->hi[0] - t->lo[0]) / t->nx, dy = (t->hi[1] - t->lo[1]) / t->ny, dz = (t->hi[2] - t->lo[2]) / t->nz;
    for (int k = 0; k < t->nz; k++)
        for (int j = 0; j < t->ny; j++)
            for (int i = 0; i < t->nx; i++) {
                double xc = t->lo[0] + (i + 0.5) * dx, yc = t->lo[1] + (j + 0.5) * dy, zc = t->lo[2] + (k + 0.5) * dz;
                m[((int64_t)k * t->ny + j) * t->nx + i] = (unsigned char)fvc_tube_inside(t, xc, yc, zc);
            }
    return m;
}

/* node (i,j,k) -> id, -1 when unused; i fastest */
static int32_t *tube_node_ids(const fvc_tube *t, const unsigned char *m, int64_t *nn_out) {
    int NX = t->nx + 1, NY = t->ny + 1, NZ = t->nz + 1;
    int64_t nnode = (int64_t)NX * NY * NZ;
    int32_t *id = (int32_t *)malloc(sizeof(int32_t) * (size_t)nnode);
    if (!id) return NULL;
    int64_t count = 0;
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                int used = 0;
                for (int c = 0; c < 8 && !used; c++) {
                    int vi = i - (c & 1), vj = j - ((c >> 1) & 1), vk = k - ((c >> 2) & 1);
                    if (vi < 0 || vj < 0 || vk < 0 || vi >= t->nx || vj >= t->ny || vk >= t->nz) continue;
                    used = m[((int64_t)vk * t->ny + vj) * t->nx + vi];
                }
                id[((int64_t)k * NY + j) * NX + i] = used ? (int32_t)count++ : -1;
            }
    *nn_out = count;
    return id;
}

int fvc_tube_count(const fvc_tube *t, int64_t *nn, int64_t *ne) {
    unsigned char *m = tube_voxel_mask(t);
    if (!m) return -2;
    int64_t nv = (int64_t)t->nx * t->ny * t->nz, kept = 0;
    for (int64_t v = 0; v < nv; v++) kept += m[v];
    int32_t *id = tube_node_ids(t, m, nn);
    free(m);
    if (!id) return -2;
    free(id);
    *ne = 6 * kept;
    return 0;
}

/* Kuhn split along the (0,0,0)-(1,1,1) diagonal: one tet per axis permutation. */
static const int kuhn_perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

static inline uint64_t face_key(int32_t a, int32_t b, int32_t c) {
    int32_t t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    return ((uint64_t)(uint32_t)a << 42) | ((uint64_t)(uint32_t)b << 21) | (uint64_t)(uint32_t)c;
}

int fvc_tube_fill(const fvc_tube *t, double *xyz, int32_t *conn, int32_t *adj) {
    unsigned char *m = tube_voxel_mask(t);
    if (!m) return -2;
    int64_t nn;
    int32_t *id = tube_node_ids(t, m, &nn);
    if (!id) {
        free(m);
        return -2;
    }
    if (nn >= (1 << 21)) { /* face keys pack 21 bits per node id */
        free(m);
        free(id);
        return -3;
    }
    int NX = t->nx + 1, NY = t->ny + 1, NZ = t->nz + 1;
    double dx = (t->hi[0] - t->lo[0]) / t->nx, dy = (t->hi[1] - t->lo[1]) / t->ny, dz = (t->hi[2] - t->lo[2]) / t->nz;
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                int32_t n = id[((int64_t)k * NY + j) * NX + i];
                if (n < 0) continue;
                xyz[3 * (int64_t)n + 0] = t->lo[0] + i * dx;
                xyz[3 * (int64_t)n + 1] = t->lo[1] + j * dy;
                xyz[3 * (int64_t)n + 2] = t->lo[2] + k * dz;
            }
    int64_t e = 0;
    for (int k = 0; k < t->nz; k++)
        for (int j = 0; j < t->ny; j++)
            for (int i = 0; i < t->nx; i++) {
                if (!m[((int64_t)k * t->ny + j) * t->nx + i]) continue;
                for (int p = 0; p < 6; p++) {
                    int c[3] = {0, 0, 0};
                    int32_t nodes[4];
                    for (int s = 0; s < 4; s++) {
                        if (s == 3) { c[0] = c[1] = c[2] = 1; }
                        else if (s > 0) c[kuhn_perm[p][s - 1]] = 1;
                        nodes[s] = id[((int64_t)(k + c[2]) * NY + (j + c[1])) * NX + (i + c[0])];
                    }
                    memcpy(conn + 4 * e, nodes, sizeof nodes);
                    e++;
                }
            }
    free(id);
    /* Face adjacency; slot s holds the neighbour across the face opposite node
     * (s+1)%4.  A Kuhn tet's face lies inside its voxel or on one voxel face, so
     * the matching tet is among the 6 tets of the same voxel or of the voxel
     * across that face: no global face table is needed. */
    int64_t ne = e;
    int64_t nvox = (int64_t)t->nx * t->ny * t->nz;
    int64_t *vbase = (int64_t *)malloc(sizeof(int64_t) * (size_t)nvox);
    if (!vbase) {
        free(m);
        return -2;
    }
    {
        int64_t q = 0;
        for (int64_t v = 0; v < nvox; v++) vbase[v] = m[v] ? 6 * q++ : -1;
    }
    free(m);
    for (int64_t q = 0; q < 4 * ne; q++) adj[q] = -1;
    for (int k = 0; k < t->nz; k++)
        for (int j = 0; j < t->ny; j++)
            for (int i = 0; i < t->nx; i++) {
                int64_t v = ((int64_t)k * t->ny + j) * t->nx + i;
                if (vbase[v] < 0) continue;
                for (int p = 0; p < 6; p++) {
                    /* local corner coordinates of the 4 nodes */
                    int lc[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}};
                    lc[1][kuhn_perm[p][0]] = 1;
                    lc[2][kuhn_perm[p][0]] = 1;
                    lc[2][kuhn_perm[p][1]] = 1;
                    int64_t el = vbase[v] + p;
                    const int32_t *nd = conn + 4 * el;
                    for (int s = 0; s < 4; s++) {
                        if (adj[4 * el + s] >= 0) continue;
                        int opp = (s + 1) & 3, fq[3], w = 0;
                        for (int q = 0; q < 4; q++)
                            if (q != opp) fq[w++] = q;
                        uint64_t key = face_key(nd[fq[0]], nd[fq[1]], nd[fq[2]]);
                        /* which voxel holds the other side */
                        int64_t vn = v;
                        for (int a = 0; a < 3; a++) {
                            int c0 = lc[fq[0]][a];
                            if (lc[fq[1]][a] == c0 && lc[fq[2]][a] == c0) {
                                int ii = i, jj = j, kk = k, dlt = c0 ? 1 : -1;
                                if (a == 0) ii += dlt; else if (a == 1) jj += dlt; else kk += dlt;
                                if (ii < 0 || jj < 0 || kk < 0 || ii >= t->nx || jj >= t->ny || kk >= t->nz)
                                    vn = -1;
                                else
                                    vn = ((int64_t)kk * t->ny + jj) * t->nx + ii;
                                break;
                            }
                        }
                        if (vn < 0 || vbase[vn] < 0) continue;
                        for (int p2 = 0; p2 < 6; p2++) {
                            int64_t e2 = vbase[vn] + p2;
                            if (e2 == el) continue;
                            const int32_t *n2 = conn + 4 * e2;
                            for (int s2 = 0; s2 < 4; s2++) {
                                int o2 = (s2 + 1) & 3, g[3], w2 = 0;
                                for (int q = 0; q < 4; q++)
                                    if (q != o2) g[w2++] = n2[q];
                                if (face_key(g[0], g[1], g[2]) == key) {
                                    adj[4 * el + s] = (int32_t)e2;
                                    adj[4 * e2 + s2] = (int32_t)el;
                                    p2 = 6;
                                    break;
                                }
                            }
                        }
                    }
                }
            }
    free(vbase);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Velocity                                                                  */
/* ------------------------------------------------------------------------ */
static void ybif_branch_velocity(const fvc_tube *t, double x, double y, double z, double *u) {
    /* Poiseuille along each branch axis with peak 2*(R_parent/r_branch)^2 scaled
     * so that flux splits evenly; blended by smooth weights of axis distance. */
    double a = t->ang_deg * M_PI / 180.0;
    double dirs[3][3] = {{0.0, 0.0, 1.0}, {sin(a), 0.0, cos(a)}, {-sin(a), 0.0, cos(a)}};
    double orig[3][3] = {{0.0, 0.0, -t->L}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double rad[3] = {t->R, t->rd, t->rd};
    double lens[3] = {t->L, t->Ld, t->Ld};
    double peak[3] = {2.0, 2.0 * 0.5 * (t->R * t->R) / (t->rd * t->rd), 0.0};
    peak[2] = peak[1];
    double wsum = 0.0, acc[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < 3; b++) {
        double px = x - orig[b][0], py = y - orig[b][1], pz = z - orig[b][2];
        double s = px * dirs[b][0] + py * dirs[b][1] + pz * dirs[b][2];
        double sc = s < 0.0 ? 0.0 : (s > lens[b] ? lens[b] : s);
        double qx = px - sc * dirs[b][0], qy = py - sc * dirs[b][1], qz = pz - sc * dirs[b][2];
        double r2 = (qx * qx + qy * qy + qz * qz) / (rad[b] * rad[b]);
        double prof = 1.0 - r2;
        if (prof < 0.0) prof = 0.0;
        double w = exp(-4.0 * r2);
        for (int c = 0; c < 3; c++) acc[c] += w * peak[b] * prof * dirs[b][c];
        wsum += w;
    }
    for (int c = 0; c < 3; c++) u[c] = wsum > 0.0 ? acc[c] / wsum : 0.0;
}

void fvc_tube_velocity(const fvc_tube *t, int64_t nn, const double *xyz, double time, double sigma,
                       uint64_t seed, int32_t slice, double *uvw) {
    fvc_rng rng;
    fvc_rng_seed(&rng, seed ^ (0xA24BAED4963EE407ull * (uint64_t)(slice + 1)));
    double pulse_pipe = 1.0 + 0.5 * sin(2.0 * M_PI * time);
    double pulse_y = 1.0 + 0.6 * sin(2.0 * M_PI * time);
    for (int64_t n = 0; n < nn; n++) {
        double x = xyz[3 * n], y = xyz[3 * n + 1], z = xyz[3 * n + 2];
        double u[3];
        if (t->shape == 0) {
            double r2 = (x * x + y * y) / (t->R * t->R);
            double g = 1.0 - r2;
            double sw = 0.1 * g * sin(M_PI * z / t->L); /* u_theta / r */
            u[0] = -sw * y;
            u[1] = sw * x;
            u[2] = 2.0 * g * pulse_pipe;
        } else {
            ybif_branch_velocity(t, x, y, z, u);
            for (int c = 0; c < 3; c++) u[c] *= pulse_y;
        }
        for (int c = 0; c < 3; c++) uvw[3 * n + c] = u[c] + (sigma > 0.0 ? sigma * fvc_rng_normal(&rng) : 0.0);
    }
}

int fvc_tube_seeds(const fvc_tube *t, int64_t n, double z0, double z1, uint64_t seed, double *xyz) {
    fvc_rng rng;
    fvc_rng_seed(&rng, seed);
    for (int64_t q = 0; q < n; q++) {
        double x, y, z;
        if (t->shape == 0) {
            do {
                x = (2.0 * fvc_rng_uniform(&rng) - 1.0) * t->R;
                y = (2.0 * fvc_rng_uniform(&rng) - 1.0) * t->R;
            } while (x * x + y * y >= t->R * t->R);
            z = z0 + (z1 - z0) * fvc_rng_uniform(&rng);
        } else {
            do {
                x = t->lo[0] + (t->hi[0] - t->lo[0]) * fvc_rng_uniform(&rng);
                y = t->lo[1] + (t->hi[1] - t->lo[1]) * fvc_rng_uniform(&rng);
                z = z0 + (z1 - z0) * fvc_rng_uniform(&rng);
            } while (!fvc_tube_inside(t, x, y, z));
        }
        xyz[3 * q] = x;
        xyz[3 * q + 1] = y;
        xyz[3 * q + 2] = z;
    }
    return 0;
}
