/* Z-order renumbering of a tetrahedral mesh: CPU restatement of
 * cudaVC/ZOrderSort/main.cpp (Work(), lines 103-299).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker of the device
 * renumbering (fvc_mesh_zorder); the product never links, loads or calls it.
 *
 * Steps as in the reference:
 *   - per axis, the sorted distinct coordinate values (std::sort +
 *     std::unique, main.cpp:32-35) and every point's rank = lower_bound index
 *     among them (main.cpp:111-124);
 *   - points ordered by the comparator of main.cpp:86-101 on the rank triples;
 *   - cell centroids summed in connectivity order from 0 and divided by 4
 *     (main.cpp:168-183, 187-202), ranked and ordered the same way;
 *   - new coordinates, connectivity (through the new point ids) and adjacency
 *     (through the new cell ids, -1 kept) written in the new orders
 *     (main.cpp:140-151, 214-259).
 * The reference sorts with std::sort, whose order among equal rank triples is
 * unspecified; here equal triples keep their input order.
 *
 * Pinned by: the reference program built from its source (oracle/ref.mk) on
 * the fixtures under tests/golden/zorder_* (bit-identical files), and by a
 * brute-force check of the comparator on small integer sets. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fvc_oracle.h"

static int cmp_dbl(const void *pa, const void *pb) {
    const double a = *(const double *)pa, b = *(const double *)pb;
    return a < b ? -1 : (b < a ? 1 : 0);
}

/* rank of every value among the distinct values (main.cpp:32-35, 120-124) */
static int rank_values(int64_t n, const double *v, int stride, int32_t *rank) {
    double *u = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!u) return OC_E_NOMEM;
    for (int64_t i = 0; i < n; i++) u[i] = v[i * stride];
    qsort(u, (size_t)n, sizeof(double), cmp_dbl);
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
        if (m == 0 || !(u[m - 1] == u[i])) u[m++] = u[i];
    for (int64_t i = 0; i < n; i++) {
        const double x = v[i * stride];
        int64_t lo = 0, hi = m; /* first u[k] not less than x */
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (u[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        rank[i] = (int32_t)lo;
    }
    free(u);
    return OC_OK;
}

static const int32_t *g_rank; /* rank triples for the comparator (single-threaded) */

/* main.cpp:86-101: the axis whose XOR has the highest set bit decides; equal
 * highest bits go to the lower axis */
static int zless(const int32_t *a, const int32_t *b) {
    int j = 0, x = 0;
    for (int k = 0; k < 3; k++) {
        const int y = a[k] ^ b[k];
        if (x < y && x < (x ^ y)) {
            j = k;
            x = y;
        }
    }
    return a[j] < b[j];
}

static int cmp_z(const void *pa, const void *pb) {
    const int32_t i = *(const int32_t *)pa, j = *(const int32_t *)pb;
    const int32_t *a = g_rank + 3 * (int64_t)i, *b = g_rank + 3 * (int64_t)j;
    if (zless(a, b)) return -1;
    if (zless(b, a)) return 1;
    return i < j ? -1 : (i > j ? 1 : 0);
}

static int order_by_z(int64_t n, const double *xyz, int32_t *order) {
    int32_t *rank = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(n > 0 ? n : 1));
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!rank || !tmp) {
        free(rank);
        free(tmp);
        return OC_E_NOMEM;
    }
    int rc = OC_OK;
    for (int a = 0; a < 3 && rc == OC_OK; a++) {
        rc = rank_values(n, xyz + a, 3, tmp);
        for (int64_t i = 0; i < n; i++) rank[3 * i + a] = tmp[i];
    }
    if (rc == OC_OK) {
        for (int64_t i = 0; i < n; i++) order[i] = (int32_t)i;
        g_rank = rank;
        qsort(order, (size_t)n, sizeof(int32_t), cmp_z);
        g_rank = NULL;
    }
    free(rank);
    free(tmp);
    return rc;
}

int oc_mesh_zorder(int64_t nn, const double *xyz, int64_t ne, const int32_t *conn, const int32_t *adj,
                   int32_t *node_order, int32_t *elem_order, double *xyz_out, int32_t *conn_out, int32_t *adj_out) {
    for (int64_t e = 0; e < 4 * ne; e++) {
        if (conn[e] < 0 || conn[e] >= nn) return OC_E_BADMESH;
        if (adj && (adj[e] < -1 || adj[e] >= ne)) return OC_E_BADMESH;
    }
    int rc = order_by_z(nn, xyz, node_order);
    if (rc) return rc;
    double *cen = (double *)malloc(sizeof(double) * 3 * (size_t)(ne > 0 ? ne : 1));
    int32_t *new_node = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nn > 0 ? nn : 1));
    int32_t *new_elem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    if (!cen || !new_node || !new_elem) {
        free(cen);
        free(new_node);
        free(new_elem);
        return OC_E_NOMEM;
    }
    for (int64_t e = 0; e < ne; e++) /* main.cpp:168-183 */
        for (int a = 0; a < 3; a++) {
            double s = 0;
            for (int q = 0; q < 4; q++) s += xyz[3 * (int64_t)conn[4 * e + q] + a];
            cen[3 * e + a] = s / 4;
        }
    rc = order_by_z(ne, cen, elem_order);
    if (rc == OC_OK) {
        for (int64_t i = 0; i < nn; i++) new_node[node_order[i]] = (int32_t)i;
        for (int64_t i = 0; i < ne; i++) new_elem[elem_order[i]] = (int32_t)i;
        if (xyz_out)
            for (int64_t i = 0; i < nn; i++) memcpy(xyz_out + 3 * i, xyz + 3 * (int64_t)node_order[i], 3 * sizeof(double));
        for (int64_t i = 0; i < ne; i++) {
            const int64_t id = elem_order[i];
            for (int q = 0; q < 4; q++) {
                if (conn_out) conn_out[4 * i + q] = new_node[conn[4 * id + q]];
                if (adj_out) {
                    const int32_t a = adj[4 * id + q];
                    adj_out[4 * i + q] = a == -1 ? -1 : new_elem[a];
                }
            }
        }
    }
    free(cen);
    free(new_node);
    free(new_elem);
    return rc;
}
