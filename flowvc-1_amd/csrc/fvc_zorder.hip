/* Z-order renumbering of a tetrahedral mesh on the device: the step
 * cudaVC/ZOrderSort/main.cpp:103-299 performs on the CPU before a mesh is
 * used, so that elements and nodes close in space are close in memory (the
 * walk's and the interpolation's gathers then share cache lines).
 *
 *   1. per axis, the rank of every coordinate among the distinct values
 *      (main.cpp:32-35 sort+unique, 111-124 lower_bound): radix sort of the
 *      values' order-preserving bit images, then a segmented count;
 *   2. points ordered by the comparator of main.cpp:86-101 on the rank
 *      triples, which is the order of Morton keys interleaving the ranks with
 *      x the most significant bit of each level (one 64-bit key for ranks
 *      below 2^21, else 96 bits sorted low word first, then high word);
 *   3. the same for the element centroids, summed in connectivity order from
 *      0 and divided by 4 (main.cpp:168-183);
 *   4. coordinates gathered in the new node order, connectivity through the
 *      new node ids, adjacency through the new element ids with -1 kept
 *      (main.cpp:140-151, 214-259).
 * Equal rank triples keep their input order (the reference's std::sort leaves
 * their order unspecified).  Integer and copy work only: HBM-bound sorts. */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "fvc.h"
#include "fvc_internal.h"

namespace fvc {
namespace {

constexpr int ZB = 256;      /* threads per block */
constexpr int ZT = 4096;     /* elements per rank tile */

inline unsigned zgrid(int64_t n) { return (unsigned)((n + ZB - 1) / ZB); }

/* order-preserving unsigned image of a double; -0 equals +0 as under == */
__device__ inline uint64_t ord_key(double v) {
    if (v == 0.0) v = 0.0;
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ inline uint64_t spread21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ __launch_bounds__(ZB) void k_zo_check(const int32_t *__restrict__ conn, const int32_t *__restrict__ adj,
                                                 int64_t ne, int64_t nn, int *__restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x;
    if (i >= 4 * ne) return;
    const int32_t c = conn[i];
    const int32_t a = adj ? adj[i] : -1;
    if (c < 0 || c >= nn || a < -1 || a >= ne) atomicOr(bad, 1);
}

__global__ __launch_bounds__(ZB) void k_zo_centroid(const double *__restrict__ xyz, const int32_t *__restrict__ conn,
                                                    int64_t ne, double *__restrict__ cen) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * ZB + threadIdx.x;
    if (e >= ne) return;
    const int4 nd = reinterpret_cast<const int4 *>(conn)[e];
    const int q[4] = {nd.x, nd.y, nd.z, nd.w};
    for (int a = 0; a < 3; a++) {
        double s = 0.0;
        for (int k = 0; k < 4; k++) s += xyz[3 * (int64_t)q[k] + a];
        cen[3 * e + a] = s / 4.0;
    }
}

__global__ __launch_bounds__(ZB) void k_zo_keys(const double *__restrict__ v, int64_t n, int a,
                                                uint64_t *__restrict__ keys, int *__restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x;
    if (i >= n) return;
    keys[i] = ord_key(v[3 * i + a]);
    vals[i] = (int)i;
}

/* distinct-value starts per tile of sorted keys */
__global__ __launch_bounds__(ZB) void k_zo_rank_count(const uint64_t *__restrict__ keys, int64_t n,
                                                      int *__restrict__ tiles) {
    const int64_t base = (int64_t)blockIdx.x * ZT;
    int cnt = 0;
    for (int q = threadIdx.x; q < ZT; q += ZB) {
        const int64_t j = base + q;
        if (j < n && (j == 0 || keys[j] != keys[j - 1])) cnt++;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    __shared__ int red[ZB / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < ZB / 64; w++) t += red[w];
        tiles[blockIdx.x] = t;
    }
}

/* rank of sorted element j = number of distinct starts at or before j, minus 1 */
__global__ __launch_bounds__(ZB) void k_zo_rank_apply(const uint64_t *__restrict__ keys, const int *__restrict__ vals,
                                                      int64_t n, const int *__restrict__ tile_offs, int a,
                                                      int32_t *__restrict__ rank, int *__restrict__ maxrank) {
    __shared__ int wsum[ZB / 64];
    __shared__ int run;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * ZT;
    if (threadIdx.x == 0) run = tile_offs[blockIdx.x];
    __syncthreads();
    for (int r = 0; r < ZT; r += ZB) {
        const int64_t j = base + r + threadIdx.x;
        const bool valid = j < n;
        const bool start = valid && (j == 0 || keys[j] != keys[j - 1]);
        const unsigned long long bal = __ballot(start);
        const int incl = __popcll(bal & (lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull)));
        if (lane == 0) wsum[w] = __popcll(bal);
        __syncthreads();
        int pre = 0;
        for (int q = 0; q < w; q++) pre += wsum[q];
        if (valid) {
            const int rk = run + pre + incl - 1;
            rank[3 * (int64_t)vals[j] + a] = rk;
            if (j == n - 1) atomicMax(maxrank, rk);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int q = 0; q < ZB / 64; q++) t += wsum[q];
            run += t;
        }
        __syncthreads();
    }
}

/* Morton keys of rank triples, x the most significant bit of each level */
__global__ __launch_bounds__(ZB) void k_zo_morton(const int32_t *__restrict__ rank, int64_t n, int wide,
                                                  uint64_t *__restrict__ lo, uint64_t *__restrict__ hi,
                                                  int *__restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x;
    if (i >= n) return;
    const uint64_t x = (uint32_t)rank[3 * i], y = (uint32_t)rank[3 * i + 1], z = (uint32_t)rank[3 * i + 2];
    vals[i] = (int)i;
    if (!wide) {
        lo[i] = (spread21(x) << 2) | (spread21(y) << 1) | spread21(z);
        return;
    }
    uint64_t l = 0, h = 0;
    for (int b = 0; b < 31; b++) {
        const uint64_t bits[3] = {(x >> b) & 1ull, (y >> b) & 1ull, (z >> b) & 1ull};
        for (int a = 0; a < 3; a++) {
            const int p = 3 * b + 2 - a;
            if (p < 64) l |= bits[a] << p;
            else h |= bits[a] << (p - 64);
        }
    }
    lo[i] = l;
    hi[i] = h;
}

__global__ __launch_bounds__(ZB) void k_zo_gather_u64(const uint64_t *__restrict__ src, const int *__restrict__ idx,
                                                      int64_t n, uint64_t *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ __launch_bounds__(ZB) void k_zo_inverse(const int *__restrict__ order, int64_t n, int *__restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x;
    if (i < n) inv[order[i]] = (int)i;
}

__global__ __launch_bounds__(ZB) void k_zo_rows(const double *__restrict__ in, const int *__restrict__ order, int64_t n,
                                                int width, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x;
    if (i >= n * width) return;
    const int64_t r = i / width, c = i - r * width;
    out[i] = in[(int64_t)order[r] * width + c];
}

__global__ __launch_bounds__(ZB) void k_zo_cells(const int32_t *__restrict__ conn, const int32_t *__restrict__ adj,
                                                 const int *__restrict__ eorder, const int *__restrict__ new_node,
                                                 const int *__restrict__ new_elem, int64_t ne,
                                                 int32_t *__restrict__ conn_out, int32_t *__restrict__ adj_out) {
    const int64_t i = (int64_t)blockIdx.x * ZB + threadIdx.x;
    if (i >= ne) return;
    const int64_t id = eorder[i];
    const int4 c = reinterpret_cast<const int4 *>(conn)[id];
    reinterpret_cast<int4 *>(conn_out)[i] = make_int4(new_node[c.x], new_node[c.y], new_node[c.z], new_node[c.w]);
    if (adj && adj_out) {
        const int4 a = reinterpret_cast<const int4 *>(adj)[id];
        reinterpret_cast<int4 *>(adj_out)[i] =
            make_int4(a.x == -1 ? -1 : new_elem[a.x], a.y == -1 ? -1 : new_elem[a.y], a.z == -1 ? -1 : new_elem[a.z],
                      a.w == -1 ? -1 : new_elem[a.w]);
    }
}

/* device buffers of one renumbering call, released together */
struct ZoBufs {
    hipStream_t s = nullptr;
    static constexpr int CAP = 64;
    void *p[CAP] = {};
    int np = 0;
    ~ZoBufs() {
        if (s) hipStreamSynchronize(s);
        for (int i = 0; i < np; i++) hipFree(p[i]);
        if (s) hipStreamDestroy(s);
    }
    template <class T>
    T *get(size_t count) {
        void *q = nullptr;
        if (np >= CAP || hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        p[np++] = q;
        return (T *)q;
    }
};

#define ZO_OK(call)                                                                                                    \
    do {                                                                                                               \
        if ((call) != hipSuccess) return FVC_E_HIP;                                                                    \
    } while (0)

/* order[0..n) = indices of the n points (xyz, stride 3, device) in Z-order */
int zorder_points(ZoBufs &B, const double *d_xyz, int64_t n, int *d_order) {
    if (n == 0) return FVC_OK;
    hipStream_t s = B.s;
    uint64_t *k0 = B.get<uint64_t>(n), *k1 = B.get<uint64_t>(n);
    int *v0 = B.get<int>(n), *v1 = B.get<int>(n);
    int32_t *rank = B.get<int32_t>(3 * (size_t)n);
    const int nt = (int)((n + ZT - 1) / ZT);
    int *tiles = B.get<int>(nt + 1);
    int *maxr = B.get<int>(1);
    if (!k0 || !k1 || !v0 || !v1 || !rank || !tiles || !maxr) return FVC_E_NOMEM;
    ZO_OK(hipMemsetAsync(maxr, 0, sizeof(int), s));
    for (int a = 0; a < 3; a++) {
        k_zo_keys<<<zgrid(n), ZB, 0, s>>>(d_xyz, n, a, k0, v0);
        int which = 0, rc = radix_sort_pairs(s, k0, v0, k1, v1, n, 64, &which);
        if (rc) return rc;
        const uint64_t *ks = which ? k1 : k0;
        const int *vs = which ? v1 : v0;
        k_zo_rank_count<<<nt, ZB, 0, s>>>(ks, n, tiles);
        if ((rc = scan_exclusive_small(s, tiles, nt))) return rc;
        k_zo_rank_apply<<<nt, ZB, 0, s>>>(ks, vs, n, tiles, a, rank, maxr);
    }
    int mr = 0;
    ZO_OK(hipMemcpyAsync(&mr, maxr, sizeof(int), hipMemcpyDeviceToHost, s));
    ZO_OK(hipStreamSynchronize(s));
    int B1 = 1;
    while (B1 < 31 && ((int64_t)1 << B1) <= mr) B1++; /* bits per axis */
    const int wide = B1 > 21;
    uint64_t *hi = wide ? B.get<uint64_t>(n) : nullptr;
    if (wide && !hi) return FVC_E_NOMEM;
    k_zo_morton<<<zgrid(n), ZB, 0, s>>>(rank, n, wide, k0, hi, v0);
    int which = 0, rc = radix_sort_pairs(s, k0, v0, k1, v1, n, wide ? 64 : 3 * B1, &which);
    if (rc) return rc;
    int *vs = which ? v1 : v0;
    if (wide) { /* then stable by the high word */
        uint64_t *kh = which ? k0 : k1; /* the free key buffer */
        int *vo = which ? v0 : v1;
        k_zo_gather_u64<<<zgrid(n), ZB, 0, s>>>(hi, vs, n, kh);
        int w2 = 0;
        rc = radix_sort_pairs(s, kh, vs, which ? k1 : k0, vo, n, 3 * B1 - 64, &w2);
        if (rc) return rc;
        vs = w2 ? vo : vs;
    }
    ZO_OK(hipMemcpyAsync(d_order, vs, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    return hipGetLastError() == hipSuccess ? FVC_OK : FVC_E_HIP;
}

} // namespace
} // namespace fvc

using namespace fvc;

extern "C" int fvc_mesh_zorder(int64_t num_nodes, const double *coords, int64_t num_elems, const int32_t *conn,
                               const int32_t *adj, int32_t *node_order, int32_t *elem_order, double *coords_out,
                               int32_t *conn_out, int32_t *adj_out) {
    if (num_nodes < 0 || num_elems < 0 || num_nodes > INT32_MAX - 1 || num_elems > INT32_MAX - 1) return FVC_E_INVALID;
    if ((num_nodes > 0 && !coords) || (num_elems > 0 && !conn)) return FVC_E_INVALID;
    if (num_elems > 0 && num_nodes == 0) return FVC_E_INVALID;
    ZoBufs B;
    ZO_OK(hipStreamCreate(&B.s));
    hipStream_t s = B.s;
    const int64_t nn = num_nodes, ne = num_elems;
    double *xyz = B.get<double>(3 * (size_t)nn);
    int32_t *dconn = B.get<int32_t>(4 * (size_t)ne);
    int32_t *dadj = adj ? B.get<int32_t>(4 * (size_t)ne) : nullptr;
    int *norder = B.get<int>(nn), *eorder = B.get<int>(ne), *nnew = B.get<int>(nn), *enew = B.get<int>(ne);
    int *bad = B.get<int>(1);
    if (!xyz || !dconn || (adj && !dadj) || !norder || !eorder || !nnew || !enew || !bad) return FVC_E_NOMEM;
    ZO_OK(hipMemcpyAsync(xyz, coords, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, s));
    ZO_OK(hipMemcpyAsync(dconn, conn, sizeof(int32_t) * 4 * ne, hipMemcpyHostToDevice, s));
    if (adj) ZO_OK(hipMemcpyAsync(dadj, adj, sizeof(int32_t) * 4 * ne, hipMemcpyHostToDevice, s));
    ZO_OK(hipMemsetAsync(bad, 0, sizeof(int), s));
    if (ne > 0) k_zo_check<<<zgrid(4 * ne), ZB, 0, s>>>(dconn, dadj, ne, nn, bad);
    int hbad = 0;
    ZO_OK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    ZO_OK(hipStreamSynchronize(s));
    if (hbad) return FVC_E_INVALID; /* node or neighbour id out of range */
    int rc = zorder_points(B, xyz, nn, norder);
    if (rc) return rc;
    if (ne > 0) {
        double *cen = B.get<double>(3 * (size_t)ne);
        if (!cen) return FVC_E_NOMEM;
        k_zo_centroid<<<zgrid(ne), ZB, 0, s>>>(xyz, dconn, ne, cen);
        if ((rc = zorder_points(B, cen, ne, eorder))) return rc;
    }
    if (nn > 0) k_zo_inverse<<<zgrid(nn), ZB, 0, s>>>(norder, nn, nnew);
    if (ne > 0) k_zo_inverse<<<zgrid(ne), ZB, 0, s>>>(eorder, ne, enew);
    if (node_order) ZO_OK(hipMemcpyAsync(node_order, norder, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, s));
    if (elem_order) ZO_OK(hipMemcpyAsync(elem_order, eorder, sizeof(int32_t) * ne, hipMemcpyDeviceToHost, s));
    if (coords_out && nn > 0) {
        double *xo = B.get<double>(3 * (size_t)nn);
        if (!xo) return FVC_E_NOMEM;
        k_zo_rows<<<zgrid(3 * nn), ZB, 0, s>>>(xyz, norder, nn, 3, xo);
        ZO_OK(hipMemcpyAsync(coords_out, xo, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
    }
    if ((conn_out || adj_out) && ne > 0) {
        int32_t *co = B.get<int32_t>(4 * (size_t)ne), *ao = B.get<int32_t>(4 * (size_t)ne);
        if (!co || !ao) return FVC_E_NOMEM;
        k_zo_cells<<<zgrid(ne), ZB, 0, s>>>(dconn, dadj, eorder, nnew, enew, ne, co, adj_out ? ao : nullptr);
        if (conn_out) ZO_OK(hipMemcpyAsync(conn_out, co, sizeof(int32_t) * 4 * ne, hipMemcpyDeviceToHost, s));
        if (adj_out && adj) ZO_OK(hipMemcpyAsync(adj_out, ao, sizeof(int32_t) * 4 * ne, hipMemcpyDeviceToHost, s));
    }
    ZO_OK(hipGetLastError());
    ZO_OK(hipStreamSynchronize(s));
    return FVC_OK;
}

extern "C" int fvc_gather_rows(int64_t n, int width, const int32_t *order, const double *in, double *out) {
    if (n < 0 || width < 1 || (n > 0 && (!order || !in || !out))) return FVC_E_INVALID;
    if (n == 0) return FVC_OK;
    for (int64_t i = 0; i < n; i++) /* the device gathers only through valid rows */
        if (order[i] < 0 || order[i] >= n) return FVC_E_INVALID;
    ZoBufs B;
    ZO_OK(hipStreamCreate(&B.s));
    hipStream_t s = B.s;
    double *din = B.get<double>((size_t)n * width), *dout = B.get<double>((size_t)n * width);
    int *dord = B.get<int>(n);
    if (!din || !dout || !dord) return FVC_E_NOMEM;
    ZO_OK(hipMemcpyAsync(din, in, sizeof(double) * n * width, hipMemcpyHostToDevice, s));
    ZO_OK(hipMemcpyAsync(dord, order, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    k_zo_rows<<<zgrid(n * width), ZB, 0, s>>>(din, dord, n, width, dout);
    ZO_OK(hipMemcpyAsync(out, dout, sizeof(double) * n * width, hipMemcpyDeviceToHost, s));
    ZO_OK(hipGetLastError());
    ZO_OK(hipStreamSynchronize(s));
    return FVC_OK;
}
