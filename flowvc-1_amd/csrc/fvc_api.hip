/*
 * fvc_api.hip - the C ABI of include/fvc.h.
 *
 * Device layout (per context):
 *   X64  double4[num_nodes]  node coordinates (FP64, exact walk decisions)
 *   ER   float4[4 num_elems] FP32 affine frame of each element (FP32 walk tests)
 *   conn int4[num_elems], adj int4[num_elems]
 *   spatial hash: uniform cells over the padded node bounding box, each cell
 *     listing (ascending) the elements whose padded bounding box meets it
 *   slices: one buffer per resident slot; double4 / float4 / 4 x 16-bit per node
 * Particles are SoA (X,Y,Z,V*,elem,left,left_time,...) for coalesced state I/O.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#include "fvc_internal.h"

using namespace fvc;

/* ------------------------------------------------------------------------ */
/* helpers                                                                   */
/* ------------------------------------------------------------------------ */
namespace fvc {

int check_err(fvc_ctx *c, const char *what) {
    /* one copy of the whole status block (error, location, step counters)
     * into pinned memory; callers read the counters from h_status after it */
    hipError_t e = hipMemcpyAsync(c->h_status, c->d_status, sizeof(fvc_status), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->last_error = std::string(what) + ": " + hipGetErrorString(e);
        return FVC_E_HIP;
    }
    const int code = c->h_status->err;
    if (code) {
        const int64_t where = c->h_status->where;
        int zero = 0;
        hipMemcpy(c->d_err, &zero, sizeof zero, hipMemcpyHostToDevice);
        char buf[256];
        snprintf(buf, sizeof buf, "%s: %s at point %lld", what, fvc_strerror(code), (long long)where);
        c->last_error = buf;
    }
    return code;
}

int side_join(fvc_ctx *c) {
    if (!c->side_pending) return FVC_OK;
    if (hipStreamWaitEvent(c->stream, c->side_done, 0) != hipSuccess) {
        c->last_error = "side stream join failed";
        return FVC_E_HIP;
    }
    c->side_pending = 0;
    return FVC_OK;
}

int side_sync(fvc_ctx *c) {
    if (!c->side) return FVC_OK;
    c->side_pending = 0;
    return hipStreamSynchronize(c->side) == hipSuccess ? FVC_OK : FVC_E_HIP;
}

/* Device -> host copy of a large block into caller (pageable) memory.  The
 * runtime moves pageable destinations through its own small staging buffer,
 * about 12 GB/s on a 168 MB copy; DMA into pinned memory runs at the link's
 * rate.  So the block goes in chunks into the context's pinned staging
 * buffer (grown to the copy's size, at most STAGE_MAX; larger copies go in
 * rounds), an event after each chunk, and host threads copy each chunk out,
 * each its own stripe, as soon as its event has completed: the host copy
 * overlaps the DMA of the chunks behind it; dst is complete on return.
 * Blocks under STAGE_MIN, or all blocks when the pinned buffer cannot be
 * had, are only enqueued on the stream as a direct copy: the caller
 * synchronises the stream before reading them. */
int copy_to_host(fvc_ctx *c, void *dst, const void *src, size_t bytes) {
    constexpr size_t STAGE_MIN = (size_t)4 << 20, STAGE_MAX = (size_t)64 << 20, CHUNK = (size_t)8 << 20;
    if (bytes == 0) return FVC_OK;
    if (bytes >= STAGE_MIN && c->stage_bytes < std::min(bytes, STAGE_MAX)) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) return FVC_E_HIP;
        if (c->stage) hipHostFree(c->stage);
        c->stage = nullptr;
        c->stage_bytes = 0;
        const size_t want = std::min(bytes, STAGE_MAX);
        if (hipHostMalloc((void **)&c->stage, want, hipHostMallocDefault) == hipSuccess)
            c->stage_bytes = want;
        else
            c->stage = nullptr;
    }
    if (bytes < STAGE_MIN || !c->stage) {
        FVC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        return FVC_OK;
    }
    const size_t max_chunks = (c->stage_bytes + CHUNK - 1) / CHUNK;
    while (c->stage_ev.size() < max_chunks) {
        hipEvent_t ev;
        FVC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        c->stage_ev.push_back(ev);
    }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nthr = (int)std::min(8u, std::max(1u, hw / 2));
    for (size_t done = 0; done < bytes;) {
        const size_t round = std::min(bytes - done, c->stage_bytes);
        const size_t nch = (round + CHUNK - 1) / CHUNK;
        for (size_t k = 0; k < nch; k++) {
            const size_t off = k * CHUNK, len = std::min(CHUNK, round - off);
            FVC_HIP(hipMemcpyAsync(c->stage + off, (const char *)src + done + off, len, hipMemcpyDeviceToHost,
                                   c->stream));
            FVC_HIP(hipEventRecord(c->stage_ev[k], c->stream));
        }
        std::atomic<int> failed{0};
        auto worker = [&](int t) {
            for (size_t k = 0; k < nch; k++) {
                const size_t off = k * CHUNK, len = std::min(CHUNK, round - off);
                const size_t lo = (len * t / nthr) & ~(size_t)63, hi = t + 1 == nthr ? len : (len * (t + 1) / nthr) & ~(size_t)63;
                if (hipEventSynchronize(c->stage_ev[k]) != hipSuccess) {
                    failed = 1;
                    return;
                }
                if (hi > lo) memcpy((char *)dst + done + off + lo, c->stage + off + lo, hi - lo);
            }
        };
        std::vector<std::thread> pool;
        try {
            for (int t = 1; t < nthr; t++) pool.emplace_back(worker, t);
        } catch (...) { /* fewer threads: the calling thread copies the rest */
        }
        const int started = (int)pool.size() + 1;
        worker(0);
        for (auto &th : pool) th.join();
        for (int t = started; t < nthr; t++) worker(t);
        if (failed) {
            c->last_error = "device to host copy failed";
            return FVC_E_HIP;
        }
        done += round;
    }
    return FVC_OK;
}

void *scratch(fvc_ctx *c, size_t bytes) {
    if (bytes <= c->scratch_bytes) return c->scratch;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return nullptr;
    if (c->scratch) hipFree(c->scratch);
    c->scratch = nullptr;
    size_t nb = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc(&c->scratch, nb) != hipSuccess) {
        c->scratch_bytes = 0;
        return nullptr;
    }
    c->scratch_bytes = nb;
    return c->scratch;
}

DevField bind_field(fvc_ctx *c, int lo, int hi, double tmin, double tmax) {
    DevField f;
    f.lo = c->slices[lo];
    f.hi = c->slices[hi];
    f.scale_lo = c->slice_scale[lo];
    f.scale_hi = c->slice_scale[hi];
    f.tmin = tmin;
    f.tmax = tmax;
    return f;
}

} // namespace fvc

static int set_err(fvc_ctx *c, int code, const char *msg) {
    if (c) c->last_error = msg;
    return code;
}

extern "C" const char *fvc_strerror(int code) {
    switch (code) {
    case FVC_OK: return "ok";
    case FVC_E_INVALID: return "invalid argument";
    case FVC_E_NOMEM: return "out of memory";
    case FVC_E_HIP: return "HIP runtime error";
    case FVC_E_INDETERMINATE_NEIGHBOR: return "Indeterminate neighbor";
    case FVC_E_TLOC_RANGE: return "tloc must be between 0 and 1";
    case FVC_E_ELEM_NEG1: return "Attempting to interpolate velocity at point with Element_Index = -1";
    case FVC_E_WALK_LIMIT: return "walk exceeded its hop limit";
    case FVC_E_EIG: return "Too many iterations in eigenvalue solver";
    case FVC_E_ADMISSION: return "request does not fit free device memory";
    case FVC_E_STATE: return "call out of sequence";
    case FVC_E_LEFT: return "Attempting to integrate point that already left domain";
    case FVC_E_NODEVICE: return "no usable GPU device";
    case FVC_E_CET_OUTSIDE: return "Point X0 passed to ComputeExposureTime appears to be outside of velocity domain";
    default: return "unknown error";
    }
}

extern "C" const char *fvc_last_error(const fvc_ctx *c) { return c ? c->last_error.c_str() : "no context"; }

template <class T>
static int dev_alloc(fvc_ctx *c, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, sizeof(T) * count) != hipSuccess) {
        c->last_error = "hipMalloc failed";
        return FVC_E_NOMEM;
    }
    c->bytes_used += (int64_t)(sizeof(T) * count);
    return FVC_OK;
}

/* ------------------------------------------------------------------------ */
/* spatial hash (host build, ascending element lists)                        */
/* ------------------------------------------------------------------------ */
static int build_hash(fvc_ctx *c, const fvc_mesh_desc *md) {
    const int dims = md->dims;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t n = 0; n < md->num_nodes; n++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], md->coords[3 * n + a]);
            hi[a] = std::max(hi[a], md->coords[3 * n + a]);
        }
    double ext = 0.0;
    for (int a = 0; a < dims; a++) ext = std::max(ext, hi[a] - lo[a]);
    const double pad = 1e-6 * ext + 1e-12;
    for (int a = 0; a < 3; a++) {
        lo[a] -= pad;
        hi[a] += pad;
    }
    if (dims == 2) {
        lo[2] = -1.0;
        hi[2] = 1.0;
    }
    double vol = 1.0;
    for (int a = 0; a < dims; a++) vol *= (hi[a] - lo[a]);
    /* cell edge ~ mean element size, offset from integer multiples to avoid
     * aligning with structured meshes */
    double h = 1.17 * pow(vol / (double)std::max<int64_t>(md->num_elems, 1), 1.0 / dims);
    int res[3] = {1, 1, 1};
    for (int a = 0; a < dims; a++) {
        res[a] = (int)std::max(1.0, std::min(4096.0, ceil((hi[a] - lo[a]) / h)));
    }
    if (c->mesh.cres[0] < 0) { /* unused */
    }
    int64_t ncells = (int64_t)res[0] * res[1] * res[2];
    double inv[3];
    for (int a = 0; a < 3; a++) inv[a] = res[a] / (hi[a] - lo[a]);
    const int nn = dims == 3 ? 4 : 3;
    auto cell_range = [&](int64_t e, int clo[3], int chi[3]) {
        double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
        for (int q = 0; q < nn; q++) {
            int64_t v = md->conn[4 * e + q];
            for (int a = 0; a < 3; a++) {
                blo[a] = std::min(blo[a], md->coords[3 * v + a]);
                bhi[a] = std::max(bhi[a], md->coords[3 * v + a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            if (a >= dims) {
                clo[a] = chi[a] = 0;
                continue;
            }
            clo[a] = std::max(0, (int)floor((blo[a] - pad - lo[a]) * inv[a]));
            chi[a] = std::min(res[a] - 1, (int)floor((bhi[a] + pad - lo[a]) * inv[a]));
        }
    };
    std::vector<int> start(ncells + 1, 0);
    int clo[3], chi[3];
    for (int64_t e = 0; e < md->num_elems; e++) {
        cell_range(e, clo, chi);
        for (int k = clo[2]; k <= chi[2]; k++)
            for (int j = clo[1]; j <= chi[1]; j++)
                for (int i = clo[0]; i <= chi[0]; i++) start[((int64_t)k * res[1] + j) * res[0] + i + 1]++;
    }
    for (int64_t q = 0; q < ncells; q++) start[q + 1] += start[q];
    int64_t total = start[ncells];
    if (total > INT32_MAX) return set_err(c, FVC_E_NOMEM, "spatial hash too large");
    std::vector<int> fill(start.begin(), start.end() - 1);
    std::vector<int> elems(total);
    for (int64_t e = 0; e < md->num_elems; e++) { /* ascending e -> ascending lists */
        cell_range(e, clo, chi);
        for (int k = clo[2]; k <= chi[2]; k++)
            for (int j = clo[1]; j <= chi[1]; j++)
                for (int i = clo[0]; i <= chi[0]; i++) elems[fill[((int64_t)k * res[1] + j) * res[0] + i]++] = (int)e;
    }
    int rc;
    if ((rc = dev_alloc(c, &c->cell_start, ncells + 1))) return rc;
    if ((rc = dev_alloc(c, &c->cell_elems, total))) return rc;
    if (hipMemcpy(c->cell_start, start.data(), sizeof(int) * (ncells + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->cell_elems, elems.data(), sizeof(int) * std::max<int64_t>(total, 1), hipMemcpyHostToDevice) !=
            hipSuccess)
        return set_err(c, FVC_E_HIP, "hash upload failed");
    c->ncells = ncells;
    c->nentries = total;
    c->mesh.cell_start = c->cell_start;
    c->mesh.cell_elems = c->cell_elems;
    for (int a = 0; a < 3; a++) {
        c->mesh.hlo[a] = lo[a];
        c->mesh.hinv[a] = inv[a];
        c->mesh.hres[a] = res[a];
    }
    return FVC_OK;
}

/* FP32 affine frame of element e (the layout fvc_device.h's ld_frame
 * reads: rows (M[a][0..2], o[a]), then (lo[0..2], g)) into R[0..3], from the
 * FP64 nodes relative to the mesh centre; returns the row-sum norm of M (the
 * 1-r-s-t row included), or 0 for an element without a finite frame. */
static double frame_of(const fvc_mesh_desc *md, const double centre[3], int64_t e, float4 *R) {
    const int dims = md->dims;
    const double u = ldexp(1.0, -24);
    const int32_t *nd = md->conn + 4 * e;
    double P[4][3] = {{0}};
    for (int q = 0; q <= dims; q++)
        for (int a = 0; a < dims; a++) P[q][a] = md->coords[3 * nd[q] + a] - centre[a];
    double E[3][3] = {{0}}, M[3][3] = {{0}}; /* E[a][q] = (node q+1 - node 0)[a] */
    for (int q = 0; q < dims; q++)
        for (int a = 0; a < dims; a++) E[a][q] = P[q + 1][a] - P[0][a];
    double det;
    if (dims == 3) {
        det = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) - E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
              E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
        M[0][0] = (E[1][1] * E[2][2] - E[1][2] * E[2][1]) / det;
        M[0][1] = (E[0][2] * E[2][1] - E[0][1] * E[2][2]) / det;
        M[0][2] = (E[0][1] * E[1][2] - E[0][2] * E[1][1]) / det;
        M[1][0] = (E[1][2] * E[2][0] - E[1][0] * E[2][2]) / det;
        M[1][1] = (E[0][0] * E[2][2] - E[0][2] * E[2][0]) / det;
        M[1][2] = (E[0][2] * E[1][0] - E[0][0] * E[1][2]) / det;
        M[2][0] = (E[1][0] * E[2][1] - E[1][1] * E[2][0]) / det;
        M[2][1] = (E[0][1] * E[2][0] - E[0][0] * E[2][1]) / det;
        M[2][2] = (E[0][0] * E[1][1] - E[0][1] * E[1][0]) / det;
    } else {
        det = E[0][0] * E[1][1] - E[0][1] * E[1][0];
        M[0][0] = E[1][1] / det;
        M[0][1] = -E[0][1] / det;
        M[1][0] = -E[1][0] / det;
        M[1][1] = E[0][0] / det;
    }
    double rown = 0.0, enorm = 0.0, l0row = 0.0;
    for (int i = 0; i < dims; i++) {
        double rs = 0.0, es = 0.0;
        for (int j = 0; j < dims; j++) {
            rs += fabs(M[i][j]);
            es += fabs(E[i][j]);
        }
        rown = std::max(rown, rs);
        enorm = std::max(enorm, es);
    }
    for (int j = 0; j < dims; j++) {
        double cs = 0.0;
        for (int i = 0; i < dims; i++) cs += M[i][j];
        l0row += fabs(cs);
    }
    rown = std::max(rown, l0row);
    float o[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < dims; a++) o[a] = (float)P[0][a];
    double lo[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < dims; i++)
        for (int a = 0; a < dims; a++) lo[i] += M[i][a] * ((double)o[a] - P[0][a]);
    const double g = 4.0 * (18.0 * rown * enorm + 12.0) * u;
    bool ok = det != 0.0 && std::isfinite(g) && g < 0.25;
    for (int i = 0; i < dims && ok; i++) {
        ok = std::isfinite(lo[i]) && fabs(lo[i]) < 1e30;
        for (int j = 0; j < dims; j++) ok = ok && fabs(M[i][j]) < 1e30;
    }
    if (!ok) { /* every test on this element is decided in FP64 */
        R[3].w = INFINITY;
        return 0.0;
    }
    R[0] = make_float4((float)M[0][0], (float)M[0][1], (float)M[0][2], o[0]);
    R[1] = make_float4((float)M[1][0], (float)M[1][1], (float)M[1][2], o[1]);
    R[2] = make_float4((float)M[2][0], (float)M[2][1], (float)M[2][2], o[2]);
    R[3] = make_float4((float)lo[0], (float)lo[1], (float)lo[2], (float)g);
    return rown;
}

/* all frames, over up to 16 host threads; returns the largest row norm */
static double build_frames(const fvc_mesh_desc *md, const double centre[3], std::vector<float4> &ER) {
    ER.assign((size_t)4 * md->num_elems, make_float4(0.f, 0.f, 0.f, 0.f));
    const int64_t ne = md->num_elems;
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)std::max(1u, std::thread::hardware_concurrency()),
                                                                16, ne / 65536 + 1}));
    std::vector<double> part(nt, 0.0);
    auto work = [&](int t) {
        double mx = 0.0;
        for (int64_t e = ne * t / nt; e < ne * (t + 1) / nt; e++)
            mx = std::max(mx, frame_of(md, centre, e, ER.data() + 4 * e));
        part[t] = mx;
    };
    std::vector<std::thread> pool;
    int started = 1;
    try {
        for (; started < nt; started++) pool.emplace_back(work, started);
    } catch (...) { /* no more threads: the rest runs here */
    }
    for (int t = started; t < nt; t++) work(t);
    work(0);
    for (auto &th : pool) th.join();
    return *std::max_element(part.begin(), part.end());
}

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */
static int noalloc_scratch(int device, void **ptr);

extern "C" int fvc_create(const fvc_mesh_desc *md, const fvc_params *pp, fvc_ctx **out) {
    if (!md || !pp || !out) return FVC_E_INVALID;
    *out = nullptr;
    if ((md->dims != 2 && md->dims != 3) || (md->mesh_type != 0 && md->mesh_type != 1)) return FVC_E_INVALID;
    if (pp->precision < 0 || pp->precision > 3 || pp->max_slices < 2) return FVC_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= pp->device || pp->device < 0) return FVC_E_NODEVICE;
    if (hipSetDevice(pp->device) != hipSuccess) return FVC_E_NODEVICE;
    fvc_ctx *c = new fvc_ctx();
    c->device = pp->device;
    c->precision = pp->precision;
    c->max_slices = pp->max_slices;
    c->slices.assign(pp->max_slices, nullptr);
    c->slice_scale.assign(pp->max_slices, 1.0f);
    c->slice_time.assign(pp->max_slices, 0.0);
    c->sr_slices.assign(pp->max_slices, nullptr);
    memset(&c->mesh, 0, sizeof c->mesh);
    DevMesh &m = c->mesh;
    m.dims = md->dims;
    m.mesh_type = md->mesh_type;
    for (int a = 0; a < 6; a++) m.bounds[a] = md->bounds[a];
    m.walk_limit = pp->walk_limit > 0 ? pp->walk_limit : (1 << 20);
    int rc = FVC_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return FVC_E_HIP;
    }
    if ((rc = dev_alloc(c, &c->d_status, 1))) {
        fvc_destroy(c);
        return rc;
    }
    c->d_err = &c->d_status->err;
    c->d_where = &c->d_status->where;
    c->d_steps = c->d_status->steps;
    if (hipHostMalloc((void **)&c->h_status, sizeof(fvc_status), hipHostMallocDefault) != hipSuccess) {
        c->h_status = nullptr;
        fvc_destroy(c);
        return FVC_E_NOMEM;
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->side_ready, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->side_done, hipEventDisableTiming) != hipSuccess) {
        fvc_destroy(c);
        return FVC_E_HIP;
    }
    {
        void *sp = nullptr; /* the no-allocation staging of this device (fvc_locate_noalloc) */
        if (noalloc_scratch(pp->device, &sp)) {
            fvc_destroy(c);
            return FVC_E_HIP;
        }
    }
    if (hipMemset(c->d_status, 0, sizeof(fvc_status)) != hipSuccess) {
        fvc_destroy(c);
        return FVC_E_HIP;
    }
    if (md->mesh_type == 0) {
        /* Cartesian grid (mesh.c:96-121) */
        for (int a = 0; a < 3; a++) {
            m.cmin[a] = md->cart_min[a];
            m.cres[a] = md->cart_res[a];
        }
        if (m.cres[0] < 2 || m.cres[1] < 2 || md->cart_min[0] >= md->cart_max[0] || md->cart_min[1] >= md->cart_max[1] ||
            (md->dims == 3 && (m.cres[2] < 2 || md->cart_min[2] >= md->cart_max[2]))) {
            fvc_destroy(c);
            return FVC_E_INVALID;
        }
        m.cdel[0] = (md->cart_max[0] - md->cart_min[0]) / (md->cart_res[0] - 1);
        m.cdel[1] = (md->cart_max[1] - md->cart_min[1]) / (md->cart_res[1] - 1);
        if (md->dims == 3)
            m.cdel[2] = (md->cart_max[2] - md->cart_min[2]) / (md->cart_res[2] - 1);
        else {
            m.cmin[2] = 0;
            m.cdel[2] = 0;
            m.cres[2] = 1;
        }
        c->nvel = (int64_t)m.cres[0] * m.cres[1] * m.cres[2];
        /* strain-rate cells (strainrate.c:86-91): kmax = 1 in 2-D or for a single z layer */
        c->nsr = (int64_t)(m.cres[0] - 1) * (m.cres[1] - 1) * ((md->dims == 2 || m.cres[2] == 1) ? 1 : m.cres[2] - 1);
    } else {
        /* device gathers address node arrays (32 B per node at most) and
         * element arrays (64 B per element at most) with 32-bit byte offsets */
        if (md->num_nodes <= 0 || md->num_elems <= 0 || !md->coords || !md->conn || !md->adj ||
            md->num_elems > (INT64_C(1) << 32) / 64 || md->num_nodes > (INT64_C(1) << 32) / 32) {
            fvc_destroy(c);
            return set_err(nullptr, FVC_E_INVALID, "");
        }
        const int nn = md->dims == 3 ? 4 : 3;
        for (int64_t e = 0; e < md->num_elems; e++)
            for (int q = 0; q < 4; q++) {
                int32_t v = md->conn[4 * e + q], a = md->adj[4 * e + q];
                if ((q < nn && (v < 0 || v >= md->num_nodes)) || a < -1 || a >= md->num_elems) {
                    fvc_destroy(c);
                    return set_err(nullptr, FVC_E_INVALID, "");
                }
            }
        m.nn = md->num_nodes;
        m.ne = md->num_elems;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t n = 0; n < md->num_nodes; n++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], md->coords[3 * n + a]);
                hi[a] = std::max(hi[a], md->coords[3 * n + a]);
            }
        for (int a = 0; a < 3; a++) m.centre[a] = 0.5 * (lo[a] + hi[a]);
        std::vector<double4> x64(md->num_nodes);
        for (int64_t n = 0; n < md->num_nodes; n++) {
            const double *p = md->coords + 3 * n;
            x64[n] = make_double4(p[0], p[1], p[2], 0.0);
        }
        std::vector<float4> er;
        const double maxrow = build_frames(md, m.centre, er);
        std::vector<int4> cn(md->num_elems), ad(md->num_elems);
        for (int64_t e = 0; e < md->num_elems; e++) {
            const int32_t *q = md->conn + 4 * e, *a = md->adj + 4 * e;
            /* 2-D meshes store -1 in the fourth slot; keep a valid id for loads */
            cn[e] = make_int4(q[0], q[1], q[2], nn == 4 ? q[3] : q[0]);
            ad[e] = make_int4(a[0], a[1], a[2], nn == 4 ? a[3] : -1);
        }
        if ((rc = dev_alloc(c, &c->X64, md->num_nodes)) || (rc = dev_alloc(c, &c->ER, 4 * md->num_elems)) ||
            (rc = dev_alloc(c, &c->conn, md->num_elems)) || (rc = dev_alloc(c, &c->adj, md->num_elems))) {
            fvc_destroy(c);
            return rc;
        }
        if (hipMemcpy(c->X64, x64.data(), sizeof(double4) * md->num_nodes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->ER, er.data(), sizeof(float4) * 4 * md->num_elems, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->conn, cn.data(), sizeof(int4) * md->num_elems, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->adj, ad.data(), sizeof(int4) * md->num_elems, hipMemcpyHostToDevice) != hipSuccess) {
            fvc_destroy(c);
            return FVC_E_HIP;
        }
        m.X64 = c->X64;
        m.ER = c->ER;
        m.conn = c->conn;
        m.adj = c->adj;
        m.gq = (float)(12.0 * ldexp(1.0, -24) * maxrow);
        if ((rc = build_hash(c, md))) {
            fvc_destroy(c);
            return rc;
        }
        c->nvel = md->num_nodes;
        c->nsr = md->num_elems;
    }
    c->vel_stride = c->precision == FVC_PREC_F64 ? 32 : (c->precision == FVC_PREC_F32 ? 16 : 8);
    if (hipDeviceSynchronize() != hipSuccess) {
        fvc_destroy(c);
        return FVC_E_HIP;
    }
    *out = c;
    return FVC_OK;
}

extern "C" void fvc_destroy(fvc_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    side_sync(c);
    if (c->stream) hipStreamSynchronize(c->stream);
    /* objects still alive on the context: servers and batches first (they
     * hold particle sets), then FTLE slides, then particle sets.  Each release
     * drops the object (and the objects it owns) from these sets. */
    while (!c->live_servers.empty()) server_release(*c->live_servers.begin());
    while (!c->live_batches.empty()) batch_release(*c->live_batches.begin());
    while (!c->live_ftle.empty()) ftle_release(*c->live_ftle.begin());
    while (!c->live_particles.empty()) particles_release(*c->live_particles.begin());
    for (void *s : c->slices)
        if (s) hipFree(s);
    for (double *s : c->sr_slices)
        if (s) hipFree(s);
    if (c->cet_lo) hipFree(c->cet_lo);
    if (c->cet_enc) hipFree(c->cet_enc);
    hipFree(c->X64);
    hipFree(c->ER);
    hipFree(c->conn);
    hipFree(c->adj);
    hipFree(c->cell_start);
    hipFree(c->cell_elems);
    hipFree(c->d_status);
    if (c->h_status) hipHostFree(c->h_status);
    if (c->stage) hipHostFree(c->stage);
    for (hipEvent_t ev : c->stage_ev) hipEventDestroy(ev);
    if (c->scratch) hipFree(c->scratch);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->side_ready) hipEventDestroy(c->side_ready);
    if (c->side_done) hipEventDestroy(c->side_done);
    if (c->side) hipStreamDestroy(c->side);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int fvc_set_timing(fvc_ctx *c, int32_t enable) {
    if (!c) return FVC_E_INVALID;
    c->timing = enable ? 1 : 0;
    return FVC_OK;
}

extern "C" int fvc_set_chain_mode(fvc_ctx *c, int32_t mode) {
    if (!c || mode < 0 || mode > 1) return FVC_E_INVALID;
    c->chain_replay = mode;
    return FVC_OK;
}

extern "C" int fvc_kernel_time(fvc_ctx *c, double *ms, int64_t *launches, int64_t *steps, int32_t reset) {
    if (!c) return FVC_E_INVALID;
    if (ms) *ms = c->t_ms;
    if (launches) *launches = c->t_launches;
    if (steps) *steps = c->t_steps;
    if (reset) {
        c->t_ms = 0.0;
        c->t_launches = 0;
        c->t_steps = 0;
    }
    return FVC_OK;
}

extern "C" int fvc_copy_probe(fvc_ctx *c, int64_t bytes) {
    if (!c || bytes <= 0) return FVC_E_INVALID;
    void *a = nullptr, *b = nullptr;
    FVC_HIP(hipMalloc(&a, (size_t)bytes));
    if (hipMalloc(&b, (size_t)bytes) != hipSuccess) {
        hipFree(a);
        return FVC_E_NOMEM;
    }
    hipMemsetAsync(a, 1, (size_t)bytes, c->stream);
    int rc = launch_copy_probe(c, b, a, bytes);
    hipFree(a);
    hipFree(b);
    return rc;
}

extern "C" int fvc_sync(fvc_ctx *c) {
    if (!c) return FVC_E_INVALID;
    if (side_join(c)) return FVC_E_HIP;
    return hipStreamSynchronize(c->stream) == hipSuccess ? FVC_OK : FVC_E_HIP;
}

extern "C" int fvc_memory(fvc_ctx *c, int64_t *used, int64_t *free_bytes) {
    if (!c) return FVC_E_INVALID;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return FVC_E_HIP;
    if (used) *used = c->bytes_used;
    if (free_bytes) *free_bytes = (int64_t)fr;
    return FVC_OK;
}

/* ------------------------------------------------------------------------ */
/* slices                                                                    */
/* ------------------------------------------------------------------------ */
static inline uint16_t to_bf16(float v) { /* round to nearest even */
    uint32_t u;
    memcpy(&u, &v, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static inline uint16_t to_f16(float v) {
    _Float16 h = (_Float16)v;
    uint16_t r;
    memcpy(&r, &h, 2);
    return r;
}

extern "C" int fvc_load_slice(fvc_ctx *c, int32_t slot, double timestamp, const double *uvw, int64_t n) {
    if (!c || !uvw || slot < 0 || slot >= c->max_slices) return FVC_E_INVALID;
    if (n != c->nvel) return set_err(c, FVC_E_INVALID, "slice length differs from the node count");
    if (side_join(c)) return FVC_E_HIP; /* an early FTLE may still read the slot */
    if (!c->slices[slot]) {
        void *p = nullptr;
        if (hipMalloc(&p, c->vel_stride * (size_t)n) != hipSuccess) return set_err(c, FVC_E_NOMEM, "slice alloc");
        c->bytes_used += (int64_t)(c->vel_stride * (size_t)n);
        c->slices[slot] = p;
    }
    std::vector<unsigned char> host(c->vel_stride * (size_t)n);
    float scale = 1.0f;
    if (c->precision == FVC_PREC_F64) {
        double4 *h = (double4 *)host.data();
        for (int64_t i = 0; i < n; i++) h[i] = make_double4(uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2], 0.0);
    } else if (c->precision == FVC_PREC_F32) {
        float4 *h = (float4 *)host.data();
        for (int64_t i = 0; i < n; i++)
            h[i] = make_float4((float)uvw[3 * i], (float)uvw[3 * i + 1], (float)uvw[3 * i + 2], 0.0f);
    } else {
        double mx = 0.0;
        for (int64_t i = 0; i < 3 * n; i++) mx = std::max(mx, fabs(uvw[i]));
        scale = mx > 0.0 ? (float)mx : 1.0f;
        const double inv = 1.0 / (double)scale;
        uint16_t *h = (uint16_t *)host.data();
        for (int64_t i = 0; i < n; i++)
            for (int q = 0; q < 4; q++) {
                float v = q < 3 ? (float)(uvw[3 * i + q] * inv) : 0.0f;
                h[4 * i + q] = c->precision == FVC_PREC_BF16 ? to_bf16(v) : to_f16(v);
            }
    }
    if (hipMemcpyAsync(c->slices[slot], host.data(), host.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return set_err(c, FVC_E_HIP, "slice upload");
    c->slice_scale[slot] = scale;
    c->slice_time[slot] = timestamp;
    return FVC_OK;
}

extern "C" int fvc_load_strain_slice(fvc_ctx *c, int32_t slot, double timestamp, const double *sr, int64_t n) {
    (void)timestamp; /* read but unused, as the reference's (strainrate.c:112-115) */
    if (!c || !sr || slot < 0 || slot >= c->max_slices) return FVC_E_INVALID;
    if (n != c->nsr) return set_err(c, FVC_E_INVALID, "strain-rate slice length differs from the element (cell) count");
    if (side_join(c)) return FVC_E_HIP;
    if (!c->sr_slices[slot]) {
        if (hipMalloc((void **)&c->sr_slices[slot], sizeof(double) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess)
            return set_err(c, FVC_E_NOMEM, "strain-rate slice alloc");
        c->bytes_used += (int64_t)(sizeof(double) * (size_t)n);
    }
    if (hipMemcpyAsync(c->sr_slices[slot], sr, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return set_err(c, FVC_E_HIP, "strain-rate slice upload");
    return FVC_OK;
}

extern "C" int fvc_cet_reset(fvc_ctx *c) {
    if (!c) return FVC_E_INVALID;
    if (c->mesh.mesh_type != 1) return set_err(c, FVC_E_INVALID, "exposure times need an unstructured mesh");
    const size_t ne = (size_t)c->mesh.ne;
    if (!c->cet_lo) {
        if (hipMalloc((void **)&c->cet_lo, 2 * sizeof(unsigned long long) * ne) != hipSuccess ||
            hipMalloc((void **)&c->cet_enc, sizeof(int) * ne) != hipSuccess)
            return set_err(c, FVC_E_NOMEM, "exposure-time arrays");
        c->cet_hi = c->cet_lo + ne;
        c->bytes_used += (int64_t)(20 * ne);
    }
    FVC_HIP(hipMemsetAsync(c->cet_lo, 0, 2 * sizeof(unsigned long long) * ne, c->stream));
    FVC_HIP(hipMemsetAsync(c->cet_enc, 0, sizeof(int) * ne, c->stream));
    FVC_HIP(hipStreamSynchronize(c->stream));
    return FVC_OK;
}

extern "C" int fvc_cet_launch(fvc_ctx *c, fvc_particles *p, int64_t begin, int64_t count) {
    if (!c || !p || p->ctx != c || begin < 0 || count < 0 || begin + count > p->cap) return FVC_E_INVALID;
    if (!c->cet_lo) return set_err(c, FVC_E_STATE, "fvc_cet_reset not called");
    int rc = launch_cet_encounters(c, ptrs(p), begin, count, c->cet_enc);
    return rc ? rc : check_err(c, "fvc_cet_launch");
}

extern "C" int fvc_cet_get(fvc_ctx *c, double *cetsum, int32_t *encounters) {
    if (!c) return FVC_E_INVALID;
    if (!c->cet_lo) return set_err(c, FVC_E_STATE, "fvc_cet_reset not called");
    const size_t ne = (size_t)c->mesh.ne;
    if (cetsum) {
        std::vector<unsigned long long> w(2 * ne);
        FVC_HIP(hipMemcpy(w.data(), c->cet_lo, sizeof(unsigned long long) * 2 * ne, hipMemcpyDeviceToHost));
        for (size_t e = 0; e < ne; e++) { /* exact 128-bit sum, rounded once */
            const unsigned __int128 v = ((unsigned __int128)w[ne + e] << 64) | w[e];
            cetsum[e] = ldexp((double)v, -CET_FRAC);
        }
    }
    if (encounters) FVC_HIP(hipMemcpy(encounters, c->cet_enc, sizeof(int) * ne, hipMemcpyDeviceToHost));
    return FVC_OK;
}

/* strain-rate slices lo, hi (activation potential) and the exposure-time
 * accumulators into the field when requested */
static int bind_strain(fvc_ctx *c, int lo, int hi, const fvc_int_params *ip, DevField &f, PtrSet &ps,
                       fvc_particles *p) {
    if (ip->exposure_substeps) {
        if (ip->exposure_substeps < 2) return set_err(c, FVC_E_INVALID, "Trace_CETSubsteps must be greater than 1");
        if (c->mesh.mesh_type != 1) return set_err(c, FVC_E_INVALID, "exposure times need an unstructured mesh");
        if (!c->cet_lo) return set_err(c, FVC_E_STATE, "fvc_cet_reset not called");
        f.cet_lo = c->cet_lo;
        f.cet_hi = c->cet_hi;
        f.cet_enc = c->cet_enc;
        f.cet_sub = ip->exposure_substeps;
    }
    if (!ip->activation_potential) return FVC_OK;
    if (!c->sr_slices[lo] || !c->sr_slices[hi]) return set_err(c, FVC_E_STATE, "strain-rate slot not loaded");
    f.sr_lo = c->sr_slices[lo];
    f.sr_hi = c->sr_slices[hi];
    ps.scalar = p->scalar;
    return FVC_OK;
}

static int check_slots(fvc_ctx *c, int lo, int hi, double tmin, double tmax) {
    if (lo < 0 || hi < 0 || lo >= c->max_slices || hi >= c->max_slices || !c->slices[lo] || !c->slices[hi])
        return set_err(c, FVC_E_STATE, "velocity slot not loaded");
    if (!(tmax > tmin)) return set_err(c, FVC_E_INVALID, "empty loaded interval");
    return FVC_OK;
}

/* ------------------------------------------------------------------------ */
/* particles                                                                 */
/* ------------------------------------------------------------------------ */
extern "C" int fvc_particles_create(fvc_ctx *c, int64_t cap, fvc_particles **out) {
    if (!c || !out || cap <= 0) return FVC_E_INVALID;
    *out = nullptr;
    fvc_particles *p = new fvc_particles();
    p->ctx = c;
    p->cap = cap;
    size_t bytes = (size_t)cap * (14 * 8 + 4 * 4) + 256;
    if (hipMalloc(&p->buf, bytes) != hipSuccess) {
        delete p;
        return set_err(c, FVC_E_NOMEM, "particle alloc");
    }
    c->live_particles.insert(p);
    c->bytes_used += (int64_t)bytes;
    double *d = (double *)p->buf;
    for (int a = 0; a < 3; a++) p->X[a] = d + a * cap;
    for (int a = 0; a < 3; a++) p->V[a] = d + (3 + a) * cap;
    p->left_time = d + 6 * cap;
    p->scalar = d + 7 * cap;
    for (int a = 0; a < 6; a++) p->H[a] = d + (8 + a) * cap;
    int *iv = (int *)(d + 14 * cap);
    p->elem = iv;
    p->aux = iv + cap;
    p->left = iv + 2 * cap;
    p->nhist = iv + 3 * cap;
    if (hipMemsetAsync(p->buf, 0, bytes, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        fvc_particles_destroy(p);
        return set_err(c, FVC_E_HIP, "particle clear");
    }
    *out = p;
    return FVC_OK;
}

void fvc::particles_release(fvc_particles *p) {
    fvc_ctx *c = p->ctx;
    if (!c) return;
    batches_release_on(c, p); /* a batch advances p: it goes first */
    hipStreamSynchronize(c->stream);
    hipFree(p->buf);
    if (p->act) hipFree(p->act);
    p->buf = nullptr;
    p->act = nullptr;
    p->act_begin = -1;
    c->live_particles.erase(p);
    p->ctx = nullptr;
}

extern "C" void fvc_particles_destroy(fvc_particles *p) {
    if (!p) return;
    particles_release(p);
    delete p;
}

extern "C" int fvc_particles_active(const fvc_particles *p, int64_t *listed) {
    if (!p || !listed) return FVC_E_INVALID;
    *listed = p->act_begin >= 0 ? p->act_n : -1;
    return FVC_OK;
}

extern "C" int64_t fvc_particles_capacity(const fvc_particles *p) { return p ? p->cap : 0; }

extern "C" int fvc_particles_set(fvc_particles *p, int64_t begin, const fvc_point *recs, int64_t n) {
    if (!p || !p->ctx || !recs || begin < 0 || n < 0 || begin + n > p->cap) return FVC_E_INVALID;
    fvc_ctx *c = p->ctx;
    /* element ids start walks on the device: outside [-1, ne) they would
     * address memory beyond the mesh (-1 is ompVC's "no element" and raises
     * FVC_E_ELEM_NEG1 when such a point is integrated).  On points flagged as
     * left an id outside that range is stored as -1: ompVC leaves the id of
     * points outside the domain unset in its stored grids (ftle.c:425-455). */
    const bool uns = c->mesh.mesh_type == 1;
    if (uns)
        for (int64_t i = 0; i < n; i++)
            if (!recs[i].LeftDomain && (recs[i].ElementIndex < -1 || recs[i].ElementIndex >= c->mesh.ne))
                return set_err(c, FVC_E_INVALID, "particle record with an element id outside the mesh");
    p->act_begin = -1; /* written state may bring particles back */
    std::vector<double> d(8 * (size_t)n);
    std::vector<int> iv(3 * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            d[a * n + i] = recs[i].X[a];
            d[(3 + a) * n + i] = recs[i].V[a];
        }
        d[6 * n + i] = recs[i].LeftDomainTime;
        d[7 * n + i] = recs[i].Scalar;
        const int e = recs[i].ElementIndex;
        iv[i] = (uns && recs[i].LeftDomain && (e < -1 || e >= c->mesh.ne)) ? -1 : e;
        iv[n + i] = recs[i].AuxElementIndex;
        iv[2 * n + i] = recs[i].LeftDomain;
    }
    double *dst[8] = {p->X[0], p->X[1], p->X[2], p->V[0], p->V[1], p->V[2], p->left_time, p->scalar};
    for (int a = 0; a < 8; a++)
        FVC_HIP(hipMemcpyAsync(dst[a] + begin, d.data() + a * n, 8 * n, hipMemcpyHostToDevice, c->stream));
    int *idst[3] = {p->elem, p->aux, p->left};
    for (int a = 0; a < 3; a++)
        FVC_HIP(hipMemcpyAsync(idst[a] + begin, iv.data() + a * n, 4 * n, hipMemcpyHostToDevice, c->stream));
    /* new states start without multistep history */
    FVC_HIP(hipMemsetAsync(p->nhist + begin, 0, 4 * n, c->stream));
    FVC_HIP(hipStreamSynchronize(c->stream));
    return FVC_OK;
}

extern "C" int fvc_particles_get(fvc_particles *p, int64_t begin, fvc_point *recs, int64_t n) {
    if (!p || !p->ctx || !recs || begin < 0 || n < 0 || begin + n > p->cap) return FVC_E_INVALID;
    fvc_ctx *c = p->ctx;
    std::vector<double> d(8 * (size_t)n);
    std::vector<int> iv(3 * (size_t)n);
    double *src[8] = {p->X[0], p->X[1], p->X[2], p->V[0], p->V[1], p->V[2], p->left_time, p->scalar};
    int rc = FVC_OK;
    for (int a = 0; a < 8 && !rc; a++) rc = copy_to_host(c, d.data() + a * n, src[a] + begin, 8 * (size_t)n);
    int *isrc[3] = {p->elem, p->aux, p->left};
    for (int a = 0; a < 3 && !rc; a++) rc = copy_to_host(c, iv.data() + a * n, isrc[a] + begin, 4 * (size_t)n);
    if (rc) return rc;
    FVC_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            recs[i].X[a] = d[a * n + i];
            recs[i].V[a] = d[(3 + a) * n + i];
        }
        recs[i].LeftDomainTime = d[6 * n + i];
        recs[i].Scalar = d[7 * n + i];
        recs[i].ElementIndex = iv[i];
        recs[i].AuxElementIndex = iv[n + i];
        recs[i].LeftDomain = iv[2 * n + i];
        recs[i].pad_ = 0;
    }
    return FVC_OK;
}

static int particles_d2d(fvc_particles *dst, int64_t db, const fvc_particles *src, int64_t sb, int64_t n) {
    fvc_ctx *c = dst->ctx;
    dst->act_begin = -1; /* copied-in state may bring particles back */
    double *ds[8] = {dst->X[0], dst->X[1], dst->X[2], dst->V[0], dst->V[1], dst->V[2], dst->left_time, dst->scalar};
    double *ss[8] = {src->X[0], src->X[1], src->X[2], src->V[0], src->V[1], src->V[2], src->left_time, src->scalar};
    for (int a = 0; a < 8; a++)
        FVC_HIP(hipMemcpyAsync(ds[a] + db, ss[a] + sb, 8 * n, hipMemcpyDeviceToDevice, c->stream));
    int *di[4] = {dst->elem, dst->aux, dst->left, dst->nhist};
    int *si[4] = {src->elem, src->aux, src->left, src->nhist};
    for (int a = 0; a < 4; a++)
        FVC_HIP(hipMemcpyAsync(di[a] + db, si[a] + sb, 4 * n, hipMemcpyDeviceToDevice, c->stream));
    for (int a = 0; a < 6; a++)
        FVC_HIP(hipMemcpyAsync(dst->H[a] + db, src->H[a] + sb, 8 * n, hipMemcpyDeviceToDevice, c->stream));
    FVC_HIP(hipStreamSynchronize(c->stream));
    return FVC_OK;
}

extern "C" int fvc_particles_copy(fvc_particles *dst, int64_t db, const fvc_particles *src, int64_t sb, int64_t n) {
    if (!dst || !src || !dst->ctx || dst->ctx != src->ctx || n < 0 || db < 0 || sb < 0 || db + n > dst->cap || sb + n > src->cap)
        return FVC_E_INVALID;
    if (n == 0) return FVC_OK;
    return particles_d2d(dst, db, src, sb, n);
}

extern "C" int fvc_particles_get_xyz(fvc_particles *p, int64_t begin, double *xyz, int32_t *left, int64_t n) {
    if (!p || !p->ctx || !xyz || begin < 0 || n < 0 || begin + n > p->cap) return FVC_E_INVALID;
    fvc_ctx *c = p->ctx;
    std::vector<double> d(3 * (size_t)n);
    for (int a = 0; a < 3; a++)
        FVC_HIP(hipMemcpyAsync(d.data() + a * n, p->X[a] + begin, 8 * n, hipMemcpyDeviceToHost, c->stream));
    if (left) FVC_HIP(hipMemcpyAsync(left, p->left + begin, 4 * n, hipMemcpyDeviceToHost, c->stream));
    FVC_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) xyz[3 * i + a] = d[a * n + i];
    return FVC_OK;
}

/* ------------------------------------------------------------------------ */
/* integration                                                               */
/* ------------------------------------------------------------------------ */
extern "C" int fvc_launch_velocity(fvc_ctx *c, fvc_particles *p, int64_t begin, int64_t count, double t, int32_t lo,
                                   int32_t hi, double tmin, double tmax) {
    if (!c || !p || p->ctx != c || begin < 0 || count < 0 || begin + count > p->cap) return FVC_E_INVALID;
    int rc = check_slots(c, lo, hi, tmin, tmax);
    if (rc) return rc;
    return launch_velocity(c, ptrs(p), begin, count, t, bind_field(c, lo, hi, tmin, tmax));
}

/* Runge-Kutta-Fehlberg parameters: positive (parameters.c:462-463) */
static int check_rkf(fvc_ctx *c, const fvc_int_params *ip) {
    if (ip->int_type != FVC_INT_RKF) return FVC_OK;
    if (!(ip->rkf_min_step >= 1.0e-10) || !(ip->rkf_max_step >= 1.0e-10) || !(ip->rkf_accuracy >= 1.0e-10))
        return set_err(c, FVC_E_INVALID, "RKF integration parameters must be positive");
    return FVC_OK;
}

static void bind_rkf(DevField &f, const fvc_int_params *ip) {
    f.rkf_min = ip->rkf_min_step;
    f.rkf_max = ip->rkf_max_step;
    f.rkf_tol = ip->rkf_accuracy;
}

extern "C" int fvc_advect(fvc_ctx *c, fvc_particles *p, int64_t begin, int64_t count, double t1, double t2, int32_t lo,
                          int32_t hi, double tmin, double tmax, const fvc_int_params *ip) {
    if (!c || !p || p->ctx != c || !ip || begin < 0 || count < 0 || begin + count > p->cap) return FVC_E_INVALID;
    if (ip->int_type != FVC_INT_RK4 && ip->int_type != FVC_INT_EULER && ip->int_type != FVC_INT_AB2 &&
        ip->int_type != FVC_INT_AB3 && ip->int_type != FVC_INT_RKF)
        return set_err(c, FVC_E_INVALID, "integrator not available in fvc_advect");
    int rc = check_rkf(c, ip);
    if (rc) return rc;
    if (ip->int_type != FVC_INT_RKF && !(ip->time_step > 1.0e-10))
        return set_err(c, FVC_E_INVALID, "Int_TimeStep must be positive");
    if ((rc = check_slots(c, lo, hi, tmin, tmax))) return rc;
    DevField f = bind_field(c, lo, hi, tmin, tmax);
    bind_rkf(f, ip);
    PtrSet ps = ptrs(p);
    if ((rc = bind_strain(c, lo, hi, ip, f, ps, p))) return rc;
    /* Particles that left the domain stay in their slots and their lanes
     * idle.  Once more than an eighth of a range's lanes idle, the range's
     * particles still to integrate are listed and later calls on the same
     * range launch over the list only (every particle is integrated exactly
     * as without the list). */
    const bool listed = p->act_begin == begin && p->act_count == count && p->act_extrap == ip->extrapolate;
    const int64_t launched = listed ? p->act_n : count;
    if ((rc = launch_advect(c, ps, listed ? p->act : nullptr, begin, launched, t1, t2, f, *ip))) return rc;
    constexpr int64_t LIST_MIN = 4096; /* smaller ranges launch whole */
    if (count >= LIST_MIN && c->last_idle * 8 > launched) {
        if (!p->act && dev_alloc(c, &p->act, (size_t)p->cap)) return set_err(c, FVC_E_NOMEM, "active list alloc");
        const int64_t steps = c->last_steps;
        int64_t n = 0;
        if ((rc = launch_active_list(c, ps, begin, count, ip->extrapolate, p->act, &n))) return rc;
        p->act_begin = begin;
        p->act_count = count;
        p->act_n = n;
        p->act_extrap = ip->extrapolate;
        c->last_steps = steps;
    }
    return FVC_OK;
}

extern "C" int fvc_advect_times(fvc_ctx *c, fvc_particles *p, int64_t begin, int64_t count, const double *t12,
                                int32_t lo, int32_t hi, double tmin, double tmax, const fvc_int_params *ip) {
    if (!c || !p || p->ctx != c || !ip || begin < 0 || count < 0 || begin + count > p->cap || (count > 0 && !t12))
        return FVC_E_INVALID;
    if (ip->int_type != FVC_INT_RK4 && ip->int_type != FVC_INT_EULER && ip->int_type != FVC_INT_RKF)
        return set_err(c, FVC_E_INVALID, "integrator not available in fvc_advect_times");
    int rc0 = check_rkf(c, ip);
    if (rc0) return rc0;
    if (ip->int_type != FVC_INT_RKF && !(ip->time_step > 1.0e-10))
        return set_err(c, FVC_E_INVALID, "Int_TimeStep must be positive");
    int rc = check_slots(c, lo, hi, tmin, tmax);
    if (rc) return rc;
    for (int64_t q = 0; q < count; q++) { /* each segment inside the loaded interval, t1 <= t2 */
        const double a = t12[2 * q], b = t12[2 * q + 1];
        if (a != a) continue;
        if (!(b >= a) || a < tmin - TINY || b > tmax + TINY)
            return set_err(c, FVC_E_INVALID, "fvc_advect_times: segment outside the loaded interval");
    }
    if (count == 0) {
        c->last_steps = 0;
        return FVC_OK;
    }
    DevField f = bind_field(c, lo, hi, tmin, tmax);
    bind_rkf(f, ip);
    PtrSet ps = ptrs(p);
    if ((rc = bind_strain(c, lo, hi, ip, f, ps, p))) return rc;
    double *d = nullptr;
    FVC_HIP(hipMallocAsync((void **)&d, sizeof(double) * 2 * count, c->stream));
    FVC_HIP(hipMemcpyAsync(d, t12, sizeof(double) * 2 * count, hipMemcpyHostToDevice, c->stream));
    rc = launch_advect_times(c, ps, begin, count, d, f, *ip);
    (void)hipFreeAsync(d, c->stream);
    return rc;
}

extern "C" int fvc_last_step_count(fvc_ctx *c, int64_t *n) {
    if (!c || !n) return FVC_E_INVALID;
    *n = c->last_steps;
    return FVC_OK;
}

/* ------------------------------------------------------------------------ */
/* location                                                                   */
/* ------------------------------------------------------------------------ */
extern "C" int fvc_locate(fvc_ctx *c, const double *xyz, const int32_t *guess, int64_t n, int32_t mode, int32_t *elem) {
    if (!c || !xyz || !elem || n < 0 || mode < 0 || mode > 2 || (mode != FVC_LOCATE_GLOBAL && !guess))
        return FVC_E_INVALID;
    if (c->mesh.mesh_type != 1) return set_err(c, FVC_E_INVALID, "location needs an unstructured mesh");
    if (n == 0) return FVC_OK;
    double *dx = nullptr;
    int *dg = nullptr, *dout = nullptr;
    FVC_HIP(hipMallocAsync((void **)&dx, sizeof(double) * 3 * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&dg, sizeof(int) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&dout, sizeof(int) * n, c->stream));
    FVC_HIP(hipMemcpyAsync(dx, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    if (guess) FVC_HIP(hipMemcpyAsync(dg, guess, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    int rc = launch_locate(c, dx, guess ? dg : nullptr, n, mode, dout);
    if (!rc) {
        FVC_HIP(hipMemcpyAsync(elem, dout, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipStreamSynchronize(c->stream));
    }
    hipFreeAsync(dx, c->stream);
    hipFreeAsync(dg, c->stream);
    hipFreeAsync(dout, c->stream);
    hipStreamSynchronize(c->stream);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* point location without allocation                                         */
/* ------------------------------------------------------------------------ */
/* Staging for fvc_locate_noalloc: FVC_NOALLOC_MAX_POINTS query points (24 B),
 * guesses (4 B) and results (4 B), each region 256-byte aligned.  Every GPU
 * the library runs on holds its own copy until the process ends; calls on
 * one GPU share it. */
static __device__ __align__(256) unsigned char scratch[(size_t)FVC_NOALLOC_MAX_POINTS * 32 + 1024];

/* scratch's device address on each GPU (its host-side name is only a
 * stand-in: kernels and copies take these addresses) */
static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr;

static int noalloc_scratch(int device, void **ptr) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0) return FVC_E_INVALID;
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) return FVC_E_HIP;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)) != hipSuccess || !p) return FVC_E_HIP;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return FVC_OK;
}

extern "C" int fvc_locate_noalloc(fvc_ctx *c, const double *xyz, const int32_t *guess, int64_t n, int32_t mode,
                                  int32_t *elem) {
    if (!c || !xyz || !elem || n < 0 || mode < 0 || mode > 2 || (mode != FVC_LOCATE_GLOBAL && !guess))
        return FVC_E_INVALID;
    if (c->mesh.mesh_type != 1) return set_err(c, FVC_E_INVALID, "location needs an unstructured mesh");
    if (n > FVC_NOALLOC_MAX_POINTS)
        return set_err(c, FVC_E_INVALID, "fvc_locate_noalloc: more points than FVC_NOALLOC_MAX_POINTS");
    if (n == 0) return FVC_OK;
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, FVC_E_HIP, "fvc_locate_noalloc: hipSetDevice");
    void *base = nullptr;
    if (noalloc_scratch(c->device, &base)) return set_err(c, FVC_E_HIP, "fvc_locate_noalloc: scratch address");
    char *b = (char *)base;
    const size_t xbytes = ((size_t)n * 24 + 255) / 256 * 256, ibytes = ((size_t)n * 4 + 255) / 256 * 256;
    double *dx = (double *)b;
    int *dg = (int *)(b + xbytes);
    int *dout = (int *)(b + xbytes + ibytes);
    FVC_HIP(hipMemcpyAsync(dx, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    if (guess) FVC_HIP(hipMemcpyAsync(dg, guess, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    int rc = launch_locate(c, dx, guess ? dg : nullptr, n, mode, dout); /* every result is written */
    if (rc) return rc;
    FVC_HIP(hipMemcpyAsync(elem, dout, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    FVC_HIP(hipStreamSynchronize(c->stream));
    return FVC_OK;
}

static int seed_common(fvc_ctx *c, int chain, const double *xyz, int64_t n, const double mn[3], const double del[3],
                       const int32_t res[3], int lsc, fvc_point *out) {
#pragma clang fp contract(off)
    double *dx = nullptr, *dlt = nullptr;
    int *de = nullptr, *dl = nullptr;
    if (chain == 0) {
        FVC_HIP(hipMallocAsync((void **)&dx, sizeof(double) * 3 * n, c->stream));
        FVC_HIP(hipMemcpyAsync(dx, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    }
    FVC_HIP(hipMallocAsync((void **)&dlt, sizeof(double) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&de, sizeof(int) * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&dl, sizeof(int) * n, c->stream));
    int rc = launch_seed_chain(c, chain, n, dx, mn, del, res, lsc, de, dl, dlt);
    if (!rc) {
        std::vector<int> e(n), l(n);
        std::vector<double> lt(n);
        FVC_HIP(hipMemcpyAsync(e.data(), de, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipMemcpyAsync(l.data(), dl, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipMemcpyAsync(lt.data(), dlt, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < n; i++) {
            fvc_point &p = out[i];
            memset(&p, 0, sizeof p);
            if (chain == 0) {
                p.X[0] = xyz[3 * i];
                p.X[1] = xyz[3 * i + 1];
                p.X[2] = c->mesh.dims == 2 ? 0.0 : xyz[3 * i + 2];
            } else {
                int ii = (int)(i % res[0]), jj = (int)((i / res[0]) % res[1]), kk = (int)(i / ((int64_t)res[0] * res[1]));
                p.X[0] = mn[0] + ii * del[0];
                p.X[1] = mn[1] + jj * del[1];
                p.X[2] = mn[2] + kk * del[2];
            }
            p.ElementIndex = c->mesh.mesh_type == 1 ? e[i] : 0;
            p.AuxElementIndex = chain == 0 && !l[i] ? e[i] : 0;
            p.LeftDomain = l[i];
            p.LeftDomainTime = lt[i];
        }
    }
    if (dx) hipFreeAsync(dx, c->stream);
    hipFreeAsync(dlt, c->stream);
    hipFreeAsync(de, c->stream);
    hipFreeAsync(dl, c->stream);
    hipStreamSynchronize(c->stream);
    return rc;
}

extern "C" int fvc_seed_list(fvc_ctx *c, const double *xyz, int64_t n, int32_t lsc, fvc_point *out) {
    if (!c || !xyz || !out || n < 0) return FVC_E_INVALID;
    if (n == 0) return FVC_OK;
    std::vector<double> x(3 * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        x[3 * i] = xyz[3 * i];
        x[3 * i + 1] = xyz[3 * i + 1];
        x[3 * i + 2] = c->mesh.dims == 2 ? 0.0 : xyz[3 * i + 2]; /* tracers.c:1033 */
    }
    return seed_common(c, 0, x.data(), n, nullptr, nullptr, nullptr, lsc, out);
}

extern "C" int fvc_seed_grid(fvc_ctx *c, const double mn[3], const double del[3], const int32_t res[3], int32_t lsc,
                             fvc_point *out) {
    if (!c || !mn || !del || !res || !out || res[0] < 1 || res[1] < 1 || res[2] < 1) return FVC_E_INVALID;
    int64_t n = (int64_t)res[0] * res[1] * res[2];
    return seed_common(c, 1, nullptr, n, mn, del, res, lsc, out);
}

extern "C" int fvc_seed_velout_grid(fvc_ctx *c, const double mn[3], const double del[3], const int32_t res[3],
                                    int32_t lsc, fvc_point *out) {
    if (!c || !mn || !del || !res || !out || res[0] < 1 || res[1] < 1 || res[2] < 1) return FVC_E_INVALID;
    int64_t n = (int64_t)res[0] * res[1] * res[2];
    return seed_common(c, 3, nullptr, n, mn, del, res, lsc, out);
}

/* ------------------------------------------------------------------------ */
/* FTLE slides                                                               */
/* ------------------------------------------------------------------------ */
__global__ void k_ftle_grid(double *X0, double *X1, double *X2, int *elem, int *left, double *lt, int64_t n, double mn0,
                            double mn1, double mn2, double d0, double d1, double d2, int r1, int r2,
                            const int *celem, const int *cleft, const double *clt) {
#pragma clang fp contract(off)
    int64_t P = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= n) return;
    int k = (int)(P % r2), j = (int)((P / r2) % r1), i = (int)(P / ((int64_t)r2 * r1));
    X0[P] = mn0 + i * d0;
    X1[P] = mn1 + j * d1;
    X2[P] = mn2 + k * d2;
    elem[P] = celem[P];
    left[P] = cleft[P];
    lt[P] = clt[P];
}

/* Storage of one slide of res[0]*res[1]*res[2] grid points: particle
 * records, per-point FTLE values and flags, one frame's staging and the
 * grid's Z-order. */
static int ftle_alloc(fvc_ctx *c, const double mn[3], const double del[3], const int32_t res[3], double start_time,
                      double stop_time, fvc_ftle **out) {
    fvc_ftle *f = new fvc_ftle();
    f->ctx = c;
    c->live_ftle.insert(f);
    for (int a = 0; a < 3; a++) {
        f->mn[a] = mn[a];
        f->del[a] = del[a];
        f->res[a] = res[a];
        f->mx[a] = mn[a] + (res[a] - 1) * del[a];
    }
    f->n = (int64_t)f->res[0] * f->res[1] * f->res[2];
    f->start_time = start_time;
    f->stop_time = stop_time;
    int rc = fvc_particles_create(c, f->n, &f->pts);
    if (rc) {
        fvc_ftle_destroy(f);
        return rc;
    }
    size_t n = (size_t)f->n;
    size_t bytes = n * (3 * 4 + 3 * 8 + 3 * 8 + 3 * 8 + 4 + 4 + 4 + 8 + 4) + 1024;
    if (hipMalloc(&f->buf, bytes) != hipSuccess) {
        fvc_ftle_destroy(f);
        return set_err(c, FVC_E_NOMEM, "FTLE alloc");
    }
    c->bytes_used += (int64_t)bytes;
    double *d = (double *)f->buf;
    f->wt = d;
    f->wot = d + n;
    for (int a = 0; a < 3; a++) {
        f->Xn[a] = d + (2 + a) * n;
        f->Vn[a] = d + (5 + a) * n;
    }
    f->exit_time = d + 8 * n;
    int *iv = (int *)(d + 9 * n);
    f->have = iv;
    f->elem_new = iv + n;
    f->exited = iv + 2 * n;
    f->advected = iv + 3 * n;
    f->active = iv + 4 * n;
    f->zperm = iv + 5 * n;
    if ((rc = launch_grid_zorder(c, f->res, f->n, f->zperm))) {
        fvc_ftle_destroy(f);
        return rc;
    }
    *out = f;
    return FVC_OK;
}

/* Snapshot of the initial state (for fvc_ftle_reset) and the count of
 * points inside. */
static int ftle_finish(fvc_ftle *f) {
    fvc_ctx *c = f->ctx;
    fvc_particles *p = f->pts;
    size_t n = (size_t)f->n;
    int rc;
    if ((rc = fvc_particles_create(c, f->n, &f->init_pts)) || (rc = particles_d2d(f->init_pts, 0, p, 0, f->n)))
        return rc;
    if (hipMalloc((void **)&f->init_have, sizeof(int) * n) != hipSuccess ||
        hipMalloc((void **)&f->init_w, sizeof(double) * 2 * n) != hipSuccess)
        return set_err(c, FVC_E_NOMEM, "FTLE snapshot alloc");
    c->bytes_used += (int64_t)(n * (4 + 16));
    FVC_HIP(hipMemcpyAsync(f->init_have, f->have, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
    FVC_HIP(hipMemcpyAsync(f->init_w, f->wt, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, c->stream));
    /* points inside = not left */
    FVC_HIP(hipStreamSynchronize(c->stream));
    std::vector<int> lh(n);
    FVC_HIP(hipMemcpy(lh.data(), p->left, sizeof(int) * n, hipMemcpyDeviceToHost));
    int64_t inside = 0;
    for (size_t q = 0; q < n; q++) inside += !lh[q];
    f->inside = inside;
    return FVC_OK;
}

extern "C" int fvc_ftle_create(fvc_ctx *c, const double mn[3], const double mx[3], const int32_t res[3], int32_t lsc,
                               double start_time, double stop_time, fvc_ftle **out) {
    if (!c || !mn || !mx || !res || !out) return FVC_E_INVALID;
    *out = nullptr;
    const int dims = c->mesh.dims;
    if (res[0] < 2 || res[1] < 2 || (dims == 3 && res[2] < 2)) return FVC_E_INVALID;
    double g_mn[3] = {mn[0], mn[1], mn[2]}, del[3];
    int32_t g_res[3] = {res[0], res[1], res[2]};
    /* parameters.c:310-320 */
    del[0] = (mx[0] - mn[0]) / (res[0] - 1);
    del[1] = (mx[1] - mn[1]) / (res[1] - 1);
    if (dims == 3)
        del[2] = (mx[2] - mn[2]) / (res[2] - 1);
    else {
        g_mn[2] = 0.0;
        del[2] = 0.0;
        g_res[2] = 1;
    }
    fvc_ftle *f = nullptr;
    int rc = ftle_alloc(c, g_mn, del, g_res, start_time, stop_time, &f);
    if (rc) return rc;
    /* location chain of InitializeFTLEArray, then the grid records */
    int *ce = f->elem_new, *cl = f->exited; /* reuse staging as chain output */
    double *clt = f->exit_time;
    rc = launch_seed_chain(c, 2, f->n, nullptr, f->mn, f->del, f->res, lsc, ce, cl, clt);
    if (rc) {
        fvc_ftle_destroy(f);
        return rc;
    }
    fvc_particles *p = f->pts;
    k_ftle_grid<<<(unsigned)((f->n + 255) / 256), 256, 0, c->stream>>>(
        p->X[0], p->X[1], p->X[2], p->elem, p->left, p->left_time, f->n, f->mn[0], f->mn[1], f->mn[2], f->del[0],
        f->del[1], f->del[2], f->res[1], f->res[2], ce, cl, clt);
    if ((rc = launch_ftle_mask(f)) || (rc = ftle_finish(f))) {
        fvc_ftle_destroy(f);
        return rc;
    }
    *out = f;
    return check_err(c, "fvc_ftle_create");
}

extern "C" int fvc_ftle_create_from(fvc_ctx *c, const double mn[3], const double del[3], const int32_t res[3],
                                    const fvc_point *recs, const int32_t *have, const double *ftle_wt,
                                    const double *ftle_wot, double start_time, double stop_time, fvc_ftle **out) {
    if (!c || !mn || !del || !res || !recs || !have || !ftle_wt || !ftle_wot || !out) return FVC_E_INVALID;
    *out = nullptr;
    const int dims = c->mesh.dims;
    if (res[0] < 1 || res[1] < 1 || res[2] < 1 || (dims == 2 && res[2] != 1)) return FVC_E_INVALID;
    const int64_t n = (int64_t)res[0] * res[1] * res[2];
    if (n > 0x7fffffffLL) return FVC_E_INVALID;
    fvc_ftle *f = nullptr;
    int rc = ftle_alloc(c, mn, del, res, start_time, stop_time, &f);
    if (rc) return rc;
    std::vector<int> hv((size_t)n);
    for (int64_t q = 0; q < n; q++) hv[q] = have[q] != 0;
    if ((rc = fvc_particles_set(f->pts, 0, recs, n))) {
        fvc_ftle_destroy(f);
        return rc;
    }
    if (hipMemcpy(f->have, hv.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->wt, ftle_wt, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->wot, ftle_wot, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
        fvc_ftle_destroy(f);
        return set_err(c, FVC_E_HIP, "FTLE state upload failed");
    }
    if ((rc = ftle_finish(f))) {
        fvc_ftle_destroy(f);
        return rc;
    }
    *out = f;
    return check_err(c, "fvc_ftle_create_from");
}

void fvc::ftle_release(fvc_ftle *f) {
    fvc_ctx *c = f->ctx;
    if (!c) return;
    side_sync(c);
    hipStreamSynchronize(c->stream);
    for (void *&sb : f->snap)
        if (sb) {
            hipFree(sb);
            sb = nullptr;
        }
    if (f->buf) hipFree(f->buf);
    if (f->init_have) hipFree(f->init_have);
    if (f->init_w) hipFree(f->init_w);
    f->buf = nullptr;
    f->init_have = nullptr;
    f->init_w = nullptr;
    fvc_particles_destroy(f->init_pts);
    fvc_particles_destroy(f->pts);
    f->init_pts = f->pts = nullptr;
    c->live_ftle.erase(f);
    f->ctx = nullptr;
}

extern "C" void fvc_ftle_destroy(fvc_ftle *f) {
    if (!f) return;
    ftle_release(f);
    delete f;
}

extern "C" int fvc_ftle_reset(fvc_ftle *f) {
    if (!f || !f->ctx || !f->init_pts) return FVC_E_INVALID;
    fvc_ctx *c = f->ctx;
    size_t n = (size_t)f->n;
    int rc = side_join(c);
    if (rc) return rc;
    rc = particles_d2d(f->pts, 0, f->init_pts, 0, f->n);
    if (rc) return rc;
    FVC_HIP(hipMemcpyAsync(f->have, f->init_have, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
    FVC_HIP(hipMemcpyAsync(f->wt, f->init_w, sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, c->stream));
    FVC_HIP(hipMemsetAsync(f->exited, 0, sizeof(int) * n, c->stream));
    FVC_HIP(hipMemsetAsync(f->advected, 0, sizeof(int) * n, c->stream));
    FVC_HIP(hipStreamSynchronize(c->stream));
    f->status = 0;
    return FVC_OK;
}

extern "C" int fvc_ftle_frame(fvc_ftle *f, int32_t dir, int32_t lo, int32_t hi, double tmin, double tmax,
                              const fvc_int_params *ip, int32_t *done) {
    if (!f || !f->ctx || !ip || (dir != 1 && dir != -1)) return FVC_E_INVALID;
    fvc_ctx *c = f->ctx;
    if (done) *done = 0;
    if (f->status == 2) return FVC_OK;
    const double T = 1.0e-10;
    if (!(f->status == 1 || (f->start_time <= tmax + T && f->start_time >= tmin - T))) return FVC_OK;
    int rc = check_slots(c, lo, hi, tmin, tmax);
    if (rc) return rc;
    if (ip->int_type != FVC_INT_RK4 && ip->int_type != FVC_INT_EULER && ip->int_type != FVC_INT_RKF)
        return FVC_E_INVALID;
    if ((rc = check_rkf(c, ip))) return rc;
    double t1;
    if (f->status == 0) {
        f->status = 1;
        t1 = f->start_time;
        rc = launch_velocity(c, ptrs(f->pts), 0, f->n, f->start_time, bind_field(c, lo, hi, tmin, tmax));
        if (rc) return rc;
    } else {
        t1 = dir > 0 ? tmin : tmax;
    }
    double t2 = dir > 0 ? (tmax < f->stop_time ? tmax : f->stop_time) : (tmin > f->stop_time ? tmin : f->stop_time);
    rc = launch_ftle_frame(f, lo, hi, tmin, tmax, t1, t2, *ip);
    if (rc) return rc;
    if (((dir > 0) ? (f->stop_time - tmax) : (tmin - f->stop_time)) < T) f->status = 2;
    if (done) *done = f->status == 2 ? 2 : 1;
    return FVC_OK;
}

/* i-major storage (P = (i*r1 + j)*r2 + k) -> k, j, i file order (q = (k*r1 + j)*r0 + i) */
__global__ void k_ftle_kji(const double *__restrict__ wt, const double *__restrict__ wot, int r0, int r1, int r2,
                           double *__restrict__ out) {
    const int64_t n = (int64_t)r0 * r1 * r2;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = (int)(q % r0), j = (int)((q / r0) % r1), k = (int)(q / ((int64_t)r0 * r1));
    const int64_t P = ((int64_t)i * r1 + j) * r2 + k;
    out[q] = wt[P];
    out[n + q] = wot[P];
}

/* FTLE values of a grid to host memory in the file order of ftle.c:1302-1315 */
int fvc::copy_ftle_kji(fvc_ctx *c, const double *d_wt, const double *d_wot, const int32_t res[3], double *wt,
                       double *wot) {
    const size_t n = (size_t)res[0] * res[1] * res[2];
    double *tmp = nullptr;
    FVC_HIP(hipMallocAsync((void **)&tmp, sizeof(double) * 2 * n, c->stream));
    k_ftle_kji<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(d_wt, d_wot, res[0], res[1], res[2], tmp);
    hipError_t e = hipGetLastError();
    int rc = e == hipSuccess ? FVC_OK : FVC_E_HIP;
    if (!rc && wt) rc = copy_to_host(c, wt, tmp, sizeof(double) * n);
    if (!rc && wot) rc = copy_to_host(c, wot, tmp + n, sizeof(double) * n);
    (void)hipFreeAsync(tmp, c->stream);
    if (!rc) rc = hipStreamSynchronize(c->stream) == hipSuccess ? FVC_OK : FVC_E_HIP;
    if (rc) c->last_error = std::string("FTLE output copy: ") + hipGetErrorString(hipGetLastError());
    return rc;
}

extern "C" int fvc_ftle_output(fvc_ftle *f, double int_t, double *wt, double *wot) {
    if (!f || !f->ctx) return FVC_E_INVALID;
    int rc = side_join(f->ctx);
    if (!rc) rc = launch_ftle_fill(f, int_t);
    return rc ? rc : copy_ftle_kji(f->ctx, f->wt, f->wot, f->res, wt, wot);
}

extern "C" int fvc_ftle_points(fvc_ftle *f, int64_t *n, int64_t *inside) {
    if (!f || !f->ctx) return FVC_E_INVALID;
    if (n) *n = f->n;
    if (inside) *inside = f->inside;
    return FVC_OK;
}

extern "C" int fvc_ftle_records(fvc_ftle *f, fvc_point *recs) {
    if (!f || !f->ctx || !recs) return FVC_E_INVALID;
    return fvc_particles_get(f->pts, 0, recs, f->n);
}

extern "C" int fvc_ftle_initial(fvc_ftle *f, fvc_point *recs, int32_t *have, double *ftle_wt, double *ftle_wot) {
    if (!f || !f->ctx || !f->init_pts) return FVC_E_INVALID;
    fvc_ctx *c = f->ctx;
    const size_t n = (size_t)f->n;
    int rc;
    if (recs && (rc = fvc_particles_get(f->init_pts, 0, recs, f->n))) return rc;
    std::vector<double> w(2 * n);
    FVC_HIP(hipMemcpy(w.data(), f->init_w, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
    if (have) FVC_HIP(hipMemcpy(have, f->init_have, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (ftle_wt) memcpy(ftle_wt, w.data(), sizeof(double) * n);
    if (ftle_wot) memcpy(ftle_wot, w.data() + n, sizeof(double) * n);
    return check_err(c, "fvc_ftle_initial");
}

/* ------------------------------------------------------------------------ */
/* Z-order                                                                    */
/* ------------------------------------------------------------------------ */
extern "C" int fvc_zsort(fvc_ctx *c, const double *xyz, int64_t n, const double lo[3], const double hi[3],
                         int32_t *perm) {
    if (!c || !xyz || !lo || !hi || !perm || n < 0) return FVC_E_INVALID;
    if (n == 0) return FVC_OK;
    double *dx = nullptr;
    int *dp = nullptr;
    FVC_HIP(hipMallocAsync((void **)&dx, sizeof(double) * 3 * n, c->stream));
    FVC_HIP(hipMallocAsync((void **)&dp, sizeof(int) * n, c->stream));
    FVC_HIP(hipMemcpyAsync(dx, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    int rc = launch_zsort(c, dx, n, lo, hi, dp);
    if (!rc) {
        FVC_HIP(hipMemcpyAsync(perm, dp, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
        FVC_HIP(hipStreamSynchronize(c->stream));
    }
    hipFreeAsync(dx, c->stream);
    hipFreeAsync(dp, c->stream);
    hipStreamSynchronize(c->stream);
    return rc;
}
