// hegj_api.cpp -- the extern "C" boundary of libhegjoin (include/hegjoin.h).
//
// Host orchestration of the device pipeline that replaces the reference's
// gridIndexingGPU / sortByWorkLoad / distanceTableNDGridBatches.  Host code
// here only sizes buffers, launches kernels on the handle's stream and reads
// back scalars; every per-point and per-pair step runs on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hegj_internal.h"

namespace hegj {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

hegj_status fail(hegj_status s, const std::string &msg) {
    g_last_error = msg;
    return s;
}

static thread_local Arena *g_arena = nullptr;

void set_thread_arena(Arena *a) { g_arena = a; }

static hipError_t arena_alloc(void **p, size_t bytes) {
    const size_t off = (g_arena->used + 255) & ~(size_t)255;
    if (off > g_arena->cap || bytes > g_arena->cap - off) return hipErrorOutOfMemory;
    *p = g_arena->base + off;
    g_arena->used = off + bytes;
    return hipSuccess;
}

static bool in_arena(const void *p) {
    return g_arena && (const char *)p >= g_arena->base && (const char *)p < g_arena->base + g_arena->cap;
}

hipError_t Scratch::alloc(void **p, size_t bytes) {
    if (g_arena) return arena_alloc(p, bytes);
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
}

Scratch::~Scratch() {
    for (void *p : ptrs) (void)hipFree(p);
}

namespace {
std::mutex g_cache_mu;
std::map<void *, std::pair<int, size_t>> g_cache_live;           // ptr -> (device, size)
std::multimap<std::pair<int, size_t>, void *> g_cache_free;      // (device, size) -> ptr
size_t g_cache_bytes = 0;
constexpr size_t kCacheCap = 16ull << 30;

size_t cache_round(size_t b) {
    if (b <= ((size_t)1 << 20)) {
        size_t s = 256;
        while (s < b) s <<= 1;
        return s;
    }
    const size_t g = (size_t)2 << 20;
    return (b + g - 1) / g * g;
}

void cache_release(int dev) {
    std::vector<void *> rel;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        for (auto it = g_cache_free.begin(); it != g_cache_free.end();) {
            if (it->first.first == dev) {
                rel.push_back(it->second);
                g_cache_bytes -= it->first.second;
                it = g_cache_free.erase(it);
            } else {
                ++it;
            }
        }
    }
    for (void *p : rel) (void)hipFree(p);
}
}  // namespace

void cache_release_device(int dev) { cache_release(dev); }

hipError_t dev_malloc(void **p, size_t bytes) {
    if (g_arena) return arena_alloc(p, bytes);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t rs = cache_round(bytes);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_cache_free.find({dev, rs});
        if (it != g_cache_free.end()) {
            *p = it->second;
            g_cache_free.erase(it);
            g_cache_bytes -= rs;
            g_cache_live[*p] = {dev, rs};
            return hipSuccess;
        }
    }
    e = hipMalloc(p, rs);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        cache_release(dev);
        e = hipMalloc(p, rs);
    }
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_cache_live[*p] = {dev, rs};
    }
    return e;
}

void dev_free(void *p) {
    if (!p || in_arena(p)) return;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_cache_live.find(p);
        if (it != g_cache_live.end()) {
            const auto key = it->second;
            g_cache_live.erase(it);
            if (g_cache_bytes + key.second <= kCacheCap) {
                g_cache_free.insert({key, p});
                g_cache_bytes += key.second;
                return;
            }
        }
    }
    (void)hipFree(p);
}

// S7: largest s with RN(sqrt(s)) <= eps.  sqrt is monotone, so
// sqrt(acc) <= eps  <=>  acc <= T, and the device needs no sqrt.
double sqrt_threshold(double eps) {
    double s = eps * eps;
    const double inf = std::numeric_limits<double>::infinity();
    while (s > 0 && std::sqrt(s) > eps) s = std::nextafter(s, 0.0);
    while (std::sqrt(std::nextafter(s, inf)) <= eps) s = std::nextafter(s, inf);
    return s;
}

// DTYPE=float (kernel.cu:363 with DTYPE float: sqrtf, correctly rounded).
float sqrt_threshold_f(float eps) {
    float s = eps * eps;
    const float inf = std::numeric_limits<float>::infinity();
    while (s > 0 && std::sqrt(s) > eps) s = std::nextafter(s, 0.0f);
    while (std::sqrt(std::nextafter(s, inf)) <= eps) s = std::nextafter(s, inf);
    return s;
}

struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit EventTimer(hipStream_t st) : s(st) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, s);
    }
    float stop() {
        float ms = 0;
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~EventTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

struct GraphPlan;
void free_graph_plan(GraphPlan *g);

}  // namespace hegj

using namespace hegj;

struct hegj_index {
    int device = 0;
    hipStream_t stream = nullptr;
    GridGeom g{};
    double maxArr[kMaxK]{};
    uint64_t totalCells = 0;
    DevIndex ix;
    uint32_t n_dead = 0;
    double build_ms = 0;
    // workload sort (lazily computed)
    uint32_t *d_work = nullptr;   // [C] candidates per cell
    uint32_t *d_order = nullptr;  // [C] cells by workload
    uint32_t *d_morton = nullptr; // [C] cells in Morton order (lazily)
    bool morton_ok = false;
    uint64_t total_candidates = 0;
    double workload_ms = 0;
    bool have_workload = false;
    // split-BF16 operands of the MFMA refine (lazily computed)
    uint16_t *gb_A = nullptr, *gb_B = nullptr;
    float *gb_w = nullptr;
    uint32_t gb_kp = 0;
    double gb_ms = 0;
    // FP64 join plan + captured batch graph, reused while the options match
    hegj::GraphPlan *gplan = nullptr;
    // centred FP32 copy for the FP32 filter (lazily computed)
    float *soa32 = nullptr;
    float *aos32 = nullptr;        // point-major, dim <= kAosMaxDim
    double *f32_centre = nullptr;  // device, dim doubles
    double f32_radius = 0;
};

struct hegj_result {
    hegj_index *idx = nullptr;
    std::vector<uint32_t *> batches;
    std::vector<uint64_t> bases;  // global offset of each batch
    uint64_t *d_row_off = nullptr;
    uint32_t *d_row_cnt = nullptr;
    hegj_result_info info{};
    bool hybrid = false;        // made by hegj_selfjoin_hybrid
    hegj_hybrid_info hinfo{};
    hipEvent_t done = nullptr;  // recorded when the join handed the result out
    ~hegj_result() {
        // the buffers go back to the cache: nothing may still be writing them
        if (done) {
            (void)hipEventSynchronize(done);
            (void)hipEventDestroy(done);
        } else if (idx && idx->stream) {
            (void)hipStreamSynchronize(idx->stream);  // failed inside the join
        }
        for (uint32_t *b : batches) dev_free(b);
        dev_free(d_row_off);
        dev_free(d_row_cnt);
    }
};

namespace {

hegj_status set_device(int dev) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(HEGJ_ERR_NO_DEVICE, "no HIP device visible");
    if (dev < 0 || dev >= count) return fail(HEGJ_ERR_INVALID_ARG, "device ordinal out of range");
    HIP_TRY(hipSetDevice(dev));
    return HEGJ_OK;
}

hegj_status check_opts(const hegj_index_opts *o, uint32_t n, double eps) {
    if (!o) return fail(HEGJ_ERR_INVALID_ARG, "opts is NULL");
    if (n == 0) return fail(HEGJ_ERR_INVALID_ARG, "empty dataset");
    if (n > HEGJ_MAX_POINTS) return fail(HEGJ_ERR_UNSUPPORTED, "more than HEGJ_MAX_POINTS (2^29 - 1) points");
    if (o->dim == 0) return fail(HEGJ_ERR_INVALID_ARG, "dim must be >= 1");
    if (o->indexed_dims == 0 || o->indexed_dims > HEGJ_MAX_INDEXED_DIMS || o->indexed_dims > o->dim)
        return fail(HEGJ_ERR_INVALID_ARG, "indexed_dims must be in [1, min(dim, 6)]");
    if (o->dtype != HEGJ_DTYPE_F64 && o->dtype != HEGJ_DTYPE_F32)
        return fail(HEGJ_ERR_INVALID_ARG, "dtype must be HEGJ_DTYPE_F64 or HEGJ_DTYPE_F32");
    if (!(eps > 0.0) || eps > 1.0) return fail(HEGJ_ERR_INVALID_ARG, "epsilon must be in ]0, 1]");
    return HEGJ_OK;
}

hegj_status build_common(hegj_index *h, uint32_t n, double eps, const hegj_index_opts *o) {
    h->g.n = n;
    h->g.dim = o->dim;
    h->g.k = o->indexed_dims;
    h->g.f32dt = o->dtype == HEGJ_DTYPE_F32 ? 1u : 0u;
    if (h->g.f32dt) {  // DTYPE epsilon = atof(...) (main.cu:56) rounds to float
        const float epsf = (float)eps;
        h->g.eps = (double)epsf;
        h->g.thresh = (double)sqrt_threshold_f(epsf);
    } else {
        h->g.eps = eps;
        h->g.thresh = sqrt_threshold(eps);
    }
    h->g.emulate_cell0 = o->emulate_cell0 ? 1u : 0u;
    EventTimer t(h->stream);
    HEGJ_TRY(build_index(&h->g, h->maxArr, &h->totalCells, &h->ix, &h->n_dead, h->stream));
    h->build_ms = t.stop();
    return HEGJ_OK;
}

template <typename F>
hegj_status new_index(const hegj_index_opts *o, hegj_index **out, F fill) {
    std::unique_ptr<hegj_index> h(new hegj_index());
    h->device = o->device;
    HEGJ_TRY(set_device(o->device));
    HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    hegj_status s = fill(h.get());
    if (s != HEGJ_OK) {
        std::string msg = g_last_error;
        hegj_index_free(h.release());
        g_last_error = msg;
        return s;
    }
    *out = h.release();
    return HEGJ_OK;
}

hegj_status ensure_workload(hegj_index *h) {
    if (h->have_workload) return HEGJ_OK;
    const uint32_t C = h->ix.C;
    HIP_TRY(hipMalloc(&h->d_work, sizeof(uint32_t) * (C + 1)));
    HIP_TRY(hipMalloc(&h->d_order, sizeof(uint32_t) * (C + 1)));
    EventTimer t(h->stream);
    HEGJ_TRY(launch_workload(h->g, h->ix, h->d_work, h->stream));
    HEGJ_TRY(workload_order(h->ix, h->d_work, h->d_order, &h->total_candidates, h->stream));
    h->workload_ms = t.stop();
    h->have_workload = true;
    return HEGJ_OK;
}

// Default queries per work unit by dimensionality (tuned on MI355X).
uint32_t default_qb(uint32_t dim) {
    if (dim <= 2) return 16;
    if (dim <= 8) return 4;
    return 8;
}

struct Counters {
    uint32_t queue_heads[kMaxParts];
    uint32_t n_deferred;
    uint32_t pad;
    unsigned long long cursor;
    unsigned long long hash;
    unsigned long long evals;
    unsigned long long pairs;
    unsigned long long band;  // FP32 filter: pairs rechecked in FP64
    unsigned long long mismatch;  // units whose write pass disagreed with their count
};

hegj_status ensure_morton(hegj_index *h) {
    if (h->d_morton) return HEGJ_OK;
    HIP_TRY(hipMalloc(&h->d_morton, sizeof(uint32_t) * (h->ix.C + 1)));
    return morton_order(h->g, h->ix, h->d_morton, &h->morton_ok, h->stream);
}

// Bound on |acc32 - acc| for a pair whose exact squared distance is S: acc32
// is the FP32 fma chain over d^ = fl32(x^_c - x^_q), x^ = fl32(fl64(x - c))
// with |x - c| <= R, acc the reference's FP64 fma chain over fl64(x_c - x_q).
//   |x^ - (x - c)| <= a R,  a = u + v + uv          (u = 2^-24, v = 2^-53)
//   |d^_l - d_l| <= A + u |d_l|,  A = 2 a R (1 + u)
//   |sum d^^2 - S| <= 2.0001 A sqrt(D S) + D A^2 + 2.0001 u S   (Cauchy-Schwarz)
//   |acc32 - sum d^^2| <= 1.0001 D u sum d^^2        (D roundings of nonnegative sums)
//   |acc - S| <= 1.0001 (D + 2) v S
// plus D * 2^-126 for FP32 values flushed to zero.
double f32_bound(double S, double D, double R) {
    const double u = std::ldexp(1.0, -24), v = std::ldexp(1.0, -53);
    const double a = u + v + u * v;
    const double A = 2.0 * a * R * (1.0 + u);
    const double sumdiff = 2.0001 * A * std::sqrt(D * S) + D * A * A + 2.0001 * u * S;
    return (1.0 + 1.0001 * D * u) * sumdiff + 1.0001 * D * u * S + 1.0001 * (D + 2.0) * v * S +
           D * std::ldexp(1.0, -126);
}

// Thresholds of the FP32 filter: acc32 <= tacc => acc <= T, acc32 > trej =>
// acc > T (twice the bound at S = 1.001 T, which covers every S whose acc or
// acc32 can reach the thresholds).  False when the band would be wider than
// 1% of T (then the filter would recheck too much to pay off).
}  // namespace

bool hegj::f32_thresholds(double T, uint32_t dim, double R, float *tacc, float *trej) {
    if (!(T > 1e-30) || !(R < 1e30)) return false;
    const double B = 2.0 * f32_bound(1.001 * T, (double)dim, R);
    if (!(B <= 0.01 * T)) return false;
    const double lo = T - B, hi = T + B;
    float fl = (float)lo, fh = (float)hi;
    if ((double)fl > lo) fl = std::nextafter(fl, -std::numeric_limits<float>::infinity());
    if ((double)fh < hi) fh = std::nextafter(fh, std::numeric_limits<float>::infinity());
    *tacc = fl;
    *trej = fh;
    return true;
}

namespace {

hegj_status ensure_soa32(hegj_index *h) {
    if (h->soa32) return HEGJ_OK;
    return make_soa32(h->g, h->ix, &h->soa32, &h->aos32, &h->f32_radius, &h->f32_centre, h->stream);
}

// Everything a FP64 join needs besides its result: work units in the chosen
// order, their split into per-XCD ranges, the kernel's scratch and the batch
// capacity (estimated on the device unless the caller fixed it).
struct JoinPlan {
    uint2 *units = nullptr;
    uint2 *units_next = nullptr, *units_tmp = nullptr;  // deferred units (ping-pong)
    uint32_t U = 0;
    uint32_t *deferred = nullptr;
    Counters *d_ctr = nullptr;
    Counters *d_ctr2 = nullptr;  // count-only passes over deferred units
    uint64_t *gmask = nullptr;
    uint32_t gmask_slots = 8192;
    uint32_t max_blocks = 0;
    uint32_t waves = 4;  // waves per workgroup
    uint32_t parts = 1;
    uint32_t part_end[kMaxParts] = {};
    uint64_t capacity = 0;
    double estimate_ms = 0;
    uint64_t estimated_pairs = 0;  // the estimator's extrapolated pair count
    bool f32 = false;  // FP32 filter with FP64 band recheck
    float tacc32 = 0, trej32 = 0;
    hipStream_t stream = nullptr;  // the stream its kernels run on
    JoinPlan() = default;
    JoinPlan(const JoinPlan &) = delete;
    JoinPlan &operator=(const JoinPlan &) = delete;
    ~JoinPlan() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : {(void *)units, (void *)units_next, (void *)units_tmp, (void *)deferred,
                        (void *)d_ctr, (void *)d_ctr2, (void *)gmask})
            dev_free(p);
    }
};

// About 32 grabs per wave keeps the queue atomics off the critical path.
uint32_t grab_for(const JoinPlan &P, uint32_t nu) {
    const uint64_t waves = (uint64_t)P.max_blocks * P.waves;
    const uint64_t g = nu / (waves * 32);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, 64));
}

bool valid_waves_batches(const hegj_join_opts &o) {
    return (o.block_waves == 0 || o.block_waves == 1 || o.block_waves == 2 || o.block_waves == 4) &&
           o.batches <= 32;  // leaves room under the 64-batch limit for deferrals
}

uint32_t blocks_for(const JoinPlan &P, uint32_t nu, uint32_t grab) {
    const uint64_t need = ((uint64_t)nu + grab * P.waves - 1) / (grab * P.waves);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(need, P.max_blocks));
}

// Workgroups of `block_waves` waves: as many waves resident as with 4-wave
// workgroups (occupancy is set by registers; workgroups that LDS keeps off a
// CU wait and find the queue drained), so the block count scales by 4 / waves.
void set_waves(JoinPlan &P, const hegj_join_opts &opts) {
    P.waves = opts.block_waves ? opts.block_waves : 4u;
    P.max_blocks = P.max_blocks * 4 / P.waves;
}

void set_parts(JoinArgs &x, uint32_t parts, const uint32_t *ends) {
    x.n_parts = parts;
    for (uint32_t p = 0; p < parts; p++) {
        x.part_begin[p] = p == 0 ? 0u : ends[p - 1];
        x.part_end[p] = ends[p];
    }
}

JoinArgs base_args(const hegj_index *h, const hegj_join_opts &opts, const JoinPlan &P) {
    JoinArgs a{};
    a.g = h->g;
    a.ix = h->ix;
    a.queue_heads = P.d_ctr->queue_heads;
    a.n_deferred = &P.d_ctr->n_deferred;
    a.cursor = &P.d_ctr->cursor;
    a.hash = &P.d_ctr->hash;
    a.evals = &P.d_ctr->evals;
    a.pairs = &P.d_ctr->pairs;
    a.deferred = P.deferred;
    a.gmask = P.gmask;
    a.gmask_slots = P.gmask_slots;
    a.block_waves = P.waves;
    a.part_by_xcc = opts.part_by_block ? 0u : 1u;
    a.prune = opts.no_prune ? 0u : 1u;
    a.band = &P.d_ctr->band;
    a.mismatch = &P.d_ctr->mismatch;
    if (P.f32) {
        a.f32 = 1;
        a.soa32 = h->soa32;
        a.aos32 = h->aos32;
        a.tacc32 = P.tacc32;
        a.trej32 = P.trej32;
    }
    return a;
}

// all_cells: units also for the cells the cell-0 wrap silences (the hybrid
// queue covers every query).
hegj_status make_plan(hegj_index *h, const hegj_join_opts &opts, uint32_t qb, const uint32_t *order,
                      JoinPlan &P, bool all_cells = false) {
    hipStream_t s = h->stream;
    if (opts.refine == HEGJ_REFINE_F32 || opts.refine == HEGJ_REFINE_AUTO) {
        bool ok = f32_supported(h->g.dim, qb, h->g.n);
        if (ok) {
            HEGJ_TRY(ensure_soa32(h));
            if (h->g.f32dt)  // the float chain itself: one threshold, no band
                P.tacc32 = P.trej32 = (float)h->g.thresh;
            else
                ok = f32_thresholds(h->g.thresh, h->g.dim, h->f32_radius, &P.tacc32, &P.trej32);
        }
        if (!ok && opts.refine == HEGJ_REFINE_F32)
            return fail(HEGJ_ERR_UNSUPPORTED,
                        "FP32 filter: built for dim 2 (4/8/16 queries per unit) and 6 (4/8), "
                        "up to 2^32 bytes of FP32 coordinates, and needs a band below 1% of eps^2");
        P.f32 = ok;
    }
    P.stream = s;
    HEGJ_TRY(build_units(h->g, h->ix, qb, order, &P.units, &P.U, s, all_cells));
    const uint32_t U = P.U;
    HIP_TRY(dev_malloc_n(&P.units_next, (size_t)U + 1));
    HIP_TRY(dev_malloc_n(&P.deferred, (size_t)U + 1));
    HIP_TRY(dev_malloc_n(&P.d_ctr, 1));
    HIP_TRY(dev_malloc_n(&P.d_ctr2, 1));
    HEGJ_TRY(join_occupancy(h->g.dim, qb, h->device, &P.max_blocks,
                            P.f32 ? JoinKind::f32 : h->g.f32dt ? JoinKind::flt : JoinKind::fp64));
    // per-wave hit-mask scratch for units whose masks exceed LDS
    set_waves(P, opts);
    HIP_TRY(dev_malloc_n(&P.gmask, (size_t)P.max_blocks * P.waves * P.gmask_slots));
    // unit ranges per XCD, split by estimated work
    P.parts = opts.xcd_parts ? std::min<uint32_t>(opts.xcd_parts, kMaxParts) : kMaxParts;
    if (U < P.parts * 256) P.parts = 1;
    HEGJ_TRY(split_units(h->ix, P.units, U, qb, h->d_work, P.parts, P.part_end, s));
    // batch estimate: count-only pass over a strided sample of ~8192 units
    // (a few per cent of the join's work at most)
    P.capacity = opts.result_capacity;
    if (P.capacity == 0 && U > 0) {
        EventTimer t(s);
        Scratch sc;
        const uint32_t stride = std::max<uint32_t>(1, U / 8192);
        const uint32_t ns = (U + stride - 1) / stride;
        uint2 *sample;
        HIP_TRY(sc.get(&sample, ns));
        HEGJ_TRY(stride_units(P.units, ns, stride, sample, s));
        HIP_TRY(hipMemsetAsync(P.d_ctr, 0, sizeof(Counters), s));
        JoinArgs e = base_args(h, opts, P);
        e.units = sample;
        e.n_units = ns;
        e.units_per_grab = grab_for(P, ns);
        e.count_only = 1;
        e.gmask_slots = 0;
        uint32_t one_end = ns;
        set_parts(e, 1, &one_end);
        HEGJ_TRY(launch_join(e, qb, blocks_for(P, ns, e.units_per_grab), s));
        Counters c;
        HIP_TRY(hipMemcpyAsync(&c, P.d_ctr, sizeof(c), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const double est = (double)c.pairs * ((double)U / (double)ns);
        P.estimated_pairs = (uint64_t)est;
        const uint32_t nb = opts.batches ? opts.batches : 1u;
        P.capacity = (uint64_t)(est * 1.10 / nb) + (1u << 20);
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        P.capacity = std::min<uint64_t>(P.capacity, (uint64_t)(free_b * 0.6) / sizeof(uint32_t));
        P.estimate_ms = t.stop();
    }
    if (P.capacity == 0) P.capacity = 1;
    return HEGJ_OK;
}

// Runs lists of work units into a result's batches.  A unit whose rows do
// not fit in the open batch is deferred; deferred units go to a new batch,
// sized exactly by a count-only pass over them (at most the first batch's
// size when the caller asked for several batches).  A batch may stay open across
// run() calls (the hybrid modes feed the GPU its share in chunks): its pair
// hash is taken when it closes, once every row it holds is written.  With
// `last` the batch closes right after the launch (no further units will
// come, or the launch defers and the batch is full anyway), so the hash pass
// follows the join without a host round trip.
struct BatchRunner {
    hegj_index *h;
    const hegj_join_opts &opts;
    uint32_t qb;
    JoinPlan &P;
    hegj_result *r;
    JoinArgs a{};
    uint64_t base = 0, capacity = 0, used = 0;
    bool open = false;
    double join_ms = 0;
    uint32_t launches = 0;
    Counters c{};

    BatchRunner(hegj_index *h_, const hegj_join_opts &o, uint32_t qb_, JoinPlan &P_, hegj_result *r_)
        : h(h_), opts(o), qb(qb_), P(P_), r(r_) {}

    hegj_status init() {
        hipStream_t s = h->stream;
        const uint32_t n = h->g.n;
        r->info.n_units = P.U;
        r->info.estimate_ms = P.estimate_ms;
        r->info.estimated_pairs = P.estimated_pairs;
        HIP_TRY(dev_malloc_n(&r->d_row_off, n));
        HIP_TRY(dev_malloc_n(&r->d_row_cnt, n));
        HIP_TRY(hipMemsetAsync(r->d_row_off, 0, sizeof(uint64_t) * n, s));
        HIP_TRY(hipMemsetAsync(r->d_row_cnt, 0, sizeof(uint32_t) * n, s));
        a = base_args(h, opts, P);
        a.row_off = r->d_row_off;
        a.row_cnt = r->d_row_cnt;
        HIP_TRY(hipMemsetAsync(P.d_ctr, 0, sizeof(Counters), s));
        capacity = P.capacity;
        return HEGJ_OK;
    }

    hegj_status open_batch() {
        if (r->batches.size() >= (size_t)kMaxBatches)
            return fail(HEGJ_ERR_CAPACITY, "result needs more than 64 batches");
        uint32_t *buf = nullptr;
        HIP_TRY(dev_malloc_n(&buf, capacity));
        r->batches.push_back(buf);
        r->bases.push_back(base);
        // the batch cursor starts at zero (hash / evals / pairs accumulate)
        HIP_TRY(hipMemsetAsync(&P.d_ctr->cursor, 0, sizeof(P.d_ctr->cursor), h->stream));
        a.out = buf;
        a.out_base = base;
        a.capacity = capacity;
        open = true;
        used = 0;
        return HEGJ_OK;
    }

    hegj_status close_batch(bool hashed) {
        if (!hashed) HEGJ_TRY(launch_pair_hash(a, capacity, h->stream));
        base += capacity;
        open = false;
        return HEGJ_OK;
    }

    hegj_status run(const uint2 *units, uint32_t nu, bool parts, bool last) {
        hipStream_t s = h->stream;
        const uint2 *cur = units;
        uint32_t cur_n = nu;
        while (cur_n > 0) {
            if (!open) HEGJ_TRY(open_batch());
            const uint64_t before = used;
            // queue heads and the deferral count restart; the cursor continues
            HIP_TRY(hipMemsetAsync(P.d_ctr, 0, offsetof(Counters, cursor), s));
            a.units = cur;
            a.n_units = cur_n;
            a.units_per_grab = opts.units_per_grab ? opts.units_per_grab : grab_for(P, cur_n);
            if (parts && cur == units && cur_n == P.U) {
                set_parts(a, P.parts, P.part_end);
            } else {
                uint32_t one_end = cur_n;
                set_parts(a, 1, &one_end);
            }
            a.count_only = 0;
            {
                EventTimer t(s);
                HEGJ_TRY(launch_join(a, qb, blocks_for(P, cur_n, a.units_per_grab), s));
                join_ms += t.stop();
                launches++;
            }
            if (last) HEGJ_TRY(launch_pair_hash(a, capacity, s));
            HIP_TRY(hipMemcpyAsync(&c, P.d_ctr, sizeof(c), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (c.mismatch)
                return fail(HEGJ_ERR_HIP, "join: the write pass of " + std::to_string(c.mismatch) +
                                              " work units disagreed with their counts");
            used = c.cursor;
            if (c.n_deferred == 0) {
                if (last) HEGJ_TRY(close_batch(true));
                break;
            }
            if (c.n_deferred == cur_n && before == 0) {
                // nothing fit in an empty batch: one unit alone exceeds it; grow
                dev_free(r->batches.back());
                r->batches.pop_back();
                r->bases.pop_back();
                open = false;
                capacity *= 2;
                continue;
            }
            HEGJ_TRY(close_batch(last));  // full
            r->info.n_deferred += c.n_deferred;
            // deferred[] indexes the current list; gather into a buffer that
            // is neither the current list nor the caller's
            uint2 *nxt = (cur == P.units_next) ? P.units_tmp : P.units_next;
            if (!nxt) {
                HIP_TRY(dev_malloc_n(&P.units_tmp, (size_t)P.U + 1));
                nxt = P.units_tmp;
            }
            HEGJ_TRY(gather_units(cur, P.deferred, c.n_deferred, nxt, s));
            cur = nxt;
            cur_n = c.n_deferred;
            // size the next batch exactly: count-only pass over what is left
            JoinArgs e = a;
            e.units = cur;
            e.n_units = cur_n;
            e.units_per_grab = grab_for(P, e.n_units);
            uint32_t one_end = e.n_units;
            set_parts(e, 1, &one_end);
            e.count_only = 1;
            e.gmask_slots = 0;
            e.queue_heads = P.d_ctr2->queue_heads;
            e.n_deferred = &P.d_ctr2->n_deferred;
            e.cursor = &P.d_ctr2->cursor;
            e.hash = &P.d_ctr2->hash;
            e.evals = &P.d_ctr2->evals;
            e.pairs = &P.d_ctr2->pairs;
            HIP_TRY(hipMemsetAsync(P.d_ctr2, 0, sizeof(Counters), s));
            HEGJ_TRY(launch_join(e, qb, blocks_for(P, e.n_units, e.units_per_grab), s));
            Counters c2;
            HIP_TRY(hipMemcpyAsync(&c2, P.d_ctr2, sizeof(c2), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (!opts.result_capacity) {
                capacity = std::max<uint64_t>(c2.pairs, 1);
                // a requested batch count keeps every batch at the first one's size
                if (opts.batches > 1) capacity = std::min<uint64_t>(capacity, P.capacity);
            }
        }
        return HEGJ_OK;
    }

    // Closes the open batch and fills the result's figures.
    hegj_status finish() {
        hipStream_t s = h->stream;
        if (open) {
            HEGJ_TRY(close_batch(false));
            HIP_TRY(hipMemcpyAsync(&c, P.d_ctr, sizeof(c), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        r->info.total_neighbors = c.pairs;
        r->info.pair_hash = c.hash;
        r->info.evals_executed = c.evals;
        r->info.band_rechecks = c.band;
        r->info.refine_used = P.f32 ? HEGJ_REFINE_F32 : HEGJ_REFINE_FP64;
        r->info.n_batches = (uint32_t)r->batches.size();
        r->info.join_ms = join_ms;
        r->info.join_launches = launches;
        r->info.join_kernel_avg_ms = launches ? join_ms / launches : 0.0;
        return HEGJ_OK;
    }
};

// Algorithm 0: every unit in one run (per-XCD ranges on the first launch).
hegj_status run_batches(hegj_index *h, const hegj_join_opts &opts, uint32_t qb, JoinPlan &P,
                        hegj_result *r) {
    BatchRunner B(h, opts, qb, P, r);
    HEGJ_TRY(B.init());
    HEGJ_TRY(B.run(P.units, P.U, true, true));
    return B.finish();
}

// ---------------------------------------------------------------------------
// Hybrid modes (algorithms 1-3).  The work queue is the unit list in workload
// order; `head` is the GPU's next unit, `tail` one past the CPU's last free
// unit, both under one host lock (WorkQueue.cpp:31-118 keeps queueIndex /
// queueIndexCPU the same way).  The GPU side runs its chunks through the
// batch runner on the calling thread; the CPU threads first build the
// Super-EGO order, then take batches of whole units from the tail.
// ---------------------------------------------------------------------------
double g_model_gpu[4] = {-0.9589, 0.4342, 0.3532, 0.2686};   // StaticPartition.cpp:42-46
double g_model_cpu[4] = {-7.2411, 0.9499, 0.7669, 0.4269};   // StaticPartition.cpp:55-59
std::mutex g_model_mu;

struct CpuRows {  // rows one CPU thread answered
    std::vector<uint32_t> qid, cnt, nbrs;
    uint64_t scanned = 0;
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

hegj_status selfjoin_hybrid(hegj_index *h, const hegj_join_opts &opts, uint32_t qb,
                            const hegj_hybrid_opts &ho, hegj_result *r) {
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = h->stream;
    const uint32_t n = h->g.n, dim = h->g.dim, C = h->ix.C;
    hegj_hybrid_info &hi = r->hinfo;
    hi.mode = ho.mode;
    hi.cpu_threads = ho.cpu_threads ? ho.cpu_threads : std::max(1u, std::thread::hardware_concurrency());
    const uint32_t cpu_batch = ho.cpu_batch ? ho.cpu_batch : 1024u;
    // the queue: every cell's units in workload order
    JoinPlan P;
    HEGJ_TRY(make_plan(h, opts, qb, h->d_order, P, true));
    const uint32_t U = P.U;
    std::vector<uint2> units(U);
    std::vector<uint32_t> cstart(C + 1), ila(n);
    std::vector<double> pts((size_t)n * dim);
    HIP_TRY(hipMemcpyAsync(units.data(), P.units, sizeof(uint2) * U, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(cstart.data(), h->ix.cell_start, sizeof(uint32_t) * (C + 1), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ila.data(), h->ix.ila, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(pts.data(), h->ix.pts, sizeof(double) * pts.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // queries before each unit
    std::vector<uint64_t> qpre(U + 1, 0);
    for (uint32_t u = 0; u < U; u++) {
        const uint32_t cell = units[u].x, size = cstart[cell + 1] - cstart[cell];
        qpre[u + 1] = qpre[u] + std::min(qb, size - units[u].y);
    }
    // GPU share: everything (dynamic), the static split, or nothing
    uint32_t gpu_limit = U;
    if (ho.mode == HEGJ_MODE_CPU) {
        gpu_limit = 0;
    } else if (ho.mode == HEGJ_MODE_STATIC) {
        double frac = ho.gpu_fraction;
        if (!(frac > 0.0)) frac = hegj_static_partition_model(n, dim, h->g.eps);
        hi.gpu_fraction = frac;
        // partitionedDBSIZE = DBSIZE * staticPartition in float (GPU.cu:437),
        // moved up to the next unit boundary
        const uint64_t target = (uint64_t)((float)n * (float)frac);
        gpu_limit = (uint32_t)(std::lower_bound(qpre.begin(), qpre.end(), target) - qpre.begin());
        gpu_limit = std::min(gpu_limit, U);
    }
    const bool dynamic = ho.mode == HEGJ_MODE_HYBRID;
    std::mutex mu;
    uint32_t head = 0, tail = U;

    // CPU side
    std::vector<CpuRows> rows(hi.cpu_threads);
    EgoIndex *ego = nullptr;
    hegj_status cpu_status = HEGJ_OK;
    std::string cpu_error;
    std::thread cpu([&]() {
        double ro = 0, so = 0;
        const hegj_status st = ego_build(pts.data(), n, dim, h->g.eps, hi.cpu_threads, &ego, &ro, &so);
        if (st != HEGJ_OK) {
            cpu_status = st;
            cpu_error = hegj_last_error();
            return;
        }
        hi.reorder_s = ro;
        hi.sort_s = so;
        const auto tc = std::chrono::steady_clock::now();
        std::vector<std::thread> workers;
        for (uint32_t t = 0; t < hi.cpu_threads; t++)
            workers.emplace_back([&, t]() {
                CpuRows &mine = rows[t];
                std::vector<uint32_t> row;
                while (true) {
                    uint32_t a, b;
                    {
                        std::lock_guard<std::mutex> lk(mu);
                        const uint32_t floor_u = dynamic ? head : gpu_limit;
                        if (tail <= floor_u) break;
                        b = tail;
                        a = b;
                        while (a > floor_u && qpre[b] - qpre[a] < cpu_batch) a--;
                        tail = a;
                    }
                    for (uint32_t u = a; u < b; u++) {
                        const uint32_t p0 = cstart[units[u].x] + units[u].y;
                        const uint32_t nq = (uint32_t)(qpre[u + 1] - qpre[u]);
                        for (uint32_t i = 0; i < nq; i++) {
                            const uint32_t q = ila[p0 + i];
                            row.clear();
                            mine.scanned += ego_query(*ego, q, row);
                            mine.qid.push_back(q);
                            mine.cnt.push_back((uint32_t)row.size());
                            mine.nbrs.insert(mine.nbrs.end(), row.begin(), row.end());
                        }
                    }
                }
            });
        for (auto &w : workers) w.join();
        hi.cpu_s = seconds_since(tc);
        hi.cpu_end_s = seconds_since(t0);
    });

    // GPU side, on this thread
    BatchRunner B(h, opts, qb, P, r);
    hegj_status gst = B.init();
    const auto tg = std::chrono::steady_clock::now();
    const uint32_t chunk = ho.gpu_chunk ? ho.gpu_chunk : std::max<uint32_t>(1, (U + 15) / 16);
    uint64_t gq = 0;
    while (gst == HEGJ_OK) {
        uint32_t a, b;
        {
            std::lock_guard<std::mutex> lk(mu);
            const uint32_t end = std::min(tail, gpu_limit);
            if (head >= end) break;
            a = head;
            b = std::min(end, head + chunk);
            head = b;
        }
        gst = B.run(P.units + a, b - a, false, false);
        gq += qpre[b] - qpre[a];
    }
    if (gst == HEGJ_OK) gst = B.finish();
    hi.gpu_s = seconds_since(tg);
    hi.gpu_end_s = seconds_since(t0);
    if (gst != HEGJ_OK) {  // stop the CPU side from taking more, then report
        const std::string msg = hegj_last_error();
        {
            std::lock_guard<std::mutex> lk(mu);
            tail = 0;
            gpu_limit = 0;
        }
        cpu.join();
        ego_free(ego);
        return fail(gst, msg);
    }
    cpu.join();
    ego_free(ego);
    if (cpu_status != HEGJ_OK) return fail(cpu_status, cpu_error);
    hi.n_queries_gpu = (uint32_t)gq;
    hi.pairs_gpu = r->info.total_neighbors;

    // the CPU rows join the result as one more batch
    uint64_t cpu_pairs = 0, cpu_hash = 0, m = 0;
    for (const CpuRows &c : rows) {
        cpu_pairs += c.nbrs.size();
        m += c.qid.size();
        hi.candidates_cpu += c.scanned;
    }
    hi.n_queries_cpu = (uint32_t)m;
    hi.pairs_cpu = cpu_pairs;
    if (m > 0) {
        std::vector<uint32_t> qid, cnt, nb;
        std::vector<uint64_t> off;
        qid.reserve(m);
        cnt.reserve(m);
        off.reserve(m);
        nb.reserve(cpu_pairs);
        for (const CpuRows &c : rows) {
            size_t k = 0;
            for (size_t i = 0; i < c.qid.size(); i++) {
                qid.push_back(c.qid[i]);
                cnt.push_back(c.cnt[i]);
                off.push_back(nb.size());
                for (uint32_t j = 0; j < c.cnt[i]; j++, k++) {
                    nb.push_back(c.nbrs[k]);
                    cpu_hash += mix64(((uint64_t)c.qid[i] << 32) | c.nbrs[k]);
                }
            }
        }
        if (r->batches.size() >= (size_t)kMaxBatches)
            return fail(HEGJ_ERR_CAPACITY, "result needs more than 64 batches");
        uint32_t *buf = nullptr;
        HIP_TRY(dev_malloc_n(&buf, std::max<uint64_t>(cpu_pairs, 1)));
        r->batches.push_back(buf);
        r->bases.push_back(B.base);
        Scratch sc;
        uint32_t *d_qid, *d_cnt;
        uint64_t *d_off;
        HIP_TRY(sc.get(&d_qid, m));
        HIP_TRY(sc.get(&d_cnt, m));
        HIP_TRY(sc.get(&d_off, m));
        if (cpu_pairs)
            HIP_TRY(hipMemcpyAsync(buf, nb.data(), sizeof(uint32_t) * cpu_pairs, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_qid, qid.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_cnt, cnt.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_off, off.data(), sizeof(uint64_t) * m, hipMemcpyHostToDevice, s));
        HEGJ_TRY(scatter_rows(d_qid, d_off, d_cnt, (uint32_t)m, B.base, r->d_row_off, r->d_row_cnt, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    r->info.total_neighbors += cpu_pairs;
    r->info.pair_hash += cpu_hash;
    r->info.n_batches = (uint32_t)r->batches.size();
    r->hybrid = true;
    return HEGJ_OK;
}

// A FP64 join plan kept by the index together with a captured graph of one
// batch: argument upload (from pinned host memory) -> counter reset (a fill
// kernel) -> join kernel reading its arguments from device memory -> counter
// download.  Replays rewrite the pinned argument block and launch the graph.
}  // namespace

struct hegj::GraphPlan {
    uint32_t key[11] = {};
    JoinPlan P;
    uint32_t qb = 0, blocks = 0, grab = 1;
    JoinArgs *d_args = nullptr, *h_args = nullptr;
    Counters *h_ctr = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (d_args) (void)hipFree(d_args);
        if (h_args) (void)hipHostFree(h_args);
        if (h_ctr) (void)hipHostFree(h_ctr);
    }
};

void hegj::free_graph_plan(GraphPlan *g) { delete g; }

namespace {

hegj_status capture_graph(hegj_index *h, GraphPlan &G) {
    hipStream_t s = h->stream;
    JoinArgs proto{};
    proto.g = h->g;  // with f32: selects the kernel instantiation
    proto.f32 = G.P.f32 ? 1u : 0u;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hegj_status st = HEGJ_OK;
    if (hipMemcpyAsync(G.d_args, G.h_args, sizeof(JoinArgs), hipMemcpyHostToDevice, s) != hipSuccess)
        st = fail(HEGJ_ERR_HIP, "graph capture: argument upload");
    if (st == HEGJ_OK) st = zero_words(G.P.d_ctr, sizeof(Counters), s);
    if (st == HEGJ_OK) st = launch_join(proto, G.qb, G.blocks, s, G.d_args);
    if (st == HEGJ_OK) st = launch_pair_hash(proto, G.P.capacity, s, G.d_args);
    if (st == HEGJ_OK &&
        hipMemcpyAsync(G.h_ctr, G.P.d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess)
        st = fail(HEGJ_ERR_HIP, "graph capture: counter download");
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (st != HEGJ_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    HIP_TRY(e);
    const hipError_t ie = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(ie);
    return HEGJ_OK;
}

hegj_status selfjoin_graph(hegj_index *h, const hegj_join_opts &opts, uint32_t qb,
                           const uint32_t *order, hegj_result *r, bool *fits) {
    hipStream_t s = h->stream;
    const uint32_t n = h->g.n;
    const uint32_t key[11] = {qb, opts.unit_order, opts.xcd_parts, opts.units_per_grab,
                              opts.part_by_block, (uint32_t)opts.result_capacity,
                              (uint32_t)(opts.result_capacity >> 32), opts.refine, opts.no_prune,
                              opts.block_waves, opts.batches};
    if (h->gplan && std::memcmp(h->gplan->key, key, sizeof(key)) != 0) {
        free_graph_plan(h->gplan);
        h->gplan = nullptr;
    }
    if (!h->gplan) {
        std::unique_ptr<GraphPlan> G(new GraphPlan());
        std::memcpy(G->key, key, sizeof(key));
        G->qb = qb;
        HEGJ_TRY(make_plan(h, opts, qb, order, G->P));
        G->grab = opts.units_per_grab ? opts.units_per_grab : grab_for(G->P, G->P.U);
        G->blocks = blocks_for(G->P, G->P.U, G->grab);
        HIP_TRY(hipMalloc(&G->d_args, sizeof(JoinArgs)));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&G->h_args), sizeof(JoinArgs), hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&G->h_ctr), sizeof(Counters), hipHostMallocDefault));
        std::memset(G->h_args, 0, sizeof(JoinArgs));
        HEGJ_TRY(capture_graph(h, *G));
        h->gplan = G.release();
    }
    GraphPlan &G = *h->gplan;
    r->info.n_units = G.P.U;
    r->info.estimate_ms = G.P.estimate_ms;
    r->info.estimated_pairs = G.P.estimated_pairs;
    HIP_TRY(dev_malloc_n(&r->d_row_off, n));
    HIP_TRY(dev_malloc_n(&r->d_row_cnt, n));
    HIP_TRY(hipMemsetAsync(r->d_row_off, 0, sizeof(uint64_t) * n, s));
    HIP_TRY(hipMemsetAsync(r->d_row_cnt, 0, sizeof(uint32_t) * n, s));
    uint32_t *buf = nullptr;
    HIP_TRY(dev_malloc_n(&buf, G.P.capacity));
    r->batches.push_back(buf);
    r->bases.push_back(0);
    JoinArgs a = base_args(h, opts, G.P);
    a.row_off = r->d_row_off;
    a.row_cnt = r->d_row_cnt;
    a.units = G.P.units;
    a.n_units = G.P.U;
    a.units_per_grab = G.grab;
    set_parts(a, G.P.parts, G.P.part_end);
    a.out = buf;
    a.out_base = 0;
    a.capacity = G.P.capacity;
    a.count_only = 0;
    HIP_TRY(hipStreamSynchronize(s));  // the previous replay has read the block
    *G.h_args = a;
    {
        EventTimer t(s);
        HIP_TRY(hipGraphLaunch(G.exec, s));
        r->info.join_ms = t.stop();  // synchronises the stream
    }
    const Counters c = *G.h_ctr;
    if (c.mismatch)
        return fail(HEGJ_ERR_HIP, "join: the write pass of " + std::to_string(c.mismatch) +
                                      " work units disagreed with their counts");
    if (c.n_deferred) {
        // grow the cached capacity for later replays; this call runs eagerly
        G.P.capacity = std::max<uint64_t>(G.P.capacity * 2, c.cursor + (1u << 20));
        *fits = false;
        return HEGJ_OK;
    }
    *fits = true;
    r->info.total_neighbors = c.pairs;
    r->info.pair_hash = c.hash;
    r->info.evals_executed = c.evals;
    r->info.band_rechecks = c.band;
    r->info.refine_used = G.P.f32 ? HEGJ_REFINE_F32 : HEGJ_REFINE_FP64;
    r->info.n_batches = 1;
    r->info.join_launches = 1;
    r->info.join_kernel_avg_ms = r->info.join_ms;
    r->info.graph_replay = 1;
    return HEGJ_OK;
}

hegj_status ensure_gb(hegj_index *h) {
    if (h->gb_A) return HEGJ_OK;
    const uint32_t kb = gb_kblocks(h->g.dim);
    if (kb == 0) return fail(HEGJ_ERR_UNSUPPORTED, "no MFMA refine for this dimensionality");
    EventTimer t(h->stream);
    HEGJ_TRY(gb_prepare(h->g, h->ix, kb * 32, &h->gb_A, &h->gb_B, &h->gb_w, h->stream));
    h->gb_kp = kb * 32;
    h->gb_ms = t.stop();
    return HEGJ_OK;
}

// Join through the MFMA guard-band refine: one pass, the pair list sized on
// the device, the band rechecked in FP64.  Same pair set as the FP64 kernel.
// result_capacity (neighbour entries) sizes the first attempt's pair list.
hegj_status selfjoin_mfma(hegj_index *h, const uint32_t *order, uint64_t result_capacity, uint32_t parts,
                          hegj_result *r) {
    hipStream_t s = h->stream;
    const uint32_t n = h->g.n;
    HEGJ_TRY(ensure_gb(h));
    Scratch sc;
    uint2 *units = nullptr;
    uint32_t U = 0;
    struct UnitsGuard {
        hipStream_t s;
        uint2 *&p;
        ~UnitsGuard() {
            (void)hipStreamSynchronize(s);
            dev_free(p);
        }
    } units_guard{s, units};
    HEGJ_TRY(build_units(h->g, h->ix, gb_unit_queries(h->gb_kp / 32), order, &units, &U, s));
    r->info.n_units = U;
    HIP_TRY(dev_malloc_n(&r->d_row_off, n));
    HIP_TRY(dev_malloc_n(&r->d_row_cnt, n));
    HIP_TRY(hipMemsetAsync(r->d_row_off, 0, sizeof(uint64_t) * n, s));
    HIP_TRY(hipMemsetAsync(r->d_row_cnt, 0, sizeof(uint32_t) * n, s));
    unsigned long long *d_hash;
    HIP_TRY(sc.get(&d_hash, 1));
    HIP_TRY(hipMemsetAsync(d_hash, 0, sizeof(*d_hash), s));
    GbOutput o;
    o.row_off = r->d_row_off;
    o.row_cnt = r->d_row_cnt;
    o.hash = d_hash;
    HEGJ_TRY(gb_join(h->g, h->ix, units, U, h->gb_A, h->gb_B, h->gb_w, h->gb_kp, h->device, s, &o,
                     result_capacity, parts));
    r->batches.push_back(o.buf);
    r->bases.push_back(0);
    unsigned long long hash = 0;
    HIP_TRY(hipMemcpyAsync(&hash, d_hash, sizeof(hash), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    r->info.total_neighbors = o.pairs;
    r->info.estimated_pairs = o.pairs;
    r->info.pair_hash = hash;
    r->info.evals_executed = o.evals;
    r->info.band_rechecks = o.band;
    r->info.n_batches = 1;
    r->info.estimate_ms = h->gb_ms;
    r->info.join_ms = o.kernel_ms;
    r->info.join_launches = 1;
    r->info.join_kernel_avg_ms = o.kernel_ms;
    r->info.refine_used = HEGJ_REFINE_MFMA;
    return HEGJ_OK;
}

}  // namespace

namespace hegj {
const GridGeom &index_geom(const hegj_index *h) { return h->g; }
const DevIndex &index_dev(const hegj_index *h) { return h->ix; }
hegj_status select_device(int device) { return set_device(device); }
hegj_status index_f32(hegj_index *h, const float **soa32, const float **aos32, double *radius,
                      const double **centre) {
    HEGJ_TRY(ensure_soa32(h));
    *soa32 = h->soa32;
    *aos32 = h->aos32;
    *radius = h->f32_radius;
    *centre = h->f32_centre;
    return HEGJ_OK;
}
}  // namespace hegj

extern "C" {

const char *hegj_status_string(hegj_status s) {
    switch (s) {
        case HEGJ_OK: return "ok";
        case HEGJ_ERR_INVALID_ARG: return "invalid argument";
        case HEGJ_ERR_UNSUPPORTED: return "unsupported";
        case HEGJ_ERR_NO_DEVICE: return "no device";
        case HEGJ_ERR_OUT_OF_MEMORY: return "out of device memory";
        case HEGJ_ERR_HIP: return "HIP error";
        case HEGJ_ERR_CAPACITY: return "result capacity exceeded";
        case HEGJ_ERR_IO: return "I/O error";
        case HEGJ_ERR_BUSY: return "busy";
    }
    return "unknown";
}

const char *hegj_last_error(void) { return g_last_error.c_str(); }

const char *hegj_version(void) { return "hegjoin-mi355x 0.1 (gfx950)"; }

int hegj_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

void hegj_release_cached_memory(int device) {
    if (device >= 0) {
        hegj::cache_release_device(device);
        return;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return;
    for (int d = 0; d < n; d++) hegj::cache_release_device(d);
}

hegj_status hegj_stream_copy_bandwidth(uint64_t bytes, int device, double *gbps) {
    if (!gbps || bytes < 4096) return fail(HEGJ_ERR_INVALID_ARG, "bytes must be >= 4096");
    HEGJ_TRY(set_device(device));
    double ms = 0;
    HEGJ_TRY(stream_copy(bytes, 10, &ms));
    *gbps = 2.0 * (double)(bytes / 16 * 16) / (ms * 1e-3) / 1e9;  // read + write
    return HEGJ_OK;
}

void hegj_index_opts_default(hegj_index_opts *o, uint32_t dim) {
    o->dim = dim;
    o->indexed_dims = dim < (uint32_t)HEGJ_MAX_INDEXED_DIMS ? dim : HEGJ_MAX_INDEXED_DIMS;
    o->dtype = HEGJ_DTYPE_F64;
    o->emulate_cell0 = 1;
    o->device = 0;
}

hegj_status hegj_index_build(const double *database, uint32_t n, double epsilon,
                             const hegj_index_opts *opts, hegj_index **out) {
    if (!database || !out) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    HEGJ_TRY(check_opts(opts, n, epsilon));
    return new_index(opts, out, [&](hegj_index *h) -> hegj_status {
        size_t bytes = sizeof(double) * (size_t)n * opts->dim;
        HIP_TRY(hipMalloc(&h->ix.pts, bytes));
        HIP_TRY(hipMemcpyAsync(h->ix.pts, database, bytes, hipMemcpyHostToDevice, h->stream));
        return build_common(h, n, epsilon, opts);
    });
}

hegj_status hegj_index_build_device(const double *dev_database, uint32_t n, double epsilon,
                                    const hegj_index_opts *opts, hegj_index **out) {
    if (!dev_database || !out) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    HEGJ_TRY(check_opts(opts, n, epsilon));
    return new_index(opts, out, [&](hegj_index *h) -> hegj_status {
        size_t bytes = sizeof(double) * (size_t)n * opts->dim;
        HIP_TRY(hipMalloc(&h->ix.pts, bytes));
        HIP_TRY(hipMemcpyAsync(h->ix.pts, dev_database, bytes, hipMemcpyDeviceToDevice, h->stream));
        return build_common(h, n, epsilon, opts);
    });
}

hegj_status hegj_index_free(hegj_index *h) {
    if (!h) return HEGJ_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_index(&h->ix);
    if (h->d_work) (void)hipFree(h->d_work);
    if (h->d_order) (void)hipFree(h->d_order);
    if (h->d_morton) (void)hipFree(h->d_morton);
    if (h->gb_A) (void)hipFree(h->gb_A);
    if (h->gb_B) (void)hipFree(h->gb_B);
    if (h->gb_w) (void)hipFree(h->gb_w);
    if (h->gplan) free_graph_plan(h->gplan);
    if (h->soa32) (void)hipFree(h->soa32);
    if (h->aos32) (void)hipFree(h->aos32);
    if (h->f32_centre) (void)hipFree(h->f32_centre);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return HEGJ_OK;
}

hegj_status hegj_index_get_info(const hegj_index *h, hegj_index_info *info) {
    if (!h || !info) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    std::memset(info, 0, sizeof(*info));
    info->n = h->g.n;
    info->dim = h->g.dim;
    info->indexed_dims = h->g.k;
    info->dtype = h->g.f32dt ? HEGJ_DTYPE_F32 : HEGJ_DTYPE_F64;
    info->nNonEmptyCells = h->ix.C;
    info->n_rows = h->ix.R;
    info->n_cell0_queries = h->n_dead;
    info->emulate_cell0 = h->g.emulate_cell0;
    info->totalCells = h->totalCells;
    info->epsilon = h->g.eps;
    info->threshold = h->g.thresh;
    for (uint32_t j = 0; j < h->g.k; j++) {
        info->minArr[j] = h->g.minArr[j];
        info->maxArr[j] = h->maxArr[j];
        info->nCells[j] = h->g.nCells[j];
    }
    info->build_ms = h->build_ms;
    return HEGJ_OK;
}

hegj_status hegj_index_copy_grid(const hegj_index *h, hegj_grid *index,
                                 hegj_grid_cell_lookup *lookup, uint32_t *ila) {
    if (!h) return fail(HEGJ_ERR_INVALID_ARG, "NULL index");
    HEGJ_TRY(set_device(h->device));
    const uint32_t C = h->ix.C, n = h->g.n;
    if (index || lookup) {
        std::vector<uint32_t> start(C + 1);
        std::vector<uint64_t> keys(C);
        HIP_TRY(hipMemcpyAsync(start.data(), h->ix.cell_start, sizeof(uint32_t) * (C + 1),
                               hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipMemcpyAsync(keys.data(), h->ix.cell_key, sizeof(uint64_t) * C,
                               hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (uint32_t c = 0; c < C; c++) {
            if (index) index[c] = hegj_grid{(int32_t)start[c], (int32_t)start[c + 1] - 1};
            if (lookup) lookup[c] = hegj_grid_cell_lookup{c, 0u, keys[c]};
        }
    }
    if (ila) {
        HIP_TRY(hipMemcpyAsync(ila, h->ix.ila, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    return HEGJ_OK;
}

hegj_status hegj_index_copy_cell_ids(const hegj_index *h, uint64_t *ids) {
    if (!h || !ids) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    HEGJ_TRY(set_device(h->device));
    HIP_TRY(hipMemcpyAsync(ids, h->ix.point_key, sizeof(uint64_t) * h->g.n, hipMemcpyDeviceToHost,
                           h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return HEGJ_OK;
}

hegj_status hegj_sort_by_workload(hegj_index *h, hegj_scheduling_cell *sorted, uint32_t *opi,
                                  uint64_t *total_candidates) {
    if (!h) return fail(HEGJ_ERR_INVALID_ARG, "NULL index");
    HEGJ_TRY(set_device(h->device));
    HEGJ_TRY(ensure_workload(h));
    const uint32_t C = h->ix.C, n = h->g.n;
    if (total_candidates) *total_candidates = h->total_candidates;
    if (sorted) {
        std::vector<uint32_t> order(C), work(C);
        HIP_TRY(hipMemcpyAsync(order.data(), h->d_order, sizeof(uint32_t) * C, hipMemcpyDeviceToHost,
                               h->stream));
        HIP_TRY(hipMemcpyAsync(work.data(), h->d_work, sizeof(uint32_t) * C, hipMemcpyDeviceToHost,
                               h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (uint32_t i = 0; i < C; i++)
            sorted[i] = hegj_scheduling_cell{(int32_t)work[order[i]], (int32_t)order[i]};
    }
    if (opi) {
        Scratch sc;
        uint32_t *d_opi;
        HIP_TRY(sc.get(&d_opi, n));
        HEGJ_TRY(origin_point_index(h->ix, h->d_order, d_opi, h->stream));
        HIP_TRY(hipMemcpyAsync(opi, d_opi, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    return HEGJ_OK;
}

void hegj_join_opts_default(hegj_join_opts *o) {
    std::memset(o, 0, sizeof(*o));
    o->unit_order = HEGJ_ORDER_MORTON;
}

static std::mutex g_tuning_mu;
static std::map<uint32_t, hegj_join_opts> g_tuning;

void hegj_join_opts_tuned(hegj_join_opts *o, uint32_t dim) {
    hegj_join_opts_default(o);
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    auto it = g_tuning.find(dim);
    if (it != g_tuning.end()) *o = it->second;
}

hegj_status hegj_tuning_set(uint32_t dim, const hegj_join_opts *opts) {
    if (!opts || dim == 0) return fail(HEGJ_ERR_INVALID_ARG, "NULL opts or dim 0");
    if (opts->queries_per_unit != 0 && opts->queries_per_unit != 4 && opts->queries_per_unit != 8 &&
        opts->queries_per_unit != 16)
        return fail(HEGJ_ERR_INVALID_ARG, "queries_per_unit must be 0, 4, 8 or 16");
    if (opts->unit_order > HEGJ_ORDER_MORTON || opts->refine > HEGJ_REFINE_F32)
        return fail(HEGJ_ERR_INVALID_ARG, "unknown unit_order or refine");
    if (!valid_waves_batches(*opts)) return fail(HEGJ_ERR_INVALID_ARG, "block_waves must be 0, 1, 2 or 4 and batches at most 32");
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    g_tuning[dim] = *opts;
    return HEGJ_OK;
}

hegj_status hegj_tuning_load(const char *path, uint32_t *n_loaded) {
    if (!path) return fail(HEGJ_ERR_INVALID_ARG, "NULL path");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(HEGJ_ERR_IO, std::string("cannot open ") + path);
    char line[512];
    uint32_t count = 0, lineno = 0;
    hegj_status st = HEGJ_OK;
    while (std::fgets(line, sizeof(line), f)) {
        lineno++;
        char *hash = std::strchr(line, '#');
        if (hash) *hash = 0;
        unsigned dim, qb, order, parts, grab, refine, waves = 0, batches = 0;
        const int got = std::sscanf(line, "%u %u %u %u %u %u %u %u", &dim, &qb, &order, &parts, &grab, &refine,
                                    &waves, &batches);
        if (got <= 0) continue;  // blank or comment
        if (got < 6) {
            st = fail(HEGJ_ERR_INVALID_ARG, "tuning table line " + std::to_string(lineno) + ": 6 to 8 fields expected");
            break;
        }
        hegj_join_opts o;
        hegj_join_opts_default(&o);
        o.queries_per_unit = qb;
        o.unit_order = order;
        o.xcd_parts = parts;
        o.units_per_grab = grab;
        o.refine = refine;
        o.block_waves = waves;
        o.batches = batches;
        st = hegj_tuning_set(dim, &o);
        if (st != HEGJ_OK) break;
        count++;
    }
    std::fclose(f);
    if (n_loaded) *n_loaded = count;
    return st;
}

static hegj_status mark_done(hegj_result *r, hipStream_t s) {
    HIP_TRY(hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(r->done, s));
    return HEGJ_OK;
}

hegj_status hegj_selfjoin(hegj_index *h, const hegj_join_opts *opts_in, hegj_result **out) {
    if (!h || !out) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    auto t_start = std::chrono::steady_clock::now();
    HEGJ_TRY(set_device(h->device));
    hegj_join_opts opts;
    if (opts_in) {
        opts = *opts_in;
    } else {
        hegj_join_opts_tuned(&opts, h->g.dim);
        // the table is tuned for DTYPE=double: float data take the grid kernels
        if (h->g.f32dt && opts.refine == HEGJ_REFINE_MFMA) opts.refine = HEGJ_REFINE_AUTO;
    }
    if (opts.refine > HEGJ_REFINE_F32) return fail(HEGJ_ERR_INVALID_ARG, "unknown refine");
    if (opts.refine == HEGJ_REFINE_MFMA && h->g.f32dt)
        return fail(HEGJ_ERR_UNSUPPORTED, "the MFMA refine is built for DTYPE=double");
    const bool mfma = opts.refine == HEGJ_REFINE_MFMA ||
                      (opts.refine == HEGJ_REFINE_AUTO && !h->g.f32dt && h->g.dim >= 16 &&
                       gb_kblocks(h->g.dim) != 0);
    if (mfma && gb_kblocks(h->g.dim) == 0)
        return fail(HEGJ_ERR_UNSUPPORTED, "no MFMA refine for this dimensionality");
    const uint32_t qb = opts.queries_per_unit ? opts.queries_per_unit : default_qb(h->g.dim);
    if (!mfma && qb != 4 && qb != 8 && qb != 16)
        return fail(HEGJ_ERR_INVALID_ARG, "queries_per_unit must be 4, 8 or 16");
    if (!valid_waves_batches(opts))
        return fail(HEGJ_ERR_INVALID_ARG, "block_waves must be 0, 1, 2 or 4 and batches at most 32");
    hipStream_t s = h->stream;
    const uint32_t n = h->g.n;

    std::unique_ptr<hegj_result> r(new hegj_result());
    r->idx = h;
    HEGJ_TRY(ensure_workload(h));
    r->info.total_candidates = h->total_candidates;
    r->info.workload_ms = h->workload_ms;
    r->info.n_queries = n - h->n_dead;

    const uint32_t *order = nullptr;
    if (opts.unit_order == HEGJ_ORDER_WORKLOAD) {
        order = h->d_order;
    } else if (opts.unit_order == HEGJ_ORDER_MORTON) {
        HEGJ_TRY(ensure_morton(h));
        if (h->morton_ok) order = h->d_morton;
    } else if (opts.unit_order != HEGJ_ORDER_SPATIAL) {
        return fail(HEGJ_ERR_INVALID_ARG, "unknown unit_order");
    }
    if (mfma) {
        // xcd_parts 0 (auto): one unit range per XCD
        HEGJ_TRY(selfjoin_mfma(h, order, opts.result_capacity, opts.xcd_parts ? opts.xcd_parts : kMaxParts,
                               r.get()));
        r->info.total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        HEGJ_TRY(mark_done(r.get(), s));
        *out = r.release();
        return HEGJ_OK;
    }
    r->info.refine_used = HEGJ_REFINE_FP64;
    if (opts.use_graph && !h->g.f32dt) {
        bool fits = false;
        HEGJ_TRY(selfjoin_graph(h, opts, qb, order, r.get(), &fits));
        if (!fits) {  // the cached capacity was too small: this call runs eagerly
            r.reset(new hegj_result());
            r->idx = h;
            r->info.total_candidates = h->total_candidates;
            r->info.workload_ms = h->workload_ms;
            r->info.n_queries = n - h->n_dead;
            r->info.refine_used = HEGJ_REFINE_FP64;
            JoinPlan P;
            HEGJ_TRY(make_plan(h, opts, qb, order, P));
            HEGJ_TRY(run_batches(h, opts, qb, P, r.get()));
        }
    } else {
        JoinPlan P;
        HEGJ_TRY(make_plan(h, opts, qb, order, P));
        HEGJ_TRY(run_batches(h, opts, qb, P, r.get()));
    }
    r->info.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    HEGJ_TRY(mark_done(r.get(), s));
    *out = r.release();
    return HEGJ_OK;
}

namespace {
std::mutex g_noalloc_mu[64];  // one per device: calls on a device run one at a time

struct ArenaGuard {
    explicit ArenaGuard(Arena *a) { set_thread_arena(a); }
    ~ArenaGuard() { set_thread_arena(nullptr); }
};
}  // namespace

hegj_status hegj_selfjoin_noalloc(hegj_index *h, const hegj_join_opts *opts_in, uint32_t *neighbors,
                                  uint64_t capacity, uint64_t *row_off, uint32_t *row_cnt,
                                  hegj_result_info *info) {
    if (!h || !neighbors || !row_off || !row_cnt || !info || capacity == 0)
        return fail(HEGJ_ERR_INVALID_ARG, "NULL argument or zero capacity");
    if (h->device < 0 || h->device >= 64) return fail(HEGJ_ERR_INVALID_ARG, "device ordinal out of range");
    const auto t_start = std::chrono::steady_clock::now();
    HEGJ_TRY(set_device(h->device));
    hegj_join_opts opts;
    if (opts_in) {
        opts = *opts_in;
    } else {
        hegj_join_opts_tuned(&opts, h->g.dim);
        if (opts.refine == HEGJ_REFINE_MFMA) opts.refine = HEGJ_REFINE_AUTO;  // the grid kernels
    }
    if (opts.refine == HEGJ_REFINE_MFMA)
        return fail(HEGJ_ERR_UNSUPPORTED, "hegj_selfjoin_noalloc runs the grid join kernels, not the MFMA refine");
    if (opts.refine > HEGJ_REFINE_F32) return fail(HEGJ_ERR_INVALID_ARG, "unknown refine");
    const uint32_t qb = opts.queries_per_unit ? opts.queries_per_unit : default_qb(h->g.dim);
    if (qb != 4 && qb != 8 && qb != 16) return fail(HEGJ_ERR_INVALID_ARG, "queries_per_unit must be 4, 8 or 16");
    if (!valid_waves_batches(opts))
        return fail(HEGJ_ERR_INVALID_ARG, "block_waves must be 0, 1, 2 or 4 and batches at most 32");
    // the index's lazily built arrays are inputs here: they must exist
    const char *prep = "the index must have run hegj_selfjoin once with these options first";
    if (!h->have_workload) return fail(HEGJ_ERR_INVALID_ARG, std::string(prep) + " (workload order)");
    const uint32_t *order = nullptr;
    if (opts.unit_order == HEGJ_ORDER_WORKLOAD) {
        order = h->d_order;
    } else if (opts.unit_order == HEGJ_ORDER_MORTON) {
        if (!h->d_morton) return fail(HEGJ_ERR_INVALID_ARG, std::string(prep) + " (Morton order)");
        if (h->morton_ok) order = h->d_morton;
    } else if (opts.unit_order != HEGJ_ORDER_SPATIAL) {
        return fail(HEGJ_ERR_INVALID_ARG, "unknown unit_order");
    }
    bool f32 = false;
    float tacc = 0, trej = 0;
    if (opts.refine == HEGJ_REFINE_F32 || opts.refine == HEGJ_REFINE_AUTO) {
        bool ok = f32_supported(h->g.dim, qb, h->g.n) && h->soa32;
        if (ok) {
            if (h->g.f32dt)
                tacc = trej = (float)h->g.thresh;
            else
                ok = f32_thresholds(h->g.thresh, h->g.dim, h->f32_radius, &tacc, &trej);
        }
        if (!ok && opts.refine == HEGJ_REFINE_F32)
            return fail(HEGJ_ERR_UNSUPPORTED, f32_supported(h->g.dim, qb, h->g.n)
                                                  ? std::string(prep) + " (FP32 copy)"
                                                  : std::string("FP32 filter not built for this dim / queries_per_unit"));
        f32 = ok;
    }
    hipStream_t s = h->stream;
    const uint32_t n = h->g.n;
    std::lock_guard<std::mutex> lk(g_noalloc_mu[h->device]);
    void *base = nullptr;
    HEGJ_TRY(noalloc_scratch(h->device, &base));
    Arena arena;
    arena.base = static_cast<char *>(base);
    arena.cap = kNoAllocScratchBytes;
    ArenaGuard guard(&arena);  // every temporary below comes from the array
    // declared after the guard, so its destructor (which hands its buffers
    // back: a no-op for arena memory) runs while the arena is installed
    JoinPlan P;
    P.f32 = f32;
    P.tacc32 = tacc;
    P.trej32 = trej;
    auto too_small = [&](hegj_status st) {
        if (st != HEGJ_OK && arena.used + (1u << 20) >= arena.cap)
            return fail(HEGJ_ERR_CAPACITY, "hegj_selfjoin_noalloc: this index needs more than the " +
                                               std::to_string(kNoAllocScratchBytes >> 20) +
                                               " MiB of fixed scratch (n = " + std::to_string(n) +
                                               ", cells = " + std::to_string(h->ix.C) + ")");
        return st;
    };
    // keep the plan's pointers out of the caller's way: JoinPlan frees them
    // (a no-op for arena memory) before the guard uninstalls the arena
    P.stream = s;
    HIP_TRY(dev_malloc_n(&P.d_ctr, 1));
    HIP_TRY(dev_malloc_n(&P.d_ctr2, 1));
    HEGJ_TRY(zero_words(P.d_ctr, sizeof(Counters), s));
    HEGJ_TRY(zero_words(row_off, sizeof(uint64_t) * n, s));
    HEGJ_TRY(zero_words(row_cnt, sizeof(uint32_t) * n, s));
    HEGJ_TRY(too_small(build_units(h->g, h->ix, qb, order, &P.units, &P.U, s)));
    const uint32_t U = P.U;
    if (arena_alloc(reinterpret_cast<void **>(&P.deferred), sizeof(uint32_t) * ((size_t)U + 1)) != hipSuccess)
        return too_small(HEGJ_ERR_OUT_OF_MEMORY);
    HEGJ_TRY(join_occupancy(h->g.dim, qb, h->device, &P.max_blocks,
                            P.f32 ? JoinKind::f32 : h->g.f32dt ? JoinKind::flt : JoinKind::fp64));
    set_waves(P, opts);
    P.gmask_slots = 0;  // units whose masks exceed LDS re-evaluate in the write pass
    P.parts = opts.xcd_parts ? std::min<uint32_t>(opts.xcd_parts, kMaxParts) : kMaxParts;
    if (U < P.parts * 256) P.parts = 1;
    HEGJ_TRY(too_small(split_units(h->ix, P.units, U, qb, h->d_work, P.parts, P.part_end, s)));
    JoinArgs a = base_args(h, opts, P);
    a.row_off = row_off;
    a.row_cnt = row_cnt;
    a.units = P.units;
    a.n_units = U;
    a.units_per_grab = opts.units_per_grab ? opts.units_per_grab : grab_for(P, U);
    set_parts(a, P.parts, P.part_end);
    a.out = neighbors;
    a.out_base = 0;
    a.capacity = capacity;
    a.count_only = 0;
    double join_ms = 0;  // host clock: no events are created here
    Counters c{};
    if (U > 0) {
        HIP_TRY(hipStreamSynchronize(s));
        const auto tj = std::chrono::steady_clock::now();
        HEGJ_TRY(launch_join(a, qb, blocks_for(P, U, a.units_per_grab), s));
        HIP_TRY(hipStreamSynchronize(s));
        join_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tj).count();
        HEGJ_TRY(launch_pair_hash(a, capacity, s));
        HIP_TRY(hipMemcpyAsync(&c, P.d_ctr, sizeof(c), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    std::memset(info, 0, sizeof(*info));
    info->total_candidates = h->total_candidates;
    info->n_queries = n - h->n_dead;
    info->n_units = U;
    info->join_ms = join_ms;
    info->join_launches = U > 0 ? 1 : 0;
    info->join_kernel_avg_ms = join_ms;
    info->workload_ms = h->workload_ms;
    info->refine_used = P.f32 ? HEGJ_REFINE_F32 : HEGJ_REFINE_FP64;
    if (c.mismatch)
        return fail(HEGJ_ERR_HIP, "join: the write pass of " + std::to_string(c.mismatch) +
                                      " work units disagreed with their counts");
    if (c.n_deferred) {
        // the rows do not fit: count them all (count-only pass) for the caller
        JoinArgs e = a;
        e.count_only = 1;
        e.queue_heads = P.d_ctr2->queue_heads;
        e.n_deferred = &P.d_ctr2->n_deferred;
        e.cursor = &P.d_ctr2->cursor;
        e.hash = &P.d_ctr2->hash;
        e.evals = &P.d_ctr2->evals;
        e.pairs = &P.d_ctr2->pairs;
        HEGJ_TRY(zero_words(P.d_ctr2, sizeof(Counters), s));
        HEGJ_TRY(launch_join(e, qb, blocks_for(P, U, e.units_per_grab), s));
        Counters c2;
        HIP_TRY(hipMemcpyAsync(&c2, P.d_ctr2, sizeof(c2), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        info->total_neighbors = c2.pairs;
        return fail(HEGJ_ERR_CAPACITY, "hegj_selfjoin_noalloc: the result has " + std::to_string(c2.pairs) +
                                           " neighbours, the buffer holds " + std::to_string(capacity) +
                                           " (info->total_neighbors has the count)");
    }
    info->total_neighbors = c.pairs;
    info->pair_hash = c.hash;
    info->evals_executed = c.evals;
    info->band_rechecks = c.band;
    info->n_batches = 1;
    info->estimated_pairs = 0;
    info->total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return HEGJ_OK;
}

void hegj_hybrid_opts_default(hegj_hybrid_opts *o, uint32_t mode) {
    std::memset(o, 0, sizeof(*o));
    o->mode = mode;
}

hegj_status hegj_selfjoin_hybrid(hegj_index *h, const hegj_join_opts *opts_in,
                                 const hegj_hybrid_opts *ho, hegj_result **out) {
    if (!h || !ho || !out) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    if (ho->mode == HEGJ_MODE_GPU) return hegj_selfjoin(h, opts_in, out);
    if (ho->mode > HEGJ_MODE_CPU) return fail(HEGJ_ERR_INVALID_ARG, "unknown search mode");
    if (ho->mode == HEGJ_MODE_STATIC && ho->gpu_fraction >= 1.0)
        return fail(HEGJ_ERR_INVALID_ARG, "the static partition must be in ]0.0, 1.0[");
    const auto t_start = std::chrono::steady_clock::now();
    HEGJ_TRY(set_device(h->device));
    hegj_join_opts opts;
    if (opts_in) {
        opts = *opts_in;
    } else {
        hegj_join_opts_tuned(&opts, h->g.dim);
        if (opts.refine == HEGJ_REFINE_MFMA) opts.refine = HEGJ_REFINE_AUTO;  // the grid kernels
    }
    if (opts.refine == HEGJ_REFINE_MFMA)
        return fail(HEGJ_ERR_UNSUPPORTED, "the hybrid modes run the grid join kernels, not the MFMA refine");
    if (opts.refine > HEGJ_REFINE_F32) return fail(HEGJ_ERR_INVALID_ARG, "unknown refine");
    const uint32_t qb = opts.queries_per_unit ? opts.queries_per_unit : default_qb(h->g.dim);
    if (qb != 4 && qb != 8 && qb != 16) return fail(HEGJ_ERR_INVALID_ARG, "queries_per_unit must be 4, 8 or 16");
    if (!valid_waves_batches(opts))
        return fail(HEGJ_ERR_INVALID_ARG, "block_waves must be 0, 1, 2 or 4 and batches at most 32");
    std::unique_ptr<hegj_result> r(new hegj_result());
    r->idx = h;
    HEGJ_TRY(ensure_workload(h));
    r->info.total_candidates = h->total_candidates;
    r->info.workload_ms = h->workload_ms;
    r->info.n_queries = h->g.n - h->n_dead;
    HEGJ_TRY(selfjoin_hybrid(h, opts, qb, *ho, r.get()));
    r->info.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    HEGJ_TRY(mark_done(r.get(), h->stream));
    *out = r.release();
    return HEGJ_OK;
}

hegj_status hegj_result_get_hybrid_info(const hegj_result *r, hegj_hybrid_info *info) {
    if (!r || !info) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    if (!r->hybrid) return fail(HEGJ_ERR_INVALID_ARG, "not a hybrid result");
    *info = r->hinfo;
    return HEGJ_OK;
}

double hegj_static_partition_model(uint32_t n, uint32_t dim, double epsilon) {
    double g[4], c[4];
    {
        std::lock_guard<std::mutex> lk(g_model_mu);
        std::memcpy(g, g_model_gpu, sizeof(g));
        std::memcpy(c, g_model_cpu, sizeof(c));
    }
    const double ld = std::log((double)dim), ln = std::log((double)n), le = std::log(epsilon);
    const double tg = std::exp(g[0] + g[1] * ld + g[2] * ln + g[3] * le);
    const double tc = std::exp(c[0] + c[1] * ld + c[2] * ln + c[3] * le);
    // int gpu_qps = DBSIZE / gpuTimeModel (main.cu:152-154)
    const double gq = std::trunc((double)n / tg), cq = std::trunc((double)n / tc);
    return gq + cq > 0 ? gq / (gq + cq) : 1.0;
}

void hegj_static_partition_set_model(const double *gpu, const double *cpu) {
    std::lock_guard<std::mutex> lk(g_model_mu);
    if (gpu) std::memcpy(g_model_gpu, gpu, sizeof(g_model_gpu));
    if (cpu) std::memcpy(g_model_cpu, cpu, sizeof(g_model_cpu));
}

hegj_status hegj_f32_thresholds(double epsilon, uint32_t dim, double radius, float *tacc,
                                float *trej) {
    if (!tacc || !trej || dim == 0 || !(epsilon > 0.0) || !(radius >= 0.0))
        return fail(HEGJ_ERR_INVALID_ARG, "bad argument");
    if (!f32_thresholds(sqrt_threshold(epsilon), dim, radius, tacc, trej))
        return fail(HEGJ_ERR_UNSUPPORTED, "FP32 band wider than 1% of eps^2");
    return HEGJ_OK;
}

hegj_status hegj_result_free(hegj_result *r) {
    if (!r) return HEGJ_OK;
    (void)hipSetDevice(r->idx->device);
    delete r;
    return HEGJ_OK;
}

hegj_status hegj_result_get_info(const hegj_result *r, hegj_result_info *info) {
    if (!r || !info) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    *info = r->info;
    return HEGJ_OK;
}

hegj_status hegj_result_copy_counts(const hegj_result *r, uint32_t *counts) {
    if (!r || !counts) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    HEGJ_TRY(set_device(r->idx->device));
    hipStream_t s = r->idx->stream;
    HIP_TRY(hipMemcpyAsync(counts, r->d_row_cnt, sizeof(uint32_t) * r->idx->g.n,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

hegj_status hegj_result_copy_csr(const hegj_result *r, uint64_t *offsets, uint32_t *nbrs,
                                 int sort_rows) {
    if (!r || !offsets) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    HEGJ_TRY(set_device(r->idx->device));
    hipStream_t s = r->idx->stream;
    const uint32_t n = r->idx->g.n;
    Scratch sc;
    uint64_t *d_off;
    HIP_TRY(sc.get(&d_off, (size_t)n + 1));
    HEGJ_TRY(csr_offsets(r->d_row_cnt, n, d_off, s));
    HIP_TRY(hipMemcpyAsync(offsets, d_off, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const uint64_t total = offsets[n];
    if (!nbrs || total == 0) return HEGJ_OK;
    uint32_t *d_nb;
    HIP_TRY(sc.get(&d_nb, total));
    HEGJ_TRY(gather_rows(r->batches, r->bases, r->d_row_off, r->d_row_cnt, d_off, n, d_nb, s));
    HIP_TRY(hipMemcpyAsync(nbrs, d_nb, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (sort_rows) {
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++)
            th.emplace_back([&, t]() {
                for (uint64_t q = t; q < n; q += nt) std::sort(nbrs + offsets[q], nbrs + offsets[q + 1]);
            });
        for (auto &x : th) x.join();
    }
    return HEGJ_OK;
}

}  // extern "C"
