// hegj_serve.cpp -- the serving loop: a device-resident index answering
// range-join requests with continuous batching and memory admission.
//
// No counterpart in the reference.  Design:
//   * the database is uploaded once and indexed at every requested grid level
//     (one hegj_index per cell side); a request runs on the smallest level
//     whose cell side covers its epsilon;
//   * `slots` worker threads, each owning one HIP stream and one device arena
//     (a bump allocator for every temporary of a request, so nothing on the
//     request path calls hipMalloc/hipFree, which synchronise the device);
//     a request that arrives while others run is taken by the next free slot
//     and its kernels overlap with theirs on the device;
//   * per request: count pass -> exact result size -> admission into the
//     result-buffer pool (first come, first served) -> fill pass writing CSR
//     rows in place.  Results stay in the pool until released.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "hegj_internal.h"

namespace hegj {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

constexpr uint64_t kAlign = 256;
uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over the result pool, with coalescing of free ranges.
struct Pool {
    std::map<uint64_t, uint64_t> free_;  // offset -> size
    uint64_t cap = 0, in_use = 0, peak = 0;
    void init(uint64_t c) {
        cap = c;
        free_.clear();
        if (c) free_[0] = c;
    }
    bool alloc(uint64_t bytes, uint64_t *off) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[*off + bytes] = rest;
            in_use += bytes;
            peak = std::max(peak, in_use);
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = align_up(bytes);
        in_use -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

enum ReqState { kPending = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
    uint64_t ticket = 0;
    std::vector<double> queries;
    uint32_t nq = 0;
    double eps = 0;
    uint32_t level = 0;
    int state = kPending;
    hegj_status status = HEGJ_OK;
    std::string error;
    bool holds_pool = false;
    uint64_t pool_off = 0, pool_bytes = 0;
    uint64_t pairs = 0, hash = 0, evals = 0, band = 0;
    bool f32 = false;
    Clock::time_point t_submit, t_start, t_counted, t_admit, t_done;
};

}  // namespace
}  // namespace hegj

using namespace hegj;

struct hegj_server {
    hegj_serve_opts opts{};
    int device = 0;
    std::vector<hegj_index *> levels;
    std::vector<QueryGeom> qgeo;
    struct LevelF32 {
        const float *soa32 = nullptr;
        const float *aos32 = nullptr;
        const double *centre = nullptr;
        double radius = 0;
    };
    std::vector<LevelF32> lf32;  // FP32 copies (dim 2 and 6)
    char *pool_base = nullptr;
    Pool pool;
    hipStream_t copy_stream = nullptr;

    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_mem;
    std::deque<Request *> pending;
    // counted requests waiting for pool memory, served first come, first
    // served (a request still counting does not hold up later ones)
    std::deque<Request *> mem_wait;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_ticket = 1;
    bool stopping = false;
    std::vector<std::thread> workers;
    std::vector<std::string> worker_errors;

    // statistics window
    std::vector<double> lat;
    uint64_t submitted = 0, completed = 0, failed = 0, pairs = 0;
    bool have_first = false;
    Clock::time_point first_submit, last_done;

    void worker(uint32_t slot);
    hegj_status run(Request *r, hipStream_t s, Arena &arena, void *ctr_mem);
};

namespace {

hegj_status admit(hegj_server *S, Request *r, uint64_t bytes) {
    std::unique_lock<std::mutex> lk(S->mu);
    if (bytes > S->pool.cap)
        return fail(HEGJ_ERR_CAPACITY, "request result (" + std::to_string(bytes) +
                                           " bytes) exceeds the result-buffer pool");
    uint64_t off = 0;
    S->mem_wait.push_back(r);
    S->cv_mem.wait(lk, [&] { return S->mem_wait.front() == r && S->pool.alloc(bytes, &off); });
    S->mem_wait.pop_front();
    r->holds_pool = true;
    r->pool_off = off;
    r->pool_bytes = bytes;
    S->cv_mem.notify_all();  // the next in line may fit too
    return HEGJ_OK;
}

}  // namespace

hegj_status hegj_server::run(Request *r, hipStream_t s, Arena &arena, void *ctr_mem) {
    arena.used = 0;
    const uint32_t nq = r->nq, dim = opts.dim;
    const QueryGeom &qg = qgeo[r->level];
    hegj_index *lv = levels[r->level];
    Scratch keep;
    double *q;
    uint32_t *counts;
    uint64_t *offsets;
    HIP_TRY(keep.get(&q, (size_t)nq * dim));
    HIP_TRY(keep.get(&counts, nq));
    HIP_TRY(keep.get(&offsets, (size_t)nq + 1));
    uint64_t *hitbits;
    HIP_TRY(keep.get(&hitbits, (size_t)nq * kQueryHitWords));
    HIP_TRY(hipMemcpyAsync(q, r->queries.data(), sizeof(double) * nq * dim, hipMemcpyHostToDevice, s));
    QueryPlan plan;
    HEGJ_TRY(query_plan(&plan, keep, qg, q, nq, s));
    struct Ctr {
        uint32_t head[2];
        unsigned long long hash, evals, band, rbits;
    };
    Ctr *ctr = static_cast<Ctr *>(ctr_mem);
    HIP_TRY(hipMemsetAsync(ctr, 0, sizeof(Ctr), s));
    // FP32 filter: queries centred like the level's copy; the band comes from
    // the larger of the level's and the queries' radius around the centre
    bool f32 = false;
    float tacc = 0, trej = 0;
    float *q32 = nullptr;
    if (!lf32.empty() && lf32[r->level].soa32) {
        const LevelF32 &L = lf32[r->level];
        HIP_TRY(keep.get(&q32, (size_t)nq * dim));
        HEGJ_TRY(query_to_f32(q, nq, dim, L.centre, q32, &ctr->rbits, s));
        unsigned long long rb = 0;
        HIP_TRY(hipMemcpyAsync(&rb, &ctr->rbits, sizeof(rb), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        double rq;
        std::memcpy(&rq, &rb, sizeof(rq));
        const double R = std::max(L.radius, rq) * (1.0 + 1e-12);
        f32 = f32_thresholds(sqrt_threshold(r->eps), dim, R, &tacc, &trej);
    }
    HIP_TRY(hipMemsetAsync(counts, 0, sizeof(uint32_t) * nq, s));
    QJoinArgs a{};
    a.qg = qg;
    a.ix = index_dev(lv);
    a.queries = q;
    a.keys = plan.keys;
    a.ids = plan.ids;
    a.starts = plan.starts;
    a.n_units = plan.n_units;
    a.queue_head = &ctr->head[0];
    a.thresh = sqrt_threshold(r->eps);
    a.counts = counts;
    a.hash = &ctr->hash;
    a.evals = &ctr->evals;
    a.band = &ctr->band;
    a.hitbits = hitbits;
    if (f32) {
        a.f32 = 1;
        a.soa32 = lf32[r->level].soa32;
        a.aos32 = lf32[r->level].aos32;
        a.q32 = q32;
        a.tacc32 = tacc;
        a.trej32 = trej;
    }
    HEGJ_TRY(query_run(a, false, device, s));
    HEGJ_TRY(query_offsets(counts, nq, offsets, s));
    uint64_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, offsets + nq, sizeof(total), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    r->t_counted = Clock::now();
    const uint64_t off_bytes = align_up(sizeof(uint64_t) * ((uint64_t)nq + 1));
    HEGJ_TRY(admit(this, r, off_bytes + sizeof(uint32_t) * std::max<uint64_t>(total, 1)));
    r->t_admit = Clock::now();
    uint64_t *res_off = reinterpret_cast<uint64_t *>(pool_base + r->pool_off);
    uint32_t *res_nb = reinterpret_cast<uint32_t *>(pool_base + r->pool_off + off_bytes);
    HIP_TRY(hipMemcpyAsync(res_off, offsets, sizeof(uint64_t) * ((size_t)nq + 1), hipMemcpyDeviceToDevice, s));
    a.queue_head = &ctr->head[1];
    a.offsets = res_off;
    a.out = res_nb;
    HEGJ_TRY(query_run(a, true, device, s));
    HEGJ_TRY(query_hash(res_off, res_nb, nq, total, &ctr->hash, s));
    Ctr c;
    HIP_TRY(hipMemcpyAsync(&c, ctr, sizeof(c), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    r->pairs = total;
    r->hash = c.hash;
    r->evals = c.evals;
    r->band = c.band;
    r->f32 = f32;
    return HEGJ_OK;
}

void hegj_server::worker(uint32_t slot) {
    (void)slot;
    hipStream_t s = nullptr;
    Arena arena;
    void *ctr = nullptr;
    std::string init_err;
    if (select_device(device) != HEGJ_OK || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        init_err = "slot: cannot create a stream";
    // one request at a time: its queries and every temporary of the pipeline
    arena.cap = (size_t)opts.max_queries * (opts.dim * 8 + 128 + kQueryHitWords * 8) + (32u << 20);
    if (init_err.empty() && hipMalloc(reinterpret_cast<void **>(&arena.base), arena.cap) != hipSuccess)
        init_err = "slot: cannot allocate its device arena";
    if (init_err.empty() && hipMalloc(&ctr, 256) != hipSuccess) init_err = "slot: cannot allocate counters";
    set_thread_arena(init_err.empty() ? &arena : nullptr);
    while (true) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stopping || !pending.empty(); });
            if (pending.empty()) break;  // stopping and drained
            r = pending.front();
            pending.pop_front();
            r->state = kRunning;
            r->t_start = Clock::now();
        }
        hegj_status st = init_err.empty() ? run(r, s, arena, ctr) : fail(HEGJ_ERR_HIP, init_err);
        std::string msg = st == HEGJ_OK ? std::string() : std::string(hegj_last_error());
        // a failed request's kernels may still be writing into its pool range:
        // drain the slot's stream before the range can go to another request
        if (st != HEGJ_OK && s) (void)hipStreamSynchronize(s);
        {
            std::lock_guard<std::mutex> lk(mu);
            r->t_done = Clock::now();
            if (st == HEGJ_OK) {
                r->state = kDone;
                completed++;
                pairs += r->pairs;
                lat.push_back(ms_between(r->t_submit, r->t_done));
                last_done = r->t_done;
            } else {
                r->state = kFailed;
                r->status = st;
                r->error = msg;
                failed++;
                if (r->holds_pool) {
                    pool.release(r->pool_off, r->pool_bytes);
                    r->holds_pool = false;
                    cv_mem.notify_all();
                }
            }
        }
        cv_done.notify_all();
    }
    set_thread_arena(nullptr);
    if (ctr) (void)hipFree(ctr);
    if (arena.base) (void)hipFree(arena.base);
    if (s) (void)hipStreamDestroy(s);
}

static hegj_status serve_teardown(hegj_server *S) {
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    S->workers.clear();
    (void)select_device(S->device);
    for (hegj_index *h : S->levels) hegj_index_free(h);
    S->levels.clear();
    if (S->pool_base) (void)hipFree(S->pool_base);
    S->pool_base = nullptr;
    if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
    S->copy_stream = nullptr;
    return HEGJ_OK;
}

extern "C" {

hegj_status hegj_serve_open(const double *database, uint32_t n, const hegj_serve_opts *o,
                            hegj_server **out) {
    if (!database || !o || !out) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    if (n == 0 || o->dim == 0) return fail(HEGJ_ERR_INVALID_ARG, "empty database");
    if (o->indexed_dims == 0 || o->indexed_dims > HEGJ_MAX_INDEXED_DIMS || o->indexed_dims > o->dim)
        return fail(HEGJ_ERR_INVALID_ARG, "indexed_dims must be in [1, min(dim, 6)]");
    if (o->n_levels == 0 || o->n_levels > HEGJ_SERVE_MAX_LEVELS)
        return fail(HEGJ_ERR_INVALID_ARG, "n_levels must be in [1, 32]");
    if (o->slots == 0 || o->slots > 16) return fail(HEGJ_ERR_INVALID_ARG, "slots must be in [1, 16]");
    if (o->max_queries == 0) return fail(HEGJ_ERR_INVALID_ARG, "max_queries must be >= 1");
    if (o->result_bytes < 4096) return fail(HEGJ_ERR_INVALID_ARG, "result_bytes must be >= 4096");
    for (uint32_t l = 0; l < o->n_levels; l++) {
        if (!(o->cell_sides[l] > 0.0) || o->cell_sides[l] > 1.0)
            return fail(HEGJ_ERR_INVALID_ARG, "cell sides must be in ]0, 1]");
        if (l && !(o->cell_sides[l] > o->cell_sides[l - 1]))
            return fail(HEGJ_ERR_INVALID_ARG, "cell sides must be strictly ascending");
    }
    HEGJ_TRY(select_device(o->device));
    std::unique_ptr<hegj_server> S(new hegj_server());
    S->opts = *o;
    S->device = o->device;
    auto bail = [&](hegj_status st) {
        std::string msg = hegj_last_error();
        serve_teardown(S.get());
        return fail(st, msg);
    };
    double *dev = nullptr;
    const size_t bytes = sizeof(double) * (size_t)n * o->dim;
    if (hipMalloc(&dev, bytes) != hipSuccess) return fail(HEGJ_ERR_OUT_OF_MEMORY, "database upload");
    if (hipMemcpy(dev, database, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dev);
        return fail(HEGJ_ERR_HIP, "database upload");
    }
    for (uint32_t l = 0; l < o->n_levels; l++) {
        hegj_index_opts io;
        hegj_index_opts_default(&io, o->dim);
        io.indexed_dims = o->indexed_dims;
        io.emulate_cell0 = 0;  // the intended stencil: queries have no reference counterpart
        io.device = o->device;
        hegj_index *h = nullptr;
        hegj_status st = hegj_index_build_device(dev, n, o->cell_sides[l], &io, &h);
        if (st != HEGJ_OK) {
            (void)hipFree(dev);
            return bail(st);
        }
        S->levels.push_back(h);
        if (o->dim == 2 || o->dim == 6) {
            hegj_server::LevelF32 L;
            st = index_f32(h, &L.soa32, &L.aos32, &L.radius, &L.centre);
            if (st != HEGJ_OK) {
                (void)hipFree(dev);
                return bail(st);
            }
            S->lf32.push_back(L);
        }
        QueryGeom qg;
        st = query_geom(index_geom(h), &qg);
        if (st != HEGJ_OK) {
            (void)hipFree(dev);
            return bail(st);
        }
        S->qgeo.push_back(qg);
    }
    (void)hipFree(dev);
    if (hipMalloc(reinterpret_cast<void **>(&S->pool_base), o->result_bytes) != hipSuccess)
        return bail(fail(HEGJ_ERR_OUT_OF_MEMORY, "result-buffer pool"));
    S->pool.init(o->result_bytes);
    if (hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(HEGJ_ERR_HIP, "copy stream"));
    for (uint32_t i = 0; i < o->slots; i++) S->workers.emplace_back(&hegj_server::worker, S.get(), i);
    *out = S.release();
    return HEGJ_OK;
}

hegj_status hegj_serve_submit(hegj_server *S, const double *queries, uint32_t nq, double eps,
                              uint64_t *ticket) {
    if (!S || !ticket || (nq && !queries)) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    if (nq > S->opts.max_queries) return fail(HEGJ_ERR_INVALID_ARG, "more queries than max_queries");
    if (!(eps > 0.0)) return fail(HEGJ_ERR_INVALID_ARG, "epsilon must be > 0");
    // smallest level whose cell side covers eps with rounding margin
    uint32_t level = S->opts.n_levels;
    for (uint32_t l = 0; l < S->opts.n_levels; l++)
        if (eps <= S->opts.cell_sides[l] * (1.0 - 0x1p-30)) {
            level = l;
            break;
        }
    if (level == S->opts.n_levels)
        return fail(HEGJ_ERR_INVALID_ARG, "epsilon larger than the largest cell side");
    std::unique_ptr<Request> r(new Request());
    r->queries.assign(queries, queries + (size_t)nq * S->opts.dim);
    r->nq = nq;
    r->eps = eps;
    r->level = level;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stopping) return fail(HEGJ_ERR_BUSY, "server is closing");
        r->ticket = S->next_ticket++;
        *ticket = r->ticket;
        S->submitted++;
        if (!S->have_first) {
            S->have_first = true;
            S->first_submit = r->t_submit;
        }
        Request *raw = r.get();
        S->reqs[raw->ticket] = std::move(r);
        if (nq == 0) {  // nothing to do on the device
            raw->state = kDone;
            raw->t_start = raw->t_counted = raw->t_admit = raw->t_done = raw->t_submit;
            S->completed++;
            S->lat.push_back(0.0);
        } else {
            S->pending.push_back(raw);
        }
    }
    S->cv_work.notify_one();
    S->cv_done.notify_all();
    return HEGJ_OK;
}

static hegj_status finished_info(const Request *r, hegj_request_info *info) {
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->ticket = r->ticket;
        info->status = r->state == kDone ? HEGJ_OK : r->status;
        info->level = r->level;
        info->n_queries = r->nq;
        info->pairs = r->pairs;
        info->pair_hash = r->hash;
        info->evals = r->evals;
        info->band_rechecks = r->band;
        info->refine_used = r->f32 ? HEGJ_REFINE_F32 : HEGJ_REFINE_FP64;
        info->result_bytes = r->holds_pool ? r->pool_bytes : 0;
        info->queue_ms = ms_between(r->t_submit, r->t_start);
        if (r->state == kDone) {
            info->count_ms = ms_between(r->t_start, r->t_counted);
            info->admit_ms = ms_between(r->t_counted, r->t_admit);
            info->fill_ms = ms_between(r->t_admit, r->t_done);
        }
        info->latency_ms = ms_between(r->t_submit, r->t_done);
    }
    if (r->state == kFailed) return fail(r->status, r->error);
    return HEGJ_OK;
}

hegj_status hegj_serve_wait(hegj_server *S, uint64_t ticket, hegj_request_info *info) {
    if (!S) return fail(HEGJ_ERR_INVALID_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return fail(HEGJ_ERR_INVALID_ARG, "unknown ticket");
    Request *r = it->second.get();
    S->cv_done.wait(lk, [&] { return r->state == kDone || r->state == kFailed; });
    return finished_info(r, info);
}

hegj_status hegj_serve_poll(hegj_server *S, uint64_t ticket, hegj_request_info *info) {
    if (!S) return fail(HEGJ_ERR_INVALID_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return fail(HEGJ_ERR_INVALID_ARG, "unknown ticket");
    const Request *r = it->second.get();
    if (r->state != kDone && r->state != kFailed) return HEGJ_ERR_BUSY;
    return finished_info(r, info);
}

hegj_status hegj_serve_copy(hegj_server *S, uint64_t ticket, uint64_t *offsets, uint32_t *neighbors) {
    if (!S || !offsets) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    Request *r;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(ticket);
        if (it == S->reqs.end()) return fail(HEGJ_ERR_INVALID_ARG, "unknown ticket");
        r = it->second.get();
        if (r->state != kDone) return fail(HEGJ_ERR_BUSY, "request has not finished");
    }
    if (r->nq == 0) {
        offsets[0] = 0;
        return HEGJ_OK;
    }
    HEGJ_TRY(select_device(S->device));
    const uint64_t off_bytes = align_up(sizeof(uint64_t) * ((uint64_t)r->nq + 1));
    HIP_TRY(hipMemcpyAsync(offsets, S->pool_base + r->pool_off, sizeof(uint64_t) * ((size_t)r->nq + 1),
                           hipMemcpyDeviceToHost, S->copy_stream));
    if (neighbors && r->pairs)
        HIP_TRY(hipMemcpyAsync(neighbors, S->pool_base + r->pool_off + off_bytes,
                               sizeof(uint32_t) * r->pairs, hipMemcpyDeviceToHost, S->copy_stream));
    HIP_TRY(hipStreamSynchronize(S->copy_stream));
    return HEGJ_OK;
}

hegj_status hegj_serve_release(hegj_server *S, uint64_t ticket) {
    if (!S) return fail(HEGJ_ERR_INVALID_ARG, "NULL server");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(ticket);
        if (it == S->reqs.end()) return fail(HEGJ_ERR_INVALID_ARG, "unknown ticket");
        Request *r = it->second.get();
        if (r->state != kDone && r->state != kFailed) return fail(HEGJ_ERR_BUSY, "request has not finished");
        if (r->holds_pool) S->pool.release(r->pool_off, r->pool_bytes);
        S->reqs.erase(it);
    }
    S->cv_mem.notify_all();
    return HEGJ_OK;
}

hegj_status hegj_serve_get_stats(hegj_server *S, hegj_serve_stats *st, int reset) {
    if (!S || !st) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = S->submitted;
    st->completed = S->completed;
    st->failed = S->failed;
    st->pairs = S->pairs;
    std::vector<double> v = S->lat;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            // nearest rank
            size_t i = (size_t)std::ceil(p * v.size());
            return v[std::min(v.size() - 1, i ? i - 1 : 0)];
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        st->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        st->mean_ms = sum / v.size();
    }
    st->pool_bytes = S->pool.cap;
    st->pool_in_use = S->pool.in_use;
    st->pool_peak = S->pool.peak;
    if (S->have_first && S->completed) st->wall_ms = ms_between(S->first_submit, S->last_done);
    if (reset) {
        S->lat.clear();
        S->submitted = S->completed = S->failed = S->pairs = 0;
        S->have_first = false;
        S->pool.peak = S->pool.in_use;
    }
    return HEGJ_OK;
}

hegj_status hegj_serve_close(hegj_server *S) {
    if (!S) return HEGJ_OK;
    {
        // queued work finishes first; requests waiting for pool memory that
        // the caller never frees would block, so results still held are freed
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto &kv : S->reqs)
            if (kv.second->holds_pool && kv.second->state == kDone) {
                S->pool.release(kv.second->pool_off, kv.second->pool_bytes);
                kv.second->holds_pool = false;
            }
    }
    S->cv_mem.notify_all();
    serve_teardown(S);
    delete S;
    return HEGJ_OK;
}

}  // extern "C"
