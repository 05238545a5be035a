// hegj_ego.cpp -- the CPU side of the hybrid modes: Super-EGO range queries.
//
// Algorithm 3 / the CPU share of algorithms 1 and 2 run the Super-EGO join
// (Kalashnikov, "Super-EGO: fast multi-dimensional similarity join", VLDB J.
// 2013) one query at a time: Util::reorderDim (Util.cpp:33-281), the EGO sort
// (main.cu:30-39, 296-318) and Util::egoJoinV2 / simpleJoin3 / simpleJoin4
// with B = {query} (Util.cpp:302-590, MultiThreadJoin.cpp:43-115).  The pair
// set each query gets is the reference's; the implementation is this
// library's own:
//   - the EGO key of every point (its (int)(x_j / eps) in each reordered
//     dimension) is computed once into an int array: sorting, pruning and
//     recursion compare integers, never divide again;
//   - the sort is a parallel merge sort of point indices by key (each thread
//     sorts a slice, slices are merged pairwise);
//   - a query walks the recursion with an explicit stack of (first, last,
//     start dimension) ranges;
//   - the leaf scan order of each start dimension m (the three intervals
//     r1, r2, r3 of Util.cpp:193-267) is precomputed as one list of
//     dimensions.
// The early exit of the reference's scans (`sum > eps2`) does not change a
// decision: each step adds a non-negative term and rounding is monotone, so
// a partial sum above eps2 means the full sum is too.  The sum is
// fma(dx, dx, sum) (the reference's Util.cpp is compiled for FMA hosts by
// g++ -O3 -march=znver2, which contracts dx * dx; sum += dx), in the scan
// order, against eps2 = eps * eps.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "hegj_internal.h"

namespace hegj {

struct EgoIndex {
    uint32_t n = 0, dim = 0;
    double eps = 0, eps2 = 0;
    std::vector<double> x;       // [n * dim] reordered dims, EGO order
    std::vector<int32_t> key;    // [n * dim] (int)(x / eps), EGO order
    std::vector<uint32_t> id;    // original point id of each EGO position
    std::vector<uint32_t> where; // original point id -> EGO position
    // scan order of the leaf distance for start dimension m: dims[off[m] ..
    // off[m + 1])
    std::vector<uint32_t> scan_dims, scan_off;
    uint32_t map[1024];          // reordered dim j = original dim map[j]
};

namespace {

constexpr int kMinLen = 32;        // MINLEN (params.h:34)
constexpr int kSample = 2000;      // Util.cpp:57

// Util::rnd (Util.cpp:597-608): the Park-Miller "minimal standard" generator
// s' = 16807 s mod (2^31 - 1) from s = 1234567, returned as s' / 2147483711.
// (16807 s < 2^46 is exact in the reference's long double, so its
// floor/multiply/subtract is this integer remainder.)
struct MinStd {
    uint64_t s = 1234567;
    double next() {
        s = (16807u * s) % 2147483647u;
        return std::fabs((double)s / 2147483711.0);
    }
};

// The dimension permutation and the per-start-dimension scan order
// (Util.cpp:57-281).  Only the mean |a_j - b_j| over 2000 sampled pairs
// orders the dimensions (descending, ties to the lower dimension); a
// dimension whose mean is below eps / 2 is "inactive".
void reorder(EgoIndex &E, const double *pts) {
    const uint32_t n = E.n, D = E.dim;
    MinStd rng;
    std::vector<double> mean(D, 0.0);
    for (int s = 0; s < kSample; s++) {
        const uint32_t ia = (uint32_t)(int)((double)(n - 1) * rng.next());
        const uint32_t ib = (uint32_t)(int)((double)(n - 1) * rng.next());
        for (uint32_t j = 0; j < D; j++)
            mean[j] += std::fabs(pts[(size_t)ia * D + j] - pts[(size_t)ib * D + j]);
    }
    for (uint32_t j = 0; j < D; j++) mean[j] /= kSample;
    // stable sort by descending mean = the reference's repeated arg-min of -mean
    std::vector<uint32_t> order(D);
    for (uint32_t j = 0; j < D; j++) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return mean[a] > mean[b]; });
    for (uint32_t j = 0; j < D; j++) E.map[j] = order[j];
    // first inactive dimension in the new order (D when none)
    const double half = E.eps / 2;
    uint32_t first_inactive = D;
    for (uint32_t j = 0; j < D; j++)
        if (mean[E.map[j]] < half) {
            first_inactive = j;
            break;
        }
    // scan order for each start dimension m in [0, D]
    E.scan_off.assign(D + 2, 0);
    E.scan_dims.clear();
    auto push = [&](uint32_t a, uint32_t b) {  // dims a..b inclusive
        for (uint32_t l = a; l <= b && l < D; l++) E.scan_dims.push_back(l);
    };
    for (uint32_t m = 0; m <= D; m++) {
        E.scan_off[m] = (uint32_t)E.scan_dims.size();
        if (m == 0 || m == D || mean[E.map[m]] < half) {
            push(0, D - 1);  // one interval
        } else if (first_inactive < D) {
            // three intervals: the active dims from m, then those before m,
            // then the inactive tail
            if (first_inactive >= 1) push(m, first_inactive - 1);
            push(0, m - 1);
            push(first_inactive, D - 1);
        } else {
            push(m, D - 1);  // two intervals
            push(0, m - 1);
        }
    }
    E.scan_off[D + 1] = (uint32_t)E.scan_dims.size();
}

bool key_less(const int32_t *a, const int32_t *b, uint32_t D) {
    for (uint32_t j = 0; j < D; j++) {
        if (a[j] != b[j]) return a[j] < b[j];
    }
    return false;
}

// EGO order: points sorted by their key, lexicographically over the
// reordered dimensions (egoSortFunction, main.cu:30-39).  Ties keep the input
// order (the reference's sample_sort leaves it open; no pair depends on it:
// equal keys give equal pruning at every position).
void ego_sort(EgoIndex &E, const std::vector<int32_t> &key0, unsigned threads) {
    const uint32_t n = E.n, D = E.dim;
    std::vector<uint32_t> perm(n), tmp(n);
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    auto less = [&](uint32_t a, uint32_t b) { return key_less(&key0[(size_t)a * D], &key0[(size_t)b * D], D); };
    unsigned parts = 1;
    while (parts * 2 <= threads && n / (parts * 2) >= 4096) parts *= 2;
    std::vector<uint32_t> cut(parts + 1);
    for (unsigned p = 0; p <= parts; p++) cut[p] = (uint32_t)((uint64_t)n * p / parts);
    {
        std::vector<std::thread> th;
        for (unsigned p = 0; p < parts; p++)
            th.emplace_back([&, p]() { std::stable_sort(perm.begin() + cut[p], perm.begin() + cut[p + 1], less); });
        for (auto &t : th) t.join();
    }
    for (unsigned w = 1; w < parts; w *= 2) {
        std::vector<std::thread> th;
        for (unsigned p = 0; p + w < parts; p += 2 * w) {
            const uint32_t a = cut[p], m = cut[p + w], b = cut[std::min(p + 2 * w, parts)];
            th.emplace_back([&, a, m, b]() {
                std::merge(perm.begin() + a, perm.begin() + m, perm.begin() + m, perm.begin() + b,
                           tmp.begin() + a, less);
                std::copy(tmp.begin() + a, tmp.begin() + b, perm.begin() + a);
            });
        }
        for (auto &t : th) t.join();
    }
    E.id = perm;
    E.where.assign(n, 0);
    for (uint32_t i = 0; i < n; i++) E.where[perm[i]] = i;
}

}  // namespace

hegj_status ego_build(const double *pts, uint32_t n, uint32_t dim, double eps, unsigned threads,
                      EgoIndex **out, double *reorder_s, double *sort_s) {
    if (dim == 0 || dim > 1024 || !(eps > 0)) return fail(HEGJ_ERR_INVALID_ARG, "Super-EGO: bad arguments");
    EgoIndex *E = new EgoIndex();
    E->n = n;
    E->dim = dim;
    E->eps = eps;
    E->eps2 = eps * eps;
    threads = std::max(1u, threads);
    auto t0 = std::chrono::steady_clock::now();
    if (n > 0) reorder(*E, pts);
    // reordered coordinates and their keys, in input order
    std::vector<double> xr((size_t)n * dim);
    std::vector<int32_t> k0((size_t)n * dim);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < threads; t++)
            th.emplace_back([&, t]() {
                for (uint64_t i = t; i < n; i += threads)
                    for (uint32_t j = 0; j < dim; j++) {
                        const double v = pts[i * dim + E->map[j]];
                        xr[i * dim + j] = v;
                        k0[i * dim + j] = (int32_t)(v / eps);
                    }
            });
        for (auto &x : th) x.join();
    }
    auto t1 = std::chrono::steady_clock::now();
    ego_sort(*E, k0, threads);
    E->x.resize((size_t)n * dim);
    E->key.resize((size_t)n * dim);
    for (uint32_t i = 0; i < n; i++) {
        std::memcpy(&E->x[(size_t)i * dim], &xr[(size_t)E->id[i] * dim], sizeof(double) * dim);
        std::memcpy(&E->key[(size_t)i * dim], &k0[(size_t)E->id[i] * dim], sizeof(int32_t) * dim);
    }
    auto t2 = std::chrono::steady_clock::now();
    if (reorder_s) *reorder_s = std::chrono::duration<double>(t1 - t0).count();
    if (sort_s) *sort_s = std::chrono::duration<double>(t2 - t1).count();
    *out = E;
    return HEGJ_OK;
}

void ego_free(EgoIndex *E) { delete E; }

// Neighbours of one query (original id q), appended to `row` as original
// ids in EGO order; returns the candidates scanned.  The walk is egoJoinV2
// with B = {q}: a range [first, last] of the EGO order is cut off when, in
// one of the dimensions from its start dimension on where both the range
// and the query sit in single cells, the cells are more than one apart; the
// first dimension where the range spans several cells becomes the start
// dimension of its halves; ranges below MINLEN points are scanned.
uint64_t ego_query(const EgoIndex &E, uint32_t q, std::vector<uint32_t> &row) {
    const uint32_t D = E.dim;
    const uint32_t qp = E.where[q];
    const double *xq = &E.x[(size_t)qp * D];
    const int32_t *kq = &E.key[(size_t)qp * D];
    struct Range {
        uint32_t first, last, start;
    };
    Range stack[96];
    int top = 0;
    stack[top++] = Range{0, E.n - 1, 0};
    uint64_t scanned = 0;
    while (top > 0) {
        const Range r = stack[--top];
        const int32_t *kf = &E.key[(size_t)r.first * D];
        const int32_t *kl = &E.key[(size_t)r.last * D];
        uint32_t i = r.start, start = r.start;
        bool cut = false;
        for (; i < D; i++) {
            if (kf[i] > kq[i] + 1 || kq[i] > kl[i] + 1) {
                cut = true;
                break;
            }
            if (kf[i] < kl[i]) {
                start = i;
                break;
            }
        }
        if (cut) continue;
        const uint32_t size = r.last - r.first + 1;
        if (size < (uint32_t)kMinLen) {
            // leaf: the scan order of dimension i (D when the range sits in
            // one cell in every dimension from its start on)
            const uint32_t *dl = &E.scan_dims[E.scan_off[i]];
            const uint32_t nd = E.scan_off[i + 1] - E.scan_off[i];
            scanned += size;
            for (uint32_t p = r.first; p <= r.last; p++) {
                const double *xp = &E.x[(size_t)p * D];
                double sum = 0.0;
                uint32_t l = 0;
                for (; l < nd; l++) {
                    const double dx = xp[dl[l]] - xq[dl[l]];
                    sum = std::fma(dx, dx, sum);
                    if (sum > E.eps2) break;
                }
                if (l == nd) row.push_back(E.id[p]);
            }
            continue;
        }
        // halves [first, first + size/2] and the rest; the first is walked
        // first (pushed last)
        const uint32_t mid = r.first + size / 2;
        stack[top++] = Range{mid + 1, r.last, start};
        stack[top++] = Range{r.first, mid, start};
    }
    return scanned;
}

}  // namespace hegj

extern "C" {

hegj_status hegj_superego_rows(const double *database, uint32_t n, uint32_t dim, double epsilon,
                               uint32_t threads, const uint32_t *qids, uint32_t nq,
                               uint32_t *counts, uint32_t **nbrs, uint64_t *total) {
    using namespace hegj;
    if (!database || !counts || !nbrs || !total || n == 0) return fail(HEGJ_ERR_INVALID_ARG, "bad argument");
    if (!qids && nq != n) return fail(HEGJ_ERR_INVALID_ARG, "without qids, nq must be n");
    for (uint32_t i = 0; qids && i < nq; i++)
        if (qids[i] >= n) return fail(HEGJ_ERR_INVALID_ARG, "query id out of range");
    const unsigned nt = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    EgoIndex *E = nullptr;
    HEGJ_TRY(ego_build(database, n, dim, epsilon, nt, &E, nullptr, nullptr));
    // each thread answers a strided share; rows are then laid out in query order
    std::vector<std::vector<uint32_t>> part(nt);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            std::vector<uint32_t> row;
            for (uint64_t i = t; i < nq; i += nt) {
                row.clear();
                ego_query(*E, qids ? qids[i] : (uint32_t)i, row);
                counts[i] = (uint32_t)row.size();
                part[t].insert(part[t].end(), row.begin(), row.end());
            }
        });
    for (auto &x : th) x.join();
    ego_free(E);
    uint64_t sum = 0;
    for (uint32_t i = 0; i < nq; i++) sum += counts[i];
    uint32_t *out = (uint32_t *)std::malloc(sizeof(uint32_t) * (sum + 1));
    if (!out) return fail(HEGJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::vector<size_t> cursor(nt, 0);
    uint64_t w = 0;
    for (uint64_t i = 0; i < nq; i++) {
        const unsigned t = (unsigned)(i % nt);
        std::memcpy(out + w, part[t].data() + cursor[t], sizeof(uint32_t) * counts[i]);
        cursor[t] += counts[i];
        w += counts[i];
    }
    *nbrs = out;
    *total = sum;
    return HEGJ_OK;
}

void hegj_free(void *p) { std::free(p); }

}  // extern "C"
