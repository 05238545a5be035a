// hegj_oracle.cpp -- CPU oracle (test infrastructure only; see hegj_oracle.h).
//
// Plain restatement of the reference's algorithm-0 and algorithm-3 semantics.
// Each function names the reference file:line it follows.  Compiled by
// oracle/Makefile into oracle/liboracle.so with g++ -O2 -mfma -fopenmp.
#include "hegj_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Grid helpers templated on DTYPE (params.h:24 DTYPE is double or float).
// ---------------------------------------------------------------------------

// getLinearID_nDimensions / getLinearID_nDimensionsGPU: main.cu:740-754,
// kernel.cu:310-325 -- dim 0 fastest, uint64 accumulation.
uint64_t linear_id(const uint32_t *idx, const uint32_t *dimLen, uint32_t k) {
    uint64_t offset = 0, mult = 1;
    for (uint32_t i = 0; i < k; i++) {
        offset += (uint64_t)idx[i] * mult;
        mult *= dimLen[i];
    }
    return offset;
}

template <typename T>
struct Grid {
    uint32_t n, dim, k;
    T eps;
    std::vector<T> pts;             // n*dim, AoS, in DTYPE
    std::vector<T> minArr, maxArr;  // k
    std::vector<uint32_t> nCells;   // k
    uint64_t totalCells = 0;
    std::vector<uint64_t> ids;      // per point linear id
    std::vector<uint32_t> indexLookupArr;
    std::vector<uint64_t> cellKeys;  // gridCellLookupArr[i].gridLinearID (idx == i)
    std::vector<og_grid> grid;

    Grid(const double *p, uint32_t n_, uint32_t dim_, uint32_t k_, double eps_)
        : n(n_), dim(dim_), k(k_), eps((T)eps_), pts((size_t)n_ * dim_) {
        for (size_t i = 0; i < pts.size(); i++) pts[i] = (T)p[i];
        dims();
        cells();
        build();
    }

    // generateNDGridDimensions, main.cu:463-518
    void dims() {
        minArr.assign(k, 0);
        maxArr.assign(k, 0);
        nCells.assign(k, 0);
        for (uint32_t j = 0; j < k; j++) {
            minArr[j] = pts[j];
            maxArr[j] = pts[j];
        }
        for (uint32_t i = 1; i < n; i++)
            for (uint32_t j = 0; j < k; j++) {
                T v = pts[(size_t)i * dim + j];
                if (v < minArr[j]) minArr[j] = v;
                if (v > maxArr[j]) maxArr[j] = v;
            }
        for (uint32_t j = 0; j < k; j++) {
            minArr[j] -= eps;
            maxArr[j] += eps;
        }
        for (uint32_t j = 0; j < k; j++) nCells[j] = (uint32_t)std::ceil((maxArr[j] - minArr[j]) / eps);
        uint64_t t = nCells[0];
        for (uint32_t j = 1; j < k; j++) t *= nCells[j];
        totalCells = t;
    }

    // S2: kernel.cu:54-57 -- DTYPE subtract, DTYPE divide, truncating conversion.
    void coords(const T *x, uint32_t *out) const {
        for (uint32_t j = 0; j < k; j++) out[j] = (uint32_t)((x[j] - minArr[j]) / eps);
    }

    void cells() {
        ids.resize(n);
        std::vector<uint32_t> c(k);
        for (uint32_t i = 0; i < n; i++) {
            coords(&pts[(size_t)i * dim], c.data());
            ids[i] = linear_id(c.data(), nCells.data(), k);
        }
    }

    // S11: GPU.cu:249-324 -- sort+unique of ids, stable sort_by_key of
    // (id, point), host run-boundary scan.
    void build() {
        indexLookupArr.resize(n);
        for (uint32_t i = 0; i < n; i++) indexLookupArr[i] = i;
        std::stable_sort(indexLookupArr.begin(), indexLookupArr.end(),
                         [&](uint32_t a, uint32_t b) { return ids[a] < ids[b]; });
        cellKeys.clear();
        grid.clear();
        for (uint32_t i = 0; i < n; i++) {
            uint64_t key = ids[indexLookupArr[i]];
            if (i == 0 || key != ids[indexLookupArr[i - 1]]) {
                if (!grid.empty()) grid.back().indexmax = (int32_t)i - 1;
                cellKeys.push_back(key);
                grid.push_back(og_grid{(int32_t)i, 0});
            }
        }
        if (!grid.empty()) grid.back().indexmax = (int32_t)n - 1;
    }

    // thrust::binary_search + lower_bound over gridCellLookupArr,
    // kernel.cu:402-406.  Returns the cell index or -1.
    int64_t find(uint64_t key) const {
        auto it = std::lower_bound(cellKeys.begin(), cellKeys.end(), key);
        if (it == cellKeys.end() || *it != key) return -1;
        return it - cellKeys.begin();
    }

    // Stencil bounds, kernel.cu:777-782 (S4).  With emulate_cell0 the lower
    // bound is CUDA's max(int 0, unsigned id-1), i.e. an unsigned max, which
    // is UINT_MAX for id == 0 (S5); otherwise the intended signed clamp.
    void stencil(const uint32_t *c, int emulate_cell0, uint32_t *lo, uint32_t *hi) const {
        for (uint32_t i = 0; i < k; i++) {
            uint32_t m1 = c[i] - 1u;  // unsigned wrap at 0, as in the reference
            if (emulate_cell0)
                lo[i] = std::max(0u, m1);
            else
                lo[i] = c[i] == 0 ? 0u : m1;
            hi[i] = std::min(nCells[i] - 1u, c[i] + 1u);
        }
    }

    // Enumerate the nested stencil loops of kernel.cu:787-789 + kernelloops.h
    // (dim 0 outermost).  fn(cellIndex) for each non-empty stencil cell.
    template <typename F>
    void for_stencil(const uint32_t *c, int emulate_cell0, F fn) const {
        std::vector<uint32_t> lo(k), hi(k), cur(k);
        stencil(c, emulate_cell0, lo.data(), hi.data());
        for (uint32_t i = 0; i < k; i++)
            if (lo[i] > hi[i]) return;  // a loop with an empty range runs nothing
        for (uint32_t i = 0; i < k; i++) cur[i] = lo[i];
        while (true) {
            int64_t ci = find(linear_id(cur.data(), nCells.data(), k));
            if (ci >= 0) fn((uint32_t)ci);
            // odometer with the innermost (last) dimension fastest
            int32_t d = (int32_t)k - 1;
            while (d >= 0) {
                if (cur[d] < hi[d]) {
                    cur[d]++;
                    break;
                }
                cur[d] = lo[d];
                d--;
            }
            if (d < 0) break;
        }
    }

    // evalPoint distance, kernel.cu:347-355: runningTotalDist += d*d with
    // nvcc's default FMA contraction, natural dimension order, no early exit;
    // predicate sqrt(acc) <= eps.
    bool within(const T *c, const T *q) const {
        T acc = 0;
        for (uint32_t l = 0; l < dim; l++) {
            T d = c[l] - q[l];
            acc = std::fma(d, d, acc);
        }
        return std::sqrt(acc) <= eps;
    }

    // kernelNDGridIndexGlobal for one query (kernel.cu:756-800).
    template <typename F>
    void join_query(uint32_t q, int emulate_cell0, F emit) const {
        const T *point = &pts[(size_t)q * dim];
        std::vector<uint32_t> c(k);
        coords(point, c.data());
        for_stencil(c.data(), emulate_cell0, [&](uint32_t ci) {
            for (int32_t m = grid[ci].indexmin; m <= grid[ci].indexmax; m++) {
                uint32_t dataIdx = indexLookupArr[m];
                if (within(&pts[(size_t)dataIdx * dim], point)) emit(dataIdx);
            }
        });
    }
};

template <typename T>
uint64_t join_counts_t(const double *p, uint32_t n, uint32_t dim, uint32_t k, double eps,
                       int emulate_cell0, uint32_t *counts) {
    Grid<T> g(p, n, dim, k, eps);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int64_t q = 0; q < (int64_t)n; q++) {
        uint32_t cnt = 0;
        g.join_query((uint32_t)q, emulate_cell0, [&](uint32_t) { cnt++; });
        counts[q] = cnt;
        total += cnt;
    }
    return total;
}

template <typename T>
void join_fill_t(const double *p, uint32_t n, uint32_t dim, uint32_t k, double eps,
                 int emulate_cell0, const uint64_t *offsets, uint32_t *nbrs) {
    Grid<T> g(p, n, dim, k, eps);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < (int64_t)n; q++) {
        uint32_t *row = nbrs + offsets[q];
        uint64_t cnt = 0;
        g.join_query((uint32_t)q, emulate_cell0, [&](uint32_t c) { row[cnt++] = c; });
        std::sort(row, row + cnt);
    }
}

template <typename T>
uint64_t join_rows_t(const double *p, uint32_t n, uint32_t dim, uint32_t k, double eps,
                     int emulate_cell0, const uint32_t *qids, uint32_t nq, uint32_t *counts,
                     const uint64_t *offsets, uint32_t *nbrs) {
    Grid<T> g(p, n, dim, k, eps);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int64_t i = 0; i < (int64_t)nq; i++) {
        std::vector<uint32_t> row;
        g.join_query(qids[i], emulate_cell0, [&](uint32_t c) { row.push_back(c); });
        std::sort(row.begin(), row.end());
        if (counts) counts[i] = (uint32_t)row.size();
        if (nbrs) std::copy(row.begin(), row.end(), nbrs + offsets[i]);
        total += row.size();
    }
    return total;
}

uint64_t mix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float threshold_of(float eps) { return og_threshold_f(eps); }
double threshold_of(double eps) { return og_threshold(eps); }

// Count and multiset hash of the algorithm-0 pair set without holding it
// (the pair sets of the full-size workloads have up to 1.2e9 pairs).  Same
// semantics as join_query, organised for speed:
//  - the stencil depends only on the query's cell (kernel.cu:763-782 derive it
//    from the cell coordinates), so it is enumerated once per non-empty cell;
//  - candidates are read from a cell-ordered copy of the points;
//  - sqrt(acc) <= eps is evaluated as acc <= T, T = the largest value whose
//    sqrt is <= eps (og_threshold; sqrt is monotone);
//  - the fma chain stops as soon as acc > T: fma(d, d, acc) >= acc for every d
//    (the exact d*d + acc is >= acc and rounding is monotone), so once above T
//    the final acc stays above T and the decision is the same.
template <typename T>
uint64_t join_stats_t(const double *p, uint32_t n, uint32_t dim, uint32_t k, double eps,
                      int emulate_cell0, uint64_t *hash, uint32_t *row_cnt, uint64_t *row_hash) {
    Grid<T> g(p, n, dim, k, eps);
    const T thr = threshold_of(g.eps);
    std::vector<T> cp((size_t)n * dim);  // points in cell order
    for (uint32_t i = 0; i < n; i++)
        std::memcpy(&cp[(size_t)i * dim], &g.pts[(size_t)g.indexLookupArr[i] * dim], sizeof(T) * dim);
    const uint32_t C = (uint32_t)g.cellKeys.size();
    uint64_t total = 0, h_all = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total, h_all)
    for (int64_t ci = 0; ci < (int64_t)C; ci++) {
        std::vector<uint32_t> c(k), st;
        const uint32_t first = (uint32_t)g.grid[ci].indexmin, last = (uint32_t)g.grid[ci].indexmax;
        g.coords(&cp[(size_t)first * dim], c.data());
        g.for_stencil(c.data(), emulate_cell0, [&](uint32_t s) { st.push_back(s); });
        for (uint32_t qp = first; qp <= last; qp++) {
            const T *xq = &cp[(size_t)qp * dim];
            const uint32_t qid = g.indexLookupArr[qp];
            uint32_t cnt = 0;
            uint64_t h = 0;
            for (uint32_t s : st) {
                for (int32_t m = g.grid[s].indexmin; m <= g.grid[s].indexmax; m++) {
                    const T *xc = &cp[(size_t)m * dim];
                    T acc = 0;
                    uint32_t l = 0;
                    for (; l < dim; l++) {
                        const T d = xc[l] - xq[l];
                        acc = std::fma(d, d, acc);
                        if (acc > thr) break;
                    }
                    if (l == dim) {
                        cnt++;
                        h += mix64(((uint64_t)qid << 32) | g.indexLookupArr[m]);
                    }
                }
            }
            if (row_cnt) row_cnt[qid] = cnt;
            if (row_hash) row_hash[qid] = h;
            total += cnt;
            h_all += h;
        }
    }
    // queries silenced by the cell-0 wrap are in cells whose stencil is empty
    // and got a zero row above
    *hash = h_all;
    return total;
}

// ---------------------------------------------------------------------------
// Super-EGO restatement (algorithm 3).  Util.cpp is (c) 2013 Kalashnikov;
// this is an independent restatement of its published algorithm.
// ---------------------------------------------------------------------------
struct SuperEgo {
    uint32_t n, dim;
    double eps, eps2;
    std::vector<int> id;        // Point::id
    std::vector<double> x;      // n*dim, reordered dims, EGO-sorted order
    std::vector<int> r1, r2, r3;  // (dim+1) x 2 scan ranges, Util.cpp:26-28
    std::vector<uint32_t> egoMapping;

    // Util::rnd, Util.cpp:597-608 (Park-Miller style LCG, seed 1234567.0).
    double seed = 1234567.0;
    double rnd() {
        const double d2p31m = 2147483647.0;
        const double d2p31 = 2147483711.0;
        long double t1 = std::floor(16807 * seed / d2p31m) * d2p31m;
        long double t2 = 16807 * seed;
        long double ns = t2 - t1;
        seed = (double)ns;
        return std::fabs(seed / d2p31);
    }

    SuperEgo(const double *p, uint32_t n_, uint32_t dim_, double eps_)
        : n(n_), dim(dim_), eps(eps_), eps2(eps_ * eps_), id(n_), x((size_t)n_ * dim_),
          r1(2 * (dim_ + 1)), r2(2 * (dim_ + 1)), r3(2 * (dim_ + 1)), egoMapping(n_) {
        for (uint32_t i = 0; i < n; i++) id[i] = (int)i;
        std::memcpy(x.data(), p, sizeof(double) * x.size());
    }

    // Util::reorderDim, Util.cpp:33-281.  Only the mean |dx| statistic
    // drives the map (g = -d, Util.cpp:115); the histograms are printed only.
    void reorder() {
        std::vector<double> d(dim, 0.0);
        const int sample_sz = 2000;
        for (int i = 0; i < sample_sz; i++) {
            int ia = (int)((n - 1) * rnd());
            int ib = (int)((n - 1) * rnd());
            for (uint32_t j = 0; j < dim; j++)
                d[j] += std::fabs(x[(size_t)ia * dim + j] - x[(size_t)ib * dim + j]);
        }
        std::vector<long double> g(dim);
        for (uint32_t i = 0; i < dim; i++) {
            d[i] = d[i] / sample_sz;
            g[i] = -d[i];
        }
        std::vector<int> map(dim);
        for (uint32_t i = 0; i < dim; i++) {
            double mn = 1000;
            int mj = -1;
            for (uint32_t j = 0; j < dim; j++)
                if (g[j] < mn) {
                    mn = (double)g[j];
                    mj = (int)j;
                }
            map[i] = mj;
            g[mj] = 2.0;
        }
        std::vector<double> tmp(dim);
        for (uint32_t i = 0; i < n; i++) {
            double *row = &x[(size_t)i * dim];
            for (uint32_t j = 0; j < dim; j++) tmp[j] = row[j];
            for (uint32_t j = 0; j < dim; j++) row[j] = tmp[map[j]];
        }
        std::vector<double> rd(dim);
        for (uint32_t j = 0; j < dim; j++) rd[j] = d[map[j]];
        auto set = [&](std::vector<int> &r, uint32_t i, int a, int b) {
            r[2 * i] = a;
            r[2 * i + 1] = b;
        };
        const int D = (int)dim;
        set(r1, 0, 0, D - 1);
        set(r2, 0, 0, -1);
        set(r3, 0, 0, -1);
        set(r1, dim, 0, D - 1);
        set(r2, dim, 0, -1);
        set(r3, dim, 0, -1);
        int kk = D;
        for (int i = 0; i < D; i++)
            if (rd[i] < eps / 2) {
                kk = i;
                break;
            }
        for (int i = 1; i < D; i++) {
            if (rd[i] < eps / 2) {
                set(r1, i, 0, D - 1);
                set(r2, i, 0, -1);
                set(r3, i, 0, -1);
                continue;
            }
            if (kk < D) {
                set(r1, i, i, kk - 1);
                set(r2, i, 0, i - 1);
                set(r3, i, kk, D - 1);
                continue;
            }
            set(r1, i, i, D - 1);
            set(r2, i, 0, i - 1);
            set(r3, i, 0, -1);
        }
    }

    int cell(size_t p, uint32_t j) const { return (int)(x[p * dim + j] / eps); }

    // EGO sort: egoSortFunction (main.cu:30-39), lexicographic on
    // (int)(x/eps) over all dims.  A stable sort fixes the tie order (the
    // reference's boost sample_sort leaves it unspecified; the pair set does
    // not depend on it).
    void ego_sort() {
        std::vector<uint32_t> perm(n);
        for (uint32_t i = 0; i < n; i++) perm[i] = i;
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
            for (uint32_t i = 0; i < dim; i++) {
                int ca = cell(a, i), cb = cell(b, i);
                if (ca < cb) return true;
                if (ca > cb) return false;
            }
            return false;
        });
        std::vector<double> nx(x.size());
        std::vector<int> nid(n);
        for (uint32_t i = 0; i < n; i++) {
            std::memcpy(&nx[(size_t)i * dim], &x[(size_t)perm[i] * dim], sizeof(double) * dim);
            nid[i] = id[perm[i]];
        }
        x.swap(nx);
        id.swap(nid);
        for (uint32_t i = 0; i < n; i++) egoMapping[id[i]] = i;  // main.cu:313-318
    }

    // Distance with early exit over [a, b], Util.cpp:419-427 / 487-506.
    bool scan(const double *p, const double *q, int a, int b, double &sum) const {
        for (int kx = a; kx <= b; kx++) {
            double dx = p[kx] - q[kx];
            sum = std::fma(dx, dx, sum);
            if (sum > eps2) return false;
        }
        return true;
    }

    // Util::simpleJoin4 (+ simpleJoin3), Util.cpp:394-589.
    void simple_join(int frA, int toA, int frB, int toB, int m, std::vector<int> &res) const {
        const int D = (int)dim;
        int r1b = r1[2 * m], r1e = r1[2 * m + 1], r2e = r2[2 * m + 1];
        int r3b = r3[2 * m], r3e = r3[2 * m + 1];
        for (int i = frB; i <= toB; i++) {
            const double *q = &x[(size_t)i * dim];
            for (int j = frA; j <= toA; j++) {
                const double *p = &x[(size_t)j * dim];
                double sum = 0;
                bool ok;
                if (r3e == -1) {
                    if (r2e == -1)
                        ok = scan(p, q, 0, D - 1, sum);  // simpleJoin3
                    else
                        ok = scan(p, q, r1b, D - 1, sum) && scan(p, q, 0, r2e, sum);
                } else {
                    ok = scan(p, q, r1b, r1e, sum) && scan(p, q, 0, r2e, sum) &&
                         scan(p, q, r3b, D - 1, sum);
                }
                if (ok) res.push_back(id[j]);
            }
        }
    }

    // Util::egoJoinV2, Util.cpp:314-390 (MINLEN = 32, params.h:28).
    void ego_join(int frA, int toA, int frB, int toB, int start_dim, std::vector<int> &res) const {
        const int MINLEN = 32;
        int i;
        for (i = start_dim; i < (int)dim; i++) {
            int lo_A = cell(frA, i), hi_B = cell(toB, i);
            if (lo_A > hi_B + 1) return;
            int lo_B = cell(frB, i), hi_A = cell(toA, i);
            if (lo_B > hi_A + 1) return;
            if (lo_A < hi_A || lo_B < hi_B) {
                start_dim = i;
                break;
            }
        }
        int A_sz = toA - frA + 1, B_sz = toB - frB + 1;
        if (A_sz < MINLEN && B_sz < MINLEN) {
            simple_join(frA, toA, frB, toB, i, res);
            return;
        }
        if (A_sz >= MINLEN && B_sz >= MINLEN) {
            ego_join(frA, frA + A_sz / 2, frB, frB + B_sz / 2, start_dim, res);
            ego_join(frA, frA + A_sz / 2, frB + B_sz / 2 + 1, toB, start_dim, res);
            ego_join(frA + A_sz / 2 + 1, toA, frB, frB + B_sz / 2, start_dim, res);
            ego_join(frA + A_sz / 2 + 1, toA, frB + B_sz / 2 + 1, toB, start_dim, res);
            return;
        }
        if (A_sz >= MINLEN && B_sz < MINLEN) {
            ego_join(frA, frA + A_sz / 2, frB, toB, start_dim, res);
            ego_join(frA + A_sz / 2 + 1, toA, frB, toB, start_dim, res);
            return;
        }
        if (A_sz < MINLEN && B_sz >= MINLEN) {
            ego_join(frA, toA, frB, frB + B_sz / 2, start_dim, res);
            ego_join(frA, toA, frB + B_sz / 2 + 1, toB, start_dim, res);
            return;
        }
    }

    // MultiThreadJoin.cpp:148-152: per query, egoJoinV2 of the whole set A
    // against the single point B[index].
    void query(uint32_t qid, std::vector<int> &res) const {
        int index = (int)egoMapping[qid];
        ego_join(0, (int)n - 1, index, index, 0, res);
    }
};

}  // namespace

extern "C" {

void og_generate_expo(uint32_t seed, uint32_t n, uint32_t dim, double lambda, double *out) {
    std::mt19937 gen(seed);
    std::exponential_distribution<double> dis(lambda);
    unsigned int length = 1;
    for (uint64_t i = 0; i < (uint64_t)n * dim; i++) {
        double val;
        do {
            val = dis(gen) * length;
        } while (val < 0 || val > 1);
        out[i] = val;
    }
}

int og_write_text(const char *path, const double *pts, uint32_t n, uint32_t dim, int precision) {
    std::ofstream f(path);
    if (!f) return -1;
    f << std::setprecision(precision);
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t j = 0; j < dim; j++) {
            f << pts[(size_t)i * dim + j];
            if (j < dim - 1) f << ", ";
        }
        f << '\n';
    }
    return f ? 0 : -1;
}

int64_t og_load_text(const char *path, uint32_t dim, double **out) {
    std::ifstream in(path);
    if (!in) return -1;
    std::vector<double> all;
    for (std::string f; std::getline(in, f, ',');) {
        double v;
        std::stringstream ss(f);
        while (ss >> v) {
            all.push_back(v);
            if (ss.peek() == ',') ss.ignore();
        }
    }
    uint64_t npts = all.size() / dim;
    double *buf = (double *)std::malloc(sizeof(double) * (npts * dim + 1));
    std::memcpy(buf, all.data(), sizeof(double) * npts * dim);
    *out = buf;
    return (int64_t)npts;
}

void og_free(void *p) { std::free(p); }

void og_grid_dims(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                  double *minArr, double *maxArr, uint32_t *nCells, uint64_t *totalCells) {
    Grid<double> g(pts, n, dim, k, eps);
    for (uint32_t j = 0; j < k; j++) {
        minArr[j] = g.minArr[j];
        maxArr[j] = g.maxArr[j];
        nCells[j] = g.nCells[j];
    }
    *totalCells = g.totalCells;
}

void og_point_cells(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                    const double *minArr, const uint32_t *nCells, uint64_t *ids) {
    std::vector<uint32_t> c(k);
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t j = 0; j < k; j++)
            c[j] = (uint32_t)((pts[(size_t)i * dim + j] - minArr[j]) / eps);
        ids[i] = linear_id(c.data(), nCells, k);
    }
}

uint32_t og_index_build(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                        const double *minArr, const uint32_t *nCells, uint32_t *indexLookupArr,
                        og_cell_lookup *lookup, og_grid *grid) {
    (void)minArr;
    (void)nCells;
    Grid<double> g(pts, n, dim, k, eps);
    std::memcpy(indexLookupArr, g.indexLookupArr.data(), sizeof(uint32_t) * n);
    for (size_t i = 0; i < g.cellKeys.size(); i++) {
        lookup[i].idx = (uint32_t)i;
        lookup[i]._pad = 0;
        lookup[i].gridLinearID = g.cellKeys[i];
        grid[i] = g.grid[i];
    }
    return (uint32_t)g.cellKeys.size();
}

void og_chain(const double *q, const double *c, uint32_t npairs, uint32_t dim, double *acc) {
    for (uint64_t i = 0; i < npairs; i++) {
        double a = 0;
        for (uint32_t l = 0; l < dim; l++) {
            const double d = c[i * dim + l] - q[i * dim + l];
            a = std::fma(d, d, a);
        }
        acc[i] = a;
    }
}

double og_threshold(double eps) {
    double s = eps * eps;
    const double inf = std::numeric_limits<double>::infinity();
    while (s > 0 && std::sqrt(s) > eps) s = std::nextafter(s, 0.0);
    while (std::sqrt(std::nextafter(s, inf)) <= eps) s = std::nextafter(s, inf);
    return s;
}

float og_threshold_f(float eps) {
    float s = eps * eps;
    const float inf = std::numeric_limits<float>::infinity();
    while (s > 0 && std::sqrt(s) > eps) s = std::nextafter(s, 0.0f);
    while (std::sqrt(std::nextafter(s, inf)) <= eps) s = std::nextafter(s, inf);
    return s;
}

uint64_t og_workload(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                     const double *minArr, const uint32_t *nCells, const uint32_t *indexLookupArr,
                     const og_cell_lookup *lookup, const og_grid *grid, uint32_t nNonEmpty,
                     int emulate_cell0, uint32_t *nbPoints, uint32_t *sortedCells,
                     uint32_t *originPointIndex) {
    (void)minArr;
    (void)nCells;
    (void)indexLookupArr;
    (void)lookup;
    (void)grid;
    Grid<double> g(pts, n, dim, k, eps);
    const uint32_t C = (uint32_t)g.cellKeys.size();
    if (C != nNonEmpty) return UINT64_MAX;
    // sortByWorkLoadGlobal, kernel.cu:95-142: the stencil of the cell's first
    // point, summing the sizes of the non-empty stencil cells.
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t tid = 0; tid < (int64_t)C; tid++) {
        uint32_t tmpId = g.indexLookupArr[g.grid[tid].indexmin];
        std::vector<uint32_t> c(k);
        g.coords(&g.pts[(size_t)tmpId * dim], c.data());
        uint32_t nb = 0;
        g.for_stencil(c.data(), emulate_cell0, [&](uint32_t ci) {
            nb += (uint32_t)(g.grid[ci].indexmax - g.grid[ci].indexmin + 1);
        });
        nbPoints[tid] = nb;
    }
    // SortByWorkload.cu:97-98 std::sort descending by nbPoints.
    for (uint32_t i = 0; i < C; i++) sortedCells[i] = i;
    std::stable_sort(sortedCells, sortedCells + C,
                     [&](uint32_t a, uint32_t b) { return nbPoints[a] > nbPoints[b]; });
    // SortByWorkload.cu:155-181 originPointIndex + accNeighbor.
    uint64_t acc = 0;
    uint32_t prec = 0;
    for (uint32_t i = 0; i < C; i++) {
        uint32_t cid = sortedCells[i];
        uint32_t nb = (uint32_t)(g.grid[cid].indexmax - g.grid[cid].indexmin + 1);
        acc += (uint64_t)nb * nbPoints[cid];
        for (uint32_t j = 0; j < nb; j++) originPointIndex[prec + j] = g.indexLookupArr[g.grid[cid].indexmin + j];
        prec += nb;
    }
    return acc;
}

uint64_t og_join_stats(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                       int emulate_cell0, int dtype_f32, uint64_t *hash, uint32_t *row_cnt,
                       uint64_t *row_hash) {
    if (dtype_f32) return join_stats_t<float>(pts, n, dim, k, eps, emulate_cell0, hash, row_cnt, row_hash);
    return join_stats_t<double>(pts, n, dim, k, eps, emulate_cell0, hash, row_cnt, row_hash);
}

uint64_t og_join_rows(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                      int emulate_cell0, int dtype_f32, const uint32_t *qids, uint32_t nq,
                      uint32_t *counts, const uint64_t *offsets, uint32_t *nbrs) {
    if (dtype_f32)
        return join_rows_t<float>(pts, n, dim, k, eps, emulate_cell0, qids, nq, counts, offsets, nbrs);
    return join_rows_t<double>(pts, n, dim, k, eps, emulate_cell0, qids, nq, counts, offsets, nbrs);
}

uint64_t og_join_counts(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                        int emulate_cell0, int dtype_f32, uint32_t *counts) {
    if (dtype_f32) return join_counts_t<float>(pts, n, dim, k, eps, emulate_cell0, counts);
    return join_counts_t<double>(pts, n, dim, k, eps, emulate_cell0, counts);
}

void og_join_fill(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                  int emulate_cell0, int dtype_f32, const uint64_t *offsets, uint32_t *nbrs) {
    if (dtype_f32)
        join_fill_t<float>(pts, n, dim, k, eps, emulate_cell0, offsets, nbrs);
    else
        join_fill_t<double>(pts, n, dim, k, eps, emulate_cell0, offsets, nbrs);
}

uint32_t og_count_cell0(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps) {
    Grid<double> g(pts, n, dim, k, eps);
    uint32_t dead = 0;
    std::vector<uint32_t> c(k);
    for (uint32_t i = 0; i < n; i++) {
        g.coords(&g.pts[(size_t)i * dim], c.data());
        for (uint32_t j = 0; j < k; j++)
            if (c[j] == 0) {
                dead++;
                break;
            }
    }
    return dead;
}

// Brute force: every ordered pair, literal fma chain + sqrt, no grid.
static bool brute_dead(const double *x, const double *minArr, uint32_t k, double eps) {
    for (uint32_t j = 0; j < k; j++)
        if ((uint32_t)((x[j] - minArr[j]) / eps) == 0) return true;
    return false;
}

uint64_t og_brute_counts(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                         int emulate_cell0, uint32_t *counts) {
    std::vector<double> mn(k), mx(k);
    std::vector<uint32_t> nc(k);
    uint64_t tc;
    og_grid_dims(pts, n, dim, k, eps, mn.data(), mx.data(), nc.data(), &tc);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t q = 0; q < (int64_t)n; q++) {
        const double *xq = pts + (size_t)q * dim;
        uint32_t cnt = 0;
        if (!(emulate_cell0 && brute_dead(xq, mn.data(), k, eps))) {
            for (uint32_t c = 0; c < n; c++) {
                const double *xc = pts + (size_t)c * dim;
                double acc = 0;
                for (uint32_t l = 0; l < dim; l++) acc = std::fma(xc[l] - xq[l], xc[l] - xq[l], acc);
                if (std::sqrt(acc) <= eps) cnt++;
            }
        }
        counts[q] = cnt;
        total += cnt;
    }
    return total;
}

void og_brute_fill(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                   int emulate_cell0, const uint64_t *offsets, uint32_t *nbrs) {
    std::vector<double> mn(k), mx(k);
    std::vector<uint32_t> nc(k);
    uint64_t tc;
    og_grid_dims(pts, n, dim, k, eps, mn.data(), mx.data(), nc.data(), &tc);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < (int64_t)n; q++) {
        const double *xq = pts + (size_t)q * dim;
        if (emulate_cell0 && brute_dead(xq, mn.data(), k, eps)) continue;
        uint64_t o = offsets[q];
        for (uint32_t c = 0; c < n; c++) {
            const double *xc = pts + (size_t)c * dim;
            double acc = 0;
            for (uint32_t l = 0; l < dim; l++) acc = std::fma(xc[l] - xq[l], xc[l] - xq[l], acc);
            if (std::sqrt(acc) <= eps) nbrs[o++] = c;
        }
    }
}

// Range join of external queries against a point set (the serving path; no
// reference counterpart): every p with sqrt(acc) <= eps, acc the fma chain of
// kernel.cu:350-355 over (p - q), neighbours in ascending id order.
uint64_t og_range_counts(const double *pts, uint32_t n, const double *qs, uint32_t nq, uint32_t dim,
                         double eps, uint32_t *counts) {
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t q = 0; q < (int64_t)nq; q++) {
        const double *xq = qs + (size_t)q * dim;
        uint32_t cnt = 0;
        for (uint32_t c = 0; c < n; c++) {
            const double *xc = pts + (size_t)c * dim;
            double acc = 0;
            for (uint32_t l = 0; l < dim; l++) acc = std::fma(xc[l] - xq[l], xc[l] - xq[l], acc);
            cnt += std::sqrt(acc) <= eps;
        }
        counts[q] = cnt;
        total += cnt;
    }
    return total;
}

void og_range_fill(const double *pts, uint32_t n, const double *qs, uint32_t nq, uint32_t dim,
                   double eps, const uint64_t *offsets, uint32_t *nbrs) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < (int64_t)nq; q++) {
        const double *xq = qs + (size_t)q * dim;
        uint64_t o = offsets[q];
        for (uint32_t c = 0; c < n; c++) {
            const double *xc = pts + (size_t)c * dim;
            double acc = 0;
            for (uint32_t l = 0; l < dim; l++) acc = std::fma(xc[l] - xq[l], xc[l] - xq[l], acc);
            if (std::sqrt(acc) <= eps) nbrs[o++] = c;
        }
    }
}

uint64_t og_pair_mix(uint32_t key, uint32_t value) {
    return mix64(((uint64_t)key << 32) | (uint64_t)value);
}

uint64_t og_csr_hash(uint32_t n, const uint64_t *offsets, const uint32_t *nbrs) {
    uint64_t h = 0;
#pragma omp parallel for reduction(+ : h) schedule(static)
    for (int64_t q = 0; q < (int64_t)n; q++)
        for (uint64_t i = offsets[q]; i < offsets[q + 1]; i++) h += og_pair_mix((uint32_t)q, nbrs[i]);
    return h;
}

uint64_t og_superego_join(const double *pts, uint32_t n, uint32_t dim, double eps, int nthreads,
                          uint32_t *counts, const uint64_t *offsets, uint32_t *nbrs,
                          double *times) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    SuperEgo se(pts, n, dim, eps);
    double t0 = omp_get_wtime();
    se.reorder();
    double t1 = omp_get_wtime();
    se.ego_sort();
    double t2 = omp_get_wtime();
    uint64_t total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
    {
        std::vector<int> res;
#pragma omp for schedule(dynamic, 64)
        for (int64_t q = 0; q < (int64_t)n; q++) {
            res.clear();
            se.query((uint32_t)q, res);
            total += res.size();
            if (counts) counts[q] = (uint32_t)res.size();
            if (nbrs) {
                std::sort(res.begin(), res.end());
                for (size_t i = 0; i < res.size(); i++) nbrs[offsets[q] + i] = (uint32_t)res[i];
            }
        }
    }
    double t3 = omp_get_wtime();
    if (times) {
        times[0] = t1 - t0;
        times[1] = t2 - t1;
        times[2] = t3 - t2;
    }
    return total;
}

uint64_t og_superego_sample(const double *pts, uint32_t n, uint32_t dim, double eps, int nthreads,
                            const uint32_t *qids, uint32_t nq, double *times) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    SuperEgo se(pts, n, dim, eps);
    double t0 = omp_get_wtime();
    se.reorder();
    double t1 = omp_get_wtime();
    se.ego_sort();
    double t2 = omp_get_wtime();
    uint64_t total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
    {
        std::vector<int> res;
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < (int64_t)nq; i++) {
            res.clear();
            se.query(qids[i], res);
            total += res.size();
        }
    }
    double t3 = omp_get_wtime();
    if (times) {
        times[0] = t1 - t0;
        times[1] = t2 - t1;
        times[2] = t3 - t2;
    }
    return total;
}

}  // extern "C"
