// ref_harness.cpp -- driver for the reference's own CPU code (test
// infrastructure only).  Linked with the reference's Util.cpp, Point.cpp,
// MultiThreadJoin.cpp, WorkQueue.cpp and import_dataset.cpp, compiled from
// their sources by oracle/Makefile into oracle/_ref/ref_harness_d<D>.
//
// The reference's own driver (main.cu) needs nvcc, a CUDA GPU and Boost, so
// this harness reproduces the algorithm-3 call sequence of main.cu:286-339
// itself: Util::reorderDim, an EGO sort with the egoSortFunction ordering
// (main.cu:30-39; std::stable_sort instead of boost::sort::sample_sort, which
// only changes the order of ties), the egoMapping, then either
//   pairs  -- Util::egoJoinV2 per query (MultiThreadJoin.cpp:152) collecting
//             the neighbour ids, written as counts + sorted rows;
//   time   -- Util::multiThreadJoinWorkQueue (MultiThreadJoin.cpp:22) in
//             mode SM_CPU over the first nq entries of a query order given
//             in a file (a seeded sample for a bounded baseline run);
//   load   -- importNDDataset (import_dataset.cpp:10) and dump the doubles;
//   rowstats -- the same per-query egoJoinV2 as `pairs`, keeping per query
//             only its neighbour count and the multiset hash
//             sum mix64(query << 32 | neighbour) (the full-size workloads have
//             up to 1.2e9 pairs);
//   rows   -- `pairs` for a list of query ids only.
//
// Usage:
//   ref_harness_dD load  <dataset.txt> <out.bin>
//   ref_harness_dD pairs <points.bin> <eps> <out.bin>
//   ref_harness_dD time  <points.bin> <eps> <order.bin> <nq>
//   ref_harness_dD rowstats <points.bin> <eps> <out.bin>   (n x u32 counts, n x u64 hashes)
//   ref_harness_dD rows <points.bin> <eps> <qids.bin> <out.bin>
// points.bin = raw row-major doubles (n*D); order.bin / qids.bin = uint32 ids.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

#include "Point.hpp"
#include "Util.hpp"
#include "WorkQueue.h"
#include "params.h"
#include "structs.h"

void importNDDataset(std::vector<std::vector<DTYPE> > *dataPoints, char *fname);

static uint64_t mix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static std::vector<double> read_bin(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        std::perror(path);
        std::exit(2);
    }
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<double> v(sz / sizeof(double));
    if (std::fread(v.data(), sizeof(double), v.size(), f) != v.size()) std::exit(2);
    std::fclose(f);
    return v;
}

// Ordering of egoSortFunction (main.cu:30-39).
static bool ego_less(Point const &p1, Point const &p2) {
    for (int i = 0; i < GPUNUMDIM; i++) {
        if ((int)(p1.x[i] / Util::eps) < (int)(p2.x[i] / Util::eps)) return true;
        if ((int)(p1.x[i] / Util::eps) > (int)(p2.x[i] / Util::eps)) return false;
    }
    return false;
}

// main.cu:101-118, 296-318.
static Point *prepare(const std::vector<double> &v, unsigned n, double eps, double *tReorder,
                      double *tSort, unsigned **mapping) {
    Util::eps = eps;
    Util::eps2 = eps * eps;
    Point *A = new Point[n + 1];
    for (unsigned i = 0; i < n; ++i) {
        A[i].id = (int)i;
        for (int j = 0; j < GPUNUMDIM; ++j) A[i].x[j] = v[(size_t)i * GPUNUMDIM + j];
    }
    double t0 = omp_get_wtime();
    Util::reorderDim(A, (int)n, A, (int)n);
    double t1 = omp_get_wtime();
    std::stable_sort(A, A + n, ego_less);
    double t2 = omp_get_wtime();
    unsigned *m = new unsigned[n];
    for (unsigned i = 0; i < n; ++i) m[A[i].id] = i;
    *tReorder = t1 - t0;
    *tSort = t2 - t1;
    *mapping = m;
    return A;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: see ref_harness.cpp\n");
        return 2;
    }
    std::string mode = argv[1];
    if (mode == "load") {
        std::vector<std::vector<DTYPE> > pts;
        importNDDataset(&pts, argv[2]);
        FILE *f = std::fopen(argv[3], "wb");
        for (auto &p : pts) std::fwrite(p.data(), sizeof(DTYPE), p.size(), f);
        std::fclose(f);
        return 0;
    }
    std::vector<double> v = read_bin(argv[2]);
    unsigned n = (unsigned)(v.size() / GPUNUMDIM);
    double eps = std::atof(argv[3]);
    double tReorder, tSort;
    unsigned *mapping;
    Point *A = prepare(v, n, eps, &tReorder, &tSort, &mapping);
    if (mode == "pairs") {
        std::vector<std::vector<int> > rows(n);
#pragma omp parallel for schedule(dynamic, 64)
        for (long q = 0; q < (long)n; q++) {
            uint64_t cand = 0;
            int index = (int)mapping[q];
            Util::egoJoinV2(A, 0, (int)n - 1, A, index, index, 0, &cand, &rows[q], nullptr);
            std::sort(rows[q].begin(), rows[q].end());
        }
        FILE *f = std::fopen(argv[4], "wb");
        for (unsigned q = 0; q < n; q++) {
            uint32_t c = (uint32_t)rows[q].size();
            std::fwrite(&c, 4, 1, f);
        }
        for (unsigned q = 0; q < n; q++) std::fwrite(rows[q].data(), 4, rows[q].size(), f);
        std::fclose(f);
        return 0;
    }
    if (mode == "rowstats") {
        std::vector<uint32_t> cnt(n);
        std::vector<uint64_t> hsh(n);
#pragma omp parallel
        {
            std::vector<int> row;
#pragma omp for schedule(dynamic, 64)
            for (long q = 0; q < (long)n; q++) {
                uint64_t cand = 0;
                int index = (int)mapping[q];
                row.clear();
                Util::egoJoinV2(A, 0, (int)n - 1, A, index, index, 0, &cand, &row, nullptr);
                uint64_t h = 0;
                for (int c : row) h += mix64(((uint64_t)q << 32) | (uint32_t)c);
                cnt[q] = (uint32_t)row.size();
                hsh[q] = h;
            }
        }
        FILE *f = std::fopen(argv[4], "wb");
        std::fwrite(cnt.data(), 4, n, f);
        std::fwrite(hsh.data(), 8, n, f);
        std::fclose(f);
        return 0;
    }
    if (mode == "rows" && argc >= 6) {
        FILE *fq = std::fopen(argv[4], "rb");
        std::fseek(fq, 0, SEEK_END);
        long nq = std::ftell(fq) / 4;
        std::fseek(fq, 0, SEEK_SET);
        std::vector<uint32_t> qids(nq);
        if (std::fread(qids.data(), 4, nq, fq) != (size_t)nq) return 2;
        std::fclose(fq);
        std::vector<std::vector<int> > rows(nq);
#pragma omp parallel for schedule(dynamic, 1)
        for (long i = 0; i < nq; i++) {
            uint64_t cand = 0;
            int index = (int)mapping[qids[i]];
            Util::egoJoinV2(A, 0, (int)n - 1, A, index, index, 0, &cand, &rows[i], nullptr);
            std::sort(rows[i].begin(), rows[i].end());
        }
        FILE *f = std::fopen(argv[5], "wb");
        for (long i = 0; i < nq; i++) {
            uint32_t c = (uint32_t)rows[i].size();
            std::fwrite(&c, 4, 1, f);
        }
        for (long i = 0; i < nq; i++) std::fwrite(rows[i].data(), 4, rows[i].size(), f);
        std::fclose(f);
        return 0;
    }
    if (mode == "time" && argc >= 6) {
        FILE *f = std::fopen(argv[4], "rb");
        std::vector<unsigned> order(n);
        if (!f || std::fread(order.data(), 4, n, f) != n) return 2;
        std::fclose(f);
        unsigned nq = (unsigned)std::atoi(argv[5]);
        if (nq > n) nq = n;
        // main.cu:97, 290: the CPU takes queries [0, queueIndexCPU) of the
        // order; limiting queueIndexCPU to nq bounds the run to the sample.
        setQueueIndexCPU(nq);
        setQueueIndex(0);
        neighborTableLookup *table = new neighborTableLookup[n];
        uint64_t *cand = new uint64_t[n]();
        double t0 = omp_get_wtime();
        uint64_t total = Util::multiThreadJoinWorkQueue(SM_CPU, A, (int)n, A, (int)n, mapping,
                                                        order.data(), table, cand);
        double t1 = omp_get_wtime();
        std::printf("\nREF_TIME {\"n\": %u, \"nq\": %u, \"threads\": %d, \"reorder_s\": %.6f, "
                    "\"sort_s\": %.6f, \"join_s\": %.6f, \"pairs\": %llu}\n",
                    n, nq, CPU_THREADS, tReorder, tSort, t1 - t0, (unsigned long long)total);
        return 0;
    }
    std::fprintf(stderr, "bad mode\n");
    return 2;
}
