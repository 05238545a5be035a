/*
 * hegj_oracle.h -- CPU oracle for the HEGJoin epsilon-grid self-join.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the CLI,
 * the Python package) links, loads or calls this code.  Only tests/, the
 * smoke() self-check and the cpu_baseline leg of bench.py use it, and there
 * only as the expected result / the timed CPU baseline.
 *
 * It restates, in plain sequential C++ (OpenMP only over independent queries),
 * the reference's algorithm-0 semantics (grid extents, cell ids, grid index,
 * workload sort, stencil join with the exact FP64 predicate and the cell-0
 * wrap) and its algorithm-3 Super-EGO join.  Every function cites the
 * reference file:line it follows.
 *
 * Parity pins (what checks each function, see tests/test_oracle.py):
 *   og_generate_expo     -- byte-identical text vs the reference generator
 *                           compiled from source (oracle/_ref), same seed.
 *   og_load_text         -- identical doubles vs the reference importNDDataset
 *                           compiled from source (oracle/_ref).
 *   og_grid_dims         -- closed forms on hand-made data (exact extents).
 *   og_point_cells       -- brute-force floor() definition on lattices.
 *   og_index_build       -- invariants (stable sort by (id, point), run
 *                           boundaries) + the unused host builder spec
 *                           (main.cu:555-737) restated independently in tests.
 *   og_threshold         -- definition: sqrt(T) <= eps < sqrt(next(T)) on
 *                           many random eps, and sqrt-literal brute force.
 *   og_workload          -- brute-force per-cell stencil sums.
 *   og_join_*            -- O(N^2) brute force with the literal sqrt predicate
 *                           (N <= 5k, D in {2,6,18,90}), lattice closed forms,
 *                           symmetry and self-pair invariants.
 *   og_superego_*        -- the reference's own Util.cpp/MultiThreadJoin.cpp
 *                           compiled from source (oracle/_ref): identical
 *                           per-query neighbour sets.
 *   og_pair_hash         -- a project checksum (not a reference semantic):
 *                           order independence only.
 */
#ifndef HEGJ_ORACLE_H
#define HEGJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* structs.h:83-98 layouts */
typedef struct { int32_t indexmin, indexmax; } og_grid;
typedef struct { uint32_t idx; uint32_t _pad; uint64_t gridLinearID; } og_cell_lookup;

/* Dataset generator: datasets/dataset_gen_fixed_len_expo_dist.cpp:57-84.
 * mt19937(seed), exponential_distribution<double>(lambda), redraw while the
 * value is < 0 or > 1, point-major order.  out has n*dim doubles. */
void og_generate_expo(uint32_t seed, uint32_t n, uint32_t dim, double lambda, double *out);

/* Writes "v, v, ...\n" per point with the given significant digits (6 is the
 * reference generator's default ostream precision; 17 round-trips doubles). */
int og_write_text(const char *path, const double *pts, uint32_t n, uint32_t dim, int precision);

/* Loader: import_dataset.cpp:10-40 (comma-separated chunks, each chunk read
 * with operator>> until it fails).  Returns the number of points (values/dim,
 * truncated) and writes them into *out (malloc'd; free with og_free). */
int64_t og_load_text(const char *path, uint32_t dim, double **out);
void og_free(void *p);

/* S1: generateNDGridDimensions, main.cu:449-520 (first k dims only). */
void og_grid_dims(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                  double *minArr, double *maxArr, uint32_t *nCells, uint64_t *totalCells);

/* S2+S3: kernelIndexComputeNonemptyCells + getLinearID_nDimensionsGPU,
 * kernel.cu:34-66, 310-325. */
void og_point_cells(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                    const double *minArr, const uint32_t *nCells, uint64_t *ids);

/* S11: gridIndexingGPU, GPU.cu:236-324.  indexLookupArr[n]; lookup and grid
 * need room for n entries.  Returns nNonEmptyCells. */
uint32_t og_index_build(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                        const double *minArr, const uint32_t *nCells, uint32_t *indexLookupArr,
                        og_cell_lookup *lookup, og_grid *grid);

/* evalPoint's distance (kernel.cu:347-355) for explicit pairs: acc[i] =
 * fma chain over dims 0..dim-1 of (c_i - q_i)^2, q and c row-major
 * npairs x dim. */
void og_chain(const double *q, const double *c, uint32_t npairs, uint32_t dim, double *acc);

/* S7: largest s with RN(sqrt(s)) <= eps (so sqrt(acc) <= eps <=> acc <= T). */
double og_threshold(double eps);
float og_threshold_f(float eps);

/* Workload: sortByWorkLoadGlobal (kernel.cu:74-144) + the host part of
 * sortByWorkLoad (SortByWorkload.cu:96-217).  nbPoints[C] is per cell in
 * gridCellLookupArr order; sortedCells[C] = cell ids sorted by nbPoints
 * descending, ties by cell id ascending (the reference's std::sort leaves tie
 * order unspecified); originPointIndex[n] = member lists in that order.
 * Returns totalCandidates = sum |c| * nbPoints(c) (uint64; the reference
 * computes it in int and overflows at 90-D, SortByWorkload.cu:158-160). */
uint64_t og_workload(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                     const double *minArr, const uint32_t *nCells, const uint32_t *indexLookupArr,
                     const og_cell_lookup *lookup, const og_grid *grid, uint32_t nNonEmpty,
                     int emulate_cell0, uint32_t *nbPoints, uint32_t *sortedCells,
                     uint32_t *originPointIndex);

/* The algorithm-0 join (kernelNDGridIndexGlobal -> evaluateCell -> evalPoint,
 * kernel.cu:727-802, 378-412, 333-370): per query its stencil (S4, S5 when
 * emulate_cell0), binary search of each stencil cell, fma distance chain in
 * dimension order (S6), sqrt(acc) <= eps (S7).  Pass 1 writes counts[n]
 * (per query, original ids) and returns the total; pass 2 fills neighbour
 * lists at offsets[n+1] (ascending neighbour id within a row).  dtype_f32 != 0
 * evaluates the reference's DTYPE=float build (coordinates given as doubles
 * are rounded to float first). */
uint64_t og_join_counts(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                        int emulate_cell0, int dtype_f32, uint32_t *counts);
void og_join_fill(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                  int emulate_cell0, int dtype_f32, const uint64_t *offsets, uint32_t *nbrs);

/* Count and multiset hash (sum of og_pair_mix(query, neighbour)) of the same
 * pair set as og_join_counts, without holding the pairs: the stencil is
 * enumerated once per cell and each fma chain stops once it exceeds the
 * threshold (the chain never decreases).  row_cnt[n] / row_hash[n] (may be
 * NULL) receive per-query counts and hashes.  Pinned by equality with
 * og_join_counts + og_csr_hash (tests/test_oracle.py). */
uint64_t og_join_stats(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                       int emulate_cell0, int dtype_f32, uint64_t *hash, uint32_t *row_cnt,
                       uint64_t *row_hash);

/* og_join_counts / og_join_fill for the queries qids[nq] only: counts[nq];
 * with nbrs, offsets[nq+1] from the counts and rows ascending. */
uint64_t og_join_rows(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                      int emulate_cell0, int dtype_f32, const uint32_t *qids, uint32_t nq,
                      uint32_t *counts, const uint64_t *offsets, uint32_t *nbrs);

/* Brute force O(n^2), no grid: pair (q, c) iff sqrt(fma-chain(q,c)) <= eps,
 * computed literally with sqrt; a query whose cell has a zero coordinate in
 * any of the first k dims emits nothing when emulate_cell0.  Same outputs as
 * og_join_counts / og_join_fill. */
uint64_t og_range_counts(const double *pts, uint32_t n, const double *qs, uint32_t nq, uint32_t dim,
                         double eps, uint32_t *counts);
void og_range_fill(const double *pts, uint32_t n, const double *qs, uint32_t nq, uint32_t dim,
                   double eps, const uint64_t *offsets, uint32_t *nbrs);
uint64_t og_brute_counts(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                         int emulate_cell0, uint32_t *counts);
void og_brute_fill(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps,
                   int emulate_cell0, const uint64_t *offsets, uint32_t *nbrs);

/* Number of queries the cell-0 wrap silences (cell coordinate 0 in any of
 * the first k dims). */
uint32_t og_count_cell0(const double *pts, uint32_t n, uint32_t dim, uint32_t k, double eps);

/* Project checksum: order-independent multiset hash of (key, value) pairs =
 * sum over pairs of mix64((key << 32) | value) mod 2^64. */
uint64_t og_pair_mix(uint32_t key, uint32_t value);
uint64_t og_csr_hash(uint32_t n, const uint64_t *offsets, const uint32_t *nbrs);

/* Super-EGO (algorithm 3): Util.cpp:33-281 (reorderDim), main.cu:30-39,306
 * (EGO sort), Util.cpp:314-589 (egoJoinV2, simpleJoin3/4),
 * MultiThreadJoin.cpp:116-203 (per query egoJoinV2 over the whole set).
 * counts[n] per original query id; if nbrs != NULL, offsets[n+1] must be
 * filled from the counts of a previous call and rows are written sorted
 * ascending.  times[3] = reorder, sort, join seconds (may be NULL).
 * nthreads <= 0 uses all OpenMP threads.  Returns the total. */
uint64_t og_superego_join(const double *pts, uint32_t n, uint32_t dim, double eps, int nthreads,
                          uint32_t *counts, const uint64_t *offsets, uint32_t *nbrs,
                          double *times);

/* Super-EGO timing over a bounded query sample: runs the full reorder + sort,
 * then joins only the nq queries qids[] (original ids).  Returns pairs found. */
uint64_t og_superego_sample(const double *pts, uint32_t n, uint32_t dim, double eps, int nthreads,
                            const uint32_t *qids, uint32_t nq, double *times);

#ifdef __cplusplus
}
#endif
#endif
