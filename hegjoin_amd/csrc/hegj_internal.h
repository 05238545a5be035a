// hegj_internal.h -- shared declarations of libhegjoin (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hegjoin.h"

namespace hegj {

// ---------------------------------------------------------------------------
// Error plumbing: every HIP call goes through HIP_TRY, which records a message
// for hegj_last_error() and returns HEGJ_ERR_HIP from the calling function.
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
hegj_status fail(hegj_status s, const std::string &msg);

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            return ::hegj::fail(_e == hipErrorOutOfMemory ? HEGJ_ERR_OUT_OF_MEMORY      \
                                                          : HEGJ_ERR_HIP,               \
                                std::string(#expr) + ": " + hipGetErrorString(_e));     \
        }                                                                               \
    } while (0)

#define HEGJ_TRY(expr)                        \
    do {                                      \
        hegj_status _s = (expr);              \
        if (_s != HEGJ_OK) return _s;         \
    } while (0)

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kMaxK = HEGJ_MAX_INDEXED_DIMS;
constexpr int kMaxRows = 243;      // 3^(kMaxK-1) stencil rows
constexpr int kMaxBatches = 64;    // result batches per join
constexpr uint32_t kAosMaxDim = 8; // point-major copies of the coordinates up to this dim
constexpr uint32_t kAos32Pad = 512; // +inf points after the n of the point-major FP32 copy

// Grid geometry passed by value to kernels.
struct GridGeom {
    double minArr[kMaxK];
    uint32_t nCells[kMaxK];
    double eps;
    double thresh;      // largest s with sqrt(s) <= eps
    uint32_t k;
    uint32_t dim;
    uint32_t n;
    uint32_t emulate_cell0;
    uint32_t f32dt;     // DTYPE=float (params.h:24): minArr, eps and thresh hold float
                        // values; cell ids and distances are computed in FP32
};

// Device-resident grid index.  Naming follows the reference: a "cell" is a
// non-empty grid cell (gridCellLookupArr entry), a "row" groups the cells that
// share indexed coordinates 1..k-1 (they differ only in coordinate 0 and are
// contiguous in linear-id order).
struct DevIndex {
    double *pts = nullptr;        // [n*dim] AoS input copy (dev_database)
    double *soa = nullptr;        // [dim][n] coordinates in cell order
    double *aos = nullptr;        // [n][dim] the same, point-major (dim <= kAosMaxDim): the
                                  // join's candidate loads touch one run of bytes per slot
    uint32_t *ila = nullptr;      // [n] indexLookupArr: cell-order position -> point id
    uint64_t *point_key = nullptr;// [n] linear cell id per point (original order)
    uint64_t *cell_key = nullptr; // [C] gridLinearID of each non-empty cell
    uint32_t *cell_start = nullptr;// [C+1] first cell-order position of each cell
    uint32_t *row_start = nullptr;// [R+1] first cell of each row
    uint64_t *row_key = nullptr;  // [R] row key = cell_key / nCells[0]
    uint64_t *hkeys = nullptr;    // row hash table keys (empty = ~0)
    uint32_t *hvals = nullptr;    // row hash table values (row index)
    uint32_t hmask = 0;
    uint32_t C = 0, R = 0;
};

// Device bump allocator owned by one host thread.  While installed for that
// thread, Scratch and dev_malloc carve temporaries from it instead of calling
// hipMalloc / hipFree (which synchronise the whole device), and dev_free of
// such a pointer does nothing; the owner resets `used`.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};
void set_thread_arena(Arena *a);  // nullptr: back to hipMalloc

// The static device scratch of hegj_selfjoin_noalloc (hegj_scratch.hip):
// its address on `device`, which must be the current device.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
hegj_status noalloc_scratch(int device, void **ptr);

// Scratch allocator for temporaries of one call (freed by the destructor).
struct Scratch {
    std::vector<void *> ptrs;
    hipError_t alloc(void **p, size_t bytes);
    template <typename T>
    hipError_t get(T **p, size_t count) {
        return alloc(reinterpret_cast<void **>(p), count * sizeof(T) + 16);
    }
    ~Scratch();
};

// Device buffers of the join (plan arrays, row tables, result batches) come
// from a cache: a freed buffer is kept for the next join of the same shape
// instead of going back through hipFree / hipMalloc, each of which costs
// ~0.1 ms and synchronises the device.  dev_free only when the work using the
// buffer has completed.  At most 16 GiB of freed buffers are kept; an
// allocation that fails first releases them.  Pointers not from dev_malloc
// are passed to hipFree.
hipError_t dev_malloc(void **p, size_t bytes);
template <typename T>
hipError_t dev_malloc_n(T **p, size_t count) {
    return dev_malloc(reinterpret_cast<void **>(p), count * sizeof(T));
}
void dev_free(void *p);
void cache_release_device(int dev);  // hipFree the cached buffers of one device

// Largest s with RN(sqrt(s)) <= eps (the distance predicate without sqrt).
double sqrt_threshold(double eps);
// The same for DTYPE=float: largest float s with RN32(sqrt(s)) <= eps.
float sqrt_threshold_f(float eps);

// ---- primitives (hegj_prims.hip) -----------------------------------------
// Exclusive scan; writes the total to *d_total when non-null.
hegj_status scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *d_total,
                     hipStream_t s);
hegj_status scan_u64(const uint64_t *in, uint64_t *out, uint32_t n, uint64_t *d_total,
                     hipStream_t s);
// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits.
// Sorted data ends in keys/vals (the alt buffers are clobbered).
hegj_status radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                           uint32_t *vals_alt, uint32_t n, int bits, hipStream_t s);
hegj_status radix_sort_u32(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt,
                           uint32_t *vals_alt, uint32_t n, int bits, hipStream_t s);

// Arguments of the join kernel (hegj_join.hip).
constexpr int kMaxParts = 8;   // one unit range (queue) per XCD
struct JoinArgs {
    GridGeom g;
    DevIndex ix;
    const uint2 *units;           // (cell, first query offset in cell)
    uint32_t n_units;
    uint32_t units_per_grab;
    uint32_t block_waves;        // waves per workgroup of the launch (host side)
    uint32_t n_parts;             // unit ranges; a workgroup starts on its XCD's range
    uint32_t part_by_xcc;         // 1: range = hardware XCC id, 0: blockIdx % n_parts
    uint32_t part_begin[kMaxParts];
    uint32_t part_end[kMaxParts];
    uint32_t *queue_heads;        // [kMaxParts] units taken per range
    uint64_t *gmask;              // per-wave hit-mask scratch (global)
    uint32_t gmask_slots;         // masks per wave in gmask
    uint32_t *out;                // this batch's neighbour buffer
    uint64_t out_base;            // global offset of out[0]
    uint64_t capacity;            // entries in out
    unsigned long long *cursor;   // entries reserved in out
    uint64_t *row_off;            // [n] global offset per query id
    uint32_t *row_cnt;            // [n]
    uint32_t *deferred;           // units that did not fit
    uint32_t *n_deferred;
    unsigned long long *hash;     // multiset hash of written pairs
    unsigned long long *evals;    // distance evaluations performed
    unsigned long long *pairs;    // pairs found (count-only) or written
    uint32_t count_only;
    uint32_t prune;               // 1: cell-box pruning of stencil cells
    // FP32 filter (f32 != 0): centred float coordinates and the two
    // thresholds that make acc32 decide exactly outside the band
    uint32_t f32;
    const float *soa32;           // [dim][n], cell order
    const float *aos32;           // [n][dim], cell order (dim <= kAosMaxDim)
    float tacc32, trej32;
    unsigned long long *band;     // pairs rechecked in FP64
    unsigned long long *mismatch; // units whose write pass disagreed with their count
};

// ---- index (hegj_index.hip) ----------------------------------------------
hegj_status build_index(GridGeom *g, double *maxArr, uint64_t *totalCells, DevIndex *ix,
                        uint32_t *n_dead, hipStream_t s);
void free_index(DevIndex *ix);

// ---- join (hegj_join.hip) ------------------------------------------------
hegj_status launch_workload(const GridGeom &g, const DevIndex &ix, uint32_t *work, hipStream_t s);
// all_cells: also the cells the cell-0 wrap silences (their queries get
// empty rows), so the units cover every point: the hybrid modes' queue.
hegj_status build_units(const GridGeom &g, const DevIndex &ix, uint32_t qb, const uint32_t *order,
                        uint2 **units, uint32_t *n_units, hipStream_t s, bool all_cells = false);
// dev_args != nullptr: the kernel reads its arguments from there (graph replay;
// `a` then only selects the instantiation by a.g.dim).
hegj_status launch_join(const JoinArgs &a, uint32_t qb, uint32_t blocks, hipStream_t s,
                        const JoinArgs *dev_args = nullptr);
// Adds the pair hash of the batch's finished rows (a.out, a.row_off,
// a.row_cnt, a.out_base, a.capacity) to *a.hash; pairs_hint (expected pairs)
// picks the lanes per row.
hegj_status launch_pair_hash(const JoinArgs &a, uint64_t pairs_hint, hipStream_t s,
                             const JoinArgs *dev_args = nullptr);
// FP32 filter support: centred float copy of the cell-ordered coordinates
// (SoA, and point-major into *aos32 when aos32 != nullptr and dim <=
// kAosMaxDim) and the largest |x - centre| (hegj_util.hip).
hegj_status make_soa32(const GridGeom &g, const DevIndex &ix, float **soa32, float **aos32, double *radius,
                       double **centre_dev, hipStream_t s);
// The index's FP32 copy (built on first use), its radius and device centre.
hegj_status index_f32(hegj_index *h, const float **soa32, const float **aos32, double *radius,
                      const double **centre);
// FP32-filter thresholds for threshold T (false: band wider than 1% of T).
bool f32_thresholds(double T, uint32_t dim, double R, float *tacc, float *trej);
// Whether the FP32 filter is instantiated for (dim, qb) and its 32-bit
// offsets cover n points.
bool f32_supported(uint32_t dim, uint32_t qb, uint32_t n);
// Zero `bytes` (a multiple of 4) with a kernel (safe inside graph capture).
hegj_status zero_words(void *p, size_t bytes, hipStream_t s);
// Workgroups of the persistent join kernel that fit on the device at once;
// kind: the FP64 kernel, the DTYPE=float generic instance, or the FP32 filter.
enum class JoinKind { fp64, flt, f32 };
hegj_status join_occupancy(uint32_t dim, uint32_t qb, int device, uint32_t *blocks,
                           JoinKind kind = JoinKind::fp64);

// ---- MFMA guard-band refine (hegj_mfma.hip) --------------------------------
struct GbOutput {
    uint32_t *buf = nullptr;          // neighbour ids, rows contiguous (allocated here)
    uint64_t *row_off = nullptr;      // [n] (caller-allocated)
    uint32_t *row_cnt = nullptr;      // [n] (caller-allocated, zeroed)
    unsigned long long *hash = nullptr;  // device counter (caller-allocated, zeroed)
    uint64_t pairs = 0, band = 0, evals = 0;
    float kernel_ms = 0;
};
uint32_t gb_kblocks(uint32_t dim);        // K blocks of 32 for dim, 0 if unsupported
uint32_t gb_unit_queries(uint32_t kb);    // queries per work unit of the MFMA kernel
double gb_coefficient(uint32_t dim, uint32_t kp);
hegj_status gb_prepare(const GridGeom &g, const DevIndex &ix, uint32_t kp, uint16_t **Av,
                       uint16_t **Bv, float **w, hipStream_t s);
// pair_cap: size (8-byte keys) of the first attempt's pair list, 0 for the
// default of 48 per point; an attempt that overflows is repeated once with a
// list sized from the count it reached.  parts: the units split into that
// many contiguous ranges (1..kMaxParts), one queue per XCD.
hegj_status gb_join(const GridGeom &g, const DevIndex &ix, const uint2 *units, uint32_t n_units,
                    const uint16_t *Av, const uint16_t *Bv, const float *w, uint32_t kp,
                    int device, hipStream_t s, GbOutput *out, uint64_t pair_cap = 0,
                    uint32_t parts = 1);

// ---- query join against a resident index (hegj_query.hip) ----------------
struct QueryGeom {
    GridGeom g;          // the index level (g.eps = cell side)
    uint64_t dead_key;   // key of queries with no cell in reach (sorts last)
    int key_bits;
};
struct QueryPlan {
    uint64_t *keys = nullptr;    // [nq] sorted widened cell keys
    uint32_t *ids = nullptr;     // [nq] query ids in key order
    uint32_t *starts = nullptr;  // [n_units + 1] first sorted slot of each unit
    uint32_t nq = 0, n_units = 0, n_live = 0;
};
struct QJoinArgs {
    QueryGeom qg;
    DevIndex ix;
    const double *queries;       // [nq * dim] row-major
    const uint64_t *keys;
    const uint32_t *ids;
    const uint32_t *starts;
    uint32_t n_units;
    uint32_t *queue_head;        // zeroed before each launch
    double thresh;               // largest s with sqrt(s) <= eps of the request
    uint32_t *counts;            // [nq] (count pass)
    const uint64_t *offsets;     // [nq + 1] (fill pass)
    uint32_t *out;               // neighbour ids (fill pass)
    unsigned long long *hash;
    unsigned long long *evals;
    // FP32 filter (f32 != 0): centred float candidates and queries
    uint32_t f32;
    const float *soa32;
    const float *aos32;          // point-major copy (dim <= kAosMaxDim)
    const float *q32;            // [nq * dim] row-major, centred like soa32
    float tacc32, trej32;
    unsigned long long *band;    // pairs rechecked in FP64 (count pass)
    // [nq * kQueryHitWords] hit-chunk bitmaps: written by the count pass,
    // read by the fill pass (null: the fill pass visits every chunk)
    uint64_t *hitbits;
};
constexpr uint32_t kQueryHitWords = 64;
hegj_status query_geom(const GridGeom &g, QueryGeom *qg);
// q32 = fl32(q - centre); *rmax_bits = bits of max |q - centre| (a positive
// double, ordered like its bit pattern; must be zeroed by the caller).
hegj_status query_to_f32(const double *q, uint32_t nq, uint32_t dim, const double *centre, float *q32,
                         unsigned long long *rmax_bits, hipStream_t s);
// Accessors of an index handle (hegj_api.cpp).
const GridGeom &index_geom(const hegj_index *h);
const DevIndex &index_dev(const hegj_index *h);
hegj_status select_device(int device);
// Sort the queries by cell and build the units; the plan's arrays come from `keep`.
hegj_status query_plan(QueryPlan *p, Scratch &keep, const QueryGeom &qg, const double *queries,
                       uint32_t nq, hipStream_t s);
hegj_status query_run(const QJoinArgs &a, bool fill, int device, hipStream_t s);
// *hash += sum of mix64(query << 32 | neighbour) over the CSR rows
hegj_status query_hash(const uint64_t *offsets, const uint32_t *nbrs, uint32_t nq, uint64_t total,
                       unsigned long long *hash, hipStream_t s);
// counts[nq] -> offsets[nq + 1] (offsets[nq] = total)
hegj_status query_offsets(const uint32_t *counts, uint32_t nq, uint64_t *offsets, hipStream_t s);

// ---- utilities (hegj_util.hip) -------------------------------------------
hegj_status workload_order(const DevIndex &ix, const uint32_t *work, uint32_t *order,
                           uint64_t *total_candidates, hipStream_t s);
hegj_status origin_point_index(const DevIndex &ix, const uint32_t *order, uint32_t *opi,
                               hipStream_t s);
hegj_status stride_units(const uint2 *in, uint32_t n_out, uint32_t stride, uint2 *out,
                         hipStream_t s);
// Morton (bit-interleaved cell coordinates) order of the cells; false in
// *ok when the coordinates need more than 64 bits.
hegj_status morton_order(const GridGeom &g, const DevIndex &ix, uint32_t *order, bool *ok,
                         hipStream_t s);
// Splits the unit list into n_parts ranges of about equal estimated work
// (queries x stencil candidates of the unit's cell).
hegj_status split_units(const DevIndex &ix, const uint2 *units, uint32_t n_units, uint32_t qb,
                        const uint32_t *work, uint32_t n_parts, uint32_t *part_end,
                        hipStream_t s);
hegj_status gather_units(const uint2 *in, const uint32_t *idx, uint32_t n, uint2 *out,
                         hipStream_t s);
hegj_status csr_offsets(const uint32_t *row_cnt, uint32_t n, uint64_t *csr_off, hipStream_t s);
hegj_status gather_rows(const std::vector<uint32_t *> &batches, const std::vector<uint64_t> &bases,
                        const uint64_t *row_off, const uint32_t *row_cnt, const uint64_t *csr_off,
                        uint32_t n, uint32_t *out, hipStream_t s);

hegj_status stream_copy(uint64_t bytes, int reps, double *ms_per_copy);
// row_off[qid[i]] = base + off[i], row_cnt[qid[i]] = cnt[i] (rows answered
// on the host, added to a device result).
hegj_status scatter_rows(const uint32_t *qid, const uint64_t *off, const uint32_t *cnt, uint32_t m,
                         uint64_t base, uint64_t *row_off, uint32_t *row_cnt, hipStream_t s);

// ---- Super-EGO on the host (hegj_ego.cpp) ---------------------------------
struct EgoIndex;
// Reorders the dimensions and EGO-sorts the points (row-major n x dim).
hegj_status ego_build(const double *pts, uint32_t n, uint32_t dim, double eps, unsigned threads,
                      EgoIndex **out, double *reorder_s, double *sort_s);
void ego_free(EgoIndex *e);
// Appends the neighbours (original ids) of original point q to row; returns
// the candidates scanned.
uint64_t ego_query(const EgoIndex &e, uint32_t q, std::vector<uint32_t> &row);

// ---- device helpers --------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t hash_slot(uint64_t key) {
    return (uint32_t)(mix64(key) >> 16);
}

// Lanes below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Id of the XCD (accelerator complex die) the wave runs on (hwreg XCC_ID).
__device__ __forceinline__ uint32_t xcc_id() {
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;
}

__device__ __forceinline__ uint32_t wave_bcast_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ uint64_t wave_bcast_u64(uint64_t v) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Row hash lookup: returns the row index or ~0u.
__device__ __forceinline__ uint32_t row_lookup(const uint64_t *__restrict__ hkeys,
                                               const uint32_t *__restrict__ hvals,
                                               uint32_t hmask, uint64_t key) {
    uint32_t s = hash_slot(key) & hmask;
    while (true) {
        uint64_t kk = hkeys[s];
        if (kk == key) return hvals[s];
        if (kk == ~0ull) return ~0u;
        s = (s + 1) & hmask;
    }
}

// First index in [lo, hi) with keys[i] >= v.
__device__ __forceinline__ uint32_t lower_bound_u64(const uint64_t *__restrict__ keys,
                                                    uint32_t lo, uint32_t hi, uint64_t v) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Decode the k cell coordinates of a linear id (inverse of S3).
__device__ __forceinline__ void decode_cell(uint64_t key, const GridGeom &g, uint32_t *c) {
    for (uint32_t j = 0; j < g.k; j++) {
        uint64_t q = key / g.nCells[j];
        c[j] = (uint32_t)(key - q * g.nCells[j]);
        key = q;
    }
}

// Stencil bounds of one indexed dim (kernel.cu:780-781).  With the cell-0
// wrap the lower bound of coordinate 0 is UINT_MAX and the range is empty.
__device__ __forceinline__ bool stencil_dim(uint32_t c, uint32_t ncell, uint32_t emulate,
                                            uint32_t *lo, uint32_t *hi) {
    uint32_t l;
    if (c == 0) {
        if (emulate) return false;
        l = 0;
    } else {
        l = c - 1;
    }
    uint32_t h = c + 1 < ncell - 1 ? c + 1 : ncell - 1;
    *lo = l;
    *hi = h;
    return l <= h;
}

// Enumerate the stencil rows of a cell: row r in [0, 3^(k-1)) maps to offsets
// in dims 1..k-1.  Writes the candidate cell range [*cfirst, *clast) of row r
// (cells with coordinate 0 in [lo0, hi0]), or an empty range.
__device__ __forceinline__ void stencil_row(const uint32_t *c, const uint32_t *lo,
                                            const uint32_t *hi, uint32_t r, const GridGeom &g,
                                            const DevIndex &ix, uint32_t *cfirst,
                                            uint32_t *clast) {
    *cfirst = 0;
    *clast = 0;
    uint64_t rkey = 0, mult = 1;
    for (uint32_t j = 1; j < g.k; j++) {
        uint32_t off = r % 3;
        r /= 3;
        uint32_t cj = c[j] + off - 1;  // wraps for c[j] == 0, rejected below
        if (cj < lo[j] || cj > hi[j]) return;
        rkey += (uint64_t)cj * mult;
        mult *= g.nCells[j];
    }
    uint32_t row = row_lookup(ix.hkeys, ix.hvals, ix.hmask, rkey);
    if (row == ~0u) return;
    uint32_t rs = ix.row_start[row], re = ix.row_start[row + 1];
    uint64_t base = rkey * g.nCells[0];
    uint32_t a, b;
    if (re - rs <= 8) {  // short rows (the common case at k >= 4): scan
        a = rs;
        while (a < re && ix.cell_key[a] < base + lo[0]) a++;
        b = a;
        while (b < re && ix.cell_key[b] <= base + hi[0]) b++;
    } else {
        a = lower_bound_u64(ix.cell_key, rs, re, base + lo[0]);
        b = lower_bound_u64(ix.cell_key, a, re, base + hi[0] + 1);
    }
    *cfirst = a;
    *clast = b;
}

__host__ __device__ inline uint32_t pow3(uint32_t e) {
    uint32_t r = 1;
    while (e--) r *= 3;
    return r;
}

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Per-lane cursor over a work unit's LDS slot list (start[s], off[s] for
// s < nslots, off[nslots] = M, offsets strictly increasing): the lane's slot,
// its position base start - off (mod 2^32) and the offset where the next
// slot begins, so a step that stays inside the slot reads no LDS.  Calls of
// pos() take t < M, never decreasing between calls.
struct SlotCursor {
    uint32_t slot, base, next;
    template <typename L>
    __device__ __forceinline__ void reset(const L &rl, uint32_t nslots = 0) {
        (void)nslots;
        slot = 0;
        base = rl.start[0] - rl.off[0];
        next = rl.off[1];
    }
    template <typename L>
    __device__ __forceinline__ uint32_t pos(const L &rl, uint32_t t) {
        while (next <= t) {
            slot++;
            base = rl.start[slot] - rl.off[slot];
            next = rl.off[slot + 1];
        }
        return base + t;
    }
};

// The same walk holding only the slot index (one register instead of three;
// each step reads the next slot's offset and its own start from LDS).
struct SlotWalk {
    uint32_t slot, n;
    template <typename L>
    __device__ __forceinline__ void reset(const L &, uint32_t nslots) {
        slot = 0;
        n = nslots;
    }
    template <typename L>
    __device__ __forceinline__ uint32_t pos(const L &rl, uint32_t t) {
        while (slot + 1 < n && rl.off[slot + 1] <= t) slot++;
        return rl.start[slot] + (t - rl.off[slot]);
    }
};

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace hegj
