/*
 * hegjoin.h -- C ABI of the MI355X-native epsilon-grid self-join
 * (libhegjoin.so).  Plain pointers and sizes only; no HIP or torch types.
 *
 * It replaces the three host entry points of the reference's algorithm 0
 * (the GPU-only mode of `./main dataset epsilon dim algorithm`):
 *
 *   hegj_index_build      <- generateNDGridDimensions  (main.cu:449-520)
 *                          + gridIndexingGPU           (GPU.h:9-27, GPU.cu:74-398)
 *   hegj_sort_by_workload <- sortByWorkLoad            (SortByWorkload.h:7-27,
 *                                                        SortByWorkload.cu:19-247)
 *   hegj_selfjoin         <- distanceTableNDGridBatches (GPU.h:29-54,
 *                                                        GPU.cu:787-1702)
 *                          + GPUBatchEst_v2            (GPU.h:73-88, GPU.cu:403-779)
 *   hegj_result_*         <- constructNeighborTableKeyValueWithPtrs
 *                                                       (GPU.h:90-95, GPU.cu:1731-1776)
 *
 * Semantics kept from the reference (params.h, structs.h):
 *   - GPUNUMDIM   -> hegj_index_opts.dim, NUMINDEXEDDIM -> .indexed_dims,
 *     DTYPE       -> .dtype (HEGJ_DTYPE_F64 is DTYPE=double);
 *   - the grid-cell index layouts: hegj_grid == struct grid,
 *     hegj_grid_cell_lookup == struct gridCellLookup, hegj_scheduling_cell ==
 *     struct schedulingCell (structs.h:83-104), byte for byte; the
 *     neighbour table (struct neighborTableLookup rows over per-batch arrays,
 *     structs.h:35-41) is exported as CSR by point id;
 *   - the pair set: ordered pairs (query, candidate) with
 *     sqrt(fma-chain distance) <= eps over the 3^k cell stencil, self-pairs
 *     included, and (by default) the reference's cell-0 wrap, which silences
 *     queries whose cell has coordinate 0 in an indexed dimension.
 *
 * Conventions: every call returns hegj_status (no exit, no printing);
 * hegj_last_error() gives a message for the calling thread.  The library owns
 * device memory; host buffers are caller-owned.  A handle is used by one
 * thread at a time.  Each handle has its own HIP stream on its device.
 */
#ifndef HEGJOIN_H
#define HEGJOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HEGJ_MAX_INDEXED_DIMS 6
/* Largest point count of an index: each dimension's coordinate array stays
 * under 4 GiB, so the join kernels address it with 32-bit byte offsets. */
#define HEGJ_MAX_POINTS 536870911u

typedef enum {
    HEGJ_OK = 0,
    HEGJ_ERR_INVALID_ARG = 1,
    HEGJ_ERR_UNSUPPORTED = 2,
    HEGJ_ERR_NO_DEVICE = 3,
    HEGJ_ERR_OUT_OF_MEMORY = 4,
    HEGJ_ERR_HIP = 5,
    HEGJ_ERR_CAPACITY = 6,
    HEGJ_ERR_IO = 7,
    HEGJ_ERR_BUSY = 8
} hegj_status;

typedef enum { HEGJ_DTYPE_F64 = 0, HEGJ_DTYPE_F32 = 1 } hegj_dtype;

/* structs.h:83-86 */
typedef struct { int32_t indexmin; int32_t indexmax; } hegj_grid;
/* structs.h:90-98 (16 bytes: idx, 4 bytes padding, gridLinearID) */
typedef struct { uint32_t idx; uint32_t _pad; uint64_t gridLinearID; } hegj_grid_cell_lookup;
/* structs.h:101-104 */
typedef struct { int32_t nbPoints; int32_t cellId; } hegj_scheduling_cell;

/* ------------------------------------------------------------------------ */
const char *hegj_status_string(hegj_status s);
const char *hegj_last_error(void);
/* Library version string. */
const char *hegj_version(void);
/* Number of visible HIP devices (0 when none). */
int hegj_device_count(void);
/* Attainable HBM bandwidth: a streaming 16-byte-vector copy of `bytes`
 * between two device buffers, (read + write bytes) / time, in GB/s.  Not part
 * of the reference; it is the measured roof the join's HBM figures are
 * compared with. */
hegj_status hegj_stream_copy_bandwidth(uint64_t bytes, int device, double *gbps);
/* The self-join keeps the device buffers of finished calls (work units, row
 * tables, result batches) for reuse by later calls, up to 16 GiB per process
 * (an allocation that fails releases them first).  This returns the unused
 * ones on `device` (-1: every device).  No reference counterpart. */
void hegj_release_cached_memory(int device);

/* ------------------------------------------------------------------------ */
/* Datasets.                                                                */
/* ------------------------------------------------------------------------ */
/* The reference generator (datasets/dataset_gen_fixed_len_expo_dist.cpp:57-84):
 * mt19937(seed), exponential(lambda), redraw outside [0, 1]; out[n*dim]
 * row-major.  seed 2137834274 reproduces the reference's own datasets. */
hegj_status hegj_dataset_generate_expo(uint32_t seed, uint32_t n, uint32_t dim, double lambda,
                                       double *out);
/* importNDDataset (import_dataset.cpp:10-40): comma/whitespace separated
 * numbers, n = values / dim; *.bin = raw row-major doubles.  *out is freed
 * with hegj_dataset_free. */
hegj_status hegj_dataset_load(const char *path, uint32_t dim, double **out, uint32_t *n);
/* hegj_dataset_load with options (0 = hegj_dataset_load):
 *   HEGJ_LOAD_F32       numbers are parsed as floats (`ss >> float` with
 *                       DTYPE=float, import_dataset.cpp:17-19), returned as
 *                       the doubles equal to those floats;
 *   HEGJ_LOAD_BIN_CACHE a text file is parsed once: its values go to the
 *                       binary cache <path>.d<dim>[f].hegjbin beside it
 *                       (header with the text file's size and mtime, then
 *                       row-major doubles), which later loads read while the
 *                       text file is unchanged.  Binary datasets
 *                       (import_dataset.cpp:63-112) are the *.bin files. */
#define HEGJ_LOAD_BIN_CACHE 1u
#define HEGJ_LOAD_F32 2u
hegj_status hegj_dataset_load_ex(const char *path, uint32_t dim, uint32_t flags, double **out,
                                 uint32_t *n);
void hegj_dataset_free(double *p);

/* ------------------------------------------------------------------------ */
/* Grid index (generateNDGridDimensions + gridIndexingGPU).                  */
/* ------------------------------------------------------------------------ */
typedef struct hegj_index hegj_index;

typedef struct {
    uint32_t dim;           /* GPUNUMDIM (params.h:5) */
    uint32_t indexed_dims;  /* NUMINDEXEDDIM (params.h:8), 1..HEGJ_MAX_INDEXED_DIMS */
    uint32_t dtype;         /* hegj_dtype; DTYPE (params.h:24).  HEGJ_DTYPE_F32: the
                               database and epsilon are rounded to float and
                               minArr, nCells, cell ids (main.cu:493-510,
                               kernel.cu:54-57) and the distance chain with its
                               sqrtf predicate (kernel.cu:347-355) are computed in
                               FP32, as the reference built with DTYPE=float.
                               The MFMA refine and graph replay are FP64 only;
                               refine FP64 then names the plain (unfiltered) kernel. */
    uint32_t emulate_cell0; /* 1 = reproduce the reference's cell-0 wrap (kernel.cu:780) */
    int32_t device;         /* HIP device ordinal */
} hegj_index_opts;

/* Fills opts with the reference defaults for a given dim (NUMINDEXEDDIM =
 * min(dim, 6), DTYPE double, cell-0 wrap emulated, device 0). */
void hegj_index_opts_default(hegj_index_opts *opts, uint32_t dim);

/* database: host, row-major n x dim doubles (main.cu:120-128 `database`).
 * epsilon must be in ]0, 1] (main.cu:67). */
hegj_status hegj_index_build(const double *database, uint32_t n, double epsilon,
                             const hegj_index_opts *opts, hegj_index **out);
/* Same, from a device pointer on opts->device (data already resident). */
hegj_status hegj_index_build_device(const double *dev_database, uint32_t n, double epsilon,
                                    const hegj_index_opts *opts, hegj_index **out);
hegj_status hegj_index_free(hegj_index *idx);

typedef struct {
    uint32_t n, dim, indexed_dims, dtype;
    uint32_t nNonEmptyCells;
    uint32_t n_rows;            /* cells grouped by indexed dims 1..k-1 */
    uint32_t n_cell0_queries;   /* queries silenced by the cell-0 wrap */
    uint32_t emulate_cell0;
    uint64_t totalCells;
    double epsilon;
    double threshold;           /* largest s with sqrt(s) <= epsilon */
    double minArr[HEGJ_MAX_INDEXED_DIMS];
    double maxArr[HEGJ_MAX_INDEXED_DIMS];
    uint32_t nCells[HEGJ_MAX_INDEXED_DIMS];
    double build_ms;            /* device time of the index build */
} hegj_index_info;
hegj_status hegj_index_get_info(const hegj_index *idx, hegj_index_info *info);

/* Host copies in the reference layouts (GPU.cu:264-324): index[C],
 * gridCellLookupArr[C], indexLookupArr[n].  Any pointer may be NULL. */
hegj_status hegj_index_copy_grid(const hegj_index *idx, hegj_grid *index,
                                 hegj_grid_cell_lookup *gridCellLookupArr,
                                 uint32_t *indexLookupArr);
/* Per-point linear cell ids in original point order (kernel.cu:34-66). */
hegj_status hegj_index_copy_cell_ids(const hegj_index *idx, uint64_t *ids);

/* ------------------------------------------------------------------------ */
/* sortByWorkLoad.                                                          */
/* ------------------------------------------------------------------------ */
/* sortedCells[C] (by candidate count descending, ties by cell id
 * ascending), originPointIndex[n] (the member lists in that order) and
 * totalCandidates (sum |cell| * candidates(cell), 64-bit).  Outputs may be
 * NULL.  Also keeps the order on the device for the join. */
hegj_status hegj_sort_by_workload(hegj_index *idx, hegj_scheduling_cell *sortedCells,
                                  uint32_t *originPointIndex, uint64_t *totalCandidates);

/* ------------------------------------------------------------------------ */
/* Self-join (distanceTableNDGridBatches + neighbour table).                */
/* ------------------------------------------------------------------------ */
typedef struct hegj_result hegj_result;

typedef enum {
    HEGJ_ORDER_SPATIAL = 0,   /* work units in cell (linear id) order */
    HEGJ_ORDER_WORKLOAD = 1,  /* work units by descending cell workload */
    HEGJ_ORDER_MORTON = 2     /* work units in Morton order of cell coordinates */
} hegj_unit_order;

typedef enum {
    HEGJ_REFINE_AUTO = 0,     /* MFMA guard band when dim >= 16, FP32 filter where it is
                                 built (dim 2 and 6) and its band is narrow, else FP64 */
    HEGJ_REFINE_FP64 = 1,     /* exact FP64 chain for every candidate */
    HEGJ_REFINE_MFMA = 2,     /* split-BF16 MFMA tiles + guard band + FP64 band recheck */
    HEGJ_REFINE_F32 = 3       /* FP32 distance filter + exact FP64 recheck of its band */
} hegj_refine;

typedef struct {
    uint64_t result_capacity; /* pairs per device result batch; 0 = estimate.  MFMA
                                 refine: the size of its first pair list (the join's
                                 pairs are held in one list, 8 bytes each; a list
                                 that overflows is rebuilt once at 1.25x the count) */
    uint32_t unit_order;      /* hegj_unit_order (default: Morton) */
    uint32_t queries_per_unit;/* FP64 kernel: 4, 8 or 16; 0 = tuned default */
    uint32_t refine;          /* hegj_refine */
    uint32_t xcd_parts;       /* unit ranges, one per XCD: 0 = auto (8), 1 = one queue */
    uint32_t units_per_grab;  /* units a wave takes per queue atomic; 0 = auto */
    uint32_t part_by_block;   /* 0: range by hardware XCD id, 1: by workgroup index */
    uint32_t use_graph;       /* FP64 refine: keep the plan in the index and run the
                                 batch as a replay of a captured hipGraph */
    uint32_t no_prune;        /* 1: evaluate every point of every stencil cell.  0
                                 (default): skip stencil cells whose box lies farther
                                 than eps from all queries of a work unit, with
                                 rounding margins; the pair set is the same */
    uint32_t block_waves;     /* waves per workgroup of the FP64 and FP32-filter
                                 joins: 1, 2 or 4; 0 = 4.  The reference's BLOCKSIZE
                                 (params.h) is the thread count of its blocks */
    uint32_t batches;         /* result_capacity 0: split the estimated result into
                                 this many device batches (0 or 1 = one, up to 32);
                                 the reference's batch count (GPU.cu) */
} hegj_join_opts;

void hegj_join_opts_default(hegj_join_opts *opts);
/* Defaults for a dimensionality: hegj_join_opts_default, then the tuned
 * entry for `dim` (hegj_tuning_set / hegj_tuning_load) when there is one.
 * hegj_selfjoin with opts == NULL uses these. */
void hegj_join_opts_tuned(hegj_join_opts *opts, uint32_t dim);
/* Register tuned join options for one dimensionality (replaces any entry). */
hegj_status hegj_tuning_set(uint32_t dim, const hegj_join_opts *opts);
/* Load a tuning table written by hegjoin_amd/autotune.py: one line per
 * dimensionality, "dim queries_per_unit unit_order xcd_parts units_per_grab
 * refine [block_waves batches]", '#' starts a comment.  *n_loaded (may be
 * NULL) = entries read. */
hegj_status hegj_tuning_load(const char *path, uint32_t *n_loaded);

hegj_status hegj_selfjoin(hegj_index *idx, const hegj_join_opts *opts, hegj_result **out);
/* Thresholds of the FP32 filter for coordinates within `radius` of their
 * centre: an FP32 fma-chain squared distance <= *tacc is within epsilon, one
 * > *trej is not, by the reference's FP64 chain; pairs in between are
 * rechecked in FP64.  HEGJ_ERR_UNSUPPORTED when the band would exceed 1% of
 * epsilon^2.  (No reference counterpart; exposed for callers and tests.) */
hegj_status hegj_f32_thresholds(double epsilon, uint32_t dim, double radius, float *tacc,
                                float *trej);
/* The operands of the MFMA refine (no reference counterpart; for callers and
 * the band's tests): for n points (row-major n x dim) centred at centre[dim],
 * the split-BF16 rows A'[n * kp] and B'[n * kp] (BF16 bit patterns) and the
 * guard w[n], by the same code the device runs.  coef_scale multiplies the
 * guard coefficient (1 = the library's).  *kp = padded K, *coef = the
 * coefficient used; with A, B and w all NULL only those two are returned.
 * A pair is rejected when its FP32 Gram value s' = A'_q . B'_c > T (rounded
 * up to FP32), accepted when s' + w_q + w_c <= T (1 - 4 * 2^-53), and
 * rechecked in FP64 otherwise.  Host only. */
hegj_status hegj_gb_operands(const double *pts, uint32_t n, uint32_t dim, const double *centre,
                             double coef_scale, uint16_t *A, uint16_t *B, float *w, uint32_t *kp,
                             double *coef);
hegj_status hegj_result_free(hegj_result *res);

typedef struct {
    uint64_t total_neighbors;  /* [RESULT] Total result set size (main.cu:368) */
    uint64_t pair_hash;        /* sum of mix64(query<<32 | neighbour) over all pairs */
    uint64_t total_candidates; /* the reference's candidate count (SortByWorkload.cu:160) */
    uint64_t evals_executed;   /* distance evaluations the kernels performed */
    uint64_t band_rechecks;    /* pairs re-evaluated in FP64 inside the guard band */
    uint32_t n_queries;        /* queries with a stencil (not silenced) */
    uint32_t n_batches;        /* result batches used */
    uint32_t n_units;          /* work units */
    uint32_t n_deferred;       /* units re-run because a batch filled up */
    double estimate_ms;        /* device time: batch estimator */
    double join_ms;            /* device time: join kernel launches */
    double workload_ms;        /* device time: workload sort */
    double total_ms;           /* host wall time of hegj_selfjoin */
    double join_kernel_avg_ms; /* average duration of one join kernel launch */
    uint32_t join_launches;
    uint32_t refine_used;      /* hegj_refine that ran: FP64 or MFMA */
    uint32_t graph_replay;     /* 1: the join ran as a replay of the captured graph */
    uint32_t reserved;
    uint64_t estimated_pairs;  /* the batch estimator's pair estimate ("[GPU | RESULT] ~
                                  Total estimated workload", GPU.cu:635); the MFMA
                                  refine sizes its output exactly, so it reports that */
} hegj_result_info;
hegj_status hegj_result_get_info(const hegj_result *res, hegj_result_info *info);

/* hegj_selfjoin for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It calls no hipMalloc / hipFree: the caller
 * passes the outputs as device buffers on the index's device --
 * neighbors[capacity] (rows back to back), row_off[n] (offset of each point's
 * row in neighbors) and row_cnt[n] (its length), by original point id -- and
 * every temporary comes from a fixed 64 MiB device array per GPU (about
 * 12 bytes per work unit plus 8 per non-empty cell: enough for the 5M-point
 * 6-D workload).  An index that needs more fails with HEGJ_ERR_CAPACITY; so
 * does a result larger than `capacity`, with info->total_neighbors set to
 * the count needed.  The index's lazily built arrays are inputs: run
 * hegj_selfjoin once with the same options first (HEGJ_ERR_INVALID_ARG
 * otherwise).  Grid kernels only (refine MFMA is refused).  The array is
 * shared by every call on its GPU and keeps the last call's contents; calls
 * on one GPU are serialised by a lock inside the library (two processes
 * each have their own array). */
hegj_status hegj_selfjoin_noalloc(hegj_index *idx, const hegj_join_opts *opts, uint32_t *neighbors,
                                  uint64_t capacity, uint64_t *row_off, uint32_t *row_cnt,
                                  hegj_result_info *info);

/* ------------------------------------------------------------------------
 * Hybrid CPU + GPU self-join: algorithms 1, 2 and 3 of
 * `./main dataset epsilon dim algorithm [static_partitioning]`
 * (distanceTableNDGridBatches' SM_HYBRID branch GPU.cu:1179-1363, WorkQueue.cpp
 * :31-118, MultiThreadJoin.cpp:21-115, main.cu:131-187, 255-362).
 *
 * The queries, in the workload order of hegj_sort_by_workload
 * (originPointIndex), form one queue of work units (up to queries_per_unit
 * queries of one cell).  The GPU takes chunks of units from the head of the
 * queue and runs each through the persistent join kernel into the result's
 * device batches; CPU threads take batches of about `cpu_batch` queries from
 * the tail (getBatchFromQueueCPU, WorkQueue.cpp:97-118) and answer each query
 * with Super-EGO over the whole dataset (Util::egoJoinV2).  The queue's head
 * and tail are host-side indices under one lock; the GPU never waits for the
 * CPU nor the reverse.  A query answered on the GPU gets its algorithm-0 row,
 * one answered on the CPU its Super-EGO row (no cell-0 wrap, Super-EGO's own
 * rounding), as in the reference.  The CPU rows are copied into the result,
 * so every hegj_result_* call sees one table.
 *   HEGJ_MODE_HYBRID  the split is dynamic (whoever reaches a unit first);
 *   HEGJ_MODE_STATIC  the GPU takes the first gpu_fraction of the queries (at
 *                     a unit boundary), the CPU the rest; gpu_fraction <= 0
 *                     uses hegj_static_partition_model;
 *   HEGJ_MODE_CPU     every query on the CPU (algorithm 3).
 * ------------------------------------------------------------------------ */
typedef enum {
    HEGJ_MODE_GPU = 0,     /* SM_GPU (params.h:58): hegj_selfjoin */
    HEGJ_MODE_HYBRID = 1,  /* SM_HYBRID */
    HEGJ_MODE_STATIC = 2,  /* SM_HYBRID_STATIC */
    HEGJ_MODE_CPU = 3      /* SM_CPU */
} hegj_search_mode;

typedef struct {
    uint32_t mode;          /* hegj_search_mode */
    uint32_t cpu_threads;   /* Super-EGO threads (CPU_THREADS, params.h:14); 0 = every CPU */
    uint32_t cpu_batch;     /* queries per CPU batch (CPU_BATCH_SIZE, params.h:17); 0 = 1024 */
    uint32_t gpu_chunk;     /* work units per GPU launch in HEGJ_MODE_HYBRID; 0 = 1/16 of them */
    double gpu_fraction;    /* HEGJ_MODE_STATIC: the GPU's share of the queries, in ]0, 1[ */
} hegj_hybrid_opts;

typedef struct {
    uint32_t mode;
    uint32_t cpu_threads;
    uint32_t n_queries_gpu;   /* "[GPU | RESULT] ~ Query points computed by the GPU" */
    uint32_t n_queries_cpu;   /* "[EGO | RESULT] ~ Query points computed by Super-EGO" */
    uint64_t pairs_gpu;       /* "[RESULT] ~ Total result set size on the GPU" */
    uint64_t pairs_cpu;       /* "[RESULT] ~ Total result set size on the CPU" */
    uint64_t candidates_cpu;  /* points Super-EGO scanned */
    double gpu_fraction;      /* HEGJ_MODE_STATIC: the split used */
    double gpu_s;             /* GPU side: first launch -> last batch done */
    double cpu_s;             /* CPU side: Super-EGO queries */
    double reorder_s, sort_s; /* CPU side: dimension reorder, EGO sort */
    double gpu_end_s, cpu_end_s; /* each side's end, from the start of the call */
} hegj_hybrid_info;

/* Defaults for a mode: every CPU thread, batches of 1024, automatic chunks,
 * gpu_fraction 0 (the model). */
void hegj_hybrid_opts_default(hegj_hybrid_opts *opts, uint32_t mode);
/* opts as for hegj_selfjoin (NULL: tuned defaults; the MFMA refine is not
 * used: the GPU runs the grid join kernels).  HEGJ_MODE_GPU is
 * hegj_selfjoin. */
hegj_status hegj_selfjoin_hybrid(hegj_index *idx, const hegj_join_opts *opts,
                                 const hegj_hybrid_opts *hopts, hegj_result **out);
/* The CPU/GPU figures of a hybrid result (HEGJ_ERR_INVALID_ARG for others). */
hegj_status hegj_result_get_hybrid_info(const hegj_result *res, hegj_hybrid_info *info);

/* Super-EGO on the host alone: the CPU share of the hybrid modes
 * (Util::reorderDim, the EGO sort of main.cu:296-318 and Util::egoJoinV2 per
 * query, MultiThreadJoin.cpp:43-115), for `threads` threads (0 = every CPU).
 * Rows of the queries qids[0..nq) (qids NULL: every point, nq = n) over all n
 * points: counts[nq], *nbrs = the rows back to back (each in EGO order;
 * freed with hegj_free), *total = pairs.  Needs no device. */
hegj_status hegj_superego_rows(const double *database, uint32_t n, uint32_t dim, double epsilon,
                               uint32_t threads, const uint32_t *qids, uint32_t nq,
                               uint32_t *counts, uint32_t **nbrs, uint64_t *total);
void hegj_free(void *p);

/* Static split model (StaticPartition.cpp:67-108, main.cu:141-168): modelled
 * times t = exp(c0 + c1 ln(dim) + c2 ln(n) + c3 ln(eps)) seconds for the GPU
 * and the CPU; the GPU share is (n / t_gpu) / (n / t_gpu + n / t_cpu), each
 * rate truncated to an integer as main.cu:152-153 does.  The coefficients
 * are the reference's (fitted on its A100 + Zen 2 host) until replaced with
 * hegj_static_partition_set_model. */
double hegj_static_partition_model(uint32_t n, uint32_t dim, double epsilon);
/* gpu[4], cpu[4] = (c0, c1, c2, c3) of the two time models. */
void hegj_static_partition_set_model(const double *gpu, const double *cpu);

/* Per-query neighbour counts, original point ids (neighborTable row sizes). */
hegj_status hegj_result_copy_counts(const hegj_result *res, uint32_t *counts);
/* The neighbour table as CSR by original point id: offsets[n+1],
 * neighbors[total].  Within a row neighbours come in cell order; rows are
 * sorted ascending when sort_rows != 0. */
hegj_status hegj_result_copy_csr(const hegj_result *res, uint64_t *offsets, uint32_t *neighbors,
                                 int sort_rows);

/* ------------------------------------------------------------------------
 * Serving: a device-resident index answering range-join requests.
 *
 * No counterpart in the reference (its only join is the self-join of
 * main.cu:255-283).  A request is a set of query points and an epsilon; the
 * answer is, for each query, every index point p with sqrt(acc) <= eps,
 * acc = fma(p_l - q_l, p_l - q_l, acc) in dimension order (the reference's
 * chain, kernel.cu:350-355), as CSR rows of original point ids.
 *
 * The index is kept at several grid levels (cell sides); a request runs on
 * the smallest cell side c with eps <= c * (1 - 2^-30).  `slots` requests are
 * on the device at once, each on its own stream: requests that arrive while
 * others run start as soon as a slot frees up (continuous batching).  Each
 * request first counts its result exactly, then is admitted into the
 * result-buffer pool of `result_bytes` when its result fits (requests that
 * wait for memory are served first come, first served; a request whose
 * count finishes early does not wait for earlier ones still counting).
 * Results stay in the pool until released, so a client should release
 * finished results (hegj_serve_poll) rather than hold them while waiting for
 * another request.  A request whose result alone is larger than the pool
 * fails with HEGJ_ERR_CAPACITY.
 * ------------------------------------------------------------------------ */
#define HEGJ_SERVE_MAX_LEVELS 32
typedef struct hegj_server hegj_server;

typedef struct {
    uint32_t dim;              /* point dimensionality */
    uint32_t indexed_dims;     /* grid dimensions k (1..6) */
    uint32_t n_levels;         /* grid levels, 1..HEGJ_SERVE_MAX_LEVELS */
    uint32_t slots;            /* requests on the device at once, 1..16 */
    double cell_sides[HEGJ_SERVE_MAX_LEVELS]; /* cell side of each level */
    uint64_t result_bytes;     /* result-buffer pool (admission budget) */
    uint32_t max_queries;      /* largest request */
    int32_t device;
} hegj_serve_opts;

typedef struct {
    uint64_t ticket;
    int32_t status;            /* hegj_status of the request */
    uint32_t level;            /* grid level used */
    uint32_t n_queries;
    uint32_t refine_used;      /* HEGJ_REFINE_F32 (dim 2, 6, narrow band) or HEGJ_REFINE_FP64 */
    uint64_t pairs;            /* result size */
    uint64_t pair_hash;        /* sum of mix64(query<<32 | point) over pairs */
    uint64_t evals;            /* distance evaluations */
    uint64_t result_bytes;     /* pool bytes held by the result */
    uint64_t band_rechecks;    /* pairs decided by the FP64 recheck (FP32 filter) */
    double queue_ms;           /* submit -> on a slot */
    double count_ms;           /* on a slot -> result size known */
    double admit_ms;           /* waiting for pool memory */
    double fill_ms;            /* admitted -> result written */
    double latency_ms;         /* submit -> result ready */
} hegj_request_info;

typedef struct {
    uint64_t submitted, completed, failed;
    uint64_t pairs;            /* over completed requests */
    double p50_ms, p99_ms, max_ms, mean_ms; /* latency over completed requests */
    uint64_t pool_bytes, pool_in_use, pool_peak;
    double wall_ms;            /* first submit -> last completion */
} hegj_serve_stats;

/* Upload `database` (n x dim, row-major) once and build every level. */
hegj_status hegj_serve_open(const double *database, uint32_t n, const hegj_serve_opts *opts,
                            hegj_server **out);
/* Queue a request; `queries` (nq x dim) is copied before the call returns. */
hegj_status hegj_serve_submit(hegj_server *srv, const double *queries, uint32_t nq, double epsilon,
                              uint64_t *ticket);
/* Block until the request has finished (or failed); fills *info. */
hegj_status hegj_serve_wait(hegj_server *srv, uint64_t ticket, hegj_request_info *info);
/* Non-blocking: HEGJ_OK and *info when the request has finished,
 * HEGJ_ERR_BUSY while it runs, its status when it failed. */
hegj_status hegj_serve_poll(hegj_server *srv, uint64_t ticket, hegj_request_info *info);
/* Copy a finished request's CSR: offsets[nq+1], neighbors[pairs]. */
hegj_status hegj_serve_copy(hegj_server *srv, uint64_t ticket, uint64_t *offsets, uint32_t *neighbors);
/* Return the request's pool memory and forget it. */
hegj_status hegj_serve_release(hegj_server *srv, uint64_t ticket);
/* Latency percentiles and pool use; reset != 0 starts a new window. */
hegj_status hegj_serve_get_stats(hegj_server *srv, hegj_serve_stats *stats, int reset);
/* Finish queued work, stop the slots and free everything. */
hegj_status hegj_serve_close(hegj_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* HEGJOIN_H */
