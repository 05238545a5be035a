// main.cpp -- `./main dataset epsilon dimensionality algorithm [static_partitioning]`
//
// Drop-in for the reference CLI (main.cu:42-446, README.md:57-74): algorithm
// 0 (GPU only, hegj_selfjoin), 1 (hybrid: the GPU takes the work queue's head,
// Super-EGO CPU threads its tail), 2 (static split: the fifth argument, or the
// model of StaticPartition.cpp) and 3 (Super-EGO alone), through
// hegj_selfjoin_hybrid.  Argument checks and exit codes follow main.cu:47-71
// and 131-140; the [RESULT] lines follow main.cu:366-407, GPU.cu:635,
// 1622-1628, MultiThreadJoin.cpp:234-236 and SortByWorkload.cu:220-221 so
// `| grep RESULT` pipelines keep working.
//
// Compile-time parameters keep the params.h semantics: -DGPUNUMDIM=<d>
// makes the CLI insist that dim == GPUNUMDIM (main.cu:61-65);
// -DNUMINDEXEDDIM=<k> sets the indexed dimensions (default 6, capped at dim).
// Extra flags after the positional arguments:
//   --indexed-dims=K    override NUMINDEXEDDIM at run time
//   --no-cell0-wrap     do not reproduce the reference's cell-0 wrap
//   --dump-csr=PATH     write the neighbour table (rows sorted) as binary:
//                       uint32 n, uint64 offsets[n+1], uint32 neighbours[]
//   --device=D          HIP device ordinal
//   --tuning=PATH       join options per dimensionality (autotuner table);
//                       by default tuning.tsv next to this executable is used
//                       when present
//   --cpu-threads=T     Super-EGO threads (CPU_THREADS) of algorithms 1-3
//   --bin-cache         keep a binary copy of a parsed text dataset beside it
//                       (<dataset>.d<dim>.hegjbin) and read that copy while
//                       the text file is unchanged
#include <cerrno>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "hegjoin.h"

#ifndef NUMINDEXEDDIM
#define NUMINDEXEDDIM 6
#endif
// DTYPE (params.h:24): -DDTYPE=float builds the CLI for float data, as the
// reference's PARAMS hook does; --dtype=float|double overrides at run time.
#ifndef DTYPE
#define DTYPE double
#endif
// CPU_THREADS (params.h:14): Super-EGO threads of algorithms 1-3; 0 = every
// CPU.  --cpu-threads=T overrides at run time.
#ifndef CPU_THREADS
#define CPU_THREADS 0
#endif

#define NB_ARGS_MAX 6

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv) {
    double tStartStart = now_s();
    std::printf("\n\n========== HEGJoin (MI355X) ==========\n\n\n");
    std::vector<char *> pos;
    uint32_t kOverride = 0;
    int emulate = 1, device = 0;
    const char *dump = nullptr, *tuning = nullptr;
    uint32_t load_flags = 0;
    bool use_float = sizeof(DTYPE) == sizeof(float);
    uint32_t cpu_threads = CPU_THREADS;
    for (int i = 1; i < argc; i++) {
        if (std::strncmp(argv[i], "--indexed-dims=", 15) == 0)
            kOverride = (uint32_t)std::atoi(argv[i] + 15);
        else if (std::strcmp(argv[i], "--no-cell0-wrap") == 0)
            emulate = 0;
        else if (std::strncmp(argv[i], "--dump-csr=", 11) == 0)
            dump = argv[i] + 11;
        else if (std::strncmp(argv[i], "--device=", 9) == 0)
            device = std::atoi(argv[i] + 9);
        else if (std::strncmp(argv[i], "--tuning=", 9) == 0)
            tuning = argv[i] + 9;
        else if (std::strcmp(argv[i], "--bin-cache") == 0)
            load_flags |= HEGJ_LOAD_BIN_CACHE;
        else if (std::strncmp(argv[i], "--cpu-threads=", 14) == 0)
            cpu_threads = (uint32_t)std::atoi(argv[i] + 14);
        else if (std::strcmp(argv[i], "--dtype=float") == 0)
            use_float = true;
        else if (std::strcmp(argv[i], "--dtype=double") == 0)
            use_float = false;
        else
            pos.push_back(argv[i]);
    }
    {
        std::string path = tuning ? tuning : "";
        if (path.empty()) {  // tuning.tsv beside the executable
            std::string self = argv[0];
            const size_t slash = self.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : self.substr(0, slash)) + "/tuning.tsv";
            FILE *f = std::fopen(path.c_str(), "r");
            if (f)
                std::fclose(f);
            else
                path.clear();
        }
        uint32_t loaded = 0;
        if (!path.empty() && hegj_tuning_load(path.c_str(), &loaded) != HEGJ_OK) {
            std::fprintf(stderr, "[MAIN] ~ Error: tuning table %s: %s\n", path.c_str(), hegj_last_error());
            return 1;
        }
    }
    int nargs = (int)pos.size() + 1;
    if (NB_ARGS_MAX != nargs && (NB_ARGS_MAX - 1) != nargs) {
        std::fprintf(stderr, "[MAIN] ~ Expected %d or %d args, found %d\n", NB_ARGS_MAX, NB_ARGS_MAX - 1, nargs);
        std::fprintf(stderr, "[MAIN] ~ Args: filename epsilon dim searchmode [CPU-GPU partitioning]\n");
        return 1;
    }
    const char *filename = pos[0];
    double epsilon = std::atof(pos[1]);
    int dim = std::atoi(pos[2]);
    int searchMode = std::atoi(pos[3]);
#ifdef GPUNUMDIM
    if (GPUNUMDIM != dim) {
        std::fprintf(stderr, "[MAIN] ~ Error: dim should be equals to GPUNUMDIM (see params.h)\n");
        return 1;
    }
#endif
    if (dim <= 0) {
        std::fprintf(stderr, "[MAIN] ~ Error: dim must be positive\n");
        return 1;
    }
    if (epsilon <= 0.0 || 1.0 < epsilon) {
        std::fprintf(stderr, "[MAIN] ~ Error: epsilon should be between ]0.0, 1.0]\n");
        return 1;
    }
    std::printf("[MAIN] ~ Dataset: %s\n", filename);
    std::printf("[MAIN] ~ Epsilon: %f\n", epsilon);
    std::printf("[MAIN] ~ Dimensionality: %d\n", dim);
    std::printf("[MAIN] ~ Search mode: %d\n", searchMode);
    if (searchMode < 0 || searchMode > 3) {
        std::fprintf(stderr, "[MAIN] ~ Error: algorithm must be 0 (GPU), 1 (hybrid), 2 (static hybrid) or 3 (CPU)\n");
        return 1;
    }
    // the static partition (main.cu:131-168): from the command line, or the model
    float staticPartition = -1.0f;
    if (searchMode == 2 && nargs == NB_ARGS_MAX) {
        staticPartition = (float)std::atof(pos[4]);
        if (staticPartition <= 0.0f || 1.0f <= staticPartition) {
            std::fprintf(stderr, "[MAIN] ~ Error: the partitioning should be between ]0.0, 1.0[");
            return 1;
        }
        std::printf("[MODEL | RESULT] ~ Statically partitioning the work, using the partitioning value %f\n",
                    staticPartition);
    }

    if (use_float) {
        load_flags |= HEGJ_LOAD_F32;
        epsilon = (double)(float)epsilon;  // DTYPE epsilon = atof(...) (main.cu:56)
    }
    double t0 = now_s();
    double *data = nullptr;
    uint32_t n = 0;
    if (hegj_dataset_load_ex(filename, (uint32_t)dim, load_flags, &data, &n) != HEGJ_OK) {
        std::fprintf(stderr, "[MAIN] ~ Error: %s\n", hegj_last_error());
        return 1;
    }
    std::printf("[MAIN] ~ Time to read the dataset: %f\n", now_s() - t0);
    std::printf("Data import: Total data points: %u\n", n);
    if (n == 0) {
        std::fprintf(stderr, "[MAIN] ~ Error: empty dataset\n");
        return 1;
    }

    hegj_index_opts o;
    hegj_index_opts_default(&o, (uint32_t)dim);
    uint32_t k = kOverride ? kOverride : (uint32_t)NUMINDEXEDDIM;
    o.indexed_dims = k < (uint32_t)dim ? k : (uint32_t)dim;
    o.emulate_cell0 = (uint32_t)emulate;
    o.device = device;
    o.dtype = use_float ? HEGJ_DTYPE_F32 : HEGJ_DTYPE_F64;

    double tIndex = now_s();
    hegj_index *idx = nullptr;
    hegj_status st = hegj_index_build(data, n, epsilon, &o, &idx);
    hegj_dataset_free(data);
    if (st != HEGJ_OK) {
        std::fprintf(stderr, "[INDEX] ~ Error: %s: %s\n", hegj_status_string(st), hegj_last_error());
        return 1;
    }
    hegj_index_info ii;
    hegj_index_get_info(idx, &ii);
    std::printf("[MAIN] ~ Total cells (including empty): %lu\n", (unsigned long)ii.totalCells);
    std::printf("[INDEX] ~ Number of full cells (non-empty): %u\n", ii.nNonEmptyCells);
    std::printf("[INDEX] ~ Queries silenced by the cell-0 wrap: %u\n", ii.n_cell0_queries);
    std::printf("[INDEX] ~ Time to index using the GPU (including allocating and transfering memory): %f\n",
                now_s() - tIndex);

    double tStartSort = now_s();
    uint64_t totalCandidates = 0;
    std::vector<hegj_scheduling_cell> cells(ii.nNonEmptyCells);
    st = hegj_sort_by_workload(idx, cells.data(), nullptr, &totalCandidates);
    if (st != HEGJ_OK) {
        std::fprintf(stderr, "[SORT] ~ Error: %s: %s\n", hegj_status_string(st), hegj_last_error());
        return 1;
    }
    double sortTime = now_s() - tStartSort;
    if (!cells.empty())
        std::printf("[SORT | RESULT] ~ Total number of candidate points to refine: %llu\n"
                    "[SORT | RESULT] ~ Number of candidates: min = %d, median = %d, max = %d, avg = %llu\n",
                    (unsigned long long)totalCandidates, cells.back().nbPoints,
                    cells[cells.size() / 2].nbPoints, cells.front().nbPoints,
                    (unsigned long long)(totalCandidates / n));
    std::printf("\n\n[MAIN] ~ Time to do everything before computing: %f\n\n\n", now_s() - tStartStart);

    if (searchMode == 2 && staticPartition < 0.0f) {
        staticPartition = (float)hegj_static_partition_model(n, (uint32_t)dim, epsilon);
        std::printf("[MODEL | RESULT] ~ Modeled GPU partition: %f, CPU partition: %f\n", staticPartition,
                    1 - staticPartition);
    }
    double tStart = now_s();
    hegj_result *res = nullptr;
    if (searchMode == 0) {
        st = hegj_selfjoin(idx, nullptr, &res);
    } else {
        hegj_hybrid_opts ho;
        hegj_hybrid_opts_default(&ho, (uint32_t)searchMode);
        ho.cpu_threads = cpu_threads;
        ho.gpu_fraction = staticPartition;
        st = hegj_selfjoin_hybrid(idx, nullptr, &ho, &res);
    }
    if (st != HEGJ_OK) {
        std::fprintf(stderr, "[GPU] ~ Error: %s: %s\n", hegj_status_string(st), hegj_last_error());
        return 1;
    }
    const double computeTime = now_s() - tStart;
    hegj_result_info ri;
    hegj_result_get_info(res, &ri);
    // the GPU / CPU split: everything on the GPU in algorithm 0, where the CPU
    // takes no work and so ends when the join starts
    hegj_hybrid_info hi{};
    if (searchMode == 0) {
        hi.n_queries_gpu = n;
        hi.pairs_gpu = ri.total_neighbors;
        hi.gpu_s = hi.gpu_end_s = computeTime;
    } else {
        hegj_result_get_hybrid_info(res, &hi);
    }
    const double gpuTime = searchMode == 3 ? 0.0 : hi.gpu_s;
    if (searchMode != 3) {
        std::printf("[GPU | RESULT] ~ Total estimated workload: %llu\n", (unsigned long long)ri.estimated_pairs);
        std::printf("[GPU] ~ Kernel execution time: %f ms over %u launch(es), %u batch(es)\n", ri.join_ms,
                    ri.join_launches, ri.n_batches);
        std::printf("[GPU | RESULT] ~ Query points computed by the GPU: %u (f: %f)\n", hi.n_queries_gpu,
                    (double)hi.n_queries_gpu / (double)n);
        std::printf("[GPU | RESULT] ~ Compute time for the GPU: %f\n", gpuTime);
    }
    if (searchMode != 0) {
        // MultiThreadJoin.cpp:234-236
        std::printf("[EGO | RESULT] ~ Query points computed by Super-EGO: %u\n", hi.n_queries_cpu);
        std::printf("[EGO | RESULT] ~ Candidate points refined by Super-EGPU: %llu\n",
                    (unsigned long long)hi.candidates_cpu);
        std::printf("[EGO | RESULT] ~ Compute time for Super-EGO: %f\n", hi.cpu_s);
    }

    std::printf("[RESULT] ~ Total result set size: %lu\n", (unsigned long)ri.total_neighbors);
    std::printf("   [RESULT] ~ Total result set size on the GPU: %lu\n", (unsigned long)hi.pairs_gpu);
    std::printf("   [RESULT] ~ Total result set size on the CPU: %lu\n", (unsigned long)hi.pairs_cpu);
    std::printf("[RESULT] ~ Total execution time: %f\n", computeTime + sortTime);
    std::printf("   [RESULT] ~ Total execution time to SortByWL: %f\n", sortTime);
    std::printf("   [RESULT] ~ Total execution time for the GPU: %f\n", gpuTime);
    std::printf("   [RESULT] ~ Total execution time for the CPU: %f (reorder: %f, sort: %f, total = %f)\n",
                hi.cpu_s, hi.reorder_s, hi.sort_s, hi.cpu_s + hi.reorder_s + hi.sort_s);
    // main.cu:400-407
    const double total = computeTime + sortTime;
    const double gpu_end = searchMode == 3 ? 0.0 : hi.gpu_end_s, cpu_end = hi.cpu_end_s;
    if (gpu_end < cpu_end)
        std::printf("[RESULT] ~ The GPU ended before the CPU, with a difference of: %f (ratio: %f)\n",
                    cpu_end - gpu_end, total > 0 ? (cpu_end - gpu_end) / total : 0.0);
    else
        std::printf("[RESULT] ~ The CPU ended before the GPU, with a difference of: %f (ratio: %f)\n",
                    gpu_end - cpu_end, total > 0 ? (gpu_end - cpu_end) / total : 0.0);
    std::printf("[RESULT] ~ Pair multiset hash: %016llx\n", (unsigned long long)ri.pair_hash);

    int rc = 0;
    if (dump) {
        std::vector<uint64_t> off(n + 1);
        std::vector<uint32_t> nb(ri.total_neighbors);
        st = hegj_result_copy_csr(res, off.data(), nb.data(), 1);
        FILE *f = st == HEGJ_OK ? std::fopen(dump, "wb") : nullptr;
        if (!f) {
            std::fprintf(stderr, "[MAIN] ~ Error: cannot write %s\n", dump);
            rc = 1;
        } else {
            std::fwrite(&n, 4, 1, f);
            std::fwrite(off.data(), 8, off.size(), f);
            std::fwrite(nb.data(), 4, nb.size(), f);
            std::fclose(f);
        }
    }
    hegj_result_free(res);
    hegj_index_free(idx);
    return rc;
}
