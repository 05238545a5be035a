/*
 * ref_config.h -- compile-time configuration for building the reference's
 * pure-C++ sources (Util.cpp, Point.cpp, MultiThreadJoin.cpp, WorkQueue.cpp,
 * import_dataset.cpp) into oracle/_ref.
 *
 * The reference is configured by editing params.h (README.md:22-31: GPUNUMDIM
 * must match the dataset, CPU_THREADS sets the Super-EGO thread count).  This
 * header is force-included (g++ -include) before each source: it includes the
 * reference's own params.h once -- its PARAMS_H guard turns every later
 * #include "params.h" into a no-op -- and then sets the two values a run needs:
 * GPUNUMDIM from -DREF_GPUNUMDIM and CPU_THREADS at run time (below).  No
 * reference source is copied or replaced.
 */
#include "params.h"
#include <cstdlib>
#include <omp.h>
#undef GPUNUMDIM
#define GPUNUMDIM REF_GPUNUMDIM
// The Super-EGO thread count is read when a run starts rather than fixed at
// compile time, so a prebuilt harness uses the cores of the host it runs on:
// $HEGJ_REF_THREADS when set, else omp_get_max_threads() (OMP_NUM_THREADS,
// or the CPUs the process may use).  CPU_THREADS is only used as a count
// (MultiThreadJoin.cpp:33-36, 45, 118, 209), so a run-time value is valid.
static inline int hegj_ref_cpu_threads() {
    const char *e = std::getenv("HEGJ_REF_THREADS");
    const int t = e ? std::atoi(e) : 0;
    return t > 0 ? t : omp_get_max_threads();
}
#undef CPU_THREADS
#define CPU_THREADS hegj_ref_cpu_threads()
