/*
 * ref_config.h -- compile-time configuration for building the reference's
 * pure-C++ sources (Util.cpp, Point.cpp, MultiThreadJoin.cpp, WorkQueue.cpp,
 * import_dataset.cpp) into oracle/_ref.
 *
 * The reference is configured by editing params.h (README.md:22-31: GPUNUMDIM
 * must match the dataset, CPU_THREADS sets the Super-EGO thread count).  This
 * header is force-included (g++ -include) before each source: it includes the
 * reference's own params.h once -- its PARAMS_H guard turns every later
 * #include "params.h" into a no-op -- and then sets the two values a run needs
 * from -DREF_GPUNUMDIM / -DREF_CPU_THREADS.  No reference source is copied or
 * replaced.
 */
#include "params.h"
#undef GPUNUMDIM
#define GPUNUMDIM REF_GPUNUMDIM
#undef CPU_THREADS
#define CPU_THREADS REF_CPU_THREADS
