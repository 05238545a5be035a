// hegj_index.hip -- grid index construction on the device.
//
// Replaces gridIndexingGPU (GPU.cu:74-398): cell ids (kernel.cu:34-66), the
// Thrust sort / unique / sort_by_key (GPU.cu:249-285) and the host
// run-boundary loop (GPU.cu:310-324).  Everything stays on the device: a
// stable LSD radix sort of (linear id, point id), run flags + scan for the
// non-empty cells, a second flag + scan for rows, an open-addressing hash
// table of rows, and the coordinates gathered into cell order as SoA.
#include "hegj_internal.h"

namespace hegj {

namespace {

constexpr int kT = 256;

// Block min/max of the first k coordinates; partials[b*2k + 2j + {0,1}].
__global__ __launch_bounds__(kT) void k_minmax(const double *__restrict__ pts, uint32_t n,
                                               uint32_t dim, uint32_t k,
                                               double *__restrict__ partials) {
    __shared__ double smin[kT], smax[kT];
    for (uint32_t j = 0; j < k; j++) {
        double mn = __builtin_inf(), mx = -__builtin_inf();
        for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < n;
             i += (uint64_t)gridDim.x * kT) {
            double v = pts[i * dim + j];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
        smin[threadIdx.x] = mn;
        smax[threadIdx.x] = mx;
        __syncthreads();
        for (int o = kT / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                double a = smin[threadIdx.x + o], b = smax[threadIdx.x + o];
                if (a < smin[threadIdx.x]) smin[threadIdx.x] = a;
                if (b > smax[threadIdx.x]) smax[threadIdx.x] = b;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            partials[(uint64_t)blockIdx.x * 2 * k + 2 * j] = smin[0];
            partials[(uint64_t)blockIdx.x * 2 * k + 2 * j + 1] = smax[0];
        }
        __syncthreads();
    }
}

// S2 + S3 (kernel.cu:54-60): FP64 subtract, IEEE FP64 divide, truncation to
// unsigned, linear id with dim 0 fastest.  Also the identity permutation.
__global__ __launch_bounds__(kT) void k_cell_ids(const double *__restrict__ pts, GridGeom g,
                                                 uint64_t *__restrict__ keys,
                                                 uint64_t *__restrict__ point_key,
                                                 uint32_t *__restrict__ vals,
                                                 uint32_t *__restrict__ dead_count) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    bool dead = false;
    if (i < g.n) {
        uint64_t key = 0, mult = 1;
        for (uint32_t j = 0; j < g.k; j++) {
            const double x = pts[(uint64_t)i * g.dim + j];
            // DTYPE=float: the same expression in FP32 (its operands hold floats)
            uint32_t c = g.f32dt ? (uint32_t)(((float)x - (float)g.minArr[j]) / (float)g.eps)
                                 : (uint32_t)((x - g.minArr[j]) / g.eps);
            dead |= (c == 0);
            key += (uint64_t)c * mult;
            mult *= g.nCells[j];
        }
        keys[i] = key;
        point_key[i] = key;
        vals[i] = i;
    }
    uint64_t m = __ballot(dead && g.emulate_cell0);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(dead_count, (uint32_t)__popcll(m));
}

// flag[i] = 1 where a new cell (or, with div > 1, a new row) starts.
__global__ __launch_bounds__(kT) void k_run_flags(const uint64_t *__restrict__ keys, uint32_t n,
                                                  uint64_t div, uint32_t *__restrict__ flags) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || keys[i] / div != keys[i - 1] / div) ? 1u : 0u;
}

// Compacts run starts: start[rank] = i, key_out[rank] = keys[i] / div.
__global__ __launch_bounds__(kT) void k_run_compact(const uint64_t *__restrict__ keys,
                                                    uint32_t n, uint64_t div,
                                                    const uint32_t *__restrict__ flags,
                                                    const uint32_t *__restrict__ rank,
                                                    uint32_t *__restrict__ start,
                                                    uint64_t *__restrict__ key_out) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < n && flags[i]) {
        start[rank[i]] = i;
        key_out[rank[i]] = keys[i] / div;
    }
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

__global__ __launch_bounds__(kT) void k_hash_fill(uint64_t *__restrict__ hkeys, uint32_t size) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < size) hkeys[i] = ~0ull;
}

__global__ __launch_bounds__(kT) void k_hash_insert(const uint64_t *__restrict__ row_key,
                                                    uint32_t R, uint64_t *__restrict__ hkeys,
                                                    uint32_t *__restrict__ hvals,
                                                    uint32_t hmask) {
    uint32_t r = blockIdx.x * kT + threadIdx.x;
    if (r >= R) return;
    uint64_t key = row_key[r];
    uint32_t s = hash_slot(key) & hmask;
    while (true) {
        unsigned long long prev = atomicCAS((unsigned long long *)&hkeys[s], ~0ull,
                                            (unsigned long long)key);
        if (prev == ~0ull || prev == key) {
            hvals[s] = r;
            return;
        }
        s = (s + 1) & hmask;
    }
}

// DTYPE=float: the database rounded to float in place (main.cu:120-128 copies
// the parsed values into a DTYPE array).
__global__ __launch_bounds__(kT) void k_round_f32(double *__restrict__ pts, uint64_t total) {
    const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (i < total) pts[i] = (double)(float)pts[i];
}

// soa[l*n + p] = pts[ila[p]*dim + l]: coordinates in cell order, SoA.
__global__ __launch_bounds__(kT) void k_gather_soa(const double *__restrict__ pts,
                                                   const uint32_t *__restrict__ ila, uint32_t n,
                                                   uint32_t dim, double *__restrict__ soa) {
    uint32_t p = blockIdx.x * kT + threadIdx.x;
    if (p >= n) return;
    const double *src = pts + (uint64_t)ila[p] * dim;
    for (uint32_t l = 0; l < dim; l++) soa[(uint64_t)l * n + p] = src[l];
}

// aos[p*dim + l] = pts[ila[p]*dim + l]: the same, point-major.
__global__ __launch_bounds__(kT) void k_gather_aos(const double *__restrict__ pts,
                                                   const uint32_t *__restrict__ ila, uint32_t n,
                                                   uint32_t dim, double *__restrict__ aos) {
    uint32_t p = blockIdx.x * kT + threadIdx.x;
    if (p >= n) return;
    const double *src = pts + (uint64_t)ila[p] * dim;
    for (uint32_t l = 0; l < dim; l++) aos[(uint64_t)p * dim + l] = src[l];
}

}  // namespace

// Host side of S1 (main.cu:493-518) from the exact device min/max.  Returns
// false when the grid cannot be represented: a dimension with 2^32 cells or
// more, or more than 2^64 - 1 cells in total (the reference's uint32 cast and
// uint64 product would wrap there and cell ids would collide).
// With DTYPE=float every step is the same expression in FP32.
bool finish_dims(const double *mn, const double *mx, uint32_t k, double eps, GridGeom *g,
                 double *maxArr, uint64_t *totalCells) {
    const float epsf = (float)eps;
    for (uint32_t j = 0; j < k; j++) {
        if (g->f32dt) {
            g->minArr[j] = (double)((float)mn[j] - epsf);
            maxArr[j] = (double)((float)mx[j] + epsf);
        } else {
            g->minArr[j] = mn[j] - eps;
            maxArr[j] = mx[j] + eps;
        }
    }
    unsigned __int128 t = 1;
    bool ok = true;
    for (uint32_t j = 0; j < k; j++) {
        const double c = g->f32dt ? (double)std::ceil(((float)maxArr[j] - (float)g->minArr[j]) / epsf)
                                  : std::ceil((maxArr[j] - g->minArr[j]) / eps);
        ok &= c < 4294967296.0;
        g->nCells[j] = ok ? (uint32_t)c : 1u;
        t *= g->nCells[j];
        ok &= t <= (unsigned __int128)~0ull;
    }
    *totalCells = ok ? (uint64_t)t : 0;
    return ok;
}

// Builds the device index from device points (already copied into ix->pts).
hegj_status build_index(GridGeom *g, double *maxArr, uint64_t *totalCells, DevIndex *ix,
                        uint32_t *n_dead, hipStream_t s) {
    const uint32_t n = g->n, k = g->k, dim = g->dim;
    Scratch sc;
    if (g->f32dt) {
        const uint64_t total = (uint64_t)n * dim;
        k_round_f32<<<ceil_div(total, kT), kT, 0, s>>>(ix->pts, total);
        HIP_TRY(hipGetLastError());
    }
    // -- S1: extents ---------------------------------------------------------
    uint32_t mm_blocks = std::min<uint32_t>(ceil_div(n, kT), 1024);
    double *partials;
    HIP_TRY(sc.get(&partials, (size_t)mm_blocks * 2 * k));
    k_minmax<<<mm_blocks, kT, 0, s>>>(ix->pts, n, dim, k, partials);
    HIP_TRY(hipGetLastError());
    std::vector<double> hp((size_t)mm_blocks * 2 * k);
    HIP_TRY(hipMemcpyAsync(hp.data(), partials, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double mn[kMaxK], mx[kMaxK];
    for (uint32_t j = 0; j < k; j++) {
        mn[j] = hp[2 * j];
        mx[j] = hp[2 * j + 1];
        for (uint32_t b = 1; b < mm_blocks; b++) {
            mn[j] = std::min(mn[j], hp[(size_t)b * 2 * k + 2 * j]);
            mx[j] = std::max(mx[j], hp[(size_t)b * 2 * k + 2 * j + 1]);
        }
    }
    if (!finish_dims(mn, mx, k, g->eps, g, maxArr, totalCells))
        return fail(HEGJ_ERR_UNSUPPORTED,
                    "grid too fine: more than 2^64 cells (or 2^32 in one dimension); index fewer "
                    "dimensions or use a larger epsilon");
    // key bits
    int bits = 0;
    while (bits < 64 && ((*totalCells - 1) >> bits) != 0) bits++;

    // -- S2/S3: cell ids, sort ------------------------------------------------
    uint64_t *keys, *keys_alt;
    uint32_t *vals_alt, *dead;
    HIP_TRY(sc.get(&keys, n));
    HIP_TRY(sc.get(&keys_alt, n));
    HIP_TRY(sc.get(&vals_alt, n));
    HIP_TRY(sc.get(&dead, 1));
    HIP_TRY(hipMalloc(&ix->point_key, sizeof(uint64_t) * n));
    HIP_TRY(hipMalloc(&ix->ila, sizeof(uint32_t) * n));
    HIP_TRY(hipMemsetAsync(dead, 0, sizeof(uint32_t), s));
    uint32_t blocks = ceil_div(n, kT);
    k_cell_ids<<<blocks, kT, 0, s>>>(ix->pts, *g, keys, ix->point_key, ix->ila, dead);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(radix_sort_u64(keys, ix->ila, keys_alt, vals_alt, n, bits, s));

    // -- S11: cells ------------------------------------------------------------
    uint32_t *flags, *rank, *d_count;
    HIP_TRY(sc.get(&flags, n));
    HIP_TRY(sc.get(&rank, n));
    HIP_TRY(sc.get(&d_count, 2));
    k_run_flags<<<blocks, kT, 0, s>>>(keys, n, 1, flags);
    HEGJ_TRY(scan_u32(flags, rank, n, d_count, s));
    uint32_t C;
    HIP_TRY(hipMemcpyAsync(&C, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(n_dead, dead, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    ix->C = C;
    HIP_TRY(hipMalloc(&ix->cell_start, sizeof(uint32_t) * (C + 1)));
    HIP_TRY(hipMalloc(&ix->cell_key, sizeof(uint64_t) * C));
    k_run_compact<<<blocks, kT, 0, s>>>(keys, n, 1, flags, rank, ix->cell_start, ix->cell_key);
    k_set_u32<<<1, 1, 0, s>>>(ix->cell_start + C, n);
    HIP_TRY(hipGetLastError());

    // -- rows: cells sharing coordinates 1..k-1 ----------------------------------
    uint32_t cb = ceil_div(C, kT);
    uint64_t div = g->nCells[0];
    k_run_flags<<<cb, kT, 0, s>>>(ix->cell_key, C, div, flags);
    HEGJ_TRY(scan_u32(flags, rank, C, d_count, s));
    uint32_t R;
    HIP_TRY(hipMemcpyAsync(&R, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    ix->R = R;
    HIP_TRY(hipMalloc(&ix->row_start, sizeof(uint32_t) * (R + 1)));
    HIP_TRY(hipMalloc(&ix->row_key, sizeof(uint64_t) * R));
    k_run_compact<<<cb, kT, 0, s>>>(ix->cell_key, C, div, flags, rank, ix->row_start, ix->row_key);
    k_set_u32<<<1, 1, 0, s>>>(ix->row_start + R, C);
    HIP_TRY(hipGetLastError());

    // -- row hash table (load factor <= 0.5) -----------------------------------
    uint32_t hsize = 1024;
    while (hsize < 2 * R) hsize <<= 1;
    ix->hmask = hsize - 1;
    HIP_TRY(hipMalloc(&ix->hkeys, sizeof(uint64_t) * hsize));
    HIP_TRY(hipMalloc(&ix->hvals, sizeof(uint32_t) * hsize));
    k_hash_fill<<<ceil_div(hsize, kT), kT, 0, s>>>(ix->hkeys, hsize);
    k_hash_insert<<<ceil_div(R, kT), kT, 0, s>>>(ix->row_key, R, ix->hkeys, ix->hvals, ix->hmask);
    HIP_TRY(hipGetLastError());

    // -- coordinates in cell order ---------------------------------------------
    HIP_TRY(hipMalloc(&ix->soa, sizeof(double) * (size_t)n * dim));
    k_gather_soa<<<blocks, kT, 0, s>>>(ix->pts, ix->ila, n, dim, ix->soa);
    HIP_TRY(hipGetLastError());
    if (dim <= kAosMaxDim) {
        HIP_TRY(hipMalloc(&ix->aos, sizeof(double) * (size_t)n * dim));
        k_gather_aos<<<blocks, kT, 0, s>>>(ix->pts, ix->ila, n, dim, ix->aos);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

void free_index(DevIndex *ix) {
    void *ps[] = {ix->pts, ix->soa, ix->aos, ix->ila, ix->point_key, ix->cell_key, ix->cell_start,
                  ix->row_start, ix->row_key, ix->hkeys, ix->hvals};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    *ix = DevIndex{};
}

}  // namespace hegj
