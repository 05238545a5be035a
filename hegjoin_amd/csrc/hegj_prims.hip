// hegj_prims.hip -- device scan and stable LSD radix sort (wave64, LDS).
//
// Replace the reference's Thrust calls of the grid build and workload sort
// (thrust::sort / unique / sort_by_key, GPU.cu:249-251, 285; std::sort,
// SortByWorkload.cu:97) with hand-written HIP.
#include "hegj_internal.h"

namespace hegj {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Block-wide exclusive scan of one value per thread; returns the block total.
template <typename T>
__device__ T block_exclusive_scan(T v, T *excl, T *sh_waves) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // inclusive wave scan via shuffles
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh_waves[wid] = x;
    __syncthreads();
    T wave_base = 0, total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        T t = sh_waves[w];
        if (w < wid) wave_base += t;
        total += t;
    }
    __syncthreads();
    *excl = wave_base + x - v;
    return total;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const T *__restrict__ in,
                                                             T *__restrict__ sums, uint32_t n) {
    __shared__ T sh[kScanThreads / 64];
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    T acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
        if (base + i < n) acc += in[base + i];
    T excl;
    T tot = block_exclusive_scan<T>(acc, &excl, sh);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single workgroup scan of the tile sums (in place, exclusive).
template <typename T>
__global__ __launch_bounds__(1024) void k_scan_sums(T *__restrict__ sums, uint32_t m,
                                                    T *__restrict__ d_total) {
    __shared__ T sh[16];
    T carry = 0;
    for (uint32_t base = 0; base < m; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        T v = i < m ? sums[i] : 0;
        T excl;
        T tot = block_exclusive_scan<T>(v, &excl, sh);
        if (i < m) sums[i] = carry + excl;
        carry += tot;
    }
    if (threadIdx.x == 0 && d_total) *d_total = carry;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T *__restrict__ in,
                                                             T *__restrict__ out,
                                                             const T *__restrict__ sums,
                                                             uint32_t n) {
    __shared__ T sh[kScanThreads / 64];
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    T v[kScanItems];
    T acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        acc += v[i];
    }
    T excl;
    block_exclusive_scan<T>(acc, &excl, sh);
    T run = sums[blockIdx.x] + excl;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
}

template <typename T>
hegj_status scan_impl(const T *in, T *out, uint32_t n, T *d_total, hipStream_t s) {
    if (n == 0) {
        if (d_total) HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(T), s));
        return HEGJ_OK;
    }
    uint32_t tiles = ceil_div(n, kScanTile);
    Scratch sc;
    T *sums;
    HIP_TRY(sc.get(&sums, tiles));
    k_scan_tiles<T><<<tiles, kScanThreads, 0, s>>>(in, sums, n);
    k_scan_sums<T><<<1, 1024, 0, s>>>(sums, tiles, d_total);
    k_scan_apply<T><<<tiles, kScanThreads, 0, s>>>(in, out, sums, n);
    HIP_TRY(hipGetLastError());
    // Scratch is freed on return: make sure the kernels are done with it.
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  Per pass: per-tile digit histograms, a
// digit-major exclusive scan, then a stable scatter in which each wave ranks
// its keys with ballot "peer masks" and the workgroup orders its waves
// through LDS counters.
// ---------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;

template <typename K>
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const K *__restrict__ keys, uint32_t n,
                                                        int shift, uint32_t *__restrict__ hist,
                                                        uint32_t tiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * kRsTile;
    for (int i = 0; i < kRsItems; i++) {
        uint64_t p = base + i * kRsThreads + threadIdx.x;
        if (p < n) atomicAdd(&h[(uint32_t)(keys[p] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

template <typename K>
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const K *__restrict__ keys,
                                                           const uint32_t *__restrict__ vals,
                                                           K *__restrict__ okeys,
                                                           uint32_t *__restrict__ ovals,
                                                           uint32_t n, int shift,
                                                           const uint32_t *__restrict__ hist,
                                                           uint32_t tiles) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kRsThreads / 64][256];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    run[threadIdx.x] = hist[(uint64_t)threadIdx.x * tiles + blockIdx.x];
    for (int w = 0; w < kRsThreads / 64; w++) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * kRsTile;
    for (int i = 0; i < kRsItems; i++) {
        uint64_t p = base + i * kRsThreads + threadIdx.x;
        bool valid = p < n;
        K key = valid ? keys[p] : K(0);
        uint32_t val = valid ? vals[p] : 0u;
        uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = lanes_below(peers);
        if (valid && rank == 0) wcnt[wid][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = run[d] + rank;
            for (int w = 0; w < wid; w++) off += wcnt[w][d];
            okeys[off] = key;
            ovals[off] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kRsThreads / 64; w++) {
            add += wcnt[w][threadIdx.x];
            wcnt[w][threadIdx.x] = 0;
        }
        run[threadIdx.x] += add;
        __syncthreads();
    }
    (void)lane;
}

template <typename K>
hegj_status radix_impl(K *keys, uint32_t *vals, K *keys_alt, uint32_t *vals_alt, uint32_t n,
                       int bits, hipStream_t s) {
    if (n <= 1 || bits <= 0) return HEGJ_OK;
    uint32_t tiles = ceil_div(n, kRsTile);
    Scratch sc;
    uint32_t *hist;
    HIP_TRY(sc.get(&hist, (size_t)tiles * 256));
    K *src_k = keys, *dst_k = keys_alt;
    uint32_t *src_v = vals, *dst_v = vals_alt;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; p++) {
        int shift = p * 8;
        k_rs_hist<K><<<tiles, kRsThreads, 0, s>>>(src_k, n, shift, hist, tiles);
        HIP_TRY(hipGetLastError());
        HEGJ_TRY(scan_u32(hist, hist, tiles * 256, nullptr, s));
        k_rs_scatter<K><<<tiles, kRsThreads, 0, s>>>(src_k, src_v, dst_k, dst_v, n, shift, hist,
                                                     tiles);
        HIP_TRY(hipGetLastError());
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
    }
    if (src_k != keys) {
        HIP_TRY(hipMemcpyAsync(keys, src_k, sizeof(K) * n, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

}  // namespace

hegj_status scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *d_total,
                     hipStream_t s) {
    return scan_impl<uint32_t>(in, out, n, d_total, s);
}

hegj_status scan_u64(const uint64_t *in, uint64_t *out, uint32_t n, uint64_t *d_total,
                     hipStream_t s) {
    return scan_impl<uint64_t>(in, out, n, d_total, s);
}

hegj_status radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                           uint32_t *vals_alt, uint32_t n, int bits, hipStream_t s) {
    return radix_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, bits, s);
}

hegj_status radix_sort_u32(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt,
                           uint32_t *vals_alt, uint32_t n, int bits, hipStream_t s) {
    return radix_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, bits, s);
}

}  // namespace hegj
