// hegj_util.hip -- small device kernels around the join: the device half of
// sortByWorkLoad (SortByWorkload.cu:96-217), unit list gathers, and the
// neighbour-table to CSR gather (the device form of
// constructNeighborTableKeyValueWithPtrs, GPU.cu:1731-1776).
#include "hegj_internal.h"

namespace hegj {

namespace {
constexpr int kT = 256;

// key = ~work so an ascending stable sort gives descending work with ties in
// ascending cell id (SortByWorkload.cu:97-98).
__global__ __launch_bounds__(kT) void k_work_keys(const uint32_t *__restrict__ work, uint32_t C,
                                                  uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ vals) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < C) {
        keys[i] = ~work[i];
        vals[i] = i;
    }
}

__global__ __launch_bounds__(kT) void k_sorted_sizes(const uint32_t *__restrict__ cell_start,
                                                     const uint32_t *__restrict__ order,
                                                     uint32_t C, uint32_t *__restrict__ sizes) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < C) {
        uint32_t c = order[i];
        sizes[i] = cell_start[c + 1] - cell_start[c];
    }
}

// originPointIndex: member lists of the cells in workload order
// (SortByWorkload.cu:155-181).  One wave per cell.
__global__ __launch_bounds__(kT) void k_origin_scatter(const uint32_t *__restrict__ cell_start,
                                                       const uint32_t *__restrict__ ila,
                                                       const uint32_t *__restrict__ order,
                                                       const uint32_t *__restrict__ offs,
                                                       uint32_t C, uint32_t *__restrict__ opi) {
    uint32_t w = (blockIdx.x * kT + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= C) return;
    uint32_t c = order[w];
    uint32_t s = cell_start[c], e = cell_start[c + 1], o = offs[w];
    for (uint32_t p = s + lane; p < e; p += 64) opi[o + (p - s)] = ila[p];
}

// totalCandidates = sum |c| * work(c) (SortByWorkload.cu:160), in 64 bits.
__global__ __launch_bounds__(kT) void k_total_candidates(const uint32_t *__restrict__ cell_start,
                                                         const uint32_t *__restrict__ work,
                                                         uint32_t C,
                                                         unsigned long long *__restrict__ out) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    uint64_t v = 0;
    if (i < C) v = (uint64_t)(cell_start[i + 1] - cell_start[i]) * work[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, (unsigned long long)v);
}

__global__ __launch_bounds__(kT) void k_stride_units(const uint2 *__restrict__ in,
                                                     uint32_t n_out, uint32_t stride,
                                                     uint2 *__restrict__ out) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < n_out) out[i] = in[(uint64_t)i * stride];
}

__global__ __launch_bounds__(kT) void k_gather_units(const uint2 *__restrict__ in,
                                                     const uint32_t *__restrict__ idx, uint32_t n,
                                                     uint2 *__restrict__ out) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}

// Copies each query's row into CSR order (one wave per query).  Batch
// buffers are virtually concatenated: batch b covers global offsets
// [base[b], base[b+1]).
struct BatchTable {
    const uint32_t *ptr[kMaxBatches];
    uint64_t base[kMaxBatches + 1];
    uint32_t nb;
};

__global__ __launch_bounds__(kT) void k_gather_rows(BatchTable bt,
                                                    const uint64_t *__restrict__ row_off,
                                                    const uint32_t *__restrict__ row_cnt,
                                                    const uint64_t *__restrict__ csr_off, uint32_t n,
                                                    uint32_t *__restrict__ out) {
    uint32_t q = (blockIdx.x * kT + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (q >= n) return;
    uint32_t cnt = row_cnt[q];
    if (cnt == 0) return;
    uint64_t go = row_off[q];
    uint32_t b = 0;
    while (b + 1 < bt.nb && bt.base[b + 1] <= go) b++;
    const uint32_t *src = bt.ptr[b] + (go - bt.base[b]);
    uint32_t *dst = out + csr_off[q];
    for (uint32_t i = lane; i < cnt; i += 64) dst[i] = src[i];
}

__global__ __launch_bounds__(kT) void k_u32_to_u64(const uint32_t *__restrict__ in, uint32_t n,
                                                   uint64_t *__restrict__ out) {
    uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < n) out[i] = in[i];
}

}  // namespace

hegj_status workload_order(const DevIndex &ix, const uint32_t *work, uint32_t *order,
                           uint64_t *total_candidates, hipStream_t s) {
    const uint32_t C = ix.C;
    Scratch sc;
    uint32_t *keys, *keys_alt, *vals_alt;
    unsigned long long *tot;
    HIP_TRY(sc.get(&keys, C));
    HIP_TRY(sc.get(&keys_alt, C));
    HIP_TRY(sc.get(&vals_alt, C));
    HIP_TRY(sc.get(&tot, 1));
    HIP_TRY(hipMemsetAsync(tot, 0, sizeof(*tot), s));
    k_work_keys<<<ceil_div(C, kT), kT, 0, s>>>(work, C, keys, order);
    k_total_candidates<<<ceil_div(C, kT), kT, 0, s>>>(ix.cell_start, work, C, tot);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(radix_sort_u32(keys, order, keys_alt, vals_alt, C, 32, s));
    HIP_TRY(hipMemcpyAsync(total_candidates, tot, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

hegj_status origin_point_index(const DevIndex &ix, const uint32_t *order, uint32_t *opi,
                               hipStream_t s) {
    const uint32_t C = ix.C;
    Scratch sc;
    uint32_t *sizes, *offs;
    HIP_TRY(sc.get(&sizes, C));
    HIP_TRY(sc.get(&offs, C));
    k_sorted_sizes<<<ceil_div(C, kT), kT, 0, s>>>(ix.cell_start, order, C, sizes);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(scan_u32(sizes, offs, C, nullptr, s));
    k_origin_scatter<<<ceil_div((uint64_t)C * 64, kT), kT, 0, s>>>(ix.cell_start, ix.ila, order, offs,
                                                                 C, opi);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

hegj_status stride_units(const uint2 *in, uint32_t n_out, uint32_t stride, uint2 *out,
                         hipStream_t s) {
    if (n_out == 0) return HEGJ_OK;
    k_stride_units<<<ceil_div(n_out, kT), kT, 0, s>>>(in, n_out, stride, out);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

hegj_status gather_units(const uint2 *in, const uint32_t *idx, uint32_t n, uint2 *out,
                         hipStream_t s) {
    if (n == 0) return HEGJ_OK;
    k_gather_units<<<ceil_div(n, kT), kT, 0, s>>>(in, idx, n, out);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

hegj_status csr_offsets(const uint32_t *row_cnt, uint32_t n, uint64_t *csr_off, hipStream_t s) {
    Scratch sc;
    uint64_t *tmp;
    HIP_TRY(sc.get(&tmp, (size_t)n + 1));
    k_u32_to_u64<<<ceil_div(n, kT), kT, 0, s>>>(row_cnt, n, tmp);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(scan_u64(tmp, csr_off, n, csr_off + n, s));
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

hegj_status gather_rows(const std::vector<uint32_t *> &batches, const std::vector<uint64_t> &bases,
                        const uint64_t *row_off, const uint32_t *row_cnt, const uint64_t *csr_off,
                        uint32_t n, uint32_t *out, hipStream_t s) {
    if (batches.size() > (size_t)kMaxBatches) return fail(HEGJ_ERR_UNSUPPORTED, "too many result batches");
    BatchTable bt{};
    bt.nb = (uint32_t)batches.size();
    for (size_t i = 0; i < batches.size(); i++) {
        bt.ptr[i] = batches[i];
        bt.base[i] = bases[i];
    }
    bt.base[batches.size()] = ~0ull;
    k_gather_rows<<<ceil_div((uint64_t)n * 64, kT), kT, 0, s>>>(bt, row_off, row_cnt, csr_off, n, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

}  // namespace hegj
