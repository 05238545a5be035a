// hegj_util.hip -- small device kernels around the join: the device half of
// sortByWorkLoad (SortByWorkload.cu:96-217), unit list gathers, and the
// neighbour-table to CSR gather (the device form of
// constructNeighborTableKeyValueWithPtrs, GPU.cu:1731-1776).
#include <algorithm>
#include <vector>

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

namespace {
__global__ __launch_bounds__(kT) void k_scatter_rows(const uint32_t *__restrict__ qid,
                                                     const uint64_t *__restrict__ off,
                                                     const uint32_t *__restrict__ cnt, uint32_t m,
                                                     uint64_t base, uint64_t *__restrict__ row_off,
                                                     uint32_t *__restrict__ row_cnt) {
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= m) return;
    row_off[qid[i]] = base + off[i];
    row_cnt[qid[i]] = cnt[i];
}
}  // namespace

hegj_status scatter_rows(const uint32_t *qid, const uint64_t *off, const uint32_t *cnt, uint32_t m,
                         uint64_t base, uint64_t *row_off, uint32_t *row_cnt, hipStream_t s) {
    if (m == 0) return HEGJ_OK;
    k_scatter_rows<<<ceil_div(m, kT), kT, 0, s>>>(qid, off, cnt, m, base, row_off, row_cnt);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

}  // namespace hegj

namespace hegj {
namespace {
// Streaming copy (16-byte vectors, grid-stride): the reference point for HBM
// counters and the attainable bandwidth of the device.
__global__ __launch_bounds__(256) void k_stream_copy(const uint4 *__restrict__ src,
                                                     uint4 *__restrict__ dst, uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * 256)
        dst[i] = src[i];
}
}  // namespace

hegj_status stream_copy(uint64_t bytes, int reps, double *ms_per_copy) {
    uint64_t n16 = bytes / 16;
    Scratch sc;
    uint4 *a, *b;
    HIP_TRY(sc.get(&a, n16));
    HIP_TRY(sc.get(&b, n16));
    HIP_TRY(hipMemset(a, 1, n16 * 16));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    k_stream_copy<<<8192, 256>>>(a, b, n16);  // warm
    HIP_TRY(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; r++) k_stream_copy<<<8192, 256>>>(a, b, n16);
    HIP_TRY(hipEventRecord(e1, nullptr));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_copy = ms / reps;
    return HEGJ_OK;
}

}  // namespace hegj

namespace hegj {
namespace {

// Morton key of each cell: bits of the k cell coordinates interleaved
// (coordinate j's bit b goes to position b*k + j).
__global__ __launch_bounds__(256) void k_morton_keys(GridGeom g, DevIndex ix, uint32_t bits,
                                                     uint64_t *__restrict__ keys,
                                                     uint32_t *__restrict__ vals) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ix.C) return;
    uint32_t c[kMaxK];
    decode_cell(ix.cell_key[i], g, c);
    uint64_t key = 0;
    for (uint32_t b = 0; b < bits; b++)
        for (uint32_t j = 0; j < g.k; j++) key |= (uint64_t)((c[j] >> b) & 1u) << (b * g.k + j);
    keys[i] = key;
    vals[i] = i;
}

// Estimated work of each unit: queries x candidates of its cell.
__global__ __launch_bounds__(256) void k_unit_work(const uint2 *__restrict__ units, uint32_t n,
                                                   const uint32_t *__restrict__ cell_start,
                                                   const uint32_t *__restrict__ work, uint32_t qb,
                                                   uint64_t *__restrict__ w) {
    uint32_t u = blockIdx.x * 256 + threadIdx.x;
    if (u >= n) return;
    uint2 un = units[u];
    uint32_t sz = cell_start[un.x + 1] - cell_start[un.x];
    uint32_t nq = min(qb, sz - un.y);
    w[u] = (uint64_t)nq * (work[un.x] + 64);  // + per-unit overhead
}

// part_end[p] = first unit whose exclusive work prefix reaches (p+1)/n of
// the total (the last part ends at n).
__global__ void k_find_splits(const uint64_t *__restrict__ prefix, uint32_t n,
                              const uint64_t *__restrict__ total, uint32_t parts,
                              uint32_t *__restrict__ part_end) {
    uint32_t p = threadIdx.x;
    if (p >= parts) return;
    if (p == parts - 1) {
        part_end[p] = n;
        return;
    }
    uint64_t target = (*total) / parts * (p + 1);
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    part_end[p] = lo;
}

}  // namespace

hegj_status morton_order(const GridGeom &g, const DevIndex &ix, uint32_t *order, bool *ok,
                         hipStream_t s) {
    uint32_t maxc = 1;
    for (uint32_t j = 0; j < g.k; j++) maxc = std::max(maxc, g.nCells[j]);
    uint32_t bits = 0;
    while (bits < 32 && ((maxc - 1) >> bits) != 0) bits++;
    if (bits * g.k > 64 || ix.C == 0) {
        *ok = false;
        return HEGJ_OK;
    }
    *ok = true;
    Scratch sc;
    uint64_t *keys, *keys_alt;
    uint32_t *vals_alt;
    HIP_TRY(sc.get(&keys, ix.C));
    HIP_TRY(sc.get(&keys_alt, ix.C));
    HIP_TRY(sc.get(&vals_alt, ix.C));
    k_morton_keys<<<ceil_div(ix.C, 256), 256, 0, s>>>(g, ix, bits, keys, order);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(radix_sort_u64(keys, order, keys_alt, vals_alt, ix.C, (int)(bits * g.k), s));
    return HEGJ_OK;
}

hegj_status split_units(const DevIndex &ix, const uint2 *units, uint32_t n_units, uint32_t qb,
                        const uint32_t *work, uint32_t n_parts, uint32_t *part_end,
                        hipStream_t s) {
    if (n_parts <= 1 || n_units == 0) {
        part_end[0] = n_units;
        return HEGJ_OK;
    }
    Scratch sc;
    uint64_t *w, *prefix, *total;
    uint32_t *d_end;
    HIP_TRY(sc.get(&w, n_units));
    HIP_TRY(sc.get(&prefix, n_units));
    HIP_TRY(sc.get(&total, 1));
    HIP_TRY(sc.get(&d_end, kMaxParts));
    k_unit_work<<<ceil_div(n_units, 256), 256, 0, s>>>(units, n_units, ix.cell_start, work, qb, w);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(scan_u64(w, prefix, n_units, total, s));
    k_find_splits<<<1, 64, 0, s>>>(prefix, n_units, total, n_parts, d_end);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(part_end, d_end, sizeof(uint32_t) * n_parts, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

namespace {

// Per-dimension min / max of the cell-ordered coordinates, one partial per block.
__global__ __launch_bounds__(256) void k_dim_minmax(const double *__restrict__ soa, uint32_t n,
                                                    uint32_t dim, double *__restrict__ part) {
    __shared__ double smin[256], smax[256];
    for (uint32_t l = 0; l < dim; l++) {
        double lo = 1.0 / 0.0, hi = -1.0 / 0.0;
        for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (uint64_t)gridDim.x * 256) {
            const double v = soa[(uint64_t)l * n + p];
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
        smin[threadIdx.x] = lo;
        smax[threadIdx.x] = hi;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + o]);
                smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            part[((uint64_t)blockIdx.x * dim + l) * 2] = smin[0];
            part[((uint64_t)blockIdx.x * dim + l) * 2 + 1] = smax[0];
        }
        __syncthreads();
    }
}

// soa32 = (float)(soa - centre): the FP64 difference rounded to FP32.
__global__ __launch_bounds__(256) void k_soa32(const double *__restrict__ soa, uint64_t total,
                                               uint32_t n, const double *__restrict__ centre,
                                               float *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < total) out[i] = (float)(soa[i] - centre[i / n]);
}

// aos32[p * dim + l] = soa32[l * n + p] (the same FP32 values, point-major),
// then kAos32Pad points at +inf (the FP32 join's padding lanes read them).
__global__ __launch_bounds__(256) void k_aos32(const float *__restrict__ soa32, uint32_t n, uint32_t dim,
                                               float *__restrict__ aos32) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n + kAos32Pad) return;
    for (uint32_t l = 0; l < dim; l++)
        aos32[(uint64_t)p * dim + l] = p < n ? soa32[(uint64_t)l * n + p] : __builtin_inff();
}

}  // namespace

hegj_status make_soa32(const GridGeom &g, const DevIndex &ix, float **soa32, float **aos32, double *radius,
                       double **centre_dev, hipStream_t s) {
    const uint32_t n = g.n, dim = g.dim;
    const uint32_t blocks = std::min<uint32_t>(ceil_div(n, 256), 512);
    Scratch sc;
    double *part, *centre;
    HIP_TRY(sc.get(&part, (size_t)blocks * dim * 2));
    HIP_TRY(hipMalloc(centre_dev, sizeof(double) * dim));
    centre = *centre_dev;
    k_dim_minmax<<<blocks, 256, 0, s>>>(ix.soa, n, dim, part);
    HIP_TRY(hipGetLastError());
    std::vector<double> hp((size_t)blocks * dim * 2), c(dim);
    HIP_TRY(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double R = 0.0;
    for (uint32_t l = 0; l < dim; l++) {
        double lo = hp[l * 2], hi = hp[l * 2 + 1];
        for (uint32_t b = 1; b < blocks; b++) {
            lo = std::min(lo, hp[((size_t)b * dim + l) * 2]);
            hi = std::max(hi, hp[((size_t)b * dim + l) * 2 + 1]);
        }
        // DTYPE=float: no centring, the copy is the float data itself
        c[l] = g.f32dt ? 0.0 : lo + (hi - lo) * 0.5;
        R = std::max(R, std::max(hi - c[l], c[l] - lo));
    }
    *radius = R * (1.0 + 1e-12);
    HIP_TRY(hipMemcpyAsync(centre, c.data(), sizeof(double) * dim, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMalloc(soa32, sizeof(float) * (size_t)n * dim));
    const uint64_t total = (uint64_t)n * dim;
    k_soa32<<<ceil_div(total, 256), 256, 0, s>>>(ix.soa, total, n, centre, *soa32);
    HIP_TRY(hipGetLastError());
    if (aos32 && dim <= kAosMaxDim) {
        HIP_TRY(hipMalloc(aos32, sizeof(float) * ((size_t)n + kAos32Pad) * dim));
        k_aos32<<<ceil_div((uint64_t)n + kAos32Pad, 256), 256, 0, s>>>(*soa32, n, dim, *aos32);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

}  // namespace hegj
