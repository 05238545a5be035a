// hegj_mfma.hip -- split-BF16 MFMA distance tiles with a proven guard band
// and an exact FP64 recheck: the refine of the join at high dimensionality.
//
// At 18-D and 90-D the k = 6 grid leaves each query a large share of the
// data as candidates (the reference evaluates ~1e12-3e12 FP64 distances,
// kernel.cu:350-355), so the refine is a dense contraction.  Each pair's
// squared distance is computed as a Gram product on the matrix cores:
//
//   points are centred (y = x - mean, FP64) and split into BF16 hi + lo
//   (y ~ hi + lo, |y - hi - lo| <= ~u_b^2 |y|, u_b = 2^-8); with
//   a = [-2y, ndn, 1], b = [y, 1, ndn] (ndn = a lowered norm, below),
//   A' = [hi(a) | hi(a) | lo(a)] and B' = [hi(b) | lo(b) | hi(b)] (K padded
//   to a multiple of 32), one v_mfma_f32_16x16x32_bf16 chain gives
//   s' = A'.B' = -2 y_q.y_c + ndn_q + ndn_c   (hi*hi + hi*lo + lo*hi terms,
//   FP32 accumulation)
//
// For every point p, with N_p = ||hi+lo||^2 and delta_p = A_c * N_p (A_c
// bounds, per unit of N_q + N_c, the MFMA accumulation error, the dropped
// lo*lo term, the BF16 split error of the coordinates and the rounding of
// the reference's own FP64 chain), ndn_p <= N_p - 2 delta_p and
// w_p >= N_p - ndn_p + 2 delta_p.  Then for the exact FP64 value acc of the
// reference (fma chain, kernel.cu:350-353):
//     s' > T                        =>  acc > T   (reject, the common case)
//     s' + w_q + w_c <= T (1-4u64)  =>  acc <= T  (accept)
// and everything else is the band: those pairs are rechecked with the exact
// FP64 chain.  The guard is twice the error bound (safety factor 2).  The
// pair set is therefore identical to the FP64 one; the band count is
// reported.
#include <algorithm>
#include <cmath>

#include "hegj_internal.h"

namespace hegj {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Waves per workgroup: the double-buffered candidate tile (64 rows of K
// bf16) takes 2 * 64 * (K + 8) * 2 bytes of LDS, so above K = 128 one
// workgroup fills a CU's LDS and its waves are all the CU runs; more waves
// then both fill the SIMDs and serve each streamed candidate to more queries.
constexpr int gb_waves(int kb) { return kb >= 6 ? 8 : 4; }
constexpr int kGbChunk = 64;         // candidates per LDS chunk (4 col tiles)
// 16-query row tiles per wave: 4 while the A fragments are small, 2 above
constexpr int gb_rt(int kb) { return kb <= 4 ? 4 : 2; }
constexpr int kGbBlk = 256;          // pair-list slots reserved per wave at a time
constexpr uint64_t kSentinel = ~0ull;

// Bit moves between float and uint32 that compile for host and device.
__host__ __device__ __forceinline__ uint32_t f2u(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}
__host__ __device__ __forceinline__ float u2f(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__host__ __device__ __forceinline__ uint16_t bf16_bits_rne(float f) {
    uint32_t u = f2u(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_to_f(uint16_t b) { return u2f((uint32_t)b << 16); }

// Neighbouring FP32 values (finite inputs), by bit arithmetic.
__host__ __device__ __forceinline__ float f32_next_up(float f) {
    uint32_t u = f2u(f);
    if (f == 0.0f) return u2f(1u);
    return u2f((u >> 31) ? u - 1 : u + 1);
}
__host__ __device__ __forceinline__ float f32_next_down(float f) {
    uint32_t u = f2u(f);
    if (f == 0.0f) return u2f(0x80000001u);
    return u2f((u >> 31) ? u + 1 : u - 1);
}

// FP64 -> FP32 rounded toward -inf.
__host__ __device__ __forceinline__ float f32_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = f32_next_down(f);
    return f;
}

// FP32 -> BF16 rounded toward -inf (finite inputs).
__host__ __device__ __forceinline__ uint16_t bf16_bits_down(float f) {
    uint32_t u = f2u(f);
    uint32_t t = u & 0xFFFF0000u;
    if ((u & 0xFFFFu) && (u >> 31)) t += 0x10000u;  // negative: away from zero
    return (uint16_t)(t >> 16);
}

// Split v (FP64) into BF16 hi + lo.
__host__ __device__ __forceinline__ void split_bf16(double v, uint16_t &hi, uint16_t &lo) {
    hi = bf16_bits_rne((float)v);
    lo = bf16_bits_rne((float)(v - (double)bf16_to_f(hi)));
}

// One point's A' and B' rows (kp bf16 each) and its guard w, from its
// coordinates x (stride `stride` between dims) and the centre.
__host__ __device__ inline void gb_prepare_point(const double *x, uint64_t stride, uint32_t dim,
                                                 uint32_t kp, const double *mean, double Acoef,
                                                 uint16_t *A, uint16_t *B, float *w) {
    const uint32_t g = dim + 2;  // group length
    double N = 0.0;
    for (uint32_t l = 0; l < dim; l++) {
        double y = x[(uint64_t)l * stride] - mean[l];
        uint16_t hi, lo;
        split_bf16(y, hi, lo);
        double yt = (double)bf16_to_f(hi) + (double)bf16_to_f(lo);
        N = __builtin_fma(yt, yt, N);
        // b: [y, 1, ndn], a: [-2y, ndn, 1]; -2 scaling is exact in BF16
        uint16_t ahi = bf16_bits_rne(-2.0f * bf16_to_f(hi));
        uint16_t alo = bf16_bits_rne(-2.0f * bf16_to_f(lo));
        B[l] = hi;
        B[g + l] = lo;
        B[2 * g + l] = hi;
        A[l] = ahi;
        A[g + l] = ahi;
        A[2 * g + l] = alo;
    }
    N *= 1.0 + 4.0 * (dim + 1) * 1.1102230246251565e-16;  // N >= ||hi+lo||^2
    const double delta = Acoef * N;
    // ndn_tilde = hi + lo <= N - 2 delta: hi rounded to nearest, lo (the
    // exact FP64 remainder) rounded down twice (to FP32, then to BF16)
    const double ndn = N - 2.0 * delta;
    const uint16_t nhi = bf16_bits_rne((float)ndn);
    const uint16_t nlo = bf16_bits_down(f32_down(ndn - (double)bf16_to_f(nhi)));
    const double nt = (double)bf16_to_f(nhi) + (double)bf16_to_f(nlo);
    const uint16_t one = 0x3F80;  // 1.0
    B[dim] = one;
    B[g + dim] = 0;
    B[2 * g + dim] = one;
    B[dim + 1] = nhi;
    B[g + dim + 1] = nlo;
    B[2 * g + dim + 1] = nhi;
    A[dim] = nhi;
    A[g + dim] = nhi;
    A[2 * g + dim] = nlo;
    A[dim + 1] = one;
    A[g + dim + 1] = one;
    A[2 * g + dim + 1] = 0;
    for (uint32_t k = 3 * g; k < kp; k++) {
        A[k] = 0;
        B[k] = 0;
    }
    // w >= N - ndn_tilde + 2 delta, rounded up to FP32
    double wd = (N - nt) + 2.0 * delta;
    float wf = (float)wd;
    if ((double)wf < wd) wf = f32_next_up(wf);
    *w = wf;
}

// Per point (cell order): A' and B' rows, and w.
__global__ __launch_bounds__(256) void k_gb_prepare(const double *__restrict__ soa, uint32_t n,
                                                    uint32_t dim, uint32_t kp,
                                                    const double *__restrict__ mean, double Acoef,
                                                    uint16_t *__restrict__ Av,
                                                    uint16_t *__restrict__ Bv,
                                                    float *__restrict__ w) {
    uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    gb_prepare_point(soa + p, n, dim, kp, mean, Acoef, Av + (uint64_t)p * kp, Bv + (uint64_t)p * kp, w + p);
}

__global__ __launch_bounds__(256) void k_dim_sums(const double *__restrict__ soa, uint32_t n,
                                                  uint32_t dim, double *__restrict__ partial) {
    // partial[b * dim + l] = sum over this block's points of coordinate l
    __shared__ double sh[256];
    for (uint32_t l = 0; l < dim; l++) {
        double s = 0;
        for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (uint64_t)gridDim.x * 256)
            s += soa[(uint64_t)l * n + p];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(uint64_t)blockIdx.x * dim + l] = sh[0];
        __syncthreads();
    }
}

struct GbArgs {
    GridGeom g;
    DevIndex ix;
    const uint2 *units;
    uint32_t n_units;
    uint32_t *queue_heads;        // [kMaxParts] units taken per range
    uint32_t n_parts;             // unit ranges: range p is [part_end[p - 1], part_end[p])
    uint32_t part_end[kMaxParts];
    const uint16_t *Av;
    const uint16_t *Bv;
    const float *w;
    uint32_t kp;
    uint64_t *pairs;              // accepted keys (qpos << 32 | cpos)
    uint64_t pair_cap;
    unsigned long long *pair_cursor;
    uint64_t *band;               // band keys
    uint64_t band_cap;
    unsigned long long *band_cursor;
    unsigned long long *n_accept;
    unsigned long long *n_band;
    unsigned long long *evals;
    uint32_t *overflow;
};

// Stencil-row list shared by the workgroup.
struct GbRows {
    uint32_t start[kMaxRows];
    uint32_t off[kMaxRows + 1];
};

__device__ __forceinline__ uint32_t gb_cand_pos(const GbRows &rl, uint32_t nslots, uint32_t t,
                                                uint32_t &slot) {
    while (slot + 1 < nslots && rl.off[slot + 1] <= t) slot++;
    return rl.start[slot] + (t - rl.off[slot]);
}

// Per-wave append buffer into a global key list (blocks of kGbBlk slots;
// unused tails are filled with the sentinel, which sorts last).
struct Appender {
    uint64_t cur, end;
};

__device__ __forceinline__ void append_keys(uint64_t key, bool pred, Appender &ap,
                                            uint64_t *list, uint64_t cap,
                                            unsigned long long *cursor, uint32_t *overflow) {
    const int lane = threadIdx.x & 63;
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t k = (uint32_t)__popcll(m);
    if (ap.cur + k > ap.end) {
        // pad the rest of the block, then take a new one
        for (uint64_t i = ap.cur + lane; i < ap.end; i += 64) list[i] = kSentinel;
        uint64_t b = 0;
        if (lane == 0) b = atomicAdd(cursor, (unsigned long long)kGbBlk);
        b = wave_bcast_u64(b);
        if (b + kGbBlk > cap) {
            if (lane == 0) atomicOr(overflow, 1u);
            ap.cur = ap.end = 0;
            return;
        }
        ap.cur = b;
        ap.end = b + kGbBlk;
    }
    if (pred) list[ap.cur + lanes_below(m)] = key;
    ap.cur += k;
}

__device__ __forceinline__ void close_appender(Appender &ap, uint64_t *list) {
    const int lane = threadIdx.x & 63;
    for (uint64_t i = ap.cur + lane; i < ap.end; i += 64) list[i] = kSentinel;
}

template <int KB>
__global__ __launch_bounds__(64 * gb_waves(KB)) void k_gb_join(GbArgs a) {
    constexpr int kGbWaves = gb_waves(KB);
    constexpr int KP = KB * 32;
    constexpr int ROW = KP + 8;  // padded LDS row (bf16) against bank conflicts
    constexpr int RT = gb_rt(KB);
    constexpr uint32_t kQueries = kGbWaves * RT * 16;
    __shared__ GbRows rl;
    __shared__ uint16_t cb[2][kGbChunk * ROW];
    __shared__ uint32_t cpos[2][kGbChunk];
    __shared__ uint2 surv[kGbWaves][256];  // survivors of one 16x16 tile
    __shared__ uint32_t sh_unit[3];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const double T = a.g.thresh;
    // reject when s' > Trej with Trej >= T (T rounded up to FP32)
    float Trej = (float)T;
    if ((double)Trej < T) Trej = f32_next_up(Trej);
    const double Tacc = T * (1.0 - 4.0 * 1.1102230246251565e-16);
    Appender apA{0, 0}, apB{0, 0};
    uint64_t my_acc = 0, my_band = 0, my_evals = 0;

    // A workgroup takes units from its XCD's range first (neighbouring cells
    // share candidates in that XCD's L2), then from the other ranges.
    const uint32_t part0 = xcc_id() % a.n_parts;
    uint32_t tried = 0;  // ranges found empty (thread 0's)
    while (true) {
        if (tid == 0) {
            uint32_t got = ~0u;
            while (tried < a.n_parts) {
                const uint32_t p = (part0 + tried) % a.n_parts;
                const uint32_t k = (p ? a.part_end[p - 1] : 0u) + atomicAdd(&a.queue_heads[p], 1u);
                if (k < a.part_end[p]) {
                    got = k;
                    break;
                }
                tried++;
            }
            sh_unit[0] = got;
        }
        __syncthreads();
        const uint32_t ui = sh_unit[0];
        if (ui >= a.n_units) break;
        const uint2 unit = a.units[ui];
        const uint32_t cell = unit.x, qoff = unit.y;
        const uint32_t cstart = a.ix.cell_start[cell];
        const uint32_t csize = a.ix.cell_start[cell + 1] - cstart;
        const uint32_t nq = min(kQueries, csize - qoff);
        const uint32_t qpos0 = cstart + qoff;

        // stencil rows of the cell (wave 0), shared through LDS
        if (wid == 0) {
            uint32_t c[kMaxK], lo[kMaxK], hi[kMaxK];
            decode_cell(a.ix.cell_key[cell], a.g, c);
            bool ok = true;
            for (uint32_t j = 0; j < a.g.k; j++)
                ok &= stencil_dim(c[j], a.g.nCells[j], a.g.emulate_cell0, &lo[j], &hi[j]);
            uint32_t slots = 0, total = 0;
            if (ok) {
                const uint32_t nrows = pow3(a.g.k - 1);
                for (uint32_t r0 = 0; r0 < nrows; r0 += 64) {
                    const uint32_t r = r0 + lane;
                    uint32_t pstart = 0, plen = 0;
                    if (r < nrows) {
                        uint32_t cf, cl;
                        stencil_row(c, lo, hi, r, a.g, a.ix, &cf, &cl);
                        if (cl > cf) {
                            pstart = a.ix.cell_start[cf];
                            plen = a.ix.cell_start[cl] - pstart;
                        }
                    }
                    const uint64_t nz = __ballot(plen > 0);
                    uint32_t x = plen;
                    for (int o = 1; o < 64; o <<= 1) {
                        uint32_t y = __shfl_up(x, o, 64);
                        if (lane >= o) x += y;
                    }
                    if (plen > 0) {
                        const uint32_t s = slots + lanes_below(nz);
                        rl.start[s] = pstart;
                        rl.off[s] = total + x - plen;
                    }
                    slots += (uint32_t)__popcll(nz);
                    total += __shfl(x, 63, 64);
                }
            }
            if (lane == 0) {
                rl.off[slots] = total;
                sh_unit[1] = total;
                sh_unit[2] = slots;
            }
        }
        __syncthreads();
        const uint32_t M = sh_unit[1], nslots = sh_unit[2];
        const uint32_t nchunks = (M + kGbChunk - 1) / kGbChunk;

        // this wave's query fragments (A operand), kept in registers
        const uint32_t wq0 = wid * RT * 16;  // first query of the wave in the unit
        bf16x8 afr[RT][KB];
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            const uint32_t q = wq0 + rt * 16 + (lane & 15);
#pragma unroll
            for (int kb = 0; kb < KB; kb++) {
                const uint32_t k0 = kb * 32 + (lane >> 4) * 8;
                if (q < nq) {
                    afr[rt][kb] = *reinterpret_cast<const bf16x8 *>(a.Av + (uint64_t)(qpos0 + q) * a.kp + k0);
                } else {
                    // an absent query: s' = 2^100 against every candidate
                    bf16x8 z;
#pragma unroll
                    for (int i = 0; i < 8; i++)
                        z[i] = (k0 + i == a.g.dim) ? (__bf16)1.2676506e30f : (__bf16)0.0f;
                    afr[rt][kb] = z;
                }
            }
        }

        // chunk loader: thread t covers candidate t % 64, row piece t / 64
        const uint32_t cand = tid & 63, piece = tid >> 6;
        constexpr int kU4 = KP * 2 / 16;                                   // 16B units per row
        constexpr int kPieceU4 = (kU4 + kGbWaves - 1) / kGbWaves;          // per thread
        uint32_t slot = 0;
        uint4 stA[kPieceU4], stB[kPieceU4];
        auto fetch = [&](uint32_t ch, uint4(&stage)[kPieceU4]) {
            const uint32_t t = ch * kGbChunk + cand;
            const bool v = t < M;
            const uint32_t pos = v ? gb_cand_pos(rl, nslots, t, slot) : 0u;
            const uint4 *src = reinterpret_cast<const uint4 *>(a.Bv + (uint64_t)pos * a.kp);
#pragma unroll
            for (int i = 0; i < kPieceU4; i++) {
                const uint32_t u4 = piece * kPieceU4 + i;
                if (u4 < (uint32_t)kU4) {
                    if (v) {
                        stage[i] = src[u4];
                    } else {
                        // an absent candidate: 2^100 at elements dim and dim+1,
                        // so s' >= 2^100 (or +inf) against any query row
                        uint32_t wd[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                        for (uint32_t x = 0; x < 2; x++) {
                            const uint32_t e = a.g.dim + x - u4 * 8;
                            if (e < 8) wd[e >> 1] |= 0x7180u << ((e & 1) * 16);
                        }
                        stage[i] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
                    }
                }
            }
            return pos;
        };
        auto store = [&](uint32_t buf, const uint4(&stage)[kPieceU4], uint32_t pos) {
            uint4 *dst = reinterpret_cast<uint4 *>(&cb[buf][cand * ROW]);
#pragma unroll
            for (int i = 0; i < kPieceU4; i++) {
                const uint32_t u4 = piece * kPieceU4 + i;
                if (u4 < (uint32_t)kU4) dst[u4] = stage[i];
            }
            if (piece == 0) cpos[buf][cand] = pos;
        };

        // Two chunks of loads in flight: chunk ch + 2 is loaded into one
        // register stage while chunk ch is computed, and stored into LDS at
        // the end of chunk ch + 1.
        auto step = [&](uint32_t ch, const uint4(&s_next)[kPieceU4], uint32_t pos_next, uint4(&s_new)[kPieceU4],
                        uint32_t &pos_new) __attribute__((always_inline)) {
            const uint32_t buf = ch & 1;
            if (ch + 2 < nchunks) pos_new = fetch(ch + 2, s_new);
            f32x4 acc[RT][4];
#pragma unroll
            for (int rt = 0; rt < RT; rt++)
#pragma unroll
                for (int ct = 0; ct < 4; ct++) acc[rt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
#pragma unroll
                for (int kb = 0; kb < KB; kb++) {
                    const bf16x8 bfr = *reinterpret_cast<const bf16x8 *>(
                        &cb[buf][(ct * 16 + (lane & 15)) * ROW + kb * 32 + (lane >> 4) * 8]);
#pragma unroll
                    for (int rt = 0; rt < RT; rt++)
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[rt][kb], bfr,
                                                                              acc[rt][ct], 0, 0, 0);
                }
            }
            // reject test: the lane's smallest s' (min3 folds), one compare
            // and one wave vote per chunk; survivors are rare.  (s' is never
            // NaN: padding rows and columns give +2^100 or +inf.)
            float mn = acc[0][0][0];
#pragma unroll
            for (int rt = 0; rt < RT; rt++)
#pragma unroll
                for (int ct = 0; ct < 4; ct++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (rt | ct | j) mn = fminf(mn, acc[rt][ct][j]);
            const uint64_t anym = __ballot(!(mn > Trej));
            if (anym) {
#pragma unroll
                for (int rt = 0; rt < RT; rt++)
#pragma unroll
                    for (int ct = 0; ct < 4; ct++) {
                        // compact this tile's survivors into LDS: (element, s')
                        uint32_t ns = 0;
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const bool sv = !(acc[rt][ct][j] > Trej);
                            const uint64_t m = __ballot(sv);
                            if (sv)
                                surv[wid][ns + lanes_below(m)] =
                                    make_uint2((uint32_t)(j * 64 + lane), f2u(acc[rt][ct][j]));
                            ns += (uint32_t)__popcll(m);
                        }
                        if (ns == 0) continue;
                        wave_lds_sync();
                        for (uint32_t i0 = 0; i0 < ns; i0 += 64) {
                            const uint32_t i = i0 + lane;
                            uint64_t key = 0;
                            bool acc_ok = false, band_ok = false;
                            const uint2 e = i < ns ? surv[wid][i] : make_uint2(0u, 0u);
                            const uint32_t j = e.x >> 6, src = e.x & 63;
                            const uint32_t q = wq0 + rt * 16 + (src >> 4) * 4 + j;
                            const uint32_t ci = ct * 16 + (src & 15);
                            // padding rows/columns never pass (s' >= 2^100); the
                            // bounds are checked here all the same
                            if (i < ns && q < nq && ch * kGbChunk + ci < M) {
                                const uint32_t qp = qpos0 + q, cp = cpos[buf][ci];
                                key = ((uint64_t)qp << 32) | cp;
                                const double hi = (double)u2f(e.y) + (double)a.w[qp] +
                                                  (double)a.w[cp];
                                acc_ok = hi <= Tacc;
                                band_ok = !acc_ok;
                            }
                            append_keys(key, acc_ok, apA, a.pairs, a.pair_cap, a.pair_cursor, a.overflow);
                            append_keys(key, band_ok, apB, a.band, a.band_cap, a.band_cursor, a.overflow);
                            my_acc += acc_ok;
                            my_band += band_ok;
                        }
                        wave_lds_sync();
                    }
            }
            if (ch + 1 < nchunks) store(buf ^ 1, s_next, pos_next);
            __syncthreads();
        };
        uint32_t posA = 0, posB = 0;
        if (nchunks > 0) store(0, stA, fetch(0, stA));
        if (nchunks > 1) posB = fetch(1, stB);
        __syncthreads();
        for (uint32_t ch = 0; ch < nchunks; ch += 2) {
            step(ch, stB, posB, stA, posA);
            if (ch + 1 >= nchunks) break;
            step(ch + 1, stA, posA, stB, posB);
        }
        if (wid == 0 && lane == 0) my_evals += (uint64_t)M * nq;
    }
    close_appender(apA, a.pairs);
    close_appender(apB, a.band);
    for (int o = 32; o > 0; o >>= 1) {
        my_acc += __shfl_xor(my_acc, o, 64);
        my_band += __shfl_xor(my_band, o, 64);
    }
    if (lane == 0) {
        if (my_acc) atomicAdd(a.n_accept, (unsigned long long)my_acc);
        if (my_band) atomicAdd(a.n_band, (unsigned long long)my_band);
        if (my_evals) atomicAdd(a.evals, (unsigned long long)my_evals);
    }
}

// Band pairs: the reference's exact FP64 chain (kernel.cu:350-355) on the
// cell-ordered coordinates; accepted pairs are appended to the pair list.
__global__ __launch_bounds__(256) void k_gb_recheck(const uint64_t *__restrict__ band, uint64_t nb,
                                                    const double *__restrict__ soa, uint32_t n,
                                                    uint32_t dim, double T,
                                                    uint64_t *__restrict__ pairs, uint64_t cap,
                                                    unsigned long long *cursor,
                                                    unsigned long long *n_accept,
                                                    uint32_t *overflow) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool ok = false;
    uint64_t key = 0;
    if (i < nb) {
        key = band[i];
        if (key != kSentinel) {
            const uint32_t qp = (uint32_t)(key >> 32), cp = (uint32_t)key;
            double acc = 0.0;
            for (uint32_t l = 0; l < dim; l++) {
                const double d = soa[(uint64_t)l * n + cp] - soa[(uint64_t)l * n + qp];
                acc = __builtin_fma(d, d, acc);
            }
            ok = acc <= T;
        }
    }
    const uint64_t m = __ballot(ok);
    if (m == 0) return;
    uint64_t b = 0;
    if (lane == 0) {
        b = atomicAdd(cursor, (unsigned long long)__popcll(m));
        atomicAdd(n_accept, (unsigned long long)__popcll(m));
    }
    b = wave_bcast_u64(b);
    if (b + __popcll(m) > cap) {
        if (lane == 0) atomicOr(overflow, 1u);
        return;
    }
    if (ok) pairs[b + lanes_below(m)] = key;
}

// Sorted pair keys -> neighbour rows: per-position counts.
__global__ __launch_bounds__(256) void k_gb_count(const uint64_t *__restrict__ keys, uint64_t np,
                                                  uint32_t *__restrict__ cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < np) atomicAdd(&cnt[(uint32_t)(keys[i] >> 32)], 1u);
}

__global__ __launch_bounds__(256) void k_gb_emit(const uint64_t *__restrict__ keys, uint64_t np,
                                                 const uint32_t *__restrict__ ila,
                                                 uint32_t *__restrict__ out,
                                                 unsigned long long *hash) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint64_t h = 0;
    if (i < np) {
        const uint64_t k = keys[i];
        const uint32_t qid = ila[(uint32_t)(k >> 32)], cid = ila[(uint32_t)k];
        out[i] = cid;
        h = mix64(((uint64_t)qid << 32) | cid);
    }
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
    if ((threadIdx.x & 63) == 0 && h) atomicAdd(hash, (unsigned long long)h);
}

__global__ __launch_bounds__(256) void k_gb_rows(const uint32_t *__restrict__ cnt,
                                                 const uint64_t *__restrict__ start, uint32_t n,
                                                 const uint32_t *__restrict__ ila, uint64_t base,
                                                 uint64_t *__restrict__ row_off,
                                                 uint32_t *__restrict__ row_cnt) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const uint32_t q = ila[p];
    row_off[q] = base + start[p];
    row_cnt[q] = cnt[p];
}

__global__ __launch_bounds__(256) void k_u32_u64(const uint32_t *__restrict__ in, uint32_t n,
                                                 uint64_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = in[i];
}

template <int KB>
hegj_status launch_gb(const GbArgs &a, uint32_t blocks, hipStream_t s) {
    k_gb_join<KB><<<blocks, 64 * gb_waves(KB), 0, s>>>(a);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

template <int KB>
hegj_status occ_gb(int device, uint32_t *blocks) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(&k_gb_join<KB>), 64 * gb_waves(KB), 0));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    *blocks = (uint32_t)std::max(per_cu, 1) * (uint32_t)prop.multiProcessorCount;
    return HEGJ_OK;
}

}  // namespace

// K blocks (of 32) instantiated for the MFMA refine.
static const int kGbKBs[] = {2, 3, 4, 6, 8, 9, 12};

uint32_t gb_unit_queries(uint32_t kb) { return gb_waves((int)kb) * gb_rt((int)kb) * 16; }

uint32_t gb_kblocks(uint32_t dim) {
    uint32_t need = (3 * (dim + 2) + 31) / 32;
    for (int kb : kGbKBs)
        if ((uint32_t)kb >= need) return (uint32_t)kb;
    return 0;  // too many dims for the instantiated kernels
}

// Error coefficient A of the guard band (see the file comment): delta_p =
// A * N_p bounds each point's share of |s' - acc| per unit of N_q + N_c.
double gb_coefficient(uint32_t dim, uint32_t kp) {
    const double ub = std::ldexp(1.0, -8), u32 = std::ldexp(1.0, -24), u64 = std::ldexp(1.0, -53);
    const double up = std::ldexp(1.0, -23);  // FP32 accumulation, any rounding mode
    const double nadd = kp + kp / 32.0 + 8.0;
    const double gamma = nadd * up / (1.0 - nadd * up);
    const double eta = 1.05 * (ub + u32) * (ub + u32) + 2.0 * u64;
    const double cll = 1.05 * (ub + u32) * (ub + u32);
    return 1.01 * (2.05 * gamma + cll + 4.04 * eta + 2.0 * eta * eta + 4.04 * (dim + 2) * u64);
}

hegj_status gb_prepare(const GridGeom &g, const DevIndex &ix, uint32_t kp, uint16_t **Av,
                       uint16_t **Bv, float **w, hipStream_t s) {
    const uint32_t n = g.n, dim = g.dim;
    // centre = per-dim mean
    uint32_t blocks = std::min<uint32_t>(ceil_div(n, 256), 1024);
    Scratch sc;
    double *partial, *mean;
    HIP_TRY(sc.get(&partial, (size_t)blocks * dim));
    HIP_TRY(sc.get(&mean, dim));
    k_dim_sums<<<blocks, 256, 0, s>>>(ix.soa, n, dim, partial);
    HIP_TRY(hipGetLastError());
    std::vector<double> hp((size_t)blocks * dim), hm(dim, 0.0);
    HIP_TRY(hipMemcpyAsync(hp.data(), partial, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (uint32_t b = 0; b < blocks; b++)
        for (uint32_t l = 0; l < dim; l++) hm[l] += hp[(size_t)b * dim + l];
    for (uint32_t l = 0; l < dim; l++) hm[l] /= n;
    HIP_TRY(hipMemcpyAsync(mean, hm.data(), sizeof(double) * dim, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMalloc(Av, sizeof(uint16_t) * (size_t)n * kp));
    HIP_TRY(hipMalloc(Bv, sizeof(uint16_t) * (size_t)n * kp));
    HIP_TRY(hipMalloc(w, sizeof(float) * n));
    k_gb_prepare<<<ceil_div(n, 256), 256, 0, s>>>(ix.soa, n, dim, kp, mean, gb_coefficient(dim, kp),
                                                 *Av, *Bv, *w);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

hegj_status gb_join(const GridGeom &g, const DevIndex &ix, const uint2 *units, uint32_t n_units,
                    const uint16_t *Av, const uint16_t *Bv, const float *w, uint32_t kp,
                    int device, hipStream_t s, GbOutput *out, uint64_t pair_cap_first, uint32_t parts) {
    const uint32_t kb = kp / 32;
    uint32_t max_blocks = 0;
    switch (kb) {
        case 2: HEGJ_TRY(occ_gb<2>(device, &max_blocks)); break;
        case 3: HEGJ_TRY(occ_gb<3>(device, &max_blocks)); break;
        case 4: HEGJ_TRY(occ_gb<4>(device, &max_blocks)); break;
        case 6: HEGJ_TRY(occ_gb<6>(device, &max_blocks)); break;
        case 8: HEGJ_TRY(occ_gb<8>(device, &max_blocks)); break;
        case 9: HEGJ_TRY(occ_gb<9>(device, &max_blocks)); break;
        case 12: HEGJ_TRY(occ_gb<12>(device, &max_blocks)); break;
        default: return fail(HEGJ_ERR_UNSUPPORTED, "no MFMA kernel for this dimensionality");
    }
    const uint32_t n = g.n;
    uint64_t pair_cap = pair_cap_first ? std::max<uint64_t>(pair_cap_first, 4096)
                                       : std::max<uint64_t>((uint64_t)n * 48, 1u << 22);
    uint64_t band_cap = std::max<uint64_t>((uint64_t)n * 8, 1u << 22);
    struct Ctr {
        uint32_t queue_heads[kMaxParts], overflow, pad;
        unsigned long long pair_cursor, band_cursor, n_accept, n_band, evals;
    };
    parts = std::max<uint32_t>(1, std::min<uint32_t>(parts, (uint32_t)kMaxParts));
    if (n_units < parts * 64) parts = 1;
    Scratch sc;
    Ctr *d;
    HIP_TRY(sc.get(&d, 1));
    Ctr c{};
    for (int attempt = 0; attempt < 4; attempt++) {
        uint64_t *pairs = nullptr, *band = nullptr;
        Scratch it;
        HIP_TRY(it.get(&pairs, pair_cap));
        HIP_TRY(it.get(&band, band_cap));
        HIP_TRY(hipMemsetAsync(d, 0, sizeof(Ctr), s));
        GbArgs a{};
        a.g = g;
        a.ix = ix;
        a.units = units;
        a.n_units = n_units;
        a.queue_heads = d->queue_heads;
        a.n_parts = parts;
        for (uint32_t p = 0; p < parts; p++)
            a.part_end[p] = (uint32_t)((uint64_t)n_units * (p + 1) / parts);
        a.Av = Av;
        a.Bv = Bv;
        a.w = w;
        a.kp = kp;
        a.pairs = pairs;
        a.pair_cap = pair_cap;
        a.pair_cursor = &d->pair_cursor;
        a.band = band;
        a.band_cap = band_cap;
        a.band_cursor = &d->band_cursor;
        a.n_accept = &d->n_accept;
        a.n_band = &d->n_band;
        a.evals = &d->evals;
        a.overflow = &d->overflow;
        const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(max_blocks, n_units));
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, s));
        switch (kb) {
            case 2: HEGJ_TRY(launch_gb<2>(a, blocks, s)); break;
            case 3: HEGJ_TRY(launch_gb<3>(a, blocks, s)); break;
            case 4: HEGJ_TRY(launch_gb<4>(a, blocks, s)); break;
            case 6: HEGJ_TRY(launch_gb<6>(a, blocks, s)); break;
            case 8: HEGJ_TRY(launch_gb<8>(a, blocks, s)); break;
            case 9: HEGJ_TRY(launch_gb<9>(a, blocks, s)); break;
            case 12: HEGJ_TRY(launch_gb<12>(a, blocks, s)); break;
        }
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipMemcpyAsync(&c, d, sizeof(Ctr), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        float kms = 0;
        HIP_TRY(hipEventElapsedTime(&kms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (c.overflow) {
            pair_cap = std::max(pair_cap * 2, (uint64_t)((c.n_accept + c.n_band) * 1.25) + (1u << 22));
            band_cap = std::max(band_cap * 2, (uint64_t)(c.n_band * 1.25) + (1u << 22));
            continue;
        }
        // exact FP64 recheck of the band
        uint64_t nb = c.band_cursor;
        if (nb) {
            k_gb_recheck<<<ceil_div(nb, 256), 256, 0, s>>>(band, nb, ix.soa, n, g.dim, g.thresh, pairs,
                                                          pair_cap, &d->pair_cursor, &d->n_accept,
                                                          &d->overflow);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(&c, d, sizeof(Ctr), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (c.overflow) {
                pair_cap = std::max(pair_cap * 2, (uint64_t)(c.n_accept * 1.25) + (1u << 22));
                continue;
            }
        }
        // sort keys (query position, candidate position): rows in cell order
        const uint64_t ntot = c.pair_cursor;
        const uint64_t nvalid = c.n_accept;
        if (ntot > 0xFFFFFFFFull) return fail(HEGJ_ERR_UNSUPPORTED, "more than 2^32 pair slots");
        int qbits = 0;
        while (qbits < 32 && ((uint64_t)n >> qbits) != 0) qbits++;
        uint64_t *keys_alt;
        uint32_t *vals, *vals_alt;
        HIP_TRY(it.get(&keys_alt, ntot + 1));
        HIP_TRY(it.get(&vals, ntot + 1));
        HIP_TRY(it.get(&vals_alt, ntot + 1));
        HIP_TRY(hipMemsetAsync(vals, 0, sizeof(uint32_t) * (ntot + 1), s));
        HEGJ_TRY(radix_sort_u64(pairs, vals, keys_alt, vals_alt, (uint32_t)ntot, 32 + qbits, s));
        // rows
        uint32_t *cnt;
        uint64_t *cnt64, *start;
        HIP_TRY(it.get(&cnt, n));
        HIP_TRY(it.get(&cnt64, n));
        HIP_TRY(it.get(&start, (size_t)n + 1));
        HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * n, s));
        if (nvalid) k_gb_count<<<ceil_div(nvalid, 256), 256, 0, s>>>(pairs, nvalid, cnt);
        k_u32_u64<<<ceil_div(n, 256), 256, 0, s>>>(cnt, n, cnt64);
        HIP_TRY(hipGetLastError());
        HEGJ_TRY(scan_u64(cnt64, start, n, start + n, s));
        HIP_TRY(hipMalloc(&out->buf, sizeof(uint32_t) * std::max<uint64_t>(nvalid, 1)));
        if (nvalid)
            k_gb_emit<<<ceil_div(nvalid, 256), 256, 0, s>>>(pairs, nvalid, ix.ila, out->buf, out->hash);
        k_gb_rows<<<ceil_div(n, 256), 256, 0, s>>>(cnt, start, n, ix.ila, 0, out->row_off, out->row_cnt);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
        out->pairs = nvalid;
        out->band = c.n_band;
        out->evals = c.evals;
        out->kernel_ms = kms;
        return HEGJ_OK;
    }
    return fail(HEGJ_ERR_CAPACITY, "MFMA refine: pair list overflowed repeatedly");
}

}  // namespace hegj

extern "C" hegj_status hegj_gb_operands(const double *pts, uint32_t n, uint32_t dim, const double *centre,
                                        double coef_scale, uint16_t *A, uint16_t *B, float *w, uint32_t *kp,
                                        double *coef) {
    using namespace hegj;
    const uint32_t kb = gb_kblocks(dim);
    if (kb == 0) return fail(HEGJ_ERR_UNSUPPORTED, "no MFMA refine for this dimensionality");
    const uint32_t kpv = kb * 32;
    const double c = gb_coefficient(dim, kpv) * coef_scale;
    if (kp) *kp = kpv;
    if (coef) *coef = c;
    if (!A && !B && !w) return HEGJ_OK;
    if (!pts || !centre || !A || !B || !w) return fail(HEGJ_ERR_INVALID_ARG, "NULL argument");
    for (uint32_t p = 0; p < n; p++)
        gb_prepare_point(pts + (uint64_t)p * dim, 1, dim, kpv, centre, c, A + (uint64_t)p * kpv,
                         B + (uint64_t)p * kpv, w + p);
    return HEGJ_OK;
}
