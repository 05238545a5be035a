// hegj_join.hip -- workload histogram and the FP64 epsilon-grid join.
//
// sortByWorkLoadGlobal (kernel.cu:74-144) becomes k_cell_workload, one thread
// per non-empty cell walking the 3^(k-1) stencil *rows* (a row is a run of
// cells that differ only in indexed coordinate 0, so the three dim-0
// neighbours of a stencil position are one contiguous range of cells and of
// points): one hash probe and two short binary searches per row instead of
// two binary searches over all cells per stencil cell.
//
// kernelNDGridIndexGlobal / evaluateCell / evalPoint (kernel.cu:727-802,
// 378-412, 333-370) become k_join: a persistent kernel in which each 64-lane
// wavefront takes work units (a non-empty cell and up to QB of its points as
// queries) from a device queue.  Per unit the wave
//   1. enumerates the stencil rows, one row per lane, into a compacted LDS
//      list of candidate point ranges (cell-order positions);
//   2. streams the candidates 64 at a time: lane t loads candidate t's
//      coordinates from the cell-ordered SoA copy (coalesced), evaluates the
//      exact FP64 chain acc = fma(c-q, c-q, acc) in dimension order against
//      every query of the unit (query coordinates are wave-uniform), and the
//      predicate acc <= T (T = largest s with sqrt(s) <= eps, so this equals
//      sqrt(acc) <= eps exactly); a ballot gives the 64-bit hit mask, which
//      is kept in LDS, its popcount is the query's running count;
//   3. reserves one contiguous segment for all its queries' neighbours with
//      a single 64-bit atomic, then writes each query's neighbours (as
//      original point ids) contiguously: lane positions come from popcounts
//      of the kept masks, so no per-pair atomics and no sort are needed.  A
//      unit that does not fit in the current result batch is deferred to the
//      next batch (exact, no estimate can overflow a buffer).
// The per-query (offset, count) table is the reference's neighborTable
// (GPU.cu:1731-1776) with a single data array per batch.
#include <type_traits>
#include <utility>

#include "hegj_internal.h"

namespace hegj {


namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kWavesPerBlockQ = kWavesPerBlock;
// dynamic LDS for staged query coordinates (with the static row lists and
// hit masks the workgroup stays within 64 KiB)
constexpr uint32_t kMaxQueryLds = 36u << 10;

// Compacted stencil-row list of one unit in LDS.
constexpr int kBandList = 64;  // FP32 filter: (query, chunk) pairs awaiting the FP64 recheck
struct RowList {
    uint32_t start[kMaxRows + 1];  // first candidate position of slot s (+1: the FP32 kernel's padding slot)
    uint32_t off[kMaxRows + 1];    // prefix of slot lengths
    float2 gap[16 * kMaxK];      // per (query, indexed dim): squared gaps to the -1 / +1 cell
    uint32_t band[kBandList];    // (chunk << 4) | query
};

// Enumerate the stencil rows of cell `cell` into rl; returns the number of
// candidates M and the number of slots through *nslots.
//
// Cell-box pruning: a stencil cell whose box lies farther than eps from every
// query of the unit cannot hold a neighbour, so it is left out.  With f_j the
// query's position inside its cell along indexed dim j (in units of eps), the
// gap to the neighbour cell at offset -1 is f_j, at +1 it is 1 - f_j; a cell
// is kept when min over the unit's queries of sum_j gap_j^2 <= T / eps^2.
// Each gap is shrunk by 1e-6 + 2e-15 (c_j + 2) (the FP64 rounding of the
// cell assignment (x - min) / eps at coordinate c_j) and the bound raised by
// 1e-5 relative (FP32 squares and sums), so a kept set can only be larger
// than the exact one: the pair set is unchanged, fewer candidates are
// evaluated.  All points of a unit lie in one cell.
template <int QB>
__device__ uint32_t build_rows(const JoinArgs &a, uint32_t cell, RowList &rl, uint32_t *nslots,
                               bool *live, uint32_t qpos0, uint32_t nq) {
    const GridGeom &g = a.g;
    const DevIndex &ix = a.ix;
    const int lane = threadIdx.x & 63;
    uint32_t c[kMaxK], lo[kMaxK], hi[kMaxK];
    decode_cell(ix.cell_key[cell], g, c);
    bool ok = true;
    for (uint32_t j = 0; j < g.k; j++) ok &= stencil_dim(c[j], g.nCells[j], g.emulate_cell0, &lo[j], &hi[j]);
    *live = ok;
    if (!ok) {
        *nslots = 0;
        return 0;
    }
    // squared gaps of query q to its -1 / +1 neighbour along indexed dim j,
    // into rl.gap[q * kMaxK + j] (read back below as wave-uniform broadcasts)
    static_assert(QB <= 16, "gap table holds 16 queries");
    const uint32_t k = g.k;
    for (uint32_t e = (uint32_t)lane; e < nq * k; e += 64) {
        const uint32_t qi = e / k, j = e - qi * k;
        uint32_t cj = 0;
#pragma unroll
        for (int jj = 0; jj < kMaxK; jj++)
            if ((uint32_t)jj == j) cj = c[jj];
        const double x = ix.soa[(uint64_t)j * g.n + qpos0 + qi];
        double f = (x - g.minArr[j]) / g.eps - (double)cj;
        f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
        // DTYPE=float assigns cells in FP32: two roundings of relative 2^-24
        // move a point by at most 1.2e-7 (c_j + 2) cells
        const double mg = g.f32dt ? 1e-6 + 3e-7 * ((double)cj + 2.0) : 1e-6 + 2e-15 * ((double)cj + 2.0);
        const double gm = f - mg > 0.0 ? f - mg : 0.0, gp = 1.0 - f - mg > 0.0 ? 1.0 - f - mg : 0.0;
        rl.gap[qi * kMaxK + j] = make_float2((float)(gm * gm), (float)(gp * gp));
    }
    wave_lds_sync();
    // DTYPE=float: the FP32 chain accepts squared distances up to about
    // (1 + (dim + 4) 2^-23) T
    const float rel = 1e-5f + (g.f32dt ? 1.2e-7f * (float)(g.dim + 4) : 0.0f);
    const float thr = a.prune ? (float)(g.thresh / (g.eps * g.eps)) * (1.0f + rel) : INFINITY;
    const uint32_t nrows = pow3(k - 1);
    // Lane l takes rows l, l + 64, l + 128, l + 192: each dependent step
    // below (hash probe, row bounds, dim-0 window, cell bounds) issues the
    // loads of all its rows at once, so a unit waits for ~6 round trips to
    // memory instead of ~6 per group of 64 rows.
    constexpr int R = (kMaxRows + 63) / 64;
    uint64_t rkey[R];
    uint32_t wlo[R], whi[R], slot[R];
    bool want[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        const uint32_t r = (uint32_t)lane + 64u * i;
        want[i] = false;
        rkey[i] = 0;
        wlo[i] = lo[0];
        whi[i] = hi[0];
        if (r >= nrows) continue;
        // offset digits of row r along dims 1..k-1 (0: -1, 1: 0, 2: +1)
        uint32_t o[kMaxK];
        uint32_t rr = r;
        bool inside = true;
        uint64_t key = 0, mult = 1;
#pragma unroll
        for (int j = 1; j < kMaxK; j++) {
            o[j] = rr % 3;
            rr /= 3;
            if ((uint32_t)j < k) {
                const uint32_t cj = c[j] + o[j] - 1;  // wraps for c[j] == 0, rejected here
                inside &= cj >= lo[j] && cj <= hi[j];
                key += (uint64_t)cj * mult;
                mult *= g.nCells[j];
            }
        }
        if (!inside) continue;
        // cell-box pruning: the row's squared box distance to the nearest
        // query (mid: dim-0 offset 0; left / right: offset -1 / +1)
        float mmid = 3.4e38f, mleft = 3.4e38f, mright = 3.4e38f;
#pragma unroll 1
        for (uint32_t qi = 0; qi < nq; qi++) {
            float rd = 0.0f;
#pragma unroll
            for (int j = 1; j < kMaxK; j++) {
                if ((uint32_t)j >= k) break;
                const float2 v = rl.gap[qi * kMaxK + j];
                rd += o[j] == 0 ? v.x : (o[j] == 2 ? v.y : 0.0f);
            }
            const float2 v0 = rl.gap[qi * kMaxK];
            mmid = fminf(mmid, rd);
            mleft = fminf(mleft, rd + v0.x);
            mright = fminf(mright, rd + v0.y);
        }
        if (!(mmid <= thr)) continue;
        if (!(mleft <= thr) && wlo[i] < c[0]) wlo[i] = c[0];
        if (!(mright <= thr) && whi[i] > c[0]) whi[i] = c[0];
        want[i] = true;
        rkey[i] = key;
    }
    // hash probes: first slots together, collisions (rare) one by one
    uint64_t hk[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        slot[i] = hash_slot(rkey[i]) & ix.hmask;
        hk[i] = want[i] ? ix.hkeys[slot[i]] : ~0ull;
    }
    uint32_t row[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        uint32_t sl = slot[i];
        uint64_t q = hk[i];
        while (want[i] && q != rkey[i] && q != ~0ull) {
            sl = (sl + 1) & ix.hmask;
            q = ix.hkeys[sl];
        }
        slot[i] = sl;
        want[i] = want[i] && q == rkey[i];
    }
#pragma unroll
    for (int i = 0; i < R; i++) row[i] = want[i] ? ix.hvals[slot[i]] : 0u;
    uint32_t rs[R], re[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        rs[i] = want[i] ? ix.row_start[row[i]] : 0u;
        re[i] = want[i] ? ix.row_start[row[i] + 1] : 0u;
    }
    // dim-0 window [wlo, whi] inside the row: rows of up to 4 cells compare
    // their keys loaded together, longer rows binary-search
    constexpr int kShort = 4;
    uint32_t ca[R], cb[R];
    uint64_t ck[R][kShort];
#pragma unroll
    for (int i = 0; i < R; i++) {
        const uint32_t len = re[i] - rs[i];
#pragma unroll
        for (int j = 0; j < kShort; j++)
            ck[i][j] = (len > 0 && len <= (uint32_t)kShort) ? ix.cell_key[rs[i] + min((uint32_t)j, len - 1)] : 0ull;
    }
#pragma unroll
    for (int i = 0; i < R; i++) {
        const uint32_t len = re[i] - rs[i];
        const uint64_t base = rkey[i] * g.nCells[0];
        uint32_t na = 0, nb = 0;
#pragma unroll
        for (int j = 0; j < kShort; j++) {
            na += ((uint32_t)j < len && ck[i][j] < base + wlo[i]) ? 1u : 0u;
            nb += ((uint32_t)j < len && ck[i][j] <= base + whi[i]) ? 1u : 0u;
        }
        ca[i] = rs[i] + na;
        cb[i] = rs[i] + nb;
        if (len > (uint32_t)kShort) {
            ca[i] = lower_bound_u64(ix.cell_key, rs[i], re[i], base + wlo[i]);
            cb[i] = lower_bound_u64(ix.cell_key, ca[i], re[i], base + whi[i] + 1);
        }
    }
    uint32_t pstart[R], plen[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        const bool any = cb[i] > ca[i];
        pstart[i] = any ? ix.cell_start[ca[i]] : 0u;
        plen[i] = any ? ix.cell_start[cb[i]] - pstart[i] : 0u;
    }
    // compact the non-empty rows, in row order, into the LDS list
    uint32_t slots = 0, total = 0;
#pragma unroll
    for (int i = 0; i < R; i++) {
        if (64u * i >= nrows) break;
        const uint64_t nz = __ballot(plen[i] > 0);
        uint32_t x = plen[i];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (plen[i] > 0) {
            const uint32_t s = slots + lanes_below(nz);
            rl.start[s] = pstart[i];
            rl.off[s] = total + x - plen[i];
        }
        slots += (uint32_t)__popcll(nz);
        total += __shfl(x, 63, 64);
    }
    if (lane == 0) rl.off[slots] = total;
    *nslots = slots;
    return total;
}

// Candidate position of flattened candidate index t (t < M), by binary search
// over the slot prefix.
__device__ __forceinline__ uint32_t cand_pos(const RowList &rl, uint32_t nslots, uint32_t t) {
    uint32_t lo = 0, hi = nslots;  // last slot with off <= t
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (rl.off[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return rl.start[lo] + (t - rl.off[lo]);
}


static_assert(HEGJ_MAX_POINTS < (1u << 29), "SoA byte offsets must fit 32 bits");

// Plain global loads and stores (the explicit address space gives global_*
// instructions: flat ones would also count against LDS waits).
template <typename T>
__device__ __forceinline__ T gld(const T *p) {
    return *(__attribute__((address_space(1))) const T *)p;
}
template <typename T>
__device__ __forceinline__ void gst(T *p, T v) {
    *(__attribute__((address_space(1))) T *)p = v;
}

// Load from global memory at a wave-uniform base plus a 32-bit byte offset:
// with the address space explicit this is one global_load with the base in
// scalar registers and no per-lane 64-bit address arithmetic.
template <typename T>
__device__ __forceinline__ T gload_off(const T *base, uint32_t byte_off) {
    typedef __attribute__((address_space(1))) const char gchar;
    typedef __attribute__((address_space(1))) const T gT;
    return *(gT *)((gchar *)base + byte_off);
}

// Candidate coordinates: D doubles from the cell-ordered SoA copy.
template <int D>
__device__ __forceinline__ void load_cand(const JoinArgs &a, uint32_t pos, bool valid, double *x) {
    // unconditional loads (invalid lanes read position 0, always in range, and
    // are masked by the caller): no branch, so the compiler can wait for
    // exactly the chunk it needs instead of every load in flight
    if constexpr (D <= (int)kAosMaxDim) {
        // point-major copy: the D coordinates of a candidate are one run of
        // bytes, and a slot of consecutive candidates one run of lines
        typedef double gd2 __attribute__((ext_vector_type(2)));
        typedef __attribute__((address_space(1))) const gd2 gd2g;
        typedef __attribute__((address_space(1))) const double gdg;
        const uint64_t p = valid ? pos : 0u;
        if constexpr (D % 2 == 0) {
            gd2g *src = (gd2g *)(a.ix.aos + p * D);
#pragma unroll
            for (int l = 0; l < D / 2; l++) {
                const gd2 v = src[l];
                x[2 * l] = v.x;
                x[2 * l + 1] = v.y;
            }
        } else {
            gdg *src = (gdg *)(a.ix.aos + p * D);
#pragma unroll
            for (int l = 0; l < D; l++) x[l] = src[l];
        }
    } else {
        // scalar per-dim base + one 32-bit byte offset for all dims (n <= HEGJ_MAX_POINTS)
        const uint64_t n = a.g.n;
        uint32_t off = (valid ? pos : 0u) * (uint32_t)sizeof(double);
        __asm__("" : "+v"(off));  // an opaque 32-bit offset: the loads take the scalar-base form
#pragma unroll
        for (int l = 0; l < D; l++) x[l] = gload_off(a.ix.soa + (uint64_t)l * n, off);
    }
}

// llvm.amdgcn.writelane (v_writelane_b32): `old` with lane `lane` replaced by
// the wave-uniform `src`.  Called as the intrinsic, not as inline asm, so
// the compiler sees the instruction and inserts the wait states it needs
// after a VALU write of `src` (v_cmp, v_readfirstlane); inline asm hides the
// read and can pick up the register's previous value.
extern "C" __device__ int hegj_writelane_i32(int src, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// v with lane `lane` replaced by the wave-uniform s.
__device__ __forceinline__ uint32_t set_lane_u32(uint32_t v, uint32_t s, uint32_t lane) {
    return (uint32_t)hegj_writelane_i32((int)s, (int)lane, (int)v);
}
__device__ __forceinline__ uint64_t set_lane_u64(uint64_t v, uint64_t s, int lane) {
    const uint32_t lo = set_lane_u32((uint32_t)v, (uint32_t)s, (uint32_t)lane);
    const uint32_t hi = set_lane_u32((uint32_t)(v >> 32), (uint32_t)(s >> 32), (uint32_t)lane);
    return ((uint64_t)hi << 32) | lo;
}

// acc[qi] for one candidate (coordinates in x) against the unit's queries:
// acc = fma(c - q, c - q, acc) in dimension order (kernel.cu:350-353 with
// FMA contraction).  q holds the unit's query coordinates, loaded once per
// unit (wave-uniform: scalar registers or LDS), so the candidate loads that
// are in flight are never waited on here.
// All QB queries are evaluated (padding queries hold zeros and are masked by
// the caller): no branch between the candidate loads and their use.
template <int D, int QB>
__device__ __forceinline__ void dist_regs(const double *x, const double (*q)[D], double *acc) {
#pragma unroll
    for (int qi = 0; qi < QB; qi++) {
        acc[qi] = 0.0;
#pragma unroll
        for (int l = 0; l < D; l++) {
            double dd = x[l] - q[qi][l];
            acc[qi] = __builtin_fma(dd, dd, acc[qi]);
        }
    }
}

// Same with the query coordinates in LDS (qs[qi * dim + l]).  The empty
// asm with a memory clobber keeps each query's LDS reads next to their use
// (hoisted, QB * D doubles of them would not fit in registers).
template <int D, int QB>
__device__ __forceinline__ void dist_regs_lds(const double *x, const double *qs, double *acc) {
#pragma unroll
    for (int qi = 0; qi < QB; qi++) {
        __asm__ volatile("" ::: "memory");
        acc[qi] = 0.0;
#pragma unroll
        for (int l = 0; l < D; l++) {
            double dd = x[l] - qs[qi * D + l];
            acc[qi] = __builtin_fma(dd, dd, acc[qi]);
        }
    }
}

// Generic dimension count (runtime or large D): candidate coordinates are
// streamed in tiles of 16 dims, acc keeps the dimension order.  The unit's
// query coordinates come from LDS (qs[qi * dim + l]).
// FLT: DTYPE=float (params.h:24).  The coordinates are floats held in
// doubles; the chain is the reference's FP32 one, d = fl32(c - q),
// acc = fmaf(d, d, acc), returned widened to double (exact).
template <int D, int QB, bool FLT = false>
__device__ __forceinline__ void dist_stream(const JoinArgs &a, uint32_t pos, bool valid,
                                            const double *qs, uint32_t nq, double *acc) {
    const uint32_t dim = D > 0 ? (uint32_t)D : a.g.dim;
    const uint64_t n = a.g.n;
    const double *__restrict__ soa = a.ix.soa;
    constexpr int DT = 16;
    if constexpr (FLT) {
        float accf[QB];
#pragma unroll
        for (int qi = 0; qi < QB; qi++) accf[qi] = 0.0f;
        for (uint32_t d0 = 0; d0 < dim; d0 += DT) {
            float x[DT];
#pragma unroll
            for (int i = 0; i < DT; i++) {
                uint32_t l = d0 + i;
                x[i] = (valid && l < dim) ? (float)soa[(uint64_t)l * n + pos] : 0.0f;
            }
#pragma unroll
            for (int qi = 0; qi < QB; qi++) {
                if ((uint32_t)qi < nq) {
#pragma unroll
                    for (int i = 0; i < DT; i++) {
                        uint32_t l = d0 + i;
                        if (l < dim) {
                            const float dd = x[i] - (float)qs[qi * dim + l];
                            accf[qi] = __builtin_fmaf(dd, dd, accf[qi]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int qi = 0; qi < QB; qi++) acc[qi] = (double)accf[qi];
        return;
    }
#pragma unroll
    for (int qi = 0; qi < QB; qi++) acc[qi] = 0.0;
    for (uint32_t d0 = 0; d0 < dim; d0 += DT) {
        double x[DT];
#pragma unroll
        for (int i = 0; i < DT; i++) {
            uint32_t l = d0 + i;
            x[i] = (valid && l < dim) ? soa[(uint64_t)l * n + pos] : 0.0;
        }
#pragma unroll
        for (int qi = 0; qi < QB; qi++) {
            if ((uint32_t)qi < nq) {
#pragma unroll
                for (int i = 0; i < DT; i++) {
                    uint32_t l = d0 + i;
                    if (l < dim) {
                        double dd = x[i] - qs[qi * dim + l];
                        acc[qi] = __builtin_fma(dd, dd, acc[qi]);
                    }
                }
            }
        }
    }
}

// Query coordinates of a unit kept in scalar registers when QB * D doubles
// leave room for the rest of the kernel's scalar state (48 of the ~100
// SGPRs; more spills them to VGPR lanes and every use becomes two
// v_readlane); otherwise they are staged in LDS, QB * dim doubles per wave.
template <int D, int QB>
constexpr bool q_in_sgprs() {
    return D > 0 && D <= 32 && QB * D <= 24;
}

template <int D, int QB>
constexpr uint32_t q_lds_bytes(uint32_t dim) {
    return q_in_sgprs<D, QB>() ? 0u : (uint32_t)(kWavesPerBlockQ * QB * (D > 0 ? D : dim) * sizeof(double));
}

// Hit masks of one unit: mask[ch * nq + qi], one 64-bit ballot per
// (query, candidate chunk).  Kept in LDS when they fit, else in a per-wave
// global scratch slice (written and read back by the same wave, so L2
// resident), else (very large units) recomputed in the write phase.
constexpr int kMaskSlots = 512;
struct HitStore {
    uint64_t w[kMaskSlots];
};

constexpr int kRegDims = 32;  // dims kept in registers (with prefetch)

// Takes the next grab of units: from the workgroup's own range first (its
// XCD's share of the space, for L2 locality), then from the others.
__device__ __forceinline__ bool next_grab(const JoinArgs &a, uint32_t &part, uint32_t &tries,
                                          uint32_t &grab, uint32_t &gend) {
    const int lane = threadIdx.x & 63;
    while (tries < a.n_parts) {
        uint32_t g = 0;
        if (lane == 0) g = atomicAdd(&a.queue_heads[part], a.units_per_grab);
        g = wave_bcast_u32(g) + a.part_begin[part];
        if (g < a.part_end[part]) {
            grab = g;
            gend = min(g + a.units_per_grab, a.part_end[part]);
            return true;
        }
        part = part + 1 == a.n_parts ? 0 : part + 1;
        tries++;
    }
    return false;
}

// IND: the arguments are read from device memory (ap) instead of the kernel
// argument block, so one captured graph serves every call (the host rewrites
// the arguments between replays).
// Waves per SIMD the register budget must allow: 4 (<= 128 VGPRs) for the
// small-D instances with their queries in scalar registers, whose loop is
// latency-bound and needs occupancy; 2 for the other small-D instances up to
// 8 queries (queries read from LDS next to their use need more vector
// registers); 1 above.
template <int D, int QB>
constexpr int join_min_waves() {
    return (D > 0 && D <= 8 && QB <= 8) ? (q_in_sgprs<D, QB>() ? 4 : 2) : 1;
}

// The exact FP64 chain for every dimensionality (the FP32 filter is
// k_join_f32 below).
// FLT: DTYPE=float in the generic instance (D == 0): the FP32 chain over the
// float coordinates held in the FP64 copies.
template <int D, int QB, bool IND = false, bool FLT = false>
__global__ __launch_bounds__(64 * kWavesPerBlock, (join_min_waves<D, QB>())) void k_join(
    JoinArgs av, const JoinArgs *__restrict__ ap) {
    const JoinArgs &a = IND ? *ap : av;
    static_assert(!FLT || D == 0, "DTYPE=float: generic instance only");
    __shared__ RowList rls[kWavesPerBlock];
    __shared__ HitStore hs_all[kWavesPerBlock];
    extern __shared__ double q_dyn[];  // QB * dim query coordinates per wave
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    RowList &rl = rls[wid];
    uint64_t *const hs_lds = hs_all[wid].w;
    uint64_t *const hs_glb = a.gmask + (uint64_t)(blockIdx.x * (blockDim.x >> 6) + wid) * a.gmask_slots;
    const double T = a.g.thresh;
    constexpr bool kRegs = (D > 0 && D <= kRegDims);
    constexpr int DR = kRegs ? D : 1;
    uint64_t my_evals = 0, my_pairs = 0;
    constexpr bool kQS = q_in_sgprs<D, QB>();
    const uint32_t dimq = D > 0 ? (uint32_t)D : a.g.dim;
    double *const qs = q_dyn + (size_t)wid * QB * dimq;
    double qreg[kQS ? QB : 1][kQS ? DR : 1];
    uint32_t part = (a.part_by_xcc ? xcc_id() : blockIdx.x) % a.n_parts;
    uint32_t tries = 0, grab = 0, gend = 0;

    while (next_grab(a, part, tries, grab, gend)) {
        for (uint32_t ui = grab; ui < gend; ui++) {
            const uint2 unit = a.units[ui];
            const uint32_t cell = wave_bcast_u32(unit.x), qoff = wave_bcast_u32(unit.y);
            const uint32_t cstart = a.ix.cell_start[cell];
            const uint32_t csize = a.ix.cell_start[cell + 1] - cstart;
            // wave-uniform by construction; readfirstlane lets the compiler
            // keep them (and the masks derived from them) in scalar registers
            const uint32_t nq = (uint32_t)__builtin_amdgcn_readfirstlane((int)min((uint32_t)QB, csize - qoff));
            const uint32_t qpos0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(cstart + qoff));
            uint32_t nslots;
            bool live;
            const uint32_t M = (uint32_t)__builtin_amdgcn_readfirstlane(
                (int)build_rows<QB>(a, cell, rl, &nslots, &live, qpos0, nq));
            nslots = (uint32_t)__builtin_amdgcn_readfirstlane((int)nslots);
            // the unit's query coordinates, once: lane i loads element i
            if constexpr (kQS) {
                constexpr int NQD = kQS ? QB * DR : 1;
#pragma unroll
                for (int i0 = 0; i0 < NQD; i0 += 64) {
                    const int i = i0 + lane;
                    const int qi = i / DR, l = i % DR;
                    double v = 0.0;
                    if (i < NQD && (uint32_t)qi < nq)
                        v = a.ix.soa[(uint64_t)l * a.g.n + qpos0 + qi];
#pragma unroll
                    for (int j = i0; j < NQD && j < i0 + 64; j++) {
                        const uint64_t u = (uint64_t)__double_as_longlong(v);
                        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, j - i0);
                        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), j - i0);
                        qreg[kQS ? j / DR : 0][kQS ? j % DR : 0] =
                            __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
                    }
                }
            } else {
                for (uint32_t i = lane; i < QB * dimq; i += 64) {
                    const uint32_t qi = i / dimq, l = i - qi * dimq;
                    qs[i] = qi < nq ? a.ix.soa[(uint64_t)l * a.g.n + qpos0 + qi] : 0.0;
                }
            }
            wave_lds_sync();
            const uint32_t nchunks = (M + 63) / 64;
            const uint32_t nmask = nq * nchunks;
            const bool keep = nmask <= a.gmask_slots || nmask <= (uint32_t)kMaskSlots;
            // masks in LDS (ds_* instructions) unless the unit has more than
            // fit; every access names its memory, so no generic (flat) access
            // enters the loops: a flat store would count against both the
            // vector-memory and the LDS counters, and the candidate ring's
            // loads could then only be waited for all at once
            const bool lds_masks = nmask <= (uint32_t)kMaskSlots;
            auto mask_put = [&](uint32_t i, uint64_t v) __attribute__((always_inline)) {
                if (lds_masks)
                    hs_lds[i] = v;
                else
                    gst(hs_glb + i, v);
            };
            auto mask_get = [&](uint32_t i) __attribute__((always_inline)) {
                return lds_masks ? hs_lds[i] : gld(hs_glb + i);
            };
            uint32_t cnt_l = 0;  // lane qi: hits of query qi so far
            uint64_t mine = 0;   // lane qi: query qi's mask of the current chunk

            // the 4-wave instances (128 registers) keep the one-register walk:
            // the cursor's two more registers spill there (6-D FP64: +4%)
            typedef typename std::conditional<(join_min_waves<D, QB>() < 4), SlotCursor, SlotWalk>::type Cursor;
            Cursor cur;
            cur.reset(rl, nslots);
            // Candidate coordinates are prefetched kDepth chunks ahead in a
            // register ring (the loop is unrolled by kDepth so ring indices
            // are static): several chunks of loads stay in flight per wave.
            constexpr int kDepth = kRegs ? (D <= 8 ? 4 : 2) : 1;
            double ring[kDepth][DR];
            if constexpr (kRegs) {
#pragma unroll
                for (int p = 0; p < kDepth; p++) {
                    const uint32_t t = p * 64 + lane;
                    const bool v = t < M;
                    load_cand<DR>(a, v ? cur.pos(rl, t) : 0u, v, ring[p]);
                }
            }
            // Whole groups of kDepth chunks: the chunks past the end of the
            // last group have no valid lane (their ballots are empty) and
            // their prefetches read position 0, so the loop body has no
            // branch around a load and the compiler waits for the oldest
            // chunk only.  Each (query, chunk) ballot goes straight into lane
            // qi of `mine` (one writelane per half), so no mask stays live in
            // scalar registers across queries.
            // The loop is built for QL queries: QB, or fewer when the unit
            // has fewer (cells of one or two points are common), so no FP
            // work is spent on padding queries.  QL only bounds the loops;
            // the query registers and masks keep their QB layout.  (Only the
            // small-D, QB <= 8 instances get the extra copies: each is a
            // fully unrolled loop, and compile time grows with them.)
            auto chunk_loop = [&](auto qlc) __attribute__((always_inline)) {
                constexpr int QL = decltype(qlc)::value;
                for (uint32_t ch0 = 0; ch0 < nchunks; ch0 += kDepth) {
#pragma unroll
                    for (int p = 0; p < kDepth; p++) {
                        const uint32_t ch = ch0 + p;
                        if (!kRegs && ch >= nchunks) break;
                        const bool valid = ch * 64 + lane < M;
                        // single compares feed the ballots; validity is applied
                        // to the scalar masks
                        const uint64_t vmask = __builtin_amdgcn_ballot_w64(valid);
                        {
                            double acc[QL];
                            if constexpr (kRegs) {
                                // use ring[p], then refill it with chunk ch + kDepth
                                // (no copy of registers that still have loads in flight)
                                if constexpr (kQS) {
                                    dist_regs<DR, QL>(ring[p], qreg, acc);
                                } else {
                                    // opaque offset: keeps the compiler from hoisting
                                    // QB * D loop-invariant LDS reads into registers
                                    uint32_t qoff = (uint32_t)wid * QB * DR;
                                    __asm__ volatile("" : "+v"(qoff));
                                    dist_regs_lds<DR, QL>(ring[p], q_dyn + qoff, acc);
                                }
                                const uint32_t tn = (ch + kDepth) * 64 + lane;
                                const bool vn = tn < M;
                                load_cand<DR>(a, vn ? cur.pos(rl, tn) : 0u, vn, ring[p]);
                            } else {
                                const uint32_t t = ch * 64 + lane;
                                const uint32_t pos = valid ? cur.pos(rl, t) : 0u;
                                uint32_t qoff = (uint32_t)wid * QB * dimq;
                                __asm__ volatile("" : "+v"(qoff));
                                dist_stream<D, QB, FLT>(a, pos, valid, q_dyn + qoff, nq, acc);
                            }
#pragma unroll
                            for (int qi = 0; qi < QL; qi++)
                                mine = set_lane_u64(mine, __builtin_amdgcn_ballot_w64(acc[qi] <= T) &
                                                              ((uint32_t)qi < nq ? vmask : 0ull), qi);
                        }
                        cnt_l += (uint32_t)__popcll(mine);
                        if (keep && (uint32_t)lane < nq && ch < nchunks) mask_put(ch * nq + lane, mine);
                    }
                }
            };
            if constexpr (kRegs && D <= 8 && QB >= 4 && QB <= 8) {
                if (nq == 1)
                    chunk_loop(std::integral_constant<int, 1>{});
                else if (nq <= (uint32_t)QB / 2)
                    chunk_loop(std::integral_constant<int, QB / 2>{});
                else
                    chunk_loop(std::integral_constant<int, QB>{});
            } else {
                chunk_loop(std::integral_constant<int, QB>{});
            }
            my_evals += (uint64_t)M * nq;
            // lane qi: first slot of query qi within the unit's segment
            uint32_t incl_l = cnt_l;
#pragma unroll
            for (int o = 1; o < QB; o <<= 1) {
                const uint32_t y = __shfl_up(incl_l, o, 64);
                if (lane >= o) incl_l += y;
            }
            const uint32_t excl_l = incl_l - cnt_l;
            const uint32_t total = __shfl(incl_l, QB - 1, 64);
            if (a.count_only) {
                my_pairs += total;
                continue;
            }
            uint64_t base = 0;
            if (total > 0) {
                if (lane == 0) base = atomicAdd(a.cursor, (unsigned long long)total);
                base = wave_bcast_u64(base);
                if (base + total > a.capacity) {
                    if (lane == 0) a.deferred[atomicAdd(a.n_deferred, 1u)] = ui;
                    continue;
                }
            }
            my_pairs += total;
            wave_lds_sync();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            // row table entries of the unit's queries
            const uint32_t qid_l = (uint32_t)lane < nq ? a.ix.ila[qpos0 + lane] : 0u;
            if ((uint32_t)lane < nq) {
                a.row_off[qid_l] = a.out_base + base + excl_l;
                a.row_cnt[qid_l] = cnt_l;
            }
            if (total == 0) continue;
            // Chunk-major write: one coalesced id load per chunk serves all of
            // the unit's queries, each query's neighbours come out in
            // candidate (cell) order.  The write cursors are wave-uniform
            // (scalar registers), the hit lanes of a mask store directly
            // under that mask.  (The pair hash is a separate pass over the
            // finished rows: here only a few lanes of each instruction would
            // be active.)
            uint64_t oq[QB];
            {
                const uint64_t o_l = base + excl_l;
#pragma unroll
                for (int qi = 0; qi < QB; qi++)
                    oq[qi] = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(o_l >> 32), qi) << 32) |
                             (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)o_l, qi);
            }
            // The loads each chunk depends on -- its neighbour ids
            // (indexLookupArr) and, when kept, its masks (lane qi: query qi's)
            // -- are issued kW chunks ahead, so no iteration waits on a
            // fresh global load.
            constexpr int kW = QB <= 8 ? 8 : 4;
            Cursor cur2;
            cur2.reset(rl, nslots);
            if (keep) {
                uint32_t idr[kW];
                uint64_t mkr[kW];
                auto fetch = [&](uint32_t c, uint32_t &id, uint64_t &mk) {
                    const uint32_t t = c * 64 + lane;
                    id = gld(a.ix.ila + (t < M ? cur2.pos(rl, t) : 0u));
                    mk = (c < nchunks && (uint32_t)lane < nq) ? mask_get(c * nq + lane) : 0ull;
                };
#pragma unroll
                for (int p = 0; p < kW; p++) fetch((uint32_t)p, idr[p], mkr[p]);
                for (uint32_t ch0 = 0; ch0 < nchunks; ch0 += kW) {
#pragma unroll
                    for (int p = 0; p < kW; p++) {
                        const uint32_t ch = ch0 + p;
                        if (ch >= nchunks) break;
                        const uint32_t id = idr[p];
                        const uint64_t mk = mkr[p];
                        fetch(ch + kW, idr[p], mkr[p]);
#pragma unroll
                        for (int qi = 0; qi < QB; qi++) {
                            // lanes >= nq of mk hold zero masks
                            const uint64_t m =
                                ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(mk >> 32), qi) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)mk, qi);
                            if (m == 0) continue;
                            if (__builtin_amdgcn_inverse_ballot_w64(m)) gst(a.out + oq[qi] + lanes_below(m), id);
                            oq[qi] += (uint64_t)__popcll(m);
                        }
                    }
                }
            } else {
                // masks did not fit anywhere (very large units): evaluate
                // each chunk again, one query at a time.  The writes stay
                // inside each query's reserved row [oq, oend); a count that
                // differs from the first pass is reported (a.mismatch), never
                // written past the row.
                uint64_t oend[QB];
                {
                    const uint64_t e_l = base + excl_l + cnt_l;
#pragma unroll
                    for (int qi = 0; qi < QB; qi++)
                        oend[qi] = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(e_l >> 32), qi) << 32) |
                                   (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)e_l, qi);
                }
                bool bad = false;
                for (uint32_t ch = 0; ch < nchunks; ch++) {
                    const uint32_t t = ch * 64 + lane;
                    const bool valid = t < M;
                    const uint32_t pos = valid ? cur2.pos(rl, t) : 0u;
                    const uint32_t id = gld(a.ix.ila + pos);
                    double x[DR];
                    if constexpr (kRegs) load_cand<DR>(a, pos, valid, x);
                    my_evals += (uint64_t)min(64u, M - ch * 64) * nq;
#pragma unroll
                    for (int qi = 0; qi < QB; qi++) {
                        if ((uint32_t)qi >= nq) continue;
                        double acc1 = 0.0;
                        if constexpr (kRegs) {
#pragma unroll
                            for (int l = 0; l < DR; l++) {
                                const double q = kQS ? qreg[kQS ? qi : 0][kQS ? l : 0] : qs[qi * DR + l];
                                const double dd = x[l] - q;
                                acc1 = __builtin_fma(dd, dd, acc1);
                            }
                        } else {
                            dist_stream<D, 1, FLT>(a, pos, valid, qs + qi * dimq, 1, &acc1);
                        }
                        const uint64_t m = __ballot(valid && acc1 <= T);
                        if (m == 0) continue;
                        const uint64_t room = oend[qi] - oq[qi];
                        const uint64_t pc = (uint64_t)__popcll(m);
                        bad |= pc > room;
                        if (__builtin_amdgcn_inverse_ballot_w64(m) && lanes_below(m) < room)
                            gst(a.out + oq[qi] + lanes_below(m), id);
                        oq[qi] += pc < room ? pc : room;
                    }
                }
#pragma unroll
                for (int qi = 0; qi < QB; qi++) bad |= (uint32_t)qi < nq && oq[qi] != oend[qi];
                if (bad && lane == 0) atomicAdd(a.mismatch, 1ull);
            }
        }
    }
    // one atomic per wave for the counters
    if (lane == 0) {
        if (my_evals) atomicAdd(a.evals, (unsigned long long)my_evals);
        if (my_pairs) atomicAdd(a.pairs, (unsigned long long)my_pairs);
    }
}

// ---------------------------------------------------------------------------
// k_join_f32: the FP32-filter join for 2-D and 6-D
// ---------------------------------------------------------------------------
// Same units, stencil rows, queue, thresholds and result layout as k_join
// with the FP32 filter; the candidate loop is built so that a chunk of 64
// candidates costs little beyond the distance arithmetic:
//  - Lane positions come from a chunk table instead of a per-lane search.
//    For each chunk of a window of kTab chunks the table holds S, the slot of
//    the chunk's lane 0, and B, a bitmap whose bit j says that a slot starts
//    at lane j + 1.  Lane l's slot is S + popcount(B & lanes below l) (one
//    v_mbcnt pair) and its byte offset one LDS read away.
//  - Padding: the FP32 copy ends with kAos32Pad points at +inf, and one more
//    slot after the unit's last maps the lanes past its last candidate onto
//    them; padding queries are +inf too.  Such lanes are neither accepted nor
//    in the band, so the loop applies no validity masks.
//  - Band detection is a scalar OR per query; only the rare chunks with band
//    lanes list their (query, chunk) pairs for the FP64 recheck.
//  - A bitmap of the chunks with hits lets the write phase skip the others.
constexpr int kRing = 4;                // candidate chunks in flight per wave
constexpr int kTab = 64;                // chunk-table entries (one per lane)
constexpr int kTabStep = kTab - kRing;  // chunks per table window: the ring's prefetches stay inside it
constexpr int kHitWords = 16;           // hit bitmap: units of up to 1024 chunks
struct F32Lds {
    uint4 tab[kTab];               // {B low, B high, S, 0} of chunk c0 + i
    uint32_t bb[kMaxRows + 1];     // slot s: byte offset of candidate t is bb[s] + t * stride (mod 2^32)
    uint64_t hitw[kHitWords];      // bit c: chunk c has a hit
};

// Orders this wave's LDS accesses (one wave's LDS instructions execute in
// order, so the compiler's ordering is all that is needed).
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Chunk table for chunks c0 .. c0 + kTab - 1.  Slot s (1 <= s <= nslots; slot
// nslots is the padding slot, off[nslots] = M) starts at flattened candidate
// off[s] and sets bit (off[s] - 1) % 64 of chunk (off[s] - 1) / 64.  S of
// chunk c is the number of slots s >= 1 with off[s] <= 64 c: the slot of its
// lane 0 (off[0] = 0 and the offsets increase strictly).
__device__ __forceinline__ void build_tab(F32Lds &f, const RowList &rl, uint32_t nslots, uint32_t c0) {
    const int lane = threadIdx.x & 63;
    wave_lds_order();
    f.tab[lane] = make_uint4(0u, 0u, 0u, 0u);
    wave_lds_order();
    const uint32_t t0 = c0 * 64u, tend = t0 + 64u * kTab;
    uint32_t before = 0;
    for (uint32_t s0 = 1; s0 <= nslots; s0 += 64) {
        const uint32_t s = s0 + (uint32_t)lane;
        const bool in = s <= nslots;
        const uint32_t o = in ? rl.off[s] : 0u;
        before += (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(in && o <= t0));
        if (in && o > t0 && o <= tend) {
            const uint32_t j = o - 1u - t0;
            atomicOr(reinterpret_cast<unsigned long long *>(&f.tab[j >> 6]), 1ull << (j & 63u));
        }
    }
    wave_lds_order();
    const uint4 e = f.tab[lane];
    const uint32_t pc = (uint32_t)__popc(e.x) + (uint32_t)__popc(e.y);
    uint32_t x = pc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    f.tab[lane].z = before + x - pc;
    wave_lds_order();
}

// Slot of the calling lane in chunk c (c0 <= c < c0 + kTab).
__device__ __forceinline__ uint32_t tab_slot(const F32Lds &f, uint32_t c, uint32_t c0) {
    const uint4 e = f.tab[c - c0];
    return __builtin_amdgcn_mbcnt_hi(e.y, __builtin_amdgcn_mbcnt_lo(e.x, e.z));
}

// v with lane L replaced by the wave-uniform s (the lane select is an
// immediate: no scalar move per write).
template <int L>
__device__ __forceinline__ uint64_t put_lane(uint64_t v, uint64_t s) {
    const uint32_t lo = (uint32_t)hegj_writelane_i32((int)(uint32_t)s, L, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)hegj_writelane_i32((int)(uint32_t)(s >> 32), L, (int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// f(std::integral_constant<int, i>) for i = 0 .. N - 1.
template <typename F, int... I>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

// D floats of the point-major FP32 copy at a 32-bit byte offset.
template <int D>
__device__ __forceinline__ void load_cand32_off(const float *aos32, uint32_t byte_off, float *x) {
    typedef float gf2 __attribute__((ext_vector_type(2)));
    const gf2 *base = reinterpret_cast<const gf2 *>(aos32);
#pragma unroll
    for (int l = 0; l < D / 2; l++) {
        const gf2 v = gload_off(base, byte_off + 8u * (uint32_t)l);
        x[2 * l] = v.x;
        x[2 * l + 1] = v.y;
    }
}

template <int D, int QB, bool IND>
__global__ __launch_bounds__(64 * kWavesPerBlock, 4) void k_join_f32(JoinArgs av, const JoinArgs *__restrict__ ap) {
    const JoinArgs &a = IND ? *ap : av;
    static_assert((D == 2 || D == 6) && QB * D <= 48 && QB <= 16, "FP32 filter: 2-D and 6-D");
    __shared__ RowList rls[kWavesPerBlock];
    __shared__ HitStore hs_all[kWavesPerBlock];
    __shared__ F32Lds fls[kWavesPerBlock];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    RowList &rl = rls[wid];
    F32Lds &fl = fls[wid];
    uint64_t *const hs_lds = hs_all[wid].w;
    uint64_t *const hs_glb = a.gmask + (uint64_t)(blockIdx.x * (blockDim.x >> 6) + wid) * a.gmask_slots;
    constexpr uint32_t kStride = D * (uint32_t)sizeof(float);
    constexpr uint32_t kChunkBytes = 64u * kStride;
    const uint32_t lane_off = (uint32_t)lane * kStride;
    const double T = a.g.thresh;
    const float tacc = a.tacc32, trej = a.trej32;
    uint64_t my_evals = 0, my_pairs = 0, my_band = 0;
    float qregf[QB][D];
    uint32_t part = (a.part_by_xcc ? xcc_id() : blockIdx.x) % a.n_parts;
    uint32_t tries = 0, grab = 0, gend = 0;

    while (next_grab(a, part, tries, grab, gend)) {
        for (uint32_t ui = grab; ui < gend; ui++) {
            const uint2 unit = a.units[ui];
            const uint32_t cell = wave_bcast_u32(unit.x), qoff = wave_bcast_u32(unit.y);
            const uint32_t cstart = a.ix.cell_start[cell];
            const uint32_t csize = a.ix.cell_start[cell + 1] - cstart;
            const uint32_t nq = (uint32_t)__builtin_amdgcn_readfirstlane((int)min((uint32_t)QB, csize - qoff));
            const uint32_t qpos0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(cstart + qoff));
            uint32_t nslots;
            bool live;
            const uint32_t M = (uint32_t)__builtin_amdgcn_readfirstlane(
                (int)build_rows<QB>(a, cell, rl, &nslots, &live, qpos0, nq));
            nslots = (uint32_t)__builtin_amdgcn_readfirstlane((int)nslots);
            // the unit's query coordinates, once (lane i loads element i);
            // padding queries are +inf
            {
                constexpr int NQD = QB * D;
                const int qi = lane / D, l = lane % D;
                float v = __builtin_inff();
                if (lane < NQD && (uint32_t)qi < nq) v = a.soa32[(uint64_t)l * a.g.n + qpos0 + qi];
#pragma unroll
                for (int j = 0; j < NQD; j++)
                    qregf[j / D][j % D] = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), j));
            }
            const uint32_t nchunks = (M + 63) / 64;
            const uint32_t nmask = nq * nchunks;
            const bool keep = nmask <= a.gmask_slots || nmask <= (uint32_t)kMaskSlots;
            const bool lds_masks = nmask <= (uint32_t)kMaskSlots;
            auto mask_put = [&](uint32_t i, uint64_t v) __attribute__((always_inline)) {
                if (lds_masks)
                    hs_lds[i] = v;
                else
                    gst(hs_glb + i, v);
            };
            auto mask_get = [&](uint32_t i) __attribute__((always_inline)) {
                return lds_masks ? hs_lds[i] : gld(hs_glb + i);
            };
            const bool hit_bits = nchunks <= 64u * (uint32_t)kHitWords;
            uint32_t cnt_l = 0;  // lane qi: hits of query qi so far
            uint64_t mine = 0;   // lane qi: query qi's mask of the current chunk
            uint32_t nband = 0;  // (query, chunk) pairs with lanes in the band
            uint32_t tab_c0 = 0; // window of the chunk table
            wave_lds_order();
            // slot byte bases; the padding slot (index nslots, off = M) maps
            // candidate t >= M to padding point n + t - M
            for (uint32_t s = (uint32_t)lane; s <= nslots; s += 64)
                fl.bb[s] = s < nslots ? (rl.start[s] - rl.off[s]) * kStride : (a.g.n - M) * kStride;

            // The loop is built for QL queries (QB, QB / 2 or 1: cells of
            // one or two points are common); the query registers and masks
            // keep their QB layout.
            auto chunk_loop = [&](auto qlc) __attribute__((always_inline)) {
                constexpr int QL = decltype(qlc)::value;
                float ring[kRing][D];
                uint64_t hw = 0;  // hit bits of the current 64 chunks
                for (uint32_t c0 = 0; c0 < nchunks; c0 += kTabStep) {
                    build_tab(fl, rl, nslots, c0);
                    tab_c0 = c0;
                    if (c0 == 0) {
#pragma unroll
                        for (int p = 0; p < kRing; p++)
                            load_cand32_off<D>(a.aos32, fl.bb[tab_slot(fl, p, 0)] + (uint32_t)p * kChunkBytes + lane_off,
                                               ring[p]);
                    }
                    const uint32_t cend = min(c0 + (uint32_t)kTabStep, nchunks);
                    for (uint32_t ch0 = c0; ch0 < cend; ch0 += kRing) {
#pragma unroll
                        for (int p = 0; p < kRing; p++) {
                            const uint32_t ch = ch0 + p;
                            float accf[QL];
#pragma unroll
                            for (int qi = 0; qi < QL; qi++) {
                                accf[qi] = 0.0f;
#pragma unroll
                                for (int l = 0; l < D; l++) {
                                    const float dd = ring[p][l] - qregf[qi][l];
                                    accf[qi] = __builtin_fmaf(dd, dd, accf[qi]);
                                }
                            }
                            {
                                // refill the ring slot with chunk ch + kRing only
                                // after the distances that read it (the loop then
                                // carries the ring without register copies)
                                if constexpr (QL > 1) {
#pragma unroll
                                    for (int qi = 0; qi < QL; qi++) __asm__ volatile("" ::"v"(accf[qi]) : "memory");
                                }
                                const uint32_t cn = ch + kRing;
                                load_cand32_off<D>(a.aos32, fl.bb[tab_slot(fl, cn, c0)] + cn * kChunkBytes + lane_off,
                                                   ring[p]);
                            }
                            uint64_t bandm = 0, anyh = 0;
                            static_for<QL>([&](auto qc) __attribute__((always_inline)) {
                                constexpr int qi = decltype(qc)::value;
                                const uint64_t ma = __builtin_amdgcn_ballot_w64(accf[qi] <= tacc);
                                const uint64_t mr = __builtin_amdgcn_ballot_w64(accf[qi] <= trej);
                                bandm |= mr & ~ma;
                                anyh |= ma;
                                mine = put_lane<qi>(mine, ma);
                            });
                            if (bandm) {
                                // decided after the loop by the exact FP64 chain
#pragma unroll
                                for (int qi = 0; qi < QL; qi++) {
                                    const uint64_t mb = __builtin_amdgcn_ballot_w64(accf[qi] > tacc && accf[qi] <= trej);
                                    if (mb) {
                                        if (lane == 0 && nband < (uint32_t)kBandList) rl.band[nband] = (ch << 4) | (uint32_t)qi;
                                        nband++;
                                        my_band += (uint64_t)__popcll(mb);
                                    }
                                }
                            }
                            cnt_l += (uint32_t)__popcll(mine);
                            if (ch < nchunks) {
                                if (keep && (uint32_t)lane < nq) mask_put(ch * nq + lane, mine);
                                hw |= (anyh ? 1ull : 0ull) << (ch & 63u);
                            }
                        }
                        if (hit_bits && (((ch0 + kRing) & 63u) == 0 || ch0 + kRing >= nchunks)) {
                            if (lane == 0) fl.hitw[ch0 >> 6] = hw;
                            hw = 0;
                        }
                    }
                }
            };
            if (nq == 1)
                chunk_loop(std::integral_constant<int, 1>{});
            else if (nq <= (uint32_t)QB / 2)
                chunk_loop(std::integral_constant<int, QB / 2>{});
            else
                chunk_loop(std::integral_constant<int, QB>{});

            // The band: the reference's exact FP64 chain decides the lanes
            // whose FP32 sum lies between the two thresholds.  Listed
            // (query, chunk) pairs, or all of them when the list overflowed.
            if (nband) {
                wave_lds_order();
                const bool all = nband > (uint32_t)kBandList;
                const uint32_t nitems = all ? nq * nchunks : nband;
                const uint64_t n = a.g.n;
                for (uint32_t e = 0; e < nitems; e++) {
                    uint32_t qi, ch;
                    if (all) {
                        ch = e / nq;
                        qi = e - ch * nq;
                    } else {
                        const uint32_t v = rl.band[e];
                        qi = v & 15u;
                        ch = v >> 4;
                    }
                    const uint32_t t = ch * 64 + lane;
                    const bool valid = t < M;
                    const uint32_t pos = valid ? cand_pos(rl, nslots, t) : 0u;
                    float af = 0.0f;
#pragma unroll
                    for (int l = 0; l < D; l++) {
                        const float dd = a.soa32[(uint64_t)l * n + pos] - a.soa32[(uint64_t)l * n + qpos0 + qi];
                        af = __builtin_fmaf(dd, dd, af);
                    }
                    const bool band = valid && !(af <= tacc) && !(af > trej);
                    bool hit = false;
                    if (band) {
                        double e64 = 0.0;
#pragma unroll
                        for (int l = 0; l < D; l++) {
                            const double dd = a.ix.soa[(uint64_t)l * n + pos] - a.ix.soa[(uint64_t)l * n + qpos0 + qi];
                            e64 = __builtin_fma(dd, dd, e64);
                        }
                        hit = e64 <= T;
                    }
                    const uint64_t madd = __builtin_amdgcn_ballot_w64(hit);
                    if (madd) {
                        if (lane == 0) {
                            if (keep) mask_put(ch * nq + qi, mask_get(ch * nq + qi) | madd);
                            if (hit_bits) fl.hitw[ch >> 6] |= 1ull << (ch & 63u);
                        }
                        cnt_l += (uint32_t)lane == qi ? (uint32_t)__popcll(madd) : 0u;
                    }
                }
            }
            my_evals += (uint64_t)M * nq;
            // lane qi: first slot of query qi within the unit's segment
            uint32_t incl_l = cnt_l;
#pragma unroll
            for (int o = 1; o < QB; o <<= 1) {
                const uint32_t y = __shfl_up(incl_l, o, 64);
                if (lane >= o) incl_l += y;
            }
            const uint32_t excl_l = incl_l - cnt_l;
            const uint32_t total = __shfl(incl_l, QB - 1, 64);
            if (a.count_only) {
                my_pairs += total;
                continue;
            }
            uint64_t base = 0;
            if (total > 0) {
                if (lane == 0) base = atomicAdd(a.cursor, (unsigned long long)total);
                base = wave_bcast_u64(base);
                if (base + total > a.capacity) {
                    if (lane == 0) a.deferred[atomicAdd(a.n_deferred, 1u)] = ui;
                    continue;
                }
            }
            my_pairs += total;
            wave_lds_sync();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            const uint32_t qid_l = (uint32_t)lane < nq ? a.ix.ila[qpos0 + lane] : 0u;
            if ((uint32_t)lane < nq) {
                a.row_off[qid_l] = a.out_base + base + excl_l;
                a.row_cnt[qid_l] = cnt_l;
            }
            if (total == 0) continue;
            uint64_t oq[QB];
            {
                const uint64_t o_l = base + excl_l;
#pragma unroll
                for (int qi = 0; qi < QB; qi++)
                    oq[qi] = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(o_l >> 32), qi) << 32) |
                             (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)o_l, qi);
            }
            if (keep) {
                // Chunks with hits (all chunks when the unit has more than the
                // bitmap holds), kW at a time: their neighbour ids and masks
                // are loaded together, then written chunk-major, each query's
                // neighbours in candidate order.
                constexpr int kW = 4;
                constexpr uint32_t kNone = ~0u;
                const uint32_t nwords = (nchunks + 63) / 64;
                uint32_t wi = 0, nextc = 0;
                uint64_t wb = hit_bits ? wave_bcast_u64(fl.hitw[0]) : 0ull;
                auto next_chunk = [&]() __attribute__((always_inline)) -> uint32_t {
                    if (!hit_bits) return nextc < nchunks ? nextc++ : kNone;
                    while (wb == 0) {
                        if (wi + 1 >= nwords) return kNone;
                        wb = wave_bcast_u64(fl.hitw[++wi]);
                    }
                    const uint32_t c = wi * 64 + (uint32_t)__builtin_ctzll(wb);
                    wb &= wb - 1;
                    return c;
                };
                for (;;) {
                    uint32_t cl[kW], idr[kW];
                    uint64_t mkr[kW];
#pragma unroll
                    for (int j = 0; j < kW; j++) cl[j] = next_chunk();
                    if (cl[0] == kNone) break;
#pragma unroll
                    for (int j = 0; j < kW; j++) {
                        idr[j] = 0;
                        mkr[j] = 0;
                        const uint32_t c = cl[j];
                        if (c != kNone) {
                            if (c < tab_c0 || c >= tab_c0 + (uint32_t)kTab) {
                                tab_c0 = c;
                                build_tab(fl, rl, nslots, c);
                            }
                            const uint32_t slot = tab_slot(fl, c, tab_c0);
                            const uint32_t t = c * 64 + lane;
                            const uint32_t pos = rl.start[slot] + t - rl.off[slot];
                            idr[j] = gld(a.ix.ila + (t < M ? pos : 0u));
                            if ((uint32_t)lane < nq) mkr[j] = mask_get(c * nq + lane);
                        }
                    }
#pragma unroll
                    for (int j = 0; j < kW; j++) {
                        if (cl[j] == kNone) continue;
#pragma unroll
                        for (int qi = 0; qi < QB; qi++) {
                            // lanes >= nq of the mask hold zero masks
                            const uint64_t m =
                                ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(mkr[j] >> 32), qi) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)mkr[j], qi);
                            if (m == 0) continue;
                            if (__builtin_amdgcn_inverse_ballot_w64(m)) gst(a.out + oq[qi] + lanes_below(m), idr[j]);
                            oq[qi] += (uint64_t)__popcll(m);
                        }
                    }
                }
            } else {
                // masks did not fit anywhere (very large units): evaluate each
                // chunk again with the exact chain, one query at a time.  The
                // writes stay inside each query's reserved row [oq, oend); a
                // count that differs from the first pass is reported
                // (a.mismatch), never written past the row.
                uint64_t oend[QB];
                {
                    const uint64_t e_l = base + excl_l + cnt_l;
#pragma unroll
                    for (int qi = 0; qi < QB; qi++)
                        oend[qi] = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(e_l >> 32), qi) << 32) |
                                   (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)e_l, qi);
                }
                bool bad = false;
                SlotWalk cur2;  // 128-register kernel: the one-register walk
                cur2.reset(rl, nslots);
                for (uint32_t ch = 0; ch < nchunks; ch++) {
                    const uint32_t t = ch * 64 + lane;
                    const bool valid = t < M;
                    const uint32_t pos = valid ? cur2.pos(rl, t) : 0u;
                    const uint32_t id = gld(a.ix.ila + pos);
                    double x[D];
                    load_cand<D>(a, pos, valid, x);
                    my_evals += (uint64_t)min(64u, M - ch * 64) * nq;
#pragma unroll
                    for (int qi = 0; qi < QB; qi++) {
                        if ((uint32_t)qi >= nq) continue;
                        double acc1 = 0.0;
                        if (a.g.f32dt) {
                            // DTYPE=float: the reference's FP32 chain
                            float af = 0.0f;
#pragma unroll
                            for (int l = 0; l < D; l++) {
                                const float dd = (float)x[l] - (float)a.ix.soa[(uint64_t)l * a.g.n + qpos0 + qi];
                                af = __builtin_fmaf(dd, dd, af);
                            }
                            acc1 = (double)af;
                        } else {
#pragma unroll
                            for (int l = 0; l < D; l++) {
                                const double dd = x[l] - a.ix.soa[(uint64_t)l * a.g.n + qpos0 + qi];
                                acc1 = __builtin_fma(dd, dd, acc1);
                            }
                        }
                        const uint64_t m = __ballot(valid && acc1 <= T);
                        if (m == 0) continue;
                        const uint64_t room = oend[qi] - oq[qi];
                        const uint64_t pc = (uint64_t)__popcll(m);
                        bad |= pc > room;
                        if (__builtin_amdgcn_inverse_ballot_w64(m) && lanes_below(m) < room)
                            gst(a.out + oq[qi] + lanes_below(m), id);
                        oq[qi] += pc < room ? pc : room;
                    }
                }
#pragma unroll
                for (int qi = 0; qi < QB; qi++) bad |= (uint32_t)qi < nq && oq[qi] != oend[qi];
                if (bad && lane == 0) atomicAdd(a.mismatch, 1ull);
            }
        }
    }
    if (lane == 0) {
        if (my_band && a.band) atomicAdd(a.band, (unsigned long long)my_band);
        if (my_evals) atomicAdd(a.evals, (unsigned long long)my_evals);
        if (my_pairs) atomicAdd(a.pairs, (unsigned long long)my_pairs);
    }
}

// Per-cell candidate count over the stencil (sortByWorkLoadGlobal).
__global__ __launch_bounds__(256) void k_cell_workload(GridGeom g, DevIndex ix,
                                                       uint32_t *__restrict__ work) {
    uint32_t cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= ix.C) return;
    uint32_t c[kMaxK], lo[kMaxK], hi[kMaxK];
    decode_cell(ix.cell_key[cell], g, c);
    bool ok = true;
    for (uint32_t j = 0; j < g.k; j++) ok &= stencil_dim(c[j], g.nCells[j], g.emulate_cell0, &lo[j], &hi[j]);
    uint32_t total = 0;
    if (ok) {
        uint32_t nrows = pow3(g.k - 1);
        for (uint32_t r = 0; r < nrows; r++) {
            uint32_t cf, cl;
            stencil_row(c, lo, hi, r, g, ix, &cf, &cl);
            if (cl > cf) total += ix.cell_start[cl] - ix.cell_start[cf];
        }
    }
    work[cell] = total;
}

// Units: cell c contributes ceil(|c| / qb) units when its stencil is live
// (every cell with all_cells: its queries then get empty rows on the GPU).
__global__ __launch_bounds__(256) void k_unit_counts(GridGeom g, DevIndex ix, uint32_t qb,
                                                     const uint32_t *__restrict__ order,
                                                     uint32_t *__restrict__ ucount,
                                                     uint32_t all_cells) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ix.C) return;
    uint32_t cell = order ? order[i] : i;
    uint32_t c[kMaxK];
    decode_cell(ix.cell_key[cell], g, c);
    bool live = true;
    for (uint32_t j = 0; j < g.k; j++)
        if (c[j] == 0 && g.emulate_cell0 && !all_cells) live = false;
    uint32_t sz = ix.cell_start[cell + 1] - ix.cell_start[cell];
    ucount[i] = live ? (sz + qb - 1) / qb : 0u;
}

__global__ __launch_bounds__(256) void k_unit_fill(DevIndex ix, uint32_t qb,
                                                   const uint32_t *__restrict__ order,
                                                   const uint32_t *__restrict__ ucount,
                                                   const uint32_t *__restrict__ uoff,
                                                   uint2 *__restrict__ units) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ix.C) return;
    uint32_t cell = order ? order[i] : i;
    for (uint32_t u = 0; u < ucount[i]; u++) units[uoff[i] + u] = make_uint2(cell, u * qb);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
hegj_status launch_workload(const GridGeom &g, const DevIndex &ix, uint32_t *work, hipStream_t s) {
    if (ix.C == 0) return HEGJ_OK;
    k_cell_workload<<<ceil_div(ix.C, 256), 256, 0, s>>>(g, ix, work);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

hegj_status build_units(const GridGeom &g, const DevIndex &ix, uint32_t qb, const uint32_t *order,
                        uint2 **units, uint32_t *n_units, hipStream_t s, bool all_cells) {
    Scratch sc;
    uint32_t *ucount, *uoff, *d_total;
    uint32_t C = ix.C;
    HIP_TRY(sc.get(&ucount, C));
    HIP_TRY(sc.get(&uoff, C));
    HIP_TRY(sc.get(&d_total, 1));
    k_unit_counts<<<ceil_div(C, 256), 256, 0, s>>>(g, ix, qb, order, ucount, all_cells ? 1u : 0u);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(scan_u32(ucount, uoff, C, d_total, s));
    uint32_t U;
    HIP_TRY(hipMemcpyAsync(&U, d_total, sizeof(U), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(dev_malloc_n(units, (size_t)U + 1));
    k_unit_fill<<<ceil_div(C, 256), 256, 0, s>>>(ix, qb, order, ucount, uoff, *units);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    *n_units = U;
    return HEGJ_OK;
}

// Threads per workgroup of a join launch: 64 per wave, `block_waves` waves
// (the kernels index their per-wave LDS and mask slices by wave).
static uint32_t block_threads(const JoinArgs &a) {
    return 64u * (a.block_waves ? a.block_waves : (uint32_t)kWavesPerBlock);
}

template <int D, int QB>
constexpr bool f32_inst() {
    return (D == 2 || D == 6) && QB * D <= 48;
}

template <int D, int QB>
static hegj_status launch_join_t(const JoinArgs &a, const JoinArgs *ap, uint32_t blocks,
                                 hipStream_t s) {
    const uint32_t lds = a.f32 ? 0u : q_lds_bytes<D, QB>(a.g.dim);
    if (lds > kMaxQueryLds)
        return fail(HEGJ_ERR_UNSUPPORTED, "queries_per_unit * dim too large for LDS; lower queries_per_unit");
    if (a.f32) {
        if constexpr (f32_inst<D, QB>()) {
            if (ap)
                k_join_f32<D, QB, true><<<blocks, block_threads(a), 0, s>>>(JoinArgs{}, ap);
            else
                k_join_f32<D, QB, false><<<blocks, block_threads(a), 0, s>>>(a, nullptr);
        } else {
            return fail(HEGJ_ERR_UNSUPPORTED, "no FP32-filter kernel for this dim / queries_per_unit");
        }
    } else if (ap) {
        k_join<D, QB, true><<<blocks, block_threads(a), lds, s>>>(JoinArgs{}, ap);
    } else {
        k_join<D, QB, false><<<blocks, block_threads(a), lds, s>>>(a, nullptr);
    }
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

// DTYPE=float outside the FP32-filter instances: the generic kernel with the
// FP32 chain (eager launches only).
template <int QB>
static hegj_status launch_join_flt(const JoinArgs &a, const JoinArgs *ap, uint32_t blocks,
                                   hipStream_t s) {
    if (ap) return fail(HEGJ_ERR_UNSUPPORTED, "DTYPE=float joins are not replayed from a graph");
    const uint32_t lds = q_lds_bytes<0, QB>(a.g.dim);
    if (lds > kMaxQueryLds)
        return fail(HEGJ_ERR_UNSUPPORTED, "queries_per_unit * dim too large for LDS; lower queries_per_unit");
    k_join<0, QB, false, true><<<blocks, block_threads(a), lds, s>>>(a, nullptr);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

template <int QB>
static hegj_status launch_join_qb(const JoinArgs &a, const JoinArgs *ap, uint32_t blocks,
                                  hipStream_t s) {
    if (a.g.f32dt && !a.f32) return launch_join_flt<QB>(a, ap, blocks, s);
    switch (a.g.dim) {
        case 2: return launch_join_t<2, QB>(a, ap, blocks, s);
        case 6: return launch_join_t<6, QB>(a, ap, blocks, s);
        case 18: return launch_join_t<18, QB>(a, ap, blocks, s);
        case 90: return launch_join_t<90, QB>(a, ap, blocks, s);
        default: return launch_join_t<0, QB>(a, ap, blocks, s);
    }
}

bool f32_supported(uint32_t dim, uint32_t qb, uint32_t n) {
    // k_join_f32 addresses the padded FP32 copy with 32-bit byte offsets
    const bool fits = ((uint64_t)n + kAos32Pad) * dim * sizeof(float) <= 0xffffffffull;
    return fits && ((dim == 2 && (qb == 4 || qb == 8 || qb == 16)) || (dim == 6 && (qb == 4 || qb == 8)));
}

hegj_status launch_join(const JoinArgs &a, uint32_t qb, uint32_t blocks, hipStream_t s,
                        const JoinArgs *dev_args) {
    switch (qb) {
        case 4: return launch_join_qb<4>(a, dev_args, blocks, s);
        case 8: return launch_join_qb<8>(a, dev_args, blocks, s);
        case 16: return launch_join_qb<16>(a, dev_args, blocks, s);
        default: return fail(HEGJ_ERR_INVALID_ARG, "queries_per_unit must be 4, 8 or 16");
    }
}

// Pair hash of one batch: sum of mix64(query << 32 | neighbour) over the rows
// the batch holds (row offsets in [out_base, out_base + capacity)).  G lanes
// per row (G = 64 for long rows), each reading 4 entries per step, so a row
// is read in runs of 4G entries; one atomic per wave.
template <bool IND, int G>
__global__ __launch_bounds__(256) void k_pair_hash(JoinArgs av, const JoinArgs *__restrict__ ap) {
    const JoinArgs &a = IND ? *ap : av;
    const uint32_t n = a.g.n;
    const uint32_t sub = threadIdx.x & (G - 1);
    const uint32_t grp = (blockIdx.x * 256 + threadIdx.x) / G, ngrp = (gridDim.x * 256) / G;
    uint64_t h = 0;
    for (uint32_t q = grp; q < n; q += ngrp) {
        const uint32_t cnt = gld(a.row_cnt + q);
        if (cnt == 0) continue;
        const uint64_t off = gld(a.row_off + q);
        if (off < a.out_base || off - a.out_base >= a.capacity) continue;
        const uint32_t *row = a.out + (off - a.out_base);
        const uint64_t qk = (uint64_t)q << 32;
        uint32_t j = sub;
        for (; j + 3 * G < cnt; j += 4 * G) {
            const uint32_t v0 = gld(row + j), v1 = gld(row + j + G), v2 = gld(row + j + 2 * G),
                           v3 = gld(row + j + 3 * G);
            h += mix64(qk | v0) + mix64(qk | v1) + mix64(qk | v2) + mix64(qk | v3);
        }
        for (; j < cnt; j += G) h += mix64(qk | gld(row + j));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
    if ((threadIdx.x & 63) == 0 && h) atomicAdd(a.hash, (unsigned long long)h);
}

hegj_status launch_pair_hash(const JoinArgs &a, uint64_t pairs_hint, hipStream_t s, const JoinArgs *dev_args) {
    const uint32_t n = std::max<uint32_t>(a.g.n, 1);
    const bool longrows = pairs_hint / n >= 256;  // average row of 256+ entries
    const uint32_t G = longrows ? 64 : 16;
    const uint32_t blocks = std::min<uint32_t>(std::max<uint32_t>(ceil_div((uint64_t)n * G, 256), 1), 4096);
    if (dev_args) {
        if (longrows)
            k_pair_hash<true, 64><<<blocks, 256, 0, s>>>(JoinArgs{}, dev_args);
        else
            k_pair_hash<true, 16><<<blocks, 256, 0, s>>>(JoinArgs{}, dev_args);
    } else {
        if (longrows)
            k_pair_hash<false, 64><<<blocks, 256, 0, s>>>(a, nullptr);
        else
            k_pair_hash<false, 16><<<blocks, 256, 0, s>>>(a, nullptr);
    }
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

__global__ void k_zero_words(uint32_t *__restrict__ p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

hegj_status zero_words(void *p, size_t bytes, hipStream_t s) {
    const uint32_t n = (uint32_t)(bytes / 4);
    k_zero_words<<<std::min<uint32_t>(ceil_div(n, 256), 1024), 256, 0, s>>>(static_cast<uint32_t *>(p), n);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

}  // namespace hegj

namespace hegj {

hegj_status join_occupancy(uint32_t dim, uint32_t qb, int device, uint32_t *blocks,
                           JoinKind kind) {
    int per_cu = 0;
    const void *fn = nullptr;
    uint32_t lds = 0;
#define HEGJ_OCC(D, Q) fn = reinterpret_cast<const void *>(&k_join<D, Q, false>), lds = q_lds_bytes<D, Q>(dim)
    uint32_t d = (dim == 2 || dim == 6 || dim == 18 || dim == 90) ? dim : 0;
    const uint32_t sel = kind == JoinKind::flt ? qb * 1000 + 999 : kind == JoinKind::f32 ? qb * 1000 + 500 + d : qb * 1000 + d;
    switch (sel) {
        case 4502: fn = reinterpret_cast<const void *>(&k_join_f32<2, 4, false>); break;
        case 8502: fn = reinterpret_cast<const void *>(&k_join_f32<2, 8, false>); break;
        case 16502: fn = reinterpret_cast<const void *>(&k_join_f32<2, 16, false>); break;
        case 4506: fn = reinterpret_cast<const void *>(&k_join_f32<6, 4, false>); break;
        case 8506: fn = reinterpret_cast<const void *>(&k_join_f32<6, 8, false>); break;
        case 4999: fn = reinterpret_cast<const void *>(&k_join<0, 4, false, true>), lds = q_lds_bytes<0, 4>(dim); break;
        case 8999: fn = reinterpret_cast<const void *>(&k_join<0, 8, false, true>), lds = q_lds_bytes<0, 8>(dim); break;
        case 16999: fn = reinterpret_cast<const void *>(&k_join<0, 16, false, true>), lds = q_lds_bytes<0, 16>(dim); break;
        case 4002: HEGJ_OCC(2, 4); break;
        case 4006: HEGJ_OCC(6, 4); break;
        case 4018: HEGJ_OCC(18, 4); break;
        case 4090: HEGJ_OCC(90, 4); break;
        case 4000: HEGJ_OCC(0, 4); break;
        case 8002: HEGJ_OCC(2, 8); break;
        case 8006: HEGJ_OCC(6, 8); break;
        case 8018: HEGJ_OCC(18, 8); break;
        case 8090: HEGJ_OCC(90, 8); break;
        case 8000: HEGJ_OCC(0, 8); break;
        case 16002: HEGJ_OCC(2, 16); break;
        case 16006: HEGJ_OCC(6, 16); break;
        case 16018: HEGJ_OCC(18, 16); break;
        case 16090: HEGJ_OCC(90, 16); break;
        case 16000: HEGJ_OCC(0, 16); break;
        default: return fail(HEGJ_ERR_INVALID_ARG, "queries_per_unit must be 4, 8 or 16");
    }
#undef HEGJ_OCC
    if (lds > kMaxQueryLds)
        return fail(HEGJ_ERR_UNSUPPORTED, "queries_per_unit * dim too large for LDS; lower queries_per_unit");
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * kWavesPerBlock, lds));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (per_cu < 1) per_cu = 1;
    *blocks = (uint32_t)per_cu * (uint32_t)prop.multiProcessorCount;
    return HEGJ_OK;
}

}  // namespace hegj
