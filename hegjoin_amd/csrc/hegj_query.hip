// hegj_query.hip -- range join of external query points against a resident
// grid index: the device side of the serving loop.
//
// The reference has no query path (its only join is the self-join of
// kernel.cu:727-802); this follows the same grid semantics with the intended
// stencil (no cell-0 wrap) and the same distance chain and predicate:
//   a query q and an index point p pair when sqrt(acc) <= eps, with
//   acc = fma(p_l - q_l, p_l - q_l, acc) over l = 0..dim-1 (kernel.cu:350-355),
//   evaluated as acc <= T, T the largest double with sqrt(T) <= eps.
// The index level's cell side c satisfies eps <= c (1 - 2^-30), so every
// point within eps lies in the 3^k cells around the query's cell even after
// the rounding of (x - min) / c.
//
// Pipeline for one request (one stream):
//   k_qcell_keys    query -> key of its cell in the grid widened by one cell
//                   on each side (coordinates -1 .. nCells), or `dead` when
//                   the query lies further out (no point can be in range);
//   radix sort      (key, query id), dead keys last;
//   k_qruns         one unit per run of equal keys (up to 16 queries of one
//                   cell); each group of 4 of them builds its stencil rows,
//                   pruned to the cells within eps of one of its queries;
//   k_qjoin<count>  per-query neighbour counts, and per group a bitmap of the
//                   candidate chunks in which one of its queries has a hit;
//   scan            CSR offsets, exact result size (admission happens here);
//   k_qjoin<fill>   each query's neighbours written at its own offset in
//                   candidate (cell) order, visiting only the hit chunks --
//                   no atomics, deterministic.
#include <type_traits>

#include "hegj_internal.h"

namespace hegj {

namespace {

constexpr int kQWaves = 4;
// Queries evaluated together against one candidate chunk, on the FP64 path
// (their coordinates fill 48 scalar registers at 6-D) and on the FP32 filter.
// 8 on the FP32 filter measured slower at 6-D (10M points, 93k queries,
// eps 0.0172): the group's pruned stencil grows (+12% distance evaluations)
// and its registers spill.
constexpr int kQB = 4;
constexpr int kQBF32 = 4;
constexpr int kMaxQB = 16;
constexpr uint32_t kHitWords = kQueryHitWords;

struct QRowList {
    uint32_t start[kMaxRows];
    uint32_t off[kMaxRows + 1];
    float2 gap[kMaxQB * kMaxK];  // per (query, indexed dim): squared gaps to the -1 / +1 cell
};

__global__ __launch_bounds__(256) void k_qcell_keys(const double *__restrict__ q, uint32_t nq,
                                                    QueryGeom qg, uint64_t *__restrict__ keys,
                                                    uint32_t *__restrict__ ids) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nq) return;
    uint64_t key = 0, mult = 1;
    bool dead = false;
    for (uint32_t j = 0; j < qg.g.k; j++) {
        const double x = q[(uint64_t)i * qg.g.dim + j];
        const double f = floor((x - qg.g.minArr[j]) / qg.g.eps);
        // widened coordinate: -1 .. nCells map to 0 .. nCells + 1
        if (!(f >= -1.0) || f > (double)qg.g.nCells[j]) {
            dead = true;
        } else {
            key += (uint64_t)(int64_t)(f + 1.0) * mult;
        }
        mult *= (uint64_t)qg.g.nCells[j] + 2;
    }
    keys[i] = dead ? qg.dead_key : key;
    ids[i] = i;
}

// Unit r starts at the r-th run of equal live keys; n_live = first dead slot.
__global__ __launch_bounds__(256) void k_qrun_flags(const uint64_t *__restrict__ keys, uint32_t nq,
                                                    uint64_t dead, uint32_t *__restrict__ flag,
                                                    uint32_t *__restrict__ n_live) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nq) return;
    const uint64_t k = keys[i];
    flag[i] = (k != dead && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
    const bool last_live = k != dead && (i + 1 == nq || keys[i + 1] == dead);
    if (last_live) *n_live = i + 1;
}

__global__ __launch_bounds__(256) void k_qrun_fill(const uint32_t *__restrict__ flag,
                                                   const uint32_t *__restrict__ rank, uint32_t nq,
                                                   uint32_t *__restrict__ starts) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < nq && flag[i]) starts[rank[i]] = i;
}

// Units of at most kUnitQueries queries: a run of queries sharing a cell is
// split so that a dense cell spreads over many waves instead of one.
constexpr uint32_t kUnitQueries = 16;

__global__ __launch_bounds__(256) void k_sub_counts(const uint32_t *__restrict__ starts, uint32_t nr,
                                                    uint32_t *__restrict__ cnt) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r < nr) cnt[r] = (starts[r + 1] - starts[r] + kUnitQueries - 1) / kUnitQueries;
}

__global__ __launch_bounds__(256) void k_sub_fill(const uint32_t *__restrict__ starts, uint32_t nr,
                                                  const uint32_t *__restrict__ first,
                                                  uint32_t *__restrict__ sub) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nr) return;
    const uint32_t b = starts[r], e = starts[r + 1];
    uint32_t o = first[r];
    for (uint32_t i = b; i < e; i += kUnitQueries) sub[o++] = i;
}

// Stencil rows of a group of queries (all in the cell whose widened key is
// `key`) into the wave's LDS list; returns the candidate count.
//
// Cell-box pruning, as in the self-join: a stencil cell whose box lies
// farther than eps from every query of the group holds no neighbour and is
// left out.  With f_j the query's position inside its cell along indexed dim
// j (in cell sides), the gap to the -1 cell is f_j, to the +1 cell 1 - f_j;
// a cell is kept when the smallest sum of squared gaps over the group is at
// most T / side^2.  Gaps are shrunk by 1e-6 + 2e-15 (|c_j| + 2) (the FP64
// rounding of the cell assignments) and the bound raised by 1e-5 relative
// (FP32 sums), so the kept set can only be larger than the exact one.
template <int QB>
__device__ uint32_t qbuild_rows(const QJoinArgs &a, uint64_t key, const uint32_t *qid, uint32_t nqg,
                                QRowList &rl, uint32_t *nslots) {
    const GridGeom &g = a.qg.g;
    const int lane = threadIdx.x & 63;
    uint32_t c[kMaxK], lo[kMaxK], hi[kMaxK];
    int64_t cs[kMaxK];
    bool ok = true;
    for (uint32_t j = 0; j < g.k; j++) {
        const uint64_t w = (uint64_t)g.nCells[j] + 2;
        const uint64_t qq = key / w;
        const int64_t cj = (int64_t)(key - qq * w) - 1;  // -1 .. nCells
        key = qq;
        const int64_t l = cj - 1 < 0 ? 0 : cj - 1;
        const int64_t h = cj + 1 > (int64_t)g.nCells[j] - 1 ? (int64_t)g.nCells[j] - 1 : cj + 1;
        ok &= l <= h;
        c[j] = (uint32_t)cj;  // -1 wraps; stencil_row adds the offset modulo 2^32
        cs[j] = cj;
        lo[j] = (uint32_t)l;
        hi[j] = (uint32_t)(h < 0 ? 0 : h);
    }
    if (!ok) {
        *nslots = 0;
        return 0;
    }
    const uint32_t k = g.k, dim = g.dim;
    const double inv_side = 1.0 / g.eps;
    wave_lds_sync();  // the previous group's reads of the list are done
    for (uint32_t e = (uint32_t)lane; e < nqg * k; e += 64) {
        const uint32_t qi = e / k, j = e - qi * k;
        uint32_t id = 0;
        int64_t cj = 0;
#pragma unroll
        for (int t = 0; t < QB; t++) id = (uint32_t)t == qi ? qid[t] : id;
#pragma unroll
        for (int jj = 0; jj < kMaxK; jj++) cj = (uint32_t)jj == j ? cs[jj] : cj;
        const double x = a.queries[(uint64_t)id * dim + j];
        double f = (x - g.minArr[j]) * inv_side - (double)cj;
        f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
        const double mg = 1e-6 + 2e-15 * ((double)(cj < 0 ? -cj : cj) + 2.0);
        const double gm = f - mg > 0.0 ? f - mg : 0.0, gp = 1.0 - f - mg > 0.0 ? 1.0 - f - mg : 0.0;
        rl.gap[qi * kMaxK + j] = make_float2((float)(gm * gm), (float)(gp * gp));
    }
    wave_lds_sync();
    const float thr = (float)(a.thresh * inv_side * inv_side) * (1.0f + 1e-5f);
    const uint32_t nrows = pow3(k - 1);
    uint32_t slots = 0, total = 0;
    for (uint32_t r0 = 0; r0 < nrows; r0 += 64) {
        const uint32_t r = r0 + lane;
        uint32_t pstart = 0, plen = 0;
        if (r < nrows) {
            // offset digits of row r along dims 1..k-1 (0: -1, 1: 0, 2: +1)
            uint32_t o[kMaxK];
            uint32_t rr = r;
#pragma unroll
            for (int j = 1; j < kMaxK; j++) {
                o[j] = rr % 3;
                rr /= 3;
            }
            float mmid = 3.4e38f, mleft = 3.4e38f, mright = 3.4e38f;
            for (uint32_t qi = 0; qi < nqg; qi++) {
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
            if (mmid <= thr) {
                // dim-0 window, without the -1 / +1 cell when no query reaches it
                int64_t wlo = lo[0], whi = hi[0];
                if (!(mleft <= thr) && cs[0] > wlo) wlo = cs[0];
                if (!(mright <= thr) && cs[0] < whi) whi = cs[0];
                if (wlo <= whi) {
                    uint32_t lo2[kMaxK], hi2[kMaxK];
#pragma unroll
                    for (int j = 0; j < kMaxK; j++) {
                        lo2[j] = lo[j];
                        hi2[j] = hi[j];
                    }
                    lo2[0] = (uint32_t)wlo;
                    hi2[0] = (uint32_t)whi;
                    uint32_t cf, cl;
                    stencil_row(c, lo2, hi2, r, g, a.ix, &cf, &cl);
                    if (cl > cf) {
                        pstart = a.ix.cell_start[cf];
                        plen = a.ix.cell_start[cl] - pstart;
                    }
                }
            }
        }
        const uint64_t nz = __ballot(plen > 0);
        uint32_t x = plen;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
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
    if (lane == 0) rl.off[slots] = total;
    *nslots = slots;
    wave_lds_sync();
    return total;
}

// D > 0: coordinates in registers; D == 0: any dimensionality, one
// dimension at a time.
// F32: candidates from the level's centred FP32 copy, queries from the
// request's (centred the same way); acc32 <= tacc32 accepts, acc32 > trej32
// rejects, the band between is decided by the exact FP64 chain.
// Held to 3 waves per SIMD (168 registers, a few spilled): at 2 waves the
// candidate loads of too few waves are in flight (6-D, 10M points, 93k
// queries: 290 -> 270 ms per request).
template <int D, bool FILL, bool F32 = false, int QB = kQB>
__global__ __launch_bounds__(64 * kQWaves) __attribute__((amdgpu_waves_per_eu(3))) void k_qjoin(QJoinArgs a) {
    static_assert(!F32 || D > 0, "FP32 filter needs a fixed dimensionality");
    __shared__ QRowList rls[kQWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    QRowList &rl = rls[wid];
    const uint32_t dim = D > 0 ? (uint32_t)D : a.qg.g.dim;
    const uint64_t n = a.qg.g.n;
    const double T = a.thresh;
    const double *__restrict__ soa = a.ix.soa;
    uint64_t my_evals = 0, my_band = 0;
    while (true) {
        uint32_t u = 0;
        if (lane == 0) u = atomicAdd(a.queue_head, 1u);
        u = __shfl(u, 0, 64);
        if (u >= a.n_units) break;
        const uint32_t s0 = a.starts[u], s1 = a.starts[u + 1];
        const uint64_t ukey = a.keys[s0];
        for (uint32_t g0 = s0; g0 < s1; g0 += QB) {
            const uint32_t nqg = min((uint32_t)QB, s1 - g0);
            uint32_t qid[QB];
            uint64_t base[QB];
            uint32_t cnt[QB];
#pragma unroll
            for (int qi = 0; qi < QB; qi++) {
                qid[qi] = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.ids[g0 + min((uint32_t)qi, nqg - 1)]);
                base[qi] = FILL ? a.offsets[qid[qi]] : 0;
                cnt[qi] = 0;
            }
            uint32_t nslots;
            const uint32_t M = qbuild_rows<QB>(a, ukey, qid, nqg, rl, &nslots);
            // the group's query coordinates, once (lane i loads element i;
            // scalar registers), so no load in the candidate loop waits on
            // anything but the candidates
            constexpr int DQ = D > 0 ? D : 1;
            double qreg[F32 ? 1 : QB][DQ];
            float qregf[F32 ? QB : 1][DQ];
            if constexpr (F32) {
#pragma unroll
                for (int h = 0; h < (QB * DQ + 63) / 64; h++) {
                    const int e = h * 64 + lane, qi = e / DQ, l = e % DQ;
                    float v = 0.0f;
                    if (e < QB * DQ) {
                        uint32_t id = 0;
#pragma unroll
                        for (int k = 0; k < QB; k++) id = k == qi ? qid[k] : id;
                        v = a.q32[(uint64_t)id * dim + l];
                    }
#pragma unroll
                    for (int j = h * 64; j < QB * DQ && j < h * 64 + 64; j++)
                        qregf[j / DQ][j % DQ] = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), j - h * 64));
                }
            } else if constexpr (D > 0) {
                static_assert(QB * D <= 64, "one load per lane");
                const int qi = lane / DQ, l = lane % DQ;
                double v = 0.0;
                if (lane < QB * DQ) {
                    uint32_t id = 0;
#pragma unroll
                    for (int k = 0; k < QB; k++) id = k == qi ? qid[k] : id;
                    v = a.queries[(uint64_t)id * dim + l];
                }
#pragma unroll
                for (int j = 0; j < QB * DQ; j++) {
                    const uint64_t u = (uint64_t)__double_as_longlong(v);
                    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, j);
                    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), j);
                    qreg[j / DQ][j % DQ] = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
                }
            }
            SlotCursor cur;
            cur.reset(rl);
            if constexpr (D > 0) {
                // Candidate chunks kDepth ahead in a register ring (coordinates,
                // positions and, when filling, point ids), so a wave keeps
                // several chunks of loads in flight instead of waiting on each.
                // A slot's position and id are copied out before the slot is
                // refilled; its coordinates are read by the distances first.
                constexpr int kDepth = 4;
                typedef typename std::conditional<F32, float, double>::type CT;
                CT ring[kDepth][DQ];
                uint32_t posr[kDepth], pidr[kDepth];
                auto load = [&](uint32_t t0, int p) __attribute__((always_inline)) {
                    const uint32_t t = t0 + lane;
                    const uint32_t pos = t < M ? cur.pos(rl, t) : 0u;
                    posr[p] = pos;
                    if constexpr (F32) {
                        if constexpr (DQ % 2 == 0) {
                            typedef float f2 __attribute__((ext_vector_type(2)));
                            const f2 *src = reinterpret_cast<const f2 *>(a.aos32 + (uint64_t)pos * DQ);
#pragma unroll
                            for (int l = 0; l < DQ / 2; l++) {
                                const f2 v = src[l];
                                ring[p][2 * l] = v.x;
                                ring[p][2 * l + 1] = v.y;
                            }
                        } else {
#pragma unroll
                            for (int l = 0; l < DQ; l++) ring[p][l] = a.soa32[(uint64_t)l * n + pos];
                        }
                    } else if constexpr (D <= (int)kAosMaxDim) {
                        const double *src = a.ix.aos + (uint64_t)pos * DQ;
#pragma unroll
                        for (int l = 0; l < DQ; l++) ring[p][l] = src[l];
                    } else {
#pragma unroll
                        for (int l = 0; l < DQ; l++) ring[p][l] = soa[(uint64_t)l * n + pos];
                    }
                    if (FILL) pidr[p] = a.ix.ila[pos];
                };
                // Hit-chunk bitmap of the group (a.hitbits: kHitWords words
                // per query, the group's at its first query), one bit per
                // `span` chunks.  The count pass sets the bits of chunks in
                // which some query of the group has a neighbour; the fill
                // pass, which does the same arithmetic, visits only those.
                const uint32_t nch = (M + 63) / 64;
                uint64_t *hb = a.hitbits ? a.hitbits + (uint64_t)g0 * kHitWords : nullptr;
                const uint32_t capb = nqg * kHitWords * 64;
                const uint32_t span = nch > capb ? (nch + capb - 1) / capb : 1u;
                const uint32_t nw = ((nch + span - 1) / span + 63) / 64;
                uint32_t cur = 0, cend = 0, wi = 0, bitpos = 0, subc = 0;
                uint64_t w = (FILL && hb && nw > 0) ? hb[0] : 0ull, hitw = 0;
                // next chunk to visit (nch when none is left)
                auto next = [&]() __attribute__((always_inline)) -> uint32_t {
                    if (!FILL || !hb) return cur < nch ? cur++ : nch;
                    if (cur + 1 < cend) return ++cur;
                    while (w == 0) {
                        if (++wi >= nw) return cur = cend = nch;
                        w = hb[wi];
                    }
                    const uint32_t b = wi * 64 + (uint32_t)__builtin_ctzll(w);
                    w &= w - 1;
                    cur = b * span;
                    cend = min(cur + span, nch);
                    return cur;
                };
                uint32_t chr[kDepth];
#pragma unroll
                for (int p = 0; p < kDepth; p++) {
                    chr[p] = next();
                    load(chr[p] * 64, p);
                }
                while (chr[0] < nch) {
#pragma unroll
                    for (int p = 0; p < kDepth; p++) {
                        const uint32_t ch = chr[p];
                        const uint32_t t0 = ch * 64;
                        const bool valid = t0 + lane < M;
                        const uint64_t vmask = __builtin_amdgcn_ballot_w64(valid);
                        const uint32_t pos = posr[p];
                        const uint32_t pid = FILL ? pidr[p] : 0u;
                        CT accv[QB];
#pragma unroll
                        for (int qi = 0; qi < QB; qi++) {
                            accv[qi] = (CT)0;
#pragma unroll
                            for (int l = 0; l < DQ; l++) {
                                if constexpr (F32) {
                                    const float d = ring[p][l] - qregf[qi][l];
                                    accv[qi] = __builtin_fmaf(d, d, accv[qi]);
                                } else {
                                    const double d = ring[p][l] - qreg[qi][l];
                                    accv[qi] = __builtin_fma(d, d, accv[qi]);
                                }
                            }
                        }
#pragma unroll
                        for (int qi = 0; qi < QB; qi++) __asm__ volatile("" ::"v"(accv[qi]) : "memory");
                        chr[p] = next();
                        load(chr[p] * 64, p);
                        uint64_t anyhit = 0;
#pragma unroll
                        for (int qi = 0; qi < QB; qi++) {
                            const uint64_t vq = (uint32_t)qi < nqg ? vmask : 0ull;
                            uint64_t m;
                            if constexpr (F32) {
                                m = __builtin_amdgcn_ballot_w64(accv[qi] <= a.tacc32) & vq;
                                const uint64_t mb = vq & ~m & ~__builtin_amdgcn_ballot_w64(accv[qi] > a.trej32);
                                if (mb) {
                                    // the band: the exact FP64 chain decides
                                    bool hit = false;
                                    if (__builtin_amdgcn_inverse_ballot_w64(mb)) {
                                        double e = 0.0;
#pragma unroll
                                        for (int l = 0; l < DQ; l++) {
                                            const double d = soa[(uint64_t)l * n + pos] - a.queries[(uint64_t)qid[qi] * dim + l];
                                            e = __builtin_fma(d, d, e);
                                        }
                                        hit = e <= T;
                                    }
                                    m |= __builtin_amdgcn_ballot_w64(hit);
                                    my_band += (uint64_t)__popcll(mb);
                                }
                            } else {
                                m = __builtin_amdgcn_ballot_w64(accv[qi] <= T) & vq;
                            }
                            if (FILL && m && __builtin_amdgcn_inverse_ballot_w64(m))
                                a.out[base[qi] + cnt[qi] + lanes_below(m)] = pid;
                            cnt[qi] += (uint32_t)__popcll(m);
                            anyhit |= m;
                        }
                        if (!FILL && hb && ch < nch) {
                            hitw |= (uint64_t)(anyhit != 0) << (bitpos & 63);
                            if (++subc == span) {
                                subc = 0;
                                if ((++bitpos & 63) == 0) {
                                    if (lane == 0) hb[(bitpos >> 6) - 1] = hitw;
                                    hitw = 0;
                                }
                            }
                        }
                    }
                }
                if (!FILL && hb && ((bitpos & 63) != 0 || subc != 0) && lane == 0) hb[bitpos >> 6] = hitw;
            } else {
                for (uint32_t t0 = 0; t0 < M; t0 += 64) {
                    const uint32_t t = t0 + lane;
                    const bool valid = t < M;
                    const uint32_t pos = valid ? cur.pos(rl, t) : 0u;
                    // one compare per ballot, validity applied to the scalar
                    // masks; the hit lanes of a fill store under the mask itself
                    const uint64_t vmask = __builtin_amdgcn_ballot_w64(valid);
                    uint32_t pid = 0;
                    if (FILL) pid = a.ix.ila[valid ? pos : 0u];
                    double acc[QB];
    #pragma unroll
                    for (int qi = 0; qi < QB; qi++) acc[qi] = 0.0;
                    if constexpr (F32) {
                        float x[DQ];
                        if constexpr (DQ % 2 == 0) {
                            typedef float f2 __attribute__((ext_vector_type(2)));
                            const f2 *src = reinterpret_cast<const f2 *>(a.aos32 + (uint64_t)pos * DQ);
    #pragma unroll
                            for (int l = 0; l < DQ / 2; l++) {
                                const f2 v = src[l];
                                x[2 * l] = v.x;
                                x[2 * l + 1] = v.y;
                            }
                        } else {
    #pragma unroll
                            for (int l = 0; l < DQ; l++) x[l] = a.soa32[(uint64_t)l * n + pos];
                        }
    #pragma unroll
                        for (int qi = 0; qi < QB; qi++) {
                            float accf = 0.0f;
    #pragma unroll
                            for (int l = 0; l < DQ; l++) {
                                const float d = x[l] - qregf[qi][l];
                                accf = __builtin_fmaf(d, d, accf);
                            }
                            const uint64_t vq = (uint32_t)qi < nqg ? vmask : 0ull;
                            uint64_t m = __builtin_amdgcn_ballot_w64(accf <= a.tacc32) & vq;
                            const uint64_t mb = vq & ~m & ~__builtin_amdgcn_ballot_w64(accf > a.trej32);
                            if (mb) {
                                // the band: the exact FP64 chain decides
                                bool hit = false;
                                if (__builtin_amdgcn_inverse_ballot_w64(mb)) {
                                    double e = 0.0;
    #pragma unroll
                                    for (int l = 0; l < DQ; l++) {
                                        const double d = soa[(uint64_t)l * n + pos] - a.queries[(uint64_t)qid[qi] * dim + l];
                                        e = __builtin_fma(d, d, e);
                                    }
                                    hit = e <= T;
                                }
                                m |= __builtin_amdgcn_ballot_w64(hit);
                                my_band += (uint64_t)__popcll(mb);
                            }
                            if (FILL && m && __builtin_amdgcn_inverse_ballot_w64(m))
                                a.out[base[qi] + cnt[qi] + lanes_below(m)] = pid;
                            cnt[qi] += (uint32_t)__popcll(m);
                        }
                        continue;
                    } else if constexpr (D > 0) {
                        double x[DQ];
                        // unconditional (invalid lanes read position 0, masked below)
                        if constexpr (D <= (int)kAosMaxDim) {
                            const double *src = a.ix.aos + (uint64_t)pos * DQ;
    #pragma unroll
                            for (int l = 0; l < DQ; l++) x[l] = src[l];
                        } else {
    #pragma unroll
                            for (int l = 0; l < DQ; l++) x[l] = soa[(uint64_t)l * n + pos];
                        }
    #pragma unroll
                        for (int qi = 0; qi < QB; qi++) {
    #pragma unroll
                            for (int l = 0; l < DQ; l++) {
                                const double d = x[l] - qreg[qi][l];
                                acc[qi] = __builtin_fma(d, d, acc[qi]);
                            }
                        }
                    } else {
                        for (uint32_t l = 0; l < dim; l++) {
                            const double xl = valid ? soa[(uint64_t)l * n + pos] : 0.0;
    #pragma unroll
                            for (int qi = 0; qi < QB; qi++) {
                                const double d = xl - a.queries[(uint64_t)qid[qi] * dim + l];
                                acc[qi] = __builtin_fma(d, d, acc[qi]);
                            }
                        }
                    }
    #pragma unroll
                    for (int qi = 0; qi < QB; qi++) {
                        const uint64_t m = __builtin_amdgcn_ballot_w64(acc[qi] <= T) & ((uint32_t)qi < nqg ? vmask : 0ull);
                        if (FILL && m && __builtin_amdgcn_inverse_ballot_w64(m))
                            a.out[base[qi] + cnt[qi] + lanes_below(m)] = pid;
                        cnt[qi] += (uint32_t)__popcll(m);
                    }
                }
            }
            if (!FILL && lane < (int)nqg) {
                uint32_t mine = 0;
#pragma unroll
                for (int qi = 0; qi < QB; qi++)
                    if (qi == lane) mine = cnt[qi];
                a.counts[qid[lane]] = mine;
            }
            if (lane == 0) my_evals += (uint64_t)M * nqg;
        }
    }
    if (lane == 0) {
        if (!FILL && my_evals) atomicAdd(a.evals, (unsigned long long)my_evals);
        if (!FILL && my_band && a.band) atomicAdd(a.band, (unsigned long long)my_band);
    }
}

// Pair hash of a request's CSR result: sum of mix64(query << 32 | neighbour)
// over its rows, G lanes per row; one atomic per wave.
template <int G>
__global__ __launch_bounds__(256) void k_csr_hash(const uint64_t *__restrict__ off, const uint32_t *__restrict__ nb,
                                                  uint32_t n, unsigned long long *hash) {
    const uint32_t sub = threadIdx.x & (G - 1);
    const uint32_t grp = (blockIdx.x * 256 + threadIdx.x) / G, ngrp = (gridDim.x * 256) / G;
    uint64_t h = 0;
    for (uint32_t q = grp; q < n; q += ngrp) {
        const uint64_t b = off[q], e = off[q + 1];
        for (uint64_t j = b + sub; j < e; j += G) h += mix64(((uint64_t)q << 32) | nb[j]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
    if ((threadIdx.x & 63) == 0 && h) atomicAdd(hash, (unsigned long long)h);
}

// q32 = fl32(fl64(q - centre)) and the largest |q - centre| (as ordered bits).
__global__ __launch_bounds__(256) void k_q32(const double *__restrict__ q, uint64_t total, uint32_t dim,
                                             const double *__restrict__ centre, float *__restrict__ q32,
                                             unsigned long long *rmax_bits) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    double r = 0.0;
    if (i < total) {
        const double d = q[i] - centre[i % dim];
        q32[i] = (float)d;
        r = fabs(d);
    }
    for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
    if ((threadIdx.x & 63) == 0 && r > 0.0)
        atomicMax(rmax_bits, (unsigned long long)__double_as_longlong(r));
}

__global__ __launch_bounds__(256) void k_q_u32_u64(const uint32_t *__restrict__ in, uint32_t n,
                                                   uint64_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = in[i];
}

template <int D, bool FILL>
hegj_status launch_q(const QJoinArgs &a, uint32_t blocks, hipStream_t s) {
    if constexpr (D > 0) {
        if (a.f32)
            k_qjoin<D, FILL, true, kQBF32><<<blocks, 64 * kQWaves, 0, s>>>(a);
        else
            k_qjoin<D, FILL><<<blocks, 64 * kQWaves, 0, s>>>(a);
    } else {
        if (a.f32) return fail(HEGJ_ERR_UNSUPPORTED, "FP32 filter needs dim 2 or 6");
        k_qjoin<D, FILL><<<blocks, 64 * kQWaves, 0, s>>>(a);
    }
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

template <bool FILL>
hegj_status launch_qd(const QJoinArgs &a, uint32_t blocks, hipStream_t s) {
    switch (a.qg.g.dim) {
        case 2: return launch_q<2, FILL>(a, blocks, s);
        case 6: return launch_q<6, FILL>(a, blocks, s);
        default: return launch_q<0, FILL>(a, blocks, s);
    }
}

}  // namespace

hegj_status query_geom(const GridGeom &g, QueryGeom *qg) {
    qg->g = g;
    uint64_t prod = 1;
    for (uint32_t j = 0; j < g.k; j++) {
        const uint64_t w = (uint64_t)g.nCells[j] + 2;
        if (prod > (~0ull >> 1) / w) return fail(HEGJ_ERR_UNSUPPORTED, "query grid needs more than 63 bits");
        prod *= w;
    }
    qg->dead_key = prod;  // one past the largest live key
    int bits = 0;
    while (bits < 64 && (prod >> bits) != 0) bits++;
    qg->key_bits = bits;
    return HEGJ_OK;
}

hegj_status query_plan(QueryPlan *p, Scratch &keep, const QueryGeom &qg, const double *queries,
                       uint32_t nq, hipStream_t s) {
    p->nq = nq;
    p->n_units = 0;
    p->n_live = 0;
    if (nq == 0) return HEGJ_OK;
    HIP_TRY(keep.get(&p->keys, nq));
    HIP_TRY(keep.get(&p->ids, nq));
    HIP_TRY(keep.get(&p->starts, (size_t)nq + 1));
    Scratch sc;
    uint64_t *keys_alt;
    uint32_t *ids_alt, *flag, *rank, *d_small;
    HIP_TRY(sc.get(&keys_alt, nq));
    HIP_TRY(sc.get(&ids_alt, nq));
    HIP_TRY(sc.get(&flag, nq));
    HIP_TRY(sc.get(&rank, nq));
    HIP_TRY(sc.get(&d_small, 2));
    k_qcell_keys<<<ceil_div(nq, 256), 256, 0, s>>>(queries, nq, qg, p->keys, p->ids);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(radix_sort_u64(p->keys, p->ids, keys_alt, ids_alt, nq, qg.key_bits, s));
    HIP_TRY(hipMemsetAsync(d_small, 0, 2 * sizeof(uint32_t), s));
    k_qrun_flags<<<ceil_div(nq, 256), 256, 0, s>>>(p->keys, nq, qg.dead_key, flag, d_small + 1);
    HIP_TRY(hipGetLastError());
    HEGJ_TRY(scan_u32(flag, rank, nq, d_small, s));
    k_qrun_fill<<<ceil_div(nq, 256), 256, 0, s>>>(flag, rank, nq, p->starts);
    HIP_TRY(hipGetLastError());
    uint32_t h[2];
    HIP_TRY(hipMemcpyAsync(h, d_small, sizeof(h), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const uint32_t runs = h[0];
    p->n_live = h[1];
    // starts[runs] = end of the live queries
    HIP_TRY(hipMemcpyAsync(p->starts + runs, &p->n_live, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    // split the runs into units of at most kUnitQueries queries
    p->n_units = 0;
    if (runs > 0) {
        uint32_t *cnt, *first, *sub_starts, *d_total;
        HIP_TRY(sc.get(&cnt, runs));
        HIP_TRY(sc.get(&first, runs));
        HIP_TRY(sc.get(&sub_starts, (size_t)nq + 1));
        HIP_TRY(sc.get(&d_total, 1));
        k_sub_counts<<<ceil_div(runs, 256), 256, 0, s>>>(p->starts, runs, cnt);
        HIP_TRY(hipGetLastError());
        HEGJ_TRY(scan_u32(cnt, first, runs, d_total, s));
        k_sub_fill<<<ceil_div(runs, 256), 256, 0, s>>>(p->starts, runs, first, sub_starts);
        HIP_TRY(hipGetLastError());
        uint32_t units = 0;
        HIP_TRY(hipMemcpyAsync(&units, d_total, sizeof(units), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpyAsync(p->starts, sub_starts, sizeof(uint32_t) * units, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(p->starts + units, &p->n_live, sizeof(uint32_t), hipMemcpyHostToDevice, s));
        p->n_units = units;
    }
    HIP_TRY(hipStreamSynchronize(s));
    return HEGJ_OK;
}

hegj_status query_run(const QJoinArgs &base, bool fill, int device, hipStream_t s) {
    if (base.n_units == 0) return HEGJ_OK;
    // resident workgroups of the join on this device (thread-safe, once)
    static const uint32_t resident = [device]() {
        hipDeviceProp_t prop;
        int per_cu = 0;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 1024u;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void *>(&k_qjoin<6, true>), 64 * kQWaves, 0) != hipSuccess)
            per_cu = 1;
        return (uint32_t)std::max(per_cu, 1) * (uint32_t)prop.multiProcessorCount;
    }();
    const uint32_t blocks =
        std::max<uint32_t>(1, std::min<uint32_t>(resident, ceil_div(base.n_units, kQWaves)));
    return fill ? launch_qd<true>(base, blocks, s) : launch_qd<false>(base, blocks, s);
}

hegj_status query_hash(const uint64_t *offsets, const uint32_t *nbrs, uint32_t nq, uint64_t total,
                       unsigned long long *hash, hipStream_t s) {
    if (nq == 0 || total == 0) return HEGJ_OK;
    const bool longrows = total / nq >= 256;
    const uint32_t G = longrows ? 64 : 16;
    const uint32_t blocks = std::min<uint32_t>(std::max<uint32_t>(ceil_div((uint64_t)nq * G, 256), 1), 4096);
    if (longrows)
        k_csr_hash<64><<<blocks, 256, 0, s>>>(offsets, nbrs, nq, hash);
    else
        k_csr_hash<16><<<blocks, 256, 0, s>>>(offsets, nbrs, nq, hash);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

hegj_status query_to_f32(const double *q, uint32_t nq, uint32_t dim, const double *centre, float *q32,
                         unsigned long long *rmax_bits, hipStream_t s) {
    const uint64_t total = (uint64_t)nq * dim;
    if (total == 0) return HEGJ_OK;
    k_q32<<<ceil_div(total, 256), 256, 0, s>>>(q, total, dim, centre, q32, rmax_bits);
    HIP_TRY(hipGetLastError());
    return HEGJ_OK;
}

hegj_status query_offsets(const uint32_t *counts, uint32_t nq, uint64_t *offsets, hipStream_t s) {
    Scratch sc;
    uint64_t *c64;
    HIP_TRY(sc.get(&c64, nq));
    k_q_u32_u64<<<ceil_div(nq, 256), 256, 0, s>>>(counts, nq, c64);
    HIP_TRY(hipGetLastError());
    return scan_u64(c64, offsets, nq, offsets + nq, s);
}

}  // namespace hegj
