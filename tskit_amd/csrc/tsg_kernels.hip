// HIP kernels for the tree-sequence statistics sweep (gfx950 / CDNA4).
//
// Work decomposition: the genome is cut into chunks of consecutive trees and
// each chunk is one 64-lane wavefront (one workgroup).  A chunk executes its
// sweep program (see tsg_treeseq.cpp): it seeds per-sample-set counts in its
// start tree, then applies the edge removals and insertions of its
// transitions in the reference order, updating each changed edge's recorded
// root path.  Lanes run over state columns (sample sets) for count updates
// and over (output, hop) pairs or outputs for summary functions and window
// sums.
//
// Node records {parent, branch length, x[0..K)}: with the hot plan only the
// nodes a chunk touches, as LDS slots; with the full plan every node of the
// chunk's start tree, in LDS when they fit, otherwise in a per-chunk HBM slab.
//
// Reference algorithm restated here:
//   branch sweep      c/tskit/trees.c:1352-1523 (update_state 1317-1327,
//                     update_running_sum 1339-1350, unpolarised 1944-1972)
//   site sweep        c/tskit/trees.c:1525-1776
//   site AFS          c/tskit/trees.c:3469-3648
//   branch AFS        c/tskit/trees.c:3650-3813
//   summary functions c/tskit/trees.c:3934-5291
// Differences that are deliberate (and exact up to rounding):
//   * running branch sums are kept per chunk relative to the chunk start and
//     re-based afterwards with an FP64 prefix over chunks (the closing
//     transition of each chunk makes the deltas chain); the first tree's sum
//     is evaluated over its nodes in parallel;
//   * summary numerators are integer polynomials of the counts, exact in
//     FP64, scaled once by the reciprocal denominator; running sums and
//     window sums are FP64;
//   * for degree-2 summaries a walk's contribution is a closed form in two
//     path moments (see k_sweep);
//   * a walk whose moved subtree carries zero weight changes no count and no
//     summary, so it is skipped; an output none of whose columns moved is not
//     re-evaluated along the walk (its terms cancel exactly).  A summary with
//     a zero denominator is NaN at every state, so the edge's own term (always
//     added) carries the reference's NaN into the sums.
#include <hip/hip_runtime.h>

#include <climits>

#include <type_traits>

#include "tsg_common.h"
#include "tsg_kernels.h"

#define WAVE 64

namespace {

// Reference summary function o.stat evaluated on counts x (or on the
// complement tot - x for the unpolarised wrapper), in FP64 as the reference
// does.  Counts are converted to FP64 first: every product of up to three
// counts below 2^17.6 (n <= ~2e5 samples per set) is exact in FP64, so the
// numerators are the exact integer polynomials, scaled once by
// 1/denominator.  (f2 / f3 / f4 multiply two exact factors, one rounding.)
//
// State columns of real-valued weights (T = double) take the weighted
// summaries, whose parameters beyond one denominator are in opar.
template <typename T, bool PRE = false>
__device__ double eval_summary(const TsgOut &o, const T *x, const T *tot, bool comp,
                               const double *opar = nullptr) {
#define XV(col) ((double) (comp ? (tot[col] - x[col]) : x[col]))
    // PRE: the four slots' values are read together, before the summary's
    // case (every slot names a valid column, unused ones the first), for
    // callers whose lanes take different cases; otherwise each case reads
    // the slots it uses (fewer live registers)
    const double xi = XV(o.i);
    double pj = 0.0, pk = 0.0, pl = 0.0;
    if constexpr (PRE) {
        pj = XV(o.j);
        pk = XV(o.k);
        pl = XV(o.l);
    }
#define XJ (PRE ? pj : XV(o.j))
#define XK (PRE ? pk : XV(o.k))
#define XL (PRE ? pl : XV(o.l))
    const double ni = (double) o.ni;
    switch (o.stat) {
    case TSG_STAT_DIVERSITY:
        return ((xi * (ni - xi)) * o.inv);
    case TSG_STAT_SEGREGATING_SITES:
        return (xi > 0) ? (ni - xi) * o.inv : 0.0;
    case TSG_STAT_Y1:
        return ((xi * (ni - xi) * (ni - xi - 1)) * o.inv);
    case TSG_STAT_DIVERGENCE:
        return ((xi * ((double) o.nj - XJ)) * o.inv);
    case TSG_STAT_RELATEDNESS: {
        double sum = 0;
        for (int c = o.cbase; c < o.cbase + o.ccount; c++) {
            sum += XV(c) / (double) tot[c];
        }
        const double mean = sum / (double) o.ccount;
        return ((xi / ni - mean) * (XJ / (double) o.nj - mean));
    }
    case TSG_STAT_RELATEDNESS_NONCENTRED:
        return ((xi * XJ) * o.inv);
    case TSG_STAT_Y2: {
        const double nj = (double) o.nj;
        return ((xi * (nj - XJ) * (nj - XJ - 1)) * o.inv);
    }
    case TSG_STAT_F2: {
        const double nj = (double) o.nj;
        const double a = xi * (nj - XJ);
        const double b = (xi - 1) * (nj - XJ - 1) - (ni - xi) * XJ;
        return (a * b * o.inv);
    }
    case TSG_STAT_Y3: {
        return ((xi * ((double) o.nj - XJ) * ((double) o.nk - XK)) * o.inv);
    }
    case TSG_STAT_F3: {
        const double nk = (double) o.nk;
        const double a = xi * ((double) o.nj - XJ);
        const double b = (xi - 1) * (nk - XK) - (ni - xi) * XK;
        return (a * b * o.inv);
    }
    case TSG_STAT_F4: {
        const double nk = (double) o.nk, nl = (double) o.nl;
        const double a = xi * ((double) o.nj - XJ);
        const double b = XK * (nl - XL) - XL * (nk - XK);
        return (a * b * o.inv);
    }
    default:
        break;
    }
    if constexpr (std::is_same<T, double>::value) {
    switch (o.stat) {
    case TSG_STAT_TRAIT_COVARIANCE:  // x^2 / (2 (n - 1)^2)
        return xi * xi * o.inv;
    case TSG_STAT_TRAIT_CORRELATION: {  // x^2 / (2 p (1 - p) n (n - 1)), p = x_l
        const double p = XL;
        return (p > 0.0 && p < 1.0) ? (xi * xi) * o.inv / (p * (1 - p)) : 0.0;
    }
    case TSG_STAT_TRAIT_LINEAR_MODEL: {
        // trait i, covariates cbase .. cbase+ccount-1 with their row of V
        // (opar + par), the sample count m = x_l of n
        const double m = XL, n = (double) o.ni;
        if (!(m > 0.0 && m < n)) return 0.0;
        const double *v = opar + o.par;
        double a = xi, denom = m;
        for (int c = 0; c < o.ccount; c++) {
            const double z = XV(o.cbase + c);
            a -= z * v[c];
            denom -= z * z;
        }
        return denom < 1e-8 ? 0.0 : (a * a) / (2 * denom * denom);
    }
    case TSG_STAT_RELATEDNESS_WEIGHTED: {  // (x_i - a x_l)(x_j - b x_l)
        const double pn = XL;
        return (xi - opar[o.par] * pn) * (XJ - opar[o.par + 1] * pn);
    }
    case TSG_STAT_RELATEDNESS_WEIGHTED_NONCENTRED:
        return xi * XJ;
    default:
        break;
    }
    }
    return 0.0;
#undef XJ
#undef XK
#undef XL
#undef XV
}

// FP64 value held in two 32-bit words (no 8-byte alignment needed)
__device__ __forceinline__ double ld_f64(const int32_t *p) { return __hiloint2double(p[1], p[0]); }
__device__ __forceinline__ void st_f64(int32_t *p, double v) {
    p[0] = __double2loint(v);
    p[1] = __double2hiint(v);
}

// Per-chunk node storage of the full plan.  rec[u*R] = parent,
// rec[u*R+1..2] = FP64 branch length, rec[u*R+3+k] = count of state column
// k below u (R = 3 + K words).
struct Nodes {
    int32_t *rec;
    int32_t R;
    __device__ int32_t parent(int u) const { return rec[(int64_t) u * R]; }
    __device__ double bl(int u) const { return ld_f64(rec + (int64_t) u * R + 1); }
    __device__ int32_t &x(int u, int k) const { return rec[(int64_t) u * R + 3 + k]; }
    __device__ int32_t get(int u, int k) const { return x(u, k); }
    __device__ void put(int u, int k, int32_t v) const { x(u, k) = v; }
    __device__ void set_bl(int u, double b) const { st_f64(rec + (int64_t) u * R + 1, b); }
    __device__ void set_edge(int u, int p, double b) const {
        rec[(int64_t) u * R] = p;
        st_f64(rec + (int64_t) u * R + 1, b);
    }
};

// Hot-plan slot storage in LDS: (BL, branch mode) rec[h*R..+1] = FP64
// branch length, then the slot's K counts, as 32-bit words or, when every
// sample set has at most 65535 samples (C16), as 16-bit halves (R =
// TSG_SLOT_WORDS(K, C16, BL)).  Site sweeps keep no branch lengths.  No
// parent pointers: the root paths are precomputed.
template <bool C16, bool BL>
struct Slots {
    static constexpr int OFF = BL ? 2 : 0;  // words before the counts
    static constexpr bool WEIGHTED = false;
    int32_t *rec;
    int32_t R;
    __device__ double bl(int h) const {
        if constexpr (BL) return ld_f64(rec + h * R);
        return 0.0;
    }
    __device__ void set_bl(int h, double b) const {
        if constexpr (BL) st_f64(rec + h * R, b);
    }
    __device__ int32_t get(int h, int k) const {
        if (C16) return reinterpret_cast<const uint16_t *>(rec + h * R + OFF)[k];
        return rec[h * R + OFF + k];
    }
    __device__ void put(int h, int k, int32_t v) const {
        if (C16) {
            reinterpret_cast<uint16_t *>(rec + h * R + OFF)[k] = (uint16_t) v;
        } else {
            rec[h * R + OFF + k] = v;
        }
    }
    // atomic += v (v >= 0); with 16-bit halves no carry crosses into the
    // neighbour, since no count exceeds its set size
    __device__ void add(int h, int k, int32_t v) const {
        if (C16) {
            atomicAdd(reinterpret_cast<unsigned *>(rec + h * R + OFF + (k >> 1)),
                      (unsigned) v << (16 * (k & 1)));
        } else {
            atomicAdd(&rec[h * R + OFF + k], v);
        }
    }
};

// The same with FP64 state columns (real-valued sample weights): rec[h*R..+1]
// = branch length (BL), then K doubles (R = TSG_SLOT_WORDS_W(K, BL), even;
// rec is 8-byte aligned).
template <bool BL>
struct SlotsW {
    static constexpr int OFF = BL ? 2 : 0;
    static constexpr bool WEIGHTED = true;
    int32_t *rec;
    int32_t R;
    __device__ double bl(int h) const {
        if constexpr (BL) return ld_f64(rec + h * R);
        return 0.0;
    }
    __device__ void set_bl(int h, double b) const {
        if constexpr (BL) st_f64(rec + h * R, b);
    }
    __device__ double *col(int h) const { return reinterpret_cast<double *>(rec + h * R + OFF); }
    __device__ double get(int h, int k) const { return col(h)[k]; }
    __device__ void put(int h, int k, double v) const { col(h)[k] = v; }
    __device__ void add(int h, int k, double v) const { atomicAdd(col(h) + k, v); }
};

// The sweep's workgroup is one wavefront.  Its lanes run in lockstep and the
// LDS executes one wavefront's instructions in issue order, so an LDS write
// is seen by every later LDS read of the wavefront without a hardware
// barrier; wave_sync only keeps the compiler from moving memory operations
// across the point.  Node records in HBM (the full plan's global store) take
// a real barrier.
#ifndef TSG_SWEEP_SYNC
#define TSG_SWEEP_SYNC 1
#endif
__device__ __forceinline__ void wave_sync() {
#if TSG_SWEEP_SYNC
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
#else
    __syncthreads();
#endif
}

// Materialise the chunk's start tree and seed the counts: each weighted
// sample walks its root path once, adding 1 to every column it belongs to.
__device__ void init_chunk(const TsgSweepArgs &a, const Nodes &nd, int64_t n0, int nloc,
                           int ncols, int lane) {
    (void) ncols;
    const int64_t words = (int64_t) nloc * nd.R;
    for (int64_t i = lane; i < words; i += WAVE) nd.rec[i] = 0;  // coalesced
    __syncthreads();
    for (int u = lane; u < nloc; u += WAVE) nd.set_edge(u, a.loc_parent0[n0 + u], a.loc_bl0[n0 + u]);
    __syncthreads();
    for (int i = lane; i < a.nseed; i += WAVE) {
        const int c0 = a.seed_off[i], c1 = a.seed_off[i + 1];
        for (int u = a.seed_sample[i]; u >= 0; u = nd.parent(u)) {
            for (int c = c0; c < c1; c++) atomicAdd(&nd.x(u, a.seed_col[c]), 1);
        }
    }
    __syncthreads();
}

// The chunk a sweep workgroup takes: workgroups go to the 8 XCDs in turn
// (workgroup b to XCD b mod 8), so XCD x is given chunks x*per, x*per + 1,
// ... in dispatch order and the chunks in flight on one XCD are neighbours
// (their sample-map lines and plan data share that XCD's L2).  The grid is
// 8 * per workgroups; those past the last chunk return at once.
__device__ __forceinline__ int xcd_chunk(int b, int C) {
    const int per = (C + 7) / 8;
    return (b % 8) * per + b / 8;
}

// Chunk c's column of the device sample map (stride TSG_SS_CB per sample).
__device__ __forceinline__ const uint16_t *sample_map(const TsgSweepArgs &a, int c) {
    return a.sample_slot + (int64_t) (c / TSG_SS_CB) * a.num_samples * TSG_SS_CB + c % TSG_SS_CB;
}

// Hot plan: zero the slots' counts, load their start-tree branch lengths,
// add every weighted sample into its nearest hot ancestor-or-self, then add
// each depth level of the hot forest into its parents, deepest level first
// (the counts of a level are complete before it is added upward).
// With real-valued weights (SlotsW) a pair adds its weight instead of 1.
template <class NS>
__device__ void init_hot(const TsgSweepArgs &a, const NS &nd, int c, int nslot, int K,
                         int lane) {
    const int64_t s0 = a.chunk_node0[c];
    const int CW = nd.R - NS::OFF;  // count words per slot
    if (a.seed_in != nullptr) {
        // the counts the batch's other run seeded for the same pairs
        if constexpr (NS::OFF != 0) {
            for (int h = lane; h < nslot; h += WAVE) nd.set_bl(h, a.loc_bl0[s0 + h]);
        }
        const int32_t *src = a.seed_in + s0 * CW;
        for (int i = lane; i < nslot * CW; i += WAVE) {
            const int h = i / CW;
            nd.rec[h * nd.R + NS::OFF + (i - h * CW)] = src[i];
        }
        wave_sync();
        return;
    }
    // parents of the slots this lane owns in the level pass (slot lane + 64 r),
    // read now so that their loads overlap the seeding
    constexpr int PR = 8;
    const bool own = nslot <= PR * WAVE;
    int par[PR];
#pragma unroll
    for (int r = 0; r < PR; r++) {
        const int h = lane + r * WAVE;
        par[r] = (own && h < nslot) ? a.slot_par[s0 + h] : -1;
    }
    const int words = nslot * nd.R;
    for (int i = lane; i < words; i += WAVE) nd.rec[i] = 0;
    wave_sync();
    if constexpr (NS::OFF != 0) {
        for (int h = lane; h < nslot; h += WAVE) nd.set_bl(h, a.loc_bl0[s0 + h]);
    }
    // (sample, column) pairs, several per lane in flight: the slot gathers
    // are independent reads
    const uint16_t *ss = sample_map(a, c);
    constexpr int U = 16;
    for (int i0 = lane; i0 < a.npairs; i0 += WAVE * U) {
        int ps[U], pc[U];
        unsigned ph[U];
#pragma unroll
        for (int r = 0; r < U; r++) {
            const int i = i0 + r * WAVE;
            ps[r] = (i < a.npairs) ? a.pair_sample[i] : -1;
            pc[r] = (i < a.npairs) ? a.pair_col[i] : 0;
        }
#pragma unroll
        for (int r = 0; r < U; r++) {
            ph[r] = (ps[r] >= 0) ? (unsigned) ss[(int64_t) ps[r] * TSG_SS_CB] : TSG_NO_SLOT;
        }
#pragma unroll
        for (int r = 0; r < U; r++) {
            if (ph[r] == TSG_NO_SLOT) continue;
            if constexpr (NS::WEIGHTED) {
                nd.add((int) ph[r], pc[r], a.pair_w[i0 + r * WAVE]);
            } else {
                nd.add((int) ph[r], pc[r], 1);
            }
        }
    }
    wave_sync();
    const int32_t *lev = a.lev_slot0 + a.chunk_lev0[c];
    const int nlev = (int) (a.chunk_lev0[c + 1] - a.chunk_lev0[c]) - 1;
    // the last level holds the roots of the hot forest: nothing above them
    for (int l = 0; l + 1 < nlev; l++) {
        const int h0 = lev[l], h1 = lev[l + 1];
        if (own) {
#pragma unroll
            for (int r = 0; r < PR; r++) {
                const int h = lane + r * WAVE;
                if (h < h0 || h >= h1 || par[r] < 0) continue;
                for (int k = 0; k < K; k++) {
                    const auto v = nd.get(h, k);
                    if (v != 0) nd.add(par[r], k, v);
                }
            }
        } else {
            for (int idx = lane; idx < (h1 - h0) * K; idx += WAVE) {
                const int h = h0 + idx / K, k = idx % K;
                const int p = a.slot_par[s0 + h];
                const auto v = nd.get(h, k);
                if (p >= 0 && v != 0) nd.add(p, k, v);
            }
        }
        wave_sync();
    }
    if (a.seed_out != nullptr) {
        int32_t *dst = a.seed_out + s0 * CW;
        for (int i = lane; i < nslot * CW; i += WAVE) {
            const int h = i / CW;
            dst[i] = nd.rec[h * nd.R + NS::OFF + (i - h * CW)];
        }
    }
}

struct LaneOut {
    int lm;         // output index within this launch (-1: none)
    int wi, wend;
    const double *win;
    int64_t slot0;  // partial index of window wfirst
    int Mq;
    uint64_t mask;  // state columns the summary reads
    double f0;      // its value on an all-zero state (unpolarised wrapper included)
    double R, A;    // running branch sum, window sum (FP64)
    // how a walk's contribution is formed (see k_sweep): TSG_KIND_POLY, the
    // summary (unpolarised wrapper included) as a polynomial in the counts
    // of four column slots v_s = x[pc[s]]:
    //   F = c0 + l1.v + v'Hv / 2 + c3 v_0^3,
    // H symmetric, its entry (s, t), s <= t, in q[qidx(s,t)];
    // TSG_KIND_SEG, segregating sites from per-column indicator sums;
    // TSG_KIND_HOP, evaluated hop by hop
    int kind;
    int pc[4];
    double c0, c3;
    double l1[4];
    double q[10];
    // TSG_KIND_SEG reuses the slots: column pc[0], 1 / denominator c0,
    // polarised pc[1] (the set size is the column total)
};

#define TSG_KIND_HOP 0
#define TSG_KIND_POLY 1
#define TSG_KIND_SEG 2

// Segregating-sites numerators for set size n (times 1/n = inv): polarised
// f(x) = [x > 0] (n - x); unpolarised f(x) + f(n - x) = [x > 0] (n - x) +
// [x < n] x, which is n for 0 < x < n and 0 otherwise (trees.c:4222-4236).
__device__ __forceinline__ double seg_pol(double x, double n) { return x > 0 ? n - x : 0.0; }
__device__ __forceinline__ double seg_unpol(double x, double n) {
    return (x > 0 ? n - x : 0.0) + (x < n ? x : 0.0);
}

// State columns a summary function reads; all ones when it cannot be told
// (more than 64 columns) or when the summary has a zero denominator (its NaN
// must reach every sum, as in the reference).
__device__ uint64_t summary_columns(const TsgOut &o, int K) {
    if (K > 64 || !isfinite(o.inv)) return ~0ull;
    uint64_t m = 1ull << o.i;
    switch (o.stat) {
    case TSG_STAT_DIVERSITY:
    case TSG_STAT_SEGREGATING_SITES:
    case TSG_STAT_Y1:
        break;
    case TSG_STAT_DIVERGENCE:
    case TSG_STAT_Y2:
    case TSG_STAT_F2:
    case TSG_STAT_RELATEDNESS_NONCENTRED:
        m |= 1ull << o.j;
        break;
    case TSG_STAT_Y3:
    case TSG_STAT_F3:
        m |= (1ull << o.j) | (1ull << o.k);
        break;
    case TSG_STAT_F4:
        m |= (1ull << o.j) | (1ull << o.k) | (1ull << o.l);
        break;
    case TSG_STAT_RELATEDNESS:
        m |= 1ull << o.j;
        for (int c = o.cbase; c < o.cbase + o.ccount; c++) m |= 1ull << c;
        break;
    case TSG_STAT_TRAIT_COVARIANCE:
        break;
    case TSG_STAT_TRAIT_CORRELATION:
        m |= 1ull << o.l;
        break;
    case TSG_STAT_TRAIT_LINEAR_MODEL:
        m |= 1ull << o.l;
        for (int c = o.cbase; c < o.cbase + o.ccount; c++) m |= 1ull << c;
        break;
    case TSG_STAT_RELATEDNESS_WEIGHTED:
        m |= (1ull << o.j) | (1ull << o.l);
        break;
    case TSG_STAT_RELATEDNESS_WEIGHTED_NONCENTRED:
        m |= 1ull << o.j;
        break;
    default:
        return ~0ull;
    }
    return m;
}

// A column mask in the sweep's count-update layout, where column k's count
// is held by lanes k G .. k G + G - 1 and the ballot of the moved counts
// sets bit k G: column k's bit moved to bit k G (all ones stay all ones).
__device__ __forceinline__ uint64_t spread_columns(uint64_t m, int K, int G) {
    if (G == 1 || m == ~0ull) return m;
    uint64_t r = 0;
    for (int k = 0; k < K; k++) {
        if ((m >> k) & 1) r |= 1ull << (k * G);
    }
    return r;
}

// index of the slot pair (s, t), s <= t, in LaneOut::q
__host__ __device__ __forceinline__ int qidx(int s, int t) {
    return s == 0 ? t : (s == 1 ? 3 + t : (s == 2 ? 5 + t : 9));
}

// Polynomial coefficients of the summaries that are polynomials of degree
// <= 2 in the counts, or <= 3 in one column, once the unpolarised wrapper
// f(x) + f(total - x) is folded in: diversity (trees.c:3935-3954), Y1
// (4249-4264) and divergence (4691-4721) and non-centred relatedness
// (4756-4790) either way; Y2 (4900-4925), f2 (4939-4965), Y3 (5178-5205),
// f3 (5220-5250) and f4 (5267-5291) unpolarised, where their cubic and
// quartic terms cancel.  The coefficients
// were derived by expanding f(x) + f(total - x).  Segregating sites
// (4222-4236) is TSG_KIND_SEG; the rest are TSG_KIND_HOP.
// Weighted summaries (real-valued columns; totw the columns' total weights,
// opar their parameters): trait covariance (trees.c:3961-3974) and weighted
// relatedness (4801-4837) are quadratic too.
template <bool WGT = false>
__host__ __device__ int moment_form(const TsgOut &o, int K, LaneOut &L,
                                    const double *totw = nullptr, const double *opar = nullptr) {
    if (K > 64 || !__builtin_isfinite(o.inv)) return TSG_KIND_HOP;
    const bool pol = (o.flags & 1) != 0;
    const double ni = (double) o.ni, nj = (double) o.nj, nk = (double) o.nk, nl = (double) o.nl;
    L.pc[0] = o.i;
    L.pc[1] = o.j;
    L.pc[2] = o.k;
    L.pc[3] = o.l;
    L.c0 = L.c3 = 0;
    for (int s = 0; s < 4; s++) L.l1[s] = 0;
    for (int s = 0; s < 10; s++) L.q[s] = 0;
    double *l1 = L.l1, *q = L.q;
    switch (o.stat) {
    case TSG_STAT_SEGREGATING_SITES:
        return TSG_KIND_SEG;
    case TSG_STAT_DIVERSITY:  // x (n - x); f(n - x) = f(x)
        l1[0] = (pol ? 1 : 2) * ni;
        q[0] = pol ? -1 : -2;
        break;
    case TSG_STAT_Y1:
        // x (n - x)(n - x - 1) = n(n-1) x - (2n-1) x^2 + x^3; unpolarised
        // + (n - x) x (x - 1): (n - 2)(n x - x^2)
        if (pol) {
            l1[0] = ni * (ni - 1);
            q[0] = -(2 * ni - 1);
            L.c3 = 1;
        } else {
            l1[0] = ni * (ni - 2);
            q[0] = -(ni - 2);
        }
        break;
    case TSG_STAT_DIVERGENCE:  // x_i (n_j - x_j) [+ (n_i - x_i) x_j]
        l1[0] = nj;
        q[qidx(0, 1)] = -1;
        if (!pol) {
            l1[1] = ni;
            q[qidx(0, 1)] = -2;
        }
        break;
    case TSG_STAT_RELATEDNESS_NONCENTRED:  // x_i x_j [+ (n_i - x_i)(n_j - x_j)]
        q[qidx(0, 1)] = 1;
        if (!pol) {
            L.c0 = ni * nj;
            l1[0] = -nj;
            l1[1] = -ni;
            q[qidx(0, 1)] = 2;
        }
        break;
    case TSG_STAT_Y2:
        if (pol) return TSG_KIND_HOP;
        l1[0] = nj * (nj - 1);
        l1[1] = -ni;
        q[qidx(0, 1)] = -2 * (nj - 1);
        q[qidx(1, 1)] = ni;
        break;
    case TSG_STAT_F2:
        if (pol) return TSG_KIND_HOP;
        l1[0] = -nj * (nj - 1);
        l1[1] = -ni * (ni - 1);
        q[qidx(0, 0)] = nj * (nj - 1);
        q[qidx(0, 1)] = -2 * (ni - 1) * (nj - 1);
        q[qidx(1, 1)] = ni * (ni - 1);
        break;
    case TSG_STAT_Y3:
        if (pol) return TSG_KIND_HOP;
        l1[0] = nj * nk;
        q[qidx(0, 1)] = -nk;
        q[qidx(0, 2)] = -nj;
        q[qidx(1, 2)] = ni;
        break;
    case TSG_STAT_F3:
        if (pol) return TSG_KIND_HOP;
        l1[0] = -nj * nk;
        q[qidx(0, 0)] = nj * nk;
        q[qidx(0, 1)] = -nk * (ni - 1);
        q[qidx(0, 2)] = -nj * (ni - 1);
        q[qidx(1, 2)] = ni * (ni - 1);
        break;
    case TSG_STAT_F4:  // (n_j x_i - n_i x_j)(n_l x_k - n_k x_l)
        if (pol) return TSG_KIND_HOP;
        q[qidx(0, 2)] = nj * nl;
        q[qidx(0, 3)] = -nj * nk;
        q[qidx(1, 2)] = -ni * nl;
        q[qidx(1, 3)] = ni * nk;
        break;
    default:
        if constexpr (!WGT) return TSG_KIND_HOP;
        break;
    }
    if constexpr (WGT) {
    if (o.stat >= TSG_STAT_TRAIT_COVARIANCE) switch (o.stat) {
    case TSG_STAT_TRAIT_COVARIANCE: {  // x^2 [+ (T - x)^2]
        const double t = totw[o.i];
        q[0] = pol ? 1 : 2;
        if (!pol) {
            L.c0 = t * t;
            l1[0] = -2 * t;
        }
        break;
    }
    case TSG_STAT_RELATEDNESS_WEIGHTED_NONCENTRED: {  // x_i x_j [+ (T_i - x_i)(T_j - x_j)]
        q[qidx(0, 1)] = pol ? 1 : 2;
        if (!pol) {
            const double ti = totw[o.i], tj = totw[o.j];
            L.c0 = ti * tj;
            l1[0] = -tj;
            l1[1] = -ti;
        }
        break;
    }
    case TSG_STAT_RELATEDNESS_WEIGHTED: {
        // slots (x_i, x_j, x_l): (x_i - a x_l)(x_j - b x_l); unpolarised adds
        // (A - x_i + a x_l)(B - x_j + b x_l), A = T_i - a T_l, B = T_j - b T_l
        L.pc[2] = o.l;
        const double ca = opar[o.par], cb = opar[o.par + 1];
        const double g = pol ? 1 : 2;
        q[qidx(0, 1)] = g;
        q[qidx(0, 2)] = -g * cb;
        q[qidx(1, 2)] = -g * ca;
        q[qidx(2, 2)] = g * ca * cb;
        if (!pol) {
            const double A = totw[o.i] - ca * totw[o.l], B = totw[o.j] - cb * totw[o.l];
            L.c0 = A * B;
            l1[0] = -B;
            l1[1] = -A;
            l1[2] = A * cb + B * ca;
        }
        break;
    }
    default:
        return TSG_KIND_HOP;
    }
    }
    // q holds the coefficients of v_s v_t (s <= t); H has the off-diagonal
    // ones and twice the diagonal ones
    const double inv = o.inv;
    L.c0 *= inv;
    L.c3 *= inv;
    for (int s = 0; s < 4; s++) l1[s] *= inv;
    for (int s = 0; s < 10; s++) q[s] *= inv;
    for (int s = 0; s < 4; s++) q[qidx(s, s)] *= 2;
    return TSG_KIND_POLY;
}

// The parts of F at the slot counts v = x[pc] (TSG_KIND_POLY) that a walk
// reuses: A = l1.v + c3 v_0^3, u = H v and B = v'Hv / 2, so F = c0 + A + B.
struct PolyParts {
    double A, B, v0;
    double u[4];
};

template <typename T>
__device__ __forceinline__ PolyParts poly_parts(const LaneOut &L, const T *x) {
    PolyParts P;
    double v[4];
#pragma unroll
    for (int s = 0; s < 4; s++) v[s] = (double) x[L.pc[s]];
    double bq = 0.0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        double us = 0.0;
#pragma unroll
        for (int t = 0; t < 4; t++) us = fma(L.q[s <= t ? qidx(s, t) : qidx(t, s)], v[t], us);
        P.u[s] = us;
        bq = fma(v[s], us, bq);
    }
    double a = L.c3 * v[0] * v[0] * v[0];
#pragma unroll
    for (int s = 0; s < 4; s++) a = fma(L.l1[s], v[s], a);
    P.A = a;
    P.B = 0.5 * bq;
    P.v0 = v[0];
    return P;
}

// F at the counts x (TSG_KIND_POLY)
template <typename T>
__device__ __forceinline__ double poly_value(const LaneOut &L, const T *x) {
    const PolyParts P = poly_parts(L, x);
    return L.c0 + P.A + P.B;
}

// sum_h bl_h (F(x_h + d) - F(x_h)) for the moved counts d = sg v (sg = +-1)
// from the path sums S0 = sum bl, S1[k] = sum bl x[k], S2[k] = sum bl x[k]^2:
// F(x + d) - F(x) = l1.d + d'Hd / 2 + x.Hd + c3 (d_0^3 + 3 x_0 d_0^2 +
// 3 x_0^2 d_0), so the sum is
//   S0 (sg A + B) + sg S1.u + 3 c3 (v_0^2 S1_0 + sg v_0 S2_0).
__device__ __forceinline__ double poly_walk(const LaneOut &L, const PolyParts &P, double sg,
                                            double s0, const double *s1v, const double *s2v,
                                            bool c3) {
    double g = 0.0;
#pragma unroll
    for (int s = 0; s < 4; s++) g = fma(s1v[L.pc[s]], P.u[s], g);
    double w = fma(s0, fma(sg, P.A, P.B), sg * g);
    if (c3) {
        const int p = L.pc[0];
        w = fma(3.0 * L.c3, P.v0 * fma(P.v0, s1v[p], sg * s2v[p]), w);
    }
    return w;
}

// v of lane + D (DPP row_shl:D, within the lane's row of 16 lanes)
template <int D>
__device__ __forceinline__ double lane_plus(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x100 + D, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x100 + D, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// S + q_0 + q_1 + ... + q_{G-1}, added one at a time in that order, where
// q_g is q of lane + g: a column's G hop groups sit in adjacent lanes (lane
// = k G + g, G <= 16 a power of two, within one row of 16), so the head lane
// of the column adds the terms of one round of G hops in hop order (lanes
// past the path's last hop hold zeros).  The whole wave runs it (uniform
// control flow); other lanes' S is not used.
__device__ __forceinline__ double hop_chain(double S, double q, int G) {
    S = S + q;
    if (G > 1) S = S + lane_plus<1>(q);
    if (G > 2) {
        S = S + lane_plus<2>(q);
        S = S + lane_plus<3>(q);
    }
    if (G > 4) {
        S = S + lane_plus<4>(q);
        S = S + lane_plus<5>(q);
        S = S + lane_plus<6>(q);
        S = S + lane_plus<7>(q);
    }
    if (G > 8) {
        S = S + lane_plus<8>(q);
        S = S + lane_plus<9>(q);
        S = S + lane_plus<10>(q);
        S = S + lane_plus<11>(q);
        S = S + lane_plus<12>(q);
        S = S + lane_plus<13>(q);
        S = S + lane_plus<14>(q);
        S = S + lane_plus<15>(q);
    }
    return S;
}

// Segregating-sites numerators as integers: polarised [x > 0] (n - x),
// unpolarised n [0 < x < n] (seg_pol, seg_unpol)
template <typename T>
__device__ __forceinline__ T iseg_pol(T x, T n) { return x > 0 ? n - x : T(0); }
template <typename T>
__device__ __forceinline__ T iseg_unpol(T x, T n) {
    return (x > 0 && x < n) ? n : T(0);
}

template <int MPL, bool WGT>
__device__ void setup_outputs(const TsgSweepArgs &a, const TsgOut *gout, int c, int lane, int G,
                              LaneOut (&lo)[MPL]) {
#pragma unroll
    for (int j = 0; j < MPL; j++) {
        LaneOut &L = lo[j];
        const int m = a.m_begin + lane + WAVE * j;
        L.lm = -1;
        L.R = L.A = 0.0;
        L.wi = L.wend = 0;
        L.mask = 0;
        L.f0 = 0.0;
        L.kind = TSG_KIND_HOP;
        if (m < a.m_end) {
            L.lm = m - a.m_begin;
            const TsgOut o = gout[L.lm];
            const TsgQuery &q = a.queries[o.q];
            const int64_t cq = (int64_t) c * a.Q + o.q;
            L.win = a.windows + q.win_off;
            L.wi = a.cq_wfirst[cq];
            L.wend = L.wi + a.cq_nwin[cq];
            L.Mq = q.M;
            L.slot0 = q.pbase + a.cq_poff[cq] * (int64_t) q.M + (m - q.mbase);
            L.mask = spread_columns(summary_columns(o, a.K), a.K, G);
            L.kind = moment_form<WGT>(o, a.K, L, a.col_total_w, a.opar);
            if (L.kind == TSG_KIND_SEG) {
                L.pc[0] = o.i;
                L.pc[1] = o.flags & 1;
                L.c0 = o.inv;
            }
        }
    }
}

__device__ __forceinline__ void flush_window(const TsgSweepArgs &a, LaneOut &L, int wfirst) {
    a.partial[L.slot0 + (int64_t) (L.wi - wfirst) * L.Mq] = L.A;
    L.A = 0.0;
}

// Unpolarised summaries add f(total - x) (trees.c:1944-1972).
template <typename T>
__device__ __forceinline__ double branch_summary(const TsgOut &o, const T *x, const T *tot,
                                                 const double *opar = nullptr) {
    double f = eval_summary(o, x, tot, false, opar);
    if (!(o.flags & 1)) f = f + eval_summary(o, x, tot, true, opar);
    return f;
}

}  // namespace

// MODE 0: branch statistics, MODE 1: site statistics.
// STORE: TSG_STORE_GLOBAL (full plan, node records in a per-chunk HBM slab),
// TSG_STORE_LDS (full plan, every local node in LDS), TSG_STORE_HOT (hot
// plan, only the chunk's hot slots, in LDS) or TSG_STORE_HOT16 (the same
// with 16-bit counts, for sample sets of at most 65535 samples).
//
// LDS: the staging block (TSG_STAGE_INTS) first, then for the LDS stores the
// node records.  The staging block holds this launch's outputs, a batch of
// trees (breakpoints, transition offsets, site ranges), a batch of events
// with their root paths, the moved subtree's counts, a path segment's counts
// before / after the update and its branch lengths, column totals, an
// all-zero state, the active-output list and one round of summary terms.
// Trees and events are read from HBM a batch at a time with coalesced loads,
// so a chunk pays the memory latency once per batch, not once per event.
//
// Branch mode, per event: the columns the moved subtree carries (a ballot)
// decide which outputs can change.  An output none of whose columns moved
// has F(old) == F(new) at every hop, so its hop terms cancel exactly and are
// not evaluated, and its child term is its all-zero value f0.  For the
// others the walk's contribution sum_h bl_h (F(x_h + d) - F(x_h)) (d = the
// moved counts, signed; trees.c:1436-1448, 1461-1472) is formed three ways:
//   * polynomial summaries of degree <= 2 in two columns or <= 3 in one
//     (diversity, divergence, non-centred relatedness, Y1): F(x + d) - F(x)
//     has degree one less in x, so the sum is a closed form in the path
//     sums S0 = sum_h bl_h, S1[k] = sum_h bl_h x_h[k] and S2[k] =
//     sum_h bl_h x_h[k]^2, which the count-update lanes accumulate;
//   * segregating sites: per column, sum_h bl_h (g(x_new) - g(x_old)) of
//     its indicator-valued numerator g, accumulated the same way;
//   * the others (f2, f3, f4, Y2, Y3, centred relatedness): (output, hop)
//     terms of a path segment are spread over the lanes (hop fastest),
//     written to LDS, and each output's owner lane adds them in hop order.
// Running sums and window sums are FP64, as in the reference.
// GC: the count-update layout's G when the launch fixes it at compile time
// (0: chosen from K at run time).
template <int MODE, int STORE, int MPL, int GC = 0>
__device__ __forceinline__ void sweep_chunk(const TsgSweepArgs &a, int c) {
    extern __shared__ int32_t smem[];
#define SWEEP_SYNC()                                     \
    do {                                                 \
        if constexpr (STORE == TSG_STORE_GLOBAL) {       \
            __syncthreads();                             \
        } else {                                         \
            wave_sync();                                 \
        }                                                \
    } while (0)
    const int lane = threadIdx.x;
    const int K = a.K;
    const int R = 3 + K;
    const int span = a.m_end - a.m_begin;
    const int64_t n0 = a.chunk_node0[c];
    const int nloc = (int) (a.chunk_node0[c + 1] - n0);
    // 8-byte items first; each lane keeps its outputs' specs in registers
    // (ow), the launch's specs stay in HBM (gout)
    const TsgOut *gout = a.outs + a.m_begin;
    double *tbp = reinterpret_cast<double *>(smem);                     // [TB+1] breakpoints
    int64_t *toff = reinterpret_cast<int64_t *>(tbp + TSG_TB + 1);       // [TB+1] tr_off
    int64_t *evp = toff + TSG_TB + 1;                                    // [EB+1] ev_path0
    double *s1v = reinterpret_cast<double *>(evp + TSG_EB + 1);          // [K] path sums
    double *s2v = s1v + K;                                               // [K]
    double *gpv = s2v + K;                                               // [K]
    double *guv = gpv + K;                                               // [K]
    // the per-hop arrays (td, sbl, xo beyond K, xn) only when some output
    // is evaluated hop by hop
    const bool hop_stage = MODE == 0 && a.no_hop == 0;
    double *td = guv + K;                                                // [WAVE] hop terms
    double *evb = td + (hop_stage ? WAVE : 0);                           // [EB] event branch length
    double *sbl = evb + TSG_EB;                                          // [PSEG] path branch lengths
    // state values: counts, or FP64 sums of real-valued weights
    constexpr bool WGT = STORE == TSG_STORE_HOTW;
    using Cnt = typename std::conditional<WGT, double, int32_t>::type;
    Cnt *xc = reinterpret_cast<Cnt *>(sbl + (hop_stage ? TSG_PSEG : 0)); // [K] moved subtree
    Cnt *xo = xc + K;                                                    // [PSEG x K] or [K]
    Cnt *xn = xo + (hop_stage ? TSG_PSEG * K : K);                       // [PSEG x K] or none
    Cnt *tot = xn + (hop_stage ? TSG_PSEG * K : 0);                      // [K]
    Cnt *zero = tot + K;                                                 // [K]
    int32_t *tnrem = reinterpret_cast<int32_t *>(zero + K);              // [TB]
    int32_t *ts0 = tnrem + TSG_TB;                                       // [TB+1] tree_site0
    int32_t *evc = ts0 + TSG_TB + 1;                                     // [EB] event child
    int32_t *pbuf = evc + TSG_EB;                                        // [PB] path entries
    int32_t *act = pbuf + TSG_PB;                                        // [span]
    constexpr bool HOTS = STORE == TSG_STORE_HOT || STORE == TSG_STORE_HOT16 || WGT;
    using NS = typename std::conditional<
        WGT, SlotsW<MODE == 0>,
        typename std::conditional<HOTS, Slots<STORE == TSG_STORE_HOT16, MODE == 0>,
                                  Nodes>::type>::type;
    NS nd;
    nd.R = WGT ? TSG_SLOT_WORDS_W(K, MODE == 0)
               : (HOTS ? TSG_SLOT_WORDS(K, STORE == TSG_STORE_HOT16, MODE == 0) : R);
    if (STORE != TSG_STORE_GLOBAL) {
        nd.rec = smem + (WGT ? TSG_STAGE_INTS_W(K, span, hop_stage)
                             : TSG_STAGE_INTS(K, span, hop_stage));
    } else {
        nd.rec = a.g_rec + (int64_t) c * a.g_stride;
    }
    const bool any_unpolarised = (a.flags & TSG_F_ANY_UNPOLARISED) != 0;
    const uint64_t below = (1ull << lane) - 1;

    for (int k = lane; k < K; k += WAVE) {
        if constexpr (WGT) {
            tot[k] = a.col_total_w[k];
        } else {
            tot[k] = a.col_total[k];
        }
        zero[k] = 0;
    }
    if constexpr (HOTS) {
        init_hot(a, nd, c, nloc, K, lane);
    } else {
        init_chunk(a, nd, n0, nloc, K, lane);
    }

    // count-update layout: with K <= 64 columns, each column has G lanes
    // (lane = k G + g, G the largest power of two <= 16 with G K <= 64, or
    // the launch's compile-time GC) that
    // update G hops of a path at once (round by round); a column's path sums
    // add the rounds' per-hop terms in hop order (hop_chain), so they do not
    // depend on K.  With more columns one group strides over them.  The
    // path-sum kinds (POLY, SEG) only occur with K <= 64.  The outputs'
    // column masks are in this layout (spread_columns).
    const bool kwide = GC > 0 ? false : (K > WAVE || K < 1);
    const int G = GC > 0 ? GC : (kwide ? 1 : min(16, 1 << (31 - __clz(WAVE / K))));
    const int uk = kwide ? lane : lane / G;
    const int ug = kwide ? 0 : lane - uk * G;
    const bool uon = kwide || uk < K;

    LaneOut lo[MPL];
    setup_outputs<MPL, WGT>(a, gout, c, lane, G, lo);
    int wfirst[MPL];
    bool any_poly = false, any_c3 = false, any_segp = false, any_segu = false, any_hop = false;
#pragma unroll
    for (int j = 0; j < MPL; j++) {
        wfirst[j] = lo[j].wi;
        if (MODE == 0 && lo[j].lm >= 0) lo[j].f0 = branch_summary(gout[lo[j].lm], zero, tot, a.opar);
        any_poly |= lo[j].kind == TSG_KIND_POLY;
        any_c3 |= lo[j].kind == TSG_KIND_POLY && lo[j].c3 != 0;
        any_segp |= lo[j].kind == TSG_KIND_SEG && lo[j].pc[1];
        any_segu |= lo[j].kind == TSG_KIND_SEG && !lo[j].pc[1];
        any_hop |= lo[j].lm >= 0 && lo[j].kind == TSG_KIND_HOP;
    }
    any_poly = MODE == 0 && __any(any_poly);
    any_c3 = MODE == 0 && __any(any_c3);
    any_segp = MODE == 0 && __any(any_segp);
    any_segu = MODE == 0 && __any(any_segu);
    // (without the per-hop arrays no output is of that kind: the host
    // decides no_hop with the same moment_form)
    any_hop = hop_stage && __any(any_hop);
    const bool any_seg = any_segp || any_segu;
    const bool any_agg = any_poly || any_seg;
    const Cnt totk = (!kwide && uon) ? tot[uk] : Cnt(0);

    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    // the chunk's events: transitions t0+1 .. t1 (the last is the closing
    // transition into the next chunk, when there is one)
    const int tlast = min(t1, a.num_trees - 1);
    const int64_t ev_end = a.tr_off[tlast + 1];

    // tree batch: entries t = tb .. tb + TB of tr_off / bp / tree_site0 and
    // tb .. tb + TB - 1 of tr_nrem (never past index num_trees)
    int tb = -1;
    auto stage_trees = [&](int t) {
        SWEEP_SYNC();
        tb = t;
        for (int l = lane; l <= TSG_TB; l += WAVE) {
            const int x = t + l;
            if (x <= a.num_trees) {
                toff[l] = a.tr_off[x];
                tbp[l] = a.bp[x];
                ts0[l] = a.tree_site0[x];
                if (l < TSG_TB) tnrem[l] = (x < a.num_trees) ? a.tr_nrem[x] : 0;
            }
        }
        SWEEP_SYNC();
    };
    auto tree = [&](int t) {  // make entries t and t + 1 resident
        if (tb < 0 || t < tb || t + 1 > tb + TSG_TB) stage_trees(t);
        return t - tb;
    };

    // event batch: events eb .. eb + neb - 1 with their paths in pbuf (or
    // read from HBM when the first event's path alone exceeds pbuf)
    int64_t eb = 0, pb0 = 0;
    int neb = 0;
    bool gpath = false;
    auto stage_events = [&](int64_t e) {
        SWEEP_SYNC();
        const int n = (int) min((int64_t) TSG_EB, ev_end - e);
        if (lane < n) {
            evc[lane] = a.ev_child[e + lane];
            evb[lane] = a.ev_bl[e + lane];
        }
        for (int l = lane; l <= n; l += WAVE) evp[l] = a.ev_path0[e + l];
        SWEEP_SYNC();
        const int64_t base = evp[0];
        // events whose paths end inside pbuf (a prefix, offsets increase)
        const bool fits = lane < n && evp[lane + 1] - base <= TSG_PB;
        const int nf = __popcll(__ballot(fits));
        neb = nf > 0 ? nf : 1;
        gpath = nf == 0;
        eb = e;
        pb0 = base;
        const int plen = gpath ? 0 : (int) (evp[neb] - base);
        for (int i = lane; i < plen; i += WAVE) pbuf[i] = a.path[base + i];
        SWEEP_SYNC();
    };

    auto apply_transition = [&](int t) {
        const int lt = tree(t);
        const int64_t e0 = toff[lt], e1 = toff[lt + 1];
        const int64_t nrem = tnrem[lt];
        for (int64_t e = e0; e < e1; e++) {
            if (e >= eb + neb) stage_events(e);
            const int le = (int) (e - eb);
            const bool ins = (e - e0) >= nrem;
            const int u = evc[le];
            const double ebl = evb[le];
            // the event's path range and the child's branch length are read
            // with its counts (independent LDS reads in flight together)
            const int64_t p0 = evp[le], p1 = evp[le + 1];
            const double ubl = MODE == 0 ? nd.bl(u) : 0.0;
            // columns the moved subtree carries: each lane reads the count of
            // its update column uk; the ballot sets column k's bit at k G
            uint64_t cm;
            Cnt xk = 0;
            if (!kwide) {
                if (uon) xk = nd.get(u, uk);
                if (uon && ug == 0) xc[uk] = xk;
                cm = __ballot(uon && ug == 0 && xk != 0);
            } else {
                bool nz = false;
                for (int k = lane; k < K; k += WAVE) {
                    const Cnt xw = nd.get(u, k);
                    xc[k] = xw;
                    nz |= (xw != 0);
                }
                cm = __any(nz) ? ~0ull : 0ull;
            }
            SWEEP_SYNC();
            // the edge's own term bl * F(x_child) (needed even for a zero-weight
            // subtree, because unpolarised summaries need not vanish at zero:
            // non-centred relatedness has F(0) = f(0) + f(total) = 1) is added
            // with the walk's, after the path update, which it does not
            // depend on (the path starts at the parent)
            const double sg = ins ? 1.0 : -1.0;
            const double b = MODE == 0 ? sg * (double) (ins ? ebl : ubl) : 0.0;
            // after seeding nothing reads parent pointers (the root paths
            // are precomputed), so only the branch length is kept current
            if (lane == 0) nd.set_bl(u, ins ? ebl : 0.0);
            if (cm == 0) {  // no count changes, no summary changes
                if (MODE == 0) {
#pragma unroll
                    for (int j = 0; j < MPL; j++) {
                        if (lo[j].lm >= 0) lo[j].R += b * lo[j].f0;
                    }
                }
                continue;
            }
            // outputs whose summaries can change, compacted into act[]
            int pos[MPL];
            int na = 0;
#pragma unroll
            for (int j = 0; j < MPL; j++) {
                const bool on = (MODE == 0) && lo[j].lm >= 0 && lo[j].kind == TSG_KIND_HOP &&
                                (lo[j].mask & cm) != 0;
                const uint64_t bal = __ballot(on);
                pos[j] = on ? na + __popcll(bal & below) : -1;
                if (on) act[pos[j]] = lo[j].lm;
                na += __popcll(bal);
            }
            // the path's hops are read from LDS: from the event batch, or,
            // for a path longer than the batch buffer (gpath: the batch is
            // this event alone), segment by segment copied into its start
            int poff = 0;
            auto path_node = [&](int i) -> int { return pbuf[poff + i]; };
            SWEEP_SYNC();
            // path sums of this lane's column uk over its hops: S0 = sum bl,
            // S1 = sum bl x, S2 = sum bl x^2 (x the count before the event),
            // and the segregating-sites numerator changes
            double s0 = 0, s1 = 0, s2 = 0, gp = 0, gu = 0;
            const Cnt dx = (!kwide && uon) ? (ins ? xk : -xk) : Cnt(0);
            for (int64_t sg0 = 0; sg0 < p1 - p0; sg0 += TSG_PSEG) {
                const int h = (int) min((int64_t) TSG_PSEG, p1 - p0 - sg0);
                if (gpath) {
                    if (lane < h) pbuf[lane] = a.path[p0 + sg0 + lane];
                    SWEEP_SYNC();
                } else {
                    poff = (int) (p0 - pb0 + sg0);
                }
                if (kwide) {
                    for (int i = 0; i < h; i++) {
                        const int node = path_node(i);
                        for (int k = lane; k < K; k += WAVE) {
                            const Cnt old = nd.get(node, k);
                            const Cnt nw = ins ? old + xc[k] : old - xc[k];
                            nd.put(node, k, nw);
                            if (MODE == 0 && any_hop) {
                                xo[i * K + k] = old;
                                xn[i * K + k] = nw;
                            }
                        }
                    }
                } else {
                    for (int r0 = 0; r0 < h; r0 += G) {
                        const int i = r0 + ug;
                        // this lane's hop's terms: bl, bl x, bl x^2 (x the
                        // count before the event), bl times the change of
                        // the segregating-sites numerators
                        double q0 = 0, q1 = 0, q2 = 0, qp = 0, qu = 0;
                        if (uon && i < h) {
                            const int node = path_node(i);
                            // the count and the branch length are read
                            // together, before the count's update
                            const Cnt old = nd.get(node, uk);
                            const double bw = (MODE == 0 && any_agg) ? nd.bl(node) : 0.0;
                            const Cnt nw = old + dx;
                            nd.put(node, uk, nw);
                            if (MODE == 0 && any_hop) {
                                xo[i * K + uk] = old;
                                xn[i * K + uk] = nw;
                            }
                            if (MODE == 0 && any_agg) {
                                q0 = bw;
                                if (any_poly) {
                                    const double xold = (double) old;
                                    q1 = bw * xold;
                                    if (any_c3) q2 = q1 * xold;
                                }
                                if (any_segp) {
                                    qp = bw * (double) (iseg_pol(nw, totk) - iseg_pol(old, totk));
                                }
                                if (any_segu) {
                                    qu = bw * (double) (iseg_unpol(nw, totk) - iseg_unpol(old, totk));
                                }
                            }
                        }
                        if (MODE == 0 && any_agg) {
                            if (any_poly) {
                                s0 = hop_chain(s0, q0, G);
                                s1 = hop_chain(s1, q1, G);
                                if (any_c3) s2 = hop_chain(s2, q2, G);
                            }
                            if (any_segp) gp = hop_chain(gp, qp, G);
                            if (any_segu) gu = hop_chain(gu, qu, G);
                        }
                    }
                }
                if (MODE == 0 && any_hop) {
                    for (int i = lane; i < h; i += WAVE) sbl[i] = nd.bl(path_node(i));
                }
                SWEEP_SYNC();
                if (MODE == 0 && na > 0) {
                    const int items = na * h;
                    for (int r0 = 0; r0 < items; r0 += WAVE) {
                        const int idx = r0 + lane;
                        if (idx < items) {
                            const int ai = idx / h, i = idx - ai * h;
                            const TsgOut o = gout[act[ai]];
                            const double fo = branch_summary(o, xo + i * K, tot, a.opar);
                            const double fn = branch_summary(o, xn + i * K, tot, a.opar);
                            td[lane] = (double) sbl[i] * (fn - fo);
                        }
                        SWEEP_SYNC();
#pragma unroll
                        for (int j = 0; j < MPL; j++) {
                            if (pos[j] < 0) continue;
                            const int q0 = max(pos[j] * h, r0), q1 = min(pos[j] * h + h, r0 + WAVE);
                            for (int q = q0; q < q1; q++) lo[j].R += td[q - r0];
                        }
                        SWEEP_SYNC();
                    }
                }
            }
            if (MODE == 0) {
                if (any_agg) {
                    // the head lane of column k (lane k G) holds its sums;
                    // S0 is column 0's, lane 0's
                    if (any_poly) {
                        s0 = __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(s0)),
                                              __builtin_amdgcn_readfirstlane(__double2loint(s0)));
                    }
                    if (uon && ug == 0) {
                        if (any_poly) s1v[uk] = s1;
                        if (any_c3) s2v[uk] = s2;
                        if (any_segp) gpv[uk] = gp;
                        if (any_segu) guv[uk] = gu;
                    }
                    SWEEP_SYNC();
                }
                // child term, then the walk's closed form (POLY, SEG)
#pragma unroll
                for (int j = 0; j < MPL; j++) {
                    LaneOut &L = lo[j];
                    if (L.lm < 0) continue;
                    if (!(L.mask & cm)) {
                        L.R += b * L.f0;
                    } else if (L.kind == TSG_KIND_POLY) {
                        const PolyParts P = poly_parts(L, xc);
                        L.R += b * (L.c0 + P.A + P.B);
                        L.R += poly_walk(L, P, sg, s0, s1v, s2v, any_c3);
                    } else if (L.kind == TSG_KIND_SEG) {
                        const int si = L.pc[0];
                        const double xa = (double) xc[si], sn = (double) tot[si];
                        L.R += b * (L.c0 * (L.pc[1] ? seg_pol(xa, sn) : seg_unpol(xa, sn)));
                        L.R += L.c0 * (L.pc[1] ? gpv[si] : guv[si]);
                    } else {
                        L.R += b * branch_summary(gout[L.lm], xc, tot, a.opar);
                    }
                }
                SWEEP_SYNC();
            }
        }
    };

    // site sweeps: each lane's output specs in registers (read once, not per
    // site), and the right end of its current window
    TsgOut ow[MPL];
    double wnext[MPL];
#pragma unroll
    for (int j = 0; j < MPL; j++) {
        wnext[j] = 0.0;
        if (MODE == 1 && lo[j].lm >= 0) {
            ow[j] = gout[lo[j].lm];
            if (lo[j].wi < lo[j].wend) wnext[j] = lo[j].win[lo[j].wi + 1];
        }
    }
    for (int t = t0; t < t1; t++) {
        if (t > t0) apply_transition(t);
        const int lt = tree(t);
        if (MODE == 0) {
            const double tl = tbp[lt], tr = tbp[lt + 1];
#pragma unroll
            for (int j = 0; j < MPL; j++) {
                LaneOut &L = lo[j];
                if (L.lm < 0) continue;
                while (L.wi < L.wend) {
                    const double wl = L.win[L.wi], wr = L.win[L.wi + 1];
                    if (!(wl < tr)) break;
                    L.A += L.R * (fmin(tr, wr) - fmax(tl, wl));
                    if (wr <= tr) {
                        flush_window(a, L, wfirst[j]);
                        L.wi++;
                    } else {
                        break;
                    }
                }
            }
        } else {
            const int s0 = ts0[lt], s1 = ts0[lt + 1];
            for (int s = s0; s < s1; s++) {
                const double pos = a.site_pos[s];
                const int na = a.site_nalleles[s];
                const int m0 = a.site_mut0[s], m1 = a.site_mut0[s + 1];
#pragma unroll
                for (int j = 0; j < MPL; j++) {
                    LaneOut &L = lo[j];
                    if (L.lm < 0) continue;
                    while (L.wi < L.wend && wnext[j] <= pos) {
                        flush_window(a, L, wfirst[j]);
                        L.wi++;
                        wnext[j] = (L.wi < L.wend) ? L.win[L.wi + 1] : 0.0;
                    }
                }
                double val[MPL];
#pragma unroll
                for (int j = 0; j < MPL; j++) val[j] = 0.0;
                // with at most 4 alleles (and one column per lane) each lane
                // forms its column's allele counts in one pass over the
                // site's mutations: w[a] = [a = 0] total + the counts below
                // the mutations to allele a - those below mutations from it
                const bool fast = na <= 4 && !kwide;
                Cnt w0 = 0, w1 = 0, w2 = 0, w3 = 0;
                if (fast && lane < K) {
                    w0 = tot[lane];
                    for (int mm = m0; mm < m1; mm++) {
                        const int node = a.mut_node[mm];
                        const int ma = a.mut_allele[mm], mp = a.mut_pallele[mm];
                        const Cnt xv = nd.get(node, lane);
                        // (added, then taken away, as the general form does)
                        w0 = (ma == 0) ? w0 + xv : w0;
                        w0 = (mp == 0) ? w0 - xv : w0;
                        w1 = (ma == 1) ? w1 + xv : w1;
                        w1 = (mp == 1) ? w1 - xv : w1;
                        w2 = (ma == 2) ? w2 + xv : w2;
                        w2 = (mp == 2) ? w2 - xv : w2;
                        w3 = (ma == 3) ? w3 + xv : w3;
                        w3 = (mp == 3) ? w3 - xv : w3;
                    }
                }
                // polarised outputs skip the ancestral allele (trees.c:1639)
                for (int al = any_unpolarised ? 0 : 1; al < na; al++) {
                    if (fast) {
                        if (lane < K) xo[lane] = al == 0 ? w0 : (al == 1 ? w1 : (al == 2 ? w2 : w3));
                    } else {
                        for (int k = lane; k < K; k += WAVE) {
                            Cnt w = (al == 0) ? tot[k] : Cnt(0);
                            for (int mm = m0; mm < m1; mm++) {
                                const int node = a.mut_node[mm];
                                if (a.mut_allele[mm] == al) w += nd.get(node, k);
                                if (a.mut_pallele[mm] == al) w -= nd.get(node, k);
                            }
                            xo[k] = w;
                        }
                    }
                    SWEEP_SYNC();
#pragma unroll
                    for (int j = 0; j < MPL; j++) {
                        if (lo[j].lm < 0) continue;
                        const TsgOut &o = ow[j];
                        if (al > 0 || !(o.flags & 1)) {
                            val[j] += eval_summary<Cnt, true>(o, xo, tot, false, a.opar);
                        }
                    }
                    SWEEP_SYNC();
                }
#pragma unroll
                for (int j = 0; j < MPL; j++) {
                    if (lo[j].lm >= 0) lo[j].A += val[j];
                }
            }
        }
    }
    if (MODE == 0) {
        if (t1 < a.num_trees) apply_transition(t1);
#pragma unroll
        for (int j = 0; j < MPL; j++) {
            if (lo[j].lm >= 0) {
                a.dR[(int64_t) c * a.M + a.m_begin + lo[j].lm] = lo[j].R;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < MPL; j++) {
        LaneOut &L = lo[j];
        if (L.lm < 0) continue;
        while (L.wi < L.wend) {
            flush_window(a, L, wfirst[j]);
            L.wi++;
        }
    }
#undef SWEEP_SYNC
}

// Waves per SIMD the one-output-per-lane sweep is compiled for (its VGPR
// budget: 3 waves allow 168)
#ifndef TSG_SWEEP_WAVES1
#define TSG_SWEEP_WAVES1 3
#endif

template <int MODE, int STORE, int MPL, int GC = 0>
__global__ void __launch_bounds__(WAVE, MPL == 1 ? TSG_SWEEP_WAVES1 : 1) k_sweep(const TsgSweepArgs a) {
    const int c = xcd_chunk(blockIdx.x, a.num_chunks);
    if (c >= a.num_chunks) return;
    sweep_chunk<MODE, STORE, MPL, GC>(a, c);
}

// Many sweeps in one launch (a batch per tree sequence): the chunks of run r
// are global chunks cpre[r] .. cpre[r+1]-1, each swept with that run's
// arguments.  All runs share MODE, STORE and MPL; the dynamic LDS is the
// largest run's.
template <int MODE, int STORE, int MPL>
__global__ void __launch_bounds__(WAVE, MPL == 1 ? TSG_SWEEP_WAVES1 : 1)
    k_sweep_multi(const TsgSweepArgs *__restrict__ args, const int32_t *__restrict__ cpre,
                  int nrun) {
    const int total = cpre[nrun];
    const int g = xcd_chunk(blockIdx.x, total);
    if (g >= total) return;
    int lo = 0, hi = nrun - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cpre[mid] <= g) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    const TsgSweepArgs &a = args[lo];
    sweep_chunk<MODE, STORE, MPL>(a, g - cpre[lo]);
}

// ---------------------------------------------------------------------------
// Allele frequency spectrum.  State columns 0..afs_K-1 are the sample sets,
// column afs_K counts all samples (trees.c:3882-3907).
namespace {

// fold (trees.c:3469-3496): coordinates are replaced by their complement
// when the allele is the minor one, with ties broken on trailing sets.
__device__ void afs_fold(int64_t *coord, const int64_t *dims, int nd) {
    double n = 0;
    int s = 0;
    for (int k = 0; k < nd; k++) {
        n += (double) dims[k] - 1;
        s += (int) coord[k];
    }
    n /= 2;
    int k = nd;
    while (s == n && k > 0) {
        k--;
        n -= ((double) (dims[k] - 1)) / 2;
        s -= (int) coord[k];
    }
    if (s > n) {
        for (k = 0; k < nd; k++) coord[k] = dims[k] - 1 - coord[k];
    }
}

#define AFS_MAX_SETS 16

__device__ int64_t afs_offset(const TsgSweepArgs &a, const int32_t *cnt, bool polarised) {
    int64_t coord[AFS_MAX_SETS];
    for (int k = 0; k < a.afs_K; k++) coord[k] = cnt[k];
    if (!polarised) afs_fold(coord, a.afs_dims, a.afs_K);
    int64_t off = 0;
    for (int k = 0; k < a.afs_K; k++) off = off * a.afs_dims[k] + coord[k];
    return off;
}

}  // namespace

// Site AFS: one histogram increment (1 or 1/2, held as half-unit integer
// counts so the result is exact) per non-fixed allele per site.
template <bool LDS>
__global__ void __launch_bounds__(WAVE) k_afs_site(const TsgSweepArgs a) {
    extern __shared__ int32_t smem[];
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int K = a.afs_K + 1;
    const int R = 3 + K;
    const int64_t n0 = a.chunk_node0[c];
    const int nloc = (int) (a.chunk_node0[c + 1] - n0);
    Nodes nd;
    nd.R = R;
    int32_t *stg;
    if (LDS) {
        nd.rec = smem;
        stg = smem + (int64_t) nloc * R;
    } else {
        nd.rec = a.g_rec + (int64_t) c * a.g_stride;
        stg = smem;
    }
    int32_t *xc = stg;
    int32_t *aw = stg + K;
    const bool polarised = (a.flags & TSG_F_POLARISED) != 0;
    init_chunk(a, nd, n0, nloc, K, lane);
    const double *win = a.windows;
    int wi = a.cq_wfirst[c];
    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    for (int t = t0; t < t1; t++) {
        if (t > t0) {
            const int64_t e0 = a.tr_off[t], e1 = a.tr_off[t + 1];
            const int64_t nrem = a.tr_nrem[t];
            for (int64_t e = e0; e < e1; e++) {
                const bool ins = (e - e0) >= nrem;
                const int u = a.ev_child[e], v = a.ev_parent[e];
                bool nz = false;
                for (int k = lane; k < K; k += WAVE) {
                    xc[k] = nd.x(u, k);
                    nz |= (xc[k] != 0);
                }
                nz = __any(nz);
                __syncthreads();
                if (lane == 0) nd.set_edge(u, ins ? v : TSG_NULL, 0.0f);
                __syncthreads();
                if (!nz) continue;
                for (int w = v; w >= 0;) {
                    const int pw = nd.parent(w);
                    for (int k = lane; k < K; k += WAVE) nd.x(w, k) += ins ? xc[k] : -xc[k];
                    w = pw;
                }
                __syncthreads();
            }
        }
        const int s0 = a.tree_site0[t], s1 = a.tree_site0[t + 1];
        for (int s = s0; s < s1; s++) {
            const double pos = a.site_pos[s];
            while (win[wi + 1] <= pos) wi++;
            const int na = a.site_nalleles[s];
            const int m0 = a.site_mut0[s], m1 = a.site_mut0[s + 1];
            for (int al = polarised ? 1 : 0; al < na; al++) {
                for (int k = lane; k < K; k += WAVE) {
                    int32_t w = (al == 0) ? a.col_total[k] : 0;
                    for (int mm = m0; mm < m1; mm++) {
                        const int node = a.mut_node[mm];
                        if (a.mut_allele[mm] == al) w += nd.x(node, k);
                        if (a.mut_pallele[mm] == al) w -= nd.x(node, k);
                    }
                    aw[k] = w;
                }
                __syncthreads();
                if (lane == 0) {
                    const int32_t all = aw[a.afs_K];
                    if (all > 0 && all < a.num_samples) {
                        const int64_t off = afs_offset(a, aw, polarised);
                        atomicAdd(&a.afs_hist_u[(int64_t) wi * a.afs_size + off],
                                  polarised ? 2ull : 1ull);
                    }
                }
                __syncthreads();
            }
        }
    }
}

// Branch AFS: each node's contribution (span since its last update times the
// time-window-clipped branch length) is added when its count changes, at
// window boundaries, and at the chunk end.
template <bool LDS>
__global__ void __launch_bounds__(WAVE) k_afs_branch(const TsgSweepArgs a) {
    extern __shared__ int32_t smem[];
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int K = a.afs_K + 1;
    const int R = 3 + K;
    const int64_t n0 = a.chunk_node0[c];
    const int nloc = (int) (a.chunk_node0[c + 1] - n0);
    Nodes nd;
    nd.R = R;
    int32_t *stg;
    double *last;
    if (LDS) {
        last = reinterpret_cast<double *>(smem);
        nd.rec = smem + 2 * (int64_t) nloc;
        stg = nd.rec + (int64_t) nloc * R;
    } else {
        nd.rec = a.g_rec + (int64_t) c * a.g_stride;
        last = a.g_last + (int64_t) c * a.g_last_stride;
        stg = smem;
    }
    int32_t *xc = stg;
    const bool polarised = (a.flags & TSG_F_POLARISED) != 0;
    const double *ltime = a.loc_time + n0;
    init_chunk(a, nd, n0, nloc, K, lane);
    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    const double cstart = a.bp[t0], cend = a.bp[t1];
    const double *win = a.windows;
    int wi = a.cq_wfirst[c];
    // The reference keeps last_update of a parentless node at its last
    // removal or the last window boundary, whichever is later; a node with a
    // parent at the chunk start has had its earlier span flushed by the
    // previous chunk.
    for (int u = lane; u < nloc; u += WAVE) {
        last[u] = (a.loc_parent0[n0 + u] >= 0) ? cstart : fmax(a.loc_lastrem[n0 + u], win[wi]);
    }
    __syncthreads();

    // contribution of node u over [last[u], pos) into window wi
    auto flush = [&](int u, double pos) {
        const int p = nd.parent(u);
        if (p >= 0) {
            int32_t cnt[AFS_MAX_SETS + 1];
            for (int k = 0; k < K; k++) cnt[k] = nd.x(u, k);
            const int32_t all = cnt[a.afs_K];
            if (all > 0 && all < a.num_samples) {
                const double tu = ltime[u], tv = ltime[p];
                const int64_t off = afs_offset(a, cnt, polarised);
                const float span = (float) (pos - last[u]);
                for (int tw = 0; tw < a.TW && a.time_windows[tw] < tv; tw++) {
                    const double hi = fmin(a.time_windows[tw + 1], tv);
                    const double lo = fmax(a.time_windows[tw], tu);
                    const float twbl = (float) fmax(0.0, hi - lo);
                    const float x = __fmul_rn(span, twbl);
                    atomicAdd(&a.afs_hist_d[((int64_t) wi * a.TW + tw) * a.afs_size + off],
                              (double) x);
                }
            }
        }
        last[u] = pos;
    };

    for (int t = t0; t < t1; t++) {
        if (t > t0) {
            const double x = a.bp[t];
            const int64_t e0 = a.tr_off[t], e1 = a.tr_off[t + 1];
            const int64_t nrem = a.tr_nrem[t];
            for (int64_t e = e0; e < e1; e++) {
                const bool ins = (e - e0) >= nrem;
                const int u = a.ev_child[e], v = a.ev_parent[e];
                for (int k = lane; k < K; k += WAVE) xc[k] = nd.x(u, k);
                __syncthreads();
                if (lane == 0) {
                    if (!ins) flush(u, x);
                    // walk: flush each ancestor before its counts change
                    for (int w = v; w >= 0;) {
                        flush(w, x);
                        w = nd.parent(w);
                    }
                }
                __syncthreads();
                for (int w = v; w >= 0;) {
                    const int pw = nd.parent(w);
                    for (int k = lane; k < K; k += WAVE) nd.x(w, k) += ins ? xc[k] : -xc[k];
                    w = pw;
                }
                __syncthreads();
                if (lane == 0) nd.set_edge(u, ins ? v : TSG_NULL, 0.0f);
                __syncthreads();
            }
        }
        const double tr = a.bp[t + 1];
        // boundaries strictly inside the chunk; one at the chunk end is
        // handled by the final flush
        while (win[wi + 1] <= tr && win[wi + 1] < cend) {
            const double wr = win[wi + 1];
            for (int u = lane; u < nloc; u += WAVE) flush(u, wr);
            __syncthreads();
            wi++;
        }
    }
    // flush everything up to the chunk end into the current window
    for (int u = lane; u < nloc; u += WAVE) flush(u, cend);
}

// ---------------------------------------------------------------------------
// Allele frequency spectrum on the hot plan: a chunk holds only the nodes
// its events and mutations touch (slots), with counts for the afs_K sample
// sets and, in column afs_K, all samples (seeded from slot_nsd).
//
// Branch mode follows the reference's flush rule (trees.c:3650-3690,
// 3738-3770): a node's contribution since its last update is added when it
// is updated -- as the child of a removed edge, or as a node on the path
// above an inserted or removed edge -- and at every window boundary.  Here a
// node is flushed only when an event updates it; the interval since its
// previous update is split at the window boundaries it crosses, and each
// piece is valued with the node's state at the piece's end, as the
// reference's boundary flush would see it: the counts are those of the whole
// interval (only updates change them), and the node has a parent at a
// boundary q when it has one now and its edge was inserted before q (an
// inserted edge's child is not updated at insertion, so the piece that ends
// at the next update includes the span since the previous one).  Intervals
// still open at the end are flushed from the last tree (k_afs_tl_*).
namespace {

// window holding position x: windows[w] <= x < windows[w + 1], 0 <= w < W
__device__ int afs_window_of(const double *win, int W, double x) {
    int lo = 0, hi = W - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (win[mid] <= x) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

// The branch AFS contributions of one node with counts cnt, time tu and
// parent time tp (-inf: no parent; the edge inserted at y) over [a, x).
__device__ void afs_flush_pieces(const TsgSweepArgs &a, const int32_t *cnt, double tu, double tp,
                                 double y, double from, double x) {
    if (!(tp > -INFINITY)) return;
    const int32_t all = cnt[a.afs_K];
    if (all <= 0 || all >= a.num_samples) return;
    const int64_t off = afs_offset(a, cnt, (a.flags & TSG_F_POLARISED) != 0);
    const double *win = a.windows;
    const int W = a.afs_W;
    auto add = [&](int wi, double span) {
        for (int tw = 0; tw < a.TW && a.time_windows[tw] < tp; tw++) {
            const double twbl = fmax(0.0, fmin(a.time_windows[tw + 1], tp) -
                                              fmax(a.time_windows[tw], tu));
            atomicAdd(&a.afs_hist_d[((int64_t) wi * a.TW + tw) * a.afs_size + off], span * twbl);
        }
    };
    int wi = afs_window_of(win, W, from);
    double p = from;
    while (wi < W && win[wi + 1] <= x) {
        const double q = win[wi + 1];
        if (y < q) add(wi, q - p);
        p = q;
        wi++;
    }
    if (wi < W && x > p) add(wi, x - p);
}

}  // namespace

// One wave per chunk.  LDS: (branch) per slot tu, tp, last, y as doubles,
// then the counts [S x (afs_K + 1)] and a staging row.
template <bool BRANCH>
__global__ void __launch_bounds__(WAVE) k_afs_hot(const TsgSweepArgs a) {
    extern __shared__ double dsm[];
    const int c = xcd_chunk(blockIdx.x, a.num_chunks);
    if (c >= a.num_chunks) return;
    const int lane = threadIdx.x;
    const int K1 = a.afs_K + 1;
    const int64_t s0 = a.chunk_node0[c];
    const int S = (int) (a.chunk_node0[c + 1] - s0);
    double *tu = dsm, *tp = dsm + S, *last = dsm + 2 * S, *yins = dsm + 3 * S;
    int32_t *X = reinterpret_cast<int32_t *>(BRANCH ? dsm + 4 * S : dsm);
    int32_t *aw = X + (int64_t) S * K1;
    const bool polarised = (a.flags & TSG_F_POLARISED) != 0;

    for (int i = lane; i < S * K1; i += WAVE) X[i] = 0;
    __syncthreads();
    for (int h = lane; h < S; h += WAVE) {
        X[h * K1 + a.afs_K] = a.slot_nsd[s0 + h];
        if (BRANCH) {
            tu[h] = a.slot_tu[s0 + h];
            tp[h] = a.slot_tp0[s0 + h];
            last[h] = a.slot_last0[s0 + h];
            yins[h] = a.slot_ins0[s0 + h];
        }
    }
    const uint16_t *ss = sample_map(a, c);
    for (int i = lane; i < a.npairs; i += WAVE) {
        const unsigned h = ss[(int64_t) a.pair_sample[i] * TSG_SS_CB];
        if (h != TSG_NO_SLOT) atomicAdd(&X[h * K1 + a.pair_col[i]], 1);
    }
    __syncthreads();
    const int32_t *lev = a.lev_slot0 + a.chunk_lev0[c];
    const int nlev = (int) (a.chunk_lev0[c + 1] - a.chunk_lev0[c]) - 1;
    for (int l = 0; l + 1 < nlev; l++) {
        const int h0 = lev[l], h1 = lev[l + 1];
        for (int idx = lane; idx < (h1 - h0) * K1; idx += WAVE) {
            const int h = h0 + idx / K1, k = idx % K1;
            const int p = a.slot_par[s0 + h];
            const int32_t v = X[h * K1 + k];
            if (p >= 0 && v != 0) atomicAdd(&X[p * K1 + k], v);
        }
        __syncthreads();
    }

    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    // branch mode also applies the closing transition (its updates happen
    // at the chunk end, before the next chunk's start state)
    const int tend = (BRANCH && t1 < a.num_trees) ? t1 + 1 : t1;
    int wi = BRANCH ? 0 : afs_window_of(a.windows, a.afs_W, a.bp[t0]);
    for (int t = t0; t < tend; t++) {
        if (t > t0) {
            const double x = a.bp[t];
            const int64_t e0 = a.tr_off[t], e1 = a.tr_off[t + 1];
            const int64_t nrem = a.tr_nrem[t];
            for (int64_t e = e0; e < e1; e++) {
                const bool ins = (e - e0) >= nrem;
                const int u = a.ev_child[e];
                const int64_t p0 = a.ev_path0[e];
                const int plen = (int) (a.ev_path0[e + 1] - p0);
                if (BRANCH) {
                    // the path nodes and, for a removal, the child: distinct
                    // nodes, one lane each
                    const int items = plen + (ins ? 0 : 1);
                    for (int j = lane; j < items; j += WAVE) {
                        const int w = (j < plen) ? a.path[p0 + j] : u;
                        afs_flush_pieces(a, X + w * K1, tu[w], tp[w], yins[w], last[w], x);
                        last[w] = x;
                    }
                    __syncthreads();
                }
                for (int idx = lane; idx < plen * K1; idx += WAVE) {
                    const int j = idx / K1, k = idx % K1;
                    const int w = a.path[p0 + j];
                    const int32_t v = X[u * K1 + k];
                    X[w * K1 + k] += ins ? v : -v;
                }
                if (BRANCH && lane == 0) {
                    if (ins) {
                        tp[u] = tu[a.path[p0]];
                        yins[u] = x;
                    } else {
                        tp[u] = -INFINITY;
                    }
                }
                __syncthreads();
            }
        }
        if (!BRANCH && t < t1) {
            const int s0s = a.tree_site0[t], s1s = a.tree_site0[t + 1];
            for (int s = s0s; s < s1s; s++) {
                const double pos = a.site_pos[s];
                while (a.windows[wi + 1] <= pos) wi++;
                const int na = a.site_nalleles[s];
                const int m0 = a.site_mut0[s], m1 = a.site_mut0[s + 1];
                for (int al = polarised ? 1 : 0; al < na; al++) {
                    for (int k = lane; k < K1; k += WAVE) {
                        int32_t w = (al == 0) ? a.col_total[k] : 0;
                        for (int mm = m0; mm < m1; mm++) {
                            const int h = a.mut_node[mm];
                            if (a.mut_allele[mm] == al) w += X[h * K1 + k];
                            if (a.mut_pallele[mm] == al) w -= X[h * K1 + k];
                        }
                        aw[k] = w;
                    }
                    __syncthreads();
                    if (lane == 0) {
                        const int32_t all = aw[a.afs_K];
                        if (all > 0 && all < a.num_samples) {
                            const int64_t off = afs_offset(a, aw, polarised);
                            atomicAdd(&a.afs_hist_u[(int64_t) wi * a.afs_size + off],
                                      polarised ? 2ull : 1ull);
                        }
                    }
                    __syncthreads();
                }
            }
        }
    }
}

// Counts of the last tree: thread i < npairs adds its (sample, set) pair,
// thread npairs + s adds sample s to the all-samples column, each walking
// its root path.
__global__ void k_afs_tl_counts(const TsgSweepArgs a, const TsgLastTree lt) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int K1 = a.afs_K + 1;
    if (i >= (int64_t) a.npairs + a.num_samples) return;
    int u, k;
    if (i < a.npairs) {
        u = a.pair_sample[i];
        k = a.pair_col[i];
    } else {
        u = (int) (i - a.npairs);
        k = a.afs_K;
    }
    for (; u >= 0; u = lt.parent[u]) atomicAdd(&lt.counts[(int64_t) u * K1 + k], 1);
}

// The reference's flush at the last window boundary (the sequence end) of
// every node with a parent in the last tree, one thread per node.
__global__ void k_afs_tl_flush(const TsgSweepArgs a, const TsgLastTree lt, double L) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= lt.nl) return;
    const int p = lt.parent[q];
    if (p < 0) return;
    afs_flush_pieces(a, lt.counts + (int64_t) q * (a.afs_K + 1), lt.time[q], lt.time[p], lt.ins[q],
                     lt.last[q], L);
}

// ---------------------------------------------------------------------------
// Node statistics (trees.c:1788-1918): for every node u and output m, the
// summary f_m(x(u)) of the node's own state, integrated along the genome in
// each window.  A node's state changes only where an event's root path
// passes through it, so f is constant between its updates; the value is
// added at each update over the span since the previous one, cut at the
// window boundaries that span crosses.  The reference adds the same value
// at the same updates and at every window boundary: the same pieces, and a
// zero-length update adds f * 0 in both (NaN when f is not finite).  The
// flush points are the branch AFS's (the path nodes, and a removed edge's
// child, whose state does not change there: an extra cut of a constant
// piece), so the planner's last-flush positions per slot apply; intervals
// still open at the end are added from the last tree.  Results go to
// res[(window * N + node) * M + m] by FP64 atomics (the order of the adds,
// and so the last bits, varies between runs).
namespace {

__device__ void node_flush(const TsgSweepArgs &a, double *res, int64_t N, int64_t node, int m,
                           double f, double from, double x) {
    const double *win = a.windows;
    const int W = a.afs_W;
    int wi = afs_window_of(win, W, from);
    double p = from;
    while (wi < W && win[wi + 1] <= x) {
        const double q = win[wi + 1];
        atomicAdd(&res[((int64_t) wi * N + node) * a.M + m], f * (q - p));
        p = q;
        wi++;
    }
    if (wi < W) atomicAdd(&res[((int64_t) wi * N + node) * a.M + m], f * (x - p));
}

}  // namespace

// One wave per chunk.  LDS: per slot its last flush position (double), then
// the counts [S x K].
__global__ void __launch_bounds__(WAVE) k_node_hot(const TsgSweepArgs a,
                                                   const int32_t *__restrict__ slot_node,
                                                   double *res, int64_t N) {
    extern __shared__ double dsm[];
    const int c = xcd_chunk(blockIdx.x, a.num_chunks);
    if (c >= a.num_chunks) return;
    const int lane = threadIdx.x;
    const int K = a.K, M = a.M;
    const int64_t s0 = a.chunk_node0[c];
    const int S = (int) (a.chunk_node0[c + 1] - s0);
    double *last = dsm;
    int32_t *X = reinterpret_cast<int32_t *>(dsm + S);
    for (int i = lane; i < S * K; i += WAVE) X[i] = 0;
    for (int h = lane; h < S; h += WAVE) last[h] = a.slot_last0[s0 + h];
    __syncthreads();
    const uint16_t *ss = sample_map(a, c);
    for (int i = lane; i < a.npairs; i += WAVE) {
        const unsigned h = ss[(int64_t) a.pair_sample[i] * TSG_SS_CB];
        if (h != TSG_NO_SLOT) atomicAdd(&X[h * K + a.pair_col[i]], 1);
    }
    __syncthreads();
    const int32_t *lev = a.lev_slot0 + a.chunk_lev0[c];
    const int nlev = (int) (a.chunk_lev0[c + 1] - a.chunk_lev0[c]) - 1;
    for (int l = 0; l + 1 < nlev; l++) {
        const int h0 = lev[l], h1 = lev[l + 1];
        for (int idx = lane; idx < (h1 - h0) * K; idx += WAVE) {
            const int h = h0 + idx / K, k = idx % K;
            const int p = a.slot_par[s0 + h];
            const int32_t v = X[h * K + k];
            if (p >= 0 && v != 0) atomicAdd(&X[p * K + k], v);
        }
        __syncthreads();
    }
    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    // the closing transition too: its updates happen at the chunk end
    const int tend = (t1 < a.num_trees) ? t1 + 1 : t1;
    for (int t = t0 + 1; t < tend; t++) {
        const double x = a.bp[t];
        const int64_t e0 = a.tr_off[t], e1 = a.tr_off[t + 1];
        const int64_t nrem = a.tr_nrem[t];
        for (int64_t e = e0; e < e1; e++) {
            const bool ins = (e - e0) >= nrem;
            const int u = a.ev_child[e];
            const int64_t p0 = a.ev_path0[e];
            const int plen = (int) (a.ev_path0[e + 1] - p0);
            const int nn = plen + (ins ? 0 : 1);
            for (int idx = lane; idx < nn * M; idx += WAVE) {
                const int j = idx / M, m = idx % M;
                const int w = (j < plen) ? a.path[p0 + j] : u;
                const double f = branch_summary(a.outs[m], X + w * K, a.col_total);
                node_flush(a, res, N, slot_node[s0 + w], m, f, last[w], x);
            }
            __syncthreads();
            for (int j = lane; j < nn; j += WAVE) last[(j < plen) ? a.path[p0 + j] : u] = x;
            for (int idx = lane; idx < plen * K; idx += WAVE) {
                const int j = idx / K, k = idx % K;
                const int w = a.path[p0 + j];
                const int32_t v = X[u * K + k];
                X[w * K + k] += ins ? v : -v;
            }
            __syncthreads();
        }
    }
}

// Counts of the last tree over the K state columns: thread i adds its
// (sample, column) pair along the sample's root path.
__global__ void k_node_tl_counts(const TsgSweepArgs a, const TsgLastTree lt) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t) a.npairs) return;
    const int k = a.pair_col[i];
    for (int u = a.pair_sample[i]; u >= 0; u = lt.parent[u]) {
        atomicAdd(&lt.counts[(int64_t) u * a.K + k], 1);
    }
}

// The final piece [last flush, L) of every node u and output m, one thread
// each; a node not in the last tree has state 0 there.
__global__ void k_node_tl_flush(const TsgSweepArgs a, const TsgLastTree lt,
                                const double *__restrict__ node_last,
                                const int32_t *__restrict__ node_lid, double *res, int64_t N,
                                const int32_t *zeros, double L) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * a.M) return;
    const int64_t u = i / a.M;
    const int m = (int) (i % a.M);
    const int32_t q = node_lid[u];
    const int32_t *x = (q >= 0) ? lt.counts + (int64_t) q * a.K : zeros;
    const double f = branch_summary(a.outs[m], x, a.col_total);
    node_flush(a, res, N, u, m, f, node_last[u], L);
}

// ---------------------------------------------------------------------------
// Running branch sum of the first tree (chunk 0's start tree), evaluated in
// parallel instead of by replaying its edge insertions one at a time:
//   R(tree 0) = sum over nodes u with a parent of bl(u) * F(x(u)),
// which is what the reference's insertion loop accumulates (trees.c:1426-
// 1480, update_running_sum 1339-1350), in FP64, per block of nodes, in a
// fixed order (deterministic).  The tree is given samples first (local id
// j = sample j): t0_parent, t0_bl.
//
// Counts: one thread per weighted sample walks its root path in the
// snapshot, adding 1 to each of its columns.
__device__ __forceinline__ void tree0_counts(const TsgSweepArgs &a, const int32_t *t0_parent,
                                             int32_t *x0, int bx) {
    const int i = bx * 256 + threadIdx.x;
    if (i >= a.nseed) return;
    const int c0 = a.seed_off[i], c1 = a.seed_off[i + 1];
    for (int u = a.seed_sample[i]; u >= 0; u = t0_parent[u]) {
        for (int c = c0; c < c1; c++) atomicAdd(&x0[(int64_t) u * a.K + a.seed_col[c]], 1);
    }
}

__global__ void __launch_bounds__(256) k_tree0_counts(const TsgSweepArgs a,
                                                      const int32_t *t0_parent, int32_t *x0) {
    tree0_counts(a, t0_parent, x0, blockIdx.x);
}

#define TSG_T0_NODES 64
// column counts up to which k_tree0_sum stages its nodes in LDS
#define TSG_T0_STAGE_K 64

// grid (ceil(nloc / TSG_T0_NODES), ceil(M / 64)); lane = output.
// (T = double: the first tree's sums of real-valued weights, x0 computed on
// the host)
template <typename T>
__device__ __forceinline__ void tree0_sum(const TsgSweepArgs &a, const int32_t *t0_parent,
                                          const double *t0_bl, const T *x0, int nloc,
                                          double *part, int bx, int by, bool stage) {
    const int m = by * WAVE + threadIdx.x;
    const int u0 = bx * TSG_T0_NODES;
    const int u1 = min(nloc, u0 + TSG_T0_NODES);
    // stage (the launch gives TSG_T0_NODES x (12 + K sizeof(T)) bytes of
    // dynamic LDS) and K <= 64: the block's nodes (branch length, counts,
    // whether they have a parent) are first read into LDS together,
    // coalesced; the sum then reads them from there, in the same node order
    const bool staged = stage && a.K <= TSG_T0_STAGE_K;
    extern __shared__ int32_t smem[];
    double *sbl = reinterpret_cast<double *>(smem);                     // [T0_NODES]
    T *sx = reinterpret_cast<T *>(sbl + TSG_T0_NODES);                   // [T0_NODES x K]
    int32_t *sroot = reinterpret_cast<int32_t *>(sx + TSG_T0_NODES * a.K);  // [T0_NODES]
    if (staged) {
        const int n = u1 - u0;
        for (int i = threadIdx.x; i < n * a.K; i += WAVE) sx[i] = x0[(int64_t) u0 * a.K + i];
        if ((int) threadIdx.x < n) {
            sbl[threadIdx.x] = t0_bl[u0 + threadIdx.x];
            sroot[threadIdx.x] = t0_parent[u0 + threadIdx.x] < 0;
        }
        __syncthreads();
    }
    if (m >= a.M) return;
    const TsgOut o = a.outs[m];
    LaneOut L;
    const bool mom = moment_form<std::is_same<T, double>::value>(o, a.K, L, a.col_total_w,
                                                                  a.opar) == TSG_KIND_POLY;
    const T *tot;
    if constexpr (std::is_same<T, double>::value) {
        tot = a.col_total_w;
    } else {
        tot = a.col_total;
    }
    double acc = 0;
    for (int u = u0; u < u1; u++) {
        const int j = u - u0;
        if (staged ? sroot[j] != 0 : t0_parent[u] < 0) continue;
        const T *x = staged ? sx + (int64_t) j * a.K : x0 + (int64_t) u * a.K;
        double f;
        if (mom) {  // the polynomial form, as in the sweep
            f = poly_value(L, x);
        } else {
            f = branch_summary(o, x, tot, a.opar);
        }
        acc += (staged ? sbl[j] : t0_bl[u]) * f;
    }
    part[(int64_t) bx * a.M + m] = acc;
}

template <typename T>
__global__ void __launch_bounds__(WAVE) k_tree0_sum(const TsgSweepArgs a, const int32_t *t0_parent,
                                                    const double *t0_bl, const T *x0,
                                                    int nloc, double *part) {
    tree0_sum(a, t0_parent, t0_bl, x0, nloc, part, blockIdx.x, blockIdx.y, true);
}

// ---------------------------------------------------------------------------
// Combine: per (query, window, output), sum the chunk partials of the chunks
// that overlap the window, re-base branch sums by the chunk start value, and
// span-normalise.  FP64, deterministic chunk order.
//
// Rstart[c] = R(first tree) + sum of the chunk deltas before c, as a
// two-level FP64 scan in a fixed order (deterministic): chunk blocks of
// TSG_SCAN_B sum their deltas, one pass per output turns the block sums into
// exclusive prefixes (starting from the first tree's nb per-block partials,
// k_tree0_sum), then every block writes its chunks' prefixes.  Grid of the
// block kernels: (blocks, ceil(M / 64)), lane = output.
#define TSG_SCAN_B 128

__device__ __forceinline__ void scan_blocks(const double *dR, double *bsum, int C, int M, int bx,
                                            int by) {
    const int m = by * 64 + threadIdx.x;
    if (m >= M) return;
    const int c0 = bx * TSG_SCAN_B, c1 = min(C, c0 + TSG_SCAN_B);
    double acc = 0;
    for (int c = c0; c < c1; c++) acc += dR[(int64_t) c * M + m];
    bsum[(int64_t) bx * M + m] = acc;
}

__global__ void __launch_bounds__(64) k_scan_blocks(const double *dR, double *bsum, int C, int M) {
    scan_blocks(dR, bsum, C, M, blockIdx.x, blockIdx.y);
}

// one workgroup of TSG_SCAN_TOP threads per output: the first tree's
// partials are summed by a fixed-shape tree (deterministic), then one lane
// scans the block sums
#define TSG_SCAN_TOP 256
__device__ __forceinline__ void scan_top(double *bsum, int nblk, int M, const double *r0part,
                                         int nb, int m, double *red) {
    const int t = threadIdx.x;
    double acc = 0;
    for (int b = t; b < nb; b += TSG_SCAN_TOP) acc += r0part[(int64_t) b * M + m];
    red[t] = acc;
    __syncthreads();
    for (int w = TSG_SCAN_TOP / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t != 0) return;
    acc = red[0];
    for (int b = 0; b < nblk; b++) {
        const double x = bsum[(int64_t) b * M + m];
        bsum[(int64_t) b * M + m] = acc;
        acc += x;
    }
}

__global__ void __launch_bounds__(TSG_SCAN_TOP)
k_scan_top(double *bsum, int nblk, int M, const double *r0part, int nb) {
    __shared__ double red[TSG_SCAN_TOP];
    scan_top(bsum, nblk, M, r0part, nb, blockIdx.x, red);
}

__device__ __forceinline__ void scan_apply(const double *dR, const double *bsum, double *Rstart,
                                           int C, int M, int bx, int by) {
    const int m = by * 64 + threadIdx.x;
    if (m >= M) return;
    const int c0 = bx * TSG_SCAN_B, c1 = min(C, c0 + TSG_SCAN_B);
    double acc = bsum[(int64_t) bx * M + m];
    for (int c = c0; c < c1; c++) {
        Rstart[(int64_t) c * M + m] = acc;
        acc += dR[(int64_t) c * M + m];
    }
}

__global__ void __launch_bounds__(64) k_scan_apply(const double *dR, const double *bsum,
                                                   double *Rstart, int C, int M) {
    scan_apply(dR, bsum, Rstart, C, M, blockIdx.x, blockIdx.y);
}

// One group of Qr.lanes lanes per (query, window, output), aligned within
// the wave: lane j of the group sums the window's chunks j, j + lanes, ...,
// then the group adds its lanes by a fixed butterfly (deterministic).  A
// window spanning many chunks (few windows) gets up to a wave.
__device__ __forceinline__ void combine(const TsgCombineArgs &a, int64_t total, int64_t bx) {
    const int64_t gid = bx * 256 + threadIdx.x;
    if (gid >= total) return;  // total is a multiple of 64: whole waves leave
    int qlo = 0, qhi = a.Q - 1;
    while (qlo < qhi) {
        const int mid = (qlo + qhi + 1) / 2;
        if (a.lanebase[mid] <= gid) {
            qlo = mid;
        } else {
            qhi = mid - 1;
        }
    }
    const int q = qlo;
    const TsgQuery &Qr = a.queries[q];
    const int G = Qr.lanes;
    const int64_t rem = gid - a.lanebase[q];
    const int64_t item = rem / G;
    const int sub = (int) (rem % G);
    const bool valid = item < (int64_t) Qr.W * Qr.M;
    double acc = 0;
    int w = 0, ml = 0;
    double wl = 0, wr = 1;
    if (valid) {
        w = (int) (item / Qr.M);
        ml = (int) (item % Qr.M);
        const double *win = a.windows + Qr.win_off;
        wl = win[w];
        wr = win[w + 1];
        // chunk c spans [bp[chunk_tree0[c]], bp[chunk_tree0[c+1]])
        int lo = 0, hi = a.num_chunks - 1;
        while (lo < hi) {  // first chunk whose end is > wl
            const int mid = (lo + hi) / 2;
            if (a.bp[a.chunk_tree0[mid + 1]] > wl) {
                hi = mid;
            } else {
                lo = mid + 1;
            }
        }
        for (int c = lo + sub; c < a.num_chunks; c += G) {
            const double cs = a.bp[a.chunk_tree0[c]], ce = a.bp[a.chunk_tree0[c + 1]];
            if (!(cs < wr)) break;
            const int64_t cq = (int64_t) c * a.Q + q;
            const int slot = w - a.cq_wfirst[cq];
            if (slot >= 0 && slot < a.cq_nwin[cq]) {
                acc += a.partial[Qr.pbase + (a.cq_poff[cq] + slot) * (int64_t) Qr.M + ml];
            }
            if (a.Rstart != nullptr) {
                const double ov = fmin(ce, wr) - fmax(cs, wl);
                if (ov > 0) acc += a.Rstart[(int64_t) c * a.M + Qr.mbase + ml] * ov;
            }
        }
    }
    // every lane of the wave takes part (groups are aligned, G divides 64)
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(acc, off);
        if (off < G) acc += o;
    }
    if (valid && sub == 0) {
        if (Qr.flags & TSG_F_SPAN_NORMALISE) acc /= (wr - wl);
        a.raw[a.rawbase[q] + (int64_t) w * Qr.M + ml] = acc;
    }
}

__global__ void __launch_bounds__(256) k_combine(const TsgCombineArgs a, int64_t total) {
    combine(a, total, blockIdx.x);
}

// FP64 composites of the reference's Python layer.
struct TsgFinal {
    int32_t kind;       // tsg_final_id
    int32_t Kq;         // number of sample sets (TajD: sets; Fst: sets)
    int32_t P;          // number of final outputs per window
    int32_t W;
    int64_t rawbase, outbase;
    int64_t pbase;      // params: Fst -> index pairs (int32 pairs); TajD -> per-set (n,h,g)
    int32_t Mraw;
};

__device__ __forceinline__ void finals(const TsgFinal *fin, int nq, const double *raw,
                                       const int32_t *ipar, const double *dpar, double *out,
                                       int64_t total, int64_t bx) {
    const int64_t gid = bx * 256 + threadIdx.x;
    if (gid >= total) return;
    int q = 0;
    int64_t rem = gid;
    while (q < nq) {
        const int64_t n = (int64_t) fin[q].W * fin[q].P;
        if (rem < n) break;
        rem -= n;
        q++;
    }
    const TsgFinal &F = fin[q];
    const int w = (int) (rem / F.P);
    const int p = (int) (rem % F.P);
    const double *r = raw + F.rawbase + (int64_t) w * F.Mraw;
    double v;
    switch (F.kind) {
    case TSG_FINAL_FST: {  // trees.py:10113-10125
        const int u = ipar[F.pbase + 2 * p], x = ipar[F.pbase + 2 * p + 1];
        const double du = r[u], dv = r[x], duv = r[F.Kq + p];
        const double denom = du + dv + 2 * duv;
        v = 1.0 - 2 * (du + dv) / denom;
        break;
    }
    case TSG_FINAL_TAJIMAS_D: {  // trees.py:10008-10018
        const double n = dpar[F.pbase + 3 * p], h = dpar[F.pbase + 3 * p + 1];
        const double g = dpar[F.pbase + 3 * p + 2];
        const double T = r[p], S = r[F.Kq + p];
        const double aa = (n + 1) / (3 * (n - 1) * h) - 1 / (h * h);
        const double bb = 2 * (n * n + n + 3) / (9 * n * (n - 1)) - (n + 2) / (h * n) + g / (h * h);
        v = (T - S / h) / sqrt(aa * S + (bb / (h * h + g)) * S * (S - 1));
        break;
    }
    case TSG_FINAL_PROPORTION:  // trees.py:8958-8974
        v = r[p] / r[F.P];
        break;
    default:
        v = r[p];
    }
    out[F.outbase + (int64_t) w * F.P + p] = v;
}

__global__ void __launch_bounds__(256) k_finals(const TsgFinal *fin, int nq, const double *raw,
                                                const int32_t *ipar, const double *dpar,
                                                double *out, int64_t total) {
    finals(fin, nq, raw, ipar, dpar, out, total, blockIdx.x);
}

// Site AFS finalise: half-unit counts to doubles (exact), span normalise.
__global__ void k_afs_site_final(const unsigned long long *h, double *out, const double *win,
                                 int W, int64_t afs_size, uint32_t flags) {
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t) W * afs_size) return;
    const int w = (int) (gid / afs_size);
    double v = (double) h[gid] * 0.5;
    if (flags & TSG_F_SPAN_NORMALISE) v /= (win[w + 1] - win[w]);
    out[gid] = v;
}

__global__ void k_afs_branch_final(double *h, const double *win, int W, int64_t per_window,
                                   uint32_t flags) {
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t) W * per_window) return;
    const int w = (int) (gid / per_window);
    if (flags & TSG_F_SPAN_NORMALISE) h[gid] /= (win[w + 1] - win[w]);
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_fill_f64(double *p, int64_t n) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_fill_i32(int32_t *p, int64_t n) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Host launchers (C++ linkage, used by tsg_api.cpp).
#include "tsg_launch.h"

namespace {
// Dynamic LDS a workgroup may be given (the device's per-workgroup limit).
// Kernels that take more than 64 KiB get this limit as their attribute --
// one value for every launch, so host threads launching concurrently never
// undo each other's setting.
int g_lds_max = 65536;

template <int MODE, int STORE, int MPL, int GC>
hipError_t launch_gc(const TsgSweepArgs &a, size_t shmem, hipStream_t s) {
    const void *fn = reinterpret_cast<const void *>(&k_sweep<MODE, STORE, MPL, GC>);
    if (shmem > 65536) {
        if ((int) shmem > g_lds_max) return hipErrorInvalidValue;
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           g_lds_max);
        if (e != hipSuccess) return e;
    }
    const unsigned grid = 8u * (unsigned) ((a.num_chunks + 7) / 8);
    hipLaunchKernelGGL((k_sweep<MODE, STORE, MPL, GC>), dim3(grid), dim3(WAVE), shmem, s, a);
    return hipGetLastError();
}

// Sweeps of one output per lane on the hot plan run a build with the
// count-update layout's G (hop groups per column: 16 for up to 4 columns, 8
// for 5-8, 4 for 9-16, 2 for 17-32, 1 for 33-64) fixed at compile time: the hop chains'
// steps and the lane arithmetic need no run-time tests (measured on the
// scale workload, K = 8: branch sweep 3.33 -> 2.81 ms).  Branch sweeps of up
// to 4 columns take G = 8 as well (any G with G K <= 64 is a valid layout,
// and the sums do not depend on it): with G = 16 the 15 unrolled chain
// steps spill registers.
template <int MODE, int STORE, int MPL>
hipError_t launch_one(const TsgSweepArgs &a, size_t shmem, hipStream_t s) {
    if constexpr (MPL == 1 && (STORE == TSG_STORE_HOT16 || STORE == TSG_STORE_HOT ||
                               STORE == TSG_STORE_HOTW)) {
        if constexpr (MODE == 1) {
            if (a.K >= 1 && a.K <= 4) return launch_gc<MODE, STORE, MPL, 16>(a, shmem, s);
        }
        if (a.K >= 1 && a.K <= 8) return launch_gc<MODE, STORE, MPL, 8>(a, shmem, s);
        if (a.K >= 9 && a.K <= 16) return launch_gc<MODE, STORE, MPL, 4>(a, shmem, s);
        if (a.K >= 17 && a.K <= 32) return launch_gc<MODE, STORE, MPL, 2>(a, shmem, s);
        if (a.K >= 33 && a.K <= 64) return launch_gc<MODE, STORE, MPL, 1>(a, shmem, s);
    }
    return launch_gc<MODE, STORE, MPL, 0>(a, shmem, s);
}

template <int MODE, int STORE>
hipError_t launch_sweep_mpl(const TsgSweepArgs &a, int mpl, size_t shmem, hipStream_t s) {
    switch (mpl) {
    case 1:
        return launch_one<MODE, STORE, 1>(a, shmem, s);
    case 2:
        return launch_one<MODE, STORE, 2>(a, shmem, s);
    default:
        return launch_one<MODE, STORE, 4>(a, shmem, s);
    }
}

template <int MODE>
hipError_t launch_sweep_store(const TsgSweepArgs &a, int store, int mpl, size_t shmem,
                              hipStream_t s) {
    switch (store) {
    case TSG_STORE_HOT:
        return launch_sweep_mpl<MODE, TSG_STORE_HOT>(a, mpl, shmem, s);
    case TSG_STORE_HOT16:
        return launch_sweep_mpl<MODE, TSG_STORE_HOT16>(a, mpl, shmem, s);
    case TSG_STORE_HOTW:
        return launch_sweep_mpl<MODE, TSG_STORE_HOTW>(a, mpl, shmem, s);
    case TSG_STORE_LDS:
        return launch_sweep_mpl<MODE, TSG_STORE_LDS>(a, mpl, shmem, s);
    default:
        return launch_sweep_mpl<MODE, TSG_STORE_GLOBAL>(a, mpl, shmem, s);
    }
}
}  // namespace

namespace {
template <int MODE, int STORE, int MPL>
hipError_t launch_multi_one(const TsgSweepArgs *args, const int32_t *cpre, int nrun, int total,
                            size_t shmem, hipStream_t s) {
    const void *fn = reinterpret_cast<const void *>(&k_sweep_multi<MODE, STORE, MPL>);
    if (shmem > 65536) {
        if ((int) shmem > g_lds_max) return hipErrorInvalidValue;
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           g_lds_max);
        if (e != hipSuccess) return e;
    }
    const unsigned grid = 8u * (unsigned) ((total + 7) / 8);
    hipLaunchKernelGGL((k_sweep_multi<MODE, STORE, MPL>), dim3(grid), dim3(WAVE), shmem, s, args,
                       cpre, nrun);
    return hipGetLastError();
}

template <int MODE, int STORE>
hipError_t launch_multi_mpl(const TsgSweepArgs *args, const int32_t *cpre, int nrun, int total,
                            int mpl, size_t shmem, hipStream_t s) {
    switch (mpl) {
    case 1:
        return launch_multi_one<MODE, STORE, 1>(args, cpre, nrun, total, shmem, s);
    case 2:
        return launch_multi_one<MODE, STORE, 2>(args, cpre, nrun, total, shmem, s);
    default:
        return launch_multi_one<MODE, STORE, 4>(args, cpre, nrun, total, shmem, s);
    }
}
}  // namespace

hipError_t tsg_launch_sweep_multi(const TsgSweepArgs *args, const int32_t *cpre, int nrun,
                                  int total_chunks, int mode, int store, int mpl, size_t shmem,
                                  hipStream_t s) {
    if (store == TSG_STORE_HOT16) {
        return mode == 0 ? launch_multi_mpl<0, TSG_STORE_HOT16>(args, cpre, nrun, total_chunks,
                                                                  mpl, shmem, s)
                         : launch_multi_mpl<1, TSG_STORE_HOT16>(args, cpre, nrun, total_chunks,
                                                                  mpl, shmem, s);
    }
    if (store == TSG_STORE_HOT) {
        return mode == 0 ? launch_multi_mpl<0, TSG_STORE_HOT>(args, cpre, nrun, total_chunks, mpl,
                                                                shmem, s)
                         : launch_multi_mpl<1, TSG_STORE_HOT>(args, cpre, nrun, total_chunks, mpl,
                                                                shmem, s);
    }
    return hipErrorInvalidValue;
}

void tsg_set_lds_max(int bytes) { g_lds_max = std::max(bytes, 65536); }
int tsg_lds_max() { return g_lds_max; }

hipError_t tsg_launch_sweep(const TsgSweepArgs &a, int mode, int store, int mpl, size_t shmem,
                            hipStream_t s) {
    return (mode == 0) ? launch_sweep_store<0>(a, store, mpl, shmem, s)
                       : launch_sweep_store<1>(a, store, mpl, shmem, s);
}

int tsg_sweep_needs_hop(const TsgOut *outs, int64_t m, int K, int weighted, const double *totw,
                        const double *opar) {
    for (int64_t i = 0; i < m; i++) {
        LaneOut L;
        const int kind = weighted ? moment_form<true>(outs[i], K, L, totw, opar)
                                  : moment_form<false>(outs[i], K, L);
        if (kind == TSG_KIND_HOP) return 1;
    }
    return 0;
}

namespace {
template <typename F>
hipError_t allow_lds(F *fn, size_t shmem) {
    if (shmem <= 65536) return hipSuccess;
    if ((int) shmem > g_lds_max) return hipErrorInvalidValue;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                               hipFuncAttributeMaxDynamicSharedMemorySize, g_lds_max);
}
}  // namespace

hipError_t tsg_launch_afs(const TsgSweepArgs &a, bool branch, bool lds, size_t shmem,
                          hipStream_t s) {
    dim3 grid(a.num_chunks), block(WAVE);
    hipError_t e;
    if (branch) {
        if (lds) {
            if ((e = allow_lds(&k_afs_branch<true>, shmem)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_afs_branch<true>), grid, block, shmem, s, a);
        } else {
            hipLaunchKernelGGL((k_afs_branch<false>), grid, block, shmem, s, a);
        }
    } else {
        if (lds) {
            if ((e = allow_lds(&k_afs_site<true>, shmem)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_afs_site<true>), grid, block, shmem, s, a);
        } else {
            hipLaunchKernelGGL((k_afs_site<false>), grid, block, shmem, s, a);
        }
    }
    return hipGetLastError();
}

hipError_t tsg_launch_afs_hot(const TsgSweepArgs &a, bool branch, size_t shmem, hipStream_t s) {
    dim3 grid(8u * (unsigned) ((a.num_chunks + 7) / 8)), block(WAVE);
    hipError_t e;
    if (branch) {
        if ((e = allow_lds(&k_afs_hot<true>, shmem)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_afs_hot<true>), grid, block, shmem, s, a);
    } else {
        if ((e = allow_lds(&k_afs_hot<false>, shmem)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_afs_hot<false>), grid, block, shmem, s, a);
    }
    return hipGetLastError();
}

hipError_t tsg_launch_afs_last_tree(const TsgSweepArgs &a, const TsgLastTree &lt, double L,
                                    hipStream_t s) {
    const int64_t words = (int64_t) lt.nl * (a.afs_K + 1);
    if (words > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned) ((words + 255) / 256)), dim3(256), 0, s,
                           lt.counts, words);
    }
    const int64_t walkers = (int64_t) a.npairs + a.num_samples;
    if (walkers > 0) {
        hipLaunchKernelGGL(k_afs_tl_counts, dim3((unsigned) ((walkers + 255) / 256)), dim3(256), 0,
                           s, a, lt);
    }
    if (lt.nl > 0) {
        hipLaunchKernelGGL(k_afs_tl_flush, dim3((unsigned) ((lt.nl + 255) / 256)), dim3(256), 0, s,
                           a, lt, L);
    }
    return hipGetLastError();
}

int tsg_tree0_blocks(int nloc) { return (nloc + TSG_T0_NODES - 1) / TSG_T0_NODES; }

hipError_t tsg_launch_tree0(const TsgSweepArgs &a, const int32_t *t0_parent, const double *t0_bl,
                            int32_t *x0, int nloc, double *part, hipStream_t s) {
    const int64_t words = (int64_t) nloc * a.K;
    if (words > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned) ((words + 255) / 256)), dim3(256), 0, s, x0,
                           words);
    }
    if (a.nseed > 0) {
        hipLaunchKernelGGL(k_tree0_counts, dim3((a.nseed + 255) / 256), dim3(256), 0, s, a,
                           t0_parent, x0);
    }
    const int nb = tsg_tree0_blocks(nloc);
    if (nb > 0 && a.M > 0) {
        const size_t sh = a.K <= TSG_T0_STAGE_K
                              ? (size_t) TSG_T0_NODES * (8 + 4 + 4 * (size_t) a.K) : 0;
        hipLaunchKernelGGL(k_tree0_sum<int32_t>, dim3(nb, (a.M + WAVE - 1) / WAVE), dim3(WAVE), sh,
                           s, a, t0_parent, t0_bl, (const int32_t *) x0, nloc, part);
    }
    return hipGetLastError();
}

hipError_t tsg_launch_tree0_w(const TsgSweepArgs &a, const int32_t *t0_parent,
                              const double *t0_bl, const double *x0, int nloc, double *part,
                              hipStream_t s) {
    const int nb = tsg_tree0_blocks(nloc);
    if (nb > 0 && a.M > 0) {
        const size_t sh = a.K <= TSG_T0_STAGE_K
                              ? (size_t) TSG_T0_NODES * (8 + 4 + 8 * (size_t) a.K) : 0;
        hipLaunchKernelGGL(k_tree0_sum<double>, dim3(nb, (a.M + WAVE - 1) / WAVE), dim3(WAVE), sh,
                           s, a, t0_parent, t0_bl, x0, nloc, part);
    }
    return hipGetLastError();
}

int tsg_scan_blocks(int num_chunks) { return (num_chunks + TSG_SCAN_B - 1) / TSG_SCAN_B; }

hipError_t tsg_launch_combine(const TsgCombineArgs &a, const double *dR, double *Rstart,
                              double *bsum, const double *r0part, int nb, int64_t total,
                              hipStream_t s) {
    if (dR != nullptr) {
        const int nblk = tsg_scan_blocks(a.num_chunks);
        const dim3 grid(nblk, (a.M + 63) / 64);
        hipLaunchKernelGGL(k_scan_blocks, grid, dim3(64), 0, s, dR, bsum, a.num_chunks, a.M);
        hipLaunchKernelGGL(k_scan_top, dim3(a.M), dim3(TSG_SCAN_TOP), 0, s, bsum, nblk, a.M,
                           r0part, nb);
        hipLaunchKernelGGL(k_scan_apply, grid, dim3(64), 0, s, dR, bsum, Rstart, a.num_chunks,
                           a.M);
    }
    const int64_t blocks = (total + 255) / 256;
    hipLaunchKernelGGL(k_combine, dim3((unsigned) blocks), dim3(256), 0, s, a, total);
    return hipGetLastError();
}

hipError_t tsg_launch_finals(const void *fin, int nq, const double *raw, const int32_t *ipar,
                             const double *dpar, double *out, int64_t total, hipStream_t s) {
    const int64_t blocks = (total + 255) / 256;
    hipLaunchKernelGGL(k_finals, dim3((unsigned) blocks), dim3(256), 0, s,
                       reinterpret_cast<const TsgFinal *>(fin), nq, raw, ipar, dpar, out, total);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The finish phases of many runs, one launch per phase: the grid covers the
// blocks of every run (prefix pre[r]), a block finds its run by binary search
// and does what the single-run kernel's block of the same index does, so the
// sums are the same, in the same order.
namespace {
__device__ __forceinline__ int run_of(const int64_t *pre, int nrun, int64_t b) {
    int lo = 0, hi = nrun - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= b) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}
}  // namespace

__global__ void __launch_bounds__(256) k_finish_multi_256(const TsgFinishRun *runs,
                                                          const int64_t *pre, int nrun,
                                                          int phase) {
    const int64_t *p = pre + (int64_t) phase * (nrun + 1);
    const int64_t b = blockIdx.x;
    if (b >= p[nrun]) return;
    const int r = run_of(p, nrun, b);
    const TsgFinishRun &R = runs[r];
    const int64_t lb = b - p[r];
    switch (phase) {
    case 0: {  // clear the first tree's counts
        const int64_t i = lb * 256 + threadIdx.x;
        if (i < (int64_t) R.nloc * R.a.K) R.x0[i] = 0;
        break;
    }
    case 1:
        tree0_counts(R.a, R.t0_parent, R.x0, (int) lb);
        break;
    case 6:
        combine(R.ca, R.lanecount, lb);
        break;
    case 7:
        finals(reinterpret_cast<const TsgFinal *>(R.fin), R.nq, R.raw, R.ipar, R.dpar, R.out,
               R.outcount, lb);
        break;
    default:
        break;
    }
}

__global__ void __launch_bounds__(TSG_SCAN_TOP) k_finish_multi_top(const TsgFinishRun *runs,
                                                                   const int64_t *pre, int nrun) {
    __shared__ double red[TSG_SCAN_TOP];
    const int64_t *p = pre + (int64_t) 4 * (nrun + 1);
    const int64_t b = blockIdx.x;
    if (b >= p[nrun]) return;
    const int r = run_of(p, nrun, b);
    const TsgFinishRun &R = runs[r];
    scan_top(R.bsum, R.nblk, R.a.M, R.r0part, R.nb, (int) (b - p[r]), red);
}

__global__ void __launch_bounds__(64) k_finish_multi_64(const TsgFinishRun *runs,
                                                        const int64_t *pre, int nrun, int phase) {
    const int64_t *p = pre + (int64_t) phase * (nrun + 1);
    const int64_t b = blockIdx.x;
    if (b >= p[nrun]) return;
    const int r = run_of(p, nrun, b);
    const TsgFinishRun &R = runs[r];
    const int64_t lb = b - p[r];
    switch (phase) {
    case 2: {  // tree0 sums: grid (nb, ceil(M / 64))
        const int gx = R.nb;
        tree0_sum(R.a, R.t0_parent, R.t0_bl, R.x0, R.nloc, R.r0part, (int) (lb % gx),
                  (int) (lb / gx), false);
        break;
    }
    case 3: {
        const int gx = R.nblk;
        scan_blocks(R.dR, R.bsum, R.a.num_chunks, R.a.M, (int) (lb % gx), (int) (lb / gx));
        break;
    }
    case 5: {
        const int gx = R.nblk;
        scan_apply(R.dR, R.bsum, R.Rstart, R.a.num_chunks, R.a.M, (int) (lb % gx),
                   (int) (lb / gx));
        break;
    }
    default:
        break;
    }
}

// Blocks of run r in each phase (the single-run launchers' grids).
void tsg_finish_blocks(const TsgFinishRun &R, int64_t *blocks) {
    const bool br = R.dR != nullptr;
    const int64_t mb = (R.a.M + 63) / 64;
    blocks[0] = br ? ((int64_t) R.nloc * R.a.K + 255) / 256 : 0;
    blocks[1] = br ? ((int64_t) R.a.nseed + 255) / 256 : 0;
    blocks[2] = (br && R.a.M > 0) ? (int64_t) R.nb * mb : 0;
    blocks[3] = br ? (int64_t) R.nblk * mb : 0;
    blocks[4] = br ? (int64_t) R.a.M : 0;
    blocks[5] = br ? (int64_t) R.nblk * mb : 0;
    blocks[6] = (R.lanecount + 255) / 256;
    blocks[7] = (R.outcount + 255) / 256;
}

hipError_t tsg_launch_finish_multi(const TsgFinishRun *runs, const int64_t *pre,
                                   const int64_t *totals, int nrun, hipStream_t s) {
    for (int ph = 0; ph < TSG_FINISH_PHASES; ph++) {
        const int64_t nb = totals[ph];
        if (nb <= 0) continue;
        if (nb > INT32_MAX) return hipErrorInvalidValue;
        if (ph == 4) {
            hipLaunchKernelGGL(k_finish_multi_top, dim3((unsigned) nb), dim3(TSG_SCAN_TOP), 0, s,
                               runs, pre, nrun);
        } else if (ph == 2 || ph == 3 || ph == 5) {
            hipLaunchKernelGGL(k_finish_multi_64, dim3((unsigned) nb), dim3(64), 0, s, runs, pre,
                               nrun, ph);
        } else {
            hipLaunchKernelGGL(k_finish_multi_256, dim3((unsigned) nb), dim3(256), 0, s, runs, pre,
                               nrun, ph);
        }
    }
    return hipGetLastError();
}

hipError_t tsg_launch_fill_u64(unsigned long long *p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t tsg_launch_fill_f64(double *p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t tsg_launch_afs_site_final(const unsigned long long *h, double *out, const double *win,
                                     int W, int64_t afs_size, uint32_t flags, hipStream_t s) {
    const int64_t n = (int64_t) W * afs_size;
    hipLaunchKernelGGL(k_afs_site_final, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, h,
                       out, win, W, afs_size, flags);
    return hipGetLastError();
}

hipError_t tsg_launch_node(const TsgSweepArgs &a, const int32_t *slot_node, double *res,
                           int64_t N, size_t shmem, const TsgLastTree &lt, const double *node_last,
                           const int32_t *node_lid, double L, hipStream_t s) {
    hipError_t e;
    if ((e = allow_lds(&k_node_hot, shmem)) != hipSuccess) return e;
    if (a.num_chunks > 0) {
        const unsigned grid = 8u * (unsigned) ((a.num_chunks + 7) / 8);
        hipLaunchKernelGGL(k_node_hot, dim3(grid), dim3(WAVE), shmem, s, a, slot_node, res, N);
    }
    // counts of the nl last-tree nodes, then one row of zeros
    const int64_t words = ((int64_t) lt.nl + 1) * a.K;
    const int32_t *zeros = lt.counts + (int64_t) lt.nl * a.K;
    if (words > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned) ((words + 255) / 256)), dim3(256), 0, s,
                           lt.counts, words);
    }
    if (a.npairs > 0) {
        hipLaunchKernelGGL(k_node_tl_counts, dim3((unsigned) ((a.npairs + 255) / 256)), dim3(256),
                           0, s, a, lt);
    }
    const int64_t items = N * a.M;
    if (items > 0) {
        hipLaunchKernelGGL(k_node_tl_flush, dim3((unsigned) ((items + 255) / 256)), dim3(256), 0,
                           s, a, lt, node_last, node_lid, res, N, zeros, L);
    }
    return hipGetLastError();
}

hipError_t tsg_launch_afs_branch_final(double *h, const double *win, int W, int64_t per_window,
                                       uint32_t flags, hipStream_t s) {
    const int64_t n = (int64_t) W * per_window;
    hipLaunchKernelGGL(k_afs_branch_final, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s,
                       h, win, W, per_window, flags);
    return hipGetLastError();
}

size_t tsg_final_struct_size() { return sizeof(TsgFinal); }
