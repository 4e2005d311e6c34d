// HIP kernels for the tree-sequence statistics sweep (gfx950 / CDNA4).
//
// Work decomposition: the genome is cut into chunks of consecutive trees and
// each chunk is one 64-lane wavefront (one workgroup).  A chunk executes its
// sweep program (see tsg_treeseq.cpp): it materialises its start tree, seeds
// per-sample-set counts up the start tree, then applies the edge removals and
// insertions of its transitions in the reference order, walking each changed
// edge's parent path.  Lanes run over state columns (sample sets) for count
// updates and over outputs for summary functions and window sums.
//
// Node records {parent, branch length, x[0..K)} live in LDS when the chunk's
// local node count fits, otherwise in a per-chunk HBM slab.
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
//     transition of each chunk makes the deltas chain);
//   * summary numerators are integer polynomials of the counts and are
//     evaluated exactly in int64, then scaled once in FP64 and rounded to FP32;
//   * running sums and window sums are FP32 TwoSum (value + error) pairs;
//   * a walk whose moved subtree carries zero weight changes no count and no
//     summary, so it is skipped (unless a summary has a zero denominator, in
//     which case the reference's NaN propagation is kept by not skipping).
#include <hip/hip_runtime.h>

#include "tsg_common.h"
#include "tsg_kernels.h"

#define WAVE 64

namespace {

__device__ __forceinline__ void tsum(float &s, float &c, float x) {
    const float t = s + x;
    const float bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
}

// Reference summary function o.stat evaluated on counts x (or on the
// complement tot - x for the unpolarised wrapper).
__device__ float eval_summary(const TsgOut &o, const int32_t *x, const int32_t *tot, bool comp) {
#define XV(col) ((int64_t) (comp ? (tot[col] - x[col]) : x[col]))
    const int64_t xi = XV(o.i);
    switch (o.stat) {
    case TSG_STAT_DIVERSITY:
        return (float) ((double) (xi * (o.ni - xi)) * o.inv);
    case TSG_STAT_SEGREGATING_SITES:
        return (xi > 0) ? (float) ((double) (o.ni - xi) * o.inv) : 0.0f;
    case TSG_STAT_Y1:
        return (float) ((double) (xi * (o.ni - xi) * (o.ni - xi - 1)) * o.inv);
    case TSG_STAT_DIVERGENCE: {
        const int64_t xj = XV(o.j);
        return (float) ((double) (xi * (o.nj - xj)) * o.inv);
    }
    case TSG_STAT_RELATEDNESS: {
        double sum = 0;
        for (int c = o.cbase; c < o.cbase + o.ccount; c++) {
            sum += (double) XV(c) / (double) tot[c];
        }
        const double mean = sum / (double) o.ccount;
        const int64_t xj = XV(o.j);
        return (float) (((double) xi / o.ni - mean) * ((double) xj / o.nj - mean));
    }
    case TSG_STAT_RELATEDNESS_NONCENTRED: {
        const int64_t xj = XV(o.j);
        return (float) ((double) (xi * xj) * o.inv);
    }
    case TSG_STAT_Y2: {
        const int64_t xj = XV(o.j);
        return (float) ((double) (xi * (o.nj - xj) * (o.nj - xj - 1)) * o.inv);
    }
    case TSG_STAT_F2: {
        const int64_t xj = XV(o.j);
        const int64_t a = xi * (o.nj - xj);
        const int64_t b = (xi - 1) * (o.nj - xj - 1) - (o.ni - xi) * xj;
        return (float) ((double) a * (double) b * o.inv);
    }
    case TSG_STAT_Y3: {
        const int64_t xj = XV(o.j), xk = XV(o.k);
        return (float) ((double) (xi * (o.nj - xj) * (o.nk - xk)) * o.inv);
    }
    case TSG_STAT_F3: {
        const int64_t xj = XV(o.j), xk = XV(o.k);
        const int64_t a = xi * (o.nj - xj);
        const int64_t b = (xi - 1) * (o.nk - xk) - (o.ni - xi) * xk;
        return (float) ((double) a * (double) b * o.inv);
    }
    case TSG_STAT_F4: {
        const int64_t xj = XV(o.j), xk = XV(o.k), xl = XV(o.l);
        const int64_t a = xi * (o.nj - xj);
        const int64_t b = xk * (o.nl - xl) - xl * (o.nk - xk);
        return (float) ((double) a * (double) b * o.inv);
    }
    default:
        return 0.0f;
    }
#undef XV
}

// Per-chunk node storage.  rec[u*R] = parent, rec[u*R+1] = branch length
// bits, rec[u*R+2+k] = count of state column k below u.
struct Nodes {
    int32_t *rec;
    int32_t R;
    __device__ int32_t parent(int u) const { return rec[(int64_t) u * R]; }
    __device__ float bl(int u) const { return __int_as_float(rec[(int64_t) u * R + 1]); }
    __device__ int32_t &x(int u, int k) const { return rec[(int64_t) u * R + 2 + k]; }
    __device__ void set_edge(int u, int p, float b) const {
        rec[(int64_t) u * R] = p;
        rec[(int64_t) u * R + 1] = __float_as_int(b);
    }
};

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

struct LaneOut {
    TsgOut o;
    int m;
    int wi, wend;
    const double *win;
    int64_t slot0;  // partial index of window wfirst
    int Mq;
    float Rs, Rc, As, Ac;
};

template <int MPL>
__device__ void setup_outputs(const TsgSweepArgs &a, int c, int lane, LaneOut (&lo)[MPL]) {
#pragma unroll
    for (int j = 0; j < MPL; j++) {
        LaneOut &L = lo[j];
        L.m = a.m_begin + lane + WAVE * j;
        L.Rs = L.Rc = L.As = L.Ac = 0.0f;
        L.wi = L.wend = 0;
        if (L.m < a.m_end) {
            L.o = a.outs[L.m];
            const TsgQuery &q = a.queries[L.o.q];
            const int64_t cq = (int64_t) c * a.Q + L.o.q;
            L.win = a.windows + q.win_off;
            L.wi = a.cq_wfirst[cq];
            L.wend = L.wi + a.cq_nwin[cq];
            L.Mq = q.M;
            L.slot0 = q.pbase + a.cq_poff[cq] * (int64_t) q.M + (L.m - q.mbase);
        }
    }
}

__device__ __forceinline__ void flush_window(const TsgSweepArgs &a, LaneOut &L, int wfirst) {
    a.partial[L.slot0 + (int64_t) (L.wi - wfirst) * L.Mq] = (double) L.As + (double) L.Ac;
    L.As = L.Ac = 0.0f;
}

// Unpolarised summaries add f(total - x) (trees.c:1944-1972).
__device__ __forceinline__ float branch_summary(const TsgOut &o, const int32_t *x,
                                                const int32_t *tot) {
    float f = eval_summary(o, x, tot, false);
    if (!(o.flags & 1)) f = f + eval_summary(o, x, tot, true);
    return f;
}

}  // namespace

// MODE 0: branch statistics, MODE 1: site statistics.
template <int MODE, bool LDS, int MPL>
__global__ void __launch_bounds__(WAVE) k_sweep(const TsgSweepArgs a) {
    extern __shared__ int32_t smem[];
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int K = a.K;
    const int R = 2 + K;
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
    int32_t *xc = stg;                              // [K] counts of the moved subtree
    int32_t *xo = stg + K;                          // [PSEG x K] counts before the update
    int32_t *xn = xo + TSG_PSEG * K;                // [PSEG x K] counts after the update
    float *sbl = reinterpret_cast<float *>(xn + TSG_PSEG * K);  // [PSEG] branch lengths
    const bool any_unpolarised = (a.flags & TSG_F_ANY_UNPOLARISED) != 0;
    const bool noskip = (MODE == 0) && (a.flags & TSG_F_NOSKIP) != 0;
    const int32_t *tot = a.col_total;

    init_chunk(a, nd, n0, nloc, K, lane);

    LaneOut lo[MPL];
    setup_outputs<MPL>(a, c, lane, lo);
    int wfirst[MPL];
#pragma unroll
    for (int j = 0; j < MPL; j++) wfirst[j] = lo[j].wi;

    auto apply_transition = [&](int t) {
        const int64_t e0 = a.tr_off[t], e1 = a.tr_off[t + 1];
        const int64_t nrem = a.tr_nrem[t];
        for (int64_t e = e0; e < e1; e++) {
            const bool ins = (e - e0) >= nrem;
            const int u = a.ev_child[e];
            const int v = a.ev_parent[e];
            bool nz = false;
            for (int k = lane; k < K; k += WAVE) {
                const int32_t xk = nd.x(u, k);
                xc[k] = xk;
                nz |= (xk != 0);
            }
            nz = __any(nz);
            const bool walk = nz || noskip;
            __syncthreads();
            // the edge's own term bl * F(x_child): needed even for a zero-weight
            // subtree, because unpolarised summaries need not vanish at zero
            // (non-centred relatedness: F(0) = f(0) + f(total) = 1)
            if (MODE == 0) {
                const float b = ins ? a.ev_bl[e] : nd.bl(u);
                const float sg = ins ? 1.0f : -1.0f;
#pragma unroll
                for (int j = 0; j < MPL; j++) {
                    if (lo[j].m < a.m_end) {
                        const float f = branch_summary(lo[j].o, xc, tot);
                        tsum(lo[j].Rs, lo[j].Rc, sg * __fmul_rn(b, f));
                    }
                }
            }
            __syncthreads();
            if (lane == 0) {
                if (ins) {
                    nd.set_edge(u, v, a.ev_bl[e]);
                } else {
                    nd.set_edge(u, TSG_NULL, 0.0f);
                }
            }
            __syncthreads();
            if (!walk) continue;
            // The event's root path is precomputed (v first).  All (hop,
            // column) counts of a path segment are updated in parallel, then
            // every output sums -bl F(old) + bl F(new) over the segment's
            // hops -- the same terms as the reference's hop-by-hop walk
            // (trees.c:1436-1448, 1461-1472), since a path visits each node once.
            const int64_t p0 = a.ev_path0[e], p1 = a.ev_path0[e + 1];
            (void) v;
            for (int64_t sg0 = p0; sg0 < p1; sg0 += TSG_PSEG) {
                const int h = (int) min((int64_t) TSG_PSEG, p1 - sg0);
                for (int idx = lane; idx < h * K; idx += WAVE) {
                    const int i = idx / K, k = idx - i * K;
                    const int node = a.path[sg0 + i];
                    const int32_t old = nd.x(node, k);
                    const int32_t nw = ins ? old + xc[k] : old - xc[k];
                    nd.x(node, k) = nw;
                    if (MODE == 0) {
                        xo[idx] = old;
                        xn[idx] = nw;
                    }
                }
                if (MODE == 0) {
                    for (int i = lane; i < h; i += WAVE) sbl[i] = nd.bl(a.path[sg0 + i]);
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < MPL; j++) {
                        if (lo[j].m < a.m_end) {
                            for (int i = 0; i < h; i++) {
                                const float bw = sbl[i];
                                const float fo = branch_summary(lo[j].o, xo + i * K, tot);
                                const float fn = branch_summary(lo[j].o, xn + i * K, tot);
                                tsum(lo[j].Rs, lo[j].Rc, -__fmul_rn(bw, fo));
                                tsum(lo[j].Rs, lo[j].Rc, __fmul_rn(bw, fn));
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
    };

    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    for (int t = t0; t < t1; t++) {
        if (t > t0) apply_transition(t);
        if (MODE == 0) {
            const double tl = a.bp[t], tr = a.bp[t + 1];
#pragma unroll
            for (int j = 0; j < MPL; j++) {
                LaneOut &L = lo[j];
                if (L.m >= a.m_end) continue;
                while (L.wi < L.wend) {
                    const double wl = L.win[L.wi], wr = L.win[L.wi + 1];
                    if (!(wl < tr)) break;
                    const float ov = (float) (fmin(tr, wr) - fmax(tl, wl));
                    tsum(L.As, L.Ac, __fmul_rn(L.Rs, ov));
                    tsum(L.As, L.Ac, __fmul_rn(L.Rc, ov));
                    if (wr <= tr) {
                        flush_window(a, L, wfirst[j]);
                        L.wi++;
                    } else {
                        break;
                    }
                }
            }
        } else {
            const int s0 = a.tree_site0[t], s1 = a.tree_site0[t + 1];
            for (int s = s0; s < s1; s++) {
                const double pos = a.site_pos[s];
#pragma unroll
                for (int j = 0; j < MPL; j++) {
                    LaneOut &L = lo[j];
                    if (L.m >= a.m_end) continue;
                    while (L.wi < L.wend && L.win[L.wi + 1] <= pos) {
                        flush_window(a, L, wfirst[j]);
                        L.wi++;
                    }
                }
                const int na = a.site_nalleles[s];
                const int m0 = a.site_mut0[s], m1 = a.site_mut0[s + 1];
                float val[MPL];
#pragma unroll
                for (int j = 0; j < MPL; j++) val[j] = 0.0f;
                // polarised outputs skip the ancestral allele (trees.c:1639)
                for (int al = any_unpolarised ? 0 : 1; al < na; al++) {
                    for (int k = lane; k < K; k += WAVE) {
                        int32_t w = (al == 0) ? tot[k] : 0;
                        for (int mm = m0; mm < m1; mm++) {
                            const int node = a.mut_node[mm];
                            if (a.mut_allele[mm] == al) w += nd.x(node, k);
                            if (a.mut_pallele[mm] == al) w -= nd.x(node, k);
                        }
                        xo[k] = w;
                    }
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < MPL; j++) {
                        if (lo[j].m < a.m_end && (al > 0 || !(lo[j].o.flags & 1))) {
                            val[j] += eval_summary(lo[j].o, xo, tot, false);
                        }
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int j = 0; j < MPL; j++) {
                    if (lo[j].m < a.m_end) tsum(lo[j].As, lo[j].Ac, val[j]);
                }
            }
        }
    }
    if (MODE == 0) {
        if (t1 < a.num_trees) apply_transition(t1);
#pragma unroll
        for (int j = 0; j < MPL; j++) {
            if (lo[j].m < a.m_end) {
                a.dR[(int64_t) c * a.M + lo[j].m] = (double) lo[j].Rs + (double) lo[j].Rc;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < MPL; j++) {
        LaneOut &L = lo[j];
        if (L.m >= a.m_end) continue;
        while (L.wi < L.wend) {
            flush_window(a, L, wfirst[j]);
            L.wi++;
        }
    }
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
    const int R = 2 + K;
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
    const int R = 2 + K;
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
// Running branch sum of the first tree (chunk 0's start tree), evaluated in
// parallel instead of by replaying its edge insertions one at a time:
//   R(tree 0) = sum over nodes u with a parent of bl(u) * F(x(u)),
// which is what the reference's insertion loop accumulates (trees.c:1426-
// 1480, update_running_sum 1339-1350).  Each term is the FP32 product the
// sweep adds for the same edge; the sum is FP64, per block of nodes, in a
// fixed order (deterministic).
//
// Counts: one thread per weighted sample walks its root path in the
// snapshot, adding 1 to each of its columns.
__global__ void k_tree0_counts(const TsgSweepArgs a, int32_t *x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nseed) return;
    const int c0 = a.seed_off[i], c1 = a.seed_off[i + 1];
    for (int u = a.seed_sample[i]; u >= 0; u = a.loc_parent0[u]) {
        for (int c = c0; c < c1; c++) atomicAdd(&x0[(int64_t) u * a.K + a.seed_col[c]], 1);
    }
}

#define TSG_T0_NODES 256

// grid (ceil(nloc / TSG_T0_NODES), ceil(M / 64)); lane = output.
__global__ void __launch_bounds__(WAVE) k_tree0_sum(const TsgSweepArgs a, const int32_t *x0,
                                                    int nloc, double *part) {
    const int m = blockIdx.y * WAVE + threadIdx.x;
    if (m >= a.M) return;
    const int u0 = blockIdx.x * TSG_T0_NODES;
    const int u1 = min(nloc, u0 + TSG_T0_NODES);
    const TsgOut o = a.outs[m];
    double acc = 0;
    for (int u = u0; u < u1; u++) {
        if (a.loc_parent0[u] < 0) continue;
        const float f = branch_summary(o, x0 + (int64_t) u * a.K, a.col_total);
        acc += (double) __fmul_rn(a.loc_bl0[u], f);
    }
    part[(int64_t) blockIdx.x * a.M + m] = acc;
}

// ---------------------------------------------------------------------------
// Combine: per (query, window, output), sum the chunk partials of the chunks
// that overlap the window, re-base branch sums by the chunk start value, and
// span-normalise.  FP64, deterministic chunk order.
// Rstart[c] = R(first tree) + sum of the chunk deltas before c; the first
// tree's running sum arrives as nb per-block partials (k_tree0_sum).
__global__ void k_scan_dR(const double *dR, double *Rstart, int C, int M, const double *r0part,
                          int nb) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    double acc = 0;
    for (int b = 0; b < nb; b++) acc += r0part[(int64_t) b * M + m];
    for (int c = 0; c < C; c++) {
        Rstart[(int64_t) c * M + m] = acc;
        acc += dR[(int64_t) c * M + m];
    }
}

__global__ void k_combine(const TsgCombineArgs a, int64_t total) {
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    // locate (q, w, ml): rawbase is the prefix sum of W_q * M_q
    int qlo = 0, qhi = a.Q - 1;
    while (qlo < qhi) {
        const int mid = (qlo + qhi + 1) / 2;
        if (a.rawbase[mid] <= gid) {
            qlo = mid;
        } else {
            qhi = mid - 1;
        }
    }
    const int q = qlo;
    const int64_t rem = gid - a.rawbase[q];
    const TsgQuery &Qr = a.queries[q];
    const int w = (int) (rem / Qr.M);
    const int ml = (int) (rem % Qr.M);
    const double *win = a.windows + Qr.win_off;
    const double wl = win[w], wr = win[w + 1];
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
    double acc = 0;
    for (int c = lo; c < a.num_chunks; c++) {
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
    if (Qr.flags & TSG_F_SPAN_NORMALISE) acc /= (wr - wl);
    a.raw[a.rawbase[q] + (int64_t) w * Qr.M + ml] = acc;
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

__global__ void k_finals(const TsgFinal *fin, int nq, const double *raw, const int32_t *ipar,
                         const double *dpar, double *out, int64_t total) {
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
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
template <int MODE, bool LDS>
hipError_t launch_sweep_mpl(const TsgSweepArgs &a, int mpl, size_t shmem, hipStream_t s) {
    dim3 grid(a.num_chunks), block(WAVE);
    switch (mpl) {
    case 1:
        hipLaunchKernelGGL((k_sweep<MODE, LDS, 1>), grid, block, shmem, s, a);
        break;
    case 2:
        hipLaunchKernelGGL((k_sweep<MODE, LDS, 2>), grid, block, shmem, s, a);
        break;
    case 4:
        hipLaunchKernelGGL((k_sweep<MODE, LDS, 4>), grid, block, shmem, s, a);
        break;
    default:
        hipLaunchKernelGGL((k_sweep<MODE, LDS, 8>), grid, block, shmem, s, a);
        break;
    }
    return hipGetLastError();
}
}  // namespace

hipError_t tsg_launch_sweep(const TsgSweepArgs &a, int mode, bool lds, int mpl, size_t shmem,
                            hipStream_t s) {
    if (mode == 0) {
        return lds ? launch_sweep_mpl<0, true>(a, mpl, shmem, s)
                   : launch_sweep_mpl<0, false>(a, mpl, shmem, s);
    }
    return lds ? launch_sweep_mpl<1, true>(a, mpl, shmem, s)
               : launch_sweep_mpl<1, false>(a, mpl, shmem, s);
}

hipError_t tsg_launch_afs(const TsgSweepArgs &a, bool branch, bool lds, size_t shmem,
                          hipStream_t s) {
    dim3 grid(a.num_chunks), block(WAVE);
    if (branch) {
        if (lds) {
            hipLaunchKernelGGL((k_afs_branch<true>), grid, block, shmem, s, a);
        } else {
            hipLaunchKernelGGL((k_afs_branch<false>), grid, block, shmem, s, a);
        }
    } else {
        if (lds) {
            hipLaunchKernelGGL((k_afs_site<true>), grid, block, shmem, s, a);
        } else {
            hipLaunchKernelGGL((k_afs_site<false>), grid, block, shmem, s, a);
        }
    }
    return hipGetLastError();
}

int tsg_tree0_blocks(int nloc) { return (nloc + TSG_T0_NODES - 1) / TSG_T0_NODES; }

hipError_t tsg_launch_tree0(const TsgSweepArgs &a, int32_t *x0, int nloc, double *part,
                            hipStream_t s) {
    const int64_t words = (int64_t) nloc * a.K;
    if (words > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned) ((words + 255) / 256)), dim3(256), 0, s, x0,
                           words);
    }
    if (a.nseed > 0) {
        hipLaunchKernelGGL(k_tree0_counts, dim3((a.nseed + 255) / 256), dim3(256), 0, s, a, x0);
    }
    const int nb = tsg_tree0_blocks(nloc);
    if (nb > 0 && a.M > 0) {
        hipLaunchKernelGGL(k_tree0_sum, dim3(nb, (a.M + WAVE - 1) / WAVE), dim3(WAVE), 0, s, a, x0,
                           nloc, part);
    }
    return hipGetLastError();
}

hipError_t tsg_launch_combine(const TsgCombineArgs &a, const double *dR, double *Rstart,
                              const double *r0part, int nb, int64_t total, hipStream_t s) {
    if (dR != nullptr) {
        hipLaunchKernelGGL(k_scan_dR, dim3((a.M + 63) / 64), dim3(64), 0, s, dR, Rstart,
                           a.num_chunks, a.M, r0part, nb);
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

hipError_t tsg_launch_afs_branch_final(double *h, const double *win, int W, int64_t per_window,
                                       uint32_t flags, hipStream_t s) {
    const int64_t n = (int64_t) W * per_window;
    hipLaunchKernelGGL(k_afs_branch_final, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s,
                       h, win, W, per_window, flags);
    return hipGetLastError();
}

size_t tsg_final_struct_size() { return sizeof(TsgFinal); }
