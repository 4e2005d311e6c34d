// Site-mode genetic relatedness matrix on the matrix cores (gfx950).
//
// For sites with at most two alleles the reference's
// genetic_relatedness_matrix(mode="site") (python/tskit/trees.py:8978-9102,
// from the divergence matrix of c/tskit/trees.c:8753-8826) reduces to
//     K = J C J / span,   C = G^T G,   J = I - 11^T / n,
// where G is the site x sample 0/1 matrix of derived alleles: the divergence
// matrix is D = a 1^T + 1 a^T - 2C (a = derived counts per sample) and the
// double centring -1/2 J D J removes the a terms exactly.  A site with more
// alleles differs between samples i and j iff d_i + d_j - d_i d_j -
// sum_a B_ai B_aj = 1 (d its any-derived row, B_a its row of allele a >= 1),
// so its rows d and B_a go into a second cross product C2 and, with
// P = 2C + C2, D = (P_ii + P_jj)/2 - P_ij and K = J P J / 2.  C and C2 are
// exact integer cross products (entries <= number of rows < 2^31), computed
// here with int8 MFMA and int32 accumulation; the centring is FP64.
//
// Kernels:
//   k_geno_bits  one wavefront per chunk: the chunk sweep with a per-node
//                bitset of the samples below it (insert: OR, remove: AND-NOT);
//                each site's row is the bitset of its mutation's node.
//   k_bits_to_gt bit rows (site-major) -> int8 Gt (sample-major, each
//                window's sites zero-padded to a multiple of 64).
//   k_syrk_i8    C = Gt Gt^T on v_mfma_i32_32x32x32_i8, upper-triangular
//                128x128 tiles, mirrored on store.
//   k_centre     row sums, then K = (C - r_i/n - r_j/n + t/n^2) / span (FP64).
#include <hip/hip_runtime.h>

#include "tsg_common.h"
#include "tsg_kernels.h"
#include "tsg_launch.h"

#define WAVE 64

// --- genotype bitsets ------------------------------------------------------
// Word w of site s's bit rows from the bitsets of its mutations' nodes
// (below(node) = samples below it), following tsk_variant_decode's rule
// (genotypes.c:473-595): samples start at allele 0 and each mutation, in
// order, sets the samples below its node to its allele.  A site with at most
// two alleles gets one row, its derived allele; a site with three or more
// gets its any-derived row in gbits[s] and one row per derived allele at
// gbits[site_row2[s] + allele - 1].
template <typename Below>
__device__ __forceinline__ void site_rows(const TsgGenoArgs &a, int s, int w, int NW,
                                          Below below) {
    const int m0 = a.site_mut0[s], m1 = a.site_mut0[s + 1];
    const int r2 = a.site_row2 ? a.site_row2[s] : -1;
    if (r2 < 0) {
        uint32_t b1 = 0u;
        for (int m = m0; m < m1; m++) {
            const uint32_t b = below(a.mut_node[m]);
            b1 = (a.mut_allele[m] != 0) ? (b1 | b) : (b1 & ~b);
        }
        a.gbits[(int64_t) s * NW + w] = b1;
        return;
    }
    const int na = a.site_nalleles[s];
    uint32_t *rows = a.gbits + (int64_t) r2 * NW + w;
    for (int al = 1; al < na; al++) rows[(int64_t) (al - 1) * NW] = 0u;
    for (int m = m0; m < m1; m++) {
        const uint32_t b = below(a.mut_node[m]);
        const int al = a.mut_allele[m];
        for (int q = 1; q < na; q++) {
            uint32_t &r = rows[(int64_t) (q - 1) * NW];
            r = (q == al) ? (r | b) : (r & ~b);
        }
    }
    uint32_t d = 0u;
    for (int al = 1; al < na; al++) d |= rows[(int64_t) (al - 1) * NW];
    a.gbits[(int64_t) s * NW + w] = d;
}

__device__ void geno_bits_chunk(const TsgGenoArgs &a, int c, uint32_t *rec);

// Each workgroup owns one slab and loops over chunks blockIdx.x,
// blockIdx.x + gridDim.x, ... so the slab memory is bounded by the grid.
__global__ void __launch_bounds__(WAVE) k_geno_bits(const TsgGenoArgs a) {
    for (int c = blockIdx.x; c < a.num_chunks; c += gridDim.x) {
        __syncthreads();
        geno_bits_chunk(a, c, a.g_rec + (int64_t) blockIdx.x * a.g_stride);
    }
}

__device__ void geno_bits_chunk(const TsgGenoArgs &a, int c, uint32_t *rec) {
    const int lane = threadIdx.x;
    const int NW = a.NW;
    const int R = 1 + NW;  // parent, then NW bitset words
    const int64_t n0 = a.chunk_node0[c];
    const int nloc = (int) (a.chunk_node0[c + 1] - n0);
    auto par = [&](int u) -> int32_t & { return reinterpret_cast<int32_t &>(rec[(int64_t) u * R]); };
    auto word = [&](int u, int w) -> uint32_t & { return rec[(int64_t) u * R + 1 + w]; };

    for (int u = lane; u < nloc; u += WAVE) {
        par(u) = a.loc_parent0[n0 + u];
        for (int w = 0; w < NW; w++) word(u, w) = 0u;
    }
    __syncthreads();
    // every sample's bit, up its root path in the start tree
    for (int j = lane; j < a.num_samples; j += WAVE) {
        const uint32_t bit = 1u << (j & 31);
        const int w = j >> 5;
        for (int u = j; u >= 0; u = par(u)) atomicOr(&word(u, w), bit);
    }
    __syncthreads();
    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    for (int t = t0; t < t1; t++) {
        if (t > t0) {
            const int64_t e0 = a.tr_off[t], e1 = a.tr_off[t + 1];
            const int64_t nrem = a.tr_nrem[t];
            for (int64_t e = e0; e < e1; e++) {
                const bool ins = (e - e0) >= nrem;
                const int u = a.ev_child[e], v = a.ev_parent[e];
                // lanes own words w = lane, lane + 64, ...; a chunk's child
                // bitset is re-read per ancestor from the same lane's word
                if (lane == 0) par(u) = ins ? v : TSG_NULL;
                __syncthreads();
                for (int w = lane; w < NW; w += WAVE) {
                    const uint32_t xc = word(u, w);
                    if (xc == 0u) continue;
                    for (int p = v; p >= 0; p = par(p)) {
                        word(p, w) = ins ? (word(p, w) | xc) : (word(p, w) & ~xc);
                    }
                }
                __syncthreads();
            }
        }
        const int s0 = a.tree_site0[t], s1 = a.tree_site0[t + 1];
        for (int s = s0; s < s1; s++) {
            for (int w = lane; w < NW; w += WAVE) {
                site_rows(a, s, w, NW, [&](int node) { return word(node, w); });
            }
        }
    }
}

// --- genotype bitsets on the hot plan ---------------------------------------
// One workgroup per (chunk, block of WB bitset words = 32*WB samples): the
// chunk's hot slots hold bitsets of the samples below them in LDS.  Seeding:
// each sample of the block sets its bit in its nearest hot ancestor-or-self,
// then the hot forest's levels OR upward, deepest first.  Events: lane w owns
// word w of every slot, so an edge insertion ORs the child's word into each
// node of its root path and a removal clears it (the child's samples lie
// below every path node and nowhere else among them) -- no lane reads
// another lane's words, so the event loop needs no barrier.  Each site's row
// rows come from its mutations' node bitsets (site_rows).  Grid: 8 * ceil(C / 8) x blocks.
//
// GENO: the genotype decode instead (tsk_variant_decode, genotypes.c:473-595):
// every sample starts at allele 0 and each mutation of the site, in order,
// sets the samples below its node to its allele index; lane w writes the 32
// int32 genotypes of its word.  Missing data is marked afterwards from the
// host's list of isolated samples (k_mark_missing).
template <bool GENO>
__global__ void __launch_bounds__(WAVE) k_geno_hot(const TsgGenoArgs a) {
    extern __shared__ uint32_t bits[];
    const int C = a.num_chunks;
    const int per = (C + 7) / 8;
    const int c = (blockIdx.x % 8) * per + blockIdx.x / 8;  // XCD-contiguous chunks
    if (c >= C) return;
    const int lane = threadIdx.x;
    const int WB = a.WB;
    const int w0 = blockIdx.y * WB;                // first word of the block
    const int nwb = min(WB, a.NW - w0);            // words in this block
    const int64_t s0 = a.chunk_node0[c];
    const int S = (int) (a.chunk_node0[c + 1] - s0);
    for (int i = lane; i < S * WB; i += WAVE) bits[i] = 0u;
    __syncthreads();
    const uint16_t *ss = a.sample_slot + (int64_t) (c / TSG_SS_CB) * a.num_samples * TSG_SS_CB +
                         c % TSG_SS_CB;
    const int j0 = w0 * 32, j1 = min(a.num_samples, (w0 + nwb) * 32);
    for (int j = j0 + lane; j < j1; j += WAVE) {
        const unsigned h = ss[(int64_t) j * TSG_SS_CB];
        if (h != TSG_NO_SLOT) atomicOr(&bits[h * WB + ((j >> 5) - w0)], 1u << (j & 31));
    }
    __syncthreads();
    const int32_t *lev = a.lev_slot0 + a.chunk_lev0[c];
    const int nlev = (int) (a.chunk_lev0[c + 1] - a.chunk_lev0[c]) - 1;
    for (int l = 0; l + 1 < nlev; l++) {
        const int h0 = lev[l], h1 = lev[l + 1];
        for (int idx = lane; idx < (h1 - h0) * nwb; idx += WAVE) {
            const int h = h0 + idx / nwb, w = idx % nwb;
            const int p = a.slot_par[s0 + h];
            const uint32_t v = bits[h * WB + w];
            if (p >= 0 && v != 0u) atomicOr(&bits[p * WB + w], v);
        }
        __syncthreads();
    }
    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    for (int t = t0; t < t1; t++) {
        if (t > t0) {
            const int64_t e0 = a.tr_off[t], e1 = a.tr_off[t + 1];
            const int64_t nrem = a.tr_nrem[t];
            for (int64_t e = e0; e < e1; e++) {
                const bool ins = (e - e0) >= nrem;
                const int u = a.ev_child[e];
                const int64_t p0 = a.ev_path0[e], p1 = a.ev_path0[e + 1];
                if (lane < nwb) {
                    const uint32_t xc = bits[u * WB + lane];
                    if (xc != 0u) {
                        for (int64_t q = p0; q < p1; q++) {
                            uint32_t &b = bits[a.path[q] * WB + lane];
                            b = ins ? (b | xc) : (b & ~xc);
                        }
                    }
                }
            }
        }
        const int sa = a.tree_site0[t], sb = a.tree_site0[t + 1];
        for (int s = sa; s < sb; s++) {
            const int m0 = a.site_mut0[s], m1 = a.site_mut0[s + 1];
            if constexpr (GENO) {
                if (lane < nwb) {
                    int32_t g[32];
#pragma unroll
                    for (int b = 0; b < 32; b++) g[b] = 0;
                    for (int m = m0; m < m1; m++) {
                        const uint32_t mask = bits[a.mut_node[m] * WB + lane];
                        const int32_t al = a.mut_allele[m];
#pragma unroll
                        for (int b = 0; b < 32; b++) {
                            if ((mask >> b) & 1u) g[b] = al;
                        }
                    }
                    const int j0w = (w0 + lane) * 32;
                    int32_t *row = a.geno + (int64_t) s * a.num_samples + j0w;
                    const int nb = min(32, a.num_samples - j0w);
#pragma unroll
                    for (int b = 0; b < 32; b++) {
                        if (b < nb) row[b] = g[b];
                    }
                }
            } else {
                if (lane < nwb) {
                    site_rows(a, s, w0 + lane, a.NW,
                              [&](int node) { return bits[node * WB + lane]; });
                }
            }
        }
    }
}

// --- bits (site-major) -> int8 Gt (sample-major, K padded per window) -------
// grid: (ceil(S_pad_total / 64), n_pad / 64); block 256.  Tile: 64 sites x
// 64 samples.  site_of_col[k] maps a padded column to its site (-1 = pad).
__global__ void __launch_bounds__(256) k_bits_to_gt(const uint32_t *gbits, int NW,
                                                    const int32_t *site_of_col, int64_t Kp,
                                                    int n, int8_t *gt) {
    __shared__ uint32_t tile[64][2];  // 64 sites x 64 sample bits
    const int64_t k0 = (int64_t) blockIdx.x * 64;
    const int j0 = blockIdx.y * 64;  // first sample of the tile
    const int tid = threadIdx.x;
    if (tid < 128) {
        const int r = tid >> 1, h = tid & 1;
        const int64_t k = k0 + r;
        uint32_t v = 0u;
        const int w = (j0 >> 5) + h;
        if (k < Kp) {
            const int s = site_of_col[k];
            if (s >= 0 && w < NW) v = gbits[(int64_t) s * NW + w];
        }
        tile[r][h] = v;
    }
    __syncthreads();
    // each thread writes 16 consecutive columns (sites) of one sample row
    const int jj = tid >> 2;          // sample within tile, 0..63
    const int kk = (tid & 3) * 16;    // 16-column group
    const int j = j0 + jj;
    if (k0 + kk >= Kp) return;
    int8_t out[16];
    for (int q = 0; q < 16; q++) {
        const uint32_t wv = tile[kk + q][jj >> 5];
        out[q] = (j < n) ? (int8_t) ((wv >> (jj & 31)) & 1u) : (int8_t) 0;
    }
    int4 v;
    __builtin_memcpy(&v, out, 16);
    *reinterpret_cast<int4 *>(gt + (int64_t) j * Kp + k0 + kk) = v;
}

// --- SYRK on int8 MFMA --------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

#define BT 128        // output tile
#define BK 64         // k per LDS stage
#define LDSP 80       // padded LDS row (bytes): 64 + 16

// One workgroup (4 waves) per upper-triangular 128x128 tile of one window.
// Wave (wy, wx) computes the 64x64 quadrant as 2x2 tiles of 32x32.
__global__ void __launch_bounds__(256) k_syrk_i8(const int8_t *gt, int64_t Kp, int n_pad,
                                                 const int64_t *win_k0, const int64_t *win_k1,
                                                 int tiles_per_dim, int32_t *C) {
    __shared__ __attribute__((aligned(16))) int8_t As[BT * LDSP];
    __shared__ __attribute__((aligned(16))) int8_t Bs[BT * LDSP];
    const int w = blockIdx.y;  // window
    // invert the upper-triangular tile index
    int t = blockIdx.x, ti = 0;
    int rowlen = tiles_per_dim;
    while (t >= rowlen) {
        t -= rowlen;
        ti++;
        rowlen--;
    }
    const int tj = ti + t;
    const int i0 = ti * BT, j0 = tj * BT;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wy = wave >> 1, wx = wave & 1;
    v16i acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int r = 0; r < 16; r++) acc[a][b][r] = 0;
    const int64_t kbeg = win_k0[w], kend = win_k1[w];
    for (int64_t k = kbeg; k < kend; k += BK) {
        // stage 128 rows x 64 bytes of A (rows i0..) and B (rows j0..)
        for (int q = tid; q < BT * 4; q += 256) {
            const int r = q >> 2, seg = (q & 3) * 16;
            *reinterpret_cast<int4 *>(As + r * LDSP + seg) =
                *reinterpret_cast<const int4 *>(gt + (int64_t) (i0 + r) * Kp + k + seg);
            *reinterpret_cast<int4 *>(Bs + r * LDSP + seg) =
                *reinterpret_cast<const int4 *>(gt + (int64_t) (j0 + r) * Kp + k + seg);
        }
        __syncthreads();
        for (int kk = 0; kk < BK; kk += 32) {
            const int ko = kk + 16 * (lane >> 5);
            v4i af[2], bf[2];
            for (int a = 0; a < 2; a++) {
                const int r = wy * 64 + a * 32 + (lane & 31);
                af[a] = *reinterpret_cast<const v4i *>(As + r * LDSP + ko);
            }
            for (int b = 0; b < 2; b++) {
                const int r = wx * 64 + b * 32 + (lane & 31);
                bf[b] = *reinterpret_cast<const v4i *>(Bs + r * LDSP + ko);
            }
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[a], bf[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    int32_t *Cw = C + (int64_t) w * n_pad * n_pad;
    for (int a = 0; a < 2; a++) {
        for (int b = 0; b < 2; b++) {
            const int col = j0 + wx * 64 + b * 32 + (lane & 31);
            for (int r = 0; r < 16; r++) {
                const int row = i0 + wy * 64 + a * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                const int32_t v = acc[a][b][r];
                Cw[(int64_t) row * n_pad + col] = v;
                Cw[(int64_t) col * n_pad + row] = v;
            }
        }
    }
}

// --- centring (FP64) ----------------------------------------------------------
// The n selected samples (sample indexes sel[0..n)) form the matrix.
// rowsum[w][a] = sum_b C[w][sel_a][sel_b] (exact in int64); block per (a, w).
// With sites of three or more alleles the matrices take P = 2 C + C2 (C
// over the rows of the other sites, C2 over the any-derived and per-allele
// rows of those sites, C2 null when there are none): two samples differ at
// a site in (P_ii + P_jj) / 2 - P_ij sites, and the relatedness is
// J P J / 2; without C2, P = 2C and these reduce to the biallelic forms.
__device__ __forceinline__ int64_t cross(const int32_t *C, const int32_t *C2, int64_t idx) {
    return C2 ? 2 * (int64_t) C[idx] + (int64_t) C2[idx] : (int64_t) C[idx];
}

__global__ void __launch_bounds__(256) k_rowsum(const int32_t *C, const int32_t *C2,
                                                const int32_t *sel, int n, int n_pad,
                                                int64_t *rowsum) {
    __shared__ long long part[256];
    const int i = blockIdx.x, w = blockIdx.y;
    const int64_t row = ((int64_t) w * n_pad + sel[i]) * n_pad;
    long long s = 0;
    for (int j = threadIdx.x; j < n; j += 256) s += cross(C, C2, row + sel[j]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int) threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) rowsum[(int64_t) w * n + i] = part[0];
}

// K[w][i][j] = (C_ij - r_i/n - r_j/n + t/n^2) * scale[w], t = sum_i r_i.
__global__ void __launch_bounds__(256) k_centre(const int32_t *C, const int32_t *C2,
                                                const int32_t *sel, const int64_t *rowsum, int n,
                                                int n_pad, const double *scale, double *K) {
    const int64_t gid = (int64_t) blockIdx.x * 256 + threadIdx.x;
    const int w = blockIdx.y;
    if (gid >= (int64_t) n * n) return;
    const int i = (int) (gid / n), j = (int) (gid % n);
    const int64_t *rs = rowsum + (int64_t) w * n;
    // grand total, recomputed per thread from the row sums would be O(n);
    // it is passed through rowsum[n * W + w] by the host-side launcher
    const double t = (double) rowsum[(int64_t) gridDim.y * n + w];
    const double dn = (double) n;
    const double cij = (double) cross(C, C2, ((int64_t) w * n_pad + sel[i]) * n_pad + sel[j]);
    const double v = cij - (double) rs[i] / dn - (double) rs[j] / dn + t / (dn * dn);
    K[(int64_t) w * n * n + gid] = (C2 ? 0.5 * v : v) * scale[w];
}

// Site divergence of singleton sets from the same cross products: samples i
// and j differ at a 0/1 site iff g_i + g_j - 2 g_i g_j = 1, so the count of
// differing sites is C_ii + C_jj - 2 C_ij (0 on the diagonal), times the
// window's span scale (trees.c:8753-8900 with sets of one sample).
__global__ void __launch_bounds__(256) k_divergence(const int32_t *C, const int32_t *C2,
                                                    const int32_t *sel, int n, int n_pad,
                                                    const double *scale, double *D) {
    const int64_t gid = (int64_t) blockIdx.x * 256 + threadIdx.x;
    const int w = blockIdx.y;
    if (gid >= (int64_t) n * n) return;
    const int i = (int) (gid / n), j = (int) (gid % n);
    const int64_t base = (int64_t) w * n_pad * n_pad;
    const int64_t si = sel[i], sj = sel[j];
    double v = 0.0;
    if (i != j) {
        const int64_t pii = cross(C, C2, base + si * n_pad + si);
        const int64_t pjj = cross(C, C2, base + sj * n_pad + sj);
        const int64_t pij = cross(C, C2, base + si * n_pad + sj);
        // without C2: C_ii + C_jj - 2 C_ij; with it (P_ii + P_jj) / 2 - P_ij
        // (P_ii is even: a site's any-derived row and its allele rows count
        // each carrier once each)
        v = C2 ? (double) ((pii + pjj) / 2 - pij) : (double) (pii + pjj - 2 * pij);
    }
    D[(int64_t) w * n * n + gid] = v * scale[w];
}

// Divergence between sample sets from the pairwise matrix D [W x n x n] of
// their samples (set k = rows off[k] .. off[k + 1]): one workgroup per
// (A, B, window) sums D over A x B in a fixed order, then divides by the
// number of pairs, |A||B| (A != B) or |A|(|A| - 1) when that is not zero
// (trees.c:8876-8899; the reference counts a pair within one set twice, as
// D's two triangles do).
__global__ void __launch_bounds__(256) k_set_divergence(const double *D, int n,
                                                        const int32_t *off, int nsets,
                                                        double *out) {
    __shared__ double part[256];
    const int A = blockIdx.x / nsets, B = blockIdx.x % nsets;
    const int w = blockIdx.y;
    const int a0 = off[A], a1 = off[A + 1], b0 = off[B], b1 = off[B + 1];
    const int nb = b1 - b0;
    const int64_t cnt = (int64_t) (a1 - a0) * nb;
    const double *Dw = D + (int64_t) w * n * n;
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < cnt; t += 256) {
        const int i = a0 + (int) (t / nb), j = b0 + (int) (t % nb);
        s += Dw[(int64_t) i * n + j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int) threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double na = (double) (a1 - a0), nbd = (double) nb;
        const double denom = (A == B) ? na * (na - 1) : na * nbd;
        const double v = part[0];
        out[((int64_t) w * nsets + A) * nsets + B] = (A == B && denom == 0) ? v : v / denom;
    }
}

// grand totals: one workgroup per window sums its n row sums (exact int64,
// so the order of the additions does not matter)
__global__ void __launch_bounds__(256) k_total(int64_t *rowsum, int n, int W) {
    __shared__ long long part[256];
    const int w = blockIdx.x;
    long long s = 0;
    for (int i = threadIdx.x; i < n; i += 256) s += rowsum[(int64_t) w * n + i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int) threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) rowsum[(int64_t) W * n + w] = part[0];
}

// --- launchers ------------------------------------------------------------------
hipError_t tsg_launch_geno_bits(const TsgGenoArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_geno_bits, dim3(a.slots), dim3(WAVE), 0, s, a);
    return hipGetLastError();
}

template <bool GENO>
hipError_t launch_geno_hot(const TsgGenoArgs &a, size_t shmem, hipStream_t s) {
    if (shmem > 65536) {
        if ((int) shmem > tsg_lds_max()) return hipErrorInvalidValue;
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_geno_hot<GENO>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           tsg_lds_max());
        if (e != hipSuccess) return e;
    }
    const dim3 grid(8u * (unsigned) ((a.num_chunks + 7) / 8),
                    (unsigned) ((a.NW + a.WB - 1) / a.WB));
    hipLaunchKernelGGL((k_geno_hot<GENO>), grid, dim3(WAVE), shmem, s, a);
    return hipGetLastError();
}

hipError_t tsg_launch_geno_hot(const TsgGenoArgs &a, size_t shmem, hipStream_t s) {
    return launch_geno_hot<false>(a, shmem, s);
}

// Missing data: geno[site * n + sample] = -1 for each listed pair.
__global__ void k_mark_missing(int32_t *geno, int n, const int32_t *site, const int32_t *sample,
                               int64_t npairs) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npairs) geno[(int64_t) site[i] * n + sample[i]] = -1;
}

hipError_t tsg_launch_genotypes(const TsgGenoArgs &a, size_t shmem, const int32_t *miss_site,
                                const int32_t *miss_sample, int64_t nmiss, hipStream_t s) {
    hipError_t e = launch_geno_hot<true>(a, shmem, s);
    if (e != hipSuccess || nmiss == 0) return e;
    hipLaunchKernelGGL(k_mark_missing, dim3((unsigned) ((nmiss + 255) / 256)), dim3(256), 0, s,
                       a.geno, a.num_samples, miss_site, miss_sample, nmiss);
    return hipGetLastError();
}

hipError_t tsg_launch_bits_to_gt(const uint32_t *gbits, int NW, const int32_t *site_of_col,
                                 int64_t Kp, int n, int n_pad, int8_t *gt, hipStream_t s) {
    dim3 grid((unsigned) ((Kp + 63) / 64), (unsigned) (n_pad / 64));
    hipLaunchKernelGGL(k_bits_to_gt, grid, dim3(256), 0, s, gbits, NW, site_of_col, Kp, n, gt);
    return hipGetLastError();
}

hipError_t tsg_launch_syrk_i8(const int8_t *gt, int64_t Kp, int n_pad, const int64_t *win_k0,
                              const int64_t *win_k1, int W, int32_t *C, hipStream_t s) {
    const int tpd = n_pad / BT;
    dim3 grid((unsigned) (tpd * (tpd + 1) / 2), (unsigned) W);
    hipLaunchKernelGGL(k_syrk_i8, grid, dim3(256), 0, s, gt, Kp, n_pad, win_k0, win_k1, tpd, C);
    return hipGetLastError();
}

hipError_t tsg_launch_divergence(const int32_t *C, const int32_t *C2, const int32_t *sel, int n,
                                 int n_pad, int W, const double *scale, double *D,
                                 hipStream_t s) {
    const int64_t nn = (int64_t) n * n;
    hipLaunchKernelGGL(k_divergence, dim3((unsigned) ((nn + 255) / 256), (unsigned) W), dim3(256),
                       0, s, C, C2, sel, n, n_pad, scale, D);
    return hipGetLastError();
}

hipError_t tsg_launch_set_divergence(const double *D, int n, int W, const int32_t *off, int nsets,
                                     double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_set_divergence, dim3((unsigned) nsets * (unsigned) nsets, (unsigned) W),
                       dim3(256), 0, s, D, n, off, nsets, out);
    return hipGetLastError();
}

hipError_t tsg_launch_centre(const int32_t *C, const int32_t *C2, const int32_t *sel,
                             int64_t *rowsum, int n, int n_pad, int W, const double *scale,
                             double *K, hipStream_t s) {
    hipLaunchKernelGGL(k_rowsum, dim3((unsigned) n, (unsigned) W), dim3(256), 0, s, C, C2, sel, n,
                       n_pad, rowsum);
    hipLaunchKernelGGL(k_total, dim3((unsigned) W), dim3(256), 0, s, rowsum, n, W);
    const int64_t nn = (int64_t) n * n;
    hipLaunchKernelGGL(k_centre, dim3((unsigned) ((nn + 255) / 256), (unsigned) W), dim3(256), 0, s,
                       C, C2, sel, rowsum, n, n_pad, scale, K);
    return hipGetLastError();
}
