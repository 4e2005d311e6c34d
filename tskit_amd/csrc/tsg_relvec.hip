// genetic_relatedness_vector in branch mode (c/tskit/trees.c:10402-10816) on
// the hot plan.
//
// For window w, focal sample f and weight column k the reference computes
//   y[w][f][k] = integral over the window of D_x(f),
//   D_x(f) = sum over the ancestors u of f at x (f included) of
//            bl_x(u) * Z_x(u, k),
// Z_x(u, k) the (centred) weights of the samples below u.  It runs one lazy
// push-down sweep along the genome.  Here the genome is cut into the hot
// plan's chunks.  Within chunk c (start a, end b) every node whose parent
// edge changes, and every node on an event's root path, is a "topological"
// hot slot; the start-tree parent of such a slot is again one (no static
// node lies between two of them), so for a focal sample f whose nearest
// topological hot ancestor-or-self is h:
//   D_x(f) = S(f) + D_x(h),   S(f) constant over the chunk,
//   integral over a segment s of the chunk of D_x(f)
//       = |s| * D_a(f) + I_s(h),   I_s(h) = integral over s of D_x(h) - D_a(h),
//   D_b(f) = D_a(f) + E(h),        E(h) = D_b(h) - D_a(h).
// k_relvec_chunk runs the reference's lazy push-down (trees.c:10553-10605)
// over the chunk's slots only, from the slots' start-tree weights, and
// writes E(h) and I_s(h) for every slot and window segment s of the chunk.
// D_0(f) of the first tree comes from k_relvec_tree0_*; then
// D_a(f) of chunk c is D_0(f) + sum of E over the chunks before c, and
// k_relvec_assemble accumulates |s| * D_a(f) + I_s(h) into y.  FP64
// throughout; the weight sums of a chunk's slots and the assembly over chunk
// ranges use FP64 atomic adds, so the last bits of a result may vary between
// runs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tsg_common.h"
#include "tsg_kernels.h"
#include "tsg_relvec.h"

#define WAVE 64

namespace {

__device__ __forceinline__ int relvec_chunk_of(int b, int C) {
    const int per = (C + 7) / 8;
    return (b % 8) * per + b / 8;
}

__device__ __forceinline__ unsigned map_slot(const TsgRelvecArgs &a, int c, int s) {
    return a.sample_slot[((int64_t) (c / TSG_SS_CB) * a.num_samples + s) * TSG_SS_CB +
                         c % TSG_SS_CB];
}

}  // namespace

// One wave per chunk.  LDS per slot: w[Kb], v[Kb], d0[Kb] (FP64), x and bl
// (FP64), parent (int32) and a topological-hot flag (int32).
__global__ void __launch_bounds__(WAVE) k_relvec_chunk(const TsgRelvecArgs a) {
    extern __shared__ double rsm[];
    const int c = relvec_chunk_of(blockIdx.x, a.num_chunks);
    if (c >= a.num_chunks) return;
    const int lane = threadIdx.x;
    const int Kb = a.Kb;
    const int64_t s0 = a.chunk_slot0[c];
    const int S = (int) (a.chunk_slot0[c + 1] - s0);
    double *w = rsm;
    double *v = w + (int64_t) S * Kb;
    double *d0 = v + (int64_t) S * Kb;
    double *xs = d0 + (int64_t) S * Kb;
    double *bl = xs + S;
    int32_t *par = reinterpret_cast<int32_t *>(bl + S);
    int32_t *top = par + S;
    const int t0 = a.chunk_tree0[c], t1 = a.chunk_tree0[c + 1];
    const double ca = a.bp[t0], cb = a.bp[t1];
    for (int i = lane; i < S * Kb; i += WAVE) {
        w[i] = 0.0;
        v[i] = 0.0;
    }
    for (int h = lane; h < S; h += WAVE) {
        xs[h] = ca;
        bl[h] = a.slot_bl0[s0 + h];
        par[h] = a.slot_par[s0 + h];
        top[h] = 0;
    }
    __syncthreads();
    // topological hot slots: the children and root paths of the chunk's
    // events (the closing transition included)
    const int tlast = min(t1, a.num_trees - 1);
    const int64_t e_begin = a.tr_off[t0 + 1 <= tlast ? t0 + 1 : tlast + 1];
    const int64_t e_end = a.tr_off[tlast + 1];
    for (int64_t e = e_begin + lane; e < e_end; e += WAVE) top[a.ev_child[e]] = 1;
    if (e_end > e_begin) {
        const int64_t p0 = a.ev_path0[e_begin], p1 = a.ev_path0[e_end];
        for (int64_t i = p0 + lane; i < p1; i += WAVE) top[a.path[i]] = 1;
    }
    // each sample's weights into its nearest hot ancestor-or-self
    if (a.ident != nullptr) {
        for (int k = lane; k < Kb; k += WAVE) {
            const unsigned h = map_slot(a, c, a.ident[a.k0 + k]);
            if (h != TSG_NO_SLOT) w[(int) h * Kb + k] += 1.0;
        }
    } else {
        for (int s = lane; s < a.num_samples; s += WAVE) {
            const unsigned h = map_slot(a, c, s);
            if (h == TSG_NO_SLOT) continue;
            for (int k = 0; k < Kb; k++) {
                atomicAdd(&w[(int) h * Kb + k], a.weights[(int64_t) s * a.K + a.k0 + k]);
            }
        }
    }
    __syncthreads();
    // then up the hot forest, deepest level first
    const int32_t *lev = a.lev_slot0 + a.chunk_lev0[c];
    const int nlev = (int) (a.chunk_lev0[c + 1] - a.chunk_lev0[c]) - 1;
    for (int l = 0; l + 1 < nlev; l++) {
        const int h0 = lev[l], h1 = lev[l + 1];
        for (int idx = lane; idx < (h1 - h0) * Kb; idx += WAVE) {
            const int h = h0 + idx / Kb, k = idx % Kb;
            const int p = par[h];
            if (p >= 0) atomicAdd(&w[p * Kb + k], w[h * Kb + k]);
        }
        __syncthreads();
    }
    // D_a over the hot forest, roots first: d0[h] = d0[parent] + bl * w
    for (int l = nlev - 1; l >= 0; l--) {
        const int h0 = lev[l], h1 = lev[l + 1];
        for (int idx = lane; idx < (h1 - h0) * Kb; idx += WAVE) {
            const int h = h0 + idx / Kb, k = idx % Kb;
            const int p = par[h];
            d0[h * Kb + k] = (p >= 0 ? d0[p * Kb + k] : 0.0) + bl[h] * w[h * Kb + k];
        }
        __syncthreads();
    }

    // the lazy sweep; lanes over weight columns
    const double *win = a.windows;
    const int W = a.num_windows;
    double pos = ca, seg0 = ca;
    // first window boundary after the chunk start
    int bi = 0;
    while (bi <= W && win[bi] <= ca) bi++;
    double *Iout = a.I + a.ioff[c];
    const int wf = a.wfirst[c];

    // v[u] += bl(u) (x - last(u)) w(u); last(u) = x, for every slot
    auto add_z_all = [&](double x) {
        for (int i = lane; i < S * Kb; i += WAVE) {
            const int h = i / Kb;
            v[i] += bl[h] * (x - xs[h]) * w[i];
        }
        __syncthreads();
        for (int h = lane; h < S; h += WAVE) xs[h] = x;
        __syncthreads();
    };
    // the segment [seg0, x) ends: path sums of v per slot, minus the
    // segment's span times D_a, stored when the segment lies in a window
    auto readout = [&](double x, int m) {
        add_z_all(x);
        if (m >= 0 && m < W && x > seg0) {
            double *dst = Iout + (int64_t) (m - wf) * S * Kb;
            for (int i = lane; i < S * Kb; i += WAVE) {
                const int h = i / Kb, k = i % Kb;
                double r = 0.0;
                // (a forest: at most S steps; the bound only guards the loop)
                for (int u = h, g = 0; u >= 0 && g < S; u = par[u], g++) r += v[u * Kb + k];
                dst[i] = r - (x - seg0) * d0[i];
            }
        }
        __syncthreads();
        for (int i = lane; i < S * Kb; i += WAVE) v[i] = 0.0;
        __syncthreads();
        seg0 = x;
    };
    // window boundaries in (pos, x]
    auto boundaries_to = [&](double x) {
        while (bi <= W && win[bi] <= x) {
            readout(win[bi], bi - 1);
            bi++;
        }
    };
    auto transition = [&](int t) {
        const double x = a.bp[t];
        const int64_t e0 = a.tr_off[t], e1 = a.tr_off[t + 1];
        const int64_t nrem = a.tr_nrem[t];
        for (int64_t e = e0; e < e1; e++) {
            const bool ins = (e - e0) >= nrem;
            const int u = a.ev_child[e];
            const int64_t p0 = a.ev_path0[e], p1 = a.ev_path0[e + 1];
            const double sg = ins ? 1.0 : -1.0;
            if (!ins) {
                // remove_edge: bring the child up to x, detach it
                for (int k = lane; k < Kb; k += WAVE) {
                    v[u * Kb + k] += bl[u] * (x - xs[u]) * w[u * Kb + k];
                }
                __syncthreads();
                if (lane == 0) {
                    xs[u] = x;
                    bl[u] = 0.0;
                    par[u] = -1;
                }
                __syncthreads();
            }
            // adjust_path_up along the event's root path (parent first)
            for (int64_t j = p0; j < p1; j++) {
                const int p = a.path[j];
                for (int k = lane; k < Kb; k += WAVE) {
                    double vp = v[p * Kb + k] + bl[p] * (x - xs[p]) * w[p * Kb + k];
                    v[p * Kb + k] = vp;
                    v[u * Kb + k] -= sg * vp;
                    w[p * Kb + k] += sg * w[u * Kb + k];
                }
                __syncthreads();
                if (lane == 0) xs[p] = x;
                __syncthreads();
            }
            if (ins && lane == 0) {
                xs[u] = x;
                par[u] = a.path[p0];
                bl[u] = a.ev_bl[e];
            }
            __syncthreads();
        }
    };

    for (int t = t0; t < t1; t++) {
        if (t > t0) {
            boundaries_to(a.bp[t]);
            transition(t);
        }
        pos = a.bp[t];
    }
    boundaries_to(cb);
    // the last segment of the chunk, when it lies in a window
    if (seg0 < cb) readout(cb, bi - 1);
    (void) pos;
    // the closing transition, then E = D_b - D_a per slot
    if (t1 < a.num_trees) transition(t1);
    double *Eout = a.E + s0 * Kb;
    for (int i = lane; i < S * Kb; i += WAVE) {
        const int h = i / Kb, k = i % Kb;
        double r = 0.0;
        for (int u = h, g = 0; u >= 0 && g < S; u = par[u], g++) r += bl[u] * w[u * Kb + k];
        Eout[i] = r - d0[i];
    }
    __syncthreads();
    // a slot that is hot only for a mutation takes the values of its nearest
    // topological hot ancestor (its static chain adds constants only), or 0
    for (int i = lane; i < S * Kb; i += WAVE) {
        const int h = i / Kb, k = i % Kb;
        if (top[h]) continue;
        int q = a.slot_par[s0 + h];
        for (int g = 0; q >= 0 && !top[q] && g < S; g++) q = a.slot_par[s0 + q];
        if (q >= 0 && !top[q]) q = -1;
        Eout[i] = q >= 0 ? Eout[q * Kb + k] : 0.0;
        for (int m = 0; m < a.nwin[c]; m++) {
            double *dst = Iout + (int64_t) m * S * Kb;
            dst[i] = q >= 0 ? dst[q * Kb + k] : 0.0;
        }
    }
}

// First tree (samples first, local id j = sample j): Z0 by each sample's
// walk up its root path (FP64 atomics), then D_0 of each focal sample by
// its walk.
__global__ void k_relvec_tree0_z(const TsgRelvecArgs a, const int32_t *t0_parent, double *z0) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (a.ident != nullptr) {  // one thread per column: its sample's walk
        if (i >= a.Kb) return;
        const int k = (int) i;
        for (int u = a.ident[a.k0 + k], g = 0; u >= 0 && g < a.nloc0; u = t0_parent[u], g++) {
            z0[(int64_t) u * a.Kb + k] += 1.0;
        }
        return;
    }
    if (i >= (int64_t) a.num_samples * a.Kb) return;
    const int s = (int) (i / a.Kb), k = (int) (i % a.Kb);
    const double ws = a.weights[(int64_t) s * a.K + a.k0 + k];
    for (int u = s, g = 0; u >= 0 && g < a.nloc0; u = t0_parent[u], g++) {
        atomicAdd(&z0[(int64_t) u * a.Kb + k], ws);
    }
}

__global__ void k_relvec_tree0_d(const TsgRelvecArgs a, const int32_t *t0_parent,
                                 const double *t0_bl, const double *z0, double *d0f) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t) a.nf * a.Kb) return;
    const int j = (int) (i / a.Kb), k = (int) (i % a.Kb);
    double r = 0.0;
    for (int u = a.focal[j], g = 0; u >= 0 && t0_parent[u] >= 0 && g < a.nloc0;
         u = t0_parent[u], g++) {
        r += t0_bl[u] * z0[(int64_t) u * a.Kb + k];
    }
    d0f[i] = r;
}

// Sum of E over the chunks of range r for focal j: thread (r, j).
__global__ void k_relvec_esum(const TsgRelvecArgs a, double *esum) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t) a.nranges * a.nf) return;
    const int r = (int) (i / a.nf), j = (int) (i % a.nf);
    const int c0 = (int) ((int64_t) a.num_chunks * r / a.nranges);
    const int c1 = (int) ((int64_t) a.num_chunks * (r + 1) / a.nranges);
    const int s = a.focal[j];
    double acc[TSG_RELVEC_KMAX];
    for (int k = 0; k < a.Kb; k++) acc[k] = 0.0;
    for (int c = c0; c < c1; c++) {
        const unsigned h = map_slot(a, c, s);
        if (h == TSG_NO_SLOT) continue;
        const double *e = a.E + (a.chunk_slot0[c] + (int64_t) h) * a.Kb;
        for (int k = 0; k < a.Kb; k++) acc[k] += e[k];
    }
    for (int k = 0; k < a.Kb; k++) esum[((int64_t) r * a.nf + j) * a.Kb + k] = acc[k];
}

// Thread (r, j): D_a(f) at the start of range r (D_0 plus the sums of the
// ranges before), then per chunk of the range and window segment of the
// chunk, |s| * D_a(f) + I_s(h) into y [W x nf x K] (FP64 atomic adds: a
// window may span several ranges).
__global__ void k_relvec_assemble(const TsgRelvecArgs a, const double *d0f, const double *esum,
                                  double *y) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t) a.nranges * a.nf) return;
    const int r = (int) (i / a.nf), j = (int) (i % a.nf);
    const int c0 = (int) ((int64_t) a.num_chunks * r / a.nranges);
    const int c1 = (int) ((int64_t) a.num_chunks * (r + 1) / a.nranges);
    const int s = a.focal[j];
    const int Kb = a.Kb;
    double P[TSG_RELVEC_KMAX], acc[TSG_RELVEC_KMAX];
    for (int k = 0; k < Kb; k++) P[k] = d0f[(int64_t) j * Kb + k];
    for (int q = 0; q < r; q++) {
        for (int k = 0; k < Kb; k++) P[k] += esum[((int64_t) q * a.nf + j) * Kb + k];
    }
    const double *win = a.windows;
    int cur = -1;  // window whose sum acc holds
    for (int k = 0; k < Kb; k++) acc[k] = 0.0;
    for (int c = c0; c < c1; c++) {
        const int nw = a.nwin[c];
        if (nw > 0) {
            const unsigned h = map_slot(a, c, s);
            const int S = (int) (a.chunk_slot0[c + 1] - a.chunk_slot0[c]);
            const double ca = a.bp[a.chunk_tree0[c]], cb = a.bp[a.chunk_tree0[c + 1]];
            for (int m = a.wfirst[c]; m < a.wfirst[c] + nw; m++) {
                const double sa = fmax(ca, win[m]), sb = fmin(cb, win[m + 1]);
                if (m != cur) {
                    if (cur >= 0) {
                        for (int k = 0; k < Kb; k++) {
                            atomicAdd(&y[((int64_t) cur * a.nf + j) * a.K + a.k0 + k], acc[k]);
                            acc[k] = 0.0;
                        }
                    }
                    cur = m;
                }
                const double *I = (h == TSG_NO_SLOT)
                                      ? nullptr
                                      : a.I + a.ioff[c] + ((int64_t) (m - a.wfirst[c]) * S + h) * Kb;
                for (int k = 0; k < Kb; k++) acc[k] += (sb - sa) * P[k] + (I ? I[k] : 0.0);
            }
        }
        const unsigned h = map_slot(a, c, s);
        if (h != TSG_NO_SLOT) {
            const double *e = a.E + (a.chunk_slot0[c] + (int64_t) h) * Kb;
            for (int k = 0; k < Kb; k++) P[k] += e[k];
        }
    }
    if (cur >= 0) {
        for (int k = 0; k < Kb; k++) atomicAdd(&y[((int64_t) cur * a.nf + j) * a.K + a.k0 + k], acc[k]);
    }
}

__global__ void k_relvec_fill(double *p, int64_t n) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Per (window, column): subtract the mean over the focal nodes (unless
// noncentred, trees.c:10676-10697), then divide by the window span (with
// span normalisation, trees.c:10763-10765).  One workgroup of 256 threads
// per (window, column); fixed-order reduction.
__global__ void __launch_bounds__(256) k_relvec_final(double *y, const double *win, int nf, int K,
                                                      int centre, int span_normalise) {
    __shared__ double red[256];
    const int m = blockIdx.x / K, k = blockIdx.x % K;
    double *row = y + (int64_t) m * nf * K + k;
    double mean = 0.0;
    if (centre && nf > 0) {
        double acc = 0.0;
        for (int j = threadIdx.x; j < nf; j += 256) acc += row[(int64_t) j * K];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int) threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        mean = red[0] / (double) nf;
    }
    const double scale = span_normalise ? 1.0 / (win[m + 1] - win[m]) : 1.0;
    for (int j = threadIdx.x; j < nf; j += 256) {
        row[(int64_t) j * K] = (row[(int64_t) j * K] - mean) * scale;
    }
}

hipError_t tsg_launch_relvec(const TsgRelvecArgs &a, size_t shmem, const int32_t *t0_parent,
                             const double *t0_bl, int nloc0, double *z0, double *d0f,
                             double *esum, double *y, hipStream_t s) {
    const int grid = ((a.num_chunks + 7) / 8) * 8;
    hipError_t err = hipFuncSetAttribute((const void *) k_relvec_chunk,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int) std::max<size_t>(shmem, 65536));
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_relvec_chunk, dim3((unsigned) grid), dim3(WAVE), shmem, s, a);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    const int64_t nz = (int64_t) nloc0 * a.Kb;
    hipLaunchKernelGGL(k_relvec_fill, dim3((unsigned) ((nz + 255) / 256)), dim3(256), 0, s, z0, nz);
    const int64_t n0 = (int64_t) a.num_samples * a.Kb;
    hipLaunchKernelGGL(k_relvec_tree0_z, dim3((unsigned) ((n0 + 255) / 256)), dim3(256), 0, s, a,
                       t0_parent, z0);
    const int64_t nd = (int64_t) a.nf * a.Kb;
    hipLaunchKernelGGL(k_relvec_tree0_d, dim3((unsigned) ((nd + 255) / 256)), dim3(256), 0, s, a,
                       t0_parent, t0_bl, z0, d0f);
    const int64_t nt = (int64_t) a.nranges * a.nf;
    hipLaunchKernelGGL(k_relvec_esum, dim3((unsigned) ((nt + 255) / 256)), dim3(256), 0, s, a, esum);
    hipLaunchKernelGGL(k_relvec_assemble, dim3((unsigned) ((nt + 255) / 256)), dim3(256), 0, s, a,
                       d0f, esum, y);
    return hipGetLastError();
}

hipError_t tsg_launch_relvec_fill(double *p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_relvec_fill, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t tsg_launch_relvec_final(double *y, const double *win, int W, int nf, int K, int centre,
                                   int span_normalise, hipStream_t s) {
    if (W <= 0 || K <= 0 || nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_relvec_final, dim3((unsigned) (W * K)), dim3(256), 0, s, y, win, nf, K,
                       centre, span_normalise);
    return hipGetLastError();
}
