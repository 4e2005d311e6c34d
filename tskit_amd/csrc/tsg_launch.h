// Host-callable launchers for the kernels in tsg_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "tsg_kernels.h"

// the device's dynamic LDS limit per workgroup (set once, before launches)
void tsg_set_lds_max(int bytes);
int tsg_lds_max();
// one launch over the chunks of nrun sweeps (hot-plan stores only): args
// and cpre (nrun + 1 chunk prefix) in device memory
hipError_t tsg_launch_sweep_multi(const TsgSweepArgs *args, const int32_t *cpre, int nrun,
                                  int total_chunks, int mode, int store, int mpl, size_t shmem,
                                  hipStream_t s);
// store: TSG_STORE_GLOBAL / TSG_STORE_LDS / TSG_STORE_HOT
hipError_t tsg_launch_sweep(const TsgSweepArgs &a, int mode, int store, int mpl, size_t shmem,
                            hipStream_t s);
// 1 when some output among outs[0..m) (host memory) is a summary the branch
// sweep evaluates hop by hop, so that its LDS staging block must hold the
// per-hop arrays (TSG_STAGE_INTS(K, span, 1)); decided by the same code the
// kernel runs.  weighted: FP64 state columns with totals totw and summary
// parameters opar (host memory).
int tsg_sweep_needs_hop(const TsgOut *outs, int64_t m, int K, int weighted, const double *totw,
                        const double *opar);
hipError_t tsg_launch_afs(const TsgSweepArgs &a, bool branch, bool lds, size_t shmem,
                          hipStream_t s);
// the AFS on the hot plan: one wave per chunk, shmem bytes of LDS
hipError_t tsg_launch_afs_hot(const TsgSweepArgs &a, bool branch, size_t shmem, hipStream_t s);
// the branch AFS's flush of the intervals open at the sequence end L
hipError_t tsg_launch_afs_last_tree(const TsgSweepArgs &a, const TsgLastTree &lt, double L,
                                    hipStream_t s);
// number of per-block partials tsg_launch_tree0 writes (each a.M doubles)
int tsg_tree0_blocks(int nloc);
// counts of the first tree (t0_parent / t0_bl, samples first, nloc nodes)
// into x0 [nloc x K] and its running branch sum as per-block partials
// part [tsg_tree0_blocks(nloc) x M]
hipError_t tsg_launch_tree0(const TsgSweepArgs &a, const int32_t *t0_parent, const double *t0_bl,
                            int32_t *x0, int nloc, double *part, hipStream_t s);
// the same for FP64 state columns (real-valued weights), the first tree's
// sums x0 [nloc x K] given
hipError_t tsg_launch_tree0_w(const TsgSweepArgs &a, const int32_t *t0_parent,
                              const double *t0_bl, const double *x0, int nloc, double *part,
                              hipStream_t s);
// chunk blocks of the running-sum scan (bsum holds blocks x M doubles)
int tsg_scan_blocks(int num_chunks);
hipError_t tsg_launch_combine(const TsgCombineArgs &a, const double *dR, double *Rstart,
                              double *bsum, const double *r0part, int nb, int64_t total,
                              hipStream_t s);
hipError_t tsg_launch_finals(const void *fin, int nq, const double *raw, const int32_t *ipar,
                             const double *dpar, double *out, int64_t total, hipStream_t s);
hipError_t tsg_launch_fill_u64(unsigned long long *p, int64_t n, hipStream_t s);
hipError_t tsg_launch_fill_f64(double *p, int64_t n, hipStream_t s);
hipError_t tsg_launch_afs_site_final(const unsigned long long *h, double *out, const double *win,
                                     int W, int64_t afs_size, uint32_t flags, hipStream_t s);
// node statistics on the hot plan: the chunk sweep, the last tree's counts
// and the final flush of every node (res zeroed by the caller; lt.counts
// holds (lt.nl + 1) x K ints)
hipError_t tsg_launch_node(const TsgSweepArgs &a, const int32_t *slot_node, double *res,
                           int64_t N, size_t shmem, const TsgLastTree &lt, const double *node_last,
                           const int32_t *node_lid, double L, hipStream_t s);
hipError_t tsg_launch_afs_branch_final(double *h, const double *win, int W, int64_t per_window,
                                       uint32_t flags, hipStream_t s);
size_t tsg_final_struct_size();
hipError_t tsg_launch_geno_bits(const TsgGenoArgs &a, hipStream_t s);
// genotype bitsets on the hot plan (shmem = max slots x WB x 4 bytes)
hipError_t tsg_launch_geno_hot(const TsgGenoArgs &a, size_t shmem, hipStream_t s);
// genotype decode on the hot plan into a.geno [S x n] int32, then -1 at the
// listed (site, sample) pairs
hipError_t tsg_launch_genotypes(const TsgGenoArgs &a, size_t shmem, const int32_t *miss_site,
                                const int32_t *miss_sample, int64_t nmiss, hipStream_t s);
hipError_t tsg_launch_bits_to_gt(const uint32_t *gbits, int NW, const int32_t *site_of_col,
                                 int64_t Kp, int n, int n_pad, int8_t *gt, hipStream_t s);
hipError_t tsg_launch_syrk_i8(const int8_t *gt, int64_t Kp, int n_pad, const int64_t *win_k0,
                              const int64_t *win_k1, int W, int32_t *C, hipStream_t s);
// D = (C_ii + C_jj - 2 C_ij) * scale[w] over the selected samples
hipError_t tsg_launch_divergence(const int32_t *C, const int32_t *C2, const int32_t *sel, int n,
                                 int n_pad, int W, const double *scale, double *D,
                                 hipStream_t s);
// divergence between sample sets (rows off[k] .. off[k + 1] of the pairwise
// matrix D [W x n x n]), count-normalised: out [W x nsets x nsets]
hipError_t tsg_launch_set_divergence(const double *D, int n, int W, const int32_t *off, int nsets,
                                     double *out, hipStream_t s);
hipError_t tsg_launch_centre(const int32_t *C, const int32_t *C2, const int32_t *sel,
                             int64_t *rowsum, int n, int n_pad, int W, const double *scale,
                             double *K, hipStream_t s);

// Finish phases of many runs, one launch per phase: runs [nrun] and the
// per-phase block prefixes pre [TSG_FINISH_PHASES x (nrun + 1)] in device
// memory, totals[phase] = pre[phase][nrun] on the host.
void tsg_finish_blocks(const TsgFinishRun &run, int64_t *blocks);
hipError_t tsg_launch_finish_multi(const TsgFinishRun *runs, const int64_t *pre,
                                   const int64_t *totals, int nrun, hipStream_t s);

// Mirrors TsgFinal in tsg_kernels.hip.
struct TsgFinalHost {
    int32_t kind;
    int32_t Kq;
    int32_t P;
    int32_t W;
    int64_t rawbase, outbase;
    int64_t pbase;
    int32_t Mraw;
};
