// Arguments and launchers of the genetic_relatedness_vector kernels
// (tsg_relvec.hip).  Plain structs of pointers and sizes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

// weight columns one launch handles (register arrays of the assembly)
#define TSG_RELVEC_KMAX 16

struct TsgRelvecArgs {
    // hot plan
    const int32_t *chunk_tree0;
    const int64_t *chunk_slot0;
    const int32_t *slot_par;
    const double *slot_bl0;
    const int64_t *chunk_lev0;
    const int32_t *lev_slot0;
    const uint16_t *sample_slot;  // blocked by TSG_SS_CB chunks, as the sweep's
    const int64_t *tr_off;
    const int32_t *tr_nrem;
    const int32_t *ev_child;
    const double *ev_bl;
    const int64_t *ev_path0;
    const int32_t *path;
    const double *bp;
    int32_t num_trees, num_chunks, num_samples;
    // weights [num_samples x K] (centred on the host unless noncentred); this
    // launch handles columns k0 .. k0 + Kb - 1
    const double *weights;
    int32_t K, k0, Kb;
    // identity weights instead (null: the dense weights): column k weighs 1
    // at sample ident[k] and 0 elsewhere
    const int32_t *ident;
    // windows [num_windows + 1]; per chunk the first window it overlaps and
    // how many, and the offset of its segment sums in I
    const double *windows;
    int32_t num_windows;
    const int32_t *wfirst, *nwin;
    const int64_t *ioff;
    // focal samples (sample indexes), and the chunk ranges of the assembly
    const int32_t *focal;
    int32_t nf, nranges;
    int32_t nloc0;                // nodes of the first tree (t0 arrays)
    // per slot [slots x Kb]: E = D_b - D_a; per (chunk, window segment, slot)
    // [.. x Kb]: I = integral of D_x - D_a over the segment
    double *E;
    double *I;
};

hipError_t tsg_launch_relvec(const TsgRelvecArgs &a, size_t shmem, const int32_t *t0_parent,
                             const double *t0_bl, int nloc0, double *z0, double *d0f,
                             double *esum, double *y, hipStream_t s);
hipError_t tsg_launch_relvec_fill(double *p, int64_t n, hipStream_t s);
hipError_t tsg_launch_relvec_final(double *y, const double *win, int W, int nf, int K, int centre,
                                   int span_normalise, hipStream_t s);
// LDS bytes of the chunk kernel for S slots and Kb columns
inline size_t tsg_relvec_shmem(int S, int Kb) {
    return (size_t) S * ((size_t) 8 * (3 * (size_t) Kb + 2) + 8);
}
