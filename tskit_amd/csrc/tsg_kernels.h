// Device-side argument blocks shared by the HIP kernels and their host
// launchers.  Plain structs of pointers and sizes; no torch types.
#pragma once
#include <cstdint>

// One output of a fused query: a reference summary function applied to the
// state columns named by i..l (columns of the fused state vector).
struct TsgOut {
    int32_t stat;        // tsg_stat_id
    int32_t q;           // query this output belongs to
    int32_t i, j, k, l;  // state columns
    int32_t ni, nj, nk, nl;  // sample-set sizes of those columns
    int32_t cbase, ccount;   // the query's own columns (relatedness mean)
    int32_t flags;       // bit 0: polarised
    int32_t par;         // weighted summaries: offset of their parameters in opar
    double inv;          // 1 / denominator of the summary function
};

// One query of a batch: its windows and where its outputs and partial sums go.
struct TsgQuery {
    int32_t W;           // number of windows
    int32_t M;           // number of raw outputs
    int32_t mbase;       // first output index in the batch
    int32_t cbase;       // first state column
    uint32_t flags;      // TSG_F_SPAN_NORMALISE
    int32_t lanes;       // combine: lanes per (window, output), a power of 2 <= 64
    int64_t win_off;     // offset of its windows in the concatenated array
    int64_t pbase;       // offset (in doubles) of its partial-sum block
};

// Chunks per block of the device sample map: one sample's slots in 64
// consecutive chunks share a 128-byte line, which the chunks' workgroups --
// placed on one XCD in chunk order -- read through that XCD's L2.
#ifndef TSG_SS_CB
#define TSG_SS_CB 64
#endif

// Hops of an event's root path staged in LDS at once by the sweep kernel.
#define TSG_PSEG 16

// 32-bit words of one TsgOut.
#define TSG_OUT_INTS ((int) (sizeof(TsgOut) / 4))
// Trees, events and path entries the sweep kernel stages in LDS per batch
// (small: the sweep's occupancy is set by its LDS per workgroup).
#define TSG_TB 16
#define TSG_EB 32
#define TSG_PB 192
// 32-bit words of the sweep kernel's LDS staging block for K state columns
// and `span` outputs in the launch (layout in k_sweep); node records follow.
// (an even count: the node records that follow may hold FP64 values)
// `hop`: the block also holds the per-hop counts and terms of summaries
// evaluated hop by hop (branch sweeps with such outputs only).
#define TSG_STAGE_INTS(K, span, hop)                                                        \
    (((int64_t) (span) + 4 * (TSG_TB + 1) + 2 * (TSG_EB + 1) + TSG_TB + (TSG_TB + 1) +      \
      3 * TSG_EB + TSG_PB + 4 * (int64_t) (K) + 8 * (int64_t) (K) + 3 +                    \
      ((hop) ? 2 * TSG_PSEG * (int64_t) (K) + 2 * TSG_PSEG + 2 * 64 : 0)) & ~(int64_t) 1)
// ... with FP64 state columns: the moved counts, the counts before (and,
// with `hop`, after) each path hop, the column totals and the zero state take
// two words each
#define TSG_STAGE_INTS_W(K, span, hop) \
    (TSG_STAGE_INTS(K, span, hop) + 4 * (int64_t) (K) + ((hop) ? 2 * TSG_PSEG * (int64_t) (K) : 0))

// Where a sweep keeps its node records (see k_sweep).
#define TSG_STORE_GLOBAL 0
#define TSG_STORE_LDS 1
#define TSG_STORE_HOT 2
#define TSG_STORE_HOT16 3
// the hot plan with FP64 state columns: real-valued sample weights
#define TSG_STORE_HOTW 4
// 32-bit words per hot slot: FP64 branch length (branch sweeps, bl) + K
// counts (16-bit when c16)
#define TSG_SLOT_WORDS(K, c16, bl) (((bl) ? 2 : 0) + ((c16) ? ((K) + 1) / 2 : (K)))
// ... and with FP64 state (TSG_STORE_HOTW): the branch length, K doubles
#define TSG_SLOT_WORDS_W(K, bl) (((bl) ? 2 : 0) + 2 * (K))
// 32-bit words per full-plan node record: parent, FP64 branch length, K counts
#define TSG_NODE_WORDS(K) (3 + (K))

#define TSG_F_POLARISED 1u
#define TSG_F_NOSKIP 2u       // a summary has a zero denominator: never skip walks
#define TSG_F_SPAN_NORMALISE 4u
#define TSG_F_ANY_UNPOLARISED 8u  // some output sums over the ancestral allele

struct TsgSweepArgs {
    // tree sequence program
    const int32_t *chunk_tree0;
    const int64_t *chunk_node0;
    const int32_t *loc_parent0;
    const double *loc_bl0;
    const double *loc_time;
    const double *loc_lastrem;   // last removal position before the chunk start
    const int64_t *tr_off;
    const int32_t *tr_nrem;
    const int32_t *ev_child, *ev_parent;
    const double *ev_bl;
    const double *bp;
    const int32_t *tree_site0;
    const double *site_pos;
    const int32_t *site_mut0, *mut_node, *mut_allele, *mut_pallele, *site_nalleles;
    int32_t num_trees, num_chunks, num_samples;
    // batch
    int32_t K, M, Q, nseed;
    int32_t m_begin, m_end;      // outputs handled by this launch
    // weighted samples: seed_sample[i] belongs to the columns
    // seed_col[seed_off[i] .. seed_off[i+1])
    const int32_t *seed_sample, *seed_off, *seed_col;
    // the same weights as a flat (sample, column) list, by sample
    const int32_t *pair_sample, *pair_col;
    int32_t npairs;
    const int64_t *ev_path0;     // per event: offset of its root path
    const int32_t *path;
    // hot plan only (node ids are slots; chunk_node0 holds slot ranges and
    // loc_bl0 the slots' start-tree branch lengths)
    const int32_t *slot_par;     // nearest hot proper ancestor, or -1
    const int64_t *chunk_lev0;   // [C+1] into lev_slot0
    const int32_t *lev_slot0;    // per chunk: level boundaries, deepest first
    // per block of TSG_SS_CB chunks, per sample, the block's chunks: entry
    // ((c / CB) * num_samples + s) * CB + c % CB
    const uint16_t *sample_slot;
    // the start-tree counts of the hot slots, shared by a batch's two runs
    // (branch, then site) with the same (sample, column) pairs: the first
    // writes its seeded counts (per slot its count words, at the chunk's
    // first slot x count words), the second reads them instead of seeding;
    // null otherwise
    int32_t *seed_out;
    const int32_t *seed_in;
    const int32_t *col_total;    // [K] total weight (set size) per column
    // real-valued weights (TSG_STORE_HOTW): each (sample, column) pair's
    // weight, the columns' total weights, the weighted summaries' parameters
    const double *pair_w;
    const double *col_total_w;
    const double *opar;
    const TsgOut *outs;          // [M]
    const TsgQuery *queries;     // [Q]
    const double *windows;       // concatenated
    const int32_t *cq_wfirst;    // [C*Q] first window touched by chunk c
    const int32_t *cq_nwin;      // [C*Q] number of windows touched
    const int64_t *cq_poff;      // [C*Q] window-slot offset within query block
    double *partial;             // per (query, chunk, window) partial sums
    double *dR;                  // [C*M] branch running-sum change over chunk
    uint32_t flags;
    // 1: no output of the batch is evaluated hop by hop, and the LDS staging
    // block leaves out the per-hop arrays (TSG_STAGE_INTS(K, span, 0))
    int32_t no_hop;
    // global-memory node records (when they do not fit in LDS)
    int32_t *g_rec;
    int64_t g_stride;            // ints per chunk slab
    double *g_last;              // AFS branch: last update position per node
    int64_t g_last_stride;
    // AFS
    int32_t afs_K;               // number of AFS sample sets (state has afs_K+1 cols)
    int64_t afs_size;            // prod(n_k + 1)
    const int64_t *afs_dims;     // [afs_K]
    int32_t TW;
    const double *time_windows;  // [TW+1]
    unsigned long long *afs_hist_u;  // site AFS: half-unit counts
    double *afs_hist_d;          // branch AFS
    // AFS on the hot plan: windows, and per slot the samples it is the
    // nearest hot ancestor-or-self of, its node's time, its parent's time at
    // the chunk start (-inf: none), the position of the node's last flush
    // before the chunk and of its parent edge's insertion
    int32_t afs_W;
    const int32_t *slot_nsd;
    const double *slot_tu, *slot_tp0, *slot_last0, *slot_ins0;
};

// The last tree of the sequence, samples first, for the branch AFS's final
// flush of the intervals still open at its end.
struct TsgLastTree {
    const int32_t *parent;
    const double *time, *last, *ins;
    int32_t nl;
    int32_t *counts;             // [nl x (afs_K + 1)] scratch
};

// Genotype bitset sweep (relatedness matrix route).
struct TsgGenoArgs {
    const int32_t *chunk_tree0;
    const int64_t *chunk_node0;
    const int32_t *loc_parent0;
    const int64_t *tr_off;
    const int32_t *tr_nrem;
    const int32_t *ev_child, *ev_parent;
    const int32_t *tree_site0;
    const int32_t *site_mut0, *mut_node;
    int32_t num_chunks, num_samples;
    int32_t NW;                  // 32-bit words per bitset
    int32_t slots;               // resident workgroups (slabs); they loop over chunks
    uint32_t *g_rec;             // per slot: max_nloc x (1 + NW) words
    int64_t g_stride;
    uint32_t *gbits;             // [S x NW] derived-allele bitsets per site
    // hot plan (k_geno_hot): slot ids in events, paths and mutations;
    // chunk_node0 holds slot ranges; WB bitset words per workgroup
    const int64_t *ev_path0;
    const int32_t *path;
    const int32_t *slot_par;
    const int64_t *chunk_lev0;
    const int32_t *lev_slot0;
    const uint16_t *sample_slot;   // blocked, as TsgSweepArgs::sample_slot
    int32_t WB;
    // genotype decode: allele index per mutation, [S x num_samples] output
    const int32_t *mut_allele;
    int32_t *geno;
    // bit rows of sites with three or more alleles (the divergence and
    // relatedness matrices): such a site's row in gbits is its any-derived
    // row, and rows site_row2[s] .. + nalleles - 2 of gbits hold the rows of
    // its derived alleles 1, 2, ...; -1 for sites with at most two alleles
    // (null: every site has at most two)
    const int32_t *site_row2;
    const int32_t *site_nalleles;
};

struct TsgCombineArgs {
    const double *bp;
    const int32_t *chunk_tree0;
    int32_t num_chunks;
    int32_t Q, M;
    const TsgQuery *queries;
    const double *windows;
    const int32_t *cq_wfirst, *cq_nwin;
    const int64_t *cq_poff;
    const double *partial;
    const double *Rstart;        // [C*M] (branch) or null
    double *raw;                 // per query [W x M_q], offset rawbase[q]
    const int64_t *rawbase;
    const int64_t *lanebase;     // per query: first combine lane (multiples of 64)
    uint32_t flags;
};

// One run's share of a finish phase launched over many runs at once
// (tsg_launch_finish_multi): the first tree's running sum (branch), the
// chunk scan and window combine, the finals.
struct TsgFinishRun {
    TsgSweepArgs a;
    TsgCombineArgs ca;
    const int32_t *t0_parent;
    const double *t0_bl;
    int32_t *x0;
    double *r0part, *bsum, *Rstart;
    const double *dR;         // null: site mode (no running sums)
    int32_t nloc, nb, nblk;
    int64_t lanecount;        // combine lanes
    const void *fin;          // TsgFinal per query
    int32_t nq;
    const double *raw;
    const int32_t *ipar;
    const double *dpar;
    double *out;
    int64_t outcount;
};

// Block-count prefixes of the phases of tsg_launch_finish_multi, per run
// (each [nrun + 1]): fill, tree0 counts, tree0 sums, scan blocks, scan top,
// scan apply, combine, finals.
#define TSG_FINISH_PHASES 8
