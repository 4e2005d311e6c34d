/*
 * tsgpu -- C ABI of the MI355X tree-sequence statistics library
 * (tskit_amd/libtsgpu.so).
 *
 * The statistics entry points take the reference's arguments in the
 * reference's order (c/tskit/trees.h:1098-1239) with the
 * `const tsk_treeseq_t *` replaced by an opaque, device-resident
 * `tsg_treeseq_t *`.  Types: sizes are uint64_t (tsk_size_t), ids int32_t
 * (tsk_id_t), options uint32_t (tsk_flags_t); option bits and error numbers
 * have the reference's values (c/tskit/trees.h:47-56, c/tskit/core.h).
 * Results are written to caller-owned buffers, C order, W x M doubles
 * (AFS: W x TW x prod(n_k + 1)): host memory, or device memory, which is
 * then filled on the device (every entry point but the branch-mode
 * divergence matrix, whose result is formed on the host).  Statistics calls
 * on one handle may run
 * concurrently from several host threads (each takes its own device
 * workspace and stream) and never mutate it; tsg_treeseq_set_num_chunks and
 * tsg_treeseq_free wait until the calls in flight on the handle have
 * returned, and calls that start meanwhile wait for them.
 */
#ifndef TSGPU_H
#define TSGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tsg_treeseq tsg_treeseq_t;
typedef struct tsg_tables tsg_tables_t;

/* option bits: c/tskit/trees.h:47-56 */
#define TSG_STAT_SITE (1u << 0)
#define TSG_STAT_BRANCH (1u << 1)
#define TSG_STAT_NODE (1u << 2)
#define TSG_STAT_POLARISED (1u << 10)
#define TSG_STAT_SPAN_NORMALISE (1u << 11)
#define TSG_STAT_ALLOW_TIME_UNCALIBRATED (1u << 12)
#define TSG_STAT_NONCENTRED (1u << 14)
/* tree-sequence option: time_units == "uncalibrated" (trees.c:537-543) */
#define TSG_TS_TIME_UNCALIBRATED (1u << 0)

/* Error message for an error number (replaces tsk_strerror, core.c). */
const char *tsg_strerror(int err);

/* ---- tree sequence lifecycle ------------------------------------------- */

/* Validates the tables, builds the edge indexes if I/O are NULL, finds the
 * trees and per-tree sites and builds the chunk plan (host side).  The
 * device copy is made on first use.  Replaces tsk_treeseq_init
 * (trees.c:455-545) with TSK_TS_INIT_BUILD_INDEXES and
 * tsk_table_collection_build_index (tables.c:11393-11461).  String columns
 * are (data, offsets) as in the reference tables; NULL states mean "0"
 * ancestral / "1" derived. */
int tsg_treeseq_init(tsg_treeseq_t **ts, double sequence_length, uint64_t num_nodes,
    const uint32_t *node_flags, const double *node_time, uint64_t num_edges,
    const double *edge_left, const double *edge_right, const int32_t *edge_parent,
    const int32_t *edge_child, uint64_t num_sites, const double *site_position,
    const char *ancestral_state, const uint64_t *ancestral_state_offset,
    uint64_t num_mutations, const int32_t *mutation_site, const int32_t *mutation_node,
    const int32_t *mutation_parent, const char *derived_state,
    const uint64_t *derived_state_offset, const int32_t *edge_insertion_order,
    const int32_t *edge_removal_order, uint32_t options);
/* Replaces tsk_treeseq_free (trees.c:107). */
int tsg_treeseq_free(tsg_treeseq_t *ts);
/* Copies the tree sequence to the current GPU now (otherwise on first use). */
int tsg_treeseq_upload(tsg_treeseq_t *ts);
/* Rebuilds the chunk plan with num_chunks chunks (autotuner hook). */
int tsg_treeseq_set_num_chunks(tsg_treeseq_t *ts, uint32_t num_chunks);

/* accessors: tsk_treeseq_get_num_trees / _samples / ... (trees.h:560-700) */
uint64_t tsg_treeseq_get_num_trees(const tsg_treeseq_t *ts);
uint64_t tsg_treeseq_get_num_samples(const tsg_treeseq_t *ts);
uint64_t tsg_treeseq_get_num_nodes(const tsg_treeseq_t *ts);
uint64_t tsg_treeseq_get_num_edges(const tsg_treeseq_t *ts);
uint64_t tsg_treeseq_get_num_sites(const tsg_treeseq_t *ts);
uint64_t tsg_treeseq_get_num_mutations(const tsg_treeseq_t *ts);
double tsg_treeseq_get_sequence_length(const tsg_treeseq_t *ts);
/* tsk_treeseq_get_breakpoints (trees.h:692): num_trees + 1 values */
int tsg_treeseq_get_breakpoints(const tsg_treeseq_t *ts, double *breakpoints);
/* tsk_treeseq_get_samples (trees.h:706) */
int tsg_treeseq_get_samples(const tsg_treeseq_t *ts, int32_t *samples);
/* indexes.edge_insertion_order / edge_removal_order (tables.h:560-601) */
int tsg_treeseq_get_index(const tsg_treeseq_t *ts, int32_t *insertion, int32_t *removal);
/* device-side facts: chunk count, device bytes held, largest chunk */
uint64_t tsg_treeseq_get_num_chunks(const tsg_treeseq_t *ts);
uint64_t tsg_treeseq_get_device_bytes(const tsg_treeseq_t *ts);
uint64_t tsg_treeseq_get_max_chunk_nodes(const tsg_treeseq_t *ts);
/* host copy of the chunk plan; sizes = {chunks, local nodes, events, trees,
 * sites, mutations} */
int tsg_treeseq_get_plan_sizes(const tsg_treeseq_t *ts, uint64_t *sizes);
int tsg_treeseq_copy_plan(const tsg_treeseq_t *ts, int32_t *chunk_tree0, int64_t *chunk_node0,
    int32_t *loc_parent0, double *loc_bl0, double *loc_time, double *loc_lastrem,
    int64_t *tr_off, int32_t *tr_nrem, int32_t *ev_child, int32_t *ev_parent, double *ev_bl,
    int32_t *tree_site0, int32_t *site_mut0, int32_t *mut_node_loc, int32_t *mut_allele,
    int32_t *mut_pallele, int32_t *site_nalleles);
/* tsk_treeseq_divergence_matrix (c/tskit/trees.h; trees.c:8902-9001).
 * Branch mode (trees.c:8579-8745): from the branch length G_ij every two
 * samples share (the relatedness vector of identity weights on the hot
 * plan), path length G_ii + G_jj - 2 G_ij, per pair of sets and
 * count-normalised as below.  Site mode for disjoint sample sets (sample_set_sizes NULL: one sample
 * each; sample_sets NULL: every sample): result = [W x K x K], per pair of
 * sets the number of (site, sample pair) with different alleles in the
 * window, over the pairs of the two sets, divided by the number of pairs
 * (trees.c:8876-8899) and, with TSG_STAT_SPAN_NORMALISE, by the window span;
 * from the same int8 cross products as the relatedness matrix: per window
 * P = 2 G1'G1 + G2'G2 (G1 the derived-allele rows of the sites with at most
 * two alleles, G2 the any-derived row and one row per derived allele of the
 * others), samples i and j differing at (P_ii + P_jj) / 2 - P_ij sites
 * (trees.c:8684-8826 groups the alleles the same way). */
int tsg_treeseq_divergence_matrix(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);
/* Genotypes of every sample at every site, out = [num_sites x num_samples]
 * int32 allele indexes (0 = ancestral, then derived states in order of first
 * appearance; -1 = missing: an isolated sample without a mutation on it,
 * unless options has TSG_ISOLATED_NOT_MISSING = 1 << 1).  Replaces the
 * per-site tsk_variant_decode loop behind TreeSequence.genotype_matrix
 * (c/tskit/genotypes.c:473-595, python/tskit/trees.py:5563-5640). */
int tsg_treeseq_genotype_matrix(tsg_treeseq_t *ts, uint32_t options, int32_t *out);
/* host copy of each event's root path (local ids, the edge's parent first);
 * ev_path0 has (events + 1) entries */
uint64_t tsg_treeseq_get_path_size(const tsg_treeseq_t *ts);
int tsg_treeseq_copy_paths(const tsg_treeseq_t *ts, int64_t *ev_path0, int32_t *path);
/* host copy of the hot plan (chunks keep only the nodes they touch, as LDS
 * slots); sizes = {usable, chunks, slots, level entries, events, path
 * entries, first-tree nodes, max slots per chunk, max levels per chunk,
 * last-tree nodes} */
int tsg_treeseq_get_hot_plan_sizes(const tsg_treeseq_t *ts, uint64_t *sizes);
int tsg_treeseq_copy_hot_plan(const tsg_treeseq_t *ts, int32_t *chunk_tree0,
    int64_t *chunk_slot0, int32_t *slot_par, double *slot_bl0, int64_t *chunk_lev0,
    int32_t *lev_slot0, uint16_t *sample_slot, int64_t *tr_off, int32_t *tr_nrem,
    int32_t *ev_child, double *ev_bl, int64_t *ev_path0, int32_t *path, int32_t *mut_slot,
    int32_t *t0_parent, double *t0_bl);
/* host copy of the hot plan's branch-AFS tables: per slot its node, the
 * samples it is the nearest hot ancestor-or-self of, the position of its
 * node's last flush by an event before the chunk and of the insertion of its
 * current parent edge; the last tree (samples first: parent, node time and
 * the same two positions per node) */
int tsg_treeseq_copy_hot_afs_plan(const tsg_treeseq_t *ts, int32_t *slot_node,
    int32_t *slot_nsd, double *slot_last0, double *slot_ins0, double *slot_tp0,
    int32_t *tl_parent, double *tl_time, double *tl_last, double *tl_ins);
/* plan = 1: run statistics sweeps on the full plan; 2: on the hot plan with
 * 32-bit counts; 0: the library chooses (the hot plan whenever its slots fit
 * in LDS, 16-bit counts when every set has <= 65535 samples) -- a testing
 * hook */
int tsg_treeseq_set_sweep_plan(tsg_treeseq_t *ts, int plan);
/* node store of the last statistics sweep: 0 HBM slabs, 1 LDS (full plan),
 * 2 LDS hot slots, 3 LDS hot slots with 16-bit counts, -1 before the first
 * sweep */
int tsg_treeseq_get_last_store(const tsg_treeseq_t *ts);
/* HIP-event timing of the sweep kernel on the library's stream */
int tsg_treeseq_set_profiling(tsg_treeseq_t *ts, int enable);
int tsg_treeseq_get_profile(const tsg_treeseq_t *ts, double *sweep_ms, uint64_t *launches);
/* the same split by mode: [0] branch sweeps, [1] site sweeps */
int tsg_treeseq_get_profile_modes(const tsg_treeseq_t *ts, double *ms, uint64_t *launches);

/* ---- statistics ----------------------------------------------------------
 * Every statistics call takes the arguments of its reference counterpart.
 * options: TSG_STAT_SITE / TSG_STAT_BRANCH / TSG_STAT_NODE (one of them;
 * none = site), TSG_STAT_POLARISED, TSG_STAT_SPAN_NORMALISE.  result is
 * [W x P], or [W x num_nodes x P] in node mode (trees.c:1788-1918, on the
 * hot plan; in a batch it runs on its own after the fused sweeps).
 * ---- one-way statistics ------------------------------------------------ */
/* tsk_treeseq_diversity (trees.h:1103, trees.c:3950-3958) */
int tsg_treeseq_diversity(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);
/* tsk_treeseq_segregating_sites (trees.h:1106, trees.c:4238-4246) */
int tsg_treeseq_segregating_sites(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);
/* tsk_treeseq_Y1 (trees.h:1109, trees.c:4266-4274) */
int tsg_treeseq_Y1(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);
/* Tajima's D, the Python composite of trees.py:9967-10022 (diversity and
 * segregating sites, not span normalised, one sweep, FP64 final). */
int tsg_treeseq_Tajimas_D(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);

/* ---- k-way statistics -------------------------------------------------- */
/* tsk_treeseq_divergence (trees.h:1189, trees.c:4711-4727) */
int tsg_treeseq_divergence(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
/* tsk_treeseq_Y2 (trees.h:1193, trees.c:4920-4936) */
int tsg_treeseq_Y2(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
/* tsk_treeseq_f2 (trees.h:1197, trees.c:4961-4977) */
int tsg_treeseq_f2(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
/* tsk_treeseq_Y3 (trees.h:1226, trees.c:5201-5217) */
int tsg_treeseq_Y3(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
/* tsk_treeseq_f3 (trees.h:1230, trees.c:5244-5260) */
int tsg_treeseq_f3(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
/* tsk_treeseq_f4 (trees.h:1236, trees.c:5293-5309) */
int tsg_treeseq_f4(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
/* tsk_treeseq_genetic_relatedness (trees.h:1201, trees.c:4775-4797);
 * TSG_STAT_NONCENTRED selects the non-centred summary. */
int tsg_treeseq_genetic_relatedness(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
/* genetic_relatedness(proportion=True), the Python composite of
 * trees.py:8830-8976: relatedness divided by the unpolarised segregating
 * sites of the union of the sample sets, one sweep, FP64 final. */
int tsg_treeseq_genetic_relatedness_proportion(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
/* Fst, the Python composite of trees.py:10024-10136: diversity of every set
 * and divergence of each pair in one sweep, FP64 final. */
int tsg_treeseq_Fst(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);

/* ---- relatedness vector -------------------------------------------------
 * tsk_treeseq_genetic_relatedness_vector (trees.h; trees.c:10402-10816,
 * entry 10783), same arguments in the same order (result before options):
 * weights [num_samples x num_weights] (rows in sample order), result
 * [num_windows x num_focal_nodes x num_weights] = per window the product of
 * the branch relatedness matrix and the weights, centred (weights over the
 * samples, result over the focal nodes) unless TSG_STAT_NONCENTRED.  Branch
 * mode only (TSG_STAT_SITE / TSG_STAT_NODE: TSG_ERR_UNSUPPORTED_STAT_MODE);
 * windows need not span the sequence.  Focal nodes must be samples here
 * (otherwise TSG_ERR_UNSUPPORTED_INPUT). */
int tsg_treeseq_genetic_relatedness_vector(tsg_treeseq_t *ts, uint64_t num_weights,
    const double *weights, uint64_t num_windows, const double *windows,
    uint64_t num_focal_nodes, const int32_t *focal_nodes, double *result, uint32_t options);

/* ---- general_stat with real-valued weights ----------------------------
 * tsk_treeseq_general_stat (trees.h:1035, trees.c:2036-2095) with the summary
 * function taken from a registry by id instead of a general_stat_func_t and
 * its f_params (trees.h:1032-1037): sample_weights [num_samples x state_dim]
 * (rows in sample order), result [num_windows x result_dim].  Branch or site
 * mode (TSG_STAT_NODE: TSG_ERR_UNSUPPORTED_STAT_MODE), TSG_STAT_POLARISED,
 * TSG_STAT_SPAN_NORMALISE; windows must span the sequence.  Summaries (and
 * their parameters, the reference's param structs):
 *   TSG_SUMMARY_TRAIT_COVARIANCE (11)   trees.c:3961-3974; result_dim =
 *       state_dim; x^2 / (2 (n - 1)^2), n = number of samples.
 *   TSG_SUMMARY_TRAIT_CORRELATION (12)  trees.c:4026-4045; result_dim =
 *       state_dim - 1, the last column the frequency.
 *   TSG_SUMMARY_TRAIT_LINEAR_MODEL (13) trees.c:4108-4152; state = [traits
 *       (result_dim) | covariates | 1]; dparams = V [result_dim x covariates].
 *   TSG_SUMMARY_RELATEDNESS_WEIGHTED (14) trees.c:4801-4820 and
 *   TSG_SUMMARY_RELATEDNESS_WEIGHTED_NONCENTRED (15) trees.c:4822-4837;
 *       index_tuples [result_dim x 2], the last column the frequency;
 *       dparams = the total weights [state_dim] (centred only).
 * State columns are FP64 in LDS on the hot plan: a tree sequence whose
 * largest chunk's slots do not fit returns TSG_ERR_UNSUPPORTED_INPUT.  Not
 * inside a batch (TSG_ERR_UNSUPPORTED_INPUT). */
#define TSG_SUMMARY_TRAIT_COVARIANCE 11
#define TSG_SUMMARY_TRAIT_CORRELATION 12
#define TSG_SUMMARY_TRAIT_LINEAR_MODEL 13
#define TSG_SUMMARY_RELATEDNESS_WEIGHTED 14
#define TSG_SUMMARY_RELATEDNESS_WEIGHTED_NONCENTRED 15
int tsg_treeseq_general_stat(tsg_treeseq_t *ts, uint64_t state_dim,
    const double *sample_weights, uint64_t result_dim, int32_t summary,
    const int32_t *index_tuples, const double *dparams, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);

/* tsk_treeseq_trait_covariance (trees.h; trees.c:3976-4022): the weights
 * centred on the host, then general_stat.  Same arguments and order. */
int tsg_treeseq_trait_covariance(tsg_treeseq_t *ts, uint64_t num_weights,
    const double *weights, uint64_t num_windows, const double *windows, uint32_t options,
    double *result);
/* tsk_treeseq_trait_correlation (trees.c:4047-4103): centred and scaled,
 * plus a column of 1/n. */
int tsg_treeseq_trait_correlation(tsg_treeseq_t *ts, uint64_t num_weights,
    const double *weights, uint64_t num_windows, const double *windows, uint32_t options,
    double *result);
/* tsk_treeseq_trait_linear_model (trees.c:4154-4220): V = W^T Z on the host,
 * state [W | Z | 1]; the covariates Z are taken as standardised. */
int tsg_treeseq_trait_linear_model(tsg_treeseq_t *ts, uint64_t num_weights,
    const double *weights, uint64_t num_covariates, const double *covariates,
    uint64_t num_windows, const double *windows, uint32_t options, double *result);
/* tsk_treeseq_genetic_relatedness_weighted (trees.c:4839-4895), result
 * before options as there; TSG_STAT_NONCENTRED selects the non-centred
 * summary. */
int tsg_treeseq_genetic_relatedness_weighted(tsg_treeseq_t *ts, uint64_t num_weights,
    const double *weights, uint64_t num_index_tuples, const int32_t *index_tuples,
    uint64_t num_windows, const double *windows, double *result, uint32_t options);

/* ---- allele frequency spectrum ----------------------------------------- */
/* tsk_treeseq_allele_frequency_spectrum (trees.h:1112, trees.c:3815-3928);
 * at most 16 sample sets. */
int tsg_treeseq_allele_frequency_spectrum(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint64_t num_time_windows, const double *time_windows,
    uint32_t options, double *result);

/* ---- relatedness matrix ---------------------------------------------------
 * genetic_relatedness_matrix(mode="site") of python/tskit/trees.py:8978-9102
 * (the divergence matrix of c/tskit/trees.c:8900-9001 double-centred) for
 * singleton sample sets (sample_sets NULL = every sample), computed as
 * K = J P J / 2 with P the int8 matrix-core cross product of the divergence
 * matrix above (exact int32; K = J G^T G J when every site has at most two
 * alleles) and FP64 centring.  Windows need not span
 * the genome (as the reference's divergence matrix).  result: W x n x n. */
int tsg_treeseq_genetic_relatedness_matrix(tsg_treeseq_t *ts, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);

/* ---- batched form (no counterpart in the reference) ---------------------
 * A batch collects queries on one tree sequence; tsg_batch_run executes them
 * as one fused sweep per mode (branch / site), sharing the path walks, then
 * its node-mode queries one by one (not fused).  Each
 * tsg_batch_<stat> takes the arguments of tsg_treeseq_<stat>, validates them
 * the same way (same error numbers) and records the query; `result` must
 * stay valid until tsg_batch_run returns. */
typedef struct tsg_batch tsg_batch_t;
int tsg_batch_init(tsg_batch_t **batch, tsg_treeseq_t *ts);
int tsg_batch_free(tsg_batch_t *batch);
int tsg_batch_clear(tsg_batch_t *batch);
uint64_t tsg_batch_get_num_queries(const tsg_batch_t *batch);
int tsg_batch_run(tsg_batch_t *batch);
/* Runs n batches (on one tree sequence or several, e.g. simulation
 * replicates) as one sweep launch per mode over all their trees, on one
 * stream; a batch whose outputs need several sweep slices, or that does not
 * fit the hot plan, launches on its own.  Results equal tsg_batch_run's on
 * each batch.  Returns when every result has been written. */
int tsg_batch_run_multi(tsg_batch_t **batches, uint64_t n);
/* tsg_batch_run without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc, hipFree, hipHostMalloc or other allocating device routine:
 * every device buffer of the run comes from a 64 MiB device array that each
 * GPU the library runs on holds until the process ends, and host arrays go
 * over from ordinary host memory.  The tree sequence must already be
 * resident on the current device (tsg_treeseq_upload; otherwise
 * TSG_ERR_NOT_RESIDENT = -30007), and its queries must run on the hot plan
 * (otherwise TSG_ERR_UNSUPPORTED_INPUT) and be of branch or site mode
 * (node mode: TSG_ERR_UNSUPPORTED_STAT_MODE); a batch needing more than the
 * 64 MiB returns TSG_ERR_SCRATCH_TOO_SMALL (-30006) with no result written.
 * Modes run one after the other on the null stream; results equal
 * tsg_batch_run's.  Calls on one GPU share its array: calls that overlap
 * on one GPU race, and the caller serialises them. */
int tsg_batch_run_noalloc(tsg_batch_t *batch);
int tsg_batch_diversity(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);
int tsg_batch_segregating_sites(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);
int tsg_batch_Y1(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);
int tsg_batch_Tajimas_D(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_windows,
    const double *windows, uint32_t options, double *result);
int tsg_batch_divergence(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
int tsg_batch_Y2(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
int tsg_batch_f2(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
int tsg_batch_Y3(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
int tsg_batch_f3(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
int tsg_batch_f4(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
int tsg_batch_genetic_relatedness(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
int tsg_batch_genetic_relatedness_proportion(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);
int tsg_batch_Fst(tsg_batch_t *batch, uint64_t num_sample_sets,
    const uint64_t *sample_set_sizes, const int32_t *sample_sets, uint64_t num_index_tuples,
    const int32_t *index_tuples, uint64_t num_windows, const double *windows,
    uint32_t options, double *result);

/* ---- synthetic inputs (no counterpart in the reference; msprime's role
 *      in its tests and benchmarks) --------------------------------------- */
/* Hudson coalescent with recombination, discrete genome, haploid samples. */
int tsg_sim_ancestry(uint32_t num_samples, double sequence_length, double population_size,
    double recombination_rate, uint64_t seed, tsg_tables_t **tables);
/* Infinite-sites 0/1 mutations at rate mu per unit genome per generation. */
int tsg_sim_mutations(tsg_tables_t *tables, double rate, uint64_t seed);
/* sizes = {nodes, edges, sites, mutations} */
int tsg_tables_get_sizes(const tsg_tables_t *tables, uint64_t *sizes);
int tsg_tables_copy_out(const tsg_tables_t *tables, uint32_t *node_flags, double *node_time,
    double *edge_left, double *edge_right, int32_t *edge_parent, int32_t *edge_child,
    double *site_position, int32_t *mutation_site, int32_t *mutation_node,
    int32_t *mutation_parent, double *mutation_time);
void tsg_tables_free(tsg_tables_t *tables);

#ifdef __cplusplus
}
#endif

#endif /* TSGPU_H */
