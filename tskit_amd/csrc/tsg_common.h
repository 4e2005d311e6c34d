// Shared constants for the tsgpu host library and its HIP kernels.
//
// Option bits and error numbers carry the same values as the reference C
// library so that a caller switching from tskit sees identical codes:
//   option bits     c/tskit/trees.h:47-56
//   error numbers   c/tskit/core.h:255-759
#pragma once
#include <cstdint>

#define TSG_NULL (-1)
#define TSG_NODE_IS_SAMPLE 1u
// sample without a hot ancestor in a hot-plan chunk (16-bit slot ids)
#define TSG_NO_SLOT 0xFFFFu

// --- statistic option bits (trees.h:47-56) ---------------------------------
#define TSG_STAT_SITE (1u << 0)
#define TSG_STAT_BRANCH (1u << 1)
#define TSG_STAT_NODE (1u << 2)
#define TSG_STAT_POLARISED (1u << 10)
#define TSG_STAT_SPAN_NORMALISE (1u << 11)
#define TSG_STAT_ALLOW_TIME_UNCALIBRATED (1u << 12)
#define TSG_STAT_NONCENTRED (1u << 14)

// --- tree-sequence creation options ---------------------------------------
// Marks the node times as uncalibrated (time_units == "uncalibrated" in the
// reference, trees.c:537-543); branch statistics then need
// TSG_STAT_ALLOW_TIME_UNCALIBRATED.
#define TSG_TS_TIME_UNCALIBRATED (1u << 0)
// genotype decode: isolated samples take the ancestral allele, not -1
// (TSK_ISOLATED_NOT_MISSING, genotypes.h:35)
#define TSG_ISOLATED_NOT_MISSING (1u << 1)

// --- error numbers (core.h) -------------------------------------------------
#define TSG_ERR_GENERIC (-1)
#define TSG_ERR_NO_MEMORY (-2)
#define TSG_ERR_BAD_PARAM_VALUE (-4)
#define TSG_ERR_NODE_OUT_OF_BOUNDS (-202)
#define TSG_ERR_EDGE_OUT_OF_BOUNDS (-203)
#define TSG_ERR_SITE_OUT_OF_BOUNDS (-205)
#define TSG_ERR_MUTATION_OUT_OF_BOUNDS (-206)
#define TSG_ERR_TIME_NONFINITE (-210)
#define TSG_ERR_GENOME_COORDS_NONFINITE (-211)
#define TSG_ERR_NULL_PARENT (-300)
#define TSG_ERR_NULL_CHILD (-301)
#define TSG_ERR_EDGES_NOT_SORTED_PARENT_TIME (-302)
#define TSG_ERR_EDGES_NONCONTIGUOUS_PARENTS (-303)
#define TSG_ERR_EDGES_NOT_SORTED_CHILD (-304)
#define TSG_ERR_EDGES_NOT_SORTED_LEFT (-305)
#define TSG_ERR_BAD_NODE_TIME_ORDERING (-306)
#define TSG_ERR_BAD_EDGE_INTERVAL (-307)
#define TSG_ERR_RIGHT_GREATER_SEQ_LENGTH (-309)
#define TSG_ERR_LEFT_LESS_ZERO (-310)
#define TSG_ERR_BAD_EDGES_CONTRADICTORY_CHILDREN (-311)
#define TSG_ERR_UNSORTED_SITES (-400)
#define TSG_ERR_DUPLICATE_SITE_POSITION (-401)
#define TSG_ERR_BAD_SITE_POSITION (-402)
#define TSG_ERR_MUTATION_PARENT_DIFFERENT_SITE (-500)
#define TSG_ERR_MUTATION_PARENT_EQUAL (-501)
#define TSG_ERR_MUTATION_PARENT_AFTER_CHILD (-502)
#define TSG_ERR_UNSORTED_MUTATIONS (-504)
#define TSG_ERR_DUPLICATE_SAMPLE (-600)
#define TSG_ERR_BAD_SAMPLES (-601)
#define TSG_ERR_BAD_SEQUENCE_LENGTH (-701)
#define TSG_ERR_BAD_NUM_WINDOWS (-900)
#define TSG_ERR_BAD_WINDOWS (-901)
#define TSG_ERR_MULTIPLE_STAT_MODES (-902)
#define TSG_ERR_BAD_STATE_DIMS (-903)
#define TSG_ERR_BAD_RESULT_DIMS (-904)
#define TSG_ERR_INSUFFICIENT_SAMPLE_SETS (-905)
#define TSG_ERR_INSUFFICIENT_INDEX_TUPLES (-906)
#define TSG_ERR_BAD_SAMPLE_SET_INDEX (-907)
#define TSG_ERR_EMPTY_SAMPLE_SET (-908)
#define TSG_ERR_UNSUPPORTED_STAT_MODE (-909)
#define TSG_ERR_TIME_UNCALIBRATED (-910)
#define TSG_ERR_STAT_POLARISED_UNSUPPORTED (-911)
#define TSG_ERR_STAT_SPAN_NORMALISE_UNSUPPORTED (-912)
#define TSG_ERR_INSUFFICIENT_WEIGHTS (-913)
#define TSG_ERR_BAD_TIME_WINDOWS_DIM (-922)
#define TSG_ERR_BAD_TIME_WINDOWS (-924)

// Errors with no counterpart in the reference: device-side failures.
#define TSG_ERR_DEVICE (-30001)          // a HIP runtime call failed
#define TSG_ERR_NO_DEVICE (-30002)       // no usable GPU
#define TSG_ERR_DEVICE_MEMORY (-30003)   // device allocation failed / over budget
#define TSG_ERR_UNSUPPORTED_INPUT (-30004)  // input shape this path does not take
#define TSG_ERR_SCRATCH_TOO_SMALL (-30006)  // no-allocation run: scratch used up
#define TSG_ERR_NOT_RESIDENT (-30007)       // no-allocation run: handle not uploaded

// --- summary-function registry ----------------------------------------------
// One id per reference summary function (trees.c:3934-5291).  Each stat
// block in a query names one of these plus its index tuples.
enum tsg_stat_id {
    TSG_STAT_DIVERSITY = 0,       // trees.c:3934-3948
    TSG_STAT_SEGREGATING_SITES,   // trees.c:4221-4236
    TSG_STAT_Y1,                  // trees.c:4248-4264
    TSG_STAT_DIVERGENCE,          // trees.c:4690-4709   (pairs)
    TSG_STAT_RELATEDNESS,         // trees.c:4729-4753   (pairs, centred)
    TSG_STAT_RELATEDNESS_NONCENTRED,  // trees.c:4755-4773 (pairs)
    TSG_STAT_Y2,                  // trees.c:4899-4918   (pairs)
    TSG_STAT_F2,                  // trees.c:4938-4959   (pairs)
    TSG_STAT_Y3,                  // trees.c:5177-5199   (triples)
    TSG_STAT_F3,                  // trees.c:5219-5242   (triples)
    TSG_STAT_F4,                  // trees.c:5266-5291   (quads)
    // summaries of real-valued sample weights (tsg_treeseq_general_stat)
    TSG_STAT_TRAIT_COVARIANCE,    // trees.c:3961-3974
    TSG_STAT_TRAIT_CORRELATION,   // trees.c:4026-4045
    TSG_STAT_TRAIT_LINEAR_MODEL,  // trees.c:4108-4152
    TSG_STAT_RELATEDNESS_WEIGHTED,             // trees.c:4801-4820 (pairs)
    TSG_STAT_RELATEDNESS_WEIGHTED_NONCENTRED,  // trees.c:4822-4837 (pairs)
    TSG_NUM_STAT_IDS
};

// Final (FP64) post-processing applied after the window combine.  These are
// the Python composites of the reference (trees.py:10024-10136, 9967-10022,
// 8958-8974), evaluated on the device in FP64.
enum tsg_final_id {
    TSG_FINAL_NONE = 0,
    TSG_FINAL_FST = 1,
    TSG_FINAL_TAJIMAS_D = 2,
    TSG_FINAL_PROPORTION = 3,
};
