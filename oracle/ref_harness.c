/*
 * ORACLE -- test infrastructure only.  A thin C harness over the reference
 * tskit C library, compiled together with the reference sources (as they lie
 * under the reference checkout) into oracle/_ref/libtskit_ref.so by
 * oracle/Makefile.  It turns plain column arrays into a reference
 * tsk_table_collection_t / tsk_treeseq_t and forwards to the reference
 * statistics functions, so tests and the CPU baseline can run the real
 * reference code on the same inputs as the GPU path.
 *
 * Reference calls used: tsk_table_collection_init / *_set_columns
 * (c/tskit/tables.h:1627,1974,2696,3084,3864), build_index (tables.c:11393),
 * tsk_treeseq_init (trees.c:455), tsk_treeseq_<stat> (trees.h:1098-1239).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <tskit.h>

typedef struct {
    tsk_table_collection_t tables;
    tsk_treeseq_t ts;
} ref_ts_t;

int ref_ts_init(ref_ts_t **out, double L, uint64_t N, const uint32_t *flags,
    const double *time, uint64_t E, const double *left, const double *right,
    const int32_t *parent, const int32_t *child, uint64_t S, const double *pos,
    const char *anc, const uint64_t *anc_off, uint64_t Mu, const int32_t *mut_site,
    const int32_t *mut_node, const int32_t *mut_parent, const char *der,
    const uint64_t *der_off, int time_uncalibrated) {
    ref_ts_t *r = calloc(1, sizeof(ref_ts_t));
    int ret;
    *out = NULL;
    ret = tsk_table_collection_init(&r->tables, 0);
    if (ret != 0) goto fail;
    r->tables.sequence_length = L;
    ret = tsk_node_table_set_columns(&r->tables.nodes, N, flags, time, NULL, NULL, NULL, NULL);
    if (ret != 0) goto fail;
    ret = tsk_edge_table_set_columns(&r->tables.edges, E, left, right, parent, child, NULL, NULL);
    if (ret != 0) goto fail;
    ret = tsk_site_table_set_columns(&r->tables.sites, S, pos, anc, anc_off, NULL, NULL);
    if (ret != 0) goto fail;
    ret = tsk_mutation_table_set_columns(&r->tables.mutations, Mu, mut_site, mut_node,
        mut_parent, NULL, der, der_off, NULL, NULL);
    if (ret != 0) goto fail;
    if (time_uncalibrated) {
        ret = tsk_table_collection_set_time_units(&r->tables, TSK_TIME_UNITS_UNCALIBRATED,
            strlen(TSK_TIME_UNITS_UNCALIBRATED));
        if (ret != 0) goto fail;
    }
    ret = tsk_table_collection_build_index(&r->tables, 0);
    if (ret != 0) goto fail;
    ret = tsk_treeseq_init(&r->ts, &r->tables, 0);
    if (ret != 0) {
        tsk_table_collection_free(&r->tables);
        free(r);
        return ret;
    }
    *out = r;
    return 0;
fail:
    tsk_table_collection_free(&r->tables);
    free(r);
    return ret;
}

void ref_ts_free(ref_ts_t *r) {
    if (r == NULL) return;
    tsk_treeseq_free(&r->ts);
    tsk_table_collection_free(&r->tables);
    free(r);
}

uint64_t ref_ts_num_trees(const ref_ts_t *r) { return tsk_treeseq_get_num_trees(&r->ts); }

void ref_ts_index(const ref_ts_t *r, int32_t *I, int32_t *O) {
    const tsk_size_t E = r->ts.tables->edges.num_rows;
    memcpy(I, r->ts.tables->indexes.edge_insertion_order, E * sizeof(int32_t));
    memcpy(O, r->ts.tables->indexes.edge_removal_order, E * sizeof(int32_t));
}

void ref_ts_breakpoints(const ref_ts_t *r, double *b) {
    const double *bp = tsk_treeseq_get_breakpoints(&r->ts);
    memcpy(b, bp, (tsk_treeseq_get_num_trees(&r->ts) + 1) * sizeof(double));
}

/* stat ids match the build's summary-function registry */
int ref_sample_count_stat(const ref_ts_t *r, int stat, uint64_t K, const uint64_t *sizes,
    const int32_t *sets, uint64_t M, const int32_t *idx, uint64_t nw,
    const double *windows, uint32_t options, double *result) {
    const tsk_treeseq_t *ts = &r->ts;
    switch (stat) {
    case 0:
        return tsk_treeseq_diversity(ts, K, sizes, sets, nw, windows, options, result);
    case 1:
        return tsk_treeseq_segregating_sites(ts, K, sizes, sets, nw, windows, options, result);
    case 2:
        return tsk_treeseq_Y1(ts, K, sizes, sets, nw, windows, options, result);
    case 3:
        return tsk_treeseq_divergence(ts, K, sizes, sets, M, idx, nw, windows, options, result);
    case 4:
        return tsk_treeseq_genetic_relatedness(
            ts, K, sizes, sets, M, idx, nw, windows, options, result);
    case 5:
        return tsk_treeseq_genetic_relatedness(
            ts, K, sizes, sets, M, idx, nw, windows, options | TSK_STAT_NONCENTRED, result);
    case 6:
        return tsk_treeseq_Y2(ts, K, sizes, sets, M, idx, nw, windows, options, result);
    case 7:
        return tsk_treeseq_f2(ts, K, sizes, sets, M, idx, nw, windows, options, result);
    case 8:
        return tsk_treeseq_Y3(ts, K, sizes, sets, M, idx, nw, windows, options, result);
    case 9:
        return tsk_treeseq_f3(ts, K, sizes, sets, M, idx, nw, windows, options, result);
    case 10:
        return tsk_treeseq_f4(ts, K, sizes, sets, M, idx, nw, windows, options, result);
    default:
        return TSK_ERR_BAD_PARAM_VALUE;
    }
}

int ref_afs(const ref_ts_t *r, uint64_t K, const uint64_t *sizes, const int32_t *sets,
    uint64_t nw, const double *windows, uint64_t ntw, const double *time_windows,
    uint32_t options, double *result) {
    return tsk_treeseq_allele_frequency_spectrum(
        &r->ts, K, sizes, sets, nw, windows, ntw, time_windows, options, result);
}

/* Site-mode divergence matrix (trees.c:8753-8826); used to cross-check the
 * relatedness-matrix closed form on sub-blocks. */
/* A tree sequence from a .trees file through the reference's own loader
 * (tsk_table_collection_load, tables.c:11800ff), indexed as for ref_ts_init. */
int ref_ts_load(ref_ts_t **out, const char *path) {
    ref_ts_t *r = calloc(1, sizeof(ref_ts_t));
    int ret;
    *out = NULL;
    ret = tsk_table_collection_load(&r->tables, path, 0);
    if (ret != 0) goto fail;
    ret = tsk_table_collection_build_index(&r->tables, 0);
    if (ret != 0) goto fail;
    ret = tsk_treeseq_init(&r->ts, &r->tables, 0);
    if (ret != 0) goto fail;
    *out = r;
    return 0;
fail:
    tsk_table_collection_free(&r->tables);
    free(r);
    return ret;
}

/* Genotypes of every sample at every site through the reference's variant
 * decoder (tsk_variant_init / tsk_variant_decode, genotypes.h:120,157). */
int ref_genotype_matrix(const ref_ts_t *r, int isolated_as_missing, int32_t *out) {
    tsk_variant_t var;
    const tsk_size_t n = tsk_treeseq_get_num_samples(&r->ts);
    const tsk_size_t S = tsk_treeseq_get_num_sites(&r->ts);
    tsk_size_t s;
    int ret = tsk_variant_init(&var, &r->ts, NULL, 0, NULL,
        isolated_as_missing ? 0 : TSK_ISOLATED_NOT_MISSING);
    if (ret != 0) {
        tsk_variant_free(&var);
        return ret;
    }
    for (s = 0; s < S; s++) {
        ret = tsk_variant_decode(&var, (tsk_id_t) s, 0);
        if (ret != 0) break;
        memcpy(out + s * n, var.genotypes, n * sizeof(int32_t));
    }
    tsk_variant_free(&var);
    return ret;
}

int ref_divergence_matrix(const ref_ts_t *r, uint64_t num_samples, const int32_t *samples,
    uint64_t nw, const double *windows, uint32_t options, double *result) {
    return tsk_treeseq_divergence_matrix(&r->ts, num_samples, NULL, samples, nw, windows,
        options, result);
}

int ref_divergence_matrix_sets(const ref_ts_t *r, uint64_t num_sets, const uint64_t *sizes,
    const int32_t *sets, uint64_t nw, const double *windows, uint32_t options,
    double *result) {
    return tsk_treeseq_divergence_matrix(&r->ts, num_sets, sizes, sets, nw, windows, options,
        result);
}

/* tsk_treeseq_genetic_relatedness_vector (trees.c:10783): weights [n x K],
 * result [nw x nf x K] */
int ref_relatedness_vector(const ref_ts_t *r, uint64_t K, const double *weights, uint64_t nw,
    const double *windows, uint64_t nf, const int32_t *focal, uint32_t options,
    double *result) {
    return tsk_treeseq_genetic_relatedness_vector(
        &r->ts, K, weights, nw, windows, nf, focal, result, options);
}

/* The reference's weighted statistics: which = 0 trait_covariance
 * (trees.c:3976), 1 trait_correlation (4047), 2 trait_linear_model (4154,
 * covariates Z [n x nc]), 3 genetic_relatedness_weighted (4839, idx
 * [nidx x 2]); weights [n x K]. */
int ref_weighted_stat(const ref_ts_t *r, int which, uint64_t K, const double *weights,
    uint64_t nc, const double *Z, uint64_t nidx, const int32_t *idx, uint64_t nw,
    const double *windows, uint32_t options, double *result) {
    switch (which) {
    case 0:
        return tsk_treeseq_trait_covariance(&r->ts, K, weights, nw, windows, options, result);
    case 1:
        return tsk_treeseq_trait_correlation(&r->ts, K, weights, nw, windows, options, result);
    case 2:
        return tsk_treeseq_trait_linear_model(
            &r->ts, K, weights, nc, Z, nw, windows, options, result);
    default:
        return tsk_treeseq_genetic_relatedness_weighted(
            &r->ts, K, weights, nidx, idx, nw, windows, result, options);
    }
}
