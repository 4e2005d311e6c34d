/*
 * ORACLE -- test infrastructure only.  A plain-C, FP64, single-threaded
 * restatement of the reference library's statistics path, used by the tests,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (tskit_amd/) never links, loads or calls anything here.
 *
 * Each function follows the reference file:line it cites.  It is pinned by
 *   - the reference C compiled from source (oracle/_ref, see Makefile) on
 *     generated tree sequences (tests/test_oracle.py),
 *   - golden vectors the reference wrote (tests/golden/: the small workload's
 *     statistics and AFS, and node statistics of node_stat),
 *   - the reference's own hand-computed fixtures (tests/golden/),
 *   - brute-force pairwise definitions on tiny trees (tests/test_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NUL (-1)

enum {
    ST_DIVERSITY = 0, ST_SEGSITES, ST_Y1, ST_DIVERGENCE, ST_RELATEDNESS,
    ST_RELATEDNESS_NC, ST_Y2, ST_F2, ST_Y3, ST_F3, ST_F4,
    /* summaries of real-valued weights (oracle_general_stat) */
    ST_TRAIT_COV, ST_TRAIT_CORR, ST_TRAIT_LM, ST_REL_W, ST_REL_W_NC
};
#define OPT_SITE (1u << 0)
#define OPT_BRANCH (1u << 1)
#define OPT_NODE (1u << 2)
#define OPT_POLARISED (1u << 10)
#define OPT_SPAN_NORMALISE (1u << 11)
#define OPT_ALLOW_UNCAL (1u << 12)

typedef struct {
    double L;
    int32_t N, E, S, Mu, n, T;
    const double *time;
    const double *left, *right;
    const int32_t *parent, *child;
    int32_t *I, *O;
    const double *pos;
    int32_t *site_mut0;  /* [S+1] */
    int32_t *mut_allele, *mut_pallele, *site_nalleles;
    const int32_t *mut_node;
    int32_t *samples, *sample_index_map;
    double *breakpoints; /* [T+1] */
    int32_t *tree_site0; /* [T+1] */
    int time_uncalibrated;
} ots_t;

/* ---- index build: tables.c:10286-10310, 11393-11461 -------------------- */
typedef struct {
    int32_t index;
    double first, second;
    int32_t third, fourth;
} okey_t;

static int cmp_key(const void *a, const void *b) {
    const okey_t *x = a, *y = b;
    if (x->first != y->first) return (x->first > y->first) - (x->first < y->first);
    if (x->second != y->second) return (x->second > y->second) - (x->second < y->second);
    if (x->third != y->third) return (x->third > y->third) - (x->third < y->third);
    return (x->fourth > y->fourth) - (x->fourth < y->fourth);
}

static void build_index(ots_t *ts) {
    okey_t *k = malloc((size_t) (ts->E > 0 ? ts->E : 1) * sizeof(okey_t));
    int32_t j;
    for (j = 0; j < ts->E; j++) {
        k[j].index = j;
        k[j].first = ts->left[j];
        k[j].second = ts->time[ts->parent[j]];
        k[j].third = ts->parent[j];
        k[j].fourth = ts->child[j];
    }
    qsort(k, (size_t) ts->E, sizeof(okey_t), cmp_key);
    for (j = 0; j < ts->E; j++) ts->I[j] = k[j].index;
    for (j = 0; j < ts->E; j++) {
        k[j].index = j;
        k[j].first = ts->right[j];
        k[j].second = -ts->time[ts->parent[j]];
        k[j].third = -ts->parent[j];
        k[j].fourth = -ts->child[j];
    }
    qsort(k, (size_t) ts->E, sizeof(okey_t), cmp_key);
    for (j = 0; j < ts->E; j++) ts->O[j] = k[j].index;
    free(k);
}

static int same_str(const char *a, uint64_t la, const char *b, uint64_t lb) {
    return la == lb && memcmp(a, b, la) == 0;
}

/* ---- tree sequence init: trees.c:404-453 (samples), 243-359 (trees),
 *      128-179 (site mutations); allele order as get_allele_weights
 *      trees.c:1525-1612 (first seen) ---------------------------------- */
int oracle_ts_init(ots_t **out, double L, int32_t N, const uint32_t *flags,
    const double *time, int32_t E, const double *left, const double *right,
    const int32_t *parent, const int32_t *child, int32_t S, const double *pos,
    const char *anc, const uint64_t *anc_off, int32_t Mu, const int32_t *mut_site,
    const int32_t *mut_node, const int32_t *mut_parent, const char *der,
    const uint64_t *der_off, int time_uncalibrated) {
    ots_t *ts = calloc(1, sizeof(ots_t));
    int32_t j, k, t, s, m, cap;
    double tl, tr;
    const char **al;
    uint64_t *all;
    ts->L = L;
    ts->N = N;
    ts->E = E;
    ts->S = S;
    ts->Mu = Mu;
    ts->time = time;
    ts->left = left;
    ts->right = right;
    ts->parent = parent;
    ts->child = child;
    ts->pos = pos;
    ts->mut_node = mut_node;
    ts->time_uncalibrated = time_uncalibrated;
    ts->I = malloc((size_t) (E + 1) * sizeof(int32_t));
    ts->O = malloc((size_t) (E + 1) * sizeof(int32_t));
    build_index(ts);
    ts->sample_index_map = malloc((size_t) (N + 1) * sizeof(int32_t));
    ts->samples = malloc((size_t) (N + 1) * sizeof(int32_t));
    ts->n = 0;
    for (j = 0; j < N; j++) {
        ts->sample_index_map[j] = NUL;
        if (flags[j] & 1u) {
            ts->sample_index_map[j] = ts->n;
            ts->samples[ts->n++] = j;
        }
    }
    /* trees */
    cap = 16;
    ts->breakpoints = malloc((size_t) cap * sizeof(double));
    ts->tree_site0 = malloc((size_t) cap * sizeof(int32_t));
    j = 0;
    k = 0;
    s = 0;
    t = 0;
    tl = 0;
    while (j < E || tl < L) {
        if (t + 2 >= cap) {
            cap *= 2;
            ts->breakpoints = realloc(ts->breakpoints, (size_t) cap * sizeof(double));
            ts->tree_site0 = realloc(ts->tree_site0, (size_t) cap * sizeof(int32_t));
        }
        ts->breakpoints[t] = tl;
        ts->tree_site0[t] = s;
        while (k < E && right[ts->O[k]] == tl) k++;
        while (j < E && left[ts->I[j]] == tl) j++;
        tr = L;
        if (j < E && left[ts->I[j]] < tr) tr = left[ts->I[j]];
        if (k < E && right[ts->O[k]] < tr) tr = right[ts->O[k]];
        while (s < S && pos[s] < tr) s++;
        tl = tr;
        t++;
    }
    ts->T = t;
    ts->breakpoints[t] = L;
    ts->tree_site0[t] = s;
    /* alleles */
    ts->site_mut0 = calloc((size_t) S + 1, sizeof(int32_t));
    for (m = 0; m < Mu; m++) ts->site_mut0[mut_site[m] + 1]++;
    for (s = 0; s < S; s++) ts->site_mut0[s + 1] += ts->site_mut0[s];
    ts->mut_allele = malloc((size_t) (Mu + 1) * sizeof(int32_t));
    ts->mut_pallele = malloc((size_t) (Mu + 1) * sizeof(int32_t));
    ts->site_nalleles = malloc((size_t) (S + 1) * sizeof(int32_t));
    al = malloc((size_t) (Mu + 2) * sizeof(char *));
    all = malloc((size_t) (Mu + 2) * sizeof(uint64_t));
    for (s = 0; s < S; s++) {
        int32_t na = 1, a;
        al[0] = anc + anc_off[s];
        all[0] = anc_off[s + 1] - anc_off[s];
        for (m = ts->site_mut0[s]; m < ts->site_mut0[s + 1]; m++) {
            const char *d = der + der_off[m];
            uint64_t dl = der_off[m + 1] - der_off[m];
            for (a = 0; a < na; a++) {
                if (same_str(d, dl, al[a], all[a])) break;
            }
            if (a == na) {
                al[na] = d;
                all[na] = dl;
                na++;
            }
            ts->mut_allele[m] = a;
            if (mut_parent[m] == NUL) {
                ts->mut_pallele[m] = 0;
            } else {
                const int32_t pm = mut_parent[m];
                const char *pd = der + der_off[pm];
                uint64_t pl = der_off[pm + 1] - der_off[pm];
                for (a = 0; a < na; a++) {
                    if (same_str(pd, pl, al[a], all[a])) break;
                }
                ts->mut_pallele[m] = a;
            }
        }
        ts->site_nalleles[s] = na;
    }
    free(al);
    free(all);
    *out = ts;
    return 0;
}

void oracle_ts_free(ots_t *ts) {
    if (ts == NULL) return;
    free(ts->I);
    free(ts->O);
    free(ts->samples);
    free(ts->sample_index_map);
    free(ts->breakpoints);
    free(ts->tree_site0);
    free(ts->site_mut0);
    free(ts->mut_allele);
    free(ts->mut_pallele);
    free(ts->site_nalleles);
    free(ts);
}

int32_t oracle_ts_num_trees(const ots_t *ts) { return ts->T; }
int32_t oracle_ts_num_samples(const ots_t *ts) { return ts->n; }
void oracle_ts_index(const ots_t *ts, int32_t *I, int32_t *O) {
    memcpy(I, ts->I, (size_t) ts->E * sizeof(int32_t));
    memcpy(O, ts->O, (size_t) ts->E * sizeof(int32_t));
}
void oracle_ts_breakpoints(const ots_t *ts, double *b) {
    memcpy(b, ts->breakpoints, (size_t) (ts->T + 1) * sizeof(double));
}

/* ---- summary functions: trees.c:3934-5291 ------------------------------ */
typedef struct {
    int stat;
    int32_t K, M, tuple;
    const double *n;      /* sample set sizes as doubles */
    const int32_t *idx;   /* [M x tuple] */
    /* weighted summaries: number of samples, FP64 parameters (V rows of the
     * linear model, total weights of weighted relatedness) */
    double nsamp;
    const double *dpar;
} sf_t;

static void summary(const sf_t *p, const double *x, double *y) {
    int32_t m, i, j, k, l;
    double ni, nj, nk, nl, sumx, meanx;
    switch (p->stat) {
    case ST_DIVERSITY: /* trees.c:3934-3948 */
        for (m = 0; m < p->K; m++) {
            ni = p->n[m];
            y[m] = x[m] * (ni - x[m]) / (ni * (ni - 1));
        }
        break;
    case ST_SEGSITES: /* trees.c:4221-4236 */
        for (m = 0; m < p->K; m++) {
            ni = p->n[m];
            y[m] = (x[m] > 0) * (1 - x[m] / ni);
        }
        break;
    case ST_Y1: /* trees.c:4248-4264 */
        for (m = 0; m < p->K; m++) {
            ni = p->n[m];
            y[m] = x[m] * (ni - x[m]) * (ni - x[m] - 1) / (ni * (ni - 1) * (ni - 2));
        }
        break;
    case ST_DIVERGENCE: /* trees.c:4690-4709 */
        for (m = 0; m < p->M; m++) {
            i = p->idx[2 * m];
            j = p->idx[2 * m + 1];
            ni = p->n[i];
            nj = p->n[j];
            y[m] = x[i] * (nj - x[j]) / (ni * (nj - (i == j)));
        }
        break;
    case ST_RELATEDNESS: /* trees.c:4729-4753 */
        sumx = 0;
        for (k = 0; k < p->K; k++) sumx += x[k] / p->n[k];
        meanx = sumx / (double) p->K;
        for (m = 0; m < p->M; m++) {
            i = p->idx[2 * m];
            j = p->idx[2 * m + 1];
            y[m] = (x[i] / p->n[i] - meanx) * (x[j] / p->n[j] - meanx);
        }
        break;
    case ST_RELATEDNESS_NC: /* trees.c:4755-4773 */
        for (m = 0; m < p->M; m++) {
            i = p->idx[2 * m];
            j = p->idx[2 * m + 1];
            y[m] = x[i] * x[j] / (p->n[i] * p->n[j]);
        }
        break;
    case ST_Y2: /* trees.c:4899-4918 */
        for (m = 0; m < p->M; m++) {
            i = p->idx[2 * m];
            j = p->idx[2 * m + 1];
            ni = p->n[i];
            nj = p->n[j];
            y[m] = x[i] * (nj - x[j]) * (nj - x[j] - 1) / (ni * nj * (nj - 1));
        }
        break;
    case ST_F2: /* trees.c:4938-4959 */
        for (m = 0; m < p->M; m++) {
            i = p->idx[2 * m];
            j = p->idx[2 * m + 1];
            ni = p->n[i];
            nj = p->n[j];
            y[m] = (x[i] * (x[i] - 1) * (nj - x[j]) * (nj - x[j] - 1)
                       - x[i] * (ni - x[i]) * (nj - x[j]) * x[j])
                   / (ni * (ni - 1) * nj * (nj - 1));
        }
        break;
    case ST_Y3: /* trees.c:5177-5199 */
        for (m = 0; m < p->M; m++) {
            i = p->idx[3 * m];
            j = p->idx[3 * m + 1];
            k = p->idx[3 * m + 2];
            y[m] = x[i] * (p->n[j] - x[j]) * (p->n[k] - x[k]) / (p->n[i] * p->n[j] * p->n[k]);
        }
        break;
    case ST_F3: /* trees.c:5219-5242 */
        for (m = 0; m < p->M; m++) {
            i = p->idx[3 * m];
            j = p->idx[3 * m + 1];
            k = p->idx[3 * m + 2];
            ni = p->n[i];
            nj = p->n[j];
            nk = p->n[k];
            y[m] = (x[i] * (x[i] - 1) * (nj - x[j]) * (nk - x[k])
                       - x[i] * (ni - x[i]) * (nj - x[j]) * x[k])
                   / (ni * (ni - 1) * nj * nk);
        }
        break;
    case ST_TRAIT_COV: /* trees.c:3961-3974 */
        for (m = 0; m < p->K; m++) {
            y[m] = (x[m] * x[m]) / (2 * (p->nsamp - 1) * (p->nsamp - 1));
        }
        break;
    case ST_TRAIT_CORR: { /* trees.c:4026-4045 */
        const double q = x[p->K - 1];
        for (m = 0; m < p->K - 1; m++) {
            y[m] = (q > 0.0 && q < 1.0)
                       ? (x[m] * x[m]) / (2 * (q * (1 - q)) * p->nsamp * (p->nsamp - 1))
                       : 0.0;
        }
        break;
    }
    case ST_TRAIT_LM: { /* trees.c:4108-4152 */
        const int32_t nc = p->K - p->M - 1;
        const double mm = x[p->K - 1];
        for (i = 0; i < p->M; i++) {
            y[i] = 0.0;
            if (mm > 0.0 && mm < p->nsamp) {
                const double *v = p->dpar + (size_t) i * nc;
                double a = x[i], denom = mm, z;
                for (j = 0; j < nc; j++) {
                    z = x[p->M + j];
                    a -= z * v[j];
                    denom -= z * z;
                }
                y[i] = (denom < 1e-8) ? 0.0 : (a * a) / (2 * denom * denom);
            }
        }
        break;
    }
    case ST_REL_W: { /* trees.c:4801-4820 */
        const double pn = x[p->K - 1];
        for (m = 0; m < p->M; m++) {
            i = p->idx[2 * m];
            j = p->idx[2 * m + 1];
            y[m] = (x[i] - p->dpar[i] * pn) * (x[j] - p->dpar[j] * pn);
        }
        break;
    }
    case ST_REL_W_NC: /* trees.c:4822-4837 */
        for (m = 0; m < p->M; m++) {
            i = p->idx[2 * m];
            j = p->idx[2 * m + 1];
            y[m] = x[i] * x[j];
        }
        break;
    case ST_F4: /* trees.c:5266-5291 */
        for (m = 0; m < p->M; m++) {
            i = p->idx[4 * m];
            j = p->idx[4 * m + 1];
            k = p->idx[4 * m + 2];
            l = p->idx[4 * m + 3];
            ni = p->n[i];
            nj = p->n[j];
            nk = p->n[k];
            nl = p->n[l];
            y[m] = (x[i] * x[k] * (nj - x[j]) * (nl - x[l])
                       - x[i] * x[l] * (nj - x[j]) * (nk - x[k]))
                   / (ni * nj * nk * nl);
        }
        break;
    }
}

/* unpolarised wrapper: trees.c:1944-1972 */
typedef struct {
    const sf_t *p;
    int polarised;
    const double *total;
    double *tmp_x, *tmp_y;
} wf_t;

static void wsummary(const wf_t *w, const double *x, double *y) {
    int32_t k;
    summary(w->p, x, y);
    if (!w->polarised) {
        for (k = 0; k < w->p->K; k++) w->tmp_x[k] = w->total[k] - x[k];
        summary(w->p, w->tmp_x, w->tmp_y);
        for (k = 0; k < w->p->M; k++) y[k] += w->tmp_y[k];
    }
}

/* ---- branch sweep: trees.c:1352-1523 ------------------------------------ */
static int branch_stat(const ots_t *ts, const double *W, const wf_t *f, int32_t nw,
    const double *windows, double *result) {
    const int32_t N = ts->N, E = ts->E, K = f->p->K, M = f->p->M;
    int32_t *par = malloc((size_t) (N + 1) * sizeof(int32_t));
    double *bl = calloc((size_t) N + 1, sizeof(double));
    double *X = calloc((size_t) N * K + 1, sizeof(double));
    double *Y = calloc((size_t) N * M + 1, sizeof(double));
    double *rs = calloc((size_t) M + 1, sizeof(double));
    double *z = calloc((size_t) K + 1, sizeof(double));
    int32_t j, k, m, u, v, tj = 0, tk = 0, h, wi = 0;
    double tl = 0, tr, wl, wr, a, b;
    for (j = 0; j < N; j++) par[j] = NUL;
    wsummary(f, z, Y); /* f(0) for every node */
    for (j = 1; j < N; j++) memcpy(Y + (size_t) j * M, Y, (size_t) M * sizeof(double));
    for (j = 0; j < ts->n; j++) {
        u = ts->samples[j];
        memcpy(X + (size_t) u * K, W + (size_t) j * K, (size_t) K * sizeof(double));
        wsummary(f, X + (size_t) u * K, Y + (size_t) u * M);
    }
    memset(result, 0, (size_t) nw * M * sizeof(double));
#define RS(node, sign)                                                          \
    for (m = 0; m < M; m++) rs[m] += (sign) * bl[node] * Y[(size_t) (node) * M + m];
    while (tj < E || tl < ts->L) {
        while (tk < E && ts->right[ts->O[tk]] == tl) {
            h = ts->O[tk++];
            u = ts->child[h];
            RS(u, -1);
            par[u] = NUL;
            bl[u] = 0;
            u = ts->parent[h];
            while (u != NUL) {
                RS(u, -1);
                for (k = 0; k < K; k++) X[(size_t) u * K + k] -= X[(size_t) ts->child[h] * K + k];
                wsummary(f, X + (size_t) u * K, Y + (size_t) u * M);
                RS(u, +1);
                u = par[u];
            }
        }
        while (tj < E && ts->left[ts->I[tj]] == tl) {
            h = ts->I[tj++];
            u = ts->child[h];
            v = ts->parent[h];
            par[u] = v;
            bl[u] = ts->time[v] - ts->time[u];
            RS(u, +1);
            u = v;
            while (u != NUL) {
                RS(u, -1);
                for (k = 0; k < K; k++) X[(size_t) u * K + k] += X[(size_t) ts->child[h] * K + k];
                wsummary(f, X + (size_t) u * K, Y + (size_t) u * M);
                RS(u, +1);
                u = par[u];
            }
        }
        tr = ts->L;
        if (tj < E && ts->left[ts->I[tj]] < tr) tr = ts->left[ts->I[tj]];
        if (tk < E && ts->right[ts->O[tk]] < tr) tr = ts->right[ts->O[tk]];
        while (windows[wi] < tr) {
            wl = windows[wi];
            wr = windows[wi + 1];
            a = tl > wl ? tl : wl;
            b = tr < wr ? tr : wr;
            for (m = 0; m < M; m++) result[(size_t) wi * M + m] += rs[m] * (b - a);
            if (wr <= tr) {
                wi++;
            } else {
                break;
            }
        }
        tl = tr;
    }
#undef RS
    free(par);
    free(bl);
    free(X);
    free(Y);
    free(rs);
    free(z);
    return 0;
}

/* ---- node statistics: trees.c:1788-1918 ---------------------------------
 * Each node's summary of its own state integrated along the genome: the
 * value held since the node's last update is added, over the span since
 * then, whenever its state changes (it is on the root path above an inserted
 * or removed edge) and at every window boundary.  result: [nw x N x M]. */
static int node_stat(const ots_t *ts, const double *W, const wf_t *f, int32_t nw,
    const double *windows, double *result) {
    const int32_t N = ts->N, E = ts->E, K = f->p->K, M = f->p->M;
    int32_t *par = malloc((size_t) (N + 1) * sizeof(int32_t));
    double *X = calloc((size_t) N * K + 1, sizeof(double));
    double *Y = calloc((size_t) N * M + 1, sizeof(double));
    double *last = calloc((size_t) N + 1, sizeof(double));
    double *z = calloc((size_t) K + 1, sizeof(double));
    int32_t j, k, m, u, c, tj = 0, tk = 0, h, wi = 0;
    double tl = 0, tr, wr;
    for (j = 0; j < N; j++) par[j] = NUL;
    wsummary(f, z, Y); /* f(0) for every node */
    for (j = 1; j < N; j++) memcpy(Y + (size_t) j * M, Y, (size_t) M * sizeof(double));
    for (j = 0; j < ts->n; j++) {
        u = ts->samples[j];
        memcpy(X + (size_t) u * K, W + (size_t) j * K, (size_t) K * sizeof(double));
        wsummary(f, X + (size_t) u * K, Y + (size_t) u * M);
    }
    memset(result, 0, (size_t) nw * N * M * sizeof(double));
#define FLUSH(node, pos)                                                              \
    do {                                                                              \
        for (m = 0; m < M; m++) {                                                     \
            result[((size_t) wi * N + (node)) * M + m] +=                             \
                ((pos) - last[node]) * Y[(size_t) (node) * M + m];                    \
        }                                                                             \
        last[node] = (pos);                                                           \
    } while (0)
    while (tj < E || tl < ts->L) {
        while (tk < E && ts->right[ts->O[tk]] == tl) {
            h = ts->O[tk++];
            c = ts->child[h];
            for (u = ts->parent[h]; u != NUL; u = par[u]) {
                FLUSH(u, tl);
                for (k = 0; k < K; k++) X[(size_t) u * K + k] -= X[(size_t) c * K + k];
                wsummary(f, X + (size_t) u * K, Y + (size_t) u * M);
            }
            par[c] = NUL;
        }
        while (tj < E && ts->left[ts->I[tj]] == tl) {
            h = ts->I[tj++];
            c = ts->child[h];
            par[c] = ts->parent[h];
            for (u = ts->parent[h]; u != NUL; u = par[u]) {
                FLUSH(u, tl);
                for (k = 0; k < K; k++) X[(size_t) u * K + k] += X[(size_t) c * K + k];
                wsummary(f, X + (size_t) u * K, Y + (size_t) u * M);
            }
        }
        tr = ts->L;
        if (tj < E && ts->left[ts->I[tj]] < tr) tr = ts->left[ts->I[tj]];
        if (tk < E && ts->right[ts->O[tk]] < tr) tr = ts->right[ts->O[tk]];
        while (wi < nw && windows[wi + 1] <= tr) {
            wr = windows[wi + 1];
            for (u = 0; u < N; u++) FLUSH(u, wr);
            wi++;
        }
        tl = tr;
    }
#undef FLUSH
    free(par);
    free(X);
    free(Y);
    free(last);
    free(z);
    return 0;
}

/* ---- site sweep: trees.c:1525-1776 -------------------------------------- */
static int site_stat(const ots_t *ts, const double *W, const sf_t *p, int polarised,
    int32_t nw, const double *windows, double *result) {
    const int32_t N = ts->N, E = ts->E, K = p->K, M = p->M;
    int32_t *par = malloc((size_t) (N + 1) * sizeof(int32_t));
    double *X = calloc((size_t) N * K + 1, sizeof(double));
    double *total = calloc((size_t) K + 1, sizeof(double));
    double *aw = NULL, *y = calloc((size_t) M + 1, sizeof(double));
    double *sr = calloc((size_t) M + 1, sizeof(double));
    int32_t j, k, m, u, v, tj = 0, tk = 0, h, wi = 0, t = 0, s, a, mm;
    int32_t maxa = 1;
    double tl = 0, tr;
    for (s = 0; s < ts->S; s++) {
        if (ts->site_nalleles[s] > maxa) maxa = ts->site_nalleles[s];
    }
    aw = calloc((size_t) maxa * K + 1, sizeof(double));
    for (j = 0; j < N; j++) par[j] = NUL;
    for (j = 0; j < ts->n; j++) {
        u = ts->samples[j];
        for (k = 0; k < K; k++) {
            X[(size_t) u * K + k] = W[(size_t) j * K + k];
            total[k] += W[(size_t) j * K + k];
        }
    }
    memset(result, 0, (size_t) nw * M * sizeof(double));
    while (tj < E || tl < ts->L) {
        while (tk < E && ts->right[ts->O[tk]] == tl) {
            h = ts->O[tk++];
            u = ts->child[h];
            v = ts->parent[h];
            while (v != NUL) {
                for (k = 0; k < K; k++) X[(size_t) v * K + k] -= X[(size_t) u * K + k];
                v = par[v];
            }
            par[u] = NUL;
        }
        while (tj < E && ts->left[ts->I[tj]] == tl) {
            h = ts->I[tj++];
            u = ts->child[h];
            v = ts->parent[h];
            par[u] = v;
            while (v != NUL) {
                for (k = 0; k < K; k++) X[(size_t) v * K + k] += X[(size_t) u * K + k];
                v = par[v];
            }
        }
        tr = ts->L;
        if (tj < E && ts->left[ts->I[tj]] < tr) tr = ts->left[ts->I[tj]];
        if (tk < E && ts->right[ts->O[tk]] < tr) tr = ts->right[ts->O[tk]];
        for (s = ts->tree_site0[t]; s < ts->tree_site0[t + 1]; s++) {
            const int32_t na = ts->site_nalleles[s];
            memset(aw, 0, (size_t) na * K * sizeof(double));
            memcpy(aw, total, (size_t) K * sizeof(double));
            for (mm = ts->site_mut0[s]; mm < ts->site_mut0[s + 1]; mm++) {
                const double *xr = X + (size_t) ts->mut_node[mm] * K;
                for (k = 0; k < K; k++) {
                    aw[(size_t) ts->mut_allele[mm] * K + k] += xr[k];
                    aw[(size_t) ts->mut_pallele[mm] * K + k] -= xr[k];
                }
            }
            memset(sr, 0, (size_t) M * sizeof(double));
            for (a = polarised ? 1 : 0; a < na; a++) {
                summary(p, aw + (size_t) a * K, y);
                for (m = 0; m < M; m++) sr[m] += y[m];
            }
            while (windows[wi + 1] <= ts->pos[s]) wi++;
            for (m = 0; m < M; m++) result[(size_t) wi * M + m] += sr[m];
        }
        t++;
        tl = tr;
    }
    free(par);
    free(X);
    free(total);
    free(aw);
    free(y);
    free(sr);
    return 0;
}

static int check_windows(const ots_t *ts, int32_t nw, const double *w) {
    int32_t j;
    if (nw < 1) return -900;
    if (w[0] != 0 || w[nw] != ts->L) return -901;
    for (j = 0; j < nw; j++) {
        if (w[j] >= w[j + 1]) return -901;
    }
    return 0;
}

/* sample_count_stat + general_stat: trees.c:2035-2095, 2114-2220, 4667-4688 */
int oracle_sample_count_stat(const ots_t *ts, int stat, int32_t K, const uint64_t *sizes,
    const int32_t *sets, int32_t tuple, int32_t M, const int32_t *idx, int32_t nw,
    const double *windows, uint32_t options, double *result) {
    int site = !!(options & OPT_SITE), branch = !!(options & OPT_BRANCH);
    int node = !!(options & OPT_NODE);
    double dw[2];
    double *W = NULL, *n = NULL, *total = NULL, *tx = NULL, *ty = NULL;
    int32_t j, k, l, u, si, ret = 0;
    sf_t p;
    wf_t wf;
    if (tuple > 0) {
        if (K < 1) return -905;
        if (M < 1) return -906;
        for (j = 0; j < M * tuple; j++) {
            if (idx[j] < 0 || idx[j] >= K) return -907;
        }
    }
    if (K == 0) return -905;
    j = 0;
    for (k = 0; k < K; k++) {
        if (sizes[k] == 0) return -908;
        for (l = 0; l < (int32_t) sizes[k]; l++) {
            u = sets[j++];
            if (u < 0 || u >= ts->N) return -202;
            if (ts->sample_index_map[u] == NUL) return -601;
        }
    }
    W = calloc((size_t) ts->n * K + 1, sizeof(double));
    j = 0;
    for (k = 0; k < K; k++) {
        for (l = 0; l < (int32_t) sizes[k]; l++) {
            si = ts->sample_index_map[sets[j++]];
            if (W[(size_t) si * K + k] != 0) {
                free(W);
                return -600;
            }
            W[(size_t) si * K + k] = 1;
        }
    }
    /* general_stat checks: trees.c:2048-2078 */
    if (!(site || branch || node)) site = 1;
    if (site + branch + node > 1) {
        free(W);
        return -902;
    }
    if (windows == NULL) {
        nw = 1;
        dw[0] = 0;
        dw[1] = ts->L;
        windows = dw;
    } else {
        ret = check_windows(ts, nw, windows);
        if (ret != 0) {
            free(W);
            return ret;
        }
    }
    n = malloc((size_t) K * sizeof(double));
    total = calloc((size_t) K, sizeof(double));
    for (k = 0; k < K; k++) n[k] = (double) sizes[k];
    for (j = 0; j < ts->n; j++) {
        for (k = 0; k < K; k++) total[k] += W[(size_t) j * K + k];
    }
    p.stat = stat;
    p.K = K;
    p.M = (tuple > 0) ? M : K;
    p.tuple = tuple;
    p.n = n;
    p.idx = idx;
    if (branch || node) {
        if (branch && ts->time_uncalibrated && !(options & OPT_ALLOW_UNCAL)) {
            ret = -910;
            goto out;
        }
        tx = calloc((size_t) K + 1, sizeof(double));
        ty = calloc((size_t) p.M + 1, sizeof(double));
        wf.p = &p;
        wf.polarised = !!(options & OPT_POLARISED);
        wf.total = total;
        wf.tmp_x = tx;
        wf.tmp_y = ty;
        ret = node ? node_stat(ts, W, &wf, nw, windows, result)
                   : branch_stat(ts, W, &wf, nw, windows, result);
    } else {
        ret = site_stat(ts, W, &p, !!(options & OPT_POLARISED), nw, windows, result);
    }
    if (ret == 0 && (options & OPT_SPAN_NORMALISE)) { /* trees.c:1920-1934, 2085-2091 */
        const size_t row = (size_t) p.M * (node ? (size_t) ts->N : 1);
        for (j = 0; j < nw; j++) {
            const double span = windows[j + 1] - windows[j];
            for (size_t r = 0; r < row; r++) result[(size_t) j * row + r] /= span;
        }
    }
out:
    free(W);
    free(n);
    free(total);
    free(tx);
    free(ty);
    return ret;
}

/* general_stat of real-valued weights W [n x K] (trees.c:2036-2095) with a
 * weighted summary (ST_TRAIT_COV .. ST_REL_W_NC): M outputs, idx [M x 2]
 * for the relatedness summaries, dpar their parameters. */
int oracle_general_stat(const ots_t *ts, int stat, int32_t K, const double *W, int32_t M,
    const int32_t *idx, const double *dpar, int32_t nw, const double *windows,
    uint32_t options, double *result) {
    int site = !!(options & OPT_SITE), branch = !!(options & OPT_BRANCH);
    int node = !!(options & OPT_NODE);
    double dw[2];
    double *total = NULL, *tx = NULL, *ty = NULL;
    int32_t j, k, ret = 0;
    sf_t p;
    wf_t wf;
    if (!(site || branch || node)) site = 1;
    if (site + branch + node > 1) return -902;
    if (K < 1) return -903;
    if (M < 1) return -904;
    if (windows == NULL) {
        nw = 1;
        dw[0] = 0;
        dw[1] = ts->L;
        windows = dw;
    } else if ((ret = check_windows(ts, nw, windows)) != 0) {
        return ret;
    }
    memset(&p, 0, sizeof(p));
    p.stat = stat;
    p.K = K;
    p.M = M;
    p.tuple = 2;
    p.idx = idx;
    p.nsamp = (double) ts->n;
    p.dpar = dpar;
    total = calloc((size_t) K, sizeof(double));
    for (j = 0; j < ts->n; j++) {
        for (k = 0; k < K; k++) total[k] += W[(size_t) j * K + k];
    }
    if (branch || node) {
        tx = calloc((size_t) K + 1, sizeof(double));
        ty = calloc((size_t) M + 1, sizeof(double));
        wf.p = &p;
        wf.polarised = !!(options & OPT_POLARISED);
        wf.total = total;
        wf.tmp_x = tx;
        wf.tmp_y = ty;
        ret = node ? node_stat(ts, W, &wf, nw, windows, result)
                   : branch_stat(ts, W, &wf, nw, windows, result);
    } else {
        ret = site_stat(ts, W, &p, !!(options & OPT_POLARISED), nw, windows, result);
    }
    if (ret == 0 && (options & OPT_SPAN_NORMALISE)) {
        const size_t row = (size_t) M * (node ? (size_t) ts->N : 1);
        for (j = 0; j < nw; j++) {
            const double span = windows[j + 1] - windows[j];
            for (size_t r = 0; r < row; r++) result[(size_t) j * row + r] /= span;
        }
    }
    free(total);
    free(tx);
    free(ty);
    return ret;
}

/* ---- genotypes: 0/1 derived-allele matrix [S x n] for sites with at most one
 * mutation (a sample carries the derived allele iff the mutation's node is on
 * its root path in the tree covering the site); the role tsk_variant_decode
 * plays for such sites (genotypes.c:473-595). ---------------------------- */
int oracle_genotypes(const ots_t *ts, int8_t *G) {
    const int32_t N = ts->N, E = ts->E;
    int32_t *par = malloc((size_t) (N + 1) * sizeof(int32_t));
    int32_t j, h, tj = 0, tk = 0, t = 0, s, u;
    double tl = 0, tr;
    for (j = 0; j < N; j++) par[j] = NUL;
    memset(G, 0, (size_t) ts->S * ts->n);
    while (tj < E || tl < ts->L) {
        while (tk < E && ts->right[ts->O[tk]] == tl) {
            h = ts->O[tk++];
            par[ts->child[h]] = NUL;
        }
        while (tj < E && ts->left[ts->I[tj]] == tl) {
            h = ts->I[tj++];
            par[ts->child[h]] = ts->parent[h];
        }
        tr = ts->L;
        if (tj < E && ts->left[ts->I[tj]] < tr) tr = ts->left[ts->I[tj]];
        if (tk < E && ts->right[ts->O[tk]] < tr) tr = ts->right[ts->O[tk]];
        for (s = ts->tree_site0[t]; s < ts->tree_site0[t + 1]; s++) {
            const int32_t m0 = ts->site_mut0[s], m1 = ts->site_mut0[s + 1];
            if (m1 - m0 > 1) {
                free(par);
                return -30005;
            }
            if (m1 == m0) continue;
            for (j = 0; j < ts->n; j++) {
                for (u = ts->samples[j]; u != NUL; u = par[u]) {
                    if (u == ts->mut_node[m0]) {
                        G[(size_t) s * ts->n + j] = 1;
                        break;
                    }
                }
            }
        }
        t++;
        tl = tr;
    }
    free(par);
    return 0;
}

/* ---- genotype decode (tsk_variant_decode, genotypes.c:473-595; missing
 * data genotypes.c:413-436): per site, every sample starts at allele 0; with
 * isolated_as_missing a sample that is a root without children in the tree
 * covering the site gets -1; then each mutation in order sets the samples
 * below its node to its allele index (first-seen order, as mut_allele).
 * G is [S x n] int32. ------------------------------------------------------ */
int oracle_genotype_matrix(const ots_t *ts, int isolated_as_missing, int32_t *G) {
    const int32_t N = ts->N, E = ts->E;
    int32_t *par = malloc((size_t) (N + 1) * sizeof(int32_t));
    int32_t *nch = calloc((size_t) N + 1, sizeof(int32_t));
    int32_t j, h, tj = 0, tk = 0, t = 0, s, u, m;
    double tl = 0, tr;
    for (j = 0; j < N; j++) par[j] = NUL;
    while (tj < E || tl < ts->L) {
        while (tk < E && ts->right[ts->O[tk]] == tl) {
            h = ts->O[tk++];
            par[ts->child[h]] = NUL;
            nch[ts->parent[h]]--;
        }
        while (tj < E && ts->left[ts->I[tj]] == tl) {
            h = ts->I[tj++];
            par[ts->child[h]] = ts->parent[h];
            nch[ts->parent[h]]++;
        }
        tr = ts->L;
        if (tj < E && ts->left[ts->I[tj]] < tr) tr = ts->left[ts->I[tj]];
        if (tk < E && ts->right[ts->O[tk]] < tr) tr = ts->right[ts->O[tk]];
        for (s = ts->tree_site0[t]; s < ts->tree_site0[t + 1]; s++) {
            int32_t *row = G + (size_t) s * ts->n;
            for (j = 0; j < ts->n; j++) {
                u = ts->samples[j];
                row[j] = (isolated_as_missing && par[u] == NUL && nch[u] == 0) ? -1 : 0;
            }
            for (m = ts->site_mut0[s]; m < ts->site_mut0[s + 1]; m++) {
                for (j = 0; j < ts->n; j++) {
                    for (u = ts->samples[j]; u != NUL; u = par[u]) {
                        if (u == ts->mut_node[m]) {
                            row[j] = ts->mut_allele[m];
                            break;
                        }
                    }
                }
            }
        }
        t++;
        tl = tr;
    }
    free(par);
    free(nch);
    return 0;
}

/* ---- AFS: trees.c:3469-3928 --------------------------------------------- */
static void fold(int64_t *coord, const int64_t *dims, int32_t nd) { /* trees.c:3469-3496 */
    int32_t k;
    double nn = 0;
    int s = 0;
    for (k = 0; k < nd; k++) {
        nn += (double) dims[k] - 1;
        s += (int) coord[k];
    }
    nn /= 2;
    k = nd;
    while (s == nn && k > 0) {
        k--;
        nn -= ((double) (dims[k] - 1)) / 2;
        s -= (int) coord[k];
    }
    if (s > nn) {
        for (k = 0; k < nd; k++) coord[k] = dims[k] - 1 - coord[k];
    }
}

static int64_t afs_index(const int64_t *coord, const int64_t *dims, int32_t nd) {
    int64_t off = 0;
    int32_t k;
    for (k = 0; k < nd; k++) off = off * dims[k] + coord[k];
    return off;
}

int oracle_afs(const ots_t *ts, int32_t K, const uint64_t *sizes, const int32_t *sets,
    int32_t nw, const double *windows, int32_t ntw, const double *time_windows,
    uint32_t options, double *result) {
    int site = !!(options & OPT_SITE), branch = !!(options & OPT_BRANCH);
    int polarised = !!(options & OPT_POLARISED);
    double dw[2], dtw[2];
    const int32_t N = ts->N, E = ts->E, KK = K + 1;
    int64_t *dims, *coord, afs_size = 1;
    double *X, *total;
    int32_t *par;
    double *last;
    int32_t j, k, l, u, v, h, tj = 0, tk = 0, wi = 0, t = 0, s, mm, a, twi;
    double tl = 0, tr;
    if (options & OPT_NODE) return -909;
    if (!(site || branch)) site = 1;
    if (site + branch > 1) return -902;
    if (windows == NULL) {
        nw = 1;
        dw[0] = 0;
        dw[1] = ts->L;
        windows = dw;
    } else {
        int r = check_windows(ts, nw, windows);
        if (r != 0) return r;
    }
    if (time_windows == NULL) {
        ntw = 1;
        dtw[0] = 0;
        dtw[1] = INFINITY;
        time_windows = dtw;
    } else {
        if (ntw < 1) return -922;
        if (time_windows[0] != 0) return -924;
        for (j = 0; j < ntw; j++) {
            if (time_windows[j] >= time_windows[j + 1]) return -924;
        }
        if (site && !(time_windows[0] == 0.0 && isinf((float) time_windows[1]))) return -909;
    }
    if (K == 0) return -905;
    j = 0;
    for (k = 0; k < K; k++) {
        if (sizes[k] == 0) return -908;
        for (l = 0; l < (int32_t) sizes[k]; l++) {
            u = sets[j++];
            if (u < 0 || u >= N) return -202;
            if (ts->sample_index_map[u] == NUL) return -601;
        }
    }
    dims = malloc((size_t) KK * sizeof(int64_t));
    coord = malloc((size_t) KK * sizeof(int64_t));
    X = calloc((size_t) N * KK + 1, sizeof(double));
    total = calloc((size_t) KK, sizeof(double));
    j = 0;
    for (k = 0; k < K; k++) {
        dims[k] = 1 + (int64_t) sizes[k];
        afs_size *= dims[k];
        total[k] = (double) sizes[k];
        for (l = 0; l < (int32_t) sizes[k]; l++) {
            u = sets[j++];
            if (X[(size_t) u * KK + k] != 0) {
                free(dims);
                free(coord);
                free(X);
                free(total);
                return -600;
            }
            X[(size_t) u * KK + k] = 1;
        }
    }
    total[K] = ts->n;
    for (j = 0; j < ts->n; j++) X[(size_t) ts->samples[j] * KK + K] = 1;
    memset(result, 0, (size_t) (nw * ntw * afs_size) * sizeof(double));
    par = malloc((size_t) (N + 1) * sizeof(int32_t));
    last = calloc((size_t) N + 1, sizeof(double));
    for (j = 0; j < N; j++) par[j] = NUL;

    if (site) { /* trees.c:3498-3648 */
        int32_t maxa = 1;
        double *aw;
        for (s = 0; s < ts->S; s++) {
            if (ts->site_nalleles[s] > maxa) maxa = ts->site_nalleles[s];
        }
        aw = calloc((size_t) maxa * KK + 1, sizeof(double));
        while (tj < E || tl < ts->L) {
            while (tk < E && ts->right[ts->O[tk]] == tl) {
                h = ts->O[tk++];
                u = ts->child[h];
                v = ts->parent[h];
                while (v != NUL) {
                    for (k = 0; k < KK; k++) X[(size_t) v * KK + k] -= X[(size_t) u * KK + k];
                    v = par[v];
                }
                par[u] = NUL;
            }
            while (tj < E && ts->left[ts->I[tj]] == tl) {
                h = ts->I[tj++];
                u = ts->child[h];
                v = ts->parent[h];
                par[u] = v;
                while (v != NUL) {
                    for (k = 0; k < KK; k++) X[(size_t) v * KK + k] += X[(size_t) u * KK + k];
                    v = par[v];
                }
            }
            tr = ts->L;
            if (tj < E && ts->left[ts->I[tj]] < tr) tr = ts->left[ts->I[tj]];
            if (tk < E && ts->right[ts->O[tk]] < tr) tr = ts->right[ts->O[tk]];
            for (s = ts->tree_site0[t]; s < ts->tree_site0[t + 1]; s++) {
                const int32_t na = ts->site_nalleles[s];
                while (windows[wi + 1] <= ts->pos[s]) wi++;
                memset(aw, 0, (size_t) na * KK * sizeof(double));
                memcpy(aw, total, (size_t) KK * sizeof(double));
                for (mm = ts->site_mut0[s]; mm < ts->site_mut0[s + 1]; mm++) {
                    const double *xr = X + (size_t) ts->mut_node[mm] * KK;
                    for (k = 0; k < KK; k++) {
                        aw[(size_t) ts->mut_allele[mm] * KK + k] += xr[k];
                        aw[(size_t) ts->mut_pallele[mm] * KK + k] -= xr[k];
                    }
                }
                for (a = polarised ? 1 : 0; a < na; a++) {
                    const double *row = aw + (size_t) a * KK;
                    const int64_t alls = (int64_t) row[K];
                    if (alls > 0 && alls < ts->n) {
                        for (k = 0; k < K; k++) coord[k] = (int64_t) row[k];
                        if (!polarised) fold(coord, dims, K);
                        result[(size_t) wi * afs_size + afs_index(coord, dims, K)]
                            += polarised ? 1 : 0.5;
                    }
                }
            }
            t++;
            tl = tr;
        }
        free(aw);
    } else { /* trees.c:3650-3813 */
#define FLUSH(node, pos_right)                                                          \
    do {                                                                                \
        const double *cr = X + (size_t) (node) * KK;                                    \
        const int64_t alls = (int64_t) cr[K];                                           \
        if (par[node] != NUL) {                                                         \
            const double tu = ts->time[node], tv = ts->time[par[node]];                 \
            if (0 < alls && alls < ts->n) {                                             \
                twi = 0;                                                                \
                while (twi < ntw && time_windows[twi] < tv) {                           \
                    double hi = time_windows[twi + 1] < tv ? time_windows[twi + 1] : tv;\
                    double lo = time_windows[twi] > tu ? time_windows[twi] : tu;        \
                    double twbl = hi - lo > 0 ? hi - lo : 0.0;                          \
                    for (k = 0; k < K; k++) coord[k] = (int64_t) cr[k];                 \
                    if (!polarised) fold(coord, dims, K);                               \
                    result[(size_t) (wi * ntw + twi) * afs_size                         \
                           + afs_index(coord, dims, K)]                                 \
                        += ((pos_right) - last[node]) * twbl;                           \
                    twi++;                                                              \
                }                                                                       \
            }                                                                           \
        }                                                                               \
        last[node] = (pos_right);                                                       \
    } while (0)
        if (ts->time_uncalibrated && !(options & OPT_ALLOW_UNCAL)) {
            free(dims);
            free(coord);
            free(X);
            free(total);
            free(par);
            free(last);
            return -910;
        }
        while (tj < E || tl < ts->L) {
            while (tk < E && ts->right[ts->O[tk]] == tl) {
                h = ts->O[tk++];
                u = ts->child[h];
                v = ts->parent[h];
                FLUSH(u, tl);
                while (v != NUL) {
                    FLUSH(v, tl);
                    for (k = 0; k < KK; k++) X[(size_t) v * KK + k] -= X[(size_t) u * KK + k];
                    v = par[v];
                }
                par[u] = NUL;
            }
            while (tj < E && ts->left[ts->I[tj]] == tl) {
                h = ts->I[tj++];
                u = ts->child[h];
                v = ts->parent[h];
                par[u] = v;
                while (v != NUL) {
                    FLUSH(v, tl);
                    for (k = 0; k < KK; k++) X[(size_t) v * KK + k] += X[(size_t) u * KK + k];
                    v = par[v];
                }
            }
            tr = ts->L;
            if (tj < E && ts->left[ts->I[tj]] < tr) tr = ts->left[ts->I[tj]];
            if (tk < E && ts->right[ts->O[tk]] < tr) tr = ts->right[ts->O[tk]];
            while (wi < nw && windows[wi + 1] <= tr) {
                const double wr = windows[wi + 1];
                for (u = 0; u < N; u++) FLUSH(u, wr);
                wi++;
            }
            tl = tr;
        }
#undef FLUSH
    }
    if (options & OPT_SPAN_NORMALISE) {
        for (j = 0; j < nw; j++) {
            const double span = windows[j + 1] - windows[j];
            int64_t q;
            for (q = 0; q < afs_size * ntw; q++) result[(size_t) j * afs_size * ntw + q] /= span;
        }
    }
    free(dims);
    free(coord);
    free(X);
    free(total);
    free(par);
    free(last);
    return 0;
}

/* ---- genetic_relatedness_vector, branch mode: trees.c:10402-10816 --------
 * For window w, focal node f and weight column k:
 *   y[w][f][k] = integral over the window of sum over the ancestors u of f
 *                (f included) of branch_length(u) * Z(u, k),
 * Z(u, k) the sum of the weights of the samples below u, the weights first
 * centred over the samples (trees.c:10517-10540, unless NONCENTRED), the
 * result then centred over the focal nodes (trees.c:10676-10697, unless
 * NONCENTRED) and span-normalised (trees.c:10763-10765).  Stated directly:
 * every tree is rebuilt from its edges and Z summed by walking each sample's
 * root path, so the reference's lazy push-down (trees.c:10553-10605) is not
 * followed.  Windows need not span the sequence (check_windows with no
 * full-span rule, trees.c:1244-1286).  result: [nw x nf x K]. */
int oracle_relatedness_vector(const ots_t *ts, int32_t K, const double *weights, int32_t nw,
    const double *windows, int32_t nf, const int32_t *focal, uint32_t options,
    double *result) {
    const int32_t N = ts->N, E = ts->E, n = ts->n;
    int32_t *par;
    double *Z, *wc, *mean;
    int32_t j, k, u, h, tj = 0, tk = 0, wi;
    double tl = 0, tr;
    if (options & (OPT_SITE | OPT_NODE)) return -909;
    if (nw < 1) return -900;
    if (windows[0] < 0 || windows[nw] > ts->L) return -901;
    for (j = 0; j < nw; j++) {
        if (windows[j] >= windows[j + 1]) return -901;
    }
    for (j = 0; j < nf; j++) {
        if (focal[j] < 0 || focal[j] >= N) return -202;
    }
    par = malloc((size_t) (N + 1) * sizeof(int32_t));
    Z = calloc((size_t) N * K + 1, sizeof(double));
    wc = malloc((size_t) n * K * sizeof(double) + 8);
    mean = calloc((size_t) K + 1, sizeof(double));
    for (j = 0; j < N; j++) par[j] = NUL;
    memcpy(wc, weights, (size_t) n * K * sizeof(double));
    if (!(options & (1u << 14))) {
        for (j = 0; j < n; j++) {
            for (k = 0; k < K; k++) mean[k] += wc[(size_t) j * K + k];
        }
        for (k = 0; k < K; k++) mean[k] /= (double) n;
        for (j = 0; j < n; j++) {
            for (k = 0; k < K; k++) wc[(size_t) j * K + k] -= mean[k];
        }
    }
    memset(result, 0, (size_t) nw * nf * K * sizeof(double));
    while (tl < ts->L) {
        while (tk < E && ts->right[ts->O[tk]] == tl) {
            par[ts->child[ts->O[tk]]] = NUL;
            tk++;
        }
        while (tj < E && ts->left[ts->I[tj]] == tl) {
            h = ts->I[tj++];
            par[ts->child[h]] = ts->parent[h];
        }
        tr = ts->L;
        if (tj < E && ts->left[ts->I[tj]] < tr) tr = ts->left[ts->I[tj]];
        if (tk < E && ts->right[ts->O[tk]] < tr) tr = ts->right[ts->O[tk]];
        /* this tree's Z: each sample's weights up its root path */
        memset(Z, 0, (size_t) N * K * sizeof(double));
        for (j = 0; j < n; j++) {
            for (u = ts->samples[j]; u != NUL; u = par[u]) {
                for (k = 0; k < K; k++) Z[(size_t) u * K + k] += wc[(size_t) j * K + k];
            }
        }
        for (wi = 0; wi < nw; wi++) {
            const double a = tl > windows[wi] ? tl : windows[wi];
            const double b = tr < windows[wi + 1] ? tr : windows[wi + 1];
            if (!(a < b)) continue;
            for (j = 0; j < nf; j++) {
                double *row = result + ((size_t) wi * nf + j) * K;
                for (u = focal[j]; u != NUL && par[u] != NUL; u = par[u]) {
                    const double bl = ts->time[par[u]] - ts->time[u];
                    for (k = 0; k < K; k++) row[k] += (b - a) * bl * Z[(size_t) u * K + k];
                }
            }
        }
        tl = tr;
    }
    if (!(options & (1u << 14)) && nf > 0) {
        for (wi = 0; wi < nw; wi++) {
            for (k = 0; k < K; k++) {
                double s = 0;
                for (j = 0; j < nf; j++) s += result[((size_t) wi * nf + j) * K + k];
                s /= (double) nf;
                for (j = 0; j < nf; j++) result[((size_t) wi * nf + j) * K + k] -= s;
            }
        }
    }
    if (options & OPT_SPAN_NORMALISE) {
        for (wi = 0; wi < nw; wi++) {
            const double span = windows[wi + 1] - windows[wi];
            for (size_t r = 0; r < (size_t) nf * K; r++) result[(size_t) wi * nf * K + r] /= span;
        }
    }
    free(par);
    free(Z);
    free(wc);
    free(mean);
    return 0;
}
