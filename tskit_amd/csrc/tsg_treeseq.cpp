// Host side of the device-resident tree sequence: validation, edge indexes,
// tree breakpoints and per-tree site ranges, allele encoding, and the chunk
// plan (per-chunk "sweep programs") that the HIP kernels execute.
//
// Reference behaviour followed (restated, not copied):
//   integrity checks     c/tskit/tables.c:10397-10498 (edges), 10508-10545
//                        (sites), 10548-10640 (mutations), 10790-10850 (trees)
//   index build          c/tskit/tables.c:10286-10310, 11393-11461
//   samples              c/tskit/trees.c:404-453
//   breakpoints / sites  c/tskit/trees.c:243-359
//   allele table         c/tskit/trees.c:1525-1612 (allele order = first seen)
#include "tsg_treeseq.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string_view>
#include <thread>

namespace {

int check_tables(const tsg_treeseq *ts) {
    const int32_t N = ts->num_nodes;
    for (int32_t u = 0; u < N; u++) {
        if (!std::isfinite(ts->time[u])) return TSG_ERR_TIME_NONFINITE;
    }
    std::vector<char> parent_seen(N, 0);
    for (int32_t j = 0; j < ts->num_edges; j++) {
        const int32_t p = ts->parent[j], c = ts->child[j];
        const double l = ts->left[j], r = ts->right[j];
        if (p == TSG_NULL) return TSG_ERR_NULL_PARENT;
        if (p < 0 || p >= N) return TSG_ERR_NODE_OUT_OF_BOUNDS;
        if (c == TSG_NULL) return TSG_ERR_NULL_CHILD;
        if (c < 0 || c >= N) return TSG_ERR_NODE_OUT_OF_BOUNDS;
        if (!(std::isfinite(l) && std::isfinite(r))) return TSG_ERR_GENOME_COORDS_NONFINITE;
        if (l < 0) return TSG_ERR_LEFT_LESS_ZERO;
        if (r > ts->L) return TSG_ERR_RIGHT_GREATER_SEQ_LENGTH;
        if (l >= r) return TSG_ERR_BAD_EDGE_INTERVAL;
        if (ts->time[c] >= ts->time[p]) return TSG_ERR_BAD_NODE_TIME_ORDERING;
        if (parent_seen[p]) return TSG_ERR_EDGES_NONCONTIGUOUS_PARENTS;
        if (j > 0) {
            const int32_t lp = ts->parent[j - 1], lc = ts->child[j - 1];
            if (ts->time[p] < ts->time[lp]) return TSG_ERR_EDGES_NOT_SORTED_PARENT_TIME;
            if (ts->time[p] == ts->time[lp]) {
                if (p == lp) {
                    if (c < lc) return TSG_ERR_EDGES_NOT_SORTED_CHILD;
                    if (c == lc && l <= ts->left[j - 1]) return TSG_ERR_EDGES_NOT_SORTED_LEFT;
                } else {
                    parent_seen[lp] = 1;
                }
            }
        }
    }
    for (int32_t j = 0; j < ts->num_sites; j++) {
        const double x = ts->site_pos[j];
        if (!std::isfinite(x) || x < 0 || x >= ts->L) return TSG_ERR_BAD_SITE_POSITION;
        if (j > 0) {
            if (ts->site_pos[j - 1] == x) return TSG_ERR_DUPLICATE_SITE_POSITION;
            if (ts->site_pos[j - 1] > x) return TSG_ERR_UNSORTED_SITES;
        }
    }
    for (int32_t j = 0; j < ts->num_mutations; j++) {
        const int32_t s = ts->mut_site[j], u = ts->mut_node[j], pm = ts->mut_parent[j];
        if (s < 0 || s >= ts->num_sites) return TSG_ERR_SITE_OUT_OF_BOUNDS;
        if (u < 0 || u >= N) return TSG_ERR_NODE_OUT_OF_BOUNDS;
        if (pm < TSG_NULL || pm >= ts->num_mutations) return TSG_ERR_MUTATION_OUT_OF_BOUNDS;
        if (pm == j) return TSG_ERR_MUTATION_PARENT_EQUAL;
        if (pm != TSG_NULL) {
            if (ts->mut_site[pm] != s) return TSG_ERR_MUTATION_PARENT_DIFFERENT_SITE;
            if (pm > j) return TSG_ERR_MUTATION_PARENT_AFTER_CHILD;
        }
        if (j > 0 && ts->mut_site[j - 1] > s) return TSG_ERR_UNSORTED_MUTATIONS;
    }
    return 0;
}

}  // namespace

// Insertion / removal orders (tables.c:11393-11461).
void tsg_build_index(const double *left, const double *right, const int32_t *parent,
                     const int32_t *child, const double *time, int32_t E, int32_t *I,
                     int32_t *O) {
    std::vector<int32_t> idx(E);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
        if (left[a] != left[b]) return left[a] < left[b];
        const double ta = time[parent[a]], tb = time[parent[b]];
        if (ta != tb) return ta < tb;
        if (parent[a] != parent[b]) return parent[a] < parent[b];
        return child[a] < child[b];
    });
    std::copy(idx.begin(), idx.end(), I);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
        if (right[a] != right[b]) return right[a] < right[b];
        const double ta = time[parent[a]], tb = time[parent[b]];
        if (ta != tb) return ta > tb;
        if (parent[a] != parent[b]) return parent[a] > parent[b];
        return child[a] > child[b];
    });
    std::copy(idx.begin(), idx.end(), O);
}

// Fills samples, breakpoints, per-tree site ranges and allele encoding, and
// checks the index describes a valid sequence of trees.
int tsg_init_trees(tsg_treeseq *ts, const char *ancestral_state,
                   const uint64_t *ancestral_state_offset, const char *derived_state,
                   const uint64_t *derived_state_offset) {
    int ret = check_tables(ts);
    if (ret != 0) return ret;
    const int32_t N = ts->num_nodes, E = ts->num_edges;
    ts->sample_index_map.assign(N, TSG_NULL);
    ts->samples.clear();
    for (int32_t u = 0; u < N; u++) {
        if (ts->flags[u] & TSG_NODE_IS_SAMPLE) {
            ts->sample_index_map[u] = (int32_t) ts->samples.size();
            ts->samples.push_back(u);
        }
    }
    ts->num_samples = (int32_t) ts->samples.size();
    for (int32_t j = 0; j < E; j++) {
        if (ts->I[j] < 0 || ts->I[j] >= E || ts->O[j] < 0 || ts->O[j] >= E) {
            return TSG_ERR_EDGE_OUT_OF_BOUNDS;
        }
    }
    // tree sweep: breakpoints + contradictory-children check
    std::vector<int32_t> par(N, TSG_NULL);
    std::vector<uint8_t> used(E, 0);
    ts->breakpoints.clear();
    ts->tree_site0.clear();
    int32_t j = 0, k = 0, site = 0;
    double tl = 0;
    while (j < E || tl < ts->L) {
        ts->breakpoints.push_back(tl);
        ts->tree_site0.push_back(site);
        while (k < E && ts->right[ts->O[k]] == tl) {
            const int32_t e = ts->O[k];
            if (used[e] != 1) return -707;  // TSK_ERR_TABLES_BAD_INDEXES
            used[e] = 2;
            par[ts->child[e]] = TSG_NULL;
            k++;
        }
        while (j < E && ts->left[ts->I[j]] == tl) {
            const int32_t e = ts->I[j];
            if (used[e] != 0) return -707;
            used[e] = 1;
            if (par[ts->child[e]] != TSG_NULL) return TSG_ERR_BAD_EDGES_CONTRADICTORY_CHILDREN;
            par[ts->child[e]] = ts->parent[e];
            j++;
        }
        double tr = ts->L;
        if (j < E) tr = std::min(tr, ts->left[ts->I[j]]);
        if (k < E) tr = std::min(tr, ts->right[ts->O[k]]);
        while (site < ts->num_sites && ts->site_pos[site] < tr) site++;
        tl = tr;
    }
    ts->num_trees = (int32_t) ts->breakpoints.size();
    ts->breakpoints.push_back(ts->L);
    ts->tree_site0.push_back(site);

    // per-site mutation ranges and allele indexes
    const int32_t S = ts->num_sites, Mu = ts->num_mutations;
    ts->site_mut0.assign(S + 1, 0);
    for (int32_t m = 0; m < Mu; m++) ts->site_mut0[ts->mut_site[m] + 1]++;
    for (int32_t s = 0; s < S; s++) ts->site_mut0[s + 1] += ts->site_mut0[s];
    ts->mut_allele.assign(Mu, 0);
    ts->mut_pallele.assign(Mu, 0);
    ts->site_nalleles.assign(S, 1);
    ts->max_alleles = 1;
    std::vector<std::string_view> alleles;
    auto dstate = [&](int32_t m) {
        if (derived_state == nullptr) return std::string_view("1", 1);
        return std::string_view(derived_state + derived_state_offset[m],
                                derived_state_offset[m + 1] - derived_state_offset[m]);
    };
    for (int32_t s = 0; s < S; s++) {
        alleles.clear();
        if (ancestral_state == nullptr) {
            alleles.emplace_back("0", 1);
        } else {
            alleles.emplace_back(ancestral_state + ancestral_state_offset[s],
                                 ancestral_state_offset[s + 1] - ancestral_state_offset[s]);
        }
        auto find = [&](std::string_view v) -> int32_t {
            for (size_t a = 0; a < alleles.size(); a++) {
                if (alleles[a] == v) return (int32_t) a;
            }
            return -1;
        };
        for (int32_t m = ts->site_mut0[s]; m < ts->site_mut0[s + 1]; m++) {
            std::string_view d = dstate(m);
            int32_t a = find(d);
            if (a < 0) {
                a = (int32_t) alleles.size();
                alleles.push_back(d);
            }
            ts->mut_allele[m] = a;
            int32_t pa = 0;
            if (ts->mut_parent[m] != TSG_NULL) pa = find(dstate(ts->mut_parent[m]));
            ts->mut_pallele[m] = pa;
        }
        ts->site_nalleles[s] = (int32_t) alleles.size();
        ts->max_alleles = std::max(ts->max_alleles, (int32_t) alleles.size());
    }
    return 0;
}

namespace {

// Cuts the trees into at most num_chunks runs of consecutive trees with about
// the same number of replayed edge events (the opening transition -- every
// edge with left == 0 -- is applied by the planner and not replayed).
std::vector<int32_t> chunk_bounds(const tsg_treeseq *ts, int32_t num_chunks) {
    const int32_t T = ts->num_trees, E = ts->num_edges;
    std::vector<int64_t> tr_events(T + 1, 0);
    {
        int32_t j = 0, k = 0;
        for (int32_t t = 0; t < T; t++) {
            const double x = ts->breakpoints[t];
            while (k < E && ts->right[ts->O[k]] == x) {
                tr_events[t]++;
                k++;
            }
            while (j < E && ts->left[ts->I[j]] == x) {
                tr_events[t]++;
                j++;
            }
        }
    }
    tr_events[0] = 0;
    int64_t total_events = 0;
    for (int32_t t = 0; t < T; t++) total_events += tr_events[t];
    const int32_t C = std::max(1, std::min(num_chunks, T));
    std::vector<int32_t> bounds{0};
    const double per = (double) total_events / C;
    int64_t acc = 0;
    int32_t c = 1;
    for (int32_t t = 0; t < T && c < C; t++) {
        acc += tr_events[t];
        if ((double) acc >= per * c && t + 1 < T) {
            if (bounds.back() < t + 1) bounds.push_back(t + 1);
            c++;
        }
    }
    bounds.push_back(T);
    return bounds;
}

}  // namespace

// Builds the per-chunk sweep programs.  Chunk c covers trees
// [chunk_tree0[c], chunk_tree0[c+1]).  Its program is
//   * a local numbering of every node it touches (samples first, so local id
//     j is sample index j in every chunk),
//   * the parent / branch length of each local node in its start tree
//     (every chunk, chunk 0 included, starts from a snapshot: the opening
//     transition -- the insertion of every edge with left == 0 -- is applied
//     here and stores no events, so no chunk replays a whole tree; the
//     running branch sum of the first tree is evaluated separately, in
//     parallel over its nodes),
//   * the transitions it applies: into each of its trees after the first,
//     and the closing transition into the next chunk's first tree (so the
//     running branch sum at the end of chunk c equals the one at the start
//     of chunk c+1).
int tsg_build_plan(tsg_treeseq *ts, int32_t num_chunks) {
    const int32_t T = ts->num_trees, E = ts->num_edges, N = ts->num_nodes;
    const int32_t n = ts->num_samples;
    ts->chunk_tree0 = chunk_bounds(ts, num_chunks);
    const int32_t C = (int32_t) ts->chunk_tree0.size() - 1;
    ts->num_chunks = C;

    ts->tr_off.assign(T + 1, 0);
    ts->tr_nrem.assign(T, 0);
    ts->ev_child.clear();
    ts->ev_parent.clear();
    ts->ev_bl.clear();
    ts->path.clear();
    ts->ev_path0.assign(1, 0);
    ts->max_path = 0;
    ts->loc_parent0.clear();
    ts->loc_bl0.clear();
    ts->loc_time.clear();
    ts->loc_lastrem.clear();
    ts->chunk_node0.assign(C + 1, 0);
    ts->mut_node_loc.assign(ts->num_mutations, TSG_NULL);
    ts->max_nloc = 0;
    ts->max_chunk_events = 0;

    std::vector<int32_t> gpar(N, TSG_NULL);
    std::vector<double> glastrem(N, 0.0);  // last removal position per node
    std::vector<double> loc_lastrem;
    std::vector<int32_t> alive;          // children with a parent
    std::vector<int32_t> alive_pos(N, -1);
    std::vector<int32_t> stamp(N, -1), locid(N, -1);
    std::vector<int32_t> loc2glob;
    std::vector<int32_t> snap_parent;
    std::vector<double> snap_bl;
    int32_t j = 0, k = 0;  // index cursors (insertion / removal)
    // opening transition: the first tree's edges (no removals at x = 0)
    while (j < E && ts->left[ts->I[j]] == ts->breakpoints[0]) {
        const int32_t e = ts->I[j];
        gpar[ts->child[e]] = ts->parent[e];
        alive_pos[ts->child[e]] = (int32_t) alive.size();
        alive.push_back(ts->child[e]);
        j++;
    }

    for (int32_t c = 0; c < C; c++) {
        loc2glob.clear();
        snap_parent.clear();
        snap_bl.clear();
        loc_lastrem.clear();
        auto ensure = [&](int32_t u) -> int32_t {
            if (stamp[u] != c) {
                stamp[u] = c;
                locid[u] = (int32_t) loc2glob.size();
                loc2glob.push_back(u);
                loc_lastrem.push_back(glastrem[u]);
                snap_parent.push_back(TSG_NULL);
                snap_bl.push_back(0.0f);
            }
            return locid[u];
        };
        // root path from p in the current tree, local ids, p first
        auto record_path = [&](int32_t p) {
            int32_t len = 0;
            for (int32_t w = p; w != TSG_NULL; w = gpar[w]) {
                ts->path.push_back(ensure(w));
                len++;
            }
            ts->ev_path0.push_back((int64_t) ts->path.size());
            ts->max_path = std::max(ts->max_path, len);
        };
        for (int32_t s = 0; s < n; s++) ensure(ts->samples[s]);
        for (int32_t u : alive) {
            const int32_t lu = ensure(u);
            const int32_t lp = ensure(gpar[u]);
            snap_parent[lu] = lp;
            snap_bl[lu] = ts->time[gpar[u]] - ts->time[u];
        }
        const int32_t t0 = ts->chunk_tree0[c], t1 = ts->chunk_tree0[c + 1];
        // transitions owned by this chunk (transition 0 is the opening one)
        int32_t tfirst = t0 + 1;
        int32_t tlast = std::min(t1, T - 1);  // inclusive; t1 is the closing one
        int64_t chunk_events = 0;
        for (int32_t t = tfirst; t <= tlast; t++) {
            ts->tr_off[t] = (int64_t) ts->ev_child.size();
            const double x = ts->breakpoints[t];
            int32_t nrem = 0;
            while (k < E && ts->right[ts->O[k]] == x) {
                const int32_t e = ts->O[k];
                const int32_t u = ts->child[e];
                ts->ev_child.push_back(ensure(u));
                ts->ev_parent.push_back(ensure(ts->parent[e]));
                ts->ev_bl.push_back(0.0f);
                record_path(ts->parent[e]);
                gpar[u] = TSG_NULL;
                glastrem[u] = x;
                // swap-remove from alive
                const int32_t pos = alive_pos[u];
                alive[pos] = alive.back();
                alive_pos[alive[pos]] = pos;
                alive.pop_back();
                alive_pos[u] = -1;
                nrem++;
                k++;
            }
            while (j < E && ts->left[ts->I[j]] == x) {
                const int32_t e = ts->I[j];
                const int32_t u = ts->child[e], p = ts->parent[e];
                ts->ev_child.push_back(ensure(u));
                ts->ev_parent.push_back(ensure(p));
                ts->ev_bl.push_back(ts->time[p] - ts->time[u]);
                gpar[u] = p;
                record_path(p);
                alive_pos[u] = (int32_t) alive.size();
                alive.push_back(u);
                j++;
            }
            ts->tr_nrem[t] = nrem;
            chunk_events += (int64_t) ts->ev_child.size() - ts->tr_off[t];
        }
        ts->max_chunk_events = (int32_t) std::max<int64_t>(ts->max_chunk_events, chunk_events);
        // mutation nodes of this chunk's sites
        for (int32_t s = ts->tree_site0[t0]; s < ts->tree_site0[t1]; s++) {
            for (int32_t m = ts->site_mut0[s]; m < ts->site_mut0[s + 1]; m++) {
                ts->mut_node_loc[m] = ensure(ts->mut_node[m]);
            }
        }
        const int32_t nloc = (int32_t) loc2glob.size();
        ts->chunk_node0[c + 1] = ts->chunk_node0[c] + nloc;
        ts->max_nloc = std::max(ts->max_nloc, nloc);
        ts->loc_parent0.insert(ts->loc_parent0.end(), snap_parent.begin(), snap_parent.end());
        ts->loc_bl0.insert(ts->loc_bl0.end(), snap_bl.begin(), snap_bl.end());
        for (int32_t i = 0; i < nloc; i++) ts->loc_time.push_back(ts->time[loc2glob[i]]);
        ts->loc_lastrem.insert(ts->loc_lastrem.end(), loc_lastrem.begin(), loc_lastrem.end());
    }
    // Every transition 0..T-1 is owned by exactly one chunk and they were
    // emitted in order, so the offsets are already monotone.
    ts->tr_off[T] = (int64_t) ts->ev_child.size();
    ts->full_built = true;
    return 0;
}

namespace {

// What one planner thread produces for a contiguous run of chunks; the
// parts are concatenated in chunk order afterwards.  Transition offsets
// (tr_off) are written part-relative into the plan and rebased on merge;
// tr_nrem, mut_slot and sample_slot entries are disjoint per part and are
// written in place.
struct HotPart {
    std::vector<int32_t> ev_child, path, slot_par, lev_slot0, slot_node, slot_nsd;
    std::vector<double> ev_bl, slot_bl0, slot_tp0;
    std::vector<double> slot_last0;    // -1: before this part (filled on merge)
    std::vector<double> slot_ins0;     // -1: before this part (filled on merge)
    std::vector<int32_t> touched;      // nodes an event of this part flushed
    std::vector<double> touched_last;  // ... and the position of the last one
    std::vector<int32_t> inserted;     // nodes an insertion of this part gave a parent
    std::vector<double> inserted_at;   // ... and the position of the last one
    std::vector<int32_t> end_par;      // parents after the part's last chunk
    std::vector<int64_t> ev_end;       // per event: end of its path (part-relative)
    std::vector<int64_t> chunk_slots;  // per chunk
    std::vector<int64_t> chunk_levs;   // per chunk: level-boundary entries
    int32_t max_slots = 0, max_path = 0, max_levels = 0;
    bool too_big = false;
};

void plan_hot_range(const tsg_treeseq *ts, tsg_hotplan &hp, int32_t cbeg, int32_t cend,
                    HotPart &out) {
    const int32_t T = ts->num_trees, E = ts->num_edges, N = ts->num_nodes;
    const int32_t n = ts->num_samples;
    std::vector<int32_t> gpar(N, TSG_NULL);
    std::vector<int32_t> hot_stamp(N, -1), hot_idx(N, -1);    // hot in chunk c, index in `hot`
    std::vector<int32_t> memo_stamp(N, -1), memo_nha(N, -1);  // nearest hot ancestor-or-self
    std::vector<int32_t> hot, ev_u, path_g, hpar, depth, slot_of, stack, lev_count, next;
    std::vector<int64_t> ev_p0;
    std::vector<double> ev_x;                                // position of each event
    std::vector<char> ev_ins;                                // insertion (1) or removal (0)
    std::vector<std::pair<int32_t, int32_t>> undo, redo;     // (child, parent before / after)
    // position of each node's last flush by an event in this part (a node on
    // an event's root path, or a removed edge's child; -1: none), and of the
    // insertion that gave it its current parent (-1: none in this part)
    std::vector<double> last_touch(N, -1.0), ins_at(N, -1.0);
    // replay up to the start tree of chunk cbeg
    int32_t j = 0, k = 0;
    for (int32_t t = 0; t <= hp.chunk_tree0[cbeg]; t++) {
        const double x = ts->breakpoints[t];
        while (k < E && ts->right[ts->O[k]] == x) gpar[ts->child[ts->O[k++]]] = TSG_NULL;
        while (j < E && ts->left[ts->I[j]] == x) {
            gpar[ts->child[ts->I[j]]] = ts->parent[ts->I[j]];
            j++;
        }
    }
    for (int32_t c = cbeg; c < cend; c++) {
        hot.clear();
        ev_u.clear();
        path_g.clear();
        ev_p0.assign(1, 0);
        ev_x.clear();
        ev_ins.clear();
        undo.clear();
        redo.clear();
        auto touch = [&](int32_t u) {
            if (hot_stamp[u] != c) {
                hot_stamp[u] = c;
                hot_idx[u] = (int32_t) hot.size();
                hot.push_back(u);
            }
        };
        auto record_path = [&](int32_t p) {
            for (int32_t w = p; w != TSG_NULL; w = gpar[w]) {
                path_g.push_back(w);
                touch(w);
            }
            ev_p0.push_back((int64_t) path_g.size());
        };
        const int32_t t0 = hp.chunk_tree0[c], t1 = hp.chunk_tree0[c + 1];
        const int32_t tlast = std::min(t1, T - 1);  // t1 is the closing transition
        const int64_t ev_base = (int64_t) out.ev_child.size();
        for (int32_t t = t0 + 1; t <= tlast; t++) {
            hp.tr_off[t] = ev_base + (int64_t) ev_u.size();
            const double x = ts->breakpoints[t];
            int32_t nrem = 0;
            while (k < E && ts->right[ts->O[k]] == x) {
                const int32_t e = ts->O[k];
                const int32_t u = ts->child[e];
                ev_u.push_back(u);
                ev_x.push_back(x);
                ev_ins.push_back(0);
                out.ev_bl.push_back(0.0f);
                touch(u);
                undo.emplace_back(u, gpar[u]);
                gpar[u] = TSG_NULL;
                redo.emplace_back(u, TSG_NULL);
                record_path(ts->parent[e]);
                nrem++;
                k++;
            }
            while (j < E && ts->left[ts->I[j]] == x) {
                const int32_t e = ts->I[j];
                const int32_t u = ts->child[e], p = ts->parent[e];
                ev_u.push_back(u);
                ev_x.push_back(x);
                ev_ins.push_back(1);
                out.ev_bl.push_back(ts->time[p] - ts->time[u]);
                touch(u);
                undo.emplace_back(u, gpar[u]);
                gpar[u] = p;
                redo.emplace_back(u, p);
                record_path(p);
                j++;
            }
            hp.tr_nrem[t] = nrem;
        }
        for (int32_t s = ts->tree_site0[t0]; s < ts->tree_site0[t1]; s++) {
            for (int32_t m = ts->site_mut0[s]; m < ts->site_mut0[s + 1]; m++) touch(ts->mut_node[m]);
        }
        // back to the start tree
        for (size_t i = undo.size(); i-- > 0;) gpar[undo[i].first] = undo[i].second;

        const int32_t H = (int32_t) hot.size();
        if (H >= (int32_t) TSG_NO_SLOT) {  // slot ids are 16-bit in sample_slot
            out.too_big = true;
            return;
        }
        // hot forest: nearest hot proper ancestor in the start tree
        hpar.assign(H, -1);
        for (int32_t i = 0; i < H; i++) {
            int32_t w = gpar[hot[i]];
            while (w != TSG_NULL && hot_stamp[w] != c) w = gpar[w];
            hpar[i] = (w == TSG_NULL) ? -1 : hot_idx[w];
        }
        depth.assign(H, -1);
        int32_t maxd = 0;
        for (int32_t i = 0; i < H; i++) {
            stack.clear();
            int32_t w = i;
            while (w >= 0 && depth[w] < 0) {
                stack.push_back(w);
                w = hpar[w];
            }
            int32_t d = (w >= 0) ? depth[w] : -1;
            for (size_t q = stack.size(); q-- > 0;) depth[stack[q]] = ++d;
            maxd = std::max(maxd, d);
        }
        // slots by decreasing depth (level 0 = deepest), stable within a level
        const int32_t nlev = (H > 0) ? maxd + 1 : 0;
        lev_count.assign(nlev + 1, 0);
        for (int32_t i = 0; i < H; i++) lev_count[maxd - depth[i] + 1]++;
        for (int32_t l = 0; l < nlev; l++) lev_count[l + 1] += lev_count[l];
        out.lev_slot0.insert(out.lev_slot0.end(), lev_count.begin(), lev_count.end());
        out.chunk_levs.push_back(nlev + 1);
        slot_of.assign(H, -1);
        next.assign(lev_count.begin(), lev_count.end() - 1);
        for (int32_t i = 0; i < H; i++) slot_of[i] = next[maxd - depth[i]]++;
        const size_t s0 = out.slot_par.size();
        out.slot_par.resize(s0 + H);
        out.slot_bl0.resize(s0 + H);
        out.slot_tp0.resize(s0 + H);
        out.slot_node.resize(s0 + H);
        out.slot_last0.resize(s0 + H);
        out.slot_ins0.resize(s0 + H);
        out.slot_nsd.resize(s0 + H, 0);
        out.chunk_slots.push_back(H);
        for (int32_t i = 0; i < H; i++) {
            const int32_t u = hot[i];
            const size_t sl = s0 + slot_of[i];
            out.slot_par[sl] = (hpar[i] < 0) ? -1 : slot_of[hpar[i]];
            out.slot_bl0[sl] = (gpar[u] == TSG_NULL) ? 0.0 : ts->time[gpar[u]] - ts->time[u];
            out.slot_tp0[sl] = (gpar[u] == TSG_NULL) ? -INFINITY : ts->time[gpar[u]];
            out.slot_node[sl] = u;
            // the node's state (parent, subtree) has been constant since its
            // last touch by an event, before this chunk
            out.slot_last0[sl] = last_touch[u];
            out.slot_ins0[sl] = ins_at[u];
        }
        for (int32_t u : ev_u) out.ev_child.push_back(slot_of[hot_idx[u]]);
        const int64_t pbase = (int64_t) out.path.size();
        for (size_t e = 1; e < ev_p0.size(); e++) {
            out.ev_end.push_back(pbase + ev_p0[e]);
            out.max_path = std::max(out.max_path, (int32_t) (ev_p0[e] - ev_p0[e - 1]));
        }
        for (int32_t w : path_g) out.path.push_back(slot_of[hot_idx[w]]);
        for (int32_t s = ts->tree_site0[t0]; s < ts->tree_site0[t1]; s++) {
            for (int32_t m = ts->site_mut0[s]; m < ts->site_mut0[s + 1]; m++) {
                hp.mut_slot[m] = slot_of[hot_idx[ts->mut_node[m]]];
            }
        }
        // each sample's nearest hot ancestor-or-self, memoised along the walk
        uint16_t *ss = hp.sample_slot.data() + (size_t) c * n;
        for (int32_t s = 0; s < n; s++) {
            stack.clear();
            int32_t w = ts->samples[s];
            while (w != TSG_NULL && hot_stamp[w] != c && memo_stamp[w] != c) {
                stack.push_back(w);
                w = gpar[w];
            }
            const int32_t res = (w == TSG_NULL) ? TSG_NULL : (hot_stamp[w] == c ? w : memo_nha[w]);
            for (int32_t v : stack) {
                memo_stamp[v] = c;
                memo_nha[v] = res;
            }
            ss[s] = (res == TSG_NULL) ? (uint16_t) TSG_NO_SLOT : (uint16_t) slot_of[hot_idx[res]];
            if (res != TSG_NULL) out.slot_nsd[s0 + ss[s]]++;
        }
        // this chunk's flushes and insertions, in event order (the last
        // one wins); an inserted edge's child is not flushed (trees.c:3765)
        for (size_t e = 0; e < ev_u.size(); e++) {
            const double x = ev_x[e];
            auto note = [&](int32_t w) {
                if (last_touch[w] < 0) out.touched.push_back(w);
                last_touch[w] = x;
            };
            if (ev_ins[e]) {
                if (ins_at[ev_u[e]] < 0) out.inserted.push_back(ev_u[e]);
                ins_at[ev_u[e]] = x;
            } else {
                note(ev_u[e]);
            }
            for (int64_t q = ev_p0[e]; q < ev_p0[e + 1]; q++) note(path_g[q]);
        }
        out.max_slots = std::max(out.max_slots, H);
        out.max_levels = std::max(out.max_levels, nlev);
        // forward to the next chunk's start tree
        for (const auto &r : redo) gpar[r.first] = r.second;
    }
    out.touched_last.resize(out.touched.size());
    for (size_t i = 0; i < out.touched.size(); i++) out.touched_last[i] = last_touch[out.touched[i]];
    out.inserted_at.resize(out.inserted.size());
    for (size_t i = 0; i < out.inserted.size(); i++) out.inserted_at[i] = ins_at[out.inserted[i]];
    if (cend == hp.num_chunks) out.end_par = gpar;
}

}  // namespace

// Builds the hot plan (tsg_hotplan) over the same kind of chunks.  Each chunk
// is planned in three steps on the live tree: record its events and root
// paths while applying them, undo them to return to the chunk's start tree
// and derive the slot structure there, then apply them again.  Runs of
// chunks are planned by separate threads, each replaying the edges up to its
// first chunk.
namespace {
int build_hot_plan_bounds(tsg_treeseq *ts, std::vector<int32_t> bounds) {
    tsg_hotplan &hp = ts->hot;
    hp = tsg_hotplan();
    const int32_t T = ts->num_trees, E = ts->num_edges, N = ts->num_nodes;
    const int32_t n = ts->num_samples;
    hp.chunk_tree0 = std::move(bounds);
    const int32_t C = (int32_t) hp.chunk_tree0.size() - 1;
    hp.num_chunks = C;
    hp.tr_off.assign(T + 1, 0);
    hp.tr_nrem.assign(T, 0);
    hp.mut_slot.assign(ts->num_mutations, -1);
    hp.sample_slot.assign((size_t) C * n, (uint16_t) TSG_NO_SLOT);

    // the first tree, samples first
    {
        std::vector<int32_t> gpar(N, TSG_NULL);
        for (int32_t j = 0; j < E && ts->left[ts->I[j]] == ts->breakpoints[0]; j++) {
            gpar[ts->child[ts->I[j]]] = ts->parent[ts->I[j]];
        }
        std::vector<int32_t> lid(N, -1), order;
        for (int32_t s = 0; s < n; s++) {
            lid[ts->samples[s]] = (int32_t) order.size();
            order.push_back(ts->samples[s]);
        }
        for (int32_t u = 0; u < N; u++) {
            if (gpar[u] == TSG_NULL) continue;
            for (int32_t w : {u, gpar[u]}) {
                if (lid[w] < 0) {
                    lid[w] = (int32_t) order.size();
                    order.push_back(w);
                }
            }
        }
        hp.t0_parent.assign(order.size(), TSG_NULL);
        hp.t0_bl.assign(order.size(), 0.0);
        for (size_t i = 0; i < order.size(); i++) {
            const int32_t u = order[i];
            if (gpar[u] != TSG_NULL) {
                hp.t0_parent[i] = lid[gpar[u]];
                hp.t0_bl[i] = ts->time[gpar[u]] - ts->time[u];
            }
        }
    }

    // runs of chunks per thread: enough work per thread to amortise its
    // replay of the edges before its first chunk
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int32_t P = (int32_t) std::max<int64_t>(
        1, std::min<int64_t>({(int64_t) std::min(hw, 16u), (int64_t) C, (int64_t) C / 16 + 1}));
    std::vector<int32_t> cb(P + 1);
    for (int32_t p = 0; p <= P; p++) cb[p] = (int32_t) ((int64_t) C * p / P);
    std::vector<HotPart> parts(P);
    {
        std::vector<std::thread> pool;
        for (int32_t p = 1; p < P; p++) {
            pool.emplace_back(plan_hot_range, ts, std::ref(hp), cb[p], cb[p + 1], std::ref(parts[p]));
        }
        plan_hot_range(ts, hp, cb[0], cb[1], parts[0]);
        for (auto &th : pool) th.join();
    }
    for (const HotPart &part : parts) {
        if (part.too_big) return 0;  // not representable: hp.usable stays false
    }
    // merge in chunk order
    std::vector<double> glast(N, 0.0);  // last event flush of each node so far
    std::vector<double> gins(N, 0.0);   // position of its current edge's insertion
    hp.chunk_slot0.assign(C + 1, 0);
    hp.chunk_lev0.assign(C + 1, 0);
    hp.ev_path0.assign(1, 0);
    for (int32_t p = 0; p < P; p++) {
        const HotPart &part = parts[p];
        const int64_t ev_base = (int64_t) hp.ev_child.size();
        const int64_t path_base = (int64_t) hp.path.size();
        const int32_t tfirst = hp.chunk_tree0[cb[p]] + 1;
        const int32_t tlast = std::min(hp.chunk_tree0[cb[p + 1]], T - 1);
        for (int32_t t = tfirst; t <= tlast; t++) hp.tr_off[t] += ev_base;
        hp.ev_child.insert(hp.ev_child.end(), part.ev_child.begin(), part.ev_child.end());
        hp.ev_bl.insert(hp.ev_bl.end(), part.ev_bl.begin(), part.ev_bl.end());
        for (int64_t e : part.ev_end) hp.ev_path0.push_back(path_base + e);
        hp.path.insert(hp.path.end(), part.path.begin(), part.path.end());
        hp.slot_par.insert(hp.slot_par.end(), part.slot_par.begin(), part.slot_par.end());
        hp.slot_bl0.insert(hp.slot_bl0.end(), part.slot_bl0.begin(), part.slot_bl0.end());
        hp.slot_tp0.insert(hp.slot_tp0.end(), part.slot_tp0.begin(), part.slot_tp0.end());
        hp.slot_node.insert(hp.slot_node.end(), part.slot_node.begin(), part.slot_node.end());
        hp.slot_nsd.insert(hp.slot_nsd.end(), part.slot_nsd.begin(), part.slot_nsd.end());
        // a slot whose node was last touched before this part: its last
        // touch in the earlier parts (0: never touched, alive since x = 0)
        for (size_t q = 0; q < part.slot_last0.size(); q++) {
            const double l = part.slot_last0[q], y = part.slot_ins0[q];
            hp.slot_last0.push_back(l >= 0 ? l : glast[part.slot_node[q]]);
            hp.slot_ins0.push_back(y >= 0 ? y : gins[part.slot_node[q]]);
        }
        for (size_t q = 0; q < part.touched.size(); q++) glast[part.touched[q]] = part.touched_last[q];
        for (size_t q = 0; q < part.inserted.size(); q++) gins[part.inserted[q]] = part.inserted_at[q];
        hp.lev_slot0.insert(hp.lev_slot0.end(), part.lev_slot0.begin(), part.lev_slot0.end());
        for (int32_t c = cb[p]; c < cb[p + 1]; c++) {
            hp.chunk_slot0[c + 1] = hp.chunk_slot0[c] + part.chunk_slots[c - cb[p]];
            hp.chunk_lev0[c + 1] = hp.chunk_lev0[c] + part.chunk_levs[c - cb[p]];
        }
        hp.max_slots = std::max(hp.max_slots, part.max_slots);
        hp.max_path = std::max(hp.max_path, part.max_path);
        hp.max_levels = std::max(hp.max_levels, part.max_levels);
    }
    hp.tr_off[T] = (int64_t) hp.ev_child.size();
    // the last tree, samples first, with each node's last touch: the state
    // intervals still open at the end of the sequence
    {
        const std::vector<int32_t> &gpar = parts[P - 1].end_par;
        std::vector<int32_t> lid(N, -1), order;
        for (int32_t s2 = 0; s2 < n; s2++) {
            lid[ts->samples[s2]] = (int32_t) order.size();
            order.push_back(ts->samples[s2]);
        }
        for (int32_t u = 0; u < N; u++) {
            if (gpar[u] == TSG_NULL) continue;
            for (int32_t w : {u, gpar[u]}) {
                if (lid[w] < 0) {
                    lid[w] = (int32_t) order.size();
                    order.push_back(w);
                }
            }
        }
        const size_t nl = order.size();
        hp.tl_parent.assign(nl, TSG_NULL);
        hp.tl_time.assign(nl, 0.0);
        hp.tl_last.assign(nl, 0.0);
        hp.tl_ins.assign(nl, 0.0);
        for (size_t q = 0; q < nl; q++) {
            const int32_t u = order[q];
            hp.tl_time[q] = ts->time[u];
            hp.tl_last[q] = glast[u];
            hp.tl_ins[q] = gins[u];
            if (gpar[u] != TSG_NULL) hp.tl_parent[q] = lid[gpar[u]];
        }
        hp.node_last = glast;
        hp.node_lid = lid;
    }
    hp.usable = true;
    return 0;
}
}  // namespace

// Chunks are first cut by edge events (about equal); with balance, chunks
// whose hot slots exceed 1.15x the mean are cut again (at tree boundaries,
// by events) and the plan rebuilt: the sweep's LDS per workgroup is set by
// the largest chunk, so trimming the tail of the slot distribution lets more
// chunks share a CU.
int tsg_build_hot_plan(tsg_treeseq *ts, int32_t num_chunks, bool balance) {
    int ret = build_hot_plan_bounds(ts, chunk_bounds(ts, num_chunks));
    tsg_hotplan &hp = ts->hot;
    if (ret != 0 || !balance || !hp.usable || hp.num_chunks < 2) return ret;
    const int32_t C = hp.num_chunks, T = ts->num_trees;
    const double mean = (double) hp.chunk_slot0[C] / C;
    const double target = std::max(1.15 * mean, 64.0);
    if ((double) hp.max_slots <= 1.05 * target) return 0;
    std::vector<int32_t> nb{0};
    for (int32_t c = 0; c < C; c++) {
        const int32_t t0 = hp.chunk_tree0[c], t1 = hp.chunk_tree0[c + 1];
        const double slots = (double) (hp.chunk_slot0[c + 1] - hp.chunk_slot0[c]);
        const int32_t parts = std::min<int32_t>(t1 - t0, (int32_t) std::ceil(slots / target));
        if (parts > 1) {
            // events of the chunk's transitions t0+1 .. t1 (t1: the closing one)
            const int32_t tl = std::min(t1, T - 1);
            const int64_t ev = hp.tr_off[tl + 1] - hp.tr_off[t0 + 1];
            int32_t made = 1;
            for (int32_t t = t0 + 1; t < t1 && made < parts; t++) {
                // cut before tree t when the events up to transition t reach
                // the next share
                if ((hp.tr_off[t + 1] - hp.tr_off[t0 + 1]) * parts >= ev * made) {
                    nb.push_back(t);
                    made++;
                }
            }
        }
        nb.push_back(t1);
    }
    const size_t nc = nb.size() - 1;
    if ((int64_t) nc == C || (double) nc * ts->num_samples * 2 > (double) (5ll << 30)) return 0;
    return build_hot_plan_bounds(ts, std::move(nb));
}
