// Host-side tree sequence and the per-chunk sweep programs uploaded to the GPU.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "tsg_common.h"

// Query-independent data the sweep kernels read, resident in HBM.
struct TsgDevice {
    int device = -1;
    // per chunk [C+1]
    int32_t *chunk_tree0 = nullptr;   // first tree of each chunk (last entry = T)
    int64_t *chunk_node0 = nullptr;   // offset of each chunk's local-node block
    // per local node [sum nloc]
    int32_t *loc_parent0 = nullptr;   // parent (local id) in the chunk's start tree
    double *loc_bl0 = nullptr;         // branch length in the start tree (0 if root)
    double *loc_time = nullptr;       // node time
    double *loc_lastrem = nullptr;    // position of the node's last edge removal
    // per transition [T+1]: events of transition t are [tr_off[t], tr_off[t+1]),
    // the first tr_nrem[t] of them removals (in removal-index order), then
    // insertions (in insertion-index order)
    int64_t *tr_off = nullptr;
    int32_t *tr_nrem = nullptr;
    int32_t *ev_child = nullptr;      // local ids in the owning chunk
    int32_t *ev_parent = nullptr;
    double *ev_bl = nullptr;          // branch length of inserted edge
    int64_t *ev_path0 = nullptr;      // [events + 1] offsets into path
    int32_t *path = nullptr;          // ancestors of each event's parent end
    double *breakpoints = nullptr;    // [T+1]
    int32_t *tree_site0 = nullptr;    // [T+1]
    double *site_pos = nullptr;       // [S]
    int32_t *site_mut0 = nullptr;     // [S+1]
    int32_t *mut_node = nullptr;      // [Mu] local id in owning chunk
    int32_t *mut_allele = nullptr;    // [Mu] allele index of the derived state
    int32_t *mut_pallele = nullptr;   // [Mu] allele index it replaces
    int32_t *site_nalleles = nullptr; // [S]
    size_t bytes = 0;
};

// Device copy of a hot plan (see tsg_hotplan).
struct TsgHotDevice {
    int32_t *chunk_tree0 = nullptr;
    int64_t *chunk_slot0 = nullptr;
    int32_t *slot_par = nullptr;
    double *slot_bl0 = nullptr;
    int64_t *chunk_lev0 = nullptr;
    int32_t *lev_slot0 = nullptr;
    uint16_t *sample_slot = nullptr;
    int64_t *tr_off = nullptr;
    int32_t *tr_nrem = nullptr;
    int32_t *ev_child = nullptr;
    double *ev_bl = nullptr;
    int64_t *ev_path0 = nullptr;
    int32_t *path = nullptr;
    int32_t *mut_slot = nullptr;
    int32_t *t0_parent = nullptr;
    double *t0_bl = nullptr;
    // allele frequency spectrum tables (uploaded on first use)
    int32_t *slot_nsd = nullptr;
    double *slot_tu = nullptr;
    double *slot_last0 = nullptr;
    double *slot_ins0 = nullptr;
    double *slot_tp0 = nullptr;
    int32_t *tl_parent = nullptr;
    double *tl_time = nullptr;
    double *tl_last = nullptr;
    double *tl_ins = nullptr;
    // node statistics tables (uploaded on first use)
    int32_t *slot_node = nullptr;
    double *node_last = nullptr;
    int32_t *node_lid = nullptr;
    size_t bytes = 0;
};

// Sweep program whose chunks keep only the nodes they touch.  A node is hot
// in chunk c when an event of c names it (edge child, or a node on the root
// path of the edge's parent end) or a mutation of c's sites sits on it; only
// hot nodes are ever read or written by c's sweep, so the kernel holds their
// counts and branch lengths in LDS ("slots").  The start tree's counts of the
// hot nodes are rebuilt from the samples without walking the tree:
//   * sample_slot[c*n + s]: slot of sample s's nearest hot ancestor-or-self
//     in c's start tree (TSG_NO_SLOT if none) -- every sample adds its weight
//     there;
//   * slot_par: the slot's nearest hot proper ancestor in the start tree
//     (-1 if none); slots are numbered by decreasing depth in that hot forest,
//     one contiguous range per depth level (lev_slot0), so adding each level
//     into its parents, deepest level first, completes the subtree counts.
// Event, path and mutation node ids are slot ids.
struct tsg_hotplan {
    bool usable = false;                  // false: a chunk exceeds the slot id range
    int32_t num_chunks = 0;
    std::vector<int32_t> chunk_tree0;     // [C+1]
    std::vector<int64_t> chunk_slot0;     // [C+1]
    std::vector<int32_t> slot_par;        // [slots] chunk-local slot or -1
    std::vector<double> slot_bl0;         // [slots] start-tree branch length (0 if root)
    std::vector<int64_t> chunk_lev0;      // [C+1] offsets into lev_slot0
    std::vector<int32_t> lev_slot0;       // per chunk: level boundaries (nlev + 1 entries)
    std::vector<uint16_t> sample_slot;    // [C * n]
    std::vector<int64_t> tr_off;          // [T+1]
    std::vector<int32_t> tr_nrem;         // [T]
    std::vector<int32_t> ev_child;        // slot ids
    std::vector<double> ev_bl;
    std::vector<int64_t> ev_path0;        // [events + 1]
    std::vector<int32_t> path;            // slot ids, parent end first
    std::vector<int32_t> mut_slot;        // [Mu]
    // the first tree, samples first (local id j = sample j): its running
    // branch sum is evaluated over all of its nodes
    std::vector<int32_t> t0_parent;
    std::vector<double> t0_bl;
    // for the branch allele frequency spectrum, per slot: its node, the
    // number of samples whose nearest hot ancestor-or-self it is, the
    // position of the node's last flush by an event before the chunk (as a
    // node on an event's root path or a removed edge's child: its state is
    // constant since then) and of the insertion of its current parent edge;
    // and the last tree, samples first, with node times and the same two
    // positions per node
    std::vector<int32_t> slot_node, slot_nsd;
    std::vector<double> slot_last0, slot_ins0;
    std::vector<double> slot_tp0;  // time of the node's parent, -inf: none
    std::vector<int32_t> tl_parent;
    std::vector<double> tl_time, tl_last, tl_ins;
    // every node: its last flush position (as slot_last0) over the whole
    // sequence, and its id in the last tree's arrays (-1: not in it)
    std::vector<double> node_last;
    std::vector<int32_t> node_lid;
    int32_t max_slots = 0, max_path = 0, max_levels = 0;
    TsgHotDevice dev;
};

struct tsg_treeseq {
    // tables (host copies)
    double L = 0;
    uint32_t options = 0;
    int32_t num_nodes = 0, num_edges = 0, num_sites = 0, num_mutations = 0;
    int32_t num_samples = 0, num_trees = 0;
    std::vector<uint32_t> flags;
    std::vector<double> time;
    std::vector<double> left, right;
    std::vector<int32_t> parent, child;
    std::vector<int32_t> I, O;
    std::vector<double> site_pos;
    std::vector<int32_t> mut_site, mut_node, mut_parent;
    std::vector<int32_t> samples, sample_index_map;
    std::vector<double> breakpoints;      // [T+1]
    std::vector<int32_t> tree_site0;      // [T+1]
    std::vector<int32_t> site_mut0;       // [S+1]
    std::vector<int32_t> mut_allele, mut_pallele, site_nalleles;
    int32_t max_alleles = 1;
    bool time_uncalibrated = false;

    // chunk plan
    int32_t num_chunks = 0;
    std::vector<int32_t> chunk_tree0;     // [C+1]
    std::vector<int64_t> chunk_node0;     // [C+1]
    int32_t max_nloc = 0;
    int32_t max_chunk_events = 0;
    std::vector<int32_t> loc_parent0;
    std::vector<double> loc_bl0;
    std::vector<double> loc_time;
    std::vector<double> loc_lastrem;
    std::vector<int64_t> tr_off;          // [T+1]
    std::vector<int32_t> tr_nrem;         // [T]
    std::vector<int32_t> ev_child, ev_parent;
    std::vector<double> ev_bl;
    // the root path of each event's parent edge end at the time of the event
    // (local ids, parent first): events of a sweep touch exactly these nodes
    std::vector<int64_t> ev_path0;        // [events + 1]
    std::vector<int32_t> path;
    int32_t max_path = 0;
    std::vector<int32_t> mut_node_loc;
    int64_t total_path_hops = 0;          // topology-only measure: sum of path lengths

    bool full_built = false;              // the full plan above is built on first need
    int32_t full_chunks = 1;              // chunk count it is built with
    tsg_hotplan hot;

    TsgDevice dev;
    void *workspace = nullptr;            // per-handle device scratch (tsg_api.cpp)
    // statistics calls in flight on the handle, and whether a call that
    // rebuilds or frees its plans is waiting for them (guarded by the
    // library's mutex, tsg_api.cpp)
    mutable int32_t calls = 0;
    mutable bool exclusive = false;
};

int tsg_build_plan(tsg_treeseq *ts, int32_t num_chunks);
int tsg_build_hot_plan(tsg_treeseq *ts, int32_t num_chunks, bool balance);
int tsg_upload(tsg_treeseq *ts);
void tsg_device_free(TsgDevice *d);
