// Synthetic tree-sequence generator: Hudson's coalescent with recombination
// (the model msprime's sim_ancestry simulates) plus infinite-sites 0/1
// mutations.  It exists because the statistics path needs realistic inputs at
// the sizes of the benchmark workloads and msprime is not available.
//
// Model and conventions
//   * n haploid samples (nodes 0..n-1, time 0, flag SAMPLE).
//   * A pair of lineages coalesces at rate 1/(2 Ne) per generation (the
//     diploid time scale of msprime's default ploidy 2); recombination happens
//     at rate r per unit of genome per generation between integer positions
//     (a discrete genome: breakpoints are integers, as sim_ancestry's
//     default discrete_genome=True).
//   * Only ancestral material is tracked; a genome region is dropped once all
//     of its sample lineages have coalesced, so the output is simplified
//     (no unary nodes, edges squashed, sorted as the reference's table sort
//     requires: (time[parent], parent, child, left)).
//   * Mutations: for each edge, a Poisson process along the edge's span with
//     intensity mu * branch_length (positions continuous and uniform), one
//     site per mutation, ancestral "0" / derived "1".
//
// Determinism across hosts: all randomness comes from xoshiro256** seeded by
// splitmix64; exponential draws use det_log() below (basic IEEE operations
// only, no libm), and recombination mass is integer, so the same seed gives
// bit-identical tables on any x86-64 host.
#pragma clang fp contract(off)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "tsg_common.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            s[i] = x ^ (x >> 31);
        }
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // uniform in [0, 1) with 53 random bits
    double uniform() { return (double) (next() >> 11) * 0x1.0p-53; }
    // uniform integer in [0, n), n > 0 (Lemire's multiply-shift with rejection)
    uint64_t below(uint64_t n) {
        unsigned __int128 m = (unsigned __int128) next() * n;
        uint64_t lo = (uint64_t) m;
        if (lo < n) {
            uint64_t t = (0 - n) % n;
            while (lo < t) {
                m = (unsigned __int128) next() * n;
                lo = (uint64_t) m;
            }
        }
        return (uint64_t) (m >> 64);
    }
};

// Natural log for x in (0, 1], from frexp + atanh series; only +,-,*,/ so the
// result is the same on every IEEE-754 host.
double det_log(double x) {
    int e;
    double m = std::frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double poly = 1.0 / 23.0;
    for (int k = 21; k >= 1; k -= 2) {
        poly = poly * s2 + 1.0 / (double) k;
    }
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return (double) e * ln2_hi + ((double) e * ln2_lo + 2.0 * s * poly);
}

double exp1(Rng &rng) { return -det_log(1.0 - rng.uniform()); }

struct Fenwick {
    std::vector<int64_t> tree;  // 1-based
    std::vector<int64_t> value;
    size_t size = 0;
    int64_t total = 0;
    void resize(size_t n) {
        std::vector<int64_t> old = value;
        size = 1;
        while (size < n) size <<= 1;
        tree.assign(size + 1, 0);
        value.assign(size, 0);
        total = 0;
        for (size_t i = 0; i < old.size(); i++) {
            if (old[i] != 0) set(i, old[i]);
        }
    }
    void set(size_t i, int64_t v) {
        int64_t d = v - value[i];
        if (d == 0) return;
        value[i] = v;
        total += d;
        for (size_t j = i + 1; j <= size; j += j & (~j + 1)) tree[j] += d;
    }
    // smallest index i with prefix(i) >= h (h in [1, total]); rem in [1, value[i]]
    size_t find(int64_t h, int64_t *rem) const {
        size_t pos = 0;
        for (size_t step = size; step > 0; step >>= 1) {
            if (pos + step <= size && tree[pos + step] < h) {
                pos += step;
                h -= tree[pos];
            }
        }
        *rem = h;
        return pos;
    }
};

struct Segment {
    int64_t left, right;
    int32_t node, prev, next;
};

struct RawEdge {
    int64_t left, right;
    int32_t parent, child;
};

struct Hudson {
    Rng rng;
    int64_t L;
    double coal_rate_per_pair;
    double rec_rate;
    std::vector<Segment> seg;
    std::vector<int32_t> free_list;
    Fenwick mass;
    std::vector<int32_t> lineages;
    std::map<int64_t, int32_t> overlap;
    std::vector<double> node_time;
    std::vector<RawEdge> edges;

    Hudson(uint64_t seed, int64_t L_, double Ne, double r)
        : rng(seed), L(L_), coal_rate_per_pair(1.0 / (2.0 * Ne)), rec_rate(r) {}

    int32_t alloc(int64_t l, int64_t r, int32_t node) {
        int32_t s;
        if (!free_list.empty()) {
            s = free_list.back();
            free_list.pop_back();
        } else {
            s = (int32_t) seg.size();
            seg.push_back(Segment{});
            if (seg.size() > mass.size) mass.resize(seg.size() * 2);
        }
        seg[s] = Segment{l, r, node, TSG_NULL, TSG_NULL};
        return s;
    }
    void release(int32_t s) {
        mass.set((size_t) s, 0);
        free_list.push_back(s);
    }
    // number of breakpoints k (split into [<k] and [>=k]) owned by segment s
    void set_mass(int32_t s) {
        const Segment &x = seg[s];
        int64_t m = (x.prev == TSG_NULL) ? (x.right - x.left - 1) : (x.right - seg[x.prev].right);
        mass.set((size_t) s, m);
    }

    void recombine() {
        int64_t h = (int64_t) rng.below((uint64_t) mass.total) + 1;
        int64_t j;
        int32_t x = (int32_t) mass.find(h, &j);
        int64_t k;
        if (seg[x].prev == TSG_NULL) {
            k = seg[x].left + j;
        } else {
            k = seg[seg[x].prev].right + j - 1;
        }
        if (k > seg[x].left) {
            int32_t z = alloc(k, seg[x].right, seg[x].node);
            seg[z].next = seg[x].next;
            if (seg[z].next != TSG_NULL) seg[seg[z].next].prev = z;
            seg[x].right = k;
            seg[x].next = TSG_NULL;
            set_mass(x);
            set_mass(z);
            lineages.push_back(z);
        } else {
            int32_t y = seg[x].prev;
            seg[y].next = TSG_NULL;
            seg[x].prev = TSG_NULL;
            set_mass(x);
            lineages.push_back(x);
        }
    }

    size_t event_edge0 = 0;
    // Records edge (l, r, parent, child), extending an abutting edge of the
    // same child made earlier in this coalescence event.
    void add_edge(int64_t l, int64_t r, int32_t parent, int32_t child) {
        for (size_t i = edges.size(); i-- > event_edge0;) {
            if (edges[i].child == child && edges[i].right == l) {
                edges[i].right = r;
                return;
            }
        }
        edges.push_back(RawEdge{l, r, parent, child});
    }

    // Merge two lineages' segment chains at time t (one coalescence event).
    int32_t merge(int32_t x, int32_t y, double t) {
        int32_t v = TSG_NULL;
        event_edge0 = edges.size();
        int32_t head = TSG_NULL, tail = TSG_NULL;
        auto append = [&](int32_t s) {
            if (tail != TSG_NULL && seg[tail].right == seg[s].left && seg[tail].node == seg[s].node) {
                seg[tail].right = seg[s].right;
                release(s);
                return;
            }
            seg[s].prev = tail;
            seg[s].next = TSG_NULL;
            if (tail == TSG_NULL) {
                head = s;
            } else {
                seg[tail].next = s;
            }
            tail = s;
        };
        while (x != TSG_NULL || y != TSG_NULL) {
            if (x == TSG_NULL || y == TSG_NULL) {
                int32_t z = (x != TSG_NULL) ? x : y;
                while (z != TSG_NULL) {
                    int32_t nz = seg[z].next;
                    append(z);
                    z = nz;
                }
                break;
            }
            if (seg[y].left < seg[x].left) std::swap(x, y);
            if (seg[x].right <= seg[y].left) {
                int32_t nx = seg[x].next;
                append(x);
                x = nx;
            } else if (seg[x].left < seg[y].left) {
                int32_t a = alloc(seg[x].left, seg[y].left, seg[x].node);
                seg[x].left = seg[y].left;
                append(a);
            } else {
                const int64_t l = seg[x].left;
                const int64_t rmax = std::min(seg[x].right, seg[y].right);
                if (v == TSG_NULL) {
                    v = (int32_t) node_time.size();
                    node_time.push_back(t);
                }
                const int64_t r = rmax;
                add_edge(l, r, v, seg[x].node);
                add_edge(l, r, v, seg[y].node);
                // Decrement the lineage count over [l, r).  Sub-intervals whose
                // count drops to one have found their MRCA and are dropped;
                // the rest is carried forward by the new node v.
                auto it = overlap.upper_bound(l);
                --it;
                if (it->first < l) {
                    it = overlap.emplace_hint(std::next(it), l, it->second);
                }
                {
                    auto e = overlap.lower_bound(r);
                    if (e->first > r) {
                        overlap.emplace_hint(e, r, std::prev(e)->second);
                    }
                }
                while (it->first < r) {
                    auto nxt = std::next(it);
                    const int32_t c = it->second;
                    if (c == 2) {
                        it->second = 0;
                    } else {
                        it->second = c - 1;
                        append(alloc(it->first, nxt->first, v));
                    }
                    // keep the overlap map small: drop a key equal to its predecessor
                    if (it != overlap.begin() && std::prev(it)->second == it->second) {
                        overlap.erase(it);
                    }
                    it = nxt;
                }
                if (it != overlap.begin() && it->first < L &&
                    std::prev(it)->second == it->second) {
                    overlap.erase(it);
                }
                if (seg[x].right == r) {
                    int32_t nx = seg[x].next;
                    release(x);
                    x = nx;
                } else {
                    seg[x].left = r;
                }
                if (seg[y].right == r) {
                    int32_t ny = seg[y].next;
                    release(y);
                    y = ny;
                } else {
                    seg[y].left = r;
                }
            }
        }
        for (int32_t s = head; s != TSG_NULL; s = seg[s].next) set_mass(s);
        return head;
    }

    void coalesce(double t) {
        const uint64_t k = lineages.size();
        uint64_t i = rng.below(k);
        uint64_t j = rng.below(k - 1);
        if (j >= i) j++;
        int32_t a = lineages[i], b = lineages[j];
        // remove i and j (larger index first)
        if (i < j) std::swap(i, j);
        lineages[i] = lineages.back();
        lineages.pop_back();
        lineages[j] = lineages.back();
        lineages.pop_back();
        int32_t h = merge(a, b, t);
        if (h != TSG_NULL) lineages.push_back(h);
    }

    int run(uint32_t n) {
        mass.resize(2 * (size_t) n + 16);
        seg.reserve(4 * (size_t) n);
        for (uint32_t j = 0; j < n; j++) {
            node_time.push_back(0.0);
            int32_t s = alloc(0, L, (int32_t) j);
            set_mass(s);
            lineages.push_back(s);
        }
        overlap[0] = (int32_t) n;
        overlap[L] = -1;
        double t = 0;
        while (lineages.size() > 1) {
            const double k = (double) lineages.size();
            const double rate_c = coal_rate_per_pair * k * (k - 1) / 2.0;
            const double rate_r = rec_rate * (double) mass.total;
            const double total = rate_c + rate_r;
            t += exp1(rng) / total;
            if (rng.uniform() * total < rate_r) {
                recombine();
            } else {
                coalesce(t);
            }
        }
        if (!lineages.empty()) return TSG_ERR_GENERIC;  // unreachable for n >= 2
        return 0;
    }
};

}  // namespace

struct tsg_tables {
    double sequence_length;
    std::vector<uint32_t> node_flags;
    std::vector<double> node_time;
    std::vector<double> left, right;
    std::vector<int32_t> parent, child;
    std::vector<double> site_position;
    std::vector<int32_t> mut_site, mut_node, mut_parent;
    std::vector<double> mut_time;
};

extern "C" {

// Simulates ancestry; writes a new table set to *out (free with
// tsg_tables_free).  Returns 0 or a negative error number.
int tsg_sim_ancestry(uint32_t num_samples, double sequence_length, double population_size,
                     double recombination_rate, uint64_t seed, tsg_tables **out) {
    *out = nullptr;
    if (num_samples < 2 || !(sequence_length >= 2) || sequence_length != std::floor(sequence_length)
        || !(population_size > 0) || !(recombination_rate >= 0)) {
        return TSG_ERR_BAD_PARAM_VALUE;
    }
    Hudson sim(seed, (int64_t) sequence_length, population_size, recombination_rate);
    int ret = sim.run(num_samples);
    if (ret != 0) return ret;

    // sort by (parent, child, left) -- parent ids are in time order -- then squash
    std::vector<RawEdge> &E = sim.edges;
    std::sort(E.begin(), E.end(), [](const RawEdge &a, const RawEdge &b) {
        if (a.parent != b.parent) return a.parent < b.parent;
        if (a.child != b.child) return a.child < b.child;
        return a.left < b.left;
    });
    tsg_tables *t = new tsg_tables();
    t->sequence_length = sequence_length;
    t->node_time = sim.node_time;
    t->node_flags.assign(sim.node_time.size(), 0u);
    for (uint32_t j = 0; j < num_samples; j++) t->node_flags[j] = TSG_NODE_IS_SAMPLE;
    for (size_t j = 0; j < E.size(); j++) {
        const RawEdge &e = E[j];
        size_t m = t->left.size();
        if (m > 0 && t->parent[m - 1] == e.parent && t->child[m - 1] == e.child
            && t->right[m - 1] == (double) e.left) {
            t->right[m - 1] = (double) e.right;
        } else {
            t->left.push_back((double) e.left);
            t->right.push_back((double) e.right);
            t->parent.push_back(e.parent);
            t->child.push_back(e.child);
        }
    }
    *out = t;
    return 0;
}

// Adds infinite-sites 0/1 mutations at rate mu per unit genome per
// generation.  Replaces any existing sites and mutations.
int tsg_sim_mutations(tsg_tables *t, double rate, uint64_t seed) {
    if (t == nullptr || !(rate >= 0)) return TSG_ERR_BAD_PARAM_VALUE;
    Rng rng(seed);
    struct M {
        double pos, time;
        int32_t node;
    };
    std::vector<M> muts;
    for (size_t j = 0; j < t->left.size(); j++) {
        const double tc = t->node_time[t->child[j]];
        const double tp = t->node_time[t->parent[j]];
        const double intensity = rate * (tp - tc);
        if (!(intensity > 0)) continue;
        double pos = t->left[j];
        while (true) {
            pos += exp1(rng) / intensity;
            if (!(pos < t->right[j])) break;
            muts.push_back(M{pos, tc + rng.uniform() * (tp - tc), t->child[j]});
        }
    }
    std::sort(muts.begin(), muts.end(), [](const M &a, const M &b) {
        if (a.pos != b.pos) return a.pos < b.pos;
        return a.node < b.node;
    });
    t->site_position.clear();
    t->mut_site.clear();
    t->mut_node.clear();
    t->mut_parent.clear();
    t->mut_time.clear();
    for (size_t j = 0; j < muts.size(); j++) {
        if (!t->site_position.empty() && t->site_position.back() == muts[j].pos) continue;
        t->mut_site.push_back((int32_t) t->site_position.size());
        t->site_position.push_back(muts[j].pos);
        t->mut_node.push_back(muts[j].node);
        t->mut_parent.push_back(TSG_NULL);
        t->mut_time.push_back(muts[j].time);
    }
    return 0;
}

// sizes[0..4) = nodes, edges, sites, mutations
int tsg_tables_get_sizes(const tsg_tables *t, uint64_t *sizes) {
    if (t == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
    sizes[0] = t->node_time.size();
    sizes[1] = t->left.size();
    sizes[2] = t->site_position.size();
    sizes[3] = t->mut_site.size();
    return 0;
}

// Copies the columns into caller buffers sized from tsg_tables_get_sizes.
int tsg_tables_copy_out(const tsg_tables *t, uint32_t *node_flags, double *node_time,
                        double *edge_left, double *edge_right, int32_t *edge_parent,
                        int32_t *edge_child, double *site_position, int32_t *mutation_site,
                        int32_t *mutation_node, int32_t *mutation_parent,
                        double *mutation_time) {
    if (t == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
    auto cp = [](auto *dst, const auto &v) {
        if (!v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(node_flags, t->node_flags);
    cp(node_time, t->node_time);
    cp(edge_left, t->left);
    cp(edge_right, t->right);
    cp(edge_parent, t->parent);
    cp(edge_child, t->child);
    cp(site_position, t->site_position);
    cp(mutation_site, t->mut_site);
    cp(mutation_node, t->mut_node);
    cp(mutation_parent, t->mut_parent);
    cp(mutation_time, t->mut_time);
    return 0;
}

void tsg_tables_free(tsg_tables *t) { delete t; }

}  // extern "C"
