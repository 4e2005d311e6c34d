// C ABI of the tsgpu library: device-resident tree sequences and the
// general_stat-family statistics computed by the HIP kernels.
//
// Every statistics entry point takes the same arguments, in the same order
// and with the same meaning, as the reference function it replaces
// (c/tskit/trees.h:1098-1239), with the const tsk_treeseq_t* replaced by an
// opaque device-resident handle.  Inputs are validated in the reference's
// order and the same error numbers are returned (trees.c:2035-2220,
// 3815-3928, 4667-4688).  The caller owns the result buffer (host memory,
// W x M doubles, C order); the callee zero-fills nothing the caller sees
// until the result is complete.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <unordered_set>
#include <thread>
#include <tuple>
#include <vector>

#include "tsg_common.h"
#include "tsg_launch.h"
#include "tsg_relvec.h"
#include "tsg_treeseq.h"

void tsg_build_index(const double *left, const double *right, const int32_t *parent,
                     const int32_t *child, const double *time, int32_t E, int32_t *I, int32_t *O);
int tsg_init_trees(tsg_treeseq *ts, const char *ancestral_state,
                   const uint64_t *ancestral_state_offset, const char *derived_state,
                   const uint64_t *derived_state_offset);

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Workspace {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf seed_off, seed_ps;
    DevBuf seedbuf;                 // seeded slot counts shared with the next run
    DevBuf pair_s, pair_c, tot, outs, queries, windows, cq_wfirst, cq_nwin, cq_poff, partial, dR,
        Rstart, raw, rawbase, lanebase, out, fin, ipar, dpar, slab, last, hist, dims, twin;
    // first tree: counts and per-block running-sum partials
    DevBuf x0, r0part, bsum;
    // one launch over several batches' sweeps: their argument blocks and
    // chunk prefixes; their finish phases (run blocks, block prefixes) and
    // results, copied back in one transfer
    DevBuf margs, mcpre, mfin, mpre, mout;
    void *mout_h = nullptr;
    size_t mout_h_bytes = 0;
    // a batch's host arrays, packed into one pinned buffer and sent with one
    // copy into one device buffer
    DevBuf arena;
    void *stage_h = nullptr;
    size_t stage_bytes = 0;
    // a batch's results, copied back in one transfer (pinned)
    void *out_h = nullptr;
    size_t out_h_bytes = 0;
    // large results into pageable host memory: two pinned pieces in turn
    // (copy_out)
    void *cp_h[2] = {nullptr, nullptr};
    hipEvent_t cp_ev[2] = {nullptr, nullptr};
    // relatedness-vector route
    DevBuf rv_E, rv_I, rv_z0, rv_d0, rv_esum, rv_y;
    // relatedness-matrix route
    DevBuf gbits, sitecol, gt, cmat, rowsum, scale, kout, wk0, wk1, sel, gslab, geno;
    // ... the rows of sites with three or more alleles
    DevBuf sitecol2, gt2, cmat2, wk0b, wk1b, row2;
    // no-allocation runs (tsg_batch_run_fixed): device buffers are carved
    // from a caller-provided region instead of allocated, and host arrays go
    // over from ordinary host memory
    char *fixed = nullptr;
    size_t fixed_cap = 0, fixed_used = 0;
    std::vector<char> fixed_stage;
    int lds_limit = 65536;          // dynamic LDS a sweep workgroup may use
    bool force_full = false;        // run sweeps on the full plan (testing hook)
    bool force_c32 = false;         // hot plan with 32-bit counts (testing hook)
    int last_store = -1;            // node store of the last sweep (TSG_STORE_*)
    // profiling of the dominant (sweep) kernel
    bool profile = false;
    double sweep_ms = 0;
    uint64_t sweep_launches = 0;
    double mode_ms[2] = {0, 0};     // branch, site sweeps
    uint64_t mode_launches[2] = {0, 0};
};

// Per resident tree sequence: a pool of workspaces, each with its own
// stream, so that queries on one tree sequence (or on several) issued from
// different host threads run concurrently on the device; and the settings
// every workspace takes when it is leased.
struct Resident {
    std::vector<Workspace *> all, idle;
    int lds_limit = 65536;
    bool force_full = false, force_c32 = false, profile = false;
    int last_store = -1;
};

// Guards device residency (uploads, plan builds), the workspace pools and
// settings, and the per-handle call counts; never held while a query runs.
std::mutex g_mutex;

// Statistics calls on one handle may run concurrently from several host
// threads (each leases its own workspace and stream).  A call holds a Use
// for its whole duration; the calls that rebuild or free the handle's plans
// (tsg_treeseq_set_num_chunks, tsg_treeseq_free) take it exclusively: they
// wait until no call is in flight on the handle, and calls that start
// meanwhile wait for them.
std::condition_variable g_calls_cv;
// handles this thread is inside (a call made from within another on the same
// handle is counted once)
thread_local std::vector<const tsg_treeseq *> t_using;

struct Use {
    const tsg_treeseq *ts = nullptr;
    explicit Use(const tsg_treeseq *t) {
        if (t == nullptr || std::find(t_using.begin(), t_using.end(), t) != t_using.end()) {
            return;
        }
        std::unique_lock<std::mutex> lock(g_mutex);
        g_calls_cv.wait(lock, [&] { return !t->exclusive; });
        t->calls++;
        ts = t;
        t_using.push_back(t);
    }
    ~Use() {
        if (ts == nullptr) return;
        t_using.erase(std::find(t_using.begin(), t_using.end(), ts));
        std::lock_guard<std::mutex> lock(g_mutex);
        ts->calls--;
        g_calls_cv.notify_all();
    }
    Use(const Use &) = delete;
    Use &operator=(const Use &) = delete;
};

// Waits, holding g_mutex, until no call is in flight on the handle; new
// calls on it wait until release().
std::unique_lock<std::mutex> exclusive(tsg_treeseq *ts) {
    std::unique_lock<std::mutex> lock(g_mutex);
    g_calls_cv.wait(lock, [&] { return !ts->exclusive; });
    ts->exclusive = true;
    g_calls_cv.wait(lock, [&] { return ts->calls == 0; });
    return lock;
}
void release(tsg_treeseq *ts) {
    ts->exclusive = false;
    g_calls_cv.notify_all();
}

int hip_check(hipError_t e) { return e == hipSuccess ? 0 : TSG_ERR_DEVICE; }

int ensure(DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.bytes >= bytes) return 0;
    if (b.p != nullptr) {
        (void) hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max(bytes, b.bytes + b.bytes / 2);
    if (hipMalloc(&b.p, want) != hipSuccess) {
        if (hipMalloc(&b.p, bytes) != hipSuccess) {
            (void) hipGetLastError();
            return TSG_ERR_DEVICE_MEMORY;
        }
        want = bytes;
    }
    b.bytes = want;
    return 0;
}

// The workspace's buffer b of at least `bytes`: allocated (or grown), or,
// in a no-allocation run, the next 256-byte-aligned piece of the fixed
// region (TSG_ERR_SCRATCH_TOO_SMALL when it is used up).
int ensure(Workspace *ws, DevBuf &b, size_t bytes) {
    if (ws->fixed == nullptr) return ensure(b, bytes);
    if (bytes == 0) bytes = 8;
    const size_t off = (ws->fixed_used + 255) & ~(size_t) 255;
    if (off > ws->fixed_cap || bytes > ws->fixed_cap - off) return TSG_ERR_SCRATCH_TOO_SMALL;
    b.p = ws->fixed + off;
    b.bytes = bytes;
    ws->fixed_used = off + bytes;
    return 0;
}

template <typename T>
int upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    int ret = ensure(b, v.size() * sizeof(T));
    if (ret != 0) return ret;
    if (!v.empty()) {
        ret = hip_check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    return ret;
}

template <typename T>
int upload_ptr(T **dst, const std::vector<T> &v, size_t &bytes) {
    const size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
    if (hipMalloc((void **) dst, n) != hipSuccess) {
        (void) hipGetLastError();
        return TSG_ERR_DEVICE_MEMORY;
    }
    bytes += n;
    if (!v.empty()) return hip_check(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

Resident *res_of(const tsg_treeseq *ts) { return reinterpret_cast<Resident *>(ts->workspace); }

// A workspace leased from the tree sequence's pool for one query or batch
// (created on demand; returned, with its last node store recorded, when the
// lease ends).  device_upload must have run.
struct Lease {
    Resident *res;
    Workspace *ws = nullptr;
    explicit Lease(tsg_treeseq *ts) : res(res_of(ts)) {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (res == nullptr) return;
        if (res->idle.empty()) {
            Workspace *w = new (std::nothrow) Workspace();
            if (w == nullptr) return;
            if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreate(&w->ev0) != hipSuccess || hipEventCreate(&w->ev1) != hipSuccess) {
                (void) hipGetLastError();
                if (w->ev0) (void) hipEventDestroy(w->ev0);
                if (w->stream) (void) hipStreamDestroy(w->stream);
                delete w;
                return;
            }
            w->lds_limit = res->lds_limit;
            w->profile = res->profile;
            res->all.push_back(w);
            res->idle.push_back(w);
        }
        ws = res->idle.back();
        res->idle.pop_back();
        ws->force_full = res->force_full;
        ws->force_c32 = res->force_c32;
        ws->last_store = -1;
    }
    ~Lease() {
        if (ws == nullptr) return;
        // an error path may return with copies or kernels still queued: the
        // next lease must find the stream (and its staging buffers) idle
        (void) hipStreamSynchronize(ws->stream);
        std::lock_guard<std::mutex> lock(g_mutex);
        if (ws->last_store >= 0) res->last_store = ws->last_store;
        res->idle.push_back(ws);
    }
    Lease(const Lease &) = delete;
    Lease &operator=(const Lease &) = delete;
};

// Groups (sample, column) pairs by sample so each weighted sample walks its
// root path once: seed_sample[i] has columns seed_col[seed_off[i]..seed_off[i+1]).
struct Seeds {
    std::vector<int32_t> seed_sample, seed_off, col, flat_s;
    std::vector<double> w;  // the pairs' weights, in flat order (weighted runs)
};

// pairs by sample (columns of one sample in their original order): the CSR
// form (one entry per weighted sample) and the flat pair list
void seeds_host(const std::vector<int32_t> &pair_s, const std::vector<int32_t> &pair_c,
                Seeds &sd, const std::vector<double> *pair_w = nullptr) {
    // the pairs' order by sample, stable: a least-significant-digit radix
    // sort over 11-bit digits of the (non-negative) sample index
    const size_t np = pair_s.size();
    std::vector<int32_t> order(np), tmp(np);
    std::iota(order.begin(), order.end(), 0);
    int32_t smax = 0;
    for (int32_t v : pair_s) smax = std::max(smax, v);
    for (int shift = 0; shift == 0 || (smax >> shift) != 0; shift += 11) {
        size_t count[2049] = {0};
        for (size_t i = 0; i < np; i++) count[((pair_s[order[i]] >> shift) & 2047) + 1]++;
        for (int d = 0; d < 2048; d++) count[d + 1] += count[d];
        for (size_t i = 0; i < np; i++) tmp[count[(pair_s[order[i]] >> shift) & 2047]++] = order[i];
        order.swap(tmp);
    }
    sd.seed_sample.clear();
    sd.seed_off.assign(1, 0);
    sd.col.resize(order.size());
    sd.flat_s.resize(order.size());
    if (pair_w != nullptr) sd.w.resize(order.size());
    for (size_t i = 0; i < order.size(); i++) {
        const int32_t si = pair_s[order[i]];
        sd.col[i] = pair_c[order[i]];
        sd.flat_s[i] = si;
        if (pair_w != nullptr) sd.w[i] = (*pair_w)[order[i]];
        if (sd.seed_sample.empty() || sd.seed_sample.back() != si) {
            if (!sd.seed_sample.empty()) sd.seed_off.push_back((int32_t) i);
            sd.seed_sample.push_back(si);
        }
    }
    sd.seed_off.push_back((int32_t) order.size());
    if (sd.seed_sample.empty()) sd.seed_off.assign(1, 0);
}

// The host arrays of a batch at 256-byte offsets of one pinned buffer, sent
// with one copy into the workspace's arena (a batch of many small arrays
// otherwise costs one transfer each).  The workspace is leased: no earlier
// copy from its staging buffer is still in flight.
struct Pack {
    struct Part {
        const void *p;
        size_t bytes, off;
    };
    std::vector<Part> parts;
    size_t total = 0;
    template <typename T>
    int add(const std::vector<T> &v) {
        const size_t off = (total + 255) & ~(size_t) 255;
        const size_t bytes = v.size() * sizeof(T);
        parts.push_back({v.data(), bytes, off});
        total = off + std::max<size_t>(bytes, 8);
        return (int) parts.size() - 1;
    }
    int send(Workspace *ws, hipStream_t s) const {
        int ret = ensure(ws, ws->arena, total);
        if (ret != 0) return ret;
        if (ws->fixed != nullptr) {
            // no pinned staging: one copy from host memory the workspace
            // keeps until the run has synchronised
            ws->fixed_stage.assign(total, 0);
            for (const Part &pt : parts) {
                if (pt.bytes) std::memcpy(ws->fixed_stage.data() + pt.off, pt.p, pt.bytes);
            }
            return hip_check(hipMemcpyAsync(ws->arena.p, ws->fixed_stage.data(), total,
                                            hipMemcpyHostToDevice, s));
        }
        if (ws->stage_bytes < total) {
            if (ws->stage_h != nullptr) (void) hipHostFree(ws->stage_h);
            ws->stage_h = nullptr;
            ws->stage_bytes = 0;
            const size_t want = std::max(total + total / 2, (size_t) 1 << 16);
            if (hipHostMalloc(&ws->stage_h, want) != hipSuccess) {
                (void) hipGetLastError();
                ws->stage_h = nullptr;
                return TSG_ERR_NO_MEMORY;
            }
            ws->stage_bytes = want;
        }
        for (const Part &pt : parts) {
            if (pt.bytes) std::memcpy((char *) ws->stage_h + pt.off, pt.p, pt.bytes);
        }
        return hip_check(
            hipMemcpyAsync(ws->arena.p, ws->stage_h, total, hipMemcpyHostToDevice, s));
    }
    template <typename T>
    T *dev(const Workspace *ws, int i) const {
        return reinterpret_cast<T *>((char *) ws->arena.p + parts[i].off);
    }
};

int upload_seeds(Workspace *ws, const std::vector<int32_t> &pair_s,
                 const std::vector<int32_t> &pair_c, int32_t num_samples, hipStream_t s,
                 TsgSweepArgs &a) {
    (void) num_samples;
    Seeds sd;
    seeds_host(pair_s, pair_c, sd);
    const std::vector<int32_t> &seed_sample = sd.seed_sample, &seed_off = sd.seed_off;
    const std::vector<int32_t> &col = sd.col, &flat_s = sd.flat_s;
    int ret;
    if ((ret = upload(ws->pair_s, seed_sample, s)) || (ret = upload(ws->pair_c, col, s)) ||
        (ret = upload(ws->seed_off, seed_off, s)) || (ret = upload(ws->seed_ps, flat_s, s))) {
        return ret;
    }
    a.npairs = (int32_t) flat_s.size();
    a.pair_sample = (const int32_t *) ws->seed_ps.p;
    a.pair_col = (const int32_t *) ws->pair_c.p;
    a.nseed = (int32_t) seed_sample.size();
    a.seed_sample = (const int32_t *) ws->pair_s.p;
    a.seed_col = (const int32_t *) ws->pair_c.p;
    a.seed_off = (const int32_t *) ws->seed_off.p;
    return 0;
}

// ---- query description -------------------------------------------------
struct Block {
    int stat;
    int tuple;                  // 0 for one-way (one output per set)
    std::vector<int32_t> idx;   // M x tuple, indexes into the query's sets
    bool polarised;
};

struct Query {
    uint32_t mode = TSG_STAT_SITE;
    bool span_normalise = false;
    std::vector<std::vector<int32_t>> sets;  // sample indexes per column
    int ncentre = 0;                         // columns used by the relatedness mean
    std::vector<Block> blocks;
    std::vector<double> windows;
    int final_kind = TSG_FINAL_NONE;
    int P = 0;                               // final outputs per window
    std::vector<int32_t> fin_ipar;
    std::vector<double> fin_dpar;
    double *result = nullptr;
    // real-valued sample weights (tsg_treeseq_general_stat): K columns,
    // [n x K] by sample index, the outputs over them and their parameters
    bool weighted = false;
    int wK = 0;
    std::vector<double> wts;
    std::vector<TsgOut> wouts;
    std::vector<double> opar;
    int M() const {
        if (weighted) return (int) wouts.size();
        int m = 0;
        for (const Block &b : blocks) m += (b.tuple == 0) ? (int) sets.size() : (int) (b.idx.size() / b.tuple);
        return m;
    }
};

int check_windows(const tsg_treeseq *ts, uint64_t nw, const double *w) {  // trees.c:1244-1286
    if (nw < 1) return TSG_ERR_BAD_NUM_WINDOWS;
    if (w[0] != 0 || w[nw] != ts->L) return TSG_ERR_BAD_WINDOWS;
    for (uint64_t j = 0; j < nw; j++) {
        if (w[j] >= w[j + 1]) return TSG_ERR_BAD_WINDOWS;
    }
    return 0;
}

// trees.c:2114-2149, then the duplicate check of trees.c:2196-2214.
int parse_sample_sets(const tsg_treeseq *ts, uint64_t K, const uint64_t *sizes, const int32_t *sets,
                      std::vector<std::vector<int32_t>> &out) {
    if (K == 0) return TSG_ERR_INSUFFICIENT_SAMPLE_SETS;
    // every sample checked (and mapped to its sample index) before any set
    // is checked for duplicates, as the reference orders its errors
    out.assign(K, {});
    uint64_t j = 0;
    for (uint64_t k = 0; k < K; k++) {
        if (sizes[k] == 0) return TSG_ERR_EMPTY_SAMPLE_SET;
        out[k].resize(sizes[k]);
        for (uint64_t l = 0; l < sizes[k]; l++) {
            const int32_t u = sets[j++];
            if (u < 0 || u >= ts->num_nodes) return TSG_ERR_NODE_OUT_OF_BOUNDS;
            const int32_t si = ts->sample_index_map[u];
            if (si == TSG_NULL) return TSG_ERR_BAD_SAMPLES;
            out[k][l] = si;
        }
    }
    // a sample twice in one set: marks in a per-sample table, cleared after
    // each set (one table per thread, reused: all zeros between calls)
    thread_local std::vector<uint8_t> seen;
    if (seen.size() < (size_t) ts->num_samples) seen.assign((size_t) ts->num_samples, 0);
    for (uint64_t k = 0; k < K; k++) {
        bool dup = false;
        for (int32_t si : out[k]) {
            dup |= seen[si] != 0;
            seen[si] = 1;
        }
        for (int32_t si : out[k]) seen[si] = 0;
        if (dup) return TSG_ERR_DUPLICATE_SAMPLE;
    }
    return 0;
}

// general_stat checks (trees.c:2048-2078) + mode/time checks.
int parse_mode_windows(const tsg_treeseq *ts, uint32_t options, uint64_t nw, const double *windows,
                       Query &q) {
    bool site = options & TSG_STAT_SITE, branch = options & TSG_STAT_BRANCH;
    const bool node = options & TSG_STAT_NODE;
    if (!(site || branch || node)) site = true;
    if ((int) site + (int) branch + (int) node > 1) return TSG_ERR_MULTIPLE_STAT_MODES;
    if (windows == nullptr) {
        q.windows = {0.0, ts->L};
    } else {
        int ret = check_windows(ts, nw, windows);
        if (ret != 0) return ret;
        q.windows.assign(windows, windows + nw + 1);
    }
    q.mode = node ? TSG_STAT_NODE : branch ? TSG_STAT_BRANCH : TSG_STAT_SITE;
    if (branch && ts->time_uncalibrated && !(options & TSG_STAT_ALLOW_TIME_UNCALIBRATED)) {
        return TSG_ERR_TIME_UNCALIBRATED;
    }
    q.span_normalise = (options & TSG_STAT_SPAN_NORMALISE) != 0;
    return 0;
}

int check_tuples(uint64_t K, int tuple, uint64_t M, const int32_t *idx) {  // trees.c:4667-4688
    if (K < 1) return TSG_ERR_INSUFFICIENT_SAMPLE_SETS;
    if (M < 1) return TSG_ERR_INSUFFICIENT_INDEX_TUPLES;
    for (uint64_t j = 0; j < M * (uint64_t) tuple; j++) {
        if (idx[j] < 0 || idx[j] >= (int32_t) K) return TSG_ERR_BAD_SAMPLE_SET_INDEX;
    }
    return 0;
}

// c: state columns of the summary's sets; same01: whether the first two
// sets are the same sample set *index* of the query (divergence's
// denominator, trees.c:4286-4311, depends on the index, not the content)
double summary_inv(int stat, const int32_t *n, const int32_t *c, bool same01) {
    auto N = [&](int t) { return (double) n[c[t]]; };
    switch (stat) {
    case TSG_STAT_DIVERSITY:
        return 1.0 / (N(0) * (N(0) - 1));
    case TSG_STAT_SEGREGATING_SITES:
        return 1.0 / N(0);
    case TSG_STAT_Y1:
        return 1.0 / (N(0) * (N(0) - 1) * (N(0) - 2));
    case TSG_STAT_DIVERGENCE:
        return 1.0 / (N(0) * (N(1) - (same01 ? 1 : 0)));
    case TSG_STAT_RELATEDNESS:
        return 1.0;
    case TSG_STAT_RELATEDNESS_NONCENTRED:
        return 1.0 / (N(0) * N(1));
    case TSG_STAT_Y2:
        return 1.0 / (N(0) * N(1) * (N(1) - 1));
    case TSG_STAT_F2:
        return 1.0 / (N(0) * (N(0) - 1) * N(1) * (N(1) - 1));
    case TSG_STAT_Y3:
        return 1.0 / (N(0) * N(1) * N(2));
    case TSG_STAT_F3:
        return 1.0 / (N(0) * (N(0) - 1) * N(1) * N(2));
    case TSG_STAT_F4:
        return 1.0 / (N(0) * N(1) * N(2) * N(3));
    }
    return 0.0;
}

int device_upload(tsg_treeseq *ts);
int ensure_full(tsg_treeseq *ts);

// A batch between its phases: prepared (inputs uploaded, sweep arguments
// set), swept, finished (combine, finals, results copied out, not yet
// synchronised).  A workspace and a stream carry it.
struct BatchRun {
    tsg_treeseq *ts = nullptr;
    std::vector<Query *> *batch = nullptr;
    Workspace *ws = nullptr;
    hipStream_t s = nullptr;
    uint32_t mode = 0;
    int mode_id = 0, store = 0, M = 0, Q = 0, C = 0, slice = 256;
    size_t shmem = 0;
    TsgSweepArgs a{};
    const int32_t *t0_parent = nullptr;
    const double *t0_bl = nullptr;
    int nloc0 = 0, nb0 = 0;
    int64_t lanecount = 0, outcount = 0;
    TsgCombineArgs ca{};
    std::vector<TsgFinalHost> fin;
    const void *fin_dev = nullptr;
    const int32_t *ipar_dev = nullptr;
    const double *dpar_dev = nullptr;
    // results in host memory after the run has synchronised: the workspace's
    // pinned buffer, or (several runs finished together) this run's part of
    // a shared one
    const double *results_h = nullptr;
    // weighted runs: the first tree's sums of the weights [nloc0 x K] (device)
    const double *x0w_dev = nullptr;
    // per query: its result buffer is device memory (filled by a device
    // copy on the run's stream, not from the pinned host copy)
    std::vector<char> on_dev;
};

// Whether p is device memory (a caller may hand a device buffer for the
// results: it is then filled on the device, no copy to the host).
bool is_device_ptr(const void *p) {
    hipPointerAttribute_t attr;
    if (p == nullptr || hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void) hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

// `bytes` of device memory src into the caller's dst, waiting for stream s
// to finish: a device copy when dst is device memory, one transfer when it
// is pinned host memory or small; otherwise through two pinned pieces in
// turn, the transfer of one piece running while host threads copy the
// previous one into dst (pageable memory takes its transfers slowly and its
// first touch is paid in the copy; spread over threads, both go faster).
constexpr size_t TSG_COPY_PIECE = (size_t) 64 << 20;

int copy_out(Workspace *ws, hipStream_t s, void *dst, const void *src, size_t bytes) {
    int ret;
    if (bytes == 0) return hip_check(hipStreamSynchronize(s));
    hipPointerAttribute_t attr;
    hipMemoryType type = hipMemoryTypeUnregistered;
    if (hipPointerGetAttributes(&attr, dst) == hipSuccess) {
        type = attr.type;
    } else {
        (void) hipGetLastError();
    }
    if (type == hipMemoryTypeDevice || type == hipMemoryTypeHost || bytes <= TSG_COPY_PIECE ||
        ws->fixed != nullptr) {
        const hipMemcpyKind kind =
            type == hipMemoryTypeDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        if ((ret = hip_check(hipMemcpyAsync(dst, src, bytes, kind, s)))) return ret;
        return hip_check(hipStreamSynchronize(s));
    }
    for (int b = 0; b < 2; b++) {
        if (ws->cp_h[b] == nullptr && hipHostMalloc(&ws->cp_h[b], TSG_COPY_PIECE) != hipSuccess) {
            (void) hipGetLastError();
            ws->cp_h[b] = nullptr;
            return TSG_ERR_NO_MEMORY;
        }
        if (ws->cp_ev[b] == nullptr &&
            (ret = hip_check(hipEventCreateWithFlags(&ws->cp_ev[b], hipEventDisableTiming)))) {
            return ret;
        }
    }
    const unsigned hc = std::thread::hardware_concurrency();
    const size_t T = std::max<size_t>(1, std::min<size_t>(8, hc == 0 ? 1 : hc));
    auto host_copy = [&](char *to, const char *from, size_t n) {
        const size_t per = (n + T - 1) / T;
        std::vector<std::thread> pool;
        for (size_t t = 1; t < T && t * per < n; t++) {
            pool.emplace_back([=]() { std::memcpy(to + t * per, from + t * per, std::min(per, n - t * per)); });
        }
        std::memcpy(to, from, std::min(per, n));
        for (auto &th : pool) th.join();
    };
    const size_t np = (bytes + TSG_COPY_PIECE - 1) / TSG_COPY_PIECE;
    for (size_t i = 0; i <= np; i++) {
        if (i < np) {
            const size_t off = i * TSG_COPY_PIECE, n = std::min(TSG_COPY_PIECE, bytes - off);
            const int b = (int) (i & 1);
            if ((ret = hip_check(hipMemcpyAsync(ws->cp_h[b], (const char *) src + off, n,
                                                hipMemcpyDeviceToHost, s))) ||
                (ret = hip_check(hipEventRecord(ws->cp_ev[b], s)))) {
                (void) hipStreamSynchronize(s);
                return ret;
            }
        }
        if (i >= 1) {  // piece i - 1 has arrived in its pinned buffer
            const size_t off = (i - 1) * TSG_COPY_PIECE, n = std::min(TSG_COPY_PIECE, bytes - off);
            const int b = (int) ((i - 1) & 1);
            if ((ret = hip_check(hipEventSynchronize(ws->cp_ev[b])))) {
                (void) hipStreamSynchronize(s);
                return ret;
            }
            host_copy((char *) dst + off, (const char *) ws->cp_h[b], n);
        }
    }
    return hip_check(hipStreamSynchronize(s));
}

// Device copies of the results of the queries whose buffers are device
// memory, from the run's result block `out` (queued on stream s).
int copy_device_results(BatchRun &r, const double *out, hipStream_t s) {
    r.on_dev.assign(r.Q, 0);
    for (int qi = 0; qi < r.Q; qi++) {
        double *dst = (*r.batch)[qi]->result;
        if (!is_device_ptr(dst)) continue;
        r.on_dev[qi] = 1;
        const TsgFinalHost &f = r.fin[qi];
        const size_t n = (size_t) f.W * f.P * 8;
        if (n == 0) continue;
        int ret = hip_check(hipMemcpyAsync(dst, out + f.outbase, n, hipMemcpyDeviceToDevice, s));
        if (ret != 0) return ret;
    }
    return 0;
}

// The fused state columns and outputs of a batch: one column per distinct
// sample set, one TsgOut per raw output, per-query windows and offsets.
struct Fused {
    std::vector<int32_t> pair_s, pair_c, col_total;
    std::vector<TsgOut> outs;
    std::vector<TsgQuery> qs;
    std::vector<double> wins;
    bool noskip = false, any_unpolarised = false;
    int K = 0, M = 0;
    // real-valued weights: per pair its weight, column totals, parameters
    bool weighted = false;
    std::vector<double> pair_w, col_total_w, opar;
};

// A query of real-valued weights runs alone: its K columns as given (pairs
// of the nonzero weights, by sample), totals summed in sample order as the
// reference does (trees.c:1990-1998).
void fuse_weighted(const Query &q, Fused &f) {
    const int K = q.wK;
    const size_t n = K > 0 ? q.wts.size() / (size_t) K : 0;
    f.weighted = true;
    f.K = K;
    f.col_total_w.assign(K, 0.0);
    f.col_total.assign(K, 0);
    for (size_t s = 0; s < n; s++) {
        for (int k = 0; k < K; k++) {
            const double w = q.wts[s * K + k];
            f.col_total_w[k] += w;
            if (w == 0.0) continue;
            f.pair_s.push_back((int32_t) s);
            f.pair_c.push_back(k);
            f.pair_w.push_back(w);
        }
    }
    f.outs = q.wouts;
    f.M = (int) f.outs.size();
    f.opar = q.opar;
    for (TsgOut &o : f.outs) {
        o.q = 0;
        if (!std::isfinite(o.inv)) f.noskip = true;
        if (!(o.flags & 1)) f.any_unpolarised = true;
    }
    f.qs.assign(1, TsgQuery{});
    TsgQuery &tq = f.qs[0];
    tq.W = (int32_t) q.windows.size() - 1;
    tq.M = f.M;
    tq.flags = q.span_normalise ? TSG_F_SPAN_NORMALISE : 0u;
    f.wins = q.windows;
}

void fuse_queries(std::vector<Query *> &batch, Fused &f) {
    const int Q = (int) batch.size();
    if (Q == 1 && batch[0]->weighted) {
        fuse_weighted(*batch[0], f);
        return;
    }
    std::vector<int32_t> &pair_s = f.pair_s, &pair_c = f.pair_c, &col_total = f.col_total;
    std::vector<TsgOut> &outs = f.outs;
    std::vector<TsgQuery> &qs = f.qs;
    qs.assign(Q, TsgQuery{});
    std::vector<double> &wins = f.wins;
    bool &noskip = f.noskip, &any_unpolarised = f.any_unpolarised;
    int &K = f.K, &M = f.M;
    // Identical sample sets (same samples, any order) share one state column
    // across the batch.  A query with a centred relatedness block keeps its
    // own contiguous columns: its summary averages over them.
    std::map<std::vector<int32_t>, int32_t> shared_col;
    // the same set given again in the same order (queries of one batch
    // usually share their sets) is matched without sorting it again
    std::map<std::vector<int32_t>, int32_t> seen_as_given;
    for (int qi = 0; qi < Q; qi++) {
        Query &q = *batch[qi];
        const int cbase = K;
        bool own = false;
        for (const Block &b : q.blocks) own |= (b.stat == TSG_STAT_RELATEDNESS);
        std::vector<int32_t> gcol(q.sets.size());
        for (size_t t = 0; t < q.sets.size(); t++) {
            const auto &s = q.sets[t];
            if (!own) {
                auto given = seen_as_given.find(s);
                if (given != seen_as_given.end()) {
                    gcol[t] = given->second;
                    continue;
                }
                std::vector<int32_t> key(s);
                std::sort(key.begin(), key.end());
                auto it = shared_col.find(key);
                if (it != shared_col.end()) {
                    gcol[t] = it->second;
                    seen_as_given.emplace(s, it->second);
                    continue;
                }
                shared_col.emplace(std::move(key), K);
                seen_as_given.emplace(s, K);
            }
            for (int32_t si : s) {
                pair_s.push_back(si);
                pair_c.push_back(K);
            }
            col_total.push_back((int32_t) s.size());
            gcol[t] = K++;
        }
        TsgQuery &tq = qs[qi];
        tq.W = (int32_t) q.windows.size() - 1;
        tq.mbase = M;
        tq.cbase = cbase;
        tq.flags = q.span_normalise ? TSG_F_SPAN_NORMALISE : 0u;
        tq.win_off = (int64_t) wins.size();
        wins.insert(wins.end(), q.windows.begin(), q.windows.end());
        for (const Block &b : q.blocks) {
            const int nout = (b.tuple == 0) ? (int) q.sets.size() : (int) (b.idx.size() / b.tuple);
            for (int m = 0; m < nout; m++) {
                TsgOut o{};
                o.stat = b.stat;
                o.q = qi;
                int32_t cols[4] = {0, 0, 0, 0};
                if (b.tuple == 0) {
                    cols[0] = m;
                } else {
                    for (int t = 0; t < b.tuple; t++) cols[t] = b.idx[(size_t) m * b.tuple + t];
                }
                int32_t gc[4];
                for (int t = 0; t < 4; t++) gc[t] = gcol[cols[t]];
                o.i = gc[0];
                o.j = gc[1];
                o.k = gc[2];
                o.l = gc[3];
                o.ni = col_total[gc[0]];
                o.nj = col_total[gc[1]];
                o.nk = col_total[gc[2]];
                o.nl = col_total[gc[3]];
                o.cbase = cbase;
                o.ccount = q.ncentre;
                o.flags = b.polarised ? 1 : 0;
                o.inv = summary_inv(b.stat, col_total.data(), gc, cols[0] == cols[1]);
                if (!std::isfinite(o.inv)) noskip = true;
                if (!b.polarised) any_unpolarised = true;
                outs.push_back(o);
                M++;
            }
        }
        tq.M = M - tq.mbase;
    }
    // Columns no output reads (sets of a k-way query that its indexes never
    // name) are dropped: the sweep's cost grows with the column count.
    {
        std::vector<char> used(K, 0);
        for (const TsgOut &o : outs) {
            used[o.i] = 1;
            switch (o.stat) {
            case TSG_STAT_DIVERSITY:
            case TSG_STAT_SEGREGATING_SITES:
            case TSG_STAT_Y1:
                break;
            case TSG_STAT_DIVERGENCE:
            case TSG_STAT_Y2:
            case TSG_STAT_F2:
            case TSG_STAT_RELATEDNESS_NONCENTRED:
                used[o.j] = 1;
                break;
            case TSG_STAT_Y3:
            case TSG_STAT_F3:
                used[o.j] = used[o.k] = 1;
                break;
            case TSG_STAT_RELATEDNESS:
                used[o.j] = 1;
                for (int c = o.cbase; c < o.cbase + o.ccount; c++) used[c] = 1;
                break;
            default:
                used[o.j] = used[o.k] = used[o.l] = 1;
                break;
            }
        }
        std::vector<int32_t> nid(K, -1);
        int K2 = 0;
        for (int c = 0; c < K; c++) {
            if (used[c]) nid[c] = K2++;
        }
        if (K2 < K) {
            size_t w = 0;
            for (size_t i = 0; i < pair_c.size(); i++) {
                if (!used[pair_c[i]]) continue;
                pair_s[w] = pair_s[i];
                pair_c[w] = nid[pair_c[i]];
                w++;
            }
            pair_s.resize(w);
            pair_c.resize(w);
            std::vector<int32_t> tot2(K2);
            for (int c = 0; c < K; c++) {
                if (used[c]) tot2[nid[c]] = col_total[c];
            }
            col_total.swap(tot2);
            // the relatedness mean's columns are all kept, so a query's
            // remaining columns stay contiguous and in order
            for (TsgOut &o : outs) {
                const int32_t i2 = nid[o.i];
                o.j = used[o.j] ? nid[o.j] : i2;
                o.k = used[o.k] ? nid[o.k] : i2;
                o.l = used[o.l] ? nid[o.l] : i2;
                o.i = i2;
                o.cbase = (o.stat == TSG_STAT_RELATEDNESS) ? nid[o.cbase] : 0;
            }
            for (TsgQuery &tq : qs) tq.cbase = 0;  // (not read by the kernels)
            K = K2;
        }
    }
}

// The seeded start-tree counts of a batch's first run, for its second run
// when both sweep the same (sample, column) pairs on the same node store.
struct SeedShare {
    int32_t *buf = nullptr;
    int store = -1, K = 0;
    std::vector<int32_t> pair_s, pair_c;
};

// Host work of a batch and its uploads on stream s: columns, outputs,
// per-(chunk, query) window ranges, node store, sweep arguments, finals.
// share_out: this run writes its seeded counts for a later run (share_in).
int batch_prepare(tsg_treeseq *ts, std::vector<Query *> &batch, Workspace *ws, hipStream_t s,
                  BatchRun &r, SeedShare *share_out = nullptr,
                  const SeedShare *share_in = nullptr) {
    int ret;
    const int Q = (int) batch.size();
    const uint32_t mode = batch[0]->mode;
    Fused fz;
    fuse_queries(batch, fz);
    std::vector<int32_t> &pair_s = fz.pair_s, &pair_c = fz.pair_c, &col_total = fz.col_total;
    std::vector<TsgOut> &outs = fz.outs;
    std::vector<TsgQuery> &qs = fz.qs;
    std::vector<double> &wins = fz.wins;
    const bool noskip = fz.noskip, any_unpolarised = fz.any_unpolarised;
    const int K = fz.K, M = fz.M;
    // Plan and node store: the hot plan with its slots in LDS when the
    // largest chunk's slots fit; otherwise the full plan, in LDS when its
    // largest chunk fits, else in per-chunk HBM slabs.
    const int R = TSG_NODE_WORDS(K);
    // the sweep's LDS staging block for the widest launch (512 outputs)
    // outputs per launch: up to 256 (four per lane; wider launches would
    // spill the per-output polynomial coefficients out of registers)
    const int slice = 256;
    // real-valued weights: FP64 state columns, on the hot plan only
    const bool wgt = fz.weighted;
    // the per-hop staging arrays only for branch sweeps with hop-by-hop
    // summaries (smaller LDS blocks: more resident workgroups)
    const bool hop = mode == TSG_STAT_BRANCH &&
                     tsg_sweep_needs_hop(outs.data(), (int64_t) outs.size(), K, wgt ? 1 : 0,
                                         wgt ? fz.col_total_w.data() : nullptr,
                                         wgt ? fz.opar.data() : nullptr) != 0;
    const size_t stage = (size_t) (wgt ? TSG_STAGE_INTS_W(K, std::min(M, slice), hop)
                                       : TSG_STAGE_INTS(K, std::min(M, slice), hop)) * 4;
    // 16-bit slot counts when no sample set exceeds 65535 samples
    const bool c16 = !wgt && !ws->force_c32 &&
                     *std::max_element(col_total.begin(), col_total.end()) <= 65535;
    const bool bl = mode == TSG_STAT_BRANCH;
    const size_t hot_need =
        (size_t) ts->hot.max_slots * (wgt ? TSG_SLOT_WORDS_W(K, bl) : TSG_SLOT_WORDS(K, c16, bl)) *
            4 + stage;
    const bool use_hot = ts->hot.usable && (wgt || !ws->force_full) &&
                         hot_need <= (size_t) ws->lds_limit;
    if (wgt && !use_hot) return TSG_ERR_UNSUPPORTED_INPUT;
    if (!use_hot) {
        // the full plan is built and uploaded on first need: not in a
        // no-allocation run
        if (ws->fixed != nullptr) return TSG_ERR_UNSUPPORTED_INPUT;
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = ensure_full(ts))) return ret;
    }
    const int C = use_hot ? ts->hot.num_chunks : ts->num_chunks;
    const int32_t *ct0 = use_hot ? ts->hot.chunk_tree0.data() : ts->chunk_tree0.data();
    // per (chunk, query) window ranges; the chunks' spans are read once, and
    // queries with the same windows as an earlier query of the batch take
    // its ranges
    std::vector<int32_t> cq_wfirst((size_t) C * Q), cq_nwin((size_t) C * Q);
    std::vector<int64_t> cq_poff((size_t) C * Q);
    std::vector<double> cspan((size_t) C + 1);
    for (int c = 0; c <= C; c++) cspan[c] = ts->breakpoints[ct0[c]];
    std::vector<int64_t> qslots(Q, 0);
    int64_t pcount = 0;
    for (int qi = 0; qi < Q; qi++) {
        const double *w = wins.data() + qs[qi].win_off;
        const int W = qs[qi].W;
        qs[qi].pbase = pcount;
        int same = -1;
        for (int pj = 0; pj < qi && same < 0; pj++) {
            if (qs[pj].W == W && std::equal(w, w + W + 1, wins.data() + qs[pj].win_off)) same = pj;
        }
        int64_t slots = 0;
        if (same >= 0) {
            for (int c = 0; c < C; c++) {
                const size_t from = (size_t) c * Q + same, to = (size_t) c * Q + qi;
                cq_wfirst[to] = cq_wfirst[from];
                cq_nwin[to] = cq_nwin[from];
                cq_poff[to] = cq_poff[from];
            }
            slots = qslots[same];
        } else {
            // windows touched by each chunk: both ends move forward with c
            int wf = 0, wl = 0;
            for (int c = 0; c < C; c++) {
                const double cs = cspan[c], ce = cspan[c + 1];
                while (wf + 1 < W && w[wf + 1] <= cs) wf++;  // last window starting <= cs
                if (wl < wf) wl = wf;
                while (wl + 1 < W && w[wl + 1] < ce) wl++;   // last window starting < ce
                const size_t at = (size_t) c * Q + qi;
                cq_wfirst[at] = wf;
                cq_nwin[at] = wl - wf + 1;
                cq_poff[at] = slots;
                slots += wl - wf + 1;
            }
        }
        qslots[qi] = slots;
        pcount += slots * qs[qi].M;
    }
    // combine: lanes per (window, output) ~ chunks per window, a power of 2
    // <= 64; each query's lanes start on a wave
    std::vector<int64_t> rawbase(Q), lanebase(Q);
    int64_t rawcount = 0, lanecount = 0;
    for (int qi = 0; qi < Q; qi++) {
        const int64_t items = (int64_t) qs[qi].W * qs[qi].M;
        rawbase[qi] = rawcount;
        rawcount += items;
        const int64_t per = ((int64_t) C + qs[qi].W - 1) / qs[qi].W;
        int g = 1;
        while (g < per && g < 64) g <<= 1;
        qs[qi].lanes = g;
        lanebase[qi] = lanecount;
        lanecount += (items * g + 63) / 64 * 64;
    }
    TsgSweepArgs &a = r.a;
    a = TsgSweepArgs{};
    // the host arrays go over in one copy at the end (pk); their device
    // pointers are set then
    Seeds sd;
    seeds_host(pair_s, pair_c, sd, wgt ? &fz.pair_w : nullptr);
    // weighted branch runs: the first tree's sums of the weights, each
    // sample's weights added up its root path in sample order (the hot
    // plan's first tree numbers samples first), so the sums are the same on
    // every run
    std::vector<double> x0w;
    if (wgt && mode == TSG_STAT_BRANCH) {
        const std::vector<int32_t> &par0 = ts->hot.t0_parent;
        x0w.assign(par0.size() * (size_t) K, 0.0);
        for (size_t i = 0; i < sd.flat_s.size(); i++) {
            const int32_t k = sd.col[i];
            const double w = sd.w[i];
            for (int32_t u = sd.flat_s[i]; u >= 0; u = par0[u]) x0w[(size_t) u * K + k] += w;
        }
    }
    Pack pk;
    const int p_ss = pk.add(sd.seed_sample), p_so = pk.add(sd.seed_off);
    const int p_col = pk.add(sd.col), p_fs = pk.add(sd.flat_s);
    const int p_tot = pk.add(col_total), p_outs = pk.add(outs), p_qs = pk.add(qs);
    const int p_win = pk.add(wins), p_wf = pk.add(cq_wfirst), p_nw = pk.add(cq_nwin);
    const int p_po = pk.add(cq_poff), p_rb = pk.add(rawbase), p_lb = pk.add(lanebase);
    if ((ret = ensure(ws, ws->partial, (size_t) pcount * 8)) ||
        (ret = ensure(ws, ws->raw, (size_t) rawcount * 8))) {
        return ret;
    }
    // the first tree, samples first (its running branch sum is evaluated
    // separately, over all of its nodes)
    const int nloc0 = use_hot ? (int) ts->hot.t0_parent.size()
                              : (int) (ts->chunk_node0[1] - ts->chunk_node0[0]);
    const int nb0 = tsg_tree0_blocks(nloc0);
    if (mode == TSG_STAT_BRANCH) {
        if ((ret = ensure(ws, ws->dR, (size_t) C * M * 8)) ||
            (ret = ensure(ws, ws->Rstart, (size_t) C * M * 8)) ||
            (ret = ensure(ws, ws->x0, (size_t) nloc0 * K * 4)) ||
            (ret = ensure(ws, ws->r0part, (size_t) nb0 * M * 8)) ||
            (ret = ensure(ws, ws->bsum, (size_t) tsg_scan_blocks(C) * M * 8))) {
            return ret;
        }
    }
    int store = wgt ? TSG_STORE_HOTW : c16 ? TSG_STORE_HOT16 : TSG_STORE_HOT;
    size_t shmem = hot_need;
    const TsgDevice &d = ts->dev;
    const TsgHotDevice &h = ts->hot.dev;
    const int32_t *t0_parent = d.loc_parent0;
    const double *t0_bl = d.loc_bl0;
    if (use_hot) {
        a.chunk_tree0 = h.chunk_tree0;
        a.chunk_node0 = h.chunk_slot0;
        a.loc_bl0 = h.slot_bl0;
        a.tr_off = h.tr_off;
        a.tr_nrem = h.tr_nrem;
        a.ev_child = h.ev_child;
        a.ev_bl = h.ev_bl;
        a.ev_path0 = h.ev_path0;
        a.path = h.path;
        a.mut_node = h.mut_slot;
        a.slot_par = h.slot_par;
        a.chunk_lev0 = h.chunk_lev0;
        a.lev_slot0 = h.lev_slot0;
        a.sample_slot = h.sample_slot;
        t0_parent = h.t0_parent;
        t0_bl = h.t0_bl;
    } else {
        const size_t lds_need = (size_t) ts->max_nloc * R * 4 + stage;
        if (lds_need <= (size_t) ws->lds_limit) {
            store = TSG_STORE_LDS;
            shmem = lds_need;
        } else {
            store = TSG_STORE_GLOBAL;
            shmem = stage;
            const size_t slab = (size_t) C * ts->max_nloc * R * 4;
            if ((ret = ensure(ws->slab, slab))) return ret;
        }
        a.ev_path0 = d.ev_path0;
        a.path = d.path;
        a.chunk_tree0 = d.chunk_tree0;
        a.chunk_node0 = d.chunk_node0;
        a.loc_parent0 = d.loc_parent0;
        a.loc_bl0 = d.loc_bl0;
        a.loc_time = d.loc_time;
        a.loc_lastrem = d.loc_lastrem;
        a.tr_off = d.tr_off;
        a.tr_nrem = d.tr_nrem;
        a.ev_child = d.ev_child;
        a.ev_parent = d.ev_parent;
        a.ev_bl = d.ev_bl;
        a.mut_node = d.mut_node;
    }
    ws->last_store = store;
    a.bp = d.breakpoints;
    a.tree_site0 = d.tree_site0;
    a.site_pos = d.site_pos;
    a.site_mut0 = d.site_mut0;
    a.mut_allele = d.mut_allele;
    a.mut_pallele = d.mut_pallele;
    a.site_nalleles = d.site_nalleles;
    a.num_trees = ts->num_trees;
    a.num_chunks = C;
    a.num_samples = ts->num_samples;
    a.K = K;
    a.M = M;
    a.Q = Q;
    a.partial = (double *) ws->partial.p;
    a.dR = (double *) ws->dR.p;
    a.flags = (noskip ? TSG_F_NOSKIP : 0u) | (any_unpolarised ? TSG_F_ANY_UNPOLARISED : 0u);
    a.no_hop = hop ? 0 : 1;
    a.g_rec = (int32_t *) ws->slab.p;
    a.g_stride = (int64_t) ts->max_nloc * R;
    r.ts = ts;
    r.batch = &batch;
    r.ws = ws;
    r.s = s;
    r.mode = mode;
    r.mode_id = (mode == TSG_STAT_BRANCH) ? 0 : 1;
    r.store = store;
    r.shmem = shmem;
    r.M = M;
    r.Q = Q;
    r.C = C;
    r.slice = slice;
    r.t0_parent = t0_parent;
    r.t0_bl = t0_bl;
    r.nloc0 = nloc0;
    r.nb0 = nb0;
    r.lanecount = lanecount;
    TsgCombineArgs &ca = r.ca;
    ca = TsgCombineArgs{};
    ca.bp = d.breakpoints;
    ca.chunk_tree0 = a.chunk_tree0;
    ca.num_chunks = C;
    ca.Q = Q;
    ca.M = M;
    ca.partial = a.partial;
    ca.Rstart = (mode == TSG_STAT_BRANCH) ? (const double *) ws->Rstart.p : nullptr;
    ca.raw = (double *) ws->raw.p;
    ca.flags = 0;
    // finals (FP64 composites), set up now: they read only the combined raw
    // sums
    std::vector<TsgFinalHost> &fin = r.fin;
    fin.assign(Q, TsgFinalHost{});
    std::vector<int32_t> ipar;
    std::vector<double> dpar;
    int64_t outcount = 0;
    for (int qi = 0; qi < Q; qi++) {
        Query &q = *batch[qi];
        TsgFinalHost &f = fin[qi];
        f.kind = q.final_kind;
        f.Kq = (int32_t) q.sets.size();
        f.P = q.P;
        f.W = qs[qi].W;
        f.rawbase = rawbase[qi];
        f.outbase = outcount;
        f.Mraw = qs[qi].M;
        if (q.final_kind == TSG_FINAL_TAJIMAS_D) {
            f.pbase = (int64_t) dpar.size();
            dpar.insert(dpar.end(), q.fin_dpar.begin(), q.fin_dpar.end());
        } else {
            f.pbase = (int64_t) ipar.size();
            ipar.insert(ipar.end(), q.fin_ipar.begin(), q.fin_ipar.end());
        }
        outcount += (int64_t) f.W * f.P;
    }
    r.outcount = outcount;
    const int p_fin = pk.add(fin), p_ip = pk.add(ipar), p_dp = pk.add(dpar);
    int p_pw = -1, p_tw = -1, p_op = -1, p_x0 = -1;
    if (wgt) {
        p_pw = pk.add(sd.w);
        p_tw = pk.add(fz.col_total_w);
        p_op = pk.add(fz.opar);
        if (!x0w.empty()) p_x0 = pk.add(x0w);
    }
    if ((ret = ensure(ws, ws->out, (size_t) outcount * 8)) || (ret = pk.send(ws, s))) return ret;
    a.npairs = (int32_t) sd.flat_s.size();
    a.pair_sample = pk.dev<const int32_t>(ws, p_fs);
    a.pair_col = pk.dev<const int32_t>(ws, p_col);
    a.seed_out = nullptr;
    a.seed_in = nullptr;
    if (use_hot && !wgt && ws->fixed == nullptr) {
        if (share_out != nullptr) {
            const int cw = (store == TSG_STORE_HOT16) ? (K + 1) / 2 : K;
            if (ensure(ws->seedbuf, (size_t) ts->hot.chunk_slot0[C] * cw * 4) == 0) {
                a.seed_out = (int32_t *) ws->seedbuf.p;
                share_out->buf = a.seed_out;
                share_out->store = store;
                share_out->K = K;
                share_out->pair_s = fz.pair_s;
                share_out->pair_c = fz.pair_c;
            }
        } else if (share_in != nullptr && share_in->buf != nullptr && share_in->store == store &&
                   share_in->K == K && share_in->pair_s == fz.pair_s &&
                   share_in->pair_c == fz.pair_c) {
            a.seed_in = share_in->buf;
        }
    }
    a.nseed = (int32_t) sd.seed_sample.size();
    a.seed_sample = pk.dev<const int32_t>(ws, p_ss);
    a.seed_col = pk.dev<const int32_t>(ws, p_col);
    a.seed_off = pk.dev<const int32_t>(ws, p_so);
    a.col_total = pk.dev<const int32_t>(ws, p_tot);
    a.outs = pk.dev<const TsgOut>(ws, p_outs);
    a.queries = pk.dev<const TsgQuery>(ws, p_qs);
    a.windows = pk.dev<const double>(ws, p_win);
    a.cq_wfirst = pk.dev<const int32_t>(ws, p_wf);
    a.cq_nwin = pk.dev<const int32_t>(ws, p_nw);
    a.cq_poff = pk.dev<const int64_t>(ws, p_po);
    ca.queries = a.queries;
    ca.windows = a.windows;
    ca.cq_wfirst = a.cq_wfirst;
    ca.cq_nwin = a.cq_nwin;
    ca.cq_poff = a.cq_poff;
    ca.rawbase = pk.dev<const int64_t>(ws, p_rb);
    ca.lanebase = pk.dev<const int64_t>(ws, p_lb);
    r.fin_dev = pk.dev<const void>(ws, p_fin);
    r.ipar_dev = pk.dev<const int32_t>(ws, p_ip);
    r.dpar_dev = pk.dev<const double>(ws, p_dp);
    if (wgt) {
        a.pair_w = pk.dev<const double>(ws, p_pw);
        a.col_total_w = pk.dev<const double>(ws, p_tw);
        a.opar = pk.dev<const double>(ws, p_op);
        r.x0w_dev = p_x0 >= 0 ? pk.dev<const double>(ws, p_x0) : nullptr;
    }
    return 0;
}

// The sweep launches of a prepared batch: slices of at most r.slice outputs.
int batch_sweep(BatchRun &r) {
    TsgSweepArgs a = r.a;
    for (int m0 = 0; m0 < r.M; m0 += r.slice) {
        a.m_begin = m0;
        a.m_end = std::min(r.M, m0 + r.slice);
        const int span = a.m_end - a.m_begin;
        const int mpl = span <= 64 ? 1 : span <= 128 ? 2 : 4;
        int ret = hip_check(tsg_launch_sweep(a, r.mode_id, r.store, mpl, r.shmem, r.s));
        if (ret != 0) return ret;
    }
    r.a.m_begin = 0;
    r.a.m_end = r.M;
    return 0;
}

// After the sweep: the first tree's running sum (branch), the chunk scan and
// window combine, the finals and the copies of the results to the host (the
// caller synchronises r.s).
int batch_finish(BatchRun &r) {
    int ret;
    Workspace *ws = r.ws;
    hipStream_t s = r.s;
    const TsgSweepArgs &a = r.a;
    if (r.mode == TSG_STAT_BRANCH) {
        if ((ret = hip_check(
                 r.x0w_dev != nullptr
                     ? tsg_launch_tree0_w(a, r.t0_parent, r.t0_bl, r.x0w_dev, r.nloc0,
                                          (double *) ws->r0part.p, s)
                     : tsg_launch_tree0(a, r.t0_parent, r.t0_bl, (int32_t *) ws->x0.p, r.nloc0,
                                        (double *) ws->r0part.p, s)))) {
            return ret;
        }
    }
    if ((ret = hip_check(tsg_launch_combine(r.ca, (r.mode == TSG_STAT_BRANCH) ? a.dR : nullptr,
                                            (double *) ws->Rstart.p, (double *) ws->bsum.p,
                                            (const double *) ws->r0part.p, r.nb0, r.lanecount,
                                            s)))) {
        return ret;
    }
    if ((ret = hip_check(tsg_launch_finals(r.fin_dev, r.Q, (const double *) ws->raw.p,
                                           r.ipar_dev, r.dpar_dev, (double *) ws->out.p,
                                           r.outcount, s)))) {
        return ret;
    }
    if ((ret = copy_device_results(r, (const double *) ws->out.p, s))) return ret;
    if (ws->fixed != nullptr) {
        // no pinned buffer: each query's results straight into its buffer
        for (int qi = 0; qi < r.Q; qi++) {
            if (r.on_dev[qi]) continue;
            const TsgFinalHost &f = r.fin[qi];
            const size_t n = (size_t) f.W * f.P * 8;
            if (n == 0) continue;
            if ((ret = hip_check(hipMemcpyAsync((*r.batch)[qi]->result,
                                                (const double *) ws->out.p + f.outbase, n,
                                                hipMemcpyDeviceToHost, s)))) {
                return ret;
            }
        }
        return 0;
    }
    // all results in one copy into the workspace's pinned buffer; each
    // query's part goes to its result buffer after the stream has finished
    // (batch_deliver)
    const size_t bytes = (size_t) std::max<int64_t>(r.outcount, 1) * 8;
    if (ws->out_h_bytes < bytes) {
        if (ws->out_h != nullptr) (void) hipHostFree(ws->out_h);
        ws->out_h = nullptr;
        ws->out_h_bytes = 0;
        const size_t want = std::max(bytes + bytes / 2, (size_t) 1 << 16);
        if (hipHostMalloc(&ws->out_h, want) != hipSuccess) {
            (void) hipGetLastError();
            ws->out_h = nullptr;
            return TSG_ERR_NO_MEMORY;
        }
        ws->out_h_bytes = want;
    }
    return hip_check(hipMemcpyAsync(ws->out_h, ws->out.p, (size_t) r.outcount * 8,
                                    hipMemcpyDeviceToHost, s));
}

// After r.s has finished: each query's results from the pinned copy.
void batch_deliver(const BatchRun &r) {
    auto dev = [&](int qi) { return qi < (int) r.on_dev.size() && r.on_dev[qi]; };
    if (r.results_h != nullptr) {
        for (int qi = 0; qi < r.Q; qi++) {
            if (dev(qi)) continue;
            const TsgFinalHost &f = r.fin[qi];
            std::memcpy((*r.batch)[qi]->result, r.results_h + f.outbase, (size_t) f.W * f.P * 8);
        }
        return;
    }
    if (r.ws->fixed != nullptr) return;  // copied by batch_finish
    for (int qi = 0; qi < r.Q; qi++) {
        if (dev(qi)) continue;
        const TsgFinalHost &f = r.fin[qi];
        std::memcpy((*r.batch)[qi]->result, (const double *) r.ws->out_h + f.outbase,
                    (size_t) f.W * f.P * 8);
    }
}

// Runs a batch of queries (same tree sequence, same mode) in one sweep.
int run_batch(tsg_treeseq *ts, std::vector<Query *> &batch) {
    int ret;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = device_upload(ts))) return ret;
    }
    Lease lease(ts);
    Workspace *ws = lease.ws;
    if (ws == nullptr) return TSG_ERR_DEVICE;
    BatchRun r;
    if ((ret = batch_prepare(ts, batch, ws, ws->stream, r))) return ret;
    if (ws->profile) (void) hipEventRecord(ws->ev0, r.s);
    if ((ret = batch_sweep(r))) return ret;
    if (ws->profile) (void) hipEventRecord(ws->ev1, r.s);
    if ((ret = batch_finish(r)) || (ret = hip_check(hipStreamSynchronize(r.s)))) return ret;
    batch_deliver(r);
    if (ws->profile) {
        float ms = 0;
        (void) hipEventElapsedTime(&ms, ws->ev0, ws->ev1);
        ws->sweep_ms += ms;
        ws->sweep_launches += 1;
        ws->mode_ms[r.mode_id] += ms;
        ws->mode_launches[r.mode_id] += 1;
    }
    return 0;
}

// While a batch is being filled, the validated query is handed to it
// instead of being run (see tsg_batch_*).
thread_local std::vector<std::unique_ptr<Query>> *g_capture = nullptr;

int run_node(tsg_treeseq *ts, Query &q);

int run_one(tsg_treeseq *ts, Query &q) {
    Use use(ts);
    if (q.mode == TSG_STAT_NODE) {
        // node statistics are not fused: a batch runs them on their own,
        // after its fused sweeps
        if (g_capture != nullptr) {
            g_capture->push_back(std::make_unique<Query>(std::move(q)));
            return 0;
        }
        return run_node(ts, q);
    }
    if (g_capture != nullptr) {
        g_capture->push_back(std::make_unique<Query>(std::move(q)));
        return 0;
    }
    std::vector<Query *> b{&q};
    return run_batch(ts, b);
}

// ---- device residency -----------------------------------------------------
// The device layout of the hot plan's sample map (host: [C x n], chunk
// major): per block of TSG_SS_CB chunks, per sample, the block's chunks --
// ((c / CB) * n + s) * CB + c % CB, the last block padded.  Cache-blocked
// transposition over threads.
std::vector<uint16_t> block_sample_map(const std::vector<uint16_t> &map, int32_t C, int32_t n) {
    const int64_t CB = TSG_SS_CB, nb = (C + CB - 1) / CB;
    std::vector<uint16_t> out((size_t) (nb * CB * n), (uint16_t) TSG_NO_SLOT);
    auto work = [&](int64_t b0, int64_t b1) {
        constexpr int64_t SB = 256;
        for (int64_t b = b0; b < b1; b++) {
            const int64_t c0 = b * CB, c1 = std::min<int64_t>(C, c0 + CB);
            for (int64_t s0 = 0; s0 < n; s0 += SB) {
                const int64_t s1 = std::min<int64_t>(n, s0 + SB);
                for (int64_t c = c0; c < c1; c++) {
                    const uint16_t *src = map.data() + c * n;
                    uint16_t *dst = out.data() + (b * n) * CB + (c - c0);
                    for (int64_t s = s0; s < s1; s++) dst[s * CB] = src[s];
                }
            }
        }
    };
    const int64_t P = std::max<int64_t>(
        1, std::min<int64_t>({(int64_t) std::thread::hardware_concurrency(), 16, nb}));
    std::vector<std::thread> pool;
    for (int64_t p = 1; p < P; p++) pool.emplace_back(work, nb * p / P, nb * (p + 1) / P);
    work(0, nb / P);
    for (auto &th : pool) th.join();
    return out;
}

// Called with g_mutex held.
int device_upload(tsg_treeseq *ts) {
    if (ts->dev.device >= 0) return 0;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        (void) hipGetLastError();
        return TSG_ERR_NO_DEVICE;
    }
    int dev = 0;
    if (hip_check(hipGetDevice(&dev))) return TSG_ERR_NO_DEVICE;
    Resident *res = res_of(ts);
    if (res == nullptr) {
        res = new (std::nothrow) Resident();
        if (res == nullptr) return TSG_ERR_NO_MEMORY;
        ts->workspace = res;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            // gfx950 has 160 KiB of LDS per CU, all of it addressable by one
            // workgroup
            res->lds_limit = (int) std::min<size_t>(prop.sharedMemPerBlock, 160 * 1024);
        }
        tsg_set_lds_max(res->lds_limit);
    }
    int ret;
    TsgDevice &d = ts->dev;
    size_t &by = d.bytes;
    std::vector<int32_t> site_nalleles = ts->site_nalleles;
    if ((ret = upload_ptr(&d.breakpoints, ts->breakpoints, by)) ||
        (ret = upload_ptr(&d.tree_site0, ts->tree_site0, by)) ||
        (ret = upload_ptr(&d.site_pos, ts->site_pos, by)) ||
        (ret = upload_ptr(&d.site_mut0, ts->site_mut0, by)) ||
        (ret = upload_ptr(&d.mut_allele, ts->mut_allele, by)) ||
        (ret = upload_ptr(&d.mut_pallele, ts->mut_pallele, by)) ||
        (ret = upload_ptr(&d.site_nalleles, site_nalleles, by))) {
        return ret;
    }
    if (ts->hot.usable) {
        const tsg_hotplan &hp = ts->hot;
        TsgHotDevice &h = ts->hot.dev;
        size_t &hb = h.bytes;
        if ((ret = upload_ptr(&h.chunk_tree0, hp.chunk_tree0, hb)) ||
            (ret = upload_ptr(&h.chunk_slot0, hp.chunk_slot0, hb)) ||
            (ret = upload_ptr(&h.slot_par, hp.slot_par, hb)) ||
            (ret = upload_ptr(&h.slot_bl0, hp.slot_bl0, hb)) ||
            (ret = upload_ptr(&h.chunk_lev0, hp.chunk_lev0, hb)) ||
            (ret = upload_ptr(&h.lev_slot0, hp.lev_slot0, hb)) ||
            (ret = upload_ptr(&h.sample_slot, block_sample_map(hp.sample_slot, hp.num_chunks,
                                                              ts->num_samples), hb)) ||
            (ret = upload_ptr(&h.tr_off, hp.tr_off, hb)) ||
            (ret = upload_ptr(&h.tr_nrem, hp.tr_nrem, hb)) ||
            (ret = upload_ptr(&h.ev_child, hp.ev_child, hb)) ||
            (ret = upload_ptr(&h.ev_bl, hp.ev_bl, hb)) ||
            (ret = upload_ptr(&h.ev_path0, hp.ev_path0, hb)) ||
            (ret = upload_ptr(&h.path, hp.path, hb)) ||
            (ret = upload_ptr(&h.mut_slot, hp.mut_slot, hb)) ||
            (ret = upload_ptr(&h.t0_parent, hp.t0_parent, hb)) ||
            (ret = upload_ptr(&h.t0_bl, hp.t0_bl, hb))) {
            return ret;
        }
    }
    d.device = dev;
    return 0;
}

// The full plan (every node of each chunk's start tree) is built and
// uploaded on first need: the allele frequency spectrum, the genotype
// route, and sweeps whose hot slots do not fit in LDS.
int ensure_full(tsg_treeseq *ts) {
    int ret = device_upload(ts);
    if (ret != 0) return ret;
    if (!ts->full_built && (ret = tsg_build_plan(ts, ts->full_chunks))) return ret;
    TsgDevice &d = ts->dev;
    if (d.chunk_tree0 != nullptr) return 0;
    size_t &by = d.bytes;
    if ((ret = upload_ptr(&d.chunk_tree0, ts->chunk_tree0, by)) ||
        (ret = upload_ptr(&d.chunk_node0, ts->chunk_node0, by)) ||
        (ret = upload_ptr(&d.loc_parent0, ts->loc_parent0, by)) ||
        (ret = upload_ptr(&d.loc_bl0, ts->loc_bl0, by)) ||
        (ret = upload_ptr(&d.loc_time, ts->loc_time, by)) ||
        (ret = upload_ptr(&d.loc_lastrem, ts->loc_lastrem, by)) ||
        (ret = upload_ptr(&d.tr_off, ts->tr_off, by)) ||
        (ret = upload_ptr(&d.tr_nrem, ts->tr_nrem, by)) ||
        (ret = upload_ptr(&d.ev_child, ts->ev_child, by)) ||
        (ret = upload_ptr(&d.ev_parent, ts->ev_parent, by)) ||
        (ret = upload_ptr(&d.ev_bl, ts->ev_bl, by)) ||
        (ret = upload_ptr(&d.ev_path0, ts->ev_path0, by)) ||
        (ret = upload_ptr(&d.path, ts->path, by)) ||
        (ret = upload_ptr(&d.mut_node, ts->mut_node_loc, by))) {
        return ret;
    }
    return 0;
}

// The hot plan's allele-frequency-spectrum tables, uploaded on first need.
int ensure_hot_afs(tsg_treeseq *ts) {
    TsgHotDevice &h = ts->hot.dev;
    if (h.slot_nsd != nullptr) return 0;
    const tsg_hotplan &hp = ts->hot;
    std::vector<double> slot_tu(hp.slot_node.size());
    for (size_t q = 0; q < slot_tu.size(); q++) slot_tu[q] = ts->time[hp.slot_node[q]];
    size_t &hb = h.bytes;
    int ret;
    if ((ret = upload_ptr(&h.slot_tu, slot_tu, hb)) ||
        (ret = upload_ptr(&h.slot_last0, hp.slot_last0, hb)) ||
        (ret = upload_ptr(&h.slot_ins0, hp.slot_ins0, hb)) ||
        (ret = upload_ptr(&h.slot_tp0, hp.slot_tp0, hb)) ||
        (ret = upload_ptr(&h.tl_parent, hp.tl_parent, hb)) ||
        (ret = upload_ptr(&h.tl_time, hp.tl_time, hb)) ||
        (ret = upload_ptr(&h.tl_last, hp.tl_last, hb)) ||
        (ret = upload_ptr(&h.tl_ins, hp.tl_ins, hb)) ||
        (ret = upload_ptr(&h.slot_nsd, hp.slot_nsd, hb))) {
        return ret;
    }
    return 0;
}

// Node statistics tables of the hot plan (with the AFS tables they share).
int ensure_hot_node(tsg_treeseq *ts) {
    int ret = ensure_hot_afs(ts);
    if (ret != 0) return ret;
    TsgHotDevice &h = ts->hot.dev;
    if (h.node_lid != nullptr) return 0;
    const tsg_hotplan &hp = ts->hot;
    size_t &hb = h.bytes;
    if ((ret = upload_ptr(&h.slot_node, hp.slot_node, hb)) ||
        (ret = upload_ptr(&h.node_last, hp.node_last, hb)) ||
        (ret = upload_ptr(&h.node_lid, hp.node_lid, hb))) {
        return ret;
    }
    return 0;
}

// Node statistics (trees.c:1788-1918, with the unpolarised wrapper of
// trees.c:1944-1972 and the span normalisation of trees.c:2085-2091) on the
// hot plan; the result is W x num_nodes x P, composed per (window, node) row
// as the batch path composes per window.
int run_node(tsg_treeseq *ts, Query &q) {
    int ret;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = device_upload(ts))) return ret;
    }
    const tsg_hotplan &hp = ts->hot;
    if (!hp.usable) return TSG_ERR_UNSUPPORTED_INPUT;
    std::vector<Query *> one{&q};
    Fused fz;
    fuse_queries(one, fz);
    const int K = fz.K, M = fz.M;
    const int W = (int) q.windows.size() - 1;
    const int64_t N = ts->num_nodes;
    if ((double) W * (double) N >= (double) INT32_MAX) return TSG_ERR_UNSUPPORTED_INPUT;
    // per slot: last flush position and K counts
    const size_t shmem = (size_t) hp.max_slots * (8 + 4 * (size_t) K) + 64;
    Lease lease(ts);
    Workspace *ws = lease.ws;
    if (ws == nullptr) return TSG_ERR_DEVICE;
    if (shmem > (size_t) ws->lds_limit) return TSG_ERR_UNSUPPORTED_INPUT;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = ensure_hot_node(ts))) return ret;
    }
    hipStream_t s = ws->stream;
    const TsgHotDevice &h = hp.dev;
    const TsgDevice &d = ts->dev;
    const int64_t nres = (int64_t) W * N * M;
    const int64_t nout = (int64_t) W * N * q.P;
    const int nl = (int) hp.tl_parent.size();
    TsgSweepArgs a{};
    if ((ret = upload_seeds(ws, fz.pair_s, fz.pair_c, ts->num_samples, s, a)) ||
        (ret = upload(ws->tot, fz.col_total, s)) || (ret = upload(ws->outs, fz.outs, s)) ||
        (ret = upload(ws->windows, q.windows, s)) || (ret = ensure(ws->raw, (size_t) nres * 8)) ||
        (ret = ensure(ws->out, (size_t) nout * 8)) ||
        (ret = ensure(ws->x0, ((size_t) nl + 1) * K * 4))) {
        return ret;
    }
    a.chunk_tree0 = h.chunk_tree0;
    a.chunk_node0 = h.chunk_slot0;
    a.tr_off = h.tr_off;
    a.tr_nrem = h.tr_nrem;
    a.ev_child = h.ev_child;
    a.ev_path0 = h.ev_path0;
    a.path = h.path;
    a.slot_par = h.slot_par;
    a.chunk_lev0 = h.chunk_lev0;
    a.lev_slot0 = h.lev_slot0;
    a.sample_slot = h.sample_slot;
    a.slot_last0 = h.slot_last0;
    a.bp = d.breakpoints;
    a.num_trees = ts->num_trees;
    a.num_chunks = hp.num_chunks;
    a.num_samples = ts->num_samples;
    a.K = K;
    a.M = M;
    a.Q = 1;
    a.col_total = (const int32_t *) ws->tot.p;
    a.outs = (const TsgOut *) ws->outs.p;
    a.windows = (const double *) ws->windows.p;
    a.afs_W = W;
    TsgLastTree lt{};
    lt.parent = h.tl_parent;
    lt.nl = nl;
    lt.counts = (int32_t *) ws->x0.p;
    double *raw = (double *) ws->raw.p;
    if ((ret = hip_check(tsg_launch_fill_f64(raw, nres, s)))) return ret;
    if (ws->profile) (void) hipEventRecord(ws->ev0, s);
    if ((ret = hip_check(tsg_launch_node(a, h.slot_node, raw, N, shmem, lt, h.node_last,
                                         h.node_lid, ts->L, s)))) {
        return ret;
    }
    if (ws->profile) (void) hipEventRecord(ws->ev1, s);
    ws->last_store = TSG_STORE_HOT;
    const uint32_t fflags = q.span_normalise ? TSG_F_SPAN_NORMALISE : 0u;
    if ((ret = hip_check(tsg_launch_afs_branch_final(raw, a.windows, W, N * M, fflags, s)))) {
        return ret;
    }
    std::vector<TsgFinalHost> fin(1);
    fin[0].kind = q.final_kind;
    fin[0].Kq = (int32_t) q.sets.size();
    fin[0].P = q.P;
    fin[0].W = (int32_t) (W * N);
    fin[0].rawbase = 0;
    fin[0].outbase = 0;
    fin[0].pbase = 0;
    fin[0].Mraw = M;
    if ((ret = upload(ws->fin, fin, s)) || (ret = upload(ws->ipar, q.fin_ipar, s)) ||
        (ret = upload(ws->dpar, q.fin_dpar, s))) {
        return ret;
    }
    if ((ret = hip_check(tsg_launch_finals(ws->fin.p, 1, raw, (const int32_t *) ws->ipar.p,
                                           (const double *) ws->dpar.p, (double *) ws->out.p,
                                           nout, s))) ||
        (ret = copy_out(ws, s, q.result, ws->out.p, (size_t) nout * 8))) {
        return ret;
    }
    if (ws->profile) {
        float ms = 0;
        (void) hipEventElapsedTime(&ms, ws->ev0, ws->ev1);
        ws->sweep_ms += ms;
        ws->sweep_launches += 1;
    }
    return 0;
}

void free_one_workspace(Workspace *ws) {
    DevBuf *bufs[] = {&ws->seed_off, &ws->seed_ps, &ws->seedbuf, &ws->pair_s, &ws->pair_c, &ws->tot, &ws->outs, &ws->queries, &ws->windows,
                      &ws->cq_wfirst, &ws->cq_nwin, &ws->cq_poff, &ws->partial, &ws->dR,
                      &ws->Rstart, &ws->raw, &ws->rawbase, &ws->lanebase, &ws->out, &ws->fin, &ws->ipar,
                      &ws->dpar, &ws->slab, &ws->last, &ws->hist, &ws->dims, &ws->twin,
                      &ws->gbits, &ws->sitecol, &ws->gt, &ws->cmat, &ws->rowsum,
                      &ws->scale,
                      &ws->kout, &ws->wk0, &ws->wk1, &ws->sel, &ws->gslab, &ws->geno, &ws->x0,
                      &ws->r0part, &ws->bsum, &ws->margs, &ws->mcpre, &ws->mfin, &ws->mpre,
                      &ws->mout, &ws->arena,
                      &ws->rv_E, &ws->rv_I, &ws->rv_z0, &ws->rv_d0, &ws->rv_esum, &ws->rv_y,
                      &ws->sitecol2, &ws->gt2, &ws->cmat2, &ws->wk0b, &ws->wk1b, &ws->row2};
    for (DevBuf *b : bufs) {
        if (b->p != nullptr) (void) hipFree(b->p);
    }
    if (ws->stage_h != nullptr) (void) hipHostFree(ws->stage_h);
    if (ws->out_h != nullptr) (void) hipHostFree(ws->out_h);
    if (ws->mout_h != nullptr) (void) hipHostFree(ws->mout_h);
    for (int b = 0; b < 2; b++) {
        if (ws->cp_h[b] != nullptr) (void) hipHostFree(ws->cp_h[b]);
        if (ws->cp_ev[b]) (void) hipEventDestroy(ws->cp_ev[b]);
    }
    if (ws->ev0) (void) hipEventDestroy(ws->ev0);
    if (ws->ev1) (void) hipEventDestroy(ws->ev1);
    if (ws->stream) (void) hipStreamDestroy(ws->stream);
    delete ws;
}

// Frees every workspace of the tree sequence; no query may be running.
void free_workspace(tsg_treeseq *ts) {
    Resident *res = res_of(ts);
    if (res == nullptr) return;
    for (Workspace *ws : res->all) free_one_workspace(ws);
    delete res;
    ts->workspace = nullptr;
}

}  // namespace

void tsg_device_free(TsgDevice *d) {
    void *ptrs[] = {d->chunk_tree0, d->chunk_node0, d->loc_parent0, d->loc_bl0, d->loc_time,
                    d->loc_lastrem, d->tr_off, d->tr_nrem, d->ev_child, d->ev_parent, d->ev_bl,
                    d->ev_path0, d->path,
                    d->breakpoints, d->tree_site0, d->site_pos, d->site_mut0, d->mut_node,
                    d->mut_allele, d->mut_pallele, d->site_nalleles};
    for (void *p : ptrs) {
        if (p != nullptr) (void) hipFree(p);
    }
    *d = TsgDevice();
}

void tsg_hot_device_free(TsgHotDevice *h) {
    void *ptrs[] = {h->chunk_tree0, h->chunk_slot0, h->slot_par, h->slot_bl0, h->chunk_lev0,
                    h->lev_slot0, h->sample_slot, h->tr_off, h->tr_nrem, h->ev_child, h->ev_bl,
                    h->ev_path0, h->path, h->mut_slot, h->t0_parent, h->t0_bl,
                    h->slot_nsd, h->slot_tu, h->slot_last0, h->slot_ins0, h->slot_tp0,
                    h->tl_parent, h->tl_time, h->tl_last, h->tl_ins, h->slot_node,
                    h->node_last, h->node_lid};
    for (void *p : ptrs) {
        if (p != nullptr) (void) hipFree(p);
    }
    *h = TsgHotDevice();
}

namespace {
int default_num_chunks(const tsg_treeseq *ts) {
    // Enough chunks to fill the device (about 4 waves per CU), each with a
    // few dozen edge events; the plan's node snapshots cost ~24 bytes per
    // local node, so cap the total near 4 GiB.
    const int64_t events = 2 * (int64_t) ts->num_edges;
    int64_t c = std::max<int64_t>(1, events / 48);
    c = std::min<int64_t>(c, 4096);
    const int64_t per_chunk = 2 * (int64_t) ts->num_samples + 64;
    const int64_t budget = (int64_t) 4 << 30;
    c = std::min<int64_t>(c, std::max<int64_t>(1, budget / (per_chunk * 24)));
    c = std::min<int64_t>(c, std::max(1, ts->num_trees));
    return (int) c;
}

int default_hot_chunks(const tsg_treeseq *ts) {
    // About 256 replayed edge events per chunk: the sweep is a serial chain
    // per chunk, so more, shorter chunks run faster as long as their hot
    // slots stay small (a few hundred, ~25 KiB of LDS at 8 sample sets, five
    // or six chunks per CU) -- but at least 1024 chunks of >= 64 events to
    // fill 256 CUs.  Each chunk stores one 16-bit slot per sample: the total
    // is capped near 4 GiB.
    const int64_t events = 2 * (int64_t) ts->num_edges;
    int64_t c = std::max<int64_t>(events / 256, std::min<int64_t>(1024, events / 64));
    c = std::max<int64_t>(c, 1);
    const int64_t per_chunk = 2 * (int64_t) std::max(1, ts->num_samples);
    c = std::min<int64_t>(c, std::max<int64_t>(1, ((int64_t) 4 << 30) / per_chunk));
    c = std::min<int64_t>(c, std::max(1, ts->num_trees));
    return (int) c;
}

// The full plan's host arrays, built on first use.
const tsg_treeseq *full_plan(const tsg_treeseq *ts) {
    if (!ts->full_built) {
        tsg_treeseq *m = const_cast<tsg_treeseq *>(ts);
        (void) tsg_build_plan(m, m->full_chunks);
    }
    return ts;
}
}  // namespace

extern "C" {

const char *tsg_strerror(int err) {
    switch (err) {
    case 0: return "Normal exit condition. This is not an error!";
    case TSG_ERR_NO_MEMORY: return "Out of memory";
    case TSG_ERR_BAD_PARAM_VALUE: return "Bad parameter value provided";
    case TSG_ERR_NODE_OUT_OF_BOUNDS: return "Node out of bounds";
    case TSG_ERR_EDGE_OUT_OF_BOUNDS: return "Edge out of bounds";
    case TSG_ERR_DUPLICATE_SAMPLE: return "Duplicate sample value";
    case TSG_ERR_BAD_SAMPLES: return "The nodes provided are not samples";
    case TSG_ERR_BAD_NUM_WINDOWS: return "Must have at least one window, [0, L]";
    case TSG_ERR_BAD_WINDOWS: return "Windows must be increasing list [0, ..., L]";
    case TSG_ERR_MULTIPLE_STAT_MODES: return "Cannot specify more than one stats mode";
    case TSG_ERR_BAD_STATE_DIMS: return "Must have state dimension >= 1";
    case TSG_ERR_BAD_RESULT_DIMS: return "Must have result dimension >= 1";
    case TSG_ERR_INSUFFICIENT_SAMPLE_SETS: return "Insufficient sample sets provided";
    case TSG_ERR_INSUFFICIENT_INDEX_TUPLES: return "Insufficient sample set index tuples provided";
    case TSG_ERR_BAD_SAMPLE_SET_INDEX: return "Sample set index out of bounds";
    case TSG_ERR_EMPTY_SAMPLE_SET: return "Samples cannot be empty";
    case TSG_ERR_UNSUPPORTED_STAT_MODE: return "Requested statistics mode not supported for this method";
    case TSG_ERR_TIME_UNCALIBRATED: return "Statistics using branch lengths cannot be calculated when time_units is 'uncalibrated'";
    case TSG_ERR_STAT_POLARISED_UNSUPPORTED: return "The polarised option is not supported by this statistic";
    case TSG_ERR_STAT_SPAN_NORMALISE_UNSUPPORTED: return "The span_normalise option is not supported by this statistic";
    case TSG_ERR_INSUFFICIENT_WEIGHTS: return "Insufficient weights provided (at least 1 required)";
    case TSG_ERR_BAD_TIME_WINDOWS_DIM: return "Must have at least one time window";
    case TSG_ERR_BAD_TIME_WINDOWS: return "Time windows must be strictly increasing and begin at 0";
    case TSG_ERR_DEVICE: return "A HIP runtime call failed";
    case TSG_ERR_NO_DEVICE: return "No usable GPU device";
    case TSG_ERR_DEVICE_MEMORY: return "Device memory allocation failed";
    case TSG_ERR_UNSUPPORTED_INPUT: return "Input shape not supported by the device path";
    case TSG_ERR_SCRATCH_TOO_SMALL:
        return "The batch needs more device scratch than the no-allocation entry point holds";
    case TSG_ERR_NOT_RESIDENT:
        return "The tree sequence is not resident on the current device (call tsg_treeseq_upload)";
    default: return "Error in tree sequence input or statistics arguments";
    }
}

int tsg_treeseq_init(tsg_treeseq **out, double sequence_length, uint64_t num_nodes,
                     const uint32_t *node_flags, const double *node_time, uint64_t num_edges,
                     const double *edge_left, const double *edge_right, const int32_t *edge_parent,
                     const int32_t *edge_child, uint64_t num_sites, const double *site_position,
                     const char *ancestral_state, const uint64_t *ancestral_state_offset,
                     uint64_t num_mutations, const int32_t *mutation_site,
                     const int32_t *mutation_node, const int32_t *mutation_parent,
                     const char *derived_state, const uint64_t *derived_state_offset,
                     const int32_t *edge_insertion_order, const int32_t *edge_removal_order,
                     uint32_t options) {
    *out = nullptr;
    if (!(sequence_length > 0) || !std::isfinite(sequence_length)) return TSG_ERR_BAD_SEQUENCE_LENGTH;
    if (num_nodes > INT32_MAX || num_edges > INT32_MAX || num_sites > INT32_MAX ||
        num_mutations > INT32_MAX) {
        return TSG_ERR_BAD_PARAM_VALUE;
    }
    std::unique_ptr<tsg_treeseq> ts(new (std::nothrow) tsg_treeseq());
    if (!ts) return TSG_ERR_NO_MEMORY;
    ts->L = sequence_length;
    ts->options = options;
    ts->time_uncalibrated = (options & TSG_TS_TIME_UNCALIBRATED) != 0;
    ts->num_nodes = (int32_t) num_nodes;
    ts->num_edges = (int32_t) num_edges;
    ts->num_sites = (int32_t) num_sites;
    ts->num_mutations = (int32_t) num_mutations;
    ts->flags.assign(node_flags, node_flags + num_nodes);
    ts->time.assign(node_time, node_time + num_nodes);
    ts->left.assign(edge_left, edge_left + num_edges);
    ts->right.assign(edge_right, edge_right + num_edges);
    ts->parent.assign(edge_parent, edge_parent + num_edges);
    ts->child.assign(edge_child, edge_child + num_edges);
    ts->site_pos.assign(site_position, site_position + num_sites);
    ts->mut_site.assign(mutation_site, mutation_site + num_mutations);
    ts->mut_node.assign(mutation_node, mutation_node + num_mutations);
    ts->mut_parent.assign(mutation_parent, mutation_parent + num_mutations);
    ts->I.resize(num_edges);
    ts->O.resize(num_edges);
    // validate node references before the index build reads time[parent]
    for (uint64_t j = 0; j < num_edges; j++) {
        if (edge_parent[j] < 0 || edge_parent[j] >= (int32_t) num_nodes || edge_child[j] < 0 ||
            edge_child[j] >= (int32_t) num_nodes) {
            return (edge_parent[j] == TSG_NULL) ? TSG_ERR_NULL_PARENT
                   : (edge_child[j] == TSG_NULL) ? TSG_ERR_NULL_CHILD
                                                 : TSG_ERR_NODE_OUT_OF_BOUNDS;
        }
    }
    if (edge_insertion_order != nullptr && edge_removal_order != nullptr) {
        std::copy(edge_insertion_order, edge_insertion_order + num_edges, ts->I.begin());
        std::copy(edge_removal_order, edge_removal_order + num_edges, ts->O.begin());
    } else {
        tsg_build_index(edge_left, edge_right, edge_parent, edge_child, node_time,
                        (int32_t) num_edges, ts->I.data(), ts->O.data());
    }
    int ret = tsg_init_trees(ts.get(), ancestral_state, ancestral_state_offset, derived_state,
                             derived_state_offset);
    if (ret != 0) return ret;
    ts->full_chunks = default_num_chunks(ts.get());
    ret = tsg_build_hot_plan(ts.get(), default_hot_chunks(ts.get()), true);
    if (ret != 0) return ret;
    *out = ts.release();
    return 0;
}

int tsg_treeseq_free(tsg_treeseq *ts) {
    if (ts == nullptr) return 0;
    {
        std::unique_lock<std::mutex> lock = exclusive(ts);
        free_workspace(ts);
        tsg_device_free(&ts->dev);
        tsg_hot_device_free(&ts->hot.dev);
    }
    delete ts;
    return 0;
}

// Rebuilds the chunk plans with a given chunk count (autotuner hook); the
// full plan is rebuilt on its next use.
int tsg_treeseq_set_num_chunks(tsg_treeseq *ts, uint32_t num_chunks) {
    if (ts == nullptr || num_chunks == 0) return TSG_ERR_BAD_PARAM_VALUE;
    std::unique_lock<std::mutex> lock = exclusive(ts);
    free_workspace(ts);
    tsg_device_free(&ts->dev);
    tsg_hot_device_free(&ts->hot.dev);
    ts->full_chunks = (int32_t) std::min<uint32_t>(num_chunks, INT32_MAX);
    ts->full_built = false;
    const int ret = tsg_build_hot_plan(ts, ts->full_chunks, false);
    release(ts);
    return ret;
}

// Testing hook: plan = 1 runs every sweep on the full plan, 2 on the hot
// plan with 32-bit counts, 0 lets the library choose (the hot plan whenever
// its slots fit in LDS, with 16-bit counts when every set has <= 65535
// samples).
int tsg_treeseq_set_sweep_plan(tsg_treeseq *ts, int plan) {
    Use use(ts);
    std::lock_guard<std::mutex> lock(g_mutex);
    int ret = device_upload(ts);
    if (ret != 0) return ret;
    res_of(ts)->force_full = plan == 1;
    res_of(ts)->force_c32 = plan == 2;
    return 0;
}

// Node store of the last statistics sweep: 0 HBM slabs (full plan), 1 LDS
// (full plan), 2 LDS hot slots, 3 LDS hot slots with 16-bit counts; -1
// before the first sweep.
int tsg_treeseq_get_last_store(const tsg_treeseq *ts) {
    std::lock_guard<std::mutex> lock(g_mutex);
    const Resident *res = res_of(ts);
    return res == nullptr ? -1 : res->last_store;
}

int tsg_treeseq_upload(tsg_treeseq *ts) {
    Use use(ts);
    std::lock_guard<std::mutex> lock(g_mutex);
    return device_upload(ts);
}

uint64_t tsg_treeseq_get_num_trees(const tsg_treeseq *ts) { return (uint64_t) ts->num_trees; }
uint64_t tsg_treeseq_get_num_samples(const tsg_treeseq *ts) { return (uint64_t) ts->num_samples; }
uint64_t tsg_treeseq_get_num_nodes(const tsg_treeseq *ts) { return (uint64_t) ts->num_nodes; }
uint64_t tsg_treeseq_get_num_edges(const tsg_treeseq *ts) { return (uint64_t) ts->num_edges; }
uint64_t tsg_treeseq_get_num_sites(const tsg_treeseq *ts) { return (uint64_t) ts->num_sites; }
uint64_t tsg_treeseq_get_num_mutations(const tsg_treeseq *ts) { return (uint64_t) ts->num_mutations; }
uint64_t tsg_treeseq_get_num_chunks(const tsg_treeseq *ts) {
    return (uint64_t) full_plan(ts)->num_chunks;
}
double tsg_treeseq_get_sequence_length(const tsg_treeseq *ts) { return ts->L; }
uint64_t tsg_treeseq_get_device_bytes(const tsg_treeseq *ts) {
    return (uint64_t) (ts->dev.bytes + ts->hot.dev.bytes);
}
uint64_t tsg_treeseq_get_max_chunk_nodes(const tsg_treeseq *ts) {
    return (uint64_t) full_plan(ts)->max_nloc;
}

int tsg_treeseq_get_breakpoints(const tsg_treeseq *ts, double *out) {
    std::copy(ts->breakpoints.begin(), ts->breakpoints.end(), out);
    return 0;
}
int tsg_treeseq_get_samples(const tsg_treeseq *ts, int32_t *out) {
    std::copy(ts->samples.begin(), ts->samples.end(), out);
    return 0;
}
int tsg_treeseq_get_index(const tsg_treeseq *ts, int32_t *I, int32_t *O) {
    std::copy(ts->I.begin(), ts->I.end(), I);
    std::copy(ts->O.begin(), ts->O.end(), O);
    return 0;
}

// Introspection of the chunk plan (host copies), used to check the plan on
// machines without a GPU.  sizes = {chunks, local nodes, events, trees,
// sites, mutations}.
int tsg_treeseq_get_plan_sizes(const tsg_treeseq *ts, uint64_t *sizes) {
    Use use(ts);
    ts = full_plan(ts);
    sizes[0] = (uint64_t) ts->num_chunks;
    sizes[1] = ts->loc_parent0.size();
    sizes[2] = ts->ev_child.size();
    sizes[3] = (uint64_t) ts->num_trees;
    sizes[4] = (uint64_t) ts->num_sites;
    sizes[5] = (uint64_t) ts->num_mutations;
    return 0;
}

int tsg_treeseq_copy_plan(const tsg_treeseq *ts, int32_t *chunk_tree0, int64_t *chunk_node0,
                          int32_t *loc_parent0, double *loc_bl0, double *loc_time,
                          double *loc_lastrem, int64_t *tr_off, int32_t *tr_nrem,
                          int32_t *ev_child, int32_t *ev_parent, double *ev_bl,
                          int32_t *tree_site0, int32_t *site_mut0, int32_t *mut_node_loc,
                          int32_t *mut_allele, int32_t *mut_pallele, int32_t *site_nalleles) {
    Use use(ts);
    ts = full_plan(ts);
    auto cp = [](auto *dst, const auto &v) { std::copy(v.begin(), v.end(), dst); };
    cp(chunk_tree0, ts->chunk_tree0);
    cp(chunk_node0, ts->chunk_node0);
    cp(loc_parent0, ts->loc_parent0);
    cp(loc_bl0, ts->loc_bl0);
    cp(loc_time, ts->loc_time);
    cp(loc_lastrem, ts->loc_lastrem);
    cp(tr_off, ts->tr_off);
    cp(tr_nrem, ts->tr_nrem);
    cp(ev_child, ts->ev_child);
    cp(ev_parent, ts->ev_parent);
    cp(ev_bl, ts->ev_bl);
    cp(tree_site0, ts->tree_site0);
    cp(site_mut0, ts->site_mut0);
    cp(mut_node_loc, ts->mut_node_loc);
    cp(mut_allele, ts->mut_allele);
    cp(mut_pallele, ts->mut_pallele);
    cp(site_nalleles, ts->site_nalleles);
    return 0;
}

// Root paths of the events (host copies): ev_path0 has events + 1 entries.
uint64_t tsg_treeseq_get_path_size(const tsg_treeseq *ts) { return full_plan(ts)->path.size(); }

int tsg_treeseq_copy_paths(const tsg_treeseq *ts, int64_t *ev_path0, int32_t *path) {
    Use use(ts);
    ts = full_plan(ts);
    std::copy(ts->ev_path0.begin(), ts->ev_path0.end(), ev_path0);
    std::copy(ts->path.begin(), ts->path.end(), path);
    return 0;
}

// Introspection of the hot plan (host copies).  sizes = {usable, chunks,
// slots, level entries, events, path entries, first-tree nodes, max slots
// per chunk, max levels per chunk}.
int tsg_treeseq_get_hot_plan_sizes(const tsg_treeseq *ts, uint64_t *sizes) {
    Use use(ts);
    const tsg_hotplan &hp = ts->hot;
    sizes[0] = hp.usable ? 1 : 0;
    sizes[1] = (uint64_t) hp.num_chunks;
    sizes[2] = hp.slot_par.size();
    sizes[3] = hp.lev_slot0.size();
    sizes[4] = hp.ev_child.size();
    sizes[5] = hp.path.size();
    sizes[6] = hp.t0_parent.size();
    sizes[7] = (uint64_t) hp.max_slots;
    sizes[8] = (uint64_t) hp.max_levels;
    sizes[9] = hp.tl_parent.size();
    return 0;
}

int tsg_treeseq_copy_hot_plan(const tsg_treeseq *ts, int32_t *chunk_tree0, int64_t *chunk_slot0,
                              int32_t *slot_par, double *slot_bl0, int64_t *chunk_lev0,
                              int32_t *lev_slot0, uint16_t *sample_slot, int64_t *tr_off,
                              int32_t *tr_nrem, int32_t *ev_child, double *ev_bl,
                              int64_t *ev_path0, int32_t *path, int32_t *mut_slot,
                              int32_t *t0_parent, double *t0_bl) {
    Use use(ts);
    const tsg_hotplan &hp = ts->hot;
    if (!hp.usable) return TSG_ERR_UNSUPPORTED_INPUT;
    auto cp = [](auto *dst, const auto &v) { std::copy(v.begin(), v.end(), dst); };
    cp(chunk_tree0, hp.chunk_tree0);
    cp(chunk_slot0, hp.chunk_slot0);
    cp(slot_par, hp.slot_par);
    cp(slot_bl0, hp.slot_bl0);
    cp(chunk_lev0, hp.chunk_lev0);
    cp(lev_slot0, hp.lev_slot0);
    cp(sample_slot, hp.sample_slot);
    cp(tr_off, hp.tr_off);
    cp(tr_nrem, hp.tr_nrem);
    cp(ev_child, hp.ev_child);
    cp(ev_bl, hp.ev_bl);
    cp(ev_path0, hp.ev_path0);
    cp(path, hp.path);
    cp(mut_slot, hp.mut_slot);
    cp(t0_parent, hp.t0_parent);
    cp(t0_bl, hp.t0_bl);
    return 0;
}

int tsg_treeseq_copy_hot_afs_plan(const tsg_treeseq *ts, int32_t *slot_node, int32_t *slot_nsd,
                                  double *slot_last0, double *slot_ins0, double *slot_tp0,
                                  int32_t *tl_parent, double *tl_time, double *tl_last,
                                  double *tl_ins) {
    Use use(ts);
    const tsg_hotplan &hp = ts->hot;
    if (!hp.usable) return TSG_ERR_UNSUPPORTED_INPUT;
    auto cp = [](auto *dst, const auto &v) { std::copy(v.begin(), v.end(), dst); };
    cp(slot_node, hp.slot_node);
    cp(slot_nsd, hp.slot_nsd);
    cp(slot_last0, hp.slot_last0);
    cp(slot_ins0, hp.slot_ins0);
    cp(slot_tp0, hp.slot_tp0);
    cp(tl_parent, hp.tl_parent);
    cp(tl_time, hp.tl_time);
    cp(tl_last, hp.tl_last);
    cp(tl_ins, hp.tl_ins);
    return 0;
}

// Profiling of the sweep kernel: HIP events on the library's stream around
// every sweep launch.
int tsg_treeseq_set_profiling(tsg_treeseq *ts, int enable) {
    Use use(ts);
    int ret = tsg_treeseq_upload(ts);
    if (ret != 0) return ret;
    std::lock_guard<std::mutex> lock(g_mutex);
    Resident *res = res_of(ts);
    res->profile = enable != 0;
    for (Workspace *ws : res->all) {
        ws->profile = enable != 0;
        ws->sweep_ms = 0;
        ws->sweep_launches = 0;
        ws->mode_ms[0] = ws->mode_ms[1] = 0;
        ws->mode_launches[0] = ws->mode_launches[1] = 0;
    }
    return 0;
}
int tsg_treeseq_get_profile(const tsg_treeseq *ts, double *sweep_ms, uint64_t *launches) {
    Use use(ts);
    std::lock_guard<std::mutex> lock(g_mutex);
    const Resident *res = res_of(ts);
    *sweep_ms = 0.0;
    *launches = 0;
    if (res == nullptr) return 0;
    for (const Workspace *ws : res->all) {
        *sweep_ms += ws->sweep_ms;
        *launches += ws->sweep_launches;
    }
    return 0;
}

// The same per statistics mode: ms[0] / launches[0] branch sweeps, [1] site.
int tsg_treeseq_get_profile_modes(const tsg_treeseq *ts, double *ms, uint64_t *launches) {
    Use use(ts);
    std::lock_guard<std::mutex> lock(g_mutex);
    const Resident *res = res_of(ts);
    for (int m = 0; m < 2; m++) {
        ms[m] = 0.0;
        launches[m] = 0;
        if (res == nullptr) continue;
        for (const Workspace *ws : res->all) {
            ms[m] += ws->mode_ms[m];
            launches[m] += ws->mode_launches[m];
        }
    }
    return 0;
}

// ---- one-way statistics (trees.h:1098-1110) -----------------------------
static int one_way(tsg_treeseq *ts, int stat, uint64_t K, const uint64_t *sizes,
                   const int32_t *sets, uint64_t nw, const double *windows, uint32_t options,
                   double *result) {
    Query q;
    int ret = parse_sample_sets(ts, K, sizes, sets, q.sets);
    if (ret != 0) return ret;
    ret = parse_mode_windows(ts, options, nw, windows, q);
    if (ret != 0) return ret;
    const bool pol = (options & TSG_STAT_POLARISED) != 0;
    q.blocks.push_back(Block{stat, 0, {}, pol});
    q.P = (int) K;
    q.result = result;
    return run_one(ts, q);
}

int tsg_treeseq_diversity(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes, const int32_t *sets,
                          uint64_t nw, const double *windows, uint32_t options, double *result) {
    return one_way(ts, TSG_STAT_DIVERSITY, K, sizes, sets, nw, windows, options, result);
}
int tsg_treeseq_segregating_sites(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes,
                                  const int32_t *sets, uint64_t nw, const double *windows,
                                  uint32_t options, double *result) {
    return one_way(ts, TSG_STAT_SEGREGATING_SITES, K, sizes, sets, nw, windows, options, result);
}
int tsg_treeseq_Y1(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes, const int32_t *sets,
                   uint64_t nw, const double *windows, uint32_t options, double *result) {
    return one_way(ts, TSG_STAT_Y1, K, sizes, sets, nw, windows, options, result);
}

// ---- k-way statistics (trees.h:1112-1239) -------------------------------
static int k_way(tsg_treeseq *ts, int stat, int tuple, uint64_t K, const uint64_t *sizes,
                 const int32_t *sets, uint64_t M, const int32_t *idx, uint64_t nw,
                 const double *windows, uint32_t options, double *result) {
    int ret = check_tuples(K, tuple, M, idx);
    if (ret != 0) return ret;
    Query q;
    ret = parse_sample_sets(ts, K, sizes, sets, q.sets);
    if (ret != 0) return ret;
    ret = parse_mode_windows(ts, options, nw, windows, q);
    if (ret != 0) return ret;
    const bool pol = (options & TSG_STAT_POLARISED) != 0;
    q.ncentre = (int) K;
    q.blocks.push_back(Block{stat, tuple, std::vector<int32_t>(idx, idx + M * tuple), pol});
    q.P = (int) M;
    q.result = result;
    return run_one(ts, q);
}

#define KWAY(name, stat, tuple)                                                                 \
    int name(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes, const int32_t *sets, uint64_t M, \
             const int32_t *idx, uint64_t nw, const double *windows, uint32_t options,          \
             double *result) {                                                                  \
        return k_way(ts, stat, tuple, K, sizes, sets, M, idx, nw, windows, options, result);    \
    }
KWAY(tsg_treeseq_divergence, TSG_STAT_DIVERGENCE, 2)
KWAY(tsg_treeseq_Y2, TSG_STAT_Y2, 2)
KWAY(tsg_treeseq_f2, TSG_STAT_F2, 2)
KWAY(tsg_treeseq_Y3, TSG_STAT_Y3, 3)
KWAY(tsg_treeseq_f3, TSG_STAT_F3, 3)
KWAY(tsg_treeseq_f4, TSG_STAT_F4, 4)
#undef KWAY

int tsg_treeseq_genetic_relatedness(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes,
                                    const int32_t *sets, uint64_t M, const int32_t *idx,
                                    uint64_t nw, const double *windows, uint32_t options,
                                    double *result) {
    const int stat = (options & TSG_STAT_NONCENTRED) ? TSG_STAT_RELATEDNESS_NONCENTRED
                                                     : TSG_STAT_RELATEDNESS;
    return k_way(ts, stat, 2, K, sizes, sets, M, idx, nw, windows, options, result);
}

// ---- composites of the reference's Python layer, fused into one sweep ----
// Fst (trees.py:10024-10136): diversity of every set and divergence of each
// pair in one sweep, then 1 - 2(d_u + d_v) / (d_u + d_v + 2 d_uv) in FP64.
int tsg_treeseq_Fst(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes, const int32_t *sets,
                    uint64_t M, const int32_t *idx, uint64_t nw, const double *windows,
                    uint32_t options, double *result) {
    int ret = check_tuples(K, 2, M, idx);
    if (ret != 0) return ret;
    Query q;
    ret = parse_sample_sets(ts, K, sizes, sets, q.sets);
    if (ret != 0) return ret;
    ret = parse_mode_windows(ts, options, nw, windows, q);
    if (ret != 0) return ret;
    const bool pol = (options & TSG_STAT_POLARISED) != 0;
    q.blocks.push_back(Block{TSG_STAT_DIVERSITY, 0, {}, pol});
    q.blocks.push_back(Block{TSG_STAT_DIVERGENCE, 2, std::vector<int32_t>(idx, idx + 2 * M), pol});
    q.final_kind = TSG_FINAL_FST;
    q.fin_ipar.assign(idx, idx + 2 * M);
    q.P = (int) M;
    q.result = result;
    return run_one(ts, q);
}

// Tajima's D (trees.py:9967-10022): diversity and segregating sites, not span
// normalised, in one sweep, then the closed form in FP64.
int tsg_treeseq_Tajimas_D(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes, const int32_t *sets,
                          uint64_t nw, const double *windows, uint32_t options, double *result) {
    Query q;
    int ret = parse_sample_sets(ts, K, sizes, sets, q.sets);
    if (ret != 0) return ret;
    ret = parse_mode_windows(ts, options & ~TSG_STAT_SPAN_NORMALISE, nw, windows, q);
    if (ret != 0) return ret;
    q.blocks.push_back(Block{TSG_STAT_DIVERSITY, 0, {}, false});
    q.blocks.push_back(Block{TSG_STAT_SEGREGATING_SITES, 0, {}, false});
    q.final_kind = TSG_FINAL_TAJIMAS_D;
    // h(n) = sum_{i<n} 1/i and g(n) = sum_{i<n} 1/i^2, accumulated in
    // increasing i: prefix sums kept per thread and extended on demand (the
    // same additions in the same order as a fresh loop)
    thread_local std::vector<double> hsum{0.0}, gsum{0.0};
    for (uint64_t k = 0; k < K; k++) {
        const uint64_t nk = sizes[k];
        while (hsum.size() < std::max<uint64_t>(nk, 1)) {
            const double i = (double) hsum.size();
            hsum.push_back(hsum.back() + 1.0 / i);
            gsum.push_back(gsum.back() + 1.0 / (i * i));
        }
        const size_t at = (size_t) (nk > 0 ? nk - 1 : 0);
        q.fin_dpar.push_back((double) nk);
        q.fin_dpar.push_back(hsum[at]);
        q.fin_dpar.push_back(gsum[at]);
    }
    q.P = (int) K;
    q.result = result;
    return run_one(ts, q);
}

// genetic_relatedness with proportion=True (trees.py:8830-8976): the
// relatedness of each pair divided by the (unpolarised) segregating sites of
// the union of all sets, both from one sweep.
int tsg_treeseq_genetic_relatedness_proportion(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes,
                                               const int32_t *sets, uint64_t M,
                                               const int32_t *idx, uint64_t nw,
                                               const double *windows, uint32_t options,
                                               double *result) {
    int ret = check_tuples(K, 2, M, idx);
    if (ret != 0) return ret;
    Query q;
    ret = parse_sample_sets(ts, K, sizes, sets, q.sets);
    if (ret != 0) return ret;
    ret = parse_mode_windows(ts, options, nw, windows, q);
    if (ret != 0) return ret;
    const bool pol = (options & TSG_STAT_POLARISED) != 0;
    const int stat = (options & TSG_STAT_NONCENTRED) ? TSG_STAT_RELATEDNESS_NONCENTRED
                                                     : TSG_STAT_RELATEDNESS;
    q.ncentre = (int) K;
    std::vector<char> inunion(ts->num_samples, 0);
    std::vector<int32_t> uni;
    for (const auto &s : q.sets) {
        for (int32_t si : s) {
            if (!inunion[si]) {
                inunion[si] = 1;
                uni.push_back(si);
            }
        }
    }
    q.sets.push_back(uni);
    q.blocks.push_back(Block{stat, 2, std::vector<int32_t>(idx, idx + 2 * M), pol});
    // segregating sites of the union, a one-output block on column K
    q.blocks.push_back(Block{TSG_STAT_SEGREGATING_SITES, 1, {(int32_t) K}, false});
    q.final_kind = TSG_FINAL_PROPORTION;
    q.P = (int) M;
    q.result = result;
    return run_one(ts, q);
}

// ---- relatedness matrix on the matrix cores --------------------------------
// genetic_relatedness_matrix(mode="site") of trees.py:8978-9102 (divergence
// matrix trees.c:8900-9001, then the double centring of trees.py:9086-9095)
// for singleton sample sets and sites with at most one mutation: K = J C J,
// C = G^T G on int8 MFMA.  sample_sets may be NULL (all samples); result is
// W x n x n doubles.
// Site-mode matrices of singleton sample sets from C = G^T G (int8 matrix
// cores): the centred relatedness K = J C J / span, or with `divergence` the
// pairwise count of differing sites C_ii + C_jj - 2 C_ij / span.
// tsk_treeseq_divergence_matrix in branch mode (trees.c:8579-8745, which
// sums the path length between every two samples over the trees, from the
// samples to their MRCA or, in different subtrees, to their own roots).
// Here from the shared branch length
//   G_ij = integral over the window of sum_u bl(u) [i below u] [j below u],
// the branch-mode noncentred relatedness of singleton sets (summary x_i x_j,
// trees.c:4729-4753 with weight 1), computed by the fused sweep in output
// slices: the path length is G_ii + G_jj - 2 G_ij in every tree (G_ij = 0
// for samples in different subtrees).  Windows that do not span the
// sequence are padded at both ends for the sweep.  The pairs are then summed
// per pair of sets and count-normalised (trees.c:8876-8899) on the host, an
// O(W n^2) pass over the sweep's result.
namespace {
int relvec_run(tsg_treeseq *ts, int K, const std::vector<double> &wc,
               const std::vector<int32_t> &ident, const std::vector<double> &win,
               const std::vector<int32_t> &fs, bool centre, bool span_normalise,
               double *result);
}  // namespace

// Branch divergence matrix (trees.c:8579-8745): per window the branch length
// G_ij that samples i and j share, integrated along the genome, gives their
// path length G_ii + G_jj - 2 G_ij.  G is the relatedness vector of the
// identity weights (sample j weighs 1 in column j; tsg_relvec.hip), noncentred;
// then per pair of sets the sum over their samples' pairs, count-normalised
// (trees.c:8876-8899).  A tree sequence without a usable hot plan takes the
// fused sweep of the noncentred relatedness of singleton sets instead
// (n(n+1)/2 outputs, at most 2^22 of them).
static int branch_divergence_matrix(tsg_treeseq *ts, const std::vector<int32_t> &sel,
                                    const std::vector<int32_t> &set_off,
                                    const std::vector<double> &win, bool span_normalise,
                                    double *result) {
    const int n = (int) sel.size();
    const int nsets = (int) set_off.size() - 1;
    const int W = (int) win.size() - 1;
    // G [W x n x n]
    std::vector<double> G;
    int ret;
    if (ts->hot.usable) {
        if ((double) W * n * (double) n >= 1e9) return TSG_ERR_UNSUPPORTED_INPUT;
        G.assign((size_t) W * n * n, 0.0);
        if ((ret = relvec_run(ts, n, {}, sel, win, sel, false, false, G.data()))) return ret;
    } else {
        const int64_t M = (int64_t) n * (n + 1) / 2;
        if (M > (1 << 22)) return TSG_ERR_UNSUPPORTED_INPUT;
        std::vector<double> fw;
        int wfirst = 0;
        if (win[0] > 0) {
            fw.push_back(0.0);
            wfirst = 1;
        }
        fw.insert(fw.end(), win.begin(), win.end());
        if (win.back() < ts->L) fw.push_back(ts->L);
        Query q;
        q.mode = TSG_STAT_BRANCH;
        q.span_normalise = false;
        q.sets.resize(n);
        for (int i = 0; i < n; i++) q.sets[i] = {sel[i]};
        Block b;
        b.stat = TSG_STAT_RELATEDNESS_NONCENTRED;
        b.tuple = 2;
        b.polarised = true;
        b.idx.reserve((size_t) 2 * M);
        for (int i = 0; i < n; i++) {
            for (int j = i; j < n; j++) {
                b.idx.push_back(i);
                b.idx.push_back(j);
            }
        }
        q.blocks.push_back(std::move(b));
        q.windows = fw;
        q.final_kind = TSG_FINAL_NONE;
        q.P = (int) M;
        std::vector<double> Gp((size_t) (fw.size() - 1) * M);
        q.result = Gp.data();
        std::vector<Query *> batch{&q};
        if ((ret = run_batch(ts, batch))) return ret;
        auto pair = [&](int i, int j) -> int64_t {  // (i <= j) in the output order above
            return (int64_t) i * n - (int64_t) i * (i - 1) / 2 + (j - i);
        };
        G.assign((size_t) W * n * n, 0.0);
        for (int w = 0; w < W; w++) {
            const double *g = Gp.data() + (size_t) (w + wfirst) * M;
            for (int i = 0; i < n; i++) {
                for (int j = 0; j < n; j++) {
                    G[((size_t) w * n + i) * n + j] = g[pair(std::min(i, j), std::max(i, j))];
                }
            }
        }
    }
    std::vector<double> D((size_t) n * n);
    for (int w = 0; w < W; w++) {
        const double *g = G.data() + (size_t) w * n * n;
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) {
                // G from the relatedness vector is symmetric only to rounding:
                // both off-diagonal entries enter, so D is exactly symmetric
                D[(size_t) i * n + j] = (i == j) ? 0.0
                                                 : g[(size_t) i * n + i] + g[(size_t) j * n + j] -
                                                       (g[(size_t) i * n + j] + g[(size_t) j * n + i]);
            }
        }
        const double span = win[w + 1] - win[w];
        double *out = result + (size_t) w * nsets * nsets;
        for (int A = 0; A < nsets; A++) {
            for (int B = A; B < nsets; B++) {  // (B, A) mirrors (A, B)
                double v = 0.0;
                for (int i = set_off[A]; i < set_off[A + 1]; i++) {
                    for (int j = set_off[B]; j < set_off[B + 1]; j++) v += D[(size_t) i * n + j];
                }
                const double na = set_off[A + 1] - set_off[A], nb = set_off[B + 1] - set_off[B];
                const double denom = (A == B) ? na * (na - 1) : na * nb;
                if (!(A == B && denom == 0)) v /= denom;
                out[(size_t) A * nsets + B] = out[(size_t) B * nsets + A] =
                    span_normalise ? v / span : v;
            }
        }
    }
    return 0;
}

static int site_gram_matrix(tsg_treeseq *ts, uint64_t num_sample_sets,
                            const uint64_t *sample_set_sizes, const int32_t *sample_sets,
                            uint64_t nw, const double *windows, uint32_t options, bool divergence,
                            double *result) {
    bool site = options & TSG_STAT_SITE, branch = options & TSG_STAT_BRANCH;
    if (options & TSG_STAT_NODE) return TSG_ERR_UNSUPPORTED_STAT_MODE;
    if (!(site || branch)) site = true;
    if (site && branch) return TSG_ERR_MULTIPLE_STAT_MODES;
    if (options & TSG_STAT_POLARISED) return -911;  // TSK_ERR_STAT_POLARISED_UNSUPPORTED
    std::vector<double> win;
    if (windows == nullptr) {
        win = {0.0, ts->L};
    } else {  // tsk_treeseq_check_windows without the full-span requirement
        if (nw < 1) return TSG_ERR_BAD_NUM_WINDOWS;
        if (windows[0] < 0 || windows[nw] > ts->L) return TSG_ERR_BAD_WINDOWS;
        for (uint64_t j = 0; j < nw; j++) {
            if (windows[j] >= windows[j + 1]) return TSG_ERR_BAD_WINDOWS;
        }
        win.assign(windows, windows + nw + 1);
    }
    // sample sets (trees.c:8833-8873): disjoint sets of samples; the
    // relatedness matrix takes singletons only.  sel lists the sets' samples
    // in order, set k holding sel[set_off[k] .. set_off[k + 1]).
    std::vector<int32_t> sel, set_off{0};
    bool singletons = true;
    if (sample_sets == nullptr) {
        for (int32_t j = 0; j < ts->num_samples; j++) {
            sel.push_back(j);
            set_off.push_back(j + 1);
        }
    } else {
        std::vector<char> seen(ts->num_samples, 0);
        uint64_t i = 0;
        for (uint64_t k = 0; k < num_sample_sets; k++) {
            const uint64_t sz = sample_set_sizes ? sample_set_sizes[k] : 1;
            for (uint64_t l = 0; l < sz; l++) {
                const int32_t u = sample_sets[i++];
                if (u < 0 || u >= ts->num_nodes) return TSG_ERR_NODE_OUT_OF_BOUNDS;
                const int32_t si = ts->sample_index_map[u];
                if (si == TSG_NULL) return TSG_ERR_BAD_SAMPLES;
                if (seen[si]) return TSG_ERR_DUPLICATE_SAMPLE;
                seen[si] = 1;
                sel.push_back(si);
            }
            singletons &= (sz == 1);
            set_off.push_back((int32_t) sel.size());
        }
        if (!singletons && !divergence) return TSG_ERR_UNSUPPORTED_INPUT;
    }
    const int nsets = (int) set_off.size() - 1;
    if (branch) {
        if (!divergence) return TSG_ERR_UNSUPPORTED_STAT_MODE;  // relatedness: site mode only
        if (ts->time_uncalibrated && !(options & TSG_STAT_ALLOW_TIME_UNCALIBRATED)) {
            return TSG_ERR_TIME_UNCALIBRATED;
        }
        if (sel.empty() || nsets == 0) return 0;
        return branch_divergence_matrix(ts, sel, set_off, win,
                                        (options & TSG_STAT_SPAN_NORMALISE) != 0, result);
    }
    // sites with three or more alleles get extra bit rows: their any-derived
    // row (row s) and one row per derived allele (rows site_row2[s] ..)
    std::vector<int32_t> site_row2(std::max(ts->num_sites, 1), -1);
    int64_t nrows = ts->num_sites;
    bool multi = false;
    for (int32_t s = 0; s < ts->num_sites; s++) {
        if (ts->site_nalleles[s] >= 3) {
            site_row2[s] = (int32_t) nrows;
            nrows += ts->site_nalleles[s] - 1;
            multi = true;
        }
    }
    if (nrows >= INT32_MAX) return TSG_ERR_UNSUPPORTED_INPUT;
    const int W = (int) win.size() - 1;
    const int n = (int) sel.size();
    if (n == 0 || nsets == 0) return 0;
    // grids: windows on y, set pairs on x
    if (W > 65535 || (int64_t) nsets * nsets > INT32_MAX) return TSG_ERR_UNSUPPORTED_INPUT;
    const int ns = ts->num_samples;
    const int n_pad = ((ns + 127) / 128) * 128;
    const int NW = (ns + 31) / 32;
    // per window: the bit rows of its sites, padded to a multiple of 64
    // columns of Gt -- sites with at most two alleles in Gt, the rows of the
    // others in Gt2
    std::vector<int64_t> wk0(W), wk1(W), wk0b(W), wk1b(W);
    std::vector<int32_t> sitecol, sitecol2;
    for (int w = 0; w < W; w++) {
        const int32_t s0 = (int32_t) (std::lower_bound(ts->site_pos.begin(), ts->site_pos.end(), win[w]) -
                                      ts->site_pos.begin());
        const int32_t s1 = (int32_t) (std::lower_bound(ts->site_pos.begin(), ts->site_pos.end(), win[w + 1]) -
                                      ts->site_pos.begin());
        wk0[w] = (int64_t) sitecol.size();
        wk0b[w] = (int64_t) sitecol2.size();
        for (int32_t s = s0; s < s1; s++) {
            if (site_row2[s] < 0) {
                sitecol.push_back(s);
            } else {
                sitecol2.push_back(s);
                for (int al = 1; al < ts->site_nalleles[s]; al++) {
                    sitecol2.push_back(site_row2[s] + al - 1);
                }
            }
        }
        while (sitecol.size() % 64 != 0) sitecol.push_back(-1);
        while (sitecol2.size() % 64 != 0) sitecol2.push_back(-1);
        wk1[w] = (int64_t) sitecol.size();
        wk1b[w] = (int64_t) sitecol2.size();
    }
    if (sitecol.empty()) sitecol.assign(64, -1);
    if (sitecol2.empty()) sitecol2.assign(64, -1);
    const int64_t Kp = (int64_t) sitecol.size(), Kp2 = (int64_t) sitecol2.size();
    std::vector<double> scale(W, 1.0);
    if (options & TSG_STAT_SPAN_NORMALISE) {
        for (int w = 0; w < W; w++) scale[w] = 1.0 / (win[w + 1] - win[w]);
    }

    int ret;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = device_upload(ts))) return ret;
    }
    Lease lease(ts);
    Workspace *ws = lease.ws;
    if (ws == nullptr) return TSG_ERR_DEVICE;
    // genotype bitsets on the hot plan when a block of WB words per slot
    // fits in LDS (WB = 64, 32 or 16 words: a workgroup's share of samples)
    int WB = 0;
    if (ts->hot.usable && !ws->force_full) {
        for (int wb : {64, 32, 16}) {
            if ((size_t) ts->hot.max_slots * wb * 4 <= (size_t) 48 * 1024) {
                WB = wb;
                break;
            }
        }
    }
    if (WB == 0) {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = ensure_full(ts))) return ret;
    }
    hipStream_t s = ws->stream;
    const int slots = std::min(ts->num_chunks, 512);
    const int64_t gstride = WB ? 1 : (int64_t) ts->max_nloc * (1 + NW);
    if (multi && ((ret = upload(ws->sitecol2, sitecol2, s)) || (ret = upload(ws->wk0b, wk0b, s)) ||
                  (ret = upload(ws->wk1b, wk1b, s)) || (ret = upload(ws->row2, site_row2, s)) ||
                  (ret = ensure(ws->gt2, (size_t) n_pad * Kp2)) ||
                  (ret = ensure(ws->cmat2, (size_t) W * n_pad * n_pad * 4)))) {
        return ret;
    }
    if ((ret = ensure(ws->gbits, (size_t) std::max<int64_t>(nrows, 1) * NW * 4)) ||
        (ret = ensure(ws->gslab, (size_t) (WB ? 1 : slots) * gstride * 4)) ||
        (ret = upload(ws->sitecol, sitecol, s)) || (ret = upload(ws->wk0, wk0, s)) ||
        (ret = upload(ws->wk1, wk1, s)) || (ret = upload(ws->scale, scale, s)) ||
        (ret = upload(ws->sel, sel, s)) || (ret = ensure(ws->gt, (size_t) n_pad * Kp)) ||
        (ret = ensure(ws->cmat, (size_t) W * n_pad * n_pad * 4)) ||
        (ret = ensure(ws->rowsum, (size_t) (W * (int64_t) n + W) * 8)) ||
        (ret = ensure(ws->kout, (size_t) W * n * (size_t) n * 8))) {
        return ret;
    }
    if (!singletons) {
        if ((ret = ensure(ws->out, (size_t) W * nsets * (size_t) nsets * 8)) ||
            (ret = upload(ws->ipar, set_off, s))) {
            return ret;
        }
    }
    const TsgDevice &d = ts->dev;
    const TsgHotDevice &hd = ts->hot.dev;
    TsgGenoArgs g{};
    g.tree_site0 = d.tree_site0;
    g.site_mut0 = d.site_mut0;
    g.num_samples = ns;
    g.NW = NW;
    g.gbits = (uint32_t *) ws->gbits.p;
    g.mut_allele = d.mut_allele;
    g.site_nalleles = d.site_nalleles;
    g.site_row2 = multi ? (const int32_t *) ws->row2.p : nullptr;
    if (WB) {
        g.chunk_tree0 = hd.chunk_tree0;
        g.chunk_node0 = hd.chunk_slot0;
        g.tr_off = hd.tr_off;
        g.tr_nrem = hd.tr_nrem;
        g.ev_child = hd.ev_child;
        g.ev_path0 = hd.ev_path0;
        g.path = hd.path;
        g.mut_node = hd.mut_slot;
        g.slot_par = hd.slot_par;
        g.chunk_lev0 = hd.chunk_lev0;
        g.lev_slot0 = hd.lev_slot0;
        g.sample_slot = hd.sample_slot;
        g.num_chunks = ts->hot.num_chunks;
        g.WB = WB;
        ws->last_store = TSG_STORE_HOT;
    } else {
        g.chunk_tree0 = d.chunk_tree0;
        g.chunk_node0 = d.chunk_node0;
        g.loc_parent0 = d.loc_parent0;
        g.tr_off = d.tr_off;
        g.tr_nrem = d.tr_nrem;
        g.ev_child = d.ev_child;
        g.ev_parent = d.ev_parent;
        g.mut_node = d.mut_node;
        g.num_chunks = ts->num_chunks;
        g.slots = slots;
        g.g_rec = (uint32_t *) ws->gslab.p;
        g.g_stride = gstride;
        ws->last_store = TSG_STORE_GLOBAL;
    }
    if (ws->profile) (void) hipEventRecord(ws->ev0, s);
    if ((ret = hip_check(WB ? tsg_launch_geno_hot(g, (size_t) ts->hot.max_slots * WB * 4, s)
                            : tsg_launch_geno_bits(g, s))) ||
        (ret = hip_check(tsg_launch_bits_to_gt((const uint32_t *) ws->gbits.p, NW,
                                               (const int32_t *) ws->sitecol.p, Kp, ns, n_pad,
                                               (int8_t *) ws->gt.p, s))) ||
        (ret = hip_check(tsg_launch_syrk_i8((const int8_t *) ws->gt.p, Kp, n_pad,
                                            (const int64_t *) ws->wk0.p,
                                            (const int64_t *) ws->wk1.p, W,
                                            (int32_t *) ws->cmat.p, s)))) {
        return ret;
    }
    const int32_t *C2 = nullptr;
    if (multi) {
        if ((ret = hip_check(tsg_launch_bits_to_gt((const uint32_t *) ws->gbits.p, NW,
                                                   (const int32_t *) ws->sitecol2.p, Kp2, ns,
                                                   n_pad, (int8_t *) ws->gt2.p, s))) ||
            (ret = hip_check(tsg_launch_syrk_i8((const int8_t *) ws->gt2.p, Kp2, n_pad,
                                                (const int64_t *) ws->wk0b.p,
                                                (const int64_t *) ws->wk1b.p, W,
                                                (int32_t *) ws->cmat2.p, s)))) {
            return ret;
        }
        C2 = (const int32_t *) ws->cmat2.p;
    }
    if ((ret = hip_check(divergence
                             ? tsg_launch_divergence((const int32_t *) ws->cmat.p, C2,
                                                     (const int32_t *) ws->sel.p, n, n_pad, W,
                                                     (const double *) ws->scale.p,
                                                     (double *) ws->kout.p, s)
                             : tsg_launch_centre((const int32_t *) ws->cmat.p, C2,
                                                 (const int32_t *) ws->sel.p,
                                                 (int64_t *) ws->rowsum.p, n, n_pad, W,
                                                 (const double *) ws->scale.p,
                                                 (double *) ws->kout.p, s)))) {
        return ret;
    }
    // sample sets of several samples: per pair of sets the sum over their
    // samples' pairs, count-normalised (trees.c:8876-8899)
    if (!singletons &&
        (ret = hip_check(tsg_launch_set_divergence((const double *) ws->kout.p, n, W,
                                                   (const int32_t *) ws->ipar.p, nsets,
                                                   (double *) ws->out.p, s)))) {
        return ret;
    }
    if (ws->profile) (void) hipEventRecord(ws->ev1, s);
    const void *res_dev = singletons ? ws->kout.p : ws->out.p;
    const int nr = singletons ? n : nsets;
    if ((ret = copy_out(ws, s, result, res_dev, (size_t) W * nr * (size_t) nr * 8))) return ret;
    if (ws->profile) {
        float ms = 0;
        (void) hipEventElapsedTime(&ms, ws->ev0, ws->ev1);
        ws->sweep_ms += ms;
        ws->sweep_launches += 1;
    }
    return 0;
}

int tsg_treeseq_genetic_relatedness_matrix(tsg_treeseq *ts, uint64_t num_sample_sets,
                                           const uint64_t *sample_set_sizes,
                                           const int32_t *sample_sets, uint64_t nw,
                                           const double *windows, uint32_t options,
                                           double *result) {
    Use use(ts);
    return site_gram_matrix(ts, num_sample_sets, sample_set_sizes, sample_sets, nw, windows,
                            options, false, result);
}

// tsk_treeseq_divergence_matrix (trees.c:8902-9001), site mode, sample sets
// of one sample each (the normalisation of fill_lower_triangle_count_normalise
// then divides off-diagonal entries by 1 and leaves the zero diagonal).
int tsg_treeseq_divergence_matrix(tsg_treeseq *ts, uint64_t num_sample_sets,
                                  const uint64_t *sample_set_sizes, const int32_t *sample_sets,
                                  uint64_t nw, const double *windows, uint32_t options,
                                  double *result) {
    Use use(ts);
    return site_gram_matrix(ts, num_sample_sets, sample_set_sizes, sample_sets, nw, windows,
                            options, true, result);
}

// ---- genotype decode ------------------------------------------------------
// (site, sample index) pairs where a sample is isolated -- no edge names it as
// parent or child over the site's position -- and no mutation of the site
// sits on it: the reference marks such samples missing before applying the
// site's mutations, and only a mutation on the node itself reaches an
// isolated node (genotypes.c:413-436, 473-595).
static void isolated_pairs(const tsg_treeseq *ts, std::vector<int32_t> &ms,
                           std::vector<int32_t> &mj) {
    const int32_t n = ts->num_samples;
    std::vector<std::vector<std::pair<double, double>>> cover(n);
    for (int32_t e = 0; e < ts->num_edges; e++) {
        for (int32_t u : {ts->parent[e], ts->child[e]}) {
            const int32_t j = ts->sample_index_map[u];
            if (j != TSG_NULL) cover[j].emplace_back(ts->left[e], ts->right[e]);
        }
    }
    auto sites_in = [&](double a, double b, int32_t j) {
        const auto &pos = ts->site_pos;
        size_t s = std::lower_bound(pos.begin(), pos.end(), a) - pos.begin();
        const size_t s1 = std::lower_bound(pos.begin(), pos.end(), b) - pos.begin();
        for (; s < s1; s++) {
            bool mut_on = false;
            for (int32_t m = ts->site_mut0[s]; m < ts->site_mut0[s + 1]; m++) {
                mut_on |= ts->mut_node[m] == ts->samples[j];
            }
            if (!mut_on) {
                ms.push_back((int32_t) s);
                mj.push_back(j);
            }
        }
    };
    for (int32_t j = 0; j < n; j++) {
        auto &iv = cover[j];
        std::sort(iv.begin(), iv.end());
        double x = 0.0;  // covered up to x so far
        for (const auto &[l, r] : iv) {
            if (l > x) sites_in(x, l, j);
            x = std::max(x, r);
        }
        if (x < ts->L) sites_in(x, ts->L, j);
    }
}

// Genotypes of every sample at every site, [S x n] int32, as
// TreeSequence.genotype_matrix (python/tskit/trees.py:5563-5640) builds them
// from tsk_variant_decode (genotypes.c:473-595): allele indexes in order of
// first appearance (0 = ancestral), -1 for isolated samples unless options
// has TSG_ISOLATED_NOT_MISSING.  Runs on the hot plan.
namespace {
// The relatedness-vector program (tsg_relvec.hip) for K weight columns --
// dense weights wc [num_samples x K] (already centred or not), or identity
// columns (ident[k] = the sample column k weighs 1 at) -- over windows win
// (need not span the sequence), focal samples fs (sample indexes): result
// [W x nf x K] on the host, centred over the focal nodes when `centre`,
// span-normalised when `span_normalise`.  The hot plan must be usable.
int relvec_run(tsg_treeseq *ts, int K, const std::vector<double> &wc,
               const std::vector<int32_t> &ident, const std::vector<double> &win,
               const std::vector<int32_t> &fs, bool centre, bool span_normalise,
               double *result) {
    const int W = (int) win.size() - 1;
    const int64_t nf = (int64_t) fs.size();
    const int n = ts->num_samples;
    int ret;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = device_upload(ts))) return ret;
    }
    const tsg_hotplan &hp = ts->hot;
    if (!hp.usable) return TSG_ERR_UNSUPPORTED_INPUT;
    Lease lease(ts);
    Workspace *ws = lease.ws;
    if (ws == nullptr) return TSG_ERR_DEVICE;
    int Kb = std::min(K, TSG_RELVEC_KMAX);
    while (Kb > 1 && tsg_relvec_shmem(hp.max_slots, Kb) > (size_t) ws->lds_limit) Kb--;
    if (tsg_relvec_shmem(hp.max_slots, Kb) > (size_t) ws->lds_limit) {
        return TSG_ERR_UNSUPPORTED_INPUT;
    }
    const int C = hp.num_chunks;
    // per chunk: the windows it overlaps and where its segment sums go
    std::vector<int32_t> wfirst(C), nwin(C);
    std::vector<int64_t> ioff(C + 1, 0);
    {
        int m = 0;
        for (int c = 0; c < C; c++) {
            const double ca = ts->breakpoints[hp.chunk_tree0[c]];
            const double cb = ts->breakpoints[hp.chunk_tree0[c + 1]];
            while (m < W && win[m + 1] <= ca) m++;
            int m1 = m;
            while (m1 < W && win[m1] < cb) m1++;
            wfirst[c] = m;
            nwin[c] = std::max(0, m1 - m);
            const int64_t S = hp.chunk_slot0[c + 1] - hp.chunk_slot0[c];
            ioff[c + 1] = ioff[c] + (int64_t) nwin[c] * S * Kb;
        }
    }
    const int nranges = (int) std::max<int64_t>(
        1, std::min<int64_t>({(65536 + nf - 1) / nf, 256, (int64_t) C}));
    const int nloc0 = (int) hp.t0_parent.size();
    hipStream_t s = ws->stream;
    // (the packer keeps pointers to the vectors until it sends them)
    Pack pk;
    const int p_w = pk.add(wc), p_id = pk.add(ident), p_win = pk.add(win);
    const int p_wf = pk.add(wfirst), p_nw = pk.add(nwin), p_io = pk.add(ioff);
    const int p_fs = pk.add(fs);
    if ((ret = ensure(ws, ws->rv_E, (size_t) hp.chunk_slot0[C] * Kb * 8)) ||
        (ret = ensure(ws, ws->rv_I, (size_t) std::max<int64_t>(ioff[C], 1) * 8)) ||
        (ret = ensure(ws, ws->rv_z0, (size_t) std::max(nloc0, 1) * Kb * 8)) ||
        (ret = ensure(ws, ws->rv_d0, (size_t) nf * Kb * 8)) ||
        (ret = ensure(ws, ws->rv_esum, (size_t) nranges * nf * Kb * 8)) ||
        (ret = ensure(ws, ws->rv_y, (size_t) W * nf * K * 8)) || (ret = pk.send(ws, s))) {
        return ret;
    }
    const TsgHotDevice &h = hp.dev;
    TsgRelvecArgs a{};
    a.chunk_tree0 = h.chunk_tree0;
    a.chunk_slot0 = h.chunk_slot0;
    a.slot_par = h.slot_par;
    a.slot_bl0 = h.slot_bl0;
    a.chunk_lev0 = h.chunk_lev0;
    a.lev_slot0 = h.lev_slot0;
    a.sample_slot = h.sample_slot;
    a.tr_off = h.tr_off;
    a.tr_nrem = h.tr_nrem;
    a.ev_child = h.ev_child;
    a.ev_bl = h.ev_bl;
    a.ev_path0 = h.ev_path0;
    a.path = h.path;
    a.bp = ts->dev.breakpoints;
    a.num_trees = ts->num_trees;
    a.num_chunks = C;
    a.num_samples = n;
    a.weights = wc.empty() ? nullptr : pk.dev<const double>(ws, p_w);
    a.ident = ident.empty() ? nullptr : pk.dev<const int32_t>(ws, p_id);
    a.K = K;
    a.windows = pk.dev<const double>(ws, p_win);
    a.num_windows = W;
    a.wfirst = pk.dev<const int32_t>(ws, p_wf);
    a.nwin = pk.dev<const int32_t>(ws, p_nw);
    a.ioff = pk.dev<const int64_t>(ws, p_io);
    a.focal = pk.dev<const int32_t>(ws, p_fs);
    a.nf = (int32_t) nf;
    a.nranges = nranges;
    a.nloc0 = nloc0;
    a.E = (double *) ws->rv_E.p;
    a.I = (double *) ws->rv_I.p;
    double *y = (double *) ws->rv_y.p;
    if ((ret = hip_check(tsg_launch_relvec_fill(y, (int64_t) W * nf * K, s)))) return ret;
    if (ws->profile) (void) hipEventRecord(ws->ev0, s);
    for (int k0 = 0; k0 < K; k0 += Kb) {
        a.k0 = k0;
        a.Kb = std::min(Kb, K - k0);
        if ((ret = hip_check(tsg_launch_relvec(a, tsg_relvec_shmem(hp.max_slots, a.Kb),
                                               h.t0_parent, h.t0_bl, nloc0,
                                               (double *) ws->rv_z0.p, (double *) ws->rv_d0.p,
                                               (double *) ws->rv_esum.p, y, s)))) {
            return ret;
        }
    }
    if (ws->profile) (void) hipEventRecord(ws->ev1, s);
    if ((ret = hip_check(tsg_launch_relvec_final(y, a.windows, W, (int) nf, K, centre,
                                                 span_normalise, s))) ||
        (ret = copy_out(ws, s, result, y, (size_t) W * nf * K * 8))) {
        return ret;
    }
    if (ws->profile) {
        float ms = 0;
        (void) hipEventElapsedTime(&ms, ws->ev0, ws->ev1);
        ws->sweep_ms += ms;
        ws->sweep_launches += 1;
    }
    return 0;
}
}  // namespace

// genetic_relatedness_vector, branch mode (trees.c:10402-10816, entry
// 10783): arguments in the reference's order (result before options).
// Weights [num_samples x num_weights], centred over the samples unless
// TSG_STAT_NONCENTRED; windows need not span the sequence; result
// [num_windows x num_focal_nodes x num_weights], centred over the focal
// nodes unless noncentred and span-normalised with TSG_STAT_SPAN_NORMALISE.
// On the hot plan (tsg_relvec.hip); focal nodes must be samples.
int tsg_treeseq_genetic_relatedness_vector(tsg_treeseq *ts, uint64_t num_weights,
                                           const double *weights, uint64_t num_windows,
                                           const double *windows, uint64_t num_focal_nodes,
                                           const int32_t *focal_nodes, double *result,
                                           uint32_t options) {
    Use use(ts);
    if (options & (TSG_STAT_SITE | TSG_STAT_NODE)) return TSG_ERR_UNSUPPORTED_STAT_MODE;
    // tsk_treeseq_check_windows without the full-span rule (trees.c:1244-1286)
    if (num_windows < 1) return TSG_ERR_BAD_NUM_WINDOWS;
    if (windows[0] < 0 || windows[num_windows] > ts->L) return TSG_ERR_BAD_WINDOWS;
    for (uint64_t j = 0; j < num_windows; j++) {
        if (!(windows[j] < windows[j + 1])) return TSG_ERR_BAD_WINDOWS;
    }
    const int64_t nf = (int64_t) num_focal_nodes;
    std::vector<int32_t> fs((size_t) nf);
    for (int64_t j = 0; j < nf; j++) {  // trees.c:10491-10496
        if (focal_nodes[j] < 0 || focal_nodes[j] >= ts->num_nodes) {
            return TSG_ERR_NODE_OUT_OF_BOUNDS;
        }
    }
    for (int64_t j = 0; j < nf; j++) {
        fs[(size_t) j] = ts->sample_index_map[focal_nodes[j]];
        if (fs[(size_t) j] < 0) return TSG_ERR_UNSUPPORTED_INPUT;
    }
    const int W = (int) num_windows;
    const int K = (int) num_weights;
    const int n = ts->num_samples;
    if (K == 0 || nf == 0) return 0;
    if ((double) W * (double) nf * (double) K >= 1e9) return TSG_ERR_UNSUPPORTED_INPUT;
    // centre the weights (trees.c:10517-10540)
    std::vector<double> wc(weights, weights + (size_t) n * K);
    if (!(options & TSG_STAT_NONCENTRED)) {
        std::vector<double> mean(K, 0.0);
        for (int j = 0; j < n; j++) {
            for (int k = 0; k < K; k++) mean[k] += wc[(size_t) j * K + k];
        }
        for (int k = 0; k < K; k++) mean[k] /= (double) n;
        for (int j = 0; j < n; j++) {
            for (int k = 0; k < K; k++) wc[(size_t) j * K + k] -= mean[k];
        }
    }
    return relvec_run(ts, K, wc, {}, std::vector<double>(windows, windows + W + 1), fs,
                      !(options & TSG_STAT_NONCENTRED),
                      (options & TSG_STAT_SPAN_NORMALISE) != 0, result);
}

// ---- general_stat with real-valued weights (trees.c:2036-2095) -----------
// The summary function is chosen from the registry by id (tsg_stat_id, the
// weighted ones) with its parameters, in place of general_stat_func_t and
// f_params (trees.h:1032-1037):
//   TSG_STAT_TRAIT_COVARIANCE        result_dim = state_dim (trees.c:3961)
//   TSG_STAT_TRAIT_CORRELATION       result_dim = state_dim - 1, the last
//                                    column the frequency (trees.c:4026)
//   TSG_STAT_TRAIT_LINEAR_MODEL      state = [traits | covariates | 1],
//                                    dparams = V [result_dim x covariates]
//                                    (trees.c:4108)
//   TSG_STAT_RELATEDNESS_WEIGHTED(_NONCENTRED)  index_tuples [result_dim x 2],
//                                    dparams = total weights [state_dim], the
//                                    last column the frequency (trees.c:4801)
// Branch and site mode on the hot plan (FP64 state columns); node mode and
// statistics inside a batch are not taken.
int tsg_treeseq_general_stat(tsg_treeseq *ts, uint64_t state_dim, const double *sample_weights,
                             uint64_t result_dim, int32_t summary, const int32_t *index_tuples,
                             const double *dparams, uint64_t num_windows, const double *windows,
                             uint32_t options, double *result) {
    // the reference's order of checks (trees.c:2048-2075)
    const int nmodes = !!(options & TSG_STAT_SITE) + !!(options & TSG_STAT_BRANCH) +
                       !!(options & TSG_STAT_NODE);
    if (nmodes > 1) return TSG_ERR_MULTIPLE_STAT_MODES;
    if (state_dim < 1) return TSG_ERR_BAD_STATE_DIMS;
    if (result_dim < 1) return TSG_ERR_BAD_RESULT_DIMS;
    Query q;
    int ret = parse_mode_windows(ts, options, num_windows, windows, q);
    if (ret != 0) return ret;
    if (options & TSG_STAT_NODE) return TSG_ERR_UNSUPPORTED_STAT_MODE;
    if (g_capture != nullptr) return TSG_ERR_UNSUPPORTED_INPUT;
    if (state_dim > 4096 || result_dim > (1u << 20)) return TSG_ERR_UNSUPPORTED_INPUT;
    const int K = (int) state_dim, P = (int) result_dim;
    const int n = ts->num_samples;
    const bool pol = (options & TSG_STAT_POLARISED) != 0;
    q.weighted = true;
    q.wK = K;
    q.wts.assign(sample_weights, sample_weights + (size_t) n * K);
    q.wouts.resize(P);
    for (int m = 0; m < P; m++) {
        TsgOut &o = q.wouts[m];
        o = TsgOut{};
        o.stat = summary;
        o.flags = pol ? 1 : 0;
        o.inv = 1.0;
        o.i = o.j = o.k = o.l = m < K ? m : 0;
        o.ni = n;
        o.par = 0;
    }
    switch (summary) {
    case TSG_STAT_TRAIT_COVARIANCE:
        if (P != K) return TSG_ERR_BAD_RESULT_DIMS;
        for (TsgOut &o : q.wouts) o.inv = 1.0 / (2 * ((double) n - 1) * ((double) n - 1));
        break;
    case TSG_STAT_TRAIT_CORRELATION:
        if (P != K - 1) return TSG_ERR_BAD_RESULT_DIMS;
        for (TsgOut &o : q.wouts) {
            o.l = K - 1;
            o.inv = 1.0 / (2 * (double) n * ((double) n - 1));
        }
        break;
    case TSG_STAT_TRAIT_LINEAR_MODEL: {
        if (P > K - 1) return TSG_ERR_BAD_RESULT_DIMS;
        const int nc = K - P - 1;
        if (nc > 0 && dparams == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
        if (nc > 0) q.opar.assign(dparams, dparams + (size_t) P * nc);
        for (int m = 0; m < P; m++) {
            TsgOut &o = q.wouts[m];
            o.cbase = P;
            o.ccount = nc;
            o.l = K - 1;
            o.par = m * nc;
        }
        break;
    }
    case TSG_STAT_RELATEDNESS_WEIGHTED:
    case TSG_STAT_RELATEDNESS_WEIGHTED_NONCENTRED: {
        if (index_tuples == nullptr) return TSG_ERR_INSUFFICIENT_INDEX_TUPLES;
        const bool centred = summary == TSG_STAT_RELATEDNESS_WEIGHTED;
        if (centred && dparams == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
        for (int m = 0; m < P; m++) {
            const int32_t i = index_tuples[2 * m], j = index_tuples[2 * m + 1];
            if (i < 0 || i >= K || j < 0 || j >= K) return TSG_ERR_BAD_SAMPLE_SET_INDEX;
            TsgOut &o = q.wouts[m];
            o.i = i;
            o.j = j;
            o.l = K - 1;
            if (centred) {
                o.par = (int32_t) q.opar.size();
                q.opar.push_back(dparams[i]);
                q.opar.push_back(dparams[j]);
            }
        }
        break;
    }
    default:
        return TSG_ERR_BAD_PARAM_VALUE;
    }
    if (q.opar.empty()) q.opar.push_back(0.0);
    q.P = P;
    q.result = result;
    return run_one(ts, q);
}

// The reference's weighted statistics (trees.c:3976-4220, 4839-4895): their
// weight preparation on the host, then general_stat.
int tsg_treeseq_trait_covariance(tsg_treeseq *ts, uint64_t num_weights, const double *weights,
                                 uint64_t num_windows, const double *windows, uint32_t options,
                                 double *result) {
    if (num_weights == 0) return TSG_ERR_INSUFFICIENT_WEIGHTS;
    const size_t n = (size_t) ts->num_samples, K = num_weights;
    std::vector<double> mean(K, 0.0), w(n * K);
    for (size_t j = 0; j < n; j++) {
        for (size_t k = 0; k < K; k++) mean[k] += weights[j * K + k];
    }
    for (size_t k = 0; k < K; k++) mean[k] /= (double) n;
    for (size_t j = 0; j < n; j++) {
        for (size_t k = 0; k < K; k++) w[j * K + k] = weights[j * K + k] - mean[k];
    }
    return tsg_treeseq_general_stat(ts, K, w.data(), K, TSG_STAT_TRAIT_COVARIANCE, nullptr,
                                    nullptr, num_windows, windows, options, result);
}

int tsg_treeseq_trait_correlation(tsg_treeseq *ts, uint64_t num_weights, const double *weights,
                                  uint64_t num_windows, const double *windows, uint32_t options,
                                  double *result) {
    if (num_weights < 1) return TSG_ERR_INSUFFICIENT_WEIGHTS;
    const size_t n = (size_t) ts->num_samples, K = num_weights;
    std::vector<double> mean(K, 0.0), msq(K, 0.0), sd(K), w(n * (K + 1));
    for (size_t j = 0; j < n; j++) {
        for (size_t k = 0; k < K; k++) {
            const double x = weights[j * K + k];
            mean[k] += x;
            msq[k] += x * x;
        }
    }
    for (size_t k = 0; k < K; k++) {
        mean[k] /= (double) n;
        msq[k] -= mean[k] * mean[k] * (double) n;
        msq[k] /= (double) (n - 1);
        sd[k] = std::sqrt(msq[k]);
    }
    for (size_t j = 0; j < n; j++) {
        for (size_t k = 0; k < K; k++) w[j * (K + 1) + k] = (weights[j * K + k] - mean[k]) / sd[k];
        w[j * (K + 1) + K] = 1.0 / (double) n;
    }
    return tsg_treeseq_general_stat(ts, K + 1, w.data(), K, TSG_STAT_TRAIT_CORRELATION, nullptr,
                                    nullptr, num_windows, windows, options, result);
}

int tsg_treeseq_trait_linear_model(tsg_treeseq *ts, uint64_t num_weights, const double *weights,
                                   uint64_t num_covariates, const double *covariates,
                                   uint64_t num_windows, const double *windows, uint32_t options,
                                   double *result) {
    if (num_weights < 1) return TSG_ERR_INSUFFICIENT_WEIGHTS;
    const size_t n = (size_t) ts->num_samples, K = num_weights, C = num_covariates;
    // V = W^T Z, in the reference's order (trees.c:4175-4185); the covariates
    // are taken as already standardised
    std::vector<double> V(K * C, 0.0), w(n * (K + C + 1));
    for (size_t s = 0; s < n; s++) {
        for (size_t i = 0; i < K; i++) {
            for (size_t j = 0; j < C; j++) V[i * C + j] += weights[s * K + i] * covariates[s * C + j];
        }
    }
    for (size_t s = 0; s < n; s++) {
        double *row = w.data() + s * (K + C + 1);
        for (size_t i = 0; i < K; i++) row[i] = weights[s * K + i];
        for (size_t i = 0; i < C; i++) row[K + i] = covariates[s * C + i];
        row[K + C] = 1.0;
    }
    return tsg_treeseq_general_stat(ts, K + C + 1, w.data(), K, TSG_STAT_TRAIT_LINEAR_MODEL,
                                    nullptr, V.empty() ? nullptr : V.data(), num_windows, windows,
                                    options, result);
}

int tsg_treeseq_genetic_relatedness_weighted(tsg_treeseq *ts, uint64_t num_weights,
                                             const double *weights, uint64_t num_index_tuples,
                                             const int32_t *index_tuples, uint64_t num_windows,
                                             const double *windows, double *result,
                                             uint32_t options) {
    if (num_weights == 0) return TSG_ERR_INSUFFICIENT_WEIGHTS;
    const size_t n = (size_t) ts->num_samples, K = num_weights;
    // W with a column of 1/n (trees.c:4862-4872)
    std::vector<double> tot(K + 1, 0.0), w(n * (K + 1));
    for (size_t j = 0; j < n; j++) {
        for (size_t k = 0; k < K; k++) {
            w[j * (K + 1) + k] = weights[j * K + k];
            tot[k] += weights[j * K + k];
        }
        w[j * (K + 1) + K] = 1.0 / (double) n;
    }
    tot[K] = 1.0;
    const int32_t id = (options & TSG_STAT_NONCENTRED) ? TSG_STAT_RELATEDNESS_WEIGHTED_NONCENTRED
                                                       : TSG_STAT_RELATEDNESS_WEIGHTED;
    return tsg_treeseq_general_stat(ts, K + 1, w.data(), num_index_tuples, id, index_tuples,
                                    tot.data(), num_windows, windows,
                                    options & ~TSG_STAT_NONCENTRED, result);
}

int tsg_treeseq_genotype_matrix(tsg_treeseq *ts, uint32_t options, int32_t *out) {
    Use use(ts);
    if (ts->num_sites == 0) return 0;
    int ret;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = device_upload(ts))) return ret;
    }
    if (!ts->hot.usable) return TSG_ERR_UNSUPPORTED_INPUT;
    int WB = 0;
    for (int wb : {64, 32, 16}) {
        if ((size_t) ts->hot.max_slots * wb * 4 <= (size_t) 48 * 1024) {
            WB = wb;
            break;
        }
    }
    if (WB == 0) return TSG_ERR_UNSUPPORTED_INPUT;
    std::vector<int32_t> ms, mj;
    if (!(options & TSG_ISOLATED_NOT_MISSING)) isolated_pairs(ts, ms, mj);
    Lease lease(ts);
    Workspace *ws = lease.ws;
    if (ws == nullptr) return TSG_ERR_DEVICE;
    hipStream_t s = ws->stream;
    const int ns = ts->num_samples;
    const size_t cells = (size_t) ts->num_sites * ns;
    if ((ret = ensure(ws->geno, cells * 4)) || (ret = upload(ws->wk0, ms, s)) ||
        (ret = upload(ws->wk1, mj, s))) {
        return ret;
    }
    const TsgDevice &d = ts->dev;
    const TsgHotDevice &hd = ts->hot.dev;
    TsgGenoArgs g{};
    g.tree_site0 = d.tree_site0;
    g.site_mut0 = d.site_mut0;
    g.num_samples = ns;
    g.NW = (ns + 31) / 32;
    g.chunk_tree0 = hd.chunk_tree0;
    g.chunk_node0 = hd.chunk_slot0;
    g.tr_off = hd.tr_off;
    g.tr_nrem = hd.tr_nrem;
    g.ev_child = hd.ev_child;
    g.ev_path0 = hd.ev_path0;
    g.path = hd.path;
    g.mut_node = hd.mut_slot;
    g.slot_par = hd.slot_par;
    g.chunk_lev0 = hd.chunk_lev0;
    g.lev_slot0 = hd.lev_slot0;
    g.sample_slot = hd.sample_slot;
    g.num_chunks = ts->hot.num_chunks;
    g.WB = WB;
    g.mut_allele = d.mut_allele;
    g.geno = (int32_t *) ws->geno.p;
    ws->last_store = TSG_STORE_HOT;
    if (ws->profile) (void) hipEventRecord(ws->ev0, s);
    if ((ret = hip_check(tsg_launch_genotypes(g, (size_t) ts->hot.max_slots * WB * 4,
                                              (const int32_t *) ws->wk0.p,
                                              (const int32_t *) ws->wk1.p, (int64_t) ms.size(),
                                              s)))) {
        return ret;
    }
    if (ws->profile) (void) hipEventRecord(ws->ev1, s);
    if ((ret = copy_out(ws, s, out, ws->geno.p, cells * 4))) return ret;
    if (ws->profile) {
        float ms_ = 0;
        (void) hipEventElapsedTime(&ms_, ws->ev0, ws->ev1);
        ws->sweep_ms += ms_;
        ws->sweep_launches += 1;
    }
    return 0;
}

// ---- batched form (no counterpart in the reference) ----------------------
// A batch collects validated queries on one tree sequence and runs them as
// one fused sweep per mode: the path walks are shared, the state vector is
// the concatenation of every query's sample sets and the outputs the
// concatenation of their summary blocks, each with its own windows.
struct tsg_batch {
    tsg_treeseq *ts;
    std::vector<std::unique_ptr<Query>> queries;
};

namespace {
struct Capture {
    explicit Capture(tsg_batch *b) { g_capture = &b->queries; }
    ~Capture() { g_capture = nullptr; }
};
}  // namespace

int tsg_batch_init(tsg_batch **out, tsg_treeseq *ts) {
    *out = nullptr;
    if (ts == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
    tsg_batch *b = new (std::nothrow) tsg_batch();
    if (b == nullptr) return TSG_ERR_NO_MEMORY;
    b->ts = ts;
    *out = b;
    return 0;
}

int tsg_batch_free(tsg_batch *b) {
    delete b;
    return 0;
}

int tsg_batch_clear(tsg_batch *b) {
    b->queries.clear();
    return 0;
}

uint64_t tsg_batch_get_num_queries(const tsg_batch *b) { return b->queries.size(); }

// Runs every query of the batch: one fused sweep for the branch-mode
// queries and one for the site-mode queries.  Results land in the buffers
// given when the queries were added.
namespace {
// Both modes of a batch: each mode's run on its own leased workspace and
// stream, the second mode's host work and uploads done while the first
// mode's sweep runs; the second sweep starts when the first has ended (the
// two do not share the device), then both finishes.
int run_two_modes(tsg_treeseq *ts, std::vector<std::vector<Query *>> &groups,
                  std::vector<std::unique_ptr<Lease>> &leases, std::vector<BatchRun> &runs) {
    int ret;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = device_upload(ts))) return ret;
    }
    runs.resize(groups.size());
    SeedShare share;
    for (size_t i = 0; i < groups.size(); i++) {
        leases.push_back(std::make_unique<Lease>(ts));
        Workspace *ws = leases.back()->ws;
        if (ws == nullptr) return TSG_ERR_DEVICE;
        BatchRun &r = runs[i];
        // the second run (it waits for the first sweep's end) takes the
        // first run's seeded counts when its pairs are the same
        if ((ret = batch_prepare(ts, groups[i], ws, ws->stream, r, i == 0 ? &share : nullptr,
                                 i == 1 ? &share : nullptr))) {
            return ret;
        }
        if (i > 0 && (ret = hip_check(hipStreamWaitEvent(r.s, leases[i - 1]->ws->ev1, 0)))) {
            return ret;
        }
        if (ws->profile) (void) hipEventRecord(ws->ev0, r.s);
        if ((ret = batch_sweep(r))) return ret;
        if ((ret = hip_check(hipEventRecord(ws->ev1, r.s)))) return ret;
    }
    for (BatchRun &r : runs) {
        if ((ret = batch_finish(r))) return ret;
    }
    return 0;
}

// The node-mode queries of a batch, each on its own (they are not fused).
int run_node_queries(tsg_batch *b) {
    for (auto &q : b->queries) {
        if (q->mode != TSG_STAT_NODE) continue;
        const int ret = run_node(b->ts, *q);
        if (ret != 0) return ret;
    }
    return 0;
}

// The branch and site queries of a batch: one fused sweep per mode.
int run_fused_modes(tsg_batch *b) {
    Use use(b->ts);
    std::vector<std::vector<Query *>> groups;
    for (uint32_t mode : {TSG_STAT_BRANCH, TSG_STAT_SITE}) {
        std::vector<Query *> group;
        for (auto &q : b->queries) {
            if (q->mode == mode) group.push_back(q.get());
        }
        if (!group.empty()) groups.push_back(std::move(group));
    }
    if (groups.empty()) return 0;
    if (groups.size() == 1) return run_batch(b->ts, groups[0]);
    std::vector<std::unique_ptr<Lease>> leases;
    std::vector<BatchRun> runs;
    int ret = run_two_modes(b->ts, groups, leases, runs);
    // every stream drains before its workspace returns to the pool
    for (size_t i = 0; i < leases.size(); i++) {
        Workspace *ws = leases[i]->ws;
        if (ws == nullptr) continue;
        const int sret = hip_check(hipStreamSynchronize(ws->stream));
        if (ret == 0) ret = sret;
        if (ret == 0 && ws->profile && i < runs.size()) {
            float ms = 0;
            (void) hipEventElapsedTime(&ms, ws->ev0, ws->ev1);
            ws->sweep_ms += ms;
            ws->sweep_launches += 1;
            ws->mode_ms[runs[i].mode_id] += ms;
            ws->mode_launches[runs[i].mode_id] += 1;
        }
    }
    if (ret == 0) {
        for (const BatchRun &r : runs) batch_deliver(r);
    }
    // last lease first: the pool hands the same workspace to the same mode
    // next time, so its buffers are already sized for it
    while (!leases.empty()) leases.pop_back();
    return ret;
}
}  // namespace

int tsg_batch_run(tsg_batch *b) {
    if (b == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
    Use use(b->ts);
    int ret = run_fused_modes(b);
    if (ret == 0) ret = run_node_queries(b);
    return ret;
}

// A batch run with no device or pinned allocation (tsg_batch_run_noalloc,
// tsg_noalloc.hip): every device buffer of the run is carved from
// [scratch, scratch + cap), a region of device memory that stays allocated
// for the life of the process.  The tree sequence must already be resident
// on the current device and its statistics must run on the hot plan.  The
// modes run one after the other on the null stream; the call returns when
// every result has been written.
__attribute__((visibility("hidden"))) int tsg_batch_run_fixed(tsg_batch *b, void *scratch,
                                                             size_t cap) {
    if (b == nullptr || scratch == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
    tsg_treeseq *ts = b->ts;
    Use use(ts);
    int dev = 0;
    if (hip_check(hipGetDevice(&dev))) return TSG_ERR_DEVICE;
    Workspace w;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        const Resident *res = res_of(ts);
        if (res == nullptr || ts->dev.device != dev || !ts->hot.usable) {
            return TSG_ERR_NOT_RESIDENT;
        }
        w.lds_limit = res->lds_limit;
        w.force_c32 = res->force_c32;
        w.force_full = res->force_full;
    }
    w.fixed = (char *) scratch;
    w.fixed_cap = cap;
    w.stream = nullptr;
    // node statistics allocate their per-node buffers: not in this form
    for (auto &q : b->queries) {
        if (q->mode == TSG_STAT_NODE) return TSG_ERR_UNSUPPORTED_STAT_MODE;
    }
    int ret = 0;
    for (uint32_t mode : {TSG_STAT_BRANCH, TSG_STAT_SITE}) {
        std::vector<Query *> group;
        for (auto &q : b->queries) {
            if (q->mode == mode) group.push_back(q.get());
        }
        if (group.empty()) continue;
        w.fixed_used = 0;
        BatchRun r;
        ret = batch_prepare(ts, group, &w, w.stream, r);
        if (ret == 0) ret = batch_sweep(r);
        if (ret == 0) ret = batch_finish(r);
        const int sret = hip_check(hipStreamSynchronize(w.stream));
        if (ret == 0) ret = sret;
        if (ret != 0) break;
    }
    return ret;
}

namespace {
// The runs of tsg_batch_run_multi: prepared in parallel, each on its
// workspace's stream; runs on the hot plan whose outputs fit one slice share
// one sweep launch per (mode, node store, outputs per lane) on the first
// run's stream, the rest launch on their own there; then each run finishes
// on its own stream.
int run_multi(tsg_batch **bs, uint64_t n, std::vector<std::unique_ptr<Lease>> &leases,
              std::deque<std::vector<Query *>> &groups, std::deque<BatchRun> &runs) {
    int ret;
    std::vector<tsg_treeseq *> run_ts;
    std::vector<Workspace *> run_ws;
    for (uint64_t i = 0; i < n; i++) {
        tsg_batch *b = bs[i];
        if (b == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
        {
            std::lock_guard<std::mutex> lock(g_mutex);
            if ((ret = device_upload(b->ts))) return ret;
        }
        for (uint32_t mode : {TSG_STAT_BRANCH, TSG_STAT_SITE}) {
            std::vector<Query *> g;
            for (auto &q : b->queries) {
                if (q->mode == mode) g.push_back(q.get());
            }
            if (g.empty()) continue;
            leases.push_back(std::make_unique<Lease>(b->ts));
            Workspace *ws = leases.back()->ws;
            if (ws == nullptr) return TSG_ERR_DEVICE;
            groups.push_back(std::move(g));
            runs.emplace_back();
            run_ts.push_back(b->ts);
            run_ws.push_back(ws);
        }
    }
    const size_t R = runs.size();
    if (R == 0) return 0;
    // host work and uploads of the runs on up to 8 host threads, each run on
    // its workspace's own stream; the shared launches wait for all of them
    std::vector<int> rets(R, 0);
    {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i; (i = next.fetch_add(1)) < R;) {
                Workspace *ws = run_ws[i];
                rets[i] = batch_prepare(run_ts[i], groups[i], ws, ws->stream, runs[i]);
                if (rets[i] == 0) rets[i] = hip_check(hipEventRecord(ws->ev0, ws->stream));
            }
        };
        std::vector<std::thread> pool;
        const size_t T = std::min<size_t>(R, 8);
        for (size_t t = 1; t < T; t++) pool.emplace_back(work);
        work();
        for (std::thread &th : pool) th.join();
    }
    for (size_t i = 0; i < R; i++) {
        if (rets[i] != 0) return rets[i];
    }
    hipStream_t s0 = run_ws[0]->stream;
    for (size_t i = 1; i < R; i++) {
        if ((ret = hip_check(hipStreamWaitEvent(s0, run_ws[i]->ev0, 0)))) return ret;
    }
    for (BatchRun &r : runs) r.s = s0;
    std::map<std::tuple<int, int, int>, std::vector<BatchRun *>> shared;
    for (BatchRun &r : runs) {
        const bool hot = r.store == TSG_STORE_HOT || r.store == TSG_STORE_HOT16;
        if (!hot || r.M > r.slice || r.M <= 0 || r.C <= 0) {
            if ((ret = batch_sweep(r))) return ret;
            continue;
        }
        const int mpl = r.M <= 64 ? 1 : r.M <= 128 ? 2 : 4;
        shared[{r.mode_id, r.store, mpl}].push_back(&r);
    }
    if (!shared.empty()) {
        std::vector<TsgSweepArgs> args;
        std::vector<int32_t> cpre;
        for (auto &kv : shared) {
            int64_t total = 0;
            cpre.push_back(0);
            for (BatchRun *r : kv.second) {
                args.push_back(r->a);
                args.back().m_begin = 0;
                args.back().m_end = r->M;
                total += r->C;
                if (total > INT32_MAX / 2) return TSG_ERR_BAD_PARAM_VALUE;
                cpre.push_back((int32_t) total);
            }
        }
        Workspace *ws0 = runs.front().ws;
        if ((ret = upload(ws0->margs, args, s0)) || (ret = upload(ws0->mcpre, cpre, s0))) {
            return ret;
        }
        size_t aoff = 0, coff = 0;
        for (auto &kv : shared) {
            const int nrun = (int) kv.second.size();
            size_t shmem = 0;
            for (BatchRun *r : kv.second) shmem = std::max(shmem, r->shmem);
            const int32_t total = cpre[coff + nrun];
            const int mode_id = std::get<0>(kv.first), store = std::get<1>(kv.first);
            const int mpl = std::get<2>(kv.first);
            if ((ret = hip_check(tsg_launch_sweep_multi(
                     (const TsgSweepArgs *) ws0->margs.p + aoff,
                     (const int32_t *) ws0->mcpre.p + coff, nrun, total, mode_id, store, mpl,
                     shmem, s0)))) {
                return ret;
            }
            aoff += nrun;
            coff += nrun + 1;
        }
    }
    // the finish phases of every run (first-tree sums, chunk scan, window
    // combine, finals), one launch per phase over all runs, after the
    // sweeps; every run's results into one device buffer, copied back in one
    // transfer
    Workspace *ws0 = run_ws[0];
    std::vector<TsgFinishRun> fr(R);
    std::vector<int64_t> pre((size_t) TSG_FINISH_PHASES * (R + 1), 0), totals(TSG_FINISH_PHASES);
    std::vector<int64_t> outoff(R + 1, 0);
    for (size_t i = 0; i < R; i++) outoff[i + 1] = outoff[i] + std::max<int64_t>(runs[i].outcount, 0);
    if ((ret = ensure(ws0->mout, (size_t) std::max<int64_t>(outoff[R], 1) * 8))) return ret;
    for (size_t i = 0; i < R; i++) {
        const BatchRun &r = runs[i];
        Workspace *ws = run_ws[i];
        TsgFinishRun &f = fr[i];
        const bool branch = r.mode == TSG_STAT_BRANCH;
        f.a = r.a;
        f.ca = r.ca;
        f.t0_parent = r.t0_parent;
        f.t0_bl = r.t0_bl;
        f.x0 = (int32_t *) ws->x0.p;
        f.r0part = (double *) ws->r0part.p;
        f.bsum = (double *) ws->bsum.p;
        f.Rstart = (double *) ws->Rstart.p;
        f.dR = branch ? r.a.dR : nullptr;
        f.nloc = r.nloc0;
        f.nb = r.nb0;
        f.nblk = tsg_scan_blocks(r.C);
        f.lanecount = r.lanecount;
        f.fin = r.fin_dev;
        f.nq = r.Q;
        f.raw = (const double *) ws->raw.p;
        f.ipar = r.ipar_dev;
        f.dpar = r.dpar_dev;
        f.out = (double *) ws0->mout.p + outoff[i];
        f.outcount = r.outcount;
        int64_t blocks[TSG_FINISH_PHASES];
        tsg_finish_blocks(f, blocks);
        for (int ph = 0; ph < TSG_FINISH_PHASES; ph++) {
            int64_t *p = pre.data() + (size_t) ph * (R + 1);
            p[i + 1] = p[i] + blocks[ph];
        }
    }
    for (int ph = 0; ph < TSG_FINISH_PHASES; ph++) {
        totals[ph] = pre[(size_t) ph * (R + 1) + R];
    }
    const size_t bytes = (size_t) std::max<int64_t>(outoff[R], 1) * 8;
    if (ws0->mout_h_bytes < bytes) {
        if (ws0->mout_h != nullptr) (void) hipHostFree(ws0->mout_h);
        ws0->mout_h = nullptr;
        ws0->mout_h_bytes = 0;
        const size_t want = std::max(bytes + bytes / 2, (size_t) 1 << 16);
        if (hipHostMalloc(&ws0->mout_h, want) != hipSuccess) {
            (void) hipGetLastError();
            return TSG_ERR_NO_MEMORY;
        }
        ws0->mout_h_bytes = want;
    }
    if ((ret = upload(ws0->mfin, fr, s0)) || (ret = upload(ws0->mpre, pre, s0)) ||
        (ret = hip_check(tsg_launch_finish_multi((const TsgFinishRun *) ws0->mfin.p,
                                                 (const int64_t *) ws0->mpre.p, totals.data(),
                                                 (int) R, s0))) ||
        (ret = hip_check(hipMemcpyAsync(ws0->mout_h, ws0->mout.p, (size_t) outoff[R] * 8,
                                        hipMemcpyDeviceToHost, s0)))) {
        return ret;
    }
    for (size_t i = 0; i < R; i++) {
        runs[i].s = s0;
        runs[i].results_h = (const double *) ws0->mout_h + outoff[i];
        if ((ret = copy_device_results(runs[i], (const double *) ws0->mout.p + outoff[i], s0))) {
            return ret;
        }
    }
    return 0;
}
}  // namespace

int tsg_batch_run_multi(tsg_batch **batches, uint64_t n) {
    if (n > 0 && batches == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
    // every tree sequence of the call stays in use until its leases end
    std::vector<std::unique_ptr<Use>> uses;
    for (uint64_t i = 0; i < n; i++) {
        if (batches[i] == nullptr) return TSG_ERR_BAD_PARAM_VALUE;
        uses.push_back(std::make_unique<Use>(batches[i]->ts));
    }
    std::vector<std::unique_ptr<Lease>> leases;
    std::deque<std::vector<Query *>> groups;
    std::deque<BatchRun> runs;
    int ret = run_multi(batches, n, leases, groups, runs);
    // the leases end (workspaces return to their pools) only once all work
    // queued on their streams has finished
    for (auto &l : leases) {
        if (l->ws == nullptr) continue;
        const int sret = hip_check(hipStreamSynchronize(l->ws->stream));
        if (ret == 0) ret = sret;
    }
    if (ret == 0) {
        for (const BatchRun &r : runs) batch_deliver(r);
    }
    for (uint64_t i = 0; i < n && ret == 0; i++) ret = run_node_queries(batches[i]);
    return ret;
}

#define BATCH_ONE_WAY(name, fn)                                                               \
    int name(tsg_batch *b, uint64_t K, const uint64_t *sizes, const int32_t *sets,            \
             uint64_t nw, const double *windows, uint32_t options, double *result) {           \
        Capture cap(b);                                                                       \
        return fn(b->ts, K, sizes, sets, nw, windows, options, result);                       \
    }
#define BATCH_K_WAY(name, fn)                                                                 \
    int name(tsg_batch *b, uint64_t K, const uint64_t *sizes, const int32_t *sets,            \
             uint64_t M, const int32_t *idx, uint64_t nw, const double *windows,               \
             uint32_t options, double *result) {                                              \
        Capture cap(b);                                                                       \
        return fn(b->ts, K, sizes, sets, M, idx, nw, windows, options, result);                \
    }
BATCH_ONE_WAY(tsg_batch_diversity, tsg_treeseq_diversity)
BATCH_ONE_WAY(tsg_batch_segregating_sites, tsg_treeseq_segregating_sites)
BATCH_ONE_WAY(tsg_batch_Y1, tsg_treeseq_Y1)
BATCH_ONE_WAY(tsg_batch_Tajimas_D, tsg_treeseq_Tajimas_D)
BATCH_K_WAY(tsg_batch_divergence, tsg_treeseq_divergence)
BATCH_K_WAY(tsg_batch_Y2, tsg_treeseq_Y2)
BATCH_K_WAY(tsg_batch_f2, tsg_treeseq_f2)
BATCH_K_WAY(tsg_batch_Y3, tsg_treeseq_Y3)
BATCH_K_WAY(tsg_batch_f3, tsg_treeseq_f3)
BATCH_K_WAY(tsg_batch_f4, tsg_treeseq_f4)
BATCH_K_WAY(tsg_batch_genetic_relatedness, tsg_treeseq_genetic_relatedness)
BATCH_K_WAY(tsg_batch_genetic_relatedness_proportion, tsg_treeseq_genetic_relatedness_proportion)
BATCH_K_WAY(tsg_batch_Fst, tsg_treeseq_Fst)
#undef BATCH_ONE_WAY
#undef BATCH_K_WAY

// ---- allele frequency spectrum (trees.c:3815-3928) -----------------------
#define TSG_AFS_MAX_SETS 16

int tsg_treeseq_allele_frequency_spectrum(tsg_treeseq *ts, uint64_t K, const uint64_t *sizes,
                                          const int32_t *sets, uint64_t nw,
                                          const double *windows, uint64_t ntw,
                                          const double *time_windows, uint32_t options,
                                          double *result) {
    Use use(ts);
    bool site = options & TSG_STAT_SITE, branch = options & TSG_STAT_BRANCH;
    if (options & TSG_STAT_NODE) return TSG_ERR_UNSUPPORTED_STAT_MODE;
    if (!(site || branch)) site = true;
    if (site && branch) return TSG_ERR_MULTIPLE_STAT_MODES;
    std::vector<double> win;
    if (windows == nullptr) {
        win = {0.0, ts->L};
    } else {
        int r = check_windows(ts, nw, windows);
        if (r != 0) return r;
        win.assign(windows, windows + nw + 1);
    }
    std::vector<double> tw;
    if (time_windows == nullptr) {
        tw = {0.0, INFINITY};
    } else {
        if (ntw < 1) return TSG_ERR_BAD_TIME_WINDOWS_DIM;
        if (time_windows[0] != 0.0) return TSG_ERR_BAD_TIME_WINDOWS;
        for (uint64_t j = 0; j < ntw; j++) {
            if (time_windows[j] >= time_windows[j + 1]) return TSG_ERR_BAD_TIME_WINDOWS;
        }
        if (site && !(time_windows[0] == 0.0 && std::isinf((float) time_windows[1]))) {
            return TSG_ERR_UNSUPPORTED_STAT_MODE;
        }
        tw.assign(time_windows, time_windows + ntw + 1);
    }
    std::vector<std::vector<int32_t>> ss;
    int ret = parse_sample_sets(ts, K, sizes, sets, ss);
    if (ret != 0) return ret;
    if (branch && ts->time_uncalibrated && !(options & TSG_STAT_ALLOW_TIME_UNCALIBRATED)) {
        return TSG_ERR_TIME_UNCALIBRATED;
    }
    if (K > TSG_AFS_MAX_SETS) return TSG_ERR_UNSUPPORTED_INPUT;
    const int W = (int) win.size() - 1, TW = (int) tw.size() - 1;
    std::vector<int64_t> dims(K);
    double afs_size_d = 1;
    for (uint64_t k = 0; k < K; k++) {
        dims[k] = (int64_t) sizes[k] + 1;
        afs_size_d *= (double) dims[k];
    }
    if (afs_size_d * W * TW > 4e9) return TSG_ERR_UNSUPPORTED_INPUT;
    const int64_t afs_size = (int64_t) afs_size_d;

    bool use_hot;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = device_upload(ts))) return ret;
    }
    Lease lease(ts);
    Workspace *ws = lease.ws;
    if (ws == nullptr) return TSG_ERR_DEVICE;
    hipStream_t s = ws->stream;
    const int KK = (int) K + 1;
    // The hot plan when its slots fit in LDS: per slot KK counts and, for
    // branch lengths, four doubles (node time, parent time, last flush,
    // insertion); then one staging row.
    const size_t hot_need = (size_t) ts->hot.max_slots * ((size_t) KK * 4 + (branch ? 32 : 0)) +
                            (size_t) KK * 4 + 64;
    use_hot = ts->hot.usable && !ws->force_full && hot_need <= (size_t) ws->lds_limit;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if ((ret = use_hot ? ensure_hot_afs(ts) : ensure_full(ts))) return ret;
    }
    const int C = use_hot ? ts->hot.num_chunks : ts->num_chunks;
    std::vector<int32_t> pair_s, pair_c, col_total;
    for (uint64_t k = 0; k < K; k++) {
        for (int32_t si : ss[k]) {
            pair_s.push_back(si);
            pair_c.push_back((int32_t) k);
        }
        col_total.push_back((int32_t) ss[k].size());
    }
    if (!use_hot) {  // the hot plan seeds the all-samples column from slot_nsd
        for (int32_t si = 0; si < ts->num_samples; si++) {
            pair_s.push_back(si);
            pair_c.push_back((int32_t) K);
        }
    }
    col_total.push_back(ts->num_samples);
    const int64_t hist_n = (int64_t) W * TW * afs_size;
    TsgSweepArgs a{};
    if ((ret = upload_seeds(ws, pair_s, pair_c, ts->num_samples, s, a)) ||
        (ret = upload(ws->tot, col_total, s)) || (ret = upload(ws->windows, win, s)) ||
        (ret = upload(ws->twin, tw, s)) || (ret = upload(ws->dims, dims, s)) ||
        (ret = ensure(ws->hist, (size_t) hist_n * 8)) ||
        (ret = ensure(ws->out, (size_t) hist_n * 8))) {
        return ret;
    }
    const TsgDevice &d = ts->dev;
    const TsgHotDevice &h = ts->hot.dev;
    const int R = TSG_NODE_WORDS(KK);
    const size_t stage = (size_t) 2 * KK * 4 + 64;
    size_t lds_need = (size_t) ts->max_nloc * R * 4 + stage;
    if (branch) lds_need += (size_t) ts->max_nloc * 8;
    const bool lds = use_hot || lds_need <= (size_t) ws->lds_limit;
    if (use_hot) {
        a.chunk_tree0 = h.chunk_tree0;
        a.chunk_node0 = h.chunk_slot0;
        a.tr_off = h.tr_off;
        a.tr_nrem = h.tr_nrem;
        a.ev_child = h.ev_child;
        a.ev_path0 = h.ev_path0;
        a.path = h.path;
        a.mut_node = h.mut_slot;
        a.slot_par = h.slot_par;
        a.chunk_lev0 = h.chunk_lev0;
        a.lev_slot0 = h.lev_slot0;
        a.sample_slot = h.sample_slot;
        a.slot_nsd = h.slot_nsd;
        a.slot_tu = h.slot_tu;
        a.slot_tp0 = h.slot_tp0;
        a.slot_last0 = h.slot_last0;
        a.slot_ins0 = h.slot_ins0;
        ws->last_store = TSG_STORE_HOT;
    } else {
        std::vector<int32_t> cq_wfirst(C), cq_nwin(C);
        std::vector<int64_t> cq_poff(C, 0);
        for (int c = 0; c < C; c++) {
            const double cs = ts->breakpoints[ts->chunk_tree0[c]];
            const double ce = ts->breakpoints[ts->chunk_tree0[c + 1]];
            const int wf = (int) (std::upper_bound(win.begin(), win.end(), cs) - win.begin()) - 1;
            const int wl = (int) (std::lower_bound(win.begin(), win.end(), ce) - win.begin()) - 1;
            cq_wfirst[c] = wf;
            cq_nwin[c] = wl - wf + 1;
        }
        if ((ret = upload(ws->cq_wfirst, cq_wfirst, s)) ||
            (ret = upload(ws->cq_nwin, cq_nwin, s)) || (ret = upload(ws->cq_poff, cq_poff, s))) {
            return ret;
        }
        if (!lds) {
            if ((ret = ensure(ws->slab, (size_t) C * ts->max_nloc * R * 4))) return ret;
            if (branch && (ret = ensure(ws->last, (size_t) C * ts->max_nloc * 8))) return ret;
        }
        a.chunk_tree0 = d.chunk_tree0;
        a.chunk_node0 = d.chunk_node0;
        a.loc_parent0 = d.loc_parent0;
        a.loc_bl0 = d.loc_bl0;
        a.loc_time = d.loc_time;
        a.loc_lastrem = d.loc_lastrem;
        a.tr_off = d.tr_off;
        a.tr_nrem = d.tr_nrem;
        a.ev_child = d.ev_child;
        a.ev_parent = d.ev_parent;
        a.ev_bl = d.ev_bl;
        a.ev_path0 = d.ev_path0;
        a.path = d.path;
        a.mut_node = d.mut_node;
        a.cq_wfirst = (const int32_t *) ws->cq_wfirst.p;
        a.cq_nwin = (const int32_t *) ws->cq_nwin.p;
        a.cq_poff = (const int64_t *) ws->cq_poff.p;
        a.g_rec = (int32_t *) ws->slab.p;
        a.g_stride = (int64_t) ts->max_nloc * R;
        a.g_last = (double *) ws->last.p;
        a.g_last_stride = ts->max_nloc;
        ws->last_store = lds ? TSG_STORE_LDS : TSG_STORE_GLOBAL;
    }
    a.bp = d.breakpoints;
    a.tree_site0 = d.tree_site0;
    a.site_pos = d.site_pos;
    a.site_mut0 = d.site_mut0;
    a.mut_allele = d.mut_allele;
    a.mut_pallele = d.mut_pallele;
    a.site_nalleles = d.site_nalleles;
    a.num_trees = ts->num_trees;
    a.num_chunks = C;
    a.num_samples = ts->num_samples;
    a.K = KK;
    a.Q = 1;
    a.col_total = (const int32_t *) ws->tot.p;
    a.windows = (const double *) ws->windows.p;
    a.flags = (options & TSG_STAT_POLARISED) ? TSG_F_POLARISED : 0u;
    a.afs_K = (int32_t) K;
    a.afs_size = afs_size;
    a.afs_dims = (const int64_t *) ws->dims.p;
    a.TW = TW;
    a.time_windows = (const double *) ws->twin.p;
    a.afs_hist_u = (unsigned long long *) ws->hist.p;
    a.afs_hist_d = (double *) ws->hist.p;
    a.afs_W = W;
    const uint32_t fflags = (options & TSG_STAT_SPAN_NORMALISE) ? TSG_F_SPAN_NORMALISE : 0u;
    if (branch) {
        if ((ret = hip_check(tsg_launch_fill_f64((double *) ws->hist.p, hist_n, s)))) return ret;
    } else {
        if ((ret = hip_check(tsg_launch_fill_u64((unsigned long long *) ws->hist.p, hist_n, s)))) {
            return ret;
        }
    }
    if (ws->profile) (void) hipEventRecord(ws->ev0, s);
    if (use_hot) {
        if ((ret = hip_check(tsg_launch_afs_hot(a, branch, hot_need, s)))) return ret;
        if (branch) {
            const tsg_hotplan &hp = ts->hot;
            const int nl = (int) hp.tl_parent.size();
            if ((ret = ensure(ws->x0, (size_t) nl * KK * 4 + 4))) return ret;
            TsgLastTree lt{};
            lt.parent = h.tl_parent;
            lt.time = h.tl_time;
            lt.last = h.tl_last;
            lt.ins = h.tl_ins;
            lt.nl = nl;
            lt.counts = (int32_t *) ws->x0.p;
            if ((ret = hip_check(tsg_launch_afs_last_tree(a, lt, ts->L, s)))) return ret;
        }
    } else {
        if ((ret = hip_check(tsg_launch_afs(a, branch, lds, lds ? lds_need : stage, s)))) {
            return ret;
        }
    }
    if (ws->profile) (void) hipEventRecord(ws->ev1, s);
    double *res_dev;
    if (branch) {
        if ((ret = hip_check(tsg_launch_afs_branch_final((double *) ws->hist.p, a.windows, W,
                                                         (int64_t) TW * afs_size, fflags, s)))) {
            return ret;
        }
        res_dev = (double *) ws->hist.p;
    } else {
        if ((ret = hip_check(tsg_launch_afs_site_final((const unsigned long long *) ws->hist.p,
                                                       (double *) ws->out.p, a.windows, W,
                                                       afs_size, fflags, s)))) {
            return ret;
        }
        res_dev = (double *) ws->out.p;
    }
    if ((ret = copy_out(ws, s, result, res_dev, (size_t) hist_n * 8))) return ret;
    if (ws->profile) {
        float ms = 0;
        (void) hipEventElapsedTime(&ms, ws->ev0, ws->ev1);
        ws->sweep_ms += ms;
        ws->sweep_launches += 1;
    }
    return 0;
}

}  // extern "C"
