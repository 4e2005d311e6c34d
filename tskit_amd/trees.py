"""Device-resident tree sequences and the general_stat-family statistics.

Two layers, as in the reference:

* ``LowLevelTreeSequence`` mirrors the statistics methods of the reference's
  CPython extension (python/_tskitmodule.c:6588-7194): same keyword lists,
  same argument parsing and the same ValueError messages, forwarding to the
  C ABI of ``libtsgpu.so`` (include/tsgpu.h).  C errors raise
  ``LibraryError`` with the reference's error number.
* ``TreeSequence`` restates the Python wrappers (python/tskit/trees.py
  :7917-10428): sample-set and index parsing, ``"trees"``/``"sites"`` window
  shortcuts, dimension dropping, and the composites Fst, Tajima's D and
  relatedness ``proportion`` (which run fused in one device sweep with their
  FP64 finals on the device).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import as_c, ptr
from .exceptions import LibraryError, check

STAT_SITE = 1 << 0
STAT_BRANCH = 1 << 1
STAT_NODE = 1 << 2
STAT_POLARISED = 1 << 10
STAT_SPAN_NORMALISE = 1 << 11
STAT_ALLOW_TIME_UNCALIBRATED = 1 << 12
STAT_NONCENTRED = 1 << 14
TS_TIME_UNCALIBRATED = 1 << 0


def _parse_stats_mode(mode):
    """_tskitmodule.c:6588-6607"""
    if mode is None or mode == "site":
        return STAT_SITE
    if mode == "branch":
        return STAT_BRANCH
    if mode == "node":
        return STAT_NODE
    raise ValueError("Unrecognised stats mode")


def _parse_sample_sets(sample_set_sizes, sample_sets):
    """_tskitmodule.c:799-848"""
    sizes = np.asarray(sample_set_sizes)
    if sizes.ndim != 1:
        raise ValueError("sample_set_sizes must be one-dimensional")
    sizes = as_c(sizes, np.uint64)
    sets = np.asarray(sample_sets)
    if sets.ndim != 1:
        raise ValueError("sample_sets must be one-dimensional")
    sets = as_c(sets, np.int32)
    total = 0
    for s in sizes.tolist():
        total += s
    if total != len(sets):
        raise ValueError("Sum of sample_set_sizes must equal length of sample_sets array")
    return sizes, sets


def _parse_windows(windows):
    """_tskitmodule.c:6609-6636"""
    w = np.asarray(windows, dtype=np.float64)
    if w.ndim != 1:
        raise ValueError("Windows must be a one-dimensional array")
    if w.shape[0] < 2:
        raise ValueError("Windows arrays must have at least 2 elements")
    return as_c(w, np.float64)


def _parse_indexes(indexes, k):
    idx = np.asarray(indexes)
    if idx.ndim != 2:
        raise ValueError("indexes must be two-dimensional")
    idx = as_c(idx, np.int32)
    if idx.shape[0] < 1 or idx.shape[1] != k:
        raise ValueError(f"indexes must be a k x {k} array.")
    return idx


def safe_np_int_cast(int_array, dtype):
    """Casts with bounds checks (restates python/tskit/util.py:98-130)."""
    if not isinstance(int_array, np.ndarray):
        int_array = np.array(int_array)
    if int_array.size == 0:
        return int_array.astype(dtype)
    src = int_array.dtype
    if src.kind in "iu" and not np.can_cast(src, dtype, casting="safe"):
        # integers of a wider type: the bounds check without the failed safe
        # cast's exception (same result, same errors)
        bounds = np.iinfo(dtype)
        if int_array.min() < bounds.min or int_array.max() > bounds.max:
            raise OverflowError(f"Cannot convert safely to {dtype} type")
        return int_array.astype(dtype, casting="unsafe")
    try:
        return int_array.astype(dtype, casting="safe")
    except TypeError:
        if int_array.dtype == np.dtype("O"):
            raise TypeError("Cannot convert to a rectangular array.")
        bounds = np.iinfo(dtype)
        if np.any(int_array < bounds.min) or np.any(int_array > bounds.max):
            raise OverflowError(f"Cannot convert safely to {dtype} type")
        return int_array.astype(dtype, casting="unsafe")


class LowLevelTreeSequence:
    """The statistics surface of the reference's ``_tskit.TreeSequence``."""

    def __init__(self, tables):
        L = _lib.lib()
        t = tables
        handle = ctypes.c_void_p()
        opts = TS_TIME_UNCALIBRATED if t.time_units == "uncalibrated" else 0
        self._keep = [
            as_c(t.node_flags, np.uint32), as_c(t.node_time, np.float64),
            as_c(t.edge_left, np.float64), as_c(t.edge_right, np.float64),
            as_c(t.edge_parent, np.int32), as_c(t.edge_child, np.int32),
            as_c(t.site_position, np.float64), as_c(t.ancestral_state, np.int8),
            as_c(t.ancestral_state_offset, np.uint64), as_c(t.mutation_site, np.int32),
            as_c(t.mutation_node, np.int32), as_c(t.mutation_parent, np.int32),
            as_c(t.derived_state, np.int8), as_c(t.derived_state_offset, np.uint64),
        ]
        k = self._keep
        err = L.tsg_treeseq_init(
            ctypes.byref(handle), t.sequence_length,
            len(k[1]), ptr(k[0], ctypes.c_uint32), ptr(k[1], ctypes.c_double),
            len(k[2]), ptr(k[2], ctypes.c_double), ptr(k[3], ctypes.c_double),
            ptr(k[4], ctypes.c_int32), ptr(k[5], ctypes.c_int32),
            len(k[6]), ptr(k[6], ctypes.c_double),
            k[7].ctypes.data_as(ctypes.c_char_p), ptr(k[8], ctypes.c_uint64),
            len(k[9]), ptr(k[9], ctypes.c_int32), ptr(k[10], ctypes.c_int32),
            ptr(k[11], ctypes.c_int32), k[12].ctypes.data_as(ctypes.c_char_p),
            ptr(k[13], ctypes.c_uint64), None, None, opts,
        )
        check(err)
        self._keep = None
        self._h = handle
        self._lib = L

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tsg_treeseq_free(h)
            self._h = None

    # -- accessors -------------------------------------------------------------
    def get_num_trees(self):
        return int(self._lib.tsg_treeseq_get_num_trees(self._h))

    def get_num_samples(self):
        return int(self._lib.tsg_treeseq_get_num_samples(self._h))

    def get_num_nodes(self):
        return int(self._lib.tsg_treeseq_get_num_nodes(self._h))

    def get_num_edges(self):
        return int(self._lib.tsg_treeseq_get_num_edges(self._h))

    def get_num_sites(self):
        return int(self._lib.tsg_treeseq_get_num_sites(self._h))

    def get_num_mutations(self):
        return int(self._lib.tsg_treeseq_get_num_mutations(self._h))

    def get_num_chunks(self):
        return int(self._lib.tsg_treeseq_get_num_chunks(self._h))

    def get_max_chunk_nodes(self):
        return int(self._lib.tsg_treeseq_get_max_chunk_nodes(self._h))

    def get_device_bytes(self):
        return int(self._lib.tsg_treeseq_get_device_bytes(self._h))

    def get_sequence_length(self):
        return float(self._lib.tsg_treeseq_get_sequence_length(self._h))

    def get_breakpoints(self):
        out = np.zeros(self.get_num_trees() + 1, dtype=np.float64)
        self._lib.tsg_treeseq_get_breakpoints(self._h, ptr(out, ctypes.c_double))
        return out

    def get_samples(self):
        out = np.zeros(self.get_num_samples(), dtype=np.int32)
        self._lib.tsg_treeseq_get_samples(self._h, ptr(out, ctypes.c_int32))
        return out

    def get_index(self):
        E = self.get_num_edges()
        I = np.zeros(E, dtype=np.int32)
        O = np.zeros(E, dtype=np.int32)
        self._lib.tsg_treeseq_get_index(self._h, ptr(I, ctypes.c_int32), ptr(O, ctypes.c_int32))
        return I, O

    def get_plan(self):
        """Host copy of the chunk plan (the per-chunk sweep programs)."""
        sizes = (ctypes.c_uint64 * 6)()
        self._lib.tsg_treeseq_get_plan_sizes(self._h, sizes)
        C, nl, ne, T, S, Mu = (int(x) for x in sizes)
        p = {
            "chunk_tree0": np.zeros(C + 1, np.int32), "chunk_node0": np.zeros(C + 1, np.int64),
            "loc_parent0": np.zeros(nl, np.int32), "loc_bl0": np.zeros(nl, np.float64),
            "loc_time": np.zeros(nl, np.float64), "loc_lastrem": np.zeros(nl, np.float64),
            "tr_off": np.zeros(T + 1, np.int64), "tr_nrem": np.zeros(T, np.int32),
            "ev_child": np.zeros(ne, np.int32), "ev_parent": np.zeros(ne, np.int32),
            "ev_bl": np.zeros(ne, np.float64), "tree_site0": np.zeros(T + 1, np.int32),
            "site_mut0": np.zeros(S + 1, np.int32), "mut_node": np.zeros(Mu, np.int32),
            "mut_allele": np.zeros(Mu, np.int32), "mut_pallele": np.zeros(Mu, np.int32),
            "site_nalleles": np.zeros(S, np.int32),
        }
        self._lib.tsg_treeseq_copy_plan(self._h, *[a.ctypes.data_as(ctypes.c_void_p)
                                                   for a in p.values()])
        p["ev_path0"] = np.zeros(ne + 1, np.int64)
        p["path"] = np.zeros(int(self._lib.tsg_treeseq_get_path_size(self._h)), np.int32)
        self._lib.tsg_treeseq_copy_paths(self._h, p["ev_path0"].ctypes.data_as(ctypes.c_void_p),
                                         p["path"].ctypes.data_as(ctypes.c_void_p))
        return p

    def get_hot_plan_sizes(self):
        """Sizes of the hot plan (no copies)."""
        sizes = (ctypes.c_uint64 * 10)()
        self._lib.tsg_treeseq_get_hot_plan_sizes(self._h, sizes)
        keys = ["usable", "chunks", "slots", "level_entries", "events", "path_entries",
                "first_tree_nodes", "max_slots", "max_levels", "last_tree_nodes"]
        return dict(zip(keys, (int(x) for x in sizes)))

    def get_hot_plan(self):
        """Host copy of the hot plan (chunks keep only the nodes they touch),
        or None when the tree sequence has a chunk it cannot represent."""
        sizes = (ctypes.c_uint64 * 10)()
        self._lib.tsg_treeseq_get_hot_plan_sizes(self._h, sizes)
        usable, C, ns, nlev, ne, npath, n0, max_slots, max_levels, nl = (int(x) for x in sizes)
        if not usable:
            return None
        T = self.get_num_trees()
        Mu = self.get_num_mutations()
        n = self.get_num_samples()
        p = {
            "chunk_tree0": np.zeros(C + 1, np.int32), "chunk_slot0": np.zeros(C + 1, np.int64),
            "slot_par": np.zeros(ns, np.int32), "slot_bl0": np.zeros(ns, np.float64),
            "chunk_lev0": np.zeros(C + 1, np.int64), "lev_slot0": np.zeros(nlev, np.int32),
            "sample_slot": np.zeros(C * n, np.uint16), "tr_off": np.zeros(T + 1, np.int64),
            "tr_nrem": np.zeros(T, np.int32), "ev_child": np.zeros(ne, np.int32),
            "ev_bl": np.zeros(ne, np.float64), "ev_path0": np.zeros(ne + 1, np.int64),
            "path": np.zeros(npath, np.int32), "mut_slot": np.zeros(Mu, np.int32),
            "t0_parent": np.zeros(n0, np.int32), "t0_bl": np.zeros(n0, np.float64),
        }
        check(self._lib.tsg_treeseq_copy_hot_plan(
            self._h, *[a.ctypes.data_as(ctypes.c_void_p) for a in p.values()]))
        afs = {
            "slot_node": np.zeros(ns, np.int32), "slot_nsd": np.zeros(ns, np.int32),
            "slot_last0": np.zeros(ns, np.float64), "slot_ins0": np.zeros(ns, np.float64),
            "slot_tp0": np.zeros(ns, np.float64),
            "tl_parent": np.zeros(nl, np.int32), "tl_time": np.zeros(nl, np.float64),
            "tl_last": np.zeros(nl, np.float64), "tl_ins": np.zeros(nl, np.float64),
        }
        check(self._lib.tsg_treeseq_copy_hot_afs_plan(
            self._h, *[a.ctypes.data_as(ctypes.c_void_p) for a in afs.values()]))
        p.update(afs)
        p["max_slots"] = max_slots
        p["max_levels"] = max_levels
        return p

    def set_sweep_plan(self, plan):
        """1: run statistics sweeps on the full plan; 2: on the hot plan with
        32-bit counts; 0: the library chooses."""
        check(self._lib.tsg_treeseq_set_sweep_plan(self._h, int(plan)))

    def get_last_store(self):
        """Node store of the last sweep: 0 HBM slabs, 1 LDS full plan, 2 LDS hot
        slots, 3 LDS hot slots with 16-bit counts."""
        return int(self._lib.tsg_treeseq_get_last_store(self._h))

    def set_num_chunks(self, num_chunks):
        check(self._lib.tsg_treeseq_set_num_chunks(self._h, int(num_chunks)))

    def upload(self):
        check(self._lib.tsg_treeseq_upload(self._h))

    def set_profiling(self, enable):
        check(self._lib.tsg_treeseq_set_profiling(self._h, int(enable)))

    def get_profile_modes(self):
        """{mode: (sweep ms, launches)} for "branch" and "site" sweeps."""
        ms = (ctypes.c_double * 2)()
        n = (ctypes.c_uint64 * 2)()
        self._lib.tsg_treeseq_get_profile_modes(self._h, ms, n)
        return {"branch": (ms[0], int(n[0])), "site": (ms[1], int(n[1]))}

    def get_profile(self):
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        self._lib.tsg_treeseq_get_profile(self._h, ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    # -- statistics ------------------------------------------------------------
    def _route(self, fn):
        """The C function and handle a statistics call goes to."""
        return fn, self._h

    def _options(self, mode, span_normalise, polarised, centre=True):
        opts = _parse_stats_mode(mode)
        if span_normalise:
            opts |= STAT_SPAN_NORMALISE
        if polarised:
            opts |= STAT_POLARISED
        if not centre:
            opts |= STAT_NONCENTRED
        return opts

    def _result_shape(self, num_windows, opts, P):
        """(windows, P), or (windows, nodes, P) for node statistics
        (_tskitmodule.c:6638-6660)."""
        if opts & STAT_NODE:
            return (num_windows, self.get_num_nodes(), P)
        return (num_windows, P)

    def _one_way(self, fn, sample_set_sizes, sample_sets, windows, mode, span_normalise,
                 polarised):
        opts = self._options(mode, span_normalise, polarised)
        sizes, sets = _parse_sample_sets(sample_set_sizes, sample_sets)
        w = _parse_windows(windows)
        # every element is written by the library (no need to clear)
        out = np.empty(self._result_shape(len(w) - 1, opts, len(sizes)), dtype=np.float64)
        fn, h = self._route(fn)
        err = fn(h, len(sizes), ptr(sizes, ctypes.c_uint64), ptr(sets, ctypes.c_int32),
                 len(w) - 1, ptr(w, ctypes.c_double), opts, ptr(out, ctypes.c_double))
        check(err)
        return out

    def _k_way(self, fn, k, sample_set_sizes, sample_sets, indexes, windows, mode,
               span_normalise, polarised, centre):
        opts = self._options(mode, span_normalise, polarised, centre)
        sizes, sets = _parse_sample_sets(sample_set_sizes, sample_sets)
        w = _parse_windows(windows)
        idx = _parse_indexes(indexes, k)
        out = np.empty(self._result_shape(len(w) - 1, opts, idx.shape[0]), dtype=np.float64)
        fn, h = self._route(fn)
        err = fn(h, len(sizes), ptr(sizes, ctypes.c_uint64), ptr(sets, ctypes.c_int32),
                 idx.shape[0], ptr(idx, ctypes.c_int32), len(w) - 1, ptr(w, ctypes.c_double),
                 opts, ptr(out, ctypes.c_double))
        check(err)
        return out

    def diversity(self, sample_set_sizes, sample_sets, windows, mode=None, span_normalise=1,
                  polarised=0):
        return self._one_way(self._lib.tsg_treeseq_diversity, sample_set_sizes, sample_sets,
                             windows, mode, span_normalise, polarised)

    def segregating_sites(self, sample_set_sizes, sample_sets, windows, mode=None,
                          span_normalise=1, polarised=0):
        return self._one_way(self._lib.tsg_treeseq_segregating_sites, sample_set_sizes,
                             sample_sets, windows, mode, span_normalise, polarised)

    def Y1(self, sample_set_sizes, sample_sets, windows, mode=None, span_normalise=1,
           polarised=0):
        return self._one_way(self._lib.tsg_treeseq_Y1, sample_set_sizes, sample_sets, windows,
                             mode, span_normalise, polarised)

    def Tajimas_D(self, sample_set_sizes, sample_sets, windows, mode=None, span_normalise=0,
                  polarised=0):
        return self._one_way(self._lib.tsg_treeseq_Tajimas_D, sample_set_sizes, sample_sets,
                             windows, mode, 0, 0)

    def divergence(self, sample_set_sizes, sample_sets, indexes, windows, mode=None,
                   span_normalise=1, polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_divergence, 2, sample_set_sizes, sample_sets,
                           indexes, windows, mode, span_normalise, polarised, centre)

    def Y2(self, sample_set_sizes, sample_sets, indexes, windows, mode=None, span_normalise=1,
           polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_Y2, 2, sample_set_sizes, sample_sets, indexes,
                           windows, mode, span_normalise, polarised, centre)

    def f2(self, sample_set_sizes, sample_sets, indexes, windows, mode=None, span_normalise=1,
           polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_f2, 2, sample_set_sizes, sample_sets, indexes,
                           windows, mode, span_normalise, polarised, centre)

    def Y3(self, sample_set_sizes, sample_sets, indexes, windows, mode=None, span_normalise=1,
           polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_Y3, 3, sample_set_sizes, sample_sets, indexes,
                           windows, mode, span_normalise, polarised, centre)

    def f3(self, sample_set_sizes, sample_sets, indexes, windows, mode=None, span_normalise=1,
           polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_f3, 3, sample_set_sizes, sample_sets, indexes,
                           windows, mode, span_normalise, polarised, centre)

    def f4(self, sample_set_sizes, sample_sets, indexes, windows, mode=None, span_normalise=1,
           polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_f4, 4, sample_set_sizes, sample_sets, indexes,
                           windows, mode, span_normalise, polarised, centre)

    def genetic_relatedness(self, sample_set_sizes, sample_sets, indexes, windows, mode=None,
                            span_normalise=1, polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_genetic_relatedness, 2, sample_set_sizes,
                           sample_sets, indexes, windows, mode, span_normalise, polarised, centre)

    def genetic_relatedness_proportion(self, sample_set_sizes, sample_sets, indexes, windows,
                                       mode=None, span_normalise=1, polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_genetic_relatedness_proportion, 2,
                           sample_set_sizes, sample_sets, indexes, windows, mode,
                           span_normalise, polarised, centre)

    def Fst(self, sample_set_sizes, sample_sets, indexes, windows, mode=None, span_normalise=1,
            polarised=0, centre=1):
        return self._k_way(self._lib.tsg_treeseq_Fst, 2, sample_set_sizes, sample_sets, indexes,
                           windows, mode, span_normalise, polarised, centre)

    def genotype_matrix(self, isolated_as_missing=True):
        """Allele index of every sample at every site, [S, n] int32 (-1 for
        missing data), decoded on the device (genotypes.c:473-595)."""
        S, n = self.get_num_sites(), self.get_num_samples()
        out = np.zeros((S, n), dtype=np.int32)
        if S == 0:
            return out
        opts = 0 if isolated_as_missing else 2  # TSK_ISOLATED_NOT_MISSING
        check(self._lib.tsg_treeseq_genotype_matrix(self._h, opts, ptr(out, ctypes.c_int32)))
        return out

    def divergence_matrix(self, windows, sample_sets=None, sample_set_sizes=None, mode=None,
                          span_normalise=1):
        """_tskitmodule.c divergence_matrix: [W, K, K] over disjoint sample
        sets (site mode, int8 matrix cores)."""
        opts = self._options(mode, span_normalise, False)
        w = _parse_windows(windows)
        if sample_sets is None:
            n = self.get_num_samples()
            sets_p = None
            sizes_p = None
        else:
            sets = as_c(np.asarray(sample_sets), np.int32)
            n = len(sets)
            sets_p = ptr(sets, ctypes.c_int32)
            sizes_p = None
            if sample_set_sizes is not None:
                sizes = as_c(np.asarray(sample_set_sizes), np.uint64)
                n = len(sizes)
                sizes_p = ptr(sizes, ctypes.c_uint64)
        out = np.zeros((len(w) - 1, n, n), dtype=np.float64)
        err = self._lib.tsg_treeseq_divergence_matrix(
            self._h, n, sizes_p, sets_p, len(w) - 1, ptr(w, ctypes.c_double), opts,
            ptr(out, ctypes.c_double))
        check(err)
        return out

    def genetic_relatedness_matrix(self, sample_sets, windows, mode=None, span_normalise=1):
        """Relatedness of singleton sample sets on the int8 matrix-core route.
        ``sample_sets`` is None (every sample) or a list of node ids, each its
        own set.  Returns [W, n, n]."""
        opts = self._options(mode, span_normalise, False)
        w = _parse_windows(windows)
        if sample_sets is None:
            n = self.get_num_samples()
            sets_p = None
        else:
            sets = as_c(np.asarray(sample_sets), np.int32)
            n = len(sets)
            sets_p = ptr(sets, ctypes.c_int32)
        out = np.zeros((len(w) - 1, n, n), dtype=np.float64)
        err = self._lib.tsg_treeseq_genetic_relatedness_matrix(
            self._h, n, None, sets_p, len(w) - 1, ptr(w, ctypes.c_double), opts,
            ptr(out, ctypes.c_double))
        check(err)
        return out

    def genetic_relatedness_vector(self, weights, windows, mode=None, span_normalise=1, centre=1,
                                   nodes=None):
        """_tskitmodule.c:7197-7283 (kwlist weights, windows, mode,
        span_normalise, centre, nodes): [W, len(nodes), num_weights]."""
        opts = self._options(mode, span_normalise, False, centre)
        w = _parse_windows(windows)
        W = np.asarray(weights, dtype=np.float64)
        if W.ndim != 2:
            raise ValueError("weights must be a 2D array")
        W = as_c(W, np.float64)
        if W.shape[0] != self.get_num_samples():
            raise ValueError("First dimension must be num_samples")
        focal = as_c(np.asarray(nodes if nodes is not None else self.get_samples()).reshape(-1),
                     np.int32)
        out = np.zeros((len(w) - 1, len(focal), W.shape[1]), dtype=np.float64)
        err = self._lib.tsg_treeseq_genetic_relatedness_vector(
            self._h, W.shape[1], ptr(W, ctypes.c_double), len(w) - 1, ptr(w, ctypes.c_double),
            len(focal), ptr(focal, ctypes.c_int32), ptr(out, ctypes.c_double), opts)
        check(err)
        return out

    def _weights(self, W):
        W = np.asarray(W, dtype=np.float64)
        if W.ndim != 2:
            raise ValueError("weights must be a 2D array")
        if W.shape[0] != self.get_num_samples():
            raise ValueError("First dimension must be num_samples")
        return as_c(W, np.float64)

    def general_stat(self, weights, summary, output_dim, windows, mode=None, polarised=0,
                     span_normalise=0):
        """_tskitmodule.c:6665-6745, with ``summary`` a registry entry
        (tskit_amd.summaries) in place of the Python callable: [W, output_dim]."""
        opts = self._options(mode, span_normalise, polarised)
        W = self._weights(weights)
        w = _parse_windows(windows)
        idx = summary.index_tuples
        dp = summary.dparams
        out = np.zeros((len(w) - 1, int(output_dim)), dtype=np.float64)
        err = self._lib.tsg_treeseq_general_stat(
            self._h, W.shape[1], ptr(W, ctypes.c_double), int(output_dim), summary.stat_id,
            None if idx is None else ptr(idx, ctypes.c_int32),
            None if dp is None else ptr(dp, ctypes.c_double), len(w) - 1,
            ptr(w, ctypes.c_double), opts, ptr(out, ctypes.c_double))
        check(err)
        return out

    def _weighted(self, fn, weights, windows, mode, polarised, span_normalise, extra=()):
        opts = self._options(mode, span_normalise, polarised)
        W = self._weights(weights)
        w = _parse_windows(windows)
        out = np.zeros((len(w) - 1, W.shape[1]), dtype=np.float64)
        err = fn(self._h, W.shape[1], ptr(W, ctypes.c_double), *extra, len(w) - 1,
                 ptr(w, ctypes.c_double), opts, ptr(out, ctypes.c_double))
        check(err)
        return out

    def trait_covariance(self, weights, windows, mode=None, polarised=0, span_normalise=1):
        """_tskitmodule.c:6748-6819 (kwlist weights, windows, mode, polarised,
        span_normalise): [W, num_weights]"""
        return self._weighted(self._lib.tsg_treeseq_trait_covariance, weights, windows, mode,
                              polarised, span_normalise)

    def trait_correlation(self, weights, windows, mode=None, polarised=0, span_normalise=1):
        """_tskitmodule.c:6748-6819: [W, num_weights]"""
        return self._weighted(self._lib.tsg_treeseq_trait_correlation, weights, windows, mode,
                              polarised, span_normalise)

    def trait_linear_model(self, weights, covariates, windows, mode=None, polarised=0,
                           span_normalise=1):
        """_tskitmodule.c:6821-6905: [W, num_weights]"""
        Z = self._weights(covariates)
        return self._weighted(self._lib.tsg_treeseq_trait_linear_model, weights, windows, mode,
                              polarised, span_normalise, (Z.shape[1], ptr(Z, ctypes.c_double)))

    def genetic_relatedness_weighted(self, weights, indexes, windows, mode=None,
                                     span_normalise=1, polarised=0, centre=1):
        """_tskitmodule.c genetic_relatedness_weighted: [W, num_indexes]"""
        opts = self._options(mode, span_normalise, polarised, centre)
        W = self._weights(weights)
        idx = as_c(np.asarray(indexes, dtype=np.int32).reshape(-1, 2), np.int32)
        w = _parse_windows(windows)
        out = np.zeros((len(w) - 1, idx.shape[0]), dtype=np.float64)
        err = self._lib.tsg_treeseq_genetic_relatedness_weighted(
            self._h, W.shape[1], ptr(W, ctypes.c_double), idx.shape[0], ptr(idx, ctypes.c_int32),
            len(w) - 1, ptr(w, ctypes.c_double), ptr(out, ctypes.c_double), opts)
        check(err)
        return out

    def allele_frequency_spectrum(self, sample_set_sizes, sample_sets, windows, time_windows,
                                  mode=None, span_normalise=1, polarised=0):
        """_tskitmodule.c:6977-7062"""
        opts = self._options(mode, span_normalise, polarised)
        sizes, sets = _parse_sample_sets(sample_set_sizes, sample_sets)
        w = _parse_windows(windows)
        tw = _parse_windows(time_windows)
        shape = [len(w) - 1, len(tw) - 1] + [int(s) + 1 for s in sizes]
        out = np.zeros(shape, dtype=np.float64)
        err = self._lib.tsg_treeseq_allele_frequency_spectrum(
            self._h, len(sizes), ptr(sizes, ctypes.c_uint64), ptr(sets, ctypes.c_int32),
            len(w) - 1, ptr(w, ctypes.c_double), len(tw) - 1, ptr(tw, ctypes.c_double), opts,
            ptr(out, ctypes.c_double))
        check(err)
        return out


class _BatchLowLevel(LowLevelTreeSequence):
    """Routes the statistics calls of a LowLevelTreeSequence into a batch:
    arguments are parsed and validated exactly as for a direct call, the
    result buffer is allocated, and the query is recorded for one fused
    sweep."""

    def __init__(self, ll):  # noqa: D107 -- shares the tree sequence handle
        self._lib = ll._lib
        self._h = ll._h
        self._parent = ll
        b = ctypes.c_void_p()
        check(self._lib.tsg_batch_init(ctypes.byref(b), ll._h))
        self._b = b
        self.last_result = None

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            self._lib.tsg_batch_free(b)
            self._b = None

    def _route(self, fn):
        name = fn.__name__.replace("tsg_treeseq_", "tsg_batch_")
        return getattr(self._lib, name), self._b

    def _one_way(self, *args, **kwargs):
        out = super()._one_way(*args, **kwargs)
        self.last_result = out
        return out

    def _k_way(self, *args, **kwargs):
        out = super()._k_way(*args, **kwargs)
        self.last_result = out
        return out

    def allele_frequency_spectrum(self, *args, **kwargs):
        raise ValueError("the allele frequency spectrum is not batched; call it directly")

    def run(self):
        check(self._lib.tsg_batch_run(self._b))
        check(self._lib.tsg_batch_clear(self._b))


class PendingStat:
    """The result of a batched query; available after StatBatch.run()."""

    def __init__(self, value, buffer, batch):
        self._value = value
        self._buffer = buffer
        self._batch = batch

    def result(self):
        if not self._batch.done:
            raise RuntimeError("StatBatch.run() has not been called")
        if isinstance(self._value, np.ndarray):
            return self._value
        return self._buffer.reshape(-1)[0]


class StatBatch:
    """Statistics queries on one tree sequence, executed together.

    Every statistics method of TreeSequence (except the allele frequency
    spectrum) is available with the same arguments; it returns a PendingStat.
    ``run()`` executes all recorded queries as one fused device sweep per
    mode, sharing the edge walks among them.
    """

    def __init__(self, ts):
        self._ll = _BatchLowLevel(ts.ll_tree_sequence)
        self._facade = TreeSequence(self._ll, ts.tables)
        self._pending = []
        self.done = False

    def __getattr__(self, name):
        method = getattr(self._facade, name)

        def call(*args, **kwargs):
            if self.done:
                raise RuntimeError("this batch has already run")
            self._ll.last_result = None
            value = method(*args, **kwargs)
            p = PendingStat(value, self._ll.last_result, self)
            self._pending.append(p)
            return p

        return call

    def __len__(self):
        return len(self._pending)

    def run(self):
        self._ll.run()
        self.done = True
        return [p.result() for p in self._pending]


def run_batches(batches):
    """Runs several StatBatch objects (on one tree sequence or several, e.g.
    simulation replicates) as one device sweep launch per mode over all of
    them; returns each batch's results, as StatBatch.run() would."""
    batches = list(batches)
    if not batches:
        return []
    for b in batches:
        if not isinstance(b, StatBatch):
            raise TypeError("run_batches takes StatBatch objects")
        if b.done:
            raise RuntimeError("this batch has already run")
    if len({id(b) for b in batches}) != len(batches):
        raise ValueError("a batch appears twice")
    lib = batches[0]._ll._lib
    handles = (ctypes.c_void_p * len(batches))(*[b._ll._b.value for b in batches])
    check(lib.tsg_batch_run_multi(handles, len(batches)))
    out = []
    for b in batches:
        check(lib.tsg_batch_clear(b._ll._b))
        b.done = True
        out.append([p.result() for p in b._pending])
    return out


class TreeSequence:
    """A tree sequence resident on the GPU, with the reference's statistics API."""

    def __init__(self, ll_tree_sequence, tables):
        self._ll_tree_sequence = ll_tree_sequence
        self._tables = tables

    @classmethod
    def load_tables(cls, tables):
        return cls(LowLevelTreeSequence(tables), tables)

    @property
    def ll_tree_sequence(self):
        return self._ll_tree_sequence

    def batch(self):
        """A StatBatch of queries on this tree sequence (one fused sweep)."""
        return StatBatch(self)

    @property
    def tables(self):
        return self._tables

    def dump(self, path):
        """Writes the tree sequence as a ``.trees`` file."""
        self._tables.dump(path)

    @property
    def sequence_length(self):
        return self._tables.sequence_length

    @property
    def num_trees(self):
        return self._ll_tree_sequence.get_num_trees()

    @property
    def num_samples(self):
        return self._ll_tree_sequence.get_num_samples()

    @property
    def num_nodes(self):
        return self._ll_tree_sequence.get_num_nodes()

    @property
    def num_edges(self):
        return self._ll_tree_sequence.get_num_edges()

    @property
    def num_sites(self):
        return self._ll_tree_sequence.get_num_sites()

    @property
    def num_mutations(self):
        return self._ll_tree_sequence.get_num_mutations()

    @property
    def sites_position(self):
        return self._tables.site_position

    def samples(self):
        return self._ll_tree_sequence.get_samples()

    def breakpoints(self, as_array=False):
        b = self._ll_tree_sequence.get_breakpoints()
        return b if as_array else iter(b)

    # -- window handling (trees.py:8108-8159) ----------------------------------
    def parse_windows(self, windows):
        if windows is None:
            windows = [0.0, self.sequence_length]
        elif isinstance(windows, str):
            if windows == "trees":
                windows = self.breakpoints(as_array=True)
            elif windows == "sites":
                windows = np.concatenate(
                    [
                        [] if self.num_sites > 0 else [0.0],
                        self.sites_position,
                        [self.sequence_length],
                    ]
                )
                windows[0] = 0.0
            else:
                raise ValueError(
                    f"Unrecognized window specification {windows}:",
                    "the only allowed strings are 'sites' or 'trees'",
                )
        return np.array(windows)

    def parse_time_windows(self, time_windows):
        if time_windows is None:
            time_windows = [0.0, np.inf]
        return np.array(time_windows)

    def _run_windowed_stat(self, windows, method, *args, **kwargs):
        strip_win = windows is None
        windows = self.parse_windows(windows)
        stat = method(*args, **kwargs, windows=windows)
        if strip_win:
            stat = stat[0]
        return stat

    def _run_windowed_stat_tw(self, windows, time_windows, method, *args, **kwargs):
        strip_win = windows is None
        strip_timewin = time_windows is None
        windows = self.parse_windows(windows)
        time_windows = self.parse_time_windows(time_windows)
        stat = method(*args, **kwargs, windows=windows, time_windows=time_windows)
        if strip_win and strip_timewin:
            stat = stat[0, 0, :]
        elif strip_win:
            stat = stat[0, :, :]
        elif strip_timewin:
            stat = stat[:, 0, :]
        return stat

    # -- sample-set handling (trees.py:8161-8221, 8359-8412) -------------------
    def _one_way_sample_set_stat(self, ll_method, sample_sets, windows=None, time_windows=None,
                                 mode=None, span_normalise=True, polarised=False, use_tw=False):
        if sample_sets is None:
            sample_sets = self.samples()
        drop_dimension = False
        try:
            sample_sets = np.array(sample_sets, dtype=np.uint64)
        except ValueError:
            pass
        else:
            if len(sample_sets.shape) == 1:
                sample_sets = [sample_sets]
                drop_dimension = True
        sample_set_sizes = np.array([len(s) for s in sample_sets], dtype=np.uint32)
        if np.any(sample_set_sizes == 0):
            raise ValueError("Sample sets must contain at least one element")
        flattened = safe_np_int_cast(np.hstack(sample_sets), np.int32)
        if use_tw:
            stat = self._run_windowed_stat_tw(
                windows, time_windows, ll_method, sample_set_sizes, flattened, mode=mode,
                span_normalise=span_normalise, polarised=polarised)
        else:
            stat = self._run_windowed_stat(
                windows, ll_method, sample_set_sizes, flattened, mode=mode,
                span_normalise=span_normalise, polarised=polarised)
        if drop_dimension:
            stat = stat.reshape(stat.shape[:-1])
            if stat.shape == () and windows is None and time_windows is None:
                stat = stat[()]
        return stat

    def _k_way_sample_set_stat(self, ll_method, k, sample_sets, indexes=None, windows=None,
                               mode=None, span_normalise=True, polarised=False, centre=True):
        sample_set_sizes = np.array([len(s) for s in sample_sets], dtype=np.uint32)
        if np.any(sample_set_sizes == 0):
            raise ValueError("Sample sets must contain at least one element")
        flattened = safe_np_int_cast(np.hstack(sample_sets), np.int32)
        drop_based_on_index = False
        if indexes is None:
            drop_based_on_index = True
            if len(sample_sets) != k:
                raise ValueError(
                    "Must specify indexes if there are not exactly {} sample sets.".format(k))
            indexes = np.arange(k, dtype=np.int32)
        drop_dimension = False
        indexes = safe_np_int_cast(indexes, np.int32)
        if len(indexes.shape) == 1:
            indexes = indexes.reshape((1, indexes.shape[0]))
            drop_dimension = True
        if len(indexes.shape) != 2 or indexes.shape[1] != k:
            raise ValueError(
                "Indexes must be convertable to a 2D numpy array with {} columns".format(k))
        stat = self._run_windowed_stat(
            windows, ll_method, sample_set_sizes, flattened, indexes, mode=mode,
            span_normalise=span_normalise, polarised=polarised, centre=centre)
        if drop_dimension:
            stat = stat.reshape(stat.shape[:-1])
            if stat.shape == () and windows is None and drop_based_on_index:
                stat = stat[()]
        return stat

    # -- one-way statistics ------------------------------------------------------
    def diversity(self, sample_sets=None, windows=None, mode="site", span_normalise=True):
        """Mean pairwise diversity (trees.py:8496-8556)."""
        return self._one_way_sample_set_stat(
            self._ll_tree_sequence.diversity, sample_sets, windows=windows, mode=mode,
            span_normalise=span_normalise)

    def segregating_sites(self, sample_sets=None, windows=None, mode="site",
                          span_normalise=True):
        """Segregating sites (trees.py:9790-9842)."""
        return self._one_way_sample_set_stat(
            self._ll_tree_sequence.segregating_sites, sample_sets, windows=windows, mode=mode,
            span_normalise=span_normalise)

    def Y1(self, sample_sets, windows=None, mode="site", span_normalise=True):
        """Y1 (trees.py:10226-10270)."""
        return self._one_way_sample_set_stat(
            self._ll_tree_sequence.Y1, sample_sets, windows=windows, mode=mode,
            span_normalise=span_normalise)

    def Tajimas_D(self, sample_sets=None, windows=None, mode="site"):
        """Tajima's D (trees.py:9967-10022): diversity and segregating sites
        from one fused sweep, composed in FP64 on the device."""
        return self._one_way_sample_set_stat(
            self._ll_tree_sequence.Tajimas_D, sample_sets, windows=windows, mode=mode,
            span_normalise=False)

    def allele_frequency_spectrum(self, sample_sets=None, windows=None, time_windows=None,
                                  mode="site", span_normalise=True, polarised=False):
        """Allele frequency spectrum (trees.py:9844-9965)."""
        if sample_sets is None:
            sample_sets = [self.samples()]
        try:
            sample_sets = np.array(sample_sets, dtype=np.uint64)
        except ValueError:
            pass
        else:
            if len(sample_sets.shape) == 1:
                sample_sets = [sample_sets]
        return self._one_way_sample_set_stat(
            self._ll_tree_sequence.allele_frequency_spectrum, sample_sets, windows=windows,
            time_windows=time_windows, mode=mode, span_normalise=span_normalise,
            polarised=polarised, use_tw=True)

    # -- multi-way statistics ----------------------------------------------------
    def divergence(self, sample_sets, indexes=None, windows=None, mode="site",
                   span_normalise=True):
        """Mean pairwise divergence (trees.py:8558-8627)."""
        return self._k_way_sample_set_stat(
            self._ll_tree_sequence.divergence, 2, sample_sets, indexes=indexes, windows=windows,
            mode=mode, span_normalise=span_normalise)

    def Y2(self, sample_sets, indexes=None, windows=None, mode="site", span_normalise=True):
        return self._k_way_sample_set_stat(
            self._ll_tree_sequence.Y2, 2, sample_sets, indexes=indexes, windows=windows,
            mode=mode, span_normalise=span_normalise)

    def f2(self, sample_sets, indexes=None, windows=None, mode="site", span_normalise=True):
        """Patterson's f2 (trees.py:10384-10428)."""
        return self._k_way_sample_set_stat(
            self._ll_tree_sequence.f2, 2, sample_sets, indexes=indexes, windows=windows,
            mode=mode, span_normalise=span_normalise)

    def Y3(self, sample_sets, indexes=None, windows=None, mode="site", span_normalise=True):
        return self._k_way_sample_set_stat(
            self._ll_tree_sequence.Y3, 3, sample_sets, indexes=indexes, windows=windows,
            mode=mode, span_normalise=span_normalise)

    def f3(self, sample_sets, indexes=None, windows=None, mode="site", span_normalise=True):
        """Patterson's f3 (trees.py:10330-10382)."""
        return self._k_way_sample_set_stat(
            self._ll_tree_sequence.f3, 3, sample_sets, indexes=indexes, windows=windows,
            mode=mode, span_normalise=span_normalise)

    def f4(self, sample_sets, indexes=None, windows=None, mode="site", span_normalise=True):
        """Patterson's f4 (trees.py:10272-10328)."""
        return self._k_way_sample_set_stat(
            self._ll_tree_sequence.f4, 4, sample_sets, indexes=indexes, windows=windows,
            mode=mode, span_normalise=span_normalise)

    def Fst(self, sample_sets, indexes=None, windows=None, mode="site", span_normalise=True):
        """Fst (trees.py:10024-10136): one fused device sweep of diversity and
        divergence, composed in FP64 on the device."""
        return self._k_way_sample_set_stat(
            self._ll_tree_sequence.Fst, 2, sample_sets, indexes=indexes, windows=windows,
            mode=mode, span_normalise=span_normalise)

    def genotype_matrix(self, *, samples=None, isolated_as_missing=None, alleles=None,
                        impute_missing_data=None):
        """The genotypes of trees.py:5563-5640: an [S, n] int32 matrix of allele
        indexes (0 = ancestral, then derived states in order of first
        appearance at the site), -1 for an isolated sample without a mutation
        on it when ``isolated_as_missing`` (the default).  ``samples`` selects
        columns (sample nodes only); ``alleles`` must be None."""
        if alleles is not None:
            raise ValueError("alleles: only the default encoding is supported on this path")
        if isolated_as_missing is None:
            isolated_as_missing = not impute_missing_data
        G = self._ll_tree_sequence.genotype_matrix(bool(isolated_as_missing))
        if samples is None:
            return G
        index = {int(u): j for j, u in enumerate(self.samples())}
        cols = []
        for u in np.atleast_1d(np.asarray(samples)):
            if int(u) not in index:
                raise ValueError("samples: only sample nodes are supported on this path")
            cols.append(index[int(u)])
        return G[:, cols]

    def divergence_matrix(self, sample_sets=None, *, windows=None, num_threads=0, mode=None,
                          span_normalise=True):
        """trees.py:8739-8828: mean pairwise divergence between disjoint sample
        sets, [W, K, K] (or [K, K] without windows).  On this path: site mode,
        from G^T G on the int8 matrix cores, the pairs of each two sets summed
        on the device (num_threads is accepted and ignored)."""
        drop = windows is None
        w = self.parse_windows(windows)
        if sample_sets is None:
            flat, sizes = None, None
        else:
            flat, sizes = [], []
            for s in sample_sets:
                s = np.atleast_1d(np.asarray(s))
                flat.extend(int(x) for x in s)
                sizes.append(len(s))
        D = self._ll_tree_sequence.divergence_matrix(
            w, sample_sets=flat, sample_set_sizes=sizes, mode=mode or "site",
            span_normalise=span_normalise)
        return D[0] if drop else D

    def genetic_relatedness_vector(self, W, windows=None, mode="site", span_normalise=True,
                                   centre=True, nodes=None):
        """Product of the branch relatedness matrix and the weights W (one row
        per sample), trees.py:9157-9224 and 8459-8490: [W, n, K], or [n, K]
        without windows; ``nodes`` (samples here) needs centre=False.  Branch
        mode only, as the reference's C (mode="site" raises its error)."""
        if (not hasattr(W, "__len__")) or (len(W) != self.num_samples):
            raise ValueError("First weight dimension must be equal to number of samples.")
        W = np.asarray(W)
        if len(W.shape) == 1:
            W = W.reshape(W.shape[0], 1)
        if nodes is None:
            nodes = list(self.samples())
        elif centre:
            raise ValueError("If `nodes` is provided, must have centre=False.")
        try:
            nodes = safe_np_int_cast(nodes, np.int32)
        except Exception:
            raise ValueError("Could not interpret `nodes` as a list of node IDs.")
        return self._run_windowed_stat(
            windows, self._ll_tree_sequence.genetic_relatedness_vector, W, mode=mode,
            span_normalise=span_normalise, centre=centre, nodes=nodes)

    def general_stat(self, W, f, output_dim, windows=None, polarised=False, mode=None,
                     span_normalise=True, strict=True):
        """Windowed statistic of weights W (one row per sample) and a summary
        function (trees.py:7917-8004).  ``f`` is an entry of the device
        summary registry (``tskit_amd.summaries``), not an arbitrary Python
        callable; the strict check of f(0) and f(total weight) runs as in the
        reference.  Branch and site mode."""
        from . import summaries

        if not isinstance(f, summaries.Summary):
            raise TypeError("f must be a tskit_amd.summaries entry: the summary runs on the "
                            "GPU, which cannot call a Python function")
        if mode is None:
            mode = "site"
        if strict:
            total_weights = np.sum(W, axis=0)
            for x in [total_weights, total_weights * 0.0]:
                with np.errstate(invalid="ignore", divide="ignore"):
                    fx = np.array(f(x))
                fx[np.isnan(fx)] = 0.0
                if not np.allclose(fx, np.zeros((output_dim,))):
                    raise ValueError("Summary function does not return zero for both "
                                     "zero weight and total weight.")
        return self._run_windowed_stat(
            windows, self._ll_tree_sequence.general_stat, W, f, output_dim,
            polarised=polarised, span_normalise=span_normalise, mode=mode)

    def trait_covariance(self, W, windows=None, mode="site", span_normalise=True):
        """Covariance of each trait (column of W) with the genotypes
        (trees.py:9559-9619): [W, num_traits], or [num_traits] without
        windows."""
        if W.shape[0] != self.num_samples:
            raise ValueError("First trait dimension must be equal to number of samples.")
        return self._run_windowed_stat(windows, self._ll_tree_sequence.trait_covariance, W,
                                       mode=mode, span_normalise=span_normalise)

    def trait_correlation(self, W, windows=None, mode="site", span_normalise=True):
        """Correlation of each trait with the genotypes (trees.py:9621-9700)."""
        if W.shape[0] != self.num_samples:
            raise ValueError("First trait dimension must be equal to number of samples.")
        sds = np.std(W, axis=0)
        if np.any(sds == 0):
            raise ValueError("Weight columns must have positive variance", sds)
        return self._run_windowed_stat(windows, self._ll_tree_sequence.trait_correlation, W,
                                       mode=mode, span_normalise=span_normalise)

    def trait_regression(self, *args, **kwargs):
        """Deprecated alias of trait_linear_model (trees.py:9702)."""
        return self.trait_linear_model(*args, **kwargs)

    def trait_linear_model(self, W, Z=None, windows=None, mode="site", span_normalise=True):
        """Squared coefficient of the genotypes in a linear model of each
        trait on the genotypes and the covariates Z (trees.py:9705-9788); Z
        gets an intercept column and is orthonormalised as there."""
        if W.shape[0] != self.num_samples:
            raise ValueError("First trait dimension must be equal to number of samples.")
        if Z is None:
            Z = np.ones((self.num_samples, 1))
        else:
            tZ = np.column_stack([Z, np.ones((Z.shape[0], 1))])
            if np.linalg.matrix_rank(tZ) == tZ.shape[1]:
                Z = tZ
        if Z.shape[0] != self.num_samples:
            raise ValueError("First dimension of Z must equal the number of samples.")
        if np.linalg.matrix_rank(Z) < Z.shape[1]:
            raise ValueError("Matrix of covariates is computationally singular.")
        K = np.linalg.cholesky(np.matmul(Z.T, Z)).T
        Z = np.matmul(Z, np.linalg.inv(K))
        return self._run_windowed_stat(windows, self._ll_tree_sequence.trait_linear_model, W,
                                       Z, mode=mode, span_normalise=span_normalise)

    def genetic_relatedness_weighted(self, W, indexes=None, windows=None, mode="site",
                                     span_normalise=True, polarised=False, centre=True):
        """Weighted genetic relatedness W_i^T C W_j for each pair (i, j) of
        weight columns (trees.py:9104-9155 and 8414-8457)."""
        if len(W) != self.num_samples:
            raise ValueError("First trait dimension must be equal to number of samples.")
        k = 2
        W = np.asarray(W)
        if indexes is None:
            if W.shape[1] != k:
                raise ValueError("Must specify indexes if there are not exactly {} columns "
                                 "in W.".format(k))
            indexes = np.arange(k, dtype=np.int32)
        drop_dimension = False
        indexes = safe_np_int_cast(indexes, np.int32)
        if len(indexes.shape) == 1:
            indexes = indexes.reshape((1, indexes.shape[0]))
            drop_dimension = True
        if len(indexes.shape) != 2 or indexes.shape[1] != k:
            raise ValueError("Indexes must be convertable to a 2D numpy array with {} "
                             "columns".format(k))
        stat = self._run_windowed_stat(
            windows, self._ll_tree_sequence.genetic_relatedness_weighted, W, indexes,
            mode=mode, span_normalise=span_normalise, polarised=polarised, centre=centre)
        if drop_dimension:
            stat = stat.reshape(stat.shape[:-1])
        return stat

    def genetic_relatedness_matrix(self, sample_sets=None, windows=None, mode="site",
                                   span_normalise=True):
        """Relatedness matrix (trees.py:8978-9102).  Site mode with singleton
        sample sets: G^T G on the int8 matrix cores, FP64 double centring on
        the device.  Otherwise (branch mode, or sets of several samples), as
        the reference: the divergence matrix of the sets, its diagonal
        corrected to include self comparisons and double-centred
        (trees.py:9080-9102).  Returns [K, K], or [W, K, K] with windows."""
        single = sample_sets is None or all(
            len(np.atleast_1d(np.asarray(s))) == 1 for s in sample_sets)
        if mode not in (None, "site") or not single:
            D = self.divergence_matrix(sample_sets, windows=windows, mode=mode,
                                       span_normalise=span_normalise)
            if sample_sets is None:
                n = np.ones(self.num_samples)
            else:
                n = np.array([len(np.atleast_1d(np.asarray(x))) for x in sample_sets])

            def _normalise(B):
                if len(B) == 0:
                    return B
                B = B.copy()
                np.fill_diagonal(B, np.diag(B) * (n - 1) / n)
                K = B + np.mean(B)
                y = np.mean(B, axis=0)
                K -= y[:, np.newaxis] + y[np.newaxis, :]
                return K / -2

            if windows is None:
                return _normalise(D)
            return np.array([_normalise(D[j]) for j in range(D.shape[0])])
        if sample_sets is not None:
            sample_sets = [int(np.atleast_1d(np.asarray(s))[0]) for s in sample_sets]
        drop = windows is None
        w = self.parse_windows(windows)
        out = self._ll_tree_sequence.genetic_relatedness_matrix(
            sample_sets, w, mode=mode, span_normalise=span_normalise)
        return out[0] if drop else out

    def genetic_relatedness(self, sample_sets, indexes=None, windows=None, mode="site",
                            span_normalise=True, polarised=True, proportion=True, centre=True):
        """Genetic relatedness (trees.py:8830-8976).  With ``proportion`` the
        division by the segregating sites of the union of the sample sets
        happens on the device, from the same sweep."""
        ll = self._ll_tree_sequence
        method = ll.genetic_relatedness_proportion if proportion else ll.genetic_relatedness
        out = self._k_way_sample_set_stat(
            method, 2, sample_sets, indexes=indexes, windows=windows, mode=mode,
            span_normalise=span_normalise, polarised=polarised, centre=centre)
        return out
