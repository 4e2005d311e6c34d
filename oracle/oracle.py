"""ORACLE -- test infrastructure only.

ctypes wrappers over
  * liboracle.so: the FP64 C restatement of the reference statistics path
    (tsk_oracle.c), and
  * _ref/libtskit_ref.so: the reference C library compiled from source with
    a thin harness (ref_harness.c), present only where the recipe could run.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker.  Both wrappers take a tskit_amd.TableCollection and
return results shaped like the reference's low-level methods.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libtskit_ref.so")

STAT_IDS = {
    "diversity": 0, "segregating_sites": 1, "Y1": 2, "divergence": 3,
    "genetic_relatedness": 4, "genetic_relatedness_noncentred": 5, "Y2": 6, "f2": 7,
    "Y3": 8, "f3": 9, "f4": 10,
}
TUPLE = {0: 0, 1: 0, 2: 0, 3: 2, 4: 2, 5: 2, 6: 2, 7: 2, 8: 3, 9: 3, 10: 4}

SITE, BRANCH, NODE = 1, 2, 4
POLARISED, SPAN_NORMALISE, ALLOW_UNCAL = 1 << 10, 1 << 11, 1 << 12
NONCENTRED = 1 << 14

_vp = ctypes.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


class OracleError(Exception):
    def __init__(self, err):
        self.err = err
        super().__init__(f"oracle error {err}")


def _table_args(tc):
    cols = [
        np.ascontiguousarray(tc.node_flags, np.uint32),
        np.ascontiguousarray(tc.node_time, np.float64),
        np.ascontiguousarray(tc.edge_left, np.float64),
        np.ascontiguousarray(tc.edge_right, np.float64),
        np.ascontiguousarray(tc.edge_parent, np.int32),
        np.ascontiguousarray(tc.edge_child, np.int32),
        np.ascontiguousarray(tc.site_position, np.float64),
        np.ascontiguousarray(tc.ancestral_state, np.int8),
        np.ascontiguousarray(tc.ancestral_state_offset, np.uint64),
        np.ascontiguousarray(tc.mutation_site, np.int32),
        np.ascontiguousarray(tc.mutation_node, np.int32),
        np.ascontiguousarray(tc.mutation_parent, np.int32),
        np.ascontiguousarray(tc.derived_state, np.int8),
        np.ascontiguousarray(tc.derived_state_offset, np.uint64),
    ]
    return cols


class _Base:
    _init = None
    _free = None

    def __init__(self, tc):
        self.tc = tc
        self._cols = _table_args(tc)
        c = self._cols
        h = _vp()
        err = self._init(
            ctypes.byref(h), ctypes.c_double(tc.sequence_length), len(c[1]), _p(c[0]), _p(c[1]),
            len(c[2]), _p(c[2]), _p(c[3]), _p(c[4]), _p(c[5]), len(c[6]), _p(c[6]), _p(c[7]),
            _p(c[8]), len(c[9]), _p(c[9]), _p(c[10]), _p(c[11]), _p(c[12]), _p(c[13]),
            1 if tc.time_units == "uncalibrated" else 0)
        if err != 0:
            raise OracleError(err)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self._free(self.h)
            self.h = None

    # stat(name, sizes, flat_sets, indexes, windows, options) -> [W, M]
    def stat(self, name, sample_sets, indexes=None, windows=None, mode="site",
             span_normalise=True, polarised=False, centre=True):
        sid = STAT_IDS[name]
        if name == "genetic_relatedness" and not centre:
            sid = STAT_IDS["genetic_relatedness_noncentred"]
        sizes = np.array([len(s) for s in sample_sets], dtype=np.uint64)
        flat = np.ascontiguousarray(np.concatenate(sample_sets).astype(np.int32))
        if windows is None:
            windows = np.array([0.0, self.tc.sequence_length])
        windows = np.ascontiguousarray(windows, np.float64)
        opts = {"site": SITE, "branch": BRANCH, "node": NODE}[mode]
        if span_normalise:
            opts |= SPAN_NORMALISE
        if polarised:
            opts |= POLARISED
        t = TUPLE[sid]
        if t == 0:
            idx = None
            M = len(sample_sets)
        else:
            idx = np.ascontiguousarray(np.asarray(indexes, dtype=np.int32).reshape(-1, t))
            M = idx.shape[0]
        shape = (len(windows) - 1, M)
        if mode == "node":
            shape = (len(windows) - 1, len(self.tc.node_time), M)
        out = np.zeros(shape, dtype=np.float64)
        err = self._stat(sid, sizes, flat, t, M, idx, windows, opts, out)
        if err != 0:
            raise OracleError(err)
        return out

    def relatedness_vector(self, W, windows=None, mode="branch", span_normalise=True,
                           centre=True, nodes=None):
        """genetic_relatedness_vector (trees.c:10402-10816): W [n, K] weights of
        the samples; result [nw, len(nodes), K] (nodes default: the samples)."""
        W = np.ascontiguousarray(np.asarray(W, dtype=np.float64).reshape(len(W), -1))
        if windows is None:
            windows = np.array([0.0, self.tc.sequence_length])
        windows = np.ascontiguousarray(windows, np.float64)
        if nodes is None:
            nodes = np.flatnonzero(self.tc.node_flags & 1)
        nodes = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32))
        opts = {"site": SITE, "branch": BRANCH, "node": NODE}[mode]
        if span_normalise:
            opts |= SPAN_NORMALISE
        if not centre:
            opts |= NONCENTRED
        out = np.zeros((len(windows) - 1, len(nodes), W.shape[1]))
        err = self._relvec(W, windows, nodes, opts, out)
        if err != 0:
            raise OracleError(err)
        return out

    def weighted_stat(self, name, W, windows=None, mode="site", span_normalise=True,
                      polarised=False, Z=None, indexes=None, centre=True):
        """The reference's weighted statistics (trees.c:3976-4220, 4839-4895):
        name in trait_covariance, trait_correlation, trait_linear_model (Z the
        covariates, taken as standardised), genetic_relatedness_weighted
        (indexes [m x 2]); W [n, K].  Result [nw, outputs]."""
        W = np.ascontiguousarray(np.asarray(W, dtype=np.float64).reshape(len(W), -1))
        if windows is None:
            windows = np.array([0.0, self.tc.sequence_length])
        windows = np.ascontiguousarray(windows, np.float64)
        opts = {"site": SITE, "branch": BRANCH, "node": NODE}[mode]
        if span_normalise:
            opts |= SPAN_NORMALISE
        if polarised:
            opts |= POLARISED
        if not centre:
            opts |= NONCENTRED
        idx = None
        if indexes is not None:
            idx = np.ascontiguousarray(np.asarray(indexes, dtype=np.int32).reshape(-1, 2))
        M = len(idx) if idx is not None else W.shape[1]
        shape = (len(windows) - 1, M) if mode != "node" else (
            len(windows) - 1, len(self.tc.node_time), M)
        out = np.zeros(shape)
        Zc = None if Z is None else np.ascontiguousarray(np.asarray(Z, dtype=np.float64))
        err = self._weighted(name, W, Zc, idx, windows, opts, out)
        if err != 0:
            raise OracleError(err)
        return out

    def afs(self, sample_sets, windows=None, time_windows=None, mode="site",
            span_normalise=True, polarised=False):
        sizes = np.array([len(s) for s in sample_sets], dtype=np.uint64)
        flat = np.ascontiguousarray(np.concatenate(sample_sets).astype(np.int32))
        if windows is None:
            windows = np.array([0.0, self.tc.sequence_length])
        if time_windows is None:
            time_windows = np.array([0.0, np.inf])
        windows = np.ascontiguousarray(windows, np.float64)
        time_windows = np.ascontiguousarray(time_windows, np.float64)
        opts = {"site": SITE, "branch": BRANCH, "node": NODE}[mode]
        if span_normalise:
            opts |= SPAN_NORMALISE
        if polarised:
            opts |= POLARISED
        shape = [len(windows) - 1, len(time_windows) - 1] + [int(s) + 1 for s in sizes]
        out = np.zeros(shape, dtype=np.float64)
        err = self._afs(sizes, flat, windows, time_windows, opts, out)
        if err != 0:
            raise OracleError(err)
        return out


class Oracle(_Base):
    """The FP64 C restatement (tsk_oracle.c)."""

    def __init__(self, tc):
        lib = ctypes.CDLL(ORACLE_LIB)
        self._lib = lib
        self._init = lib.oracle_ts_init
        self._free = lib.oracle_ts_free
        lib.oracle_ts_num_trees.restype = ctypes.c_int32
        super().__init__(tc)

    @property
    def num_trees(self):
        return int(self._lib.oracle_ts_num_trees(self.h))

    def breakpoints(self):
        out = np.zeros(self.num_trees + 1)
        self._lib.oracle_ts_breakpoints(self.h, _p(out))
        return out

    def index(self):
        E = len(self.tc.edge_left)
        I = np.zeros(E, np.int32)
        O = np.zeros(E, np.int32)
        self._lib.oracle_ts_index(self.h, _p(I), _p(O))
        return I, O

    def genotypes(self):
        """0/1 derived-allele matrix [S, n] (sites with <= 1 mutation)."""
        G = np.zeros((len(self.tc.site_position), int(np.sum(self.tc.node_flags & 1))),
                     dtype=np.int8)
        err = self._lib.oracle_genotypes(self.h, _p(G))
        if err != 0:
            raise OracleError(err)
        return G

    def genotype_matrix(self, isolated_as_missing=True):
        """Allele index of every sample at every site, [S, n] int32 (-1:
        missing)."""
        G = np.zeros((len(self.tc.site_position), int(np.sum(self.tc.node_flags & 1))),
                     dtype=np.int32)
        err = self._lib.oracle_genotype_matrix(self.h, int(bool(isolated_as_missing)), _p(G))
        if err != 0:
            raise OracleError(err)
        return G

    def relatedness_matrix(self, windows=None, span_normalise=True):
        """Site-mode relatedness of every sample, J (G^T G) J per window (exact
        integer cross product, FP64 centring)."""
        G = self.genotypes().astype(np.int64)
        pos = self.tc.site_position
        if windows is None:
            windows = np.array([0.0, self.tc.sequence_length])
        n = G.shape[1]
        out = np.zeros((len(windows) - 1, n, n))
        for w in range(len(windows) - 1):
            sel = (pos >= windows[w]) & (pos < windows[w + 1])
            C = (G[sel].T @ G[sel]).astype(np.float64)
            r = C.sum(axis=1)
            K = C - r[:, None] / n - r[None, :] / n + r.sum() / n**2
            if span_normalise:
                K /= windows[w + 1] - windows[w]
            out[w] = K
        return out

    def _stat(self, sid, sizes, flat, t, M, idx, windows, opts, out):
        return self._lib.oracle_sample_count_stat(
            self.h, sid, len(sizes), _p(sizes), _p(flat), t, M, _p(idx), len(windows) - 1,
            _p(windows), ctypes.c_uint32(opts), _p(out))

    def _relvec(self, W, windows, nodes, opts, out):
        return self._lib.oracle_relatedness_vector(
            self.h, W.shape[1], _p(W), len(windows) - 1, _p(windows), len(nodes), _p(nodes),
            ctypes.c_uint32(opts), _p(out))

    def _weighted(self, name, W, Z, idx, windows, opts, out):
        # the reference's weight preparation (trees.c:3976-4220, 4839-4895),
        # restated in numpy, then the restated general_stat
        n, K = W.shape
        dpar = None
        if name == "trait_covariance":
            stat, Wn, M = 11, W - W.mean(axis=0), K
        elif name == "trait_correlation":
            mean = W.mean(axis=0)
            msq = (np.sum(W * W, axis=0) - mean * mean * n) / (n - 1)
            Wn = np.column_stack([(W - mean) / np.sqrt(msq), np.full(n, 1.0 / n)])
            stat, M = 12, K
        elif name == "trait_linear_model":
            V = W.T @ Z
            Wn = np.column_stack([W, Z, np.ones(n)])
            stat, M, dpar = 13, K, V
        else:
            Wn = np.column_stack([W, np.full(n, 1.0 / n)])
            dpar = np.append(W.sum(axis=0), 1.0)
            stat = 15 if opts & NONCENTRED else 14
            M = len(idx)
        Wn = np.ascontiguousarray(Wn, dtype=np.float64)
        dpar = None if dpar is None else np.ascontiguousarray(dpar, dtype=np.float64)
        return self._lib.oracle_general_stat(
            self.h, stat, Wn.shape[1], _p(Wn), M, None if idx is None else _p(idx),
            None if dpar is None else _p(dpar), len(windows) - 1, _p(windows),
            ctypes.c_uint32(opts & ~NONCENTRED), _p(out))

    def _afs(self, sizes, flat, windows, time_windows, opts, out):
        return self._lib.oracle_afs(
            self.h, len(sizes), _p(sizes), _p(flat), len(windows) - 1, _p(windows),
            len(time_windows) - 1, _p(time_windows), ctypes.c_uint32(opts), _p(out))


def reference_available():
    return os.path.exists(REF_LIB)


class Reference(_Base):
    """The reference C library compiled from source (oracle/_ref).  With
    ``path`` the tree sequence comes from the reference's own ``.trees``
    loader (``tc`` then only supplies shapes: site positions, sample count)."""

    def __init__(self, tc, path=None):
        lib = ctypes.CDLL(REF_LIB)
        self._lib = lib
        self._init = lib.ref_ts_init
        self._init.argtypes = (
            [ctypes.POINTER(_vp), ctypes.c_double, ctypes.c_uint64, _vp, _vp, ctypes.c_uint64]
            + [_vp] * 4 + [ctypes.c_uint64, _vp, _vp, _vp, ctypes.c_uint64] + [_vp] * 5
            + [ctypes.c_int])
        self._free = lib.ref_ts_free
        lib.ref_ts_num_trees.restype = ctypes.c_uint64
        lib.ref_sample_count_stat.argtypes = (
            [_vp, ctypes.c_int, ctypes.c_uint64, _vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64,
             _vp, ctypes.c_uint32, _vp])
        lib.ref_afs.argtypes = (
            [_vp, ctypes.c_uint64, _vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64, _vp,
             ctypes.c_uint32, _vp])
        lib.ref_genotype_matrix.argtypes = [_vp, ctypes.c_int, _vp]
        lib.ref_divergence_matrix.argtypes = (
            [_vp, ctypes.c_uint64, _vp, ctypes.c_uint64, _vp, ctypes.c_uint32, _vp])
        lib.ref_divergence_matrix_sets.argtypes = (
            [_vp, ctypes.c_uint64, _vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint32, _vp])
        lib.ref_ts_load.argtypes = [ctypes.POINTER(_vp), ctypes.c_char_p]
        if path is None:
            super().__init__(tc)
            return
        self.tc = tc
        h = _vp()
        err = lib.ref_ts_load(ctypes.byref(h), os.fsencode(path))
        if err != 0:
            raise OracleError(err)
        self.h = h

    @property
    def num_trees(self):
        return int(self._lib.ref_ts_num_trees(self.h))

    def breakpoints(self):
        out = np.zeros(self.num_trees + 1)
        self._lib.ref_ts_breakpoints(self.h, _p(out))
        return out

    def index(self):
        E = len(self.tc.edge_left)
        I = np.zeros(E, np.int32)
        O = np.zeros(E, np.int32)
        self._lib.ref_ts_index(self.h, _p(I), _p(O))
        return I, O

    def _stat(self, sid, sizes, flat, t, M, idx, windows, opts, out):
        return self._lib.ref_sample_count_stat(
            self.h, sid, len(sizes), _p(sizes), _p(flat), M, _p(idx), len(windows) - 1,
            _p(windows), opts, _p(out))

    def _relvec(self, W, windows, nodes, opts, out):
        return self._lib.ref_relatedness_vector(
            self.h, ctypes.c_uint64(W.shape[1]), _p(W), ctypes.c_uint64(len(windows) - 1),
            _p(windows), ctypes.c_uint64(len(nodes)), _p(nodes), ctypes.c_uint32(opts), _p(out))

    def _weighted(self, name, W, Z, idx, windows, opts, out):
        which = ["trait_covariance", "trait_correlation", "trait_linear_model",
                 "genetic_relatedness_weighted"].index(name)
        nc = 0 if Z is None else Z.shape[1]
        return self._lib.ref_weighted_stat(
            self.h, which, ctypes.c_uint64(W.shape[1]), _p(W), ctypes.c_uint64(nc),
            None if Z is None else _p(Z), ctypes.c_uint64(0 if idx is None else len(idx)),
            None if idx is None else _p(idx), ctypes.c_uint64(len(windows) - 1), _p(windows),
            ctypes.c_uint32(opts), _p(out))

    def _afs(self, sizes, flat, windows, time_windows, opts, out):
        return self._lib.ref_afs(
            self.h, len(sizes), _p(sizes), _p(flat), len(windows) - 1, _p(windows),
            len(time_windows) - 1, _p(time_windows), opts, _p(out))

    def genotype_matrix(self, isolated_as_missing=True):
        G = np.zeros((len(self.tc.site_position), int(np.sum(self.tc.node_flags & 1))),
                     dtype=np.int32)
        err = self._lib.ref_genotype_matrix(self.h, int(bool(isolated_as_missing)), _p(G))
        if err != 0:
            raise OracleError(err)
        return G

    def divergence_matrix(self, samples, windows=None, mode="site", span_normalise=True):
        samples = np.ascontiguousarray(samples, np.int32)
        if windows is None:
            windows = np.array([0.0, self.tc.sequence_length])
        windows = np.ascontiguousarray(windows, np.float64)
        opts = {"site": SITE, "branch": BRANCH}[mode]
        if span_normalise:
            opts |= SPAN_NORMALISE
        n = len(samples)
        out = np.zeros((len(windows) - 1, n, n))
        err = self._lib.ref_divergence_matrix(self.h, n, _p(samples), len(windows) - 1,
                                              _p(windows), opts, _p(out))
        if err != 0:
            raise OracleError(err)
        return out

    def divergence_matrix_sets(self, sample_sets, windows=None, mode="site",
                               span_normalise=True):
        """tsk_treeseq_divergence_matrix over sample sets: [W, K, K]."""
        sizes = np.array([len(s) for s in sample_sets], dtype=np.uint64)
        flat = np.ascontiguousarray(np.concatenate(sample_sets).astype(np.int32))
        if windows is None:
            windows = np.array([0.0, self.tc.sequence_length])
        windows = np.ascontiguousarray(windows, np.float64)
        opts = {"site": SITE, "branch": BRANCH}[mode]
        if span_normalise:
            opts |= SPAN_NORMALISE
        k = len(sizes)
        out = np.zeros((len(windows) - 1, k, k))
        err = self._lib.ref_divergence_matrix_sets(self.h, k, _p(sizes), _p(flat),
                                                   len(windows) - 1, _p(windows), opts, _p(out))
        if err != 0:
            raise OracleError(err)
        return out


# composites, restated from python/tskit/trees.py (FP64 numpy)
def fst(div, dvg, indexes):
    """trees.py:10113-10125: div [W, K], dvg [W, M] -> [W, M]"""
    out = np.ones_like(dvg)
    for i, (u, v) in enumerate(indexes):
        denom = div[:, u] + div[:, v] + 2 * dvg[:, i]
        with np.errstate(divide="ignore", invalid="ignore"):
            out[:, i] -= 2 * (div[:, u] + div[:, v]) / denom
    return out


def tajimas_d(T, S, sizes):
    """trees.py:10008-10018: T, S [W, K] (not span normalised)."""
    n = np.asarray(sizes, dtype=np.float64)
    h = np.array([np.sum(1 / np.arange(1, nn)) for nn in sizes])
    g = np.array([np.sum(1 / np.arange(1, nn) ** 2) for nn in sizes])
    with np.errstate(invalid="ignore", divide="ignore"):
        a = (n + 1) / (3 * (n - 1) * h) - 1 / h**2
        b = 2 * (n**2 + n + 3) / (9 * n * (n - 1)) - (n + 2) / (h * n) + g / h**2
        return (T - S / h) / np.sqrt(a * S + (b / (h**2 + g)) * S * (S - 1))
