"""CPU tests of the host side of the product: the C ABI library loads and
exports every symbol include/tsgpu.h declares, tree-sequence construction
(validation, indexes, breakpoints) agrees with the oracle, errors carry the
reference's numbers, and the chunk plan is in-bounds and executes to the
oracle's answer (plan_emulator) -- all without calling a GPU kernel."""
import ctypes
import os
import re

import numpy as np
import pytest

import cases
import plan_emulator
import tskit_amd as ta
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tsgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsg_[a-zA-Z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    from tskit_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    for name in names:
        assert hasattr(lib, name), name
    # and the binding declares a signature for each of them
    assert set(names) <= set(_lib.SIGNATURES)


def test_generator_deterministic():
    a = ta.simulate_tables(300, 3e5, 1e4, 1e-8, 1e-8, 42)
    b = ta.simulate_tables(300, 3e5, 1e4, 1e-8, 1e-8, 42)
    c = ta.simulate_tables(300, 3e5, 1e4, 1e-8, 1e-8, 43)
    assert a.column_digest() == b.column_digest()
    assert a.column_digest() != c.column_digest()


def test_config1_shape():
    tc = cases.config1()
    assert tc.num_nodes == 3895 and tc.num_edges == 11726 and tc.num_sites == 2898
    ts = tc.tree_sequence()
    assert ts.num_trees == 2625
    assert ts.num_samples == 1000


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_host_init_matches_oracle(seed):
    tc = cases.generated(150, 2e5, seed)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    assert ts.num_trees == orc.num_trees
    assert np.array_equal(ts.breakpoints(as_array=True), orc.breakpoints())
    I, O = ts.ll_tree_sequence.get_index()
    Io, Oo = orc.index()
    assert np.array_equal(I, Io) and np.array_equal(O, Oo)
    assert np.array_equal(ts.samples(), np.arange(150))


@pytest.mark.parametrize("fixture", ["paper_ex", "single_tree_ex", "nonbinary_ex"])
def test_fixture_trees(fixture):
    tc = getattr(cases, fixture)()
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    assert np.array_equal(ts.breakpoints(as_array=True), orc.breakpoints())


def test_paper_ex_breakpoints():
    ts = cases.paper_ex().tree_sequence()
    assert list(ts.breakpoints(as_array=True)) == [0, 2, 7, 10]


# ---- table validation errors (same numbers as c/tskit/core.h) ---------------
def _bad(mutate):
    tc = cases.paper_ex()
    mutate(tc)
    with pytest.raises(ta.LibraryError) as e:
        tc.tree_sequence()
    return e.value.err


def test_error_bad_edge_interval():
    def m(tc):
        tc.edge_right[0] = tc.edge_left[0]
    assert _bad(m) == -307


def test_error_node_out_of_bounds():
    def m(tc):
        tc.edge_child[0] = 100
    assert _bad(m) == -202


def test_error_time_ordering():
    def m(tc):
        tc.node_time[4] = 1.0
    assert _bad(m) in (-306, -302)


def test_error_right_greater_than_L():
    def m(tc):
        tc.edge_right[-1] = 11
    assert _bad(m) == -309


def test_error_unsorted_sites():
    def m(tc):
        tc.site_position[:] = [4.5, 1.0, 8.5]
    assert _bad(m) == -400


def test_error_bad_site_position():
    def m(tc):
        tc.site_position[2] = 10.0
    assert _bad(m) == -402


def test_error_contradictory_children():
    tc = cases.paper_ex()
    # a second parent for node 0 over [0, 7) (parent 7 is older than 6)
    tc.set_edges(
        np.append(tc.edge_left, 0), np.append(tc.edge_right, 7),
        np.append(tc.edge_parent, 7), np.append(tc.edge_child, 0))
    tc.sort_edges()
    with pytest.raises(ta.LibraryError) as e:
        tc.tree_sequence()
    assert e.value.err == -311


# ---- statistics argument errors: raised before any device work -------------
@pytest.fixture(scope="module")
def paper():
    return cases.paper_ex().tree_sequence()


def _err(fn):
    with pytest.raises(ta.LibraryError) as e:
        fn()
    return e.value.err


def test_stat_errors_match_reference_numbers(paper):
    ll = paper.ll_tree_sequence
    w = [0, 10]
    assert _err(lambda: ll.diversity([2, 0], [0, 1], w)) == -908          # empty set
    assert _err(lambda: ll.diversity([2], [0, 100], w)) == -202           # node OOB
    assert _err(lambda: ll.diversity([2], [0, 5], w)) == -601             # not a sample
    assert _err(lambda: ll.diversity([2], [0, 0], w)) == -600             # duplicate
    assert _err(lambda: ll.diversity([2], [0, 1], [0, 5])) == -901        # windows end
    assert _err(lambda: ll.diversity([2], [0, 1], [1, 10])) == -901       # windows start
    assert _err(lambda: ll.diversity([2], [0, 1], [0, 5, 5, 10])) == -901  # not increasing
    assert _err(lambda: ll.diversity([2], [0, 1], [0, 5], mode="node")) == -901  # node mode checks too
    assert _err(lambda: ll.divergence([2, 2], [0, 1, 2, 3], [(0, 2)], w)) == -907
    assert _err(lambda: ll.f4([1, 1, 1, 1], [0, 1, 2, 3], [(0, 1, 2, 4)], w)) == -907
    assert _err(lambda: ll.allele_frequency_spectrum(
        [2], [0, 1], w, [0, 1, np.inf], mode="site")) == -909
    assert _err(lambda: ll.allele_frequency_spectrum(
        [2], [0, 1], w, [1, np.inf], mode="branch")) == -924
    assert _err(lambda: ll.allele_frequency_spectrum(
        [2], [0, 1], w, [0, np.inf], mode="node")) == -909


def test_stat_value_errors(paper):
    ll = paper.ll_tree_sequence
    with pytest.raises(ValueError):
        ll.diversity([2], [0, 1], [0, 10], mode="bad")
    with pytest.raises(ValueError):
        ll.diversity([2], [0, 1], [0])
    with pytest.raises(ValueError):
        ll.diversity([3], [0, 1], [0, 10])
    with pytest.raises(ValueError):
        ll.divergence([2, 2], [0, 1, 2, 3], [(0, 1, 1)], [0, 10])
    with pytest.raises(ValueError):
        paper.divergence([[0, 1], [2, 3], [0]])  # 3 sets, no indexes
    with pytest.raises(ValueError):
        paper.diversity([[0, 1], []])
    with pytest.raises(ValueError):
        paper.diversity(windows="bad")


def test_time_uncalibrated(paper):
    tc = cases.paper_ex()
    tc.time_units = "uncalibrated"
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    assert _err(lambda: ll.diversity([2], [0, 1], [0, 10], mode="branch")) == -910
    assert _err(lambda: ll.allele_frequency_spectrum(
        [2], [0, 1], [0, 10], [0, np.inf], mode="branch")) == -910


# ---- the chunk plan ----------------------------------------------------------
@pytest.mark.parametrize("n,L,seed,chunks", [
    (200, 2e5, 3, None), (200, 2e5, 3, 7), (60, 5e4, 5, 1), (300, 3e5, 9, 50),
])
def test_plan_executes_to_oracle(n, L, seed, chunks):
    tc = cases.generated(n, L, seed)
    ts = tc.tree_sequence()
    if chunks:
        ts.ll_tree_sequence.set_num_chunks(chunks)
    plan = ts.ll_tree_sequence.get_plan()
    plan["_site_pos"] = tc.site_position
    bp = ts.breakpoints(as_array=True)
    plan_emulator.check_plan_bounds(plan, ts.num_trees, ts.num_samples)
    orc = Oracle(tc)
    sets = cases.sample_sets(n, 2, n // 3, seed)
    w = np.linspace(0, L, 6)
    for mode in ["branch", "site"]:
        for stat, idx in [("diversity", None), ("divergence", [(0, 1), (1, 1)])]:
            for pol in [False, True]:
                got = plan_emulator.run(plan, bp, w, sets, n, stat, mode, idx, pol)
                exp = orc.stat(stat, sets, idx, w, mode, False, pol)
                # FP64 plan and emulator: rounding only
                assert np.max(np.abs(got - exp)) / np.max(np.abs(exp)) < 1e-10


@pytest.mark.parametrize("n,L,seed,chunks", [
    (200, 2e5, 3, None), (200, 2e5, 3, 7), (60, 5e4, 5, 1), (300, 3e5, 9, 50),
    (120, 1e5, 11, 100000),
])
def test_hot_plan_executes_to_oracle(n, L, seed, chunks):
    """The hot plan (slots seeded from the samples' nearest hot ancestors and
    the hot forest's levels) reproduces the oracle, and every index the
    kernel reads is in range."""
    tc = cases.generated(n, L, seed)
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    if chunks:
        ll.set_num_chunks(chunks)
    hp = ll.get_hot_plan()
    assert hp is not None
    plan_emulator.check_hot_plan_bounds(hp, ts.num_trees, ts.num_samples,
                                        ll.get_num_mutations())
    full = ll.get_plan()
    sites = {k: full[k] for k in ("tree_site0", "site_mut0", "mut_allele", "mut_pallele",
                                  "site_nalleles")}
    sites["_site_pos"] = tc.site_position
    bp = ts.breakpoints(as_array=True)
    orc = Oracle(tc)
    sets = cases.sample_sets(n, 2, n // 3, seed)
    w = np.linspace(0, L, 6)
    for mode in ["branch", "site"]:
        for stat, idx in [("diversity", None), ("divergence", [(0, 1), (1, 1)])]:
            for pol in [False, True]:
                got = plan_emulator.run_hot(hp, sites, bp, w, sets, n, stat, mode, idx, pol)
                exp = orc.stat(stat, sets, idx, w, mode, False, pol)
                assert np.max(np.abs(got - exp)) / np.max(np.abs(exp)) < 1e-10


@pytest.mark.parametrize("case,n,L,seed,chunks", [
    ("generated", 60, 5e4, 5, 1), ("generated", 200, 2e5, 3, 7),
    ("generated", 120, 1e5, 11, 100000), ("gapped", 200, 2e5, 3, 7),
    ("gapped", 120, 1e5, 11, None),
])
def test_hot_plan_afs_executes_to_oracle(case, n, L, seed, chunks):
    """The hot-plan AFS program (flush on update only, pieces split at window
    boundaries, last-tree pass) reproduces the oracle; the gapped cases have
    nodes that lose and regain their parent edge, where the reference's
    update rule for an inserted edge's child matters."""
    tc = getattr(cases, case)(n, L, seed)
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    if chunks:
        ll.set_num_chunks(chunks)
    hp = ll.get_hot_plan()
    full = ll.get_plan()
    sites = {k: full[k] for k in ("tree_site0", "site_mut0", "mut_allele", "mut_pallele",
                                  "site_nalleles")}
    sites["_site_pos"] = tc.site_position
    bp = ts.breakpoints(as_array=True)
    orc = Oracle(tc)
    sets = cases.sample_sets(n, 2, n // 3, seed)
    w = np.linspace(0, L, 23)
    for mode, tw in [("site", None), ("branch", None), ("branch", [0, 500, 3000, np.inf])]:
        for pol in [False, True]:
            got = plan_emulator.run_hot_afs(hp, sites, bp, w, tw, sets, n, tc.node_time, mode,
                                            pol)
            exp = orc.afs(sets, w, tw, mode, True, pol)
            if mode == "site":
                assert np.array_equal(got, exp)
            else:
                assert np.max(np.abs(got - exp)) / np.max(np.abs(exp)) < 1e-12


def test_hot_plan_afs_tables():
    """slot_nsd counts each sample at its nearest hot ancestor-or-self, and
    the last tree's parent pointers and times match the oracle's last tree."""
    tc = cases.gapped(120, 1e5, 11)
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    ll.set_num_chunks(9)
    hp = ll.get_hot_plan()
    n = ts.num_samples
    cs0 = hp["chunk_slot0"]
    for c in range(len(cs0) - 1):
        ss = hp["sample_slot"][c * n:(c + 1) * n]
        nsd = np.bincount(ss[ss != plan_emulator.NO_SLOT], minlength=cs0[c + 1] - cs0[c])
        assert np.array_equal(nsd, hp["slot_nsd"][cs0[c]:cs0[c + 1]])
    tlp, tlt = hp["tl_parent"], hp["tl_time"]
    assert np.all(tlt[:n] == tc.node_time[:n])
    assert np.all((tlp == -1) | (tlt[np.maximum(tlp, 0)] > tlt))
    assert np.all(hp["tl_last"] >= 0) and np.all(hp["tl_last"] < tc.sequence_length)
    assert np.all(hp["slot_last0"] >= 0) and np.all(hp["slot_ins0"] >= 0)


def test_hot_plan_bounds_config1():
    tc = cases.config1()
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    hp = ll.get_hot_plan()
    plan_emulator.check_hot_plan_bounds(hp, ts.num_trees, ts.num_samples,
                                        ll.get_num_mutations())


def test_plan_bounds_config1():
    tc = cases.config1()
    ts = tc.tree_sequence()
    plan = ts.ll_tree_sequence.get_plan()
    plan_emulator.check_plan_bounds(plan, ts.num_trees, ts.num_samples)


def test_matrix_and_batch_errors(paper):
    """Argument errors of the divergence matrix (trees.c:8902-8960, 8833-8873)
    and run_batches, raised before any device work, with the reference's
    numbers where it has them; node statistics are recorded in a batch."""
    import tskit_amd as tam

    ll = paper.ll_tree_sequence
    w = [0, 10]
    assert _err(lambda: ll.divergence_matrix(w, mode="node")) == -909
    assert _err(lambda: ll.divergence_matrix(w, sample_sets=[0, 0])) == -600
    assert _err(lambda: ll.divergence_matrix(w, sample_sets=[0, 5])) == -601
    assert _err(lambda: ll.divergence_matrix(w, sample_sets=[0, 100])) == -202
    assert _err(lambda: ll.divergence_matrix(
        w, sample_sets=[0, 1, 1, 2], sample_set_sizes=[2, 2])) == -600
    assert _err(lambda: ll.divergence_matrix([0, 5, 5], mode="site")) == -901
    b = paper.batch()
    b.diversity([[0, 1]], mode="node")  # recorded: runs on its own at b.run()
    assert len(b) == 1
    assert tam.run_batches([]) == []
    with pytest.raises(TypeError):
        tam.run_batches([object()])
    lib = paper.ll_tree_sequence._lib
    assert lib.tsg_batch_run_multi(None, 0) == 0
    assert lib.tsg_batch_run_multi(None, 1) == -4


def test_serve_request_normalise_and_pack():
    """The serving loop's request bookkeeping on the host: no sample sets
    means every sample, a flat list is one set; identical sets share a
    state column when requests are packed."""
    from tskit_amd import serve

    class FakeTS:
        def samples(self):
            return np.arange(10, dtype=np.int32)

    ts = FakeTS()
    a = serve.Request(0, 0, "diversity", {"mode": "site"})
    a.normalise(ts)
    assert len(a.sets) == 1 and np.array_equal(a.sets[0], np.arange(10))
    assert a.shape() == (1, 1, 1)
    b = serve.Request(1, 0, "diversity", {"sample_sets": [3, 1, 2], "mode": "site"})
    b.normalise(ts)
    assert len(b.sets) == 1 and b.shape()[0] == 1
    c = serve.Request(2, 0, "divergence", {"sample_sets": [[1, 2, 3], [4, 5]],
                                           "indexes": [(0, 1)], "windows": [0, 5, 10]})
    c.normalise(ts)
    assert c.shape() == (2, 1, 2)
    srv = serve.StatServer.__new__(serve.StatServer)
    srv.max_columns, srv.max_outputs, srv.launch_budget = 64, 4096, 1 << 30
    srv.one_per_launch = False
    launches = srv._pack([a, b, c], 10)
    assert len(launches) == 1 and len(launches[0]) == 3


@pytest.mark.parametrize("case,chunks", [("generated", 13), ("generated", 1), ("gapped", 9),
                                         ("paper_ex", 2)])
def test_hot_plan_relvec_executes_to_oracle(case, chunks):
    """The relatedness-vector program of tsg_relvec.hip (per-chunk lazy
    push-down over the hot slots, first-tree D_0, prefix over chunks),
    executed on the CPU by the plan emulator, equals the oracle's
    genetic_relatedness_vector: windows spanning the sequence, windows that
    do not, a window boundary inside a chunk."""
    tc = {"generated": lambda: cases.generated(60, 1e5, 3),
          "gapped": lambda: cases.gapped(60, 2e5, 3, count=10),
          "paper_ex": cases.paper_ex}[case]()
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    ll.set_num_chunks(chunks)
    hp = ll.get_hot_plan()
    bp = ts.breakpoints(as_array=True)
    n = ts.num_samples
    weights = np.random.default_rng(3).normal(size=(n, 2))
    L = tc.sequence_length
    orc = Oracle(tc)
    smp = np.flatnonzero(tc.node_flags & 1)
    for win in [np.array([0, L]), np.array([0, 0.3 * L, 0.31 * L, L]),
                np.array([0.1 * L, 0.5 * L, 0.77 * L])]:
        focal = np.arange(n)[::-1]
        y = plan_emulator.run_hot_relvec(hp, bp, win, weights, focal)
        exp = orc.relatedness_vector(weights, win, "branch", False, False, smp[focal])
        assert np.allclose(y, exp, rtol=1e-10, atol=1e-12 * np.max(np.abs(exp)))


def test_relatedness_vector_errors(paper):
    """Argument errors of genetic_relatedness_vector, in the reference's order
    and numbers (trees.c:10783-10816, 1244-1286, 10491-10496)."""
    ll = paper.ll_tree_sequence
    W = np.ones((4, 1))
    assert _err(lambda: ll.genetic_relatedness_vector(W, [0, 10], mode="site")) == -909
    assert _err(lambda: ll.genetic_relatedness_vector(W, [0, 10], mode="node")) == -909
    assert _err(lambda: ll.genetic_relatedness_vector(W, [0, 11], mode="branch")) == -901
    assert _err(lambda: ll.genetic_relatedness_vector(W, [-1, 5], mode="branch")) == -901
    assert _err(lambda: ll.genetic_relatedness_vector(W, [0, 5, 5], mode="branch")) == -901
    assert _err(lambda: ll.genetic_relatedness_vector(W, [0, 10], mode="branch",
                                                      nodes=[0, 99])) == -202
    with pytest.raises(ValueError):
        ll.genetic_relatedness_vector(np.ones((3, 1)), [0, 10], mode="branch")
    with pytest.raises(ValueError):
        paper.genetic_relatedness_vector(np.ones(4), mode="branch", nodes=[0, 1])
    with pytest.raises(ta.LibraryError):
        paper.genetic_relatedness_vector(np.ones(4))  # mode="site", as the reference


def test_weighted_stat_errors(paper):
    """Argument errors of general_stat and the weighted statistics, in the
    reference's order and numbers (trees.c:2048-2075, 3976-3995,
    4839-4860), raised before any device work."""
    from tskit_amd import summaries

    ll = paper.ll_tree_sequence
    W = np.ones((4, 2))
    w = [0, 10]
    cov = summaries.trait_covariance(4)
    assert _err(lambda: ll.general_stat(np.ones((4, 0)), cov, 1, w)) == -903
    assert _err(lambda: ll.general_stat(W, cov, 0, w)) == -904
    assert _err(lambda: ll.general_stat(W, cov, 2, [0, 5])) == -901
    assert _err(lambda: ll.general_stat(W, cov, 3, w)) == -904      # result_dim != state_dim
    assert _err(lambda: ll.general_stat(W, cov, 2, w, mode="node")) == -909
    bad = summaries.relatedness_weighted_noncentred([(0, 2)])
    assert _err(lambda: ll.general_stat(W, bad, 1, w)) == -907
    assert _err(lambda: ll.trait_covariance(np.ones((4, 0)), w)) == -913
    assert _err(lambda: ll.trait_correlation(np.ones((4, 0)), w)) == -913
    assert _err(lambda: ll.genetic_relatedness_weighted(np.ones((4, 0)), [(0, 0)], w)) == -913
    with pytest.raises(ValueError):
        ll.trait_covariance(np.ones((3, 2)), w)
    with pytest.raises(ValueError):
        paper.trait_covariance(np.ones((3, 2)))
    with pytest.raises(ValueError):
        paper.trait_correlation(np.ones((4, 2)))  # zero variance
    with pytest.raises(ValueError):
        paper.genetic_relatedness_weighted(np.ones((4, 3)))  # three columns, no indexes
    with pytest.raises(ValueError):
        paper.trait_linear_model(np.ones((4, 1)), Z=np.ones((3, 1)))


def test_serve_memory_admission():
    """A job starts only while the launches in flight and its own fit the
    device-memory budget, or when nothing is in flight."""
    from tskit_amd import serve

    class FakeTS:
        def samples(self):
            return np.arange(1000, dtype=np.int32)

    srv = serve.StatServer.__new__(serve.StatServer)
    srv.chunks = [100]
    r = serve.Request(0, 0, "diversity", {"sample_sets": [np.arange(500)], "windows": [0, 1, 2]})
    r.normalise(FakeTS())
    job = [[r]]
    need = srv.job_bytes(job)
    assert need == serve.launch_bytes(1, 1, 2, 100, 500)
    srv.memory_budget, srv.inflight_bytes = need + 10, 0
    assert srv.admits(job, 0) and srv.admits(job, 1)
    srv.inflight_bytes = 11
    assert not srv.admits(job, 1)
    assert srv.admits(job, 0)  # nothing in flight: always
    srv.memory_budget = None
    assert srv.admits(job, 5)
