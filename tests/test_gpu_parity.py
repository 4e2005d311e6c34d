"""GPU parity: every statistic computed by the HIP kernels (through the C ABI)
against the FP64 oracle on the same inputs.

Tolerances (stated per statistic, the bar of the hot path):
  * non-negative statistics (diversity, segregating sites, divergence, Y1,
    Y2, Y3, non-centred relatedness, branch AFS): elementwise relative error
    <= 1e-6;
  * statistics that change sign (f2, f3, f4, centred relatedness): error
    <= 1e-6 of the largest |value| of that output over the windows;
  * FP64 composites (Fst, Tajima's D, relatedness proportion), which change
    sign too (Fst and D of sets from one population are near 0): error
    <= 1e-6 of the largest |value| of that output over the windows;
  * site-mode AFS: bit-exact.
"""
import numpy as np
import pytest

import cases
from cases import rel_err
from oracle.oracle import Oracle, fst, tajimas_d

pytestmark = pytest.mark.gpu

TOL = 1e-6
POSITIVE = {"diversity", "segregating_sites", "Y1", "divergence", "Y2", "Y3"}


def _say(*args):
    print("[case]", *args, flush=True)


def _check(name, got, exp):
    if name in POSITIVE or name == "genetic_relatedness_nc":
        err = rel_err(got, exp)
    else:
        fin = np.where(np.isfinite(exp), np.abs(exp), 0.0)
        scale = np.max(fin, axis=tuple(range(exp.ndim - 1)), keepdims=True)
        err = rel_err(got, exp, scale=np.maximum(scale, 1e-300))
    assert err <= TOL, f"{name}: relative error {err:.3e}"
    return err


@pytest.fixture(scope="module")
def small(require_gpu):
    tc = cases.generated(200, 2e5, 7)
    return tc, tc.tree_sequence(), Oracle(tc)


@pytest.fixture(scope="module")
def cfg1(require_gpu):
    tc = cases.config1()
    return tc, tc.tree_sequence(), Oracle(tc)


ONE_WAY = ["diversity", "segregating_sites", "Y1"]
K_WAY = [
    ("divergence", [(0, 1), (0, 0), (2, 3), (3, 1)]),
    ("Y2", [(0, 1), (2, 3)]),
    ("f2", [(0, 1), (1, 2)]),
    ("Y3", [(0, 1, 2), (3, 2, 1)]),
    ("f3", [(0, 1, 2), (2, 0, 3)]),
    ("f4", [(0, 1, 2, 3), (3, 2, 0, 1)]),
]


@pytest.mark.parametrize("mode", ["branch", "site"])
@pytest.mark.parametrize("windows", ["none", "ten", "trees", "sites"])
def test_one_way(small, mode, windows):
    tc, ts, orc = small
    sets = cases.sample_sets(200, 4, 40, 1)
    w = {"none": None, "ten": np.linspace(0, tc.sequence_length, 11),
         "trees": ts.breakpoints(as_array=True), "sites": ts.parse_windows("sites")}[windows]
    for name in ONE_WAY:
        _say(name, mode, windows)
        got = getattr(ts, name)(sets, windows=w, mode=mode)
        exp = orc.stat(name, sets, None, ts.parse_windows(w), mode)
        if w is None:
            exp = exp[0]
        _check(name, got, exp)


@pytest.mark.parametrize("mode", ["branch", "site"])
@pytest.mark.parametrize("span_normalise", [True, False])
def test_k_way(small, mode, span_normalise):
    tc, ts, orc = small
    sets = cases.sample_sets(200, 4, 30, 2)
    w = np.array([0, 1.5e4, 6e4, 6.00001e4, 1.3e5, 2e5])
    for name, idx in K_WAY:
        _say(name, mode, span_normalise)
        got = getattr(ts, name)(sets, indexes=idx, windows=w, mode=mode,
                                span_normalise=span_normalise)
        exp = orc.stat(name, sets, idx, w, mode, span_normalise)
        _check(name, got, exp)


@pytest.mark.parametrize("mode", ["branch", "site"])
def test_k_way_polarised(small, mode):
    """Polarised k-way statistics through the low-level interface: in branch
    mode these are the summaries evaluated hop by hop (cubic and quartic),
    unlike their unpolarised forms, which are quadratic."""
    tc, ts, orc = small
    ll = ts.ll_tree_sequence
    sets = cases.sample_sets(200, 4, 30, 12)
    sizes = np.array([len(s) for s in sets], dtype=np.uint64)
    flat = np.concatenate(sets).astype(np.int32)
    w = np.array([0, 3e4, 1.2e5, 2e5])
    for name, idx in K_WAY + [("Y1", None)]:
        _say(name, "polarised", mode)
        if idx is None:
            got = ll.Y1(sizes, flat, w, mode=mode, span_normalise=1, polarised=1)
        else:
            got = getattr(ll, name)(sizes, flat, np.array(idx, dtype=np.int32), w, mode=mode,
                                    span_normalise=1, polarised=1)
        exp = orc.stat(name, sets, idx, w, mode, True, True)
        _check(name, got, exp)


@pytest.mark.parametrize("mode", ["branch", "site"])
def test_relatedness(small, mode):
    tc, ts, orc = small
    sets = cases.sample_sets(200, 5, 12, 3)
    idx = [(0, 1), (1, 1), (2, 4), (3, 0)]
    w = np.linspace(0, tc.sequence_length, 4)
    for polarised in [True, False]:
        for centre in [True, False]:
            _say("relatedness", mode, polarised, centre)
            got = ts.genetic_relatedness(sets, indexes=idx, windows=w, mode=mode,
                                         polarised=polarised, centre=centre, proportion=False)
            exp = orc.stat("genetic_relatedness", sets, idx, w, mode, True, polarised, centre)
            _check("genetic_relatedness" if centre else "genetic_relatedness_nc", got, exp)
            # proportion: divide by segregating sites of the union (unpolarised)
            got = ts.genetic_relatedness(sets, indexes=idx, windows=w, mode=mode,
                                         polarised=polarised, centre=centre, proportion=True)
            union = np.array(sorted({int(u) for s in sets for u in s}), dtype=np.int32)
            seg = orc.stat("segregating_sites", [union], None, w, mode, True, False)
            _check("genetic_relatedness_proportion", got, exp / seg)


@pytest.mark.parametrize("mode", ["branch", "site"])
def test_composites(small, mode):
    tc, ts, orc = small
    sets = cases.sample_sets(200, 3, 50, 4)
    idx = [(0, 1), (1, 2), (0, 2)]
    w = np.linspace(0, tc.sequence_length, 8)
    _say("Fst", mode)
    got = ts.Fst(sets, indexes=idx, windows=w, mode=mode)
    div = orc.stat("diversity", sets, None, w, mode)
    dvg = orc.stat("divergence", sets, idx, w, mode)
    _check("Fst", got, fst(div, dvg, idx))
    _say("Tajimas_D", mode)
    got = ts.Tajimas_D(sets, windows=w, mode=mode)
    T = orc.stat("diversity", sets, None, w, mode, False)
    S = orc.stat("segregating_sites", sets, None, w, mode, False)
    _check("Tajimas_D", got, tajimas_d(T, S, [len(s) for s in sets]))


def test_config1_headline(cfg1):
    """The small workload: diversity, divergence and Fst on 2 disjoint sets of
    500 over 10 windows, both modes."""
    tc, ts, orc = cfg1
    sets = cases.sample_sets(1000, 2, 500, 1)
    w = np.linspace(0, tc.sequence_length, 11)
    for mode in ["branch", "site"]:
        _say("config1", mode)
        _check("diversity", ts.diversity(sets, windows=w, mode=mode),
               orc.stat("diversity", sets, None, w, mode))
        _check("divergence", ts.divergence(sets, windows=w, mode=mode),
               orc.stat("divergence", sets, [(0, 1)], w, mode)[:, 0])
        div = orc.stat("diversity", sets, None, w, mode)
        dvg = orc.stat("divergence", sets, [(0, 1)], w, mode)
        _check("Fst", ts.Fst(sets, windows=w, mode=mode), fst(div, dvg, [(0, 1)])[:, 0])


@pytest.mark.parametrize("num_chunks", [1, 2, 37, 100000])
def test_chunk_counts(require_gpu, num_chunks):
    """Results do not depend on how the genome is cut into chunks."""
    tc = cases.generated(120, 1e5, 11)
    ts = tc.tree_sequence()
    ts.ll_tree_sequence.set_num_chunks(num_chunks)
    orc = Oracle(tc)
    sets = cases.sample_sets(120, 3, 30, 5)
    w = np.linspace(0, tc.sequence_length, 7)
    for mode in ["branch", "site"]:
        _say("chunks", num_chunks, mode)
        _check("diversity", ts.diversity(sets, windows=w, mode=mode),
               orc.stat("diversity", sets, None, w, mode))
        _check("f4", ts.f4(sets + [sets[0]], windows=w, mode=mode),
               orc.stat("f4", sets + [sets[0]], [(0, 1, 2, 3)], w, mode)[:, 0])


@pytest.mark.parametrize("plan", ["hot16", "hot32", "full"])
def test_node_stores(require_gpu, plan):
    """The same statistics from the hot plan (slots in LDS, 16- or 32-bit
    counts) and from the full plan, whose chunks are too large for LDS here
    and run from per-chunk HBM slabs."""
    tc = cases.generated(3000, 1e6, 13)
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    ll.set_sweep_plan({"hot16": 0, "full": 1, "hot32": 2}[plan])
    orc = Oracle(tc)
    sets = cases.sample_sets(3000, 8, 100, 6)
    w = np.linspace(0, tc.sequence_length, 5)
    idx = [(i, j) for i in range(8) for j in range(i + 1, 8)]
    for mode in ["branch", "site"]:
        _say("store", plan, mode)
        _check("divergence", ts.divergence(sets, indexes=idx, windows=w, mode=mode),
               orc.stat("divergence", sets, idx, w, mode))
        assert ll.get_last_store() == {"hot16": 3, "full": 0, "hot32": 2}[plan]
        _check("f4", ts.f4(sets[:4], windows=w, mode=mode),
               orc.stat("f4", sets[:4], [(0, 1, 2, 3)], w, mode)[:, 0])


@pytest.mark.parametrize("mode", ["branch", "site"])
def test_gapped_stats(require_gpu, mode):
    """Statistics on trees with several roots and samples missing over gaps
    (nodes that lose their parent edge and regain it later)."""
    tc = cases.gapped(200, 2e5, 3)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    sets = cases.sample_sets(200, 4, 40, 3)
    w = np.linspace(0, tc.sequence_length, 9)
    _say("gapped stats", mode)
    _check("diversity", ts.diversity(sets, windows=w, mode=mode),
           orc.stat("diversity", sets, None, w, mode))
    idx = [(0, 1), (2, 3), (1, 1)]
    _check("divergence", ts.divergence(sets, indexes=idx, windows=w, mode=mode),
           orc.stat("divergence", sets, idx, w, mode))
    _check("f4", ts.f4(sets, indexes=[(0, 1, 2, 3)], windows=w, mode=mode),
           orc.stat("f4", sets, [(0, 1, 2, 3)], w, mode))


def test_unused_sets_dropped(small):
    """Sets that no index tuple names do not become state columns; the
    remaining columns are renumbered (results as the oracle's)."""
    tc, ts, orc = small
    sets = cases.sample_sets(200, 7, 20, 11)
    w = np.linspace(0, tc.sequence_length, 4)
    for mode in ["branch", "site"]:
        _say("unused sets", mode)
        for stat, idx in [("divergence", [(6, 2), (2, 2)]), ("f3", [(5, 1, 3)]),
                          ("f4", [(6, 4, 2, 0)]), ("Y2", [(3, 5)])]:
            got = getattr(ts, stat)(sets, indexes=idx, windows=w, mode=mode)
            _check(stat, np.asarray(got).reshape(len(w) - 1, -1),
                   orc.stat(stat, sets, idx, w, mode))
        b = ts.batch()
        p1 = b.f2(sets, indexes=[(4, 1)], windows=w, mode=mode)
        p2 = b.diversity(sets[2:4], windows=w, mode=mode)
        b.run()
        _check("f2", np.asarray(p1.result()).reshape(len(w) - 1, -1),
               orc.stat("f2", sets, [(4, 1)], w, mode))
        _check("diversity", p2.result(), orc.stat("diversity", sets[2:4], None, w, mode))


@pytest.mark.parametrize("name", ["basic_tree_seq", "minimal-example", "single-locus-example",
                                  "tskit-0.3.3"])
def test_trees_files(require_gpu, name):
    """The reference's own .trees data files (tests/golden/trees: a docs
    example, two SLiM outputs, a tskit 0.3.3 file), read by the kastore
    reader: GPU statistics and genotypes against the oracle."""
    import os

    import tskit_amd as ta

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trees",
                        name + ".trees")
    tc = ta.load_tables(path)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    smp = ts.samples()
    half = [smp[: len(smp) // 2], smp[len(smp) // 2:]]
    w = np.linspace(0, tc.sequence_length, 3)
    for mode in (["branch", "site"] if tc.num_sites else ["branch"]):
        _say("trees file", name, mode)
        got = ts.diversity(half, windows=w, mode=mode)
        exp = orc.stat("diversity", half, None, w, mode)
        assert np.allclose(got, exp, rtol=1e-9, atol=1e-12 * max(1.0, np.max(np.abs(exp))))
        got = ts.divergence(half, indexes=[(0, 1)], windows=w, mode=mode)
        exp = orc.stat("divergence", half, [(0, 1)], w, mode)
        assert np.allclose(np.asarray(got).reshape(exp.shape), exp, rtol=1e-9,
                           atol=1e-12 * max(1.0, np.max(np.abs(exp))))
    if tc.num_sites:
        assert np.array_equal(ts.genotype_matrix(), orc.genotype_matrix())


@pytest.mark.parametrize("fixture", ["paper_ex", "single_tree_ex", "nonbinary_ex"])
def test_reference_fixtures(require_gpu, fixture):
    tc = getattr(cases, fixture)()
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    n = ts.num_samples
    samples = list(range(n))
    sets = [samples[: n // 2], samples[n // 2:]]
    for mode in ["branch", "site"]:
        for pol in [False, True]:
            _say(fixture, mode, pol)
            got = ts.ll_tree_sequence.diversity(
                [len(s) for s in sets], np.concatenate(sets), [0, tc.sequence_length],
                mode=mode, span_normalise=False, polarised=pol)
            exp = orc.stat("diversity", sets, None, None, mode, False, pol)
            _check("diversity", got, exp)
            got = ts.ll_tree_sequence.divergence(
                [len(s) for s in sets], np.concatenate(sets), [(0, 1)], [0, tc.sequence_length],
                mode=mode, span_normalise=False, polarised=pol)
            exp = orc.stat("divergence", sets, [(0, 1)], None, mode, False, pol)
            _check("divergence", got, exp)


def test_paper_ex_known_answers(require_gpu):
    """c/tests/test_stats.c:1718-1742, 1872-1897, 2518-2547."""
    ts = cases.paper_ex().tree_sequence()
    ll = ts.ll_tree_sequence
    pi = ll.diversity([4], [0, 1, 2, 3], [0, 10], mode="site", span_normalise=False)
    assert abs(pi[0, 0] - 1.5) < 1e-6
    pi1 = ll.diversity([1], [0], [0, 10], mode="site", span_normalise=False)
    assert np.isnan(pi1[0, 0])
    seg = ll.segregating_sites([4], [0, 1, 2, 3], [0, 10], mode="site", span_normalise=False)
    assert abs(seg[0, 0] - 3.0) < 1e-6
    seg1 = ll.segregating_sites([1], [0], [0, 10], mode="site", span_normalise=False)
    assert abs(seg1[0, 0]) < 1e-6
    afs = ll.allele_frequency_spectrum([4], [0, 1, 2, 3], [0, 10], [0, np.inf], mode="site",
                                       span_normalise=False)
    assert list(afs[0, 0]) == [0, 3, 0, 0, 0]
    afs = ll.allele_frequency_spectrum([4], [0, 1, 2, 3], [0, 10], [0, np.inf], mode="site",
                                       span_normalise=False, polarised=True)
    assert list(afs[0, 0]) == [0, 2, 0, 1, 0]


def _afs_plan(ll, plan):
    """1: the full plan (every node of each chunk); 0: the library's choice,
    the hot plan here."""
    ll.set_sweep_plan(1 if plan == "full" else 0)


def _afs_store_ok(ll, plan):
    store = ll.get_last_store()
    assert (store == 2) if plan == "hot" else store in (0, 1)


@pytest.mark.parametrize("plan", ["hot", "full"])
@pytest.mark.parametrize("mode", ["site", "branch"])
@pytest.mark.parametrize("nsets", [1, 2, 3])
@pytest.mark.parametrize("polarised", [False, True])
def test_afs(small, plan, mode, nsets, polarised):
    tc, ts, orc = small
    ll = ts.ll_tree_sequence
    sets = cases.sample_sets(200, nsets, 13, 8)
    w = np.linspace(0, tc.sequence_length, 5)
    _say("afs", plan, mode, nsets, polarised)
    _afs_plan(ll, plan)
    try:
        got = ts.allele_frequency_spectrum(sets, windows=w, mode=mode, polarised=polarised)
        _afs_store_ok(ll, plan)
    finally:
        ll.set_sweep_plan(0)
    exp = orc.afs(sets, w, None, mode, True, polarised)[:, 0]
    if mode == "site":
        assert np.array_equal(got, exp)
    else:
        scale = np.max(np.abs(exp), axis=tuple(range(1, exp.ndim)), keepdims=True)
        assert rel_err(got, exp, np.maximum(scale, 1e-300)) <= TOL


@pytest.mark.parametrize("plan", ["hot", "full"])
def test_afs_time_windows(small, plan):
    tc, ts, orc = small
    ll = ts.ll_tree_sequence
    sets = cases.sample_sets(200, 2, 20, 9)
    w = np.linspace(0, tc.sequence_length, 3)
    tw = np.array([0, 2000, 10000, np.inf])
    _say("afs time windows", plan)
    _afs_plan(ll, plan)
    try:
        got = ts.allele_frequency_spectrum(sets, windows=w, time_windows=tw, mode="branch")
        _afs_store_ok(ll, plan)
    finally:
        ll.set_sweep_plan(0)
    exp = orc.afs(sets, w, tw, "branch", True, False)
    scale = np.max(np.abs(exp))
    assert rel_err(got, exp, scale) <= TOL


@pytest.mark.parametrize("plan", ["hot", "full"])
@pytest.mark.parametrize("polarised", [False, True])
def test_afs_gapped(require_gpu, plan, polarised):
    """Trees with several roots and samples missing over gaps: nodes lose
    their parent edge and regain it later, the case where the reference's
    update rule (an inserted edge's child keeps its last update position)
    shapes the branch spectrum; many windows and chunks."""
    tc = cases.gapped(200, 2e5, 3)
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    ll.set_num_chunks(7)
    orc = Oracle(tc)
    sets = cases.sample_sets(200, 2, 66, 3)
    w = np.linspace(0, tc.sequence_length, 23)
    tw = np.array([0, 500, 3000, np.inf])
    _afs_plan(ll, plan)
    for mode, t in [("site", None), ("branch", None), ("branch", tw)]:
        _say("afs gapped", plan, mode, t is not None, polarised)
        got = ts.allele_frequency_spectrum(sets, windows=w, time_windows=t, mode=mode,
                                           polarised=polarised)
        _afs_store_ok(ll, plan)
        exp = orc.afs(sets, w, t, mode, True, polarised)
        if t is None:
            exp = exp[:, 0]
        if mode == "site":
            assert np.array_equal(got, exp)
        else:
            assert rel_err(got, exp, np.max(np.abs(exp))) <= TOL


@pytest.mark.parametrize("windows", [None, "three"])
def test_relatedness_matrix(require_gpu, windows):
    """int8 MFMA cross products are exact, so the GPU matrix equals the FP64
    closed form to rounding of the FP64 centring."""
    tc = cases.generated(300, 5e5, 17)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    w = None if windows is None else np.array([0, 1e5, 3.5e5, 5e5])
    _say("relatedness matrix", windows)
    got = ts.genetic_relatedness_matrix(windows=w)
    exp = orc.relatedness_matrix(w, True)
    if w is None:
        exp = exp[0]
    assert got.shape == exp.shape
    assert np.allclose(got, exp, rtol=1e-10, atol=1e-12 * np.max(np.abs(exp)))


@pytest.mark.parametrize("plan", ["hot", "full"])
def test_relatedness_matrix_wide(require_gpu, plan):
    """3000 samples: several workgroups per chunk split the sample bitsets
    (hot plan, 2048 samples each); the hot and full genotype sweeps give the
    same matrix, the oracle's closed form to FP64 rounding."""
    tc = cases.generated(3000, 3e5, 23)
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    orc = Oracle(tc)
    w = np.array([0, 1.2e5, 3e5])
    _say("relatedness matrix wide", plan)
    ll.set_sweep_plan(1 if plan == "full" else 0)
    try:
        got = ts.genetic_relatedness_matrix(windows=w)
        assert ll.get_last_store() == (2 if plan == "hot" else 0)
    finally:
        ll.set_sweep_plan(0)
    exp = orc.relatedness_matrix(w, True)
    assert np.allclose(got, exp, rtol=1e-10, atol=1e-12 * np.max(np.abs(exp)))


@pytest.mark.parametrize("case", ["paper_ex", "single_tree_ex", "nonbinary_ex", "generated",
                                  "gapped", "wide"])
@pytest.mark.parametrize("missing", [True, False])
def test_genotype_matrix(require_gpu, case, missing):
    """Device genotype decode equals the oracle's (pinned to the reference's
    tsk_variant_decode): multi-allelic sites with back and recurrent
    mutations (single_tree_ex), isolated samples as missing or ancestral
    (gapped), several 2048-sample blocks per chunk (wide).  Bit-exact."""
    if case == "generated":
        tc = cases.generated(150, 2e5, 29)
    elif case == "gapped":
        tc = cases.gapped(200, 2e5, 3)
    elif case == "wide":
        tc = cases.generated(5000, 1e5, 33)
    else:
        tc = getattr(cases, case)()
    ts = tc.tree_sequence()
    _say("genotype matrix", case, missing)
    got = ts.genotype_matrix(isolated_as_missing=missing)
    exp = Oracle(tc).genotype_matrix(missing)
    assert got.dtype == np.int32 and np.array_equal(got, exp)
    if case == "generated":
        sel = ts.samples()[[5, 1, 77]]
        assert np.array_equal(ts.genotype_matrix(samples=sel), exp[:, [5, 1, 77]])


@pytest.mark.parametrize("windows", [None, "three"])
def test_divergence_matrix(require_gpu, windows):
    """Site divergence matrix of singleton sets: C_ii + C_jj - 2 C_ij from the
    int8 cross products, against the reference's own tsk_treeseq_divergence_
    matrix when oracle/_ref is present, else the oracle's genotypes."""
    from oracle import oracle as om

    tc = cases.generated(300, 5e5, 17)
    ts = tc.tree_sequence()
    w = None if windows is None else np.array([0, 1e5, 3.5e5, 5e5])
    _say("divergence matrix", windows)
    got = ts.divergence_matrix(windows=w)
    wn = np.array([0, 5e5]) if w is None else w
    if om.reference_available():
        exp = om.Reference(tc).divergence_matrix(np.arange(300, dtype=np.int32), wn, "site")
    else:
        G = Oracle(tc).genotypes().astype(np.int64)
        pos = tc.site_position
        exp = []
        for j in range(len(wn) - 1):
            g = G[(pos >= wn[j]) & (pos < wn[j + 1])]
            C = g.T @ g
            a = np.diag(C)
            exp.append((a[:, None] + a[None, :] - 2 * C) / (wn[j + 1] - wn[j]))
        exp = np.array(exp, dtype=np.float64)
    if w is None:
        exp = exp[0]
    assert got.shape == exp.shape
    assert np.allclose(got, exp, rtol=1e-12, atol=1e-15)
    sel = [[7], [3], [150]]
    sub = ts.divergence_matrix(sample_sets=sel, windows=w)
    full = got if w is None else got
    idx = [7, 3, 150]
    assert np.allclose(sub, full[..., idx, :][..., :, idx], rtol=1e-12, atol=1e-15)


def test_relatedness_matrix_subset(require_gpu):
    tc = cases.generated(200, 3e5, 19)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    sel = np.array([5, 17, 3, 150, 99, 42, 0, 7], dtype=np.int32)
    got = ts.genetic_relatedness_matrix(sample_sets=[[s] for s in sel])
    G = orc.genotypes().astype(np.int64)[:, sel]
    C = (G.T @ G).astype(np.float64)
    r = C.sum(axis=1)
    n = len(sel)
    exp = (C - r[:, None] / n - r[None, :] / n + r.sum() / n**2) / tc.sequence_length
    assert np.allclose(got, exp, rtol=1e-10, atol=1e-12 * np.max(np.abs(exp)))


@pytest.mark.parametrize("mode", ["branch", "site"])
def test_batch_shared_columns(small, mode):
    """A fused batch whose queries share sample sets (one state column each)
    gives the single-query results and the oracle's, including divergence
    between two indices holding the same samples (its denominator depends on
    the index, not the content) and a centred relatedness query, which keeps
    its own columns."""
    tc, ts, orc = small
    sets = cases.sample_sets(200, 4, 30, 11)
    same = [sets[0], sets[0][::-1].copy(), sets[1]]
    w = np.linspace(0, tc.sequence_length, 6)
    div_idx = [(0, 1), (0, 0), (1, 2)]
    b = ts.batch()
    p = [
        b.diversity(sets, windows=w, mode=mode),
        b.divergence(same, indexes=div_idx, windows=w, mode=mode),
        b.Fst(sets, indexes=[(0, 1), (2, 3)], windows=w, mode=mode),
        b.genetic_relatedness(sets, indexes=[(0, 1), (2, 2)], windows=w, mode=mode,
                              polarised=True, proportion=False),
        b.f4(sets, indexes=[(0, 1, 2, 3)], windows=w, mode=mode),
    ]
    res = b.run()
    _say("batch", mode)
    singles = [
        ts.diversity(sets, windows=w, mode=mode),
        ts.divergence(same, indexes=div_idx, windows=w, mode=mode),
        ts.Fst(sets, indexes=[(0, 1), (2, 3)], windows=w, mode=mode),
        ts.genetic_relatedness(sets, indexes=[(0, 1), (2, 2)], windows=w, mode=mode,
                               polarised=True, proportion=False),
        ts.f4(sets, indexes=[(0, 1, 2, 3)], windows=w, mode=mode),
    ]
    assert len(p) == len(res)
    for got, one in zip(res, singles):
        scale = np.maximum(np.max(np.abs(one)), 1e-300)
        assert np.max(np.abs(got - one)) <= 1e-6 * scale
    _check("diversity", res[0], orc.stat("diversity", sets, None, w, mode))
    _check("divergence", res[1], orc.stat("divergence", same, div_idx, w, mode))
    _check("genetic_relatedness", res[3],
           orc.stat("genetic_relatedness", sets, [(0, 1), (2, 2)], w, mode, True, True))
    _check("f4", res[4], orc.stat("f4", sets, [(0, 1, 2, 3)], w, mode))


def test_repeatable(small):
    """Two launches give bit-identical results (deterministic combine)."""
    tc, ts, orc = small
    sets = cases.sample_sets(200, 3, 30, 10)
    a = ts.diversity(sets, windows="trees", mode="branch")
    b = ts.diversity(sets, windows="trees", mode="branch")
    assert np.array_equal(a, b)


def _replicate_queries(b, tc, seed, wide=False):
    n = int(np.sum(tc.node_flags & 1))
    sets = cases.sample_sets(n, 25 if wide else 3, n // 25 if wide else min(20, n // 4), seed)
    w = np.linspace(0, tc.sequence_length, 4 + seed % 3)
    if wide:
        pairs = [(i, j) for i in range(25) for j in range(i + 1, 25)]  # 300 outputs
    else:
        pairs = [(0, 1), (1, 2)]
    out = []
    for mode in ("branch", "site"):
        out.append(b.diversity(sets, windows=w, mode=mode))
        out.append(b.divergence(sets, indexes=pairs, windows=w, mode=mode))
        if not wide:
            out.append(b.Tajimas_D(sets, windows=w, mode=mode))
            out.append(b.f3(sets, indexes=[(0, 1, 2)], windows=w, mode=mode))
    return out


def test_run_batches_multi(require_gpu):
    """Batches on several tree sequences run as one sweep launch per mode give
    bit-identical results to each batch run alone: replicates of different
    shapes, a gapped sequence, two batches on one tree sequence, a batch too
    wide for one sweep slice and one forced onto the full plan (both launch
    on their own inside the same call)."""
    import tskit_amd as ta

    tcs = [cases.generated(60 + 20 * i, 1e5 + 5e4 * i, 40 + i) for i in range(5)]
    tcs.append(cases.gapped(80, 2e5, 7))
    tss = [tc.tree_sequence() for tc in tcs]
    tss[4].ll_tree_sequence.set_sweep_plan(1)
    specs = [(i, i, False) for i in range(len(tss))] + [(0, 9, False), (2, 11, True)]
    alone = []
    for ti, seed, wide in specs:
        b = tss[ti].batch()
        p = _replicate_queries(b, tcs[ti], seed, wide)
        b.run()
        alone.append([x.result() for x in p])
    batches, pend = [], []
    for ti, seed, wide in specs:
        b = tss[ti].batch()
        pend.append(_replicate_queries(b, tcs[ti], seed, wide))
        batches.append(b)
    res = ta.run_batches(batches)
    _say("run_batches", len(batches))
    assert len(res) == len(specs)
    for k in range(len(specs)):
        assert len(res[k]) == len(alone[k])
        for got, exp, p in zip(res[k], alone[k], pend[k]):
            assert np.array_equal(np.asarray(got), np.asarray(exp), equal_nan=True), specs[k]
            assert np.array_equal(np.asarray(p.result()), np.asarray(exp), equal_nan=True)
    orc = Oracle(tcs[5])
    sets = cases.sample_sets(80, 3, 20, 5)
    w = np.linspace(0, tcs[5].sequence_length, 4 + 5 % 3)
    _check("diversity", res[5][0], orc.stat("diversity", sets, None, w, "branch"))
    _check("divergence", res[5][5], orc.stat("divergence", sets, [(0, 1), (1, 2)], w, "site"))
    with pytest.raises(RuntimeError):
        ta.run_batches(batches[:1])
    assert ta.run_batches([]) == []


def _node_case(name):
    if name == "generated":
        return cases.generated(120, 2e5, 23)
    if name == "gapped":
        return cases.gapped(120, 2e5, 23, count=30)
    return getattr(cases, name)()


@pytest.mark.parametrize("name", ["generated", "gapped", "paper_ex", "nonbinary_ex",
                                  "single_tree_ex"])
def test_node_mode(require_gpu, name):
    """Node statistics (trees.c:1788-1918) against the oracle, which is pinned
    to the reference's node outputs: one- and k-way summaries, polarised and
    not, span-normalised and not, several windows and none, and the composed
    Fst and Tajima's D; shapes as the reference's (windows, nodes, outputs)."""
    tc = _node_case(name)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    n = int(np.sum(tc.node_flags & 1))
    N = tc.num_nodes
    k = 3 if n >= 6 else 2
    size = max(1, min(25, n // k))
    sets = cases.sample_sets(n, k, size, 31)
    L = tc.sequence_length
    w = np.array([0, 0.23 * L, 0.5 * L, L])
    _say("node", name, N)
    pairs = [(0, 1), (1, 0)] if k == 2 else [(0, 1), (1, 2), (2, 2)]
    for sn in (True, False):
        got = ts.diversity(sets, windows=w, mode="node", span_normalise=sn)
        assert got.shape == (3, N, k)
        _check("diversity", got, orc.stat("diversity", sets, None, w, "node", sn))
        got = ts.divergence(sets, indexes=pairs, windows=w, mode="node", span_normalise=sn)
        assert got.shape == (3, N, len(pairs))
        _check("divergence", got, orc.stat("divergence", sets, pairs, w, "node", sn))
        got = ts.segregating_sites(sets, windows=w, mode="node", span_normalise=sn)
        _check("segregating_sites", got, orc.stat("segregating_sites", sets, None, w, "node", sn))
    got = ts.genetic_relatedness(sets, indexes=pairs, windows=w, mode="node", polarised=True,
                                 proportion=False)
    _check("genetic_relatedness", got,
           orc.stat("genetic_relatedness", sets, pairs, w, "node", True, True))
    got = ts.genetic_relatedness(sets, indexes=pairs, windows=w, mode="node", polarised=False,
                                 proportion=False, centre=False)
    _check("genetic_relatedness noncentred", got,
           orc.stat("genetic_relatedness", sets, pairs, w, "node", True, False, False))
    if k >= 3 and size >= 2:
        got = ts.f3(sets, indexes=[(0, 1, 2)], windows=w, mode="node")
        _check("f3", got, orc.stat("f3", sets, [(0, 1, 2)], w, "node"))
    # no windows: (nodes, outputs); a single set: (nodes,)
    got = ts.diversity(sets, mode="node")
    assert got.shape == (N, k)
    _check("diversity whole", got, orc.stat("diversity", sets, None, None, "node")[0])
    assert ts.diversity(sets[0], mode="node").shape == (N,)
    # composed statistics, per (window, node)
    div = orc.stat("diversity", sets, None, w, "node").reshape(-1, k)
    dvg = orc.stat("divergence", sets, pairs, w, "node").reshape(-1, len(pairs))
    got = ts.Fst(sets, indexes=pairs, windows=w, mode="node")
    exp = fst(div, dvg, pairs).reshape(3, N, len(pairs))
    # Fst is a ratio of sums near 1 - 1 at many nodes: compared with an
    # absolute floor of 1e-12 (its inputs are sums in another order)
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    assert np.allclose(got, exp, rtol=TOL, atol=1e-12, equal_nan=True)
    T = orc.stat("diversity", sets, None, w, "node", False).reshape(-1, k)
    S = orc.stat("segregating_sites", sets, None, w, "node", False).reshape(-1, k)
    got = ts.Tajimas_D(sets, windows=w, mode="node")
    _check("Tajimas_D", got, tajimas_d(T, S, [len(s) for s in sets]).reshape(3, N, k))


def test_node_mode_in_batch(small):
    """A node statistic recorded in a batch runs on its own after the fused
    sweeps: the same values as the direct call (node sums are FP64 atomics,
    so to rounding), and the batch's fused results are unchanged."""
    tc, ts, orc = small
    sets = cases.sample_sets(200, 2, 10, 1)
    w = np.linspace(0, tc.sequence_length, 4)
    b = ts.batch()
    p_node = b.diversity(sets, windows=w, mode="node")
    p_br = b.divergence(sets, windows=w, mode="branch")
    p_site = b.Tajimas_D(sets, windows=w, mode="site")
    b.run()
    node = ts.diversity(sets, windows=w, mode="node")
    assert p_node.result().shape == node.shape
    assert np.allclose(p_node.result(), node, rtol=1e-12, atol=1e-15)
    assert np.array_equal(p_br.result(), ts.divergence(sets, windows=w, mode="branch"))
    assert np.array_equal(p_site.result(), ts.Tajimas_D(sets, windows=w, mode="site"),
                          equal_nan=True)


@pytest.mark.parametrize("windows", [None, "three"])
def test_set_divergence_matrix(require_gpu, windows):
    """Site divergence matrix between disjoint sample sets of several samples
    (one a singleton): the pairwise matrix from the int8 cross products summed
    per pair of sets and count-normalised, against the closed form on the
    oracle's genotypes (pinned to the reference's tsk_treeseq_divergence_
    matrix) and the reference itself when oracle/_ref is present."""
    import test_oracle as to
    from oracle import oracle as om

    tc = cases.generated(300, 5e5, 17)
    ts = tc.tree_sequence()
    w = None if windows is None else np.array([0, 1e5, 3.5e5, 5e5])
    wn = np.array([0, 5e5]) if w is None else w
    rng = np.random.default_rng(9)
    perm = rng.permutation(300).astype(np.int32)
    sets = [perm[:40], perm[40:41], perm[41:140], perm[140:300]]
    _say("set divergence matrix", windows)
    got = ts.divergence_matrix(sample_sets=sets, windows=w)
    G = Oracle(tc).genotypes().astype(np.int64)
    exp = to._set_divergence(G, tc.site_position, wn, sets)
    if w is None:
        exp = exp[0]
    assert got.shape == exp.shape
    assert np.allclose(got, exp, rtol=1e-12, atol=0)
    if om.reference_available():
        ref = om.Reference(tc).divergence_matrix_sets(sets, wn, "site")
        assert np.allclose(got, ref if w is not None else ref[0], rtol=1e-12, atol=0)
    # a set's samples in another order give the same matrix
    again = ts.divergence_matrix(sample_sets=[s[::-1] for s in sets], windows=w)
    assert np.allclose(again, got, rtol=1e-13, atol=0)


@pytest.mark.parametrize("case", ["generated", "gapped"])
def test_branch_divergence_matrix(require_gpu, case):
    """Branch divergence matrix (trees.c:8579-8745) between sample sets and of
    singleton sets, over windows that span the sequence and windows that do
    not: against the shared-branch-length closed form on the oracle (pinned
    to the reference's MRCA route) and the reference itself when oracle/_ref
    is present."""
    import test_oracle as to
    from oracle import oracle as om

    tc = cases.generated(40, 2e5, 53) if case == "generated" else cases.gapped(40, 2e5, 53, 10)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    n = 40
    L = tc.sequence_length
    rng = np.random.default_rng(5)
    perm = rng.permutation(n).astype(np.int32)
    sets = [perm[:9], perm[9:10], perm[10:25], perm[25:40]]
    pairs = [(i, j) for i in range(n) for j in range(i, n)]
    single = [np.array([i], np.int32) for i in range(n)]
    fw = np.array([0, 2e4, 7e4, 1.5e5, L])
    G = orc.stat("genetic_relatedness", single, pairs, fw, "branch", False, True, False)
    exp_all = to._branch_divergence_from_shared(G, n, fw, sets)
    exp_one = to._branch_divergence_from_shared(G, n, fw, [[i] for i in range(n)])
    _say("branch divergence matrix", case)
    got = ts.divergence_matrix(sample_sets=sets, windows=fw, mode="branch")
    assert np.allclose(got, exp_all, rtol=1e-9, atol=0)
    part = np.array([2e4, 7e4, 1.5e5])  # inner windows only
    got = ts.divergence_matrix(sample_sets=sets, windows=part, mode="branch")
    assert np.allclose(got, exp_all[1:3], rtol=1e-9, atol=0)
    got = ts.divergence_matrix(windows=part, mode="branch")
    assert np.allclose(got, exp_one[1:3], rtol=1e-9, atol=1e-300)
    whole = ts.divergence_matrix(sample_sets=sets, mode="branch")
    exp_whole = to._branch_divergence_from_shared(
        orc.stat("genetic_relatedness", single, pairs, None, "branch", False, True, False),
        n, np.array([0, L]), sets)[0]
    assert np.allclose(whole, exp_whole, rtol=1e-9, atol=0)
    if om.reference_available():
        ref = om.Reference(tc)
        assert np.allclose(ts.divergence_matrix(sample_sets=sets, windows=part, mode="branch"),
                           ref.divergence_matrix_sets(sets, part, "branch"), rtol=1e-9, atol=0)


def _relatedness_from_divergence(D, n):
    """trees.py:9080-9102 restated: self comparisons into the diagonal, then
    double centring."""
    B = D.copy()
    np.fill_diagonal(B, np.diag(B) * (n - 1) / n)
    K = B + np.mean(B)
    y = np.mean(B, axis=0)
    K -= y[:, None] + y[None, :]
    return K / -2


def test_relatedness_matrix_branch_and_sets(require_gpu):
    """genetic_relatedness_matrix in branch mode and over sets of several
    samples: the reference's composition of the divergence matrix, which is
    checked against the closed forms pinned to the reference."""
    import test_oracle as to

    tc = cases.generated(40, 2e5, 53)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    n = 40
    L = tc.sequence_length
    rng = np.random.default_rng(5)
    perm = rng.permutation(n).astype(np.int32)
    sets = [perm[:9], perm[9:10], perm[10:25], perm[25:40]]
    sizes = np.array([len(s) for s in sets])
    fw = np.array([0, 7e4, L])
    pairs = [(i, j) for i in range(n) for j in range(i, n)]
    single = [np.array([i], np.int32) for i in range(n)]
    G = orc.stat("genetic_relatedness", single, pairs, fw, "branch", False, True, False)
    Db = to._branch_divergence_from_shared(G, n, fw, sets)
    got = ts.genetic_relatedness_matrix(sample_sets=sets, windows=fw, mode="branch")
    exp = np.array([_relatedness_from_divergence(Db[j], sizes) for j in range(2)])
    assert np.allclose(got, exp, rtol=1e-9, atol=1e-12 * np.max(np.abs(exp)))
    Ds = to._set_divergence(orc.genotypes().astype(np.int64), tc.site_position, fw, sets)
    got = ts.genetic_relatedness_matrix(sample_sets=sets, windows=fw, mode="site")
    exp = np.array([_relatedness_from_divergence(Ds[j], sizes) for j in range(2)])
    assert np.allclose(got, exp, rtol=1e-9, atol=1e-12 * np.max(np.abs(exp)))
    got = ts.genetic_relatedness_matrix(mode="branch")
    D1 = to._branch_divergence_from_shared(
        orc.stat("genetic_relatedness", single, pairs, None, "branch", False, True, False),
        n, np.array([0, L]), [[i] for i in range(n)])[0]
    exp = _relatedness_from_divergence(D1, np.ones(n))
    assert np.allclose(got, exp, rtol=1e-9, atol=1e-12 * np.max(np.abs(exp)))




@pytest.mark.parametrize("case", ["generated", "gapped", "paper_ex", "wide"])
def test_relatedness_vector(require_gpu, case):
    """genetic_relatedness_vector (trees.c:10402-10816) on the hot plan
    against the compiled reference when oracle/_ref is present and the
    oracle (pinned to it and to the closed form G w): windows spanning the
    sequence and not, centred and not, span-normalised and not, a subset of
    samples as focal nodes, 20 weight columns (two column groups), one and
    several chunks.  Tolerance: <= 1e-6 of the largest |value| of each
    window (the result is centred, so it changes sign)."""
    from oracle import oracle as om

    tc = {"generated": lambda: cases.generated(150, 3e5, 41),
          "gapped": lambda: cases.gapped(120, 2e5, 43, count=20),
          "paper_ex": cases.paper_ex,
          "wide": lambda: cases.generated(3000, 1e6, 13)}[case]()
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    ref = om.Reference(tc) if om.reference_available() else None
    n = ts.num_samples
    L = tc.sequence_length
    rng = np.random.default_rng(6)
    smp = ts.samples()

    def check(got, W, win, sn, centre, nodes=None):
        for impl in [orc] + ([ref] if ref is not None else []):
            exp = impl.relatedness_vector(W, win, "branch", sn, centre, nodes)
            if win is None:
                exp = exp[0]
            assert got.shape == exp.shape
            scale = np.max(np.abs(exp), axis=tuple(range(1, exp.ndim)) if win is not None
                           else None, keepdims=True)
            err = rel_err(got, exp, np.maximum(scale, 1e-300))
            assert err <= TOL, (type(impl).__name__, err)

    for chunks in ([None, 1] if case != "wide" else [None]):
        if chunks:
            ts.ll_tree_sequence.set_num_chunks(chunks)
        for win in [None, np.linspace(0, L, 6), np.array([0.05 * L, 0.4 * L, 0.9 * L])]:
            W = rng.normal(size=(n, 3))
            _say("relatedness vector", case, chunks, None if win is None else len(win))
            for sn in (True, False):
                for centre in (True, False):
                    got = ts.genetic_relatedness_vector(W, windows=win, mode="branch",
                                                        span_normalise=sn, centre=centre)
                    check(got, W, win, sn, centre)
        sub = smp[::-1][: max(2, n // 3)]
        got = ts.genetic_relatedness_vector(W, windows=np.linspace(0, L, 3), mode="branch",
                                            centre=False, nodes=sub)
        check(got, W, np.linspace(0, L, 3), True, False, sub)
    W20 = rng.normal(size=(n, 20))
    _say("relatedness vector 20 columns", case)
    got = ts.genetic_relatedness_vector(W20, windows=np.linspace(0, L, 4), mode="branch")
    check(got, W20, np.linspace(0, L, 4), True, True)


@pytest.mark.parametrize("case", ["multiallelic", "single_tree_ex", "wide"])
@pytest.mark.parametrize("plan", ["hot", "full"])
def test_multiallelic_matrices(require_gpu, case, plan):
    """Site divergence and relatedness matrices with back, recurrent and
    stacked mutations (2 to 4 alleles per site): the extra allele rows go
    into a second int8 cross product.  Against the genotype closed form on
    the oracle's genotypes (pinned to tsk_variant_decode) and the reference's
    tsk_treeseq_divergence_matrix when oracle/_ref is present; the
    relatedness matrix against the reference's composition of the divergence
    matrix (trees.py:9080-9102).  Integer cross products: rtol 1e-12."""
    import test_oracle as to
    from oracle import oracle as om

    tc = {"multiallelic": lambda: cases.multiallelic(150, 3e5, 5),
          "single_tree_ex": cases.single_tree_ex,
          "wide": lambda: cases.multiallelic(3000, 2e5, 9, every=2)}[case]()
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    G = Oracle(tc).genotype_matrix(False)
    n = G.shape[1]
    L = tc.sequence_length
    w = np.array([0, 0.4 * L, L])
    _say("multiallelic matrices", case, plan)
    ll.set_sweep_plan(1 if plan == "full" else 0)
    try:
        got = ts.divergence_matrix(windows=w)
        exp = to._allele_divergence(G, tc.site_position, w) / np.diff(w)[:, None, None]
        assert np.allclose(got, exp, rtol=1e-12, atol=0)
        if om.reference_available() and n <= 300:
            ref = om.Reference(tc).divergence_matrix(np.arange(n, dtype=np.int32), w, "site")
            assert np.allclose(got, ref, rtol=1e-12, atol=0)
        if n >= 4:
            sets = [np.arange(0, n // 2, dtype=np.int32),
                    np.arange(n // 2, n - 1, dtype=np.int32), np.array([n - 1], np.int32)]
            got = ts.divergence_matrix(sample_sets=sets, windows=w)
            exp = to._allele_divergence(G, tc.site_position, w, [list(s) for s in sets])
            assert np.allclose(got, exp / np.diff(w)[:, None, None], rtol=1e-12, atol=0)
        K = ts.genetic_relatedness_matrix(windows=w)
        D = to._allele_divergence(G, tc.site_position, w) / np.diff(w)[:, None, None]
        expK = np.array([_relatedness_from_divergence(D[j], np.ones(n)) for j in range(2)])
        assert np.allclose(K, expK, rtol=1e-9, atol=1e-12 * np.max(np.abs(expK)))
    finally:
        ll.set_sweep_plan(0)


def test_branch_divergence_matrix_wide(require_gpu):
    """The branch divergence matrix of 3000 singleton samples (4.5e6 sample
    pairs, beyond the fused-sweep route's 2^22): symmetric with a zero
    diagonal, and a block of 30 samples equal to the shared-branch-length
    closed form on the oracle (pinned to the reference's MRCA route), over
    windows that do not span the sequence."""
    import test_oracle as to

    tc = cases.generated(3000, 1e5, 61)
    ts = tc.tree_sequence()
    orc = Oracle(tc)
    L = tc.sequence_length
    w = np.array([0.1 * L, 0.6 * L, 0.95 * L])
    _say("branch divergence matrix wide")
    D = ts.divergence_matrix(windows=w, mode="branch")
    assert D.shape == (2, 3000, 3000)
    assert np.array_equal(D, np.transpose(D, (0, 2, 1)))
    assert np.all(D[:, np.arange(3000), np.arange(3000)] == 0)
    pick = np.sort(np.random.default_rng(4).choice(3000, 30, replace=False))
    single = [np.array([i], np.int32) for i in pick]
    pairs = [(i, j) for i in range(30) for j in range(i, 30)]
    fw = np.concatenate([[0.0], w, [L]])
    G = orc.stat("genetic_relatedness", single, pairs, fw, "branch", False, True, False)
    exp = to._branch_divergence_from_shared(G, 30, fw, [[i] for i in range(30)])[1:3]
    got = D[:, pick][:, :, pick]
    assert np.allclose(got, exp, rtol=1e-9, atol=1e-12 * np.max(np.abs(exp)))


def test_autotune(require_gpu, tmp_path):
    """The chunk-count autotuner times its candidates on a tree sequence,
    records the fastest per shape in its JSON cache and applies it to
    another tree sequence of the same shape; statistics on the tuned plan
    equal the oracle's."""
    import json

    from tskit_amd import autotune

    path = str(tmp_path / "tune.json")
    tc = cases.generated(300, 3e5, 71)
    ts = tc.tree_sequence()
    _say("autotune")
    got = autotune.tune(ts, candidates=(64, 256), reps=2, path=path)
    assert got["events_per_chunk"] in (64, 256)
    assert ts.ll_tree_sequence.get_hot_plan_sizes()["chunks"] == got["chunks"]
    table = json.load(open(path))
    assert autotune.shape_key(ts) in table
    ts2 = cases.generated(300, 3e5, 71).tree_sequence()
    assert autotune.apply(ts2, path=path) == got["events_per_chunk"]
    assert ts2.ll_tree_sequence.get_hot_plan_sizes()["chunks"] == got["chunks"]
    orc = Oracle(tc)
    sets = cases.sample_sets(300, 3, 40, 2)
    w = np.linspace(0, 3e5, 9)
    for mode in ("branch", "site"):
        _check("diversity", ts2.diversity(sets, windows=w, mode=mode),
               orc.stat("diversity", sets, None, w, mode))
    # one setting for tree sequences swept together (run_batches)
    import tskit_amd as ta

    def batches():
        out = []
        for t in (ts, ts2):
            b = t.batch()
            b.diversity(sets, windows=w, mode="branch")
            out.append(b)
        return out

    joint = autotune.tune_batch([ts, ts2], lambda: ta.run_batches(batches()),
                                candidates=(64, 256), reps=1)
    assert joint["events_per_chunk"] in (64, 256)
    res = ta.run_batches(batches())
    exp = orc.stat("diversity", sets, None, w, "branch")
    for r in res:
        _check("diversity", r[0], exp)


@pytest.mark.parametrize("case", ["generated", "gapped", "multiallelic", "paper_ex"])
def test_weighted_stats(require_gpu, case):
    """general_stat of real-valued weights (FP64 state columns on the hot
    plan): trait covariance, correlation and linear model, weighted genetic
    relatedness centred and not, both modes and polarisations, windows
    spanning the sequence and in pieces, one and several chunks, against the
    oracle (pinned to the compiled reference and to closed forms) and the
    compiled reference when present.  Tolerance: <= 1e-9 of the largest
    |value| of each window (the traits are centred: values change sign).
    A repeated call is bit-identical."""
    from oracle import oracle as om

    tc = {"generated": lambda: cases.generated(150, 3e5, 41),
          "gapped": lambda: cases.gapped(120, 2e5, 43, count=20),
          "multiallelic": lambda: cases.multiallelic(100, 2e5, 5),
          "paper_ex": cases.paper_ex}[case]()
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    orc = Oracle(tc)
    ref = om.Reference(tc) if om.reference_available() else None
    n = ts.num_samples
    L = tc.sequence_length
    rng = np.random.default_rng(9)
    W = rng.normal(size=(n, 3))
    Z = rng.normal(size=(n, 2))
    calls = [("trait_covariance", {}, lambda w, m, p: ll.trait_covariance(W, w, m, p)),
             ("trait_correlation", {}, lambda w, m, p: ll.trait_correlation(W, w, m, p)),
             ("trait_linear_model", {"Z": Z},
              lambda w, m, p: ll.trait_linear_model(W, Z, w, m, p)),
             ("genetic_relatedness_weighted", {"indexes": [(0, 1), (1, 2), (2, 2)]},
              lambda w, m, p: ll.genetic_relatedness_weighted(W, [(0, 1), (1, 2), (2, 2)], w,
                                                              m, 1, p, 1)),
             ("genetic_relatedness_weighted", {"indexes": [(0, 2)], "centre": False},
              lambda w, m, p: ll.genetic_relatedness_weighted(W, [(0, 2)], w, m, 1, p, 0))]
    for chunks in (None, 1):
        if chunks:
            ll.set_num_chunks(chunks)
        for w in [np.array([0.0, L]), np.array([0, 0.3 * L, 0.35 * L, L])]:
            for mode in ("branch", "site"):
                for pol in (False, True):
                    _say("weighted", case, chunks, len(w), mode, pol)
                    for name, kw, gpu in calls:
                        got = gpu(w, mode, pol)
                        for impl in [orc] + ([ref] if ref is not None else []):
                            exp = impl.weighted_stat(name, W, w, mode, True, pol, **kw)
                            scale = np.max(np.abs(exp), axis=1, keepdims=True)
                            err = rel_err(got, exp, np.maximum(scale, 1e-300))
                            assert err <= 1e-9, (name, type(impl).__name__, err)
                        assert np.array_equal(got, gpu(w, mode, pol)), name


def test_general_stat_registry(require_gpu):
    """TreeSequence.general_stat with a registry summary equals the
    statistic it defines (trait covariance of centred weights), the strict
    check rejects a summary that is not zero at the total weight, a Python
    callable is refused, node mode is not taken, and the Python wrappers
    drop dimensions as the reference's."""
    import tskit_amd as ta
    from tskit_amd import summaries

    tc = cases.generated(120, 2e5, 17)
    ts = tc.tree_sequence()
    n = ts.num_samples
    rng = np.random.default_rng(4)
    W = rng.normal(size=(n, 2))
    Wc = W - W.mean(axis=0)
    w = np.linspace(0, tc.sequence_length, 4)
    for mode in ("branch", "site"):
        a = ts.general_stat(Wc, summaries.trait_covariance(n), 2, windows=w, mode=mode)
        b = ts.trait_covariance(W, windows=w, mode=mode)
        assert a.shape == b.shape == (3, 2)
        assert rel_err(a, b, np.max(np.abs(b))) <= 1e-12
    assert ts.trait_covariance(W, mode="branch").shape == (2,)
    assert ts.genetic_relatedness_weighted(W, mode="branch").shape == ()
    assert ts.genetic_relatedness_weighted(W, indexes=[(0, 1), (1, 1)],
                                           mode="branch").shape == (2,)
    assert ts.trait_linear_model(W, mode="site").shape == (2,)
    with pytest.raises(ValueError):
        ts.general_stat(W, summaries.trait_covariance(n), 2)  # not centred: f(total) != 0
    with pytest.raises(TypeError):
        ts.general_stat(Wc, lambda x: x, 2)
    with pytest.raises(ta.LibraryError):
        ts.general_stat(Wc, summaries.trait_covariance(n), 2, mode="node")
    with pytest.raises(ta.LibraryError):
        ts.ll_tree_sequence.trait_covariance(np.zeros((n, 0)), [0, tc.sequence_length])


def test_device_results(require_gpu):
    """A result buffer in device memory is filled on the device (no copy to
    the host): a direct call, a batch and a multi-batch run write the same
    values into torch device tensors as into host buffers."""
    import ctypes

    import torch

    from tskit_amd._lib import as_c, ptr

    tc = cases.generated(150, 2e5, 23)
    ts = tc.tree_sequence()
    ll = ts.ll_tree_sequence
    lib, h = ll._lib, ll._h
    sets = cases.sample_sets(150, 3, 20, 8)
    sizes = as_c(np.array([len(s) for s in sets], dtype=np.uint64), np.uint64)
    flat = as_c(np.concatenate(sets).astype(np.int32), np.int32)
    w = as_c(np.linspace(0, tc.sequence_length, 6), np.float64)
    c_f64p = ctypes.POINTER(ctypes.c_double)
    BRANCH, SITE, SPAN = 1 << 1, 1 << 0, 1 << 11
    for opts in (BRANCH | SPAN, SITE | SPAN):
        host = np.zeros((5, 3))
        assert lib.tsg_treeseq_diversity(h, 3, ptr(sizes, ctypes.c_uint64),
                                         ptr(flat, ctypes.c_int32), 5, ptr(w, ctypes.c_double),
                                         opts, ptr(host, ctypes.c_double)) == 0
        dev = torch.full((5, 3), -1.0, dtype=torch.float64, device="cuda")
        _say("device result", opts)
        assert lib.tsg_treeseq_diversity(h, 3, ptr(sizes, ctypes.c_uint64),
                                         ptr(flat, ctypes.c_int32), 5, ptr(w, ctypes.c_double),
                                         opts, ctypes.cast(dev.data_ptr(), c_f64p)) == 0
        assert np.array_equal(dev.cpu().numpy(), host)
        # in a batch (one device, one host result), and two batches run together
        devs = [torch.full((5, 3), -1.0, dtype=torch.float64, device="cuda") for _ in range(2)]
        hosts = [np.zeros((5, 3)) for _ in range(2)]
        bs = []
        for k in range(2):
            b = ctypes.c_void_p()
            assert lib.tsg_batch_init(ctypes.byref(b), h) == 0
            for out in (ctypes.cast(devs[k].data_ptr(), c_f64p), ptr(hosts[k], ctypes.c_double)):
                assert lib.tsg_batch_diversity(b, 3, ptr(sizes, ctypes.c_uint64),
                                               ptr(flat, ctypes.c_int32), 5,
                                               ptr(w, ctypes.c_double), opts, out) == 0
            bs.append(b)
        assert lib.tsg_batch_run(bs[0]) == 0
        handles = (ctypes.c_void_p * 1)(bs[1].value)
        assert lib.tsg_batch_run_multi(handles, 1) == 0
        for k in range(2):
            assert np.array_equal(hosts[k], host)
            assert np.array_equal(devs[k].cpu().numpy(), host)
            lib.tsg_batch_free(bs[k])


@pytest.mark.parametrize("K", [3, 6, 12, 24, 40])
def test_column_layouts(small, K):
    """Every compile-time count-update layout (hop groups per column 8, 4, 2
    and 1 in branch mode; 16, 8, 4, 2, 1 in site mode): diversity and
    divergence of K sets against the oracle, and a query on two of the sets
    gives the same bits alone as fused with the others."""
    tc, ts, orc = small
    sets = cases.sample_sets(200, K, 4, 40 + K)
    w = np.linspace(0, tc.sequence_length, 4)
    idx = [(0, 1), (K - 2, K - 1)]
    for mode in ["branch", "site"]:
        _say("layout", K, mode)
        _check("diversity", ts.diversity(sets, windows=w, mode=mode),
               orc.stat("diversity", sets, None, w, mode))
        _check("divergence", ts.divergence(sets, indexes=idx, windows=w, mode=mode),
               orc.stat("divergence", sets, idx, w, mode))
        alone = ts.divergence(sets[:2], windows=w, mode=mode)
        b = ts.batch()
        p = b.divergence(sets[:2], windows=w, mode=mode)
        b.diversity(sets, windows=w, mode=mode)
        b.run()
        assert np.array_equal(p.result(), alone)


@pytest.mark.parametrize("same_sets", [True, False])
def test_two_mode_batch_seed_sharing(require_gpu, same_sets):
    """A batch of branch and site queries on the same sample sets seeds the
    site sweep's start counts from the branch sweep's (integer counts, so
    the site results are the same bits as a site-only batch); with other
    sets per mode each run seeds its own."""
    tc = cases.generated(3000, 1e6, 13)
    ts = tc.tree_sequence()
    sets = cases.sample_sets(3000, 6, 200, 21)
    site_sets = sets if same_sets else cases.sample_sets(3000, 6, 200, 22)
    w = np.linspace(0, tc.sequence_length, 6)
    idx = [(0, 1), (2, 5)]

    def queries(b, mode, ss):
        return [b.diversity(ss, windows=w, mode=mode),
                b.divergence(ss, indexes=idx, windows=w, mode=mode),
                b.Tajimas_D(ss, windows=w, mode=mode),
                b.f4(ss[:4], windows=w, mode=mode)]

    b = ts.batch()
    both_branch = queries(b, "branch", sets)
    both_site = queries(b, "site", site_sets)
    b.run()
    for mode, ss, got in (("branch", sets, both_branch), ("site", site_sets, both_site)):
        one = ts.batch()
        alone = queries(one, mode, ss)
        one.run()
        _say("seed sharing", same_sets, mode)
        for g, a in zip(got, alone):
            assert np.array_equal(g.result(), a.result(), equal_nan=True)
    orc = Oracle(tc)
    _check("diversity", both_site[0].result(), orc.stat("diversity", site_sets, None, w, "site"))


@pytest.mark.parametrize("mode", ["branch", "site"])
def test_fused_results_bit_identical(small, mode):
    """A query's results are the same bits alone and fused with unrelated
    queries into one sweep with more state columns (K = 2 against K = 13):
    the path sums add each hop's terms in hop order whatever the lane
    layout, which depends on K.  Covers the closed forms (diversity, f4,
    non-centred relatedness, polarised Y1), the segregating-sites form
    (both polarisations) and the hop-by-hop form (polarised f2)."""
    tc, ts, orc = small
    ll = ts.ll_tree_sequence
    sets = cases.sample_sets(200, 13, 12, 31)
    two = sets[:2]
    w = np.linspace(0, tc.sequence_length, 5)
    sizes2 = np.array([len(s) for s in two], dtype=np.uint64)
    flat2 = np.concatenate(two).astype(np.int32)

    alone = [ts.diversity(two, windows=w, mode=mode),
             ts.segregating_sites(two, windows=w, mode=mode),
             ts.f4([two[0], two[1], two[0], two[1]], windows=w, mode=mode),
             ts.genetic_relatedness(two, windows=w, mode=mode, centre=False, proportion=False),
             ll.Y1(sizes2, flat2, w, mode=mode, span_normalise=1, polarised=1),
             ll.segregating_sites(sizes2, flat2, w, mode=mode, span_normalise=1, polarised=1),
             ll.f2(sizes2, flat2, np.array([(0, 1)], dtype=np.int32), w, mode=mode,
                   span_normalise=1, polarised=1)]
    b = ts.batch()
    p = [b.diversity(two, windows=w, mode=mode),
         b.segregating_sites(two, windows=w, mode=mode),
         b.f4([two[0], two[1], two[0], two[1]], windows=w, mode=mode),
         b.genetic_relatedness(two, windows=w, mode=mode, centre=False, proportion=False),
         b.diversity(sets[2:], windows=w, mode=mode),
         b.divergence(sets[2:], indexes=[(0, 1), (3, 7)], windows=w, mode=mode)]
    b.run()
    _say("fused bit-identical", mode)
    for k in range(4):
        assert np.array_equal(p[k].result(), alone[k], equal_nan=True), k
    # the polarised ones through a batch of the low-level calls
    bl = ts.batch()
    bl.diversity(sets[2:], windows=w, mode=mode)
    bll = bl._ll
    r1 = bll.Y1(sizes2, flat2, w, mode=mode, span_normalise=1, polarised=1)
    r2 = bll.segregating_sites(sizes2, flat2, w, mode=mode, span_normalise=1, polarised=1)
    r3 = bll.f2(sizes2, flat2, np.array([(0, 1)], dtype=np.int32), w, mode=mode,
                span_normalise=1, polarised=1)
    bl.run()
    for k, r in enumerate((r1, r2, r3)):
        assert np.array_equal(r, alone[4 + k], equal_nan=True), 4 + k
