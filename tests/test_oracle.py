"""Pins the oracle (oracle/tsk_oracle.c, the FP64 CPU restatement) so the GPU
parity tests can trust it:

  * against golden vectors written by the reference C library itself
    (tests/golden/config1_reference.npz, see make_golden.py);
  * against the reference C library live, where oracle/_ref was built;
  * against brute-force definitions on tiny trees (pairwise path lengths,
    genotype pairs, textbook Tajima's D and Fst from pairwise differences),
    which pin the mathematics independently of the reference's sweep.
"""
import itertools
import os

import numpy as np
import pytest

import cases
from oracle import oracle as om
from oracle.oracle import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "config1_reference.npz")


@pytest.fixture(scope="module")
def golden():
    g = np.load(GOLDEN)
    tc = cases.config1()
    digest = bytes(g["digest"]).decode()
    assert tc.column_digest() == digest, "generator output changed"
    return g, tc, Oracle(tc)


def _sets(flat, k):
    return list(np.asarray(flat).reshape(k, -1))


def test_golden_trees(golden):
    g, tc, orc = golden
    assert orc.num_trees == int(g["num_trees"][0])
    assert np.array_equal(orc.breakpoints(), g["breakpoints"])
    I, O = orc.index()
    assert np.array_equal(I, g["index_I"]) and np.array_equal(O, g["index_O"])


@pytest.mark.parametrize("mode", ["branch", "site"])
def test_golden_stats(golden, mode):
    g, tc, orc = golden
    s2, s4 = _sets(g["sets2"], 2), _sets(g["sets4"], 4)
    w = g["windows"]
    checks = [
        ("diversity", s2, None, True, False, True, "diversity"),
        ("divergence", s2, [(0, 1)], True, False, True, "divergence"),
        ("segregating_sites", s2, None, True, False, True, "segsites"),
        ("f4", s4, [(0, 1, 2, 3), (0, 2, 1, 3)], True, False, True, "f4"),
        ("f2", s4, [(0, 1), (2, 3)], True, False, True, "f2"),
        ("genetic_relatedness", s4, [(0, 1), (2, 2)], True, True, True, "relatedness"),
        ("diversity", s2, None, False, False, True, "diversity_raw"),
        ("segregating_sites", s2, None, False, False, True, "segsites_raw"),
    ]
    for name, sets, idx, sn, pol, centre, key in checks:
        got = orc.stat(name, sets, idx, w, mode, sn, pol, centre)
        exp = g[f"{key}_{mode}"]
        assert np.allclose(got, exp, rtol=1e-12, atol=0), key


@pytest.mark.parametrize("mode", ["branch", "site"])
def test_golden_afs(golden, mode):
    g, tc, orc = golden
    sets = _sets(g["afs_sets"], 2)
    w = np.linspace(0, 1e6, 5)
    for pol, key in [(False, "afs"), (True, "afs_polarised")]:
        got = orc.afs(sets, w, None, mode, True, pol)
        exp = g[f"{key}_{mode}"]
        if mode == "site":
            assert np.array_equal(got, exp)
        else:
            assert np.allclose(got, exp, rtol=1e-12, atol=1e-300)
    got = orc.afs(sets, np.linspace(0, 1e6, 3), np.array([0, 5000, 20000, np.inf]), "branch")
    assert np.allclose(got, g["afs_tw_branch"], rtol=1e-12, atol=1e-300)


# ---- the reference, live ----------------------------------------------------
needs_ref = pytest.mark.skipif(not om.reference_available(), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("seed", [21, 22])
def test_oracle_vs_reference_generated(seed):
    tc = cases.generated(80, 1e5, seed)
    orc, ref = Oracle(tc), om.Reference(tc)
    sets = cases.sample_sets(80, 4, 15, seed)
    w = np.linspace(0, 1e5, 4)
    for mode in ["branch", "site"]:
        for name in ["diversity", "segregating_sites", "Y1"]:
            for pol in [False, True]:
                assert np.allclose(orc.stat(name, sets, None, w, mode, True, pol),
                                   ref.stat(name, sets, None, w, mode, True, pol),
                                   rtol=1e-13, atol=0)
        for name, idx in [("divergence", [(0, 1), (2, 2)]), ("Y2", [(1, 3)]), ("f2", [(0, 3)]),
                          ("Y3", [(0, 1, 2)]), ("f3", [(2, 1, 0)]), ("f4", [(0, 1, 2, 3)]),
                          ("genetic_relatedness", [(0, 1), (3, 3)])]:
            for pol in [False, True]:
                for centre in ([True, False] if name == "genetic_relatedness" else [True]):
                    a = orc.stat(name, sets, idx, w, mode, True, pol, centre)
                    b = ref.stat(name, sets, idx, w, mode, True, pol, centre)
                    assert np.allclose(a, b, rtol=1e-12, atol=1e-15 * np.max(np.abs(b))), name
        for pol in [False, True]:
            assert np.allclose(orc.afs(sets[:3], w, None, mode, True, pol),
                               ref.afs(sets[:3], w, None, mode, True, pol), rtol=1e-12, atol=0)


@needs_ref
@pytest.mark.parametrize("fixture", ["paper_ex", "single_tree_ex", "nonbinary_ex"])
def test_oracle_vs_reference_fixtures(fixture):
    tc = getattr(cases, fixture)()
    orc, ref = Oracle(tc), om.Reference(tc)
    n = int(np.sum(tc.node_flags & 1))
    sets = [list(range(n // 2)), list(range(n // 2, n))]
    for mode in ["branch", "site"]:
        for pol in [False, True]:
            for sn in [False, True]:
                assert np.allclose(orc.stat("diversity", sets, None, None, mode, sn, pol),
                                   ref.stat("diversity", sets, None, None, mode, sn, pol))
                assert np.allclose(orc.stat("divergence", sets, [(0, 1)], None, mode, sn, pol),
                                   ref.stat("divergence", sets, [(0, 1)], None, mode, sn, pol))
            assert np.array_equal(orc.afs(sets, None, None, mode, True, pol),
                                  ref.afs(sets, None, None, mode, True, pol))


@needs_ref
def test_oracle_error_numbers_match_reference():
    tc = cases.paper_ex()
    orc, ref = Oracle(tc), om.Reference(tc)
    bad = [
        (("diversity", [[0, 1], []], None, None), {}),
        (("diversity", [[0, 100]], None, None), {}),
        (("diversity", [[0, 5]], None, None), {}),
        (("diversity", [[0, 0]], None, None), {}),
        (("diversity", [[0, 1]], None, np.array([0, 5.0])), {}),
        (("divergence", [[0, 1], [2, 3]], [(0, 2)], None), {}),
    ]
    for (name, sets, idx, w), kw in bad:
        codes = []
        for impl in (orc, ref):
            try:
                impl.stat(name, sets, idx, w)
                codes.append(0)
            except om.OracleError as e:
                codes.append(e.err)
        assert codes[0] == codes[1] != 0, (name, codes)


# ---- brute force on tiny trees ----------------------------------------------
def _trees(tc):
    """[(left, right, parent array)] for every tree."""
    bps = sorted(set(tc.edge_left) | set(tc.edge_right) | {0.0, tc.sequence_length})
    out = []
    for a, b in zip(bps[:-1], bps[1:]):
        par = np.full(tc.num_nodes, -1)
        on = (tc.edge_left <= a) & (tc.edge_right >= b)
        par[tc.edge_child[on]] = tc.edge_parent[on]
        out.append((a, b, par))
    return out


def _path(par, u):
    p = [u]
    while par[p[-1]] != -1:
        p.append(par[p[-1]])
    return p


def _branch_distance(tc, par, u, v):
    pu, pv = _path(par, u), _path(par, v)
    common = set(pu) & set(pv)
    t = tc.node_time
    mrca = min(common, key=lambda x: t[x]) if common else None
    if mrca is None:
        return np.inf
    return (t[mrca] - t[u]) + (t[mrca] - t[v])


def _genotypes(tc):
    """Derived-allele indicator per (site, sample) for 0/1 single-mutation sites."""
    trees = _trees(tc)
    n = int(np.sum(tc.node_flags & 1))
    G = np.zeros((tc.num_sites, n), dtype=int)
    for s, x in enumerate(tc.site_position):
        par = next(p for a, b, p in trees if a <= x < b)
        for m in np.where(tc.mutation_site == s)[0]:
            node = tc.mutation_node[m]
            for j in range(n):
                if node in _path(par, j):
                    G[s, j] = 1
    return G


@pytest.fixture(scope="module")
def tiny():
    tc = cases.generated(8, 3e4, 31, mu=3e-8)
    assert tc.num_sites > 3 and len(_trees(tc)) > 2
    return tc, Oracle(tc)


def test_brute_branch_diversity_divergence(tiny):
    tc, orc = tiny
    A, B = [0, 1, 2, 3], [4, 5, 6, 7]
    L = tc.sequence_length
    div_a = div_ab = 0.0
    for a, b, par in _trees(tc):
        span = b - a
        div_a += span * np.mean([_branch_distance(tc, par, u, v)
                                 for u, v in itertools.combinations(A, 2)])
        div_ab += span * np.mean([_branch_distance(tc, par, u, v) for u in A for v in B])
    got = orc.stat("diversity", [A], None, None, "branch", True)[0, 0]
    assert np.isclose(got, div_a / L, rtol=1e-12)
    got = orc.stat("divergence", [A, B], [(0, 1)], None, "branch", True)[0, 0]
    assert np.isclose(got, div_ab / L, rtol=1e-12)


def test_brute_site_stats(tiny):
    tc, orc = tiny
    G = _genotypes(tc)
    A, B = [0, 1, 2], [3, 4, 5, 6]
    pi = sum(np.mean([G[s, u] != G[s, v] for u, v in itertools.combinations(A, 2)])
             for s in range(G.shape[0]))
    assert np.isclose(orc.stat("diversity", [A], None, None, "site", False)[0, 0], pi)
    dxy = sum(np.mean([G[s, u] != G[s, v] for u in A for v in B]) for s in range(G.shape[0]))
    assert np.isclose(orc.stat("divergence", [A, B], [(0, 1)], None, "site", False)[0, 0], dxy)
    seg = sum(0 < G[s, A].sum() < len(A) for s in range(G.shape[0]))
    assert np.isclose(orc.stat("segregating_sites", [A], None, None, "site", False)[0, 0], seg)
    # f4(A,B;C,D) = mean over quartets of (g_a - g_b)(g_c - g_d) for 0/1 sites
    C, D = [6, 7], [0, 7]
    f4 = sum(np.mean([(G[s, a] - G[s, b]) * (G[s, c] - G[s, d])
                      for a in A for b in B for c in C for d in D]) for s in range(G.shape[0]))
    got = orc.stat("f4", [A, B, C, D], [(0, 1, 2, 3)], None, "site", False)[0, 0]
    assert np.isclose(got, f4, atol=1e-12)


def test_brute_tajimas_d(tiny):
    """Tajima (1989) eq. 38 written with a1, a2, b1, b2, c1, c2, e1, e2 -- a
    different algebraic form from trees.py:10008-10018."""
    tc, orc = tiny
    G = _genotypes(tc)
    S_ = list(range(8))
    n = len(S_)
    T = orc.stat("diversity", [S_], None, None, "site", False)[0, 0]
    S = orc.stat("segregating_sites", [S_], None, None, "site", False)[0, 0]
    assert S == sum(0 < G[s].sum() < n for s in range(G.shape[0]))
    a1 = sum(1 / i for i in range(1, n))
    a2 = sum(1 / i**2 for i in range(1, n))
    b1 = (n + 1) / (3 * (n - 1))
    b2 = 2 * (n**2 + n + 3) / (9 * n * (n - 1))
    c1 = b1 - 1 / a1
    c2 = b2 - (n + 2) / (a1 * n) + a2 / a1**2
    e1, e2 = c1 / a1, c2 / (a1**2 + a2)
    D = (T - S / a1) / np.sqrt(e1 * S + e2 * S * (S - 1))
    assert np.isclose(om.tajimas_d(np.array([[T]]), np.array([[S]]), [n])[0, 0], D, rtol=1e-12)


def test_brute_fst(tiny):
    """Slatkin's Fst = (pi_T - pi_S) / pi_T from mean within- and all-pair
    differences, against trees.py:10113-10125's form."""
    tc, orc = tiny
    A, B = [0, 1, 2, 3], [4, 5, 6, 7]
    du = orc.stat("diversity", [A], None, None, "site")[0, 0]
    dv = orc.stat("diversity", [B], None, None, "site")[0, 0]
    duv = orc.stat("divergence", [A, B], [(0, 1)], None, "site")[0, 0]
    pi_S = (du + dv) / 2
    pi_T = (du + dv + 2 * duv) / 4
    exp = (pi_T - pi_S) / pi_T
    got = om.fst(np.array([[du, dv]]), np.array([[duv]]), [(0, 1)])[0, 0]
    assert np.isclose(got, exp, rtol=1e-12)


def test_brute_afs_site(tiny):
    tc, orc = tiny
    G = _genotypes(tc)
    A = [0, 1, 2, 3, 4]
    exp = np.zeros(len(A) + 1)
    for s in range(G.shape[0]):
        k = G[s, A].sum()
        if 0 < G[s].sum() < 8:
            exp[k] += 1
    got = orc.afs([A], None, None, "site", False, True)[0, 0]
    assert np.array_equal(got, exp)


def test_oracle_genotypes_brute_force(tiny):
    tc, orc = tiny
    assert np.array_equal(orc.genotypes(), _genotypes(tc))
    # the decoder agrees on these single-mutation sites
    assert np.array_equal(orc.genotype_matrix(), _genotypes(tc).astype(np.int32))


@needs_ref
@pytest.mark.parametrize("case", ["paper_ex", "single_tree_ex", "nonbinary_ex", "generated",
                                  "gapped"])
@pytest.mark.parametrize("missing", [True, False])
def test_oracle_genotype_matrix_vs_reference(case, missing):
    """The oracle's decoder equals the reference's tsk_variant_decode, with
    back and recurrent mutations (single_tree_ex) and isolated samples
    (gapped: samples without edges over stretches) as missing or not."""
    if case == "generated":
        tc = cases.generated(70, 1e5, 31)
    elif case == "gapped":
        tc = cases.gapped(120, 1e5, 11)
    else:
        tc = getattr(cases, case)()
    orc, ref = Oracle(tc), om.Reference(tc)
    got = orc.genotype_matrix(missing)
    exp = ref.genotype_matrix(missing)
    assert np.array_equal(got, exp)
    if case == "gapped" and missing:
        assert np.any(exp == -1)  # the case exercises missing data


def _normalise_divergence(D):
    """The double centring of python/tskit/trees.py:9086-9095 for singleton
    sets (n = 1 each, so the diagonal correction zeroes the diagonal)."""
    B = D.copy()
    np.fill_diagonal(B, 0.0)
    K = B + np.mean(B)
    y = np.mean(B, axis=0)
    K -= y[:, None] + y[None, :]
    return K / -2


@needs_ref
def test_relatedness_matrix_closed_form_vs_reference():
    """K = J (G^T G) J equals the reference's genetic_relatedness_matrix route
    (its C divergence matrix, double-centred) for single-mutation sites."""
    tc = cases.generated(60, 2e5, 41)
    orc, ref = Oracle(tc), om.Reference(tc)
    samples = np.arange(60, dtype=np.int32)
    w = np.array([0, 5e4, 1.2e5, 2e5])
    D = ref.divergence_matrix(samples, w, "site", True)
    exp = np.array([_normalise_divergence(D[j]) for j in range(D.shape[0])])
    got = orc.relatedness_matrix(w, True)
    assert np.allclose(got, exp, rtol=1e-10, atol=1e-12 * np.max(np.abs(exp)))


@needs_ref
def test_divergence_matrix_closed_form_vs_reference():
    """The site divergence matrix of singleton sets is C_ii + C_jj - 2 C_ij
    (C = G^T G of the 0/1 genotypes) per window span: the form the GPU route
    computes, against the reference's tsk_treeseq_divergence_matrix."""
    tc = cases.generated(60, 2e5, 43)
    orc, ref = Oracle(tc), om.Reference(tc)
    w = np.array([0, 7e4, 2e5])
    exp = ref.divergence_matrix(np.arange(60, dtype=np.int32), w, "site", True)
    G = orc.genotypes().astype(np.int64)
    pos = tc.site_position
    for j in range(2):
        g = G[(pos >= w[j]) & (pos < w[j + 1])]
        C = g.T @ g
        a = np.diag(C)
        got = (a[:, None] + a[None, :] - 2 * C) / (w[j + 1] - w[j])
        assert np.allclose(got, exp[j], rtol=1e-13, atol=0)


def test_identity_invariant():
    """test_stats.c:1061-1124: branch diversity-style identity -- the branch
    segregating-sites of all samples with span normalisation off equals the
    total branch length integrated over the genome (each branch subtends
    0 < x < n samples, summary (1 - x/n) + x/n = 1 unpolarised)."""
    tc = cases.generated(10, 2e4, 33)
    orc = Oracle(tc)
    total = 0.0
    for a, b, par in _trees(tc):
        for u in range(tc.num_nodes):
            if par[u] != -1:
                total += (b - a) * (tc.node_time[par[u]] - tc.node_time[u])
    got = orc.stat("segregating_sites", [list(range(10))], None, None, "branch", False)[0, 0]
    assert np.isclose(got, total, rtol=1e-12)


NODE_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "node_reference.npz")


@pytest.mark.parametrize("key", ["generated", "gapped"])
def test_golden_node_stats(key):
    """Node statistics of the oracle against the reference's own outputs
    (tests/golden/make_node_golden.py)."""
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_node_golden as mk

    g = np.load(NODE_GOLDEN)
    tc = mk.inputs()[key]
    assert tc.column_digest() == bytes(g[f"{key}_digest"]).decode(), "generator output changed"
    orc = Oracle(tc)
    sets = _sets(g["sets"], 4)
    for name, stat, idx, pol, sn in mk.CHECKS:
        got = orc.stat(stat, sets, idx, g["windows"], "node", sn, pol)
        exp = g[f"{key}_{name}"]
        assert got.shape == exp.shape == (3, tc.num_nodes, exp.shape[2])
        assert np.allclose(got, exp, rtol=1e-12, atol=1e-12 * np.max(np.abs(exp))), name


@needs_ref
@pytest.mark.parametrize("fixture", ["paper_ex", "nonbinary_ex", "single_tree_ex"])
def test_oracle_node_vs_reference_fixtures(fixture):
    tc = getattr(cases, fixture)()
    orc, ref = Oracle(tc), om.Reference(tc)
    n = int(np.sum(tc.node_flags & 1))
    sets = [np.arange(0, n // 2, dtype=np.int32), np.arange(n // 2, n, dtype=np.int32)]
    w = np.linspace(0, tc.sequence_length, 4)
    for name, idx in [("diversity", None), ("segregating_sites", None),
                      ("divergence", [(0, 1), (1, 1)]), ("f2", [(0, 1)])]:
        for pol in (False, True):
            for sn in (False, True):
                a = orc.stat(name, sets, idx, w, "node", sn, pol)
                b = ref.stat(name, sets, idx, w, "node", sn, pol)
                assert np.allclose(a, b, rtol=1e-13, atol=1e-13, equal_nan=True), name


def test_node_single_tree_identity():
    """On one tree the branch statistic is the node statistic weighted by
    each node's branch length (trees.c:1352-1523 vs 1788-1918): a dropped
    piece, a wrong window or a missed update fails this."""
    tc = cases.single_tree_ex()
    orc = Oracle(tc)
    sets = [np.array([0, 1, 2], np.int32), np.array([1, 2, 3], np.int32)]
    par = np.full(tc.num_nodes, -1)
    par[tc.edge_child] = tc.edge_parent
    bl = np.where(par >= 0, tc.node_time[par] - tc.node_time, 0.0)
    for name, idx in [("diversity", None), ("divergence", [(0, 1)]), ("Y1", None)]:
        for pol in (False, True):
            node = orc.stat(name, sets, idx, None, "node", False, pol)
            branch = orc.stat(name, sets, idx, None, "branch", False, pol)
            assert node.shape == (1, tc.num_nodes, node.shape[2])
            assert np.all(np.isfinite(node))
            assert np.allclose(np.einsum("wnm,n->wm", node, bl), branch, rtol=1e-13), name


def _set_divergence(G, pos, w, sets):
    """Closed form of the site divergence matrix between disjoint sample sets
    for 0/1 genotypes: per window the count of differing genotype pairs over
    A x B (a pair within one set counted in both orders), divided by |A||B|,
    or |A|(|A|-1) on the diagonal when not zero, and by the window span."""
    out = []
    for j in range(len(w) - 1):
        g = G[(pos >= w[j]) & (pos < w[j + 1])]
        C = g.T @ g
        a = np.diag(C)
        P = (a[:, None] + a[None, :] - 2 * C).astype(np.float64)
        K = len(sets)
        D = np.zeros((K, K))
        for x in range(K):
            for y in range(K):
                v = P[np.ix_(sets[x], sets[y])].sum()
                na, nb = len(sets[x]), len(sets[y])
                den = na * (na - 1) if x == y else na * nb
                D[x, y] = v / den if den != 0 else v
        out.append(D / (w[j + 1] - w[j]))
    return np.array(out)


@needs_ref
def test_set_divergence_matrix_closed_form_vs_reference():
    """Sample sets of several samples (and one singleton) in the reference's
    tsk_treeseq_divergence_matrix: the closed form the GPU route computes."""
    tc = cases.generated(60, 2e5, 43)
    orc, ref = Oracle(tc), om.Reference(tc)
    w = np.array([0, 7e4, 2e5])
    rng = np.random.default_rng(3)
    perm = rng.permutation(60).astype(np.int32)
    sets = [perm[:10], perm[10:11], perm[11:30], perm[30:52]]
    exp = ref.divergence_matrix_sets(sets, w, "site", True)
    G = orc.genotypes().astype(np.int64)
    got = _set_divergence(G, tc.site_position, w, sets)
    assert np.allclose(got, exp, rtol=1e-13, atol=0)


def _branch_divergence_from_shared(G, n, w, sets):
    """Branch divergence matrix between disjoint sample sets from the shared
    branch length of every two samples (G [W, n(n+1)/2] in (i <= j) order):
    path length G_ii + G_jj - 2 G_ij, summed per pair of sets, divided by the
    number of pairs and by the window span."""
    iu = np.triu_indices(n)
    out = []
    for j in range(len(w) - 1):
        S = np.zeros((n, n))
        S[iu] = G[j]
        S = S + np.triu(S, 1).T
        d = np.diag(S)
        P = d[:, None] + d[None, :] - 2 * S
        np.fill_diagonal(P, 0.0)
        K = len(sets)
        D = np.zeros((K, K))
        for x in range(K):
            for y in range(K):
                na, nb = len(sets[x]), len(sets[y])
                den = na * (na - 1) if x == y else na * nb
                v = P[np.ix_(sets[x], sets[y])].sum()
                D[x, y] = v / den if den != 0 else v
        out.append(D / (w[j + 1] - w[j]))
    return np.array(out)


@needs_ref
@pytest.mark.parametrize("case", ["generated", "gapped"])
def test_branch_divergence_matrix_shared_length_vs_reference(case):
    """The reference's branch divergence matrix (per-tree MRCA path lengths,
    or paths to the own roots across subtrees) equals G_ii + G_jj - 2 G_ij of
    the branch-mode noncentred relatedness of singleton sets, per pair of
    sets: the route the GPU takes."""
    tc = cases.generated(30, 1e5, 47) if case == "generated" else cases.gapped(30, 1e5, 47, 8)
    orc, ref = Oracle(tc), om.Reference(tc)
    n = 30
    w = np.array([0, 3e4, 1e5])
    rng = np.random.default_rng(4)
    perm = rng.permutation(n).astype(np.int32)
    sets = [perm[:6], perm[6:7], perm[7:20], perm[20:30]]
    pairs = [(i, j) for i in range(n) for j in range(i, n)]
    single = [np.array([i], np.int32) for i in range(n)]
    G = orc.stat("genetic_relatedness", single, pairs, w, "branch", False, True, False)
    # sets hold node ids; samples 0..n-1 are nodes 0..n-1 here
    got = _branch_divergence_from_shared(G, n, w, sets)
    exp = ref.divergence_matrix_sets(sets, w, "branch", True)
    assert np.allclose(got, exp, rtol=1e-11, atol=0)
    exp1 = ref.divergence_matrix(np.arange(n, dtype=np.int32), w, "branch", True)
    got1 = _branch_divergence_from_shared(G, n, w, [[i] for i in range(n)])
    assert np.allclose(got1, exp1, rtol=1e-11, atol=1e-300)


RELVEC_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "relvec_reference.npz")


@pytest.mark.parametrize("key", ["generated", "gapped"])
def test_golden_relatedness_vector(key):
    """genetic_relatedness_vector of the oracle against the reference's own
    outputs (tests/golden/make_relvec_golden.py): windows spanning the
    sequence and not, centred and not, every node as a focal node."""
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_relvec_golden as mk

    g = np.load(RELVEC_GOLDEN)
    tc = mk.inputs()[key]
    assert tc.column_digest() == bytes(g[f"{key}_digest"]).decode(), "generator output changed"
    orc = Oracle(tc)
    W = mk.weights(int(np.sum(tc.node_flags & 1)))
    for name, wk, centre, sn, fk in mk.CHECKS:
        got = orc.relatedness_vector(W, mk.windows()[wk], "branch", sn, centre, mk.focal(tc, fk))
        exp = g[f"{key}_{name}"]
        assert got.shape == exp.shape
        assert np.allclose(got, exp, rtol=1e-12, atol=1e-12 * np.max(np.abs(exp))), name


@needs_ref
@pytest.mark.parametrize("fixture", ["paper_ex", "single_tree_ex", "nonbinary_ex", "generated"])
def test_relatedness_vector_vs_reference(fixture):
    tc = cases.generated(40, 3e4, 12) if fixture == "generated" else getattr(cases, fixture)()
    orc, ref = Oracle(tc), om.Reference(tc)
    n = int(np.sum(tc.node_flags & 1))
    W = np.random.default_rng(2).normal(size=(n, 2))
    L = tc.sequence_length
    for w in [None, np.array([0.2 * L, 0.5 * L, 0.9 * L])]:
        for centre in (True, False):
            a = orc.relatedness_vector(W, w, "branch", True, centre)
            b = ref.relatedness_vector(W, w, "branch", True, centre)
            assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b)))
    for mode in ("site", "node"):
        with pytest.raises(om.OracleError):
            orc.relatedness_vector(W, None, mode)
        with pytest.raises(om.OracleError):
            ref.relatedness_vector(W, None, mode)


@pytest.mark.parametrize("case", ["generated", "gapped", "paper_ex"])
def test_relatedness_vector_closed_form(case):
    """Uncentred, the relatedness vector is G w with G the branch length each
    pair of samples shares (the noncentred relatedness of singleton sets,
    trees.c:4755-4773, itself checked against the reference); centred, it is
    J G J w with J = I - 11'/n (trees.py:9157-9225)."""
    tc = {"generated": lambda: cases.generated(30, 4e4, 8),
          "gapped": lambda: cases.gapped(30, 4e4, 8, count=6),
          "paper_ex": cases.paper_ex}[case]()
    orc = Oracle(tc)
    smp = np.flatnonzero(tc.node_flags & 1)
    n = len(smp)
    single = [np.array([i], np.int32) for i in smp]
    pairs = [(i, j) for i in range(n) for j in range(n)]
    L = tc.sequence_length
    w = np.array([0, 0.4 * L, L])
    G = orc.stat("genetic_relatedness", single, pairs, w, "branch", False, True, False)
    W = np.random.default_rng(7).normal(size=(n, 2))
    J = np.eye(n) - 1.0 / n
    raw = orc.relatedness_vector(W, w, "branch", False, False)
    cen = orc.relatedness_vector(W, w, "branch", False, True)
    for j in range(2):
        Gj = G[j].reshape(n, n)
        assert np.allclose(raw[j], Gj @ W, rtol=1e-12, atol=1e-12 * np.max(np.abs(raw[j])))
        assert np.allclose(cen[j], J @ Gj @ J @ W, rtol=1e-10,
                           atol=1e-12 * np.max(np.abs(Gj @ W)))


WEIGHTED = ["trait_covariance", "trait_correlation", "trait_linear_model",
            "genetic_relatedness_weighted"]


def _weighted_calls(n, rng):
    """(name, keyword arguments) of each weighted statistic on random
    weights of n samples."""
    Z = rng.normal(size=(n, 2))
    return [("trait_covariance", {}), ("trait_correlation", {}),
            ("trait_linear_model", {"Z": Z}),
            ("genetic_relatedness_weighted", {"indexes": [(0, 1), (1, 2), (2, 2)]}),
            ("genetic_relatedness_weighted", {"indexes": [(0, 2), (1, 1)], "centre": False})]


@needs_ref
@pytest.mark.parametrize("fixture", ["paper_ex", "nonbinary_ex", "generated", "multiallelic"])
def test_weighted_stats_vs_reference(fixture):
    """The restated weighted statistics (general_stat of real-valued weights
    with the trait and weighted-relatedness summaries, trees.c:3961-4220,
    4801-4895) against the compiled reference, both modes and
    polarisations, windows spanning the sequence in several pieces."""
    tc = {"generated": lambda: cases.generated(40, 3e4, 12),
          "multiallelic": lambda: cases.multiallelic(30, 2e4, 3)}.get(
              fixture, getattr(cases, fixture, None))()
    orc, ref = Oracle(tc), om.Reference(tc)
    n = int(np.sum(tc.node_flags & 1))
    rng = np.random.default_rng(3)
    W = rng.normal(size=(n, 3))
    L = tc.sequence_length
    for w in [None, np.array([0, 0.3 * L, 0.35 * L, L])]:
        for mode in ("branch", "site"):
            for pol in (False, True):
                for name, kw in _weighted_calls(n, rng):
                    a = orc.weighted_stat(name, W, w, mode, True, pol, **kw)
                    b = ref.weighted_stat(name, W, w, mode, True, pol, **kw)
                    assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b))), name
    with pytest.raises(om.OracleError):
        orc.weighted_stat("trait_covariance", W, np.array([0.1, L]))
    with pytest.raises(om.OracleError):
        ref.weighted_stat("trait_covariance", W, np.array([0.1, L]))


@pytest.mark.parametrize("case", ["generated", "gapped", "paper_ex"])
def test_weighted_stats_closed_form(case):
    """Branch mode, polarised, unnormalised: with G the branch length each
    pair of samples shares (the non-centred relatedness of singleton sets,
    itself checked against the reference), trait covariance is
    w'Gw / (2 (n-1)^2) for the centred trait w, and weighted relatedness is
    v_i'G v_j with v the centred weights (or the raw weights, non-centred):
    the summaries' sums over the nodes, sum bl x x', written as quadratic
    forms."""
    tc = {"generated": lambda: cases.generated(30, 4e4, 8),
          "gapped": lambda: cases.gapped(30, 4e4, 8, count=6),
          "paper_ex": cases.paper_ex}[case]()
    orc = Oracle(tc)
    smp = np.flatnonzero(tc.node_flags & 1)
    n = len(smp)
    single = [np.array([i], np.int32) for i in smp]
    pairs = [(i, j) for i in range(n) for j in range(n)]
    L = tc.sequence_length
    w = np.array([0, 0.4 * L, L])
    G = orc.stat("genetic_relatedness", single, pairs, w, "branch", False, True, False)
    W = np.random.default_rng(7).normal(size=(n, 3))
    Wc = W - W.mean(axis=0)
    cov = orc.weighted_stat("trait_covariance", W, w, "branch", False, True)
    idx = [(0, 1), (2, 2), (1, 0)]
    rel = orc.weighted_stat("genetic_relatedness_weighted", W, w, "branch", False, True,
                            indexes=idx)
    reln = orc.weighted_stat("genetic_relatedness_weighted", W, w, "branch", False, True,
                             indexes=idx, centre=False)
    for j in range(2):
        Gj = G[j].reshape(n, n)
        exp = np.einsum("sk,st,tk->k", Wc, Gj, Wc) / (2 * (n - 1) ** 2)
        assert np.allclose(cov[j], exp, rtol=1e-10, atol=1e-12 * np.max(np.abs(exp)))
        for m, (a, b) in enumerate(idx):
            e1 = Wc[:, a] @ Gj @ Wc[:, b]
            e2 = W[:, a] @ Gj @ W[:, b]
            assert np.isclose(rel[j, m], e1, rtol=1e-10, atol=1e-12 * np.max(np.abs(Gj)))
            assert np.isclose(reln[j, m], e2, rtol=1e-10, atol=1e-12 * np.max(np.abs(Gj)))


@needs_ref
def test_window_partitioned_reference_equals_whole():
    """The CPU baseline's window-partitioned reference (each process clips
    edges and sites to its block of windows, trees.c:1255-1266 forcing the
    outer windows) gives the single-process reference's values."""
    import functools

    from oracle import partitioned

    tc = cases.generated(300, 4e5, 21)
    sets = cases.sample_sets(300, 4, 40, 2)
    w = np.linspace(0, 4e5, 11)
    pairs = [(0, 1), (2, 3), (1, 3)]
    quads = [(0, 1, 2, 3)]
    step = functools.partial(partitioned.scale_step, sets=sets, pairs=pairs, quads=quads)
    whole = step(om.Reference(tc), w)
    got, wall, t_init, t_stat, P = partitioned.run(tc, w, step, processes=3)
    assert P == 3 and wall > 0 and t_init > 0 and t_stat > 0
    for k, v in whole.items():
        fin = np.isfinite(v)
        assert np.array_equal(fin, np.isfinite(got[k])), k
        scale = np.max(np.abs(v[fin]))
        assert np.max(np.abs(got[k][fin] - v[fin])) <= 1e-10 * scale, k


def _allele_divergence(G, pos, windows, sets=None):
    """Site divergence matrix from genotypes G [S, n] (isolated samples
    ancestral): per window the number of (site, sample pair) with different
    alleles, summed over each pair of sets and divided by its number of pairs
    (trees.c:8684-8826, 8876-8899)."""
    n = G.shape[1]
    if sets is None:
        sets = [[i] for i in range(n)]
    out = []
    for j in range(len(windows) - 1):
        g = G[(pos >= windows[j]) & (pos < windows[j + 1])]
        # sites minus the pairs carrying the same allele
        D = np.full((n, n), len(g), dtype=np.int64)
        for a in range(int(g.max()) + 1 if len(g) else 0):
            B = (g == a).astype(np.int64)
            D -= B.T @ B
        D = D.astype(np.float64)
        K = len(sets)
        R = np.zeros((K, K))
        for a in range(K):
            for b in range(K):
                v = D[np.ix_(sets[a], sets[b])].sum()
                na, nb = len(sets[a]), len(sets[b])
                den = na * (na - 1) if a == b else na * nb
                R[a, b] = v / den if den else v
        out.append(R)
    return np.array(out)


@pytest.mark.parametrize("case", ["multiallelic", "single_tree_ex"])
def test_multiallelic_rows_identity(case):
    """The row decomposition the matrix kernels use for sites with three or
    more alleles: with P = 2 G1'G1 + G2'G2 (G1: derived rows of sites with at
    most two alleles; G2: any-derived row and one row per derived allele of
    the others), (P_ii + P_jj)/2 - P_ij counts the sites where samples i and
    j carry different alleles, on the oracle's genotypes (pinned to
    tsk_variant_decode)."""
    tc = cases.multiallelic(40, 6e4, 4) if case == "multiallelic" else cases.single_tree_ex()
    G = Oracle(tc).genotype_matrix(False)
    na = G.max(axis=1) + 1
    # single_tree_ex: two alleles per site, but back and recurrent mutations
    assert np.any(na >= 3) or case == "single_tree_ex"
    rows1 = [(G[s] != 0).astype(np.int64) for s in range(len(G)) if na[s] <= 2]
    rows2 = []
    for s in np.flatnonzero(na >= 3):
        rows2.append((G[s] != 0).astype(np.int64))
        rows2.extend((G[s] == a).astype(np.int64) for a in range(1, na[s]))
    G1 = np.array(rows1).reshape(-1, G.shape[1])
    G2 = np.array(rows2).reshape(-1, G.shape[1])
    P = 2 * G1.T @ G1 + G2.T @ G2
    d = np.diag(P)
    assert np.all(d % 2 == 0)
    D = (d[:, None] + d[None, :]) // 2 - P
    exp = (G[:, :, None] != G[:, None, :]).sum(axis=0)
    assert np.array_equal(D, exp)


@needs_ref
@pytest.mark.parametrize("case", ["multiallelic", "single_tree_ex"])
def test_multiallelic_divergence_matrix_vs_reference(case):
    """The genotype closed form of the site divergence matrix against the
    reference's tsk_treeseq_divergence_matrix on sites with back, recurrent
    and stacked mutations (2 to 4 alleles), singletons and sets."""
    tc = cases.multiallelic(40, 6e4, 4) if case == "multiallelic" else cases.single_tree_ex()
    ref = om.Reference(tc)
    G = Oracle(tc).genotype_matrix(False)
    n = G.shape[1]
    L = tc.sequence_length
    w = np.array([0, 0.35 * L, L])
    got = ref.divergence_matrix(np.arange(n, dtype=np.int32), w, "site")
    exp = _allele_divergence(G, tc.site_position, w) / np.diff(w)[:, None, None]
    assert np.allclose(got, exp, rtol=1e-12, atol=0)
    if n >= 4:
        sets = [np.arange(0, n // 2, dtype=np.int32), np.arange(n // 2, n - 1, dtype=np.int32),
                np.array([n - 1], dtype=np.int32)]
        got = ref.divergence_matrix_sets(sets, w, "site")
        exp = _allele_divergence(G, tc.site_position, w, [list(s) for s in sets])
        assert np.allclose(got, exp / np.diff(w)[:, None, None], rtol=1e-12, atol=0)
