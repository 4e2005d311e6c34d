"""Executes the device chunk plan on the CPU, in FP64, with the same indexing
as the HIP sweep kernels (tskit_amd/csrc/tsg_kernels.hip) and explicit bounds
checks.  Used by the CPU tests to show that the plan the host builds is
in-bounds and reproduces the oracle before any kernel runs it.
"""
import numpy as np


def check_plan_bounds(plan, num_trees, num_samples):
    ct0, cn0 = plan["chunk_tree0"], plan["chunk_node0"]
    C = len(ct0) - 1
    assert ct0[0] == 0 and ct0[-1] == num_trees
    assert np.all(np.diff(ct0) >= 1)
    assert cn0[0] == 0 and cn0[-1] == len(plan["loc_parent0"])
    tr_off, tr_nrem = plan["tr_off"], plan["tr_nrem"]
    assert np.all(np.diff(tr_off) >= 0)
    assert tr_off[-1] == len(plan["ev_child"])
    ep = plan["ev_path0"]
    assert len(ep) == len(plan["ev_child"]) + 1 and ep[0] == 0 and ep[-1] == len(plan["path"])
    assert np.all(np.diff(ep) >= 1)  # every path starts at the edge's parent
    assert np.array_equal(plan["path"][ep[:-1]], plan["ev_parent"])
    assert np.all(tr_nrem >= 0) and np.all(tr_nrem <= np.diff(tr_off))
    # the opening transition is applied by the planner: no stored events
    assert tr_off[0] == 0 and tr_off[1] == 0
    for c in range(C):
        nloc = cn0[c + 1] - cn0[c]
        assert nloc >= num_samples
        par = plan["loc_parent0"][cn0[c]:cn0[c + 1]]
        assert np.all((par >= -1) & (par < nloc))
        t0, t1 = ct0[c], ct0[c + 1]
        first = t0 + 1
        last = min(t1, num_trees - 1)
        if first <= last:
            e0, e1 = tr_off[first], tr_off[last + 1]
            for a in ("ev_child", "ev_parent"):
                ids = plan[a][e0:e1]
                assert np.all((ids >= 0) & (ids < nloc)), (c, a)
            pids = plan["path"][ep[e0]:ep[e1]]
            assert np.all((pids >= 0) & (pids < nloc)), (c, "path")
        s0, s1 = plan["tree_site0"][t0], plan["tree_site0"][t1]
        m0, m1 = plan["site_mut0"][s0], plan["site_mut0"][s1]
        mn = plan["mut_node"][m0:m1]
        assert np.all((mn >= 0) & (mn < nloc))
    return True


def _summary(stat, x, n, idx, polarised, total):
    def f(xx):
        if stat == "diversity":
            return xx * (n - xx) / (n * (n - 1))
        if stat == "divergence":
            i, j = idx[:, 0], idx[:, 1]
            return xx[i] * (n[j] - xx[j]) / (n[i] * (n[j] - (i == j)))
        raise ValueError(stat)

    y = f(x)
    if not polarised:
        y = y + f(total - x)
    return y


def run(plan, bp, windows, sets, num_samples, stat, mode, idx=None, polarised=False):
    """Returns the raw (not span-normalised) [W, M] statistic."""
    ct0, cn0 = plan["chunk_tree0"], plan["chunk_node0"]
    C = len(ct0) - 1
    T = len(bp) - 1
    K = len(sets)
    n = np.array([len(s) for s in sets], dtype=np.float64)
    total = n.copy()
    if idx is not None:
        idx = np.asarray(idx)
    M = K if idx is None else len(idx)
    W = len(windows) - 1
    out = np.zeros((W, M))
    dR = np.zeros((C, M))
    Rstart_terms = []
    for c in range(C):
        nloc = cn0[c + 1] - cn0[c]
        par = plan["loc_parent0"][cn0[c]:cn0[c + 1]].astype(np.int64).copy()
        bl = plan["loc_bl0"][cn0[c]:cn0[c + 1]].astype(np.float64).copy()
        X = np.zeros((nloc, K))
        for k, s in enumerate(sets):
            for si in s:
                u = si
                while u >= 0:
                    X[u, k] += 1
                    u = par[u]
        R = np.zeros(M)
        A = np.zeros((W, M))

        def F(x):
            return _summary(stat, x, n, idx, polarised, total)

        if c == 0:
            # running sum of the first tree, evaluated over its nodes
            R0 = np.zeros(M)
            for u in range(nloc):
                if par[u] >= 0:
                    R0 += bl[u] * F(X[u])

        def transition(t):
            nonlocal R
            e0, e1 = plan["tr_off"][t], plan["tr_off"][t + 1]
            nrem = plan["tr_nrem"][t]
            for e in range(e0, e1):
                ins = (e - e0) >= nrem
                u, v = plan["ev_child"][e], plan["ev_parent"][e]
                assert 0 <= u < nloc and 0 <= v < nloc
                xc = X[u].copy()
                if mode == "branch":
                    b = plan["ev_bl"][e] if ins else bl[u]
                    R += (1 if ins else -1) * b * F(xc)
                if ins:
                    par[u], bl[u] = v, plan["ev_bl"][e]
                else:
                    par[u], bl[u] = -1, 0.0
                # the recorded root path must be the live parent walk from v
                p0, p1 = plan["ev_path0"][e], plan["ev_path0"][e + 1]
                path = plan["path"][p0:p1]
                walk = []
                w = v
                while w >= 0:
                    walk.append(w)
                    w = par[w]
                    assert len(walk) <= nloc
                assert list(path) == walk, (e, list(path), walk)
                for w in path:
                    if mode == "branch":
                        R -= bl[w] * F(X[w])
                    X[w] = X[w] + xc if ins else X[w] - xc
                    if mode == "branch":
                        R += bl[w] * F(X[w])

        t0, t1 = ct0[c], ct0[c + 1]
        for t in range(t0, t1):
            if t > t0:
                transition(t)
            if mode == "branch":
                tl, tr = bp[t], bp[t + 1]
                for w in range(W):
                    lo, hi = max(tl, windows[w]), min(tr, windows[w + 1])
                    if hi > lo:
                        A[w] += R * (hi - lo)
            else:
                s0, s1 = plan["tree_site0"][t], plan["tree_site0"][t + 1]
                for s in range(s0, s1):
                    pos_w = None
                    na = plan["site_nalleles"][s]
                    m0, m1 = plan["site_mut0"][s], plan["site_mut0"][s + 1]
                    val = np.zeros(M)
                    for al in range(1 if polarised else 0, na):
                        aw = total.copy() if al == 0 else np.zeros(K)
                        for mm in range(m0, m1):
                            node = plan["mut_node"][mm]
                            if plan["mut_allele"][mm] == al:
                                aw += X[node]
                            if plan["mut_pallele"][mm] == al:
                                aw -= X[node]
                        y = F(aw) if not polarised else F(aw)
                        if stat == "diversity":
                            y = aw * (n - aw) / (n * (n - 1))
                        else:
                            i, j = idx[:, 0], idx[:, 1]
                            y = aw[i] * (n[j] - aw[j]) / (n[i] * (n[j] - (i == j)))
                        val += y
                    pos_w = s  # site window found by the caller below
                    A[np.searchsorted(windows, plan["_site_pos"][pos_w], side="right") - 1] += val
        if mode == "branch" and t1 < T:
            transition(t1)
        dR[c] = R
        out += A
        Rstart_terms.append((bp[t0], bp[t1]))
    if mode == "branch":
        Rs = R0 + np.vstack([np.zeros(M), np.cumsum(dR, axis=0)[:-1]])
        for c, (cs, ce) in enumerate(Rstart_terms):
            for w in range(W):
                ov = min(ce, windows[w + 1]) - max(cs, windows[w])
                if ov > 0:
                    out[w] += Rs[c] * ov
    return out


NO_SLOT = 0xFFFF


def check_hot_plan_bounds(hp, num_trees, num_samples, num_mutations):
    """Every index the hot sweep kernel (k_sweep with the hot store) reads is
    inside its chunk's slot range, and the level order is valid for the
    bottom-up seeding."""
    ct0, cs0, cl0 = hp["chunk_tree0"], hp["chunk_slot0"], hp["chunk_lev0"]
    C = len(ct0) - 1
    n = num_samples
    assert ct0[0] == 0 and ct0[-1] == num_trees and np.all(np.diff(ct0) >= 1)
    assert cs0[0] == 0 and cs0[-1] == len(hp["slot_par"]) == len(hp["slot_bl0"])
    assert cl0[0] == 0 and cl0[-1] == len(hp["lev_slot0"])
    assert len(hp["sample_slot"]) == C * n
    tr_off, tr_nrem, ep = hp["tr_off"], hp["tr_nrem"], hp["ev_path0"]
    assert tr_off[0] == 0 and tr_off[1] == 0 and tr_off[-1] == len(hp["ev_child"])
    assert np.all(np.diff(tr_off) >= 0) and np.all(tr_nrem <= np.diff(tr_off))
    assert ep[0] == 0 and ep[-1] == len(hp["path"]) and np.all(np.diff(ep) >= 1)
    assert np.all(hp["mut_slot"] >= 0) or num_mutations == 0
    t0p = hp["t0_parent"]
    assert len(t0p) >= n and np.all((t0p >= -1) & (t0p < len(t0p)))
    for c in range(C):
        S = cs0[c + 1] - cs0[c]
        assert S <= hp["max_slots"]
        par = hp["slot_par"][cs0[c]:cs0[c + 1]]
        assert np.all((par >= -1) & (par < S))
        lev = hp["lev_slot0"][cl0[c]:cl0[c + 1]]
        assert lev[0] == 0 and lev[-1] == S and np.all(np.diff(lev) >= 0)
        level = np.zeros(S, np.int64)
        for j in range(len(lev) - 1):
            level[lev[j]:lev[j + 1]] = j
        nlev = len(lev) - 1
        # a parent is exactly one level up; the roots of the hot forest, and
        # only they, form the last level (which is not propagated)
        has_par = par >= 0
        assert np.all(level[par[has_par]] == level[has_par] + 1)
        assert np.array_equal(~has_par, level == nlev - 1)
        ss = hp["sample_slot"][c * n:(c + 1) * n].astype(np.int64)
        assert np.all((ss < S) | (ss == NO_SLOT))
        t0, t1 = ct0[c], ct0[c + 1]
        first, last = t0 + 1, min(t1, num_trees - 1)
        if first <= last:
            e0, e1 = tr_off[first], tr_off[last + 1]
            ids = hp["ev_child"][e0:e1]
            assert np.all((ids >= 0) & (ids < S)), c
            pids = hp["path"][ep[e0]:ep[e1]]
            assert np.all((pids >= 0) & (pids < S)), c
    return True


def run_hot(hp, plan_sites, bp, windows, sets, num_samples, stat, mode, idx=None,
            polarised=False):
    """The hot sweep program executed in FP64 as the kernel does it: slots
    seeded from sample_slot and the level order, events through the recorded
    paths, the first tree's running sum over t0_parent.  plan_sites holds the
    site arrays (tree_site0, site_mut0, mut_allele, mut_pallele,
    site_nalleles, _site_pos).  Returns the raw [W, M] statistic."""
    ct0, cs0, cl0 = hp["chunk_tree0"], hp["chunk_slot0"], hp["chunk_lev0"]
    C = len(ct0) - 1
    T = len(bp) - 1
    K = len(sets)
    n = np.array([len(s) for s in sets], dtype=np.float64)
    total = n.copy()
    if idx is not None:
        idx = np.asarray(idx)
    M = K if idx is None else len(idx)
    W = len(windows) - 1
    nsamp = num_samples
    weight = np.zeros((nsamp, K))
    for k, s in enumerate(sets):
        weight[np.asarray(s), k] = 1

    def F(x):
        return _summary(stat, x, n, idx, polarised, total)

    # the first tree's running sum
    t0p, t0b = hp["t0_parent"], hp["t0_bl"].astype(np.float64)
    X0 = np.zeros((len(t0p), K))
    for s in range(nsamp):
        u = s
        while u >= 0:
            X0[u] += weight[s]
            u = t0p[u]
    R0 = np.zeros(M)
    for u in range(len(t0p)):
        if t0p[u] >= 0:
            R0 += t0b[u] * F(X0[u])

    out = np.zeros((W, M))
    dR = np.zeros((C, M))
    spans = []
    for c in range(C):
        S = cs0[c + 1] - cs0[c]
        par = hp["slot_par"][cs0[c]:cs0[c + 1]]
        bl = hp["slot_bl0"][cs0[c]:cs0[c + 1]].astype(np.float64).copy()
        X = np.zeros((S, K))
        ss = hp["sample_slot"][c * nsamp:(c + 1) * nsamp]
        for s in range(nsamp):
            if ss[s] != NO_SLOT:
                X[ss[s]] += weight[s]
        lev = hp["lev_slot0"][cl0[c]:cl0[c + 1]]
        for j in range(len(lev) - 2):
            for h in range(lev[j], lev[j + 1]):
                if par[h] >= 0:
                    X[par[h]] += X[h]
        R = np.zeros(M)
        A = np.zeros((W, M))

        def transition(t):
            nonlocal R
            e0, e1 = hp["tr_off"][t], hp["tr_off"][t + 1]
            nrem = hp["tr_nrem"][t]
            for e in range(e0, e1):
                ins = (e - e0) >= nrem
                u = hp["ev_child"][e]
                xc = X[u].copy()
                if mode == "branch":
                    b = hp["ev_bl"][e] if ins else bl[u]
                    R += (1 if ins else -1) * b * F(xc)
                bl[u] = hp["ev_bl"][e] if ins else 0.0
                for w in hp["path"][hp["ev_path0"][e]:hp["ev_path0"][e + 1]]:
                    if mode == "branch":
                        R -= bl[w] * F(X[w])
                    X[w] = X[w] + xc if ins else X[w] - xc
                    if mode == "branch":
                        R += bl[w] * F(X[w])

        t0, t1 = ct0[c], ct0[c + 1]
        for t in range(t0, t1):
            if t > t0:
                transition(t)
            if mode == "branch":
                tl, tr = bp[t], bp[t + 1]
                for w in range(W):
                    lo, hi = max(tl, windows[w]), min(tr, windows[w + 1])
                    if hi > lo:
                        A[w] += R * (hi - lo)
            else:
                ps = plan_sites
                for s in range(ps["tree_site0"][t], ps["tree_site0"][t + 1]):
                    m0, m1 = ps["site_mut0"][s], ps["site_mut0"][s + 1]
                    val = np.zeros(M)
                    for al in range(1 if polarised else 0, ps["site_nalleles"][s]):
                        aw = total.copy() if al == 0 else np.zeros(K)
                        for mm in range(m0, m1):
                            h = hp["mut_slot"][mm]
                            if ps["mut_allele"][mm] == al:
                                aw += X[h]
                            if ps["mut_pallele"][mm] == al:
                                aw -= X[h]
                        if stat == "diversity":
                            val += aw * (n - aw) / (n * (n - 1))
                        else:
                            i, j = idx[:, 0], idx[:, 1]
                            val += aw[i] * (n[j] - aw[j]) / (n[i] * (n[j] - (i == j)))
                    A[np.searchsorted(windows, ps["_site_pos"][s], side="right") - 1] += val
        if mode == "branch" and t1 < T:
            transition(t1)
        dR[c] = R
        out += A
        spans.append((bp[t0], bp[t1]))
    if mode == "branch":
        Rs = R0 + np.vstack([np.zeros(M), np.cumsum(dR, axis=0)[:-1]])
        for c, (cs, ce) in enumerate(spans):
            for w in range(W):
                ov = min(ce, windows[w + 1]) - max(cs, windows[w])
                if ov > 0:
                    out[w] += Rs[c] * ov
    return out


def _fold(coord, dims):
    """Minor-allele fold of an AFS coordinate (trees.c:3469-3496)."""
    n = sum(d - 1 for d in dims) / 2
    s = sum(coord)
    k = len(dims)
    while s == n and k > 0:
        k -= 1
        n -= (dims[k] - 1) / 2
        s -= coord[k]
    if s > n:
        return [d - 1 - c for c, d in zip(coord, dims)]
    return list(coord)


def run_hot_afs(hp, plan_sites, bp, windows, time_windows, sets, num_samples, node_time, mode,
                polarised=False):
    """The hot-plan AFS program as the kernels run it, in FP64: slot counts
    (sample sets, then all samples seeded from slot_nsd), events through the
    recorded paths; branch mode flushes a node only when an event changes it
    (as a removed edge's child or a node on an event's path), splitting the
    interval since its last flush at window boundaries, and a last-tree pass
    flushes the intervals still open at the end.  Returns the span-normalised
    [W, TW, dims...] spectrum (TW = 1 for sites)."""
    ct0, cs0, cl0 = hp["chunk_tree0"], hp["chunk_slot0"], hp["chunk_lev0"]
    C = len(ct0) - 1
    T = len(bp) - 1
    L = bp[-1]
    Ks = len(sets)
    K1 = Ks + 1
    dims = [len(s) + 1 for s in sets]
    W = len(windows) - 1
    tw = np.asarray([0, np.inf] if time_windows is None else time_windows, np.float64)
    TW = len(tw) - 1 if mode == "branch" else 1
    hist = np.zeros((W, TW) + tuple(dims))
    n = num_samples
    weight = np.zeros((n, K1))
    for k, s in enumerate(sets):
        weight[np.asarray(s), k] += 1
    weight[:, Ks] = 1
    total = weight.sum(axis=0)

    def coord(x):
        c = [int(v) for v in x[:Ks]]
        return tuple(c if polarised else _fold(c, dims))

    def flush_pieces(x_counts, tu, tp, a, y, x):
        """Contribution of a node with counts x_counts, own time tu and
        parent time tp (-inf: none; its edge inserted at y) over [a, x)."""
        if not (tp > -np.inf) or not (0 < x_counts[Ks] < n):
            return
        off = coord(x_counts)
        wi = int(np.searchsorted(windows, a, side="right")) - 1
        p = a
        pieces = []
        while wi < W and windows[wi + 1] <= x:
            q = windows[wi + 1]
            if y < q:
                pieces.append((wi, q - p))
            p = q
            wi += 1
        if wi < W and x > p:
            pieces.append((wi, x - p))
        for w, span in pieces:
            j = 0
            while j < TW and tw[j] < tp:
                twbl = max(0.0, min(tw[j + 1], tp) - max(tw[j], tu))
                hist[(w, j) + off] += span * twbl
                j += 1

    for c in range(C):
        S = cs0[c + 1] - cs0[c]
        sl = slice(cs0[c], cs0[c + 1])
        par = hp["slot_par"][sl]
        X = np.zeros((S, K1))
        ss = hp["sample_slot"][c * n:(c + 1) * n]
        for s in range(n):
            if ss[s] != NO_SLOT:
                X[ss[s], :Ks] += weight[s, :Ks]
        X[:, Ks] = hp["slot_nsd"][sl]
        lev = hp["lev_slot0"][cl0[c]:cl0[c + 1]]
        for j in range(len(lev) - 2):
            for h in range(lev[j], lev[j + 1]):
                if par[h] >= 0:
                    X[par[h]] += X[h]
        tu = node_time[hp["slot_node"][sl]]
        tp = hp["slot_tp0"][sl].copy()
        last = hp["slot_last0"][sl].copy()
        ins = hp["slot_ins0"][sl].copy()

        def flush(h, x):
            flush_pieces(X[h], tu[h], tp[h], last[h], ins[h], x)
            last[h] = x

        def transition(t):
            x = bp[t]
            e0, e1 = hp["tr_off"][t], hp["tr_off"][t + 1]
            nrem = hp["tr_nrem"][t]
            for e in range(e0, e1):
                is_ins = (e - e0) >= nrem
                u = hp["ev_child"][e]
                path = hp["path"][hp["ev_path0"][e]:hp["ev_path0"][e + 1]]
                xc = X[u].copy()
                if mode == "branch" and not is_ins:
                    flush(u, x)
                for w in path:
                    if mode == "branch":
                        flush(w, x)
                    X[w] = X[w] + xc if is_ins else X[w] - xc
                if is_ins:
                    tp[u] = tu[path[0]]
                    ins[u] = x
                else:
                    tp[u] = -np.inf

        t0, t1 = ct0[c], ct0[c + 1]
        for t in range(t0, t1):
            if t > t0:
                transition(t)
            if mode == "site":
                ps = plan_sites
                for s in range(ps["tree_site0"][t], ps["tree_site0"][t + 1]):
                    wi = np.searchsorted(windows, ps["_site_pos"][s], side="right") - 1
                    m0, m1 = ps["site_mut0"][s], ps["site_mut0"][s + 1]
                    for al in range(1 if polarised else 0, ps["site_nalleles"][s]):
                        aw = total.copy() if al == 0 else np.zeros(K1)
                        for mm in range(m0, m1):
                            h = hp["mut_slot"][mm]
                            if ps["mut_allele"][mm] == al:
                                aw += X[h]
                            if ps["mut_pallele"][mm] == al:
                                aw -= X[h]
                        if 0 < aw[Ks] < n:
                            hist[(wi, 0) + coord(aw)] += 1.0 if polarised else 0.5
        if mode == "branch" and t1 < T:
            transition(t1)
    if mode == "branch":
        tlp = hp["tl_parent"]
        XL = np.zeros((len(tlp), K1))
        for s in range(n):
            u = s
            while u >= 0:
                XL[u] += weight[s]
                u = tlp[u]
        for q in range(len(tlp)):
            if tlp[q] >= 0:
                flush_pieces(XL[q], hp["tl_time"][q], hp["tl_time"][tlp[q]], hp["tl_last"][q],
                             hp["tl_ins"][q], L)
    span = np.diff(windows)
    return hist / span.reshape((W,) + (1,) * (hist.ndim - 1))


def run_hot_relvec(hp, bp, windows, weights, focal):
    """The genetic_relatedness_vector program of tsg_relvec.hip executed on
    the CPU, step by step as the kernels do it: per chunk the lazy push-down
    over the hot slots (k_relvec_chunk), writing E and the window-segment
    integrals I per slot; the first tree's D_0 (k_relvec_tree0_*); then per
    focal sample the prefix over chunks (k_relvec_esum, k_relvec_assemble).
    weights [n, K] as the kernels take them (already centred or not); focal
    are sample indexes.  Returns y [W, nf, K], before the final centring and
    span normalisation."""
    ct0, cs0, cl0 = hp["chunk_tree0"], hp["chunk_slot0"], hp["chunk_lev0"]
    C = len(ct0) - 1
    T = len(bp) - 1
    n, K = weights.shape
    W = len(windows) - 1
    E = [None] * C
    I = [dict() for _ in range(C)]
    for c in range(C):
        s0, s1 = cs0[c], cs0[c + 1]
        S = s1 - s0
        sp = hp["slot_par"][s0:s1]
        par = sp.astype(np.int64).copy()
        bl = hp["slot_bl0"][s0:s1].astype(np.float64).copy()
        t0, t1 = ct0[c], ct0[c + 1]
        ca, cb = bp[t0], bp[t1]
        xs = np.full(S, ca)
        w = np.zeros((S, K))
        v = np.zeros((S, K))
        top = np.zeros(S, bool)
        tlast = min(t1, T - 1)
        if t0 + 1 <= tlast:
            e0, e1 = hp["tr_off"][t0 + 1], hp["tr_off"][tlast + 1]
            top[hp["ev_child"][e0:e1]] = True
            top[hp["path"][hp["ev_path0"][e0]:hp["ev_path0"][e1]]] = True
        ss = hp["sample_slot"][c * n:(c + 1) * n]
        for s in range(n):
            if ss[s] != NO_SLOT:
                w[ss[s]] += weights[s]
        lev = hp["lev_slot0"][cl0[c]:cl0[c + 1]]
        nlev = len(lev) - 1
        for j in range(nlev - 1):
            for h in range(lev[j], lev[j + 1]):
                if par[h] >= 0:
                    w[par[h]] += w[h]
        d0 = np.zeros((S, K))
        for j in range(nlev - 1, -1, -1):
            for h in range(lev[j], lev[j + 1]):
                d0[h] = (d0[par[h]] if par[h] >= 0 else 0.0) + bl[h] * w[h]
        seg0 = ca
        bi = 0
        while bi <= W and windows[bi] <= ca:
            bi += 1

        def readout(x, m):
            nonlocal seg0, v
            v += (bl * (x - xs))[:, None] * w
            xs[:] = x
            if 0 <= m < W and x > seg0:
                r = np.zeros((S, K))
                for h in range(S):
                    u = h
                    while u >= 0:
                        r[h] += v[u]
                        u = par[u]
                I[c][m] = r - (x - seg0) * d0
            v = np.zeros((S, K))
            seg0 = x

        def boundaries_to(x):
            nonlocal bi
            while bi <= W and windows[bi] <= x:
                readout(windows[bi], bi - 1)
                bi += 1

        def transition(t):
            x = bp[t]
            e0, e1 = hp["tr_off"][t], hp["tr_off"][t + 1]
            nrem = hp["tr_nrem"][t]
            for e in range(e0, e1):
                ins = (e - e0) >= nrem
                u = hp["ev_child"][e]
                sg = 1.0 if ins else -1.0
                if not ins:
                    v[u] += bl[u] * (x - xs[u]) * w[u]
                    xs[u] = x
                    bl[u] = 0.0
                    par[u] = -1
                path = hp["path"][hp["ev_path0"][e]:hp["ev_path0"][e + 1]]
                for p in path:
                    v[p] += bl[p] * (x - xs[p]) * w[p]
                    xs[p] = x
                    v[u] -= sg * v[p]
                    w[p] += sg * w[u]
                if ins:
                    xs[u] = x
                    par[u] = path[0]
                    bl[u] = hp["ev_bl"][e]

        for t in range(t0 + 1, t1):
            boundaries_to(bp[t])
            transition(t)
        boundaries_to(cb)
        if seg0 < cb:
            readout(cb, bi - 1)
        if t1 < T:
            transition(t1)
        Ec = np.zeros((S, K))
        for h in range(S):
            u = h
            while u >= 0:
                Ec[h] += bl[u] * w[u]
                u = par[u]
        Ec -= d0
        for h in range(S):
            if top[h]:
                continue
            q = sp[h]
            while q >= 0 and not top[q]:
                q = sp[q]
            Ec[h] = Ec[q] if q >= 0 else 0.0
            for m in I[c]:
                I[c][m][h] = I[c][m][q] if q >= 0 else 0.0
        E[c] = Ec
    # first tree
    t0p, t0b = hp["t0_parent"], hp["t0_bl"]
    z0 = np.zeros((len(t0p), K))
    for s in range(n):
        u = s
        while u >= 0:
            z0[u] += weights[s]
            u = t0p[u]
    y = np.zeros((W, len(focal), K))
    for j, s in enumerate(focal):
        P = np.zeros(K)
        u = s
        while u >= 0 and t0p[u] >= 0:
            P += t0b[u] * z0[u]
            u = t0p[u]
        for c in range(C):
            h = hp["sample_slot"][c * n + s]
            ca, cb = bp[ct0[c]], bp[ct0[c + 1]]
            for m in range(W):
                sa, sb = max(ca, windows[m]), min(cb, windows[m + 1])
                if not sa < sb:
                    continue
                y[m, j] += (sb - sa) * P
                if h != NO_SLOT:
                    y[m, j] += I[c][m][h]
            if h != NO_SLOT:
                P += E[c][h]
    return y
