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
