"""GPU parity at the full sizes of the benchmark workloads.

* scale workload (1e5 samples, L=1e8, seed 2; K=8 sets of 1000, 1000
  windows): diversity, all-pairs divergence, Fst over the 28 pairs, Tajima's
  D and f4, both modes, against the oracle;
* replicate batch (256 generated tree sequences, 1-32 sets, 1-2000 windows):
  one sweep launch per mode over all of them equals every replicate run
  alone, bit for bit, and sampled replicates equal the oracle;
* dense workload (5000 samples, ~2e5 biallelic sites, seed 3): the site-mode
  relatedness matrix (symmetry, zero row sums of the double centring, 64x64
  blocks against the FP64 closed form) and the joint AFS of sets of 200 and
  300 samples over 100 windows (site bit-exact, branch <= 1e-6);
* served mix (8 resident tree sequences, 1e3..1e5 samples): every request's
  served result equals its result alone (the branch AFS, whose bins take
  FP64 atomics, to 1e-12), and sampled requests -- Fst and the AFS among
  them -- equal the oracle.

Tolerances: as tests/test_gpu_parity.py states them -- non-negative
statistics elementwise relative <= 1e-6; sign-changing statistics and the
composites Fst and Tajima's D (which change sign: random sets of one
population have Fst and D near 0) <= 1e-6 of the output's largest |value|
over the windows; the per-statistic errors are printed.
"""
import numpy as np
import pytest

import cases
from cases import rel_err
from oracle.oracle import Oracle, fst, tajimas_d

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _say(*args):
    print("[case]", *args, flush=True)


def _per_output_err(got, exp):
    """max |got - exp| / max over windows |exp|, per output (last axis)."""
    exp = np.asarray(exp, dtype=np.float64)
    fin = np.where(np.isfinite(exp), np.abs(exp), 0.0)
    scale = np.max(fin, axis=tuple(range(exp.ndim - 1)), keepdims=True)
    return rel_err(got, exp, scale=np.maximum(scale, 1e-300))


@pytest.fixture(scope="module")
def scale(require_gpu):
    tc = cases.generated(100000, 1e8, 2)
    ts = tc.tree_sequence()
    return tc, ts, Oracle(tc)


def test_scale_workload(scale):
    """Diversity and all-pairs divergence, both modes, elementwise <= 1e-6."""
    tc, ts, orc = scale
    sets = cases.sample_sets(100000, 8, 1000, 2)
    w = np.linspace(0, 1e8, 1001)
    pairs = [(i, j) for i in range(8) for j in range(i + 1, 8)]
    for mode in ["branch", "site"]:
        _say("scale workload", mode)
        e1 = rel_err(ts.diversity(sets, windows=w, mode=mode),
                     orc.stat("diversity", sets, None, w, mode))
        e2 = rel_err(ts.divergence(sets, indexes=pairs, windows=w, mode=mode),
                     orc.stat("divergence", sets, pairs, w, mode))
        _say(f"  diversity {e1:.2e} divergence {e2:.2e}")
        assert e1 <= TOL and e2 <= TOL


def test_scale_workload_composites(scale):
    """The benchmark step itself (one batch: Fst over the 28 pairs, Tajima's D
    of every set and f4 over (0,1;2,3), both modes) against the oracle's
    composites, formed as the reference's Python layer forms them
    (trees.py:10113-10125, 10008-10018)."""
    tc, ts, orc = scale
    sets = cases.sample_sets(100000, 8, 1000, 2)
    w = np.linspace(0, 1e8, 1001)
    pairs = [(i, j) for i in range(8) for j in range(i + 1, 8)]
    quads = [(0, 1, 2, 3)]
    b = ts.batch()
    pend = {m: (b.Fst(sets, indexes=pairs, windows=w, mode=m),
                b.Tajimas_D(sets, windows=w, mode=m),
                b.f4(sets, indexes=quads, windows=w, mode=m)) for m in ("branch", "site")}
    b.run()
    for mode in ("branch", "site"):
        _say("scale composites", mode)
        div = orc.stat("diversity", sets, None, w, mode)
        dvg = orc.stat("divergence", sets, pairs, w, mode)
        T = orc.stat("diversity", sets, None, w, mode, False)
        S = orc.stat("segregating_sites", sets, None, w, mode, False)
        exp = {"Fst": fst(div, dvg, pairs), "Tajimas_D": tajimas_d(T, S, [1000] * 8),
               "f4": orc.stat("f4", sets, quads, w, mode)}
        for name, p in zip(("Fst", "Tajimas_D", "f4"), pend[mode]):
            got = np.asarray(p.result(), dtype=np.float64).reshape(exp[name].shape)
            err = _per_output_err(got, exp[name])
            _say(f"  {name} {mode}: error / max|value| per output {err:.2e}")
            assert err <= TOL, (name, mode, err)


def test_replicate_batch():
    """The 256-replicate batch through run_batches (one sweep launch per mode
    over all replicates): bit-identical to each replicate's batch run alone;
    every 32nd replicate against the oracle."""
    import bench
    import tskit_amd as ta

    reps = bench.replicate_workload()
    _say("replicates generated", len(reps))
    for r in reps:
        r["ts"] = r["tc"].tree_sequence()

    def queue(r):
        b = r["ts"].batch()
        pend = {}
        for mode in ("branch", "site"):
            pend[("diversity", mode)] = b.diversity(r["sets"], windows=r["windows"], mode=mode)
            if r["pairs"]:
                pend[("divergence", mode)] = b.divergence(r["sets"], indexes=r["pairs"],
                                                          windows=r["windows"], mode=mode)
        return b, pend

    alone = []
    for r in reps:
        b, pend = queue(r)
        b.run()
        alone.append({k: np.asarray(p.result()) for k, p in pend.items()})
    _say("replicates run alone")
    queued = [queue(r) for r in reps]
    ta.run_batches([b for b, _ in queued])
    _say("replicates run as one launch per mode")
    for i, (_, pend) in enumerate(queued):
        for k, p in pend.items():
            assert np.array_equal(np.asarray(p.result()), alone[i][k], equal_nan=True), (i, k)
    worst = 0.0
    for i in range(0, len(reps), 32):
        r = reps[i]
        orc = Oracle(r["tc"])
        for (stat, mode), got in alone[i].items():
            idx = r["pairs"] if stat == "divergence" else None
            exp = orc.stat(stat, r["sets"], idx, r["windows"], mode)
            err = rel_err(got.reshape(exp.shape), exp)
            worst = max(worst, err)
            assert err <= TOL, (i, stat, mode, err)
    _say(f"  sampled replicates vs oracle: max elementwise rel err {worst:.2e}")


@pytest.fixture(scope="module")
def dense(require_gpu):
    import bench
    import tskit_amd as ta

    d = bench.DENSE
    tc = ta.simulate_tables(d["n"], d["L"], 1e4, 1e-8, d["mu"], d["seed"])
    return d, tc, tc.tree_sequence()


def test_dense_relatedness_matrix(dense):
    """5000 x 5000 site-mode relatedness over ~2e5 sites: symmetric, rows of
    the double-centred matrix sum to 0, and three 64x64 blocks (diagonal,
    off-diagonal, the last corner) equal the closed form
    K = (C - r 1'/n - 1 r'/n + k.k/n^2) / L in FP64 (C = G'G, r = G'k, k the
    allele counts), which is pinned to the reference's divergence-matrix
    route in tests/test_oracle.py."""
    d, tc, ts = dense
    n = d["n"]
    _say("dense relatedness", n, tc.num_sites)
    K = ts.genetic_relatedness_matrix()
    assert K.shape == (n, n)
    assert np.array_equal(K, K.T)
    rows = np.abs(K.sum(axis=1))
    assert np.max(rows) <= 1e-9 * np.max(np.abs(K)) * n, np.max(rows)
    G = Oracle(tc).genotypes()
    assert G.shape == (tc.num_sites, n)
    k = G.sum(axis=1, dtype=np.int64).astype(np.float64)
    r = np.zeros(n)
    for s0 in range(0, G.shape[0], 16384):
        r += G[s0:s0 + 16384].T.astype(np.float64) @ k[s0:s0 + 16384]
    for i0, j0 in [(0, 0), (640, 3200), (n - 64, n - 64)]:
        a, b = np.arange(i0, i0 + 64), np.arange(j0, j0 + 64)
        Ga, Gb = G[:, a].astype(np.float64), G[:, b].astype(np.float64)
        blk = (Ga.T @ Gb - r[a, None] / n - r[None, b] / n + (k @ k) / n**2) / d["L"]
        err = float(np.max(np.abs(K[np.ix_(a, b)] - blk)) / np.max(np.abs(blk)))
        _say(f"  block ({i0},{j0}) error / max|value| {err:.2e}")
        assert err <= 1e-10


def test_dense_joint_afs(dense):
    """Joint AFS of sets of 200 and 300 samples over 100 windows: site mode
    bit-exact, branch mode <= 1e-6 of the largest bin."""
    d, tc, ts = dense
    perm = np.random.default_rng(d["seed"]).permutation(d["n"])
    a, b = d["afs_sets"]
    sets = [perm[:a], perm[a:a + b]]
    w = np.linspace(0, d["L"], d["afs_windows"] + 1)
    orc = Oracle(tc)
    _say("dense joint AFS")
    got = ts.allele_frequency_spectrum(sets, windows=w, mode="site")
    exp = orc.afs(sets, w, None, "site", True, False)[:, 0]
    assert got.shape == (100, 201, 301)
    assert np.array_equal(got, exp)
    got = ts.allele_frequency_spectrum(sets, windows=w, mode="branch")
    exp = orc.afs(sets, w, None, "branch", True, False)[:, 0]
    err = rel_err(got, exp, np.max(np.abs(exp)))
    _say(f"  branch AFS error / max bin {err:.2e}")
    assert err <= TOL


def test_served_mix(scale):
    """The served workload: 8 resident tree sequences (the scale one, one of
    1000 samples and six of 2e3..7e4), requests drawn uniformly over
    diversity, divergence, Fst, f2, f3, f4 and the AFS with 1-64 sets of
    10-1e4 samples and 1-1e4 windows.  Served in fused launches, 8 in flight,
    each result equals the same request served alone; sampled requests on
    the tree sequences of <= 2e4 samples equal the oracle."""
    import bench
    import tskit_amd as ta
    from tskit_amd import serve

    tc_big, big, _ = scale
    tcs = [ta.simulate_tables(n, L, 1e4, 1e-8, 1e-8, sd) for n, L, sd in bench.SERVE_SHAPES]
    pool = [tc.tree_sequence() for tc in tcs] + [big]
    tcs.append(tc_big)
    reqs = serve.random_requests(pool, 120, 7)
    _say("served mix", len(reqs), "requests")
    srv = serve.StatServer(pool, max_inflight=8)
    for r in reqs:
        srv.submit(r)
    srv.step()
    srv.close()
    assert srv.launches < len(reqs)
    mixed = {r.rid: np.asarray(r.result) for r in reqs}
    alone = serve.random_requests(pool, 120, 7)
    srv = serve.StatServer(pool, one_per_launch=True, max_inflight=1)
    for r in alone:
        srv.submit(r)
    srv.step()
    srv.close()
    stats = set()
    for r in alone:
        got, exp = mixed[r.rid], np.asarray(r.result)
        stats.add(r.stat)
        if r.stat == "allele_frequency_spectrum" and r.mode == "branch":
            scale_ = max(float(np.max(np.abs(exp))), 1e-300)
            assert np.max(np.abs(got - exp)) <= 1e-12 * scale_, r.rid
        else:
            assert np.array_equal(got, exp, equal_nan=True), (r.rid, r.stat)
    assert {"Fst", "allele_frequency_spectrum", "f4"} <= stats
    _say("  mixed equals alone for", len(alone), "requests")
    orcs = {}
    checked = {}
    for r in alone:
        if pool[r.ts_id].num_samples > 20000 or checked.get(r.stat, 0) >= 3:
            continue
        checked[r.stat] = checked.get(r.stat, 0) + 1
        orc = orcs.setdefault(r.ts_id, Oracle(tcs[r.ts_id]))
        kw = r.kwargs
        sets, w, mode = kw["sample_sets"], kw["windows"], kw["mode"]
        got = np.asarray(mixed[r.rid], dtype=np.float64)
        if r.stat == "allele_frequency_spectrum":
            exp = orc.afs(sets, w, None, mode, True, False)[:, 0].reshape(got.shape)
            if mode == "site":
                assert np.array_equal(got, exp)
                continue
            err = rel_err(got, exp, np.max(np.abs(exp)))
        elif r.stat == "Fst":
            div = orc.stat("diversity", sets, None, w, mode)
            dvg = orc.stat("divergence", sets, kw["indexes"], w, mode)
            exp = fst(div, dvg, kw["indexes"])
            err = _per_output_err(got.reshape(exp.shape), exp)
        else:
            exp = orc.stat(r.stat, sets, kw.get("indexes"), w, mode)
            err = _per_output_err(got.reshape(exp.shape), exp)
        _say(f"  request {r.rid} {r.stat} {mode}: error {err:.2e}")
        assert err <= TOL, (r.rid, r.stat, err)
    assert {"Fst", "allele_frequency_spectrum"} <= set(checked)
