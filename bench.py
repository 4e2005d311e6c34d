"""Benchmark of the tree-sequence statistics hot path on one MI355X.

Workload (the scale configuration): a Hudson-coalescent tree sequence with
1e5 haploid samples, L = 1e8, Ne = 1e4, r = mu = 1e-8 (seed 2), K = 8
disjoint random sample sets of 1000 samples and 1000 equal windows.  One
step computes, for branch and site mode, Fst over all 28 pairs, Tajima's D
of every set and f4 over sets (0,1;2,3), as one fused sweep per mode (one
batch; the site sweep's host work overlaps the branch sweep).

Metric: edge x sample-set updates per second.  Every statistic query of a
step counts 2 * E * K_q updates (each edge inserted and removed once, for
each of the K_q sample sets the query takes), whatever the implementation
fuses or repeats internally; the CPU baseline counts the same.

Dense phase (unless --no-dense): the site-mode relatedness matrix of 5000
samples over ~2e5 sites on the int8 matrix cores and the joint 2-set allele
frequency spectrum (200 + 300 samples, 100 windows), timed, with errors
against the CPU when the CPU leg runs.

Replicate phase (unless --no-replicates): 256 generated tree sequences,
plain, batched per replicate, and one sweep launch per mode over all of them
(run_batches).  Node phase: node-mode diversity of the 8 sets on the scale
tree sequence.

Served metric (unless --no-serve): 8 resident tree sequences -- the one
above, one of the 1000-sample shape (L = 1e6, seed 1) and six of 2e3..7e4
samples (L = 2e6) -- and requests drawn uniformly (a tree sequence;
diversity, divergence, Fst, f2, f3, f4 or the AFS; 1-64 sample sets of
10-1e4 samples; 1-1e4 windows; a mode).  Reported, on the 8 and on the
1e5-sample sequence alone: queries/s of the served form (fused launches, up
to 8 in flight, admitted by device memory) and of the plain form (one
request per launch, one at a time) with every request waiting, and p50/p99
latency of both under Poisson arrivals at 0.25, 0.5, 1 and 2 times the plain
capacity; also the capacity of the multi-sequence form (each job one sweep
launch per mode over several tree sequences).

CPU baseline (unless --no-cpu): the reference C library (oracle/_ref, built
from the reference's sources) on the host's cores as window-partitioned
processes -- each clips the tree sequence to its block of windows and runs
the step's statistics there -- with the clip + init and statistics times
reported separately, and the same step in one process beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--small] [--no-serve]
       [--no-dense] [--no-cpu]
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(small):
    if small:
        return dict(name="1000 samples, L=1e6 (seed 1), K=2x500, W=10", n=1000, L=1e6, seed=1,
                    K=2, set_size=500, W=10)
    return dict(name="1e5 samples, L=1e8, Ne=1e4, r=mu=1e-8 (seed 2), K=8x1000, W=1000",
                n=100000, L=1e8, seed=2, K=8, set_size=1000, W=1000)


def step_queries(cfg):
    K = cfg["K"]
    pairs = [(i, j) for i in range(K) for j in range(i + 1, K)]
    quads = [(0, 1, 2, 3)] if K >= 4 else [(0, 1, 0, 1)]
    return pairs, quads


def algorithmic_bytes(E, N, S, n, queries, W):
    """Input bytes one fused sweep must read plus output it writes:
    24 B/edge (left, right, parent, child) + 8 B/edge (insertion/removal
    index) + 8 B/node (time) + 12 B/site (position + mutation node), and per
    query n * ceil(K/8) (sample-set bitmask) + 8 (W+1) + 8 W M."""
    b = 24 * E + 8 * E + 8 * N + 12 * S
    for K, M in queries:
        b += n * ((K + 7) // 8) + 8 * (W + 1) + 8 * W * M
    return b


def run_gpu_step(ts, sets, windows, pairs, quads):
    """One step: Fst (all pairs), Tajima's D and f4, fused per mode (one
    batch holding both modes: the site sweep's host work overlaps the
    branch sweep)."""
    b = ts.batch()
    pend = {}
    for mode in ("branch", "site"):
        pend[mode] = (b.Fst(sets, indexes=pairs, windows=windows, mode=mode),
                      b.Tajimas_D(sets, windows=windows, mode=mode),
                      b.f4(sets, indexes=quads, windows=windows, mode=mode))
    b.run()
    return {mode: tuple(p.result() for p in ps) for mode, ps in pend.items()}


def max_rel_err(gpu, cpu):
    """Per statistic and mode (keys "Fst_branch", ...): (a) the largest
    |gpu - cpu| of an output over the windows divided by that output's
    largest |cpu| (the normaliser of sign-changing statistics: Fst and
    Tajima's D of sets drawn from one population, and f4, scatter around 0),
    and (b) the largest elementwise |gpu - cpu| / |cpu| over the values with
    |cpu| >= 1e-3 of that output's largest (below that the relative error
    measures the cancellation of 1 - ratio, not the sweep)."""
    per_output, elementwise = {}, {}
    for key, c in cpu.items():
        c = np.asarray(c, dtype=np.float64)
        g = np.asarray(gpu[key], dtype=np.float64).reshape(c.shape)
        fin = np.isfinite(c)
        a = np.where(fin, np.abs(c), 0.0)
        scale = np.maximum(np.max(a, axis=0, keepdims=True), 1e-300)
        d = np.where(fin, np.abs(g - c), 0.0)
        per_output[key] = float(np.max(d / scale))
        big = fin & (a >= 1e-3 * scale)
        elementwise[key] = float(np.max(d[big] / a[big])) if np.any(big) else 0.0
    return per_output, elementwise


SERVE_SHAPES = [(1000, 1e6, 1), (2000, 2e6, 51), (5000, 2e6, 52), (10000, 2e6, 53),
                (20000, 2e6, 54), (40000, 2e6, 55), (70000, 2e6, 56)]


def served_phase(big_ts, count, seed):
    """Capacity and latency of the serving loop (see the module docstring):
    the served form (fused launches, up to 8 in flight on their own streams)
    against the plain path (one request per launch, one launch at a time),
    on the 8 resident tree sequences and on the 1e5-sample one alone."""
    import tskit_amd as ta
    from tskit_amd import serve

    t0 = time.perf_counter()
    tss = [ta.simulate_tables(n, L, 1e4, 1e-8, 1e-8, sd).tree_sequence()
           for n, L, sd in SERVE_SHAPES]
    tss.append(big_ts)
    log(f"serve: 8 resident tree sequences generated in {time.perf_counter() - t0:.1f}s")
    forms = {"served": dict(max_inflight=8), "multi": dict(max_inflight=8, multi=True),
             "plain": dict(one_per_launch=True, max_inflight=1)}

    def capacity(pool, form, n_req, sd):
        srv = serve.StatServer(pool, **forms[form])
        reqs = serve.random_requests(pool, n_req, sd)
        t = time.perf_counter()
        for r in reqs:
            r.t_arrival = t
            srv.submit(r)
        srv.step()
        wall = time.perf_counter() - t
        srv.close()
        return n_req / wall, srv.launches, srv.resident_bytes

    out = {"requests": count, "resident": len(tss)}
    for name, pool, n_req in [("mix8", tss, count), ("big", [big_ts], max(count // 3, 40))]:
        capacity(pool, "served", max(n_req // 4, 20), seed + 100)  # warm buffers
        qs, ls, resident = capacity(pool, "served", n_req, seed)
        qm, lm, _ = capacity(pool, "multi", n_req, seed)
        qp, lp, _ = capacity(pool, "plain", n_req, seed)
        log(f"serve {name}: capacity served {qs:.1f} q/s ({ls} launches), multi {qm:.1f} q/s "
            f"({lm} jobs), plain {qp:.1f} q/s ({lp} launches): {qs / qp:.2f}x, "
            f"{qm / qp:.2f}x")
        res = {"requests": n_req, "resident_bytes": resident, "qps_served": qs,
               "qps_plain": qp, "speedup": qs / qp, "launches_served": ls,
               "launches_plain": lp, "qps_multi": qm, "speedup_multi": qm / qp,
               "jobs_multi": lm}
        # Poisson arrivals at a sweep of offered rates (fractions of the plain
        # capacity): p50 / p99 latency and achieved rate of both forms
        sweep = []
        for frac in (0.25, 0.5, 1.0, 2.0):
            offered = frac * qp
            rng = np.random.default_rng(seed + 1)
            arrivals = np.cumsum(rng.exponential(1.0 / offered, n_req))
            point = {"offered_qps": offered, "offered_frac_of_plain": frac}
            for form in ["served", "plain"]:
                srv = serve.StatServer(pool, **forms[form])
                reqs = serve.random_requests(pool, n_req, seed + 2)
                wall = srv.run_open_loop(reqs, arrivals)
                srv.close()
                st = serve.latency_stats(reqs)
                point[f"p50_ms_{form}"] = st["p50_ms"]
                point[f"p99_ms_{form}"] = st["p99_ms"]
                point[f"achieved_qps_{form}"] = n_req / wall
                log(f"serve {name}: offered {offered:.1f} q/s, {form}: p50 {st['p50_ms']:.2f} ms "
                    f"p99 {st['p99_ms']:.2f} ms, achieved {n_req / wall:.1f} q/s")
            sweep.append(point)
        res["latency_sweep"] = sweep
        out[name] = res
    return out, tss


def served_accuracy(tss, om, kind, count=16, seed=9):
    """Served results (fused launches) against the CPU implementation on the
    smaller resident tree sequences: max |error| / max |value| per output."""
    from tskit_amd import serve

    small = [i for i, t in enumerate(tss) if t.num_samples <= 20000]
    sub = [tss[i] for i in small]
    reqs = serve.random_requests(sub, count, seed, stats=("diversity", "divergence", "f2",
                                                          "f3", "f4"))
    srv = serve.StatServer(sub)
    for r in reqs:
        srv.submit(r)
    srv.step()
    impls = {}
    worst = 0.0
    for r in reqs:
        if r.ts_id not in impls:
            tc = sub[r.ts_id].tables
            impls[r.ts_id] = om.Reference(tc) if kind == "reference" else om.Oracle(tc)
        kw = r.kwargs
        exp = impls[r.ts_id].stat(r.stat, kw["sample_sets"], kw.get("indexes"), kw["windows"],
                                  kw["mode"])
        got = np.asarray(r.result, dtype=np.float64).reshape(exp.shape)
        fin = np.isfinite(exp)
        scale = np.broadcast_to(np.maximum(np.max(np.abs(np.where(fin, exp, 0)), axis=0), 1e-300),
                                exp.shape)
        if np.any(fin):
            worst = max(worst, float(np.max(np.abs(got - exp)[fin] / scale[fin])))
    return worst


DENSE = dict(n=5000, L=2e7, seed=3, mu=2.75e-8, afs_sets=(200, 300), afs_windows=100)
INT8_PEAK_TOPS = 5000.0


def dense_phase(om, with_cpu):
    """The dense workload: the site-mode relatedness matrix of 5000 samples
    (each its own set) over ~2e5 biallelic sites (L = 2e7, seed 3, mu scaled
    to give 2e5 sites) on the int8 matrix cores, and the joint 2-set allele
    frequency spectrum (sets of 200 and 300 samples, 100 windows), site and
    branch mode."""
    import torch

    import tskit_amd as ta

    d = DENSE
    tc = ta.simulate_tables(d["n"], d["L"], 1e4, 1e-8, d["mu"], d["seed"])
    ts = tc.tree_sequence()
    ts.ll_tree_sequence.upload()
    S = tc.num_sites
    out = {"workload": f"{d['n']} samples, L={d['L']:g}, {S} sites (seed {d['seed']}); "
                       f"AFS sets {d['afs_sets']}, W={d['afs_windows']}",
           "sites": S}

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps, r

    t_rel, rel = timed(lambda: ts.genetic_relatedness_matrix())
    # GEMM-equivalent ops (2 n^2 S) and the ops the int8 SYRK executes: the
    # upper-triangular 128 x 128 tiles (diagonal ones whole) over the sites
    # padded to a multiple of 64
    ops = 2.0 * d["n"] ** 2 * S
    T = (d["n"] + 127) // 128
    executed = 2.0 * (T * (T + 1) // 2) * 128 * 128 * ((S + 63) // 64 * 64)
    out["relatedness_matrix_ms"] = t_rel * 1e3
    out["relatedness_int_tops_end_to_end"] = ops / t_rel / 1e12
    out["relatedness_frac_of_int8_peak_end_to_end"] = ops / t_rel / 1e12 / INT8_PEAK_TOPS
    out["relatedness_syrk_executed_ops"] = executed
    t_geno, geno = timed(lambda: ts.genotype_matrix(), reps=2)
    out["genotype_matrix_ms"] = t_geno * 1e3
    out["genotype_matrix_bytes"] = int(geno.nbytes)
    perm = np.random.default_rng(d["seed"]).permutation(d["n"])
    a, b = d["afs_sets"]
    sets = [perm[:a], perm[a:a + b]]
    w = np.linspace(0, d["L"], d["afs_windows"] + 1)
    afs = {}
    for mode in ("site", "branch"):
        t_afs, afs[mode] = timed(lambda: ts.allele_frequency_spectrum(sets, windows=w, mode=mode))
        out[f"afs_{mode}_ms"] = t_afs * 1e3
    log(f"dense: relatedness {d['n']}^2 x {S} sites {t_rel * 1e3:.1f} ms "
        f"({out['relatedness_int_tops_end_to_end']:.0f} int TOP/s end to end); AFS site "
        f"{out['afs_site_ms']:.2f} ms, branch {out['afs_branch_ms']:.2f} ms")
    if with_cpu:
        impl = om.Reference(tc) if om.reference_available() else om.Oracle(tc)
        exp = impl.afs(sets, w, None, "site", True, False)[:, 0]
        out["afs_site_bit_exact"] = bool(np.array_equal(afs["site"], exp))
        exp = impl.afs(sets, w, None, "branch", True, False)[:, 0]
        out["afs_branch_max_rel_err"] = float(np.max(np.abs(afs["branch"] - exp)) /
                                              np.max(np.abs(exp)))
        # relatedness: a 64 x 64 block of the centred closed form
        # K = C - r 1^T / n - 1 r^T / n + sum(r) / n^2 (C = G^T G, r = G^T k,
        # k the allele counts), span-normalised, in FP64 on the CPU from the
        # oracle's genotype matrix
        G = om.Oracle(tc).genotypes()
        n = G.shape[1]
        k = G.sum(axis=1, dtype=np.int64).astype(np.float64)
        r = np.zeros(n)
        for s0 in range(0, G.shape[0], 16384):
            r += G[s0:s0 + 16384].T.astype(np.float64) @ k[s0:s0 + 16384]
        idx = np.arange(64)
        Gb = G[:, idx].astype(np.float64)
        blk = (Gb.T @ Gb - r[idx, None] / n - r[None, idx] / n + (k @ k) / n**2) / d["L"]
        out["relatedness_block_max_rel_err"] = float(
            np.max(np.abs(rel[np.ix_(idx, idx)] - blk)) / np.max(np.abs(blk)))
        del G
        if om.reference_available():
            ref = om.Reference(tc)
            t = time.perf_counter()
            gexp = ref.genotype_matrix(True)
            out["genotype_matrix_cpu_reference_s"] = time.perf_counter() - t
            out["genotype_matrix_bit_exact"] = bool(np.array_equal(geno, gexp))
            del gexp
            log(f"dense: genotype matrix {out['genotype_matrix_ms']:.0f} ms (reference decode "
                f"{out['genotype_matrix_cpu_reference_s']:.1f} s, 1 core), bit-exact "
                f"{out['genotype_matrix_bit_exact']}")
        del geno
        log(f"dense: AFS site bit-exact {out['afs_site_bit_exact']}, branch max rel err "
            f"{out['afs_branch_max_rel_err']:.2e}, relatedness 64x64 block max rel err "
            f"{out['relatedness_block_max_rel_err']:.2e}")
    return out


def replicate_workload(count=256, seed0=1000):
    """The replicate batch: `count` tree sequences with log-uniform 1e2..1e4
    samples and L 1e5..1e7 (r = mu = 1e-8, seeds seed0..), each with 1-32
    random disjoint sample sets of 10-5000 samples (clipped to fit) and
    1-2000 equal windows; per replicate and mode the queries are diversity of
    every set and divergence of consecutive sets."""
    import tskit_amd as ta

    rng = np.random.default_rng(seed0)
    reps = []
    for i in range(count):
        n = int(round(10 ** rng.uniform(2, 4)))
        L = float(round(10 ** rng.uniform(5, 7)))
        tc = ta.simulate_tables(n, L, 1e4, 1e-8, 1e-8, seed0 + i)
        k = int(rng.integers(1, 33))
        size = int(rng.integers(10, 5001))
        size = max(2, min(size, n // k))
        k = max(1, min(k, n // size))
        perm = rng.permutation(n)
        sets = [perm[j * size:(j + 1) * size] for j in range(k)]
        w = np.linspace(0, L, int(rng.integers(1, 2001)) + 1)
        pairs = [(j, j + 1) for j in range(k - 1)]
        reps.append(dict(tc=tc, sets=sets, windows=w, pairs=pairs))
    return reps


def replicate_phase(om, with_cpu, inflight=4):
    """The replicate batch, plain (every query its own launch, one launch at a
    time) and batched (one fused launch per replicate and mode, `inflight`
    launches in flight on their own streams)."""
    import concurrent.futures

    import torch

    import tskit_amd as ta

    from tskit_amd import autotune

    t0 = time.perf_counter()
    reps = replicate_workload()
    for r in reps:
        r["ts"] = r["tc"].tree_sequence()
    t_setup = time.perf_counter() - t0
    # chunk length by the autotuner: the first replicate of each shape bucket
    # is tuned, the others take its setting
    t0 = time.perf_counter()
    table = {}
    for r in reps:
        if autotune.shape_key(r["ts"]) not in table:
            autotune.tune(r["ts"], reps=2, table=table)
        autotune.apply(r["ts"], table=table)
        r["ts"].ll_tree_sequence.upload()
    t_tune = time.perf_counter() - t0
    log(f"replicates: autotuned {len(table)} shape buckets in {t_tune:.1f}s: "
        + ", ".join(f"{k}:{v['events_per_chunk']}" for k, v in sorted(table.items())))
    edges = sum(r["tc"].num_edges for r in reps)
    updates = 0
    for r in reps:
        k = len(r["sets"])
        updates += 2 * (2 * r["tc"].num_edges * k) * (2 if r["pairs"] else 1)

    def plain(r):
        ts, out = r["ts"], {}
        for mode in ("branch", "site"):
            out[("diversity", mode)] = ts.diversity(r["sets"], windows=r["windows"], mode=mode)
            if r["pairs"]:
                out[("divergence", mode)] = ts.divergence(r["sets"], indexes=r["pairs"],
                                                          windows=r["windows"], mode=mode)
        return out

    def fused(r, mode):
        b = r["ts"].batch()
        p1 = b.diversity(r["sets"], windows=r["windows"], mode=mode)
        p2 = b.divergence(r["sets"], indexes=r["pairs"], windows=r["windows"],
                          mode=mode) if r["pairs"] else None
        b.run()
        return {("diversity", mode): p1.result(),
                ("divergence", mode): p2.result() if p2 else None}

    for r in reps[:8]:  # warm buffers
        plain(r)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for r in reps:
        plain(r)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t
    results = [dict() for _ in reps]
    with concurrent.futures.ThreadPoolExecutor(max_workers=inflight) as pool:
        list(pool.map(lambda r: fused(r, "branch"), reps[:8]))
        t = time.perf_counter()
        futs = [(i, pool.submit(fused, r, mode)) for i, r in enumerate(reps)
                for mode in ("branch", "site")]
        for i, f in futs:
            results[i].update(f.result())
        torch.cuda.synchronize()
        t_batched = time.perf_counter() - t

    def multi_batches():
        bs, pend = [], []
        for r in reps:
            b = r["ts"].batch()
            pm = {}
            for mode in ("branch", "site"):
                pm[("diversity", mode)] = b.diversity(r["sets"], windows=r["windows"], mode=mode)
                if r["pairs"]:
                    pm[("divergence", mode)] = b.divergence(r["sets"], indexes=r["pairs"],
                                                            windows=r["windows"], mode=mode)
            bs.append(b)
            pend.append(pm)
        return bs, pend

    # the single launch over all replicates is tuned jointly (its total work
    # decides, not one replicate's latency); a different chunking rounds
    # differently, so it is compared with the batched results to a tolerance
    t0 = time.perf_counter()
    joint = autotune.tune_batch([r["ts"] for r in reps],
                                lambda: ta.run_batches(multi_batches()[0]))
    t_tune_multi = time.perf_counter() - t0
    bs, pend = multi_batches()
    torch.cuda.synchronize()
    t = time.perf_counter()
    ta.run_batches(bs)
    torch.cuda.synchronize()
    t_multi = time.perf_counter() - t
    multi_diff = 0.0
    for i in range(len(reps)):
        for k in pend[i]:
            a = np.asarray(pend[i][k].result(), dtype=np.float64)
            b = np.asarray(results[i][k], dtype=np.float64)
            fin = np.isfinite(b)
            scale = max(float(np.max(np.abs(b[fin]))), 1e-300) if np.any(fin) else 1.0
            if not np.array_equal(np.isfinite(a), fin):
                multi_diff = float("inf")
            elif np.any(fin):
                multi_diff = max(multi_diff, float(np.max(np.abs(a[fin] - b[fin]))) / scale)
    out = {"replicates": len(reps), "edges": edges,
           "samples_min_max": [int(min(len(r["ts"].samples()) for r in reps)),
                               int(max(len(r["ts"].samples()) for r in reps))],
           "setup_s": t_setup, "plain_s": t_plain, "batched_s": t_batched,
           "speedup": t_plain / t_batched, "updates_per_s_batched": updates / t_batched,
           "queries_per_s_batched": sum(2 * (2 if r["pairs"] else 1) for r in reps) / t_batched,
           "form": f"one fused launch per replicate and mode, {inflight} in flight",
           "multi_s": t_multi, "multi_speedup": t_plain / t_multi,
           "multi_form": "one sweep launch per mode over all replicates, one launch per "
                         "finish phase over all of them, one copy back",
           "tuned_events_per_chunk": {k: v["events_per_chunk"] for k, v in table.items()},
           "tune_s": round(t_tune, 2),
           "multi_events_per_chunk": joint["events_per_chunk"],
           "multi_tune_seconds": joint["seconds"], "multi_tune_s": round(t_tune_multi, 2),
           "multi_max_diff_vs_batched": multi_diff}
    log(f"replicates: {len(reps)} tree sequences ({edges} edges) plain {t_plain * 1e3:.0f} ms, "
        f"batched {t_batched * 1e3:.0f} ms ({t_plain / t_batched:.2f}x), one launch per mode "
        f"{t_multi * 1e3:.0f} ms ({t_plain / t_multi:.2f}x; jointly tuned "
        f"{joint['events_per_chunk']} events per chunk: "
        + ", ".join(f"{k}: {v * 1e3:.0f} ms" for k, v in joint["seconds"].items())
        + f"; max diff vs batched {multi_diff:.2e})")
    if with_cpu:
        worst = 0.0
        for i in range(0, len(reps), 32):
            r = reps[i]
            impl = om.Reference(r["tc"]) if om.reference_available() else om.Oracle(r["tc"])
            for (stat, mode), got in results[i].items():
                if got is None:
                    continue
                idx = r["pairs"] if stat == "divergence" else None
                exp = impl.stat(stat, r["sets"], idx, r["windows"], mode)
                got = np.asarray(got, dtype=np.float64).reshape(exp.shape)
                fin = np.isfinite(exp)
                scale = max(float(np.max(np.abs(exp[fin]))), 1e-300) if np.any(fin) else 1.0
                if np.any(fin):
                    worst = max(worst, float(np.max(np.abs(got[fin] - exp[fin]))) / scale)
        out["max_rel_err_vs_cpu_sampled"] = worst
        log(f"replicates: max rel err of 8 sampled replicates vs CPU {worst:.2e}")
    return out


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "k_sweep_traffic.json")
# the branch sweep of the scale workload: hot slots with 16-bit counts, one
# output per lane, 8 hop groups per column (K = 8)
SWEEP_KERNEL = "k_sweep<0, 3, 1, 8>"


def pmc_summary(fetch_csv, write_csv):
    """HBM bytes per dispatch of the branch sweep from rocprofv3 counter
    CSVs: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) averaged over its
    timed dispatches (the first, the warm-up, first touches the workspace),
    each scaled by bytes / counter of the calibration kernel (torch.neg over
    COPY_BYTES: reads and writes COPY_BYTES)."""
    import csv

    def means(path):
        acc = {}
        for r in csv.DictReader(open(path)):
            per = acc.setdefault(r["Kernel_Name"], {})
            d = int(r["Dispatch_Id"])
            per[d] = per.get(d, 0.0) + float(r["Counter_Value"])
        out = {}
        for k, per in acc.items():
            vals = [per[d] for d in sorted(per)]
            if len(vals) > 1 and "neg" not in k:
                vals = vals[1:]
            out[k] = sum(vals) / len(vals)
        return out

    out = {"workload": workload(False)["name"], "kernel": SWEEP_KERNEL}
    for tag, path in [("fetch", fetch_csv), ("write", write_csv)]:
        m = means(path)
        sweep = [v for k, v in m.items() if SWEEP_KERNEL in k]
        calib = max((v for k, v in m.items() if "neg" in k), default=None)
        if not sweep or calib is None:
            raise SystemExit(f"{path}: no {SWEEP_KERNEL} or calibration dispatches")
        scale = COPY_BYTES / (calib * 1024.0)
        out[f"{tag}_kib_per_dispatch"] = sweep[0]
        out[f"calibration_{tag}_kib"] = calib
        out[f"{tag}_scale"] = scale
        out[f"{tag}_bytes"] = sweep[0] * 1024.0 * scale
    out["traffic_bytes"] = out["fetch_bytes"] + out["write_bytes"]
    os.makedirs(os.path.dirname(TRAFFIC_JSON), exist_ok=True)
    with open(TRAFFIC_JSON, "w") as f:
        json.dump(out, f, indent=1)
    return out


def measured_traffic(cfg_name):
    """traffic_bytes of the committed counter summary for this workload."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("traffic_bytes") if t.get("workload") == cfg_name else None


def node_phase(ts, tc, sets, om, with_cpu, reps=3):
    """Node-mode diversity of the K sets over the whole sequence (one value per
    node and set): end-to-end time per call, its node kernels' time from HIP
    events, and the largest relative error against the CPU."""
    import torch

    ll = ts.ll_tree_sequence
    got = ts.diversity(sets, mode="node")  # first call uploads the node tables
    torch.cuda.synchronize()
    ll.set_profiling(True)
    t = time.perf_counter()
    for _ in range(reps):
        got = ts.diversity(sets, mode="node")
    torch.cuda.synchronize()
    t_call = (time.perf_counter() - t) / reps
    kern_ms, launches = ll.get_profile()
    ll.set_profiling(False)
    out = {"stat": "diversity", "mode": "node", "sets": len(sets), "shape": list(got.shape),
           "call_ms": round(t_call * 1e3, 3),
           "kernels_ms": round(kern_ms / max(launches, 1), 3)}
    msg = (f"node: diversity of {len(sets)} sets at {got.shape[0]} nodes {t_call * 1e3:.1f} ms "
           f"per call, node kernels {out['kernels_ms']:.2f} ms")
    if with_cpu:
        kind = "reference" if om.reference_available() else "port"
        impl = om.Reference(tc) if kind == "reference" else om.Oracle(tc)
        t = time.perf_counter()
        exp = impl.stat("diversity", sets, None, None, "node")[0]
        out["cpu_s"] = round(time.perf_counter() - t, 3)
        out["cpu_kind"] = kind
        out["max_rel_err"] = float(np.max(np.abs(got - exp)) / np.max(np.abs(exp)))
        msg += f"; CPU ({kind}) {out['cpu_s']:.2f} s, max rel err {out['max_rel_err']:.2e}"
    log(msg)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--small", action="store_true", help="the 1000-sample workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--chunks", type=int, default=0, help="override the chunk count")
    ap.add_argument("--no-serve", action="store_true", help="skip the served phase")
    ap.add_argument("--no-replicates", action="store_true",
                    help="skip the 256-replicate batch phase")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the relatedness-matrix / joint AFS phase")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of "
                         "`bench.py --no-serve --no-cpu --copy-calibration` into "
                         "profiles/k_sweep_traffic.json and exit")
    ap.add_argument("--inflight", type=int, default=1,
                    help="steps in flight from that many host threads (default 1: one after "
                         "the other)")
    ap.add_argument("--dump", metavar="NPZ",
                    help="write the last timed step's statistics to an .npz file")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the steps, copy 1 GiB device to device three times (the "
                         "reference kernel that scales FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary), indent=1))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import tskit_amd as ta
    from oracle import oracle as om

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    cfg = workload(args.small)
    t0 = time.perf_counter()
    tc = ta.simulate_tables(cfg["n"], cfg["L"], 1e4, 1e-8, 1e-8, cfg["seed"])
    t_gen = time.perf_counter() - t0
    log(f"generated {tc.num_nodes} nodes {tc.num_edges} edges {tc.num_sites} sites "
        f"in {t_gen:.1f}s digest {tc.column_digest()[:16]}")
    t0 = time.perf_counter()
    ts = tc.tree_sequence()
    if args.chunks:
        ts.ll_tree_sequence.set_num_chunks(args.chunks)
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    ts.ll_tree_sequence.upload()
    torch.cuda.synchronize()
    t_upload = time.perf_counter() - t0
    ll = ts.ll_tree_sequence
    hp = ll.get_hot_plan_sizes()
    props = torch.cuda.get_device_properties(0)
    log(f"host init {t_init:.1f}s upload {t_upload:.2f}s trees {ts.num_trees} "
        f"hot chunks {hp['chunks']} max slots {hp['max_slots']} max levels {hp['max_levels']} "
        f"device bytes {ll.get_device_bytes() / 2**20:.0f} MiB "
        f"LDS/block {getattr(props, 'shared_memory_per_block', 'n/a')}")

    sets = ta.random_sample_sets(cfg["n"], cfg["K"], cfg["set_size"], cfg["seed"])
    windows = np.linspace(0, cfg["L"], cfg["W"] + 1)
    pairs, quads = step_queries(cfg)
    K = cfg["K"]
    E, N, S = tc.num_edges, tc.num_nodes, tc.num_sites
    # queries per step (per mode): Fst, Tajima's D and f4, each over the K sets
    kq = [K, K, K]
    updates_per_step = 2 * sum(2 * E * k for k in kq)
    # the fused sweep of a mode updates each of the K distinct state columns
    # once per edge event: 2 modes x 2E x K column updates per step
    column_updates_per_step = 2 * 2 * E * K
    queries_per_step = 2 * len(kq)

    for i in range(args.warmup):
        run_gpu_step(ts, sets, windows, pairs, quads)
        log(f"warmup {i + 1}/{args.warmup} done (node store {ll.get_last_store()})")
    ll.set_profiling(True)
    # Python's cyclic garbage collector is off while the steps are timed (as
    # timeit does): its occasional pauses are tens of milliseconds
    gc.collect()
    gc.disable()
    torch.cuda.synchronize()
    if args.inflight > 1:
        # steps submitted from `inflight` host threads: one step's host work
        # (building and preparing its batch) overlaps another's device work;
        # each step is the same whole batch, with its own workspaces
        import concurrent.futures

        with concurrent.futures.ThreadPoolExecutor(max_workers=args.inflight) as pool:
            t0 = time.perf_counter()
            futs = [pool.submit(run_gpu_step, ts, sets, windows, pairs, quads)
                    for _ in range(args.steps)]
            for f in futs:
                gpu_out = f.result()
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - t0
    else:
        t0 = time.perf_counter()
        for i in range(args.steps):
            gpu_out = run_gpu_step(ts, sets, windows, pairs, quads)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    gc.enable()
    if args.dump:
        np.savez(args.dump, **{f"{name}_{mode}": np.asarray(v)
                               for mode, vals in gpu_out.items()
                               for name, v in zip(("Fst", "Tajimas_D", "f4"), vals)})
    sweep_ms, launches = ll.get_profile()
    per_mode = ll.get_profile_modes()
    ll.set_profiling(False)
    # the same steps two in flight (one step's host work overlapping the
    # other's device work): throughput only -- co-running kernels stretch
    # each other, so the roofline keeps the one-at-a-time launch times
    pipelined = None
    if args.inflight == 1 and not args.small:
        import concurrent.futures

        gc.disable()
        torch.cuda.synchronize()
        with concurrent.futures.ThreadPoolExecutor(max_workers=2) as pool:
            t1 = time.perf_counter()
            futs = [pool.submit(run_gpu_step, ts, sets, windows, pairs, quads)
                    for _ in range(args.steps)]
            for f in futs:
                f.result()
            torch.cuda.synchronize()
            e2 = time.perf_counter() - t1
        gc.enable()
        pipelined = {"ms_per_step": 1e3 * e2 / args.steps,
                     "value": updates_per_step * args.steps / e2, "steps_in_flight": 2}
        log(f"two steps in flight: {1e3 * e2 / args.steps:.2f} ms/step")
    ms_per_step = 1e3 * elapsed / args.steps
    value = updates_per_step * args.steps / elapsed
    mode_ms = {m: t / max(n, 1) for m, (t, n) in per_mode.items()}
    log(f"{args.steps} steps in {elapsed:.3f}s ({ms_per_step:.2f} ms/step), sweep kernel "
        f"branch {mode_ms['branch']:.3f} ms, site {mode_ms['site']:.3f} ms "
        f"({launches} launches)")

    # roofline of the dominant kernel (the fused branch-mode sweep,
    # k_sweep<0,...>): algorithmic bytes per launch / its average launch time
    # from HIP events on the library's stream
    per_launch_queries = [(K, K + len(pairs)), (K, 2 * K), (K, len(quads))]
    alg_bytes = algorithmic_bytes(E, N, S, cfg["n"], per_launch_queries, cfg["W"])
    avg_launch_s = mode_ms["branch"] / 1e3
    achieved = alg_bytes / avg_launch_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": None if args.small else measured_traffic(cfg["name"]),
                "kernel": "k_sweep<0, 3, 1, 8> (branch, fused batch)", "bytes_per_launch": alg_bytes,
                "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                "site_sweep_ms": round(mode_ms["site"], 4)}

    if args.copy_calibration:
        # an elementwise kernel reading and writing COPY_BYTES each
        src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log(f"copy calibration: 3 x {COPY_BYTES} bytes read and written")

    dense = None
    if not args.no_dense and not args.small:
        dense = dense_phase(om, not args.no_cpu)
    replicates = None
    if not args.no_replicates and not args.small:
        replicates = replicate_phase(om, not args.no_cpu)

    node = None
    if not args.small:
        node = node_phase(ts, tc, sets, om, not args.no_cpu)

    served, serve_tss = None, None
    if not args.no_serve and not args.small:
        served, serve_tss = served_phase(ts, args.serve_requests, 7)

    cpu = None
    errs, errs_elem = {}, {}
    if not args.no_cpu:
        import functools

        from oracle import partitioned

        kind = "reference" if om.reference_available() else "port"
        step = functools.partial(partitioned.scale_step, sets=sets, pairs=pairs, quads=quads)
        if kind == "reference":
            # the reference as P window-partitioned processes on the host's
            # cores (general_stat needs windows spanning the sequence:
            # each process clips to its block and drops the outer windows)
            cpu_out, t_cpu, t_init, t_stat, P = partitioned.run(tc, windows, step)
        else:
            t0 = time.perf_counter()
            cpu_out = step(om.Oracle(tc), windows)
            t_cpu, t_init, P = time.perf_counter() - t0, 0.0, 1
            t_stat = t_cpu
        t0 = time.perf_counter()
        impl = om.Reference(tc) if kind == "reference" else om.Oracle(tc)
        step(impl, windows)
        t_one = time.perf_counter() - t0
        cpu = {"value": updates_per_step / t_cpu, "unit": "edge×sample-set updates/s",
               "cores": P,
               "kind": ("reference, window-partitioned" if kind == "reference" else "port"),
               "seconds": round(t_cpu, 3), "clip_init_seconds_max": round(t_init, 3),
               "stat_seconds_max": round(t_stat, 3), "single_thread_seconds": round(t_one, 3),
               "sample": f"one full step (Fst over 28 pairs, Tajima's D, f4; branch and site) on "
                         f"the same tree sequence: {P} processes, each the reference on its "
                         f"block of windows (edges and sites clipped to it); wall time of "
                         f"the parallel run"}
        gpu_dict = {f"{name}_{mode}": v for mode, vals in gpu_out.items()
                    for name, v in zip(("Fst", "Tajimas_D", "f4"), vals)}
        errs, errs_elem = max_rel_err(gpu_dict, cpu_out)
        log(f"cpu step {t_cpu:.2f}s wall on {P} processes (clip+init max {t_init:.2f}s, stats "
            f"max {t_stat:.2f}s; one process {t_one:.2f}s, {kind}); max error / max|value| per "
            f"output {max(errs.values()):.2e}, elementwise rel err {max(errs_elem.values()):.2e}")
        if served is not None:
            served["max_rel_err_vs_cpu"] = served_accuracy(serve_tss, om, kind)
            log(f"serve: max rel err of sampled served results vs CPU "
                f"{served['max_rel_err_vs_cpu']:.2e}")

    line = {
        "metric": "edge×sample-set updates/s; served stat queries/s + p99 at 1e-6 rel err",
        "value": value,
        "unit": "edge×sample-set updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32/uint16 counts; fp64 sums, combine and finals",
        "data": "synthetic (own Hudson coalescent + infinite-sites generator)",
        "config": {"workload": cfg["name"], "edges": E, "trees": ts.num_trees, "sites": S,
                   "nodes": N, "queries_per_step": queries_per_step,
                   "chunks": hp["chunks"], "node_store": ["hbm", "lds", "lds-hot", "lds-hot16"][
                       ll.get_last_store()], "steps_in_flight": args.inflight},
        "pipelined": pipelined,
        "queries_per_s": queries_per_step * args.steps / elapsed,
        "column_updates_per_s": column_updates_per_step * args.steps / elapsed,
        "value_note": "value counts 2 E K_q updates for each of the 3 queries of a mode (the "
                      "reference's sweeps, one per query); column_updates_per_s counts each "
                      "distinct state column once per edge event",
        "roofline": roofline,
        "cpu_baseline": cpu,
        "max_rel_err_vs_cpu": errs,
        "max_rel_err_vs_cpu_note": "per output: max |gpu - cpu| over windows / max |cpu| "
                                   "of that output",
        "max_elementwise_rel_err_vs_cpu": errs_elem,
        "served": served,
        "dense": dense,
        "node": node,
        "replicates": replicates,
        "setup_s": {"generate": round(t_gen, 2), "host_init": round(t_init, 2),
                    "upload": round(t_upload, 3)},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
