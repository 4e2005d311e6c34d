"""The serving loop: fused launches give the results of the same requests
served alone (bit for bit: a fused sweep adds every output's terms in the
same order; the branch AFS, whose bins take FP64 atomic adds in no fixed
order, to 1e-12 of its largest bin), the oracle's within the parity
tolerance, and the loop reports latencies."""
import numpy as np
import pytest

import cases
from oracle.oracle import Oracle

from tskit_amd import serve

pytestmark = pytest.mark.gpu


def _same(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


@pytest.fixture(scope="module")
def resident(require_gpu):
    tcs = [cases.generated(n, L, seed) for n, L, seed in [(200, 2e5, 21), (300, 1e5, 22),
                                                           (120, 3e5, 23)]]
    return tcs, [tc.tree_sequence() for tc in tcs]


@pytest.mark.parametrize("multi", [False, True])
def test_mixed_equals_alone(resident, multi):
    tcs, tss = resident
    reqs = serve.random_requests(tss, 60, seed=5, max_sets=8, max_set_size=60, max_windows=12)
    alone = serve.random_requests(tss, 60, seed=5, max_sets=8, max_set_size=60, max_windows=12)
    # fused, concurrent launches; with multi, one sweep launch per mode over
    # several tree sequences
    mixed = serve.StatServer(tss, max_inflight=4, multi=multi)
    for r in reqs:
        mixed.submit(r)
    done = mixed.step()
    mixed.close()
    print("[case] mixed launches", mixed.launches, flush=True)
    single = serve.StatServer(tss, one_per_launch=True, max_inflight=1)
    for r in alone:
        single.submit(r)
    single.step()
    assert len(done) == len(reqs)
    assert mixed.launches < single.launches == len(reqs)
    for a, b in zip(reqs, alone):
        if a.stat == "allele_frequency_spectrum" and a.mode == "branch":
            x, y = np.asarray(a.result), np.asarray(b.result)
            assert x.shape == y.shape
            assert np.max(np.abs(x - y)) <= 1e-12 * max(np.max(np.abs(y)), 1e-300)
        else:
            assert _same(a.result, b.result), (a.stat, a.kwargs.get("mode"))


def test_served_results_match_oracle(resident):
    tcs, tss = resident
    reqs = serve.random_requests(tss, 24, seed=6, max_sets=6, max_set_size=50, max_windows=8,
                                 stats=("diversity", "divergence", "f2", "f3", "f4"))
    srv = serve.StatServer(tss)
    for r in reqs:
        srv.submit(r)
    srv.step()
    orcs = [Oracle(tc) for tc in tcs]
    for r in reqs:
        kw = r.kwargs
        idx = kw.get("indexes")
        exp = orcs[r.ts_id].stat(r.stat, kw["sample_sets"], idx, kw["windows"], kw["mode"])
        got = np.asarray(r.result).reshape(exp.shape)
        scale = np.maximum(np.max(np.abs(exp), axis=0, keepdims=True), 1e-300)
        fin = np.isfinite(exp)
        assert np.array_equal(fin, np.isfinite(got))
        assert np.max(np.abs(got - exp)[fin] / np.broadcast_to(scale, exp.shape)[fin]) <= 1e-6


@pytest.mark.parametrize("multi", [False, True])
def test_open_loop_latencies(resident, multi):
    tcs, tss = resident
    reqs = serve.random_requests(tss, 40, seed=7, max_sets=4, max_set_size=40, max_windows=6)
    srv = serve.StatServer(tss, multi=multi)
    arrivals = np.cumsum(np.random.default_rng(1).exponential(1 / 400.0, len(reqs)))
    wall = srv.run_open_loop(reqs, arrivals)
    st = serve.latency_stats(reqs)
    assert wall > 0 and all(r.result is not None for r in reqs)
    assert 0 < st["p50_ms"] <= st["p99_ms"]


def test_default_and_flat_sets_and_errors(resident):
    """Requests with no sample sets (every sample) and with one flat set are
    packed like any other; a request the library rejects (an index out of
    range) gets its error, and the requests of the other launches their
    results."""
    tcs, tss = resident
    w = np.linspace(0, tss[0].sequence_length, 4)
    flat = tss[0].samples()[:50]
    reqs = [serve.Request(0, 0, "diversity", {"windows": w, "mode": "branch"}),
            serve.Request(1, 0, "diversity", {"sample_sets": flat, "windows": w, "mode": "site"}),
            serve.Request(2, 1, "divergence", {"sample_sets": [tss[1].samples()[:9],
                                                               tss[1].samples()[9:20]],
                                               "indexes": [(0, 5)], "windows": None,
                                               "mode": "site"}),
            serve.Request(3, 2, "diversity", {"sample_sets": [tss[2].samples()[:30]],
                                              "mode": "branch"})]
    srv = serve.StatServer(tss, max_inflight=2)
    for r in reqs:
        srv.submit(r)
    srv.step()
    srv.close()
    assert _same(reqs[0].result, tss[0].diversity(windows=w, mode="branch"))
    assert _same(reqs[1].result, tss[0].diversity(flat, windows=w, mode="site"))
    assert reqs[2].result is None and reqs[2].error is not None
    assert _same(reqs[3].result, tss[2].diversity([tss[2].samples()[:30]], mode="branch"))


def test_memory_admission_defers_launches(resident):
    """With a device-memory budget below two launches, launches wait for
    each other (counted in `deferred`) and every request is still served,
    with the results of an unconstrained server."""
    tcs, tss = resident
    reqs = serve.random_requests(tss, 30, seed=9, max_sets=6, max_set_size=50, max_windows=8,
                                 stats=("diversity", "divergence", "f2"))
    free = serve.StatServer(tss, max_inflight=4)
    for r in reqs:
        free.submit(r)
    free.step()
    free.close()
    tight = serve.random_requests(tss, 30, seed=9, max_sets=6, max_set_size=50, max_windows=8,
                                  stats=("diversity", "divergence", "f2"))
    srv = serve.StatServer(tss, max_inflight=4, memory_budget=1)
    for r in tight:
        srv.submit(r)
    done = srv.step()
    srv.close()
    assert len(done) == len(tight) and srv.deferred > 0 and srv.inflight_bytes == 0
    for a, b in zip(reqs, tight):
        assert _same(a.result, b.result)
