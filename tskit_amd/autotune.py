"""Chunk-count autotuner for the sweep kernels.

The sweep cuts a tree sequence's trees into chunks, one wavefront each; a
chunk is a serial chain of edge events, so more, shorter chunks run faster
as long as their hot slots stay small enough for several chunks per CU.
The best events-per-chunk depends on the tree sequence's shape, and on the
host work per chunk that every call also pays (window ranges per chunk and
query).  ``tune(ts)`` times a probe query (diversity and divergence of two
sample sets over ten windows, branch mode, one fused batch) end to end at a
few candidate chunk counts, keeps the fastest and records it,
keyed by the shape (log2 buckets of samples and edges), in a JSON cache;
``apply(ts)`` sets a tree sequence's chunk count from the cache.  The cache
file is ``$TSKIT_AMD_TUNE_CACHE`` or ``~/.cache/tskit_amd/tune.json``.
"""
import json
import math
import os

import numpy as np

CANDIDATES = (32, 64, 128, 256, 512)  # edge events per chunk


def cache_path():
    return os.environ.get("TSKIT_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "tskit_amd",
                                       "tune.json"))


def shape_key(ts):
    """The shape bucket a tuned setting applies to."""
    n = max(ts.num_samples, 1)
    e = max(ts.num_edges, 1)
    return f"n2^{int(round(math.log2(n)))}_e2^{int(round(math.log2(e)))}"


def load(path=None):
    try:
        with open(path or cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(table, path=None):
    path = path or cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def chunks_for(ts, events_per_chunk):
    """Chunk count giving about `events_per_chunk` edge events per chunk
    (each edge is inserted and removed once), at least one."""
    return max(1, min(ts.num_trees, int(round(2 * ts.num_edges / events_per_chunk))))


def _probe(ts, reps):
    """Wall-clock milliseconds of the probe batch (host work, sweep and
    finish), the fastest of `reps` runs after a warm one."""
    import time

    n = ts.num_samples
    smp = ts.samples()
    half = max(1, n // 2)
    sets = [smp[:half], smp[half:]] if n >= 2 else [smp, smp]
    w = np.linspace(0, ts.sequence_length, 11)

    def run():
        b = ts.batch()
        b.diversity(sets, windows=w, mode="branch")
        b.divergence(sets, windows=w, mode="branch")
        b.run()

    run()  # warm the workspace
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        run()
        best = min(best, time.perf_counter() - t)
    return best * 1e3


def tune(ts, candidates=CANDIDATES, reps=3, path=None, table=None):
    """Times the probe at each candidate events-per-chunk, applies the
    fastest to ts and records it in the cache (or in `table`, when given,
    without writing).  Returns {"events_per_chunk": best, "chunks": ...,
    "probe_ms": {candidate: ms}}."""
    own = table is None
    if own:
        table = load(path)
    times = {}
    for epc in candidates:
        ts.ll_tree_sequence.set_num_chunks(chunks_for(ts, epc))
        times[epc] = _probe(ts, reps)
    best = min(times, key=times.get)
    ts.ll_tree_sequence.set_num_chunks(chunks_for(ts, best))
    entry = {"events_per_chunk": int(best), "probe_ms": {str(k): v for k, v in times.items()}}
    table[shape_key(ts)] = entry
    if own:
        save(table, path)
    return dict(entry, chunks=chunks_for(ts, best))


def tune_batch(tss, run, candidates=CANDIDATES, reps=2):
    """One setting for tree sequences that are swept together (one launch
    per mode over all of them, run_batches): there the GPU is full whatever
    the chunk count, so the total work of the launch, which grows with the
    chunks (each seeds its start counts from every sample), decides, not one
    tree sequence's latency.  Sets every tree sequence of `tss` to each
    candidate events-per-chunk, times `run()` (which sweeps them together)
    by wall clock with the device synchronised, keeps the fastest and applies
    it.  Returns {"events_per_chunk": best, "seconds": {candidate: s}}."""
    import time

    import torch

    times = {}
    for epc in candidates:
        for ts in tss:
            ts.ll_tree_sequence.set_num_chunks(chunks_for(ts, epc))
            ts.ll_tree_sequence.upload()
        run()  # warm the workspaces at this plan
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            run()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t)
        times[epc] = best
    pick = min(times, key=times.get)
    for ts in tss:
        ts.ll_tree_sequence.set_num_chunks(chunks_for(ts, pick))
        ts.ll_tree_sequence.upload()
    run()  # warm the workspaces at the chosen plan
    torch.cuda.synchronize()
    return {"events_per_chunk": int(pick), "seconds": {str(k): v for k, v in times.items()}}


def apply(ts, table=None, path=None):
    """Sets ts's chunk count from the cache; returns the events per chunk
    used, or None when the shape was never tuned (the library's default
    plan stays)."""
    table = load(path) if table is None else table
    entry = table.get(shape_key(ts))
    if entry is None:
        return None
    ts.ll_tree_sequence.set_num_chunks(chunks_for(ts, entry["events_per_chunk"]))
    return entry["events_per_chunk"]
