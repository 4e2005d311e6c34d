"""ORACLE -- test infrastructure only: the CPU baseline of bench.py.

The reference C library (oracle/_ref, compiled from the reference's sources)
run as P window-partitioned processes on the host's cores.  The reference's
general_stat requires windows that span the whole sequence
(c/tskit/trees.c:1255-1266), so each process takes a contiguous block
[a, b) of the windows, clips the edges and sites to it (what
TreeSequence.keep_intervals([[a, b]]) keeps, python/tskit/tables.py), builds
the reference tree sequence (indexes included, tables.c:11393-11461), runs
the statistics with windows [0, a, ..., b, L] and drops the two outer
windows.  Clip plus tree-sequence init, and the statistics, are timed
separately.  The processes are started fresh (spawn): they share nothing
with the parent but the clipped columns they are sent.
"""
import multiprocessing
import os
import time
import types

import numpy as np


def clip(tc, a, b):
    """The columns of tc with edges clipped to [a, b) and the sites (and
    their mutations) in [a, b); nodes unchanged."""
    left, right = np.asarray(tc.edge_left), np.asarray(tc.edge_right)
    keep = (right > a) & (left < b)
    pos = np.asarray(tc.site_position)
    skeep = (pos >= a) & (pos < b)
    site_new = np.cumsum(skeep) - 1
    msite = np.asarray(tc.mutation_site)
    mkeep = skeep[msite]
    mut_new = np.cumsum(mkeep) - 1
    mpar = np.asarray(tc.mutation_parent)[mkeep]
    mpar = np.where(mpar >= 0, mut_new[np.maximum(mpar, 0)], -1)
    anc, anc_off = np.asarray(tc.ancestral_state), np.asarray(tc.ancestral_state_offset)
    der, der_off = np.asarray(tc.derived_state), np.asarray(tc.derived_state_offset)

    def ragged(data, off, sel):
        lens = np.diff(off)[sel]
        starts = off[:-1][sel]
        parts = [data[s:s + n] for s, n in zip(starts, lens)]
        out = np.concatenate(parts) if parts else np.zeros(0, data.dtype)
        return out, np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)

    a_state, a_off = ragged(anc, anc_off, skeep)
    d_state, d_off = ragged(der, der_off, mkeep)
    return dict(
        sequence_length=float(tc.sequence_length), time_units=tc.time_units,
        node_flags=np.asarray(tc.node_flags), node_time=np.asarray(tc.node_time),
        edge_left=np.maximum(left[keep], a), edge_right=np.minimum(right[keep], b),
        edge_parent=np.asarray(tc.edge_parent)[keep], edge_child=np.asarray(tc.edge_child)[keep],
        site_position=pos[skeep], ancestral_state=a_state, ancestral_state_offset=a_off,
        mutation_site=site_new[msite[mkeep]].astype(np.int32),
        mutation_node=np.asarray(tc.mutation_node)[mkeep], mutation_parent=mpar.astype(np.int32),
        derived_state=d_state, derived_state_offset=d_off)


def block_windows(windows, i0, i1, L):
    """Windows [0, a, ..., b, L] for the block of windows i0 .. i1 - 1, and
    the rows of the result that belong to the block."""
    inner = list(windows[i0:i1 + 1])
    w = ([0.0] if inner[0] > 0 else []) + inner + ([L] if inner[-1] < L else [])
    first = 1 if inner[0] > 0 else 0
    return np.array(w, dtype=np.float64), first, first + (i1 - i0)


def _block(args):
    """One process: clip, init, then the statistics of `step` on its block."""
    from oracle import oracle as om

    cols, wins, r0, r1, step = args
    t0 = time.perf_counter()
    ref = om.Reference(types.SimpleNamespace(**cols))
    t1 = time.perf_counter()
    out = step(ref, wins)
    t2 = time.perf_counter()
    trimmed = {k: np.asarray(v)[r0:r1] for k, v in out.items()}
    return trimmed, t1 - t0, t2 - t1


def run(tc, windows, step, processes=None):
    """step(impl, windows) -> dict of [W, ...] arrays (one row per window),
    run on P window blocks in P processes.  Returns (dict of [W, ...]
    arrays, wall seconds, max clip + init seconds, max statistics seconds,
    P)."""
    W = len(windows) - 1
    P = int(processes or min(16, os.cpu_count() or 1))
    P = max(1, min(P, W))
    L = float(tc.sequence_length)
    bounds = [round(W * p / P) for p in range(P + 1)]
    jobs = []
    for p in range(P):
        i0, i1 = bounds[p], bounds[p + 1]
        wins, r0, r1 = block_windows(windows, i0, i1, L)
        jobs.append((clip(tc, windows[i0], windows[i1]), wins, r0, r1, step))
    ctx = multiprocessing.get_context("spawn")
    t = time.perf_counter()
    with ctx.Pool(P) as pool:
        parts = pool.map(_block, jobs)
    wall = time.perf_counter() - t
    out = {k: np.concatenate([p[0][k] for p in parts]) for k in parts[0][0]}
    return out, wall, max(p[1] for p in parts), max(p[2] for p in parts), P


def scale_step(impl, windows, sets, pairs, quads):
    """The benchmark step on the CPU, composites as the reference's Python
    layer forms them (two sweeps each, python/tskit/trees.py:10113-10125,
    10008-10018): Fst over `pairs`, Tajima's D of every set and f4 over
    `quads`, branch and site mode."""
    from oracle.oracle import fst, tajimas_d

    out = {}
    for mode in ("branch", "site"):
        div = impl.stat("diversity", sets, None, windows, mode)
        dvg = impl.stat("divergence", sets, pairs, windows, mode)
        T = impl.stat("diversity", sets, None, windows, mode, False)
        S = impl.stat("segregating_sites", sets, None, windows, mode, False)
        out[f"Fst_{mode}"] = fst(div, dvg, pairs)
        out[f"Tajimas_D_{mode}"] = tajimas_d(T, S, [len(s) for s in sets])
        out[f"f4_{mode}"] = impl.stat("f4", sets, quads, windows, mode)
    return out
