"""Serving loop: tree sequences resident on the GPU, statistics requests
batched into fused sweeps.

``StatServer`` keeps a set of tree sequences uploaded and answers requests
(``Request``: a tree-sequence id, a statistic name and its keyword
arguments, as for ``TreeSequence``).  Waiting requests are grouped by (tree
sequence, mode) and packed into launches of at most ``max_columns`` distinct
sample sets and ``max_outputs`` outputs, admitted while the launch's
estimated device memory stays within ``launch_budget`` bytes; a request that
does not fit starts the next launch.  Allele frequency spectra run one per
launch (they are not fused).  Up to ``max_inflight`` launches run
concurrently on their own streams, and a launch is admitted only while the
device memory the launches in flight are estimated to need, plus its own,
stays within ``memory_budget`` (by default the device's free memory when the
server starts, less ``reserve_bytes``); a launch that does not fit waits
for others to finish (one always runs, so the loop makes progress).  With ``multi=True`` the fused launches
of a planning round are dealt into at most ``max_inflight`` jobs, and each
job runs all of its launches -- on any of the tree sequences -- as one sweep
launch per mode (``run_batches``).  ``one_per_launch=True, max_inflight=1``
serves every request alone, one launch at a time: the plain path the served
loop is measured against.

Results do not depend on the company a request keeps: a fused sweep computes
every output from its own columns in the same order as a sweep of that
request alone.
"""
import collections
import concurrent.futures
import time

import numpy as np

FUSABLE = {"diversity", "divergence", "Fst", "Tajimas_D", "f2", "f3", "f4", "Y1", "Y2", "Y3",
           "segregating_sites", "genetic_relatedness"}


class Request:
    """One statistics request: ``getattr(ts, stat)(**kwargs)`` on tree
    sequence ``ts_id``."""

    __slots__ = ("rid", "ts_id", "stat", "kwargs", "t_arrival", "t_start", "t_done", "result",
                 "error", "launch", "sets")

    def __init__(self, rid, ts_id, stat, kwargs, t_arrival=0.0):
        self.rid = rid
        self.ts_id = ts_id
        self.stat = stat
        self.kwargs = kwargs
        self.t_arrival = t_arrival
        self.t_start = None
        self.t_done = None
        self.result = None
        self.error = None  # the exception the request's launch raised, if any
        self.launch = None
        self.sets = None   # the sample sets as a list of arrays (set on submit)

    @property
    def mode(self):
        return self.kwargs.get("mode", "site")

    def normalise(self, ts):
        """The request's sample sets as a list of arrays, as the statistic
        reads them: none means every sample, a flat list one set."""
        sets = self.kwargs.get("sample_sets")
        if sets is None:
            sets = [ts.samples()]
        elif len(sets) > 0 and np.ndim(sets[0]) == 0:
            sets = [sets]
        self.sets = [np.asarray(x, dtype=np.int64) for x in sets]

    def shape(self):
        """(sample sets, outputs, windows) of the request."""
        sets = self.sets if self.sets is not None else self.kwargs.get("sample_sets")
        k = len(sets) if sets is not None else 1
        idx = self.kwargs.get("indexes")
        if self.stat in ("diversity", "segregating_sites", "Y1", "Tajimas_D"):
            m = k
        elif self.stat == "Fst":
            m = k + (len(idx) if idx is not None else 1)
        else:
            m = len(idx) if idx is not None else 1
        w = self.kwargs.get("windows")
        nw = 1 if w is None or isinstance(w, str) else len(w) - 1
        return k, m, nw


def launch_bytes(k, m, nw, chunks, samples=0):
    """Device bytes one fused launch needs for a query of m outputs over nw
    windows on a plan of `chunks` chunks: window partials (chunks + windows)
    x m, the combined result and finals, the per-chunk branch sums, and the
    (sample, column) seeds and per-(chunk, query) window tables."""
    return 8 * m * (3 * chunks + 3 * nw) + 64 * k + 16 * samples + 16 * chunks


def device_free_bytes():
    """Free device memory of the current GPU (bytes), or None without one."""
    try:
        import torch

        if torch.cuda.is_available():
            return int(torch.cuda.mem_get_info()[0])
    except Exception:  # noqa: BLE001 -- no device: no memory admission
        return None
    return None


class StatServer:
    """Continuous batching over resident tree sequences.

    Up to ``max_inflight`` launches run at once, each from its own host
    thread on its own device stream (the library leases every call a
    workspace with its own stream), so launches on different tree sequences
    -- or on the same one -- overlap on the device.  Requests that arrive
    while launches are in flight are admitted as soon as a slot frees."""

    def __init__(self, tree_sequences, launch_budget=2 << 30, max_columns=64, max_outputs=4096,
                 one_per_launch=False, max_inflight=8, multi=False, memory_budget=None,
                 reserve_bytes=1 << 30):
        self.ts = list(tree_sequences)
        for ts in self.ts:
            ts.ll_tree_sequence.upload()
        self.resident_bytes = sum(ts.ll_tree_sequence.get_device_bytes() for ts in self.ts)
        if memory_budget is None:
            free = device_free_bytes()
            memory_budget = None if free is None else max(0, free - int(reserve_bytes))
        self.memory_budget = memory_budget
        self.inflight_bytes = 0
        self.deferred = 0  # launches that waited for device memory
        self.chunks = [ts.ll_tree_sequence.get_hot_plan_sizes()["chunks"] for ts in self.ts]
        self.launch_budget = launch_budget
        self.max_columns = max_columns
        self.max_outputs = max_outputs
        self.one_per_launch = one_per_launch
        self.max_inflight = max(1, int(max_inflight))
        self.multi = bool(multi) and not one_per_launch
        self.pool = (concurrent.futures.ThreadPoolExecutor(max_workers=self.max_inflight)
                     if self.max_inflight > 1 else None)
        self.queue = collections.deque()
        self.launches = 0

    def close(self):
        if self.pool is not None:
            self.pool.shutdown(wait=True)
            self.pool = None

    def submit(self, req):
        req.normalise(self.ts[req.ts_id])
        self.queue.append(req)

    def _pack(self, reqs, chunks):
        """Greedy packing of one group into launches under the column,
        output and memory limits; requests keep their arrival order."""
        launches, cur, cols, outs, mem = [], [], set(), 0, 0
        for r in reqs:
            k, m, nw = r.shape()
            # a set's identity: its sorted samples, as bytes (hashed in C)
            keys = {np.sort(s).tobytes() for s in r.sets}
            b = launch_bytes(k, m, nw, chunks, sum(len(x) for x in r.sets))
            full = cur and (len(cols | keys) > self.max_columns or outs + m > self.max_outputs
                            or mem + b > self.launch_budget)
            if full or self.one_per_launch and cur:
                launches.append(cur)
                cur, cols, outs, mem = [], set(), 0, 0
            cur.append(r)
            cols |= keys
            outs += m
            mem += b
        if cur:
            launches.append(cur)
        return launches

    def _plan(self, pending, njobs=None):
        """Jobs for the pending requests, each a list of launches: fused
        groups per (tree sequence, mode), each allele frequency spectrum
        alone; with ``multi`` the fused launches are dealt into at most
        ``njobs`` jobs (default ``max_inflight``), else one launch a job."""
        groups = collections.OrderedDict()
        alone = []
        for r in pending:
            if r.stat in FUSABLE:
                groups.setdefault((r.ts_id, r.mode), []).append(r)
            else:
                alone.append([r])
        fused = []
        for (tid, _), reqs in groups.items():
            fused.extend(self._pack(reqs, self.chunks[tid]))
        if not self.multi:
            return [[launch] for launch in fused] + [[a] for a in alone]
        nj = max(1, min(njobs or self.max_inflight, len(fused)))
        jobs = [fused[j::nj] for j in range(nj)]
        return [j for j in jobs if j] + [[a] for a in alone]

    def _run(self, job):
        """Runs one job; an error (a request the library rejects) is stored
        on the job's requests, whose results stay None, and does not stop
        the other jobs of the round."""
        from .trees import run_batches

        t0 = time.perf_counter()
        reqs = [r for launch in job for r in launch]
        for r in reqs:
            r.t_start = t0
        try:
            if len(job) == 1 and len(job[0]) == 1 and job[0][0].stat not in FUSABLE:
                r = job[0][0]
                r.result = getattr(self.ts[r.ts_id], r.stat)(**r.kwargs)
            else:
                batches, pend = [], []
                for launch in job:
                    b = self.ts[launch[0].ts_id].batch()
                    pend.extend(getattr(b, r.stat)(**r.kwargs) for r in launch)
                    batches.append(b)
                if len(batches) == 1:
                    batches[0].run()
                else:
                    run_batches(batches)
                for r, p in zip(reqs, pend):
                    r.result = p.result()
        except Exception as exc:  # noqa: BLE001 -- reported per request
            for r in reqs:
                r.result = None
                r.error = exc
        t1 = time.perf_counter()
        for r in reqs:
            r.t_done = t1
        return reqs

    def job_bytes(self, job):
        """Estimated device bytes of a job's launches."""
        total = 0
        for launch in job:
            chunks = self.chunks[launch[0].ts_id]
            for r in launch:
                k, m, nw = r.shape()
                total += launch_bytes(k, m, nw, chunks, sum(len(x) for x in r.sets))
        return total

    def admits(self, job, inflight):
        """A job may start when its memory fits beside the jobs in flight (or
        nothing is in flight)."""
        if self.memory_budget is None or inflight == 0:
            return True
        return self.inflight_bytes + self.job_bytes(job) <= self.memory_budget

    def _start(self, job):
        self.launches += 1
        for launch in job:
            for r in launch:
                r.launch = self.launches
        need = self.job_bytes(job)
        self.inflight_bytes += need
        if self.pool is None:
            f = concurrent.futures.Future()
            f.set_result(self._run(job))
        else:
            f = self.pool.submit(self._run, job)
        f.bytes = need
        return f

    def _finish(self, f):
        self.inflight_bytes -= getattr(f, "bytes", 0)
        return f.result()

    def step(self):
        """Serves every request queued now, admitting jobs by slot and device
        memory; returns the requests completed."""
        if not self.queue:
            return []
        pending = list(self.queue)
        self.queue.clear()
        work = collections.deque(self._plan(pending))
        inflight, done = set(), []
        while work or inflight:
            while work and len(inflight) < self.max_inflight and self.admits(work[0],
                                                                             len(inflight)):
                inflight.add(self._start(work.popleft()))
            if work and len(inflight) < self.max_inflight:
                self.deferred += 1
            fin, inflight = concurrent.futures.wait(
                inflight, return_when=concurrent.futures.FIRST_COMPLETED)
            for f in fin:
                done.extend(self._finish(f))
        return done

    def run_open_loop(self, requests, arrivals):
        """Replays `requests` arriving at `arrivals` (seconds from now) with
        continuous batching: whenever a launch slot is free, the requests
        that have arrived are packed into launches and started.  Returns the
        wall time."""
        order = np.argsort(arrivals, kind="stable")
        t_begin = time.perf_counter()
        i = 0
        n = len(order)
        served = 0
        inflight = set()
        while served < n:
            now = time.perf_counter() - t_begin
            while i < n and arrivals[order[i]] <= now:
                r = requests[order[i]]
                r.t_arrival = t_begin + arrivals[order[i]]
                self.submit(r)
                i += 1
            free = self.max_inflight - len(inflight)
            if free > 0 and self.queue:
                work = self._plan(list(self.queue), free)
                self.queue.clear()
                started = 0
                for job in work:
                    if started < free and self.admits(job, len(inflight)):
                        inflight.add(self._start(job))
                        started += 1
                    else:  # no slot or no memory yet: wait in the queue
                        if started < free:
                            self.deferred += 1
                        for launch in job:
                            self.queue.extend(launch)
                free -= started
            if inflight:
                wait = None if i >= n else max(0.0, arrivals[order[i]] - now)
                if free > 0 and wait is not None:
                    wait = min(wait, 1e-3)
                done, inflight = concurrent.futures.wait(
                    inflight, timeout=wait, return_when=concurrent.futures.FIRST_COMPLETED)
                for f in done:
                    served += len(self._finish(f))
            elif i < n:
                time.sleep(max(0.0, min(1e-3, arrivals[order[i]] - now)))
        return time.perf_counter() - t_begin


def latency_stats(requests):
    """p50 / p99 / mean latency (ms) from arrival to completion."""
    lat = np.array([(r.t_done - r.t_arrival) * 1e3 for r in requests])
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean())}


def random_requests(tree_sequences, count, seed, max_sets=64, max_set_size=10000,
                    max_windows=10000, stats=("diversity", "divergence", "Fst", "f2", "f3", "f4",
                                              "allele_frequency_spectrum")):
    """Requests drawn uniformly: a tree sequence, a statistic, K sample sets
    (disjoint, 10 .. max_set_size samples, clipped to fit), windows and a
    mode.  f2/f3/f4 and divergence take random index tuples, Fst all pairs
    of its first (up to 8) sets; the AFS takes 1 or 2 sets of at most 300
    samples over at most 10 windows."""
    rng = np.random.default_rng(seed)
    out = []
    for rid in range(count):
        tid = int(rng.integers(len(tree_sequences)))
        ts = tree_sequences[tid]
        n = ts.num_samples
        stat = stats[int(rng.integers(len(stats)))]
        need = {"f2": 2, "f3": 3, "f4": 4, "divergence": 2, "Fst": 2}.get(stat, 1)
        k = int(rng.integers(need, max_sets + 1))
        if stat == "allele_frequency_spectrum":
            k = int(rng.integers(1, 3))
        if stat == "Fst":
            k = min(k, 8)
        size = int(rng.integers(10, max_set_size + 1))
        if stat == "allele_frequency_spectrum":  # prod(n_k + 1) bins per window
            size = min(size, 300)
        size = max(2, min(size, n // k))
        k = min(k, n // size)
        perm = rng.permutation(ts.samples())
        sets = [perm[j * size:(j + 1) * size] for j in range(k)]
        L = ts.sequence_length
        nw = int(rng.integers(1, max_windows + 1))
        windows = np.linspace(0, L, nw + 1)
        mode = "branch" if rng.integers(2) else "site"
        kw = {"sample_sets": sets, "windows": windows, "mode": mode}
        if stat in ("divergence", "f2", "f3", "f4"):
            t = {"divergence": 2, "f2": 2, "f3": 3, "f4": 4}[stat]
            m = int(rng.integers(1, 9))
            kw["indexes"] = [tuple(int(v) for v in rng.choice(k, t, replace=k < t))
                             for _ in range(m)]
        elif stat == "Fst":
            kw["indexes"] = [(a, b) for a in range(k) for b in range(a + 1, k)]
        if stat == "allele_frequency_spectrum":
            kw["windows"] = np.linspace(0, L, min(nw, 10) + 1)
        out.append(Request(rid, tid, stat, kw))
    return out
