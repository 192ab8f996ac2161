"""Continuous-batching sort server on one GPU.

Sort jobs arrive over time.  The server admits a job into a device JobPool as
soon as its HBM footprint fits (its characters plus one u64 offset and one u32
length per string: the pool's capacities are the admission limit), keeps it
resident until its batch completes, and runs up to `inflight` batches at once,
each on its own HIP stream and host thread.  A worker takes every admitted job
that is waiting (first come first served, up to `max_batch_strings` strings,
at least one job), sorts all of them in one sequence of kernel launches (job k
of the batch is virtual PE k), computes each job's D and releases the job's
memory, which lets the admission thread bring in the next arrivals while the
other batch is still on the device.

The job stream follows the mixed workload of the batched/served benchmark:
sizes log-uniform, lengths uniform, D/N uniform in [0.1, 0.9], every job drawn
from the D/N-ratio generator (src/util/random_string_generator.hpp:166-287) or
its skewed variant (long/short strings plus exact duplicates).
"""
from __future__ import annotations

import math
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from ._lib import DssError
from .pool import JobPool

DSS_ENOMEM = -2


@dataclass
class JobSpec:
    jid: int
    kind: str            # "dton" or "skewed"
    n: int               # strings
    length: int          # string length (long strings for "skewed")
    dton: float          # D/N of the (long) strings
    seed: int
    arrival: float = 0.0  # seconds after the stream starts
    short_len: int = 0   # "skewed": length and D/N of the short strings
    short_dn: float = 0.0

    def add_to(self, pool: JobPool) -> int:
        if self.kind == "dton":
            return pool.add_dton(self.n, self.length, self.dton, self.seed)
        return pool.add_skewed(self.n, self.seed, long_len=self.length, short_len=self.short_len,
                               long_dn=self.dton, short_dn=self.short_dn)

    def max_bytes(self) -> int:
        """Upper bound of the job's characters (NULs included)."""
        if self.kind == "dton":
            k = max(int(self.length * self.dton), math.ceil(math.log(max(self.n, 2)) / math.log(26)))
            return self.n * (max(self.length, k) + 1)
        k = math.ceil(math.log(max(self.n, 2)) / math.log(255))
        return self.n * (max(self.length, int(self.length * self.dton), k) + 1)


def job_stream(num_jobs: int, n_min: int, n_max: int, len_min: int = 10, len_max: int = 1000,
               rate: float = 0.0, seed: int = 0x5EED0005) -> list[JobSpec]:
    """Seeded mixed job stream; arrivals form a Poisson process of `rate`
    jobs/s (rate 0: every job arrives at time 0)."""
    rng = np.random.default_rng(seed)
    arrivals = np.random.default_rng([seed, 1])  # jobs do not depend on the rate
    specs = []
    t = 0.0
    for j in range(num_jobs):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        n = min(max(n, n_min), n_max)
        length = int(rng.integers(len_min, len_max + 1))
        d = float(rng.uniform(0.1, 0.9))
        kind = ("dton", "dton", "skewed")[int(rng.integers(0, 3))]
        if rate > 0:
            t += float(arrivals.exponential(1.0 / rate))
        s = JobSpec(jid=j, kind=kind, n=n, length=length, dton=d, seed=seed * 7919 + j, arrival=t)
        if kind == "skewed":
            s.short_len = max(len_min, length // 10)
            s.short_dn = d / 3.0
        specs.append(s)
    return specs


@dataclass
class JobResult:
    jid: int
    n: int
    D: int = 0
    arrival: float = 0.0    # seconds since the run started
    admitted: float = 0.0
    started: float = 0.0
    done: float = 0.0
    batch_jobs: int = 0     # jobs in the batch that sorted this one
    perm: np.ndarray | None = field(default=None, repr=False)
    lcp: np.ndarray | None = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class Server:
    """`inflight` workers, each with its own batch set and stream.  Worker 0
    takes the oldest waiting large job (at least `large_strings` strings) on
    its own, else a batch of small jobs; the other workers only batch small
    jobs, so small jobs never wait behind a large one in the same batch.

    Admission is in arrival order, with backfill: while an arrived job does
    not fit the pool, later arrivals whose characters take at most
    `backfill_bytes` (default: 1/16 of the pool's character capacity) are
    admitted past it, so a large job waiting for memory does not hold up the
    small jobs behind it; the large job is admitted as soon as it fits."""

    def __init__(self, pool: JobPool, inflight: int = 2, max_batch_strings: int = 32 << 20,
                 large_strings: int | None = 4 << 20, reserve_strings: int = 0,
                 backfill_bytes: int | None = None):
        self.pool = pool
        self.inflight = inflight
        self.max_batch_strings = max_batch_strings
        self.large_strings = large_strings
        self.backfill_bytes = backfill_bytes
        # device buffers of every worker's batch set are grown to this many
        # strings when the set is created, so batches allocate nothing later
        self.reserve_strings = reserve_strings
        self._batches = {}  # worker -> Batch, kept across runs (same pool)

    def prepare(self, job: int):
        """Creates every worker's batch set (on any live job of the pool) and
        grows its buffers, so that no allocation falls into a timed run."""
        for w in range(self.inflight):
            if w not in self._batches:
                b = self.pool.batch([job])
                if self.reserve_strings:
                    b.set.reserve(self.reserve_strings, 65536)
                self._batches[w] = b

    def close(self):
        for b in self._batches.values():
            b.close()
        self._batches = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _take(self, ready, worker):
        """Removes and returns the next batch for `worker` from `ready`, or []
        if it has nothing to take.  Without a large-job threshold (or with
        one worker) batches are FIFO prefixes of the queue."""
        large = self.large_strings
        if large is None or self.inflight == 1:
            take, strings = [], 0
            while ready and (not take or (strings + ready[0][0].n <= self.max_batch_strings
                                          and len(take) < 65536)):
                if take and large is not None and ready[0][0].n >= large:
                    break
                item = ready.popleft()
                take.append(item)
                strings += item[0].n
                if large is not None and item[0].n >= large:
                    break
            return take
        if worker == 0:
            for k, (s, _) in enumerate(ready):
                if s.n >= large:
                    item = ready[k]
                    del ready[k]
                    return [item]
        take, strings, keep = [], 0, deque()
        while ready:
            item = ready.popleft()
            s = item[0]
            if s.n < large and (not take or (strings + s.n <= self.max_batch_strings and len(take) < 65536)):
                take.append(item)
                strings += s.n
            else:
                keep.append(item)
        keep.extend(ready)
        ready.clear()
        ready.extend(keep)
        return take

    def run(self, specs, realtime: bool = True, keep_results: bool = False,
            resident: dict | None = None) -> list[JobResult]:
        """Serves the stream; with realtime=False every job arrives at once.

        With `resident` (jid -> pool job id) the jobs are already in the pool:
        they are ready at once, nothing is admitted and nothing is freed (the
        batched form of jobs that are resident in HBM)."""
        pool = self.pool
        order = sorted(specs, key=lambda s: s.arrival)
        res = {s.jid: JobResult(jid=s.jid, n=s.n) for s in order}
        cv = threading.Condition()
        ready = deque()          # (spec, pool job id)
        state = {"admitting": True, "inflight_jobs": 0, "error": None}
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        backfill = self.backfill_bytes
        if backfill is None and resident is None:
            backfill = pool.usage()["bytes_capacity"] // 16

        def admit():
            try:
                pending = []  # arrived, not yet admitted, in arrival order
                k = 0
                while k < len(order) or pending:
                    # arrivals up to now join the pending list
                    while k < len(order) and (not realtime or order[k].arrival <= now()):
                        s = order[k]
                        res[s.jid].arrival = s.arrival if realtime else 0.0
                        pending.append(s)
                        k += 1
                    if resident is not None:
                        for s in pending:
                            res[s.jid].admitted = now()
                            with cv:
                                ready.append((s, resident[s.jid]))
                                cv.notify_all()
                        pending = []
                        continue
                    # admit in arrival order; behind a job that does not fit,
                    # only jobs of at most `backfill` characters
                    blocked, keep = None, []
                    for s in pending:
                        if blocked is not None and s.max_bytes() > backfill:
                            keep.append(s)
                            continue
                        try:
                            job = s.add_to(pool)
                        except DssError as e:
                            if e.code != DSS_ENOMEM:
                                raise
                            with cv:
                                if not ready and state["inflight_jobs"] == 0 and blocked is None \
                                        and s is pending[0]:
                                    raise DssError(e.code, f"job {s.jid} does not fit the empty pool") from e
                            if blocked is None:
                                blocked = s
                            keep.append(s)
                            continue
                        res[s.jid].admitted = now()
                        with cv:
                            ready.append((s, job))
                            cv.notify_all()
                    pending = keep
                    # wait for the next arrival, or for memory to free
                    nxt = order[k].arrival - now() if (realtime and k < len(order)) else None
                    if pending or (nxt is not None and nxt > 0):
                        timeout = 0.05 if pending else nxt
                        if pending and nxt is not None:
                            timeout = min(timeout, max(nxt, 0.0))
                        if pending:
                            with cv:
                                cv.wait(timeout=timeout)
                        else:
                            time.sleep(timeout)
            except BaseException as e:  # noqa: BLE001 -- reported by run()
                state["error"] = e
            finally:
                with cv:
                    state["admitting"] = False
                    cv.notify_all()

        def work(worker):
            batch = self._batches.get(worker)
            if batch is not None and batch.pool is not pool:
                batch.close()
                batch = None
            try:
                while True:
                    with cv:
                        while True:
                            if state["error"] is not None:
                                return
                            take = self._take(ready, worker)
                            if take:
                                break
                            if not state["admitting"] and not ready:
                                return
                            if not state["admitting"] and worker != 0 and self.large_strings is not None \
                                    and self.inflight > 1 and all(s.n >= self.large_strings for s, _ in ready):
                                return  # only large jobs remain: worker 0 runs them
                            cv.wait()
                        state["inflight_jobs"] += len(take)
                    started = now()
                    ids = [job for _, job in take]
                    if batch is None:
                        batch = pool.batch(ids)
                        if self.reserve_strings:
                            batch.set.reserve(self.reserve_strings, 65536)
                    else:
                        batch.rebatch(ids)
                    batch.sort()
                    D = batch.dist_sizes()
                    outs = [batch.result(k) for k in range(len(take))] if keep_results else None
                    if resident is None:
                        for _, job in take:
                            pool.free(job)
                    done = now()
                    with cv:
                        for k, (s, _) in enumerate(take):
                            r = res[s.jid]
                            r.D = int(D[k])
                            r.started, r.done, r.batch_jobs = started, done, len(take)
                            if outs is not None:
                                r.perm, r.lcp = outs[k]
                        state["inflight_jobs"] -= len(take)
                        cv.notify_all()
            except BaseException as e:  # noqa: BLE001 -- reported by run()
                with cv:
                    state["error"] = e
                    cv.notify_all()
            finally:
                if batch is not None:
                    self._batches[worker] = batch

        threads = [threading.Thread(target=admit, name="dss-admit")]
        threads += [threading.Thread(target=work, args=(k,), name=f"dss-worker{k}") for k in range(self.inflight)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if state["error"] is not None:
            raise state["error"]
        return [res[s.jid] for s in specs]


def run_plain(pool: JobPool, specs, resident: dict | None = None, keep_results: bool = False):
    """Runs every job alone (a batch of one), one after the other.

    With `resident` (jid -> pool job id of jobs already admitted) only the sort
    is timed; otherwise each job is admitted first (untimed ingestion).
    Returns per job (D, seconds of sort, seconds of ingestion, results)."""
    out = {}
    batch = None
    largest = max((s.n for s in specs), default=0)
    try:
        for s in specs:
            g0 = time.perf_counter()
            job = resident[s.jid] if resident is not None else s.add_to(pool)
            g1 = time.perf_counter()
            if batch is None:
                batch = pool.batch([job])
                batch.set.reserve(largest, 1)  # no allocation inside the timed sorts
                g1 = time.perf_counter()
            else:
                batch.rebatch([job])
            batch.sort()
            D = int(batch.dist_sizes()[0])
            t1 = time.perf_counter()
            r = batch.result(0) if keep_results else None
            if resident is None:
                pool.free(job)
            out[s.jid] = (D, t1 - g1, g1 - g0, r)
    finally:
        if batch is not None:
            batch.close()
    return out


def fifo_latencies(specs, service: dict) -> np.ndarray:
    """Latencies of a single server that runs jobs one at a time in arrival
    order, given each job's service time (seconds)."""
    t = 0.0
    lat = []
    for s in sorted(specs, key=lambda s: s.arrival):
        t = max(t, s.arrival) + service[s.jid]
        lat.append(t - s.arrival)
    return np.array(lat)


def percentiles(x) -> dict:
    x = np.asarray(x, dtype=np.float64)
    if x.size == 0:
        return {"p50_ms": None, "p99_ms": None}
    return {"p50_ms": round(float(np.percentile(x, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(x, 99)) * 1e3, 3)}
