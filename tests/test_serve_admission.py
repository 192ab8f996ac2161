"""Admission and batching logic of the continuous-batching server on the CPU,
over a stand-in pool: footprints come from the job specs, a sort takes time
proportional to its strings, and the pool raises the library's ENOMEM error
when a job's characters do not fit.  No GPU work is involved."""
import threading
import time

import numpy as np
import pytest

from dss_amd._lib import DssError
from dss_amd.serve import DSS_ENOMEM, JobSpec, Server


class _Set:
    def reserve(self, *_):
        pass


class _Batch:
    def __init__(self, pool, jobs):
        self.pool = pool
        self.set = _Set()
        self.jobs = list(jobs)

    def rebatch(self, jobs):
        self.jobs = list(jobs)

    def sort(self):
        time.sleep(sum(self.pool.n[j] for j in self.jobs) * self.pool.sec_per_string)

    def dist_sizes(self):
        return np.array([self.pool.n[j] for j in self.jobs], dtype=np.uint64)

    def close(self):
        pass


class _Pool:
    """Characters only: a job takes n * (length + 1) bytes."""

    def __init__(self, cap, sec_per_string=1e-6):
        self.cap = cap
        self.used = 0
        self.n = {}
        self.size = {}
        self.next = 0
        self.sec_per_string = sec_per_string
        self.lock = threading.Lock()
        self.admitted = []  # job ids in admission order

    def _add(self, n, nbytes):
        with self.lock:
            if self.used + nbytes > self.cap:
                raise DssError(DSS_ENOMEM, "no room")
            j = self.next
            self.next += 1
            self.used += nbytes
            self.n[j] = n
            self.size[j] = nbytes
            self.admitted.append(j)
            return j

    def add_dton(self, n, length, dton, seed):
        return self._add(n, n * (length + 1))

    def free(self, j):
        with self.lock:
            self.used -= self.size.pop(j)

    def usage(self):
        return dict(bytes_used=self.used, strings_used=0, bytes_capacity=self.cap, strings_capacity=0)

    def batch(self, jobs):
        return _Batch(self, jobs)


def _spec(jid, n, arrival):
    return JobSpec(jid=jid, kind="dton", n=n, length=99, dton=0.5, seed=jid, arrival=arrival)


def _stream():
    # job 0 holds most of the pool; job 1 fits only after job 0 is freed;
    # jobs 2.. are small and arrive right behind job 1
    specs = [_spec(0, 60_000, 0.0), _spec(1, 60_000, 0.01)]
    specs += [_spec(2 + k, 1_000, 0.02 + 0.001 * k) for k in range(8)]
    return specs


@pytest.mark.parametrize("backfill", [None, 0])
def test_admission_backfill(backfill):
    pool = _Pool(cap=8_000_000, sec_per_string=2e-6)  # job 0 sorts for 0.12 s
    srv = Server(pool, inflight=2, max_batch_strings=1 << 20, large_strings=50_000, backfill_bytes=backfill)
    res = srv.run(_stream(), realtime=True)
    by = {r.jid: r for r in res}
    assert sorted(by) == list(range(10))
    assert all(int(by[j].D) == s.n for j, s in enumerate(_stream()))
    small = [by[j] for j in range(2, 10)]
    if backfill is None:
        # the small jobs pass job 1, which waits for job 0's memory
        assert all(r.admitted < by[1].admitted for r in small)
        assert max(r.done for r in small) < by[1].admitted
    else:
        # strict arrival order: everything behind job 1 waits for it
        assert all(r.admitted >= by[1].admitted for r in small)
    assert by[1].admitted >= by[0].done - 1e-3  # job 1 fits only after job 0 left
    assert pool.used == 0


def test_job_larger_than_the_pool_is_an_error():
    pool = _Pool(cap=1_000_000)
    srv = Server(pool, inflight=2, max_batch_strings=1 << 20, large_strings=50_000)
    with pytest.raises(DssError, match="does not fit the empty pool"):
        srv.run([_spec(0, 20_000, 0.0)], realtime=True)


def test_resident_jobs_all_ready_at_once():
    pool = _Pool(cap=1 << 40, sec_per_string=0.0)
    specs = [_spec(k, 100 + k, 0.5 * k) for k in range(20)]
    ids = {s.jid: pool.add_dton(s.n, 99, 0.5, 0) for s in specs}
    srv = Server(pool, inflight=2, max_batch_strings=1 << 20, large_strings=None)
    res = srv.run(specs, realtime=False, resident=ids)
    assert sorted(r.jid for r in res) == list(range(20))
    assert all(r.arrival == 0.0 for r in res)
    assert all(int(r.D) == s.n for r, s in zip(res, specs))
    # resident jobs stay in the pool
    assert len(pool.size) == 20
