"""Device job pool: many independent sort jobs resident in HBM, sorted in
batches (every job of a batch is one virtual PE of a single string set, so
one sequence of kernel launches sorts all of them)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _p(a, t):
    return a.ctypes.data_as(t)


class JobPool:
    """HBM arena of sort jobs with first-fit allocation of characters and
    strings; capacities are fixed at creation and act as the admission limit."""

    def __init__(self, char_capacity: int, string_capacity: int):
        self._h = _lib.pool_p()
        check(lib().dss_pool_create(char_capacity, string_capacity, C.byref(self._h)))

    def close(self):
        if self._h:
            check(lib().dss_pool_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ jobs
    def add(self, chars: np.ndarray) -> int:
        chars = np.ascontiguousarray(chars, dtype=np.uint8)
        j = C.c_uint32()
        check(lib().dss_pool_add_job(self._h, _p(chars, _lib.u8p) if chars.size else None,
                                     chars.size, C.byref(j)))
        return j.value

    def add_dton(self, n: int, length: int, dton: float, seed: int) -> int:
        j = C.c_uint32()
        check(lib().dss_pool_add_job_dton(self._h, n, length, dton, seed, C.byref(j)))
        return j.value

    def add_skewed(self, n: int, seed: int, long_len=1000, short_len=100, long_dn=0.75,
                   short_dn=0.25, long_permille=200, dup_permille=100) -> int:
        j = C.c_uint32()
        check(lib().dss_pool_add_job_skewed(self._h, n, long_len, short_len, long_dn, short_dn,
                                            long_permille, dup_permille, seed, C.byref(j)))
        return j.value

    def free(self, job: int):
        check(lib().dss_pool_free_job(self._h, job))

    def job_info(self, job: int):
        n = C.c_uint64()
        b = C.c_uint64()
        check(lib().dss_pool_job_info(self._h, job, C.byref(n), C.byref(b)))
        return n.value, b.value

    def usage(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        check(lib().dss_pool_usage(self._h, *[C.byref(x) for x in v]))
        return dict(bytes_used=v[0].value, strings_used=v[1].value, bytes_capacity=v[2].value,
                    strings_capacity=v[3].value)

    # ------------------------------------------------------------ batches
    def batch(self, jobs) -> "Batch":
        ids = np.ascontiguousarray(jobs, dtype=np.uint32)
        h = _lib.set_p()
        check(lib().dss_pool_batch(self._h, _p(ids, _lib.u32p), ids.size, C.byref(h)))
        return Batch(self, h, [int(x) for x in ids])


class Batch:
    """A string set over pool jobs; PE k of the set is jobs[k]."""

    def __init__(self, pool, handle, jobs):
        from . import DeviceStringSet
        self.pool = pool
        self.set = DeviceStringSet(handle)
        self.jobs = jobs

    def rebatch(self, jobs):
        """Points this batch at other jobs of the same pool, keeping its buffers."""
        ids = np.ascontiguousarray(jobs, dtype=np.uint32)
        check(lib().dss_pool_rebatch(self.pool._h, self.set._h, _p(ids, _lib.u32p), ids.size))
        self.jobs = [int(x) for x in ids]

    def sort(self):
        self.set.sort_lcp()

    def dist_sizes(self) -> np.ndarray:
        out = np.zeros(max(len(self.jobs), 1), dtype=np.uint64)
        check(lib().dss_pe_dist_sizes(self.set._h, _p(out, _lib.u64p)))
        return out[:len(self.jobs)]

    def result(self, k: int):
        """Sorted order of job k as job-local string indices, and its LCP array."""
        pb = self.set.pe_begin()
        cnt = int(pb[k + 1] - pb[k])
        perm = np.zeros(max(cnt, 1), dtype=np.uint32)
        lcp = np.zeros(max(cnt, 1), dtype=np.uint32)
        check(lib().dss_get_sorted_pe(self.set._h, k, _p(perm, _lib.u32p), _p(lcp, _lib.u32p)))
        return perm[:cnt], lcp[:cnt]

    def close(self):
        self.set.close()
