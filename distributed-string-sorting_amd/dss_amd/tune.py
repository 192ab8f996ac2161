"""Autotuner of the local sort's parameters, keyed by input regime.

A regime is (log2 of the strings per PE, log2 of the mean string length,
tenths of the mean LCP over the mean length, number of PEs): inputs in one
regime share their MSD pass structure and base-case sizes.  `autotune` times
the sort under each candidate setting on the given set (every setting yields
the same output, so any set of the regime serves as the sample) and records
the fastest in a cache that can be saved as JSON; `apply` sets the cached
setting for a new set of a known regime.
"""
from __future__ import annotations

import json
import math
import os
import time

DEFAULT = {"large_min": 512, "wave_class_max": 6, "skip_lanes": 8, "skip_strings": 4}

CANDIDATES = [
    {"large_min": lm, "wave_class_max": wc, "skip_lanes": sk, "skip_strings": ss}
    for lm in (256, 512, 1024) for wc in (5, 6)
    for sk, ss in ((8, 1), (8, 2), (8, 4), (16, 1), (16, 2))
]


def regime(info: dict, avg_lcp: int | None = None) -> str:
    n, nbytes, p = info["num_strings"], info["num_bytes"], info["num_pes"]
    per_pe = max(n // max(p, 1), 1)
    mean_len = max((nbytes - n) / max(n, 1), 1.0)
    lcp_part = "?" if avg_lcp is None else str(min(10, int(10 * avg_lcp / mean_len)))
    return f"n2^{int(math.log2(per_pe))}/len2^{int(math.log2(mean_len))}/lcp{lcp_part}/p{p}"


class Tuner:
    def __init__(self, path: str | None = None):
        self.path = path
        self.table: dict[str, dict] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    def save(self):
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)

    def key(self, ds) -> str:
        info = ds.info()
        avg = None
        try:
            avg = ds.avg_lcp()  # available once the set has been sorted
        except Exception:  # noqa: BLE001 -- an unsorted set has no LCPs yet
            avg = None
        return regime(info, avg)

    def autotune(self, ds, candidates=None, repeats: int = 2) -> dict:
        """Times ds.sort_lcp() under every candidate; keeps the fastest for
        the set's regime.  Returns {"regime", "best", "times_ms"}."""
        candidates = candidates or CANDIDATES
        ds.set_sort_params(DEFAULT)
        ds.sort_lcp()  # warm up buffers; also gives the LCPs for the regime key
        ds.sync()
        key = self.key(ds)
        times = []
        for cand in candidates:
            ds.set_sort_params(cand)
            best = math.inf
            for _ in range(repeats):
                ds.sync()
                t0 = time.perf_counter()
                ds.sort_lcp()
                ds.sync()
                best = min(best, time.perf_counter() - t0)
            times.append((best, cand))
        times.sort(key=lambda t: t[0])
        self.table[key] = dict(times[0][1])
        ds.set_sort_params(times[0][1])
        return {"regime": key, "best": dict(times[0][1]),
                "times_ms": [(round(t * 1e3, 3), c) for t, c in times]}

    def apply(self, ds) -> bool:
        """Sets the cached parameters of the set's regime (after one sort the
        regime includes the LCP ratio); returns False if the regime is new."""
        key = self.key(ds)
        if key in self.table:
            ds.set_sort_params(self.table[key])
            return True
        return False
