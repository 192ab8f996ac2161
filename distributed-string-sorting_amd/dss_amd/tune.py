"""Autotuner of the sort's runtime parameters, keyed by input regime.

A regime is (log2 of the strings per PE, log2 of the mean string length,
tenths of the mean LCP over the mean length, number of PEs): inputs in one
regime share their MSD pass structure, base-case sizes and merge shape.
`autotune` times the local sort under each candidate setting (base-case
threshold, one-wave base-case classes, prefix-skip lanes and strings), then
the best of them with host-driven and with device-driven passes, then, with
more than one PE, the LCP merge of the exchanged runs under each merge fan-in
(2, 4, 8, 16 runs per pass) and, with fan-in 2, each merge tile size (512,
1024, 2048 outputs); every setting yields the same output, so
any set of the regime serves as the sample.  The fastest setting is recorded
in a cache that can be saved as JSON; `apply` sets the cached setting for a
new set of a known regime.
"""
from __future__ import annotations

import json
import math
import os
import time

DEFAULT = {"large_min": 512, "wave_class_max": 6, "skip_lanes": 8, "skip_strings": 4, "merge_fan_in": 2,
           "pass_mode": 0, "merge_tile": 1024, "digit_chars": 0}

CANDIDATES = [
    {"large_min": lm, "wave_class_max": wc, "skip_lanes": sk, "skip_strings": ss}
    for lm in (256, 512, 1024) for wc in (5, 6)
    for sk, ss in ((8, 1), (8, 2), (8, 4), (16, 1), (16, 2))
]
MERGE_FANS = (2, 4, 8, 16)
PASS_MODES = (1, 2)  # host-driven passes; device-driven passes replayed from hipGraphs
MERGE_TILES = (512, 1024, 2048)  # outputs per tile of the pairwise merge
DIGIT_CHARS = (1, 2, 3)  # characters per radix digit (more than one on small alphabets only)


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

    @staticmethod
    def _time(ds, run, repeats):
        best = math.inf
        for _ in range(repeats):
            ds.sync()
            t0 = time.perf_counter()
            run()
            ds.sync()
            best = min(best, time.perf_counter() - t0)
        return best

    def autotune(self, ds, candidates=None, repeats: int = 2, fans=MERGE_FANS, modes=PASS_MODES,
                 tiles=MERGE_TILES, digits=DIGIT_CHARS) -> dict:
        """Times ds.sort_lcp() under every candidate, the best candidate under
        every radix digit width (characters per digit, host-driven passes)
        and every pass mode, then (more than one PE) ds.merge_lcp() of the
        exchanged runs under every merge fan-in and, with fan-in 2, every
        merge tile size; keeps the fastest for the set's regime.  Returns
        {"regime", "best", "times_ms", "digit_times_ms", "pass_times_ms",
        "merge_times_ms", "merge_tile_times_ms"}."""
        candidates = [dict(c, merge_fan_in=DEFAULT["merge_fan_in"], pass_mode=DEFAULT["pass_mode"],
                           merge_tile=DEFAULT["merge_tile"], digit_chars=DEFAULT["digit_chars"])
                      for c in (candidates or CANDIDATES)]
        ds.set_sort_params(DEFAULT)
        ds.sort_lcp()  # warm up buffers; also gives the LCPs for the regime key
        ds.sync()
        key = self.key(ds)
        times = []
        for cand in candidates:
            ds.set_sort_params(cand)
            times.append((self._time(ds, ds.sort_lcp, repeats), cand))
        times.sort(key=lambda t: t[0])
        best = dict(times[0][1])
        digit_times = []
        for dc in digits:
            ds.set_sort_params(dict(best, digit_chars=dc, pass_mode=1))
            digit_times.append((self._time(ds, ds.sort_lcp, repeats), dc))
        if digit_times:
            digit_times.sort()
            best["digit_chars"] = digit_times[0][1]
        pass_times = []
        for mode in modes:
            ds.set_sort_params(dict(best, pass_mode=mode))
            ds.sort_lcp()  # first sort in a mode: its capture
            pass_times.append((self._time(ds, ds.sort_lcp, repeats), mode))
        if pass_times:
            pass_times.sort()
            best["pass_mode"] = pass_times[0][1]
        merge_times = []
        tile_times = []
        if ds.p > 1 and fans:
            ds.set_sort_params(best)
            ds.sort_lcp()
            ds.partition("IndexedNumStrings", max_len=2 * 100 * (ds.avg_lcp() + 5))
            ds.exchange("whole")
            for fan in fans:
                ds.set_sort_params(dict(best, merge_fan_in=fan))
                ds.merge_lcp()  # first call of a shape: buffers
                merge_times.append((self._time(ds, ds.merge_lcp, repeats), fan))
            merge_times.sort()
            best["merge_fan_in"] = merge_times[0][1]
            if best["merge_fan_in"] == 2:
                for tile in tiles:
                    ds.set_sort_params(dict(best, merge_tile=tile))
                    ds.merge_lcp()
                    tile_times.append((self._time(ds, ds.merge_lcp, repeats), tile))
                if tile_times:
                    tile_times.sort()
                    best["merge_tile"] = tile_times[0][1]
        self.table[key] = best
        ds.set_sort_params(best)
        return {"regime": key, "best": dict(best),
                "times_ms": [(round(t * 1e3, 3), c) for t, c in times],
                "digit_times_ms": [(round(t * 1e3, 3), d) for t, d in digit_times],
                "pass_times_ms": [(round(t * 1e3, 3), m) for t, m in pass_times],
                "merge_times_ms": [(round(t * 1e3, 3), f) for t, f in merge_times],
                "merge_tile_times_ms": [(round(t * 1e3, 3), m) for t, m in tile_times]}

    def apply(self, ds) -> bool:
        """Sets the cached parameters of the set's regime (after one sort the
        regime includes the LCP ratio); returns False if the regime is new."""
        key = self.key(ds)
        if key in self.table:
            ds.set_sort_params(self.table[key])
            return True
        return False
