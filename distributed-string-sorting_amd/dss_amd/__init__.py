"""dss_amd -- MI355X-native per-PE string-sorting path.

Python host side over the C ABI (include/dss_gpu.h): a device string set with
p virtual PEs and the stages of the reference's distributed sorters
(local LCP sort, partition, exchange, LCP merge, prefix decompression,
prefix-doubling duplicate detection), all executed by gfx950 HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DssError, check, lib

__all__ = ["DeviceStringSet", "DssError", "JobPool", "WireCodec", "device_info"]


def _p(a, t):
    return a.ctypes.data_as(t)


def device_info() -> str:
    buf = C.create_string_buffer(256)
    check(lib().dss_device_info(buf, 256))
    return buf.value.decode()


class DeviceStringSet:
    """A device string set of p virtual PEs (a StringLcpContainer per PE).

    Create with from_arena / dton / skewed; stages run on the set's stream."""

    def __init__(self, handle):
        self._h = handle

    # ------------------------------------------------------------ creation
    @classmethod
    def from_arena(cls, chars: np.ndarray, num_pes: int) -> "DeviceStringSet":
        chars = np.ascontiguousarray(chars, dtype=np.uint8)
        h = _lib.set_p()
        ptr = _p(chars, _lib.u8p) if chars.size else None
        check(lib().dss_set_create(ptr, chars.size, num_pes, C.byref(h)))
        return cls(h)

    @classmethod
    def from_text(cls, text, num_pes: int, max_size: int = 0) -> "DeviceStringSet":
        """FileDistributer on a '\\n'-separated text held in host memory."""
        t = np.frombuffer(text, dtype=np.uint8) if isinstance(text, (bytes, bytearray)) else \
            np.ascontiguousarray(text, dtype=np.uint8)
        h = _lib.set_p()
        check(lib().dss_set_create_text(_p(t, _lib.u8p) if t.size else None, t.size, max_size, num_pes,
                                        C.byref(h)))
        return cls(h)

    @classmethod
    def load_file(cls, path: str, num_pes: int, max_size: int = 0) -> "DeviceStringSet":
        """FileDistributer(path) (src/util/random_string_generator.hpp:50-62)."""
        h = _lib.set_p()
        check(lib().dss_set_load_file(path.encode(), max_size, num_pes, C.byref(h)))
        return cls(h)

    @classmethod
    def from_pes(cls, pe_arenas) -> "DeviceStringSet":
        """One NUL-terminated arena per PE (the per-rank inputs)."""
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in pe_arenas]
        bb = np.zeros(len(arrs) + 1, dtype=np.uint64)
        bb[1:] = np.cumsum([a.size for a in arrs])
        chars = np.concatenate(arrs) if arrs else np.zeros(0, np.uint8)
        if chars.size == 0:
            chars = np.zeros(1, np.uint8)
        h = _lib.set_p()
        check(lib().dss_set_create_pes(_p(chars, _lib.u8p), _p(bb, _lib.u64p), len(arrs), C.byref(h)))
        return cls(h)

    @classmethod
    def dton(cls, n: int, length: int, dton: float, seed: int, num_pes: int) -> "DeviceStringSet":
        h = _lib.set_p()
        check(lib().dss_set_create_dton(n, length, dton, seed, num_pes, C.byref(h)))
        return cls(h)

    @classmethod
    def skewed(cls, n: int, seed: int, num_pes: int, long_len=1000, short_len=100,
               long_dn=0.75, short_dn=0.25, long_permille=200,
               dup_permille=100) -> "DeviceStringSet":
        h = _lib.set_p()
        check(lib().dss_set_create_skewed(n, long_len, short_len, long_dn, short_dn,
                                          long_permille, dup_permille, seed, num_pes,
                                          C.byref(h)))
        return cls(h)

    @classmethod
    def suffixes(cls, text: np.ndarray, cap: int, num_pes: int) -> "DeviceStringSet":
        """Capped suffixes of `text` as views (no materialised copies)."""
        text = np.ascontiguousarray(text, dtype=np.uint8)
        h = _lib.set_p()
        check(lib().dss_set_create_suffixes(_p(text, _lib.u8p) if text.size else None, text.size, cap,
                                            num_pes, C.byref(h)))
        return cls(h)

    @classmethod
    def markov_suffixes(cls, text_len: int, seed: int, num_pes: int, repeats=5000, rep_min=1000,
                        rep_max=10000, cap=500) -> "DeviceStringSet":
        """Capped suffixes of a seeded order-3 Markov ACGT text with planted
        repeats, generated on the device."""
        h = _lib.set_p()
        check(lib().dss_set_create_markov_suffixes(text_len, repeats, rep_min, rep_max, cap, seed, num_pes,
                                                   C.byref(h)))
        return cls(h)

    def close(self):
        if self._h:
            check(lib().dss_set_destroy(self._h))
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

    # ------------------------------------------------------------ info
    def info(self) -> dict:
        i = _lib.SetInfo()
        check(lib().dss_set_info(self._h, C.byref(i)))
        return dict(num_strings=i.num_strings, num_bytes=i.num_bytes, num_pes=i.num_pes,
                    max_len=i.max_len)

    @property
    def n(self) -> int:
        return self.info()["num_strings"]

    @property
    def p(self) -> int:
        return self.info()["num_pes"]

    def pe_begin(self) -> np.ndarray:
        out = np.zeros(self.p + 1, dtype=np.uint64)
        check(lib().dss_set_pe_begin(self._h, _p(out, _lib.u64p)))
        return out

    def download(self):
        info = self.info()
        chars = np.zeros(max(info["num_bytes"], 1), dtype=np.uint8)
        off = np.zeros(max(info["num_strings"], 1), dtype=np.uint64)
        ln = np.zeros(max(info["num_strings"], 1), dtype=np.uint32)
        check(lib().dss_set_download(self._h, _p(chars, _lib.u8p), _p(off, _lib.u64p),
                                     _p(ln, _lib.u32p)))
        return chars[:info["num_bytes"]], off[:info["num_strings"]], ln[:info["num_strings"]]

    def set_stream(self, stream: int | None):
        """Queue the set's work on a caller's HIP stream (an int handle), or
        on its own stream again (None)."""
        check(lib().dss_set_stream(self._h, stream))

    # ------------------------------------------------------------ stages
    def sort_lcp(self):
        """Local LCP sort of every PE (tlx radixsort_CI3 call sites)."""
        check(lib().dss_sort_lcp(self._h))

    def get_sorted(self):
        n = self.n
        perm = np.zeros(max(n, 1), dtype=np.uint32)
        lcp = np.zeros(max(n, 1), dtype=np.uint32)
        check(lib().dss_get_sorted(self._h, _p(perm, _lib.u32p), _p(lcp, _lib.u32p)))
        return perm[:n], lcp[:n]

    def reserve(self, num_strings: int, num_pes: int):
        """Grows the device buffers for sorts of up to num_strings strings."""
        check(lib().dss_set_reserve(self._h, num_strings, num_pes))

    def sort_params(self) -> dict:
        t = _lib.SortParams()
        check(lib().dss_get_sort_params(self._h, C.byref(t)))
        return dict(large_min=t.large_min, wave_class_max=t.wave_class_max, skip_lanes=t.skip_lanes,
                    skip_strings=t.skip_strings, merge_fan_in=t.merge_fan_in, pass_mode=t.pass_mode,
                    merge_tile=t.merge_tile, digit_chars=t.digit_chars)

    def set_sort_params(self, params=None):
        """Local-sort tuning (large_min, wave_class_max, skip_lanes,
        skip_strings, merge_fan_in, pass_mode: 0 auto, 1 host-driven
        passes, 2 device-driven passes replayed from hipGraphs, merge_tile:
        512, 1024 or 2048 outputs per pairwise merge tile, digit_chars: 0
        auto, 1, or up to 2 / 3 characters per radix digit on small
        alphabets); None restores the defaults.  Every setting gives the
        same output."""
        if params is None:
            check(lib().dss_set_sort_params(self._h, None))
            return
        cur = self.sort_params()
        cur.update(params)
        t = _lib.SortParams(cur["large_min"], cur["wave_class_max"], cur["skip_lanes"], cur["skip_strings"],
                            cur.get("merge_fan_in", 0), cur.get("pass_mode", 0), cur.get("merge_tile", 0),
                            cur.get("digit_chars", 0))
        check(lib().dss_set_sort_params(self._h, C.byref(t)))

    def _per_string(self, a, dtype, what):
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.size != self.n:
            raise ValueError(f"{what} must have one entry per string ({self.n}), got {a.size}")
        return a

    def set_local_order(self, perm, lcp):
        perm = self._per_string(perm, np.uint32, "perm")
        lcp = self._per_string(lcp, np.uint32, "lcp")
        check(lib().dss_set_local_order(self._h, _p(perm, _lib.u32p), _p(lcp, _lib.u32p)))

    def set_interval(self, interval):
        iv = np.ascontiguousarray(interval, dtype=np.uint64).reshape(-1)
        if iv.size != self.p * self.p:
            raise ValueError(f"interval must have p * p = {self.p * self.p} entries, got {iv.size}")
        check(lib().dss_set_interval(self._h, _p(iv, _lib.u64p)))

    def set_dist(self, dist):
        d = self._per_string(dist, np.uint32, "dist")
        check(lib().dss_set_dist(self._h, _p(d, _lib.u32p)))

    def avg_lcp(self) -> int:
        out = C.c_uint64()
        check(lib().dss_avg_lcp(self._h, C.byref(out)))
        return out.value

    SAMPLERS = {"NumStrings": 0, "NumChars": 1, "IndexedNumStrings": 2, "IndexedNumChars": 3}
    MODES = {"whole": 0, "lcp_compact": 1, "pd_prefix": 2, "no_lcps": 3}

    def partition(self, sampler="IndexedNumStrings", max_len=0, use_dist=False, factor=2):
        p = self.p
        out = np.zeros(p * p, dtype=np.uint64)
        check(lib().dss_partition(self._h, self.SAMPLERS.get(sampler, sampler), factor, max_len,
                                  int(use_dist), _p(out, _lib.u64p)))
        return out.reshape(p, p)

    def exchange(self, mode):
        p = self.p
        rb = np.zeros(p + 1, dtype=np.uint64)
        bb = np.zeros(p + 1, dtype=np.uint64)
        nb = C.c_uint64()
        check(lib().dss_exchange(self._h, self.MODES.get(mode, mode), _p(rb, _lib.u64p),
                                 _p(bb, _lib.u64p), C.byref(nb)))
        return rb, bb, nb.value

    def get_exchange(self):
        n = self.n
        lcp = np.zeros(max(n, 1), dtype=np.uint32)
        pe = np.zeros(max(n, 1), dtype=np.uint32)
        pos = np.zeros(max(n, 1), dtype=np.uint32)
        check(lib().dss_get_exchange(self._h, None, _p(lcp, _lib.u32p), _p(pe, _lib.u32p),
                                     _p(pos, _lib.u32p)))
        return lcp[:n], pe[:n], pos[:n]

    def get_exchange_bytes(self, nbytes):
        out = np.zeros(max(nbytes, 1), dtype=np.uint8)
        check(lib().dss_get_exchange(self._h, _p(out, _lib.u8p), None, None, None))
        return out[:nbytes]

    def merge_lcp(self):
        check(lib().dss_merge_lcp(self._h))

    def get_merged(self):
        n = self.n
        o = np.zeros(max(n, 1), dtype=np.uint32)
        l = np.zeros(max(n, 1), dtype=np.uint32)
        sv = np.zeros(max(n, 1), dtype=np.uint32)
        check(lib().dss_get_merged(self._h, _p(o, _lib.u32p), _p(l, _lib.u32p), _p(sv, _lib.u32p)))
        return o[:n], l[:n], sv[:n]

    def set_merged(self, order, lcp):
        """A caller-built merged output (the checker's input as the
        reference's checker tests build it); entries are receive indices."""
        order = self._per_string(order, np.uint32, "order")
        lcp = self._per_string(lcp, np.uint32, "lcp")
        check(lib().dss_set_merged(self._h, _p(order, _lib.u32p), _p(lcp, _lib.u32p)))

    def extend_prefix(self) -> int:
        nb = C.c_uint64()
        check(lib().dss_extend_prefix(self._h, C.byref(nb)))
        return nb.value

    def get_extended(self, nbytes):
        n = self.n
        out = np.zeros(max(nbytes, 1), dtype=np.uint8)
        offs = np.zeros(max(n, 1), dtype=np.uint64)
        check(lib().dss_get_extended(self._h, _p(out, _lib.u8p), _p(offs, _lib.u64p)))
        return out[:nbytes], offs[:n]

    def prefix_doubling(self, m=2**64 - 1, start_depth=8) -> int:
        nr = C.c_uint32()
        check(lib().dss_prefix_doubling(self._h, m, start_depth, C.byref(nr)))
        return nr.value

    def get_prefix_doubling(self):
        n = self.n
        res = np.zeros(max(n, 1), dtype=np.uint32)
        rounds = np.zeros(max(n, 1), dtype=np.uint8)
        check(lib().dss_get_prefix_doubling(self._h, _p(res, _lib.u32p), _p(rounds, _lib.u8p)))
        return res[:n], rounds[:n]

    def hash_prefixes(self, ids, depth, m=2**64 - 1):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        out = np.zeros(max(ids.size, 1), dtype=np.uint64)
        check(lib().dss_hash_prefixes(self._h, _p(ids, _lib.u32p), ids.size, depth, m,
                                      _p(out, _lib.u64p)))
        return out[:ids.size]

    # ------------------------------------------------------------ drivers
    ENCODERS = {"whole": 0, "lcp_compact": 1, "no_lcps": 2}

    def merge_sort(self, sampler="IndexedNumStrings", compressed=True):
        """DistributedMergeSort::sort (+ extendPrefix when compressed).
        compressed: True/"lcp_compact" (EmptyLcpByteEncoderMemCpy +
        extendPrefix), False/"whole" (EmptyByteEncoderMemCpy), "no_lcps"
        (the NoLcps variant with noLcpMerge)."""
        mode = self.ENCODERS[compressed] if isinstance(compressed, str) else int(bool(compressed))
        check(lib().dss_merge_sort(self._h, self.SAMPLERS.get(sampler, sampler), mode))

    def prefix_doubling_sort(self, sampler="IndexedNumStrings", m=2**64 - 1):
        """DistributedPrefixDoublingSort::sort."""
        check(lib().dss_prefix_doubling_sort(self._h, self.SAMPLERS.get(sampler, sampler), m))

    def get_permutation(self):
        n = self.n
        si = np.zeros(max(n, 1), dtype=np.uint32)
        pi = np.zeros(max(n, 1), dtype=np.uint32)
        check(lib().dss_get_permutation(self._h, _p(si, _lib.u32p), _p(pi, _lib.u32p)))
        return si[:n], pi[:n]

    def check(self) -> dict:
        """is_complete_and_sorted + exhaustive LCP check of the merged output; D."""
        r = _lib.CheckResult()
        check(lib().dss_check(self._h, C.byref(r)))
        return dict(num_strings=r.num_strings, order_errors=r.order_errors,
                    lcp_errors=r.lcp_errors, D=r.dist_prefix_size, sorted=bool(r.sorted),
                    complete=bool(r.complete))

    def exact_dist(self):
        """ExcatDistinguishingPrefix::filter_exact: exact distinguishing-prefix
        length per position (index space of get_prefix_doubling's results)."""
        out = np.zeros(max(self.n, 1), dtype=np.uint32)
        check(lib().dss_exact_dist(self._h, _p(out, _lib.u32p)))
        return out[:self.n]

    def check_dist(self) -> dict:
        """Bounds of the prefix-doubling results against the exact lengths."""
        r = _lib.DistCheck()
        check(lib().dss_check_dist(self._h, C.byref(r)))
        return dict(num_strings=r.num_strings, lower_violations=r.lower_violations,
                    upper_violations=r.upper_violations, exact_sum=r.exact_sum,
                    results_sum=r.results_sum)

    # ------------------------------------------------------------ timing
    def sync(self):
        check(lib().dss_sync(self._h))

    def set_timing(self, enable: bool):
        check(lib().dss_set_timing(self._h, int(enable)))

    def kernel_stats(self, reset: bool = False) -> dict:
        cap = 64
        arr = (_lib.KernelStat * cap)()
        cnt = C.c_uint32()
        check(lib().dss_kernel_stats(self._h, arr, cap, C.byref(cnt), int(reset)))
        out = {}
        for k in range(min(cnt.value, cap)):
            s = arr[k]
            out[s.name.decode()] = dict(launches=s.launches, ms=s.total_ms, alg_bytes=s.alg_bytes)
        return out


from .pool import Batch, JobPool  # noqa: E402
from .codec import WireCodec  # noqa: E402
