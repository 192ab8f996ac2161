"""Wire codecs of the exchange on the GPU (C ABI dss_varint_* / dss_golomb_*).

Names follow the reference:
  IntegerCompression.write_ranges / read_ranges   -- src/mpi/byte_encoder.hpp:31-78
  golomb_b                                        -- getB, src/encoding/golomb_encoding.hpp:49-58
  AllToAllHashesGolomb policies (hash intervals, duplicate indices) and
  AllToAllHashValuesPipeline framing              -- src/sorter/distributed/bloomfilter.hpp:164-331
Every coding step runs in HIP kernels; host arrays are staged to HBM by the
library.  The *_dev methods take device pointers (ints) of caller-owned HBM.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DssError, check, lib

BLOCKS = 0      # DSS_GOLOMB_BLOCKS: AllToAllHashesGolomb framing
PIPELINED = 1   # DSS_GOLOMB_PIPELINED: AllToAllHashValuesPipeline framing
ESCAPE_B = 20000000000  # b used when getB gives 1 (bloomfilter.hpp:196-197)
DSS_ERANGE = -5

u8p, u64p, u32p = _lib.u8p, _lib.u64p, _lib.u32p


def _p(a, t):
    return a.ctypes.data_as(t)


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a if a.size else np.zeros(1, np.uint64), a.size


def golomb_b(max_value: int, num_values: int) -> int:
    b = C.c_uint64()
    check(lib().dss_golomb_b(max_value, num_values, C.byref(b)))
    return b.value


def hash_interval_b(filter_size: int, num_pes: int, count: int) -> int:
    """b of AllToAllHashesGolomb for the hash intervals (bloomfilter.hpp:194-197)."""
    b = golomb_b(filter_size // num_pes, count)
    return ESCAPE_B if b == 1 else b


def index_interval_b(interval_max: int, count: int) -> int:
    """b of AllToAllHashesGolomb for duplicate indices (bloomfilter.hpp:267-269)."""
    return max(8, golomb_b(interval_max, count))


class WireCodec:
    def __init__(self):
        h = _lib.codec_p()
        check(lib().dss_codec_create(C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().dss_codec_destroy(self._h)
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

    # ------------------------------------------------------------ varint
    def write_ranges(self, values, counts):
        """(bytes, per-range byte counts) of IntegerCompression::writeRanges."""
        v, n = _u64(values)
        c, nr = _u64(counts)
        out = np.zeros(max(1, n * 10), np.uint8)
        rb = np.zeros(max(1, nr), np.uint64)
        tot = C.c_uint64()
        check(lib().dss_varint_write_ranges(self._h, _p(v, u64p), _p(c, u64p), nr, _p(out, u8p), out.size,
                                            _p(rb, u64p), C.byref(tot)))
        return out[:tot.value].copy(), rb[:nr].copy()

    def read_ranges(self, data, counts):
        """IntegerCompression::readRanges: sum(counts) values."""
        b = np.ascontiguousarray(data, dtype=np.uint8)
        c, nr = _u64(counts)
        n = int(c[:nr].sum()) if nr else 0
        out = np.zeros(max(1, n), np.uint64)
        check(lib().dss_varint_read_ranges(self._h, _p(b if b.size else np.zeros(1, np.uint8), u8p), b.size,
                                           _p(c, u64p), nr, _p(out, u64p)))
        return out[:n].copy()

    # ------------------------------------------------------------ Golomb
    def golomb_encode(self, values, counts, bs, layout=BLOCKS):
        """(words, per-range block words) in the given framing."""
        v, n = _u64(values)
        c, nr = _u64(counts)
        b, _ = _u64(bs)
        rw = np.zeros(max(1, nr), np.uint64)
        tot = C.c_uint64()
        rc = lib().dss_golomb_encode(self._h, layout, _p(v, u64p), _p(c, u64p), _p(b, u64p), nr, None, 0,
                                     _p(rw, u64p), C.byref(tot))
        if rc not in (0, DSS_ERANGE):
            check(rc)
        out = np.zeros(max(1, tot.value), np.uint64)
        check(lib().dss_golomb_encode(self._h, layout, _p(v, u64p), _p(c, u64p), _p(b, u64p), nr, _p(out, u64p),
                                      out.size, _p(rw, u64p), C.byref(tot)))
        return out[:tot.value].copy(), rw[:nr].copy()

    def golomb_decode(self, words, max_blocks, cap, layout=BLOCKS, b_pipelined=1048576):
        """(values, per-block value counts)."""
        w, nw = _u64(words)
        out = np.zeros(max(1, cap), np.uint64)
        bc = np.zeros(max(1, max_blocks), np.uint64)
        nb = C.c_uint32()
        check(lib().dss_golomb_decode(self._h, layout, _p(w, u64p), nw, max_blocks, b_pipelined, _p(out, u64p), cap,
                                      _p(bc, u64p), C.byref(nb)))
        counts = bc[:nb.value].copy()
        return out[:int(counts.sum())].copy(), counts

    # The AllToAllHashesGolomb send/receive sides of one PE.
    def send_hashes(self, hashes, interval_sizes, filter_size):
        """Encodes sorted hash intervals (one per destination PE) with
        b = getB(filter_size / p, size), 1 -> 2e10."""
        p = len(interval_sizes)
        bs = [hash_interval_b(filter_size, p, int(s)) for s in interval_sizes]
        return self.golomb_encode(hashes, interval_sizes, bs, BLOCKS)

    def send_indices(self, indices, interval_sizes, interval_range):
        """Encodes sorted duplicate indices with b = max(8, getB(range, size))."""
        bs = [index_interval_b(int(r), int(s)) for r, s in zip(interval_range, interval_sizes)]
        return self.golomb_encode(indices, interval_sizes, bs, BLOCKS)

    # ------------------------------------------------------------ device forms
    def write_ranges_dev(self, values_ptr: int, counts, out_ptr: int, out_cap: int):
        c, nr = _u64(counts)
        rb = np.zeros(max(1, nr), np.uint64)
        tot = C.c_uint64()
        check(lib().dss_varint_write_ranges_dev(self._h, values_ptr, _p(c, u64p), nr, out_ptr, out_cap,
                                                _p(rb, u64p), C.byref(tot)))
        return tot.value, rb[:nr].copy()

    def write_ranges_noalloc(self, values_ptr: int, counts, out_ptr: int, out_cap: int):
        """write_ranges_dev with the library's fixed scratch array: allocates
        no device memory (at most 2^21 values and 4096 ranges per call)."""
        c, nr = _u64(counts)
        rb = np.zeros(max(1, nr), np.uint64)
        tot = C.c_uint64()
        check(lib().dss_varint_write_ranges_noalloc(self._h, values_ptr, _p(c, u64p), nr, out_ptr, out_cap,
                                                    _p(rb, u64p), C.byref(tot)))
        return tot.value, rb[:nr].copy()

    def read_ranges_dev(self, bytes_ptr: int, nbytes: int, counts, values_ptr: int):
        c, nr = _u64(counts)
        check(lib().dss_varint_read_ranges_dev(self._h, bytes_ptr, nbytes, _p(c, u64p), nr, values_ptr))

    def golomb_encode_dev(self, values_ptr: int, counts, bs, out_ptr: int, out_cap: int, layout=BLOCKS):
        c, nr = _u64(counts)
        b, _ = _u64(bs)
        rw = np.zeros(max(1, nr), np.uint64)
        tot = C.c_uint64()
        check(lib().dss_golomb_encode_dev(self._h, layout, values_ptr, _p(c, u64p), _p(b, u64p), nr, out_ptr,
                                          out_cap, _p(rw, u64p), C.byref(tot)))
        return tot.value, rw[:nr].copy()

    def golomb_decode_dev(self, words_ptr: int, nwords: int, max_blocks: int, values_ptr: int, cap: int,
                          layout=BLOCKS, b_pipelined=1048576):
        bc = np.zeros(max(1, max_blocks), np.uint64)
        nb = C.c_uint32()
        check(lib().dss_golomb_decode_dev(self._h, layout, words_ptr, nwords, max_blocks, b_pipelined, values_ptr,
                                          cap, _p(bc, u64p), C.byref(nb)))
        return bc[:nb.value].copy()

    # ------------------------------------------------------------ timing
    def sync(self):
        check(lib().dss_codec_sync(self._h))

    def set_timing(self, enable: bool):
        check(lib().dss_codec_set_timing(self._h, int(enable)))

    def kernel_stats(self, reset: bool = False) -> dict:
        cap = 32
        arr = (_lib.KernelStat * cap)()
        cnt = C.c_uint32()
        check(lib().dss_codec_kernel_stats(self._h, arr, cap, C.byref(cnt), int(reset)))
        return {arr[k].name.decode(): dict(launches=arr[k].launches, ms=arr[k].total_ms,
                                           alg_bytes=arr[k].alg_bytes) for k in range(min(cnt.value, cap))}


__all__ = ["WireCodec", "golomb_b", "hash_interval_b", "index_interval_b", "BLOCKS", "PIPELINED", "DssError"]
