"""Python bindings of the CPU oracle (oracle/dss_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the
CPU-baseline leg of bench.py as the checker.  The product never imports it.

Each wrapper names the reference function it restates; the C file carries the
file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libdss_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdss_ref.so")

_lib = None
_ref = None

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_rng.restype = C.c_uint64
        L.orc_rng.argtypes = [C.c_uint64] * 3
        L.orc_shuffle_index.restype = C.c_uint64
        L.orc_shuffle_index.argtypes = [C.c_uint64] * 3
        L.orc_dton_params.argtypes = [C.c_uint64, C.c_uint32, C.c_double, u32p, u32p]
        L.orc_gen_dton.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_uint64, u8p]
        L.orc_gen_skewed_size.restype = C.c_uint64
        L.orc_gen_skewed_size.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_gen_skewed.argtypes = [C.c_void_p, C.c_uint64, u8p]
        L.orc_gen_markov_text.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                          C.c_uint32, u8p]
        L.orc_suffix_arena.restype = C.c_uint64
        L.orc_suffix_arena.argtypes = [u8p, C.c_uint64, C.c_uint32, u8p]
        L.orc_count_strings.restype = C.c_uint64
        L.orc_count_strings.argtypes = [u8p, C.c_uint64]
        L.orc_index.argtypes = [u8p, C.c_uint64, u64p, u32p]
        L.orc_pe_split.argtypes = [u64p, C.c_uint64, C.c_uint64, C.c_uint32, u64p]
        L.orc_local_sort.argtypes = [u8p, u64p, u64p, C.c_uint32, u32p, u32p]
        L.orc_compute_d.restype = C.c_uint64
        L.orc_compute_d.argtypes = [u32p, C.c_uint64]
        L.orc_avg_lcp.restype = C.c_uint64
        L.orc_avg_lcp.argtypes = [u32p, C.c_uint64]
        L.orc_partition.restype = C.c_int
        L.orc_partition.argtypes = [u8p, u64p, u32p, u64p, C.c_uint32, u32p, C.c_int,
                                    C.c_uint64, C.c_uint64, u32p, u64p]
        L.orc_partition_ex.restype = C.c_int
        L.orc_partition_ex.argtypes = [u8p, u64p, u32p, u64p, C.c_uint32, u32p, C.c_int,
                                       C.c_uint64, C.c_uint64, u32p, u64p, u64p, u32p, u64p, u64p, u32p]
        L.orc_pd_sort_mt.restype = C.c_uint32
        L.orc_pd_sort_mt.argtypes = [u8p, u64p, u32p, u64p, C.c_uint32, C.c_int, C.c_uint64, C.c_int,
                                     u32p, u32p, u32p, u32p, C.POINTER(C.c_double), u64p]
        L.orc_exchange.restype = C.c_uint64
        L.orc_exchange.argtypes = [C.c_int, u8p, u64p, u32p, u64p, C.c_uint32, u32p, u32p,
                                   u32p, u64p, u8p, u64p, u64p, u32p, u32p, u32p]
        L.orc_merge_losertree.argtypes = [C.POINTER(u8p), u32p, u64p, C.c_uint32, C.c_int,
                                          u64p, u32p, u32p]
        L.orc_extend_prefix.restype = C.c_uint64
        L.orc_extend_prefix.argtypes = [C.POINTER(u8p), u64p, u32p, C.c_uint64, u8p, u64p]
        L.orc_xxh64_stream.restype = C.c_uint64
        L.orc_xxh64_stream.argtypes = [u8p, C.c_uint64]
        L.orc_pd_hash.restype = C.c_uint64
        L.orc_pd_hash.argtypes = [u8p, C.c_uint64, C.c_uint64]
        L.orc_prefix_doubling.restype = C.c_uint32
        L.orc_prefix_doubling.argtypes = [u8p, u64p, u32p, u64p, C.c_uint32, u32p, u32p,
                                          C.c_uint64, C.c_uint32, u32p, u8p]
        L.orc_is_sorted.restype = C.c_int
        L.orc_is_sorted.argtypes = [C.POINTER(u8p), u64p, C.c_uint32]
        L.orc_varint_write_ranges.restype = C.c_uint64
        L.orc_varint_write_ranges.argtypes = [u64p, u64p, C.c_uint32, u8p, u64p]
        L.orc_varint_read_ranges.restype = C.c_int64
        L.orc_varint_read_ranges.argtypes = [u8p, C.c_uint64, u64p, C.c_uint32, u64p]
        L.orc_golomb_b.restype = C.c_uint64
        L.orc_golomb_b.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_golomb_encode.restype = C.c_uint64
        L.orc_golomb_encode.argtypes = [u64p, C.c_uint64, C.c_uint64, u64p]
        L.orc_golomb_decode.restype = C.c_int64
        L.orc_golomb_decode.argtypes = [u64p, C.c_uint64, C.c_uint64, u64p, C.c_uint64]
        _lib = L
    return _lib


def ref():
    """The reference-header build (oracle/_ref), or None when absent."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_xxh_stream.restype = C.c_uint64
        R.ref_xxh_stream.argtypes = [u8p, C.c_uint64]
        R.ref_merge.restype = C.c_int
        R.ref_merge.argtypes = [u8p, C.c_uint64, u64p, u64p, C.c_uint32, C.c_int,
                                u64p, u64p, u64p]
        R.ref_extend_prefix.restype = C.c_uint64
        R.ref_extend_prefix.argtypes = [u8p, C.c_uint64, u64p, u64p, C.c_uint64, u8p]
        R.ref_varint_write_ranges.restype = C.c_uint64
        R.ref_varint_write_ranges.argtypes = [u64p, u64p, C.c_uint32, u8p, u64p]
        R.ref_varint_read_ranges.restype = None
        R.ref_varint_read_ranges.argtypes = [u8p, u64p, C.c_uint32, u64p]
        R.ref_golomb_b.restype = C.c_uint64
        R.ref_golomb_b.argtypes = [C.c_uint64, C.c_uint64]
        R.ref_golomb_encode.restype = C.c_uint64
        R.ref_golomb_encode.argtypes = [u64p, C.c_uint64, C.c_uint64, u64p]
        R.ref_golomb_decode.restype = C.c_uint64
        R.ref_golomb_decode.argtypes = [u64p, C.c_uint64, C.c_uint64, u64p]
        _ref = R
    return _ref


def P(a, t):
    return a.ctypes.data_as(t)


class SkewParams(C.Structure):
    _fields_ = [("n", C.c_uint64), ("long_len", C.c_uint32), ("short_len", C.c_uint32),
                ("long_dn", C.c_double), ("short_dn", C.c_double),
                ("long_permille", C.c_uint32), ("dup_permille", C.c_uint32)]


# ----------------------------------------------------------------- generators
def dton_params(n, L, dton):
    k = C.c_uint32()
    s = C.c_uint32()
    lib().orc_dton_params(n, L, dton, C.byref(k), C.byref(s))
    return k.value, s.value


def gen_dton(n, L, dton, seed):
    k, slen = dton_params(n, L, dton)
    out = np.empty(n * (slen + 1), dtype=np.uint8)
    lib().orc_gen_dton(n, L, dton, seed, P(out, u8p))
    return out


def skew_params(n, long_len=1000, short_len=100, long_dn=0.75, short_dn=0.25,
                long_permille=200, dup_permille=100):
    return SkewParams(n, long_len, short_len, long_dn, short_dn, long_permille, dup_permille)


def gen_skewed(params, seed):
    size = lib().orc_gen_skewed_size(C.byref(params), seed)
    out = np.empty(size, dtype=np.uint8)
    lib().orc_gen_skewed(C.byref(params), seed, P(out, u8p))
    return out


def gen_markov_text(text_len, seed, repeats=5, rep_min=1000, rep_max=10000):
    text = np.empty(text_len, dtype=np.uint8)
    lib().orc_gen_markov_text(text_len, seed, repeats, rep_min, rep_max, P(text, u8p))
    return text


def gen_suffixes(text_len, seed, repeats=5, rep_min=1000, rep_max=10000, cap=500):
    text = np.empty(text_len, dtype=np.uint8)
    lib().orc_gen_markov_text(text_len, seed, repeats, rep_min, rep_max, P(text, u8p))
    size = lib().orc_suffix_arena(P(text, u8p), text_len, cap, None)
    out = np.empty(size, dtype=np.uint8)
    lib().orc_suffix_arena(P(text, u8p), text_len, cap, P(out, u8p))
    return out


# ------------------------------------------------------------------ string set
class StringSet:
    """NUL-terminated arena + offsets + lengths + p-way PE split."""

    def __init__(self, chars: np.ndarray, p: int):
        chars = np.ascontiguousarray(chars, dtype=np.uint8)
        n = int(lib().orc_count_strings(P(chars, u8p), chars.size))
        self.chars = chars
        self.n = n
        self.off = np.empty(max(n, 1), dtype=np.uint64)[:n]
        self.len = np.empty(max(n, 1), dtype=np.uint32)[:n]
        if n:
            lib().orc_index(P(chars, u8p), chars.size, P(self.off, u64p), P(self.len, u32p))
        self.p = p
        self.pe_begin = np.zeros(p + 1, dtype=np.uint64)
        offc = self.off if n else np.zeros(1, dtype=np.uint64)
        lib().orc_pe_split(P(offc, u64p), n, chars.size, p, P(self.pe_begin, u64p))

    @classmethod
    def from_pes(cls, pe_arenas):
        """One NUL-terminated arena per PE (per-rank inputs, no byte slicing)."""
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in pe_arenas]
        chars = np.concatenate(arrs) if arrs else np.zeros(0, np.uint8)
        self = cls(chars, 1)
        counts = [int(lib().orc_count_strings(P(a, u8p), a.size)) if a.size else 0 for a in arrs]
        self.p = len(arrs)
        self.pe_begin = np.zeros(self.p + 1, dtype=np.uint64)
        self.pe_begin[1:] = np.cumsum(counts)
        return self

    def string(self, i) -> bytes:
        o = int(self.off[i])
        return self.chars[o:o + int(self.len[i])].tobytes()

    def pointers(self, ids):
        base = self.chars.ctypes.data
        arr = (u8p * len(ids))()
        for k, i in enumerate(ids):
            arr[k] = C.cast(base + int(self.off[i]), u8p)
        return arr


def local_sort(ss: StringSet):
    """tlx radixsort_CI3 per PE -> (perm u32, lcp u32), both per position."""
    n = ss.n
    perm = np.empty(max(n, 1), dtype=np.uint32)
    lcp = np.empty(max(n, 1), dtype=np.uint32)
    offc = ss.off if n else np.zeros(1, dtype=np.uint64)
    lib().orc_local_sort(P(ss.chars, u8p), P(offc, u64p), P(ss.pe_begin, u64p), ss.p,
                         P(perm, u32p), P(lcp, u32p))
    return perm[:n], lcp[:n]


def compute_d(lcp: np.ndarray) -> int:
    lcp = np.ascontiguousarray(lcp, dtype=np.uint32)
    if lcp.size == 0:
        return 0
    return int(lib().orc_compute_d(P(lcp, u32p), lcp.size))


def avg_lcp(lcp: np.ndarray) -> int:
    lcp = np.ascontiguousarray(lcp, dtype=np.uint32)
    if lcp.size == 0:
        return 0
    return int(lib().orc_avg_lcp(P(lcp, u32p), lcp.size))


SAMPLERS = {"NumStrings": 0, "NumChars": 1, "IndexedNumStrings": 2, "IndexedNumChars": 3}


def partition(ss: StringSet, perm, sampler, max_len=0, dist=None, factor=2):
    p = ss.p
    out = np.zeros(p * p, dtype=np.uint64)
    offc = ss.off if ss.n else np.zeros(1, dtype=np.uint64)
    lenc = ss.len if ss.n else np.zeros(1, dtype=np.uint32)
    permc = perm if ss.n else np.zeros(1, dtype=np.uint32)
    d = None
    if dist is not None:
        d = np.ascontiguousarray(dist, dtype=np.uint32)
    rc = lib().orc_partition(P(ss.chars, u8p), P(offc, u64p), P(lenc, u32p),
                             P(ss.pe_begin, u64p), p, P(np.ascontiguousarray(permc), u32p),
                             SAMPLERS.get(sampler, sampler), factor, max_len,
                             P(d, u32p) if d is not None else None, P(out, u64p))
    assert rc == 0
    return out.reshape(p, p)


def partition_ex(ss: StringSet, perm, sampler, max_len=0, dist=None, factor=2):
    """partition() plus the samples in emission order and the splitters,
    each as (global index = PE offset + sorted position, truncated length);
    a missing splitter has index 2^64 - 1."""
    p = ss.p
    out = np.zeros(p * p, dtype=np.uint64)
    cap = max(p * factor * max(p - 1, 0), 1)
    sg = np.zeros(cap, dtype=np.uint64)
    st = np.zeros(cap, dtype=np.uint32)
    ns = C.c_uint64()
    pg = np.zeros(max(p - 1, 1), dtype=np.uint64)
    pt = np.zeros(max(p - 1, 1), dtype=np.uint32)
    offc = ss.off if ss.n else np.zeros(1, dtype=np.uint64)
    lenc = ss.len if ss.n else np.zeros(1, dtype=np.uint32)
    permc = perm if ss.n else np.zeros(1, dtype=np.uint32)
    d = None if dist is None else np.ascontiguousarray(dist, dtype=np.uint32)
    rc = lib().orc_partition_ex(P(ss.chars, u8p), P(offc, u64p), P(lenc, u32p),
                                P(ss.pe_begin, u64p), p, P(np.ascontiguousarray(permc), u32p),
                                SAMPLERS.get(sampler, sampler), factor, max_len,
                                P(d, u32p) if d is not None else None, P(out, u64p),
                                P(sg, u64p), P(st, u32p), C.byref(ns), P(pg, u64p), P(pt, u32p))
    assert rc == 0
    k = ns.value
    return out.reshape(p, p), (sg[:k], st[:k]), (pg[:p - 1], pt[:p - 1])


MODES = {"whole": 0, "lcp_compact": 1, "pd_prefix": 2}


def exchange(ss: StringSet, perm, lcp, interval, mode, dist=None):
    """Returns dict(bytes, recv_begin[p+1], byte_begin[p+1], lcp, src_pe, src_pos)."""
    p = ss.p
    m = MODES.get(mode, mode)
    interval = np.ascontiguousarray(interval, dtype=np.uint64).reshape(-1)
    n = ss.n
    offc = ss.off if n else np.zeros(1, dtype=np.uint64)
    lenc = ss.len if n else np.zeros(1, dtype=np.uint32)
    permc = np.ascontiguousarray(perm if n else np.zeros(1, np.uint32), dtype=np.uint32)
    lcpc = np.ascontiguousarray(lcp if n else np.zeros(1, np.uint32), dtype=np.uint32)
    d = np.ascontiguousarray(dist, dtype=np.uint32) if dist is not None else None
    args = [m, P(ss.chars, u8p), P(offc, u64p), P(lenc, u32p), P(ss.pe_begin, u64p), p,
            P(permc, u32p), P(lcpc, u32p), P(d, u32p) if d is not None else None,
            P(interval, u64p)]
    total = int(lib().orc_exchange(*args, None, None, None, None, None, None))
    out = np.zeros(max(total, 1), dtype=np.uint8)
    rb = np.zeros(p + 1, dtype=np.uint64)
    bb = np.zeros(p + 1, dtype=np.uint64)
    rl = np.zeros(max(n, 1), dtype=np.uint32)
    sp = np.zeros(max(n, 1), dtype=np.uint32)
    spos = np.zeros(max(n, 1), dtype=np.uint32)
    lib().orc_exchange(*args, P(out, u8p), P(rb, u64p), P(bb, u64p), P(rl, u32p),
                       P(sp, u32p), P(spos, u32p))
    return dict(bytes=out[:total], recv_begin=rb, byte_begin=bb, lcp=rl[:n], src_pe=sp[:n],
                src_pos=spos[:n])


def _arena_pointers(arena: np.ndarray):
    """Char pointers to every NUL-terminated string of an arena."""
    ends = np.flatnonzero(arena == 0)
    starts = np.concatenate([[0], ends[:-1] + 1]) if ends.size else np.zeros(0, np.int64)
    ptrs = np.zeros(max(len(starts), 1), dtype=np.uint64)
    ptrs[:len(starts)] = np.uint64(arena.ctypes.data) + starts.astype(np.uint64)
    arr = (u8p * ptrs.size).from_buffer(ptrs)
    arr._keep = ptrs  # keep the backing buffer alive with the array
    return arr, starts


def merge_losertree(arena: np.ndarray, lcp, run_sizes, compressed: bool):
    """LcpStringLoserTree_ over one PE's receive arena -> (order, lcp, saved)."""
    arena = np.ascontiguousarray(arena, dtype=np.uint8)
    ptrs, starts = _arena_pointers(arena)
    n = len(starts)
    p = len(run_sizes)
    rb = np.zeros(p + 1, dtype=np.uint64)
    rb[1:] = np.cumsum(np.asarray(run_sizes, dtype=np.uint64))
    assert int(rb[-1]) == n
    lcpc = np.ascontiguousarray(lcp, dtype=np.uint32).copy()
    # compute_ranges_and_set_lcp_at_start_of_range (misc.hpp:454-471)
    for r in range(p):
        if run_sizes[r]:
            lcpc[int(rb[r])] = 0
    order = np.zeros(max(n, 1), dtype=np.uint64)
    olcp = np.zeros(max(n, 1), dtype=np.uint32)
    saved = np.zeros(max(n, 1), dtype=np.uint32)
    if n:
        lib().orc_merge_losertree(ptrs, P(lcpc, u32p), P(rb, u64p), p, int(compressed),
                                  P(order, u64p), P(olcp, u32p),
                                  P(saved, u32p) if compressed else None)
    return order[:n], olcp[:n], saved[:n]


def extend_prefix(arena: np.ndarray, order, saved):
    arena = np.ascontiguousarray(arena, dtype=np.uint8)
    ptrs, starts = _arena_pointers(arena)
    order = np.ascontiguousarray(order, dtype=np.uint64)
    saved = np.ascontiguousarray(saved, dtype=np.uint32)
    n = order.size
    if n == 0:
        return np.zeros(0, np.uint8)
    size = int(lib().orc_extend_prefix(ptrs, P(order, u64p), P(saved, u32p), n, None, None))
    out = np.zeros(size, dtype=np.uint8)
    offs = np.zeros(n, dtype=np.uint64)
    lib().orc_extend_prefix(ptrs, P(order, u64p), P(saved, u32p), n, P(out, u8p), P(offs, u64p))
    return out


def xxh64_stream(data: bytes) -> int:
    a = np.frombuffer(data, dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    return int(lib().orc_xxh64_stream(P(a, u8p), len(data)))


def pd_hash(data: bytes, m: int = 2**64 - 1) -> int:
    a = np.frombuffer(data, dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    return int(lib().orc_pd_hash(P(a, u8p), len(data), m))


def prefix_doubling(ss: StringSet, perm, lcp, m=2**64 - 1, start_depth=8):
    """computeDistinguishingPrefixes per PE -> (results u32, rounds u8, nrounds)."""
    n = ss.n
    res = np.zeros(max(n, 1), dtype=np.uint32)
    rounds = np.zeros(max(n, 1), dtype=np.uint8)
    offc = ss.off if n else np.zeros(1, dtype=np.uint64)
    lenc = ss.len if n else np.zeros(1, dtype=np.uint32)
    permc = np.ascontiguousarray(perm if n else np.zeros(1, np.uint32), dtype=np.uint32)
    lcpc = np.ascontiguousarray(lcp if n else np.zeros(1, np.uint32), dtype=np.uint32)
    nr = lib().orc_prefix_doubling(P(ss.chars, u8p), P(offc, u64p), P(lenc, u32p),
                                   P(ss.pe_begin, u64p), ss.p, P(permc, u32p), P(lcpc, u32p),
                                   m, start_depth, P(res, u32p), P(rounds, u8p))
    return res[:n], rounds[:n], int(nr)


# ---------------------------------------------------------------- pipelines
def merge_sort(ss: StringSet, sampler="IndexedNumStrings", encoder="lcp_compact"):
    """DistributedMergeSort::sort (lcp_aware_number_string_splitting.hpp:189-289)
    plus extendPrefix (distributed_sorter.cpp:127-131) for LCP compression.
    Returns per-PE lists: (arena bytes of sorted strings, lcp array, saved lcps)."""
    perm, lcp = local_sort(ss)
    p = ss.p
    if p == 1:
        return dict(perm=perm, lcp=lcp, local=(perm, lcp), out=None)
    if encoder == "no_lcps":
        return _merge_sort_no_lcps(ss, perm, lcp, sampler)
    avg = avg_lcp(lcp) + 5
    interval = partition(ss, perm, sampler, max_len=2 * 100 * avg)
    ex = exchange(ss, perm, lcp, interval, encoder)
    outs = []
    compressed = encoder == "lcp_compact"
    for q in range(p):
        b0, b1 = int(ex["byte_begin"][q]), int(ex["byte_begin"][q + 1])
        r0, r1 = int(ex["recv_begin"][q]), int(ex["recv_begin"][q + 1])
        arena = ex["bytes"][b0:b1]
        runs = interval[:, q].astype(np.uint64)
        order, olcp, saved = merge_losertree(arena, ex["lcp"][r0:r1], runs, compressed)
        if compressed:
            full = extend_prefix(arena, order, saved)
        else:
            ptrs, starts = _arena_pointers(arena)
            full = b"".join(arena[int(starts[i]):].tobytes().split(b"\0", 1)[0] + b"\0"
                            for i in order) if len(order) else b""
            full = np.frombuffer(full, dtype=np.uint8)
        outs.append(dict(arena=np.asarray(full, dtype=np.uint8), lcp=olcp, saved=saved,
                         order=order, src_pe=ex["src_pe"][r0:r1][order.astype(np.int64)],
                         src_pos=ex["src_pos"][r0:r1][order.astype(np.int64)]))
    return dict(perm=perm, lcp=lcp, interval=interval, exchange=ex, out=outs)


def _merge_sort_no_lcps(ss: StringSet, perm, lcp, sampler):
    """The NoLcps variant of DistributedMergeSort::sort (lcp_aware_number_
    string_splitting.hpp:218-224, 276-281): the local sort computes no LCPs,
    so getAvgLcp reads the zero-initialised array (avg = 0 + 5); whole strings
    are exchanged without LCPs (alltoall.hpp:567-640) and noLcpMerge (:139-182)
    merges the runs with a plain loser tree (string_compare).  Per PE: the
    merged strings (order of equal strings is the tree's, not pinned)."""
    import heapq
    interval = partition(ss, perm, sampler, max_len=2 * 100 * 5)
    ex = exchange(ss, perm, lcp, interval, "whole")
    outs = []
    for q in range(ss.p):
        b0, b1 = int(ex["byte_begin"][q]), int(ex["byte_begin"][q + 1])
        a = ex["bytes"][b0:b1].tobytes()
        strs = a.split(b"\0")[:-1] if a else []
        runs, k = [], 0
        for r in range(ss.p):
            c = int(interval[r, q])
            runs.append(strs[k:k + c])
            k += c
        merged = list(heapq.merge(*runs))
        outs.append(dict(arena=np.frombuffer(b"".join(x + b"\0" for x in merged), dtype=np.uint8),
                         strings=merged))
    return dict(perm=perm, lcp=lcp, interval=interval, exchange=ex, out=outs)


def prefix_doubling_sort(ss: StringSet, sampler="IndexedNumStrings", m=2**64 - 1):
    """DistributedPrefixDoublingSort::sort (prefix_doubling.hpp:204-294) ->
    per-PE permutation of (stringIndex, PEIndex) plus intermediate results."""
    perm, lcp = local_sort(ss)
    p = ss.p
    if p == 1:
        return dict(perm=perm, lcp=lcp, out=None)
    res, rounds, nr = prefix_doubling(ss, perm, lcp, m=m)
    interval = partition(ss, perm, sampler, dist=res)
    ex = exchange(ss, perm, lcp, interval, "pd_prefix", dist=res)
    outs = []
    for q in range(p):
        b0, b1 = int(ex["byte_begin"][q]), int(ex["byte_begin"][q + 1])
        r0, r1 = int(ex["recv_begin"][q]), int(ex["recv_begin"][q + 1])
        arena = ex["bytes"][b0:b1]
        runs = interval[:, q].astype(np.uint64)
        order, olcp, saved = merge_losertree(arena, ex["lcp"][r0:r1], runs, True)
        oi = order.astype(np.int64)
        outs.append(dict(string_index=ex["src_pos"][r0:r1][oi], pe_index=ex["src_pe"][r0:r1][oi],
                         lcp=olcp, saved=saved))
    return dict(perm=perm, lcp=lcp, results=res, rounds=rounds, nrounds=nr,
                interval=interval, exchange=ex, out=outs)


def prefix_doubling_sort_mt(ss: StringSet, sampler="IndexedNumStrings", m=2**64 - 1, threads=0):
    """The same sort PE-parallel on `threads` CPU threads (0: OpenMP's
    default), one PE per thread at a time (orc_pd_sort_mt): per merged
    output position stringIndex, PEIndex and merged LCP (0 at receiver
    starts), results per local sorted position, rounds and per-phase wall
    seconds (local sort, prefix doubling, partition, exchange, merge,
    permutation)."""
    n = ss.n
    nn = max(n, 1)
    si = np.zeros(nn, np.uint32)
    pe = np.zeros(nn, np.uint32)
    lcp = np.zeros(nn, np.uint32)
    res = np.zeros(nn, np.uint32)
    ph = (C.c_double * 6)()
    D = np.zeros(1, np.uint64)
    offc = ss.off if n else np.zeros(1, dtype=np.uint64)
    lenc = ss.len if n else np.zeros(1, dtype=np.uint32)
    nr = lib().orc_pd_sort_mt(P(ss.chars, u8p), P(offc, u64p), P(lenc, u32p), P(ss.pe_begin, u64p), ss.p,
                              SAMPLERS.get(sampler, sampler), m, threads, P(si, u32p), P(pe, u32p),
                              P(lcp, u32p), P(res, u32p), ph, P(D, u64p))
    return dict(string_index=si[:n], pe_index=pe[:n], lcp=lcp[:n], results=res[:n], nrounds=int(nr),
                phase_s=list(ph), D=int(D[0]))


# ---------------------------------------------------------------- wire codecs

def varint_write_ranges(values, counts):
    """IntegerCompression::writeRanges (src/mpi/byte_encoder.hpp:31-53):
    (bytes, per-range byte counts)."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    out = np.zeros(max(1, v.size * 10), dtype=np.uint8)
    rb = np.zeros(max(1, c.size), dtype=np.uint64)
    n = lib().orc_varint_write_ranges(P(v, u64p), P(c, u64p), c.size, P(out, u8p), P(rb, u64p))
    return out[:n].copy(), rb[:c.size].copy()


def varint_read_ranges(data, counts):
    """IntegerCompression::readRanges (src/mpi/byte_encoder.hpp:64-78)."""
    b = np.ascontiguousarray(data, dtype=np.uint8)
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    out = np.zeros(max(1, int(c.sum())), dtype=np.uint64)
    rc = lib().orc_varint_read_ranges(P(b if b.size else np.zeros(1, np.uint8), u8p), b.size,
                                      P(c, u64p), c.size, P(out, u64p))
    if rc < 0:
        raise ValueError("malformed varint stream")
    return out[:int(c.sum())].copy()


def golomb_b(max_value, num_values):
    """getB (src/encoding/golomb_encoding.hpp:49-58)."""
    return int(lib().orc_golomb_b(max_value, num_values))


def golomb_encode(values, b):
    """getDeltaEncoding (src/encoding/golomb_encoding.hpp:60-76) of ascending
    values: words (u64)."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    # raw word + per code at most (q + 1 + 64) bits, q summed over the codes
    qsum = (int(v[-1]) - int(v[0])) // max(b, 1) if v.size else 0
    out = np.zeros(2 + (qsum + v.size * 65) // 64 + 1, dtype=np.uint64)
    n = lib().orc_golomb_encode(P(v if v.size else np.zeros(1, np.uint64), u64p), v.size, b,
                                P(out, u64p))
    return out[:n].copy()


def golomb_decode(words, b, cap):
    """getDeltaDecoding (src/encoding/golomb_encoding.hpp:93-109)."""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    out = np.zeros(max(1, cap), dtype=np.uint64)
    n = lib().orc_golomb_decode(P(w if w.size else np.zeros(1, np.uint64), u64p), w.size, b,
                                P(out, u64p), cap)
    if n < 0:
        raise ValueError("malformed Golomb stream")
    return out[:n].copy()


GOLOMB_ESCAPE_B = 20000000000  # b used when getB gives 1 (bloomfilter.hpp:196-197)


def golomb_hash_b(filter_size, p, count):
    """b of AllToAllHashesGolomb for hash intervals (bloomfilter.hpp:194-197)."""
    b = golomb_b(filter_size // p, count)
    return GOLOMB_ESCAPE_B if b == 1 else b


def golomb_index_b(interval_max, count):
    """b of AllToAllHashesGolomb for duplicate indices (bloomfilter.hpp:267-269)."""
    return max(8, golomb_b(interval_max, count))


def golomb_blocks(values, counts, bs):
    """Send buffer of AllToAllHashesGolomb::alltoallv (bloomfilter.hpp:
    164-215, 247-280): for every non-empty interval [size][b][encoding], size
    counting b and the encoding; empty intervals send nothing.  Returns
    (words, per-interval word counts)."""
    out, sizes, i = [], [], 0
    for c, b in zip(counts, bs):
        c = int(c)
        if c == 0:
            sizes.append(0)
            continue
        enc = golomb_encode(values[i:i + c], b)
        out += [1 + enc.size, b] + [int(x) for x in enc]
        sizes.append(2 + enc.size)
        i += c
    return np.array(out, dtype=np.uint64), np.array(sizes, dtype=np.uint64)


def golomb_blocks_decode(words, cap):
    """Receive side of AllToAllHashesGolomb::alltoallv (bloomfilter.hpp:
    226-240, 285-297): blocks [size][b][encoding] back to back.  Returns the
    decoded values and the value count of every block."""
    vals, counts, k = [], [], 0
    w = np.asarray(words, dtype=np.uint64)
    while k < w.size:
        size, b = int(w[k]), int(w[k + 1])
        d = golomb_decode(w[k + 2:k + 1 + size], b, cap)
        vals += [int(x) for x in d]
        counts.append(d.size)
        k += 1 + size
    return np.array(vals, dtype=np.uint64), np.array(counts, dtype=np.uint64)


def golomb_pipeline_blocks(values, counts, b=1048576):
    """AllToAllHashValuesPipeline::getEncoding per partner (bloomfilter.hpp:
    322-331): [size][encoding] for every interval, [0] for an empty one."""
    out, sizes, i = [], [], 0
    for c in counts:
        c = int(c)
        enc = golomb_encode(values[i:i + c], b) if c else np.zeros(0, np.uint64)
        out += [enc.size] + [int(x) for x in enc]
        sizes.append(1 + enc.size)
        i += c
    return np.array(out, dtype=np.uint64), np.array(sizes, dtype=np.uint64)


def exact_dist(ss: StringSet, perm):
    """ExcatDistinguishingPrefix::filter_exact (src/sorter/distributed/
    bloomfilter.hpp:847-965): all strings sorted stably (by string, then PE
    and local sorted index), every string gets 1 + the larger LCP with its
    neighbours in that order (0 when there is a single string).  Indexed by
    position pe_begin[r] + j, j = index in PE r's local sorted order."""
    n = ss.n
    keys = []
    for r in range(ss.p):
        for pos in range(int(ss.pe_begin[r]), int(ss.pe_begin[r + 1])):
            keys.append((ss.string(int(perm[pos])), pos))
    keys.sort(key=lambda t: t[0])  # stable: equal strings keep (PE, index) order
    out = np.zeros(n, dtype=np.uint32)
    for i in range(1, len(keys)):
        a, b = keys[i - 1][0], keys[i][0]
        k = 0
        while k < min(len(a), len(b)) and a[k] == b[k]:
            k += 1
        d = 1 + k
        out[keys[i - 1][1]] = max(out[keys[i - 1][1]], d)
        out[keys[i][1]] = max(out[keys[i][1]], d)
    return out
