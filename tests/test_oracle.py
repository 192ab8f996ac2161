"""CPU tests that pin the oracle (oracle/dss_oracle.c) to the reference.

Each function of the oracle is checked against what fixes it independently
of its own formula: the reference's test vectors and known-answer tests, the
reference headers compiled in place (oracle/_ref, or the golden fixtures it
wrote to tests/golden/ref_vectors.json), published XXH64 values, textbook
sorting, a literal restatement of the file-distribution procedure, and exact
prefix equality for the prefix-doubling rounds."""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import file_distributer

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))


def arena(strings):
    return np.frombuffer(b"".join(s + b"\0" for s in strings), dtype=np.uint8).copy()


def brute_lcp(a: bytes, b: bytes) -> int:
    k = 0
    while k < min(len(a), len(b)) and a[k] == b[k]:
        k += 1
    return k


# ----------------------------------------------------------------- XXH64
def test_xxh64_published_vectors():
    # XXH64 with seed 0 (Y. Collet's reference values)
    assert O.xxh64_stream(b"") == 0xEF46DB3751D8E999
    assert O.xxh64_stream(b"abc") == 0x44BC2CF5AD770999
    assert O.xxh64_stream(b"The quick brown fox jumps over the lazy dog") == 0x0B242D361FDA71BC


def test_xxh64_stream_matches_reference_fixture():
    """xxh::hash_state_t<64> digests dumped from the reference header,
    including its `total_len > 32` branch at exactly 32 bytes."""
    pattern = bytes((i * 131 + 7) % 251 + 1 for i in range(300))
    for L, hx in enumerate(GOLDEN["xxh64_stream"]["digests"]):
        assert O.xxh64_stream(pattern[:L]) == int(hx, 16), L


def test_xxh64_stream_matches_reference_build():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    rng = np.random.default_rng(2)
    for L in list(range(0, 80)) + [127, 128, 129, 255, 256, 264, 520, 1000]:
        d = rng.integers(0, 256, L, dtype=np.uint8)
        buf = d.copy() if L else np.zeros(1, np.uint8)
        assert O.xxh64_stream(d.tobytes()) == R.ref_xxh_stream(buf.ctypes.data_as(O.u8p), L), L


def test_pd_hash_range_reduction():
    # XXHasher::hash reduces % bloomFilterSize = 2^64-1
    h = O.xxh64_stream(b"abc")
    assert O.pd_hash(b"abc") == h % (2**64 - 1)
    assert O.pd_hash(b"abc", 1000) == h % 1000


# ------------------------------------------------------------ string set
def test_string_lengths_reference_case():
    # tests_old/catch/string_set_tests.cpp:8-52
    raw = b"AABCA\0BA\0BAB\0\0CCC\0CCD\0"
    ss = O.StringSet(np.frombuffer(raw, np.uint8).copy(), 1)
    assert list(map(int, ss.len)) == [5, 2, 3, 0, 3, 3]


@pytest.mark.parametrize("p", [1, 2, 3, 4, 7, 16])
def test_pe_split_matches_file_distributer(p):
    rng = np.random.default_rng(p)
    strs = [bytes(rng.integers(65, 91, int(l), dtype=np.uint8)) for l in rng.integers(0, 40, 500)]
    text = b"".join(s + b"\n" for s in strs)
    per_pe = file_distributer(text, p)
    ss = O.StringSet(arena(strs), p)
    for r in range(p):
        got = b"".join(ss.string(i) + b"\0" for i in range(int(ss.pe_begin[r]), int(ss.pe_begin[r + 1])))
        assert got == per_pe[r], r


# ------------------------------------------------------------ local sort
INPUTS = {
    "dton": lambda: O.gen_dton(3000, 60, 0.5, 7),
    "skewed": lambda: O.gen_skewed(O.skew_params(3000, long_len=100, short_len=10), 9),
    "suffix": lambda: O.gen_suffixes(4000, 3, repeats=3, rep_min=50, rep_max=200, cap=100),
    "small_alphabet": lambda: arena([bytes(np.random.default_rng(1).integers(97, 99, int(l), dtype=np.uint8))
                                     for l in np.random.default_rng(2).integers(0, 20, 70000)]),
}


@pytest.mark.parametrize("name", sorted(INPUTS))
@pytest.mark.parametrize("p", [1, 3, 4])
def test_local_sort_is_textbook_order_and_lcp(name, p):
    ss = O.StringSet(INPUTS[name](), p)
    perm, lcp = O.local_sort(ss)
    for r in range(p):
        b, e = int(ss.pe_begin[r]), int(ss.pe_begin[r + 1])
        got = [ss.string(perm[i]) for i in range(b, e)]
        assert got == sorted(ss.string(i) for i in range(b, e))
        assert sorted(int(x) for x in perm[b:e]) == list(range(b, e))
        if e > b:
            assert lcp[b] == 0
        for i in range(b + 1, e):
            assert lcp[i] == brute_lcp(got[i - b - 1], got[i - b])


# ----------------------------------------------------------------- D / D/N
def test_compute_d_hand_example():
    # sorted: a, ab, abc, b -> lcps 0,1,2,0; D = (1+1)+(2+1)+(2+1)+(0+1)
    assert O.compute_d(np.array([0, 1, 2, 0], np.uint32)) == 2 + 3 + 3 + 1
    assert O.compute_d(np.array([0], np.uint32)) == 0


@pytest.mark.parametrize("size", [1000, 10000])
@pytest.mark.parametrize("length", [100, 250, 500])
@pytest.mark.parametrize("ratio,eps", [(0.0, 0.075), (0.2, 0.01), (0.25, 0.01), (0.45, 0.01),
                                       (0.76, 0.01), (0.95, 0.01), (1.0, 0.01)])
def test_dton_generator_ratio(size, length, ratio, eps):
    """DNGenerator_tests.cpp:47-101: D / (size * stringLength) within epsilon
    of the requested ratio, all strings of the requested length."""
    chars = O.gen_dton(size, length, ratio, 0x5EED0001 + size)
    ss = O.StringSet(chars, 1)
    assert ss.n == size
    assert (ss.len == length).all()
    perm, lcp = O.local_sort(ss)
    D = O.compute_d(lcp)
    assert abs(D / (size * length) - ratio) < eps


def test_generators_are_deterministic_and_seeded():
    a = O.gen_dton(5000, 50, 0.5, 1)
    assert (a == O.gen_dton(5000, 50, 0.5, 1)).all()
    assert not (a == O.gen_dton(5000, 50, 0.5, 2)).all()
    sp = O.skew_params(2000)
    assert (O.gen_skewed(sp, 3) == O.gen_skewed(sp, 3)).all()
    # the skewed input has exact duplicates and both length classes
    ss = O.StringSet(O.gen_skewed(O.skew_params(4000, long_len=300, short_len=30), 5), 1)
    strs = [ss.string(i) for i in range(ss.n)]
    assert len(set(strs)) < len(strs)
    assert set(int(x) for x in ss.len) == {300, 30}


# ----------------------------------------------------------- exchange KAT
def kat_setup(p, size_per_pe=5, common=10, suffix=20):
    """alltoallStrings_tests.cpp:17-127: PE r holds, for each destination,
    size_per_pe strings common*(r+1) + [i] + suffix*(r+1); LCPs are `common`
    except 0 at the start; every PE sends size_per_pe strings to every PE."""
    pes = []
    for r in range(p):
        strs = []
        for _ in range(p):
            for i in range(1, size_per_pe + 1):
                strs.append(bytes([r + 1]) * common + bytes([i]) + bytes([r + 1]) * suffix)
        pes.append(arena(strs))
    ss = O.StringSet.from_pes(pes)
    n = ss.n
    perm = np.arange(n, dtype=np.uint32)
    lcp = np.full(n, common, dtype=np.uint32)
    for r in range(p):
        lcp[int(ss.pe_begin[r])] = 0
    iv = np.full((p, p), size_per_pe, dtype=np.uint64)
    return ss, perm, lcp, iv


@pytest.mark.parametrize("p", [1, 2, 4, 8])
def test_exchange_reference_vectors(p):
    size, common, suffix = 5, 10, 20
    ss, perm, lcp, iv = kat_setup(p, size, common, suffix)
    for mode in ("whole", "lcp_compact"):
        ex = O.exchange(ss, perm, lcp, iv, mode)
        for q in range(p):
            b0, b1 = int(ex["byte_begin"][q]), int(ex["byte_begin"][q + 1])
            r0, r1 = int(ex["recv_begin"][q]), int(ex["recv_begin"][q + 1])
            got = ex["bytes"][b0:b1].tobytes()
            exp = b""
            exp_lcp = []
            for src in range(p):
                c = bytes([src + 1])
                if mode == "lcp_compact":  # getExpectedRawStringsForPrefixCompression
                    exp += c * common
                    for i in range(size):
                        exp += bytes([i + 1]) + c * suffix + b"\0"
                else:  # getExpectedRawStrings
                    for i in range(size):
                        exp += c * common + bytes([i + 1]) + c * suffix + b"\0"
                exp_lcp += [0] + [common] * (size - 1)  # getExpectedLcpValues
            assert got == exp, (mode, q)
            assert list(map(int, ex["lcp"][r0:r1])) == exp_lcp
    for dup in (False, True):
        # AllToAllStringImplPrefixDoubling_test: dist = common (+1 without duplicates)
        dist = np.full(ss.n, common if dup else common + 1, dtype=np.uint32)
        ex = O.exchange(ss, perm, lcp, iv, "pd_prefix", dist=dist)
        for q in range(p):
            b0, b1 = int(ex["byte_begin"][q]), int(ex["byte_begin"][q + 1])
            exp = b""
            for src in range(p):
                exp += bytes([src + 1]) * common
                exp += b"\0" * size if dup else b"".join(bytes([i + 1]) + b"\0" for i in range(size))
            assert ex["bytes"][b0:b1].tobytes() == exp


# ------------------------------------------------- merge / extendPrefix
def test_merge_and_extend_prefix_match_reference_fixture():
    for case in GOLDEN["merges"]:
        a = np.frombuffer(bytes.fromhex(case["arena_hex"]), dtype=np.uint8).copy()
        comp = case["mode"] == "lcp_compact"
        order, lcp, saved = O.merge_losertree(a, np.array(case["recv_lcp"], np.uint32),
                                              np.array(case["runs"], np.uint64), comp)
        assert list(map(int, order)) == case["order"]
        assert list(map(int, lcp)) == case["merged_lcp"]
        if comp:
            assert list(map(int, saved)) == case["saved_lcp"]
            ext = O.extend_prefix(a, order, saved)
            assert ext.tobytes().hex() == case["extended_hex"]


def test_merge_matches_reference_build_on_fresh_inputs():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    chars = O.gen_skewed(O.skew_params(3000, long_len=80, short_len=6), 77)
    for p in (2, 5, 8):
        ss = O.StringSet(chars, p)
        perm, lcp = O.local_sort(ss)
        iv = O.partition(ss, perm, "NumStrings", max_len=1000)
        ex = O.exchange(ss, perm, lcp, iv, "lcp_compact")
        for q in range(p):
            r0, r1 = int(ex["recv_begin"][q]), int(ex["recv_begin"][q + 1])
            b0, b1 = int(ex["byte_begin"][q]), int(ex["byte_begin"][q + 1])
            n = r1 - r0
            if not n:
                continue
            a = np.ascontiguousarray(ex["bytes"][b0:b1])
            rl = ex["lcp"][r0:r1].astype(np.uint64)
            runs = iv[:, q].astype(np.uint64).copy()
            oi, ol, osv = (np.zeros(n, np.uint64) for _ in range(3))
            assert R.ref_merge(a.ctypes.data_as(O.u8p), a.size, rl.ctypes.data_as(O.u64p),
                               runs.ctypes.data_as(O.u64p), p, 1, oi.ctypes.data_as(O.u64p),
                               ol.ctypes.data_as(O.u64p), osv.ctypes.data_as(O.u64p)) == 0
            order, mlcp, saved = O.merge_losertree(a, ex["lcp"][r0:r1], iv[:, q].astype(np.uint64), True)
            assert (order == oi).all() and (mlcp == ol).all() and (saved == osv).all()


# ------------------------------------------------------- end-to-end order
@pytest.mark.parametrize("name", sorted(INPUTS))
@pytest.mark.parametrize("sampler", ["NumStrings", "NumChars", "IndexedNumStrings", "IndexedNumChars"])
def test_merge_sort_output_is_the_sorted_input(name, sampler):
    ss = O.StringSet(INPUTS[name](), 4)
    res = O.merge_sort(ss, sampler=sampler, encoder="lcp_compact")
    got = b"".join(o["arena"].tobytes() for o in res["out"]).split(b"\0")[:-1]
    assert got == sorted(ss.string(i) for i in range(ss.n))
    # merged LCPs are the LCPs of neighbours inside each PE
    for o in res["out"]:
        strs = o["arena"].tobytes().split(b"\0")[:-1]
        for i in range(1, len(strs)):
            assert int(o["lcp"][i]) == brute_lcp(strs[i - 1], strs[i])


# ------------------------------------------------------- prefix doubling
def exact_rounds(strs_by_pe, start=8):
    """Candidate sets with exact prefix equality in place of the hash (the
    semantics of the reference's debug tracker, tracker.hpp:104-110,187-240):
    a candidate stays while another candidate of any PE has the same prefix
    of the round's depth; results = depth, or the length at end of string."""
    cands = {(r, i) for r, ss in enumerate(strs_by_pe) for i in range(len(ss))}
    results = {k: 0 for k in cands}
    rounds = {k: 0 for k in cands}
    depth, step = start, 2 * start
    while cands:
        groups = {}
        nxt = set()
        for k in cands:
            s = strs_by_pe[k[0]][k[1]]
            rounds[k] += 1
            if depth > len(s):
                results[k] = len(s)
                continue
            results[k] = depth
            groups.setdefault(s[:depth], []).append(k)
        for g in groups.values():
            if len(g) > 1:
                nxt.update(g)
        cands = nxt
        depth, step = start + step, step * 2
    return results, rounds


@pytest.mark.parametrize("name", ["dton", "skewed", "suffix", "small_alphabet"])
@pytest.mark.parametrize("p", [1, 3, 5])
def test_prefix_doubling_matches_exact_prefix_semantics(name, p):
    chars = INPUTS[name]()
    if name == "small_alphabet":
        chars = chars[:20000].copy()
        chars[-1] = 0
    ss = O.StringSet(chars, p)
    perm, lcp = O.local_sort(ss)
    res, rounds, nr = O.prefix_doubling(ss, perm, lcp)
    strs = [[ss.string(perm[i]) for i in range(int(ss.pe_begin[r]), int(ss.pe_begin[r + 1]))] for r in range(p)]
    er, erounds = exact_rounds(strs)
    for r in range(p):
        b = int(ss.pe_begin[r])
        for i in range(len(strs[r])):
            assert res[b + i] == er[(r, i)], (r, i)
            assert rounds[b + i] == erounds[(r, i)], (r, i)
    # DIST <= results for non-EOS strings (the exact distinguishing prefix)
    for r in range(p):
        b = int(ss.pe_begin[r])


def test_prefix_doubling_depth_schedule():
    # every string shares 200 chars: rounds at depths 8, 24, 40, 72, 136, 264
    strs = [b"x" * 200 + bytes([65 + i % 26, 65 + (i // 26) % 26]) for i in range(100)]
    ss = O.StringSet(arena(strs), 2)
    perm, lcp = O.local_sort(ss)
    res, rounds, nr = O.prefix_doubling(ss, perm, lcp)
    assert set(int(x) for x in res) == {202}  # EOS at depth 264 > 202
    assert nr == 6


# ----------------------------------------------------------------- checker
def test_is_sorted_checker_cases():
    strs = [b"a", b"b", b"b", b"c", b"d"]
    a = arena(strs)
    ptrs, _ = O._arena_pointers(a)
    assert O.lib().orc_is_sorted(ptrs, np.array([0, 2, 2, 5], np.uint64).ctypes.data_as(O.u64p), 3)
    a2 = arena([b"a", b"c", b"b", b"d"])
    ptrs2, _ = O._arena_pointers(a2)
    assert not O.lib().orc_is_sorted(ptrs2, np.array([0, 2, 4], np.uint64).ctypes.data_as(O.u64p), 2)


# ---------------------------------------------------------------- partition
# A second, independent restatement of the partition rules in plain Python
# (sorted() for the global sample order, the reference's halving binary
# search), checked against the C oracle on small inputs with and without
# duplicates, for truncated splitters and for prefix-doubling lengths.

def _ref_binary_search(strs, key, idx_key=None, local_offset=0):
    """binarySearch / binarySearchIndexed (misc.hpp:310-374): halving search
    that returns the first probe that compares equal, else the insertion point."""
    left, right = 0, len(strs)
    while left != right:
        dist = (right - left) // 2
        cur = strs[left + dist]
        if idx_key is None:
            res = (cur > key) - (cur < key)
        else:
            a, b = (cur, left + dist + local_offset), (key, idx_key)
            res = (a > b) - (a < b)
        if res < 0:
            left = left + dist + 1
        elif res == 0:
            return left + dist
        else:
            right = left + dist
    return left


def _numchars_indices(weights, nr):
    """The NumChars walk (samplingStrategies.hpp:199-236, :243-273, :307-337,
    :345-386): the count restarts for every splitter, the splitter is the
    last string consumed, the walk stops at the end of the PE."""
    n = len(weights)
    sd = sum(weights) // (nr + 1)
    out, si = [], 0
    for _ in range(nr):
        if si >= n:
            break
        seen = 0
        while seen < sd and si < n:
            seen += weights[si]
            si += 1
        out.append(si - 1 if si else 0)  # the reference reads index -1 if nothing was consumed
    return out


def _brute_partition(pes, indexed, max_len=None, dist=None, factor=2, chars=False):
    """pes: per PE the sorted list of bytes strings."""
    p = len(pes)
    offsets = [0]
    for x in pes:
        offsets.append(offsets[-1] + len(x))
    samples = []
    for r, ss in enumerate(pes):
        n = len(ss)
        nr = min(factor * (p - 1), n)
        if chars:
            # only the indexed policy walks the distinguishing prefixes; the
            # plain one walks the lengths and truncates to dist
            w = [int(dist[r][j]) for j in range(n)] if (dist is not None and indexed) else [len(x) for x in ss]
            idxs = _numchars_indices(w, nr)
        else:
            sd = n / (nr + 1)
            idxs = [int(i * sd) for i in range(1, nr + 1)]
        for j in idxs:
            lim = dist[r][j] if dist is not None else max_len
            samples.append((ss[j][:int(lim)], offsets[r] + j))
    samples.sort(key=(lambda t: (t[0], t[1])) if indexed else (lambda t: t[0]))
    nspl = min(p - 1, len(samples))
    step = len(samples) // (nspl + 1) if nspl + 1 else 0
    spl = [samples[i * step] for i in range(1, nspl + 1)]
    out = np.zeros((p, p), dtype=np.uint64)
    for r, ss in enumerate(pes):
        pos = [_ref_binary_search(ss, s, i if indexed else None, offsets[r]) for s, i in spl]
        pos = [0] + pos + [len(ss)]
        for q in range(len(pos) - 1):  # fewer splitters than p - 1: later buckets stay empty
            out[r, q] = pos[q + 1] - pos[q]
    return out


def _sorted_pes(ss):
    perm, lcp = O.local_sort(ss)
    pes = []
    for r in range(ss.p):
        b, e = int(ss.pe_begin[r]), int(ss.pe_begin[r + 1])
        pes.append([ss.string(int(i)) for i in perm[b:e]])
    return perm, lcp, pes


PART_INPUTS = {
    "distinct": lambda: O.gen_dton(900, 12, 0.5, 21),
    "duplicates": lambda: arena([bytes([97 + (i * 7) % 3]) * (1 + (i * 5) % 4) for i in range(700)]),
    "skewed": lambda: O.gen_skewed(O.skew_params(800, long_len=30, short_len=5, dup_permille=300), 22),
}


@pytest.mark.parametrize("name", sorted(PART_INPUTS))
@pytest.mark.parametrize("p", [2, 3, 5])
@pytest.mark.parametrize("sampler", ["NumStrings", "IndexedNumStrings", "NumChars", "IndexedNumChars"])
def test_partition_matches_brute_force(name, p, sampler):
    ss = O.StringSet(PART_INPUTS[name](), p)
    perm, lcp, pes = _sorted_pes(ss)
    indexed = sampler.startswith("Indexed")
    chars = sampler.endswith("Chars")
    for max_len in (1000, 3):
        got = O.partition(ss, perm, sampler, max_len=max_len)
        exp = _brute_partition(pes, indexed, max_len=max_len, chars=chars)
        assert (got == exp).all(), (name, p, sampler, max_len, got, exp)
    # prefix-doubling form: splitters truncated to the per-position lengths
    dist = np.array([1 + (i * 13) % 9 for i in range(ss.n)], dtype=np.uint32)
    dl = [list(dist[int(ss.pe_begin[r]):int(ss.pe_begin[r + 1])]) for r in range(p)]
    got = O.partition(ss, perm, sampler, dist=dist)
    exp = _brute_partition(pes, indexed, dist=dl, chars=chars)
    assert (got == exp).all(), (name, p, sampler, "dist", got, exp)


def test_numchars_walk_counts_lengths_unless_indexed_with_dist():
    """The plain NumChars policy walks string lengths even in the
    prefix-doubling form (dist only truncates the splitter), the indexed one
    walks the distinguishing prefixes: with lengths and dist chosen so that
    the two walks pick different strings, the partitions differ."""
    strs = sorted([bytes([65 + i % 26]) * (1 + (i * 7) % 40) + bytes([66 + i % 5]) for i in range(400)])
    ss = O.StringSet(arena(strs), 2)
    perm, lcp, pes = _sorted_pes(ss)
    dist = np.array([1 + (i * 29) % 3 for i in range(ss.n)], dtype=np.uint32)
    dl = [list(dist[int(ss.pe_begin[r]):int(ss.pe_begin[r + 1])]) for r in range(2)]
    by_len = _numchars_indices([len(x) for x in pes[0]], 2)
    by_dist = _numchars_indices([int(v) for v in dl[0]], 2)
    assert by_len != by_dist
    for sampler, indexed in (("NumChars", False), ("IndexedNumChars", True)):
        got = O.partition(ss, perm, sampler, dist=dist)
        exp = _brute_partition(pes, indexed, dist=dl, chars=True)
        assert (got == exp).all(), (sampler, got, exp)


@pytest.mark.parametrize("sampler", ["NumChars", "IndexedNumChars"])
def test_partition_numchars_buckets_respect_the_order(sampler):
    """NumChars samplers: every PE's buckets split its sorted strings at the
    same splitter boundaries (strings in bucket q never exceed strings in
    bucket q+1 of any PE), and every string is sent exactly once."""
    ss = O.StringSet(O.gen_skewed(O.skew_params(1200, long_len=40, short_len=4), 23), 4)
    perm, lcp, pes = _sorted_pes(ss)
    iv = O.partition(ss, perm, sampler, max_len=1000)
    assert (iv.sum(axis=1) == np.diff(ss.pe_begin)).all()
    lo = [None] * 4
    hi = [None] * 4
    for r in range(4):
        b = np.concatenate([[0], np.cumsum(iv[r])]).astype(int)
        for q in range(4):
            part = pes[r][b[q]:b[q + 1]]
            if part:
                lo[q] = min(part) if lo[q] is None else min(lo[q], min(part))
                hi[q] = max(part) if hi[q] is None else max(hi[q], max(part))
    for q in range(3):
        if hi[q] is not None and lo[q + 1] is not None:
            assert hi[q] <= lo[q + 1]


# ------------------------------------------- partition, pinned by counting
# Checks that the mathematics fixes, with no halving search and no restated
# NumChars walk: (i) with the indexed samplers every (string, global index)
# key is distinct and each PE's sorted strings are sorted by that key, so the
# boundary before bucket q + 1 is the number of local keys smaller than
# splitter q's key, whatever search finds it; (ii) without an index it is the
# number of smaller strings when no local string equals the splitter, and lies
# inside the run of equal strings otherwise (where inside is implementation
# defined: binarySearch returns any equal probe, misc.hpp:310-333); (iii) the
# splitters are the samples at global ranks i * floor(S / (nspl + 1)) of the
# textbook-sorted sample list (getSplitters(Indexed), misc.hpp:89-189), the
# samples sit at positions floor(i * n / (nr + 1)) (NumStrings,
# samplingStrategies.hpp:20-90, 107-187) or where the running character count
# since the previous sample first reaches num_chars / (nr + 1) (NumChars,
# :189-389), truncated to the splitter length; (iv) the buckets of all PEs
# respect the global order.

MISSING = 2**64 - 1
ALL_SAMPLERS = ["NumStrings", "NumChars", "IndexedNumStrings", "IndexedNumChars"]


def _few_chars_set(p):
    """PE 1 holds 6 strings with 2 characters in total, PE 2 three strings
    with one: their NumChars walks have splitter_dist 0 (at p = 4, factor 2)."""
    rng = np.random.default_rng(41)
    pes = []
    for r in range(p):
        if r == 1:
            strs = [b"", b"b", b"", b"", b"a", b""]
        elif r == 2:
            strs = [b"", b"c", b""]
        else:
            strs = [bytes(rng.integers(97, 100, int(l), dtype=np.uint8)) for l in rng.integers(0, 6, 150)]
        pes.append(arena(strs))
    return O.StringSet.from_pes(pes)


PIN_SETS = {
    "distinct": lambda p: O.StringSet(PART_INPUTS["distinct"](), p),
    "duplicates": lambda p: O.StringSet(PART_INPUTS["duplicates"](), p),
    "skewed": lambda p: O.StringSet(PART_INPUTS["skewed"](), p),
    "few_chars": _few_chars_set,
}


# (the few-characters set is built for 4 PEs)
PIN_CASES = [(n, p) for n in sorted(PIN_SETS) for p in (2, 3, 4, 7) if n != "few_chars" or p == 4]


def _pin_case(name, p, sampler, trunc):
    ss = PIN_SETS[name](p)
    perm, lcp = O.local_sort(ss)
    strs = [ss.string(int(perm[g])) for g in range(ss.n)]  # by global sorted position
    dist = None
    kw = {}
    if trunc == "dist":
        dist = np.array([1 + (g * 13) % 9 for g in range(ss.n)], dtype=np.uint32)
        kw["dist"] = dist
    else:
        kw["max_len"] = 1000 if trunc == "long" else 3
    iv, samples, spl = O.partition_ex(ss, perm, sampler, **kw)
    return ss, strs, dist, kw, iv, samples, spl


def _bound_errors(ss, strs, iv, spl, indexed):
    """Boundaries that contradict the splitters (counting checks i and ii)."""
    errs = []
    sg, st = spl
    p = ss.p
    for r in range(p):
        b, e = int(ss.pe_begin[r]), int(ss.pe_begin[r + 1])
        if int(iv[r].sum()) != e - b:
            errs.append((r, "row sum"))
        bounds = np.cumsum(iv[r].astype(np.int64))[:-1]
        local = strs[b:e]
        for q in range(p - 1):
            if int(sg[q]) == MISSING:  # fewer splitters than PEs: later buckets empty
                if int(bounds[q]) != e - b:
                    errs.append((r, q, "missing splitter"))
                continue
            key = strs[int(sg[q])][:int(st[q])]
            if indexed:
                kk = (key, int(sg[q]))
                exp = sum(1 for j, s in enumerate(local) if (s, b + j) < kk)
                if int(bounds[q]) != exp:
                    errs.append((r, q, int(bounds[q]), exp))
            else:
                lt = sum(1 for s in local if s < key)
                le = sum(1 for s in local if s <= key)
                if not lt <= int(bounds[q]) <= le:
                    errs.append((r, q, int(bounds[q]), lt, le))
    return errs


@pytest.mark.parametrize("trunc", ["long", "short", "dist"])
@pytest.mark.parametrize("sampler", ALL_SAMPLERS)
@pytest.mark.parametrize("name,p", PIN_CASES)
def test_partition_bounds_count_smaller_keys(name, p, sampler, trunc):
    ss, strs, dist, kw, iv, samples, spl = _pin_case(name, p, sampler, trunc)
    assert _bound_errors(ss, strs, iv, spl, sampler.startswith("Indexed")) == []


def test_partition_bound_check_is_not_vacuous():
    """The counting check rejects a boundary moved by one string, and
    boundaries computed with index ties broken the wrong way (descending)."""
    ss, strs, dist, kw, iv, samples, spl = _pin_case("duplicates", 3, "IndexedNumStrings", "long")
    assert _bound_errors(ss, strs, iv, spl, True) == []
    bad = iv.copy()
    bad[1, 0] += 1
    bad[1, 1] -= 1
    assert _bound_errors(ss, strs, bad, spl, True) != []
    sg, st = spl
    wrong = np.zeros_like(iv)
    for r in range(3):
        b, e = int(ss.pe_begin[r]), int(ss.pe_begin[r + 1])
        cut = [0]
        for q in range(2):
            key = (strs[int(sg[q])][:int(st[q])], -int(sg[q]))
            cut.append(sum(1 for j, s in enumerate(strs[b:e]) if (s, -(b + j)) < key))
        cut.append(e - b)
        wrong[r] = np.diff(cut)
    assert (wrong != iv).any()
    assert _bound_errors(ss, strs, wrong, spl, True) != []


@pytest.mark.parametrize("trunc", ["long", "short", "dist"])
@pytest.mark.parametrize("sampler", ALL_SAMPLERS)
@pytest.mark.parametrize("name,p", PIN_CASES)
def test_partition_samples_and_splitters(name, p, sampler, trunc):
    ss, strs, dist, kw, iv, (sgid, stl), (pg, pt) = _pin_case(name, p, sampler, trunc)
    indexed = sampler.startswith("Indexed")
    by_chars = sampler.endswith("Chars")
    factor = 2
    # samples come PE by PE, each PE's in ascending position
    pe_of = np.searchsorted(ss.pe_begin[1:], sgid, side="right")
    assert (np.diff(pe_of) >= 0).all()
    for r in range(p):
        b, e = int(ss.pe_begin[r]), int(ss.pe_begin[r + 1])
        m = e - b
        nr = min(factor * (p - 1), m)
        own = [int(g) - b for g in sgid[pe_of == r]]
        if not by_chars:
            assert own == [int(i * (m / (nr + 1))) for i in range(1, nr + 1)]
        else:
            # the indexed policy walks distinguishing prefixes in the
            # prefix-doubling form, the plain one lengths
            w = [int(dist[b + j]) if (dist is not None and indexed) else len(strs[b + j]) for j in range(m)]
            sd = sum(w) // (nr + 1)
            if sd == 0:
                # the reference reads index -1 here (undefined); the oracle
                # takes the first string every time
                assert own == [0] * nr
                continue
            prev = -1
            for t in own:
                assert t > prev
                run = sum(w[prev + 1:t + 1])
                assert sum(w[prev + 1:t]) < sd  # no earlier index reaches sd
                assert run >= sd or t == m - 1  # t reaches it, or the PE ended
                prev = t
            assert len(own) == nr or (own and own[-1] == m - 1)
        # truncation: max_len, or the string's distinguishing prefix
        for g, tl in zip(sgid[pe_of == r], stl[pe_of == r]):
            lim = int(dist[int(g)]) if dist is not None else kw["max_len"]
            assert int(tl) == min(len(strs[int(g)]), lim)
    # splitters: ranks i * floor(S / (nspl + 1)) of the textbook-sorted samples
    smp = [(strs[int(g)][:int(t)], int(g)) for g, t in zip(sgid, stl)]
    smp.sort()
    S = len(smp)
    nspl = min(p - 1, S)
    step = S // (nspl + 1)
    for q in range(p - 1):
        if q >= nspl:
            assert int(pg[q]) == MISSING
            continue
        exp = smp[(q + 1) * step]
        got = (strs[int(pg[q])][:int(pt[q])], int(pg[q]))
        if indexed:
            assert got == exp
        else:  # the order among equal sample strings is not part of parity
            assert got[0] == exp[0]


@pytest.mark.parametrize("sampler", ALL_SAMPLERS)
@pytest.mark.parametrize("name", sorted(PIN_SETS))
def test_partition_buckets_respect_the_order_all_samplers(name, sampler):
    """Every string of bucket q (any PE) precedes every string of bucket q + 1
    (any PE): by string, and for the indexed samplers strictly by (string,
    global index)."""
    p = 4
    ss, strs, dist, kw, iv, samples, spl = _pin_case(name, p, sampler, "long")
    indexed = sampler.startswith("Indexed")
    buckets = [[] for _ in range(p)]
    for r in range(p):
        b = int(ss.pe_begin[r])
        cut = np.concatenate([[0], np.cumsum(iv[r].astype(np.int64))])
        for q in range(p):
            for j in range(int(cut[q]), int(cut[q + 1])):
                buckets[q].append((strs[b + j], b + j))
    for q in range(p - 1):
        if buckets[q] and buckets[q + 1]:
            if indexed:
                assert max(buckets[q]) < min(buckets[q + 1])
            else:
                assert max(x[0] for x in buckets[q]) <= min(x[0] for x in buckets[q + 1])


@pytest.mark.parametrize("p", [1, 3, 4])
def test_merge_sort_no_lcps_is_textbook_sort(p):
    """NoLcps variant of the oracle: every PE's merged strings concatenated
    in PE order are the textbook sorted input, and the buckets are those of
    the partition with maximum splitter length 2 * 100 * (0 + 5)."""
    chars = INPUTS["skewed"]()
    ss = O.StringSet(chars, p)
    res = O.merge_sort(ss, encoder="no_lcps")
    allstr = sorted(ss.string(i) for i in range(ss.n))
    if p == 1:
        return
    got = [s for o in res["out"] for s in o["strings"]]
    assert got == allstr
    perm, _ = O.local_sort(ss)
    assert (res["interval"] == O.partition(ss, perm, "IndexedNumStrings", max_len=1000)).all()


@pytest.mark.parametrize("name", sorted(INPUTS))
def test_exact_dist_sums_to_computed(name):
    """The exact distinguishing prefixes of all strings sum to computeD of
    the globally sorted LCP array (DNGenerator_tests.cpp:22-45, the C
    restatement), and a string equal to another has length + 1."""
    chars = INPUTS[name]()
    if O.StringSet(chars, 1).n > 20000:
        chars = chars[:int(np.flatnonzero(chars == 0)[19999]) + 1]
    for p in (1, 3):
        ss = O.StringSet(chars, p)
        perm, _ = O.local_sort(ss)
        ex = O.exact_dist(ss, perm)
        one = O.StringSet(chars, 1)
        _, glcp = O.local_sort(one)
        assert int(ex.sum()) == O.compute_d(glcp)
        strs = [ss.string(int(perm[i])) for i in range(ss.n)]
        count = {}
        for s in strs:
            count[s] = count.get(s, 0) + 1
        for i, s in enumerate(strs):
            if count[s] > 1:
                assert ex[i] == len(s) + 1
    assert list(O.exact_dist(O.StringSet(arena([b"abc"]), 1), np.zeros(1, np.uint32))) == [0]


# ------------------------------------------- PE-parallel CPU restatement
@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("p", [2, 5, 16])
@pytest.mark.parametrize("name", sorted(INPUTS))
def test_pd_sort_mt_equals_single_thread(name, p, threads):
    """orc_pd_sort_mt (the benchmark's CPU baseline) gives the single-thread
    oracle's permutation, merged LCPs, results and round count."""
    chars = INPUTS[name]()
    ss = O.StringSet(chars, p)
    ref = O.prefix_doubling_sort(ss)
    mt = O.prefix_doubling_sort_mt(ss, threads=threads)
    assert mt["nrounds"] == ref["nrounds"]
    assert (mt["results"] == ref["results"]).all()
    si = np.concatenate([o["string_index"] for o in ref["out"]])
    pe = np.concatenate([o["pe_index"] for o in ref["out"]])
    lc = np.concatenate([o["lcp"] for o in ref["out"]])
    assert (mt["string_index"] == si).all() and (mt["pe_index"] == pe).all()
    assert (mt["lcp"] == lc).all()
    # D of the global order: computeD over the textbook-sorted input
    one = O.StringSet(chars, 1)
    _, glcp = O.local_sort(one)
    assert mt["D"] == O.compute_d(glcp)
