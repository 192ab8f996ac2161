"""GPU parity of the distributed stages (average LCP, partition, exchange,
LCP merge, prefix decompression, prefix doubling) and of the two composite
drivers against the oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from helpers import canonical_ties

pytestmark = pytest.mark.gpu

SAMPLERS = ["NumStrings", "NumChars", "IndexedNumStrings", "IndexedNumChars"]


def _dss():
    import dss_amd
    return dss_amd


def arena(strings):
    return np.frombuffer(b"".join(s + b"\0" for s in strings), dtype=np.uint8).copy()


def _short_with_long(rng):
    strs = [bytes(rng.integers(97, 101, int(l), dtype=np.uint8)) for l in rng.integers(1, 9, 20000)]
    for pos in rng.choice(len(strs), 6, replace=False):
        strs[int(pos)] = bytes(rng.integers(97, 101, 5000, dtype=np.uint8))
    return strs


def _tiny_with_long(rng):
    strs = [bytes(rng.integers(97, 100, int(l), dtype=np.uint8)) for l in rng.integers(0, 8, 20000)]
    for pos in rng.choice(len(strs), 300, replace=False):
        strs[int(pos)] = bytes(rng.integers(97, 100, 3000, dtype=np.uint8))
    return strs


def inputs():
    rng = np.random.default_rng(3)
    small = [bytes(rng.integers(97, 100, int(l), dtype=np.uint8)) for l in rng.integers(0, 12, 3000)]
    return {
        "dton": O.gen_dton(6000, 80, 0.5, 0x5EED0001),
        "skewed": O.gen_skewed(O.skew_params(6000, long_len=200, short_len=20), 0x5EED0003),
        "suffix": O.gen_suffixes(5000, 0x5EED0004, repeats=10, rep_min=100, rep_max=800, cap=300),
        "tiny_alphabet": arena(small),
        # mostly 1-8 characters with a few 5000-character strings: the short
        # copy path of the exchange, where a wave's byte range can exceed its
        # LDS buffer
        "short_with_long": arena(_short_with_long(rng)),
        # 0-7 characters with 300 strings of 3000: more than 33 bytes per
        # string on average, so the exchange copy is balanced over output
        # words, where one word can hold the ends of several short strings and
        # a 4 KiB chunk holds more strings than it stages in LDS
        "tiny_with_long": arena(_tiny_with_long(np.random.default_rng(5))),
    }


INPUTS = inputs()


def eq_runs(strings, lcp):
    """Boolean mask: position i continues a run of equal strings."""
    n = len(strings)
    m = np.zeros(n, bool)
    for i in range(1, n):
        m[i] = strings[i] == strings[i - 1]
    return m


def canon(keys, eqmask):
    """Sort key tuples inside runs of equal strings."""
    keys = list(keys)
    i = 0
    n = len(keys)
    while i < n:
        j = i + 1
        while j < n and eqmask[j]:
            j += 1
        keys[i:j] = sorted(keys[i:j])
        i = j
    return keys


@pytest.mark.parametrize("name", sorted(INPUTS))
@pytest.mark.parametrize("p", [3, 8, 40])  # 40: above the exchange's staged-table limit of 32 PEs
def test_merge_sort_stages(name, p):
    dss = _dss()
    chars = INPUTS[name]
    ss = O.StringSet(chars, p)
    operm, olcp = O.local_sort(ss)
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.sort_lcp()
        avg = ds.avg_lcp()
        assert avg == O.avg_lcp(olcp)
        max_len = 2 * 100 * (avg + 5)
        for sampler in SAMPLERS:
            giv = ds.partition(sampler, max_len=max_len)
            oiv = O.partition(ss, operm, sampler, max_len=max_len)
            assert (giv == oiv).all(), sampler
        # short truncation exercises prefix samples
        giv = ds.partition("IndexedNumStrings", max_len=3)
        oiv = O.partition(ss, operm, "IndexedNumStrings", max_len=3)
        assert (giv == oiv).all()
        iv = ds.partition("IndexedNumStrings", max_len=max_len)
        for mode in ["whole", "lcp_compact"]:
            rb, bb, nbytes = ds.exchange(mode)
            ex = O.exchange(ss, operm, olcp, iv, mode)
            assert (rb == ex["recv_begin"]).all() and (bb == ex["byte_begin"]).all()
            gbytes = ds.get_exchange_bytes(nbytes)
            assert gbytes.tobytes() == ex["bytes"].tobytes(), mode
            glcp, gpe, gpos = ds.get_exchange()
            assert (glcp == ex["lcp"]).all() and (gpe == ex["src_pe"]).all() and (gpos == ex["src_pos"]).all()
            ds.merge_lcp()
            gord, gml, gsv = ds.get_merged()
            for q in range(p):
                r0, r1 = int(rb[q]), int(rb[q + 1])
                b0, b1 = int(bb[q]), int(bb[q + 1])
                runs = iv[:, q].astype(np.uint64)
                a = ex["bytes"][b0:b1]
                oo, ol, osv = O.merge_losertree(a, ex["lcp"][r0:r1], runs, mode == "lcp_compact")
                assert (gml[r0:r1] == ol).all(), (mode, q)
                # strings in merged order (receive arena is identical on both sides)
                starts = np.concatenate([[0], np.flatnonzero(a == 0)[:-1] + 1]) if r1 > r0 else []
                recv = [a[int(s):].tobytes().split(b"\0", 1)[0] for s in starts]
                go = [int(x) - r0 for x in gord[r0:r1]]
                # with prefix compression a received string is a suffix; compare
                # the sequence of (saved lcp, stored chars) after canonical ties
                gk = [(recv[i], int(v)) for i, v in zip(go, gsv[r0:r1])]
                ok = [(recv[int(i)], int(v)) for i, v in zip(oo, osv)]
                if mode == "whole":
                    # strings and LCPs in merged order are unique
                    gstr = [recv[i] for i in go]
                    ostr = [recv[int(i)] for i in oo]
                    assert gstr == ostr
                else:
                    # saved LCPs (oldLcps) exist only with prefix compression;
                    # their order among equal strings is the tie order
                    assert sorted(gk) == sorted(ok)
            if mode == "lcp_compact":
                nb = ds.extend_prefix()
                gout, goff = ds.get_extended(nb)
                exp = []
                for q in range(p):
                    r0, r1 = int(rb[q]), int(rb[q + 1])
                    b0, b1 = int(bb[q]), int(bb[q + 1])
                    a = ex["bytes"][b0:b1]
                    oo, ol, osv = O.merge_losertree(a, ex["lcp"][r0:r1], iv[:, q].astype(np.uint64), True)
                    exp.append(O.extend_prefix(a, oo, osv).tobytes())
                assert gout.tobytes() == b"".join(exp)


@pytest.mark.parametrize("name", sorted(INPUTS))
@pytest.mark.parametrize("p", [2, 5, 37])
def test_prefix_doubling_stages(name, p):
    dss = _dss()
    chars = INPUTS[name]
    ss = O.StringSet(chars, p)
    operm, olcp = O.local_sort(ss)
    ores, orounds, onr = O.prefix_doubling(ss, operm, olcp)
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.sort_lcp()
        nr = ds.prefix_doubling()
        gres, grounds = ds.get_prefix_doubling()
        assert nr == onr
        assert (gres == ores).all()
        assert (grounds == orounds).all()
        for sampler in SAMPLERS:
            giv = ds.partition(sampler, use_dist=True)
            oiv = O.partition(ss, operm, sampler, dist=ores)
            assert (giv == oiv).all(), sampler
        iv = ds.partition("IndexedNumStrings", use_dist=True)
        rb, bb, nbytes = ds.exchange("pd_prefix")
        ex = O.exchange(ss, operm, olcp, iv, "pd_prefix", dist=ores)
        assert ds.get_exchange_bytes(nbytes).tobytes() == ex["bytes"].tobytes()


@pytest.mark.parametrize("m", [2**64 - 1, 997, 64])
def test_prefix_doubling_small_hash_range(m):
    """Hash ranges far below 2^64 force collisions (the Bloom-sized variant)."""
    dss = _dss()
    chars = INPUTS["skewed"]
    ss = O.StringSet(chars, 4)
    operm, olcp = O.local_sort(ss)
    ores, orounds, onr = O.prefix_doubling(ss, operm, olcp, m=m)
    with dss.DeviceStringSet.from_arena(chars, 4) as ds:
        ds.sort_lcp()
        nr = ds.prefix_doubling(m=m)
        gres, grounds = ds.get_prefix_doubling()
    assert nr == onr and (gres == ores).all() and (grounds == orounds).all()


@pytest.mark.parametrize("name", ["dton", "skewed"])
def test_prefix_doubling_bloom_8_bits_per_string_200k(name):
    """Hash range m = 8 n (8 bits per string) on 200k strings: results,
    rounds and round count equal the oracle's with the same m, collisions
    included (with 2^64 - 1 there are none in practice)."""
    dss = _dss()
    n = 200_000
    chars = (O.gen_dton(n, 120, 0.5, 0x5EED0012) if name == "dton" else
             O.gen_skewed(O.skew_params(n, long_len=300, short_len=30), 0x5EED0013))
    ss = O.StringSet(chars, 8)
    m = 8 * ss.n
    operm, olcp = O.local_sort(ss)
    ores, orounds, onr = O.prefix_doubling(ss, operm, olcp, m=m)
    with dss.DeviceStringSet.from_arena(chars, 8) as ds:
        ds.sort_lcp()
        nr = ds.prefix_doubling(m=m)
        gres, grounds = ds.get_prefix_doubling()
    assert nr == onr and (gres == ores).all() and (grounds == orounds).all()
    # m = 8 n does collide here: some results exceed the 2^64 - 1 ones
    eres, _, _ = O.prefix_doubling(ss, operm, olcp)
    assert (ores >= eres).all() and (ores > eres).any()


def test_hash_prefixes_match_reference_hash():
    dss = _dss()
    rng = np.random.default_rng(9)
    strs = [bytes(rng.integers(1, 256, l, dtype=np.uint8)) for l in range(0, 130)]
    chars = arena(strs)
    with dss.DeviceStringSet.from_arena(chars, 1) as ds:
        ids = np.arange(len(strs), dtype=np.uint32)
        for depth in [1, 7, 8, 31, 32, 33, 64, 100, 129]:
            got = ds.hash_prefixes(ids, depth)
            exp = [O.pd_hash(s[:depth]) for s in strs]
            assert list(map(int, got)) == exp, depth


@pytest.mark.parametrize("name", sorted(INPUTS))
def test_drivers(name):
    dss = _dss()
    chars = INPUTS[name]
    p = 4
    ss = O.StringSet(chars, p)
    # merge sort with LCP compression: the rebuilt arena is unique
    res = O.merge_sort(ss, encoder="lcp_compact")
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.merge_sort(compressed=True)
        gord, gml, gsv = ds.get_merged()
        exp = b"".join(o["arena"].tobytes() for o in res["out"])
        got, _ = ds.get_extended(len(exp))
        assert got.tobytes() == exp
        assert (gml == np.concatenate([o["lcp"] for o in res["out"]])).all()
    # prefix doubling: permutation modulo ties among equal strings
    res = O.prefix_doubling_sort(ss)
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.prefix_doubling_sort()
        si, pi = ds.get_permutation()
        gperm, glcp = ds.get_sorted()
    exp_pairs = []
    for o in res["out"]:
        exp_pairs += list(zip(map(int, o["pe_index"]), map(int, o["string_index"])))
    got_pairs = list(zip(map(int, pi), map(int, si)))

    def full(pair, perm):
        pe, idx = pair
        return ss.string(perm[int(ss.pe_begin[pe]) + idx])
    gs = [full(x, gperm) for x in got_pairs]
    es = [full(x, res["perm"]) for x in exp_pairs]
    assert gs == es
    eqm = np.zeros(len(gs), bool)
    eqm[1:] = [gs[i] == gs[i - 1] for i in range(1, len(gs))]
    assert canon(got_pairs, eqm) == canon(exp_pairs, eqm)


@pytest.mark.parametrize("name", sorted(INPUTS))
@pytest.mark.parametrize("p", [5, 16, 37])
def test_merge_fan_in_gives_the_same_output(name, p):
    """The K of the K-way merge (runs merged per pass: 2, 4, 8, 16) changes
    the number of passes, and the pairwise merge's tile size (512, 1024, 2048
    outputs) its tiling, not the output: merged order, LCPs and saved LCPs
    are identical for every setting (ties go to the earlier run), the merged
    strings equal the oracle's loser tree and the checker passes."""
    dss = _dss()
    chars = INPUTS[name]
    ss = O.StringSet(chars, p)
    operm, olcp = O.local_sort(ss)
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.sort_lcp()
        iv = ds.partition("IndexedNumStrings", max_len=1000)
        for mode in ("whole", "lcp_compact"):
            rb, bb, nbytes = ds.exchange(mode)
            ex = O.exchange(ss, operm, olcp, iv, mode)
            ref = None
            for fan, tile in ((16, 1024), (8, 1024), (4, 1024), (2, 512), (2, 2048), (2, 1024)):
                ds.set_sort_params({"merge_fan_in": fan, "merge_tile": tile})
                ds.merge_lcp()
                got = ds.get_merged()
                if ref is None:
                    ref = got
                    for q in range(p):
                        r0, r1 = int(rb[q]), int(rb[q + 1])
                        b0, b1 = int(bb[q]), int(bb[q + 1])
                        oo, ol, osv = O.merge_losertree(ex["bytes"][b0:b1], ex["lcp"][r0:r1],
                                                        iv[:, q].astype(np.uint64), mode == "lcp_compact")
                        assert (got[1][r0:r1] == ol).all(), (mode, q)
                else:
                    for a, b in zip(got, ref):
                        assert (a == b).all(), (mode, fan)
                chk = ds.check()
                assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0, (mode, fan)
        ds.set_sort_params(None)


def test_caller_stream_gives_the_same_results():
    """dss_set_stream: the whole prefix-doubling sort queued on a caller's
    HIP stream gives the own-stream results; the set returns to its own
    stream with NULL and never destroys the caller's."""
    from helpers import HipMem
    dss = _dss()
    chars = INPUTS["skewed"]
    with dss.DeviceStringSet.from_arena(chars, 5) as ds, HipMem() as hm:
        ds.prefix_doubling_sort()
        ref = ds.get_permutation(), ds.get_prefix_doubling()
        st = hm.stream()
        ds.set_stream(st)
        ds.prefix_doubling_sort()
        got = ds.get_permutation(), ds.get_prefix_doubling()
        chk = ds.check()
        ds.set_stream(None)
        ds.merge_sort(compressed=True)  # back on its own stream
        assert ds.check()["sorted"]
    for a, b in zip(ref, got):
        for x, y in zip(a, b):
            assert (x == y).all()
    assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0


def test_few_strings_many_pes():
    dss = _dss()
    chars = arena([b"b", b"a", b"c", b"a"])
    ss = O.StringSet(chars, 8)
    res = O.merge_sort(ss, encoder="lcp_compact")
    with dss.DeviceStringSet.from_arena(chars, 8) as ds:
        ds.merge_sort(compressed=True)
        exp = b"".join(o["arena"].tobytes() for o in res["out"])
        got, _ = ds.get_extended(len(exp))
        assert got.tobytes() == exp


def _few_chars_pes(rng, p):
    """Per-PE inputs where PE 1 has 6 strings of 2 characters in total and
    PE 2 three strings of 1 character: with p = 4 and factor 2 their
    NumChars walks have splitter_dist 0 and emit their first string several
    times (identical samples)."""
    pes = []
    for r in range(p):
        if r == 1:
            strs = [b"", b"b", b"", b"", b"a", b""]
        elif r == 2:
            strs = [b"", b"c", b""]
        else:
            strs = [bytes(rng.integers(97, 100, int(l), dtype=np.uint8)) for l in rng.integers(0, 6, 200)]
        pes.append(arena(strs))
    return pes


@pytest.mark.parametrize("use_dist", [False, True])
def test_partition_pe_with_fewer_chars_than_splitters(use_dist):
    """Identical samples (the same string and index, emitted more than once
    by a NumChars walk) get distinct ranks in the global sample order, so
    every splitter slot is written; buckets equal the oracle's for all four
    samplers, with and without prefix-doubling lengths."""
    dss = _dss()
    p = 4
    pes = _few_chars_pes(np.random.default_rng(31), p)
    ss = O.StringSet.from_pes(pes)
    operm, olcp = O.local_sort(ss)
    ores = O.prefix_doubling(ss, operm, olcp)[0] if use_dist else None
    with dss.DeviceStringSet.from_pes(pes) as ds:
        ds.sort_lcp()
        if use_dist:
            ds.prefix_doubling()
        for rep in range(2):  # the second call reuses the sample buffers
            for sampler in SAMPLERS:
                giv = ds.partition(sampler, max_len=1000, use_dist=use_dist)
                oiv = O.partition(ss, operm, sampler, max_len=1000, dist=ores)
                assert (giv == oiv).all(), (sampler, rep, giv, oiv)
                assert (giv.sum(axis=1) == np.diff(ss.pe_begin)).all()
        # the whole driver on the same input
        ds.merge_sort(sampler="IndexedNumChars", compressed=False)
        chk = ds.check()
        assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0


@pytest.mark.parametrize("m", [2**64 - 1, 1 << 20, 4093])
def test_prefix_doubling_shared_prefix_rounds(m):
    """D/N 0.5 over 500 chars: in the depth-264 round every string is hashed
    while sharing ~246 leading characters with its predecessor, so its first
    224 characters are read from the head of its run.  With small hash ranges
    the candidate sets depend on every hash value, so any difference from the
    oracle's hashes of the strings' own bytes shows up in the results."""
    dss = _dss()
    chars = O.gen_dton(20000, 500, 0.5, 0x5EED0007)
    ss = O.StringSet(chars, 4)
    operm, olcp = O.local_sort(ss)
    ores, orounds, onr = O.prefix_doubling(ss, operm, olcp, m=m)
    assert onr >= 6  # the depth-264 round ran
    with dss.DeviceStringSet.from_arena(chars, 4) as ds:
        ds.sort_lcp()
        nr = ds.prefix_doubling(m=m)
        gres, grounds = ds.get_prefix_doubling()
    assert nr == onr and (gres == ores).all() and (grounds == orounds).all()


def _pd_trace(text):
    """(depth, hashed, shared) of every round the library's round trace
    (DSS_PD_TRACE) printed."""
    out = []
    for line in text.splitlines():
        if line.startswith("[dss pd] round"):
            f = line.split()
            out.append((int(f[5]), int(f[9]), int(f[11])))
    return out


@pytest.mark.parametrize("limit", [None, "8"])
@pytest.mark.parametrize("m", [2**64 - 1, 4093, 64])
@pytest.mark.parametrize("name", ["skewed", "dton", "dton500"])
def test_prefix_doubling_partitioned_duplicates(name, m, limit, monkeypatch, capfd):
    """Rounds with many hashed candidates find repeated hashes by partitioning
    them into buckets and checking each bucket in an LDS table; buckets with
    more distinct hashes than the table allows (a limit of 8 forces that)
    fall back to the exact global table.  Forced on for every round here,
    both must give the oracle's results, collisions of small ranges included.
    The round trace shows which hash layout each round used: shared-prefix
    rounds store hashes by position (the 500-character D/N input's depth-264
    round), the others in list order (the skewed input's rounds)."""
    dss = _dss()
    if name == "skewed":
        chars = INPUTS[name]
    elif name == "dton":
        chars = O.gen_dton(20000, 200, 0.5, 0x5EED0008)
    else:
        chars = O.gen_dton(20000, 500, 0.5, 0x5EED0007)
    ss = O.StringSet(chars, 4)
    operm, olcp = O.local_sort(ss)
    ores, orounds, onr = O.prefix_doubling(ss, operm, olcp, m=m)
    monkeypatch.setenv("DSS_PD_PARTITION_MIN", "1")
    monkeypatch.setenv("DSS_PD_TRACE", "1")
    if limit:
        monkeypatch.setenv("DSS_PD_TABLE_LIMIT", limit)
    capfd.readouterr()
    with dss.DeviceStringSet.from_arena(chars, 4) as ds:
        ds.sort_lcp()
        nr = ds.prefix_doubling(m=m)
        gres, grounds = ds.get_prefix_doubling()
    trace = _pd_trace(capfd.readouterr().err)
    assert nr == onr and (gres == ores).all() and (grounds == orounds).all()
    assert trace and len(trace) <= nr, trace  # rounds skipped up front print nothing
    layouts = {shared for _, hashed, shared in trace if hashed}
    if name == "skewed":
        assert 0 in layouts, trace
    elif name == "dton500":
        assert 1 in layouts, trace


@pytest.mark.parametrize("name", sorted(INPUTS))
@pytest.mark.parametrize("p", [4, 16])
def test_merge_sort_no_lcps(name, p):
    """The NoLcps variant: partition with the LCP average of the never-computed
    LCP array (0 + 5), whole strings sent without LCPs, the receiver's run
    LCPs derived from the received strings (equal to what mode 0 would have
    sent), merged strings per PE equal to noLcpMerge's."""
    dss = _dss()
    chars = INPUTS[name]
    ss = O.StringSet(chars, p)
    res = O.merge_sort(ss, encoder="no_lcps")
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.merge_sort(compressed="no_lcps")
        rb, bb = None, None
        glcp, gpe, gpos = ds.get_exchange()
        ex = res["exchange"]
        assert (glcp == ex["lcp"]).all() and (gpe == ex["src_pe"]).all() and (gpos == ex["src_pos"]).all()
        nbytes = int(ex["byte_begin"][-1])
        gbytes = ds.get_exchange_bytes(nbytes)
        assert gbytes.tobytes() == ex["bytes"].tobytes()
        gord, gml, _ = ds.get_merged()
        chk = ds.check()
        assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0
    for q in range(p):
        r0, r1 = int(ex["recv_begin"][q]), int(ex["recv_begin"][q + 1])
        b0, b1 = int(ex["byte_begin"][q]), int(ex["byte_begin"][q + 1])
        a = ex["bytes"][b0:b1].tobytes()
        recv = a.split(b"\0")[:-1] if a else []
        got = [recv[int(i) - r0] for i in gord[r0:r1]]
        assert got == res["out"][q]["strings"], q


@pytest.mark.parametrize("name", sorted(INPUTS))
def test_exact_dist_and_prefix_doubling_bounds(name):
    """ExcatDistinguishingPrefix on the device equals the oracle's after
    every driver; prefix doubling's results are upper bounds of the exact
    lengths and, with the full hash range, below 2 * DIST + 8; the exact
    lengths sum to the checker's D."""
    dss = _dss()
    chars = INPUTS[name]
    p = 4
    ss = O.StringSet(chars, p)
    operm, _ = O.local_sort(ss)
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        for enc in ("whole", "lcp_compact", "no_lcps"):
            ds.merge_sort(compressed=enc)
            gperm, _ = ds.get_sorted()
            gex = ds.exact_dist()
            # positions hold the same strings on both sides (ties aside)
            oex = O.exact_dist(ss, gperm)
            assert (gex == oex).all(), enc
            assert int(gex.sum()) == ds.check()["D"]
        ds.prefix_doubling_sort()
        gperm, _ = ds.get_sorted()
        gex = ds.exact_dist()
        assert (gex == O.exact_dist(ss, gperm)).all()
        res, _ = ds.get_prefix_doubling()
        c = ds.check_dist()
        assert c["lower_violations"] == 0 and c["upper_violations"] == 0, c
        assert c["exact_sum"] == int(gex.sum()) and c["results_sum"] == int(res.sum())
        # a tiny hash range makes candidates survive collisions: results stay
        # upper bounds
        ds.prefix_doubling_sort(m=64)
        c = ds.check_dist()
        assert c["lower_violations"] == 0


@pytest.mark.parametrize("p", [4, 16])
def test_prefix_doubling_exchange_from_kept_characters(p):
    """Long shared prefixes: the last rounds hash from run heads' states and
    keep each string's last stripe and tail, from which the exchange copies
    the distinguishing prefixes and the merge's cached keys.  Receive bytes,
    merged LCPs and the permutation match the oracle."""
    dss = _dss()
    chars = O.gen_dton(20000, 300, 0.5, 0x5EED0002)
    ss = O.StringSet(chars, p)
    res = O.prefix_doubling_sort(ss)
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.prefix_doubling_sort()
        gres, grounds = ds.get_prefix_doubling()
        assert (gres == res["results"]).all() and (grounds == res["rounds"]).all()
        ex = res["exchange"]
        nbytes = int(ex["byte_begin"][-1])
        assert ds.get_exchange_bytes(nbytes).tobytes() == ex["bytes"].tobytes()
        gord, gml, gsv = ds.get_merged()
        assert (gml == np.concatenate([o["lcp"] for o in res["out"]])).all()
        si, pi = ds.get_permutation()
        exp = [(int(a), int(b)) for o in res["out"] for a, b in zip(o["pe_index"], o["string_index"])]
        assert list(zip(map(int, pi), map(int, si))) == exp  # D/N strings are distinct: no ties
        chk = ds.check()
        assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0


def _rows_inputs():
    rng = np.random.default_rng(11)
    base = [bytes(s) for s in O.gen_dton(8000, 500, 0.5, 0x5EED0009).tobytes().split(b"\0")[:-1]]
    # every string twice, the copies in the other half of the PEs: a
    # duplicate on another PE keeps each string a candidate past the round
    # after its local LCPs, up to its end
    dup = base + [base[int(i)] for i in rng.permutation(len(base))]
    # D/N strings cut at 200-300 characters: some end before the round
    # after their local LCPs
    cut = [x[:int(l)] for x, l in zip(base, rng.integers(200, 301, len(base)))]
    return {"dton500": arena(base), "cross_pe_duplicates": arena(dup), "cut_200_300": arena(cut)}


ROWS_INPUTS = _rows_inputs()


@pytest.mark.parametrize("pass_mode", [1, 2])
@pytest.mark.parametrize("m", [2**64 - 1, 4093])
@pytest.mark.parametrize("name", sorted(ROWS_INPUTS))
def test_prefix_doubling_sort_kept_rows_edge_cases(name, m, pass_mode):
    """The composite driver on 500-character D/N strings, where the depth-264
    round hashes every string from its run head's state and keeps its last
    stripe and tail for the exchange: rows that are the last ones (distinct
    strings), rows overwritten by later rounds (exact duplicates on other
    PEs, small hash ranges) and strings that end before that round.  Results,
    rounds, receive bytes and merged LCPs match the oracle, twice in a row
    (the second sort reuses the first one's buffers), for host- and
    device-driven local sorts (16 PEs)."""
    dss = _dss()
    chars = ROWS_INPUTS[name]
    p = 4 if pass_mode == 1 else 16
    ss = O.StringSet(chars, p)
    res = O.prefix_doubling_sort(ss, m=m)
    assert max(res["rounds"]) >= 6  # the depth-264 round ran
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.set_sort_params({"pass_mode": pass_mode})
        for _ in range(2):
            ds.prefix_doubling_sort(m=m)
            gres, grounds = ds.get_prefix_doubling()
            assert (gres == res["results"]).all() and (grounds == res["rounds"]).all()
            ex = res["exchange"]
            nbytes = int(ex["byte_begin"][-1])
            assert ds.get_exchange_bytes(nbytes).tobytes() == ex["bytes"].tobytes()
            gord, gml, gsv = ds.get_merged()
            assert (gml == np.concatenate([o["lcp"] for o in res["out"]])).all()
            chk = ds.check()
            assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0


def test_exchange_origin_follows_each_exchange():
    """The received strings' origin (source PE, position) is computed on the
    first request after an exchange: repeated exchanges of one set with other
    send counts report the origins of the latest one (oracle: orc_exchange)."""
    dss = _dss()
    rng = np.random.default_rng(77)
    p = 5
    chars = O.gen_dton(6000, 40, 0.5, 78)
    ss = O.StringSet(chars, p)
    perm, lcp = O.local_sort(ss)
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.set_local_order(perm, lcp)
        for k in range(3):
            iv = np.zeros((p, p), np.uint64)
            for r in range(p):
                cnt = int(ss.pe_begin[r + 1] - ss.pe_begin[r])
                cuts = np.sort(rng.integers(0, cnt + 1, p - 1))
                iv[r] = np.diff(np.concatenate([[0], cuts, [cnt]]))
            ds.set_interval(iv)
            ds.exchange("whole")
            glcp, gpe, gpos = ds.get_exchange()
            ex = O.exchange(ss, perm, lcp, iv, "whole")
            assert (glcp == ex["lcp"]).all(), k
            assert (gpe == ex["src_pe"]).all() and (gpos == ex["src_pos"]).all(), k
