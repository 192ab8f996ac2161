"""GPU parity of the local LCP sort (tlx radixsort_CI3 call sites) against the
oracle's restatement, through the C ABI, on the project's seeded inputs and
on the edge cases the reference's sorter has to handle."""
import numpy as np
import pytest

import oracle as O
from helpers import canonical_ties

pytestmark = pytest.mark.gpu


def _dss():
    import dss_amd
    return dss_amd


def check_sort(chars, p):
    dss = _dss()
    ss = O.StringSet(chars, p)
    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        assert ds.n == ss.n
        assert (ds.pe_begin() == ss.pe_begin).all()
        ds.sort_lcp()
        gperm, glcp = ds.get_sorted()
    operm, olcp = O.local_sort(ss)
    assert (glcp == olcp).all(), "LCP arrays differ"
    a = canonical_ties(gperm, glcp, ss.len, ss.pe_begin)
    b = canonical_ties(operm, olcp, ss.len, ss.pe_begin)
    assert (a == b).all(), "sorted permutations differ"
    return ss


def arena(strings):
    return np.frombuffer(b"".join(s + b"\0" for s in strings), dtype=np.uint8).copy()


@pytest.mark.parametrize("p", [1, 4, 16])
def test_dton_small(p):
    check_sort(O.gen_dton(5000, 60, 0.5, 0x5EED0001), p)


def test_dton_c1_shape_scaled():
    # C1 construction (500 chars, D/N 0.5) at 40k strings, 4 PEs
    check_sort(O.gen_dton(40000, 500, 0.5, 0x5EED0001), 4)


@pytest.mark.parametrize("dn", [0.0, 0.25, 0.75, 1.0])
def test_dton_ratios(dn):
    check_sort(O.gen_dton(20000, 100, dn, 11), 4)


def test_skewed_with_duplicates():
    params = O.skew_params(30000, long_len=300, short_len=30)
    check_sort(O.gen_skewed(params, 0x5EED0003), 16)


def test_suffixes_high_lcp():
    check_sort(O.gen_suffixes(30000, 0x5EED0004, repeats=20, rep_min=200, rep_max=2000, cap=500), 4)


def test_random_large_single_pe():
    rng = np.random.default_rng(5)
    lens = rng.integers(0, 40, 300000)
    strs = [bytes(rng.integers(1, 4, l, dtype=np.uint8)) for l in lens]
    check_sort(arena(strs), 1)


def test_edge_cases():
    # empty strings, identical strings above and below the base-case size,
    # one-string PEs, fewer strings than PEs, prefix chains
    check_sort(arena([b""] * 10 + [b"a"] * 3 + [b""]), 2)
    check_sort(arena([b"same"] * 5000), 3)
    check_sort(arena([b"x" * 900] * 3000 + [b"x" * 899 + b"y"]), 1)
    check_sort(arena([b"only"]), 1)
    check_sort(arena([b"b", b"a"]), 8)
    check_sort(arena([b"a" * k for k in range(300, 0, -1)]), 2)
    check_sort(arena([bytes([255]) * 3, bytes([1]), bytes([255, 1]), bytes([128]) * 17]), 1)


def test_device_generators_match_oracle():
    dss = _dss()
    n, L, dn, seed = 7777, 123, 0.5, 99
    with dss.DeviceStringSet.dton(n, L, dn, seed, 4) as ds:
        chars, off, ln = ds.download()
    exp = O.gen_dton(n, L, dn, seed)
    assert chars.tobytes() == exp.tobytes()
    params = O.skew_params(5000, long_len=200, short_len=20)
    with dss.DeviceStringSet.skewed(5000, 42, 4, long_len=200, short_len=20) as ds:
        chars, off, ln = ds.download()
    exp = O.gen_skewed(params, 42)
    assert chars.tobytes() == exp.tobytes()


@pytest.mark.parametrize("n,L,dn,seed", [
    (1, 1, 0.5, 1), (2, 3, 0.0, 7), (100, 5, 0.5, 3), (1000, 7, 1.0, 9), (3000, 1, 0.0, 4),
    (20000, 13, 0.7, 8), (700, 4095, 0.01, 6), (9, 4096, 0.5, 10), (26 ** 3 + 5, 2, 0.0, 12)])
def test_dton_generator_shapes(n, L, dn, seed):
    # strides of 2..4097 bytes: strings shorter than a word, words across
    # several strings, strings longer than a 4096-byte chunk, a ragged tail
    dss = _dss()
    with dss.DeviceStringSet.dton(n, L, dn, seed, 1) as ds:
        chars, off, ln = ds.download()
    assert chars.tobytes() == O.gen_dton(n, L, dn, seed).tobytes()


@pytest.mark.parametrize("n,ll,sl,ldn,sdn,lp,dp", [
    (1, 9, 1, 0.5, 0.0, 500, 0), (63, 16, 3, 0.3, 0.5, 200, 100), (64, 600, 7, 0.75, 0.25, 200, 100),
    (65, 40, 8, 0.9, 0.1, 900, 300), (4097, 5000, 2, 0.5, 0.0, 20, 500), (30000, 120, 12, 0.6, 0.3, 250, 50)])
def test_skewed_generator_shapes(n, ll, sl, ldn, sdn, lp, dp):
    dss = _dss()
    seed = 0x5EED00A0 + n
    with dss.DeviceStringSet.skewed(n, seed, 1, long_len=ll, short_len=sl, long_dn=ldn, short_dn=sdn,
                                    long_permille=lp, dup_permille=dp) as ds:
        chars, off, ln = ds.download()
    exp = O.gen_skewed(O.skew_params(n, long_len=ll, short_len=sl, long_dn=ldn, short_dn=sdn,
                                     long_permille=lp, dup_permille=dp), seed)
    assert chars.tobytes() == exp.tobytes()


PARAM_SETS = [
    {"large_min": 16, "wave_class_max": 0, "skip_lanes": 8, "skip_strings": 1, "merge_fan_in": 2, "pass_mode": 1,
     "merge_tile": 512, "digit_chars": 3},
    {"large_min": 64, "wave_class_max": 6, "skip_lanes": 16, "skip_strings": 2, "merge_fan_in": 4, "pass_mode": 2,
     "merge_tile": 2048, "digit_chars": 3},
    {"large_min": 256, "wave_class_max": 5, "skip_lanes": 8, "skip_strings": 4, "merge_fan_in": 8, "pass_mode": 0,
     "merge_tile": 1024, "digit_chars": 0},
    {"large_min": 1024, "wave_class_max": 6, "skip_lanes": 16, "skip_strings": 1, "merge_fan_in": 16, "pass_mode": 2,
     "merge_tile": 512, "digit_chars": 2},
    {"large_min": 128, "wave_class_max": 4, "skip_lanes": 8, "skip_strings": 2, "merge_fan_in": 2, "pass_mode": 1,
     "merge_tile": 2048, "digit_chars": 2},
    {"large_min": 32, "wave_class_max": 6, "skip_lanes": 16, "skip_strings": 4, "merge_fan_in": 16, "pass_mode": 2,
     "merge_tile": 1024, "digit_chars": 1},
    {"large_min": 16, "wave_class_max": 3, "skip_lanes": 8, "skip_strings": 1, "merge_fan_in": 2, "pass_mode": 2,
     "merge_tile": 512, "digit_chars": 2},
]
DEFAULT_PARAMS = {"large_min": 512, "wave_class_max": 6, "skip_lanes": 8, "skip_strings": 4, "merge_fan_in": 2,
                  "pass_mode": 0, "merge_tile": 1024, "digit_chars": 0}


def _small_alphabet(n, alphabet, max_len, seed, late=b""):
    """n strings over `alphabet` of lengths 0..max_len, a tenth of them
    repeated and a tenth cut to prefixes of others (ends inside digits), then
    the strings of `late` at the end of the arena."""
    rng = np.random.default_rng(seed)
    sym = np.frombuffer(alphabet, dtype=np.uint8)
    out = []
    for _ in range(n):
        k = int(rng.integers(0, max_len + 1))
        out.append(sym[rng.integers(0, sym.size, k)].tobytes())
    for i in rng.integers(0, n, n // 10):
        out.append(out[int(i)])
    for i in rng.integers(0, n, n // 10):
        t = out[int(i)]
        out.append(t[: int(rng.integers(0, len(t) + 1))])
    order = rng.permutation(len(out))
    return arena([out[int(i)] for i in order] + [t for t in late.split(b",") if late])


MULTICHAR_INPUTS = {
    # sigma = values + 1 (the end): 5 -> 3 characters per digit, 3 -> 3, 11 -> 2
    "acgt": lambda: _small_alphabet(60000, b"ACGT", 30, 41),
    "binary": lambda: _small_alphabet(60000, b"01", 40, 42),
    "ten": lambda: _small_alphabet(60000, b"abcdefghij", 20, 43),
    "suffixes": lambda: O.gen_suffixes(30000, 44, repeats=20, rep_min=200, rep_max=2000, cap=300),
    # more than the 1 MiB the alphabet scan samples first; the two strings
    # after it add 'Z' and 'a' (sigma 7: 2 characters per digit)
    "late_symbols": lambda: _small_alphabet(90000, b"ACGT", 24, 45, late=b"ACGZT,ZZ,aAC,ACGTa"),
}


@pytest.mark.parametrize("large_min", [16, 512])
@pytest.mark.parametrize("name", sorted(MULTICHAR_INPUTS))
def test_multichar_digits_match_one_char_digits(name, large_min):
    """Radix digits of two or three characters on small alphabets (the byte
    values the arena holds, ranked, plus the end as rank 0) give the same
    permutation and LCP array as one-character digits, and the oracle's LCPs
    and order: bucket boundaries inside a digit, strings that end inside a
    digit (duplicates, prefixes, empty strings), keys re-gathered when the
    next digit leaves the 8 cached characters, and byte values that appear
    only after the scanned sample."""
    dss = _dss()
    chars = MULTICHAR_INPUTS[name]()
    assert name != "late_symbols" or chars.size > (1 << 20)
    ss = O.StringSet(chars, 4)
    operm, olcp = O.local_sort(ss)
    with dss.DeviceStringSet.from_arena(chars, 4) as ds:
        out = {}
        for dc in (1, 2, 3):
            ds.set_sort_params({"large_min": large_min, "pass_mode": 1, "digit_chars": dc})
            ds.sort_lcp()
            out[dc] = ds.get_sorted()
    for dc in (2, 3):
        assert (out[dc][0] == out[1][0]).all() and (out[dc][1] == out[1][1]).all(), dc
    assert (out[1][1] == olcp).all()
    a = canonical_ties(out[1][0], out[1][1], ss.len, ss.pe_begin)
    assert (a == canonical_ties(operm, olcp, ss.len, ss.pe_begin)).all()


@pytest.mark.parametrize("name", ["dton", "skewed", "suffix"])
def test_sort_params_do_not_change_the_output(name):
    """Every tuning of the local sort returns the same permutation and LCPs
    (ties by id), equal to the oracle's order and LCPs."""
    dss = _dss()
    chars = {"dton": lambda: O.gen_dton(30000, 80, 0.5, 31),
             "skewed": lambda: O.gen_skewed(O.skew_params(30000, long_len=200, short_len=20), 32),
             "suffix": lambda: O.gen_suffixes(30000, 33, repeats=10, rep_min=200, rep_max=2000, cap=300)}[name]()
    ss = O.StringSet(chars, 4)
    operm, olcp = O.local_sort(ss)
    with dss.DeviceStringSet.from_arena(chars, 4) as ds:
        ds.sort_lcp()
        ref = ds.get_sorted()
        for params in PARAM_SETS:
            ds.set_sort_params(params)
            assert ds.sort_params() == params
            ds.sort_lcp()
            perm, lcp = ds.get_sorted()
            assert (perm == ref[0]).all() and (lcp == ref[1]).all(), params
        ds.set_sort_params(None)
        assert ds.sort_params() == DEFAULT_PARAMS
    assert (ref[1] == olcp).all()
    assert (canonical_ties(ref[0], ref[1], ss.len, ss.pe_begin) == canonical_ties(operm, olcp, ss.len, ss.pe_begin)).all()


def test_autotuner_picks_and_applies_a_setting(tmp_path):
    dss = _dss()
    from dss_amd.tune import Tuner
    tuner = Tuner(str(tmp_path / "t.json"))
    with dss.DeviceStringSet.dton(200000, 100, 0.5, 34, 4) as ds:
        out = tuner.autotune(ds, repeats=1)
        later = ("merge_fan_in", "pass_mode", "merge_tile", "digit_chars")
        sort_best = {k: v for k, v in out["best"].items() if k not in later}
        assert sort_best in [{k: v for k, v in c.items() if k not in later} for _, c in out["times_ms"]]
        if out["best"]["merge_fan_in"] == 2:
            assert out["best"]["merge_tile"] == out["merge_tile_times_ms"][0][1]
            assert sorted(m for _, m in out["merge_tile_times_ms"]) == [512, 1024, 2048]
        assert out["best"]["digit_chars"] == out["digit_times_ms"][0][1]
        assert sorted(d for _, d in out["digit_times_ms"]) == [1, 2, 3]
        assert out["best"]["pass_mode"] == out["pass_times_ms"][0][1]
        assert sorted(m for _, m in out["pass_times_ms"]) == [1, 2]
        assert out["best"]["merge_fan_in"] == out["merge_times_ms"][0][1]
        assert sorted(f for _, f in out["merge_times_ms"]) == [2, 4, 8, 16]
        assert ds.sort_params() == out["best"]
        tuner.save()
    with dss.DeviceStringSet.dton(200000, 100, 0.5, 35, 4) as ds:
        ds.sort_lcp()
        assert Tuner(str(tmp_path / "t.json")).apply(ds)
        assert ds.sort_params() == out["best"]
    with dss.DeviceStringSet.from_arena(arena([b"a", b"b"]), 1) as ds:
        with pytest.raises(dss.DssError):
            ds.set_sort_params({"large_min": 2000})
        with pytest.raises(dss.DssError):
            ds.set_sort_params({"skip_lanes": 12})
        with pytest.raises(dss.DssError):
            ds.set_sort_params({"skip_strings": 3})
        with pytest.raises(dss.DssError):
            ds.set_sort_params({"merge_fan_in": 3})
        with pytest.raises(dss.DssError):
            ds.set_sort_params({"pass_mode": 3})
        with pytest.raises(dss.DssError):
            ds.set_sort_params({"merge_tile": 768})
        with pytest.raises(dss.DssError):
            ds.set_sort_params({"digit_chars": 4})


def _two_letter_strings(n, length, seed):
    """Strings over {a, b}: every MSD pass halves the segments, so sorting
    2^16 of them takes more passes than the device-driven form captures."""
    rng = np.random.default_rng(seed)
    return [bytes(rng.integers(97, 99, int(l), dtype=np.uint8)) for l in rng.integers(0, length + 1, n)]


DEVICE_DRIVEN_INPUTS = {
    "dton": lambda: O.gen_dton(30000, 120, 0.5, 41),
    "skewed": lambda: O.gen_skewed(O.skew_params(30000, long_len=300, short_len=30), 42),
    "suffix": lambda: O.gen_suffixes(30000, 43, repeats=20, rep_min=100, rep_max=1500, cap=400),
    "two_letters_deep": lambda: arena(_two_letter_strings(1 << 16, 40, 44)),
    "tiny": lambda: arena([b"", b"b", b"a", b"", b"ab", b"a"]),
}


@pytest.mark.parametrize("name", sorted(DEVICE_DRIVEN_INPUTS))
@pytest.mark.parametrize("p", [1, 5, 64])
def test_device_driven_passes_match_host_driven(name, p):
    """The device-driven local sort (persistent kernels, two hipGraphs
    replayed, one read-back) gives the host-driven form's permutation, LCPs
    and sorted offsets/lengths and the oracle's order; three replays in a row,
    with eager work and synchronisation between them, give the same result
    each time (p = 64 also leaves PEs empty or with one string)."""
    dss = _dss()
    chars = DEVICE_DRIVEN_INPUTS[name]()
    ss = O.StringSet(chars, p)
    operm, olcp = O.local_sort(ss)
    def exchanged(ds):
        # the exchange reads the sorted offsets/lengths the sort wrote
        ds.partition("IndexedNumStrings", max_len=2 * 100 * (ds.avg_lcp() + 5))
        rb, bb, nbytes = ds.exchange("whole")
        return ds.get_exchange_bytes(nbytes).tobytes()

    with dss.DeviceStringSet.from_arena(chars, p) as ds:
        ds.set_sort_params({"pass_mode": 1})
        ds.sort_lcp()
        hperm, hlcp = ds.get_sorted()
        hbytes = exchanged(ds)
        ds.set_sort_params({"pass_mode": 2})
        for rep in range(3):
            ds.sort_lcp()
            perm, lcp = ds.get_sorted()
            assert (perm == hperm).all() and (lcp == hlcp).all(), rep
            assert ds.avg_lcp() == O.avg_lcp(olcp)
            assert exchanged(ds) == hbytes, rep
            ds.sync()
    assert (hlcp == olcp).all()
    assert (canonical_ties(hperm, hlcp, ss.len, ss.pe_begin) == canonical_ties(operm, olcp, ss.len, ss.pe_begin)).all()


def test_device_driven_replays_across_rebatched_jobs():
    """One capture of the device-driven sort serves batches of other sizes:
    a batch set reserved once is re-pointed at different jobs (other string
    and PE counts) and each replay equals the host-driven sort of the same
    jobs and the oracle."""
    dss = _dss()
    from dss_amd import JobPool
    rng = np.random.default_rng(45)
    arenas = [O.gen_dton(int(rng.integers(500, 20000)), int(rng.integers(20, 200)), float(rng.uniform(0.2, 0.8)),
                         int(100 + k)) for k in range(12)]
    groups = [[0, 1, 2], [3], [4, 5, 6, 7, 8], [9, 10, 11], [2, 7]]
    with JobPool(1 << 28, 1 << 22) as pool:
        ids = [pool.add(a) for a in arenas]
        b = pool.batch([ids[k] for k in groups[0]])
        b.set.reserve(sum(int((a == 0).sum()) for a in arenas), 16)
        for g in groups:
            b.rebatch([ids[k] for k in g])
            b.set.set_sort_params({"pass_mode": 1})
            b.sort()
            host = [b.result(k) for k in range(len(g))]
            hd = b.dist_sizes()
            b.set.set_sort_params({"pass_mode": 2})
            b.sort()
            for k, jk in enumerate(g):
                perm, lcp = b.result(k)
                assert (perm == host[k][0]).all() and (lcp == host[k][1]).all(), (g, k)
                oss = O.StringSet(arenas[jk], 1)
                operm, olcp = O.local_sort(oss)
                assert (lcp == olcp).all()
                assert (canonical_ties(perm, lcp, oss.len, oss.pe_begin) ==
                        canonical_ties(operm, olcp, oss.len, oss.pe_begin)).all()
            assert (b.dist_sizes() == hd).all()
        b.close()
