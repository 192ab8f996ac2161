"""GPU tests at the workload sizes of the project's configurations, through
size-independent properties (global order, completeness, exact LCPs, D), and
the reference's exchange known-answer layouts through the C ABI."""
import numpy as np
import pytest

import oracle as O
from test_oracle import kat_setup

pytestmark = pytest.mark.gpu


def _dss():
    import dss_amd
    return dss_amd


def test_exchange_reference_vectors_on_device():
    """alltoallStrings_tests.cpp known answers, received by the device
    bucket scatter (caller-given local order, LCPs and send counts)."""
    dss = _dss()
    size, common, suffix = 5, 10, 20
    for p in (1, 2, 4, 8):
        ss, perm, lcp, iv = kat_setup(p, size, common, suffix)
        pes = []
        for r in range(p):
            b, e = int(ss.pe_begin[r]), int(ss.pe_begin[r + 1])
            pes.append(np.frombuffer(b"".join(ss.string(i) + b"\0" for i in range(b, e)), np.uint8).copy())
        for mode in ("whole", "lcp_compact", "pd_prefix"):
            with dss.DeviceStringSet.from_pes(pes) as ds:
                ds.set_local_order(perm, lcp)
                if mode == "pd_prefix":
                    ds.set_dist(np.full(ss.n, common + 1, np.uint32))
                ds.set_interval(iv)
                rb, bb, nb = ds.exchange(mode)
                got = ds.get_exchange_bytes(nb)
                glcp, gpe, gpos = ds.get_exchange()
            dist = np.full(ss.n, common + 1, np.uint32) if mode == "pd_prefix" else None
            ex = O.exchange(ss, perm, lcp, iv, mode, dist=dist)
            assert got.tobytes() == ex["bytes"].tobytes(), (p, mode)
            assert (glcp == ex["lcp"]).all() and (gpe == ex["src_pe"]).all() and (gpos == ex["src_pos"]).all()


def check_ok(chk):
    assert chk["sorted"] and chk["complete"] and chk["order_errors"] == 0 and chk["lcp_errors"] == 0, chk


def test_c2_dton_20m_prefix_doubling():
    """20M x 500 chars, D/N 0.5, 16 PEs, prefix doubling."""
    dss = _dss()
    n = 20_000_000
    with dss.DeviceStringSet.dton(n, 500, 0.5, 0x5EED0002, 16) as ds:
        ds.prefix_doubling_sort()
        chk = ds.check()
        check_ok(chk)
        # every string is distinguished by its 250-char digit block: D = 250 n
        assert chk["D"] == 250 * n
        res, rounds = ds.get_prefix_doubling()
        assert (res == 264).all() and (rounds == 6).all()


def test_c2_dton_20m_bloom_8_bits_per_string():
    """The same input with the hash range sized 8 bits per string (m = 8 n,
    the Bloom-filter-sized variant of the configuration; the reference uses
    2^64 - 1, bloomfilter.hpp:1135-1136): collisions keep strings candidates
    longer, so results may exceed the exact distinguishing prefixes but never
    fall below them; the sorted output is exact."""
    dss = _dss()
    n = 20_000_000
    with dss.DeviceStringSet.dton(n, 500, 0.5, 0x5EED0002, 16) as ds:
        ds.prefix_doubling_sort(m=8 * n)
        chk = ds.check()
        check_ok(chk)
        assert chk["D"] == 250 * n
        dc = ds.check_dist()
        assert dc["lower_violations"] == 0
        assert dc["exact_sum"] == chk["D"]
        res, rounds = ds.get_prefix_doubling()
        assert (res >= 250).all()


def test_c1_dton_1m_merge_sort_compressed():
    """1M x 500 chars, D/N 0.5, 4 PEs, merge sort with LCP compression."""
    dss = _dss()
    n = 1_000_000
    with dss.DeviceStringSet.dton(n, 500, 0.5, 0x5EED0001, 4) as ds:
        ds.merge_sort(compressed=True)
        chk = ds.check()
        check_ok(chk)
        assert abs(chk["D"] / (n * 500) - 0.5) < 0.01
        out, offs = ds.get_extended(n * 501)
        assert (out[offs.astype(np.int64) + 500] == 0).all()  # every rebuilt string has 500 chars


@pytest.mark.parametrize("driver", ["merge_sort", "prefix_doubling_sort"])
def test_c3_skewed_20m(driver):
    """20M strings: 20% of 1000 chars (D/N .75), 80% of 100 (D/N .25), 10%
    exact duplicates, bytes 1..255, 16 PEs."""
    dss = _dss()
    with dss.DeviceStringSet.skewed(20_000_000, 0x5EED0003, 16) as ds:
        getattr(ds, driver)()
        check_ok(ds.check())


def test_c4_suffixes_parity_slice():
    """Capped suffixes of an order-3 Markov ACGT text with planted repeats:
    parity with the oracle on a 300 KB slice, 16 PEs."""
    dss = _dss()
    chars = O.gen_suffixes(300_000, 0x5EED0004, repeats=50, rep_min=1000, rep_max=10000, cap=500)
    ss = O.StringSet(chars, 16)
    res = O.merge_sort(ss, encoder="lcp_compact")
    exp = b"".join(o["arena"].tobytes() for o in res["out"])
    with dss.DeviceStringSet.from_arena(chars, 16) as ds:
        ds.merge_sort(compressed=True)
        check_ok(ds.check())
        got, _ = ds.get_extended(len(exp))
    assert got.tobytes() == exp


def test_dton_arena_beyond_32gib():
    """40M x 900 chars (36 GB arena: more 8-byte words than one dispatch has
    work items), D/N 0.5: generated, sorted and checked as a set and as a
    pool job.  Each string is distinguished by its 450-char digit block, so
    D = 450 n exactly."""
    dss = _dss()
    n, length, k = 40_000_000, 900, 450
    with dss.DeviceStringSet.dton(n, length, 0.5, 0x5EED0006, 16) as ds:
        ds.prefix_doubling_sort()
        chk = ds.check()
        check_ok(chk)
        assert chk["D"] == k * n
    with dss.JobPool(n * (length + 1) + 64, n) as pool:
        j = pool.add_dton(n, length, 0.5, 0x5EED0006)
        b = pool.batch([j])
        b.sort()
        assert int(b.dist_sizes()[0]) == k * n
        b.close()
        pool.free(j)


def test_c4_suffix_views_match_materialised_slice():
    """The device Markov generator and suffix views reproduce the oracle's
    text and materialised capped suffixes: same PE split, same merge-sort
    output bytes (300 KB slice, 16 PEs)."""
    dss = _dss()
    n, seed = 300_000, 0x5EED0004
    text = O.gen_markov_text(n, seed, repeats=50, rep_min=1000, rep_max=10000)
    chars = O.gen_suffixes(n, seed, repeats=50, rep_min=1000, rep_max=10000, cap=500)
    ss = O.StringSet(chars, 16)
    res = O.merge_sort(ss, encoder="lcp_compact")
    exp = b"".join(o["arena"].tobytes() for o in res["out"])
    for make in (lambda: dss.DeviceStringSet.markov_suffixes(n, seed, 16, repeats=50, rep_min=1000,
                                                             rep_max=10000, cap=500),
                 lambda: dss.DeviceStringSet.suffixes(text, 500, 16)):
        with make() as ds:
            dtext, off, ln = ds.download()
            assert dtext.tobytes() == text.tobytes()
            assert (ds.pe_begin() == ss.pe_begin).all()
            assert (ln == ss.len).all()
            ds.merge_sort(compressed=True)
            check_ok(ds.check())
            got, _ = ds.get_extended(len(exp))
        assert got.tobytes() == exp


def test_c4_suffixes_full_scale():
    """Capped (500) suffixes of a 200 MB order-3 Markov ACGT text with 5,000
    planted repeats of 1-10 KB: 200M suffix views over 16 PEs through the
    prefix-doubling sort, checked for global order, completeness and exact
    LCPs; the text is the oracle generator's."""
    dss = _dss()
    n, seed = 200_000_000, 0x5EED0004
    with dss.DeviceStringSet.markov_suffixes(n, seed, 16, repeats=5000, rep_min=1000, rep_max=10000,
                                             cap=500) as ds:
        ds.prefix_doubling_sort()
        chk = ds.check()
        check_ok(chk)
        assert chk["num_strings"] == n
        dtext, _, _ = ds.download()
    text = O.gen_markov_text(n, seed, repeats=5000, rep_min=1000, rep_max=10000)
    assert dtext.tobytes() == text.tobytes()
