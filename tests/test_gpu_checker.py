"""The device checker (dss_check: is_sorted / is_complete_and_sorted,
src/mpi/is_sorted.hpp:27-198) on the reference's checker cases
(src/tests/isSorted_tests.cpp:12-143) and on outputs with a planted fault
(out of order, wrong LCP, a string missing and another repeated): the
checker must report each failure, not only pass correct outputs.

The per-PE data of a case is put through the device path unchanged: every
PE keeps its strings in the given order (caller-set local order and LCPs),
sends them to itself (diagonal send counts), and the whole-string exchange
and the LCP merge of one run per PE leave that order; dss_check then sees
the concatenation of the PEs' data as the global output, as the reference's
checker sees the ranks' data.  The verdict is compared with the oracle's
checker on the same data."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _dss():
    import dss_amd
    return dss_amd


def arena(strings):
    return np.frombuffer(b"".join(s + b"\0" for s in strings), dtype=np.uint8).copy()


def _lcp(a, b):
    k = 0
    while k < min(len(a), len(b)) and a[k] == b[k]:
        k += 1
    return k


def _local_lcps(pes):
    out = []
    for strs in pes:
        if strs:  # an empty PE has no positions
            out += [0] + [_lcp(strs[i - 1], strs[i]) for i in range(1, len(strs))]
    return np.array(out, dtype=np.uint32)


def _oracle_sorted(pes):
    flat = [s for strs in pes for s in strs]
    a = arena(flat) if flat else np.zeros(1, np.uint8)
    ptrs, _ = O._arena_pointers(a)
    pb = np.zeros(len(pes) + 1, dtype=np.uint64)
    pb[1:] = np.cumsum([len(s) for s in pes])
    return bool(O.lib().orc_is_sorted(ptrs, pb.ctypes.data_as(O.u64p), len(pes)))


def _device_check(pes, lcp_fault=None, order_fault=None):
    """dss_check of the PEs' data kept in the given order; lcp_fault(lcp)
    edits the local LCPs, order_fault(order, lcp) replaces the merged output
    through dss_set_merged."""
    dss = _dss()
    with dss.DeviceStringSet.from_pes([arena(s) for s in pes]) as ds:
        n = ds.n
        assert n == sum(len(s) for s in pes)
        pb = ds.pe_begin()
        lcp = _local_lcps(pes)
        assert lcp.size == n
        if lcp_fault:
            lcp_fault(lcp)
        ds.set_local_order(np.arange(n, dtype=np.uint32), lcp)
        ds.set_interval(np.diag(np.diff(pb).astype(np.uint64)))
        ds.exchange("whole")
        ds.merge_lcp()
        order, mlcp, _ = ds.get_merged()
        assert (order == np.arange(n)).all()  # one run per PE: the merge keeps the order
        if order_fault:
            order, mlcp = order.copy(), mlcp.copy()
            order_fault(order, mlcp)
            ds.set_merged(order, mlcp)
        return ds.check()


def _pe_data(p, rank_char, empty=()):
    return [[] if r in empty else [bytes([rank_char(r), 97, 97]), bytes([rank_char(r), 98, 98])] for r in range(p)]


P_VALUES = [4, 8]


@pytest.mark.parametrize("p", P_VALUES)
def test_all_pes_sorted(p):
    """test_AllPEHaveSortedData: one string per PE, first char rank + 1."""
    pes = [[bytes([r + 1, 97, 97])] for r in range(p)]
    chk = _device_check(pes)
    assert chk["sorted"] and chk["complete"] and chk["order_errors"] == 0 and chk["lcp_errors"] == 0
    assert _oracle_sorted(pes)


@pytest.mark.parametrize("p", P_VALUES)
@pytest.mark.parametrize("which", ["first", "middle", "last"])
def test_all_but_one_pe_sorted(p, which):
    """test_AllButOnePEHaveSortedData_1/2/3: one PE (0, p/2, p-1) empty."""
    e = {"first": 0, "middle": p // 2, "last": p - 1}[which]
    pes = _pe_data(p, lambda r: r + 65, empty=(e,))
    chk = _device_check(pes)
    assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0
    assert chk["num_strings"] == 2 * (p - 1)
    assert _oracle_sorted(pes)


@pytest.mark.parametrize("p", P_VALUES)
def test_all_but_n_pes_sorted(p):
    """test_AllButnPEHaveSortedData(n) for n = 0..p-1: n random PEs empty."""
    rng = np.random.default_rng(p)
    for k in range(p):
        empty = set(int(x) for x in rng.permutation(p)[:k])
        pes = _pe_data(p, lambda r: r + 65, empty=empty)
        chk = _device_check(pes)
        assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0, (k, empty)
        assert _oracle_sorted(pes)


@pytest.mark.parametrize("p", P_VALUES)
def test_one_pe_locally_unsorted(p):
    """test_onePELocallyUnsorted: PE 0 holds {'Aba', 'Aaa'}."""
    pes = _pe_data(p, lambda r: r + 65)
    pes[0] = [b"Aba", b"Aaa"]
    chk = _device_check(pes)
    assert not chk["sorted"] and chk["order_errors"] == 1
    assert chk["complete"]
    assert not _oracle_sorted(pes)


@pytest.mark.parametrize("p", P_VALUES)
def test_n_pes_globally_unsorted(p):
    """test_nPEGloballyUnsorted(n) for n = 1..p-1: n random PEs take first
    char p - rank + 65, so their data breaks the order across PEs while
    every PE is locally sorted."""
    rng = np.random.default_rng(100 + p)
    for k in range(1, p):
        moved = set(int(x) for x in rng.permutation(p)[:k])
        pes = _pe_data(p, lambda r: (p - r + 65) if r in moved else (r + 65))
        exp_sorted = _oracle_sorted(pes)
        chk = _device_check(pes)
        assert chk["sorted"] == exp_sorted, (k, moved)
        assert chk["complete"] and chk["lcp_errors"] == 0
        # the reference's test asserts "not sorted" for every n >= 1; a moved
        # PE can land in order by chance only if its char equals its rank's
        if not exp_sorted:
            assert chk["order_errors"] > 0


def test_corrupted_lcp_is_reported():
    """A merged LCP below the true one (the LCP array is part of the
    output: the exhaustive check compares it)."""
    pes = [[b"Xaaa", b"Xaab", b"Xabb"], [b"Ybbb", b"Ybbc"]]

    def fault(lcp):
        lcp[1] = 1  # true LCP('Xaaa', 'Xaab') = 3
        lcp[4] = 0  # true LCP('Ybbb', 'Ybbc') = 3

    chk = _device_check(pes, lcp_fault=fault)
    assert chk["sorted"] and chk["complete"]
    assert chk["lcp_errors"] == 2
    assert _device_check(pes)["lcp_errors"] == 0


def test_repeated_and_missing_string_is_incomplete():
    """A merged output that repeats one received string in place of another
    (still sorted, LCPs consistent): is_complete_and_sorted fails on the
    multiset, here the origin positions."""
    pes = [[b"Kaa", b"Kab", b"Kac"], [b"Lxa", b"Lxb"]]

    def fault(order, mlcp):
        order[2] = order[1]  # 'Kab' twice, 'Kac' missing
        mlcp[2] = 3          # LCP('Kab', 'Kab')
        mlcp[3] = 0          # LCP('Kab', 'Lxa'), a PE start anyway

    chk = _device_check(pes, order_fault=fault)
    assert chk["sorted"] and chk["lcp_errors"] == 0
    assert not chk["complete"]


def test_swapped_output_is_unsorted():
    """Two merged outputs swapped (complete, but out of order) with the
    LCPs left as they were: both the order and the LCP check fail."""
    pes = [[b"Maa", b"Mbb", b"Mcc", b"Mdd"], [b"Nxx"]]

    def fault(order, mlcp):
        order[1], order[2] = order[2], order[1]
        mlcp[1] = 1

    chk = _device_check(pes, order_fault=fault)
    assert chk["complete"]
    assert not chk["sorted"] and chk["order_errors"] == 1
    assert chk["lcp_errors"] == 0  # LCP('Maa','Mcc') = LCP('Mcc','Mbb') = 1


def test_set_merged_rejects_out_of_range_entries():
    dss = _dss()
    pes = [[b"a", b"b"], [b"c"]]
    with dss.DeviceStringSet.from_pes([arena(s) for s in pes]) as ds:
        ds.set_local_order(np.arange(3, dtype=np.uint32), np.zeros(3, np.uint32))
        ds.set_interval(np.diag([2, 1]).astype(np.uint64))
        ds.exchange("whole")
        with pytest.raises(dss.DssError):
            ds.set_merged(np.array([0, 1, 3], np.uint32), np.zeros(3, np.uint32))
        ds.set_merged(np.array([0, 1, 2], np.uint32), np.zeros(3, np.uint32))
        assert ds.check()["sorted"]
        with pytest.raises(dss.DssError):
            ds.extend_prefix()
