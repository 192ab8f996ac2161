"""GPU parity of the wire codecs (dss_varint_* / dss_golomb_*, through the C
ABI) against the oracle and the reference's own vectors: byte- and
word-exact encodings, exact round trips, the reference's known answers, and
the reference's error behaviour on malformed input."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "codec_vectors.json")))


@pytest.fixture(scope="module")
def codec():
    from dss_amd import WireCodec
    with WireCodec() as c:
        yield c


def test_varint_reference_known_answer(codec):
    nums = [1, 1 << 8, 1 << 16, 1 << 24, 1 << 32, 1 << 40, 1 << 48, 1 << 56]
    vals = np.array(nums * 2, dtype=np.uint64)
    data, rb = codec.write_ranges(vals, [8, 8])
    assert list(rb) == [37, 37]
    assert np.array_equal(codec.read_ranges(data, [8, 8]), vals)


def test_varint_matches_reference_fixture(codec):
    for case in GOLDEN["varint_ranges"]:
        vals = np.array([int(x) for x in case["values"]], dtype=np.uint64)
        data, rb = codec.write_ranges(vals, case["counts"])
        assert data.tobytes().hex() == case["bytes_hex"]
        assert [int(x) for x in rb] == case["range_bytes"]
        assert np.array_equal(codec.read_ranges(data, case["counts"]), vals)


@pytest.mark.parametrize("n,p,maxbits", [(0, 3, 8), (1, 1, 64), (5000, 16, 64), (3_000_000, 16, 12)])
def test_varint_matches_oracle(codec, n, p, maxbits):
    rng = np.random.default_rng(n + p)
    bits = rng.integers(0, maxbits + 1, size=n)
    vals = (rng.integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)) >> \
        (np.uint64(64) - bits.astype(np.uint64))
    vals[bits == 0] = 0
    cuts = np.sort(rng.integers(0, n + 1, size=p - 1))
    counts = np.diff(np.concatenate([[0], cuts, [n]])).astype(np.uint64)
    data, rb = codec.write_ranges(vals, counts)
    odata, orb = O.varint_write_ranges(vals, counts)
    assert np.array_equal(data, odata) and np.array_equal(rb, orb)
    assert np.array_equal(codec.read_ranges(data, counts), vals)


def test_varint_rejects_malformed_input(codec):
    from dss_amd import DssError
    data, _ = codec.write_ranges(np.array([1 << 40, 5], np.uint64), [2])
    with pytest.raises(DssError):
        codec.read_ranges(data[:-1], [2])  # ends inside a value / too few values
    with pytest.raises(DssError):
        codec.read_ranges(data, [3])
    with pytest.raises(DssError):  # 10th byte beyond the 64th bit (overflow_error)
        codec.read_ranges(np.array([0xFF] * 9 + [0x02], np.uint8), [1])


def test_golomb_matches_reference_fixture(codec):
    """Word-exact getDeltaEncoding vectors of the reference headers, one
    range per call, in the pipelined framing ([size][encoding])."""
    for case in GOLDEN["golomb"]:
        b = int(case["b"])
        vals = np.array([int(x) for x in case["values"]], dtype=np.uint64)
        words, rw = codec.golomb_encode(vals, [vals.size], [b], layout=1)
        ref = [int(x, 16) for x in case["words"]]
        assert int(words[0]) == len(ref) and [int(x) for x in words[1:]] == ref, b
        dec, cnt = codec.golomb_decode(words, 1, vals.size, layout=1, b_pipelined=b)
        assert list(cnt) == [vals.size] and np.array_equal(dec, vals)


def test_golomb_pipeline_reference_round_trip(codec):
    """pipelinedGolombEncoding_tests.cpp: 0..999 to every partner, b = 2^20."""
    p = 8
    sample = np.arange(1000, dtype=np.uint64)
    words, rw = codec.golomb_encode(np.tile(sample, p), [1000] * p, [1048576] * p, layout=1)
    owords, orw = O.golomb_pipeline_blocks(np.tile(sample, p), [1000] * p)
    assert np.array_equal(words, owords) and np.array_equal(rw, orw)
    dec, cnt = codec.golomb_decode(words, p, 1000 * p, layout=1, b_pipelined=1048576)
    assert list(cnt) == [1000] * p
    assert np.array_equal(dec, np.tile(sample, p))


def _hash_intervals(rng, p, n_per, m=2**64 - 1, empty=()):
    vals, counts = [], []
    for q in range(p):
        c = 0 if q in empty else int(rng.integers(max(0, n_per // 2), n_per + 1))
        lo, hi = q * (m // p), (q + 1) * (m // p) - 1
        vals.append(np.sort(rng.integers(lo, hi, size=c, dtype=np.uint64)))
        counts.append(c)
    return np.concatenate(vals), np.array(counts, np.uint64)


@pytest.mark.parametrize("p,n_per", [(1, 1), (4, 50), (16, 2000), (16, 1_250_000)])
def test_golomb_hash_blocks_match_oracle(codec, p, n_per):
    """AllToAllHashesGolomb on sorted hash intervals of one PE (b =
    getB(m/p, size), 1 -> 2e10), word-exact against the oracle's framing,
    and decoded back exactly; the last shape is the 20M x 500 B input's
    first round (1.25M hashes per PE, 16 PEs)."""
    from dss_amd.codec import hash_interval_b
    rng = np.random.default_rng(p * 7 + n_per)
    vals, counts = _hash_intervals(rng, p, n_per, empty=(2,) if p > 2 else ())
    m = 2**64 - 1
    bs = [hash_interval_b(m, p, int(c)) for c in counts]
    assert bs == [O.golomb_hash_b(m, p, int(c)) for c in counts]
    words, rw = codec.send_hashes(vals, counts, m)
    owords, orw = O.golomb_blocks(vals, counts, bs)
    assert np.array_equal(rw, orw) and np.array_equal(words, owords)
    dec, dcounts = codec.golomb_decode(words, p, vals.size)
    assert np.array_equal(dec, vals)
    assert list(dcounts) == [int(c) for c in counts if c]


def test_golomb_index_blocks_and_long_unary_runs(codec):
    """Duplicate-index intervals (b = max(8, getB(range, size))) and streams
    whose codes have unary runs longer than a word and than a decode chunk."""
    rng = np.random.default_rng(3)
    for b, gap in [(8, 5000), (3, 100), (5, 2000), (16, 70000), (1000, 10**6)]:
        n = 3000
        vals = np.cumsum(rng.integers(0, 2 * gap, size=n)).astype(np.uint64)
        counts = [n // 3, 0, n - n // 3]
        words, rw = codec.golomb_encode(vals, counts, [b] * 3)
        owords, orw = O.golomb_blocks(vals, counts, [b] * 3)
        assert np.array_equal(words, owords), b
        dec, dc = codec.golomb_decode(words, 3, n)
        assert np.array_equal(dec, vals) and list(dc) == [n // 3, n - n // 3]
    from dss_amd.codec import index_interval_b
    idx = np.sort(rng.choice(10**6, size=5000, replace=False)).astype(np.uint64)
    counts, rng_max = [2000, 3000], [10**6, 10**6]
    words, _ = codec.send_indices(idx, counts, rng_max)
    owords, _ = O.golomb_blocks(idx, counts, [index_interval_b(r, c) for r, c in zip(rng_max, counts)])
    assert np.array_equal(words, owords)


def test_golomb_rejects_malformed_input(codec):
    from dss_amd import DssError
    vals = np.arange(100, dtype=np.uint64) * np.uint64(1000)
    words, _ = codec.golomb_encode(vals, [100], [1024])
    with pytest.raises(DssError):  # block size beyond the input
        codec.golomb_decode(words[:-1], 1, 100)
    with pytest.raises(DssError):  # more blocks than allowed
        codec.golomb_decode(np.concatenate([words, words]), 1, 200)
    with pytest.raises(DssError):  # capacity too small
        codec.golomb_decode(words, 1, 50)
