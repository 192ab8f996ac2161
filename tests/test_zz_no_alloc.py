"""The allocation-free varint encoder (dss_varint_write_ranges_noalloc)
against the allocating one (dss_varint_write_ranges_dev) through the C ABI:
identical bytes, range sizes and totals on two calls in a row (the second
smaller, so stale scratch from the first cannot leak into it), both equal to
the oracle; sizes beyond the limits are rejected before any launch; repeated
calls leave the device's free memory unchanged.  Device buffers come from
the HIP runtime the library loaded (tests/helpers.py)."""
import numpy as np
import pytest

import oracle as O
from helpers import HipMem

pytestmark = pytest.mark.gpu


def _values(rng, n):
    """u64 values of every varint length 1..10 bytes."""
    v = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)
    return v >> rng.integers(0, 64, n).astype(np.uint64)


def _counts(rng, n, nr):
    cuts = np.sort(rng.integers(0, n + 1, nr - 1))
    return np.diff(np.concatenate([[0], cuts, [n]])).astype(np.uint64)


def test_noalloc_matches_allocating_encoder():
    from dss_amd import DssError, WireCodec

    rng = np.random.default_rng(11)
    with WireCodec() as c, HipMem() as hm:
        for n, nr in ((1_500_000, 37), (20_000, 5)):  # the second call is smaller
            vals = _values(rng, n)
            counts = _counts(rng, n, nr)
            print(f"[no-alloc] n={n} ranges={nr}", flush=True)
            dv = hm.upload(vals)
            cap = n * 10 + 8
            out_a = hm.alloc(cap)
            out_b = hm.alloc(cap)
            ta, rba = c.write_ranges_dev(dv, counts, out_a, cap)
            tb, rbb = c.write_ranges_noalloc(dv, counts, out_b, cap)
            c.sync()
            assert ta == tb
            assert (rba == rbb).all()
            a = hm.download(out_a, cap)
            b = hm.download(out_b, cap)
            assert a[:ta].tobytes() == b[:tb].tobytes()
            assert not b[tb:].any()  # nothing written beyond the total
            exp, exp_rb = O.varint_write_ranges(vals, counts)
            assert b[:tb].tobytes() == exp.tobytes()
            assert (rbb == exp_rb).all()
        # limits: rejected on the host, before any launch
        small = hm.alloc(64)
        out = hm.alloc(64)
        with pytest.raises(DssError):
            c.write_ranges_noalloc(small, [(1 << 21) + 1], out, 64)
        with pytest.raises(DssError):
            c.write_ranges_noalloc(small, [0] * 4097, out, 64)
        # an output buffer that is too small: the size is reported, nothing written
        with pytest.raises(DssError):
            c.write_ranges_noalloc(small, [8], out, 4)
        c.sync()
        assert not hm.download(out, 64).any()


def test_noalloc_leaves_free_memory_unchanged():
    from dss_amd import WireCodec

    rng = np.random.default_rng(12)
    n = 200_000
    vals = _values(rng, n)
    counts = _counts(rng, n, 16)
    with WireCodec() as c, HipMem() as hm:
        dv = hm.upload(vals)
        out = hm.alloc(n * 10)
        t0, _ = c.write_ranges_noalloc(dv, counts, out, n * 10)  # first call on the device
        c.sync()
        free0 = hm.free_bytes()
        for _ in range(3):
            t, _ = c.write_ranges_noalloc(dv, counts, out, n * 10)
            assert t == t0
        c.sync()
        assert hm.free_bytes() == free0
