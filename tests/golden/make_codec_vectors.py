"""Writes tests/golden/codec_vectors.json from oracle/_ref (the reference's
encoding headers compiled in place: IntegerCompression::writeRanges, getB,
getDeltaEncoding over the delta and Golomb bit streams).  Run in a checkout
next to the reference:

    bash oracle/build_ref.sh && python tests/golden/make_codec_vectors.py

The fixture holds inputs and the reference's outputs only (no reference code).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def main():
    R = O.ref()
    if R is None:
        sys.exit("oracle/_ref/libdss_ref.so is not built (bash oracle/build_ref.sh)")
    rng = np.random.default_rng(0x5EEDC0DE)
    out = {"source": "reference headers compiled by oracle/build_ref.sh"}

    # varint ranges: values of every byte length 1..10, several ranges
    vcases = []
    for t in range(6):
        nr = int(rng.integers(1, 6))
        counts = rng.integers(0, 24, size=nr).astype(np.uint64)
        bits = rng.integers(0, 65, size=int(counts.sum()))
        vals = np.array([(int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))) >> (64 - int(b))
                         if b else 0 for b in bits], dtype=np.uint64)
        buf = np.zeros(max(1, vals.size * 10), np.uint8)
        rb = np.zeros(nr, np.uint64)
        n = R.ref_varint_write_ranges(O.P(vals if vals.size else np.zeros(1, np.uint64), O.u64p),
                                      O.P(counts, O.u64p), nr, O.P(buf, O.u8p), O.P(rb, O.u64p))
        vcases.append({"values": [str(int(x)) for x in vals], "counts": [int(x) for x in counts],
                       "bytes_hex": buf[:n].tobytes().hex(), "range_bytes": [int(x) for x in rb]})
    out["varint_ranges"] = vcases

    # getB
    pairs = [(0, 0), (5, 10), (2**64 - 1, 1), (2**64 - 1, 3), (2**63, 1), (2**53 + 1, 1),
             (2**60 - 1, 1), (10**6, 7), (1152921504606846975, 1250000), (123456789, 1000)]
    out["golomb_b"] = [{"max": str(a), "count": b, "b": str(R.ref_golomb_b(a, b))} for a, b in pairs]

    # delta + Golomb encodings of ascending values
    gcases = []
    for b, n, span in [(8, 50, 4000), (3, 40, 2000), (5, 64, 300), (1000, 120, 10**6), (12345, 90, 10**7),
                       (1048576, 200, 2**28), (20000000000, 100, 2**40), (2**56, 150, 2**63),
                       (16, 30, 30 * 16 * 150)]:
        vals = np.sort(rng.integers(0, span, size=n, dtype=np.uint64))
        enc = np.zeros(8 + (span // b + n * 65) // 64, np.uint64)
        k = R.ref_golomb_encode(O.P(vals, O.u64p), n, b, O.P(enc, O.u64p))
        gcases.append({"b": str(b), "values": [str(int(x)) for x in vals],
                       "words": ["%016x" % int(x) for x in enc[:k]]})
    # the pipelined exchange test's message: 0..999 with b = 1048576
    vals = np.arange(1000, dtype=np.uint64)
    enc = np.zeros(2000, np.uint64)
    k = R.ref_golomb_encode(O.P(vals, O.u64p), 1000, 1048576, O.P(enc, O.u64p))
    gcases.append({"b": "1048576", "values": [str(int(x)) for x in vals],
                   "words": ["%016x" % int(x) for x in enc[:k]]})
    out["golomb"] = gcases
    path = os.path.join(HERE, "codec_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(f"wrote {path}: {len(vcases)} varint cases, {len(gcases)} Golomb cases")


if __name__ == "__main__":
    main()
