"""Writes tests/golden/ref_vectors.json from oracle/_ref (the reference's own
headers compiled in place: xxh::hash_state_t<64>, LcpStringLoserTree_,
StringLcpContainer::extendPrefix).  Run in a checkout next to the reference:

    bash oracle/build_ref.sh && python tests/golden/make_ref_vectors.py

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
    out = {"source": "reference headers compiled by oracle/build_ref.sh"}
    # XXH64 streaming digests of a fixed byte pattern, lengths 0..300
    pattern = np.array([(i * 131 + 7) % 251 + 1 for i in range(300)], dtype=np.uint8)
    hashes = []
    for L in range(0, 301):
        buf = pattern[:L].copy() if L else np.zeros(1, np.uint8)
        hashes.append("%016x" % R.ref_xxh_stream(buf.ctypes.data_as(O.u8p), L))
    out["xxh64_stream"] = {"pattern": "byte i = (i*131+7) % 251 + 1", "digests": hashes}

    # loser-tree merges and prefix decompression on exchanged inputs
    cases = []
    specs = [("dton", O.gen_dton(600, 40, 0.5, 11), 4),
             ("skewed", O.gen_skewed(O.skew_params(800, long_len=60, short_len=8), 12), 3),
             ("suffix", O.gen_suffixes(700, 13, repeats=4, rep_min=20, rep_max=90, cap=50), 5)]
    for name, chars, p in specs:
        ss = O.StringSet(chars, p)
        perm, lcp = O.local_sort(ss)
        iv = O.partition(ss, perm, "IndexedNumStrings", max_len=1000)
        for mode in ("whole", "lcp_compact"):
            ex = O.exchange(ss, perm, lcp, iv, mode)
            for q in range(p):
                r0, r1 = int(ex["recv_begin"][q]), int(ex["recv_begin"][q + 1])
                b0, b1 = int(ex["byte_begin"][q]), int(ex["byte_begin"][q + 1])
                n = r1 - r0
                if n == 0:
                    continue
                a = np.ascontiguousarray(ex["bytes"][b0:b1])
                rl = ex["lcp"][r0:r1].astype(np.uint64)
                runs = iv[:, q].astype(np.uint64).copy()
                oi = np.zeros(n, np.uint64)
                ol = np.zeros(n, np.uint64)
                osv = np.zeros(n, np.uint64)
                comp = int(mode == "lcp_compact")
                assert R.ref_merge(a.ctypes.data_as(O.u8p), a.size, rl.ctypes.data_as(O.u64p),
                                   runs.ctypes.data_as(O.u64p), p, comp, oi.ctypes.data_as(O.u64p),
                                   ol.ctypes.data_as(O.u64p), osv.ctypes.data_as(O.u64p)) == 0
                case = {"input": name, "pes": p, "mode": mode, "receiver": q,
                        "arena_hex": a.tobytes().hex(), "recv_lcp": [int(x) for x in rl],
                        "runs": [int(x) for x in runs], "order": [int(x) for x in oi],
                        "merged_lcp": [int(x) for x in ol]}
                if comp:
                    case["saved_lcp"] = [int(x) for x in osv]
                    size = R.ref_extend_prefix(a.ctypes.data_as(O.u8p), a.size, oi.ctypes.data_as(O.u64p),
                                               osv.ctypes.data_as(O.u64p), n, None)
                    buf = np.zeros(size, np.uint8)
                    R.ref_extend_prefix(a.ctypes.data_as(O.u8p), a.size, oi.ctypes.data_as(O.u64p),
                                        osv.ctypes.data_as(O.u64p), n, buf.ctypes.data_as(O.u8p))
                    case["extended_hex"] = buf.tobytes().hex()
                cases.append(case)
    out["merges"] = cases
    path = os.path.join(HERE, "ref_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(f"wrote {path}: {len(hashes)} digests, {len(cases)} merges")


if __name__ == "__main__":
    main()
