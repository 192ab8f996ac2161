"""Shared helpers of the test suite (tie canonicalisation, comparisons)."""
import numpy as np


def canonical_ties(perm, lcp, lengths, pe_begin):
    """Sort every run of equal strings by string id.

    Tie order among equal strings is implementation-defined in the reference
    (tlx's in-place radix permutation is unstable), so both sides are
    compared after this canonicalisation.  A run of equal strings is a
    maximal range of positions i with lcp[i] == len(s[i]) == len(s[i-1])."""
    perm = np.array(perm, dtype=np.uint32, copy=True)
    L = lengths[perm.astype(np.int64)]
    eq = np.zeros(perm.size, dtype=bool)
    if perm.size > 1:
        eq[1:] = (lcp[1:] == L[1:]) & (L[1:] == L[:-1])
    for r in range(len(pe_begin) - 1):
        b = int(pe_begin[r])
        if b < perm.size:
            eq[b] = False
    idx = np.flatnonzero(eq)
    if idx.size == 0:
        return perm
    # run starts: positions i-1 where eq[i] and not eq[i-1]
    i = 0
    while i < idx.size:
        s = idx[i] - 1
        j = i
        while j + 1 < idx.size and idx[j + 1] == idx[j] + 1:
            j += 1
        e = idx[j] + 1
        perm[s:e] = np.sort(perm[s:e])
        i = j + 1
    return perm


def file_distributer(text: bytes, p: int):
    """distribute_strings (src/mpi/readInput.hpp:40-104) done the way the
    reference does it: byte slices, '\\n' -> NUL, every PE but the first hands
    its first partial string to its left neighbour."""
    size = len(text)
    small, larger = divmod(size, p)
    slices = []
    for r in range(p):
        ln = small + (1 if r < larger else 0)
        off = r * (small + 1) if r < larger else larger * (small + 1) + (r - larger) * small
        slices.append(bytearray(text[off:off + ln]))
    firsts = []
    for r in range(p):
        s = slices[r]
        fe = s.index(b"\n") if b"\n" in s else len(s)
        firsts.append(fe)
    out = []
    for r in range(p):
        s = bytearray(slices[r].replace(b"\n", b"\0"))
        if r + 1 < p:
            nxt = slices[r + 1]
            s += nxt[:firsts[r + 1] + 1].replace(b"\n", b"\0")
        elif s and s[-1] != 0:
            s.append(0)
        if r > 0:
            s = s[firsts[r] + 1:]
        out.append(bytes(s))
    return out
