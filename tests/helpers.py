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
