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


class HipMem:
    """Device buffers for tests of the library's device-pointer entry points,
    through the HIP runtime the library itself loaded (libamdhip64.so.7,
    resolved by soname once dss_amd is imported), so no second runtime is
    brought into the process."""

    def __init__(self):
        import ctypes as C
        import dss_amd  # noqa: F401  (loads libdss_gpu.so and its HIP runtime)
        self.C = C
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.h.hipDeviceSynchronize.argtypes = []
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.bufs = []
        self.streams = []

    def stream(self) -> int:
        s = self.C.c_void_p()
        self._ok(self.h.hipStreamCreate(self.C.byref(s)), "hipStreamCreate")
        self.streams.append(s.value)
        return s.value

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def alloc(self, nbytes: int) -> int:
        p = self.C.c_void_p()
        self._ok(self.h.hipMalloc(self.C.byref(p), max(int(nbytes), 1)), "hipMalloc")
        self._ok(self.h.hipMemset(p, 0, max(int(nbytes), 1)), "hipMemset")
        self.bufs.append(p.value)
        return p.value

    def upload(self, a) -> int:
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self._ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, ptr: int, nbytes: int, dtype=np.uint8):
        out = np.zeros(int(nbytes), dtype=np.uint8)
        if nbytes:
            self._ok(self.h.hipMemcpy(out.ctypes.data, ptr, int(nbytes), 2), "hipMemcpy D2H")
        return out.view(dtype)

    def free_bytes(self) -> int:
        f, t = self.C.c_size_t(), self.C.c_size_t()
        self._ok(self.h.hipMemGetInfo(self.C.byref(f), self.C.byref(t)), "hipMemGetInfo")
        return f.value

    def sync(self):
        self._ok(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []
        for st in self.streams:
            self.h.hipStreamSynchronize(st)
            self.h.hipStreamDestroy(st)
        self.streams = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
