"""CPU checks of the C ABI: the built library loads and exports every symbol
that include/dss_gpu.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "distributed-string-sorting_amd", "dss_amd", "libdss_gpu.so")


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(dss_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    assert os.path.exists(LIB), "libdss_gpu.so not built"
    lib = ctypes.CDLL(LIB)
    names = declared_symbols()
    assert len(names) > 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_declared_symbols():
    import dss_amd._lib as L
    assert set(L.symbols()) == declared_symbols()


def test_job_stream_is_seeded_and_bounded():
    from dss_amd.serve import job_stream
    a = job_stream(300, 1000, 50_000_000, 10, 1000, rate=5.0, seed=0x5EED0005)
    b = job_stream(300, 1000, 50_000_000, 10, 1000, rate=5.0, seed=0x5EED0005)
    assert [vars(x) for x in a] == [vars(x) for x in b]
    assert all(1000 <= s.n <= 50_000_000 and 10 <= s.length <= 1000 for s in a)
    assert all(0.1 <= s.dton <= 0.9 for s in a)
    assert {s.kind for s in a} == {"dton", "skewed"}
    arr = [s.arrival for s in a]
    assert arr == sorted(arr) and arr[-1] > 0
    # log-uniform: the median size is near the geometric mean of the bounds
    med = sorted(s.n for s in a)[150]
    assert 50_000 < med < 5_000_000


def test_fifo_latencies_single_server():
    from dss_amd.serve import JobSpec, fifo_latencies
    specs = [JobSpec(0, "dton", 1, 1, 0.5, 0, arrival=0.0), JobSpec(1, "dton", 1, 1, 0.5, 0, arrival=1.0),
             JobSpec(2, "dton", 1, 1, 0.5, 0, arrival=10.0)]
    lat = fifo_latencies(specs, {0: 3.0, 1: 1.0, 2: 0.5})
    # job 1 waits for job 0 (done at 3), job 2 finds the server idle
    assert list(lat) == [3.0, 3.0, 0.5]


def test_job_stream_jobs_do_not_depend_on_rate():
    from dss_amd.serve import job_stream
    a = job_stream(50, 1000, 10**6, rate=0.0, seed=3)
    b = job_stream(50, 1000, 10**6, rate=7.5, seed=3)
    strip = lambda s: {k: v for k, v in vars(s).items() if k != "arrival"}  # noqa: E731
    assert [strip(x) for x in a] == [strip(x) for x in b]


def test_tuner_regime_and_cache(tmp_path):
    from dss_amd.tune import Tuner, regime
    info = dict(num_strings=1_000_000, num_bytes=501_000_000, num_pes=4, max_len=500)
    assert regime(info) == "n2^17/len2^8/lcp?/p4"
    assert regime(info, avg_lcp=249) == "n2^17/len2^8/lcp4/p4"
    path = tmp_path / "tune.json"
    t = Tuner(str(path))
    t.table["x"] = {"large_min": 256, "wave_class_max": 5, "skip_lanes": 16, "skip_strings": 2, "merge_fan_in": 8}
    t.save()
    assert Tuner(str(path)).table == t.table


REFERENCE = "/root/reference"
ADAPTOR_TU = r'''
#include <cstdint>
#include <vector>
#include "strings/stringcontainer.hpp"
#include "dss_gpu_adaptors.hpp"

// the shape of StringIndexPEIndex (src/util/structs.hpp:9-14, whose header
// pulls in MPI)
struct Entry {
    size_t stringIndex, PEIndex;
    Entry(size_t s, size_t p) : stringIndex(s), PEIndex(p) {}
};
using SSet = dss_schimek::UCharLengthStringSet;

int main(int argc, char**) {
    if (argc < 100) return 0;  // built and linked here, never run (no GPU)
    std::vector<unsigned char> raw = {'b', 0, 'a', 0};
    dss_schimek::StringLcpContainer<SSet> cont(std::move(raw));
    auto ptr = cont.make_string_lcp_ptr();
    dss_gpu::radixsort_CI3(ptr, 0, 0);  // the call of lcp_aware...:218-224
    std::vector<uint64_t> hashes;
    dss_gpu::hash_round(nullptr, {0, 1}, 8, hashes);
    auto perm = dss_gpu::prefix_doubling_sort<Entry>(cont.raw_strings().data(), {0, cont.char_size()});
    return (int)perm.size();
}
'''


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "src")), reason="reference checkout not present")
def test_reference_side_adaptors_compile_and_link(tmp_path):
    """include/dss_gpu_adaptors.hpp (the drop-ins of INTEGRATION.md)
    instantiated with the reference's own StringLcpContainer / StringLcpPtr
    compiles against the reference headers and links against libdss_gpu.so,
    so a change of either side's interface fails here."""
    import subprocess
    tlx = os.path.join(REFERENCE, "NeuKaDiS/KaDiS/external/RBC/external/tlx")
    src = tmp_path / "use_adaptors.cpp"
    src.write_text(ADAPTOR_TU)
    libdir = os.path.dirname(LIB)
    cmd = ["g++", "-std=c++17", "-DNDEBUG", "-w", "-I", os.path.join(REFERENCE, "src"), "-I", tlx,
           "-I", os.path.join(ROOT, "include"), str(src), "-o", str(tmp_path / "use_adaptors"),
           f"-L{libdir}", "-ldss_gpu", f"-Wl,-rpath,{libdir}", "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
