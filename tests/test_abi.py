"""CPU checks of the C ABI: the built library loads and exports every symbol
that include/dss_gpu.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

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
