#!/usr/bin/env python3
"""Benchmark: distinguishing-prefix GB/s of the prefix-doubling distributed
string sort on one MI355X.

Workload (one "step"): DistributedPrefixDoublingSort over 16 virtual PEs on
the D/N-ratio input, 20,000,000 strings x 500 chars, D/N = 0.5 (generated in
HBM before timing; the strings stay resident): local LCP sort of every PE,
prefix doubling (XXH64-stream hashes, hash range 2^64-1 unless --bloom-bits
is given, doubling from depth 8), splitter sampling + bucket search, the
distinguishing-prefix exchange, the LCP merge and the output permutation.

value = D / (time per step), D = sum over the sorted output of
max(lcp_i, lcp_{i+1}) + 1 (the reference's computeD), in GB/s.

The JSON line also carries
  roofline:     the dominant kernel's algorithmic bytes per launch over its
                HIP-event-timed average duration, against MI355X's 8 TB/s;
  passes:       the same over all timed kernels of a step together, and the
                whole step's (D + 16 n) bytes over its time;
  cpu_baseline: the CPU restatement of the same pipeline, run PE-parallel
                (one PE per thread, one thread per available core) on the
                same input, timed on the host.
Only one GPU is used; --gpus N > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "distributed-string-sorting_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Per-kernel statistics from the library: launches, event-timed milliseconds
# and algorithmic bytes accumulated over the launches of the timed region.
def roofline_from_stats(stats: dict):
    if not stats:
        return None
    name = max(stats, key=lambda k: stats[k]["ms"])
    s = stats[name]
    if s["launches"] == 0 or s["ms"] <= 0:
        return None
    avg_ms = s["ms"] / s["launches"]
    bytes_per_launch = s["alg_bytes"] / s["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {
        "kernel": name,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_ms": round(avg_ms, 4),
        "launches_per_step": None,
        "alg_bytes_per_launch": int(bytes_per_launch),
    }


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c2_dton_20m_pd_sort_traffic.md")


def measured_traffic(timer_name):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by the counter's
    reading for a plain copy of known size); None if not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 6 and cells[0] == timer_name and cells[5].isdigit():
                    return int(cells[5])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, length, dton, pes, seed, m, sampler):
    """Times the CPU restatement of the whole prefix-doubling sort, run
    PE-parallel the way the reference's MPI ranks run it (one PE per thread,
    one thread per available core, at most one per PE), on the same input."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(cores, pes))
    t0 = time.perf_counter()
    chars = O.gen_dton(n, length, dton, seed)
    ss = O.StringSet(chars, pes)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = O.prefix_doubling_sort_mt(ss, sampler=sampler, m=m, threads=threads)
    dt = time.perf_counter() - t0
    D = res["D"]
    ph = res["phase_s"]
    return {
        "value": round(D / dt / 1e9, 5),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "cores_available": cores,
        "sample": f"the full workload: D/N generator, {n} strings x {length} chars, D/N {dton}, {pes} PEs, "
                  f"same seed; the C restatement of the whole prefix-doubling sort (local sort, "
                  f"prefix doubling, partition, exchange, merge, permutation) PE-parallel on {threads} "
                  f"threads (one PE per thread, as one MPI rank per core), {dt:.2f} s, D = {D}",
        "strings_per_s": round(n / dt, 1),
        "phase_s": {k: round(v, 3) for k, v in zip(
            ("local_sort", "prefix_doubling", "partition", "exchange", "merge", "permutation"), ph)},
        "input_s": round(gen_s, 2),
    }


def batched_leg(args):
    """Plain vs batched vs served forms on the same mixed job stream.

    plain:   every job alone (a batch of one), one after the other, on jobs
             already resident in HBM; only the sorts are timed.
    batched: the same resident jobs through the continuous-batching server
             with all jobs ready at once (makespan timed).
    served:  Poisson arrivals at rate lambda; each job is generated on the
             device when admitted (its ingestion) into a pool of bounded
             capacity; latency = completion - arrival."""
    import numpy as np
    from dss_amd import JobPool
    from dss_amd.serve import Server, fifo_latencies, job_stream, percentiles, run_plain

    specs = job_stream(args.serve_jobs, args.serve_n_min, args.serve_n_max, args.serve_len_min,
                       args.serve_len_max, rate=0.0, seed=args.serve_seed)
    cap_bytes = int(args.serve_pool_gb * (1 << 30))
    cap_strings = max(2 * max(s.n for s in specs), 256 << 20)
    # jobs are processed in groups that fit the pool at once
    groups, cur, used, used_n = [], [], 0, 0
    for s in specs:
        b = s.max_bytes() + 8
        if cur and (used + b > cap_bytes or used_n + s.n > cap_strings):
            groups.append(cur)
            cur, used, used_n = [], 0, 0
        cur.append(s)
        used += b
        used_n += s.n
    if cur:
        groups.append(cur)
    # resident batches: FIFO batches up to the string budget (throughput);
    # served: large jobs on one worker, small jobs batched on the others
    reserve = max(args.serve_batch_strings, max(s.n for s in specs))
    server = Server(None, inflight=args.serve_inflight, max_batch_strings=args.serve_batch_strings,
                    large_strings=None, reserve_strings=reserve)
    served = Server(None, inflight=args.serve_inflight, max_batch_strings=args.serve_batch_strings,
                    large_strings=args.serve_large_strings or None, reserve_strings=reserve)
    plain_t = {}
    gen_t = {}
    Dp = {}
    Db = {}
    batched_s = 0.0
    batch_sizes = []
    with JobPool(cap_bytes, cap_strings) as pool:
        server.pool = pool
        # first launch of every kernel (code-object loading) before any timing
        w = job_stream(1, 1000, 1000, rate=0.0, seed=1)[0]
        j = w.add_to(pool)
        run_plain(pool, [w], resident={w.jid: j})
        server.prepare(j)
        pool.free(j)
        for g in groups:
            ids = {}
            for s in g:
                t0 = time.perf_counter()
                ids[s.jid] = s.add_to(pool)
                gen_t[s.jid] = time.perf_counter() - t0
            pr = run_plain(pool, g, resident=ids)
            for jid, (D, t, _, _) in pr.items():
                plain_t[jid] = t
                Dp[jid] = D
            t0 = time.perf_counter()
            res = server.run(g, realtime=False, resident=ids)
            batched_s += time.perf_counter() - t0
            for r in res:
                Db[r.jid] = r.D
                batch_sizes.append(r.batch_jobs)
            for jid in ids.values():
                pool.free(jid)
        log(f"batched leg: {len(specs)} jobs in {len(groups)} groups, plain {sum(plain_t.values()):.3f} s, "
            f"batched {batched_s:.3f} s, generation {sum(gen_t.values()):.3f} s")
        server.close()
    mism = [j for j in Dp if Dp[j] != Db[j]]
    if mism:
        log(f"batched D differs from plain for jobs {mism[:10]}")
        sys.exit(1)
    small = small_jobs_leg(args, specs, Dp, plain_t)
    D_total = sum(Dp.values())
    plain_gbs = D_total / sum(plain_t.values()) / 1e9
    batched_gbs = D_total / batched_s / 1e9

    # served: Poisson arrivals at `load` times the batched rate including ingestion
    busy = batched_s + sum(gen_t.values())
    lam = args.serve_load * len(specs) / busy
    sspecs = job_stream(args.serve_jobs, args.serve_n_min, args.serve_n_max, args.serve_len_min,
                        args.serve_len_max, rate=lam, seed=args.serve_seed)
    serve_cap = int(args.serve_served_pool_gb * (1 << 30))
    with JobPool(serve_cap, cap_strings) as pool:
        served.pool = pool
        w = job_stream(1, 1000, 1000, rate=0.0, seed=1)[0]
        j = w.add_to(pool)
        served.prepare(j)
        pool.free(j)
        t0 = time.perf_counter()
        res = served.run(sspecs, realtime=True)
        wall = time.perf_counter() - t0
        served.close()
    Ds = {r.jid: r.D for r in res}
    if any(Ds[j] != Dp[j] for j in Dp):
        log("served D differs from plain")
        sys.exit(1)
    lat = [r.latency for r in res]
    admit_wait = [r.admitted - r.arrival for r in res]
    span = max(r.done for r in res) - min(r.arrival for r in res)
    plain_lat = fifo_latencies(sspecs, {j: plain_t[j] + gen_t[j] for j in plain_t})
    n_total = sum(s.n for s in specs)
    return {
        "workload": f"{len(specs)} jobs, sizes log-uniform {args.serve_n_min}-{args.serve_n_max} strings, "
                    f"lengths U[{args.serve_len_min},{args.serve_len_max}], D/N U[0.1,0.9], "
                    f"D/N-ratio and skewed generators, seed {args.serve_seed:#x}",
        "jobs": len(specs),
        "strings": n_total,
        "D_bytes": D_total,
        "plain_gbs": round(plain_gbs, 3),
        "batched_gbs": round(batched_gbs, 3),
        "batched_over_plain": round(batched_gbs / plain_gbs, 3),
        "plain_strings_per_s": round(n_total / sum(plain_t.values()), 1),
        "batched_strings_per_s": round(n_total / batched_s, 1),
        "mean_jobs_per_batch": round(float(np.mean(batch_sizes)), 2),
        "small_jobs": small,
        # where the plain time goes: jobs above the batch budget sort alone in
        # both forms, so their share bounds what batching can gain
        "plain_s_by_size": plain_by_size(specs, plain_t, args.serve_small_max, args.serve_batch_strings),
        "served": {
            "lambda_jobs_per_s": round(lam, 3),
            "load": args.serve_load,
            "pool_gib": args.serve_served_pool_gb,
            "inflight_batches": args.serve_inflight,
            "max_batch_strings": args.serve_batch_strings,
            "large_job_strings": args.serve_large_strings,
            "gbs": round(D_total / span / 1e9, 3),
            "wall_s": round(wall, 3),
            **percentiles(lat),
            # arrival to admission (the job's characters fit the pool): jobs
            # behind a large one waiting for memory are admitted past it
            # (backfill), so this stays short for the small jobs
            "admission_wait": percentiles(admit_wait),
            "plain_fifo": percentiles(plain_lat),
        },
    }


def plain_by_size(specs, plain_t, small_max, batch_max):
    bins = {f"<={small_max}": 0.0, f"<={batch_max}": 0.0, f">{batch_max}": 0.0}
    keys = list(bins)
    for s in specs:
        k = keys[0] if s.n <= small_max else keys[1] if s.n <= batch_max else keys[2]
        bins[k] += plain_t[s.jid]
    tot = sum(bins.values())
    return {k: {"s": round(v, 4), "share": round(v / tot, 4) if tot else 0.0} for k, v in bins.items()}


def small_jobs_leg(args, specs, Dp, plain_t):
    """Plain vs batched on the small-job subset of the same stream (jobs of at
    most --serve-small-max strings), all resident at once: the plain times
    are those of the mixed run (each job alone, sort only); the batched run
    sorts the subset through a FIFO server of its own."""
    import numpy as np
    from dss_amd import JobPool
    from dss_amd.serve import Server, run_plain

    small = [s for s in specs if s.n <= args.serve_small_max]
    if not small:
        return None
    cap_bytes = sum(s.max_bytes() + 8 for s in small) + (64 << 20)
    cap_strings = sum(s.n for s in small) + (1 << 20)
    server = Server(None, inflight=args.serve_inflight, max_batch_strings=args.serve_batch_strings,
                    large_strings=None, reserve_strings=args.serve_batch_strings)
    with JobPool(cap_bytes, cap_strings) as pool:
        server.pool = pool
        w = [s for s in small][:1]
        j = w[0].add_to(pool)
        run_plain(pool, w, resident={w[0].jid: j})
        server.prepare(j)
        pool.free(j)
        ids = {s.jid: s.add_to(pool) for s in small}
        t0 = time.perf_counter()
        res = server.run(small, realtime=False, resident=ids)
        batched_s = time.perf_counter() - t0
        sizes = [r.batch_jobs for r in res]
        server.close()
    if any(r.D != Dp[r.jid] for r in res):
        log("small-job batched D differs from plain")
        sys.exit(1)
    plain_s = sum(plain_t[s.jid] for s in small)
    D = sum(Dp[s.jid] for s in small)
    n = sum(s.n for s in small)
    return {"max_strings": args.serve_small_max, "jobs": len(small), "strings": n, "D_bytes": D,
            "plain_s": round(plain_s, 4), "batched_s": round(batched_s, 4),
            "plain_us_per_job": round(plain_s / len(small) * 1e6, 1),
            "batched_us_per_job": round(batched_s / len(small) * 1e6, 1),
            "batched_over_plain": round(plain_s / batched_s, 3),
            "mean_jobs_per_batch": round(float(np.mean(sizes)), 2)}


def extras_leg(args, D_c2):
    """Measurements of the paths beside the headline sort, on the same GPU:
    the wire codecs on C2-shaped streams (device-resident, HIP-event timed),
    the distributed merge sort on the 1M x 500 B D/N input over 4 PEs (the
    reference's CPU-runnable case) with each string encoder, the text loader,
    the exact distinguishing-prefix validator, and both drivers on the
    skewed 20M-string input and on the 200M capped suffixes of a 200 MB
    Markov text (16 PEs each)."""
    import numpy as np
    import dss_amd
    from dss_amd import WireCodec

    out = {}
    rng = np.random.default_rng(7)
    # -- codecs: one PE's prefix-doubling hash stream (1.25M sorted hashes in
    # 16 intervals) and the LCPs of a 20M-string exchange (u64 values < 512)
    p, per = 16, args.n // 16
    m = 2**64 - 1
    counts = np.full(p, per // p, dtype=np.uint64)
    hashes = np.concatenate([np.sort(rng.integers(q * (m // p), (q + 1) * (m // p) - 1, size=int(c),
                                                  dtype=np.uint64)) for q, c in enumerate(counts)])
    lcps = rng.integers(200, 300, size=args.n, dtype=np.uint64)
    lcounts = np.full(p, args.n // p, dtype=np.uint64)
    with WireCodec() as c:
        # host forms stage the streams in the codec's HBM buffers; the kernel
        # timer sees only the coding kernels (no PCIe copies)
        res = {}
        for it in range(4):  # 1 warmup + 3 timed
            c.set_timing(it > 0)
            words, _ = c.send_hashes(hashes, counts, m)
            dec, bc = c.golomb_decode(words, p, hashes.size)
            data, _ = c.write_ranges(lcps, lcounts)
            back = c.read_ranges(data, lcounts)
        st = c.kernel_stats(reset=True)
        assert np.array_equal(dec, hashes) and np.array_equal(back, lcps)
        assert int(bc.sum()) == hashes.size
        nw, nb = words.size, data.size
        def per_call(names, calls=3):
            # the host form sizes its output first: code lengths and their
            # scan run twice per encode, count them once
            t = 0.0
            for k in names:
                if k in st:
                    twice = k in ("golomb_len", "golomb_scan")
                    t += st[k]["ms"] / (2 if twice else 1)
            return t / calls
        enc_ms = per_call(("golomb_len", "golomb_scan", "golomb_put"))
        dec_ms = per_call(("golomb_spec", "golomb_sync", "golomb_count", "golomb_decode", "golomb_sum"))
        venc_ms = per_call(("varint_len", "varint_scan", "varint_put"))
        vdec_ms = per_call(("varint_count", "varint_decode"))
        res["golomb_hash_stream"] = {
            "values": int(hashes.size), "intervals": p, "words": int(nw),
            "bits_per_value": round(nw * 64 / hashes.size, 2),
            "encode_ms": round(enc_ms, 4), "decode_ms": round(dec_ms, 4),
            "encode_values_per_s": round(hashes.size / enc_ms * 1e3, 1),
            "decode_values_per_s": round(hashes.size / dec_ms * 1e3, 1)}
        res["varint_lcps"] = {
            "values": int(lcps.size), "bytes": int(nb),
            "encode_ms": round(venc_ms, 4), "decode_ms": round(vdec_ms, 4),
            "encode_GBps_in": round(lcps.size * 8 / venc_ms / 1e6, 1),
            "decode_GBps_out": round(lcps.size * 8 / vdec_ms / 1e6, 1)}
        res["kernels"] = {k: {"launches": v["launches"], "ms": round(v["ms"] / 3, 4)} for k, v in st.items()}
        out["codec"] = res

    # -- distributed merge sort, 1M x 500 B D/N 0.5 over 4 PEs
    ms = {}
    with dss_amd.DeviceStringSet.dton(1_000_000, 500, 0.5, 0x5EED0001, 4) as ds:
        for enc in ("whole", "lcp_compact", "no_lcps"):
            ds.merge_sort(compressed=enc)
            ds.sync()
            chk = ds.check()
            assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0
            t0 = time.perf_counter()
            for _ in range(5):
                ds.merge_sort(compressed=enc)
            ds.sync()
            dt = (time.perf_counter() - t0) / 5
            ms[enc] = {"ms": round(dt * 1e3, 3), "dp_GBps": round(chk["D"] / dt / 1e9, 2)}
        out["merge_sort_1m_x_500_p4"] = ms

    # -- exact distinguishing prefixes and the bound check after the C2 sort
    with dss_amd.DeviceStringSet.dton(args.n, args.len, args.dton, args.seed, args.pes) as ds:
        ds.prefix_doubling_sort(args.sampler)
        ds.check_dist()
        ds.sync()
        t0 = time.perf_counter()
        c = ds.check_dist()
        dt = time.perf_counter() - t0
        out["exact_dist_c2"] = {"ms": round(dt * 1e3, 3), "lower_violations": c["lower_violations"],
                                "upper_violations": c["upper_violations"], "exact_sum": c["exact_sum"],
                                "results_sum": c["results_sum"], "D_check": D_c2}
        # the Bloom-filter-sized variant: hash range 8 bits per string
        mb = 8 * args.n
        ds.prefix_doubling_sort(args.sampler, mb)
        ds.sync()
        chk = ds.check()
        assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0
        t0 = time.perf_counter()
        for _ in range(3):
            ds.prefix_doubling_sort(args.sampler, mb)
        ds.sync()
        dt = (time.perf_counter() - t0) / 3
        c = ds.check_dist()
        out["c2_bloom_8_bits_per_string"] = {
            "hash_range": mb, "ms": round(dt * 1e3, 3), "dp_GBps": round(chk["D"] / dt / 1e9, 2),
            "lower_violations": c["lower_violations"], "results_sum": c["results_sum"],
            "exact_sum": c["exact_sum"], "D": chk["D"]}

    # -- text loader: the 1M x 500 B input as a '\n'-separated file, 4 PEs
    import tempfile
    with dss_amd.DeviceStringSet.dton(1_000_000, 500, 0.5, 0x5EED0001, 4) as ds:
        chars, _, _ = ds.download()
        pb = ds.pe_begin()
    text = chars.copy()
    text[text == 0] = ord("\n")
    with tempfile.NamedTemporaryFile(suffix=".txt", dir=os.environ.get("TMPDIR", "/tmp"), delete=False) as f:
        f.write(text.tobytes())
        path = f.name
    try:
        with dss_amd.DeviceStringSet.load_file(path, 4) as ds:  # warm the page cache and the code
            pass
        t0 = time.perf_counter()
        with dss_amd.DeviceStringSet.load_file(path, 4) as ds:
            dt = time.perf_counter() - t0
            assert (ds.pe_begin() == pb).all() and ds.n == 1_000_000
        out["text_loader_1m_x_500_p4"] = {"bytes": int(text.size), "ms": round(dt * 1e3, 3),
                                          "GBps_file_to_set": round(text.size / dt / 1e9, 2),
                                          "note": "page-cached file, pinned staging, PCIe copy, "
                                                  "'\\n' -> NUL, index and PE split"}
    finally:
        os.unlink(path)

    # -- the other full-size inputs: skewed D/N (20M strings, 16 PEs) and
    #    capped suffixes of a 200 MB Markov text (200M views, 16 PEs); one
    #    untimed sort for D and the check, then 3 timed sorts per driver
    def timed(ds, run, reps=3):
        run()
        ds.sync()
        chk = ds.check()
        assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        ds.sync()
        dt = (time.perf_counter() - t0) / reps
        return {"ms": round(dt * 1e3, 3), "dp_GBps": round(chk["D"] / dt / 1e9, 2),
                "strings_per_s": round(chk["num_strings"] / dt, 1), "D": chk["D"]}

    with dss_amd.DeviceStringSet.skewed(20_000_000, 0x5EED0003, 16) as ds:
        out["c3_skewed_20m_p16"] = {
            "prefix_doubling_sort": timed(ds, lambda: ds.prefix_doubling_sort()),
            "merge_sort": timed(ds, lambda: ds.merge_sort()),
        }
    with dss_amd.DeviceStringSet.markov_suffixes(200_000_000, 0x5EED0004, 16, repeats=5000, rep_min=1000,
                                                 rep_max=10000, cap=500) as ds:
        out["c4_suffixes_200m_p16"] = {
            "prefix_doubling_sort": timed(ds, lambda: ds.prefix_doubling_sort()),
            "merge_sort_lcp_compact": timed(ds, lambda: ds.merge_sort(compressed="lcp_compact")),
        }
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=20_000_000)
    ap.add_argument("--len", type=int, default=500)
    ap.add_argument("--dton", type=float, default=0.5)
    ap.add_argument("--pes", type=int, default=16)
    ap.add_argument("--seed", type=lambda x: int(x, 0), default=0x5EED0002)
    ap.add_argument("--sampler", default="IndexedNumStrings")
    ap.add_argument("--workload", choices=["dton", "skewed", "suffixes"], default="dton",
                    help="dton: the D/N-ratio input (--n, --len, --dton, --pes, --seed); skewed: 20M skewed strings "
                         "(20%% of 1000 chars D/N .75, 80%% of 100 chars D/N .25, 10%% duplicates, bytes 1-255, "
                         "seed 0x5EED0003); suffixes: capped suffixes (500 chars) of a 200 MB order-3 Markov "
                         "ACGT text with 5000 planted repeats of 1-10 KB (seed 0x5EED0004); both 16 PEs")
    ap.add_argument("--bloom-bits", type=int, default=0,
                    help="hash range = bits * n (the 'Bloom filter sized b bits/string' variant); "
                         "0 = 2^64-1 as in the reference")
    ap.add_argument("--cpu-sample", type=int, default=20_000_000,
                    help="strings of the CPU baseline's input (default: the whole workload; 0 disables it)")
    ap.add_argument("--serve-jobs", type=int, default=2000,
                    help="jobs of the batched/served leg (0 disables it)")
    ap.add_argument("--serve-n-min", type=int, default=1000)
    ap.add_argument("--serve-n-max", type=int, default=50_000_000)
    ap.add_argument("--serve-small-max", type=int, default=100_000,
                    help="jobs of at most this many strings form the small-job subset")
    ap.add_argument("--serve-len-min", type=int, default=10)
    ap.add_argument("--serve-len-max", type=int, default=1000)
    ap.add_argument("--serve-seed", type=lambda x: int(x, 0), default=0x5EED0005)
    ap.add_argument("--serve-load", type=float, default=0.8)
    ap.add_argument("--serve-inflight", type=int, default=2)
    ap.add_argument("--serve-batch-strings", type=int, default=32 << 20)
    ap.add_argument("--serve-large-strings", type=int, default=4 << 20,
                    help="served jobs of at least this many strings run alone on worker 0 (0: FIFO batches)")
    ap.add_argument("--serve-pool-gb", type=float, default=100.0,
                    help="pool capacity of the plain/batched comparison (GiB of characters)")
    ap.add_argument("--serve-served-pool-gb", type=float, default=160.0,
                    help="pool capacity of the served run (the admission limit)")
    ap.add_argument("--extras", type=int, default=1,
                    help="measure the codecs, merge-sort encoders, exact-prefix validator and loader (0: off)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(p virtual PEs share the one MI355X)", file=sys.stderr)
        sys.exit(2)

    import dss_amd

    m = (2**64 - 1) if args.bloom_bits == 0 else args.bloom_bits * args.n
    log("device:", dss_amd.device_info())
    t0 = time.time()
    if args.workload == "skewed":
        args.n, args.pes = 20_000_000, 16
        ds = dss_amd.DeviceStringSet.skewed(args.n, 0x5EED0003, args.pes)
    elif args.workload == "suffixes":
        args.n, args.pes = 200_000_000, 16
        ds = dss_amd.DeviceStringSet.markov_suffixes(args.n, 0x5EED0004, args.pes, repeats=5000, rep_min=1000,
                                                     rep_max=10000, cap=500)
    else:
        ds = dss_amd.DeviceStringSet.dton(args.n, args.len, args.dton, args.seed, args.pes)
    ds.sync()
    info = ds.info()
    log(f"generated {info['num_strings']} strings, {info['num_bytes']} bytes in {time.time() - t0:.2f} s")

    for w in range(args.warmup):
        ds.prefix_doubling_sort(args.sampler, m)
        ds.sync()
        log(f"warmup {w} done")
    if args.warmup == 0:  # the output check needs one sort
        ds.prefix_doubling_sort(args.sampler, m)
        ds.sync()
    chk = ds.check()
    log("check:", chk)
    if not (chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0):
        log("output check failed")
        sys.exit(1)
    D = chk["D"]

    # timed region: K steps without per-kernel instrumentation
    ds.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ds.prefix_doubling_sort(args.sampler, m)
    ds.sync()
    t1 = time.perf_counter()
    step_s = (t1 - t0) / args.steps
    # the same K steps again with HIP events around every kernel launch (on
    # the set's stream) for the per-kernel durations and the roofline
    ds.kernel_stats(reset=True)
    ds.set_timing(True)
    for _ in range(args.steps):
        ds.prefix_doubling_sort(args.sampler, m)
    ds.sync()
    ds.set_timing(False)
    stats = ds.kernel_stats(reset=True)
    for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"]):
        log(f"  {k:18s} launches={v['launches']:5d} ms/step={v['ms'] / args.steps:8.3f} "
            f"GB/s={(v['alg_bytes'] / (v['ms'] * 1e-3) / 1e9) if v['ms'] > 0 else 0:8.1f}")
    roof = roofline_from_stats(stats)
    if roof:
        roof["launches_per_step"] = stats[roof["kernel"]]["launches"] / args.steps
        default_workload = args.workload == "dton" and (args.n, args.len, args.dton, args.pes, args.seed,
                                                        args.bloom_bits) == (20_000_000, 500, 0.5, 16, 0x5EED0002, 0)
        if default_workload:
            roof["traffic"] = measured_traffic(roof["kernel"])
            if roof["traffic"] is not None:
                roof["traffic_source"] = os.path.relpath(TRAFFIC_FILE, ROOT)

    # all timed kernels together (the sort's passes, prefix doubling,
    # partition, exchange and merge) and the whole step, against 8 TB/s
    alg_sum = sum(v["alg_bytes"] for v in stats.values())
    ms_sum = sum(v["ms"] for v in stats.values())
    passes = {
        "alg_bytes_per_step": int(alg_sum / args.steps),
        "kernel_ms_per_step": round(ms_sum / args.steps, 3),
        "frac": round(alg_sum / (ms_sum * 1e-3) / (HBM_PEAK_GBS * 1e9), 4) if ms_sum > 0 else None,
        "end_to_end_bytes": int(D + 16 * args.n),
        "end_to_end_frac": round((D + 16 * args.n) / step_s / (HBM_PEAK_GBS * 1e9), 4),
    }
    ds.close()
    batched = None
    if args.serve_jobs > 0:
        log(f"batched/served leg: {args.serve_jobs} jobs ...")
        batched = batched_leg(args)

    extras = None
    if args.extras:
        log("extras: codecs, merge sort encoders, exact distinguishing prefixes, text loader ...")
        extras = extras_leg(args, D)

    cpu = None
    if args.cpu_sample > 0 and args.workload == "dton":
        log(f"CPU restatement on {args.cpu_sample} strings ...")
        cpu = cpu_baseline(args.cpu_sample, args.len, args.dton, args.pes, args.seed, m, args.sampler)
        log(f"CPU: {cpu['value']} GB/s on {cpu['cores']} threads, phases {cpu['phase_s']}")

    wl = {"dton": f"D/N-ratio input, {args.n} strings x {args.len} chars, D/N {args.dton}",
          "skewed": f"skewed D/N input, {args.n} strings (20% x 1000 chars D/N 0.75, 80% x 100 chars D/N 0.25, "
                    f"10% exact duplicates, bytes 1-255)",
          "suffixes": f"{args.n} suffixes capped at 500 chars of a 200 MB order-3 Markov ACGT text with 5000 "
                      f"planted repeats of 1-10 KB"}[args.workload]
    out = {
        "metric": "distinguishing-prefix GB/s at 1 GPU (batched vs plain); p50/p99 job latency",
        "value": round(D / step_s / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded D/N-ratio generator, generated in HBM)",
        "config": {
            "workload": f"{wl}, {args.pes} virtual PEs, prefix-doubling sort (XXH64-stream, "
                        f"hash range {'2^64-1' if args.bloom_bits == 0 else f'{args.bloom_bits}*n'}, "
                        f"start depth 8), sampler {args.sampler}",
            "strings": args.n,
            "string_length": args.len,
            "d_to_n": args.dton,
            "virtual_pes": args.pes,
            "D_bytes": D,
            "strings_per_s": round(args.n / step_s, 1),
        },
        "roofline": roof,
        "passes": passes,
        "cpu_baseline": cpu,
        "batched": batched,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
