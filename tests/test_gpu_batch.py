"""Batched and served sorting on the GPU: every job's result is the same when
it is sorted alone, inside a mixed batch, and by the continuous-batching
server, and matches the oracle's local sort of that job's strings."""
import numpy as np
import pytest

import oracle as O
from helpers import canonical_ties

pytestmark = pytest.mark.gpu


def _dss():
    import dss_amd
    return dss_amd


def arena(strings):
    return np.frombuffer(b"".join(s + b"\0" for s in strings), dtype=np.uint8).copy()


def mixed_jobs():
    """(kind, args, host chars) of jobs of uneven size, length and LCP profile."""
    jobs = [
        ("dton", (3000, 40, 0.5, 101)),
        ("dton", (1, 12, 0.5, 102)),
        ("skewed", (5000, 103, 120, 12, 0.75, 0.25, 200, 100)),
        ("dton", (20000, 200, 0.9, 104)),
        ("host", arena([b"b", b"a", b"", b"ab", b"a", b"abc" * 40, b"", b"b"])),
        ("dton", (700, 10, 0.1, 105)),
        ("host", arena([b"x" * k for k in range(300, 0, -7)])),
        ("skewed", (12000, 106, 400, 40, 0.5, 0.2, 300, 250)),
        ("host", np.zeros(0, np.uint8)),
        ("dton", (2, 1000, 0.9, 107)),
    ]
    out = []
    for kind, a in jobs:
        if kind == "dton":
            n, length, dn, seed = a
            chars = O.gen_dton(n, length, dn, seed)
        elif kind == "skewed":
            n, seed, ll, sl, ldn, sdn, lp, dp = a
            chars = O.gen_skewed(O.skew_params(n, long_len=ll, short_len=sl, long_dn=ldn, short_dn=sdn,
                                               long_permille=lp, dup_permille=dp), seed)
        else:
            chars = a
        out.append((kind, a, chars))
    return out


def add(pool, kind, a, chars):
    if kind == "dton":
        return pool.add_dton(*a)
    if kind == "skewed":
        n, seed, ll, sl, ldn, sdn, lp, dp = a
        return pool.add_skewed(n, seed, long_len=ll, short_len=sl, long_dn=ldn, short_dn=sdn,
                               long_permille=lp, dup_permille=dp)
    return pool.add(chars)


def oracle_job(chars):
    ss = O.StringSet(chars, 1)
    perm, lcp = O.local_sort(ss)
    return ss, perm, lcp, O.compute_d(lcp)


def test_batch_alone_vs_mixed_vs_oracle():
    dss = _dss()
    jobs = mixed_jobs()
    with dss.JobPool(64 << 20, 1 << 20) as pool:
        ids = [add(pool, k, a, c) for k, a, c in jobs]
        for (k, a, c), j in zip(jobs, ids):
            n, nb = pool.job_info(j)
            assert n == O.StringSet(c, 1).n and nb == c.size, (k, a)
        alone = []
        b = pool.batch([ids[0]])
        for j in ids:
            b.rebatch([j])
            b.sort()
            alone.append((b.result(0), int(b.dist_sizes()[0])))
        b.close()
        order = [3, 0, 7, 1, 9, 4, 2, 8, 6, 5]
        mix = pool.batch([ids[i] for i in order])
        mix.sort()
        dm = mix.dist_sizes()
        for k, i in enumerate(order):
            perm, lcp = mix.result(k)
            (ap, al), ad = alone[i]
            assert (perm == ap).all() and (lcp == al).all(), f"job {i} differs in the mix"
            assert int(dm[k]) == ad
        mix.close()
        for i, (kind, a, chars) in enumerate(jobs):
            (gp, gl), gd = alone[i]
            if chars.size == 0:
                assert gp.size == 0 and gd == 0
                continue
            ss, operm, olcp, od = oracle_job(chars)
            assert (gl == olcp).all(), f"job {i} ({kind}) LCP differs from the oracle"
            pb = np.array([0, ss.n], dtype=np.uint64)
            assert (canonical_ties(gp, gl, ss.len, pb) == canonical_ties(operm, olcp, ss.len, pb)).all()
            assert gd == od, f"job {i} D {gd} != {od}"
        for j in ids:
            pool.free(j)
        u = pool.usage()
        assert u["bytes_used"] == 0 and u["strings_used"] == 0


def test_server_results_match_alone():
    dss = _dss()
    from dss_amd.serve import Server, job_stream, run_plain
    specs = job_stream(40, 200, 30000, 10, 300, rate=400.0, seed=0x5EED0005)
    with dss.JobPool(256 << 20, 4 << 20) as pool:
        plain = run_plain(pool, specs, keep_results=True)
        assert pool.usage()["bytes_used"] == 0
        # realtime arrivals, two batches in flight, small batches; then every
        # job at once (admission outpaces the sorts, so batches fill up)
        srv = Server(pool, inflight=2, max_batch_strings=60000)
        res = srv.run(specs, realtime=True, keep_results=True)
        assert pool.usage()["bytes_used"] == 0
        res2 = srv.run(specs, realtime=False, keep_results=True)
        assert pool.usage()["bytes_used"] == 0
        assert max(r.batch_jobs for r in res2) > 1, "no batch held more than one job"
        for r in res + res2:
            D, _, _, (pp, pl) = plain[r.jid]
            assert r.D == D
            assert (r.perm == pp).all() and (r.lcp == pl).all(), f"job {r.jid} differs when served"
            assert r.done >= r.started >= r.admitted >= r.arrival - 1e-3


def test_server_admission_waits_for_memory():
    dss = _dss()
    from dss_amd.serve import Server, job_stream
    specs = job_stream(24, 1000, 20000, 50, 200, rate=0.0, seed=7)
    need = max(s.max_bytes() for s in specs)
    # room for about two of the largest jobs: admission must wait for frees
    with dss.JobPool(2 * need + 64, 60000) as pool:
        res = Server(pool, inflight=2, max_batch_strings=40000).run(specs, realtime=False)
        assert all(r.done > 0 for r in res)
        for s, r in zip(specs, res):
            chars = O.gen_dton(s.n, s.length, s.dton, s.seed) if s.kind == "dton" else None
            if chars is not None:
                assert r.D == O.compute_d(O.local_sort(O.StringSet(chars, 1))[1])
        assert pool.usage()["bytes_used"] == 0


def test_pool_errors():
    dss = _dss()
    with dss.JobPool(1 << 16, 1000) as pool:
        with pytest.raises(dss.DssError) as e:
            pool.add_dton(2000, 100, 0.5, 1)  # more strings than the pool holds
        assert e.value.code == -2
        j = pool.add(arena([b"abc", b"ab"]))
        pool.free(j)
        with pytest.raises(dss.DssError):
            pool.free(j)
        with pytest.raises(dss.DssError):
            pool.add(np.frombuffer(b"abc", dtype=np.uint8).copy())  # no final NUL
        a = pool.add(arena([b"q"]))
        with pytest.raises(dss.DssError):
            pool.batch([a, a])
        with dss.DeviceStringSet.from_arena(arena([b"z", b"y"]), 1) as ds:
            from dss_amd import _lib
            ids = np.array([a], dtype=np.uint32)
            rc = _lib.lib().dss_pool_rebatch(pool._h, ds._h, ids.ctypes.data_as(_lib.u32p), 1)
            assert rc == -1
        pool.free(a)


def test_reserved_batches_give_the_same_results():
    """Growing a batch set's buffers (dss_set_reserve) keeps its jobs'
    metadata; served runs on pre-grown sets match the plain results."""
    dss = _dss()
    from dss_amd.serve import Server, job_stream, run_plain
    specs = job_stream(30, 300, 20000, 10, 200, rate=0.0, seed=41)
    with dss.JobPool(128 << 20, 2 << 20) as pool:
        plain = run_plain(pool, specs, keep_results=True)
        j = specs[0].add_to(pool)
        b = pool.batch([j])
        b.set.reserve(1 << 20, 64)  # grows off/len after they were filled
        b.sort()
        perm, lcp = b.result(0)
        D = int(b.dist_sizes()[0])
        b.close()
        pool.free(j)
        assert D == plain[specs[0].jid][0]
        assert (perm == plain[specs[0].jid][3][0]).all() and (lcp == plain[specs[0].jid][3][1]).all()
        with Server(pool, inflight=2, max_batch_strings=50000, reserve_strings=1 << 20) as srv:
            w = specs[1].add_to(pool)
            srv.prepare(w)
            pool.free(w)
            for _ in range(2):  # the batch sets persist across runs
                res = srv.run(specs, realtime=False, keep_results=True)
                for r in res:
                    D, _, _, (pp, pl) = plain[r.jid]
                    assert r.D == D and (r.perm == pp).all() and (r.lcp == pl).all()
        assert pool.usage()["bytes_used"] == 0


def test_reserved_set_sorts_without_allocating():
    """After dss_set_reserve a batch set sorts jobs up to the reserved size
    without allocating device memory: a 1M-string job whose base cases run
    one segment per wave (where the sort packs every string's offset and
    length into 16-byte records) leaves the free device memory unchanged, as
    does re-pointing the set at a smaller job."""
    dss = _dss()
    from helpers import HipMem
    with dss.JobPool(256 << 20, 4 << 20) as pool, HipMem() as hm:
        big = pool.add_dton(1_000_000, 20, 0.1, 7)
        small = pool.add_dton(300_000, 20, 0.1, 8)
        b = pool.batch([big])
        b.set.reserve(2_000_000, 4)
        b.sort()  # code objects and the mailbox are set up by the first sort
        b.dist_sizes()
        hm.sync()
        free0 = hm.free_bytes()
        D = []
        for jobs in ([big], [small], [big, small]):
            b.rebatch(jobs)
            b.sort()
            D.append([int(x) for x in b.dist_sizes()])
        hm.sync()
        assert hm.free_bytes() == free0
        assert D[2] == D[0] + D[1]
        b.close()


def _spec_chars(s):
    """Host copy of a job spec's strings from the oracle's generators."""
    if s.kind == "dton":
        return O.gen_dton(s.n, s.length, s.dton, s.seed)
    return O.gen_skewed(O.skew_params(s.n, long_len=s.length, short_len=s.short_len, long_dn=s.dton,
                                      short_dn=s.short_dn), s.seed)


def test_server_large_jobs_alone_path_matches_plain_and_oracle():
    """With the large-job threshold lowered to 20k strings, jobs above it
    take worker 0's alone path while the smaller ones are batched on the
    other worker; every job's sorted order, LCP array and D are identical to
    its plain (alone) sort, and the large jobs' equal the oracle's."""
    dss = _dss()
    from dss_amd.serve import Server, job_stream, run_plain
    specs = job_stream(30, 1000, 150000, 10, 300, rate=300.0, seed=0x5EED0015)
    large = 20000
    assert sum(s.n >= large for s in specs) >= 3 and sum(s.n < large for s in specs) >= 3
    with dss.JobPool(512 << 20, 4 << 20) as pool:
        plain = run_plain(pool, specs, keep_results=True)
        with Server(pool, inflight=2, max_batch_strings=100000, large_strings=large) as srv:
            res = srv.run(specs, realtime=True, keep_results=True)
        assert pool.usage()["bytes_used"] == 0
    by = {s.jid: s for s in specs}
    for r in res:
        D, _, _, (pp, pl) = plain[r.jid]
        assert r.D == D, r.jid
        assert (r.perm == pp).all() and (r.lcp == pl).all(), f"job {r.jid} differs when served"
        if by[r.jid].n >= large:
            assert r.batch_jobs == 1, f"large job {r.jid} shared a batch"
            ss, operm, olcp, od = oracle_job(_spec_chars(by[r.jid]))
            pb = np.array([0, ss.n], dtype=np.uint64)
            assert (r.lcp == olcp).all() and r.D == od
            assert (canonical_ties(r.perm, r.lcp, ss.len, pb) == canonical_ties(operm, olcp, ss.len, pb)).all()
