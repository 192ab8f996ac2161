// The distributed stages after the local sort, for p virtual PEs held on one
// GPU: global average LCP (a3), splitter sampling and bucket computation
// (a4, a5), the string exchange as an in-HBM bucket scatter (a6, a7, a13),
// the LCP merge of the received runs (a8) and prefix decompression (a9).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dss_device.hpp"
#include "dss_internal.hpp"

namespace dss {

// ================================================================== scan
constexpr uint32_t kST = 256, kSI = 8, kSB = kST * kSI;

__device__ inline uint64_t block_excl_sum_u64(uint64_t v, uint64_t* tmp, uint64_t* total) {
    const uint32_t t = threadIdx.x;
    tmp[t] = v;
    __syncthreads();
    for (uint32_t o = 1; o < kST; o <<= 1) {
        const uint64_t x = t >= o ? tmp[t - o] : 0;
        __syncthreads();
        tmp[t] += x;
        __syncthreads();
    }
    const uint64_t incl = tmp[t];
    if (total) *total = tmp[kST - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kST) k_scan_reduce(const uint64_t* __restrict__ in, uint64_t n,
                                                     uint64_t* __restrict__ sums) {
    __shared__ uint64_t tmp[kST];
    const uint64_t b = (uint64_t)blockIdx.x * kSB;
    uint64_t s = 0;
    for (uint32_t k = 0; k < kSI; ++k) {
        const uint64_t i = b + (uint64_t)k * kST + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint64_t tot;
    block_excl_sum_u64(s, tmp, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_sums(uint64_t* __restrict__ v, uint64_t n,
                                                    uint64_t* __restrict__ total) {
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = t * per, e = (b + per < n) ? b + per : n;
    uint64_t s = 0;
    for (uint64_t i = b; i < e; ++i) s += v[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint64_t x = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint64_t run = t ? part[t - 1] : 0;
    for (uint64_t i = b; i < e; ++i) {
        const uint64_t x = v[i];
        v[i] = run;
        run += x;
    }
    if (t == 1023) *total = part[1023];
}

__global__ void __launch_bounds__(kST) k_scan_apply(const uint64_t* __restrict__ in, uint64_t* out,
                                                    uint64_t n, const uint64_t* __restrict__ sums) {
    __shared__ uint64_t tmp[kST];
    const uint64_t b = (uint64_t)blockIdx.x * kSB + (uint64_t)threadIdx.x * kSI;
    uint64_t v[kSI];
    uint64_t s = 0;
    for (uint32_t k = 0; k < kSI; ++k) {
        v[k] = (b + k < n) ? in[b + k] : 0;
        s += v[k];
    }
    uint64_t run = sums[blockIdx.x] + block_excl_sum_u64(s, tmp, nullptr);
    for (uint32_t k = 0; k < kSI; ++k) {
        if (b + k < n) out[b + k] = run;
        run += v[k];
    }
}

__global__ void k_store_u64(uint64_t* __restrict__ dst, uint64_t v) { *dst = v; }

uint64_t scan_scratch_words(uint64_t n) { return std::max<uint64_t>((n + kSB - 1) / kSB, 1) + 1; }

void exclusive_scan_u64_scratch(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t st, uint64_t* sums,
                                uint64_t* total_dev) {
    const uint64_t nb = scan_scratch_words(n) - 1;
    fill_small(st, sums, 0, (nb + 1) * 8);
    if (n) hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kST), 0, st, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, sums, nb, total_dev);
    if (n) hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kST), 0, st, in, out, n, sums);
    DSS_HIP(hipGetLastError());
}

void exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t st, PipeState& P,
                        uint64_t* total_host) {
    P.scan_tmp.reserve(scan_scratch_words(n) * 8);
    P.scan_tot.reserve(8);
    exclusive_scan_u64_scratch(in, out, n, st, P.scan_tmp.as<uint64_t>(), P.scan_tot.as<uint64_t>());
    if (total_host) P.mbox.fetch(st, total_host, P.scan_tot.p, 8);
}

// ================================================================== (a3)
// getAvgLcp (src/sorter/distributed/misc.hpp:67-87): sum of all local LCP
// entries over all PEs divided by the number of strings.
__global__ void k_sum_u32(const uint32_t* __restrict__ v, uint64_t n, unsigned long long* __restrict__ out) {
    uint64_t s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        s += v[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down((unsigned long long)s, o);
    if (lane_id() == 0) atomicAdd(out, (unsigned long long)s);
}

uint64_t avg_lcp(dss_set_s* s) {
    DSS_REQUIRE(s->sorted, DSS_ESTATE, "dss_sort_lcp has not run");
    if (s->n == 0) return 0;
    PipeState& P = s->pp;
    P.scan_tot.reserve(8);
    DSS_HIP(hipMemsetAsync(P.scan_tot.p, 0, 8, s->stream));
    hipLaunchKernelGGL(k_sum_u32, dim3(std::min<uint64_t>((s->n + 255) / 256, 2048)), dim3(256), 0, s->stream,
                       s->lcp.as<uint32_t>(), s->n, P.scan_tot.as<unsigned long long>());
    uint64_t tot = 0;
    P.mbox.fetch(s->stream, &tot, P.scan_tot.p, 8);
    return tot / s->n;
}

// ================================================================== (a4, a5)
// Sample = a (possibly truncated) string of a PE plus its global index.
struct Sample {
    uint64_t off;   // arena offset of the string
    uint32_t tlen;  // truncated length
    uint32_t pad;
    uint64_t gidx;  // localOffset + index in the PE's sorted order
};

// One thread per PE: the sampling policies of samplingStrategies.hpp.
//  NumStrings / IndexedNumStrings (:20-90, :107-187): sample i at
//    (size_t)(i * (double)n / (double)(nr + 1));
//  NumChars / IndexedNumChars (:189-389): walk lengths (IndexedNumChars in
//    the prefix-doubling form: distinguishing prefixes) until splitter_dist
//    = num / (nr + 1) characters were seen; done by binary searches on the
//    prefix sums w_prefix.
// nr = min(factor * (p - 1), n) (getNumberSplitter, :14-18).  A sample is
// truncated to max_len (merge sort) or to dist[idx] (prefix doubling).
__global__ void k_sample(const uint64_t* __restrict__ pe_begin, uint32_t p, const uint32_t* __restrict__ perm,
                         const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                         const uint32_t* __restrict__ dist, const uint64_t* __restrict__ wprefix, int by_chars,
                         uint64_t factor, uint64_t max_len, uint32_t nr_max, Sample* __restrict__ samples,
                         uint32_t* __restrict__ counts) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= p) return;
    const uint64_t b = pe_begin[r], m = pe_begin[r + 1] - b;
    uint64_t nr = factor * (uint64_t)(p - 1);
    if (nr > m) nr = m;
    uint32_t cnt = 0;
    auto emit = [&](uint64_t idx) {
        const uint32_t id = perm[b + idx];
        const uint64_t ml = dist ? (uint64_t)dist[b + idx] : max_len;
        const uint32_t sl = len[id];
        Sample S;
        S.off = off[id];
        S.tlen = (uint32_t)(sl > ml ? ml : sl);
        S.pad = 0;
        S.gidx = b + idx;
        samples[(uint64_t)r * nr_max + cnt++] = S;
    };
    if (!by_chars) {
        const double sd = (double)m / (double)(nr + 1);
        for (uint64_t i = 1; i <= nr; ++i) emit((uint64_t)((double)i * sd));
    } else {
        const uint64_t base = wprefix[b];
        const uint64_t num = wprefix[b + m] - base;
        const uint64_t sd = num / (nr + 1);
        uint64_t si = 0;
        for (uint64_t i = 1; i <= nr && si < m; ++i) {
            if (sd == 0) {
                emit(si ? si - 1 : 0);
                continue;
            }
            // smallest t >= si with W(si..t) >= sd, else m - 1
            const uint64_t start = wprefix[b + si];
            uint64_t lo = si, hi = m;  // first t with wprefix[b+t+1]-start >= sd
            while (lo < hi) {
                const uint64_t mid = (lo + hi) / 2;
                if (wprefix[b + mid + 1] - start >= sd) hi = mid; else lo = mid + 1;
            }
            if (lo >= m) {
                si = m;
                emit(m - 1);
            } else {
                si = lo + 1;
                emit(lo);
            }
        }
    }
    counts[r] = cnt;
}

// Compacts the per-PE sample slots in PE order (one block; thread r copies
// PE r's samples after the samples of PEs 0..r-1).
__global__ void k_sample_compact(const Sample* __restrict__ slots, const uint32_t* __restrict__ counts, uint32_t p,
                                 uint32_t nr_max, Sample* __restrict__ out, uint32_t* __restrict__ total) {
    for (uint32_t r = threadIdx.x; r < p; r += blockDim.x) {
        uint32_t k = 0;
        for (uint32_t q = 0; q < r; ++q) k += counts[q];
        for (uint32_t i = 0; i < counts[r]; ++i) out[k + i] = slots[(uint64_t)r * nr_max + i];
        if (r == p - 1) *total = k + counts[r];
    }
}

// Rank of each sample in the global sample order (the result of the sample
// sort RQuick::sort, misc.hpp:503-514): by string, then (indexed samplers,
// IndexStringComparator misc.hpp:35-53) by global index, then by sample id.
// The last key makes the ranks a permutation: a NumChars walk over a PE with
// fewer characters than splitters emits its first string several times
// (identical samples, same string and index), and tied ranks would leave
// slots of `sorted` unwritten.  Identical samples are interchangeable, so
// the order among them changes no splitter.
// One block per sample; the launch may cover more blocks than the *ns_p
// samples the compaction produced (they return at once).
__global__ void __launch_bounds__(256) k_sample_rank(const Sample* __restrict__ smp, const uint32_t* __restrict__ ns_p,
                                                     int indexed, const uint8_t* __restrict__ chars,
                                                     uint32_t* __restrict__ sorted) {
    __shared__ uint32_t red[4];
    const uint32_t ns = *ns_p;
    const uint32_t i = blockIdx.x;
    if (i >= ns) return;  // block-uniform
    const Sample a = smp[i];
    uint32_t c = 0;
    for (uint32_t j = threadIdx.x; j < ns; j += blockDim.x) {
        if (j == i) continue;
        const Sample b = smp[j];
        const int r = view_cmp32(chars, b.off, b.tlen, a.off, a.tlen, nullptr);
        if (r < 0) ++c;
        else if (r == 0 && (indexed ? (b.gidx < a.gidx || (b.gidx == a.gidx && j < i)) : (j < i))) ++c;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if (lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) sorted[red[0] + red[1] + red[2] + red[3]] = i;
}

// Per (PE r, splitter q): binarySearch / binarySearchIndexed
// (misc.hpp:310-333, :347-372) and bucket sizes (compute_interval_binary(_
// index), :376-414).  One block per PE, p-1 threads.
__global__ void k_bucket_search(const uint64_t* __restrict__ pe_begin, uint32_t p, const uint32_t* __restrict__ perm,
                                const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                                const uint8_t* __restrict__ chars, const Sample* __restrict__ smp,
                                const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ ns_p, int indexed,
                                uint64_t* __restrict__ interval) {
    extern __shared__ uint64_t pos[];
    const uint32_t r = blockIdx.x;
    const uint64_t b = pe_begin[r], m = pe_begin[r + 1] - b;
    const uint32_t ns = *ns_p;
    const uint32_t nspl = (p - 1) < ns ? (p - 1) : ns;
    const uint32_t sdist = ns / (nspl + 1);
    for (uint32_t q = threadIdx.x; q + 1 < p; q += blockDim.x) {
        uint64_t res;
        if (q >= nspl) {
            res = m;
        } else {
            const Sample sp = smp[sorted[(q + 1) * sdist]];
            uint64_t left = 0, right = m;
            res = ~0ull;
            while (left != right) {
                const uint64_t d = (right - left) / 2;
                const uint32_t id = perm[b + left + d];
                int c = view_cmp32(chars, off[id], len[id], sp.off, sp.tlen, nullptr);
                if (c == 0 && indexed) {
                    const uint64_t gi = b + left + d;
                    c = gi < sp.gidx ? -1 : (gi > sp.gidx ? 1 : 0);
                }
                if (c < 0) left = left + d + 1;
                else if (c == 0) { res = left + d; break; }
                else right = left + d;
            }
            if (res == ~0ull) res = left;
        }
        pos[q] = res;
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < p; q += blockDim.x) {
        const uint64_t hi = (q + 1 < p) ? pos[q] : m;
        const uint64_t lo = q ? pos[q - 1] : 0;
        interval[(uint64_t)r * p + q] = hi - lo;
    }
}

// The same searches with one wave per (PE r, splitter q): every string
// comparison is a 512-character wave comparison (one dependent load instead
// of one per 8 characters).  pos[r * (p-1) + q] = boundary.
__global__ void __launch_bounds__(64)
k_bucket_search_wave(const uint64_t* __restrict__ pe_begin, uint32_t p, const uint32_t* __restrict__ perm,
                     const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                     const uint8_t* __restrict__ chars, const Sample* __restrict__ smp,
                     const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ ns_p, int indexed,
                     uint64_t* __restrict__ pos) {
    const uint32_t r = blockIdx.x / (p - 1), q = blockIdx.x % (p - 1);
    const uint64_t b = pe_begin[r], m = pe_begin[r + 1] - b;
    const uint32_t ns = *ns_p;
    const uint32_t nspl = (p - 1) < ns ? (p - 1) : ns;
    uint64_t res = m;
    if (q < nspl) {  // block-uniform
        const uint32_t sdist = ns / (nspl + 1);
        const Sample sp = smp[sorted[(q + 1) * sdist]];
        uint64_t left = 0, right = m;
        res = ~0ull;
        while (left != right) {
            const uint64_t d = (right - left) / 2;
            const uint32_t id = perm[b + left + d];
            int c = wave_cmp(chars, off[id], len[id], sp.off, sp.tlen);
            if (c == 0 && indexed) {
                const uint64_t gi = b + left + d;
                c = gi < sp.gidx ? -1 : (gi > sp.gidx ? 1 : 0);
            }
            if (c < 0) left = left + d + 1;
            else if (c == 0) {
                res = left + d;
                break;
            } else right = left + d;
        }
        if (res == ~0ull) res = left;
    }
    if (threadIdx.x == 0) pos[(uint64_t)r * (p - 1) + q] = res;
}

__global__ void k_intervals(const uint64_t* __restrict__ pe_begin, uint32_t p, const uint64_t* __restrict__ pos,
                            uint64_t* __restrict__ interval) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint64_t)p * p) return;
    const uint32_t r = (uint32_t)(k / p), q = (uint32_t)(k % p);
    const uint64_t m = pe_begin[r + 1] - pe_begin[r];
    const uint64_t* pr = pos + (uint64_t)r * (p - 1);
    const uint64_t hi = (q + 1 < p) ? pr[q] : m;
    const uint64_t lo = q ? pr[q - 1] : 0;
    interval[k] = hi - lo;
}

__global__ void k_weights(const uint32_t* __restrict__ s_len, const uint32_t* __restrict__ dist, uint64_t n,
                          uint64_t* __restrict__ w) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = dist ? dist[i] : s_len[i];
}

void partition(dss_set_s* s, int sampler, uint64_t factor, uint64_t max_len, bool use_dist) {
    DSS_REQUIRE(s->sorted, DSS_ESTATE, "dss_sort_lcp has not run");
    DSS_REQUIRE(sampler >= 0 && sampler <= 3, DSS_EINVAL, "sampler must be 0..3");
    DSS_REQUIRE(!use_dist || s->pp.has_dist, DSS_ESTATE, "prefix doubling has not run");
    DSS_REQUIRE(factor >= 1 && factor * (s->p - 1) <= (1u << 20), DSS_ERANGE, "sampling factor too large");
    DSS_REQUIRE(s->p <= 4096, DSS_ERANGE, "partition supports at most 4096 PEs");
    PipeState& P = s->pp;
    const hipStream_t st = s->stream;
    const uint32_t p = s->p;
    const int indexed = sampler >= 2;
    const int by_chars = sampler == 1 || sampler == 3;
    const uint32_t nr_max = std::max<uint32_t>((uint32_t)(factor * (p - 1)), 1);
    const uint64_t nsmax = (uint64_t)nr_max * p;
    P.samples.reserve(nsmax * 2 * sizeof(Sample) + 64);
    P.sample_sorted.reserve(nsmax * 4 + 64);
    P.sample_rank.reserve((p + 2) * 4);
    P.interval_d.reserve((uint64_t)p * p * 8);
    Sample* slots = P.samples.as<Sample>();
    Sample* compact = slots + nsmax;
    uint32_t* counts = P.sample_rank.as<uint32_t>();
    uint32_t* ns_d = counts + p;
    const uint32_t* dist = use_dist ? P.dist.as<uint32_t>() : nullptr;
    uint64_t* wprefix = nullptr;
    if (by_chars && s->n) {
        P.wprefix.reserve((s->n + 1) * 8);
        wprefix = P.wprefix.as<uint64_t>();
        // the indexed policy walks distinguishing prefixes, the plain one
        // lengths (samplingStrategies.hpp:243-273, :345-386)
        hipLaunchKernelGGL(k_weights, dim3((s->n + 255) / 256), dim3(256), 0, st, s->s_len.as<uint32_t>(),
                           indexed ? dist : nullptr, s->n, wprefix);
        uint64_t tot = 0;
        exclusive_scan_u64(wprefix, wprefix, s->n, st, P, &tot);
        hipLaunchKernelGGL(k_store_u64, dim3(1), dim3(1), 0, st, wprefix + s->n, tot);
    }
    {
        // algorithmic bytes: every possible sample's id, offset and length and
        // its characters (at most max_len, at most the mean string length),
        // the splitters' binary searches over every PE (log2 of its strings
        // probes of id, offset, length and an 8-character key each) and the
        // p x p interval counts
        const double mean_len = s->n ? (double)(s->nbytes - s->n) / (double)s->n : 0.0;
        const double probes = std::log2((double)s->n / std::max<uint32_t>(p, 1) + 1.0);
        const double part_bytes = (double)nsmax * (16.0 + std::min<double>((double)max_len, mean_len)) +
                                  (double)p * (p - 1) * probes * 24.0 + (double)p * p * 8.0;
        KScope ks(s->timer, st, "partition", p > 1 && s->n ? part_bytes : 0.0);
        if (p > 1 && s->n) {
            hipLaunchKernelGGL(k_sample, dim3((p + 63) / 64), dim3(64), 0, st, s->pe_begin_d.as<uint64_t>(), p,
                               s->perm.as<uint32_t>(), s->off.as<uint64_t>(), s->len.as<uint32_t>(), dist, wprefix,
                               by_chars, factor, max_len, nr_max, slots, counts);
            hipLaunchKernelGGL(k_sample_compact, dim3(1), dim3(256), 0, st, slots, counts, p, nr_max, compact, ns_d);
            // a block per sample: up to 65536 possible samples the grid covers
            // them all and needs no read-back of the count
            uint64_t rank_blocks = nsmax;
            if (nsmax > 65536) {
                uint32_t ns = 0;
                P.mbox.fetch(st, &ns, ns_d, 4);
                rank_blocks = ns;
            }
            if (rank_blocks)
                hipLaunchKernelGGL(k_sample_rank, dim3((uint32_t)rank_blocks), dim3(256), 0, st, compact, ns_d, indexed,
                                   s->arena(), P.sample_sorted.as<uint32_t>());
            if ((uint64_t)p * (p - 1) <= (1u << 16)) {
                P.splitter_pos.reserve((uint64_t)p * (p - 1) * 8);
                hipLaunchKernelGGL(k_bucket_search_wave, dim3(p * (p - 1)), dim3(64), 0, st,
                                   s->pe_begin_d.as<uint64_t>(), p, s->perm.as<uint32_t>(), s->off.as<uint64_t>(),
                                   s->len.as<uint32_t>(), s->arena(), compact, P.sample_sorted.as<uint32_t>(), ns_d,
                                   indexed, P.splitter_pos.as<uint64_t>());
                hipLaunchKernelGGL(k_intervals, dim3(((uint64_t)p * p + 255) / 256), dim3(256), 0, st,
                                   s->pe_begin_d.as<uint64_t>(), p, P.splitter_pos.as<uint64_t>(),
                                   P.interval_d.as<uint64_t>());
            } else {
                hipLaunchKernelGGL(k_bucket_search, dim3(p), dim3(std::min<uint32_t>(std::max<uint32_t>(p, 64), 256)),
                                   (size_t)p * 8, st, s->pe_begin_d.as<uint64_t>(), p, s->perm.as<uint32_t>(),
                                   s->off.as<uint64_t>(), s->len.as<uint32_t>(), s->arena(), compact,
                                   P.sample_sorted.as<uint32_t>(), ns_d, indexed, P.interval_d.as<uint64_t>());
            }
        } else {
            // one PE (or no strings): everything stays on its PE
            std::vector<uint64_t> iv((uint64_t)p * p, 0);
            for (uint32_t r = 0; r < p; ++r) iv[(uint64_t)r * p + r] = s->pe_begin[r + 1] - s->pe_begin[r];
            DSS_HIP(hipMemcpyAsync(P.interval_d.p, iv.data(), iv.size() * 8, hipMemcpyHostToDevice, st));
        }
    }
    DSS_HIP(hipGetLastError());
    P.interval.assign((uint64_t)p * p, 0);
    P.mbox.fetch(st, P.interval.data(), P.interval_d.p, (uint64_t)p * p * 8);
    P.partitioned = true;
    P.exchanged = P.merged = P.extended = false;
}

// ================================================================== (a6, a13)
// Layout (host side, p^2 values): send_start[r*p+q] = first position of
// bucket q inside PE r; run_off[q*p+r] = receive position of run r of PE q.
__device__ inline uint32_t bucket_of(const uint64_t* __restrict__ send_start, uint32_t p, uint32_t r, uint64_t j) {
    const uint64_t* ss = send_start + (uint64_t)r * p;
    uint32_t lo = 0, hi = p;  // first q with ss[q] > j
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (ss[mid] <= j) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

// Where source position i goes: its PE r, its position j inside r, its
// bucket q, the bucket's first position st inside r and its receive
// position dst (bucket q of PE r lands as run r of receiver q).
struct ExchPos {
    uint32_t r, q;
    uint64_t j, st, dst;
};
__device__ inline ExchPos exch_pos(const uint64_t* __restrict__ pe_begin, const uint64_t* __restrict__ send_start,
                                   const uint64_t* __restrict__ run_off, uint32_t p, uint64_t i) {
    uint32_t lo = 0, hi = p;  // PE of position i: last r with pe_begin[r] <= i (non-empty)
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (pe_begin[mid] <= i) lo = mid + 1; else hi = mid;
    }
    ExchPos e;
    e.r = lo - 1;
    e.j = i - pe_begin[e.r];
    e.q = bucket_of(send_start, p, e.r, e.j);
    e.st = send_start[(uint64_t)e.r * p + e.q];
    e.dst = run_off[(uint64_t)e.q * p + e.r] + (e.j - e.st);
    return e;
}

// Origin (source PE, position inside it) of every received string, computed
// from the exchange layout when a caller asks for it; the exchange itself
// does not write it.
__global__ void k_exch_origin(const uint64_t* __restrict__ pe_begin, uint32_t p,
                              const uint64_t* __restrict__ send_start, const uint64_t* __restrict__ run_off,
                              uint64_t n, uint32_t* __restrict__ r_pe, uint32_t* __restrict__ r_pos) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ExchPos e = exch_pos(pe_begin, send_start, run_off, p, i);
    r_pe[e.dst] = e.r;
    r_pos[e.dst] = (uint32_t)e.j;
}

void recv_origin(dss_set_s* s) {
    PipeState& P = s->pp;
    DSS_REQUIRE(P.exchanged, DSS_ESTATE, "dss_exchange has not run");
    if (P.origin_valid) return;
    const uint64_t n = s->n;
    const uint32_t p = s->p;
    // both arrays were reserved by the exchange: no allocation here
    if (n) {
        const uint64_t* ls = P.layout.as<uint64_t>();
        hipLaunchKernelGGL(k_exch_origin, dim3((n + 255) / 256), dim3(256), 0, s->stream,
                           s->pe_begin_d.as<uint64_t>(), p, ls, ls + (uint64_t)p * p, n,
                           P.recv_src_pe.as<uint32_t>(), P.recv_src_pos.as<uint32_t>());
        DSS_HIP(hipGetLastError());
    }
    P.origin_valid = true;
}

constexpr uint32_t kXiP = 32;  // PEs up to which k_exch_index stages its tables

// One thread per source position: destination, size and receive metadata.
// Encoders (src/mpi/byte_encoder.hpp): mode 0 whole string incl. NUL
// (:81-107), mode 1 chars [lcp, len] incl. NUL (:109-137), mode 2 chars
// [lcp, dist) then NUL (:139-172).  The LCP at every bucket start is zero
// (setLcpAtStartOfInterval, src/mpi/alltoall.hpp:396-404).
__global__ void k_exch_index(const uint64_t* __restrict__ pe_begin, uint32_t p, const uint32_t* __restrict__ lcp,
                             const uint64_t* __restrict__ s_off, const uint32_t* __restrict__ s_len,
                             const uint32_t* __restrict__ dist,
                             const uint64_t* __restrict__ send_start, const uint64_t* __restrict__ run_off, int mode,
                             uint64_t n, uint64_t* __restrict__ size, uint32_t* __restrict__ r_lcp,
                             uint64_t* __restrict__ r_voff,
                             uint32_t* __restrict__ r_vlen, uint32_t* __restrict__ r_from,
                             const uint32_t* __restrict__ sdepth, uint32_t* __restrict__ r_stage) {
    // up to kXiP PEs the PE begins and the layout tables are staged in LDS,
    // so the two searches per position read LDS instead of global memory
    __shared__ uint64_t s_pe[kXiP + 1], s_ss[kXiP * kXiP], s_ro[kXiP * kXiP];
    const bool staged_tables = p <= kXiP;  // block-uniform
    if (staged_tables) {
        for (uint32_t t = threadIdx.x; t <= p; t += blockDim.x) s_pe[t] = pe_begin[t];
        for (uint32_t t = threadIdx.x; t < p * p; t += blockDim.x) {
            s_ss[t] = send_start[t];
            s_ro[t] = run_off[t];
        }
        __syncthreads();
    }
    const uint64_t* const PB = staged_tables ? s_pe : pe_begin;
    const uint64_t* const SS = staged_tables ? s_ss : send_start;
    const uint64_t* const RO = staged_tables ? s_ro : run_off;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ExchPos e = exch_pos(PB, SS, RO, p, i);
    const uint64_t j = e.j, st = e.st, dst = e.dst;
    const uint32_t l = (j == st) ? 0u : lcp[i];
    const uint32_t L = s_len[i];
    uint64_t sz;
    uint32_t from = 0, vlen = L;
    if (mode == 0 || mode == 3) {
        sz = (uint64_t)L + 1;
    } else if (mode == 1) {
        from = l;
        sz = (uint64_t)L + 1 - l;
    } else {
        from = l;
        vlen = dist[i];
        sz = (uint64_t)vlen - l + 1;
    }
    size[dst] = sz;
    r_lcp[dst] = mode == 3 && j != st ? kNone : l;  // mode 3 sends no LCPs
    r_voff[dst] = s_off[i];
    r_vlen[dst] = vlen;
    r_from[dst] = from;
    // mode 2: the prefix doubling kept [32(vlen/32-1), vlen) of this string
    if (r_stage) r_stage[dst] = (sdepth[i] == vlen && vlen >= 64) ? (uint32_t)i : kNone;
}

// Common prefix of all strings a PE received: min over its runs r of the
// LCP of the PE's first string with the first and last string of run r
// (one wave per receiver and run; 512 characters per wave step).
__global__ void k_recv_cp(const uint64_t* __restrict__ run_off, const uint64_t* __restrict__ interval, uint32_t p,
                          const uint64_t* __restrict__ recv_begin, const uint8_t* __restrict__ chars,
                          const uint64_t* __restrict__ voff, const uint32_t* __restrict__ vlen,
                          uint32_t* __restrict__ cp_out) {
    const uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // wave-uniform
    if (t >= (uint64_t)p * p) return;
    const uint32_t q = (uint32_t)(t / p), r = (uint32_t)(t % p);
    const uint64_t c = interval[(uint64_t)r * p + q];
    if (!c) return;
    const uint64_t s0 = recv_begin[q];
    const uint64_t f = run_off[(uint64_t)q * p + r], l = f + c - 1;
    const uint32_t x = wave_lcp(chars, voff[s0], vlen[s0], voff[f], vlen[f]);
    const uint32_t y = wave_lcp(chars, voff[s0], vlen[s0], voff[l], vlen[l]);
    if (lane_id() == 0) atomicMin(&cp_out[q], min(x, y));
}

// Receiving PE of receive position k.
__device__ inline uint32_t receiver_of(const uint64_t* __restrict__ recv_begin, uint32_t p, uint64_t k) {
    uint32_t lo = 0, hi = p;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (recv_begin[mid] <= k) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

constexpr uint32_t kXcWords = 512;  // LDS words per wave for the short-copy path
constexpr uint32_t kXcP = 64;        // receivers up to which the copy stages its tables

// Short copies (at most 33 characters on average, such as the few
// characters after the LCP that an LCP-compressed distinguishing prefix
// sends): every lane copies one received string into the receive arena, 32
// characters at a time from five aligned 8-byte loads, so a wave has 64
// strings in flight.  The lane also stores the merge's cached keys: the
// string's 16 characters at its receiver's common prefix cp (two words),
// from the lines the copy reads.
__global__ void k_exch_copy(const uint8_t* __restrict__ chars, const uint64_t* __restrict__ r_voff,
                            const uint32_t* __restrict__ r_vlen, const uint32_t* __restrict__ r_from,
                            const uint64_t* __restrict__ byte_off, uint64_t n,
                            uint8_t* __restrict__ out, const uint64_t* __restrict__ recv_begin, uint32_t p,
                            const uint32_t* __restrict__ cp, uint64_t* __restrict__ keyc,
                            uint64_t* __restrict__ keyc2,
                            const uint32_t* __restrict__ r_stage, const uint64_t* __restrict__ stage,
                            uint32_t stage_w) {
    // up to kXcP receivers their first receive positions and common prefixes
    // are staged in LDS for the receiver search of every string
    __shared__ uint64_t s_rb[kXcP + 1];
    __shared__ uint32_t s_cp[kXcP];
    const bool staged_tab = p <= kXcP;  // grid-uniform
    if (staged_tab) {
        for (uint32_t t = threadIdx.x; t <= p; t += blockDim.x) s_rb[t] = recv_begin[t];
        for (uint32_t t = threadIdx.x; t < p; t += blockDim.x) s_cp[t] = cp[t];
        __syncthreads();
    }
    const uint64_t* const RB = staged_tab ? s_rb : recv_begin;
    const uint32_t* const CP = staged_tab ? s_cp : cp;
    {
        // A wave's 64 consecutive receive positions own one contiguous byte
        // range of the receive arena.  When it fits, the lanes build it in
        // LDS and the wave stores it as aligned words (byte stores only for
        // the two edge words, which neighbouring waves share); otherwise the
        // lanes store their bytes directly.
        __shared__ uint64_t xw[4][kXcWords];
        uint64_t* const sw = xw[(threadIdx.x >> 6) & 3];
        uint8_t* const sb = reinterpret_cast<uint8_t*>(sw);
        const uint32_t lane = threadIdx.x & 63;
        const uint64_t nwv = (n + 63) / 64;
        for (uint64_t wv = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; wv < nwv;
             wv += ((uint64_t)gridDim.x * blockDim.x) >> 6) {  // wave-uniform
            const uint64_t k = wv * 64 + lane;
            const bool valid = k < n;
            const uint32_t last = (uint32_t)min<uint64_t>(63, n - 1 - wv * 64);
            uint32_t si = kNone, vlen = 0, from = 0, cpk = 0;
            uint64_t start = 0;
            if (valid) {
                cpk = CP[receiver_of(RB, p, k)];
                si = r_stage ? r_stage[k] : kNone;
                vlen = r_vlen[k];
                from = r_from[k];
                start = byte_off[k];
            }
            const uint32_t m = vlen - from;  // characters; a NUL follows
            const uint64_t A = __shfl(start, 0);
            const uint64_t E = __shfl(start + m + 1, last);
            const uint64_t A0 = A & ~7ull;
            const bool fits = E - A0 <= 8ull * kXcWords;  // wave-uniform
            uint8_t* const gdst = out + start;  // strings of a wave that does not fit: byte stores
            if (fits) {
                // the wave's buffer words start at zero (the NULs and any
                // padding), every string ORs its characters in
                const uint32_t nwords = (uint32_t)((E - A0 + 7) / 8);
                for (uint32_t w = lane; w < nwords; w += 64) sw[w] = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
            // ORs bytes [lo, hi) of the word W, whose byte 0 belongs at buffer
            // byte db (< 0 for bytes before this string's), into the buffer:
            // two word ORs instead of a store per byte
            auto or_word = [&](int64_t db, uint64_t W, int lo, int hi) {
                if (lo >= hi) return;
                W &= (hi - lo == 8 ? ~0ull : ((1ull << (8 * (hi - lo))) - 1)) << (8 * lo);
                const int64_t q = db >> 3;
                const unsigned sh = (unsigned)(db & 7) * 8;
                const uint64_t wl = W << sh, wh = sh ? W >> (64 - sh) : 0ull;
                if (wl) atomicOr(reinterpret_cast<unsigned long long*>(sw + q), (unsigned long long)wl);
                if (wh) atomicOr(reinterpret_cast<unsigned long long*>(sw + q + 1), (unsigned long long)wh);
            };
            const int64_t d0 = (int64_t)(start - A0);  // this string's first buffer byte
            if (valid) {
                bool done = false;
                if (si != kNone) {
                    // the characters were kept by the prefix doubling's last
                    // hash of this string: row = [B, dist), B = 32 (dist/32 -
                    // 1), in stage_w words (the characters beyond dist are
                    // not used)
                    const uint32_t B = 32 * (vlen / 32 - 1);
                    if (from >= B && cpk >= B) {
                        const uint64_t* rw = stage + stage_w * (uint64_t)si;
                        uint64_t row[9];
#pragma unroll
                        for (uint32_t i = 0; i < 9; ++i) row[i] = i < stage_w ? rw[i] : 0;
                        // 8 characters at row byte x, little endian
                        auto rword = [&](uint32_t x) -> uint64_t {
                            const uint32_t w = x >> 3, sh = (x & 7) * 8;
                            return sh ? (row[w] >> sh) | (row[w + 1] << (64 - sh)) : row[w];
                        };
                        uint64_t key = 0;
                        if (cpk < vlen) {
                            key = __builtin_bswap64(rword(cpk - B));
                            const uint32_t rem = vlen - cpk;
                            if (rem < 8) key &= ~0ull << (8 * (8 - rem));
                        }
                        keyc[k] = key;
                        uint64_t key2 = 0;  // the next 8 characters
                        if (cpk + 8 < vlen) {
                            key2 = __builtin_bswap64(rword(cpk + 8 - B));
                            const uint32_t rem = vlen - cpk - 8;
                            if (rem < 8) key2 &= ~0ull << (8 * (8 - rem));
                        }
                        keyc2[k] = key2;
                        // row bytes [x0, x1) are the characters to copy
                        const int x0 = (int)(from - B), x1 = x0 + (int)m;
                        if (fits) {
#pragma unroll
                            for (int j = 0; j < 9; ++j)
                                if (j < (int)stage_w)
                                    or_word(d0 - x0 + 8 * j, row[j], max(0, x0 - 8 * j), min(8, x1 - 8 * j));
                        } else {
                            for (uint32_t t = 0; t < m; t += 8) {
                                const uint64_t y = rword(from - B + t);
#pragma unroll
                                for (uint32_t b = 0; b < 8; ++b)
                                    if (t + b < m) gdst[t + b] = (uint8_t)(y >> (8 * b));
                            }
                        }
                        done = true;
                    }
                }
                if (!done) {
                    keyc[k] = load_key8(chars, r_voff[k], vlen, cpk);
                    keyc2[k] = load_key8(chars, r_voff[k], vlen, cpk + 8);
                    const uint64_t a = r_voff[k] + from;
                    if (fits) {
                        // the arena's aligned words from the one holding
                        // character `from`, four at a time
                        const uint64_t* w = reinterpret_cast<const uint64_t*>(chars + (a & ~7ull));
                        const int r = (int)(a & 7), end = r + (int)m;
                        for (int i = 0; i < end; i += 32) {
                            uint64_t x[4];
#pragma unroll
                            for (int j = 0; j < 4; ++j) x[j] = i + 8 * j < end ? w[(i >> 3) + j] : 0ull;
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int o = i + 8 * j;
                                or_word(d0 - r + o, x[j], max(0, r - o), min(8, end - o));
                            }
                        }
                    } else {
                        for (uint32_t t = 0; t < m; t += 32) {
                            const uint64_t* w = reinterpret_cast<const uint64_t*>(chars + ((a + t) & ~7ull));
                            const unsigned sh = (unsigned)((a + t) & 7) * 8;
                            uint64_t x[5], y[4];
#pragma unroll
                            for (int i = 0; i < 5; ++i) x[i] = w[i];
#pragma unroll
                            for (int i = 0; i < 4; ++i) y[i] = sh ? (x[i] >> sh) | (x[i + 1] << (64 - sh)) : x[i];
#pragma unroll
                            for (uint32_t b = 0; b < 32; ++b)
                                if (t + b < m) gdst[t + b] = (uint8_t)(y[b >> 3] >> (8 * (b & 7)));
                        }
                    }
                }
                if (!fits) gdst[m] = 0;
            }
            if (fits) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                // word w of the buffer is the arena word at A0 + 8 w
                const uint32_t nwords = (uint32_t)((E - A0 + 7) / 8);
                for (uint32_t w = lane; w < nwords; w += 64) {
                    const uint64_t g = A0 + 8ull * w;
                    if (g >= A && g + 8 <= E) {
                        reinterpret_cast<uint64_t*>(out)[g >> 3] = sw[w];
                    } else {
                        for (uint32_t b = 0; b < 8; ++b)
                            if (g + b >= A && g + b < E) out[g + b] = sb[8 * w + b];
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();  // read before the next range is built
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
        }
    }
}

// Longer copies, balanced over output words: wave c builds the receive
// arena's bytes [kXbChunk c, kXbChunk (c + 1)), lane l its words l + 64 j,
// so every store instruction writes 512 contiguous bytes and neighbouring
// lanes read neighbouring source words.  A word's owner is the last string
// that starts at or before its first byte (a max-scan over the chunk's
// words of where each string starts); the word holds that string's
// characters from there, its NUL, then the heads of the strings after it.
// The lane that walks a string whose start lies in the chunk also stores
// its 16 cached merge characters at its receiver's common prefix.
constexpr uint32_t kXbChunk = 4096;            // output bytes per wave
constexpr uint32_t kXbWords = kXbChunk / 8;    // output words per wave
constexpr uint32_t kXbPer = kXbWords / 64;     // output words per lane
constexpr uint32_t kXbMeta = 128;              // strings per chunk whose layout is staged in LDS

// chunk_first[c]: the string that holds output byte kXbChunk c.
__global__ void k_copy_chunks(const uint64_t* __restrict__ byte_off, const uint32_t* __restrict__ r_vlen,
                              const uint32_t* __restrict__ r_from, uint64_t n, uint32_t* __restrict__ chunk_first) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t o = byte_off[k], e = o + (r_vlen[k] - r_from[k]) + 1;
        for (uint64_t c = (o + kXbChunk - 1) / kXbChunk; c * kXbChunk < e; ++c) chunk_first[c] = (uint32_t)k;
    }
}

__global__ void __launch_bounds__(256)
k_exch_copy_bal(const uint8_t* __restrict__ chars, const uint64_t* __restrict__ r_voff,
                const uint32_t* __restrict__ r_vlen, const uint32_t* __restrict__ r_from,
                const uint64_t* __restrict__ byte_off, uint64_t n, uint64_t total,
                const uint32_t* __restrict__ chunk_first, uint64_t nchunks, uint8_t* __restrict__ out,
                const uint64_t* __restrict__ recv_begin, uint32_t p, const uint32_t* __restrict__ cp,
                uint64_t* __restrict__ keyc, uint64_t* __restrict__ keyc2) {
    __shared__ uint64_t s_rb[kXcP + 1];
    __shared__ uint32_t s_cp[kXcP];
    __shared__ uint32_t mark_all[4][kXbWords];
    __shared__ uint64_t mo_all[4][kXbMeta], ms_all[4][kXbMeta];  // output start, source start
    __shared__ uint32_t mm_all[4][kXbMeta];                      // characters sent
    const bool staged_tab = p <= kXcP;  // grid-uniform
    if (staged_tab) {
        for (uint32_t t = threadIdx.x; t <= p; t += blockDim.x) s_rb[t] = recv_begin[t];
        for (uint32_t t = threadIdx.x; t < p; t += blockDim.x) s_cp[t] = cp[t];
        __syncthreads();
    }
    const uint64_t* const RB = staged_tab ? s_rb : recv_begin;
    const uint32_t* const CP = staged_tab ? s_cp : cp;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wv = (threadIdx.x >> 6) & 3;
    uint32_t* const mk = mark_all[wv];
    uint64_t* const mo = mo_all[wv];
    uint64_t* const msrc = ms_all[wv];
    uint32_t* const mm = mm_all[wv];
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < nchunks; c += waves) {
        const uint64_t A = c * kXbChunk, E = min(A + kXbChunk, total);
        const uint32_t k0 = chunk_first[c];
        uint32_t k1 = (uint32_t)(n - 1);  // the last string that starts before E
        if (c + 1 < nchunks) {
            k1 = chunk_first[c + 1];
            if (byte_off[k1] >= E) --k1;
        }
        // mk[w] = the last string (relative to k0) that starts at or before
        // byte A + 8 w: marks where each string starts, then a max-scan
        for (uint32_t w = lane; w < kXbWords; w += 64) mk[w] = 0;
        // the layout of the chunk's first kXbMeta strings
        for (uint32_t r = lane; r < kXbMeta && k0 + r <= k1; r += 64) {
            const uint32_t k = k0 + r, f = r_from[k];
            mo[r] = byte_off[k];
            mm[r] = r_vlen[k] - f;
            msrc[r] = r_voff[k] + f;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t k = k0 + 1 + lane; k <= k1; k += 64) {
            const uint64_t ws = ((k - k0 < kXbMeta ? mo[k - k0] : byte_off[k]) - A + 7) / 8;
            if (ws < kXbWords) atomicMax(&mk[ws], k - k0);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        uint32_t loc[kXbPer];
#pragma unroll
        for (uint32_t i = 0; i < kXbPer; ++i) {
            loc[i] = mk[kXbPer * lane + i];
            if (i) loc[i] = max(loc[i], loc[i - 1]);
        }
        uint32_t incl = loc[kXbPer - 1];
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= o) incl = max(incl, y);
        }
        uint32_t excl = __shfl_up(incl, 1, 64);
        if (lane == 0) excl = 0;
#pragma unroll
        for (uint32_t i = 0; i < kXbPer; ++i) mk[kXbPer * lane + i] = max(loc[i], excl);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        // the cached merge keys of the strings that start in this chunk
        for (uint32_t k = k0 + lane; k <= k1; k += 64) {
            if (byte_off[k] < A) continue;  // k0 started in an earlier chunk
            const uint32_t cpk = CP[receiver_of(RB, p, k)];
            const uint32_t vlen = r_vlen[k];
            keyc[k] = load_key8(chars, r_voff[k], vlen, cpk);
            keyc2[k] = load_key8(chars, r_voff[k], vlen, cpk + 8);
        }
        // the words: characters of the owner from the word's first byte on,
        // the owner's NUL, then the strings after it
        const uint64_t nw = (E - A + 7) / 8;
#pragma unroll
        for (uint32_t j = 0; j < kXbPer; ++j) {
            const uint32_t w = lane + 64 * j;
            if (w >= nw) continue;
            const uint64_t g = A + 8ull * w;
            uint64_t k = k0 + mk[w];
            uint64_t v = 0;
            uint32_t filled = 0;
            // string k's layout: from LDS for the chunk's first kXbMeta strings
            auto layout = [&](uint64_t kk, uint64_t& o, uint32_t& m, uint64_t& src) {
                const uint64_t r = kk - k0;
                if (r < kXbMeta) {
                    o = mo[r];
                    m = mm[r];
                    src = msrc[r];
                } else {
                    const uint32_t f = r_from[kk];
                    o = byte_off[kk];
                    m = r_vlen[kk] - f;
                    src = r_voff[kk] + f;
                }
            };
            uint64_t o, src;
            uint32_t m;
            layout(k, o, m, src);
            // the owner's part (every word of a long string's middle ends
            // here), then the strings that follow while the word has room
            uint32_t at = (uint32_t)(g - o);  // output byte of k; byte m is its NUL
            uint32_t take = min(8u, m + 1 - at);
            uint32_t nch = at < m ? min(take, m - at) : 0u;  // characters before the NUL
            if (nch) {
                v = load_word_le(chars, src + at);
                if (nch < 8) v &= (1ull << (8 * nch)) - 1;
            }
            filled = take;
            while (filled < 8 && g + filled < total) {
                layout(++k, o, m, src);
                take = min(8u - filled, m + 1);
                nch = min(take, m);
                if (nch) {
                    uint64_t x = load_word_le(chars, src);
                    if (nch < 8) x &= (1ull << (8 * nch)) - 1;
                    v |= x << (8 * filled);
                }
                filled += take;
            }
            *reinterpret_cast<uint64_t*>(out + g) = v;
        }
    }
}

// Mode 3 (AllToAllStringImpl<..., NoLcps>, src/mpi/alltoall.hpp:567-640)
// sends whole strings and no LCPs; the receiver derives the LCP of every
// received string with its predecessor in the same run from the receive
// arena (8 characters per step), so the LCP merge can order the runs.
__global__ void k_recv_lcp(const uint8_t* __restrict__ bytes, const uint64_t* __restrict__ byte_off,
                           const uint32_t* __restrict__ vlen, uint64_t n, uint32_t* __restrict__ r_lcp) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        if (r_lcp[k] != kNone) continue;  // a run start (LCP 0)
        const uint64_t a = byte_off[k - 1], b = byte_off[k];
        const uint32_t m = min(vlen[k - 1], vlen[k]);
        uint32_t d = 0;
        while (d < m) {
            const uint64_t x = load_word_le(bytes, a + d) ^ load_word_le(bytes, b + d);
            if (x) {
                d += __builtin_ctzll(x) / 8;
                break;
            }
            d += 8;
        }
        r_lcp[k] = min(d, m);
    }
}

__global__ void k_byte_begin(const uint64_t* __restrict__ recv_begin, const uint64_t* __restrict__ byte_off,
                             uint64_t n, uint64_t total, uint32_t p, uint64_t* __restrict__ out) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > p) return;
    const uint64_t rb = recv_begin[q];
    out[q] = rb < n ? byte_off[rb] : total;
}

// p + 1 byte begins of the receivers, after the layout's 2 p^2 values and p + 1
// spare slots.
// Receivers' first receive positions (p+1) on the device, in the merge's
// common-prefix buffer after the p+1 common prefixes.
static inline uint64_t* recv_begin_dev(PipeState& P, uint32_t p) {
    return reinterpret_cast<uint64_t*>(P.m_cp.as<uint8_t>() + ((uint64_t)(p + 1) * 4 + 7) / 8 * 8);
}

static inline uint64_t* byte_begin_dev(PipeState& P, uint32_t p) {
    return P.layout.as<uint64_t>() + (uint64_t)2 * p * p + (p + 1);
}

void exchange(dss_set_s* s, int mode) {
    PipeState& P = s->pp;
    DSS_REQUIRE(P.partitioned, DSS_ESTATE, "dss_partition has not run");
    DSS_REQUIRE(mode >= 0 && mode <= 3, DSS_EINVAL, "mode must be 0..3");
    DSS_REQUIRE(mode != 2 || P.has_dist, DSS_ESTATE, "prefix doubling has not run");
    const hipStream_t st = s->stream;
    const uint32_t p = s->p;
    const uint64_t n = s->n;
    // layout on the host (p^2 numbers, like the reference's alltoall of counts)
    std::vector<uint64_t> lay((uint64_t)2 * p * p, 0);
    uint64_t* send_start = lay.data();
    uint64_t* run_off = lay.data() + (uint64_t)p * p;
    P.recv_begin.assign(p + 1, 0);
    for (uint32_t r = 0; r < p; ++r) {
        uint64_t acc = 0;
        for (uint32_t q = 0; q < p; ++q) {
            send_start[(uint64_t)r * p + q] = acc;
            acc += P.interval[(uint64_t)r * p + q];
        }
    }
    uint64_t acc = 0;
    for (uint32_t q = 0; q < p; ++q) {
        P.recv_begin[q] = acc;
        for (uint32_t r = 0; r < p; ++r) {
            run_off[(uint64_t)q * p + r] = acc;
            acc += P.interval[(uint64_t)r * p + q];
        }
    }
    P.recv_begin[p] = acc;
    DSS_REQUIRE(acc == n, DSS_EINVAL, "interval sizes do not cover all strings");
    P.layout.reserve(lay.size() * 8 + (uint64_t)2 * (p + 1) * 8);
    DSS_HIP(hipMemcpyAsync(P.layout.p, lay.data(), lay.size() * 8, hipMemcpyHostToDevice, st));
    const uint64_t nn = std::max<uint64_t>(n, 1);
    P.recv_size.reserve(nn * 8);
    P.recv_byte_off.reserve((nn + 1) * 8);
    P.recv_lcp.reserve(nn * 4);
    P.recv_view_off.reserve(nn * 8);
    P.recv_view_len.reserve(nn * 4);
    P.recv_from.reserve(nn * 4);
    // the origin arrays are filled only on request (recv_origin), but their
    // memory is reserved here so that a later query allocates nothing
    P.recv_src_pe.reserve(nn * 4);
    P.recv_src_pos.reserve(nn * 4);
    const bool staged = mode == 2 && P.stage_valid;
    if (staged) P.recv_stage.reserve(nn * 4);
    uint32_t* r_stage = staged ? P.recv_stage.as<uint32_t>() : nullptr;
    const uint64_t* ls = P.layout.as<uint64_t>();
    if (n) {
        // reads lcp, offset, length (+ dist in mode 2), writes size, lcp,
        // view offset, length and start (+ the kept-row index)
        KScope ks(s->timer, st, "exchange_index", n * ((mode == 2 ? 48.0 : 44.0) + (staged ? 8.0 : 0.0)));
        hipLaunchKernelGGL(k_exch_index, dim3((n + 255) / 256), dim3(256), 0, st, s->pe_begin_d.as<uint64_t>(), p,
                           s->lcp.as<uint32_t>(), s->s_off.as<uint64_t>(), s->s_len.as<uint32_t>(),
                           mode == 2 ? P.dist.as<uint32_t>() : nullptr, ls,
                           ls + (uint64_t)p * p, mode, n, P.recv_size.as<uint64_t>(), P.recv_lcp.as<uint32_t>(),
                           P.recv_view_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(),
                           P.recv_from.as<uint32_t>(), staged ? P.pd_sdepth.as<uint32_t>() : nullptr, r_stage);
    }
    // the merge's per-receiver common prefix (from the views, before the copy)
    P.m_keyc.reserve(nn * 8);
    P.m_keyc2.reserve(nn * 8);
    const uint64_t cp_bytes = ((uint64_t)(p + 1) * 4 + 7) / 8 * 8;
    P.m_cp.reserve(cp_bytes + (p + 1) * 8);
    uint32_t* cpd = P.m_cp.as<uint32_t>();
    uint64_t* rb_d = recv_begin_dev(P, p);
    DSS_HIP(hipMemcpyAsync(rb_d, P.recv_begin.data(), (p + 1) * 8, hipMemcpyHostToDevice, st));
    fill_small(st, cpd, 0xFF, (p + 1) * 4);
    if (n) {
        KScope ks(s->timer, st, "exchange_index", (double)p * p * 64.0);
        hipLaunchKernelGGL(k_recv_cp, dim3(((uint64_t)p * p * 64 + 255) / 256), dim3(256), 0, st, ls + (uint64_t)p * p,
                           P.interval_d.as<uint64_t>(), p, rb_d, s->arena(), P.recv_view_off.as<uint64_t>(),
                           P.recv_view_len.as<uint32_t>(), cpd);
    }
    uint64_t total = 0;
    exclusive_scan_u64(P.recv_size.as<uint64_t>(), P.recv_byte_off.as<uint64_t>(), n, st, P, &total);
    P.recv_nbytes = total;
    P.recv_bytes.reserve(total + kArenaPad);
    fill_small(st, P.recv_bytes.as<uint8_t>() + total, 0, kArenaPad);
    if (n) {
        KScope ks(s->timer, st, "exchange_copy", 2.0 * total + n * 40.0);
        // a lane per string for copies of at most 32 characters on average
        // (the wave assembles its range in LDS), otherwise waves over 4 KiB
        // chunks of the receive arena, a lane per output word
        if (total <= 33 * n) {
            hipLaunchKernelGGL(k_exch_copy, dim3(stride_grid(n, 256)), dim3(256), 0, st, s->arena(),
                               P.recv_view_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(),
                               P.recv_from.as<uint32_t>(), P.recv_byte_off.as<uint64_t>(), n,
                               P.recv_bytes.as<uint8_t>(), rb_d, p, cpd, P.m_keyc.as<uint64_t>(),
                               P.m_keyc2.as<uint64_t>(), r_stage, staged ? P.pd_stage.as<uint64_t>() : nullptr,
                               P.stage_w);
        } else {
            const uint64_t nchunks = (total + kXbChunk - 1) / kXbChunk;
            P.copy_chunks.reserve(nchunks * 4);
            hipLaunchKernelGGL(k_copy_chunks, dim3(stride_grid(n, 256)), dim3(256), 0, st,
                               P.recv_byte_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(),
                               P.recv_from.as<uint32_t>(), n, P.copy_chunks.as<uint32_t>());
            hipLaunchKernelGGL(k_exch_copy_bal, dim3(stride_grid(nchunks * 64, 256)), dim3(256), 0, st, s->arena(),
                               P.recv_view_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(),
                               P.recv_from.as<uint32_t>(), P.recv_byte_off.as<uint64_t>(), n, total,
                               P.copy_chunks.as<uint32_t>(), nchunks, P.recv_bytes.as<uint8_t>(), rb_d, p, cpd,
                               P.m_keyc.as<uint64_t>(), P.m_keyc2.as<uint64_t>());
        }
    }
    if (n && mode == 3) {
        KScope ks(s->timer, st, "exchange_lcp", 2.0 * total + n * 20.0);
        hipLaunchKernelGGL(k_recv_lcp, dim3(stride_grid(n, 256)), dim3(256), 0, st, P.recv_bytes.as<uint8_t>(),
                           P.recv_byte_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(), n,
                           P.recv_lcp.as<uint32_t>());
    }
    DSS_HIP(hipGetLastError());
    // byte_begin per receiver (the byte offset of its first received string),
    // on the device; dss_exchange reads it back when the caller asks for it
    hipLaunchKernelGGL(k_byte_begin, dim3((p + 1 + 255) / 256), dim3(256), 0, st, rb_d, P.recv_byte_off.as<uint64_t>(),
                       n, total, p, byte_begin_dev(P, p));
    P.mode = mode;
    P.exchanged = true;
    P.origin_valid = false;
    P.merged = P.extended = false;
}

// ================================================================== (a8)
// K-way LCP merge of the runs every PE received: choose_merge / merge<K>
// with LcpStringLoserTree_ (src/sorter/distributed/merging.hpp:23-123,
// src/merge/bingmann-lcp_losertree.hpp:204-480), K = the merge fan-in F, a
// runtime parameter: F = 2 runs the pairwise merge tree further below (one
// global pass per tree level), F = 4, 8 or 16 the K-way passes here; a PE
// with more runs than F merges them in ceil(log_F runs) passes.  Both give
// the same output (ties: the earlier run first).
//
// A K-way pass merges groups of at most F adjacent lists (a receiver's runs, or a
// previous pass's outputs) with one read and one write of every element:
//   1. samples: every S-th element of every list.  F lanes per sample find
//      its co-rank (the number of elements before it) in each list of its
//      group by binary search; their sum is its rank in the group's merged
//      order.  The order is by string, equal strings by list (an earlier
//      list first: on equality the loser tree keeps its defender).  With
//      S = floor((G - 1) / lists) + 1 consecutive sample ranks differ by at
//      most G (between two consecutive samples every list holds fewer than S
//      elements).
//   2. tiles: tile k of a group starts at B_k, the largest sample rank
//      <= k G (B_0 = 0: a list's first element is a sample), so a tile holds
//      fewer than 2G elements; the co-ranks of its two bounding samples give
//      every list's part of the tile.
//   3. one block per tile stages its elements in LDS (receive index, the
//      cached 8 characters at the receiver's common prefix, the LCP with the
//      list predecessor) and runs the tournament's levels in LDS: level r
//      merges adjacent pairs of 2^r-list units by merge-path partitioned
//      two-way LCP merges with the loser tree's node rule (bingmann-lcp_
//      losertree.hpp:280-333: the head with the larger LCP to the last output
//      is smaller; equal LCPs compare characters from that LCP on).  The
//      merged LCP array is exact; the saved LCP of an output is the run LCP
//      it was received with (oldLcps).
//   4. the LCP of every tile's first output with its predecessor.
// Character comparisons first use the cached keys, then the strings.
constexpr uint32_t kKT = 512;           // threads per tile block
constexpr uint32_t kKW = 4;             // outputs per thread and level at most
constexpr uint32_t kKTile = kKT * kKW;  // LDS capacity of a tile (elements)
constexpr uint32_t kKG = kKTile / 2;    // tile step: a tile holds fewer than 2 kKG elements
constexpr uint32_t kKMaxFan = 16;

struct KGroup {
    uint64_t out;     // position of the group's first element
    uint64_t n;       // elements of the group (>= 1)
    uint64_t tile0;   // first tile
    uint64_t ntiles;  // ceil(n / kKG)
    uint32_t list0;   // first list
    uint32_t nl;      // lists (1..F)
    uint32_t recv;    // receiving PE
    uint32_t S;       // sample stride of its lists
};

struct KList {
    uint64_t start;   // position of the first element
    uint64_t size;    // >= 1
    uint64_t samp0;   // first sample (index into the pass's samples)
    uint64_t nsamp;   // ceil(size / S)
    uint32_t group;
    uint32_t slot;    // index inside the group
};

struct ViewCtx {
    const uint8_t* chars;
    const uint64_t* voff;
    const uint32_t* vlen;
    const uint64_t* key2;  // characters [cp + 8, cp + 16) per receive index
    uint32_t cp;
};

// Compares received strings a and b (receive indices, cached keys ka, kb at
// the receiver's common prefix cp) known to agree on the first `from` >= cp
// characters; returns the sign and the LCP.  Characters up to cp + 16 come
// from the cached keys (the second word from global memory, only on a tie of
// the first), the rest from the strings.
__device__ inline int elem_cmp(const ViewCtx& V, uint32_t a, uint64_t ka, uint32_t b, uint64_t kb, uint32_t from,
                               uint32_t& lcp) {
    if (from < V.cp + 8) {
        if (ka != kb) {
            lcp = V.cp + key_lcp(ka, kb);
            return ka < kb ? -1 : 1;
        }
        if ((ka & 0xFF) == 0) {
            lcp = V.cp + key_strlen(ka);
            return 0;
        }
        from = V.cp + 8;
    }
    if (from < V.cp + 16) {
        const uint64_t xa = V.key2[a], xb = V.key2[b];
        if (xa != xb) {
            lcp = V.cp + 8 + key_lcp(xa, xb);
            return xa < xb ? -1 : 1;
        }
        if ((xa & 0xFF) == 0) {
            lcp = V.cp + 8 + key_strlen(xa);
            return 0;
        }
        from = V.cp + 16;
    }
    return view_cmp32(V.chars, V.voff[a], V.vlen[a], V.voff[b], V.vlen[b], &lcp, from);
}

// Step 1: F lanes per sample (F a power of two <= 16: a sample's lanes share
// a wave); lane `slot` finds the sample's co-rank in the group's list `slot`.
// Two launches: coarse (every kKCoarse-th sample of a list) searches the
// whole list, fine searches only between the co-ranks of the coarse samples
// around it (co-ranks grow with the sample's position in its list).
constexpr uint32_t kKCoarse = 16;

__global__ void __launch_bounds__(256)
k_kmerge_sample(const KGroup* __restrict__ groups, const KList* __restrict__ lists, uint32_t lgF, int coarse, const uint32_t* __restrict__ ord, const uint8_t* __restrict__ chars,
                const uint64_t* __restrict__ voff, const uint32_t* __restrict__ vlen, const uint64_t* __restrict__ keyc,
                const uint64_t* __restrict__ keyc2, const uint32_t* __restrict__ rcp, uint32_t* __restrict__ corank,
                uint64_t* __restrict__ rankv) {
    const uint32_t F = 1u << lgF;
    // block (x, y): list y, its samples from x * blockDim / F on (coarse:
    // its coarse samples, every kKCoarse-th); F lanes per sample
    const KList L = lists[blockIdx.y];
    const uint64_t u = (uint64_t)blockIdx.x * (blockDim.x >> lgF) + (threadIdx.x >> lgF);
    const uint64_t si = coarse ? u * kKCoarse : u;  // sample index inside the list
    const uint32_t slot = threadIdx.x & (F - 1);
    const uint64_t s = L.samp0 + si;
    uint64_t co = 0;
    const bool mine = si < L.nsamp && (coarse || si % kKCoarse != 0);  // this launch computes sample s
    if (mine) {
        const KGroup G = groups[L.group];
        const uint64_t j = si * G.S;
        if (slot == L.slot) {
            co = j;
        } else if (slot < G.nl) {
            const KList M = lists[G.list0 + slot];
            const ViewCtx V{chars, voff, vlen, keyc2, rcp[G.recv]};
            const uint32_t x = ord ? ord[L.start + j] : (uint32_t)(L.start + j);
            const uint64_t kx = keyc[x];
            const bool later = slot > L.slot;  // a later list's equal elements come after x
            uint64_t a = 0, b = M.size;
            if (!coarse) {  // bracket by the coarse samples before and after
                const uint64_t c0 = L.samp0 + si / kKCoarse * kKCoarse, c1 = c0 + kKCoarse;
                a = corank[c0 * F + slot];
                if (c1 < L.samp0 + L.nsamp) b = corank[c1 * F + slot];
            }
            while (a < b) {
                const uint64_t mid = (a + b) / 2;
                const uint32_t y = ord ? ord[M.start + mid] : (uint32_t)(M.start + mid);
                uint32_t l;
                const int c = elem_cmp(V, y, keyc[y], x, kx, V.cp, l);
                if (c < 0 || (c == 0 && !later)) a = mid + 1; else b = mid;
            }
            co = a;
        }
        corank[s * F + slot] = (uint32_t)co;
    }
    uint64_t sum = co;  // the sample's rank: the sum over its F lanes
    for (uint32_t o = 1; o < F; o <<= 1) sum += __shfl_xor((unsigned long long)sum, (int)o);
    if (mine && slot == 0) rankv[s] = sum;
}

// Step 2: the start of every tile, kKMaxFan lanes per tile: lane l finds the
// last sample of list l with rank <= k G, the largest of them bounds the
// tile; its rank and co-ranks are stored (tile_b, tile_co).
__global__ void __launch_bounds__(256)
k_kmerge_bounds(const KGroup* __restrict__ groups, uint32_t ngroups, uint64_t ntiles, const KList* __restrict__ lists,
                uint32_t lgF, const uint32_t* __restrict__ corank, const uint64_t* __restrict__ rankv,
                uint64_t* __restrict__ tile_b, uint32_t* __restrict__ tile_co, uint32_t* __restrict__ tile_g) {
    const uint32_t F = 1u << lgF;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t tb = t / kKMaxFan;
    const uint32_t l = (uint32_t)(t % kKMaxFan);
    uint64_t best = 0, bsamp = 0;  // rank + 1 of the list's last sample <= target (0: none)
    uint32_t nl = 0;
    if (tb < ntiles) {
        uint32_t lo = 0, hi = ngroups;
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (groups[mid].tile0 <= tb) lo = mid; else hi = mid;
        }
        const KGroup G = groups[lo];
        nl = G.nl;
        if (l == 0) tile_g[tb] = lo;
        const uint64_t target = (tb - G.tile0) * kKG;
        if (l < G.nl) {
            const KList L = lists[G.list0 + l];
            uint64_t a = 0, b = L.nsamp;  // samples with rank <= target form a prefix
            while (a < b) {
                const uint64_t mid = (a + b) / 2;
                if (rankv[L.samp0 + mid] <= target) a = mid + 1; else b = mid;
            }
            if (a > 0) {
                bsamp = L.samp0 + a - 1;
                best = rankv[bsamp] + 1;
            }
        }
    }
    uint64_t m = best;
    for (uint32_t o = 1; o < kKMaxFan; o <<= 1) m = max(m, (uint64_t)__shfl_xor((unsigned long long)m, (int)o));
    if (tb < ntiles && best != 0 && best == m) {  // the bounding sample (ranks are distinct)
        tile_b[tb] = best - 1;
        for (uint32_t q = 0; q < nl; ++q) tile_co[tb * F + q] = corank[bsamp * F + q];
    }
}

// One element of a tile in LDS: its cached key, receive index and the LCP
// with its predecessor in its current list (16 bytes: one LDS access).
struct KElem {
    uint64_t key;
    uint32_t id;
    uint32_t lcp;
};

// LDS slot of tile element i: one spare record after every 16, so that lanes
// touching records 2 or 4 apart (a thread's outputs are consecutive) fall on
// different banks.
__device__ inline uint32_t kpad(uint32_t i) { return i + (i >> 4); }
constexpr uint32_t kKTileP = kKTile + kKTile / 16;

// Step 3 helper: merges outputs [dl, dl + cnt) of the unit made of lists A =
// LDS positions [aS, aS + na) and B = [bS, bS + nb) (bS = aS + na) into the
// destination buffer at the same positions.
__device__ inline void kmerge_unit(const ViewCtx& V, const KElem* __restrict__ E, KElem* __restrict__ D, uint32_t aS,
                                   uint32_t na, uint32_t bS, uint32_t nb, uint32_t dl, uint32_t cnt) {
    uint32_t lo = dl > nb ? dl - nb : 0, hi = dl < na ? dl : na;
    while (lo < hi) {  // merge path: A elements among the first dl outputs (A wins ties)
        const uint32_t mid = (lo + hi) / 2;
        const KElem x = E[kpad(aS + mid)], y = E[kpad(bS + dl - mid - 1)];
        uint32_t l;
        if (elem_cmp(V, x.id, x.key, y.id, y.key, V.cp, l) <= 0) lo = mid + 1;
        else hi = mid;
    }
    uint32_t i = lo, j = dl - lo;
    uint32_t la = V.cp, lb = V.cp;  // LCPs of the heads with the last output
    KElem ha, hb;                   // the heads
    if (i < na) ha = E[kpad(aS + i)];
    if (j < nb) hb = E[kpad(bS + j)];
    if (dl > 0) {
        bool prev_is_a;
        uint32_t x;
        KElem pa, pb;
        if (i > 0) pa = E[kpad(aS + i - 1)];
        if (j > 0) pb = E[kpad(bS + j - 1)];
        if (i == 0) prev_is_a = false;
        else if (j == 0) prev_is_a = true;
        else prev_is_a = elem_cmp(V, pa.id, pa.key, pb.id, pb.key, V.cp, x) > 0;
        const KElem pp = prev_is_a ? pa : pb;
        if (i < na) {
            if (prev_is_a) la = ha.lcp;
            else elem_cmp(V, ha.id, ha.key, pp.id, pp.key, V.cp, la);
        }
        if (j < nb) {
            if (!prev_is_a) lb = hb.lcp;
            else elem_cmp(V, hb.id, hb.key, pp.id, pp.key, V.cp, lb);
        }
    }
    for (uint32_t c = 0; c < cnt; ++c) {
        const bool has_a = i < na, has_b = j < nb;
        bool take_a;
        if (!has_a) take_a = false;
        else if (!has_b) take_a = true;
        else if (la > lb) take_a = true;
        else if (lb > la) take_a = false;
        else {
            uint32_t x;
            const int cr = elem_cmp(V, ha.id, ha.key, hb.id, hb.key, la, x);
            take_a = cr <= 0;
            if (take_a) lb = x; else la = x;
        }
        KElem out;
        if (take_a) {
            out = ha;
            out.lcp = la;
            ++i;
            if (i < na) {
                ha = E[kpad(aS + i)];
                la = ha.lcp;
            }
        } else {
            out = hb;
            out.lcp = lb;
            ++j;
            if (j < nb) {
                hb = E[kpad(bS + j)];
                lb = hb.lcp;
            }
        }
        D[kpad(aS + dl + c)] = out;
    }
}

// Steps 2 and 3: one block per tile.
__global__ void __launch_bounds__(kKT)
k_kmerge_tile(const KGroup* __restrict__ groups, uint32_t ngroups, const KList* __restrict__ lists, uint32_t lgF,
              const uint64_t* __restrict__ tile_b, const uint32_t* __restrict__ tile_co,
              const uint32_t* __restrict__ tile_g,
              const uint32_t* __restrict__ ord, const uint32_t* __restrict__ lcp_in, uint32_t* __restrict__ ord_out,
              uint32_t* __restrict__ lcp_out, const uint8_t* __restrict__ chars, const uint64_t* __restrict__ voff,
              const uint32_t* __restrict__ vlen, const uint64_t* __restrict__ keyc,
              const uint64_t* __restrict__ keyc2, const uint32_t* __restrict__ rcp, const uint32_t* __restrict__ rlcp,
              uint32_t* __restrict__ saved) {
    __shared__ KElem sE[2][kKTileP];
    __shared__ uint64_t sBnd[2];
    __shared__ uint32_t sCo[2][kKMaxFan];
    __shared__ uint32_t sOff[kKMaxFan + 1];
    __shared__ uint64_t sStart[kKMaxFan];  // position of every list's first element of the tile
    const uint32_t F = 1u << lgF;
    const uint64_t tb = blockIdx.x;
    const KGroup G = groups[tile_g[tb]];
    const uint64_t k = tb - G.tile0;
    // tile bounds (step 2); the group's end bounds its last tile
    if (threadIdx.x < 2 * kKMaxFan) {
        const uint32_t which = threadIdx.x / kKMaxFan, l = threadIdx.x % kKMaxFan;
        const bool at_end = k + which >= G.ntiles;
        if (l == 0) sBnd[which] = at_end ? G.n : tile_b[tb + which];
        if (l < G.nl) sCo[which][l] = at_end ? (uint32_t)lists[G.list0 + l].size : tile_co[(tb + which) * F + l];
    }
    __syncthreads();
    const uint64_t B0 = sBnd[0], B1 = sBnd[1];
    if (threadIdx.x < G.nl) sStart[threadIdx.x] = lists[G.list0 + threadIdx.x].start + sCo[0][threadIdx.x];
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t l = 0; l < G.nl; ++l) {
            sOff[l] = acc;
            acc += sCo[1][l] - sCo[0][l];
        }
        sOff[G.nl] = acc;
    }
    __syncthreads();
    const uint32_t tsize = (uint32_t)(B1 - B0);
    if (B1 < B0 || B1 - B0 > kKTile || sOff[G.nl] != tsize) return;  // cannot happen (gap bound)
    const ViewCtx V{chars, voff, vlen, keyc2, rcp[G.recv]};
    // stage the tile: list l's part at LDS [sOff[l], sOff[l+1])
    for (uint32_t l = 0; l < G.nl; ++l) {
        const uint32_t o = sOff[l], c = sOff[l + 1] - o;
        const uint64_t p0 = sStart[l];
        for (uint32_t i = threadIdx.x; i < c; i += kKT) {
            const uint64_t pos = p0 + i;
            const uint32_t id = ord ? ord[pos] : (uint32_t)pos;
            KElem e;
            e.key = keyc[id];
            e.id = id;
            e.lcp = lcp_in[pos];
            sE[0][kpad(o + i)] = e;
        }
    }
    __syncthreads();
    // outputs per thread: as few as the tile allows, so that every thread works
    const uint32_t kw = (tsize + kKT - 1) / kKT;
    uint32_t cur = 0;
    for (uint32_t w = 1; w < G.nl; w <<= 1) {  // level: units of w lists merge pairwise
        const uint32_t o0 = threadIdx.x * kw, oe = min(o0 + kw, tsize);
        uint32_t o = o0;
        while (o < oe) {
            uint32_t u = 0;  // last unit (2w lists) starting at or before o
            while ((u + 1) * 2 * w < G.nl && sOff[(u + 1) * 2 * w] <= o) ++u;
            const uint32_t aL = u * 2 * w, bL = min(aL + w, G.nl), eL = min(aL + 2 * w, G.nl);
            const uint32_t aS = sOff[aL], bS = sOff[bL], bE = sOff[eL];
            const uint32_t cnt = min(oe, bE) - o;
            kmerge_unit(V, sE[cur], sE[cur ^ 1], aS, bS - aS, bS, bE - bS, o - aS, cnt);
            o += cnt;
        }
        __syncthreads();
        cur ^= 1;
    }
    for (uint32_t i = threadIdx.x; i < tsize; i += kKT) {
        const KElem e = sE[cur][kpad(i)];
        const uint32_t id = e.id;
        const uint64_t pos = G.out + B0 + i;
        ord_out[pos] = id;
        lcp_out[pos] = e.lcp;  // the tile's first output: step 4
        if (saved) saved[pos] = rlcp[id];
    }
}

// Step 4: the LCP of every tile's first output with the previous output (0 at
// a group's start).
__global__ void k_kmerge_fix(const KGroup* __restrict__ groups, uint32_t ngroups, uint64_t ntiles,
                             const uint64_t* __restrict__ tile_b, const uint32_t* __restrict__ ord_out,
                             uint32_t* __restrict__ lcp_out, const uint8_t* __restrict__ chars,
                             const uint64_t* __restrict__ voff, const uint32_t* __restrict__ vlen,
                             const uint64_t* __restrict__ keyc, const uint64_t* __restrict__ keyc2,
                             const uint32_t* __restrict__ rcp) {
    const uint64_t tb = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tb >= ntiles) return;
    uint32_t lo = 0, hi = ngroups;
    while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (groups[mid].tile0 <= tb) lo = mid; else hi = mid;
    }
    const KGroup G = groups[lo];
    const uint64_t b = tile_b[tb];
    if (b >= G.n) return;
    const uint64_t pos = G.out + b;
    if (b == 0) {
        lcp_out[pos] = 0;
        return;
    }
    const ViewCtx V{chars, voff, vlen, keyc2, rcp[G.recv]};
    const uint32_t a = ord_out[pos - 1], c = ord_out[pos];
    uint32_t l = 0;
    elem_cmp(V, a, keyc[a], c, keyc[c], V.cp, l);
    lcp_out[pos] = l;
}

// Fan-in 2: the binary tree of the same node merges, one global pass per
// tree level (merge<2> nested log2(runs) deep).  Every pass splits its pairs'
// outputs into tiles of kTileOut by merge-path searches in global memory
// (k_merge_split), then each tile's windows of the pair are staged in LDS and
// merged (k_merge_tile).  No sampling and exact tiles: on one MI355X this is
// the faster form where the merge is bound by the node merges' dependent LDS
// latency rather than by HBM traffic (the D/N inputs); the tuner picks the
// fan-in per input regime.
constexpr uint32_t kMT = 256;  // threads per merge block; kMW outputs per thread (a template
                               // parameter: 2, 4 or 8, i.e. 512, 1024 or 2048 outputs per tile)

struct MergePair {
    uint64_t out;      // first receive position of the pair (A then B)
    uint64_t na, nb;   // run lengths
    uint64_t tile0;    // first tile of the pair
    uint32_t recv;     // receiving PE
    uint32_t pad;
};

__device__ inline uint32_t pair_of_tile(const MergePair* __restrict__ pairs, uint32_t npairs, uint64_t t) {
    uint32_t lo = 0, hi = npairs;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (pairs[mid].tile0 <= t) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

// Merge-path split (number of A elements among the first d outputs, A wins
// ties) at the start of every tile, from global memory.  ord == nullptr: the
// first round, whose order is the receive order.  kSL lanes per tile search
// kSL-ary: lanes 0..kSL-2 each test one of the interval's kSL-1 inner points
// per step, so a split takes about log_kSL of the run length dependent steps
// (a string comparison that the cached keys leave open walks the strings in
// its lane, so fewer steps matter most where runs share long prefixes).
constexpr uint32_t kSL = 8;

template <uint32_t kMW>
__global__ void k_merge_split(const MergePair* __restrict__ pairs, uint32_t npairs, uint64_t ntiles,
                              const uint32_t* __restrict__ ord, const uint8_t* __restrict__ chars,
                              const uint64_t* __restrict__ voff, const uint32_t* __restrict__ vlen,
                              const uint64_t* __restrict__ keyc, const uint64_t* __restrict__ keyc2,
                              const uint32_t* __restrict__ rcp, uint64_t* __restrict__ split) {
    const uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kSL;  // uniform over the group
    const uint32_t l = threadIdx.x % kSL, g0 = (threadIdx.x & 63) & ~(kSL - 1);
    const bool live = t < ntiles;
    uint64_t lo = 0, hi = 0, A = 0, B = 0, d = 0;
    ViewCtx V{chars, voff, vlen, keyc2, 0};
    if (live) {
        const MergePair Pr = pairs[pair_of_tile(pairs, npairs, t)];
        V.cp = rcp[Pr.recv];
        A = Pr.out;
        B = Pr.out + Pr.na;
        constexpr uint32_t kTileOut = kMT * kMW;
        d = (t - Pr.tile0) * kTileOut;
        lo = d > Pr.nb ? d - Pr.nb : 0;
        hi = d < Pr.na ? d : Pr.na;
    }
    // the answer is the first mid in [lo, hi] with A[mid] > B[d - mid - 1]
    // (hi if none); the test is monotone in mid, so the probes that hold form
    // a prefix of lanes 0..kSL-2
    while (__ballot(lo < hi)) {  // every lane of the wave takes part in the ballot
        const uint64_t span = hi - lo;
        const uint64_t m = lo + span * (l + 1) / kSL;  // non-decreasing in l, < hi for l < kSL - 1
        bool le = false;
        if (lo < hi && l < kSL - 1) {
            const uint64_t ia = A + m, ib = B + d - m - 1;
            const uint32_t a = ord ? ord[ia] : (uint32_t)ia, b = ord ? ord[ib] : (uint32_t)ib;
            uint32_t x;
            le = elem_cmp(V, a, keyc[a], b, keyc[b], V.cp, x) <= 0;
        }
        const uint64_t held = (__ballot(le) >> g0) & ((1ull << (kSL - 1)) - 1);
        const uint32_t c = (uint32_t)__popcll(held);  // probes that held
        if (lo < hi) {
            const uint64_t mc = lo + span * (c + 1) / kSL;  // the first probe that failed (c < kSL - 1)
            const uint64_t ml = lo + span * c / kSL;        // the last probe that held (c > 0)
            if (c > 0) lo = ml + 1;
            if (c < kSL - 1) hi = mc;
        }
    }
    if (live && l == 0) split[t] = lo;
}

// A tile's strings in LDS: slot x holds the receive index o[x], the LCP l[x]
// with its predecessor in its run, and the 16 characters at `base` (k1[x],
// k2[x]).  Every two strings the tile compares agree on their first `base`
// characters.  k2s: the second keys are staged; otherwise they are read from
// the global keys at the receiver's common prefix (base is that prefix then).
struct TileView {
    const uint32_t* o;
    const uint64_t* k1;
    const uint64_t* k2;
    const uint8_t* chars;
    const uint64_t* voff;
    const uint32_t* vlen;
    const uint64_t* gk2;
    uint32_t base;
    bool k2s;
};

// Compares, for every lane of the wave with `act` set, the strings of slots
// x and y, known to agree on their first `from` characters; returns the sign
// and the LCP.  All lanes of the wave call it together.  The staged keys are
// compared per lane; the lanes they leave undecided walk their strings on
// lane groups, up to eight at a time, each lane of a group comparing 32
// characters (256 to 1024 characters per group and dependent step); the
// last undecided lane is served by the whole wave, lane l comparing the 8
// characters at from + 8 l (512 characters per dependent step).  A lane
// whose strings agree on hundreds of characters thus costs the wave one or
// two loads instead of a loop of 32-character steps that the other lanes
// wait out.
constexpr int kWideKeys = 4;  // keys (8 characters each) per lane and step of a grouped walk
__device__ inline int wave_tile_cmp(const TileView& T, bool act, uint32_t x, uint32_t y, uint32_t from,
                                    uint32_t& lcp) {
    int c = 0;
    bool pend = false;
    if (act) {
        pend = true;
        if (from < T.base + 8) {
            const uint64_t a = T.k1[x], b = T.k1[y];
            if (a != b) {
                lcp = T.base + key_lcp(a, b);
                c = a < b ? -1 : 1;
                pend = false;
            } else if ((a & 0xFF) == 0) {
                lcp = T.base + key_strlen(a);
                pend = false;
            }
            from = T.base + 8;
        }
        if (pend && from < T.base + 16) {
            const uint64_t a = T.k2s ? T.k2[x] : T.gk2[T.o[x]], b = T.k2s ? T.k2[y] : T.gk2[T.o[y]];
            if (a != b) {
                lcp = T.base + 8 + key_lcp(a, b);
                c = a < b ? -1 : 1;
                pend = false;
            } else if ((a & 0xFF) == 0) {
                lcp = T.base + 8 + key_strlen(a);
                pend = false;
            }
            from = T.base + 16;
        }
    }
    uint64_t ao = 0, bo = 0;
    uint32_t al = 0, bl = 0;
    if (pend) {
        const uint32_t oa = T.o[x], ob = T.o[y];
        ao = T.voff[oa];
        al = T.vlen[oa];
        bo = T.voff[ob];
        bl = T.vlen[ob];
    }
    const uint32_t lane = lane_id();
    uint64_t m = __ballot(pend);
    // several open lanes: up to kG = 8 at a time, each on its own group of
    // W = 64 / kG lanes whose every lane reads kWideKeys keys, so a group
    // reaches 32 W characters (256 to 1024) per dependent step; a lane still
    // open after its step is taken again, the last open lane is served by
    // the whole wave below
    while (__popcll(m) >= 2) {  // wave-uniform
        uint64_t todo = m;
        const uint32_t k = min(8u, (uint32_t)__popcll(todo));
        const uint32_t lgk = k <= 2 ? 1u : (k <= 4 ? 2u : 3u);
        const uint32_t W = 64u >> lgk, g = lane >> (6 - lgk);
        uint64_t sel = 0;
        uint32_t myld = 0;
        bool gact = false;
        for (uint32_t r = 0; r < k; ++r) {
            const uint32_t ldr = (uint32_t)__ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            sel |= 1ull << ldr;
            if (g == r) {
                myld = ldr;
                gact = true;
            }
        }
        const uint64_t pa = __shfl((unsigned long long)ao, myld), pb = __shfl((unsigned long long)bo, myld);
        const uint32_t pal = (uint32_t)__shfl((int)al, myld), pbl = (uint32_t)__shfl((int)bl, myld);
        const uint32_t pf = (uint32_t)__shfl((int)from, myld);
        const uint32_t sub = lane & (W - 1);
        uint64_t ea = 0, eb = 0;
        uint32_t ej = 0;
        bool ev = false;
        if (gact) {
            uint64_t ka[kWideKeys], kb[kWideKeys];
            load_keys_n<kWideKeys>(T.chars, pa, pal, pf + 8 * kWideKeys * sub, ka);
            load_keys_n<kWideKeys>(T.chars, pb, pbl, pf + 8 * kWideKeys * sub, kb);
#pragma unroll
            for (int j = kWideKeys - 1; j >= 0; --j) {  // the lane's first event wins
                if (ka[j] != kb[j] || (ka[j] & 0xFF) == 0) {
                    ev = true;
                    ea = ka[j];
                    eb = kb[j];
                    ej = (uint32_t)j;
                }
            }
        }
        const uint64_t wmask = W == 64 ? ~0ull : ((1ull << W) - 1);
        const uint64_t gm = (__ballot(ev) >> (g * W)) & wmask;
        const uint32_t f = g * W + (gm ? (uint32_t)__ffsll((unsigned long long)gm) - 1 : 0u);
        const uint64_t fa = __shfl((unsigned long long)ea, f), fb = __shfl((unsigned long long)eb, f);
        const uint32_t fj = (uint32_t)__shfl((int)ej, f);
        int gc = 0;
        uint32_t glcp = 0;
        if (gm) {
            gc = fa < fb ? -1 : (fa > fb ? 1 : 0);
            glcp = pf + 8 * kWideKeys * (f - g * W) + 8 * fj + (fa != fb ? key_lcp(fa, fb) : key_strlen(fa));
        }
        const uint32_t src = min((uint32_t)__popcll(sel & lanemask_lt()), (1u << lgk) - 1) * W;
        const int rc = __shfl(gc, src);
        const uint32_t rl = (uint32_t)__shfl((int)glcp, src);
        const bool rd = __shfl((int)(gm != 0), src) != 0;
        if ((sel >> lane) & 1) {
            if (rd) {
                c = rc;
                lcp = rl;
                pend = false;
            } else {
                from += 8 * kWideKeys * W;
            }
        }
        m = __ballot(pend);
    }
    for (; m; m &= m - 1) {
        const int ld = __ffsll((unsigned long long)m) - 1;
        const uint64_t pa = __shfl((unsigned long long)ao, ld), pb = __shfl((unsigned long long)bo, ld);
        const uint32_t pal = (uint32_t)__shfl((int)al, ld), pbl = (uint32_t)__shfl((int)bl, ld);
        const uint32_t pf = (uint32_t)__shfl((int)from, ld);
        for (uint32_t base = pf;; base += 512) {
            const uint32_t pos = base + 8 * lane;
            const uint64_t ka = load_key8(T.chars, pa, pal, pos), kb = load_key8(T.chars, pb, pbl, pos);
            const uint64_t mm = __ballot(ka != kb || (ka & 0xFF) == 0);
            if (mm) {
                const int f = __ffsll((unsigned long long)mm) - 1;
                const uint64_t fa = __shfl((unsigned long long)ka, f), fb = __shfl((unsigned long long)kb, f);
                if (lane == (uint32_t)ld) {
                    c = fa < fb ? -1 : (fa > fb ? 1 : 0);
                    lcp = base + 8 * (uint32_t)f + (fa != fb ? key_lcp(fa, fb) : key_strlen(fa));
                }
                break;
            }
        }
    }
    return c;
}

// One tile of kTileOut outputs per block: the tile's windows of A and B
// (plus the element before each) are staged in LDS, A's at slots 0..la_n and
// B's after them; every thread finds its start by a merge-path search in LDS
// and merges kMW outputs with the LCP rules of the loser tree
// (bingmann-lcp_losertree.hpp:280-333): the head with the larger LCP to the
// last output is smaller, equal LCPs compare characters from there on.
//
// Characters: the staged run LCPs bound the LCP of any two strings of one
// window from below (their minimum over the window's adjacent pairs); with
// the LCP of the two windows' first strings (one wave compares them) they
// bound every pair the tile compares.  When that bound T is at least 8
// characters past the receiver's common prefix cp, the cached keys at cp are
// equal throughout the tile and the strings' 16 characters at T are staged
// instead; otherwise the keys at cp are staged, the second one only when some
// adjacent pair of a window agrees beyond cp + 8.
//
// ord == nullptr: the first round (receive order); saved != nullptr: the last
// round, which also writes every output's saved LCP (the LCP it was received
// with, rlcp).
template <uint32_t kMW>
__global__ void __launch_bounds__(kMT)
k_merge_tile(const MergePair* __restrict__ pairs, uint32_t npairs, const uint64_t* __restrict__ split,
             const uint32_t* __restrict__ ord, const uint32_t* __restrict__ lcpa, uint32_t* __restrict__ ord_out,
             uint32_t* __restrict__ lcp_out, const uint8_t* __restrict__ chars, const uint64_t* __restrict__ voff,
             const uint32_t* __restrict__ vlen, const uint64_t* __restrict__ keyc,
             const uint64_t* __restrict__ keyc2, const uint32_t* __restrict__ rcp, const uint32_t* __restrict__ rlcp,
             uint32_t* __restrict__ saved) {
    constexpr uint32_t kTileOut = kMT * kMW;
    constexpr uint32_t kSlots = kTileOut + 2;
    __shared__ uint32_t sO[kSlots], sL[kSlots];
    __shared__ uint64_t sK1[kSlots], sK2[kSlots];
    __shared__ uint32_t sWMin[kMT / 64], sWMax[kMT / 64], sCross;
    const uint64_t t = blockIdx.x;
    const uint32_t pi = pair_of_tile(pairs, npairs, t);
    const MergePair Pr = pairs[pi];
    const uint32_t cp = rcp[Pr.recv];
    const uint64_t A = Pr.out, B = Pr.out + Pr.na;
    const uint64_t na = Pr.na, nb = Pr.nb, tot = na + nb;
    const uint64_t d0 = (t - Pr.tile0) * kTileOut;
    const uint64_t d1 = min(d0 + kTileOut, tot);
    const uint64_t i0 = split[t];
    const uint64_t i1 = (d1 == tot) ? na : split[t + 1];
    const uint64_t j0 = d0 - i0, j1 = d1 - i1;
    const uint32_t la_n = (uint32_t)(i1 - i0), lb_n = (uint32_t)(j1 - j0);
    const uint32_t bs = la_n + 1;  // slot of B's window position 0 (B[j0 - 1])
    // stage A[i0-1 .. i1) at slots 0..la_n and B[j0-1 .. j1) at bs..bs+lb_n;
    // the LCP of a window position with a predecessor inside the window
    // enters the bounds
    uint32_t lmin = 0xFFFFFFFFu, lmax = 0;
    for (uint32_t k = threadIdx.x; k <= la_n + 1 + lb_n; k += kMT) {
        const bool in_a = k <= la_n;
        const uint32_t w = in_a ? k : k - bs;
        const int64_t g = (int64_t)(in_a ? i0 : j0) - 1 + w;
        const uint64_t len = in_a ? na : nb;
        if (g >= 0 && (uint64_t)g < len) {
            const uint64_t pos = (in_a ? A : B) + g;
            const uint32_t o = ord ? ord[pos] : (uint32_t)pos;
            const uint32_t l = lcpa[pos];
            sO[k] = o;
            sL[k] = l;
            sK1[k] = keyc[o];
            if (w >= 1 && g >= 1) {
                lmin = min(lmin, l);
                lmax = max(lmax, l);
            }
        }
    }
    for (uint32_t m = 32; m; m >>= 1) {
        lmin = min(lmin, (uint32_t)__shfl_xor((int)lmin, m));
        lmax = max(lmax, (uint32_t)__shfl_xor((int)lmax, m));
    }
    if (lane_id() == 0) {
        sWMin[threadIdx.x / 64] = lmin;
        sWMax[threadIdx.x / 64] = lmax;
    }
    __syncthreads();
    uint32_t sMin = sWMin[0], sMax = sWMax[0];
#pragma unroll
    for (uint32_t w = 1; w < kMT / 64; ++w) {
        sMin = min(sMin, sWMin[w]);
        sMax = max(sMax, sWMax[w]);
    }
    // first staged slot of each window (none: kNone)
    const uint32_t fa = i0 > 0 ? 0u : (la_n > 0 ? 1u : kNone);
    const uint32_t fb = j0 > 0 ? bs : (lb_n > 0 ? bs + 1 : kNone);
    const bool try_restage = fa != kNone && fb != kNone && sMin != 0xFFFFFFFFu && sMin >= cp + 8;  // block-uniform
    if (try_restage) {
        if (threadIdx.x < 64) {
            const uint32_t oa = sO[fa], ob = sO[fb];
            const uint32_t x = wave_lcp(chars, voff[oa], vlen[oa], voff[ob], vlen[ob], cp);
            if (lane_id() == 0) sCross = x;
        }
        __syncthreads();
    }
    const uint32_t T = try_restage ? min(sMin, sCross) : cp;
    const bool restage = T >= cp + 8;
    const bool k2s = restage || sMax >= cp + 8;
    if (k2s) {
        for (uint32_t k = threadIdx.x; k <= la_n + 1 + lb_n; k += kMT) {
            const bool in_a = k <= la_n;
            const int64_t g = (int64_t)(in_a ? i0 : j0) - 1 + (in_a ? k : k - bs);
            if (g >= 0 && (uint64_t)g < (in_a ? na : nb)) {
                const uint32_t o = sO[k];
                if (restage) {
                    const uint64_t so = voff[o];
                    const uint32_t sl = vlen[o];
                    sK1[k] = load_key8(chars, so, sl, T);
                    sK2[k] = load_key8(chars, so, sl, T + 8);
                } else {
                    sK2[k] = keyc2[o];
                }
            }
        }
        __syncthreads();
    }
    const TileView V{sO, sK1, sK2, chars, voff, vlen, keyc2, restage ? T : cp, k2s};
    const uint32_t dl = threadIdx.x * kMW;
    const bool live = d0 + dl < d1;
    // local merge path (window positions; A's position w is slot w, B's is
    // bs + w); every lane runs the wave's iterations, the finished ones idle
    uint32_t lo = 0, hi = 0;
    if (live) {
        lo = dl > lb_n ? dl - lb_n : 0;
        hi = dl < la_n ? dl : la_n;
    }
    while (__ballot(lo < hi)) {
        const bool act = lo < hi;
        const uint32_t mid = (lo + hi) / 2;
        uint32_t x;
        const int c = wave_tile_cmp(V, act, mid + 1, act ? bs + dl - mid : 0u, cp, x);
        if (act) {
            if (c <= 0) lo = mid + 1;
            else hi = mid;
        }
    }
    uint32_t il = lo, jl = dl - lo;             // heads: slots il + 1 and bs + jl + 1
    const uint64_t gi = i0 + il, gj = j0 + jl;  // global indices in A and B
    uint32_t la = cp, lb = cp;
    // the previous output is the later of A[gi - 1] and B[gj - 1] (B on
    // equality); the head of the other list gets its LCP with it by a
    // comparison, the head of the same list has its run LCP
    const bool has_prev = live && d0 + dl > 0;
    const bool cmp_prev = has_prev && gi != 0 && gj != 0;
    uint32_t x;
    const int pc = wave_tile_cmp(V, cmp_prev, il, bs + jl, cp, x);
    const bool prev_is_a = has_prev && gi != 0 && (gj == 0 || pc > 0);
    const uint32_t ps = prev_is_a ? il : bs + jl;
    const bool cmp_a = has_prev && !prev_is_a && il < la_n;
    const bool cmp_b = has_prev && prev_is_a && jl < lb_n;
    uint32_t lh = cp;
    wave_tile_cmp(V, cmp_a || cmp_b, cmp_a ? il + 1 : bs + jl + 1, ps, cp, lh);
    if (has_prev) {
        if (il < la_n) la = prev_is_a ? sL[il + 1] : lh;
        if (jl < lb_n) lb = prev_is_a ? lh : sL[bs + jl + 1];
    }
    const uint32_t oend = live ? (uint32_t)min((uint64_t)dl + kMW, d1 - d0) : dl;
    uint64_t res[kMW];  // (receive index << 32) | LCP of the thread's outputs
    for (uint32_t q = 0; q < kMW; ++q) {
        const uint32_t o = dl + q;
        const bool valid = o < oend;
        const bool has_a = il < la_n, has_b = jl < lb_n;
        const bool tie = valid && has_a && has_b && la == lb;
        uint32_t y;
        const int cr = wave_tile_cmp(V, tie, il + 1, bs + jl + 1, la, y);
        if (!valid) continue;
        bool take_a;
        if (!has_a) take_a = false;
        else if (!has_b) take_a = true;
        else if (la > lb) take_a = true;
        else if (lb > la) take_a = false;
        else {
            take_a = cr <= 0;
            if (take_a) lb = y; else la = y;
        }
        uint32_t lout, outo;
        if (take_a) {
            lout = la;
            outo = sO[il + 1];
            ++il;
            if (il < la_n) la = sL[il + 1];
        } else {
            lout = lb;
            outo = sO[bs + jl + 1];
            ++jl;
            if (jl < lb_n) lb = sL[bs + jl + 1];
        }
        res[q] = ((uint64_t)outo << 32) | ((d0 + o == 0) ? 0u : lout);
    }
    // the tile's outputs leave through LDS, so that consecutive lanes store
    // consecutive outputs (a thread's kMW outputs are adjacent)
    __syncthreads();  // every comparison has read the staged keys
    for (uint32_t q = 0; q < kMW; ++q)
        if (dl + q < oend) sK1[dl + q] = res[q];
    __syncthreads();
    const uint32_t cnt = (uint32_t)(d1 - d0);
    for (uint32_t o = threadIdx.x; o < cnt; o += kMT) {
        const uint64_t r = sK1[o];
        const uint32_t outo = (uint32_t)(r >> 32);
        ord_out[A + d0 + o] = outo;
        lcp_out[A + d0 + o] = (uint32_t)r;
        if (saved) saved[A + d0 + o] = rlcp[outo];  // the last round: the LCP it was received with
    }
}


__global__ void k_iota_copy(uint32_t* __restrict__ ord, uint32_t* __restrict__ l, const uint32_t* __restrict__ rl,
                            uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        ord[k] = (uint32_t)k;
        l[k] = rl[k];
    }
}

__global__ void k_gather_saved(const uint32_t* __restrict__ ord, const uint32_t* __restrict__ rl, uint64_t n,
                               uint32_t* __restrict__ saved) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) saved[k] = rl[ord[k]];
}

void merge(dss_set_s* s) {
    PipeState& P = s->pp;
    DSS_REQUIRE(P.exchanged, DSS_ESTATE, "dss_exchange has not run");
    const hipStream_t st = s->stream;
    const uint32_t p = s->p;
    const uint64_t n = s->n;
    const uint64_t nn = std::max<uint64_t>(n, 1);
    P.m_order.reserve(nn * 4);
    P.m_lcp.reserve(nn * 4);
    P.m_order2.reserve(nn * 4);
    P.m_lcp2.reserve(nn * 4);
    P.m_saved.reserve(nn * 4);
    // the receivers' common prefixes and the cached keys at them were
    // written by the exchange's copy
    uint32_t* cp = P.m_cp.as<uint32_t>();
    uint32_t* ord = P.m_order.as<uint32_t>();
    uint32_t* lc = P.m_lcp.as<uint32_t>();
    uint32_t* ord2 = P.m_order2.as<uint32_t>();
    uint32_t* lc2 = P.m_lcp2.as<uint32_t>();
    const uint32_t F = s->sort_params.merge_fan;
    uint32_t lgF = 0;
    while ((1u << lgF) < F) ++lgF;
    // every receiver's non-empty runs, in source order
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> lists(p);
    bool any = false;
    for (uint32_t q = 0; q < p; ++q) {
        uint64_t pos = P.recv_begin[q];
        for (uint32_t r = 0; r < p; ++r) {
            const uint64_t c = P.interval[(uint64_t)r * p + q];
            if (c) lists[q].push_back({pos, c});
            pos += c;
        }
        any |= lists[q].size() > 1;
    }
    if (any && F == 2) {
        // fan-in 2: all rounds' pair lists are built first and uploaded
        // once, so the rounds run back to back without host synchronisation
        auto& runs = lists;
        // outputs per tile: 512, 1024 (default) or 2048 (kMW 2, 4, 8)
        const uint32_t mw = s->sort_params.merge_tile / kMT;
        const uint64_t tile_out = (uint64_t)kMT * mw;
        std::vector<MergePair> all;
        std::vector<size_t> rbeg;          // first pair of each round
        std::vector<uint64_t> rtiles, tbeg;  // tiles of each round, first split slot
        uint64_t tsum = 0;
        while (true) {
            bool more = false;
            for (uint32_t q = 0; q < p; ++q) more |= runs[q].size() > 1;
            if (!more) break;
            rbeg.push_back(all.size());
            tbeg.push_back(tsum);
            uint64_t tiles = 0;
            for (uint32_t q = 0; q < p; ++q) {
                auto& R = runs[q];
                std::vector<std::pair<uint64_t, uint64_t>> next;
                for (size_t k = 0; k < R.size(); k += 2) {
                    MergePair mp;
                    mp.out = R[k].first;
                    mp.na = R[k].second;
                    mp.nb = k + 1 < R.size() ? R[k + 1].second : 0;
                    mp.recv = q;
                    mp.pad = 0;
                    mp.tile0 = tiles;
                    tiles += (mp.na + mp.nb + tile_out - 1) / tile_out;
                    all.push_back(mp);
                    next.push_back({mp.out, mp.na + mp.nb});
                }
                R.swap(next);
            }
            rtiles.push_back(tiles);
            tsum += tiles + 1;
        }
        rbeg.push_back(all.size());
        if (!all.empty()) {
            const uint64_t pair_bytes = (all.size() * sizeof(MergePair) + 15) / 16 * 16;
            P.m_work.reserve(pair_bytes + (tsum + 1) * 8);
            MergePair* dpairs = P.m_work.as<MergePair>();
            uint64_t* splits = reinterpret_cast<uint64_t*>(P.m_work.as<uint8_t>() + pair_bytes);
            DSS_HIP(hipMemcpyAsync(dpairs, all.data(), all.size() * sizeof(MergePair), hipMemcpyHostToDevice, st));
            // round 0 reads the receive order and LCPs directly; round r writes
            // the buffer pair round r - 1 did not
            const uint32_t* in_ord = nullptr;
            const uint32_t* in_lc = P.recv_lcp.as<uint32_t>();
            for (size_t r = 0; r + 1 < rbeg.size(); ++r) {
                const uint32_t np = (uint32_t)(rbeg[r + 1] - rbeg[r]);
                const uint64_t tiles = rtiles[r];
                MergePair* rp = dpairs + rbeg[r];
                uint64_t* split = splits + tbeg[r];
                const bool last = r + 2 == rbeg.size();
                KScope ks(s->timer, st, "merge", n * 24.0);
                auto launch = [&](auto split_kern, auto tile_kern) {
                    hipLaunchKernelGGL(split_kern, dim3((tiles * kSL + 255) / 256), dim3(256), 0, st, rp, np, tiles,
                                       in_ord, s->arena(), P.recv_view_off.as<uint64_t>(),
                                       P.recv_view_len.as<uint32_t>(), P.m_keyc.as<uint64_t>(),
                                       P.m_keyc2.as<uint64_t>(), cp, split);
                    hipLaunchKernelGGL(tile_kern, dim3(tiles), dim3(kMT), 0, st, rp, np, split, in_ord, in_lc, ord,
                                       lc, s->arena(), P.recv_view_off.as<uint64_t>(),
                                       P.recv_view_len.as<uint32_t>(), P.m_keyc.as<uint64_t>(),
                                       P.m_keyc2.as<uint64_t>(), cp, P.recv_lcp.as<uint32_t>(),
                                       last ? P.m_saved.as<uint32_t>() : nullptr);
                };
                if (mw == 2) launch(k_merge_split<2>, k_merge_tile<2>);
                else if (mw == 8) launch(k_merge_split<8>, k_merge_tile<8>);
                else launch(k_merge_split<4>, k_merge_tile<4>);
                in_ord = ord;
                in_lc = lc;
                std::swap(ord, ord2);
                std::swap(lc, lc2);
            }
            DSS_HIP(hipGetLastError());
            // results in the last round's output pair
            ord = const_cast<uint32_t*>(in_ord);
            lc = const_cast<uint32_t*>(in_lc);
        }
    } else if (any) {
        // all passes' groups and lists are built first and uploaded once, so
        // the passes run back to back without host synchronisation; a
        // receiver whose lists are already merged is carried along as a
        // one-list group (a copy into the pass's output buffers)
        struct Pass {
            size_t g0, ng, l0, nl;
            uint64_t nsamp, ntiles, n, max_list_samp;
        };
        std::vector<KGroup> G;
        std::vector<KList> L;
        std::vector<Pass> passes;
        while (any) {
            Pass ps{G.size(), 0, L.size(), 0, 0, 0, 0, 0};
            any = false;
            for (uint32_t q = 0; q < p; ++q) {
                auto& R = lists[q];
                std::vector<std::pair<uint64_t, uint64_t>> next;
                for (size_t k = 0; k < R.size(); k += F) {
                    const uint32_t nl = (uint32_t)std::min<size_t>(F, R.size() - k);
                    KGroup g{};
                    g.out = R[k].first;
                    g.list0 = (uint32_t)(L.size() - ps.l0);
                    g.nl = nl;
                    g.recv = q;
                    g.S = (kKG - 1) / nl + 1;
                    g.tile0 = ps.ntiles;
                    for (uint32_t i = 0; i < nl; ++i) {
                        KList l{};
                        l.start = R[k + i].first;
                        l.size = R[k + i].second;
                        l.samp0 = (ps.nsamp + kKCoarse - 1) / kKCoarse * kKCoarse;  // coarse-aligned
                        l.nsamp = (l.size + g.S - 1) / g.S;
                        l.group = (uint32_t)(G.size() - ps.g0);
                        l.slot = i;
                        ps.nsamp = l.samp0 + l.nsamp;
                        ps.max_list_samp = std::max(ps.max_list_samp, l.nsamp);
                        g.n += l.size;
                        L.push_back(l);
                    }
                    g.ntiles = (g.n + kKG - 1) / kKG;
                    ps.ntiles += g.ntiles;
                    ps.n += g.n;
                    G.push_back(g);
                    next.push_back({g.out, g.n});
                }
                R.swap(next);
                any |= R.size() > 1;
            }
            ps.ng = G.size() - ps.g0;
            ps.nl = L.size() - ps.l0;
            passes.push_back(ps);
        }
        uint64_t max_samp = 0, max_tiles = 0;
        for (const Pass& ps : passes) {
            max_samp = std::max(max_samp, ps.nsamp);
            max_tiles = std::max(max_tiles, ps.ntiles);
        }
        const uint64_t gbytes = (G.size() * sizeof(KGroup) + 255) / 256 * 256;
        const uint64_t lbytes = (L.size() * sizeof(KList) + 255) / 256 * 256;
        const uint64_t cbytes = (max_samp * F * 4 + 255) / 256 * 256;
        const uint64_t rbytes = (max_samp * 8 + 255) / 256 * 256;
        const uint64_t tbytes = (max_tiles * 8 + 255) / 256 * 256;
        const uint64_t cobytes = (max_tiles * F * 4 + 255) / 256 * 256;
        P.m_work.reserve(gbytes + lbytes + cbytes + rbytes + tbytes + cobytes + max_tiles * 4 + 256);
        uint8_t* w = P.m_work.as<uint8_t>();
        KGroup* dG = reinterpret_cast<KGroup*>(w);
        KList* dL = reinterpret_cast<KList*>(w + gbytes);
        uint32_t* corank = reinterpret_cast<uint32_t*>(w + gbytes + lbytes);
        uint64_t* rankv = reinterpret_cast<uint64_t*>(w + gbytes + lbytes + cbytes);
        uint64_t* tile_b = reinterpret_cast<uint64_t*>(w + gbytes + lbytes + cbytes + rbytes);
        uint32_t* tile_co = reinterpret_cast<uint32_t*>(w + gbytes + lbytes + cbytes + rbytes + tbytes);
        uint32_t* tile_g = reinterpret_cast<uint32_t*>(w + gbytes + lbytes + cbytes + rbytes + tbytes + cobytes);
        DSS_HIP(hipMemcpyAsync(dG, G.data(), G.size() * sizeof(KGroup), hipMemcpyHostToDevice, st));
        DSS_HIP(hipMemcpyAsync(dL, L.data(), L.size() * sizeof(KList), hipMemcpyHostToDevice, st));
        // pass 0 reads the receive order and LCPs directly; pass r writes the
        // buffer pair pass r - 1 did not; the last also writes the saved LCPs
        const uint32_t* in_ord = nullptr;
        const uint32_t* in_lc = P.recv_lcp.as<uint32_t>();
        for (size_t r = 0; r < passes.size(); ++r) {
            const Pass& ps = passes[r];
            const bool last = r + 1 == passes.size();
            const double bytes = ps.n * (12.0 + (in_ord ? 4.0 : 0.0) + 8.0 + (last ? 8.0 : 0.0));
            KScope ks(s->timer, st, "merge", bytes);
            for (int coarse = 1; coarse >= 0; --coarse) {
                const uint64_t per_list = coarse ? (ps.max_list_samp + kKCoarse - 1) / kKCoarse : ps.max_list_samp;
                const uint64_t per_block = 256 >> lgF;
                hipLaunchKernelGGL(k_kmerge_sample, dim3((uint32_t)((per_list + per_block - 1) / per_block),
                                                         (uint32_t)ps.nl), dim3(256), 0, st,
                                   dG + ps.g0, dL + ps.l0, lgF, coarse, in_ord,
                                   s->arena(), P.recv_view_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(),
                                   P.m_keyc.as<uint64_t>(), P.m_keyc2.as<uint64_t>(), cp, corank, rankv);
            }
            hipLaunchKernelGGL(k_kmerge_bounds, dim3((uint32_t)((ps.ntiles * kKMaxFan + 255) / 256)), dim3(256), 0,
                               st, dG + ps.g0, (uint32_t)ps.ng, ps.ntiles, dL + ps.l0, lgF, corank, rankv, tile_b,
                               tile_co, tile_g);
            hipLaunchKernelGGL(k_kmerge_tile, dim3((uint32_t)ps.ntiles), dim3(kKT), 0, st, dG + ps.g0,
                               (uint32_t)ps.ng, dL + ps.l0, lgF, tile_b, tile_co, tile_g, in_ord, in_lc, ord, lc,
                               s->arena(),
                               P.recv_view_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(),
                               P.m_keyc.as<uint64_t>(), P.m_keyc2.as<uint64_t>(), cp, P.recv_lcp.as<uint32_t>(),
                               last ? P.m_saved.as<uint32_t>() : nullptr);
            hipLaunchKernelGGL(k_kmerge_fix, dim3((uint32_t)((ps.ntiles + 255) / 256)), dim3(256), 0, st, dG + ps.g0,
                               (uint32_t)ps.ng, ps.ntiles, tile_b, ord, lc, s->arena(),
                               P.recv_view_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(),
                               P.m_keyc.as<uint64_t>(), P.m_keyc2.as<uint64_t>(), cp);
            in_ord = ord;
            in_lc = lc;
            std::swap(ord, ord2);
            std::swap(lc, lc2);
        }
        DSS_HIP(hipGetLastError());
        // results in the last pass's output pair
        ord = const_cast<uint32_t*>(in_ord);
        lc = const_cast<uint32_t*>(in_lc);
    } else if (n) {
        // no receiver holds two runs: the receive order is the merged order
        KScope ks(s->timer, st, "merge_init", n * 16.0);
        hipLaunchKernelGGL(k_iota_copy, dim3((n + 255) / 256), dim3(256), 0, st, ord, lc, P.recv_lcp.as<uint32_t>(), n);
        hipLaunchKernelGGL(k_gather_saved, dim3((n + 255) / 256), dim3(256), 0, st, ord, P.recv_lcp.as<uint32_t>(),
                           n, P.m_saved.as<uint32_t>());
    }
    // results live in (ord, lc); make them the canonical buffers
    if (ord != P.m_order.as<uint32_t>()) {
        std::swap(P.m_order.p, P.m_order2.p);
        std::swap(P.m_order.bytes, P.m_order2.bytes);
        std::swap(P.m_lcp.p, P.m_lcp2.p);
        std::swap(P.m_lcp.bytes, P.m_lcp2.bytes);
    }
    DSS_HIP(hipGetLastError());
    P.merged = true;
    P.merged_by_caller = false;
    P.perm_valid = false;
    P.extended = false;
}

// StringIndexPEIndex of every merged output position: the permutation
// writeback of DistributedPrefixDoublingSort::sort (prefix_doubling.hpp:
// 282-290).  The merged order holds receive indices; receive index a lies in
// receiver q's range and in its run from source PE r (the last run starting
// at or before a: empty runs share their successor's start), so its origin
// is PE r, sorted position send_start[r][q] + (a - run_off[q][r]).  Up to
// kXiP PEs the tables are staged in LDS.
__global__ void k_permutation_out(const uint32_t* __restrict__ ord, uint64_t n, uint32_t p,
                                  const uint64_t* __restrict__ send_start, const uint64_t* __restrict__ run_off,
                                  const uint64_t* __restrict__ recv_begin, uint32_t* __restrict__ si,
                                  uint32_t* __restrict__ pi) {
    __shared__ uint64_t s_rb[kXiP + 1], s_ss[kXiP * kXiP], s_ro[kXiP * kXiP];
    const bool staged = p <= kXiP;  // block-uniform
    if (staged) {
        for (uint32_t t = threadIdx.x; t <= p; t += blockDim.x) s_rb[t] = recv_begin[t];
        for (uint32_t t = threadIdx.x; t < p * p; t += blockDim.x) {
            s_ss[t] = send_start[t];
            s_ro[t] = run_off[t];
        }
        __syncthreads();
    }
    const uint64_t* rb = staged ? s_rb : recv_begin;
    const uint64_t* ss = staged ? s_ss : send_start;
    const uint64_t* ro = staged ? s_ro : run_off;
    const uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    uint32_t a4[4] = {0, 0, 0, 0};
    if (base + 4 <= n) {
        const uint4 v = *reinterpret_cast<const uint4*>(ord + base);
        a4[0] = v.x; a4[1] = v.y; a4[2] = v.z; a4[3] = v.w;
    } else {
        for (uint32_t t = 0; t < 4 && base + t < n; ++t) a4[t] = ord[base + t];
    }
    uint32_t o_si[4], o_pe[4];
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
        const uint64_t a = a4[t];
        uint32_t lo = 0, hi = p;  // receiver: last q with rb[q] <= a
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (rb[mid] <= a) lo = mid; else hi = mid;
        }
        const uint32_t q = lo;
        const uint64_t* row = ro + (uint64_t)q * p;
        lo = 0;
        hi = p;  // run: last r with row[r] <= a
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (row[mid] <= a) lo = mid; else hi = mid;
        }
        o_pe[t] = lo;
        o_si[t] = (uint32_t)(ss[(uint64_t)lo * p + q] + (a - row[lo]));
    }
    if (base + 4 <= n) {
        *reinterpret_cast<uint4*>(si + base) = make_uint4(o_si[0], o_si[1], o_si[2], o_si[3]);
        *reinterpret_cast<uint4*>(pi + base) = make_uint4(o_pe[0], o_pe[1], o_pe[2], o_pe[3]);
    } else {
        for (uint32_t t = 0; t < 4 && base + t < n; ++t) {
            si[base + t] = o_si[t];
            pi[base + t] = o_pe[t];
        }
    }
}

void permutation(dss_set_s* s) {
    PipeState& P = s->pp;
    DSS_REQUIRE(P.merged, DSS_ESTATE, "no merged output");
    if (P.perm_valid) return;
    const uint64_t n = s->n, n4 = perm_pe_off(n);
    const uint32_t p = s->p;
    P.perm_out.reserve(n4 * 8);
    if (n) {
        const uint64_t* ls = P.layout.as<uint64_t>();
        KScope ks(s->timer, s->stream, "permutation", n * 12.0);
        hipLaunchKernelGGL(k_permutation_out, dim3((uint32_t)((n + 1023) / 1024)), dim3(256), 0, s->stream,
                           P.m_order.as<uint32_t>(), n, p, ls, ls + (uint64_t)p * p, recv_begin_dev(P, p),
                           P.perm_out.as<uint32_t>(), P.perm_out.as<uint32_t>() + n4);
        DSS_HIP(hipGetLastError());
    }
    P.perm_valid = true;
}

// ================================================================== (a9)
// StringLcpContainer::extendPrefix (src/strings/stringcontainer.hpp:247-319)
// on the merged receive arena of every PE: output string i = the first
// saved[i] characters of output string i-1, then its stored characters.
// Char c < saved[i] of string i comes from the stored part of the nearest
// k <= i-1 whose saved[k] <= c; following "previous strictly smaller saved"
// links (psv) from i visits exactly those sources, so every string is
// rebuilt independently: chars [saved[psv], cur) come from stored(psv).
// psv is computed per 1024-position block with a stack, falling back to a
// backward walk over earlier blocks (block minima) for open positions.
constexpr uint32_t kPsvB = 1024;
constexpr uint32_t kPsvG = 32;  // positions per group minimum

// Per block of kPsvB positions its minimum (bmin) and the minima of its
// groups of kPsvG positions (gmin), for the look-back of k_psv.
__global__ void __launch_bounds__(256)
k_block_min(const uint32_t* __restrict__ saved, const uint64_t* __restrict__ recv_begin, uint32_t p, uint64_t n,
            uint32_t* __restrict__ bmin, uint32_t* __restrict__ gmin) {
    __shared__ uint32_t red[4];
    const uint64_t b = (uint64_t)blockIdx.x * kPsvB;
    uint32_t m = 0xFFFFFFFFu;
    for (uint32_t t = threadIdx.x; t < kPsvB; t += blockDim.x) {
        const uint64_t i = b + t;
        uint32_t x = i < n ? saved[i] : 0xFFFFFFFFu;
        // the group's minimum over its kPsvG consecutive lanes
        for (uint32_t o = kPsvG / 2; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor(x, o, kPsvG));
        if ((t & (kPsvG - 1)) == 0) gmin[(uint64_t)blockIdx.x * (kPsvB / kPsvG) + t / kPsvG] = x;
        m = min(m, x);
    }
    for (int o = 32; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_down(m, o));
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) bmin[blockIdx.x] = min(min(red[0], red[1]), min(red[2], red[3]));
}

// psv[i] = max{k < i in the same receiver : saved[k] < saved[i]}, or ~0
// when saved[i] == 0 or no such k exists (the receiver's first string).  A
// block of kPsvB positions builds a min-tree of its values in LDS; each
// position climbs from its leaf until a left sibling holds a smaller value
// and descends into its rightmost smaller leaf (2 log2 kPsvB steps).
// Positions whose block holds none look back over earlier blocks by their
// minima and scan the first block that holds one from its end.
__global__ void __launch_bounds__(256)
k_psv(const uint32_t* __restrict__ saved, const uint64_t* __restrict__ recv_begin, uint32_t p, uint64_t n,
      const uint32_t* __restrict__ bmin, const uint32_t* __restrict__ gmin, uint64_t* __restrict__ psv) {
    __shared__ uint32_t tree[2 * kPsvB];  // node j: min of its subtree; leaves kPsvB + t
    const uint64_t b = (uint64_t)blockIdx.x * kPsvB;
    for (uint32_t t = threadIdx.x; t < kPsvB; t += blockDim.x) tree[kPsvB + t] = b + t < n ? saved[b + t] : 0u;
    __syncthreads();
    for (uint32_t lvl = kPsvB / 2; lvl >= 1; lvl >>= 1) {
        for (uint32_t j = lvl + threadIdx.x; j < 2 * lvl; j += blockDim.x) tree[j] = min(tree[2 * j], tree[2 * j + 1]);
        __syncthreads();
    }
    for (uint32_t t = threadIdx.x; t < kPsvB; t += blockDim.x) {
        const uint64_t i = b + t;
        if (i >= n) break;
        uint32_t lo = 0, hi = p;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (recv_begin[mid] <= i) lo = mid + 1; else hi = mid;
        }
        const uint64_t rb = recv_begin[lo - 1];
        const uint32_t v = tree[kPsvB + t];
        uint64_t res = ~0ull;
        if (v != 0 && i > rb) {
            uint32_t node = kPsvB + t;
            while (node > 1) {
                if ((node & 1) && tree[node - 1] < v) {  // the left sibling holds a smaller value
                    node -= 1;
                    while (node < kPsvB) node = tree[2 * node + 1] < v ? 2 * node + 1 : 2 * node;
                    res = b + (node - kPsvB);
                    break;
                }
                node >>= 1;
            }
            if (res == ~0ull) {
                // earlier blocks: skip those whose minimum is not smaller
                int64_t blk = (int64_t)(b / kPsvB) - 1;
                while (blk >= 0 && (uint64_t)(blk + 1) * kPsvB > rb) {
                    if (bmin[blk] < v) {
                        // the last group holding a smaller value, then its last such position
                        for (int32_t gi = (int32_t)(kPsvB / kPsvG) - 1; gi >= 0 && res == ~0ull; --gi) {
                            if (gmin[(uint64_t)blk * (kPsvB / kPsvG) + gi] >= v) continue;
                            const uint64_t g0 = (uint64_t)blk * kPsvB + (uint64_t)gi * kPsvG;
                            for (uint64_t e = g0 + kPsvG; e > g0;) {
                                --e;
                                if (saved[e] < v) { res = e; break; }
                            }
                        }
                        break;  // the block holds one: found
                    }
                    --blk;
                }
            }
            if (res != ~0ull && res < rb) res = ~0ull;  // nearest smaller lies in an earlier receiver
        }
        psv[i] = res;
    }
}

__global__ void k_out_size(const uint32_t* __restrict__ ord, const uint32_t* __restrict__ saved,
                           const uint64_t* __restrict__ rsize, uint64_t n, uint64_t* __restrict__ out_size) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out_size[i] = saved[i] + rsize[ord[i]];  // prefix + stored chars incl. NUL
}

// Copies n bytes from base[sa ..] to out[oa ..] with L lanes (this is lane
// l): the lanes own the aligned 8-byte words of [oa, oa + n); a word wholly
// inside is built from two aligned source loads and stored at once, the edge
// words byte by byte (their other bytes belong to other copies).
__device__ inline void copy_span(uint8_t* __restrict__ out, uint64_t oa, const uint8_t* __restrict__ base, uint64_t sa,
                                 uint64_t n, uint32_t l, uint32_t L) {
    const uint64_t e = oa + n;
    for (uint64_t w = (oa >> 3) + l; w < (e + 7) >> 3; w += L) {
        const uint64_t g = w << 3;
        if (g >= oa && g + 8 <= e) {
            *reinterpret_cast<uint64_t*>(out + g) = load_word_le(base, sa + (g - oa));
        } else {
            for (uint64_t x = max(g, oa); x < min(g + 8, e); ++x) out[x] = base[sa + (x - oa)];
        }
    }
}

// 16 lanes per output string.
__global__ void k_extend(const uint32_t* __restrict__ ord, const uint32_t* __restrict__ saved,
                         const uint64_t* __restrict__ psv, const uint64_t* __restrict__ rbyte,
                         const uint64_t* __restrict__ rsize, const uint8_t* __restrict__ rbytes,
                         const uint64_t* __restrict__ out_off, uint64_t n, uint8_t* __restrict__ out) {
    const uint32_t l = threadIdx.x & 15;
    for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; i < n;
         i += ((uint64_t)gridDim.x * blockDim.x) >> 4) {
        const uint64_t o0 = out_off[i];
        uint32_t cur = saved[i];
        // own stored characters (incl. NUL) after the prefix
        copy_span(out, o0 + cur, rbytes, rbyte[ord[i]], rsize[ord[i]], l, 16);
        uint64_t k = psv[i];
        while (cur > 0 && k != ~0ull) {
            const uint32_t sk = saved[k];
            copy_span(out, o0 + sk, rbytes, rbyte[ord[k]], cur - sk, l, 16);
            cur = sk;
            k = psv[k];
        }
    }
}

void extend_prefix(dss_set_s* s) {
    PipeState& P = s->pp;
    DSS_REQUIRE(P.merged, DSS_ESTATE, "dss_merge has not run");
    DSS_REQUIRE(!P.merged_by_caller, DSS_ESTATE, "a merged output set by dss_set_merged has no saved LCPs");
    DSS_REQUIRE(P.mode == 1, DSS_ESTATE, "prefix decompression needs the LCP-compressed exchange (mode 1)");
    const hipStream_t st = s->stream;
    const uint64_t n = s->n;
    const uint32_t p = s->p;
    const uint64_t nn = std::max<uint64_t>(n, 1);
    const uint64_t nblk = (n + kPsvB - 1) / kPsvB;
    P.out_psv.reserve(nn * 8);
    P.out_off.reserve((nn + 1) * 8);
    P.out_aux.reserve((nblk + 1) * 4 * (1 + kPsvB / kPsvG) + (p + 1) * 8 + 64);
    uint32_t* bmin = P.out_aux.as<uint32_t>();
    uint32_t* gmin = bmin + (nblk + 1);
    uint64_t* rb_d = reinterpret_cast<uint64_t*>(
        P.out_aux.as<uint8_t>() + ((nblk + 1) * 4 * (1 + kPsvB / kPsvG) + 7) / 8 * 8);
    DSS_HIP(hipMemcpyAsync(rb_d, P.recv_begin.data(), (p + 1) * 8, hipMemcpyHostToDevice, st));
    uint64_t total = 0;
    if (n) {
        KScope ks(s->timer, st, "extend_psv", n * 16.0);
        hipLaunchKernelGGL(k_block_min, dim3(nblk), dim3(256), 0, st, P.m_saved.as<uint32_t>(), rb_d, p, n, bmin,
                           gmin);
        hipLaunchKernelGGL(k_psv, dim3(nblk), dim3(256), 0, st, P.m_saved.as<uint32_t>(), rb_d, p, n, bmin, gmin,
                           P.out_psv.as<uint64_t>());
        hipLaunchKernelGGL(k_out_size, dim3((n + 255) / 256), dim3(256), 0, st, P.m_order.as<uint32_t>(),
                           P.m_saved.as<uint32_t>(), P.recv_size.as<uint64_t>(), n, P.out_off.as<uint64_t>());
    }
    exclusive_scan_u64(P.out_off.as<uint64_t>(), P.out_off.as<uint64_t>(), n, st, P, &total);
    P.out_nbytes = total;
    P.out_bytes.reserve(total + kArenaPad);
    if (n) {
        KScope ks(s->timer, st, "extend_copy", (double)total * 2.0 + n * 24.0);
        hipLaunchKernelGGL(k_extend, dim3(stride_grid(n * 16, 256)), dim3(256), 0, st, P.m_order.as<uint32_t>(),
                           P.m_saved.as<uint32_t>(), P.out_psv.as<uint64_t>(), P.recv_byte_off.as<uint64_t>(),
                           P.recv_size.as<uint64_t>(), P.recv_bytes.as<uint8_t>(), P.out_off.as<uint64_t>(), n,
                           P.out_bytes.as<uint8_t>());
    }
    DSS_HIP(hipGetLastError());
    DSS_HIP(hipStreamSynchronize(st));
    P.extended = true;
}

}  // namespace dss

using namespace dss;

extern "C" {

int dss_avg_lcp(dss_set s, uint64_t* out) {
    return guarded([&] {
        DSS_REQUIRE(s && out, DSS_EINVAL, "NULL argument");
        *out = avg_lcp(s);
    });
}

int dss_partition(dss_set s, int sampler, uint64_t sampling_factor, uint64_t max_len, int use_dist,
                  uint64_t* interval) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        partition(s, sampler, sampling_factor, max_len, use_dist != 0);
        if (interval) memcpy(interval, s->pp.interval.data(), s->pp.interval.size() * 8);
    });
}

int dss_exchange(dss_set s, int mode, uint64_t* recv_begin, uint64_t* byte_begin, uint64_t* recv_nbytes) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        exchange(s, mode);
        if (recv_begin) memcpy(recv_begin, s->pp.recv_begin.data(), (s->p + 1) * 8);
        if (byte_begin) s->pp.mbox.fetch(s->stream, byte_begin, byte_begin_dev(s->pp, s->p), (s->p + 1) * 8);
        if (recv_nbytes) *recv_nbytes = s->pp.recv_nbytes;
    });
}

int dss_get_exchange(dss_set s, uint8_t* bytes, uint32_t* lcp, uint32_t* src_pe, uint32_t* src_pos) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        PipeState& P = s->pp;
        DSS_REQUIRE(P.exchanged, DSS_ESTATE, "dss_exchange has not run");
        const hipStream_t st = s->stream;
        if (bytes && P.recv_nbytes)
            DSS_HIP(hipMemcpyAsync(bytes, P.recv_bytes.p, P.recv_nbytes, hipMemcpyDeviceToHost, st));
        if (s->n) {
            if (lcp) DSS_HIP(hipMemcpyAsync(lcp, P.recv_lcp.p, s->n * 4, hipMemcpyDeviceToHost, st));
            if (src_pe || src_pos) recv_origin(s);
            if (src_pe) DSS_HIP(hipMemcpyAsync(src_pe, P.recv_src_pe.p, s->n * 4, hipMemcpyDeviceToHost, st));
            if (src_pos) DSS_HIP(hipMemcpyAsync(src_pos, P.recv_src_pos.p, s->n * 4, hipMemcpyDeviceToHost, st));
        }
        DSS_HIP(hipStreamSynchronize(st));
    });
}

int dss_merge_lcp(dss_set s) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        merge(s);
    });
}

int dss_get_merged(dss_set s, uint32_t* order, uint32_t* lcp, uint32_t* saved) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        PipeState& P = s->pp;
        DSS_REQUIRE(P.merged, DSS_ESTATE, "dss_merge_lcp has not run");
        const hipStream_t st = s->stream;
        if (s->n) {
            if (order) DSS_HIP(hipMemcpyAsync(order, P.m_order.p, s->n * 4, hipMemcpyDeviceToHost, st));
            if (lcp) DSS_HIP(hipMemcpyAsync(lcp, P.m_lcp.p, s->n * 4, hipMemcpyDeviceToHost, st));
            if (saved) DSS_HIP(hipMemcpyAsync(saved, P.m_saved.p, s->n * 4, hipMemcpyDeviceToHost, st));
        }
        DSS_HIP(hipStreamSynchronize(st));
    });
}

int dss_extend_prefix(dss_set s, uint64_t* out_nbytes) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        extend_prefix(s);
        if (out_nbytes) *out_nbytes = s->pp.out_nbytes;
    });
}

int dss_get_extended(dss_set s, uint8_t* bytes, uint64_t* offsets) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        PipeState& P = s->pp;
        DSS_REQUIRE(P.extended, DSS_ESTATE, "dss_extend_prefix has not run");
        const hipStream_t st = s->stream;
        if (bytes && P.out_nbytes) DSS_HIP(hipMemcpyAsync(bytes, P.out_bytes.p, P.out_nbytes, hipMemcpyDeviceToHost, st));
        if (offsets && s->n) DSS_HIP(hipMemcpyAsync(offsets, P.out_off.p, s->n * 8, hipMemcpyDeviceToHost, st));
        DSS_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"

// ================================================================== (a14)
// is_complete_and_sorted / CheckerWithCompleteExchange (src/mpi/is_sorted.hpp:
// 27-198) on the merged output of all PEs, plus D of the globally sorted
// sequence (computeD, src/tests/DNGenerator_tests.cpp:22-45).  The global
// order is the concatenation of the PEs' merged outputs; the LCP across a PE
// boundary is recomputed.  Every received string must be a distinct origin
// position (completeness) and every merged LCP must equal the recomputed one.
namespace dss {

__global__ void k_check(const uint32_t* __restrict__ ord, const uint32_t* __restrict__ mlcp,
                        const uint64_t* __restrict__ voff, const uint32_t* __restrict__ vlen,
                        const uint32_t* __restrict__ rpe, const uint32_t* __restrict__ rpos,
                        const uint64_t* __restrict__ pe_begin, const uint64_t* __restrict__ recv_begin, uint32_t p,
                        const uint8_t* __restrict__ chars, uint64_t n, uint32_t* __restrict__ glcp,
                        unsigned int* __restrict__ seen, unsigned long long* __restrict__ cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t a = ord[i];
    // completeness: mark the origin position
    const uint64_t origin = pe_begin[rpe[a]] + rpos[a];
    const unsigned int bit = 1u << (origin & 31);
    const unsigned int old = atomicOr(&seen[origin >> 5], bit);
    if (old & bit) atomicAdd(&cnt[2], 1ull);  // duplicate origin
    if (i == 0) {
        glcp[0] = 0;
        return;
    }
    const uint32_t b = ord[i - 1];
    uint32_t l;
    const int c = view_cmp32(chars, voff[b], vlen[b], voff[a], vlen[a], &l);
    if (c > 0) atomicAdd(&cnt[0], 1ull);  // order violation
    uint32_t lo = 0, hi = p;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (recv_begin[mid] <= i) lo = mid + 1; else hi = mid;
    }
    const bool pe_start = recv_begin[lo - 1] == i;
    if (!pe_start && mlcp[i] != l) atomicAdd(&cnt[1], 1ull);  // LCP mismatch
    glcp[i] = l;
}

// Sum of per-thread values over a block (256 threads), one atomic per block;
// every thread of the block calls it.
__device__ inline void block_add(uint64_t t, unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[4];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down((unsigned long long)t, o);
    if (lane_id() == 0) part[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long b = part[0] + part[1] + part[2] + part[3];
        if (b) atomicAdd(out, b);
    }
    __syncthreads();  // part[] is reused by the next call
}

constexpr uint32_t kSumBlocks = 2048;  // grid of the two reductions below (256 threads each)

// Σ (max(glcp[i], glcp[i+1]) + 1): the distinguishing-prefix total implied by
// the sorted order's LCPs.
__global__ void __launch_bounds__(256) k_dsum(const uint32_t* __restrict__ glcp, uint64_t n,
                                              unsigned long long* __restrict__ out) {
    uint64_t t = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        t += (uint64_t)(i + 1 < n ? max(glcp[i], glcp[i + 1]) : glcp[i]) + 1;
    block_add(t, out);
}

__global__ void __launch_bounds__(256) k_popc(const unsigned int* __restrict__ w, uint64_t nw,
                                              unsigned long long* __restrict__ out) {
    uint64_t t = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (uint64_t)gridDim.x * blockDim.x)
        t += (uint64_t)__popc(w[i]);
    block_add(t, out);
}

// ExcatDistinguishingPrefix::filter_exact (src/sorter/distributed/
// bloomfilter.hpp:847-965) on the globally sorted output: the reference
// allgathers every candidate string, sorts them and gives each string
// 1 + the larger LCP with its neighbours in that order.  The merge already
// holds the global order, so one lane per output position compares its full
// string (through its origin, not the possibly truncated received view) with
// its predecessor, and a second pass takes the maximum of both neighbours.
__global__ void k_full_glcp(const uint32_t* __restrict__ ord, const uint32_t* __restrict__ rpe,
                            const uint32_t* __restrict__ rpos, const uint64_t* __restrict__ pe_begin,
                            const uint64_t* __restrict__ s_off, const uint32_t* __restrict__ s_len,
                            const uint8_t* __restrict__ chars, uint64_t n, uint32_t* __restrict__ glcp,
                            uint64_t* __restrict__ gpos) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t a = ord[i];
    const uint64_t pa = pe_begin[rpe[a]] + rpos[a];
    gpos[i] = pa;
    if (i == 0) {
        glcp[0] = 0;
        return;
    }
    const uint32_t b = ord[i - 1];
    const uint64_t pb = pe_begin[rpe[b]] + rpos[b];
    uint32_t l = 0;
    (void)view_cmp32(chars, s_off[pb], s_len[pb], s_off[pa], s_len[pa], &l);
    glcp[i] = l;
}

__global__ void k_exact_dist(const uint32_t* __restrict__ glcp, const uint64_t* __restrict__ gpos, uint64_t n,
                             uint32_t* __restrict__ exact) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t m = 0;
    if (i > 0) m = glcp[i];
    if (i + 1 < n) m = max(m, glcp[i + 1]);
    exact[gpos[i]] = n < 2 ? 0 : m + 1;  // one string: no neighbour pair
}

// Bounds of prefix doubling's results against the exact lengths: cnt[0]
// strings with results below DIST (below DIST - 1 for a string whose result
// is its whole length), cnt[1] strings cut before their end whose result is
// not below 2 * DIST + 8 (the doubling schedule's bound when no hashes
// collide), cnt[2] sum of DIST, cnt[3] sum of results.
__global__ void __launch_bounds__(256) k_dist_bounds(const uint32_t* __restrict__ exact,
                                                     const uint32_t* __restrict__ res,
                                                     const uint32_t* __restrict__ s_len, uint64_t n,
                                                     unsigned long long* __restrict__ cnt) {
    uint64_t lo = 0, hi = 0, sd = 0, sr = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t d = exact[i], r = res[i], L = s_len[i];
        lo += d > r + (r >= L ? 1 : 0);
        hi += r < L && r >= 2 * d + 8;
        sd += d;
        sr += r;
    }
    block_add(lo, &cnt[0]);
    block_add(hi, &cnt[1]);
    block_add(sd, &cnt[2]);
    block_add(sr, &cnt[3]);
}

static void exact_dist(dss_set_s* s) {
    PipeState& P = s->pp;
    DSS_REQUIRE(P.merged, DSS_ESTATE, "no merged output (dss_merge_sort / dss_prefix_doubling_sort)");
    const uint64_t n = s->n;
    const uint64_t nn = std::max<uint64_t>(n, 1);
    P.exact.reserve(nn * 4);
    P.exact_glcp.reserve(nn * 4);
    P.exact_pos.reserve(nn * 8);
    if (!n) return;
    const hipStream_t st = s->stream;
    recv_origin(s);
    KScope ks(s->timer, st, "exact_dist", n * 48.0);
    hipLaunchKernelGGL(k_full_glcp, dim3((n + 255) / 256), dim3(256), 0, st, P.m_order.as<uint32_t>(),
                       P.recv_src_pe.as<uint32_t>(), P.recv_src_pos.as<uint32_t>(), s->pe_begin_d.as<uint64_t>(),
                       s->s_off.as<uint64_t>(), s->s_len.as<uint32_t>(), s->arena(), n, P.exact_glcp.as<uint32_t>(),
                       P.exact_pos.as<uint64_t>());
    hipLaunchKernelGGL(k_exact_dist, dim3((n + 255) / 256), dim3(256), 0, st, P.exact_glcp.as<uint32_t>(),
                       P.exact_pos.as<uint64_t>(), n, P.exact.as<uint32_t>());
    DSS_HIP(hipGetLastError());
}

}  // namespace dss

extern "C" int dss_exact_dist(dss_set s, uint32_t* exact) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        exact_dist(s);
        if (s->n && exact)
            DSS_HIP(hipMemcpyAsync(exact, s->pp.exact.p, s->n * 4, hipMemcpyDeviceToHost, s->stream));
        DSS_HIP(hipStreamSynchronize(s->stream));
    });
}

extern "C" int dss_check_dist(dss_set s, dss_dist_check_t* out) {
    return guarded([&] {
        DSS_REQUIRE(s && out, DSS_EINVAL, "NULL argument");
        PipeState& P = s->pp;
        DSS_REQUIRE(P.has_dist, DSS_ESTATE, "no prefix-doubling results");
        exact_dist(s);
        memset(out, 0, sizeof(*out));
        const uint64_t n = s->n;
        out->num_strings = n;
        if (!n) return;
        const hipStream_t st = s->stream;
        P.exact_cnt.reserve(32);
        DSS_HIP(hipMemsetAsync(P.exact_cnt.p, 0, 32, st));
        hipLaunchKernelGGL(k_dist_bounds, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, kSumBlocks)), dim3(256), 0,
                           st, P.exact.as<uint32_t>(), P.dist.as<uint32_t>(), s->s_len.as<uint32_t>(), n,
                           P.exact_cnt.as<unsigned long long>());
        DSS_HIP(hipGetLastError());
        unsigned long long h[4];
        DSS_HIP(hipMemcpyAsync(h, P.exact_cnt.p, 32, hipMemcpyDeviceToHost, st));
        DSS_HIP(hipStreamSynchronize(st));
        out->lower_violations = h[0];
        out->upper_violations = h[1];
        out->exact_sum = h[2];
        out->results_sum = h[3];
    });
}

extern "C" int dss_check(dss_set s, dss_check_t* out) {
    return guarded([&] {
        DSS_REQUIRE(s && out, DSS_EINVAL, "NULL argument");
        PipeState& P = s->pp;
        DSS_REQUIRE(P.merged, DSS_ESTATE, "no merged output to check");
        const uint64_t n = s->n;
        const hipStream_t st = s->stream;
        memset(out, 0, sizeof(*out));
        out->num_strings = n;
        if (!n) {
            out->sorted = out->complete = 1;
            return;
        }
        const uint64_t nw = (n + 31) / 32;
        DevBuf buf;
        buf.reserve((n * 4 + nw * 4 + 7) / 8 * 8 + 64 + (s->p + 1) * 8);
        uint32_t* glcp = buf.as<uint32_t>();
        unsigned int* seen = reinterpret_cast<unsigned int*>(glcp + n);
        unsigned long long* cnt = reinterpret_cast<unsigned long long*>(buf.as<uint8_t>() + (n * 4 + nw * 4 + 7) / 8 * 8);
        uint64_t* rb = reinterpret_cast<uint64_t*>(cnt + 8);
        DSS_HIP(hipMemsetAsync(seen, 0, nw * 4, st));
        DSS_HIP(hipMemsetAsync(cnt, 0, 64, st));
        DSS_HIP(hipMemcpyAsync(rb, P.recv_begin.data(), (s->p + 1) * 8, hipMemcpyHostToDevice, st));
        recv_origin(s);
        hipLaunchKernelGGL(k_check, dim3((n + 255) / 256), dim3(256), 0, st, P.m_order.as<uint32_t>(),
                           P.m_lcp.as<uint32_t>(), P.recv_view_off.as<uint64_t>(), P.recv_view_len.as<uint32_t>(),
                           P.recv_src_pe.as<uint32_t>(), P.recv_src_pos.as<uint32_t>(), s->pe_begin_d.as<uint64_t>(),
                           rb, s->p, s->arena(), n, glcp, seen, cnt);
        hipLaunchKernelGGL(k_dsum, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, kSumBlocks)), dim3(256), 0, st,
                           glcp, n, cnt + 3);
        hipLaunchKernelGGL(k_popc, dim3((uint32_t)std::min<uint64_t>((nw + 255) / 256, kSumBlocks)), dim3(256), 0, st,
                           seen, nw, cnt + 4);
        DSS_HIP(hipGetLastError());
        unsigned long long h[5];
        DSS_HIP(hipMemcpyAsync(h, cnt, 40, hipMemcpyDeviceToHost, st));
        DSS_HIP(hipStreamSynchronize(st));
        out->order_errors = h[0];
        out->lcp_errors = h[1];
        out->sorted = h[0] == 0;
        out->complete = h[2] == 0 && h[4] == n;
        out->dist_prefix_size = h[3];
    });
}

extern "C" int dss_set_interval(dss_set s, const uint64_t* interval) {
    return guarded([&] {
        DSS_REQUIRE(s && interval, DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(s->sorted, DSS_ESTATE, "no local order (dss_sort_lcp or dss_set_local_order)");
        const uint32_t p = s->p;
        for (uint32_t r = 0; r < p; ++r) {
            uint64_t sum = 0;
            for (uint32_t q = 0; q < p; ++q) sum += interval[(uint64_t)r * p + q];
            DSS_REQUIRE(sum == s->pe_begin[r + 1] - s->pe_begin[r], DSS_EINVAL,
                        "interval row must sum to the PE's string count");
        }
        PipeState& P = s->pp;
        P.interval.assign(interval, interval + (uint64_t)p * p);
        P.interval_d.reserve((uint64_t)p * p * 8);
        DSS_HIP(hipMemcpyAsync(P.interval_d.p, interval, (uint64_t)p * p * 8, hipMemcpyHostToDevice, s->stream));
        DSS_HIP(hipStreamSynchronize(s->stream));
        P.partitioned = true;
        P.exchanged = P.merged = P.extended = false;
    });
}

extern "C" int dss_set_merged(dss_set s, const uint32_t* order, const uint32_t* lcp) {
    return guarded([&] {
        DSS_REQUIRE(s && (s->n == 0 || (order && lcp)), DSS_EINVAL, "NULL argument");
        PipeState& P = s->pp;
        DSS_REQUIRE(P.exchanged, DSS_ESTATE, "dss_exchange has not run");
        const uint64_t n = s->n, nn = std::max<uint64_t>(n, 1);
        // every entry must name a received string (the checker reads its
        // view and origin); repeats are allowed and reported as incomplete
        for (uint64_t i = 0; i < n; ++i)
            DSS_REQUIRE(order[i] < n, DSS_EINVAL, "merged order entry beyond the received strings");
        P.m_order.reserve(nn * 4);
        P.m_lcp.reserve(nn * 4);
        P.m_saved.reserve(nn * 4);
        const hipStream_t st = s->stream;
        if (n) {
            DSS_HIP(hipMemcpyAsync(P.m_order.p, order, n * 4, hipMemcpyHostToDevice, st));
            DSS_HIP(hipMemcpyAsync(P.m_lcp.p, lcp, n * 4, hipMemcpyHostToDevice, st));
            fill_small(st, P.m_saved.p, 0, n * 4);
        }
        DSS_HIP(hipStreamSynchronize(st));
        P.merged = true;
        P.merged_by_caller = true;
        P.perm_valid = false;
        P.extended = false;
    });
}

extern "C" int dss_set_dist(dss_set s, const uint32_t* dist) {
    return guarded([&] {
        DSS_REQUIRE(s && (s->n == 0 || dist), DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(s->sorted, DSS_ESTATE, "no local order (dss_sort_lcp or dss_set_local_order)");
        PipeState& P = s->pp;
        P.dist.reserve(std::max<uint64_t>(s->n, 1) * 4);
        if (s->n) DSS_HIP(hipMemcpyAsync(P.dist.p, dist, s->n * 4, hipMemcpyHostToDevice, s->stream));
        DSS_HIP(hipStreamSynchronize(s->stream));
        P.has_dist = true;
        P.stage_valid = false;  // the kept characters belong to the computed results
        P.partitioned = P.exchanged = P.merged = P.extended = false;
    });
}
