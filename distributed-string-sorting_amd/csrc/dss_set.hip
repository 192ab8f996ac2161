// Device string set: creation from host arenas and on-device generators,
// PE split, download, timing and error plumbing of the C ABI.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dss_device.hpp"
#include "dss_internal.hpp"

namespace dss {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

// ------------------------------------------------------------------ timing
hipEvent_t KernelTimer::get_event() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    DSS_HIP(hipEventCreate(&e));
    return e;
}

void KernelTimer::begin(hipStream_t st, const char* name, double alg_bytes, Rec& r) {
    r.name = name;
    r.bytes = alg_bytes;
    r.a = get_event();
    r.b = get_event();
    DSS_HIP(hipEventRecord(r.a, st));
}

void KernelTimer::end(hipStream_t st, Rec& r) {
    (void)hipEventRecord(r.b, st);
    pending.push_back(r);
}

void KernelTimer::collect() {
    for (auto& r : pending) {
        float ms = 0;
        (void)hipEventSynchronize(r.b);
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        Stat* st = nullptr;
        for (auto& s : stats)
            if (s.name == r.name) st = &s;
        if (!st) {
            stats.push_back(Stat{r.name});
            st = &stats.back();
        }
        st->launches++;
        st->ms += ms;
        st->bytes += r.bytes;
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
}

KernelTimer::~KernelTimer() {
    for (auto& r : pending) {
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
}

// ------------------------------------------------------------- read-backs
__global__ void __launch_bounds__(64)
k_mailbox(const uint32_t* __restrict__ src, uint32_t nwords, uint32_t* __restrict__ dst, uint64_t* flag,
          uint64_t seq) {
    for (uint32_t i = threadIdx.x; i < nwords; i += 64) dst[i] = src[i];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_fill_small(uint8_t* __restrict__ p, uint64_t n, uint8_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_small(hipStream_t st, void* p, uint8_t v, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_small, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 1024)), dim3(256), 0, st,
                       static_cast<uint8_t*>(p), (uint64_t)n, v);
    DSS_HIP(hipGetLastError());
}

Mailbox::~Mailbox() {
    if (flag) (void)hipHostFree(flag);
}

void Mailbox::fetch(hipStream_t st, void* dst, const void* dev, size_t bytes) {
    DSS_REQUIRE(bytes % 4 == 0, DSS_EINVAL, "read-back size");
    if (bytes > kBytes) {
        DSS_HIP(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, st));
        DSS_HIP(hipStreamSynchronize(st));
        return;
    }
    if (!flag) {
        void* h = nullptr;
        DSS_HIP(hipHostMalloc(&h, kBytes + 64, hipHostMallocCoherent | hipHostMallocMapped));
        flag = static_cast<uint64_t*>(h);
        flag[0] = 0;
    }
    const uint64_t s = ++seq;
    uint32_t* data = reinterpret_cast<uint32_t*>(flag + 8);
    hipLaunchKernelGGL(k_mailbox, dim3(1), dim3(64), 0, st, static_cast<const uint32_t*>(dev),
                       (uint32_t)(bytes / 4), data, flag, s);
    DSS_HIP(hipGetLastError());
    for (uint64_t it = 1;; ++it) {
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == s) break;
        if ((it & 1023) == 0) {
            const hipError_t e = hipStreamQuery(st);
            if (e == hipSuccess) {
                if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == s) break;
                throw Error{DSS_EHIP, "read-back: the stream finished without publishing"};
            }
            if (e != hipErrorNotReady)
                throw Error{DSS_EHIP, std::string("read-back: ") + hipGetErrorString(e)};
        }
        __builtin_ia32_pause();
    }
    if (bytes) std::memcpy(dst, data, bytes);
}

// ------------------------------------------------------------------ set
dss_set_s* new_set(uint32_t num_pes) {
    DSS_REQUIRE(num_pes >= 1 && num_pes <= (1u << 16), DSS_EINVAL, "num_pes must be in [1, 65536]");
    dss_set_s* s = new dss_set_s();
    s->p = num_pes;
    DSS_HIP(hipGetDevice(&s->device));
    hipError_t e = hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        throw Error{DSS_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
    }
    s->stream = s->own_stream;
    return s;
}

// Frees a set and its own stream (error paths of the constructors).
void free_set(dss_set_s* s) {
    hipStream_t st = s->own_stream;
    (void)hipStreamSynchronize(s->stream);
    delete s;
    (void)hipStreamDestroy(st);
}

// ------------------------------------------------------ arena indexing
constexpr uint32_t kIdxThreads = 256;
constexpr uint32_t kIdxBytesPerThread = 64;
constexpr uint64_t kIdxChunk = (uint64_t)kIdxThreads * kIdxBytesPerThread;

__global__ void k_count_nul(const uint8_t* __restrict__ chars, uint64_t nbytes,
                            uint64_t* __restrict__ block_cnt) {
    __shared__ uint32_t red[kIdxThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kIdxChunk + (uint64_t)threadIdx.x * kIdxBytesPerThread;
    uint32_t c = 0;
    for (uint32_t i = 0; i < kIdxBytesPerThread; ++i) {
        uint64_t a = base + i;
        if (a < nbytes && chars[a] == 0) ++c;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if (lane_id() == 0) red[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kIdxThreads / 64; ++w) t += red[w];
        block_cnt[blockIdx.x] = t;
    }
}

// Exclusive scan of n u64 values in place, one block of 1024 threads.
__global__ void k_scan_single(uint64_t* __restrict__ v, uint64_t n, uint64_t* __restrict__ total) {
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = t * per, e = (b + per < n) ? b + per : n;
    uint64_t s = 0;
    for (uint64_t i = b; i < e; ++i) s += v[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        uint64_t x = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint64_t run = t ? part[t - 1] : 0;
    for (uint64_t i = b; i < e; ++i) {
        uint64_t x = v[i];
        v[i] = run;
        run += x;
    }
    if (t == 1023 && total) *total = part[1023];
}

__global__ void k_write_nul_pos(const uint8_t* __restrict__ chars, uint64_t nbytes,
                                const uint64_t* __restrict__ block_off, uint64_t* __restrict__ nul_pos) {
    __shared__ uint32_t cnt[kIdxThreads];
    const uint64_t base = (uint64_t)blockIdx.x * kIdxChunk + (uint64_t)threadIdx.x * kIdxBytesPerThread;
    uint32_t c = 0;
    for (uint32_t i = 0; i < kIdxBytesPerThread; ++i) {
        uint64_t a = base + i;
        if (a < nbytes && chars[a] == 0) ++c;
    }
    cnt[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t t = 0; t < kIdxThreads; ++t) {
            uint32_t x = cnt[t];
            cnt[t] = run;
            run += x;
        }
    }
    __syncthreads();
    uint64_t k = block_off[blockIdx.x] + cnt[threadIdx.x];
    for (uint32_t i = 0; i < kIdxBytesPerThread; ++i) {
        uint64_t a = base + i;
        if (a < nbytes && chars[a] == 0) nul_pos[k++] = a;
    }
}

__global__ void k_offsets_from_nul(uint64_t* __restrict__ off, uint32_t* __restrict__ len,
                                   const uint64_t* __restrict__ nul_pos, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t s = i == 0 ? 0 : nul_pos[i - 1] + 1;
    len[i] = (uint32_t)(nul_pos[i] - s);
    // off is written in a second pass (it aliases nothing; nul_pos is separate)
    off[i] = s;
}

__global__ void k_max_len(const uint32_t* __restrict__ len, uint64_t n, uint32_t* __restrict__ out) {
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, len[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_down(m, o));
    if (lane_id() == 0) atomicMax(out, m);
}

// distribute_strings (src/mpi/readInput.hpp:57-103): the string starting at
// byte s belongs to the PE whose byte slice holds s-1 (s = 0 to PE 0), so
// pe_begin[r] = number of strings with offset <= o_r.
__global__ void k_pe_split(const uint64_t* __restrict__ off, uint64_t n, uint64_t nbytes,
                           uint32_t p, uint64_t* __restrict__ pe_begin) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > p) return;
    if (r == 0) { pe_begin[0] = 0; return; }
    if (r == p) { pe_begin[p] = n; return; }
    const uint64_t small = nbytes / p, larger = nbytes % p;
    const uint64_t o_r = r < larger ? r * (small + 1) : larger * (small + 1) + (r - larger) * small;
    uint64_t lo = 0, hi = n;  // first index with off > o_r
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (off[mid] <= o_r) lo = mid + 1; else hi = mid;
    }
    pe_begin[r] = lo;
}

uint64_t index_arena(dss_set_s* s) {
    const uint64_t nb = s->nbytes;
    if (nb == 0) {
        s->n = 0;
        return 0;
    }
    const uint64_t blocks = (nb + kIdxChunk - 1) / kIdxChunk;
    DevBuf bcnt, tot, nul;
    bcnt.reserve((blocks + 1) * sizeof(uint64_t));
    tot.reserve(sizeof(uint64_t));
    hipLaunchKernelGGL(k_count_nul, dim3(blocks), dim3(kIdxThreads), 0, s->stream,
                       s->chars.as<uint8_t>(), nb, bcnt.as<uint64_t>());
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s->stream, bcnt.as<uint64_t>(),
                       blocks, tot.as<uint64_t>());
    uint64_t n = 0;
    DSS_HIP(hipMemcpyAsync(&n, tot.p, sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
    DSS_HIP(hipStreamSynchronize(s->stream));
    DSS_REQUIRE(n < 0xFFFFFFFFull, DSS_ERANGE, "more than 2^32-1 strings");
    s->n = n;
    s->off.reserve(std::max<uint64_t>(n, 1) * sizeof(uint64_t));
    s->len.reserve(std::max<uint64_t>(n, 1) * sizeof(uint32_t));
    nul.reserve(std::max<uint64_t>(n, 1) * sizeof(uint64_t));
    if (n) {
        hipLaunchKernelGGL(k_write_nul_pos, dim3(blocks), dim3(kIdxThreads), 0, s->stream,
                           s->chars.as<uint8_t>(), nb, bcnt.as<uint64_t>(), nul.as<uint64_t>());
        hipLaunchKernelGGL(k_offsets_from_nul, dim3((n + 255) / 256), dim3(256), 0, s->stream,
                           s->off.as<uint64_t>(), s->len.as<uint32_t>(), nul.as<uint64_t>(), n);
    }
    DSS_HIP(hipGetLastError());
    DSS_HIP(hipStreamSynchronize(s->stream));
    return n;
}

void compute_pe_split(dss_set_s* s) {
    s->pe_begin_d.reserve((s->p + 1) * sizeof(uint64_t));
    DevBuf dummy;
    const uint64_t* off = s->off.as<uint64_t>();
    if (!off) {
        dummy.reserve(8);
        off = dummy.as<uint64_t>();
    }
    hipLaunchKernelGGL(k_pe_split, dim3((s->p + 1 + 255) / 256), dim3(256), 0, s->stream, off, s->n,
                       s->slice_bytes ? s->slice_bytes : s->nbytes, s->p, s->pe_begin_d.as<uint64_t>());
    s->pe_begin.assign(s->p + 1, 0);
    DSS_HIP(hipMemcpyAsync(s->pe_begin.data(), s->pe_begin_d.p, (s->p + 1) * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, s->stream));
    // max length
    DevBuf mx;
    mx.reserve(4);
    DSS_HIP(hipMemsetAsync(mx.p, 0, 4, s->stream));
    if (s->n)
        hipLaunchKernelGGL(k_max_len, dim3(std::min<uint64_t>((s->n + 255) / 256, 4096)), dim3(256), 0,
                           s->stream, s->len.as<uint32_t>(), s->n, mx.as<uint32_t>());
    DSS_HIP(hipMemcpyAsync(&s->max_len, mx.p, 4, hipMemcpyDeviceToHost, s->stream));
    DSS_HIP(hipGetLastError());
    DSS_HIP(hipStreamSynchronize(s->stream));
}

// ------------------------------------------------------ D/N generator
// DNRatioGenerator::getRawStringsTimoStyle (random_string_generator.hpp:188-
// 235) with a seeded shuffle: arena position j holds string i = shuffle(j),
// the base-26 digits of i ('A'..'Z', padded with 'A') in the first k chars,
// then slen-k copies of the fill character.
//
// First pass: the shuffled id of every arena position, one lane per string
// (the Feistel walk of one string per lane keeps every lane busy).
__global__ void k_dton_ids(uint32_t* __restrict__ ids, uint64_t n, uint64_t seed) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x)
        ids[j] = (uint32_t)shuffle_index(j, n, seed);
}

// Character at position p of a string with digit word dw (byte e = base-26
// digit e of its id, as a letter); positions before dstart = max(k - 7, 0)
// are 'A' (ids are < 2^32 < 26^7).
__device__ inline uint8_t dton_char(uint32_t p, uint32_t k, uint32_t dstart, uint32_t slen, uint8_t fill,
                                    uint64_t dw) {
    if (p == slen) return 0;
    if (p >= k) return fill;
    if (p < dstart) return 'A';
    return (uint8_t)(dw >> (8 * (k - 1 - p)));
}

// Second pass: a wave builds a 4096-byte chunk in LDS and stores it as whole
// lines.  Every word is first written as 'A's or fill characters by the
// region of its first byte; that is exact except at the digits, the 8 bytes
// after them, a string's first 8 bytes and its NUL, which one lane per string
// then rewrites byte by byte.
__global__ void __launch_bounds__(256)
k_gen_dton(uint8_t* __restrict__ out, uint64_t n, uint32_t k, uint32_t slen, uint8_t fill,
           const uint32_t* __restrict__ ids, uint64_t words) {
    const uint32_t stride = slen + 1;
    const uint64_t total = n * (uint64_t)stride;
    const uint32_t dstart = k > 7 ? k - 7 : 0;
    const uint64_t wordA = 0x4141414141414141ull, wordF = 0x0101010101010101ull * fill;
    const uint32_t inc = 512u % stride;  // position step between a lane's words
    __shared__ uint64_t stage[4][512];
    uint64_t* const sw = stage[(threadIdx.x >> 6) & 3];
    uint8_t* const sb = reinterpret_cast<uint8_t*>(sw);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nchunks = (words + 511) / 512;
    for (uint64_t ch = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; ch < nchunks;
         ch += (uint64_t)gridDim.x * blockDim.x / 64) {
        const uint64_t a0 = ch * 4096;
        const uint64_t j0 = a0 / stride;
        const uint32_t pos0 = (uint32_t)(a0 - j0 * stride);
        // lane l builds words l, l + 64, ... of the chunk
        uint32_t pos = (pos0 + 8 * lane) % stride;
        for (uint32_t q = lane; q < 512; q += 64) {
            const uint64_t ab = a0 + 8ull * q;
            uint64_t v = pos < k ? wordA : wordF;
            if (ab + 8 > total) v = ab >= total ? 0 : v & ((1ull << (8 * (total - ab))) - 1);
            sw[q] = v;
            pos += inc;
            if (pos >= stride) pos -= stride;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        // strings that overlap the chunk: j0 .. j0 + nstr - 1
        const uint64_t aend = min<uint64_t>(a0 + 4096, total);
        const uint64_t nstr = a0 < total ? (aend - 1) / stride - j0 + 1 : 0;
        for (uint64_t t = lane; t < nstr; t += 64) {
            const uint64_t j = j0 + t;
            const uint64_t sj = j * stride;
            // the chunk's part of the string: positions [lo, hi)
            const uint32_t lo = a0 > sj ? (uint32_t)(a0 - sj) : 0u;
            const uint32_t hi = (uint32_t)min<uint64_t>(aend - sj, stride);
            uint64_t dw = 0;
            if (lo < k && hi > dstart) {  // digits inside the chunk
                uint32_t x = ids[j];
                for (uint32_t e = 0; e < 7; ++e) {
                    const uint32_t qd = x / 26u;
                    dw |= (uint64_t)('A' + (x - 26u * qd)) << (8 * e);
                    x = qd;
                }
            }
            const uint32_t r1 = min(8u, stride);
            for (uint32_t p = lo; p < min(hi, r1); ++p)
                sb[sj + p - a0] = dton_char(p, k, dstart, slen, fill, dw);
            const uint32_t r2e = min(k + 8, stride);
            for (uint32_t p = max(lo, max(dstart, r1)); p < min(hi, r2e); ++p)
                sb[sj + p - a0] = dton_char(p, k, dstart, slen, fill, dw);
            if (slen >= lo && slen < hi) sb[sj + slen - a0] = 0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const uint64_t w0 = a0 >> 3;
        for (uint32_t q = lane; q < 512; q += 64)
            if (w0 + q < words) reinterpret_cast<uint64_t*>(out)[w0 + q] = sw[q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();  // read before the next chunk overwrites
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

__global__ void k_fixed_offsets(uint64_t* __restrict__ off, uint32_t* __restrict__ len, uint64_t n,
                                uint32_t slen) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    off[i] = i * (uint64_t)(slen + 1);
    len[i] = slen;
}

// ------------------------------------------------------ skewed generator
struct SkewP {
    uint64_t n;
    uint32_t long_len, short_len, k_long, k_short, long_permille, dup_permille;
};

__device__ inline void skew_id_len(const SkewP& P, uint64_t seed, uint64_t i, uint64_t& id,
                                   uint32_t& len, uint32_t& k) {
    const uint64_t d = rng(seed, 2, i) % 1000;
    const uint64_t x = (d < P.dup_permille) ? rng(seed, 3, i) % P.n : i;
    const bool is_long = (rng(seed, 4, x) % 1000) < P.long_permille;
    uint32_t L = is_long ? P.long_len : P.short_len;
    const uint32_t kk = is_long ? P.k_long : P.k_short;
    if (L < kk) L = kk;
    id = x;
    len = L;
    k = kk;
}

__global__ void k_skew_len(SkewP P, uint64_t seed, uint64_t* __restrict__ size_out,
                           uint32_t* __restrict__ len_out) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.n) return;
    const uint64_t i = shuffle_index(j, P.n, seed);
    uint64_t id;
    uint32_t L, k;
    skew_id_len(P, seed, i, id, L, k);
    size_out[j] = L + 1;
    len_out[j] = L;
}

// One wave per 64 consecutive strings (contiguous in the arena).  Lane t
// derives string t's id, digit count k, length and filler hash once; the
// strings' 8-character pieces (one filler word each) are then dealt to the
// lanes in rounds of 64, so short strings keep the wave busy.  Ids are below
// 2^32 < 255^5, so digits 5 and up are 0 (character 1); the first 8 digit
// characters are kept as one word.
__global__ void __launch_bounds__(256)
k_skew_fill(SkewP P, uint64_t seed, const uint64_t* __restrict__ off, uint8_t* __restrict__ out) {
    __shared__ uint64_t s_off[4][64], s_h[4][64], s_dw[4][64];
    __shared__ uint32_t s_len[4][64], s_k[4][64], s_pre[4][65];
    const uint32_t wv = (threadIdx.x >> 6) & 3, lane = threadIdx.x & 63;
    const uint64_t ngroups = (P.n + 63) / 64;
    for (uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; g < ngroups;
         g += (uint64_t)gridDim.x * blockDim.x / 64) {
        const uint64_t j = g * 64 + lane;
        uint32_t pieces = 0;
        if (j < P.n) {
            const uint64_t i = shuffle_index(j, P.n, seed);
            uint64_t id;
            uint32_t L, k;
            skew_id_len(P, seed, i, id, L, k);
            uint64_t dw = 0, v = id;
            for (uint32_t e = 0; e < 8; ++e) {
                dw |= (uint64_t)(1 + v % 255) << (8 * e);
                v /= 255;
            }
            s_off[wv][lane] = off[j];
            s_h[wv][lane] = rng(seed, 5, id);
            s_dw[wv][lane] = dw;
            s_len[wv][lane] = L;
            s_k[wv][lane] = k;
            pieces = L / 8 + 1;  // positions [0, L] in pieces of 8
        }
        // inclusive scan of the piece counts over the wave (all lanes active)
        uint32_t incl = pieces;
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        s_pre[wv][lane + 1] = incl;
        if (lane == 0) s_pre[wv][0] = 0;
        const uint32_t total = __shfl(incl, 63, 64);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t x = lane; x < total; x += 64) {
            // string t: the last one with s_pre[t] <= x (s_pre is non-decreasing)
            uint32_t t = 0;
            for (uint32_t step = 32; step; step >>= 1)
                if (s_pre[wv][t + step] <= x) t += step;
            const uint32_t L = s_len[wv][t], k = s_k[wv][t];
            const uint32_t t0 = 8 * (x - s_pre[wv][t]);
            const uint64_t fill = mix64(s_h[wv][t] + (t0 >> 3));
            const uint64_t dw = s_dw[wv][t];
            uint8_t* const sp = out + s_off[wv][t];
            for (uint32_t b = 0; b < 8; ++b) {
                const uint32_t tt = t0 + b;
                if (tt > L) break;
                uint8_t c;
                if (tt == L) c = 0;
                else if (tt >= k) c = (uint8_t)(1 + ((fill >> (8 * b)) & 0xFF) % 255);
                else {
                    const uint32_t e = k - 1 - tt;
                    c = e < 8 ? (uint8_t)(dw >> (8 * e)) : (uint8_t)1;
                }
                sp[tt] = c;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();  // LDS read before the next group overwrites
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// ------------------------------------------------------ suffix views
// Capped suffixes of one text without materialising them: string i is the
// view text[i .. i + min(cap, len - i)) (the strings of SuffixGenerator,
// random_string_generator.hpp:89-115, capped as in the suffix workload).
__global__ void k_suffix_views(uint64_t* __restrict__ off, uint32_t* __restrict__ len, uint64_t n, uint32_t cap) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    off[i] = i;
    len[i] = (uint32_t)min((uint64_t)cap, n - i);
}

// Materialised size of every string (length + NUL), for the byte-slice PE
// split of the materialised arena.
__global__ void k_len_plus1(const uint32_t* __restrict__ len, uint64_t n, uint64_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint64_t)len[i] + 1;
}

// Order-3 Markov text over {A,C,G,T}: the chain restarts from a seeded
// context every 4096 characters, so one thread generates one 4096-character
// block (the oracle's generator, same seed streams).
struct MarkovTable {
    uint32_t cum[64][4];
};

__global__ void k_markov_text(uint8_t* __restrict__ text, uint64_t text_len, uint64_t seed, MarkovTable T) {
    const uint64_t blk = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t base = blk * 4096;
    if (base >= text_len) return;
    const uint64_t end = min(text_len, base + 4096);
    uint32_t ctx = (uint32_t)(rng(seed, 8, base) % 64);
    for (uint64_t i = base; i < end; ++i) {
        const uint32_t r = (uint32_t)(rng(seed, 9, i) % T.cum[ctx][3]);
        uint32_t sym = 0;
        while (r >= T.cum[ctx][sym]) ++sym;
        text[i] = "ACGT"[sym];
        ctx = ((ctx << 2) | sym) & 63;
    }
}

// Planted repeats, applied in order (each copy may read earlier copies):
// one block moves each repeat with memmove semantics through LDS chunks,
// front to back when the destination lies before the source, else back to
// front.
__global__ void __launch_bounds__(1024)
k_plant_repeats(uint8_t* __restrict__ text, uint64_t text_len, uint32_t repeats, uint32_t rep_min,
                uint32_t rep_max, uint64_t seed) {
    __shared__ uint8_t buf[8192];
    for (uint32_t r = 0; r < repeats; ++r) {
        const uint64_t span = rep_max > rep_min ? (uint64_t)rep_max - rep_min + 1 : 1;
        const uint64_t len = rep_min + rng(seed, 10, r) % span;
        if (len * 2 >= text_len) continue;
        const uint64_t src = rng(seed, 11, r) % (text_len - len);
        const uint64_t dst = rng(seed, 12, r) % (text_len - len);
        const uint64_t nchunks = (len + sizeof(buf) - 1) / sizeof(buf);
        for (uint64_t q = 0; q < nchunks; ++q) {
            const uint64_t k = dst <= src ? q : nchunks - 1 - q;
            const uint64_t c0 = k * sizeof(buf), c1 = min(len, c0 + sizeof(buf));
            for (uint64_t t = c0 + threadIdx.x; t < c1; t += blockDim.x) buf[t - c0] = text[src + t];
            __syncthreads();
            for (uint64_t t = c0 + threadIdx.x; t < c1; t += blockDim.x) text[dst + t] = buf[t - c0];
            __syncthreads();
        }
    }
}

// Views over the set's arena (already holding the text), PE split and max length.
static void make_suffix_views(dss_set_s* s, uint64_t text_len, uint32_t cap) {
    const uint64_t n = text_len;
    const hipStream_t st = s->stream;
    s->n = n;
    s->nbytes = text_len;
    s->off.reserve(std::max<uint64_t>(n, 1) * 8);
    s->len.reserve(std::max<uint64_t>(n, 1) * 4);
    DevBuf voff;
    voff.reserve(std::max<uint64_t>(n, 1) * 8);
    uint64_t total = 0;
    if (n) {
        hipLaunchKernelGGL(k_suffix_views, dim3((n + 255) / 256), dim3(256), 0, st, s->off.as<uint64_t>(),
                           s->len.as<uint32_t>(), n, cap);
        hipLaunchKernelGGL(k_len_plus1, dim3((n + 255) / 256), dim3(256), 0, st, s->len.as<uint32_t>(), n,
                           voff.as<uint64_t>());
        exclusive_scan_u64(voff.as<uint64_t>(), voff.as<uint64_t>(), n, st, s->pp, &total);
    }
    s->pe_begin_d.reserve((s->p + 1) * 8);
    hipLaunchKernelGGL(k_pe_split, dim3((s->p + 1 + 255) / 256), dim3(256), 0, st, voff.as<uint64_t>(), n, total,
                       s->p, s->pe_begin_d.as<uint64_t>());
    s->pe_begin.assign(s->p + 1, 0);
    DSS_HIP(hipMemcpyAsync(s->pe_begin.data(), s->pe_begin_d.p, (s->p + 1) * 8, hipMemcpyDeviceToHost, st));
    DSS_HIP(hipGetLastError());
    DSS_HIP(hipStreamSynchronize(st));
    s->max_len = (uint32_t)std::min<uint64_t>(cap, n);
}

}  // namespace dss

using namespace dss;

extern "C" {

int dss_set_create_suffixes(const uint8_t* text, uint64_t text_len, uint32_t cap, uint32_t num_pes,
                            dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(out && (text_len == 0 || text), DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(cap >= 1, DSS_EINVAL, "cap must be >= 1");
        DSS_REQUIRE(text_len < 0xFFFFFFFFull, DSS_ERANGE, "more than 2^32-1 suffixes");
        for (uint64_t i = 0; i < text_len; ++i)
            DSS_REQUIRE(text[i] != 0, DSS_EINVAL, "the text must not contain NUL bytes");
        dss_set_s* s = new_set(num_pes);
        try {
            s->chars.reserve(text_len + kArenaPad);
            DSS_HIP(hipMemsetAsync(s->chars.p, 0, text_len + kArenaPad, s->stream));
            if (text_len) DSS_HIP(hipMemcpyAsync(s->chars.p, text, text_len, hipMemcpyHostToDevice, s->stream));
            make_suffix_views(s, text_len, cap);
        } catch (...) {
            free_set(s);
            throw;
        }
        *out = s;
    });
}

int dss_set_create_markov_suffixes(uint64_t text_len, uint32_t repeats, uint32_t rep_min, uint32_t rep_max,
                                   uint32_t cap, uint64_t seed, uint32_t num_pes, dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(out, DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(cap >= 1, DSS_EINVAL, "cap must be >= 1");
        DSS_REQUIRE(text_len < 0xFFFFFFFFull, DSS_ERANGE, "more than 2^32-1 suffixes");
        MarkovTable T;
        for (int c = 0; c < 64; ++c) {
            uint32_t acc = 0;
            for (int k = 0; k < 4; ++k) {
                acc += 1 + (uint32_t)(rng(seed, 7, (uint64_t)c * 4 + k) % 16);
                T.cum[c][k] = acc;
            }
        }
        dss_set_s* s = new_set(num_pes);
        try {
            s->chars.reserve(text_len + kArenaPad);
            DSS_HIP(hipMemsetAsync(s->chars.p, 0, text_len + kArenaPad, s->stream));
            if (text_len) {
                const uint64_t blocks = (text_len + 4095) / 4096;
                hipLaunchKernelGGL(k_markov_text, dim3((blocks + 255) / 256), dim3(256), 0, s->stream,
                                   s->chars.as<uint8_t>(), text_len, seed, T);
                hipLaunchKernelGGL(k_plant_repeats, dim3(1), dim3(1024), 0, s->stream, s->chars.as<uint8_t>(),
                                   text_len, repeats, rep_min, rep_max, seed);
            }
            make_suffix_views(s, text_len, cap);
        } catch (...) {
            free_set(s);
            throw;
        }
        *out = s;
    });
}

const char* dss_last_error(void) { return g_err.c_str(); }

int dss_device_info(char* buf, size_t buflen) {
    return guarded([&] {
        int dev = 0;
        DSS_HIP(hipGetDevice(&dev));
        hipDeviceProp_t pr;
        DSS_HIP(hipGetDeviceProperties(&pr, dev));
        snprintf(buf, buflen, "%s %s CUs=%d HBM=%.1fGiB", pr.name, pr.gcnArchName,
                 pr.multiProcessorCount, pr.totalGlobalMem / 1073741824.0);
    });
}

int dss_set_create(const uint8_t* chars, uint64_t nbytes, uint32_t num_pes, dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(out != nullptr, DSS_EINVAL, "out is NULL");
        DSS_REQUIRE(nbytes == 0 || chars != nullptr, DSS_EINVAL, "chars is NULL");
        DSS_REQUIRE(nbytes == 0 || chars[nbytes - 1] == 0, DSS_EINVAL,
                    "arena must end with a NUL terminator");
        dss_set_s* s = new_set(num_pes);
        try {
            s->nbytes = nbytes;
            s->chars.reserve(nbytes + kArenaPad);
            DSS_HIP(hipMemsetAsync(s->chars.p, 0, nbytes + kArenaPad, s->stream));
            if (nbytes)
                DSS_HIP(hipMemcpyAsync(s->chars.p, chars, nbytes, hipMemcpyHostToDevice, s->stream));
            index_arena(s);
            compute_pe_split(s);
        } catch (...) {
            free_set(s);
            throw;
        }
        *out = s;
    });
}

// ------------------------------------------------------ text input
// distribute_strings (src/mpi/readInput.hpp:40-104) on a '\n'-separated
// text: every '\n' becomes a NUL (8 bytes per lane: the bytes equal to '\n'
// are found with an exact zero-byte test on x ^ 0x0A..0A), a NUL is appended
// when the text does not end with one, and the PE split slices the text's
// bytes (not the appended NUL), as the MPI ranks slice the file.
__global__ void k_newline_to_nul(uint8_t* __restrict__ a, uint64_t nbytes) {
    const uint64_t words = nbytes / 8;
    uint64_t* w = reinterpret_cast<uint64_t*>(a);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr uint64_t lo7 = 0x7F7F7F7F7F7F7F7Full;
    for (uint64_t i = gid; i < words; i += stride) {
        const uint64_t x = w[i];
        const uint64_t t = x ^ 0x0A0A0A0A0A0A0A0Aull;
        const uint64_t z = ~(((t & lo7) + lo7) | t | lo7);  // 0x80 in every byte equal to '\n'
        if (z) w[i] = x & ~((z >> 7) * 0xFF);
    }
    for (uint64_t j = words * 8 + gid; j < nbytes; j += stride)
        if (a[j] == '\n') a[j] = 0;
}

// The first L text bytes are in the arena, which is zeroed from L on.
static void finish_text(dss_set_s* s, uint64_t L, uint8_t last) {
    if (L)
        hipLaunchKernelGGL(k_newline_to_nul, dim3(stride_grid((L + 7) / 8, 256)), dim3(256), 0, s->stream,
                           s->chars.as<uint8_t>(), L);
    DSS_HIP(hipGetLastError());
    s->nbytes = L + (L && last != '\n' && last != 0 ? 1 : 0);
    s->slice_bytes = L;
    index_arena(s);
    compute_pe_split(s);
}

int dss_set_create_text(const uint8_t* text, uint64_t text_len, uint64_t max_size, uint32_t num_pes,
                        dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(out != nullptr, DSS_EINVAL, "out is NULL");
        DSS_REQUIRE(text_len == 0 || text != nullptr, DSS_EINVAL, "text is NULL");
        const uint64_t L = max_size ? std::min(max_size, text_len) : text_len;
        dss_set_s* s = new_set(num_pes);
        try {
            s->chars.reserve(L + 1 + kArenaPad);
            DSS_HIP(hipMemsetAsync(s->chars.as<uint8_t>() + L, 0, 1 + kArenaPad, s->stream));
            if (L) DSS_HIP(hipMemcpyAsync(s->chars.p, text, L, hipMemcpyHostToDevice, s->stream));
            finish_text(s, L, L ? text[L - 1] : 0);
        } catch (...) {
            free_set(s);
            throw;
        }
        *out = s;
    });
}

int dss_set_load_file(const char* path, uint64_t max_size, uint32_t num_pes, dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(out != nullptr && path != nullptr, DSS_EINVAL, "NULL argument");
        FILE* f = fopen(path, "rb");
        DSS_REQUIRE(f != nullptr, DSS_EINVAL, std::string("cannot open ") + path);
        dss_set_s* s = nullptr;
        uint8_t* pin[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        auto cleanup = [&] {
            for (int k = 0; k < 2; ++k) {
                if (ev[k]) (void)hipEventDestroy(ev[k]);
                if (pin[k]) (void)hipHostFree(pin[k]);
            }
            fclose(f);
        };
        try {
            DSS_REQUIRE(fseeko(f, 0, SEEK_END) == 0, DSS_EINVAL, "cannot seek the input file");
            const off_t size = ftello(f);
            DSS_REQUIRE(size >= 0 && fseeko(f, 0, SEEK_SET) == 0, DSS_EINVAL, "cannot size the input file");
            const uint64_t L = max_size ? std::min<uint64_t>(max_size, (uint64_t)size) : (uint64_t)size;
            s = new_set(num_pes);
            s->chars.reserve(L + 1 + kArenaPad);
            DSS_HIP(hipMemsetAsync(s->chars.as<uint8_t>() + L, 0, 1 + kArenaPad, s->stream));
            // file -> pinned staging (two 64 MiB buffers) -> HBM, reads overlapping copies
            constexpr uint64_t kStage = 64ull << 20;
            const uint64_t stage = std::min<uint64_t>(kStage, std::max<uint64_t>(L, 1));
            for (int k = 0; k < 2; ++k) {
                DSS_HIP(hipHostMalloc((void**)&pin[k], stage, hipHostMallocDefault));
                DSS_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
            }
            uint8_t last = 0;
            uint64_t i = 0;
            for (uint64_t o = 0; o < L; o += stage, ++i) {
                const uint64_t k = std::min(stage, L - o);
                uint8_t* b = pin[i & 1];
                if (i >= 2) DSS_HIP(hipEventSynchronize(ev[i & 1]));
                DSS_REQUIRE(fread(b, 1, k, f) == k, DSS_EINVAL, "short read of the input file");
                last = b[k - 1];
                DSS_HIP(hipMemcpyAsync(s->chars.as<uint8_t>() + o, b, k, hipMemcpyHostToDevice, s->stream));
                DSS_HIP(hipEventRecord(ev[i & 1], s->stream));
            }
            DSS_HIP(hipStreamSynchronize(s->stream));
            finish_text(s, L, last);
        } catch (...) {
            cleanup();
            free_set(s);
            throw;
        }
        cleanup();
        *out = s;
    });
}

int dss_set_create_dton(uint64_t n, uint32_t len, double dton, uint64_t seed, uint32_t num_pes,
                        dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(out != nullptr, DSS_EINVAL, "out is NULL");
        DSS_REQUIRE(n >= 1 && n < 0xFFFFFFFFull, DSS_ERANGE, "n must be in [1, 2^32-1)");
        // k and stringLength exactly as random_string_generator.hpp:196-199
        const double a = (double)len * dton;
        const double b = std::ceil(std::log((double)n) / std::log(26.0));
        const uint32_t k = (uint32_t)(a > b ? a : b);
        const uint32_t slen = len > k ? len : k;
        const uint8_t fill = (uint8_t)('A' + rng(seed, 1, 0) % 26);
        dss_set_s* s = new_set(num_pes);
        try {
            s->n = n;
            s->nbytes = n * (uint64_t)(slen + 1);
            s->chars.reserve(s->nbytes + kArenaPad);
            DSS_HIP(hipMemsetAsync(s->chars.as<uint8_t>() + s->nbytes, 0, kArenaPad, s->stream));
            const uint64_t words = (s->nbytes + 7) / 8;
            s->off.reserve(n * sizeof(uint64_t));
            s->len.reserve(n * sizeof(uint32_t));
            // the lengths array holds the shuffled ids until it is filled
            hipLaunchKernelGGL(k_dton_ids, dim3(stride_grid(n, 256)), dim3(256), 0, s->stream,
                               s->len.as<uint32_t>(), n, seed);
            hipLaunchKernelGGL(k_gen_dton, dim3(stride_grid((words + 511) / 512, 4)), dim3(256), 0, s->stream,
                               s->chars.as<uint8_t>(), n, k, slen, fill, s->len.as<const uint32_t>(), words);
            hipLaunchKernelGGL(k_fixed_offsets, dim3((n + 255) / 256), dim3(256), 0, s->stream,
                               s->off.as<uint64_t>(), s->len.as<uint32_t>(), n, slen);
            DSS_HIP(hipGetLastError());
            compute_pe_split(s);
        } catch (...) {
            free_set(s);
            throw;
        }
        *out = s;
    });
}

int dss_set_create_skewed(uint64_t n, uint32_t long_len, uint32_t short_len, double long_dn,
                          double short_dn, uint32_t long_permille, uint32_t dup_permille,
                          uint64_t seed, uint32_t num_pes, dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(out != nullptr, DSS_EINVAL, "out is NULL");
        DSS_REQUIRE(n >= 1 && n < 0xFFFFFFFFull, DSS_ERANGE, "n must be in [1, 2^32-1)");
        auto kfun = [&](uint32_t L, double dn) {
            double a = (double)L * dn;
            double b = std::ceil(std::log((double)(n > 1 ? n : 2)) / std::log(255.0));
            return (uint32_t)(a > b ? a : b);
        };
        SkewP P{n, long_len, short_len, kfun(long_len, long_dn), kfun(short_len, short_dn),
                long_permille, dup_permille};
        dss_set_s* s = new_set(num_pes);
        try {
            s->n = n;
            s->off.reserve(n * sizeof(uint64_t));
            s->len.reserve(n * sizeof(uint32_t));
            hipLaunchKernelGGL(k_skew_len, dim3((n + 255) / 256), dim3(256), 0, s->stream, P, seed,
                               s->off.as<uint64_t>(), s->len.as<uint32_t>());
            exclusive_scan_u64(s->off.as<uint64_t>(), s->off.as<uint64_t>(), n, s->stream, s->pp, &s->nbytes);
            s->chars.reserve(s->nbytes + kArenaPad);
            DSS_HIP(hipMemsetAsync(s->chars.as<uint8_t>() + s->nbytes, 0, kArenaPad, s->stream));
            hipLaunchKernelGGL(k_skew_fill, dim3(stride_grid((n + 63) / 64, 4)), dim3(256), 0, s->stream, P, seed,
                               s->off.as<uint64_t>(), s->chars.as<uint8_t>());
            DSS_HIP(hipGetLastError());
            compute_pe_split(s);
        } catch (...) {
            free_set(s);
            throw;
        }
        *out = s;
    });
}

int dss_set_destroy(dss_set s) {
    return guarded([&] {
        if (!s) return;
        (void)hipStreamSynchronize(s->stream);
        hipStream_t st = s->own_stream;  // a caller's stream stays the caller's
        delete s;
        (void)hipStreamDestroy(st);
    });
}

int dss_set_stream(dss_set s, void* stream) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        DSS_HIP(hipStreamSynchronize(s->stream));  // work queued so far completes on its stream
        s->stream = stream ? static_cast<hipStream_t>(stream) : s->own_stream;
    });
}

int dss_set_info(dss_set s, dss_set_info_t* info) {
    return guarded([&] {
        DSS_REQUIRE(s && info, DSS_EINVAL, "NULL argument");
        info->num_strings = s->n;
        info->num_bytes = s->nbytes;
        info->num_pes = s->p;
        info->max_len = s->max_len;
    });
}

int dss_set_pe_begin(dss_set s, uint64_t* pe_begin) {
    return guarded([&] {
        DSS_REQUIRE(s && pe_begin, DSS_EINVAL, "NULL argument");
        memcpy(pe_begin, s->pe_begin.data(), (s->p + 1) * sizeof(uint64_t));
    });
}

int dss_set_download(dss_set s, uint8_t* chars, uint64_t* offsets, uint32_t* lengths) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        DSS_REQUIRE(!(chars && s->borrowed), DSS_EINVAL, "a batch set has no arena of its own to download");
        if (chars && s->nbytes)
            DSS_HIP(hipMemcpyAsync(chars, s->arena(), s->nbytes, hipMemcpyDeviceToHost, s->stream));
        if (offsets && s->n)
            DSS_HIP(hipMemcpyAsync(offsets, s->off.p, s->n * 8, hipMemcpyDeviceToHost, s->stream));
        if (lengths && s->n)
            DSS_HIP(hipMemcpyAsync(lengths, s->len.p, s->n * 4, hipMemcpyDeviceToHost, s->stream));
        DSS_HIP(hipStreamSynchronize(s->stream));
    });
}

int dss_sync(dss_set s) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        DSS_HIP(hipStreamSynchronize(s->stream));
        s->timer.collect();
    });
}

int dss_set_timing(dss_set s, int enable) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        s->timer.enabled = enable != 0;
    });
}

int dss_kernel_stats(dss_set s, dss_kernel_stat_t* stats, uint32_t cap, uint32_t* count, int reset) {
    return guarded([&] {
        DSS_REQUIRE(s && count, DSS_EINVAL, "NULL argument");
        DSS_HIP(hipStreamSynchronize(s->stream));
        s->timer.collect();
        uint32_t k = 0;
        for (auto& st : s->timer.stats) {
            if (k < cap && stats) {
                memset(&stats[k], 0, sizeof(dss_kernel_stat_t));
                strncpy(stats[k].name, st.name.c_str(), sizeof(stats[k].name) - 1);
                stats[k].launches = st.launches;
                stats[k].total_ms = st.ms;
                stats[k].alg_bytes = st.bytes;
            }
            ++k;
        }
        *count = k;
        if (reset) s->timer.stats.clear();
    });
}

}  // extern "C"

// ------------------------------------------------------------ explicit PEs
namespace dss {
__global__ void k_pe_from_bytes(const uint64_t* __restrict__ off, uint64_t n, const uint64_t* __restrict__ byte_begin,
                                uint32_t p, uint64_t* __restrict__ pe_begin) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > p) return;
    if (r == p) { pe_begin[p] = n; return; }
    const uint64_t b = byte_begin[r];
    uint64_t lo = 0, hi = n;  // first string starting at or after b
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (off[mid] < b) lo = mid + 1; else hi = mid;
    }
    pe_begin[r] = lo;
}
}  // namespace dss

extern "C" int dss_set_create_pes(const uint8_t* chars, const uint64_t* pe_byte_begin, uint32_t num_pes,
                                  dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(out && pe_byte_begin, DSS_EINVAL, "NULL argument");
        const uint64_t nbytes = pe_byte_begin[num_pes];
        for (uint32_t r = 0; r < num_pes; ++r)
            DSS_REQUIRE(pe_byte_begin[r] <= pe_byte_begin[r + 1], DSS_EINVAL, "PE byte ranges must be ordered");
        DSS_REQUIRE(pe_byte_begin[0] == 0, DSS_EINVAL, "PE byte ranges must start at 0");
        DSS_REQUIRE(nbytes == 0 || chars != nullptr, DSS_EINVAL, "chars is NULL");
        for (uint32_t r = 1; r <= num_pes; ++r)
            DSS_REQUIRE(pe_byte_begin[r] == pe_byte_begin[r - 1] || chars[pe_byte_begin[r] - 1] == 0, DSS_EINVAL,
                        "every PE's bytes must end with a NUL terminator");
        dss_set_s* s = new_set(num_pes);
        try {
            s->nbytes = nbytes;
            s->chars.reserve(nbytes + kArenaPad);
            DSS_HIP(hipMemsetAsync(s->chars.p, 0, nbytes + kArenaPad, s->stream));
            if (nbytes) DSS_HIP(hipMemcpyAsync(s->chars.p, chars, nbytes, hipMemcpyHostToDevice, s->stream));
            index_arena(s);
            DevBuf bb;
            bb.reserve((num_pes + 1) * 8);
            DSS_HIP(hipMemcpyAsync(bb.p, pe_byte_begin, (num_pes + 1) * 8, hipMemcpyHostToDevice, s->stream));
            s->pe_begin_d.reserve((num_pes + 1) * 8);
            DevBuf dummy;
            const uint64_t* offp = s->off.as<uint64_t>();
            if (!offp) {
                dummy.reserve(8);
                offp = dummy.as<uint64_t>();
            }
            hipLaunchKernelGGL(k_pe_from_bytes, dim3((num_pes + 1 + 255) / 256), dim3(256), 0, s->stream, offp, s->n,
                               bb.as<uint64_t>(), num_pes, s->pe_begin_d.as<uint64_t>());
            s->pe_begin.assign(num_pes + 1, 0);
            DSS_HIP(hipMemcpyAsync(s->pe_begin.data(), s->pe_begin_d.p, (num_pes + 1) * 8, hipMemcpyDeviceToHost,
                                   s->stream));
            DSS_HIP(hipStreamSynchronize(s->stream));
            uint32_t mx = 0;
            std::vector<uint32_t> lens(s->n);
            if (s->n) DSS_HIP(hipMemcpy(lens.data(), s->len.p, s->n * 4, hipMemcpyDeviceToHost));
            for (auto v : lens) mx = std::max(mx, v);
            s->max_len = mx;
        } catch (...) {
            free_set(s);
            throw;
        }
        *out = s;
    });
}

// ================================================================== job pool
// Device-resident arenas of many sort jobs for batched and served execution:
// one HBM region of characters and one of per-string offsets (absolute in the
// pool arena) and lengths, with first-fit allocation.  A batch is a string
// set over the pool whose PEs are the batch's jobs; the set borrows the pool
// arena and owns copies of the jobs' offset/length arrays.
namespace dss {

struct RangeAlloc {
    std::map<uint64_t, uint64_t> free_;  // begin -> size
    uint64_t cap = 0;
    void init(uint64_t c) {
        cap = c;
        free_.clear();
        if (c) free_[0] = c;
    }
    bool alloc(uint64_t size, uint64_t& at) {
        if (size == 0) {
            at = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= size) {
                at = it->first;
                const uint64_t rest = it->second - size;
                free_.erase(it);
                if (rest) free_[at + size] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t at, uint64_t size) {
        if (!size) return;
        auto it = free_.emplace(at, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used() const {
        uint64_t f = 0;
        for (auto& kv : free_) f += kv.second;
        return cap - f;
    }
};

__global__ void k_offsets_base(uint64_t* __restrict__ off, uint32_t* __restrict__ len, uint64_t n, uint32_t slen,
                               uint64_t base) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    off[i] = base + i * (uint64_t)(slen + 1);
    len[i] = slen;
}

__global__ void k_add_base(uint64_t* __restrict__ off, uint64_t n, uint64_t base) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) off[i] += base;
}

}  // namespace dss

struct dss_pool_s {
    int device = 0;
    hipStream_t stream = nullptr;
    dss::DevBuf chars, off, len, tot;
    dss::RangeAlloc bytes, strings;
    dss::PipeState scan;  // scratch of the device-wide scan
    struct Job {
        uint64_t byte0 = 0, nbytes = 0, alloc = 0, str0 = 0, n = 0;
        uint32_t max_len = 0;
        bool live = false;
    };
    std::vector<Job> jobs;
    std::mutex mu;
};

using namespace dss;

// Character ranges are allocated in whole 8-byte words so that every job
// starts 8-byte aligned (the generators write whole words).
static uint32_t pool_register(dss_pool_s* P, uint64_t byte0, uint64_t nbytes, uint64_t asz, uint64_t str0,
                              uint64_t n) {
    dss_pool_s::Job J;
    J.byte0 = byte0;
    J.nbytes = nbytes;
    J.alloc = asz;
    J.str0 = str0;
    J.n = n;
    J.live = true;
    for (uint32_t k = 0; k < P->jobs.size(); ++k)
        if (!P->jobs[k].live) {
            P->jobs[k] = J;
            return k;
        }
    P->jobs.push_back(J);
    return (uint32_t)P->jobs.size() - 1;
}

static uint32_t pool_new_job(dss_pool_s* P, uint64_t nbytes, uint64_t n, uint64_t& byte0, uint64_t& str0) {
    const uint64_t asz = (nbytes + 7) / 8 * 8;
    if (!P->bytes.alloc(asz, byte0)) throw Error{DSS_ENOMEM, "job pool: no room for the job's characters"};
    if (!P->strings.alloc(n, str0)) {
        P->bytes.release(byte0, asz);
        throw Error{DSS_ENOMEM, "job pool: no room for the job's strings"};
    }
    return pool_register(P, byte0, nbytes, asz, str0, n);
}

// Points s at the given pool jobs (PE k = jobs[k]); keeps s's device buffers.
static void pool_fill_batch(dss_pool_s* P, dss_set_s* s, const uint32_t* jobs, uint32_t count) {
    std::lock_guard<std::mutex> g(P->mu);
    std::vector<uint64_t> pb(count + 1, 0);
    uint64_t n = 0;
    uint32_t mx = 0;
    for (uint32_t k = 0; k < count; ++k) {
        DSS_REQUIRE(jobs[k] < P->jobs.size() && P->jobs[jobs[k]].live, DSS_EINVAL, "no such job");
        for (uint32_t q = 0; q < k; ++q) DSS_REQUIRE(jobs[q] != jobs[k], DSS_EINVAL, "job listed twice");
        n += P->jobs[jobs[k]].n;
        pb[k + 1] = n;
        mx = std::max(mx, P->jobs[jobs[k]].max_len);
    }
    DSS_REQUIRE(n < 0xFFFFFFFFull, DSS_ERANGE, "batch exceeds 2^32-1 strings");
    s->borrowed = P->chars.as<uint8_t>();
    s->nbytes = 0;
    s->p = count;
    s->pe_begin = pb;
    s->n = n;
    s->max_len = mx;
    s->sorted = false;
    s->pp.has_dist = s->pp.partitioned = s->pp.exchanged = s->pp.merged = s->pp.extended = false;
    s->off.reserve(std::max<uint64_t>(n, 1) * 8);
    s->len.reserve(std::max<uint64_t>(n, 1) * 4);
    for (uint32_t k = 0; k < count; ++k) {
        const auto& J = P->jobs[jobs[k]];
        if (!J.n) continue;
        DSS_HIP(hipMemcpyAsync(s->off.as<uint64_t>() + pb[k], P->off.as<uint64_t>() + J.str0, J.n * 8,
                               hipMemcpyDeviceToDevice, s->stream));
        DSS_HIP(hipMemcpyAsync(s->len.as<uint32_t>() + pb[k], P->len.as<uint32_t>() + J.str0, J.n * 4,
                               hipMemcpyDeviceToDevice, s->stream));
    }
    s->pe_begin_d.reserve((count + 1) * 8);
    DSS_HIP(hipMemcpyAsync(s->pe_begin_d.p, pb.data(), (count + 1) * 8, hipMemcpyHostToDevice, s->stream));
    DSS_HIP(hipStreamSynchronize(s->stream));
}

extern "C" {

int dss_pool_create(uint64_t char_capacity, uint64_t string_capacity, dss_pool* out) {
    return guarded([&] {
        DSS_REQUIRE(out, DSS_EINVAL, "out is NULL");
        dss_pool_s* P = new dss_pool_s();
        try {
            DSS_HIP(hipGetDevice(&P->device));
            DSS_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
            P->chars.reserve(char_capacity + kArenaPad);
            DSS_HIP(hipMemsetAsync(P->chars.p, 0, char_capacity + kArenaPad, P->stream));
            P->off.reserve(std::max<uint64_t>(string_capacity, 1) * 8);
            P->len.reserve(std::max<uint64_t>(string_capacity, 1) * 4);
            P->tot.reserve(8);
            // the generators' scan scratch and read-back mailbox, so that adding
            // a job allocates nothing
            P->scan.scan_tmp.reserve(scan_scratch_words(std::max<uint64_t>(string_capacity, 1)) * 8);
            P->scan.scan_tot.reserve(8);
            P->bytes.init(char_capacity);
            P->strings.init(string_capacity);
            P->scan.mbox.wait(P->stream);
        } catch (...) {
            if (P->stream) (void)hipStreamDestroy(P->stream);
            delete P;
            throw;
        }
        *out = P;
    });
}

int dss_pool_destroy(dss_pool P) {
    return guarded([&] {
        if (!P) return;
        (void)hipStreamSynchronize(P->stream);
        hipStream_t st = P->stream;
        delete P;
        (void)hipStreamDestroy(st);
    });
}

int dss_pool_add_job(dss_pool P, const uint8_t* chars, uint64_t nbytes, uint32_t* job) {
    return guarded([&] {
        DSS_REQUIRE(P && job && (nbytes == 0 || chars), DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(nbytes == 0 || chars[nbytes - 1] == 0, DSS_EINVAL, "arena must end with a NUL terminator");
        std::vector<uint64_t> off;
        std::vector<uint32_t> len;
        uint64_t s = 0;
        uint32_t mx = 0;
        for (uint64_t i = 0; i < nbytes; ++i)
            if (chars[i] == 0) {
                off.push_back(s);
                len.push_back((uint32_t)(i - s));
                mx = std::max(mx, (uint32_t)(i - s));
                s = i + 1;
            }
        std::lock_guard<std::mutex> g(P->mu);
        uint64_t b0, s0;
        const uint32_t id = pool_new_job(P, nbytes, off.size(), b0, s0);
        P->jobs[id].max_len = mx;
        if (nbytes) DSS_HIP(hipMemcpyAsync(P->chars.as<uint8_t>() + b0, chars, nbytes, hipMemcpyHostToDevice, P->stream));
        if (!off.empty()) {
            for (auto& o : off) o += b0;
            DSS_HIP(hipMemcpyAsync(P->off.as<uint64_t>() + s0, off.data(), off.size() * 8, hipMemcpyHostToDevice, P->stream));
            DSS_HIP(hipMemcpyAsync(P->len.as<uint32_t>() + s0, len.data(), len.size() * 4, hipMemcpyHostToDevice, P->stream));
        }
        DSS_HIP(hipStreamSynchronize(P->stream));
        *job = id;
    });
}

int dss_pool_add_job_dton(dss_pool P, uint64_t n, uint32_t length, double dton, uint64_t seed, uint32_t* job) {
    return guarded([&] {
        DSS_REQUIRE(P && job, DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(n >= 1 && n < 0xFFFFFFFFull, DSS_ERANGE, "n must be in [1, 2^32-1)");
        const double a = (double)length * dton;
        const double b = std::ceil(std::log((double)n) / std::log(26.0));
        const uint32_t k = (uint32_t)(a > b ? a : b);
        const uint32_t slen = length > k ? length : k;
        const uint8_t fill = (uint8_t)('A' + rng(seed, 1, 0) % 26);
        const uint64_t nbytes = n * (uint64_t)(slen + 1);
        std::lock_guard<std::mutex> g(P->mu);
        uint64_t b0, s0;
        const uint32_t id = pool_new_job(P, nbytes, n, b0, s0);
        P->jobs[id].max_len = slen;
        const uint64_t words = (nbytes + 7) / 8;
        // the job's lengths hold the shuffled ids until they are filled
        hipLaunchKernelGGL(k_dton_ids, dim3(stride_grid(n, 256)), dim3(256), 0, P->stream,
                           P->len.as<uint32_t>() + s0, n, seed);
        hipLaunchKernelGGL(k_gen_dton, dim3(stride_grid((words + 511) / 512, 4)), dim3(256), 0, P->stream,
                           P->chars.as<uint8_t>() + b0, n, k, slen, fill, P->len.as<const uint32_t>() + s0, words);
        hipLaunchKernelGGL(k_offsets_base, dim3((n + 255) / 256), dim3(256), 0, P->stream, P->off.as<uint64_t>() + s0,
                           P->len.as<uint32_t>() + s0, n, slen, b0);
        DSS_HIP(hipGetLastError());
        DSS_HIP(hipStreamSynchronize(P->stream));
        *job = id;
    });
}

int dss_pool_add_job_skewed(dss_pool P, uint64_t n, uint32_t long_len, uint32_t short_len, double long_dn,
                            double short_dn, uint32_t long_permille, uint32_t dup_permille, uint64_t seed,
                            uint32_t* job) {
    return guarded([&] {
        DSS_REQUIRE(P && job, DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(n >= 1 && n < 0xFFFFFFFFull, DSS_ERANGE, "n must be in [1, 2^32-1)");
        auto kfun = [&](uint32_t L, double dn) {
            double a = (double)L * dn;
            double b = std::ceil(std::log((double)(n > 1 ? n : 2)) / std::log(255.0));
            return (uint32_t)(a > b ? a : b);
        };
        SkewP S{n, long_len, short_len, kfun(long_len, long_dn), kfun(short_len, short_dn), long_permille,
                dup_permille};
        std::lock_guard<std::mutex> g(P->mu);
        // string slots first: sizes are generated into the job's offset array,
        // scanned in place, and the character range is allocated afterwards
        uint64_t s0;
        if (!P->strings.alloc(n, s0)) throw Error{DSS_ENOMEM, "job pool: no room for the job's strings"};
        uint64_t* off = P->off.as<uint64_t>() + s0;
        hipLaunchKernelGGL(k_skew_len, dim3((n + 255) / 256), dim3(256), 0, P->stream, S, seed, off,
                           P->len.as<uint32_t>() + s0);
        uint64_t nbytes = 0;
        exclusive_scan_u64(off, off, n, P->stream, P->scan, &nbytes);
        const uint64_t asz = (nbytes + 7) / 8 * 8;
        uint64_t b0;
        if (!P->bytes.alloc(asz, b0)) {
            P->strings.release(s0, n);
            throw Error{DSS_ENOMEM, "job pool: no room for the job's characters"};
        }
        const uint32_t id = pool_register(P, b0, nbytes, asz, s0, n);
        P->jobs[id].max_len = std::max(std::max(long_len, S.k_long), std::max(short_len, S.k_short));
        hipLaunchKernelGGL(k_add_base, dim3((n + 255) / 256), dim3(256), 0, P->stream, off, n, b0);
        hipLaunchKernelGGL(k_skew_fill, dim3(stride_grid((n + 63) / 64, 4)), dim3(256), 0, P->stream, S, seed, off,
                           P->chars.as<uint8_t>());
        DSS_HIP(hipGetLastError());
        DSS_HIP(hipStreamSynchronize(P->stream));
        *job = id;
    });
}

int dss_pool_free_job(dss_pool P, uint32_t job) {
    return guarded([&] {
        DSS_REQUIRE(P, DSS_EINVAL, "NULL pool");
        std::lock_guard<std::mutex> g(P->mu);
        DSS_REQUIRE(job < P->jobs.size() && P->jobs[job].live, DSS_EINVAL, "no such job");
        auto& J = P->jobs[job];
        P->bytes.release(J.byte0, J.alloc);
        P->strings.release(J.str0, J.n);
        J.live = false;
    });
}

int dss_pool_usage(dss_pool P, uint64_t* bytes_used, uint64_t* strings_used, uint64_t* bytes_cap,
                   uint64_t* strings_cap) {
    return guarded([&] {
        DSS_REQUIRE(P, DSS_EINVAL, "NULL pool");
        std::lock_guard<std::mutex> g(P->mu);
        if (bytes_used) *bytes_used = P->bytes.used();
        if (strings_used) *strings_used = P->strings.used();
        if (bytes_cap) *bytes_cap = P->bytes.cap;
        if (strings_cap) *strings_cap = P->strings.cap;
    });
}

int dss_pool_job_info(dss_pool P, uint32_t job, uint64_t* num_strings, uint64_t* num_bytes) {
    return guarded([&] {
        DSS_REQUIRE(P, DSS_EINVAL, "NULL pool");
        std::lock_guard<std::mutex> g(P->mu);
        DSS_REQUIRE(job < P->jobs.size() && P->jobs[job].live, DSS_EINVAL, "no such job");
        if (num_strings) *num_strings = P->jobs[job].n;
        if (num_bytes) *num_bytes = P->jobs[job].nbytes;
    });
}

int dss_pool_batch(dss_pool P, const uint32_t* jobs, uint32_t count, dss_set* out) {
    return guarded([&] {
        DSS_REQUIRE(P && out && jobs && count >= 1, DSS_EINVAL, "NULL argument or empty batch");
        dss_set_s* s = new_set(count);
        try {
            pool_fill_batch(P, s, jobs, count);
        } catch (...) {
            free_set(s);
            throw;
        }
        *out = s;
    });
}

int dss_pool_rebatch(dss_pool P, dss_set s, const uint32_t* jobs, uint32_t count) {
    return guarded([&] {
        DSS_REQUIRE(P && s && jobs && count >= 1, DSS_EINVAL, "NULL argument or empty batch");
        DSS_REQUIRE(s->borrowed == P->chars.as<uint8_t>(), DSS_EINVAL, "set is not a batch of this pool");
        DSS_REQUIRE(count <= (1u << 16), DSS_EINVAL, "at most 65536 jobs per batch");
        DSS_HIP(hipStreamSynchronize(s->stream));
        pool_fill_batch(P, s, jobs, count);
    });
}

}  // extern "C"
