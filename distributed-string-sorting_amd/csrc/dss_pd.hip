// (a10-a12) Prefix doubling: distinguishing-prefix estimates by rounds of
// prefix hashing and duplicate detection, and the two composite drivers.
//
// computeDistinguishingPrefixes (src/sorter/distributed/prefix_doubling.hpp:
// 156-198) with BloomFilter::filter (src/sorter/distributed/bloomfilter.hpp:
// 1415-1556).  Round r uses depth 8 (r = 0) or 8 + 16*2^(r-1).  For every
// candidate position c of every PE, in local sorted order:
//   depth > len(c)                         -> end of string: results = len;
//   c-1 is a candidate and lcp[c] >= depth -> local LCP duplicate (stays);
//   otherwise                              -> hashed: XXHasher::hash of the
//                                             first depth chars % m.
// A hashed candidate stays a candidate iff its hash occurs more than once
// among the hashed candidates of all PEs, or the next position is a local
// LCP duplicate (the LCP-local root).  results = depth for every candidate of
// the round.  On one GPU the range-partitioned exchange of hashes
// (SendOnlyHashesToFilter / FindDuplicates, :581-765) is a global hash table
// in HBM: an equal hash of any two hashed candidates of any PEs marks both,
// which is what the merge of the owners' sorted hash runs reports.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dss_device.hpp"
#include "dss_internal.hpp"

namespace dss {

// ------------------------------------------------------------------ XXH64
// xxh::hash_state_t<64> with seed 0 fed by one update (src/hash/xxhash.hpp:
// 508-640): the published XXH64 algorithm, except that the digest takes the
// short-input branch (v3 + PRIME5) when exactly 32 bytes were hashed
// (`total_len > 32`, :573).
constexpr uint64_t P1 = 11400714785074694791ull;
constexpr uint64_t P2 = 14029467366897019727ull;
constexpr uint64_t P3 = 1609587929392839161ull;
constexpr uint64_t P4 = 9650029242287828579ull;
constexpr uint64_t P5 = 2870177450012600261ull;

__device__ inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
__device__ inline uint64_t xround(uint64_t acc, uint64_t in) {
    acc += in * P2;
    acc = rotl64(acc, 31);
    return acc * P1;
}
__device__ inline uint64_t xmerge(uint64_t acc, uint64_t v) {
    acc ^= xround(0, v);
    return acc * P1 + P4;
}

__device__ uint64_t xxh64_stream(const uint8_t* __restrict__ chars, uint64_t off, uint32_t len) {
    uint64_t v1 = P1 + P2, v2 = P2, v3 = 0, v4 = (uint64_t)0 - P1;
    uint32_t p = 0;
    if (len >= 32) {
        do {
            v1 = xround(v1, load_word_le(chars, off + p));
            v2 = xround(v2, load_word_le(chars, off + p + 8));
            v3 = xround(v3, load_word_le(chars, off + p + 16));
            v4 = xround(v4, load_word_le(chars, off + p + 24));
            p += 32;
        } while (p + 32 <= len);
    }
    uint64_t h;
    if (len > 32) {
        h = rotl64(v1, 1) + rotl64(v2, 7) + rotl64(v3, 12) + rotl64(v4, 18);
        h = xmerge(h, v1);
        h = xmerge(h, v2);
        h = xmerge(h, v3);
        h = xmerge(h, v4);
    } else {
        h = v3 + P5;
    }
    h += len;
    while (p + 8 <= len) {
        h ^= xround(0, load_word_le(chars, off + p));
        h = rotl64(h, 27) * P1 + P4;
        p += 8;
    }
    if (p + 4 <= len) {
        h ^= (load_word_le(chars, off + p) & 0xFFFFFFFFull) * P1;
        h = rotl64(h, 23) * P2 + P3;
        p += 4;
    }
    while (p < len) {
        h ^= (uint64_t)chars[off + p] * P5;
        h = rotl64(h, 11) * P1;
        ++p;
    }
    h ^= h >> 33;
    h *= P2;
    h ^= h >> 29;
    h *= P3;
    h ^= h >> 32;
    return h;
}

// ------------------------------------------------------------------ rounds
constexpr uint8_t kHashed = 1, kLcpDup = 2;
constexpr uint64_t kEmpty = ~0ull;  // never a hash value: hash % m < m <= 2^64-1

// Every position starts as a candidate; `skipped` leading rounds (the last
// at depth `dlast`) are already applied: in them every position stayed a
// candidate with results = depth.
__global__ void k_pd_init(uint8_t* __restrict__ cand, uint8_t* __restrict__ rounds, uint32_t* __restrict__ dist,
                          uint64_t n, uint32_t skipped, uint32_t dlast, uint32_t* __restrict__ sdepth) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        cand[i] = 1;
        rounds[i] = (uint8_t)min(skipped, 255u);
        dist[i] = skipped ? dlast : 0;
        if (sdepth) sdepth[i] = 0;
    }
}

// min over all positions c of max(lcp[c], lcp[c+1]) (lcp is 0 at every PE
// start, so lcp[c+1] of a PE's last position never counts).  A round at
// depth d <= this minimum changes nothing but results and the round count:
// every candidate c has a neighbour in its PE sharing >= d characters, which
// makes it a local LCP duplicate or an LCP-local root, and both stay
// candidates (and d <= the LCP <= len(c), so no string ends).
__global__ void k_pd_min_run(const uint32_t* __restrict__ lcp, uint64_t n, unsigned int* __restrict__ out) {
    uint32_t v = 0xFFFFFFFFu;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = n / 4;  // four positions per 16-byte load
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n4; w += stride) {
        const uint4 x = reinterpret_cast<const uint4*>(lcp)[w];
        const uint32_t nx = 4 * w + 4 < n ? lcp[4 * w + 4] : 0u;
        v = min(v, min(min(max(x.x, x.y), max(x.y, x.z)), min(max(x.z, x.w), max(x.w, nx))));
    }
    for (uint64_t c = 4 * n4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride)
        v = min(v, max(lcp[c], c + 1 < n ? lcp[c + 1] : 0u));
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_down(v, o));
    __shared__ uint32_t part[4];  // one atomic per block: single-address atomics serialise
    if (lane_id() == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        v = min(min(part[0], part[1]), min(part[2], part[3]));
        if (v != 0xFFFFFFFFu) atomicMin(out, v);
    }
}

// Sum of a per-thread count over the block, added to *out with one atomic.
__device__ inline void block_count_add(uint32_t v, unsigned long long* out) {
    __shared__ uint32_t part[4];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if (lane_id() == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(out, (unsigned long long)t);
    }
}

constexpr uint32_t kPdGrid = 4096;  // blocks of the grid-stride PD kernels

// h % m; the reference's range m = 2^64-1 only maps 2^64-1 to 0.
__device__ inline uint64_t mod_range(uint64_t h, uint64_t m) {
    if (m == ~0ull) return h == ~0ull ? 0 : h;
    return h % m;
}

constexpr uint32_t kPdPos = 64;            // positions per block iteration
constexpr uint32_t kWin = 256;             // prefix bytes staged per window
constexpr uint32_t kChunks = kWin / 16 + 1;  // 16-byte chunks per staged row
constexpr uint32_t kRowW = kChunks * 2;    // row length in 8-byte words

// 8 bytes at byte offset b of the staged rows (little endian).
__device__ inline uint64_t lds_word(const uint64_t* rows, uint32_t b) {
    const uint32_t w = b >> 3, sh = (b & 7) * 8;
    const uint64_t lo = rows[w];
    return sh ? (lo >> sh) | (rows[w + 1] << (64 - sh)) : lo;
}

// Classification of the candidates of one round, four positions per thread
// (word loads and stores of the byte and 32-bit arrays): end of string,
// local LCP duplicate or hashed.  Positions to hash collect in LDS over
// kMarkIters block iterations and are appended to `list` with one global
// atomic per flush.
constexpr uint32_t kMarkIters = 4;
constexpr uint32_t kMarkPer = 4;  // positions per thread and iteration

// One position: its flag; rounds and dist are updated for candidates.
__device__ inline uint8_t pd_mark_one(uint8_t cc, uint8_t cprev, uint32_t L, uint32_t lc, uint64_t c, uint32_t depth,
                                      uint8_t& rnd, uint32_t& d, bool& wd) {
    uint8_t f = 0;
    wd = false;
    if (cc) {
        // lcp is 0 at every PE start and depth >= 1, so lcp[c] >= depth
        // implies that c-1 is in the same PE
        if (depth <= L) f = (c > 0 && lc >= depth && cprev) ? kLcpDup : kHashed;
        rnd = rnd + (rnd < 255);
        d = depth > L ? L : depth;  // end of string: results = len
        wd = true;
    }
    return f;
}

__global__ void __launch_bounds__(256)
k_pd_mark(const uint8_t* __restrict__ cand, const uint32_t* __restrict__ lcp, const uint32_t* __restrict__ s_len,
          uint32_t depth, uint64_t n, uint8_t* __restrict__ flags, uint8_t* __restrict__ rounds,
          uint32_t* __restrict__ dist, uint32_t* __restrict__ list, unsigned long long* __restrict__ nhashed) {
    __shared__ uint32_t sl[256 * kMarkPer * kMarkIters];
    __shared__ uint32_t scount, base;
    const uint32_t lane = lane_id();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * kMarkPer;
    if (threadIdx.x == 0) scount = 0;
    __syncthreads();
    uint32_t it = 0;
    for (uint64_t b0 = (uint64_t)blockIdx.x * blockDim.x * kMarkPer; b0 < n; b0 += stride) {  // block-uniform
        const uint64_t c0 = b0 + (uint64_t)kMarkPer * threadIdx.x;
        uint8_t f[kMarkPer] = {0, 0, 0, 0};
        if (c0 + kMarkPer <= n) {
            const uint32_t C = *reinterpret_cast<const uint32_t*>(cand + c0);
            const uint8_t cprev = c0 > 0 ? cand[c0 - 1] : 0;
            const uint4 L = *reinterpret_cast<const uint4*>(s_len + c0);
            const uint4 LC = *reinterpret_cast<const uint4*>(lcp + c0);
            const uint32_t Ls[4] = {L.x, L.y, L.z, L.w}, Lc[4] = {LC.x, LC.y, LC.z, LC.w};
            uint32_t R = C ? *reinterpret_cast<const uint32_t*>(rounds + c0) : 0u;
            uint32_t d[4] = {0, 0, 0, 0};
            bool wd[4];
#pragma unroll
            for (uint32_t k = 0; k < kMarkPer; ++k) {
                const uint8_t cc = (uint8_t)(C >> (8 * k));
                const uint8_t cp = k ? (uint8_t)(C >> (8 * (k - 1))) : cprev;
                uint8_t rnd = (uint8_t)(R >> (8 * k));
                f[k] = pd_mark_one(cc, cp, Ls[k], Lc[k], c0 + k, depth, rnd, d[k], wd[k]);
                R = (R & ~(0xFFu << (8 * k))) | ((uint32_t)rnd << (8 * k));
            }
            if (C) *reinterpret_cast<uint32_t*>(rounds + c0) = R;
            if (wd[0] && wd[1] && wd[2] && wd[3]) {
                *reinterpret_cast<uint4*>(dist + c0) = make_uint4(d[0], d[1], d[2], d[3]);
            } else {
#pragma unroll
                for (uint32_t k = 0; k < kMarkPer; ++k)
                    if (wd[k]) dist[c0 + k] = d[k];
            }
            *reinterpret_cast<uint32_t*>(flags + c0) =
                (uint32_t)f[0] | ((uint32_t)f[1] << 8) | ((uint32_t)f[2] << 16) | ((uint32_t)f[3] << 24);
        } else {
            for (uint32_t k = 0; k < kMarkPer && c0 + k < n; ++k) {
                const uint64_t c = c0 + k;
                uint8_t rnd = rounds[c];
                uint32_t dk = 0;
                bool w = false;
                f[k] = pd_mark_one(cand[c], c > 0 ? cand[c - 1] : 0, s_len[c], lcp[c], c, depth, rnd, dk, w);
                if (w) {
                    rounds[c] = rnd;
                    dist[c] = dk;
                }
                flags[c] = f[k];
            }
        }
        // append in position order (the hash kernel writes a row per listed
        // position; consecutive positions keep those writes in whole lines):
        // a wave prefix sum of the threads' counts
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t k = 0; k < kMarkPer; ++k) cnt += f[k] == kHashed;
        uint32_t incl = cnt;
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        const uint32_t wtot = __shfl(incl, 63, 64);
        uint32_t wbase = 0;
        if (lane == 0 && wtot) wbase = atomicAdd(&scount, wtot);
        wbase = __shfl(wbase, 0, 64) + (incl - cnt);
#pragma unroll
        for (uint32_t k = 0; k < kMarkPer; ++k)
            if (f[k] == kHashed) sl[wbase++] = (uint32_t)(c0 + k);
        if (++it == kMarkIters || b0 + stride >= n) {  // block-uniform
            it = 0;
            __syncthreads();
            const uint32_t cnt = scount;
            if (threadIdx.x == 0 && cnt) base = (uint32_t)atomicAdd(nhashed, (unsigned long long)cnt);
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) list[base + k] = sl[k];
            __syncthreads();
            if (threadIdx.x == 0) scount = 0;
            __syncthreads();
        }
    }
}

// Hashing of the listed positions.  A block takes 64 positions per
// iteration, four lanes each.  The prefixes are staged through LDS in
// 256-byte windows with 16-byte loads (consecutive lanes read consecutive
// chunks of one string); lane j of a group then accumulates XXH64 lane v_j+1
// over the window's 32-byte stripes and lane 0 merges and hashes the tail.
// All loops are block-uniform (the hashed length is the round's depth), so
// every shuffle runs with all lanes active.  Hashes are stored in list order
// (these rounds list few positions: later passes then read them as a
// stream).
__global__ void __launch_bounds__(256)
k_pd_hash(const uint32_t* __restrict__ list, const unsigned long long* __restrict__ nhashed,
          const uint64_t* __restrict__ s_off, const uint8_t* __restrict__ chars, uint32_t depth, uint64_t m,
          uint64_t* __restrict__ hashv) {
    __shared__ uint64_t rows[kPdPos * kRowW];
    __shared__ uint64_t s_so[kPdPos];
    const uint32_t j = threadIdx.x & 3, g = threadIdx.x >> 2;
    const uint32_t base_lane = threadIdx.x & ~3u;
    const uint32_t stripes = depth / 32;
    const uint32_t nwin = (depth + kWin - 1) / kWin;
    const uint64_t nh = *nhashed;
    for (uint64_t blk = (uint64_t)blockIdx.x * kPdPos; blk < nh; blk += (uint64_t)gridDim.x * kPdPos) {
        const uint64_t i = blk + g;
        const bool hashed = i < nh;
        const uint32_t c = hashed ? list[i] : 0u;
        const uint64_t so = hashed ? s_off[c] : 0ull;
        if (j == 0) s_so[g] = hashed ? so : ~0ull;
        uint64_t v = j == 0 ? P1 + P2 : (j == 1 ? P2 : (j == 2 ? 0ull : (uint64_t)0 - P1));
        uint32_t shift = 0;
        for (uint32_t w = 0; w < nwin; ++w) {
            __syncthreads();  // s_so visible; previous window's rows consumed
            const uint32_t wend = min(depth, (w + 1) * kWin);  // bytes needed: [w*kWin, wend)
            for (uint32_t k = threadIdx.x; k < kPdPos * kChunks; k += blockDim.x) {
                const uint32_t row = k / kChunks, q = k - row * kChunks;
                const uint64_t o = s_so[row];
                if (o == ~0ull) continue;
                const uint64_t A = (o + (uint64_t)w * kWin) & ~15ull;
                if (A + 16ull * q < o + wend) {
                    const uint4 x = *reinterpret_cast<const uint4*>(chars + A + 16ull * q);
                    uint64_t* dst = &rows[row * kRowW + 2 * q];
                    dst[0] = ((uint64_t)x.y << 32) | x.x;
                    dst[1] = ((uint64_t)x.w << 32) | x.z;
                }
            }
            __syncthreads();
            shift = (uint32_t)((so + (uint64_t)w * kWin) & 15);
            if (hashed) {
                const uint64_t* rw = &rows[g * kRowW];
                for (uint32_t t = 0; t < kWin / 32; ++t) {
                    const uint32_t st = w * (kWin / 32) + t;
                    if (st < stripes) v = xround(v, lds_word(rw, shift + 32 * t + 8 * j));
                }
            }
        }
        const uint64_t v1 = __shfl(v, base_lane + 0);
        const uint64_t v2 = __shfl(v, base_lane + 1);
        const uint64_t v3 = __shfl(v, base_lane + 2);
        const uint64_t v4 = __shfl(v, base_lane + 3);
        if (hashed && j == 0) {
            uint64_t h;
            if (depth > 32) {
                h = rotl64(v1, 1) + rotl64(v2, 7) + rotl64(v3, 12) + rotl64(v4, 18);
                h = xmerge(h, v1);
                h = xmerge(h, v2);
                h = xmerge(h, v3);
                h = xmerge(h, v4);
            } else {
                h = v3 + P5;  // includes depth == 32, where v3 consumed one stripe
            }
            h += depth;
            // tail bytes [stripes*32, depth) lie in the last staged window
            const uint64_t* rw = &rows[g * kRowW];
            const uint32_t wbase = (nwin - 1) * kWin;
            uint32_t q = stripes * 32;
            while (q + 8 <= depth) {
                h ^= xround(0, lds_word(rw, shift + q - wbase));
                h = rotl64(h, 27) * P1 + P4;
                q += 8;
            }
            if (q + 4 <= depth) {
                h ^= (lds_word(rw, shift + q - wbase) & 0xFFFFFFFFull) * P1;
                h = rotl64(h, 23) * P2 + P3;
                q += 4;
            }
            while (q < depth) {
                h ^= (lds_word(rw, shift + q - wbase) & 0xFF) * P5;
                h = rotl64(h, 11) * P1;
                ++q;
            }
            h ^= h >> 33;
            h *= P2;
            h ^= h >> 29;
            h *= P3;
            h ^= h >> 32;
            hashv[i] = mod_range(h, m);  // in list order
        }
        __syncthreads();  // rows and s_so are rewritten by the next iteration
    }
}

// ------------------------------------------------ shared-prefix hashing
// When most positions are hashed and share long prefixes with their
// predecessors, the XXH64 state after the first L = S-1 full 32-byte stripes
// (S = full stripes of the depth, B = 32L characters) is the same along a
// run of positions whose LCP with the predecessor is at least B.  The first
// position of every run (its head: LCP below B, every PE start included)
// computes that state once; the other hashed candidates of the run continue
// from it with their own last stripe and tail.  The digest is unchanged.
// A scan over positions gives every position the rank of its run head.
constexpr uint32_t kHsT = 256, kHsI = 8, kHsB = kHsT * kHsI;

__global__ void __launch_bounds__(kHsT)
k_head_count(const uint32_t* __restrict__ lcp, uint64_t n, uint32_t B, uint32_t* __restrict__ bcnt) {
    __shared__ uint32_t red[kHsT / 64];
    const uint64_t b = (uint64_t)blockIdx.x * kHsB + (uint64_t)threadIdx.x * kHsI;
    uint32_t cnt = 0;
    if (b + kHsI <= n) {  // two 16-byte loads
        const uint4 x = reinterpret_cast<const uint4*>(lcp + b)[0], y = reinterpret_cast<const uint4*>(lcp + b)[1];
        cnt = (x.x < B) + (x.y < B) + (x.z < B) + (x.w < B) + (y.x < B) + (y.y < B) + (y.z < B) + (y.w < B);
    } else {
        for (uint32_t k = 0; k < kHsI; ++k)
            if (b + k < n) cnt += lcp[b + k] < B;
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o);
    if (lane_id() == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) bcnt[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Exclusive sum-scan of the block counts in place (one block); the total to
// *total.
__global__ void __launch_bounds__(1024) k_head_sums(uint32_t* __restrict__ v, uint64_t nb,
                                                    unsigned long long* __restrict__ total) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (nb + 1023) / 1024;
    const uint64_t b = t * per, e = (b + per < nb) ? b + per : nb;
    uint32_t sum = 0;
    for (uint64_t i = b; i < e; ++i) sum += v[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint32_t x = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (uint64_t i = b; i < e; ++i) {
        const uint32_t x = v[i];
        v[i] = run;
        run += x;
    }
    if (t == 1023) *total = part[1023];
}

// Per position: rank of its run head (heads before or at it, minus one),
// with kHeadBit set at the heads themselves; heads record their position
// under their rank.
constexpr uint32_t kHeadBit = 0x80000000u;
__global__ void __launch_bounds__(kHsT)
k_head_rank(const uint32_t* __restrict__ lcp, uint64_t n, uint32_t B, const uint32_t* __restrict__ bcnt,
            uint32_t* __restrict__ hrank, uint32_t* __restrict__ hpos) {
    __shared__ uint32_t part[kHsT];
    const uint64_t b = (uint64_t)blockIdx.x * kHsB + (uint64_t)threadIdx.x * kHsI;
    uint32_t f[kHsI];
    uint32_t cnt = 0;
    const bool whole = b + kHsI <= n;
    if (whole) {  // two 16-byte loads
        const uint4 x = reinterpret_cast<const uint4*>(lcp + b)[0], y = reinterpret_cast<const uint4*>(lcp + b)[1];
        f[0] = x.x < B; f[1] = x.y < B; f[2] = x.z < B; f[3] = x.w < B;
        f[4] = y.x < B; f[5] = y.y < B; f[6] = y.z < B; f[7] = y.w < B;
    } else {
        for (uint32_t k = 0; k < kHsI; ++k) f[k] = b + k < n && lcp[b + k] < B;
    }
    for (uint32_t k = 0; k < kHsI; ++k) cnt += f[k];
    part[threadIdx.x] = cnt;
    __syncthreads();
    for (uint32_t o = 1; o < kHsT; o <<= 1) {
        const uint32_t x = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = bcnt[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0u);
    uint32_t r[kHsI];
    for (uint32_t k = 0; k < kHsI; ++k) {
        if (b + k >= n) break;
        if (f[k]) hpos[run++] = (uint32_t)(b + k);
        r[k] = (run - 1) | (f[k] ? kHeadBit : 0u);
    }
    if (whole) {  // two 16-byte stores
        reinterpret_cast<uint4*>(hrank + b)[0] = make_uint4(r[0], r[1], r[2], r[3]);
        reinterpret_cast<uint4*>(hrank + b)[1] = make_uint4(r[4], r[5], r[6], r[7]);
    } else {
        for (uint32_t k = 0; k < kHsI && b + k < n; ++k) hrank[b + k] = r[k];
    }
}

// XXH64 lanes v1..v4 after the first L stripes of every run head.  A head
// shorter than 32L characters heads a run of one (its successor shares
// fewer characters), so its state is never used and not computed.
__global__ void k_head_state(const uint32_t* __restrict__ hpos, uint64_t nheads, const uint64_t* __restrict__ s_off,
                             const uint32_t* __restrict__ s_len, const uint8_t* __restrict__ chars, uint32_t L,
                             uint64_t* __restrict__ state) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nheads * 4) return;
    const uint64_t r = t >> 2;
    const uint32_t j = (uint32_t)(t & 3);
    if (s_len[hpos[r]] < 32 * L) {
        state[t] = 0;
        return;
    }
    const uint64_t o = s_off[hpos[r]];
    uint64_t v = j == 0 ? P1 + P2 : (j == 1 ? P2 : (j == 2 ? 0ull : (uint64_t)0 - P1));
    for (uint32_t k = 0; k < L; ++k) v = xround(v, load_word_le(chars, o + 32 * k + 8 * j));
    state[t] = v;
}

// Digest of XXH64 (the 32-byte quirk included) from the lanes v1..v4 after
// the full stripes and the little-endian words t0..t3 of the tail (the
// depth - 32·stripes < 32 characters after them).
__device__ inline uint64_t xxh64_digest(uint64_t v1, uint64_t v2, uint64_t v3, uint64_t v4, uint32_t depth,
                                        uint64_t t0, uint64_t t1, uint64_t t2, uint64_t t3) {
    uint64_t h;
    if (depth > 32) {
        h = rotl64(v1, 1) + rotl64(v2, 7) + rotl64(v3, 12) + rotl64(v4, 18);
        h = xmerge(h, v1);
        h = xmerge(h, v2);
        h = xmerge(h, v3);
        h = xmerge(h, v4);
    } else {
        h = v3 + P5;
    }
    h += depth;
    const uint32_t tail = depth & 31;
    uint32_t q = 0;
    while (q + 8 <= tail) {
        const uint64_t w = q == 0 ? t0 : (q == 8 ? t1 : t2);
        h ^= xround(0, w);
        h = rotl64(h, 27) * P1 + P4;
        q += 8;
    }
    const uint64_t w = q == 0 ? t0 : (q == 8 ? t1 : (q == 16 ? t2 : t3));  // the partial word
    uint32_t b = 0;
    if (q + 4 <= tail) {
        h ^= (w & 0xFFFFFFFFull) * P1;
        h = rotl64(h, 23) * P2 + P3;
        q += 4;
        b = 4;
    }
    for (; q < tail; ++q, ++b) {
        h ^= ((w >> (8 * b)) & 0xFF) * P5;
        h = rotl64(h, 11) * P1;
    }
    h ^= h >> 33;
    h *= P2;
    h ^= h >> 29;
    h *= P3;
    h ^= h >> 32;
    return h;
}

// Rounds of depth at most kLaneDepth hash one string per lane: the lane
// loads the aligned words that cover the string's first depth characters
// (independent loads, issued together) and runs XXH64 in registers.  Hashes
// are stored in list order.
constexpr uint32_t kLaneDepth = 64;
constexpr uint32_t kLaneWords = kLaneDepth / 8 + 1;  // aligned words covering kLaneDepth characters

__global__ void __launch_bounds__(256)
k_pd_hash_lane(const uint32_t* __restrict__ list, const unsigned long long* __restrict__ nhashed,
               const uint64_t* __restrict__ s_off, const uint8_t* __restrict__ chars, uint32_t depth, uint64_t m,
               uint64_t* __restrict__ hashv) {
    const uint64_t nh = *nhashed;
    const uint32_t nw = (depth + 7) / 8 + 1;  // <= kLaneWords
    const uint32_t stripes = depth / 32;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t o = s_off[list[i]];
        const uint64_t* w = reinterpret_cast<const uint64_t*>(chars + (o & ~7ull));
        const uint32_t sh = (uint32_t)(o & 7) * 8;
        uint64_t x[kLaneWords + 1];
#pragma unroll
        for (uint32_t k = 0; k < kLaneWords; ++k) x[k] = k < nw ? w[k] : 0ull;
        x[kLaneWords] = 0;
        // the little-endian word at character 8 k of the string (k < nw)
        auto word = [&](uint32_t k) -> uint64_t {
            uint64_t lo = 0, hi = 0;
#pragma unroll
            for (uint32_t q = 0; q < kLaneWords; ++q)
                if (q == k) {
                    lo = x[q];
                    hi = x[q + 1];
                }
            return sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
        };
        uint64_t v1 = P1 + P2, v2 = P2, v3 = 0, v4 = (uint64_t)0 - P1;
        for (uint32_t st = 0; st < stripes; ++st) {
            v1 = xround(v1, word(4 * st));
            v2 = xround(v2, word(4 * st + 1));
            v3 = xround(v3, word(4 * st + 2));
            v4 = xround(v4, word(4 * st + 3));
        }
        // tail words past the loaded ones are never read by the digest
        // (it reads depth % 32 characters)
        const uint32_t t = 4 * stripes;
        const uint64_t t0 = word(t), t1 = word(t + 1), t2 = word(t + 2), t3 = word(t + 3);
        hashv[i] = mod_range(xxh64_digest(v1, v2, v3, v4, depth, t0, t1, t2, t3), m);
    }
}

// XXH64 of the listed candidates, four lanes per string and kHashU strings
// per lane group at a time: a run head hashes its own bytes from the start,
// any other position continues from its head's state after L stripes.  Lane
// j loads word j of every stripe it hashes and word j of the tail up front,
// so a string waits for its list entry, its offset and rank, then its words
// and head state.  Hashes are stored by position.
constexpr uint32_t kHashU = 2;

__global__ void __launch_bounds__(256)
k_pd_hash_shared(const uint32_t* __restrict__ list, const unsigned long long* __restrict__ nhashed,
                 const uint64_t* __restrict__ s_off, const uint32_t* __restrict__ hrank,
                 const uint64_t* __restrict__ state, const uint8_t* __restrict__ chars, uint32_t depth, uint32_t L,
                 uint64_t m, uint64_t* __restrict__ hashv, uint64_t* __restrict__ stage, uint32_t* __restrict__ sdepth,
                 uint32_t stage_w) {
    const uint32_t j = threadIdx.x & 3, g = threadIdx.x >> 2;
    const uint32_t base_lane = threadIdx.x & ~3u;
    const uint32_t stripes = depth / 32;
    const bool tw_in = 8 * j < (depth & 31);  // lane j's tail word holds characters of the prefix
    const uint64_t nh = *nhashed;
    const uint64_t v0 = j == 0 ? P1 + P2 : (j == 1 ? P2 : (j == 2 ? 0ull : (uint64_t)0 - P1));
    constexpr uint64_t kPer = 64 * kHashU;
    for (uint64_t blk = (uint64_t)blockIdx.x * kPer; blk < nh; blk += (uint64_t)gridDim.x * kPer) {
        uint32_t c[kHashU], rk[kHashU];
        uint64_t own[kHashU], v[kHashU], tw[kHashU], last[kHashU];  // last: word j of the last full stripe
        bool hashed[kHashU];
#pragma unroll
        for (uint32_t u = 0; u < kHashU; ++u) {
            const uint64_t i = blk + 64 * u + g;
            hashed[u] = i < nh;
            c[u] = hashed[u] ? list[i] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kHashU; ++u) {
            own[u] = hashed[u] ? s_off[c[u]] : 0ull;
            rk[u] = hashed[u] ? hrank[c[u]] : kHeadBit;
        }
#pragma unroll
        for (uint32_t u = 0; u < kHashU; ++u) {
            v[u] = v0;
            tw[u] = 0;
            last[u] = 0;
            if (hashed[u]) {
                uint32_t k0 = 0;
                if (!(rk[u] & kHeadBit)) {
                    v[u] = state[4 * (uint64_t)rk[u] + j];
                    k0 = L;
                }
                if (tw_in) tw[u] = load_word_le(chars, own[u] + 32 * stripes + 8 * j);
                for (uint32_t k = k0; k < stripes; ++k) {
                    last[u] = load_word_le(chars, own[u] + 32 * k + 8 * j);
                    v[u] = xround(v[u], last[u]);
                }
            }
        }
        if (stage) {
            // characters [32(stripes-1), depth) for the exchange, stored after
            // every string's loads are issued
#pragma unroll
            for (uint32_t u = 0; u < kHashU; ++u) {
                if (!hashed[u]) continue;
                // packed rows of stage_w words: consecutive positions fill
                // whole lines (a partly written line costs a read as well)
                uint64_t* row = stage + stage_w * (uint64_t)c[u];
                row[j] = last[u];
                if (tw_in) row[4 + j] = tw[u];
                if (j == 0) sdepth[c[u]] = depth;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kHashU; ++u) {
            const uint64_t v1 = __shfl(v[u], base_lane + 0);
            const uint64_t v2 = __shfl(v[u], base_lane + 1);
            const uint64_t v3 = __shfl(v[u], base_lane + 2);
            const uint64_t v4 = __shfl(v[u], base_lane + 3);
            const uint64_t t0 = __shfl(tw[u], base_lane + 0);
            const uint64_t t1 = __shfl(tw[u], base_lane + 1);
            const uint64_t t2 = __shfl(tw[u], base_lane + 2);
            const uint64_t t3 = __shfl(tw[u], base_lane + 3);
            if (hashed[u] && j == 0) hashv[c[u]] = mod_range(xxh64_digest(v1, v2, v3, v4, depth, t0, t1, t2, t3), m);
        }
    }
}

__global__ void k_fill_u64(uint64_t* __restrict__ a, uint64_t n, uint64_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__device__ inline uint64_t slot_of(uint64_t h, uint64_t mask) {
    h ^= h >> 31;
    h *= 0x9E3779B97F4A7C15ull;
    return (h >> 17) & mask;
}

// The exact table takes the nm candidates listed in `maybe`: entry k names
// key keyv[maybe[k]] (keyv[kidx[maybe[k]]] when kidx is given), and slot[k]
// receives that key's slot.  `empty` is a value no key takes.
__global__ void k_pd_insert(const uint32_t* __restrict__ maybe, uint64_t nm, const uint64_t* __restrict__ keyv,
                            const uint32_t* __restrict__ kidx, uint64_t empty, uint64_t* __restrict__ keys,
                            uint8_t* __restrict__ dup, uint64_t mask, uint32_t* __restrict__ slot) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const uint32_t e = maybe[k];
    const uint64_t h = keyv[kidx ? kidx[e] : e];
    uint64_t s = slot_of(h, mask);
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long old =
            atomicCAS(reinterpret_cast<unsigned long long*>(&keys[s]), (unsigned long long)empty, (unsigned long long)h);
        if (old == empty) break;
        if (old == h) {
            dup[s] = 1;
            break;
        }
        s = (s + 1) & mask;
    }
    slot[k] = (uint32_t)s;
}

// ------------------------------------------------ duplicate prefilter
// Two bitmaps over 2^bits positions of mix64(hash) (about 16 bits per hashed
// candidate, small enough to stay in the last-level cache): `seen` and
// `again`.  A hash whose bit was set twice may occur more than once and goes
// to the exact open-addressing table (k_pd_insert); every other hash is
// unique.  Equal hashes always share a bit, so no duplicate is missed.
__device__ inline uint64_t pd_bit(uint64_t h, uint32_t bits) { return mix64(h) >> (64 - bits); }

__global__ void __launch_bounds__(256)
k_pd_bits(const uint32_t* __restrict__ list, bool by_pos, uint64_t nh, const uint64_t* __restrict__ hashv,
          uint32_t bits, uint32_t* __restrict__ seen, uint32_t* __restrict__ again) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = pd_bit(hashv[by_pos ? list[i] : i], bits);
        const uint32_t b = 1u << (x & 31);
        if (atomicOr(&seen[x >> 5], b) & b) atomicOr(&again[x >> 5], b);
    }
}

// Unique hashes get dup = 0; the others' list indices are appended to
// `maybe` (collected in LDS over kMarkIters iterations, one global atomic
// per flush).
__global__ void __launch_bounds__(256)
k_pd_maybe(const uint32_t* __restrict__ list, bool by_pos, uint64_t nh, const uint64_t* __restrict__ hashv,
           uint32_t bits,
           const uint32_t* __restrict__ again, uint8_t* __restrict__ dup, uint32_t* __restrict__ maybe,
           unsigned long long* __restrict__ nmaybe) {
    __shared__ uint32_t sl[256 * kMarkIters];
    __shared__ uint32_t scount, base;
    const uint32_t lane = lane_id();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (threadIdx.x == 0) scount = 0;
    __syncthreads();
    uint32_t it = 0;
    for (uint64_t b0 = (uint64_t)blockIdx.x * blockDim.x; b0 < nh; b0 += stride) {
        const uint64_t i = b0 + threadIdx.x;
        bool m = false;
        if (i < nh) {
            const uint32_t c = list[i];
            const uint64_t x = pd_bit(hashv[by_pos ? c : (uint32_t)i], bits);
            m = (again[x >> 5] >> (x & 31)) & 1;
            if (!m) dup[c] = 0;
        }
        const uint64_t bal = __ballot(m);
        uint32_t wb = 0;
        if (lane == 0 && bal) wb = atomicAdd(&scount, (uint32_t)__popcll(bal));
        wb = __shfl(wb, 0);
        if (m) sl[wb + (uint32_t)__popcll(bal & lanemask_lt())] = (uint32_t)i;
        if (++it == kMarkIters || b0 + stride >= nh) {  // block-uniform
            it = 0;
            __syncthreads();
            const uint32_t cnt = scount;
            if (threadIdx.x == 0 && cnt) base = (uint32_t)atomicAdd(nmaybe, (unsigned long long)cnt);
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) maybe[base + k] = sl[k];
            __syncthreads();
            if (threadIdx.x == 0) scount = 0;
            __syncthreads();
        }
    }
}

// Exact result for the table's candidates: their slot's flag, written at
// position pos[maybe[k]].
__global__ void k_pd_flag(const uint32_t* __restrict__ maybe, uint64_t nm, const uint32_t* __restrict__ pos,
                          const uint32_t* __restrict__ slot, const uint8_t* __restrict__ tdup,
                          uint8_t* __restrict__ dup) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nm) dup[pos[maybe[k]]] = tdup[slot[k]];
}

// ------------------------------------------------ partitioned duplicates
// Rounds with many hashed candidates: x = mix64(hash) (a bijection, so equal
// x means equal hashes) is put in order of its top kb bits by two stable LSD
// partition passes over tiles of kHpTile candidates (a per-tile digit
// histogram, an exclusive scan in digit-major order, a scatter that ranks
// equal digits with ballots), so that equal hashes share one bucket of about
// nh / 2^kb candidates.  A block per bucket then inserts the bucket's hashes
// into an LDS open-addressing table, flags slots that receive a hash twice,
// and marks the candidates of flagged slots as duplicates.  A bucket with too
// many distinct hashes for the table hands its records (indices into the
// partitioned keys) to the exact global table instead, which then takes x as
// the key.  Every access is a streamed read or write except the
// duplicate flags, which only duplicates write.
constexpr uint32_t kHpT = 512, kHpI = 4, kHpTile = kHpT * kHpI;
constexpr uint32_t kHpMaxBits = 9;     // digit bits per pass
constexpr uint32_t kHpCap = 2048;      // LDS table slots per bucket
constexpr uint64_t kHpMin = 1u << 20;  // hashed candidates from which rounds partition

__global__ void __launch_bounds__(kHpT)
k_hp_hist(const uint32_t* __restrict__ list, bool by_pos, const uint64_t* __restrict__ hashv,
          const uint64_t* __restrict__ keys,
          uint64_t nh, uint32_t shift, uint32_t bits, uint64_t ntiles, uint64_t* __restrict__ hist) {
    __shared__ uint32_t h[1u << kHpMaxBits];
    const uint32_t nb = 1u << bits, tid = threadIdx.x;
    for (uint32_t d = tid; d < nb; d += kHpT) h[d] = 0;
    __syncthreads();
    const uint64_t t0 = (uint64_t)blockIdx.x * kHpTile;
    for (uint32_t r = 0; r < kHpI; ++r) {
        const uint64_t i = t0 + r * kHpT + tid;
        if (i < nh) {
            const uint64_t x = keys ? keys[i] : mix64(hashv[by_pos ? list[i] : (uint32_t)i]);
            atomicAdd(&h[(uint32_t)(x >> shift) & (nb - 1)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = tid; d < nb; d += kHpT) hist[(uint64_t)d * ntiles + blockIdx.x] = h[d];
}

// Stable scatter of one tile by the digit (x >> shift) & (2^bits - 1): the
// first pass reads the candidate list and hashes (a record's id is its
// position), the second the first pass's output.  The tile is first reordered by digit in LDS and then
// written out run by run, so consecutive lanes write consecutive addresses
// of a digit's run instead of one line per record.  Wave w holds the tile's
// records [w kHpI 64, (w + 1) kHpI 64), kHpI per lane in (round, lane)
// order; a record's rank among the wave's earlier records of its digit comes
// from ballots (lanes of one round) and the wave's own running digit counts
// (earlier rounds), the waves' and digits' offsets from per-digit scans.
__global__ void __launch_bounds__(kHpT)
k_hp_scatter(const uint32_t* __restrict__ list, bool by_pos, const uint64_t* __restrict__ hashv,
             const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ ids_in, uint64_t nh, uint32_t shift,
             uint32_t bits, uint64_t ntiles, const uint64_t* __restrict__ off, uint64_t* __restrict__ keys_out,
             uint32_t* __restrict__ ids_out) {
    constexpr uint32_t kW = kHpT / 64, kWR = kHpI * 64;
    constexpr uint32_t kMaxD = 1u << kHpMaxBits;
    static_assert(kMaxD <= kHpT, "one digit per thread");
    __shared__ uint64_t skey[kHpTile];
    __shared__ uint32_t sid[kHpTile];
    __shared__ uint64_t gbase[kMaxD];    // the digit's output offset for this tile
    __shared__ uint32_t lstart[kMaxD];   // the digit's first slot in the reordered tile
    __shared__ uint16_t wcnt[kW][kMaxD]; // a wave's count of a digit, then the waves before it
    __shared__ uint32_t wtot[kW];
    const uint32_t nb = 1u << bits, tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const uint64_t t0 = (uint64_t)blockIdx.x * kHpTile;
    const bool dig = tid < nb;
    if (dig) gbase[tid] = off[(uint64_t)tid * ntiles + blockIdx.x];
    for (uint32_t k = tid; k < kW * kMaxD; k += kHpT) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    const uint64_t lt = lanemask_lt();
    uint64_t x[kHpI];
    uint32_t id[kHpI], dg[kHpI], rk[kHpI];
#pragma unroll
    for (uint32_t r = 0; r < kHpI; ++r) {
        const uint64_t i = t0 + w * kWR + r * 64 + lane;
        x[r] = 0;
        id[r] = 0;
        if (i < nh) {
            if (keys_in) {
                x[r] = keys_in[i];
                id[r] = ids_in[i];
            } else {
                id[r] = list[i];
                x[r] = mix64(hashv[by_pos ? id[r] : (uint32_t)i]);
            }
        }
        dg[r] = (uint32_t)(x[r] >> shift) & (nb - 1);
    }
#pragma unroll
    for (uint32_t r = 0; r < kHpI; ++r) {
        const bool valid = t0 + w * kWR + r * 64 + lane < nh;
        uint64_t m = __ballot(valid);
        for (uint32_t b = 0; b < bits; ++b) {
            const bool bit = (dg[r] >> b) & 1;
            const uint64_t bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        // every lane reads its digit's count before the digit's first lane
        // adds the round's (one wave: its LDS accesses keep program order)
        rk[r] = valid ? wcnt[w][dg[r]] + rank : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) wcnt[w][dg[r]] = (uint16_t)(wcnt[w][dg[r]] + __popcll(m));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: the waves' counts become their exclusive prefix, and the
    // digits' totals an exclusive scan (wave shuffles, then the wave sums)
    uint32_t tot = 0;
    if (dig) {
#pragma unroll
        for (uint32_t k = 0; k < kW; ++k) {
            const uint32_t c = wcnt[k][tid];
            wcnt[k][tid] = (uint16_t)tot;
            tot += c;
        }
    }
    uint32_t inc = tot;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    if (dig) {
        uint32_t pre = 0;
        for (uint32_t k = 0; k < w; ++k) pre += wtot[k];
        lstart[tid] = pre + inc - tot;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kHpI; ++r) {
        if (t0 + w * kWR + r * 64 + lane < nh) {
            const uint32_t q = lstart[dg[r]] + wcnt[w][dg[r]] + rk[r];
            skey[q] = x[r];
            sid[q] = id[r];
        }
    }
    __syncthreads();
    // runs out: slot k of digit d goes to gbase[d] + (k - lstart[d])
    const uint32_t cnt = (uint32_t)min<uint64_t>(kHpTile, nh - t0);
    for (uint32_t k = tid; k < cnt; k += kHpT) {
        const uint64_t xv = skey[k];
        const uint32_t d = (uint32_t)(xv >> shift) & (nb - 1);
        const uint64_t q = gbase[d] + (k - lstart[d]);
        keys_out[q] = xv;
        ids_out[q] = sid[k];
    }
}

// Bucket extents [bstart, bend) of the partitioned keys (top kb bits); empty
// buckets keep 0, 0.
__global__ void k_hp_bounds(const uint64_t* __restrict__ keys, uint64_t nh, uint32_t kb, uint32_t* __restrict__ bstart,
                            uint32_t* __restrict__ bend) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    const bool valid = i < nh;
    const uint64_t x = valid ? keys[i] : 0;
    // the neighbours' keys from the adjacent lanes (loads at the wave edges)
    uint64_t prev = __shfl_up(x, 1, 64), next = __shfl_down(x, 1, 64);
    if (!valid) return;
    if (lane == 0 && i > 0) prev = keys[i - 1];
    if (lane == 63 && i + 1 < nh) next = keys[i + 1];
    const uint64_t b = x >> (64 - kb);
    if (i == 0 || (prev >> (64 - kb)) != b) bstart[b] = (uint32_t)i;
    if (i + 1 == nh || (next >> (64 - kb)) != b) bend[b] = (uint32_t)(i + 1);
}

__global__ void __launch_bounds__(256)
k_hp_dedup(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ ids, const uint32_t* __restrict__ bstart,
           const uint32_t* __restrict__ bend, uint32_t nbuckets, uint32_t kb, uint32_t limit, uint8_t* __restrict__ dup,
           uint32_t* __restrict__ maybe, unsigned long long* __restrict__ nmaybe) {
    __shared__ unsigned long long tab[kHpCap];
    __shared__ uint8_t fl[kHpCap];
    __shared__ uint32_t used, over, mbase;
    const uint32_t tid = threadIdx.x;
    for (uint32_t b = blockIdx.x; b < nbuckets; b += gridDim.x) {
        const uint32_t s = bstart[b], e = bend[b];
        if (e - s < 2) continue;  // block-uniform
        // a value outside the bucket (other top bits) marks empty slots
        const unsigned long long empty = (unsigned long long)(b ^ 1u) << (64 - kb);
        for (uint32_t k = tid; k < kHpCap; k += 256) {
            tab[k] = empty;
            fl[k] = 0;
        }
        if (tid == 0) used = over = 0;
        __syncthreads();
        for (uint32_t i = s + tid; i < e; i += 256) {
            if (over) break;
            const unsigned long long x = keys[i];
            uint32_t slot = (uint32_t)x & (kHpCap - 1);
            uint32_t probe = 0;
            for (; probe < kHpCap; ++probe) {
                const unsigned long long old = atomicCAS(&tab[slot], empty, x);
                if (old == empty) {
                    if (atomicAdd(&used, 1u) >= limit) over = 1;
                    break;
                }
                if (old == x) {
                    fl[slot] = 1;
                    break;
                }
                slot = (slot + 1) & (kHpCap - 1);
            }
            if (probe == kHpCap) over = 1;
        }
        __syncthreads();
        if (over) {
            if (tid == 0) mbase = (uint32_t)atomicAdd(nmaybe, (unsigned long long)(e - s));
            __syncthreads();
            for (uint32_t i = s + tid; i < e; i += 256) maybe[mbase + (i - s)] = i;  // key index
        } else {
            for (uint32_t i = s + tid; i < e; i += 256) {
                const unsigned long long x = keys[i];
                uint32_t slot = (uint32_t)x & (kHpCap - 1);
                while (tab[slot] != x) slot = (slot + 1) & (kHpCap - 1);
                if (fl[slot]) dup[ids[i]] = 1;
            }
        }
        __syncthreads();  // the table is reset for the next bucket
    }
}

// Next round's candidates: LCP duplicates stay, hashed positions stay if
// they are an LCP-local root (the next position, in the same PE as lcp is 0
// at PE starts, is a local LCP duplicate) or a hash duplicate.  Eight
// positions per thread and step, so the byte arrays move as 8-byte words.
__device__ inline uint32_t pd_keep(uint8_t f, uint32_t lcp_next, uint8_t cand_next, bool has_next, uint8_t d,
                                   uint32_t depth) {
    if (f == kLcpDup) return 1;
    if (f == kHashed) return (has_next && lcp_next >= depth && cand_next) ? 1u : (d ? 1u : 0u);
    return 0;
}

__global__ void __launch_bounds__(256)
k_pd_next(const uint8_t* __restrict__ cand, const uint8_t* __restrict__ flags, const uint64_t* __restrict__ pe_begin,
          uint32_t p, const uint32_t* __restrict__ lcp, uint32_t depth, uint64_t n, const uint8_t* __restrict__ dup,
          uint8_t* __restrict__ next, unsigned long long* __restrict__ ncand) {
    uint32_t kept = 0;
    const uint64_t n8 = n / 8;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n8; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c0 = 8 * w;
        const uint64_t F = reinterpret_cast<const uint64_t*>(flags)[w];
        const uint64_t Dp = reinterpret_cast<const uint64_t*>(dup)[w];
        const uint64_t Cd = reinterpret_cast<const uint64_t*>(cand)[w];
        const uint4 la = reinterpret_cast<const uint4*>(lcp)[2 * w];
        const uint4 lb = reinterpret_cast<const uint4*>(lcp)[2 * w + 1];
        const bool tail = c0 + 8 < n;
        const uint32_t l8 = tail ? lcp[c0 + 8] : 0u;
        const uint8_t c8 = tail ? cand[c0 + 8] : 0;
        const uint32_t ln[8] = {la.y, la.z, la.w, lb.x, lb.y, lb.z, lb.w, l8};  // lcp of the next position
        uint64_t out = 0;
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const uint8_t cn = b < 7 ? (uint8_t)(Cd >> (8 * (b + 1))) : c8;
            const uint32_t k = pd_keep((uint8_t)(F >> (8 * b)), ln[b], cn, b < 7 || tail, (uint8_t)(Dp >> (8 * b)),
                                       depth);
            out |= (uint64_t)k << (8 * b);
            kept += k;
        }
        reinterpret_cast<uint64_t*>(next)[w] = out;
    }
    // the last n % 8 positions
    for (uint64_t c = 8 * n8 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (uint64_t)gridDim.x * blockDim.x) {
        const bool has = c + 1 < n;
        const uint32_t k = pd_keep(flags[c], has ? lcp[c + 1] : 0u, has ? cand[c + 1] : (uint8_t)0, has, dup[c], depth);
        next[c] = (uint8_t)k;
        kept += k;
    }
    block_count_add(kept, ncand);
}

// Duplicate flags of a round's nh hashed candidates by partitioning (see
// k_hp_dedup): dup is cleared and set for the repeats found in LDS; the
// records of buckets with more than table_limit distinct hashes are listed
// in pd_maybe (indices into the partitioned keys and positions) for the
// exact table.  Returns their number.  by_pos: hashes are stored by
// position, else in list order.
static uint64_t partition_duplicates(dss_set_s* s, uint64_t nh, uint32_t table_limit, bool by_pos) {
    PipeState& P = s->pp;
    const hipStream_t st = s->stream;
    unsigned long long* ctr = P.pd_ctr.as<unsigned long long>();
    // bucket bits: at most about 768 candidates per bucket on average (the
    // LDS table holds kHpCap * 7/8), so the digit passes stay narrow and
    // their runs per tile long (short runs leave partly written lines)
    uint32_t kb = 2;
    while (kb < 2 * kHpMaxBits && (nh >> kb) > kHpCap / 8 * 3) ++kb;
    const uint32_t b1 = kb / 2, b0 = kb - b1;  // pass 0: the lower b0 of the top kb bits; pass 1: the top b1
    const uint64_t ntiles = (nh + kHpTile - 1) / kHpTile;
    const uint32_t nbk = 1u << kb;
    P.pd_pkey.reserve(nh * 16);
    P.pd_pid.reserve(nh * 8);
    P.pd_phist.reserve((1ull << b0) * ntiles * 8);
    P.pd_pbound.reserve((uint64_t)nbk * 8);
    uint64_t* kA = P.pd_pkey.as<uint64_t>();
    uint64_t* kB = kA + nh;
    uint32_t* iA = P.pd_pid.as<uint32_t>();
    uint32_t* iB = iA + nh;
    uint64_t* hist = P.pd_phist.as<uint64_t>();
    const uint32_t* list = P.pd_list.as<uint32_t>();
    const uint64_t* hash = P.pd_hash.as<uint64_t>();
    {
        // per candidate: list entry and hash (twice), two 12-byte scatters
        // (read and write) and the second histogram's key read
        KScope ks(s->timer, st, "pd_partition", (double)nh * 68.0);
        const uint32_t sh0 = 64 - kb, sh1 = 64 - b1;
        hipLaunchKernelGGL(k_hp_hist, dim3(ntiles), dim3(kHpT), 0, st, list, by_pos, hash, nullptr, nh, sh0, b0, ntiles,
                           hist);
        exclusive_scan_u64(hist, hist, (1ull << b0) * ntiles, st, P, nullptr);
        hipLaunchKernelGGL(k_hp_scatter, dim3(ntiles), dim3(kHpT), 0, st, list, by_pos, hash, nullptr, nullptr, nh,
                           sh0, b0, ntiles, hist, kA, iA);
        hipLaunchKernelGGL(k_hp_hist, dim3(ntiles), dim3(kHpT), 0, st, nullptr, false, nullptr, kA, nh, sh1, b1, ntiles,
                           hist);
        exclusive_scan_u64(hist, hist, (1ull << b1) * ntiles, st, P, nullptr);
        hipLaunchKernelGGL(k_hp_scatter, dim3(ntiles), dim3(kHpT), 0, st, nullptr, false, nullptr, kA, iA, nh, sh1,
                           b1, ntiles, hist, kB, iB);
    }
    {
        // per candidate: bucket bounds (key read), table insert and lookup
        // (key read twice, id read)
        KScope ks(s->timer, st, "pd_dedup", (double)nh * 28.0);
        uint32_t* bstart = P.pd_pbound.as<uint32_t>();
        uint32_t* bend = bstart + nbk;
        fill_small(st, bstart, 0, (uint64_t)nbk * 8);
        hipLaunchKernelGGL(k_hp_bounds, dim3((nh + 255) / 256), dim3(256), 0, st, kB, nh, kb, bstart, bend);
        DSS_HIP(hipMemsetAsync(P.pd_dup.p, 0, s->n, st));
        hipLaunchKernelGGL(k_hp_dedup, dim3(std::min<uint32_t>(nbk, 65536)), dim3(256), 0, st, kB, iB, bstart, bend,
                           nbk, kb, table_limit, P.pd_dup.as<uint8_t>(), P.pd_maybe.as<uint32_t>(), ctr + 2);
    }
    DSS_HIP(hipGetLastError());
    uint64_t nm = 0;
    s->pp.mbox.fetch(st, &nm, ctr + 2, 8);
    return nm;
}

constexpr uint64_t kStageMax = 16ull << 30;

void prefix_doubling(dss_set_s* s, uint64_t m, uint32_t start_depth) {
    DSS_REQUIRE(s->sorted, DSS_ESTATE, "dss_sort_lcp has not run");
    DSS_REQUIRE(m >= 1, DSS_EINVAL, "hash range m must be >= 1");
    DSS_REQUIRE(start_depth >= 1, DSS_EINVAL, "start depth must be >= 1");
    PipeState& P = s->pp;
    const hipStream_t st = s->stream;
    const uint64_t n = s->n;
    const uint32_t p = s->p;
    const uint64_t nn = std::max<uint64_t>(n, 1);
    P.dist.reserve(nn * 4);
    P.rounds.reserve(nn);
    P.pd_cand.reserve(nn);
    P.pd_next.reserve(nn);
    P.pd_flags.reserve(nn);
    P.pd_hash.reserve(nn * 8);
    P.pd_slot.reserve(nn * 4);
    P.pd_list.reserve(nn * 4);
    P.pd_maybe.reserve(nn * 4);
    P.pd_dup.reserve(nn);
    P.pd_ctr.reserve(64);
    uint8_t* cand = P.pd_cand.as<uint8_t>();
    uint8_t* next = P.pd_next.as<uint8_t>();
    unsigned long long* ctr = P.pd_ctr.as<unsigned long long>();
    const dim3 grid((n + 255) / 256), blk(256);
    uint64_t ncand = n;
    uint32_t round = 0;
    uint64_t depth = start_depth, step = (uint64_t)start_depth * 2;
    // leading rounds in which every string has a neighbour sharing at least
    // the round's depth keep every candidate: they are applied at once
    uint32_t skipped = 0, dlast = 0, stage_w = 0;
    // the characters the last stripe and the tail of shared-prefix hashing
    // read are kept for the exchange (at most kStageMax bytes)
    const bool stage_on = n && n * 64 <= kStageMax;
    if (stage_on) {
        P.pd_stage.reserve(n * 64);
        P.pd_sdepth.reserve(n * 4);
    }
    if (n) {
        unsigned int minrun = 0;
        {
            KScope ks(s->timer, st, "pd_skip", (double)n * 4.0);
            fill_small(st, ctr, 0xFF, 4);
            hipLaunchKernelGGL(k_pd_min_run, dim3(std::min<uint64_t>((n + 255) / 256, 1024)), blk, 0, st,
                               s->lcp.as<uint32_t>(), n, reinterpret_cast<unsigned int*>(ctr));
        }
        s->pp.mbox.fetch(st, &minrun, ctr, 4);
        while (depth <= minrun) {
            dlast = (uint32_t)depth;
            ++skipped;
            depth = start_depth + step;
            step *= 2;
        }
        round = skipped;
        hipLaunchKernelGGL(k_pd_init, grid, blk, 0, st, cand, P.rounds.as<uint8_t>(), P.dist.as<uint32_t>(), n,
                           skipped, dlast, stage_on ? P.pd_sdepth.as<uint32_t>() : nullptr);
    }
    // duplicate detection by partitioning from part_min hashed candidates on;
    // both settings can be lowered by the environment (tests cover the
    // partitioned path and its exact-table fallback on small inputs)
    uint64_t part_min = kHpMin;
    uint32_t table_limit = kHpCap / 8 * 7;
    const bool trace = getenv("DSS_PD_TRACE") != nullptr;  // per-round counts to stderr
    if (const char* e = getenv("DSS_PD_PARTITION_MIN")) part_min = std::max<uint64_t>(strtoull(e, nullptr, 10), 1);
    if (const char* e = getenv("DSS_PD_TABLE_LIMIT"))
        table_limit = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(strtoull(e, nullptr, 10), 1), kHpCap / 8 * 7);
    while (ncand > 0) {
        DSS_REQUIRE(depth < 0xFFFFFFFFull, DSS_ERANGE, "prefix doubling depth overflow");
        DSS_REQUIRE(2 * ncand < (1ull << 32), DSS_ERANGE, "too many candidates for the hash table");
        fill_small(st, ctr, 0, 32);
        {
            // every position: candidate flags of c and c-1, length, lcp, the
            // flag written; every candidate: rounds and dist
            KScope ks(s->timer, st, "pd_mark", (double)n * 11.0 + (double)ncand * 6.0);
            hipLaunchKernelGGL(k_pd_mark, dim3(std::min<uint64_t>((n + 256 * kMarkPer - 1) / (256 * kMarkPer), kPdGrid)),
                               blk, 0, st, cand,
                               s->lcp.as<uint32_t>(), s->s_len.as<uint32_t>(), (uint32_t)depth, n,
                               P.pd_flags.as<uint8_t>(), P.rounds.as<uint8_t>(), P.dist.as<uint32_t>(),
                               P.pd_list.as<uint32_t>(), ctr);
        }
        uint64_t nh = 0;
        s->pp.mbox.fetch(st, &nh, ctr, 8);
        // shared-prefix hashing: at least 2 full stripes, at least half of the
        // positions hashed, and at most one run head per 4 hashed candidates
        const uint32_t stripes = (uint32_t)(depth / 32);
        const uint32_t L = stripes >= 2 ? stripes - 1 : 0;
        const uint32_t B = 32 * L;
        bool shared = L >= 1 && nh * 2 >= n;
        uint64_t nheads = 0;
        const uint64_t nbh = (n + kHsB - 1) / kHsB;
        if (shared) {
            KScope ks(s->timer, st, "pd_heads", (double)n * 4.0);
            P.pd_hmax.reserve((nbh + 1) * 4);
            hipLaunchKernelGGL(k_head_count, dim3(nbh), dim3(kHsT), 0, st, s->lcp.as<uint32_t>(), n, B,
                               P.pd_hmax.as<uint32_t>());
            hipLaunchKernelGGL(k_head_sums, dim3(1), dim3(1024), 0, st, P.pd_hmax.as<uint32_t>(), nbh, ctr + 3);
            s->pp.mbox.fetch(st, &nheads, ctr + 3, 8);
            shared = nheads * 4 <= nh;
        }
        if (shared) {
            // ranks: lcp read, rank written, head positions; states: L stripes per head
            KScope ks(s->timer, st, "pd_heads", (double)n * 8.0 + (double)nheads * (B + 36.0));
            P.pd_head.reserve(nn * 4);
            P.pd_hpos.reserve(std::max<uint64_t>(nheads, 1) * 4);
            P.pd_hstate.reserve(std::max<uint64_t>(nheads, 1) * 32);
            hipLaunchKernelGGL(k_head_rank, dim3(nbh), dim3(kHsT), 0, st, s->lcp.as<uint32_t>(), n, B,
                               P.pd_hmax.as<uint32_t>(), P.pd_head.as<uint32_t>(), P.pd_hpos.as<uint32_t>());
            hipLaunchKernelGGL(k_head_state, dim3((nheads * 4 + 255) / 256), blk, 0, st, P.pd_hpos.as<uint32_t>(),
                               nheads, s->s_off.as<uint64_t>(), s->s_len.as<uint32_t>(), s->arena(), L,
                               P.pd_hstate.as<uint64_t>());
        }
        if (shared && stage_on) {
            // rows of another width would overlap the kept ones: drop those
            const uint32_t w = 4 + ((uint32_t)(depth & 31) + 7) / 8;
            if (stage_w && w != stage_w) DSS_HIP(hipMemsetAsync(P.pd_sdepth.p, 0, n * 4, st));
            stage_w = w;
        }
        {
            KScope ks(s->timer, st, "pd_hash", 0.0);
            const dim3 g(std::min<uint64_t>((nh + kPdPos - 1) / kPdPos, kPdGrid));
            const dim3 gs(std::min<uint64_t>((nh + 64 * kHashU - 1) / (64 * kHashU), kPdGrid));
            if (shared)
                hipLaunchKernelGGL(k_pd_hash_shared, gs, blk, 0, st, P.pd_list.as<uint32_t>(), ctr,
                                   s->s_off.as<uint64_t>(), P.pd_head.as<uint32_t>(), P.pd_hstate.as<uint64_t>(),
                                   s->arena(), (uint32_t)depth, L, m, P.pd_hash.as<uint64_t>(),
                                   stage_on ? P.pd_stage.as<uint64_t>() : nullptr,
                                   stage_on ? P.pd_sdepth.as<uint32_t>() : nullptr, stage_w);
            else if (nh && depth <= kLaneDepth)
                hipLaunchKernelGGL(k_pd_hash_lane, dim3(stride_grid(nh, 256)), blk, 0, st, P.pd_list.as<uint32_t>(),
                                   ctr, s->s_off.as<uint64_t>(), s->arena(), (uint32_t)depth, m,
                                   P.pd_hash.as<uint64_t>());
            else if (nh)
                hipLaunchKernelGGL(k_pd_hash, g, blk, 0, st, P.pd_list.as<uint32_t>(), ctr, s->s_off.as<uint64_t>(),
                                   s->arena(), (uint32_t)depth, m, P.pd_hash.as<uint64_t>());
        }
        if (trace)
            fprintf(stderr, "[dss pd] round %u depth %llu candidates %llu hashed %llu shared %d heads %llu\n", round,
                    (unsigned long long)depth, (unsigned long long)ncand, (unsigned long long)nh, (int)shared,
                    (unsigned long long)nheads);
        // every hashed candidate: list entry, offset, the characters read
        // from its own string (all of them, or beyond B), hash
        s->timer.amend_last("pd_hash", (double)nh * ((shared ? depth - B : depth) + 24.0));
        // the shared-prefix kernel stores hashes by position (its rounds hash
        // at least half of the positions), the plain kernel in list order
        const bool by_pos = shared;
        uint64_t nm = 0;  // possible repeats for the exact table
        // the exact table's keys: keyv[maybe[k]] (keyv[kidx[maybe[k]]]), at
        // position pos[maybe[k]]
        const uint64_t* keyv = P.pd_hash.as<uint64_t>();
        const uint32_t* kidx = by_pos ? P.pd_list.as<uint32_t>() : nullptr;
        const uint32_t* kpos = P.pd_list.as<uint32_t>();
        uint64_t kempty = kEmpty;
        if (nh >= part_min) {
            nm = partition_duplicates(s, nh, table_limit, by_pos);
            // keys x = mix64(hash): a bijection, and no hash is kEmpty
            keyv = P.pd_pkey.as<uint64_t>() + nh;
            kidx = nullptr;
            kpos = P.pd_pid.as<uint32_t>() + nh;
            kempty = mix64(kEmpty);
        } else if (nh) {
            uint32_t bits = 16;
            while (bits < 32 && (1ull << bits) < 16 * nh) ++bits;
            const uint64_t words = (1ull << bits) / 32;
            P.pd_bits.reserve(words * 8);
            uint32_t* seen = P.pd_bits.as<uint32_t>();
            uint32_t* again = seen + words;
            DSS_HIP(hipMemsetAsync(seen, 0, words * 8, st));
            const dim3 g(std::min<uint64_t>((nh + 255) / 256, kPdGrid));
            {
                // hashed candidates: list entry, hash, two bitmap atomics
                // (seen, rarely again), one bitmap read, the flag
                KScope ks(s->timer, st, "pd_prefilter", (double)nh * 25.0);
                hipLaunchKernelGGL(k_pd_bits, g, blk, 0, st, P.pd_list.as<uint32_t>(), by_pos, nh,
                                   P.pd_hash.as<uint64_t>(), bits, seen, again);
                hipLaunchKernelGGL(k_pd_maybe, g, blk, 0, st, P.pd_list.as<uint32_t>(), by_pos, nh,
                                   P.pd_hash.as<uint64_t>(), bits, again, P.pd_dup.as<uint8_t>(),
                                   P.pd_maybe.as<uint32_t>(), ctr + 2);
            }
            s->pp.mbox.fetch(st, &nm, ctr + 2, 8);
        }
        if (nm) {
            uint64_t T = 1024;
            while (T < 2 * nm) T <<= 1;
            P.pd_table.reserve(T * 8);
            P.pd_tflag.reserve(T);
            // table fill, per possible repeat: hash, slot, flags
            KScope ks(s->timer, st, "pd_table", (double)T * 9.0 + (double)nm * 30.0);
            hipLaunchKernelGGL(k_fill_u64, dim3((T + 255) / 256), blk, 0, st, P.pd_table.as<uint64_t>(), T, kempty);
            DSS_HIP(hipMemsetAsync(P.pd_tflag.p, 0, T, st));
            hipLaunchKernelGGL(k_pd_insert, dim3((nm + 255) / 256), blk, 0, st, P.pd_maybe.as<uint32_t>(), nm, keyv,
                               kidx, kempty, P.pd_table.as<uint64_t>(), P.pd_tflag.as<uint8_t>(), T - 1,
                               P.pd_slot.as<uint32_t>());
            hipLaunchKernelGGL(k_pd_flag, dim3((nm + 255) / 256), blk, 0, st, P.pd_maybe.as<uint32_t>(), nm, kpos,
                               P.pd_slot.as<uint32_t>(), P.pd_tflag.as<uint8_t>(), P.pd_dup.as<uint8_t>());
        }
        {
            // every position: flag, candidate flags, lcp, duplicate flag, next
            KScope ks(s->timer, st, "pd_next", (double)n * 8.0);
            hipLaunchKernelGGL(k_pd_next, dim3(std::min<uint64_t>((n + 255) / 256, kPdGrid)), blk, 0, st, cand,
                               P.pd_flags.as<uint8_t>(), s->pe_begin_d.as<uint64_t>(), p, s->lcp.as<uint32_t>(),
                               (uint32_t)depth, n, P.pd_dup.as<uint8_t>(), next, ctr + 1);
        }
        DSS_HIP(hipGetLastError());
        s->pp.mbox.fetch(st, &ncand, ctr + 1, 8);
        std::swap(cand, next);
        ++round;
        depth = start_depth + step;
        step *= 2;
    }
    P.pd_rounds = round;
    P.has_dist = true;
    P.stage_valid = stage_on && stage_w;
    P.stage_w = stage_w;
    P.partitioned = P.exchanged = P.merged = P.extended = false;
}

// ------------------------------------------------------------------ drivers
// Hash of the first min(depth, len) characters of the given strings.
__global__ void k_hash_ids(const uint32_t* __restrict__ ids, uint64_t count, const uint64_t* __restrict__ off,
                           const uint32_t* __restrict__ len, const uint8_t* __restrict__ chars, uint64_t n,
                           uint32_t depth, uint64_t m, uint64_t* __restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t id = ids[k];
    if (id >= n) {
        out[k] = 0;
        return;
    }
    const uint32_t L = min(depth, len[id]);
    out[k] = mod_range(xxh64_stream(chars, off[id], L), m);
}

}  // namespace dss

using namespace dss;

extern "C" {

int dss_prefix_doubling(dss_set s, uint64_t m, uint32_t start_depth, uint32_t* num_rounds) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        prefix_doubling(s, m, start_depth);
        if (num_rounds) *num_rounds = s->pp.pd_rounds;
    });
}

int dss_get_prefix_doubling(dss_set s, uint32_t* results, uint8_t* rounds) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        DSS_REQUIRE(s->pp.has_dist, DSS_ESTATE, "dss_prefix_doubling has not run");
        if (s->n) {
            if (results) DSS_HIP(hipMemcpyAsync(results, s->pp.dist.p, s->n * 4, hipMemcpyDeviceToHost, s->stream));
            if (rounds) DSS_HIP(hipMemcpyAsync(rounds, s->pp.rounds.p, s->n, hipMemcpyDeviceToHost, s->stream));
        }
        DSS_HIP(hipStreamSynchronize(s->stream));
    });
}

int dss_hash_prefixes(dss_set s, const uint32_t* ids, uint64_t count, uint32_t depth, uint64_t m, uint64_t* out) {
    return guarded([&] {
        DSS_REQUIRE(s && (count == 0 || (ids && out)), DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(m >= 1, DSS_EINVAL, "hash range m must be >= 1");
        if (!count) return;
        DevBuf buf;
        buf.reserve(count * 12);
        uint32_t* d_ids = reinterpret_cast<uint32_t*>(buf.as<uint8_t>() + count * 8);
        DSS_HIP(hipMemcpyAsync(d_ids, ids, count * 4, hipMemcpyHostToDevice, s->stream));
        hipLaunchKernelGGL(k_hash_ids, dim3((count + 255) / 256), dim3(256), 0, s->stream, d_ids, count,
                           s->off.as<uint64_t>(), s->len.as<uint32_t>(), s->arena(), s->n, depth, m,
                           buf.as<uint64_t>());
        DSS_HIP(hipGetLastError());
        DSS_HIP(hipMemcpyAsync(out, buf.p, count * 8, hipMemcpyDeviceToHost, s->stream));
        DSS_HIP(hipStreamSynchronize(s->stream));
    });
}

int dss_merge_sort(dss_set s, int sampler, int compressed) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        DSS_REQUIRE(compressed >= 0 && compressed <= 2, DSS_EINVAL, "compressed must be 0, 1 or 2");
        sort_lcp(s);
        if (s->p == 1) return;
        // lcpSummand (lcp_aware...:207,236); the NoLcps variant sorts without
        // LCPs (:218-224), so getAvgLcp reads the zero-initialised array
        const uint64_t avg = (compressed == 2 ? 0 : avg_lcp(s)) + 5;
        partition(s, sampler, 2, 2 * 100 * avg, false);
        exchange(s, compressed == 2 ? 3 : compressed);
        merge(s);
        if (compressed == 1) extend_prefix(s);
    });
}

int dss_prefix_doubling_sort(dss_set s, int sampler, uint64_t m) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        sort_lcp(s);
        if (s->p == 1) return;
        prefix_doubling(s, m, 8);
        partition(s, sampler, 2, 0, true);
        exchange(s, 2);
        merge(s);
        permutation(s);  // the writeback of prefix_doubling.hpp:282-290, on the device
    });
}

int dss_get_permutation(dss_set s, uint32_t* string_index, uint32_t* pe_index) {
    return guarded([&] {
        DSS_REQUIRE(s && string_index && pe_index, DSS_EINVAL, "NULL argument");
        PipeState& P = s->pp;
        DSS_REQUIRE(P.merged, DSS_ESTATE, "no merged output");
        const uint64_t n = s->n;
        if (!n) return;
        permutation(s);  // already written by dss_prefix_doubling_sort
        const uint32_t* si = P.perm_out.as<uint32_t>();
        DSS_HIP(hipMemcpyAsync(string_index, si, n * 4, hipMemcpyDeviceToHost, s->stream));
        DSS_HIP(hipMemcpyAsync(pe_index, si + perm_pe_off(n), n * 4, hipMemcpyDeviceToHost, s->stream));
        DSS_HIP(hipStreamSynchronize(s->stream));
    });
}

}  // extern "C"
