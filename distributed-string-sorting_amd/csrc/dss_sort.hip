// (a2) Local LCP string sort of all virtual PEs of a set.
//
// Replaces tlx radixsort_CI3 (TLXV/sort/strings/radix_sort.hpp:539-880) at
// src/sorter/distributed/lcp_aware_number_string_splitting.hpp:218-224 and
// src/sorter/distributed/prefix_doubling.hpp:228.  Output per position: the
// sorted string id and lcp[i] = LCP(s[i-1], s[i]) inside the PE, lcp at a PE
// start = 0.  The LCP rules are those of the reference: strings that end
// together are equal and get lcp = their length, the first string of every
// further non-empty bucket gets lcp = the bucket depth.
//
// Algorithm (MSD radix over 8-char big-endian keys, all PEs in one launch
// sequence; every PE starts as one segment):
//   large segments (> large_min strings, a runtime parameter), per pass,
//   over 2048-string tiles:
//     prefix skip  -- common prefix of the segment from its depth: groups of
//                     8 or 16 lanes compare each string with the segment's
//                     first string, 32 bytes per lane and step, loads bounded
//                     by the minimum found so far;
//     key gather   -- 8 chars at the new depth per string;
//     histogram    -- 256 digit counters per tile in LDS (a digit is one
//                     byte, or on small alphabets up to 3 ranked
//                     characters: see DigitMap);
//     scan/classify-- digit offsets (two-level over 64-tile chunks for long
//                     segments), sub-bucket lists, LCP at sub-bucket
//                     boundaries;
//     scatter      -- stable rank inside the tile by wave ballots, writes the
//                     next key/id buffers, the id array and LCPs of equal
//                     strings.  The digit passes over one gathered key
//                     cover its 8 characters (8 passes of one character, 3
//                     of up to three).
//   small segments (2..large_min strings), batched by power-of-two class:
//     up to 64 strings one segment per wave (register bitonic sort), larger
//     ones in 1024-slot LDS blocks; both sort (group, key, id) and refine
//     equal-key groups 8 characters at a time until every group is a single
//     string or a run of equal strings.  Ties are broken by id, so the order
//     of equal strings is their input order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dss_device.hpp"
#include "dss_internal.hpp"

namespace dss {

constexpr uint32_t kSmallMax = 1024;  // base-case capacity (slots per block)
constexpr uint32_t kClasses = 11;     // class c holds sizes (2^(c-1), 2^c], c = 1..10
constexpr uint32_t kTile = 2048;      // strings per tile in the global passes
constexpr uint32_t kThreads = 256;
constexpr uint32_t kChunkTiles = 64;  // tiles per chunk of the two-level digit-offset scan
constexpr uint32_t kWaveClassMax = 6;  // default: classes up to 64 strings use the one-wave base case

struct Seg {
    uint32_t begin, cnt;
    uint32_t depth;   // chars known to be common to all members
    uint32_t kdepth;  // depth of the cached keys, kNone if keys must be gathered
    uint32_t cp;      // common prefix found by the prefix skip (from depth)
    uint32_t tile0;   // first tile of the segment in the current pass
    uint32_t chunk0;  // first tile chunk (segments of more than kChunkTiles tiles)
};

struct Tile {
    uint32_t seg, begin, end;
    uint32_t keyed;  // common prefix at which the prefix skip wrote the tile's keys, or kNone
};

// A base-case segment; `cache` says where the 8-char keys of its strings
// already are: 0 none, else bits 0-2 the segment depth minus the keys' depth
// (1..7: the keys' last 8 - that characters are the segment's next ones) and
// bit 3 the key buffer (0 A, 1 B) the last MSD pass scattered them into.
struct SmallSeg {
    uint32_t begin, cnt, depth, cache;
};

struct SmallLists {
    uint64_t off[kClasses];
    uint64_t cap[kClasses];
};

// counters: [3 + c] small class c, [3 + kClasses + c] strings in small
// class c; two sets of per-pass counters at kCtrPass + 4 * (pass & 1):
// [+0] large segments of the list, [+1] tiles, [+2] tile chunks
constexpr uint32_t kCtrLarge = 0, kCtrTiles = 1, kCtrChunks = 2, kCtrSmall = 3,
                   kCtrSmallElems = 3 + kClasses, kCtrPass = 32;

__device__ inline void push_small(SmallSeg* small, const SmallLists& L, uint32_t* ctr,
                                  uint32_t begin, uint32_t cnt, uint32_t depth) {
    const uint32_t c = 32u - __builtin_clz(cnt - 1u);  // ceil(log2 cnt), cnt >= 2
    const uint32_t k = atomicAdd(&ctr[kCtrSmall + c], 1u);
    atomicAdd(&ctr[kCtrSmallElems + c], cnt);
    small[L.off[c] + k] = SmallSeg{begin, cnt, depth, 0};
}

// ------------------------------------------------------------------ init
// D of each PE's sorted strings (computeD, src/tests/DNGenerator_tests.cpp:
// 22-45) over chunks of at most kDistChunk positions: chunk k covers
// [begin, begin + cnt) of PE pe; the PE's first LCP counts as 0 and the
// position after its last string contributes no LCP.
constexpr uint64_t kDistChunk = 1 << 16;

// Block b covers positions [b kDistChunk, (b + 1) kDistChunk) and adds the
// contribution of each PE it overlaps (PEs of fewer than two strings
// contribute nothing) to that PE's sum.
__global__ void __launch_bounds__(256) k_pe_dsum(const uint32_t* __restrict__ lcp, const uint64_t* __restrict__ pe_begin,
                                                 uint32_t p, uint64_t n, unsigned long long* __restrict__ d) {
    __shared__ unsigned long long part[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * kDistChunk, b1 = min(n, b0 + kDistChunk);
    uint32_t lo = 0, hi = p;  // the last PE that begins at or before b0
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (pe_begin[mid] <= b0) lo = mid;
        else hi = mid;
    }
    for (uint32_t r = lo; r < p && pe_begin[r] < b1; ++r) {  // block-uniform
        const uint64_t pb = pe_begin[r], pe = pe_begin[r + 1];
        unsigned long long t = 0;
        if (pe - pb >= 2)
            for (uint64_t i = max(pb, b0) + threadIdx.x; i < min(pe, b1); i += blockDim.x) {
                const uint32_t x = i == pb ? 0u : lcp[i];
                const uint32_t y = i + 1 < pe ? lcp[i + 1] : 0u;
                t += (uint64_t)max(x, y) + 1;
            }
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
        if (lane_id() == 0) part[threadIdx.x >> 6] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long sum = part[0] + part[1] + part[2] + part[3];
            if (sum) atomicAdd(d + r, sum);
        }
        __syncthreads();  // part is rewritten for the next PE
    }
}

// Identity order, zero LCPs, and "no metadata yet" (kNone) in the sorted
// length array: base cases fill offset/length of the strings they place,
// k_sorted_meta the rest.
// kDev (device-driven form): the string count comes from dn and the grid
// strides over it.
template <bool kDev>
__global__ void k_sort_init(uint32_t* __restrict__ perm, uint32_t* __restrict__ lcp, uint32_t* __restrict__ out_len,
                            uint64_t n, const uint64_t* __restrict__ dn) {
    if (kDev) n = *dn;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        perm[i] = (uint32_t)i;
        lcp[i] = 0;
        out_len[i] = kNone;
    }
}

template <bool kDev>
__global__ void k_sort_seed(const uint64_t* __restrict__ pe_begin, uint32_t p, uint32_t large_min,
                            Seg* __restrict__ segs, SmallSeg* __restrict__ small, SmallLists L,
                            uint32_t* __restrict__ ctr, const uint32_t* __restrict__ dp) {
    if (kDev) p = *dp;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < p; r += gridDim.x * blockDim.x) {
        const uint32_t b = (uint32_t)pe_begin[r];
        const uint32_t cnt = (uint32_t)(pe_begin[r + 1] - pe_begin[r]);
        if (cnt > large_min) {
            const uint32_t k = atomicAdd(&ctr[kCtrPass + kCtrLarge], 1u);
            segs[k] = Seg{b, cnt, 0, kNone, kNone, 0, 0};
        } else if (cnt >= 2) {
            push_small(small, L, ctr, b, cnt, 0);
        }
    }
}

// ------------------------------------------------------------------ tiles
// Tile and chunk numbering of the large segments (one block scans the
// per-segment counts); the tiles themselves are written by k_fill_tiles.
// Segments of more than kChunkTiles tiles also get chunks of kChunkTiles
// tiles for the two-level digit-offset scan.
__device__ inline uint32_t seg_tiles(uint32_t cnt) { return (cnt + kTile - 1) / kTile; }
__device__ inline uint32_t seg_chunks(uint32_t nt) {
    return nt > kChunkTiles ? (nt + kChunkTiles - 1) / kChunkTiles : 0u;
}

// 1024 threads: wave-level shuffle scans of the tile and chunk counts, the
// 16 wave totals combined through LDS (three barriers per 1024 segments).
__global__ void __launch_bounds__(1024) k_build_tiles(Seg* __restrict__ segs, uint32_t* __restrict__ pc) {
    const uint32_t nseg = pc[kCtrLarge];  // written by the previous pass (or the seed)
    __shared__ uint32_t wsum[16], wcsum[16], woff[16], wcoff[16];
    __shared__ uint32_t carry, ccarry;
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry = ccarry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nseg; base += 1024) {  // block-uniform
        const uint32_t i = base + t;
        const uint32_t nt = i < nseg ? seg_tiles(segs[i].cnt) : 0u;
        const uint32_t nc = seg_chunks(nt);
        uint32_t a = nt, c = nc;  // inclusive scans within the wave
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t x = __shfl_up(a, o, 64), y = __shfl_up(c, o, 64);
            if (lane >= o) {
                a += x;
                c += y;
            }
        }
        if (lane == 63) {
            wsum[w] = a;
            wcsum[w] = c;
        }
        __syncthreads();
        if (t == 0) {
            uint32_t x = carry, y = ccarry;
            for (uint32_t k = 0; k < 16; ++k) {
                woff[k] = x;
                wcoff[k] = y;
                x += wsum[k];
                y += wcsum[k];
            }
            carry = x;
            ccarry = y;
        }
        __syncthreads();
        if (i < nseg) {
            Seg s = segs[i];
            s.tile0 = woff[w] + a - nt;
            s.chunk0 = wcoff[w] + c - nc;
            if (s.kdepth == kNone) s.cp = kNone;
            segs[i] = s;
        }
        __syncthreads();  // woff is rewritten for the next 1024 segments
    }
    if (t == 0) {
        pc[kCtrTiles] = carry;
        pc[kCtrChunks] = ccarry;
    }
}

// Last segment whose first tile (or first chunk) is <= x: tile0 and chunk0
// do not decrease along the segment list.
template <bool kChunk>
__device__ inline uint32_t seg_of(const Seg* __restrict__ segs, uint32_t nseg, uint32_t x) {
    uint32_t lo = 0, hi = nseg;  // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t v = kChunk ? segs[mid].chunk0 : segs[mid].tile0;
        if (v <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// kDev: segment and tile counts from the pass counters hp, grid-stride.
// pn: the next pass's counters, cleared here (before this pass's
// classification counts into them).
template <bool kDev>
__global__ void k_fill_tiles(const Seg* __restrict__ segs, uint32_t nseg, uint32_t ntiles,
                             Tile* __restrict__ tiles, const uint32_t* __restrict__ hp, uint32_t* __restrict__ pn) {
    if (blockIdx.x == 0 && threadIdx.x < 4) pn[threadIdx.x] = 0;
    if (kDev) {
        nseg = hp[kCtrLarge];
        ntiles = hp[kCtrTiles];
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
        const uint32_t i = seg_of<false>(segs, nseg, t);
        const Seg s = segs[i];
        const uint32_t b = s.begin + (t - s.tile0) * kTile;
        tiles[t] = Tile{i, b, min(s.begin + s.cnt, b + kTile), kNone};
    }
}

// Per chunk and digit: the sum of the chunk's tile histograms.
template <bool kDev>
__global__ void __launch_bounds__(kThreads)
k_chunk_sum(const Seg* __restrict__ segs, uint32_t nseg, const uint32_t* __restrict__ hist,
            uint32_t* __restrict__ csum, const uint32_t* __restrict__ hp) {
    const uint32_t d = threadIdx.x;
    if (kDev) nseg = hp[kCtrLarge];
    const uint32_t nch = kDev ? hp[kCtrChunks] : blockIdx.x + 1;
    for (uint32_t k = blockIdx.x; k < nch; k += gridDim.x) {
        const Seg s = segs[seg_of<true>(segs, nseg, k)];
        const uint32_t t0 = s.tile0 + (k - s.chunk0) * kChunkTiles;
        const uint32_t t1 = min(s.tile0 + seg_tiles(s.cnt), t0 + kChunkTiles);
        uint32_t tot = 0;
        for (uint32_t t = t0; t < t1; ++t) tot += hist[(uint64_t)t * 256 + d];
        csum[(uint64_t)k * 256 + d] = tot;
    }
}

// Per chunk and digit: tile offsets from the chunk's base offset.
template <bool kDev>
__global__ void __launch_bounds__(kThreads)
k_chunk_apply(const Seg* __restrict__ segs, uint32_t nseg, uint32_t* __restrict__ hist,
              const uint32_t* __restrict__ csum, const uint32_t* __restrict__ hp) {
    const uint32_t d = threadIdx.x;
    if (kDev) nseg = hp[kCtrLarge];
    const uint32_t nch = kDev ? hp[kCtrChunks] : blockIdx.x + 1;
    for (uint32_t k = blockIdx.x; k < nch; k += gridDim.x) {
        const Seg s = segs[seg_of<true>(segs, nseg, k)];
        const uint32_t t0 = s.tile0 + (k - s.chunk0) * kChunkTiles;
        const uint32_t t1 = min(s.tile0 + seg_tiles(s.cnt), t0 + kChunkTiles);
        uint32_t run = csum[(uint64_t)k * 256 + d];
        for (uint32_t t = t0; t < t1; ++t) {
            uint32_t* h = &hist[(uint64_t)t * 256 + d];
            const uint32_t x = *h;
            *h = run;
            run += x;
        }
    }
}

// Effective key depth and digit depth of a segment in the current pass.
__device__ inline void seg_depths(const Seg& s, uint32_t& kd, uint32_t& dep) {
    if (s.kdepth == kNone) {
        kd = s.depth + s.cp;
        dep = kd;
    } else {
        kd = s.kdepth;
        dep = s.depth;
    }
}

// ------------------------------------------------------------ digits
// A radix digit covers kc characters.  With a small alphabet (sigma values:
// the string end, rank 0, and the byte values that occur, ranked in byte
// order) kc characters c_0..c_{kc-1} form the digit sum rank(c_i)
// sigma^(kc-1-i) < 256, which orders like the characters; kc = 1 is the
// byte itself.  A digit whose last rank is 0 holds strings that end inside
// it (the ranks after an end are 0 too); digit 0 holds strings that end at
// its first character.  The map is computed on the device per sort from the
// bytes of the arena (k_alphabet, k_digit_map).
struct DigitMap {
    uint32_t kc, sigma;
    uint8_t rank[256];
};

constexpr uint64_t kDigitBitsAt = 512;   // the present-byte bits after the map in SortScratch::digit
constexpr uint64_t kDigitBytes = 1024;

struct DigitCtx {
    uint32_t kc, sigma;
    const uint8_t* rank;  // LDS copy of DigitMap::rank (kc > 1)
};

// The digit at byte `pos` of a big-endian 8-character key (pos + kc <= 8).
__device__ inline uint32_t digit_of(uint64_t key, uint32_t pos, const DigitCtx& D) {
    if (D.kc == 1) return (uint32_t)(key >> (56 - 8 * pos)) & 0xFF;
    uint32_t d = 0;
    for (uint32_t i = 0; i < D.kc; ++i) d = d * D.sigma + D.rank[(uint32_t)(key >> (56 - 8 * (pos + i))) & 0xFF];
    return d;
}

// Characters two different digits share at their start.
__device__ inline uint32_t digit_lead(uint32_t a, uint32_t b, const DigitCtx& D) {
    if (D.kc == 1) return 0;
    uint32_t pw = 1;
    for (uint32_t i = 1; i < D.kc; ++i) pw *= D.sigma;
    for (uint32_t i = 0; i < D.kc; ++i, pw /= D.sigma)
        if ((a / pw) % D.sigma != (b / pw) % D.sigma) return i;
    return D.kc;
}

// Characters of a (non-zero) digit before the end of its strings: kc unless
// they end inside it.
__device__ inline uint32_t digit_chars(uint32_t d, const DigitCtx& D) {
    if (D.kc == 1) return 1;
    uint32_t pw = 1;
    for (uint32_t i = 1; i < D.kc; ++i) pw *= D.sigma;
    for (uint32_t i = 0; i < D.kc; ++i, pw /= D.sigma)
        if ((d / pw) % D.sigma == 0) return i;
    return D.kc;
}

// A segment's digit context: the digit starting at byte pos of its keys
// covers at most the 8 - pos characters the keys still hold, so that keys
// are re-gathered every 8 characters whatever the digit width.
__device__ inline DigitCtx seg_digits(const DigitCtx& D, uint32_t pos) {
    return D.kc == 1 ? D : DigitCtx{min(D.kc, 8u - pos), D.sigma, D.rank};
}

// The block's digit context: kMulti false is one character per digit (a
// compile-time constant, so the byte digits compile as before); otherwise
// the ranks are staged in `lds` (the caller synchronises).
template <bool kMulti>
__device__ inline DigitCtx digit_ctx(const DigitMap* __restrict__ dm, uint8_t* lds) {
    if (!kMulti) return DigitCtx{1u, 256u, lds};
    for (uint32_t t = threadIdx.x; t < 256; t += blockDim.x) lds[t] = dm->rank[t];
    return DigitCtx{dm->kc, dm->sigma, lds};
}

// Byte values present in chars[lo, hi): a block marks them in LDS with plain
// stores (16 bytes per thread and step) and ORs them into bits[8].  With
// `gate`, blocks return at once when bits already holds more than 15
// non-zero byte values (more than one character per digit is then ruled
// out).
constexpr uint64_t kAlphaSample = 1ull << 20;  // bytes scanned before the gated rest

__global__ void __launch_bounds__(256)
k_alphabet(const uint8_t* __restrict__ chars, uint64_t lo, uint64_t hi, uint32_t* __restrict__ bits, int gate) {
    __shared__ uint8_t present[256];
    __shared__ int stop;
    const uint32_t t = threadIdx.x;
    present[t] = 0;
    if (t == 0) {
        int c = 0;
        if (gate) {
            for (int w = 0; w < 8; ++w) c += __popc(bits[w]);
            c -= (int)(bits[0] & 1);  // byte 0 is the end marker
        }
        stop = c > 15;
    }
    __syncthreads();
    if (stop) return;  // block-uniform
    const uint64_t a0 = (lo + 15) & ~15ull, a1 = max(a0, hi & ~15ull);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t a = a0 + 16ull * ((uint64_t)blockIdx.x * blockDim.x + t); a < a1; a += 16 * stride) {
        const uint4 x = *reinterpret_cast<const uint4*>(chars + a);
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) present[(w[i] >> (8 * k)) & 0xFF] = 1;
    }
    // the unaligned edges
    if (blockIdx.x == 0) {
        for (uint64_t a = lo + t; a < min(a0, hi); a += blockDim.x) present[chars[a]] = 1;
        for (uint64_t a = max(a1, lo) + t; a < hi; a += blockDim.x) present[chars[a]] = 1;
    }
    __syncthreads();
    const uint64_t m = __ballot(present[t] != 0);
    if ((t & 31) == 0) {
        const uint32_t word = (uint32_t)(m >> (t & 32));
        if (word) atomicOr(&bits[t >> 5], word);
    }
}

// The digit map from the present byte values: sigma = 1 + present non-zero
// values, kc the most characters (at most max_kc) with sigma^kc <= 256.
__global__ void __launch_bounds__(256)
k_digit_map(const uint32_t* __restrict__ bits, uint32_t max_kc, DigitMap* __restrict__ dm) {
    __shared__ uint32_t wc[4];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const bool pres = t > 0 && ((bits[t >> 5] >> (t & 31)) & 1);
    const uint64_t m = __ballot(pres);
    if (lane == 0) wc[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = (uint32_t)__popcll(m & lanemask_lt());
    for (uint32_t k = 0; k < w; ++k) before += wc[k];
    dm->rank[t] = pres ? (uint8_t)(1 + before) : (uint8_t)0;
    if (t == 0) {
        const uint32_t sigma = 1 + wc[0] + wc[1] + wc[2] + wc[3];
        uint32_t kc = 1;
        while (kc < max_kc) {
            uint32_t pw = 1;
            for (uint32_t i = 0; i <= kc; ++i) pw *= sigma;
            if (pw > 256) break;
            ++kc;
        }
        dm->kc = kc;
        dm->sigma = sigma;
    }
}

// ------------------------------------------------------------ prefix skip
// Groups of 8 (or 16) lanes compare kU strings at a time against the
// segment's first string, 32 bytes per lane and step (four 8-char keys from
// five aligned loads), so a wave keeps 8·kU (4·kU) strings and 32·kU
// independent key loads in flight; the reference's keys of a step are
// loaded once for the kU strings (its first-step keys stay in registers).
// The tile's offsets and lengths are staged in LDS first, so a step waits
// for one dependent load.  The minimum found so far (the block's, and every
// 8 rounds the segment's) bounds each further comparison: keys at or beyond
// it cannot lower the minimum and are not loaded, and a string stops once its
// compared prefix reaches it.  All loops that contain cross-lane operations
// have block- or wave-uniform trip counts, so every shuffle runs with all
// lanes active.  The lane that finds a string's first difference (or end)
// keeps the string's 8 characters from there (with the next lane's first key
// when the difference lies in its last one); the block then writes its
// strings' ids and their keys at the block minimum, composed from the
// reference's characters before the difference and the kept ones, without
// reading the strings again (strings without kept characters are loaded).
// The key gather keeps the keys if the block minimum is the segment's, which
// the whole grid decides.
template <uint32_t kSkipLanes, uint32_t kU>
__device__ __forceinline__ void prefix_skip_tile(uint32_t tb, Tile* __restrict__ tiles, Seg* __restrict__ segs,
                                                 const uint32_t* __restrict__ perm,
                                                 const uint8_t* __restrict__ chars, const uint64_t* __restrict__ off,
                                                 const uint32_t* __restrict__ len, uint64_t* __restrict__ key,
                                                 uint32_t* __restrict__ pb) {
    __shared__ uint64_t t_off[kTile];  // offset; after a string's event its kept characters
    __shared__ uint32_t t_len[kTile];
    __shared__ uint32_t t_res[kTile];  // event depth - segment depth if characters were kept
    __shared__ uint32_t bmin;
    const Tile T = tiles[tb];
    const Seg S = segs[T.seg];
    if (S.kdepth != kNone) return;
    const uint32_t cnt = T.end - T.begin;
    if (threadIdx.x == 0) bmin = kNone;
    // the first pass (pb == nullptr) starts from the identity order: ids are
    // positions, and neither the order is read nor the id buffer written
    for (uint32_t k = threadIdx.x; k < cnt; k += kThreads) {
        const uint32_t id = pb ? perm[T.begin + k] : T.begin + k;
        t_off[k] = off[id];
        t_len[k] = len[id];
        t_res[k] = kNone;
        if (pb) pb[T.begin + k] = id;
    }
    __syncthreads();
    constexpr uint32_t kSkipGroups = kThreads / kSkipLanes;
    constexpr uint32_t kPerRound = kSkipGroups * kU;
    const uint32_t g = threadIdx.x / kSkipLanes, l = threadIdx.x % kSkipLanes;
    const uint32_t ref = pb ? perm[S.begin] : S.begin;
    const uint64_t roff = off[ref];
    const uint32_t rlen = len[ref];
    const uint32_t d0 = S.depth;
    uint64_t rk[4];
    load_keys4(chars, roff, rlen, d0 + 32 * l, rk);
    const volatile uint32_t* cpp = &segs[T.seg].cp;
    const volatile uint32_t* vbmin = &bmin;
    uint32_t glim = kNone;
    const uint32_t rounds = (cnt + kPerRound - 1) / kPerRound;
    for (uint32_t r = 0; r < rounds; ++r) {
        uint64_t so[kU];
        uint32_t sl[kU], res[kU];
        bool done[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t k = (r * kU + u) * kSkipGroups + g;
            done[u] = k >= cnt;
            so[u] = 0;
            sl[u] = 0;
            res[u] = kNone;
            if (!done[u]) {
                so[u] = t_off[k];
                sl[u] = t_len[k];
            }
        }
        if ((r & 7) == 0) glim = *cpp;
        uint32_t lim = min(glim, *vbmin);
        uint32_t base = d0;  // the kU strings of a group step together
        for (;;) {
            bool any = false;
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u) any |= !done[u];
            if (__ballot(any) == 0) break;  // wave-uniform condition
            const uint32_t p0 = base + 32 * l;
            const bool inwin = lim == kNone || p0 - d0 < lim;
            uint64_t a[kU][4], b[4];
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u)
                if (!done[u] && inwin) load_keys4(chars, so[u], sl[u], p0, a[u]);
            if (base == d0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = rk[j];
            } else if (any && inwin) {
                load_keys4(chars, roff, rlen, p0, b);
            }
            // the next lane's window starts where this lane's ends
            const bool next_in = l + 1 < kSkipLanes && (lim == kNone || p0 + 32 - d0 < lim);
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u) {
                uint32_t m = kNone;
                uint64_t cap = 0;  // the string's 8 characters from m
                bool kept = false;
                bool want_next = false;
                if (!done[u] && inwin) {
#pragma unroll
                    for (int j = 3; j >= 0; --j) {  // the first event of the lane wins
                        if (a[u][j] != b[j]) {
                            const uint32_t e = key_lcp(a[u][j], b[j]);
                            m = p0 + 8 * j + e;
                            cap = a[u][j] << (8 * e);
                            if (e && j < 3) cap |= a[u][j + 1] >> (64 - 8 * e);
                            kept = e == 0 || j < 3;
                            want_next = !kept;
                        } else if ((a[u][j] & 0xFF) == 0) {
                            m = p0 + 8 * j + key_strlen(a[u][j]);
                            cap = 0;  // a key is zero after its string's end
                            kept = true;
                            want_next = false;
                        }
                    }
                }
                if (__ballot(want_next)) {  // the rest comes from the next lane's first key
                    const uint64_t nx = __shfl_down((unsigned long long)a[u][0], 1, kSkipLanes);
                    if (want_next && next_in) {
                        cap |= nx >> (64 - 8 * ((m - p0) & 7));
                        kept = true;
                    }
                }
                const uint32_t mine = m;
                // lanes cover consecutive 32-byte pieces of the step: the
                // minimum over the group is the first mismatch or end
                for (int o = kSkipLanes / 2; o > 0; o >>= 1)
                    m = min(m, (uint32_t)__shfl_xor(m, o, kSkipLanes));
                if (!done[u] && m != kNone) {
                    res[u] = m - d0;
                    done[u] = true;
                    if (mine == m && kept) {
                        const uint32_t k = (r * kU + u) * kSkipGroups + g;
                        t_off[k] = cap;
                        t_res[k] = m - d0;
                    }
                }
            }
            base += 32 * kSkipLanes;
            lim = min(lim, *vbmin);
            if (lim != kNone && base - d0 >= lim) {  // cannot lower the minimum
#pragma unroll
                for (uint32_t u = 0; u < kU; ++u) done[u] = true;
            }
        }
        uint32_t best = kNone;
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) best = min(best, res[u]);
        if (l == 0 && best != kNone) atomicMin(&bmin, best);
    }
    __syncthreads();
    const uint32_t bm = bmin;
    if (bm == kNone) return;
    if (threadIdx.x == 0) {
        atomicMin(&segs[T.seg].cp, bm);
        tiles[tb].keyed = bm;
    }
    // a string equals the reference before its event depth r >= bm: its key
    // at bm is the reference's r - bm characters, then its kept ones
    const uint64_t rkey = load_key8(chars, roff, rlen, d0 + bm);
    for (uint32_t k = threadIdx.x; k < cnt; k += kThreads) {
        const uint32_t r = t_res[k];
        uint64_t kk;
        if (r == kNone) {
            kk = load_key8(chars, t_off[k], t_len[k], d0 + bm);
        } else {
            const uint32_t dl = r - bm;
            kk = dl >= 8 ? rkey : dl ? (rkey & (~0ull << (64 - 8 * dl))) | (t_off[k] >> (8 * dl)) : t_off[k];
        }
        key[T.begin + k] = kk;
    }
}

// kDev (device-driven form, persistent grid): the tile count comes from the
// pass counters hp and every block loops over tiles.
template <uint32_t kSkipLanes, uint32_t kU, bool kDev>
__global__ void __launch_bounds__(kThreads)
k_prefix_skip(Tile* __restrict__ tiles, Seg* __restrict__ segs, const uint32_t* __restrict__ perm,
              const uint8_t* __restrict__ chars, const uint64_t* __restrict__ off,
              const uint32_t* __restrict__ len, uint64_t* __restrict__ key, uint32_t* __restrict__ pb,
              const uint32_t* __restrict__ hp) {
    if constexpr (!kDev) {
        prefix_skip_tile<kSkipLanes, kU>(blockIdx.x, tiles, segs, perm, chars, off, len, key, pb);
    } else {
        const uint32_t nb = hp[kCtrTiles];
        for (uint32_t tb = blockIdx.x; tb < nb; tb += gridDim.x) {
            prefix_skip_tile<kSkipLanes, kU>(tb, tiles, segs, perm, chars, off, len, key, pb);
            __syncthreads();  // LDS of this tile read before the next tile's
        }
    }
}

// ------------------------------------------------------------ key gather
__device__ __forceinline__ void gather_tile(uint32_t tb, const Tile* __restrict__ tiles, const Seg* __restrict__ segs,
                                            const uint32_t* __restrict__ perm, const uint8_t* __restrict__ chars,
                                            const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                                            uint64_t* __restrict__ key) {
    const Tile T = tiles[tb];
    const Seg S = segs[T.seg];
    // the prefix skip wrote the ids, and the keys if its block minimum is the segment's
    if (S.kdepth != kNone || T.keyed == S.cp) return;
    const uint32_t kd = S.depth + S.cp;
    for (uint32_t e = T.begin + threadIdx.x; e < T.end; e += kThreads) {
        const uint32_t id = perm[e];
        key[e] = load_key8(chars, off[id], len[id], kd);
    }
}

template <bool kDev>
__global__ void __launch_bounds__(kThreads)
k_gather(const Tile* __restrict__ tiles, const Seg* __restrict__ segs, const uint32_t* __restrict__ perm,
         const uint8_t* __restrict__ chars, const uint64_t* __restrict__ off,
         const uint32_t* __restrict__ len, uint64_t* __restrict__ key, const uint32_t* __restrict__ hp) {
    if constexpr (!kDev) {
        gather_tile(blockIdx.x, tiles, segs, perm, chars, off, len, key);
    } else {
        const uint32_t nb = hp[kCtrTiles];
        for (uint32_t tb = blockIdx.x; tb < nb; tb += gridDim.x) {
            gather_tile(tb, tiles, segs, perm, chars, off, len, key);
            __syncthreads();  // LDS of this tile read before the next tile's
        }
    }
}

// ------------------------------------------------------------ histogram
__device__ __forceinline__ void hist_tile(uint32_t tb, const Tile* __restrict__ tiles, const Seg* __restrict__ segs,
                                          const uint64_t* __restrict__ key, uint32_t* __restrict__ hist,
                                          const DigitCtx& D) {
    __shared__ uint32_t h[256];
    const Tile T = tiles[tb];
    const Seg S = segs[T.seg];
    uint32_t kd, dep;
    seg_depths(S, kd, dep);
    const DigitCtx Ds = seg_digits(D, dep - kd);
    h[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t e = T.begin + threadIdx.x; e < T.end; e += kThreads)
        atomicAdd(&h[digit_of(key[e], dep - kd, Ds)], 1u);
    __syncthreads();
    hist[(uint64_t)tb * 256 + threadIdx.x] = h[threadIdx.x];
}

// kMulti: digits from the digit map dm (else one byte each).
template <bool kDev, bool kMulti>
__global__ void __launch_bounds__(kThreads)
k_hist(const Tile* __restrict__ tiles, const Seg* __restrict__ segs, const uint64_t* __restrict__ key,
       uint32_t* __restrict__ hist, const DigitMap* __restrict__ dm, const uint32_t* __restrict__ hp) {
    __shared__ uint8_t rk[256];
    const DigitCtx D = digit_ctx<kMulti>(dm, rk);  // hist_tile synchronises before its first use
    if constexpr (!kDev) {
        hist_tile(blockIdx.x, tiles, segs, key, hist, D);
    } else {
        const uint32_t nb = hp[kCtrTiles];
        for (uint32_t tb = blockIdx.x; tb < nb; tb += gridDim.x) {
            hist_tile(tb, tiles, segs, key, hist, D);
            __syncthreads();  // LDS of this tile read before the next tile's
        }
    }
}

// ------------------------------------------------------------ scan/classify
// One block per large segment: digit totals and offsets over the segment's
// tiles, LCPs at bucket boundaries, and the sub-bucket lists.  List slots
// are reserved with one global atomic per list per block.
// kScGroups segments per block, 256 threads (one per digit) each; the
// block's small-class and large-list counts reach the global counters with
// one atomic per list (same-address atomics of many small blocks serialise).
constexpr uint32_t kScGroups = 4;

__device__ __forceinline__ void scan_classify_block(uint32_t tb, const Seg* __restrict__ segs, uint32_t nseg,
                                                    uint32_t* __restrict__ hist, uint32_t* __restrict__ csum,
                                                    Seg* __restrict__ next, SmallSeg* __restrict__ small,
                                                    const SmallLists& L, uint32_t large_min, uint32_t* __restrict__ ctr,
                                                    uint32_t* __restrict__ next_large, uint32_t* __restrict__ lcp,
                                                    uint32_t key_buf, const DigitMap* __restrict__ dm) {
    __shared__ uint32_t sc_all[kScGroups][256];
    __shared__ uint32_t pv_all[kScGroups][256];
    __shared__ uint32_t first_all[kScGroups];
    __shared__ uint32_t lcnt[kClasses + 1], lelem[kClasses], lbase[kClasses + 1];
    const uint32_t grp = threadIdx.x >> 8;
    const uint32_t d = threadIdx.x & 255;
    const uint32_t si = tb * kScGroups + grp;
    const bool live = si < nseg;  // group-uniform
    uint32_t* const sc = sc_all[grp];
    Seg S{};
    uint32_t kd = 0, dep = 0, tot = 0, nrow = 0;
    uint32_t* col = nullptr;
    if (live) {
        S = segs[si];
        seg_depths(S, kd, dep);
        const uint32_t nt = seg_tiles(S.cnt);
        const uint32_t nc = seg_chunks(nt);
        // digit totals over the tiles, or over the chunk sums of a long segment
        col = nc ? csum + (uint64_t)S.chunk0 * 256 : hist + (uint64_t)S.tile0 * 256;
        nrow = nc ? nc : nt;
        for (uint32_t t = 0; t < nrow; ++t) tot += col[(uint64_t)t * 256 + d];
    }
    sc[d] = tot;
    if (d == 0) first_all[grp] = 256;
    if (threadIdx.x <= kClasses) lcnt[threadIdx.x] = 0;
    if (threadIdx.x < kClasses) lelem[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t o = 1; o < 256; o <<= 1) {
        uint32_t x = d >= o ? sc[d - o] : 0;
        __syncthreads();
        sc[d] += x;
        __syncthreads();
    }
    const uint32_t base = sc[d] - tot;
    if (tot) atomicMin(&first_all[grp], d);
    // exclusive offsets per tile (k_chunk_apply expands chunk offsets)
    uint32_t run = S.begin + base;
    for (uint32_t t = 0; t < nrow; ++t) {
        uint32_t* hp = &col[(uint64_t)t * 256 + d];
        const uint32_t x = *hp;
        *hp = run;
        run += x;
    }
    __syncthreads();
    // digits of kc characters: a bucket's first string shares dep characters
    // and the leading characters of the two digits with the last string of
    // the previous non-empty bucket (found by a max-scan)
    const DigitCtx D0{dm ? dm->kc : 1u, dm ? dm->sigma : 256u, nullptr};
    const DigitCtx D = seg_digits(D0, dep - kd);
    uint32_t prev = 0;       // the previous non-empty digit + 1, 0 if none
    if (dm && D0.kc > 1) {   // block-uniform
        uint32_t* const pv = pv_all[grp];
        pv[d] = tot ? d + 1 : 0;
        __syncthreads();
        for (uint32_t o = 1; o < 256; o <<= 1) {
            const uint32_t x = d >= o ? pv[d - o] : 0;
            __syncthreads();
            pv[d] = max(pv[d], x);
            __syncthreads();
        }
        prev = d ? pv[d - 1] : 0u;
    }
    const uint32_t b = S.begin + base;
    if (tot && d != first_all[grp])  // boundary between non-empty buckets
        lcp[b] = dep + (prev ? digit_lead(prev - 1, d, D) : 0u);
    const uint32_t nd = dep + digit_chars(d, D);  // the sub-bucket's depth (d != 0)
    // d == 0: equal strings; tot == 1: a single string -- both done
    int list = -1;  // small class c, or kClasses for the large list
    uint32_t slot = 0;
    if (tot > 1 && d != 0) {
        list = tot <= large_min ? (int)(32u - __builtin_clz(tot - 1u)) : (int)kClasses;
        slot = atomicAdd(&lcnt[list], 1u);
        if (list < (int)kClasses) atomicAdd(&lelem[list], tot);
    }
    __syncthreads();
    if (threadIdx.x < kClasses && lcnt[threadIdx.x]) {
        lbase[threadIdx.x] = atomicAdd(&ctr[kCtrSmall + threadIdx.x], lcnt[threadIdx.x]);
        atomicAdd(&ctr[kCtrSmallElems + threadIdx.x], lelem[threadIdx.x]);
    }
    if (threadIdx.x == kClasses && lcnt[kClasses]) lbase[kClasses] = atomicAdd(next_large, lcnt[kClasses]);
    __syncthreads();
    if (list < 0) return;
    if (list < (int)kClasses) {
        // the scatter of this pass leaves the bucket's keys (depth kd) in key_buf
        const uint32_t sh = nd - kd;
        small[L.off[list] + lbase[list] + slot] = SmallSeg{b, tot, nd, sh < 8 ? (sh | (key_buf << 3)) : 0u};
    } else {
        // keys run out after 8 characters; a bucket that holds the whole
        // segment (one digit shared by all) also re-runs the prefix skip,
        // which jumps over the rest of the shared prefix at once
        const uint32_t nkd = (nd - kd >= 8 || tot == S.cnt) ? kNone : kd;
        next[lbase[kClasses] + slot] = Seg{b, tot, nd, nkd, kNone, 0, 0};
    }
}

// kDev: the segment count comes from the pass counters hp; every block loops
// over groups of kScGroups segments.  kMulti: digits from the digit map dm.
template <bool kDev, bool kMulti>
__global__ void __launch_bounds__(kThreads * kScGroups)
k_scan_classify(const Seg* __restrict__ segs, uint32_t nseg, uint32_t* __restrict__ hist, uint32_t* __restrict__ csum,
                Seg* __restrict__ next, SmallSeg* __restrict__ small, SmallLists L, uint32_t large_min,
                uint32_t* __restrict__ ctr, uint32_t* __restrict__ next_large, uint32_t* __restrict__ lcp,
                uint32_t key_buf, const DigitMap* __restrict__ dm, const uint32_t* __restrict__ hp) {
    if constexpr (!kDev) {
        scan_classify_block(blockIdx.x, segs, nseg, hist, csum, next, small, L, large_min, ctr, next_large, lcp,
                            key_buf, kMulti ? dm : nullptr);
    } else {
        nseg = hp[kCtrLarge];
        const uint32_t nb = (nseg + kScGroups - 1) / kScGroups;
        for (uint32_t tb = blockIdx.x; tb < nb; tb += gridDim.x) {
            scan_classify_block(tb, segs, nseg, hist, csum, next, small, L, large_min, ctr, next_large, lcp, key_buf,
                                kMulti ? dm : nullptr);
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------ scatter
// Stable scatter of one tile by its digit.  The tile is first put in digit
// order in LDS (wave ballots and per-wave digit counts rank equal digits, in
// index order) and then written run by run, so consecutive lanes store
// consecutive addresses of a digit's run instead of a few records of every
// run per store.  512 threads x 4 records per tile, 30 KB of LDS per 8 waves;
// four block barriers per tile.
constexpr uint32_t kScT = 512;           // scatter threads per tile
constexpr uint32_t kScI = kTile / kScT;  // records per thread
static_assert(kScI * kScT == kTile && kScT >= 256 && kScT % 64 == 0, "scatter tile shape");

__device__ __forceinline__ void scatter_tile(uint32_t tb, const Tile* __restrict__ tiles, const Seg* __restrict__ segs,
                                             const uint32_t* __restrict__ hist, const uint64_t* __restrict__ key_in,
                                             const uint32_t* __restrict__ pb_in, uint64_t* __restrict__ key_out,
                                             uint32_t* __restrict__ pb_out, uint32_t* __restrict__ perm,
                                             uint32_t* __restrict__ lcp, const DigitCtx& D) {
    constexpr uint32_t kW = kScT / 64;         // waves per tile
    constexpr uint32_t kWR = kScI * 64;        // records per wave (contiguous)
    __shared__ uint64_t skey[kTile];
    __shared__ uint32_t sid[kTile];
    __shared__ uint32_t gbase[256];   // the digit's output offset for this tile
    __shared__ uint32_t lstart[256];  // the digit's first slot in the reordered tile
    __shared__ uint16_t wcnt[kW][256];  // a wave's count of a digit, then the waves before it
    __shared__ uint32_t wtot[256 / 64];
    const Tile T = tiles[tb];
    const Seg S = segs[T.seg];
    uint32_t kd, dep;
    seg_depths(S, kd, dep);
    const uint32_t pos = dep - kd;  // the digit's first byte in the keys
    const DigitCtx Ds = seg_digits(D, pos);
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const bool dig = tid < 256;  // the threads that own a digit
    if (dig) gbase[tid] = hist[(uint64_t)tb * 256 + tid];
    for (uint32_t k = tid; k < kW * 256; k += kScT) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    // wave w holds records [w kWR, (w + 1) kWR) of the tile, kScI per lane in
    // (round, lane) order; a record's rank among the wave's earlier records
    // of its digit comes from ballots (lanes of the round) and the wave's own
    // running digit counts (earlier rounds): no block barrier between rounds
    const uint64_t lt = lanemask_lt();
    uint64_t kv[kScI];
    uint32_t idv[kScI], dg[kScI], off[kScI];
#pragma unroll
    for (uint32_t r = 0; r < kScI; ++r) {
        const uint32_t i = w * kWR + r * 64 + lane;
        const uint32_t e = T.begin + i;
        const bool valid = e < T.end;
        kv[r] = valid ? key_in[e] : 0ull;
        idv[r] = valid ? (pb_in ? pb_in[e] : e) : 0u;  // first pass: ids are positions
        dg[r] = digit_of(kv[r], pos, Ds);
    }
#pragma unroll
    for (uint32_t r = 0; r < kScI; ++r) {
        const bool valid = T.begin + w * kWR + r * 64 + lane < T.end;
        uint64_t m = __ballot(valid);
        for (uint32_t b = 0; b < 8; ++b) {
            const bool bit = (dg[r] >> b) & 1;
            const uint64_t bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        // every lane reads its digit's count before the digit's first lane
        // adds the round's (one wave: its LDS accesses keep program order)
        off[r] = valid ? wcnt[w][dg[r]] + rank : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) wcnt[w][dg[r]] = (uint16_t)(wcnt[w][dg[r]] + __popcll(m));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: the waves' counts become their exclusive prefix, and the
    // digits' totals an exclusive scan (wave shuffles, then the 4 wave sums)
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
    if (dig && lane == 63) wtot[w] = inc;
    __syncthreads();
    if (dig) {
        uint32_t pre = 0;
        for (uint32_t k = 0; k < w; ++k) pre += wtot[k];
        lstart[tid] = pre + inc - tot;
    }
    __syncthreads();
    // stable placement in LDS, in index order
#pragma unroll
    for (uint32_t r = 0; r < kScI; ++r) {
        if (T.begin + w * kWR + r * 64 + lane < T.end) {
            const uint32_t q = lstart[dg[r]] + wcnt[w][dg[r]] + off[r];
            skey[q] = kv[r];
            sid[q] = idv[r];
        }
    }
    __syncthreads();
    // runs out: slot k of digit d goes to gbase[d] + (k - lstart[d])
    const uint32_t cnt = T.end - T.begin;
    for (uint32_t k = tid; k < cnt; k += kScT) {
        const uint64_t x = skey[k];
        const uint32_t d = digit_of(x, pos, Ds);
        const uint32_t q = gbase[d] + (k - lstart[d]);
        const uint32_t id = sid[k];
        key_out[q] = x;
        pb_out[q] = id;
        perm[q] = id;
        if (d == 0 && q != S.begin) lcp[q] = dep;  // equal strings of length dep
    }
}

// kMulti: digits from the digit map dm (else one byte each).
template <bool kDev, bool kMulti>
__global__ void __launch_bounds__(kScT)
k_scatter(const Tile* __restrict__ tiles, const Seg* __restrict__ segs, const uint32_t* __restrict__ hist,
          const uint64_t* __restrict__ key_in, const uint32_t* __restrict__ pb_in,
          uint64_t* __restrict__ key_out, uint32_t* __restrict__ pb_out, uint32_t* __restrict__ perm,
          uint32_t* __restrict__ lcp, const DigitMap* __restrict__ dm, const uint32_t* __restrict__ hp) {
    __shared__ uint8_t rk[256];
    const DigitCtx D = digit_ctx<kMulti>(dm, rk);  // scatter_tile synchronises before its first use
    if constexpr (!kDev) {
        scatter_tile(blockIdx.x, tiles, segs, hist, key_in, pb_in, key_out, pb_out, perm, lcp, D);
    } else {
        const uint32_t nb = hp[kCtrTiles];
        for (uint32_t tb = blockIdx.x; tb < nb; tb += gridDim.x) {
            scatter_tile(tb, tiles, segs, hist, key_in, pb_in, key_out, pb_out, perm, lcp, D);
            __syncthreads();  // LDS of this tile read before the next tile's
        }
    }
}

// ------------------------------------------------------------ small sort
// Block-wide exclusive max-scan helper over kThreads threads.
__device__ inline uint32_t block_excl_max(uint32_t v, uint32_t* tmp) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x = max(x, y);
    }
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t k = 0; k < w; ++k) pre = max(pre, tmp[k]);
    uint32_t ex = __shfl_up(x, 1);
    if (lane == 0) ex = 0;
    __syncthreads();
    return max(pre, ex);
}

constexpr uint32_t kCap = kSmallMax;
// large_min <= kSmallMax keeps every small segment in classes 1..kClasses-1,
// and a block holds at least one segment of each of them
static_assert((kCap >> (kClasses - 1)) == 1, "the largest class fills one block");
constexpr uint32_t kPer = kCap / kThreads;  // elements per thread
constexpr uint16_t kNoSlot = 0xFFFF;        // empty slot position
static_assert(kCap < kNoSlot, "slot positions are 16-bit");
constexpr uint32_t kRankMax = 32;           // rank sort when no active group is larger

__device__ __forceinline__ void small_sort_block(uint32_t tb, const SmallSeg* __restrict__ segs, uint32_t nseg,
                                                 uint32_t cls_log2, uint32_t* __restrict__ perm,
                                                 uint32_t* __restrict__ lcp, const uint8_t* __restrict__ chars,
                                                 const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                                                 uint64_t* __restrict__ out_off, uint32_t* __restrict__ out_len) {
    __shared__ uint64_t K[kCap];
    __shared__ uint64_t OFF[kCap];  // offset and length of the string loaded at
    __shared__ uint32_t LEN[kCap];  // slot position e (gathered once per block)
    __shared__ uint16_t G[kCap];
    __shared__ uint16_t I[kCap];   // first slot position of the string now at e
    __shared__ uint32_t LO[kCap];  // lcp per slot position
    __shared__ uint32_t RD[kCap];  // group depth, indexed by group start
    __shared__ uint32_t RC[kCap];  // group common prefix / scratch
    __shared__ uint8_t ACT[kCap];  // group active, indexed by group start
    __shared__ uint32_t tmp[kThreads / 64];
    __shared__ uint32_t sMaxG;  // largest group activated by the round's step 4
    const uint32_t CLS = 1u << cls_log2;
    const uint32_t spb = kCap >> cls_log2;
    const uint32_t s0 = tb * spb;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) sMaxG = 0;

    for (uint32_t e = tid; e < kCap; e += kThreads) {
        const uint32_t slot = e >> cls_log2, j = e & (CLS - 1);
        const uint32_t sbase = slot << cls_log2;
        LO[e] = 0;
        ACT[e] = 0;
        K[e] = ~0ull;
        I[e] = kNoSlot;
        G[e] = (uint16_t)sbase;
        if (s0 + slot < nseg) {
            const SmallSeg sg = segs[s0 + slot];
            if (j < sg.cnt) {
                const uint32_t id = perm[sg.begin + j];
                OFF[e] = off[id];
                LEN[e] = len[id];
                I[e] = (uint16_t)e;
                K[e] = 0;
            } else {
                G[e] = (uint16_t)(sbase + sg.cnt);
            }
            if (j == 0) {
                RD[e] = sg.depth;
                ACT[e] = sg.cnt >= 2;
            }
        }
    }
    __syncthreads();

    uint32_t maxg = kNone;  // largest active group of the round
    bool first = true;
    for (;;) {
        // 1. prefix skip inside every active group
        for (uint32_t e = tid; e < kCap; e += kThreads)
            if (I[e] != kNoSlot && ACT[G[e]] && e == G[e]) RC[e] = kNone;
        __syncthreads();
        if (tid == 0) sMaxG = 0;  // every thread read it before the barrier
        // every member's LCP with its group's first member, from the group
        // depth: one 32-character step per member in its lane, then the
        // walks still open are served one at a time by the whole wave (512
        // characters per step), so a lane whose strings share hundreds of
        // characters does not hold the other lanes through a long loop
        {
            uint64_t wa[kPer], wb[kPer];
            uint32_t wal[kPer], wbl[kPer], wfrom[kPer], wg[kPer];
            bool open[kPer];
#pragma unroll
            for (uint32_t q = 0; q < kPer; ++q) {
                const uint32_t e = tid + q * kThreads;
                open[q] = false;
                const uint32_t sl = I[e];
                if (sl == kNoSlot) continue;
                const uint32_t g = G[e];
                if (!ACT[g] || e == g) continue;  // groups have >= 2 members
                const uint32_t r = I[g];
                const uint32_t d = RD[g];
                wa[q] = OFF[sl];
                wal[q] = LEN[sl];
                wb[q] = OFF[r];
                wbl[q] = LEN[r];
                wg[q] = g;
                uint64_t x[4], y[4];
                load_keys4(chars, wa[q], wal[q], d, x);
                load_keys4(chars, wb[q], wbl[q], d, y);
                uint32_t l = kNone;
#pragma unroll
                for (int j = 3; j >= 0; --j) {  // the first difference or end wins
                    if (x[j] != y[j]) l = 8 * j + key_lcp(x[j], y[j]);
                    else if ((x[j] & 0xFF) == 0) l = 8 * j + key_strlen(x[j]);
                }
                if (l != kNone) {
                    atomicMin(&RC[g], l);
                } else {
                    open[q] = true;
                    wfrom[q] = d + 32;
                }
            }
            const uint32_t lane = lane_id();
#pragma unroll
            for (uint32_t q = 0; q < kPer; ++q) {
                for (uint64_t m = __ballot(open[q]); m; m &= m - 1) {  // wave-uniform
                    const int ld = __ffsll((unsigned long long)m) - 1;
                    const uint64_t ao = __shfl((unsigned long long)wa[q], ld), bo = __shfl((unsigned long long)wb[q], ld);
                    const uint32_t al = (uint32_t)__shfl((int)wal[q], ld), bl = (uint32_t)__shfl((int)wbl[q], ld);
                    const uint32_t from = (uint32_t)__shfl((int)wfrom[q], ld);
                    const uint32_t l = wave_lcp(chars, ao, al, bo, bl, from);
                    if (lane == (uint32_t)ld) atomicMin(&RC[wg[q]], l - (from - 32));
                }
            }
        }
        __syncthreads();
        for (uint32_t e = tid; e < kCap; e += kThreads)
            if (I[e] != kNoSlot && e == G[e] && ACT[e]) RD[e] += RC[e];
        __syncthreads();
        // 2. keys of active groups
        for (uint32_t e = tid; e < kCap; e += kThreads) {
            const uint32_t sl = I[e];
            if (sl != kNoSlot && ACT[G[e]]) K[e] = load_key8(chars, OFF[sl], LEN[sl], RD[G[e]]);
        }
        __syncthreads();
        // 3. sort of (G, K, I) inside every aligned CLS slot.  The first slot
        // position makes the order total, so equal strings keep their order
        // on entry, however many further rounds a block runs for other
        // slots' groups.  The MSD passes are stable from the identity
        // permutation, so that entry order of equal strings is ascending id
        // order.
        if (maxg <= kRankMax) {
            // every active group is small: each member counts the members
            // before it and moves there (inactive groups stay in place)
            uint64_t nk[kPer];
            uint32_t npos[kPer], ni[kPer];
#pragma unroll
            for (uint32_t q = 0; q < kPer; ++q) {
                const uint32_t e = tid + q * kThreads;
                npos[q] = kNone;
                const uint32_t sl = I[e];
                const uint32_t g = G[e];
                if (sl == kNoSlot || !ACT[g]) continue;
                const uint64_t k = K[e];
                uint32_t rank = 0;
                for (uint32_t m = g; m < kCap && G[m] == g; ++m) {
                    const uint64_t km = K[m];
                    rank += (km < k || (km == k && I[m] < sl));
                }
                npos[q] = g + rank;
                nk[q] = k;
                ni[q] = sl;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t q = 0; q < kPer; ++q) {
                if (npos[q] == kNone) continue;
                K[npos[q]] = nk[q];
                I[npos[q]] = (uint16_t)ni[q];
            }
            __syncthreads();
        } else if (first) {
            // one group per slot: empty positions (key ~0, slot kNoSlot) end
            // every slot, so (K, I) alone orders it
            for (uint32_t k = 2; k <= CLS; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t t = tid; t < kCap / 2; t += kThreads) {
                        const uint32_t i = 2 * t - (t & (j - 1));
                        const uint32_t q = i + j;
                        const bool asc = (k == CLS) || ((i & k) == 0);
                        const uint64_t ki = K[i], kq = K[q];
                        const uint32_t ii = I[i], iq = I[q];
                        const bool gt = ki > kq || (ki == kq && ii > iq);
                        if (asc == gt && (ki != kq || ii != iq)) {
                            K[i] = kq;
                            K[q] = ki;
                            I[i] = (uint16_t)iq;
                            I[q] = (uint16_t)ii;
                        }
                    }
                    __syncthreads();
                }
            }
        } else {
            for (uint32_t k = 2; k <= CLS; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t t = tid; t < kCap / 2; t += kThreads) {
                        const uint32_t i = 2 * t - (t & (j - 1));
                        const uint32_t q = i + j;
                        const bool asc = (k == CLS) || ((i & k) == 0);
                        const uint32_t gi = G[i], gq = G[q];
                        const uint64_t ki = K[i], kq = K[q];
                        const uint32_t ii = I[i], iq = I[q];
                        const bool gt = gi > gq || (gi == gq && (ki > kq || (ki == kq && ii > iq)));
                        const bool lt = gi < gq || (gi == gq && (ki < kq || (ki == kq && ii < iq)));
                        if (asc ? gt : lt) {
                            G[i] = (uint16_t)gq;
                            G[q] = (uint16_t)gi;
                            K[i] = kq;
                            K[q] = ki;
                            I[i] = (uint16_t)iq;
                            I[q] = (uint16_t)ii;
                        }
                    }
                    __syncthreads();
                }
            }
        }
        // 4. boundaries, LCPs and new groups (thread owns kPer consecutive slots)
        uint32_t flag[kPer];
        uint32_t oldg[kPer];
        uint32_t last = 0;
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t e = tid * kPer + q;
            const uint32_t g = G[e];
            oldg[q] = g;
            uint32_t f = (e == g);
            if (!f && I[e] != kNoSlot && ACT[g]) {
                const uint64_t a = K[e - 1], b = K[e];
                const uint32_t d = RD[g];
                if (a != b) {
                    f = 1;
                    LO[e] = d + key_lcp(a, b);
                } else if ((b & 0xFF) == 0) {
                    LO[e] = d + key_strlen(b);
                }
            }
            flag[q] = f;
            if (f) last = e;
        }
        uint32_t run = block_excl_max(last, tmp);
        uint32_t newg[kPer];
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t e = tid * kPer + q;
            if (flag[q]) run = e;
            newg[q] = run;
        }
        // old group depth per element (read before any write)
        uint32_t od[kPer];
        uint8_t oact[kPer];
        for (uint32_t q = 0; q < kPer; ++q) {
            od[q] = RD[oldg[q]];
            oact[q] = ACT[oldg[q]];
        }
        __syncthreads();
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t e = tid * kPer + q;
            G[e] = (uint16_t)newg[q];
            if (flag[q]) {
                ACT[e] = 0;
                if (oact[q]) RD[e] = od[q] + 8;
            }
        }
        __syncthreads();
        // a new group stays active if it came from an active group, has a
        // second member and its key did not reach the end of the strings
        // (the condition holds for all members but the first of such a
        // group, so the last member's position gives the group size)
        int any = 0;
        uint32_t gmax = 0;
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t e = tid * kPer + q;
            if (oact[q] && !flag[q] && I[e] != kNoSlot && (K[e] & 0xFF) != 0) {
                ACT[newg[q]] = 1;
                any = 1;
                gmax = max(gmax, e - newg[q] + 1);
            }
        }
        if (gmax) atomicMax(&sMaxG, gmax);
        if (!__syncthreads_or(any)) break;
        maxg = sMaxG;
        first = false;
    }
    // write back: the ids are read from the entry permutation at their first
    // slot positions (same segment) before any position is overwritten
    uint32_t wid[kPer], wpos[kPer];
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t e = tid + q * kThreads;
        const uint32_t slot = e >> cls_log2, j = e & (CLS - 1);
        wid[q] = kNone;
        if (s0 + slot >= nseg) continue;
        const SmallSeg sg = segs[s0 + slot];
        if (j >= sg.cnt) continue;
        wpos[q] = sg.begin + j;
        wid[q] = perm[sg.begin + (I[e] & (CLS - 1))];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        if (wid[q] == kNone) continue;
        const uint32_t e = tid + q * kThreads;
        const uint32_t sl = I[e];
        perm[wpos[q]] = wid[q];
        out_off[wpos[q]] = OFF[sl];
        out_len[wpos[q]] = LEN[sl];
        if (e & (CLS - 1)) lcp[wpos[q]] = LO[e];
    }
}

// kDev: the class's segment count comes from dcnt; every block loops over
// groups of kCap >> cls_log2 segments.
template <bool kDev>
__global__ void __launch_bounds__(kThreads)
k_small_sort(const SmallSeg* __restrict__ segs, uint32_t nseg, uint32_t cls_log2,
             uint32_t* __restrict__ perm, uint32_t* __restrict__ lcp, const uint8_t* __restrict__ chars,
             const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
             uint64_t* __restrict__ out_off, uint32_t* __restrict__ out_len, const uint32_t* __restrict__ dcnt) {
    if constexpr (!kDev) {
        small_sort_block(blockIdx.x, segs, nseg, cls_log2, perm, lcp, chars, off, len, out_off, out_len);
    } else {
        nseg = *dcnt;
        const uint32_t spb = kCap >> cls_log2;
        const uint32_t nb = (nseg + spb - 1) / spb;
        for (uint32_t tb = blockIdx.x; tb < nb; tb += gridDim.x) {
            small_sort_block(tb, segs, nseg, cls_log2, perm, lcp, chars, off, len, out_off, out_len);
            __syncthreads();
        }
    }
}

// True if a key whose first w characters are valid reaches the end of its
// string within them (the key is zero after a string's end).
__device__ inline bool key_ended(uint64_t k, uint32_t w) { return ((k >> (64 - 8 * w)) & 0xFF) == 0; }

// Small segments of at most 64 strings, one wave per 64 slots (64/CLS
// segments per wave, one string per lane).  Each lane keeps its string's
// 8-char key at its group's depth, the group leader lane G and the lane the
// string came from in registers; id, offset and length stay in the lane that
// loaded them and are fetched across lanes when needed.  An iteration sorts
// (G, key, lane) with a register bitonic network inside each aligned
// CLS-lane slot, splits groups where keys differ (LCP = depth + common key
// bytes), and lets every group whose members still share the key continue
// deeper; groups of equal strings (key ended) stop.  A segment whose keys
// the last MSD pass left in a key buffer starts from those (fewer than 8
// valid characters, no string access); later keys are loaded 8 characters
// at a time.  A round in which no active group has two different keys only
// advances the depths.  Group extents come from ballots of the leader lanes;
// every loop runs while a ballot over the wave is non-zero, so all lanes
// take part in every shuffle.
// Offset and length of every string in one 16-byte record, so that the base
// case's random lookup of a string reads one line instead of two.
struct alignas(16) OffLen {
    uint64_t off;
    uint64_t len;
};

__global__ void k_pack_offlen(const uint64_t* __restrict__ off, const uint32_t* __restrict__ len, uint64_t n,
                              OffLen* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = OffLen{off[i], len[i]};
}

// The value of lane (lane ^ kMask): masks 1 and 2 as DPP quad permutations,
// 4 to 16 as LDS swizzles (no address operand), 32 by a permute.
template <uint32_t kMask>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    if constexpr (kMask == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    else if constexpr (kMask == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    else if constexpr (kMask < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (kMask << 10));
    else return (uint32_t)__shfl_xor((int)v, (int)kMask);
}
template <uint32_t kMask>
__device__ __forceinline__ uint64_t xor_lane64(uint64_t v) {
    return (uint64_t)xor_lane<kMask>((uint32_t)v) | ((uint64_t)xor_lane<kMask>((uint32_t)(v >> 32)) << 32);
}

// One compare-exchange step of the network: partner lane ^ kJ, blocks of kK.
template <uint32_t kCls, uint32_t kK, uint32_t kJ, bool kOneGroup>
__device__ __forceinline__ void slot_step(uint32_t lane, uint32_t& G, uint64_t& K, uint32_t& src) {
    const uint64_t pK = xor_lane64<kJ>(K);
    const uint32_t pS = xor_lane<kJ>(src);
    const uint32_t pG = kOneGroup ? G : xor_lane<kJ>(G);
    const bool asc = (kK == kCls) || ((lane & kK) == 0);
    const bool lower = (lane & kJ) == 0;
    const bool pless = kOneGroup ? (pK < K || (pK == K && pS < src))
                                 : (pG < G || (pG == G && (pK < K || (pK == K && pS < src))));
    // the lower lane keeps the smaller element when ascending
    if (lower == asc ? pless : !pless) {
        G = pG;
        K = pK;
        src = pS;
    }
}

template <uint32_t kCls, uint32_t kK, uint32_t kJ, bool kOneGroup>
__device__ __forceinline__ void slot_merge(uint32_t lane, uint32_t& G, uint64_t& K, uint32_t& src) {
    slot_step<kCls, kK, kJ, kOneGroup>(lane, G, K, src);
    if constexpr (kJ > 1) slot_merge<kCls, kK, kJ / 2, kOneGroup>(lane, G, K, src);
}

template <uint32_t kCls, uint32_t kK, bool kOneGroup>
__device__ __forceinline__ void slot_sort(uint32_t lane, uint32_t& G, uint64_t& K, uint32_t& src) {
    slot_merge<kCls, kK, kK / 2, kOneGroup>(lane, G, K, src);
    if constexpr (kK < kCls) slot_sort<kCls, kK * 2, kOneGroup>(lane, G, K, src);
}

// Bitonic sort of (G, K, src) inside every aligned kCls-lane slot of the
// wave, the network unrolled (constant lane masks).  With one_group every
// slot is one group of strings followed by padding lanes (key ~0, larger
// source lanes), so (K, src) decides and G stays in place.
template <uint32_t kCls>
__device__ __forceinline__ void slot_network(uint32_t lane, bool one_group, uint32_t& G, uint64_t& K, uint32_t& src) {
    if (one_group) slot_sort<kCls, 2, true>(lane, G, K, src);
    else slot_sort<kCls, 2, false>(lane, G, K, src);
}

__device__ __forceinline__ void small_sort_wave_unit(uint64_t wave, const SmallSeg* __restrict__ segs, uint32_t nseg,
                                                     uint32_t cls_log2, uint32_t* __restrict__ perm,
                                                     uint32_t* __restrict__ lcp, const uint8_t* __restrict__ chars,
                                                     const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                                                     const OffLen* __restrict__ offlen,
                                                     const uint64_t* __restrict__ keyA,
                                                     const uint64_t* __restrict__ keyB,
                                                     uint64_t* __restrict__ out_off, uint32_t* __restrict__ out_len) {
    const uint32_t lane = lane_id();
    const uint32_t CLS = 1u << cls_log2;
    const uint32_t spw = 64u >> cls_log2;
    const uint32_t slot_start = lane & ~(CLS - 1), j = lane & (CLS - 1);
    const uint64_t sidx = wave * spw + (lane >> cls_log2);
    bool valid = false;
    uint32_t sbegin = 0, depth0 = 0, scnt = 0, cache = 0;
    uint32_t my_id = 0, my_len = 0;
    uint64_t my_off = 0, K0 = 0;
    if (sidx < nseg) {
        const SmallSeg sg = segs[sidx];
        sbegin = sg.begin;
        scnt = sg.cnt;
        depth0 = sg.depth;
        cache = sg.cache;
        if (j < sg.cnt) {
            valid = true;
            my_id = perm[sg.begin + j];
            if (cache) K0 = ((cache & 8) ? keyB : keyA)[sg.begin + j] << (8 * (cache & 7));
            if (offlen) {
                const OffLen ol = offlen[my_id];
                my_off = ol.off;
                my_len = (uint32_t)ol.len;
            } else {
                my_off = off[my_id];
                my_len = len[my_id];
            }
        }
    }
    // per element (moves with the sort): G, K, src (original lane)
    uint32_t G = valid ? slot_start : lane;  // padding lanes: own inactive group, after the strings
    uint64_t K = valid ? 0 : ~0ull;
    uint32_t src = lane;
    // per position: LCP with the predecessor; per leader lane: depth, width
    // of the group's next key, active
    uint32_t lo = 0;
    uint32_t rd = depth0;
    uint32_t rw = cache ? 8 - (cache & 7) : 8;
    bool act = valid && j == 0 && scnt >= 2;
    // until the first network has run every slot is one group of its strings
    // (padding lanes after them, key ~0 and a larger source lane), so the
    // first network orders (key, source lane) and leaves G in place
    bool one_group = true;
    for (;;) {
        const uint64_t leaders = __ballot(lane == G);
        const uint32_t grd = __shfl(rd, G);
        const uint32_t gw = __shfl(rw, G);
        const bool gact = __shfl((int)act, G) != 0;
        if (__ballot(gact) == 0) break;
        if (__ballot(gact && gw == 8)) {  // some group reads its next key from the strings
            const uint64_t so = __shfl((unsigned long long)my_off, src);
            const uint32_t sl = __shfl(my_len, src);
            if (gact && gw == 8) K = load_key8(chars, so, sl, grd);
        }
        // a cached key is a segment's first (nothing is sorted yet: src == lane)
        if (gact && gw < 8) K = K0;
        // group extent [G, gend) from the leader mask
        const uint64_t above = G >= 63 ? 0ull : (leaders & (~0ull << (G + 1)));
        const uint32_t gend = above ? (uint32_t)__builtin_ctzll(above) : 64u;
        const uint64_t grp = (gend >= 64 ? ~0ull : ((1ull << gend) - 1)) & (~0ull << G);
        const uint64_t leadK = __shfl((unsigned long long)K, G);
        const bool differs = gact && K != leadK;
        const uint64_t dmask = __ballot(differs);
        if (__ballot(gact && (dmask & grp) != 0) == 0) {
            // no active group has two different keys: equal keys go deeper
            // unless they ended (then the group is a run of equal strings)
            const bool ended = key_ended(K, gw);
            if (gact && lane != G && ended) lo = grd + key_strlen(K);
            if (lane == G && gact) {
                if (ended) {
                    act = false;
                } else {
                    rd = grd + gw;
                    rw = 8;
                }
            }
            continue;
        }
        // bitonic sort of (G, K, src) inside every aligned CLS-lane slot
        switch (cls_log2) {
            case 1: slot_network<2>(lane, one_group, G, K, src); break;
            case 2: slot_network<4>(lane, one_group, G, K, src); break;
            case 3: slot_network<8>(lane, one_group, G, K, src); break;
            case 4: slot_network<16>(lane, one_group, G, K, src); break;
            case 5: slot_network<32>(lane, one_group, G, K, src); break;
            default: slot_network<64>(lane, one_group, G, K, src); break;
        }
        one_group = false;
        // boundaries and LCPs inside active groups
        const uint64_t pK = __shfl_up((unsigned long long)K, 1);
        bool flag = lane == G;
        if (!flag && gact) {
            if (pK != K) {
                flag = true;
                lo = grd + key_lcp(pK, K);
            } else if (key_ended(K, gw)) {
                lo = grd + key_strlen(K);
            }
        }
        const uint64_t fmask = __ballot(flag);
        const uint32_t newG = 63u - (uint32_t)__builtin_clzll(fmask & (lane == 63 ? ~0ull : ((2ull << lane) - 1)));
        if (flag && gact) {
            // a new group of an active group: deeper by the key's width,
            // active if it has a second member (padding lanes are always
            // leaders) and its key did not end
            const bool second = lane < 63 && !((fmask >> (lane + 1)) & 1);
            rd = grd + gw;
            rw = 8;
            act = second && !key_ended(K, gw);
        }
        G = newG;
    }
    const uint32_t fid = __shfl(my_id, src);
    const uint64_t fo = __shfl((unsigned long long)my_off, src);
    const uint32_t fl = __shfl(my_len, src);
    if (valid) {
        perm[sbegin + j] = fid;
        out_off[sbegin + j] = fo;
        out_len[sbegin + j] = fl;
        if (j) lcp[sbegin + j] = lo;
    }
}

// kDev: the class's segment count comes from dcnt; every wave loops over
// groups of 64 >> cls_log2 segments.
template <bool kDev>
__global__ void __launch_bounds__(kThreads)
k_small_sort_wave(const SmallSeg* __restrict__ segs, uint32_t nseg, uint32_t cls_log2,
                  uint32_t* __restrict__ perm, uint32_t* __restrict__ lcp, const uint8_t* __restrict__ chars,
                  const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                  const OffLen* __restrict__ offlen,
                  const uint64_t* __restrict__ keyA, const uint64_t* __restrict__ keyB,
                  uint64_t* __restrict__ out_off, uint32_t* __restrict__ out_len, const uint32_t* __restrict__ dcnt) {
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // wave-uniform
    if constexpr (!kDev) {
        small_sort_wave_unit(wave, segs, nseg, cls_log2, perm, lcp, chars, off, len, offlen, keyA, keyB, out_off,
                             out_len);
    } else {
        nseg = *dcnt;
        const uint32_t spw = 64u >> cls_log2;
        const uint64_t nw = (nseg + spw - 1) / spw;
        for (uint64_t w = wave; w < nw; w += ((uint64_t)gridDim.x * blockDim.x) >> 6)
            small_sort_wave_unit(w, segs, nseg, cls_log2, perm, lcp, chars, off, len, offlen, keyA, keyB, out_off,
                                 out_len);
    }
}

// Offset and length per sorted position: every later stage streams these
// instead of gathering through the permutation.
// With only_missing, positions the base cases already filled (s_len not
// kNone) are skipped.
template <bool kDev>
__global__ void k_sorted_meta(const uint32_t* __restrict__ perm, const uint64_t* __restrict__ off,
                              const uint32_t* __restrict__ len, uint64_t n, uint64_t* __restrict__ s_off,
                              uint32_t* __restrict__ s_len, int only_missing, const uint64_t* __restrict__ dn) {
    if (kDev) n = *dn;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (!only_missing || s_len[i] == kNone) {
            const uint32_t id = perm[i];
            s_off[i] = off[id];
            s_len[i] = len[id];
        }
    }
}

// Device-driven base cases, every class in one launch: the one-wave classes
// 1..wave_class_max share one kernel, the block classes above share another.
// A unit (a wave's or a block's group of segments) of class c is found from
// the per-class unit counts, read from the device counters.
__global__ void __launch_bounds__(kThreads)
k_small_sort_wave_all(const SmallSeg* __restrict__ small, SmallLists L, uint32_t cmax, uint32_t* __restrict__ perm,
                      uint32_t* __restrict__ lcp, const uint8_t* __restrict__ chars, const uint64_t* __restrict__ off,
                      const uint32_t* __restrict__ len, const uint64_t* __restrict__ keyA,
                      const uint64_t* __restrict__ keyB, uint64_t* __restrict__ out_off,
                      uint32_t* __restrict__ out_len, const uint32_t* __restrict__ ctr) {
    uint64_t ubeg[kWaveClassMax + 2];
    uint64_t tot = 0;
    for (uint32_t c = 1; c <= cmax; ++c) {
        ubeg[c] = tot;
        const uint32_t spw = 64u >> c;
        tot += (ctr[kCtrSmall + c] + spw - 1) / spw;
    }
    ubeg[cmax + 1] = tot;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t u = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < tot; u += nw) {  // wave-uniform
        uint32_t c = 1;
        while (u >= ubeg[c + 1]) ++c;
        small_sort_wave_unit(u - ubeg[c], small + L.off[c], ctr[kCtrSmall + c], c, perm, lcp, chars, off, len,
                             nullptr, keyA, keyB, out_off, out_len);
    }
}

__global__ void __launch_bounds__(kThreads)
k_small_sort_all(const SmallSeg* __restrict__ small, SmallLists L, uint32_t cmin, uint32_t* __restrict__ perm,
                 uint32_t* __restrict__ lcp, const uint8_t* __restrict__ chars, const uint64_t* __restrict__ off,
                 const uint32_t* __restrict__ len, uint64_t* __restrict__ out_off, uint32_t* __restrict__ out_len,
                 const uint32_t* __restrict__ ctr) {
    uint32_t ubeg[kClasses + 1];
    uint32_t tot = 0;
    for (uint32_t c = cmin; c < kClasses; ++c) {
        ubeg[c] = tot;
        const uint32_t spb = kCap >> c;
        tot += (ctr[kCtrSmall + c] + spb - 1) / spb;
    }
    ubeg[kClasses] = tot;
    for (uint32_t u = blockIdx.x; u < tot; u += gridDim.x) {  // block-uniform
        uint32_t c = cmin;
        while (u >= ubeg[c + 1]) ++c;
        small_sort_block(u - ubeg[c], small + L.off[c], ctr[kCtrSmall + c], c, perm, lcp, chars, off, len, out_off,
                         out_len);
        __syncthreads();
    }
}

// ------------------------------------------------------------------ driver
// Capacities of the sort's work lists for n strings over p PEs.
struct SortCaps {
    uint64_t seg_cap, tile_cap, chunk_cap;
    SmallLists L;
    uint64_t small_total;
};

static SortCaps sort_caps(uint64_t n, uint32_t p, const SortParams& T) {
    SortCaps c;
    c.seg_cap = n / (T.large_min + 1) + p + 16;
    c.tile_cap = n / kTile + c.seg_cap + 16;
    c.chunk_cap = c.tile_cap / kChunkTiles + 16;
    uint64_t tot = 0;
    for (uint32_t k = 0; k < kClasses; ++k) {
        const uint64_t cap = k == 0 ? 0 : n / ((1ull << (k - 1)) + 1) + 1;
        c.L.off[k] = tot;
        c.L.cap[k] = cap;
        tot += cap;
    }
    c.small_total = tot;
    return c;
}

// Grows every buffer the local sort of n strings over p PEs uses.
static SortCaps reserve_sort(dss_set_s* s, uint64_t n, uint32_t p) {
    const uint64_t nn = std::max<uint64_t>(n, 1);
    s->perm.reserve(nn * 4);
    s->lcp.reserve(nn * 4);
    s->s_off.reserve(nn * 8);
    s->s_len.reserve(nn * 4);
    SortScratch& X = s->ss;
    X.keyA.reserve(nn * 8);
    X.keyB.reserve(nn * 8);
    X.pbA.reserve(nn * 4);
    X.pbB.reserve(nn * 4);
    const SortCaps c = sort_caps(n, p, s->sort_params);
    X.segA.reserve(c.seg_cap * sizeof(Seg));
    X.segB.reserve(c.seg_cap * sizeof(Seg));
    X.tiles.reserve(c.tile_cap * sizeof(Tile));
    X.hist.reserve(c.tile_cap * 256 * 4);
    X.csum.reserve(c.chunk_cap * 256 * 4);
    X.small.reserve(std::max<uint64_t>(c.small_total, 1) * sizeof(SmallSeg));
    X.counters.reserve(64 * 4 + 16);  // 64 counters, then the device-driven form's string and PE counts
    X.digit.reserve(kDigitBytes);
    return c;
}

// The digit map of a host-driven sort, or nullptr for one character per
// digit.  digit_chars 0 (auto) scans sets of at least 2^22 strings whose
// arena holds at most 2 GiB (a sample of the first 1 MiB first: arenas with
// more than 15 byte values stop there), 2 or 3 scan any set and allow up to
// that many characters.  Batch sets over a job pool keep one character.
static const DigitMap* digit_map(dss_set_s* s) {
    const uint32_t want = s->sort_params.digit_chars;
    if (want == 1 || s->borrowed || s->nbytes == 0) return nullptr;
    if (want == 0 && (s->n < (1ull << 22) || s->nbytes > (1ull << 31))) return nullptr;
    const hipStream_t st = s->stream;
    SortScratch& X = s->ss;
    DigitMap* dm = X.digit.as<DigitMap>();
    uint32_t* bits = reinterpret_cast<uint32_t*>(X.digit.as<uint8_t>() + kDigitBitsAt);
    const uint64_t nb = s->nbytes, ns = std::min<uint64_t>(nb, kAlphaSample);
    KScope ks(s->timer, st, "alphabet", (double)nb);
    fill_small(st, bits, 0, 32);
    hipLaunchKernelGGL(k_alphabet, dim3(stride_grid(ns / 16 + 1, 256)), dim3(256), 0, st, s->arena(), (uint64_t)0, ns,
                       bits, 0);
    if (nb > ns)
        hipLaunchKernelGGL(k_alphabet, dim3(stride_grid((nb - ns) / 16 + 1, 256 * 16)), dim3(256), 0, st, s->arena(),
                           ns, nb, bits, 1);
    hipLaunchKernelGGL(k_digit_map, dim3(1), dim3(256), 0, st, bits, want ? want : 3u, dm);
    DSS_HIP(hipGetLastError());
    return dm;
}

// Pass list buffers: the current and next segment lists, key and id
// buffers, and which of the two per-pass counter sets is the current one.
struct PassBufs {
    Seg* cur;
    Seg* nxt;
    uint64_t* kc;
    uint64_t* kn;
    uint32_t* pc_ids;
    uint32_t* pn_ids;
    uint32_t par;
    const DigitMap* dm;  // characters per digit (nullptr: one)
    void swap_pass() {
        par ^= 1;
        std::swap(cur, nxt);
        std::swap(kc, kn);
        std::swap(pc_ids, pn_ids);
    }
};

static PassBufs first_pass_bufs(SortScratch& X) {
    return PassBufs{X.segA.as<Seg>(), X.segB.as<Seg>(), X.keyA.as<uint64_t>(), X.keyB.as<uint64_t>(),
                    X.pbA.as<uint32_t>(), X.pbB.as<uint32_t>(), 0, nullptr};
}

static const char* const kClassNames[kClasses] = {
    "small_sort_c0", "small_sort_c1", "small_sort_c2", "small_sort_c3", "small_sort_c4", "small_sort_c5",
    "small_sort_c6", "small_sort_c7", "small_sort_c8", "small_sort_c9", "small_sort_c10"};

// Host-driven MSD passes from the state in `b` until the segment list is
// empty: every pass sizes its launches from the counters read back after the
// previous one (h_ctr holds them on entry and on return).
// first: the passes start from the seed, where every id equals its position
// (the first pass neither copies nor reads the id buffer).
static void host_passes(dss_set_s* s, const SortCaps& caps, uint32_t* ctr, uint32_t* h_ctr, PassBufs& b,
                        bool trace, bool first) {
    const hipStream_t st = s->stream;
    SortScratch& X = s->ss;
    const SortParams& T = s->sort_params;
    const SmallLists L = caps.L;
    const uint8_t* chars = s->arena();
    const uint64_t* off = s->off.as<uint64_t>();
    const uint32_t* len = s->len.as<uint32_t>();
    uint32_t* perm = s->perm.as<uint32_t>();
    uint32_t* lcp = s->lcp.as<uint32_t>();
    for (;;) {
        uint32_t* pn = ctr + kCtrPass + 4 * (b.par ^ 1);  // the next pass
        const uint32_t* hp = h_ctr + kCtrPass + 4 * b.par;
        const uint32_t nlarge = hp[kCtrLarge];
        if (!nlarge) break;
        const uint32_t nt = hp[kCtrTiles];
        const uint32_t nchunks = hp[kCtrChunks];
        DSS_REQUIRE(nlarge <= caps.seg_cap, DSS_ERANGE, "segment list overflow");
        DSS_REQUIRE(nt <= caps.tile_cap, DSS_ERANGE, "tile list overflow");
        DSS_REQUIRE(nchunks <= caps.chunk_cap, DSS_ERANGE, "chunk list overflow");
        Seg* const cur = b.cur;
        Seg* const nxt = b.nxt;
        uint64_t* const kc = b.kc;
        uint64_t* const kn = b.kn;
        hipLaunchKernelGGL(k_fill_tiles<false>, dim3((nt + 255) / 256), dim3(256), 0, st, cur, nlarge, nt,
                           X.tiles.as<Tile>(), nullptr, pn);
        Tile* tl = X.tiles.as<Tile>();
        // elements in large segments of this pass (for the algorithmic byte counts)
        std::vector<Seg> hseg;
        uint64_t cnt_large = 0;
        if (s->timer.enabled) {
            // read in stream order: the set's stream does not block on the
            // null stream, and the first pass's segments may still be in flight
            hseg.resize(nlarge);
            DSS_HIP(hipMemcpyAsync(hseg.data(), cur, nlarge * sizeof(Seg), hipMemcpyDeviceToHost, st));
            DSS_HIP(hipStreamSynchronize(st));
            for (auto& q : hseg) cnt_large += q.cnt;
        }
        {
            KScope ks(s->timer, st, "prefix_skip", 0.0);
            decltype(&k_prefix_skip<8, 1, false>) kern = &k_prefix_skip<8, 1, false>;
            if (T.skip_lanes == 16)
                kern = T.skip_strings == 4 ? &k_prefix_skip<16, 4, false>
                       : T.skip_strings == 2 ? &k_prefix_skip<16, 2, false> : &k_prefix_skip<16, 1, false>;
            else
                kern = T.skip_strings == 4 ? &k_prefix_skip<8, 4, false>
                       : T.skip_strings == 2 ? &k_prefix_skip<8, 2, false> : &k_prefix_skip<8, 1, false>;
            hipLaunchKernelGGL(kern, dim3(nt), dim3(kThreads), 0, st, tl, cur, perm, chars, off, len, kc,
                               first ? nullptr : b.pc_ids, nullptr);
        }
        {
            KScope ks(s->timer, st, "gather", 0.0);  // bytes amended after the pass
            hipLaunchKernelGGL(k_gather<false>, dim3(nt), dim3(kThreads), 0, st, tl, cur, perm, chars, off, len, kc,
                               nullptr);
        }
        {
            KScope ks(s->timer, st, "hist", (double)cnt_large * 8.0);
            decltype(&k_hist<false, false>) kern = b.dm ? &k_hist<false, true> : &k_hist<false, false>;
            hipLaunchKernelGGL(kern, dim3(nt), dim3(kThreads), 0, st, tl, cur, kc, X.hist.as<uint32_t>(), b.dm,
                               nullptr);
        }
        {
            KScope ks(s->timer, st, "scan_classify", (double)nt * 256 * 4 * 3);
            if (nchunks)
                hipLaunchKernelGGL(k_chunk_sum<false>, dim3(nchunks), dim3(kThreads), 0, st, cur, nlarge,
                                   X.hist.as<uint32_t>(), X.csum.as<uint32_t>(), nullptr);
            decltype(&k_scan_classify<false, false>) kern =
                b.dm ? &k_scan_classify<false, true> : &k_scan_classify<false, false>;
            hipLaunchKernelGGL(kern, dim3((nlarge + kScGroups - 1) / kScGroups),
                               dim3(kThreads * kScGroups), 0, st, cur, nlarge, X.hist.as<uint32_t>(),
                               X.csum.as<uint32_t>(), nxt, X.small.as<SmallSeg>(), L, T.large_min, ctr,
                               pn + kCtrLarge, lcp, kn == X.keyB.as<uint64_t>() ? 1u : 0u, b.dm, nullptr);
            if (nchunks)
                hipLaunchKernelGGL(k_chunk_apply<false>, dim3(nchunks), dim3(kThreads), 0, st, cur, nlarge,
                                   X.hist.as<uint32_t>(), X.csum.as<uint32_t>(), nullptr);
        }
        {
            KScope ks(s->timer, st, "scatter", (double)cnt_large * 28.0);
            decltype(&k_scatter<false, false>) kern = b.dm ? &k_scatter<false, true> : &k_scatter<false, false>;
            hipLaunchKernelGGL(kern, dim3(nt), dim3(kScT), 0, st, tl, cur, X.hist.as<uint32_t>(), kc,
                               first ? nullptr : b.pc_ids, kn, b.pn_ids, perm, lcp, b.dm, nullptr);
        }
        hipLaunchKernelGGL(k_build_tiles, dim3(1), dim3(1024), 0, st, nxt, pn);
        DSS_HIP(hipGetLastError());
        s->pp.mbox.fetch(st, h_ctr, ctr, 64 * 4);
        if (s->timer.enabled) {
            // prefix skip: per string of a segment that gathers keys, its
            // characters up to the common prefix + 8 (the key), id, offset
            // and length reads, key and id writes; gather: the tile records
            // it reads, and per string of a tile whose keys are re-gathered,
            // id, offset, length, 8 chars and the key write
            std::vector<Seg> after(nlarge);
            DSS_HIP(hipMemcpy(after.data(), cur, nlarge * sizeof(Seg), hipMemcpyDeviceToHost));
            std::vector<Tile> ht(nt);
            DSS_HIP(hipMemcpy(ht.data(), tl, nt * sizeof(Tile), hipMemcpyDeviceToHost));
            double skip = 0, gath = (double)nt * sizeof(Tile);
            for (uint32_t k = 0; k < nlarge; ++k)
                if (hseg[k].kdepth == kNone) skip += (double)after[k].cnt * ((double)after[k].cp + 36.0);
            for (const Tile& q : ht)
                if (hseg[q.seg].kdepth == kNone && q.keyed != after[q.seg].cp) gath += (double)(q.end - q.begin) * 32.0;
            s->timer.amend_last("prefix_skip", skip);
            s->timer.amend_last("gather", gath);
        }
        if (trace) {
            std::vector<Seg> hs(std::min<uint32_t>(nlarge, 4));
            DSS_HIP(hipMemcpy(hs.data(), cur, hs.size() * sizeof(Seg), hipMemcpyDeviceToHost));
            fprintf(stderr, "[dss sort] pass nlarge=%u ntiles=%u next=%u", nlarge, nt,
                    h_ctr[kCtrPass + 4 * (b.par ^ 1) + kCtrLarge]);
            for (auto& q : hs)
                fprintf(stderr, " {b=%u c=%u d=%u kd=%d cp=%d}", q.begin, q.cnt, q.depth, (int)q.kdepth, (int)q.cp);
            fprintf(stderr, "\n");
        }
        b.swap_pass();
        first = false;
    }
}

// Host-driven base cases (from the small-class counts in h_ctr) and the
// sorted offset/length arrays.
static void host_base_cases(dss_set_s* s, const SortCaps& caps, const uint32_t* h_ctr, bool trace) {
    const hipStream_t st = s->stream;
    const uint64_t n = s->n;
    SortScratch& X = s->ss;
    const SortParams& T = s->sort_params;
    const SmallLists L = caps.L;
    const uint8_t* chars = s->arena();
    const uint64_t* off = s->off.as<uint64_t>();
    const uint32_t* len = s->len.as<uint32_t>();
    uint32_t* perm = s->perm.as<uint32_t>();
    uint32_t* lcp = s->lcp.as<uint32_t>();
    const uint32_t* cls = h_ctr + kCtrSmall;  // small-class segment and string counts
    if (trace) {
        fprintf(stderr, "[dss sort] small classes:");
        for (uint32_t c = 1; c < kClasses; ++c)
            if (cls[c]) fprintf(stderr, " c%u=%u segs/%u strings", c, cls[c], cls[kClasses + c]);
        fprintf(stderr, "\n");
    }
    // when most strings go through the one-wave base cases, their random
    // offset/length lookups read packed records
    const OffLen* offlen = nullptr;
    {
        uint64_t wave_strings = 0;
        for (uint32_t c = 1; c < kClasses && c <= T.wave_class_max; ++c) wave_strings += cls[kClasses + c];
        if (wave_strings >= n / 4 && wave_strings >= (1u << 16)) {
            X.offlen.reserve(n * sizeof(OffLen));
            KScope ks(s->timer, st, "offlen_pack", (double)n * 28.0);
            hipLaunchKernelGGL(k_pack_offlen, dim3(stride_grid(n, 256)), dim3(256), 0, st, off, len, n,
                               X.offlen.as<OffLen>());
            offlen = X.offlen.as<OffLen>();
        }
    }
    for (uint32_t c = 1; c < kClasses; ++c) {
        const uint32_t cnt = cls[c];
        if (!cnt) continue;
        DSS_REQUIRE(cnt <= L.cap[c], DSS_ERANGE, "small list overflow");
        // per string: id read, offset/length, one 8-char key, id + lcp writes
        KScope ks(s->timer, st, trace ? kClassNames[c] : "small_sort", (double)cls[kClasses + c] * 32.0);
        if (c <= T.wave_class_max) {
            const uint64_t waves = (cnt + (64u >> c) - 1) / (64u >> c);
            hipLaunchKernelGGL(k_small_sort_wave<false>, dim3((waves + 3) / 4), dim3(kThreads), 0, st,
                               X.small.as<SmallSeg>() + L.off[c], cnt, c, perm, lcp, chars, off, len, offlen,
                               X.keyA.as<uint64_t>(), X.keyB.as<uint64_t>(), s->s_off.as<uint64_t>(),
                               s->s_len.as<uint32_t>(), nullptr);
        } else {
            const uint32_t spb = kCap >> c;
            const uint32_t blocks = (cnt + spb - 1) / spb;
            hipLaunchKernelGGL(k_small_sort<false>, dim3(blocks), dim3(kThreads), 0, st,
                               X.small.as<SmallSeg>() + L.off[c], cnt, c, perm, lcp, chars, off, len,
                               s->s_off.as<uint64_t>(), s->s_len.as<uint32_t>(), nullptr);
        }
    }
    {
        // the sorted-length check per position; positions the base cases did
        // not place (normally few) add an id/offset/length gather
        KScope ks(s->timer, st, "sort_meta", n * 4.0);
        hipLaunchKernelGGL(k_sorted_meta<false>, dim3((n + 255) / 256), dim3(256), 0, st, perm, off, len, n,
                           s->s_off.as<uint64_t>(), s->s_len.as<uint32_t>(), 1, nullptr);
    }
    DSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ device-driven
// The device-driven form of the local sort: every kernel reads its work
// counts from the device counters and strides a fixed grid over them, so the
// launch sequence does not depend on the data and is captured once per set
// (for its buffer capacities and parameters) in two hipGraphs: (1)
// initialisation and kGraphPasses MSD passes, (2) the base cases of all
// classes (one launch for the one-wave classes, one for the block classes)
// and the sorted offset/length arrays.  Between them one read-back
// checks whether the list of large segments is empty; if not, the remaining
// passes run host-driven.  String and PE counts are device parameters written
// before each replay, and the list capacities follow the set's reservation,
// so one capture serves every batch that fits it.
constexpr uint32_t kGraphPasses = 2;
// pass_mode 0 takes the device-driven form for sets of at most
// kGraphMaxStrings strings over at least kGraphMinPes PEs (batches of jobs),
// where it measured at least as fast as the host-driven form; single-PE sorts
// of 10K-4M strings measured 5-8 % slower (a graph node costs about as much
// to dispatch as an eager launch, and the speculative empty passes add nodes).
constexpr uint64_t kGraphMaxStrings = 1ull << 22;
constexpr uint32_t kGraphMinPes = 16;
constexpr uint32_t kDevGrid = 1024;  // blocks of the persistent tile loops

__global__ void k_sort_params(uint64_t* __restrict__ dn, uint32_t* __restrict__ dp, uint64_t n, uint32_t p) {
    *dn = n;
    *dp = p;
}

static uint32_t grid_cap(uint64_t items, uint32_t per_block, uint32_t cap) {
    const uint64_t b = (items + per_block - 1) / per_block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

template <uint32_t kL, uint32_t kU>
static void launch_dev_pass(dss_set_s* s, const SortCaps& caps, uint32_t* ctr, PassBufs& b, bool first) {
    const hipStream_t st = s->stream;
    SortScratch& X = s->ss;
    const SortParams& T = s->sort_params;
    const uint8_t* chars = s->arena();
    const uint64_t* off = s->off.as<uint64_t>();
    const uint32_t* len = s->len.as<uint32_t>();
    uint32_t* perm = s->perm.as<uint32_t>();
    uint32_t* lcp = s->lcp.as<uint32_t>();
    const uint32_t* hp = ctr + kCtrPass + 4 * b.par;
    uint32_t* pn = ctr + kCtrPass + 4 * (b.par ^ 1);
    Tile* tl = X.tiles.as<Tile>();
    const uint32_t gt = grid_cap(caps.tile_cap, 1, kDevGrid);
    const uint32_t gc = grid_cap(caps.chunk_cap, 1, 256);
    hipLaunchKernelGGL(k_fill_tiles<true>, dim3(grid_cap(caps.tile_cap, 256, kDevGrid)), dim3(256), 0, st, b.cur, 0u,
                       0u, tl, hp, pn);
    hipLaunchKernelGGL((k_prefix_skip<kL, kU, true>), dim3(gt), dim3(kThreads), 0, st, tl, b.cur, perm, chars, off,
                       len, b.kc, first ? nullptr : b.pc_ids, hp);
    hipLaunchKernelGGL(k_gather<true>, dim3(gt), dim3(kThreads), 0, st, tl, b.cur, perm, chars, off, len, b.kc, hp);
    hipLaunchKernelGGL((k_hist<true, false>), dim3(gt), dim3(kThreads), 0, st, tl, b.cur, b.kc, X.hist.as<uint32_t>(), b.dm,
                       hp);
    hipLaunchKernelGGL(k_chunk_sum<true>, dim3(gc), dim3(kThreads), 0, st, b.cur, 0u, X.hist.as<uint32_t>(),
                       X.csum.as<uint32_t>(), hp);
    hipLaunchKernelGGL((k_scan_classify<true, false>), dim3(grid_cap(caps.seg_cap, kScGroups, 512)),
                       dim3(kThreads * kScGroups), 0, st, b.cur, 0u, X.hist.as<uint32_t>(), X.csum.as<uint32_t>(),
                       b.nxt, X.small.as<SmallSeg>(), caps.L, T.large_min, ctr, pn + kCtrLarge, lcp,
                       b.kn == X.keyB.as<uint64_t>() ? 1u : 0u, b.dm, hp);
    hipLaunchKernelGGL(k_chunk_apply<true>, dim3(gc), dim3(kThreads), 0, st, b.cur, 0u, X.hist.as<uint32_t>(),
                       X.csum.as<uint32_t>(), hp);
    hipLaunchKernelGGL((k_scatter<true, false>), dim3(gt), dim3(kScT), 0, st, tl, b.cur, X.hist.as<uint32_t>(), b.kc,
                       first ? nullptr : b.pc_ids, b.kn, b.pn_ids, perm, lcp, b.dm, hp);
    hipLaunchKernelGGL(k_build_tiles, dim3(1), dim3(1024), 0, st, b.nxt, pn);
    b.swap_pass();
}

// Enqueues (1): initialisation and kGraphPasses passes.
static void enqueue_dev_passes(dss_set_s* s, const SortCaps& caps, uint32_t* ctr, uint64_t ncap, uint32_t pcap) {
    const hipStream_t st = s->stream;
    SortScratch& X = s->ss;
    const SortParams& T = s->sort_params;
    uint64_t* dn = reinterpret_cast<uint64_t*>(ctr + 64);
    uint32_t* dp = ctr + 66;
    fill_small(st, ctr, 0, 64 * 4);
    hipLaunchKernelGGL(k_sort_init<true>, dim3(grid_cap(ncap, 256, 2048)), dim3(256), 0, st, s->perm.as<uint32_t>(),
                       s->lcp.as<uint32_t>(), s->s_len.as<uint32_t>(), (uint64_t)0, dn);
    hipLaunchKernelGGL(k_sort_seed<true>, dim3(grid_cap(pcap, 256, 256)), dim3(256), 0, st,
                       s->pe_begin_d.as<uint64_t>(), 0u, T.large_min, X.segA.as<Seg>(), X.small.as<SmallSeg>(),
                       caps.L, ctr, dp);
    hipLaunchKernelGGL(k_build_tiles, dim3(1), dim3(1024), 0, st, X.segA.as<Seg>(), ctr + kCtrPass);
    PassBufs b = first_pass_bufs(X);
    for (uint32_t k = 0; k < kGraphPasses; ++k) {
        const bool first = k == 0;
        if (T.skip_lanes == 16)
            T.skip_strings == 4 ? launch_dev_pass<16, 4>(s, caps, ctr, b, first)
            : T.skip_strings == 2 ? launch_dev_pass<16, 2>(s, caps, ctr, b, first)
                                  : launch_dev_pass<16, 1>(s, caps, ctr, b, first);
        else
            T.skip_strings == 4 ? launch_dev_pass<8, 4>(s, caps, ctr, b, first)
            : T.skip_strings == 2 ? launch_dev_pass<8, 2>(s, caps, ctr, b, first)
                                  : launch_dev_pass<8, 1>(s, caps, ctr, b, first);
    }
}

// Enqueues (2): the base cases of every class and the sorted offset/length
// arrays.
static void enqueue_dev_base_cases(dss_set_s* s, const SortCaps& caps, uint32_t* ctr, uint64_t ncap) {
    const hipStream_t st = s->stream;
    SortScratch& X = s->ss;
    const SortParams& T = s->sort_params;
    const SmallLists L = caps.L;
    const uint8_t* chars = s->arena();
    const uint64_t* off = s->off.as<uint64_t>();
    const uint32_t* len = s->len.as<uint32_t>();
    uint32_t* perm = s->perm.as<uint32_t>();
    uint32_t* lcp = s->lcp.as<uint32_t>();
    const uint64_t* dn = reinterpret_cast<const uint64_t*>(ctr + 64);
    // grids: the larger class capacity of each kind (every class has at most
    // ncap / 2^(c-1) segments)
    const uint32_t cmax = std::min<uint32_t>(T.wave_class_max, kClasses - 1);
    if (cmax >= 1) {
        uint64_t waves = 0;
        for (uint32_t c = 1; c <= cmax; ++c) waves = std::max<uint64_t>(waves, (L.cap[c] + (64u >> c) - 1) / (64u >> c));
        hipLaunchKernelGGL(k_small_sort_wave_all, dim3(grid_cap(waves, 4, kDevGrid)), dim3(kThreads), 0, st,
                           X.small.as<SmallSeg>(), L, cmax, perm, lcp, chars, off, len, X.keyA.as<uint64_t>(),
                           X.keyB.as<uint64_t>(), s->s_off.as<uint64_t>(), s->s_len.as<uint32_t>(), ctr);
    }
    if (cmax + 1 < kClasses) {
        uint64_t blocks = 0;
        for (uint32_t c = cmax + 1; c < kClasses; ++c)
            blocks = std::max<uint64_t>(blocks, (L.cap[c] + (kCap >> c) - 1) / (kCap >> c));
        hipLaunchKernelGGL(k_small_sort_all, dim3(grid_cap(blocks, 1, kDevGrid)), dim3(kThreads), 0, st,
                           X.small.as<SmallSeg>(), L, cmax + 1, perm, lcp, chars, off, len, s->s_off.as<uint64_t>(),
                           s->s_len.as<uint32_t>(), ctr);
    }
    hipLaunchKernelGGL(k_sorted_meta<true>, dim3(grid_cap(ncap, 256, 2048)), dim3(256), 0, st, perm, off, len,
                       (uint64_t)0, s->s_off.as<uint64_t>(), s->s_len.as<uint32_t>(), 1, dn);
}

// Captures the launches of `fill` on the set's stream into an executable graph.
template <class F>
static hipGraphExec_t capture(hipStream_t st, F&& fill) {
    hipGraph_t g = nullptr;
    DSS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        fill();
    } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        throw Error{DSS_EHIP, std::string("graph capture: ") + hipGetErrorString(e)};
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) throw Error{DSS_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ei)};
    return x;
}

static bool use_device_driven(const dss_set_s* s) {
    if (s->stream == nullptr || s->timer.enabled) return false;  // no capture on the null stream; timing is per launch
    const uint32_t mode = s->sort_params.pass_mode;
    return mode == 2 || (mode == 0 && s->n <= kGraphMaxStrings && s->p >= kGraphMinPes);
}

static void sort_lcp_device_driven(dss_set_s* s) {
    const hipStream_t st = s->stream;
    const uint64_t n = s->n;
    const uint32_t p = s->p;
    // the size class: grids and list capacities are those of the next powers
    // of two (at least 4096 strings), so small sorts replay small grids
    uint64_t ncap = 4096;
    while (ncap < n) ncap <<= 1;
    uint32_t pcap = 1;
    while (pcap < p) pcap <<= 1;
    // buffers grow to the reservation or the class, whichever is larger
    reserve_sort(s, std::max<uint64_t>(ncap, s->sort_cap_n), std::max<uint32_t>(pcap, s->sort_cap_p));
    const SortCaps caps = sort_caps(ncap, pcap, s->sort_params);
    SortScratch& X = s->ss;
    uint32_t* ctr = X.counters.as<uint32_t>();
    const SortGraphKey key{ncap, pcap, s->sort_params.large_min, s->sort_params.wave_class_max,
                           s->sort_params.skip_lanes, s->sort_params.skip_strings, s->arena(), s->off.p, s->len.p,
                           s->pe_begin_d.p, s->perm.p, s->lcp.p, s->s_off.p, s->s_len.p, X.keyA.p, X.keyB.p,
                           X.pbA.p, X.pbB.p, X.segA.p, X.segB.p, X.tiles.p, X.hist.p, X.csum.p, X.small.p,
                           X.counters.p, st};
    bool hit = false;
    SortGraph& G = X.graphs.find(key, hit);
    if (!hit) {
        G.passes = capture(st, [&] { enqueue_dev_passes(s, caps, ctr, ncap, pcap); });
        G.base = capture(st, [&] { enqueue_dev_base_cases(s, caps, ctr, ncap); });
        G.key = key;
    }
    hipLaunchKernelGGL(k_sort_params, dim3(1), dim3(1), 0, st, reinterpret_cast<uint64_t*>(ctr + 64), ctr + 66, n,
                       p);
    DSS_HIP(hipGetLastError());
    DSS_HIP(hipGraphLaunch(G.passes, st));
    uint32_t h_ctr[64];
    s->pp.mbox.fetch(st, h_ctr, ctr, sizeof(h_ctr));
    PassBufs b = first_pass_bufs(X);
    for (uint32_t k = 0; k < kGraphPasses; ++k) b.swap_pass();
    if (h_ctr[kCtrPass + 4 * b.par + kCtrLarge]) host_passes(s, caps, ctr, h_ctr, b, false, false);
    DSS_HIP(hipGraphLaunch(G.base, st));
}

void sort_lcp(dss_set_s* s) {
    const uint64_t n = s->n;
    const hipStream_t st = s->stream;
    s->sorted = false;
    if (n == 0) {
        reserve_sort(s, n, s->p);
        s->sorted = true;
        return;
    }
    if (use_device_driven(s)) {
        sort_lcp_device_driven(s);
        s->sorted = true;
        return;
    }
    const SortCaps caps = reserve_sort(s, n, s->p);
    SortScratch& X = s->ss;
    const SortParams& T = s->sort_params;
    uint32_t* ctr = X.counters.as<uint32_t>();
    fill_small(st, ctr, 0, 64 * 4);
    {
        KScope ks(s->timer, st, "sort_init", n * 8.0);
        hipLaunchKernelGGL(k_sort_init<false>, dim3((n + 255) / 256), dim3(256), 0, st, s->perm.as<uint32_t>(),
                           s->lcp.as<uint32_t>(), s->s_len.as<uint32_t>(), n, nullptr);
    }
    hipLaunchKernelGGL(k_sort_seed<false>, dim3((s->p + 255) / 256), dim3(256), 0, st,
                       s->pe_begin_d.as<uint64_t>(), s->p, T.large_min, X.segA.as<Seg>(), X.small.as<SmallSeg>(),
                       caps.L, ctr, nullptr);
    DSS_HIP(hipGetLastError());
    // every pass ends with one read-back of all counters: the next list's
    // segment count and its tile/chunk counts (k_build_tiles of the next pass
    // runs before the read-back), and the small-class counts
    uint32_t h_ctr[64];
    hipLaunchKernelGGL(k_build_tiles, dim3(1), dim3(1024), 0, st, X.segA.as<Seg>(), ctr + kCtrPass);
    // the first pass's segment, tile and chunk counts follow from the PE
    // sizes (one segment per PE above large_min); without large segments the
    // small-class counts the seed made are read back instead
    {
        uint64_t nl = 0, nt = 0, nc = 0;
        for (uint32_t r = 0; r < s->p; ++r) {
            const uint64_t cnt = s->pe_begin[r + 1] - s->pe_begin[r];
            if (cnt <= T.large_min) continue;
            const uint64_t t = (cnt + kTile - 1) / kTile;
            ++nl;
            nt += t;
            nc += t > kChunkTiles ? (t + kChunkTiles - 1) / kChunkTiles : 0;
        }
        if (nl) {
            std::memset(h_ctr, 0, sizeof(h_ctr));
            h_ctr[kCtrPass + kCtrLarge] = (uint32_t)nl;
            h_ctr[kCtrPass + kCtrTiles] = (uint32_t)nt;
            h_ctr[kCtrPass + kCtrChunks] = (uint32_t)nc;
        } else {
            s->pp.mbox.fetch(st, h_ctr, ctr, sizeof(h_ctr));
        }
    }
    const bool trace = getenv("DSS_SORT_TRACE") != nullptr;
    PassBufs b = first_pass_bufs(X);
    b.dm = digit_map(s);
    host_passes(s, caps, ctr, h_ctr, b, trace, true);
    host_base_cases(s, caps, h_ctr, trace);
    s->sorted = true;
}

}  // namespace dss

using namespace dss;

extern "C" {

int dss_sort_lcp(dss_set s) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        sort_lcp(s);
    });
}

int dss_set_reserve(dss_set s, uint64_t num_strings, uint32_t num_pes) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        DSS_REQUIRE(num_strings < 0xFFFFFFFFull && num_pes >= 1 && num_pes <= (1u << 16), DSS_ERANGE,
                    "num_strings must be < 2^32-1 and num_pes in [1, 65536]");
        const uint64_t nn = std::max<uint64_t>(num_strings, 1);
        DSS_HIP(hipStreamSynchronize(s->stream));
        reserve_sort(s, num_strings, num_pes);
        // the base cases' packed offset/length records are otherwise grown on
        // first use, an allocation inside a sort (which can wait on the
        // driver, e.g. for freed memory to be cleared)
        s->ss.offlen.reserve(nn * sizeof(OffLen));
        s->sort_cap_n = std::max<uint64_t>(s->sort_cap_n, num_strings);
        s->sort_cap_p = std::max<uint32_t>(s->sort_cap_p, num_pes);
        s->pe_begin_d.reserve_keep((uint64_t)(num_pes + 1) * 8, s->stream);
        s->aux.reserve((uint64_t)num_pes * 8 + (num_strings / 65536 + num_pes) * 16 + 8);
        if (s->borrowed) {  // a batch set: its offset/length copies too
            s->off.reserve_keep(nn * 8, s->stream);
            s->len.reserve_keep(nn * 4, s->stream);
        }
        // grown result arrays hold no results: the set must be sorted again
        s->sorted = false;
        s->pp.has_dist = s->pp.partitioned = s->pp.exchanged = s->pp.merged = s->pp.extended = false;
    });
}

int dss_set_sort_params(dss_set s, const dss_sort_params_t* params) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        SortParams T;  // NULL restores the defaults
        if (params) {
            DSS_REQUIRE(params->large_min >= 16 && params->large_min <= kSmallMax, DSS_EINVAL,
                        "large_min must be in [16, 1024]");
            DSS_REQUIRE(params->wave_class_max <= kWaveClassMax, DSS_EINVAL, "wave_class_max must be in [0, 6]");
            DSS_REQUIRE(params->skip_lanes == 8 || params->skip_lanes == 16, DSS_EINVAL, "skip_lanes must be 8 or 16");
            DSS_REQUIRE(params->skip_strings <= 4 && params->skip_strings != 3, DSS_EINVAL,
                        "skip_strings must be 0 (default), 1, 2 or 4");
            DSS_REQUIRE(params->merge_fan_in == 0 || params->merge_fan_in == 2 || params->merge_fan_in == 4 ||
                            params->merge_fan_in == 8 || params->merge_fan_in == 16,
                        DSS_EINVAL, "merge_fan_in must be 0 (default), 2, 4, 8 or 16");
            DSS_REQUIRE(params->pass_mode <= 2, DSS_EINVAL, "pass_mode must be 0 (auto), 1 or 2");
            DSS_REQUIRE(params->merge_tile == 0 || params->merge_tile == 512 || params->merge_tile == 1024 ||
                            params->merge_tile == 2048,
                        DSS_EINVAL, "merge_tile must be 0 (default), 512, 1024 or 2048");
            DSS_REQUIRE(params->digit_chars <= 3, DSS_EINVAL, "digit_chars must be 0 (auto), 1, 2 or 3");
            T.large_min = params->large_min;
            T.wave_class_max = params->wave_class_max;
            T.skip_lanes = params->skip_lanes;
            if (params->skip_strings) T.skip_strings = params->skip_strings;
            if (params->merge_fan_in) T.merge_fan = params->merge_fan_in;
            T.pass_mode = params->pass_mode;
            if (params->merge_tile) T.merge_tile = params->merge_tile;
            T.digit_chars = params->digit_chars;
        }
        s->sort_params = T;
    });
}

int dss_get_sort_params(dss_set s, dss_sort_params_t* params) {
    return guarded([&] {
        DSS_REQUIRE(s && params, DSS_EINVAL, "NULL argument");
        params->large_min = s->sort_params.large_min;
        params->wave_class_max = s->sort_params.wave_class_max;
        params->skip_lanes = s->sort_params.skip_lanes;
        params->skip_strings = s->sort_params.skip_strings;
        params->merge_fan_in = s->sort_params.merge_fan;
        params->pass_mode = s->sort_params.pass_mode;
        params->merge_tile = s->sort_params.merge_tile;
        params->digit_chars = s->sort_params.digit_chars;
    });
}

int dss_pe_dist_sizes(dss_set s, uint64_t* d_per_pe) {
    return guarded([&] {
        DSS_REQUIRE(s && d_per_pe, DSS_EINVAL, "NULL argument");
        DSS_REQUIRE(s->sorted, DSS_ESTATE, "dss_sort_lcp has not run");
        const uint32_t p = s->p;
        const uint64_t n = s->n;
        DevBuf& buf = s->aux;
        buf.reserve((uint64_t)std::max<uint32_t>(p, 1) * 8);
        unsigned long long* d = buf.as<unsigned long long>();
        fill_small(s->stream, d, 0, (size_t)p * 8);
        if (n) {
            hipLaunchKernelGGL(k_pe_dsum, dim3((uint32_t)((n + kDistChunk - 1) / kDistChunk)), dim3(256), 0,
                               s->stream, s->lcp.as<uint32_t>(), s->pe_begin_d.as<uint64_t>(), p, n, d);
            DSS_HIP(hipGetLastError());
        }
        s->pp.mbox.fetch(s->stream, d_per_pe, d, (size_t)p * 8);
    });
}

int dss_get_sorted_pe(dss_set s, uint32_t pe, uint32_t* perm, uint32_t* lcp) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        DSS_REQUIRE(s->sorted, DSS_ESTATE, "dss_sort_lcp has not run");
        DSS_REQUIRE(pe < s->p, DSS_EINVAL, "PE out of range");
        const uint64_t b = s->pe_begin[pe], cnt = s->pe_begin[pe + 1] - b;
        if (perm && cnt)
            DSS_HIP(hipMemcpyAsync(perm, s->perm.as<uint32_t>() + b, cnt * 4, hipMemcpyDeviceToHost, s->stream));
        if (lcp && cnt)
            DSS_HIP(hipMemcpyAsync(lcp, s->lcp.as<uint32_t>() + b, cnt * 4, hipMemcpyDeviceToHost, s->stream));
        DSS_HIP(hipStreamSynchronize(s->stream));
        if (perm)
            for (uint64_t i = 0; i < cnt; ++i) perm[i] -= (uint32_t)b;
    });
}

int dss_get_sorted(dss_set s, uint32_t* perm, uint32_t* lcp) {
    return guarded([&] {
        DSS_REQUIRE(s, DSS_EINVAL, "NULL set");
        DSS_REQUIRE(s->sorted, DSS_ESTATE, "dss_sort_lcp has not run");
        if (perm && s->n)
            DSS_HIP(hipMemcpyAsync(perm, s->perm.p, s->n * 4, hipMemcpyDeviceToHost, s->stream));
        if (lcp && s->n)
            DSS_HIP(hipMemcpyAsync(lcp, s->lcp.p, s->n * 4, hipMemcpyDeviceToHost, s->stream));
        DSS_HIP(hipStreamSynchronize(s->stream));
    });
}

}  // extern "C"

extern "C" int dss_set_local_order(dss_set s, const uint32_t* perm, const uint32_t* lcp) {
    return guarded([&] {
        DSS_REQUIRE(s && (s->n == 0 || (perm && lcp)), DSS_EINVAL, "NULL argument");
        const uint64_t n = s->n;
        s->perm.reserve(std::max<uint64_t>(n, 1) * 4);
        s->lcp.reserve(std::max<uint64_t>(n, 1) * 4);
        s->s_off.reserve(std::max<uint64_t>(n, 1) * 8);
        s->s_len.reserve(std::max<uint64_t>(n, 1) * 4);
        if (n) {
            for (uint32_t r = 0; r < s->p; ++r)
                for (uint64_t i = s->pe_begin[r]; i < s->pe_begin[r + 1]; ++i)
                    DSS_REQUIRE(perm[i] >= s->pe_begin[r] && perm[i] < s->pe_begin[r + 1], DSS_EINVAL,
                                "perm must permute every PE's own strings");
            DSS_HIP(hipMemcpyAsync(s->perm.p, perm, n * 4, hipMemcpyHostToDevice, s->stream));
            DSS_HIP(hipMemcpyAsync(s->lcp.p, lcp, n * 4, hipMemcpyHostToDevice, s->stream));
            hipLaunchKernelGGL(k_sorted_meta<false>, dim3((n + 255) / 256), dim3(256), 0, s->stream,
                               s->perm.as<uint32_t>(), s->off.as<uint64_t>(), s->len.as<uint32_t>(), n,
                               s->s_off.as<uint64_t>(), s->s_len.as<uint32_t>(), 0, nullptr);
            DSS_HIP(hipGetLastError());
            DSS_HIP(hipStreamSynchronize(s->stream));
        }
        s->sorted = true;
        s->pp.has_dist = s->pp.partitioned = s->pp.exchanged = s->pp.merged = s->pp.extended = false;
    });
}
