// Wire codecs of the exchange: varint LCP ranges and delta + Golomb streams
// of sorted hash values / duplicate indices.
//
// Replaces
//   IntegerCompression::writeRanges / readRanges (src/mpi/byte_encoder.hpp:
//     31-78) over PutVarint / GetVarint (src/encoding/integer_compression.hpp),
//     used by sendLcps (src/mpi/alltoall.hpp:406-425);
//   getDeltaEncoding / getDeltaDecoding (src/encoding/golomb_encoding.hpp:
//     60-109) over DeltaStreamWriter/Reader (delta_stream.hpp) and
//     GolombBitStreamWriter/Reader (golomb_bit_stream.hpp, bit_stream.hpp),
//     framed as AllToAllHashesGolomb (src/sorter/distributed/bloomfilter.hpp:
//     164-313) or AllToAllHashValuesPipeline::getEncoding (:322-331).
//
// The reference codes sequentially, one value after the other.  Here every
// stream is coded by the whole grid:
//   varint encode  -- byte length per value, scan, one lane writes a value;
//   varint decode  -- a byte without 0x80 ends a value: terminators are
//                     counted and ranked per block, and every terminator's
//                     lane assembles its value from the <= 10 bytes before it;
//   Golomb encode  -- code length per gap, scan, every code ORs its bits into
//                     the zeroed words (words inside one code's unary run are
//                     stored whole), the last word of a stream is padded with
//                     ones;
//   Golomb decode  -- code boundaries are found without a sequential walk:
//                     every 4096-bit chunk is decoded speculatively from its
//                     first bit, recording the code starts it visits; then
//                     every chunk is decoded again from its predecessor's exit
//                     until that path meets one of its speculative starts
//                     (Golomb codes resynchronise within a few codes), and the
//                     exits are iterated to a fixed point (chunk k of a stream
//                     is exact after k iterations; in practice after 1-2).
//                     Gaps are then decoded in parallel from the exact start
//                     bitmap and summed with a segmented scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "dss_device.hpp"
#include "dss_internal.hpp"

struct dss_codec_s {
    int device = 0;
    hipStream_t stream = nullptr;
    dss::PipeState pp;  // scan scratch
    dss::KernelTimer timer;
    // host-form staging and per-call scratch (grow-only)
    dss::DevBuf in, out, lenv, pos, meta, bitmap, spec, xa, xb, slots, flag;
};

namespace dss {
namespace {

constexpr uint64_t kEnd = ~0ull;
constexpr uint64_t kNone64 = ~0ull - 1;  // no exit found yet
constexpr uint32_t kChunkBits = 4096;  // speculative decode chunk (64 words)
constexpr uint32_t kChunkWords = kChunkBits / 64;
constexpr uint32_t kVBytes = 16;      // bytes per lane in the varint decode
constexpr uint32_t kVThreads = 256;

unsigned ceil_log2(uint64_t b) {  // tlx::integer_log2_ceil
    unsigned k = 0;
    while (k < 64 && (1ull << k) < b) ++k;
    return k;
}

// ---------------------------------------------------------------- varint
__global__ void k_varint_len(const uint64_t* __restrict__ v, uint64_t n, uint64_t* __restrict__ L) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == n) {
            L[i] = 0;
            continue;
        }
        const uint64_t x = v[i];
        const uint32_t bits = x ? 64 - __clzll(x) : 1;
        L[i] = (bits + 6) / 7;
    }
}

__global__ void k_varint_put(const uint64_t* __restrict__ v, uint64_t n, const uint64_t* __restrict__ O,
                             uint8_t* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t x = v[i], o = O[i];
        while (x >= 128) {
            out[o++] = (uint8_t)((x & 0x7F) | 0x80);
            x >>= 7;
        }
        out[o] = (uint8_t)x;
    }
}

// Terminators (bytes without 0x80) per block of kVThreads * kVBytes bytes.
__global__ void __launch_bounds__(kVThreads)
k_vterm_count(const uint8_t* __restrict__ b, uint64_t nbytes, uint64_t* __restrict__ cnt) {
    const uint64_t seg = (uint64_t)blockIdx.x * kVThreads * kVBytes;
    uint32_t c = 0;
    const uint64_t s = seg + (uint64_t)threadIdx.x * kVBytes;
    for (uint32_t k = 0; k < kVBytes; ++k)
        if (s + k < nbytes && !(b[s + k] & 0x80)) ++c;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    __shared__ uint32_t ws[kVThreads / 64];
    if (lane_id() == 0) ws[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kVThreads / 64; ++w) t += ws[w];
        cnt[blockIdx.x] = t;
    }
}

// Every terminator's lane assembles its value; err bit 0: a value longer than
// 10 bytes or a 10th byte with bits beyond the 64th (GetVarint's overflow).
__global__ void __launch_bounds__(kVThreads)
k_varint_decode(const uint8_t* __restrict__ b, uint64_t nbytes, const uint64_t* __restrict__ base,
                uint64_t* __restrict__ out, uint32_t* __restrict__ err) {
    __shared__ uint32_t pre[kVThreads];
    const uint64_t seg = (uint64_t)blockIdx.x * kVThreads * kVBytes;
    const uint64_t s = seg + (uint64_t)threadIdx.x * kVBytes;
    uint32_t c = 0;
    for (uint32_t k = 0; k < kVBytes; ++k)
        if (s + k < nbytes && !(b[s + k] & 0x80)) ++c;
    pre[threadIdx.x] = c;
    __syncthreads();
    for (uint32_t o = 1; o < kVThreads; o <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t x = threadIdx.x >= o ? pre[threadIdx.x - o] : 0;
        __syncthreads();
        pre[threadIdx.x] += x;
        __syncthreads();
    }
    uint64_t idx = base[blockIdx.x] + pre[threadIdx.x] - c;
    for (uint32_t k = 0; k < kVBytes; ++k) {
        const uint64_t j = s + k;
        if (j >= nbytes || (b[j] & 0x80)) continue;
        uint64_t st = j;
        uint32_t len = 1;
        while (st > 0 && (b[st - 1] & 0x80) && len <= 10) {
            --st;
            ++len;
        }
        if (len > 10 || (len == 10 && (b[j] & 0xFE))) {
            atomicOr(err, 1u);
            ++idx;
            continue;
        }
        uint64_t v = 0;
        for (uint32_t t = 0; t < len; ++t) v |= (uint64_t)(b[st + t] & 0x7F) << (7 * t);
        out[idx++] = v;
    }
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const uint64_t* __restrict__ idx, uint32_t m,
                             uint64_t* __restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) dst[i] = src[idx[i]];
}

// ---------------------------------------------------------------- Golomb
struct GRange {         // encoder: one input range
    uint64_t vbegin, cnt, b, mlv;
    uint64_t bitword0;  // first word of the range's bit stream in the output
    uint64_t nbits;     // bits of its codes (before padding)
    uint64_t hdr0;      // first word of the block (headers)
    uint32_t log2b, layout;
};

struct GBlock {         // decoder: one block found in the input
    uint64_t raw;       // word index of the raw first value (kEnd: none)
    uint64_t bitw0;     // first word of the bit stream
    uint64_t nbits;     // 64 * bit-stream words
    uint64_t b, mlv;
    uint64_t chunk0, nchunks;
    uint32_t log2b, pad;
};

__device__ inline uint32_t code_bits(uint64_t gap, uint64_t b, uint32_t log2b, uint64_t mlv, uint64_t* q_out,
                                     uint64_t* rv_out) {
    const uint64_t q = gap / b, r = gap - q * b;
    *q_out = q;
    if (r >= mlv) {
        *rv_out = r + mlv;
        return log2b;
    }
    *rv_out = r;
    return log2b - 1;
}

__device__ inline uint32_t find_range(const GRange* __restrict__ R, uint32_t nr, uint64_t i) {
    uint32_t lo = 0, hi = nr;  // last range with vbegin <= i
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (R[mid].vbegin <= i) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_golomb_len(const uint64_t* __restrict__ v, uint64_t n, const GRange* __restrict__ R, uint32_t nr,
                             uint64_t* __restrict__ L) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == n) {
            L[i] = 0;
            continue;
        }
        const GRange& g = R[find_range(R, nr, i)];
        if (i == g.vbegin) {
            L[i] = 0;  // the first value travels as a raw word
            continue;
        }
        uint64_t q, rv;
        const uint32_t rb = code_bits(v[i] - v[i - 1], g.b, g.log2b, g.mlv, &q, &rv);
        L[i] = q + 1 + rb;
    }
}

__global__ void k_golomb_head(const uint64_t* __restrict__ v, const GRange* __restrict__ R, uint32_t nr,
                              uint64_t* __restrict__ out) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const GRange g = R[r];
    const uint64_t words = (g.nbits + 63) / 64;
    if (g.layout == DSS_GOLOMB_BLOCKS) {
        if (g.cnt == 0) return;
        out[g.hdr0] = 2 + words;  // b, raw value and the bit stream
        out[g.hdr0 + 1] = g.b;
        out[g.hdr0 + 2] = v[g.vbegin];
    } else {
        out[g.hdr0] = g.cnt ? 1 + words : 0;
        if (g.cnt) out[g.hdr0 + 1] = v[g.vbegin];
    }
    if (g.nbits & 63)  // fill the last word up with ones
        atomicOr((unsigned long long*)&out[g.bitword0 + words - 1], (unsigned long long)(~0ull >> (g.nbits & 63)));
}

__device__ inline void or_bits(uint64_t* out, uint64_t pos, uint64_t val, uint32_t nb) {
    if (nb == 0) return;
    const uint64_t w = pos >> 6;
    const uint32_t off = pos & 63;
    if (off + nb <= 64) {
        atomicOr((unsigned long long*)&out[w], (unsigned long long)(val << (64 - off - nb)));
    } else {
        const uint32_t k1 = 64 - off;
        atomicOr((unsigned long long*)&out[w], (unsigned long long)(val >> (nb - k1)));
        atomicOr((unsigned long long*)&out[w + 1], (unsigned long long)(val << (64 - (nb - k1))));
    }
}

__global__ void k_golomb_put(const uint64_t* __restrict__ v, uint64_t n, const GRange* __restrict__ R, uint32_t nr,
                             const uint64_t* __restrict__ G, uint64_t* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const GRange& g = R[find_range(R, nr, i)];
        if (i == g.vbegin) continue;
        uint64_t q, rv;
        const uint32_t rb = code_bits(v[i] - v[i - 1], g.b, g.log2b, g.mlv, &q, &rv);
        uint64_t pos = g.bitword0 * 64 + (G[i] - G[g.vbegin]);
        while (q) {  // the unary part: q ones
            const uint32_t off = pos & 63;
            const uint64_t k = min<uint64_t>(64 - off, q);
            if (k == 64) {
                out[pos >> 6] = ~0ull;  // a word inside this code only
            } else {
                const uint64_t m = ((1ull << k) - 1) << (64 - off - k);
                atomicOr((unsigned long long*)&out[pos >> 6], (unsigned long long)m);
            }
            pos += k;
            q -= k;
        }
        pos += 1;  // the terminating zero
        or_bits(out, pos, rv, rb);
    }
}

__device__ inline uint64_t get_bits(const uint64_t* __restrict__ W, uint64_t pos, uint32_t k) {
    if (k == 0) return 0;
    const uint64_t w = pos >> 6;
    const uint32_t off = pos & 63;
    uint64_t hi = W[w] << off;
    if (off + k > 64) hi |= W[w + 1] >> (64 - off);
    return hi >> (64 - k);
}

// End of the code that starts at `pos` (kEnd if no zero bit follows, i.e.
// only padding remains, or the code runs past the stream); *val = its gap.
__device__ inline uint64_t golomb_next(const uint64_t* __restrict__ W, const GBlock& B, uint64_t pos, uint64_t* val) {
    uint64_t z = kEnd;
    for (uint64_t p = pos; p < B.nbits;) {
        const uint64_t w = p >> 6;
        const uint64_t x = (~W[w]) << (p & 63);
        if (x) {
            z = p + __clzll(x);
            break;
        }
        p = (w + 1) << 6;
    }
    if (z == kEnd) return kEnd;
    uint64_t p = z + 1, r = 0;
    if (B.log2b) {
        if (p + B.log2b - 1 > B.nbits) return kEnd;
        r = get_bits(W, p, B.log2b - 1);
        p += B.log2b - 1;
        if (r >= B.mlv) {
            if (p + 1 > B.nbits) return kEnd;
            r = ((r << 1) | get_bits(W, p, 1)) - B.mlv;
            p += 1;
        }
    }
    *val = (z - pos) * B.b + r;
    return p;
}

__device__ inline uint32_t find_block(const GBlock* __restrict__ blk, uint32_t nb, uint64_t c) {
    uint32_t lo = 0, hi = nb;  // last block with chunk0 <= c and chunks
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (blk[mid].chunk0 <= c) lo = mid; else hi = mid;
    }
    return lo;
}

// Code starts are recorded in a bitmap with one word per stream word;
// a lane accumulates the word it is in and stores it when it moves on (its
// chunk's words are its own, zeroed before).
struct StartBits {
    uint64_t* out;
    uint64_t w = kEnd, acc = 0;
    __device__ explicit StartBits(uint64_t* o) : out(o) {}
    __device__ void set(uint64_t pos) {
        const uint64_t w2 = pos >> 6;
        if (w2 != w) {
            flush();
            w = w2;
        }
        acc |= 1ull << (63 - (pos & 63));
    }
    __device__ void flush() {
        if (w != kEnd && acc) out[w] |= acc;
        acc = 0;
    }
};

__global__ void k_golomb_spec(const uint64_t* __restrict__ words, const GBlock* __restrict__ blk, uint32_t nb,
                              uint64_t nchunks, uint64_t* __restrict__ S, uint64_t* __restrict__ E) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const GBlock B = blk[find_block(blk, nb, c)];
    const uint64_t* W = words + B.bitw0;
    const uint64_t cs = (c - B.chunk0) * kChunkBits, ce = min<uint64_t>(cs + kChunkBits, B.nbits);
    StartBits sb(S + B.bitw0);
    uint64_t pos = cs;
    while (pos < ce) {
        uint64_t v;
        const uint64_t nx = golomb_next(W, B, pos, &v);
        if (nx == kEnd) {
            pos = kEnd;
            break;
        }
        sb.set(pos);
        pos = nx;
    }
    sb.flush();
    E[c] = pos;
}

// One round of the fixed point of chunk exits: a chunk's exit given the exit
// of its predecessor (its entry).  Entering at its first bit, the
// speculative exit holds; otherwise the chunk is decoded from the entry
// until the walk meets a speculative start (from there on both walks agree)
// or leaves the chunk.
__global__ void k_golomb_sync(const uint64_t* __restrict__ words, const GBlock* __restrict__ blk, uint32_t nb,
                              uint64_t nchunks, const uint64_t* __restrict__ S, const uint64_t* __restrict__ E,
                              const uint64_t* __restrict__ Xin, uint64_t* __restrict__ Xout,
                              uint32_t* __restrict__ changed) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const GBlock B = blk[find_block(blk, nb, c)];
    const uint64_t* W = words + B.bitw0;
    const uint64_t cs = (c - B.chunk0) * kChunkBits, ce = min<uint64_t>(cs + kChunkBits, B.nbits);
    const uint64_t* Sb = S + B.bitw0;
    const uint64_t entry = c == B.chunk0 ? 0 : Xin[c - 1];
    uint64_t x;
    if (entry == cs) {
        x = E[c];
    } else if (entry == kEnd || entry >= ce) {
        x = entry;  // no code starts in this chunk
    } else {
        uint64_t pos = entry;
        x = kNone64;
        while (pos < ce) {
            if ((Sb[pos >> 6] >> (63 - (pos & 63))) & 1) {
                x = E[c];  // joined the speculative walk
                break;
            }
            uint64_t v;
            const uint64_t nx = golomb_next(W, B, pos, &v);
            if (nx == kEnd) {
                pos = kEnd;
                break;
            }
            pos = nx;
        }
        if (x == kNone64) x = pos;
    }
    Xout[c] = x;
    if (x != Xin[c]) atomicOr(changed, 1u);
}

// The exact start bitmap T (a copy of S on entry) after the fixed point:
// a chunk entered elsewhere than at its first bit replaces its speculative
// starts before the point where the walks join by the decoded ones.
__global__ void k_golomb_starts(const uint64_t* __restrict__ words, const GBlock* __restrict__ blk, uint32_t nb,
                                uint64_t nchunks, const uint64_t* __restrict__ S, const uint64_t* __restrict__ X,
                                uint64_t* __restrict__ T) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    const GBlock B = blk[find_block(blk, nb, c)];
    const uint64_t* W = words + B.bitw0;
    const uint64_t cs = (c - B.chunk0) * kChunkBits, ce = min<uint64_t>(cs + kChunkBits, B.nbits);
    const uint64_t* Sb = S + B.bitw0;
    uint64_t* Tb = T + B.bitw0;
    const uint64_t entry = c == B.chunk0 ? 0 : X[c - 1];
    if (entry == cs) return;  // T = S here
    const uint64_t w0 = cs / 64, w1 = (ce + 63) / 64;
    if (entry == kEnd || entry >= ce) {
        for (uint64_t w = w0; w < w1; ++w) Tb[w] = 0;
        return;
    }
    uint64_t pos = entry, w = w0, acc = 0;
    bool synced = false;
    while (pos < ce) {
        if ((Sb[pos >> 6] >> (63 - (pos & 63))) & 1) {
            synced = true;
            break;
        }
        uint64_t v;
        const uint64_t nx = golomb_next(W, B, pos, &v);
        if (nx == kEnd) {
            pos = kEnd;
            break;
        }
        for (; w < (pos >> 6); ++w) {  // words before the start's word are complete
            Tb[w] = acc;
            acc = 0;
        }
        acc |= 1ull << (63 - (pos & 63));
        pos = nx;
    }
    if (synced) {
        const uint64_t pw = pos >> 6;
        for (; w < pw; ++w) {
            Tb[w] = acc;
            acc = 0;
        }
        Tb[pw] = acc | (Sb[pw] & (~0ull >> (pos & 63)));  // later words keep S
    } else {
        for (; w < w1; ++w) {
            Tb[w] = acc;
            acc = 0;
        }
    }
}

// Value slots: block j owns slot chunk0_j + j (its raw value) and one slot
// per chunk after it (the codes starting in the chunk).
__global__ void k_golomb_count(const GBlock* __restrict__ blk, uint32_t nb, uint64_t nchunks,
                               const uint64_t* __restrict__ T, uint64_t* __restrict__ A) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nchunks) {
        const uint32_t j = find_block(blk, nb, c);
        const GBlock B = blk[j];
        const uint64_t cs = (c - B.chunk0) * kChunkBits;
        uint64_t k = 0;
        for (uint32_t i = 0; i < kChunkWords && cs + 64 * i < B.nbits; ++i) k += __popcll(T[B.bitw0 + cs / 64 + i]);
        A[c + j + 1] = k;
    }
    if (c < nb) A[blk[c].chunk0 + c] = blk[c].raw != kEnd ? 1 : 0;
    if (c == 0) A[nchunks + nb] = 0;
}

__global__ void k_golomb_decode(const uint64_t* __restrict__ words, const GBlock* __restrict__ blk, uint32_t nb,
                                uint64_t nchunks, const uint64_t* __restrict__ T, const uint64_t* __restrict__ Apos,
                                uint64_t* __restrict__ vals, uint64_t* __restrict__ bstart) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nchunks) {
        const uint32_t j = find_block(blk, nb, c);
        const GBlock B = blk[j];
        const uint64_t* W = words + B.bitw0;
        const uint64_t cs = (c - B.chunk0) * kChunkBits;
        uint64_t o = Apos[c + j + 1];
        for (uint32_t i = 0; i < kChunkWords && cs + 64 * i < B.nbits; ++i) {
            uint64_t m = T[B.bitw0 + cs / 64 + i];
            while (m) {
                const uint32_t bit = __clzll(m);
                m &= ~(1ull << (63 - bit));
                uint64_t v = 0;
                (void)golomb_next(W, B, cs + 64 * i + bit, &v);
                vals[o++] = v;
            }
        }
    }
    if (c < nb) {
        const uint64_t s = Apos[blk[c].chunk0 + c];
        bstart[c] = s;
        if (blk[c].raw != kEnd) vals[s] = words[blk[c].raw];
    }
    if (c == 0) bstart[nb] = Apos[nchunks + nb];
}

// vals = raw value, then gaps per block; ex = exclusive scan of vals.
__global__ void k_golomb_sum(uint64_t* __restrict__ vals, const uint64_t* __restrict__ ex, uint64_t n,
                             const uint64_t* __restrict__ bstart, uint32_t nb) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nb;  // last block starting at or before i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (bstart[mid] <= i) lo = mid; else hi = mid;
        }
        vals[i] = ex[i] + vals[i] - ex[bstart[lo]];
    }
}

// Walks the block headers (sequential by construction: each header gives
// the size of its block).  err: 1 malformed, 2 more blocks than allowed.
__global__ void k_golomb_parse(const uint64_t* __restrict__ w, uint64_t nwords, int layout, uint32_t max_blocks,
                               uint64_t b_pipe, GBlock* __restrict__ blk, uint32_t* __restrict__ nblk_out,
                               uint32_t* __restrict__ err) {
    uint64_t pos = 0;
    uint32_t k = 0;
    const bool pipelined = layout == DSS_GOLOMB_PIPELINED;
    while (pipelined ? k < max_blocks : pos < nwords) {
        if (k >= max_blocks || pos >= nwords) {
            *err = k >= max_blocks ? 2 : 1;
            break;
        }
        const uint64_t size = w[pos];
        const uint64_t hdr = pipelined ? 1 : 2;  // words before the encoding
        if ((!pipelined && size < 1) || size > nwords - pos - 1) {
            *err = 1;
            break;
        }
        GBlock B;
        B.b = pipelined ? b_pipe : w[pos + 1];
        const uint64_t enc = pipelined ? size : size - 1;  // words of the encoding
        B.raw = enc ? pos + hdr : kEnd;
        B.bitw0 = pos + hdr + 1;
        B.nbits = enc > 1 ? (enc - 1) * 64 : 0;
        B.chunk0 = 0;
        B.nchunks = 0;
        B.log2b = 0;
        B.mlv = 0;
        B.pad = 0;
        blk[k++] = B;
        pos += 1 + size;
    }
    if (pipelined && *err == 0 && pos != nwords) *err = 1;  // words after the last block
    *nblk_out = k;
}

void check_launch() { DSS_HIP(hipGetLastError()); }

}  // namespace

// ------------------------------------------------- allocation-free varint
// dss_varint_write_ranges_noalloc keeps its scratch (per-value byte lengths
// and offsets, the scan's block sums and total, the range counts and ends)
// in one fixed device array instead of hipMalloc'd buffers.  Offsets in u64
// words, each region 256-byte aligned.
constexpr uint64_t kNaMaxValues = DSS_NOALLOC_VARINT_MAX_VALUES;
constexpr uint64_t kNaMaxRanges = DSS_NOALLOC_VARINT_MAX_RANGES;
constexpr uint64_t na_align(uint64_t w) { return (w + 31) / 32 * 32; }
constexpr uint64_t kNaScanWords = (kNaMaxValues + 1 + 2047) / 2048 + 1;  // scan_scratch_words(n + 1), n <= max
constexpr uint64_t kNaL = 0;
constexpr uint64_t kNaO = kNaL + na_align(kNaMaxValues + 1);
constexpr uint64_t kNaSums = kNaO + na_align(kNaMaxValues + 1);
constexpr uint64_t kNaTot = kNaSums + na_align(kNaScanWords);
constexpr uint64_t kNaCounts = kNaTot + 32;
constexpr uint64_t kNaEnds = kNaCounts + na_align(kNaMaxRanges);
constexpr uint64_t kNaWords = kNaEnds + na_align(kNaMaxRanges + 1);

// One copy per GPU, for the life of the process.  Host code must not pass
// this name to a kernel (on the host it is a stand-in, not a device
// address): noalloc_scratch() resolves each device's address once.
static __device__ __align__(256) uint64_t scratch[kNaWords];

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mutex;
static uint64_t* g_na_ptr[kNaMaxDevices];

// The device address of `scratch` on `device`, resolved by
// hipGetSymbolAddress the first time with that device current.
static uint64_t* noalloc_scratch(int device) {
    DSS_REQUIRE(device >= 0 && device < kNaMaxDevices, DSS_ERANGE, "device index beyond the supported range");
    std::lock_guard<std::mutex> g(g_na_mutex);
    if (!g_na_ptr[device]) {
        int cur = 0;
        DSS_HIP(hipGetDevice(&cur));
        if (cur != device) DSS_HIP(hipSetDevice(device));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr)
            throw Error{DSS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        g_na_ptr[device] = static_cast<uint64_t*>(ptr);
    }
    return g_na_ptr[device];
}

// ends[r] = O[first value of range r] for r < nr, ends[nr] = O[n]: one block,
// four ranges per thread (nr <= 4096), an LDS scan of the counts.
__global__ void __launch_bounds__(1024) k_range_ends(const uint64_t* __restrict__ counts, uint32_t nr,
                                                     const uint64_t* __restrict__ O, uint64_t* __restrict__ ends) {
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    uint64_t c[4], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t r = 4 * t + k;
        c[k] = r < nr ? counts[r] : 0;
        sum += c[k];
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
        const uint64_t x = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint64_t run = t ? part[t - 1] : 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t r = 4 * t + k;
        if (r < nr) ends[r] = O[run];
        run += c[k];
    }
    if (t == 1023) ends[nr] = O[part[1023]];
}

// ----------------------------------------------------------- entry points
static void varint_write(dss_codec_s* c, const uint64_t* values, const uint64_t* counts, uint32_t nr, uint8_t* out,
                         uint64_t out_cap, uint64_t* range_bytes, uint64_t* total_bytes) {
    hipStream_t st = c->stream;
    uint64_t n = 0;
    std::vector<uint64_t> idx;
    idx.reserve(nr + 1);
    for (uint32_t r = 0; r < nr; ++r) {
        idx.push_back(n);
        n += counts[r];
    }
    idx.push_back(n);
    c->lenv.reserve((n + 1) * 8);
    c->pos.reserve((n + 1) * 8);
    uint64_t* L = c->lenv.as<uint64_t>();
    uint64_t* O = c->pos.as<uint64_t>();
    {
        KScope ks(c->timer, st, "varint_len", n * 16.0);
        hipLaunchKernelGGL(k_varint_len, dim3(stride_grid(n + 1, 256)), dim3(256), 0, st, values, n, L);
    }
    check_launch();
    uint64_t total = 0;
    {
        KScope ks(c->timer, st, "varint_scan", n * 16.0);
        exclusive_scan_u64(L, O, n + 1, st, c->pp, nullptr);
    }
    DSS_HIP(hipMemcpyAsync(&total, O + n, 8, hipMemcpyDeviceToHost, st));
    c->meta.reserve(idx.size() * 16);
    uint64_t* di = c->meta.as<uint64_t>();
    uint64_t* dv = di + idx.size();
    DSS_HIP(hipMemcpyAsync(di, idx.data(), idx.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gather_u64, dim3(((uint32_t)idx.size() + 255) / 256), dim3(256), 0, st, O, di,
                       (uint32_t)idx.size(), dv);
    check_launch();
    std::vector<uint64_t> at(idx.size());
    DSS_HIP(hipMemcpyAsync(at.data(), dv, at.size() * 8, hipMemcpyDeviceToHost, st));
    DSS_HIP(hipStreamSynchronize(st));
    for (uint32_t r = 0; r < nr; ++r) range_bytes[r] = at[r + 1] - at[r];
    *total_bytes = total;
    DSS_REQUIRE(total <= out_cap, DSS_ERANGE,
                "varint output needs " + std::to_string(total) + " bytes, capacity " + std::to_string(out_cap));
    if (n) {
        KScope ks(c->timer, st, "varint_put", n * 16.0 + total);
        hipLaunchKernelGGL(k_varint_put, dim3(stride_grid(n, 256)), dim3(256), 0, st, values, n, O, out);
    }
    check_launch();
}

static void varint_read(dss_codec_s* c, const uint8_t* bytes, uint64_t nbytes, const uint64_t* counts, uint32_t nr,
                        uint64_t* values) {
    hipStream_t st = c->stream;
    uint64_t n = 0;
    for (uint32_t r = 0; r < nr; ++r) n += counts[r];
    if (nbytes == 0) {
        DSS_REQUIRE(n == 0, DSS_EINVAL, "varint input is empty but counts ask for values");
        return;
    }
    const uint64_t nblk = (nbytes + kVThreads * kVBytes - 1) / (kVThreads * kVBytes);
    DSS_REQUIRE(nblk < (1ull << 31), DSS_ERANGE, "varint input too large");
    c->lenv.reserve((nblk + 1) * 8);
    c->pos.reserve((nblk + 1) * 8);
    c->flag.reserve(8);
    uint64_t* cnt = c->lenv.as<uint64_t>();
    uint64_t* base = c->pos.as<uint64_t>();
    uint32_t* err = c->flag.as<uint32_t>();
    DSS_HIP(hipMemsetAsync(err, 0, 4, st));
    uint64_t total = 0;
    {
        KScope ks(c->timer, st, "varint_count", (double)nbytes + nblk * 16.0);
        hipLaunchKernelGGL(k_vterm_count, dim3((uint32_t)nblk), dim3(kVThreads), 0, st, bytes, nbytes, cnt);
        exclusive_scan_u64(cnt, base, nblk, st, c->pp, nullptr);
    }
    check_launch();
    DSS_HIP(hipMemcpyAsync(&total, c->pp.scan_tot.p, 8, hipMemcpyDeviceToHost, st));
    uint8_t last = 0;
    DSS_HIP(hipMemcpyAsync(&last, bytes + nbytes - 1, 1, hipMemcpyDeviceToHost, st));
    DSS_HIP(hipStreamSynchronize(st));
    DSS_REQUIRE(!(last & 0x80), DSS_EINVAL, "varint input ends inside a value");
    DSS_REQUIRE(total == n, DSS_EINVAL,
                "varint input holds " + std::to_string(total) + " values, counts ask for " + std::to_string(n));
    {
        KScope ks(c->timer, st, "varint_decode", (double)nbytes + n * 8.0);
        hipLaunchKernelGGL(k_varint_decode, dim3((uint32_t)nblk), dim3(kVThreads), 0, st, bytes, nbytes, base,
                           values, err);
    }
    check_launch();
    uint32_t e = 0;
    DSS_HIP(hipMemcpyAsync(&e, err, 4, hipMemcpyDeviceToHost, st));
    DSS_HIP(hipStreamSynchronize(st));
    DSS_REQUIRE(e == 0, DSS_EINVAL, "varint overflow (a value longer than 64 bits)");
}

static void golomb_encode(dss_codec_s* c, int layout, const uint64_t* values, const uint64_t* counts,
                          const uint64_t* bs, uint32_t nr, uint64_t* out, uint64_t out_cap, uint64_t* range_words,
                          uint64_t* total_words) {
    DSS_REQUIRE(layout == DSS_GOLOMB_BLOCKS || layout == DSS_GOLOMB_PIPELINED, DSS_EINVAL, "unknown Golomb layout");
    hipStream_t st = c->stream;
    std::vector<GRange> R(nr);
    std::vector<uint64_t> idx;
    uint64_t n = 0;
    for (uint32_t r = 0; r < nr; ++r) {
        DSS_REQUIRE(bs[r] >= 1, DSS_EINVAL, "Golomb parameter b must be >= 1");
        GRange& g = R[r];
        g.vbegin = n;
        g.cnt = counts[r];
        g.b = bs[r];
        g.log2b = ceil_log2(g.b);
        g.mlv = (g.log2b == 64 ? 0 : (1ull << g.log2b)) - g.b;
        g.layout = (uint32_t)layout;
        idx.push_back(n);
        n += g.cnt;
        idx.push_back(n);
    }
    c->meta.reserve(nr * sizeof(GRange) + idx.size() * 16 + 16);
    GRange* dR = c->meta.as<GRange>();
    uint64_t* di = reinterpret_cast<uint64_t*>(dR + nr);
    uint64_t* dv = di + idx.size();
    c->lenv.reserve((n + 1) * 8);
    c->pos.reserve((n + 1) * 8);
    uint64_t* L = c->lenv.as<uint64_t>();
    uint64_t* G = c->pos.as<uint64_t>();
    DSS_HIP(hipMemcpyAsync(dR, R.data(), nr * sizeof(GRange), hipMemcpyHostToDevice, st));
    if (!idx.empty()) DSS_HIP(hipMemcpyAsync(di, idx.data(), idx.size() * 8, hipMemcpyHostToDevice, st));
    {
        KScope ks(c->timer, st, "golomb_len", n * 16.0);
        hipLaunchKernelGGL(k_golomb_len, dim3(stride_grid(n + 1, 256)), dim3(256), 0, st, values, n, dR, nr, L);
    }
    check_launch();
    {
        KScope ks(c->timer, st, "golomb_scan", n * 16.0);
        exclusive_scan_u64(L, G, n + 1, st, c->pp, nullptr);
    }
    std::vector<uint64_t> at(idx.size());
    if (!idx.empty()) {
        hipLaunchKernelGGL(k_gather_u64, dim3(((uint32_t)idx.size() + 255) / 256), dim3(256), 0, st, G, di,
                           (uint32_t)idx.size(), dv);
        check_launch();
        DSS_HIP(hipMemcpyAsync(at.data(), dv, at.size() * 8, hipMemcpyDeviceToHost, st));
    }
    DSS_HIP(hipStreamSynchronize(st));
    uint64_t w = 0;
    for (uint32_t r = 0; r < nr; ++r) {
        GRange& g = R[r];
        g.nbits = at[2 * r + 1] - at[2 * r];
        const uint64_t words = (g.nbits + 63) / 64;
        g.hdr0 = w;
        uint64_t size;
        if (layout == DSS_GOLOMB_BLOCKS) {
            size = g.cnt ? 3 + words : 0;
            g.bitword0 = w + 3;
        } else {
            size = 1 + (g.cnt ? 1 + words : 0);
            g.bitword0 = w + 2;
        }
        range_words[r] = size;
        w += size;
    }
    *total_words = w;
    DSS_REQUIRE(w <= out_cap, DSS_ERANGE,
                "Golomb output needs " + std::to_string(w) + " words, capacity " + std::to_string(out_cap));
    if (w == 0) return;
    DSS_HIP(hipMemcpyAsync(dR, R.data(), nr * sizeof(GRange), hipMemcpyHostToDevice, st));
    DSS_HIP(hipMemsetAsync(out, 0, w * 8, st));
    {
        KScope ks(c->timer, st, "golomb_put", n * 24.0 + w * 8.0);
        hipLaunchKernelGGL(k_golomb_head, dim3((nr + 255) / 256), dim3(256), 0, st, values, dR, nr, out);
        if (n) hipLaunchKernelGGL(k_golomb_put, dim3(stride_grid(n, 256)), dim3(256), 0, st, values, n, dR, nr, G, out);
    }
    check_launch();
}

static void golomb_decode(dss_codec_s* c, int layout, const uint64_t* words, uint64_t nwords, uint32_t max_blocks,
                          uint64_t b_pipe, uint64_t* values, uint64_t cap, uint64_t* block_counts,
                          uint32_t* num_blocks) {
    DSS_REQUIRE(layout == DSS_GOLOMB_BLOCKS || layout == DSS_GOLOMB_PIPELINED, DSS_EINVAL, "unknown Golomb layout");
    DSS_REQUIRE(layout == DSS_GOLOMB_BLOCKS || b_pipe >= 1, DSS_EINVAL, "Golomb parameter b must be >= 1");
    hipStream_t st = c->stream;
    *num_blocks = 0;
    if (max_blocks == 0 || (nwords == 0 && layout == DSS_GOLOMB_BLOCKS)) return;
    c->meta.reserve((uint64_t)max_blocks * sizeof(GBlock) + 64);
    c->flag.reserve(16);
    GBlock* dB = c->meta.as<GBlock>();
    uint32_t* dflag = c->flag.as<uint32_t>();  // [0] blocks, [1] error, [2] changed
    DSS_HIP(hipMemsetAsync(dflag, 0, 16, st));
    hipLaunchKernelGGL(k_golomb_parse, dim3(1), dim3(1), 0, st, words, nwords, layout, max_blocks, b_pipe, dB, dflag,
                       dflag + 1);
    check_launch();
    uint32_t hf[2];
    DSS_HIP(hipMemcpyAsync(hf, dflag, 8, hipMemcpyDeviceToHost, st));
    DSS_HIP(hipStreamSynchronize(st));
    DSS_REQUIRE(hf[1] != 2, DSS_EINVAL, "Golomb input holds more blocks than allowed");
    DSS_REQUIRE(hf[1] == 0, DSS_EINVAL, "malformed Golomb input (block sizes do not fit the words)");
    const uint32_t nb = hf[0];
    std::vector<GBlock> B(nb);
    if (nb) DSS_HIP(hipMemcpyAsync(B.data(), dB, nb * sizeof(GBlock), hipMemcpyDeviceToHost, st));
    DSS_HIP(hipStreamSynchronize(st));
    uint64_t nch = 0;
    for (auto& b : B) {
        DSS_REQUIRE(b.b >= 1, DSS_EINVAL, "Golomb parameter b must be >= 1");
        b.log2b = ceil_log2(b.b);
        b.mlv = (b.log2b == 64 ? 0 : (1ull << b.log2b)) - b.b;
        b.chunk0 = nch;
        b.nchunks = (b.nbits + kChunkBits - 1) / kChunkBits;
        nch += b.nchunks;
    }
    *num_blocks = nb;
    if (nb == 0) return;
    DSS_HIP(hipMemcpyAsync(dB, B.data(), nb * sizeof(GBlock), hipMemcpyHostToDevice, st));
    c->bitmap.reserve(nwords * 16 + 16);
    c->spec.reserve(nch * 8 + 8);
    c->xa.reserve(nch * 8 + 8);
    c->xb.reserve(nch * 8 + 8);
    c->slots.reserve((nch + nb + 1) * 16);
    c->lenv.reserve(((uint64_t)nb + 1) * 8);
    uint64_t* S = c->bitmap.as<uint64_t>();
    uint64_t* T = S + nwords;
    uint64_t* E = c->spec.as<uint64_t>();
    uint64_t* Xa = c->xa.as<uint64_t>();
    uint64_t* Xb = c->xb.as<uint64_t>();
    uint64_t* A = c->slots.as<uint64_t>();
    uint64_t* Ap = A + (nch + nb + 1);
    uint64_t* bstart = c->lenv.as<uint64_t>();
    const uint32_t gch = (uint32_t)std::max<uint64_t>((nch + 255) / 256, 1);
    if (nch) {
        {
            KScope ks(c->timer, st, "golomb_spec", nwords * 16.0);
            DSS_HIP(hipMemsetAsync(S, 0, nwords * 8, st));
            hipLaunchKernelGGL(k_golomb_spec, dim3(gch), dim3(256), 0, st, words, dB, nb, nch, S, E);
        }
        check_launch();
        DSS_HIP(hipMemcpyAsync(Xa, E, nch * 8, hipMemcpyDeviceToDevice, st));
        // fixed point of the chunk exits; chunk k of a stream is exact after k rounds
        for (uint64_t it = 0;; ++it) {
            DSS_REQUIRE(it <= nch + 1, DSS_EHIP, "Golomb decode did not converge");
            DSS_HIP(hipMemsetAsync(dflag + 2, 0, 4, st));
            {
                KScope ks(c->timer, st, "golomb_sync", nch * 24.0);
                hipLaunchKernelGGL(k_golomb_sync, dim3(gch), dim3(256), 0, st, words, dB, nb, nch, S, E, Xa, Xb,
                                   dflag + 2);
            }
            check_launch();
            uint32_t ch = 0;
            DSS_HIP(hipMemcpyAsync(&ch, dflag + 2, 4, hipMemcpyDeviceToHost, st));
            DSS_HIP(hipStreamSynchronize(st));
            std::swap(Xa, Xb);
            if (!ch) break;
        }
        KScope ks(c->timer, st, "golomb_starts", nwords * 16.0);
        DSS_HIP(hipMemcpyAsync(T, S, nwords * 8, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_golomb_starts, dim3(gch), dim3(256), 0, st, words, dB, nb, nch, S, Xa, T);
    }
    const uint32_t gcb = (uint32_t)std::max<uint64_t>((std::max<uint64_t>(nch, nb) + 255) / 256, 1);
    uint64_t total = 0;
    {
        KScope ks(c->timer, st, "golomb_count", nwords * 8.0 + (nch + nb) * 24.0);
        hipLaunchKernelGGL(k_golomb_count, dim3(gcb), dim3(256), 0, st, dB, nb, nch, T, A);
        exclusive_scan_u64(A, Ap, nch + nb + 1, st, c->pp, nullptr);
    }
    check_launch();
    DSS_HIP(hipMemcpyAsync(&total, c->pp.scan_tot.p, 8, hipMemcpyDeviceToHost, st));
    DSS_HIP(hipStreamSynchronize(st));
    DSS_REQUIRE(total <= cap, DSS_ERANGE,
                "Golomb input decodes to " + std::to_string(total) + " values, capacity " + std::to_string(cap));
    {
        KScope ks(c->timer, st, "golomb_decode", nwords * 16.0 + total * 8.0);
        hipLaunchKernelGGL(k_golomb_decode, dim3(gcb), dim3(256), 0, st, words, dB, nb, nch, T, Ap, values, bstart);
    }
    check_launch();
    if (total) {
        c->pos.reserve(total * 8);
        uint64_t* ex = c->pos.as<uint64_t>();
        KScope ks(c->timer, st, "golomb_sum", total * 40.0);
        exclusive_scan_u64(values, ex, total, st, c->pp, nullptr);
        hipLaunchKernelGGL(k_golomb_sum, dim3(stride_grid(total, 256)), dim3(256), 0, st, values, ex, total, bstart, nb);
    }
    check_launch();
    std::vector<uint64_t> bs(nb + 1);
    DSS_HIP(hipMemcpyAsync(bs.data(), bstart, (nb + 1) * 8, hipMemcpyDeviceToHost, st));
    DSS_HIP(hipStreamSynchronize(st));
    for (uint32_t j = 0; j < nb; ++j) block_counts[j] = bs[j + 1] - bs[j];
}

}  // namespace dss

using namespace dss;

extern "C" {

int dss_codec_create(dss_codec* out) {
    return guarded([&] {
        DSS_REQUIRE(out, DSS_EINVAL, "NULL argument");
        dss_codec_s* c = new dss_codec_s();
        hipError_t e = hipGetDevice(&c->device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw Error{DSS_EHIP, std::string("codec stream: ") + hipGetErrorString(e)};
        }
        *out = c;
    });
}

int dss_codec_destroy(dss_codec c) {
    return guarded([&] {
        if (!c) return;
        (void)hipStreamSynchronize(c->stream);
        hipStream_t st = c->stream;
        delete c;
        (void)hipStreamDestroy(st);
    });
}

int dss_codec_sync(dss_codec c) {
    return guarded([&] {
        DSS_REQUIRE(c, DSS_EINVAL, "NULL codec");
        DSS_HIP(hipStreamSynchronize(c->stream));
    });
}

int dss_codec_set_timing(dss_codec c, int enable) {
    return guarded([&] {
        DSS_REQUIRE(c, DSS_EINVAL, "NULL codec");
        c->timer.enabled = enable != 0;
    });
}

int dss_codec_kernel_stats(dss_codec c, dss_kernel_stat_t* stats, uint32_t cap, uint32_t* count, int reset) {
    return guarded([&] {
        DSS_REQUIRE(c && count, DSS_EINVAL, "NULL argument");
        DSS_HIP(hipStreamSynchronize(c->stream));
        c->timer.collect();
        uint32_t k = 0;
        for (auto& s : c->timer.stats) {
            if (k < cap && stats) {
                memset(&stats[k], 0, sizeof(dss_kernel_stat_t));
                strncpy(stats[k].name, s.name.c_str(), sizeof(stats[k].name) - 1);
                stats[k].launches = s.launches;
                stats[k].total_ms = s.ms;
                stats[k].alg_bytes = s.bytes;
            }
            ++k;
        }
        *count = k;
        if (reset) c->timer.stats.clear();
    });
}

int dss_golomb_b(uint64_t max_value, uint64_t num_values, uint64_t* b) {
    return guarded([&] {
        DSS_REQUIRE(b, DSS_EINVAL, "NULL argument");
        if (num_values == 0 || max_value / num_values < 1) {
            *b = 1;
            return;
        }
        const uint64_t k = (uint64_t)std::floor(std::log2((double)(max_value / num_values)));
        *b = 1ull << k;
    });
}

int dss_varint_write_ranges_dev(dss_codec c, const uint64_t* values, const uint64_t* counts, uint32_t num_ranges,
                                uint8_t* out, uint64_t out_cap, uint64_t* range_bytes, uint64_t* total_bytes) {
    return guarded([&] {
        DSS_REQUIRE(c && (counts || !num_ranges) && (range_bytes || !num_ranges) && total_bytes, DSS_EINVAL,
                    "NULL argument");
        varint_write(c, values, counts, num_ranges, out, out_cap, range_bytes, total_bytes);
    });
}

int dss_varint_write_ranges_noalloc(dss_codec c, const uint64_t* values, const uint64_t* counts,
                                    uint32_t num_ranges, uint8_t* out, uint64_t out_cap, uint64_t* range_bytes,
                                    uint64_t* total_bytes) {
    return guarded([&] {
        DSS_REQUIRE(c && (counts || !num_ranges) && (range_bytes || !num_ranges) && total_bytes, DSS_EINVAL,
                    "NULL argument");
        DSS_REQUIRE(num_ranges <= kNaMaxRanges, DSS_ERANGE,
                    "dss_varint_write_ranges_noalloc takes at most " + std::to_string(kNaMaxRanges) +
                        " ranges, got " + std::to_string(num_ranges));
        uint64_t n = 0;
        for (uint32_t r = 0; r < num_ranges; ++r) {
            n += counts[r];
            DSS_REQUIRE(n <= kNaMaxValues, DSS_ERANGE,
                        "dss_varint_write_ranges_noalloc takes at most " + std::to_string(kNaMaxValues) +
                            " values per call");
        }
        DSS_REQUIRE(n == 0 || values, DSS_EINVAL, "NULL values");
        uint64_t* S = noalloc_scratch(c->device);
        uint64_t* L = S + kNaL;
        uint64_t* O = S + kNaO;
        uint64_t* ends = S + kNaEnds;
        const hipStream_t st = c->stream;
        // byte length per value (and 0 after the last), offsets by the scan
        hipLaunchKernelGGL(k_varint_len, dim3(stride_grid(n + 1, 256)), dim3(256), 0, st, values, n, L);
        check_launch();
        exclusive_scan_u64_scratch(L, O, n + 1, st, S + kNaSums, S + kNaTot);
        uint64_t last = 0;
        if (num_ranges) {
            DSS_HIP(hipMemcpyAsync(S + kNaCounts, counts, num_ranges * 8, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_range_ends, dim3(1), dim3(1024), 0, st, S + kNaCounts, num_ranges, O, ends);
            check_launch();
            DSS_HIP(hipMemcpyAsync(range_bytes, ends, num_ranges * 8, hipMemcpyDeviceToHost, st));
        }
        DSS_HIP(hipMemcpyAsync(&last, O + n, 8, hipMemcpyDeviceToHost, st));
        DSS_HIP(hipStreamSynchronize(st));
        for (uint32_t r = 0; r < num_ranges; ++r)  // range ends -> range sizes, in place
            range_bytes[r] = (r + 1 < num_ranges ? range_bytes[r + 1] : last) - range_bytes[r];
        *total_bytes = last;
        DSS_REQUIRE(last <= out_cap, DSS_ERANGE,
                    "varint output needs " + std::to_string(last) + " bytes, capacity " + std::to_string(out_cap));
        if (n) hipLaunchKernelGGL(k_varint_put, dim3(stride_grid(n, 256)), dim3(256), 0, st, values, n, O, out);
        check_launch();
    });
}

int dss_varint_read_ranges_dev(dss_codec c, const uint8_t* bytes, uint64_t nbytes, const uint64_t* counts,
                               uint32_t num_ranges, uint64_t* values) {
    return guarded([&] {
        DSS_REQUIRE(c && (counts || !num_ranges), DSS_EINVAL, "NULL argument");
        varint_read(c, bytes, nbytes, counts, num_ranges, values);
    });
}

int dss_varint_write_ranges(dss_codec c, const uint64_t* values, const uint64_t* counts, uint32_t num_ranges,
                            uint8_t* out, uint64_t out_cap, uint64_t* range_bytes, uint64_t* total_bytes) {
    return guarded([&] {
        DSS_REQUIRE(c && (counts || !num_ranges) && (range_bytes || !num_ranges) && total_bytes, DSS_EINVAL,
                    "NULL argument");
        uint64_t n = 0;
        for (uint32_t r = 0; r < num_ranges; ++r) n += counts[r];
        DSS_REQUIRE(values || !n, DSS_EINVAL, "NULL values");
        c->in.reserve(n * 8 + 8);
        c->out.reserve(n * 10 + 8);
        if (n) DSS_HIP(hipMemcpyAsync(c->in.p, values, n * 8, hipMemcpyHostToDevice, c->stream));
        varint_write(c, c->in.as<uint64_t>(), counts, num_ranges, c->out.as<uint8_t>(), n * 10, range_bytes,
                     total_bytes);
        DSS_REQUIRE(*total_bytes <= out_cap, DSS_ERANGE,
                    "varint output needs " + std::to_string(*total_bytes) + " bytes, capacity " +
                        std::to_string(out_cap));
        if (*total_bytes) {
            DSS_REQUIRE(out, DSS_EINVAL, "NULL output");
            DSS_HIP(hipMemcpyAsync(out, c->out.p, *total_bytes, hipMemcpyDeviceToHost, c->stream));
        }
        DSS_HIP(hipStreamSynchronize(c->stream));
    });
}

int dss_varint_read_ranges(dss_codec c, const uint8_t* bytes, uint64_t nbytes, const uint64_t* counts,
                           uint32_t num_ranges, uint64_t* values) {
    return guarded([&] {
        DSS_REQUIRE(c && (counts || !num_ranges) && (bytes || !nbytes), DSS_EINVAL, "NULL argument");
        uint64_t n = 0;
        for (uint32_t r = 0; r < num_ranges; ++r) n += counts[r];
        DSS_REQUIRE(values || !n, DSS_EINVAL, "NULL values");
        c->in.reserve(nbytes + 8);
        c->out.reserve(n * 8 + 8);
        if (nbytes) DSS_HIP(hipMemcpyAsync(c->in.p, bytes, nbytes, hipMemcpyHostToDevice, c->stream));
        varint_read(c, c->in.as<uint8_t>(), nbytes, counts, num_ranges, c->out.as<uint64_t>());
        if (n) DSS_HIP(hipMemcpyAsync(values, c->out.p, n * 8, hipMemcpyDeviceToHost, c->stream));
        DSS_HIP(hipStreamSynchronize(c->stream));
    });
}

int dss_golomb_encode_dev(dss_codec c, int layout, const uint64_t* values, const uint64_t* counts, const uint64_t* b,
                          uint32_t num_ranges, uint64_t* out, uint64_t out_cap, uint64_t* range_words,
                          uint64_t* total_words) {
    return guarded([&] {
        DSS_REQUIRE(c && (counts || !num_ranges) && (b || !num_ranges) && (range_words || !num_ranges) && total_words,
                    DSS_EINVAL, "NULL argument");
        golomb_encode(c, layout, values, counts, b, num_ranges, out, out_cap, range_words, total_words);
    });
}

int dss_golomb_decode_dev(dss_codec c, int layout, const uint64_t* words, uint64_t nwords, uint32_t max_blocks,
                          uint64_t b_pipelined, uint64_t* values, uint64_t cap, uint64_t* block_counts,
                          uint32_t* num_blocks) {
    return guarded([&] {
        DSS_REQUIRE(c && (block_counts || !max_blocks) && num_blocks, DSS_EINVAL, "NULL argument");
        golomb_decode(c, layout, words, nwords, max_blocks, b_pipelined, values, cap, block_counts, num_blocks);
    });
}

int dss_golomb_encode(dss_codec c, int layout, const uint64_t* values, const uint64_t* counts, const uint64_t* b,
                      uint32_t num_ranges, uint64_t* out, uint64_t out_cap, uint64_t* range_words,
                      uint64_t* total_words) {
    return guarded([&] {
        DSS_REQUIRE(c && (counts || !num_ranges) && (b || !num_ranges) && (range_words || !num_ranges) && total_words,
                    DSS_EINVAL, "NULL argument");
        uint64_t n = 0;
        for (uint32_t r = 0; r < num_ranges; ++r) n += counts[r];
        DSS_REQUIRE(values || !n, DSS_EINVAL, "NULL values");
        c->in.reserve(n * 8 + 8);
        if (n) DSS_HIP(hipMemcpyAsync(c->in.p, values, n * 8, hipMemcpyHostToDevice, c->stream));
        // the output size is known after the code lengths: size it first
        std::vector<uint64_t> rw(num_ranges);
        uint64_t need = 0;
        try {
            golomb_encode(c, layout, c->in.as<uint64_t>(), counts, b, num_ranges, nullptr, 0, rw.data(), &need);
        } catch (const Error& e) {
            if (e.code != DSS_ERANGE) throw;
        }
        *total_words = need;
        for (uint32_t r = 0; r < num_ranges; ++r) range_words[r] = rw[r];
        DSS_REQUIRE(need <= out_cap, DSS_ERANGE,
                    "Golomb output needs " + std::to_string(need) + " words, capacity " + std::to_string(out_cap));
        if (need == 0) return;
        DSS_REQUIRE(out, DSS_EINVAL, "NULL output");
        c->out.reserve(need * 8);
        golomb_encode(c, layout, c->in.as<uint64_t>(), counts, b, num_ranges, c->out.as<uint64_t>(), need, rw.data(),
                      &need);
        DSS_HIP(hipMemcpyAsync(out, c->out.p, need * 8, hipMemcpyDeviceToHost, c->stream));
        DSS_HIP(hipStreamSynchronize(c->stream));
    });
}

int dss_golomb_decode(dss_codec c, int layout, const uint64_t* words, uint64_t nwords, uint32_t max_blocks,
                      uint64_t b_pipelined, uint64_t* values, uint64_t cap, uint64_t* block_counts,
                      uint32_t* num_blocks) {
    return guarded([&] {
        DSS_REQUIRE(c && (words || !nwords) && (block_counts || !max_blocks) && num_blocks, DSS_EINVAL,
                    "NULL argument");
        c->in.reserve(nwords * 8 + 8);
        c->out.reserve(cap * 8 + 8);
        if (nwords) DSS_HIP(hipMemcpyAsync(c->in.p, words, nwords * 8, hipMemcpyHostToDevice, c->stream));
        golomb_decode(c, layout, c->in.as<uint64_t>(), nwords, max_blocks, b_pipelined, c->out.as<uint64_t>(), cap,
                      block_counts, num_blocks);
        uint64_t total = 0;
        for (uint32_t j = 0; j < *num_blocks; ++j) total += block_counts[j];
        DSS_REQUIRE(values || !total, DSS_EINVAL, "NULL values");
        if (total) DSS_HIP(hipMemcpyAsync(values, c->out.p, total * 8, hipMemcpyDeviceToHost, c->stream));
        DSS_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
