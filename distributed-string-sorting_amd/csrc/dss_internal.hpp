// Internal declarations shared by the HIP translation units of libdss_gpu.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "dss_gpu.h"

namespace dss {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define DSS_HIP(call)                                                               \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            throw ::dss::Error{DSS_EHIP, std::string(#call) + ": " +                 \
                                            hipGetErrorString(e_)};                 \
    } while (0)

#define DSS_REQUIRE(cond, code, msg)                                                \
    do {                                                                            \
        if (!(cond)) throw ::dss::Error{(code), (msg)};                             \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return DSS_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return DSS_EHIP;
    }
}

// ------------------------------------------------------------- device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grows (never shrinks) to at least n bytes; contents are not kept.
    void reserve(size_t n) {
        if (n <= bytes) return;
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{DSS_ENOMEM, "hipMalloc(" + std::to_string(n) + " B) failed: " +
                                        hipGetErrorString(e)};
        }
        bytes = n;
    }
    // Grows to at least n bytes keeping the current contents (copied on st).
    void reserve_keep(size_t n, hipStream_t st) {
        if (n <= bytes) return;
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, n);
        if (e != hipSuccess)
            throw Error{DSS_ENOMEM, "hipMalloc(" + std::to_string(n) + " B) failed: " + hipGetErrorString(e)};
        if (p && bytes) {
            e = hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                (void)hipFree(q);
                throw Error{DSS_EHIP, std::string("reserve_keep copy: ") + hipGetErrorString(e)};
            }
        }
        release();
        p = q;
        bytes = n;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Grid of a grid-stride kernel over `items` work items with `per_block`
// items per block.  A dispatch covers fewer than 2^32 work items, so kernels
// whose item count can reach that (bytes or lanes of a large arena) loop over
// the items with a grid capped at kMaxGridBlocks.
constexpr uint32_t kMaxGridBlocks = 1u << 20;
inline uint32_t stride_grid(uint64_t items, uint32_t per_block) {
    const uint64_t b = (items + per_block - 1) / per_block;
    return (uint32_t)(b < 1 ? 1 : (b > kMaxGridBlocks ? kMaxGridBlocks : b));
}

// ------------------------------------------------------------- kernel timing
struct KernelTimer {
    bool enabled = false;
    struct Rec {
        std::string name;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Rec> pending;
    struct Stat {
        std::string name;
        uint64_t launches = 0;
        double ms = 0, bytes = 0;
    };
    std::vector<Stat> stats;
    std::vector<hipEvent_t> pool;

    hipEvent_t get_event();
    void begin(hipStream_t st, const char* name, double alg_bytes, Rec& r);
    void end(hipStream_t st, Rec& r);
    void collect();  // after a stream sync
    // Replaces the algorithmic bytes of the most recent record of `name`
    // (for byte counts known only after the launch).
    void amend_last(const char* name, double bytes) {
        for (auto it = pending.rbegin(); it != pending.rend(); ++it)
            if (it->name == name) {
                it->bytes = bytes;
                return;
            }
    }
    ~KernelTimer();
};

// Scoped helper: records events around one kernel launch when enabled.
struct KScope {
    KernelTimer& t;
    hipStream_t st;
    KernelTimer::Rec r;
    bool on;
    KScope(KernelTimer& t_, hipStream_t st_, const char* name, double bytes)
        : t(t_), st(st_), on(t_.enabled) {
        if (on) t.begin(st, name, bytes, r);
    }
    ~KScope() {
        if (on) t.end(st, r);
    }
};

// ------------------------------------------------------------- read-backs
// Small device-to-host read-backs between launches (counters that size the
// next launches).  A copy into pageable memory plus a stream synchronisation
// leaves the device idle for tens of microseconds; instead a one-wave kernel
// copies the words into pinned, coherent host memory and then publishes a
// sequence number with a system-scope release store, and the host spins on
// that number.  Every 1024 polls the host queries the stream, so a failed or
// drained stream without the number raises an error instead of spinning.
struct Mailbox {
    static constexpr size_t kBytes = 4096;  // data bytes per read-back
    uint64_t* flag = nullptr;               // pinned: [0] sequence, data from byte 64
    uint64_t seq = 0;
    Mailbox() = default;
    Mailbox(const Mailbox&) = delete;
    Mailbox& operator=(const Mailbox&) = delete;
    ~Mailbox();
    // Copies `bytes` (a multiple of 4) from device memory `dev` to `dst` once
    // every earlier launch on `st` has finished; more than kBytes go through a
    // copy and a stream synchronisation.
    void fetch(hipStream_t st, void* dst, const void* dev, size_t bytes);
    // Returns once every earlier launch on `st` has finished.
    void wait(hipStream_t st) { fetch(st, nullptr, nullptr, 0); }
};

// Sets n bytes at device address p to v with one small kernel on st (for
// counters and short arrays between launches: cheaper to issue than a
// runtime memset).
void fill_small(hipStream_t st, void* p, uint8_t v, size_t n);

// ------------------------------------------------------------- sort scratch
// Key of the local sort's captured launch sequence: the capacities and
// parameters it was captured for and every buffer address it reads or writes.
struct SortGraphKey {
    uint64_t ncap;
    uint32_t pcap, large_min, wave_class_max, skip_lanes, skip_strings;
    const void* ptr[19];
    hipStream_t stream;
    SortGraphKey() : ncap(0), pcap(0), large_min(0), wave_class_max(0), skip_lanes(0), skip_strings(0), ptr{},
                     stream(nullptr) {}
    template <class... P>
    SortGraphKey(uint64_t n, uint32_t p, uint32_t lm, uint32_t wc, uint32_t sl, uint32_t ss, P... ps)
        : ncap(n), pcap(p), large_min(lm), wave_class_max(wc), skip_lanes(sl), skip_strings(ss), ptr{},
          stream(nullptr) {
        const void* a[] = {static_cast<const void*>(ps)...};
        static_assert(sizeof...(P) == 20, "19 buffers and the stream");
        for (int k = 0; k < 19; ++k) ptr[k] = a[k];
        stream = (hipStream_t)const_cast<void*>(a[19]);
    }
    bool operator==(const SortGraphKey& o) const {
        if (ncap != o.ncap || pcap != o.pcap || large_min != o.large_min || wave_class_max != o.wave_class_max ||
            skip_lanes != o.skip_lanes || skip_strings != o.skip_strings || stream != o.stream)
            return false;
        for (int k = 0; k < 19; ++k)
            if (ptr[k] != o.ptr[k]) return false;
        return true;
    }
};

// The two executable graphs of the device-driven local sort (passes; base
// cases) for one key.
struct SortGraph {
    hipGraphExec_t passes = nullptr, base = nullptr;
    SortGraphKey key;
    uint64_t last_use = 0;
    void reset() {
        if (passes) (void)hipGraphExecDestroy(passes);
        if (base) (void)hipGraphExecDestroy(base);
        passes = base = nullptr;
        key = SortGraphKey();
        last_use = 0;
    }
    SortGraph() = default;
    SortGraph(const SortGraph&) = delete;
    SortGraph& operator=(const SortGraph&) = delete;
    ~SortGraph() { reset(); }
};

// A set's captured graphs, one per size class (strings and PEs rounded up to
// powers of two), the least recently used replaced when all are taken.
struct SortGraphCache {
    static constexpr int kSlots = 6;
    SortGraph slot[kSlots];
    uint64_t clock = 0;
    SortGraph& find(const SortGraphKey& k, bool& hit) {
        SortGraph* lru = &slot[0];
        for (auto& g : slot) {
            if (g.passes && g.base && g.key == k) {
                hit = true;
                g.last_use = ++clock;
                return g;
            }
            if (g.last_use < lru->last_use) lru = &g;
        }
        hit = false;
        lru->reset();
        lru->last_use = ++clock;
        return *lru;
    }
    void reset() {
        for (auto& g : slot) g.reset();
    }
};

struct SortScratch {
    DevBuf keyA, keyB, pbA, pbB;
    DevBuf segA, segB;      // large segment lists
    DevBuf tiles;           // tile list
    DevBuf hist;            // per tile x 256 digit counters / offsets
    DevBuf csum;            // per tile chunk x 256 digit sums / offsets
    DevBuf small;           // small segment lists (all classes)
    DevBuf counters;        // device counters
    DevBuf offlen;          // packed (offset, length) per string for the base cases
    DevBuf digit;           // the radix digit map of the sort and the arena's byte values
    std::vector<uint64_t> small_off, small_cap;  // per class region in `small`
    uint64_t seg_cap = 0, tile_cap = 0;
    SortGraphCache graphs;  // the device-driven form's captured launch sequences
};

// ------------------------------------------------------------- pipeline state
struct PipeState {
    // prefix doubling (per position)
    bool has_dist = false;
    DevBuf dist;     // u32[n] distinguishing-prefix estimates (results)
    DevBuf rounds;   // u8[n]  rounds a position was a candidate in
    uint32_t pd_rounds = 0;
    DevBuf pd_cand, pd_next, pd_hash, pd_flags, pd_table, pd_tflag, pd_ctr, pd_slot, pd_list, pd_maybe, pd_dup,
        pd_bits, pd_head, pd_hmax, pd_hpos, pd_hstate;
    DevBuf pd_pkey, pd_pid, pd_phist, pd_pbound;  // partitioned duplicate detection
    // characters [32(stripes-1), depth) of strings hashed from a run head's
    // state: one packed row per position and the depth it was read for;
    // valid for the results of the last prefix doubling (stage_valid)
    DevBuf pd_stage, pd_sdepth;
    DevBuf copy_chunks;  // exchange copy: the string holding each 4 KiB chunk's first byte
    bool stage_valid = false;
    uint32_t stage_w = 0;  // words per kept row: 4 + ceil((depth % 32) / 8)
    // partition
    bool partitioned = false;
    std::vector<uint64_t> interval;  // p*p, [r*p+q] = strings PE r sends to PE q
    DevBuf interval_d;
    DevBuf samples, sample_rank, sample_sorted, splitters, wprefix, splitter_pos;
    // exchange (receive side, indexed by receive position)
    bool exchanged = false;
    int mode = 0;
    uint64_t recv_nbytes = 0;
    DevBuf recv_bytes;
    DevBuf recv_byte_off, recv_lcp, recv_src_pe, recv_src_pos, recv_view_off, recv_view_len,
        recv_from, recv_size, recv_stage;
    bool origin_valid = false;         // recv_src_pe / recv_src_pos filled (on demand)
    std::vector<uint64_t> recv_begin;  // p+1
    DevBuf layout;                                 // send_start / run offsets (u64)
    // merge (per receive-side output position)
    bool merged = false;
    bool merged_by_caller = false;  // merged output from dss_set_merged (no saved LCPs)
    bool perm_valid = false;        // perm_out holds the merged output's permutation
    DevBuf perm_out;                // u32 stringIndex[n], then u32 PEIndex[n] from perm_pe_off(n)
    DevBuf m_order, m_lcp, m_saved, m_order2, m_lcp2, m_keyc, m_keyc2, m_cp, m_work;
    // prefix decompression
    bool extended = false;
    uint64_t out_nbytes = 0;
    DevBuf out_bytes, out_off, out_psv, out_aux;
    // exact distinguishing prefixes of the merged output
    DevBuf exact, exact_glcp, exact_pos, exact_cnt;
    // generic scratch
    DevBuf scan_tmp, scan_tot;
    Mailbox mbox;  // counter read-backs on the owner's stream
};

// Device-wide exclusive scan of u64 values (out may alias in); returns the
// total when total_host != nullptr (synchronises the stream).
void exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t st,
                        PipeState& P, uint64_t* total_host);
// The same scan with caller-provided scratch: sums holds scan_scratch_words(n)
// u64 words, the total is written to the device word total_dev; no
// allocation and no read-back.
uint64_t scan_scratch_words(uint64_t n);
void exclusive_scan_u64_scratch(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t st,
                                uint64_t* sums, uint64_t* total_dev);

}  // namespace dss

// ------------------------------------------------------------------ the set
// Runtime parameters of the local sort (chosen per input regime).
struct SortParams {
    uint32_t large_min = 512;      // segments above go through the MSD passes
    uint32_t wave_class_max = 6;   // base-case classes handled one segment per wave
    uint32_t skip_lanes = 8;       // lanes per string in the prefix skip
    uint32_t skip_strings = 4;     // strings per lane group and round in the prefix skip
    uint32_t merge_fan = 2;        // runs merged per pass (2: the pairwise tree; 4, 8, 16: K-way tiles)
    uint32_t pass_mode = 0;        // local sort: 0 auto, 1 host-driven passes, 2 device-driven (graph)
    uint32_t merge_tile = 1024;    // outputs per tile of the pairwise merge (512, 1024, 2048)
    uint32_t digit_chars = 0;      // characters per radix digit: 0 auto (small alphabets), 1, or up to 2 / 3
};

struct dss_set_s {
    int device = 0;
    hipStream_t stream = nullptr;      // the stream all work is queued on
    hipStream_t own_stream = nullptr;  // the set's own stream (stream unless the caller set one)
    uint64_t n = 0;        // strings
    uint64_t nbytes = 0;   // arena bytes (incl. NULs)
    uint64_t slice_bytes = 0;  // bytes the PE split slices (0: nbytes); a text file's size
    uint32_t p = 1;        // virtual PEs
    uint32_t max_len = 0;
    std::vector<uint64_t> pe_begin;  // host copy, p+1
    uint64_t sort_cap_n = 0;  // strings and PEs the local sort's buffers are reserved for
    uint32_t sort_cap_p = 0;

    dss::DevBuf chars;  // nbytes + padding (owned arena)
    const uint8_t* borrowed = nullptr;  // arena owned by a job pool (batches)
    const uint8_t* arena() const { return borrowed ? borrowed : chars.as<uint8_t>(); }
    dss::DevBuf off;    // u64[n]
    dss::DevBuf len;    // u32[n]
    dss::DevBuf pe_begin_d;  // u64[p+1]

    // local sort results (positions)
    bool sorted = false;
    dss::DevBuf perm;   // u32[n]
    dss::DevBuf lcp;    // u32[n]
    dss::DevBuf s_off;  // u64[n] offset of the string at each sorted position
    dss::DevBuf s_len;  // u32[n] length of the string at each sorted position

    SortParams sort_params;
    dss::SortScratch ss;
    dss::PipeState pp;
    dss::DevBuf aux;    // small per-call scratch of result queries
    dss::KernelTimer timer;
};

namespace dss {

constexpr uint32_t kArenaPad = 64;  // readable bytes beyond the arena end

// Index a device arena (find NULs) into off/len; returns number of strings.
uint64_t index_arena(dss_set_s* s);
// Computes the PE split (readInput.hpp rule) from device offsets.
void compute_pe_split(dss_set_s* s);
// Allocates a set object with a stream on the current device.
dss_set_s* new_set(uint32_t num_pes);

void sort_lcp(dss_set_s* s);
uint64_t avg_lcp(dss_set_s* s);
void partition(dss_set_s* s, int sampler, uint64_t factor, uint64_t max_len, bool use_dist);
void exchange(dss_set_s* s, int mode);
void recv_origin(dss_set_s* s);  // fills recv_src_pe / recv_src_pos once per exchange
void permutation(dss_set_s* s);  // fills perm_out once per merged output
// Element offset of the PEIndex array in perm_out (16-byte aligned).
inline uint64_t perm_pe_off(uint64_t n) { return (std::max<uint64_t>(n, 1) + 3) / 4 * 4; }
void merge(dss_set_s* s);
void extend_prefix(dss_set_s* s);
void prefix_doubling(dss_set_s* s, uint64_t m, uint32_t start_depth);

}  // namespace dss
