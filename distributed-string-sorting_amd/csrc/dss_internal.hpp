// Internal declarations shared by the HIP translation units of libdss_gpu.
#pragma once

#include <hip/hip_runtime.h>

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
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

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

// ------------------------------------------------------------- sort scratch
struct SortScratch {
    DevBuf keyA, keyB, pbA, pbB;
    DevBuf segA, segB;      // large segment lists
    DevBuf tiles;           // tile list
    DevBuf hist;            // per tile x 256 digit counters / offsets
    DevBuf small;           // small segment lists (all classes)
    DevBuf counters;        // device counters
    std::vector<uint64_t> small_off, small_cap;  // per class region in `small`
    uint64_t seg_cap = 0, tile_cap = 0;
};

}  // namespace dss

// ------------------------------------------------------------------ the set
struct dss_set_s {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t n = 0;        // strings
    uint64_t nbytes = 0;   // arena bytes (incl. NULs)
    uint32_t p = 1;        // virtual PEs
    uint32_t max_len = 0;
    std::vector<uint64_t> pe_begin;  // host copy, p+1

    dss::DevBuf chars;  // nbytes + padding
    dss::DevBuf off;    // u64[n]
    dss::DevBuf len;    // u32[n]
    dss::DevBuf pe_begin_d;  // u64[p+1]

    // local sort results (positions)
    bool sorted = false;
    dss::DevBuf perm;   // u32[n]
    dss::DevBuf lcp;    // u32[n]

    dss::SortScratch ss;
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

}  // namespace dss
