/*
 * dss_gpu.h -- C ABI of the MI355X (gfx950) per-PE string-sorting path.
 *
 * The reference has no runtime plugin or FFI: its drivers take compile-time
 * policies and call header templates.  Each entry point below replaces one of
 * those call sites; the citation names the reference interface it stands in
 * for (paths relative to the reference repository root, TLXV = the vendored
 * tlx copy NeuKaDiS/KaDiS/external/RBC/external/tlx/tlx).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative DSS_E* code on
 *     failure; dss_last_error() then describes the failure (thread-local).
 *   - A dss_set is an opaque handle to a device string set holding p virtual
 *     PEs: one byte arena of NUL-terminated strings in HBM plus per-string
 *     u64 offset and u32 length, and per-PE contiguous id ranges.  All later
 *     stages keep their state in the handle (device resident), like the
 *     StringLcpContainer that the reference moves through its pipeline.
 *   - Host buffers passed in or out are caller-owned plain memory.
 *   - One handle must not be used from two host threads at once.
 *   - Positions: every per-string output is indexed by position in the
 *     concatenation of the PEs (PE r owns positions pe_begin[r]..pe_begin[r+1]).
 *   - LCP values are u32 (the reference stores size_t); strings are < 4 GiB.
 */
#ifndef DSS_GPU_H
#define DSS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSS_OK 0
#define DSS_EINVAL -1   /* invalid argument */
#define DSS_ENOMEM -2   /* device allocation failed */
#define DSS_EHIP -3     /* HIP runtime error */
#define DSS_ESTATE -4   /* stage called out of order */
#define DSS_ERANGE -5   /* size beyond supported range */

typedef struct dss_set_s* dss_set;

/* Last error message of the calling thread ("" if none). */
const char* dss_last_error(void);

/* Library / device information: writes a short description (device name,
 * gfx arch, CU count, HBM size) into buf. */
int dss_device_info(char* buf, size_t buflen);

/* ---- device string set (src/strings/stringcontainer.hpp:163-431,
 *      InitPolicy :40-70; PE split of distribute_strings,
 *      src/mpi/readInput.hpp:40-104) ---- */

/* Uploads a NUL-separated arena (every string NUL-terminated) and splits it
 * into num_pes virtual PEs by the reference's byte-slice rule. */
int dss_set_create(const uint8_t* chars, uint64_t nbytes, uint32_t num_pes, dss_set* out);

/* Builds the D/N-ratio input on the device (DNRatioGenerator,
 * src/util/random_string_generator.hpp:166-287, seeded): n strings of
 * max(len, k) chars, k = max(len*dton, ceil(log26 n)). */
int dss_set_create_dton(uint64_t n, uint32_t len, double dton, uint64_t seed,
                        uint32_t num_pes, dss_set* out);

/* Builds the skewed D/N input on the device (see the oracle's generator
 * specification in DESIGN notes of the project): long_permille of the
 * strings long, dup_permille exact duplicates, bytes 1..255. */
int dss_set_create_skewed(uint64_t n, uint32_t long_len, uint32_t short_len,
                          double long_dn, double short_dn, uint32_t long_permille,
                          uint32_t dup_permille, uint64_t seed, uint32_t num_pes,
                          dss_set* out);

int dss_set_destroy(dss_set s);

typedef struct {
    uint64_t num_strings;
    uint64_t num_bytes;   /* arena bytes incl. NULs */
    uint32_t num_pes;
    uint32_t max_len;
} dss_set_info_t;

int dss_set_info(dss_set s, dss_set_info_t* info);
/* pe_begin: num_pes+1 entries. */
int dss_set_pe_begin(dss_set s, uint64_t* pe_begin);
/* Copies the arena, per-string offsets (u64) and lengths (u32) to the host;
 * any pointer may be NULL. */
int dss_set_download(dss_set s, uint8_t* chars, uint64_t* offsets, uint32_t* lengths);

/* ---- (a2) local LCP sort: replaces tlx radixsort_CI3(StringLcpPtr, 0, 0)
 *      (TLXV/sort/strings/radix_sort.hpp:860-880) at
 *      src/sorter/distributed/lcp_aware_number_string_splitting.hpp:218-224
 *      and src/sorter/distributed/prefix_doubling.hpp:228 ---- */

/* Sorts every PE's strings; afterwards the set holds per position the sorted
 * string id and the LCP with the previous position of the same PE (0 at each
 * PE start). */
int dss_sort_lcp(dss_set s);
/* perm/lcp: num_strings entries each (either may be NULL). */
int dss_get_sorted(dss_set s, uint32_t* perm, uint32_t* lcp);

/* ---- measurement ---- */

/* Waits for all work queued on the set's stream. */
int dss_sync(dss_set s);

/* Enables per-kernel HIP-event timing on the set's stream (0/1). */
int dss_set_timing(dss_set s, int enable);

typedef struct {
    char name[32];
    uint64_t launches;
    double total_ms;      /* sum of event-measured launch durations */
    double alg_bytes;     /* algorithmic bytes summed over launches */
} dss_kernel_stat_t;

/* Fills up to cap entries, returns the number of kernels with statistics in
 * *count; reset != 0 clears the statistics afterwards. */
int dss_kernel_stats(dss_set s, dss_kernel_stat_t* stats, uint32_t cap, uint32_t* count,
                     int reset);

#ifdef __cplusplus
}
#endif

#endif /* DSS_GPU_H */
