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

/* FileDistributer / distribute_strings (src/util/random_string_generator.hpp:
 * 50-62, src/mpi/readInput.hpp:40-104): a '\n'-separated text (the first
 * max_size bytes if max_size > 0), '\n' -> NUL on the device, a NUL appended
 * to an unterminated last line, split into num_pes PEs by byte slices of the
 * text as the MPI ranks slice the file: a string belongs to the PE whose
 * slice holds the byte before its start.  (A slice without any '\n', i.e. a
 * line longer than the slice, is undefined in the reference: its shift_left
 * reads past the slice.)  _text takes a host buffer, _load_file reads the
 * file through pinned staging buffers. */
int dss_set_create_text(const uint8_t* text, uint64_t text_len, uint64_t max_size, uint32_t num_pes,
                        dss_set* out);
int dss_set_load_file(const char* path, uint64_t max_size, uint32_t num_pes, dss_set* out);

/* Builds the D/N-ratio input on the device (DNRatioGenerator,
 * src/util/random_string_generator.hpp:166-287, seeded): n strings of
 * max(len, k) chars, k = max(len*dton, ceil(log26 n)). */
int dss_set_create_dton(uint64_t n, uint32_t len, double dton, uint64_t seed,
                        uint32_t num_pes, dss_set* out);

/* Builds the skewed D/N input on the device (modelled on
 * SkewedDNRatioGenerator, src/util/random_string_generator.hpp:431-509):
 * long_permille of the strings have long_len chars, the rest short_len;
 * the first k = max(L*dn, ceil(log255 n)) chars are the base-255 digits of
 * the string's id (bytes 1..255), the rest seeded filler; dup_permille of the
 * strings repeat the id of another string (exact duplicates). */
int dss_set_create_skewed(uint64_t n, uint32_t long_len, uint32_t short_len,
                          double long_dn, double short_dn, uint32_t long_permille,
                          uint32_t dup_permille, uint64_t seed, uint32_t num_pes,
                          dss_set* out);

/* Uploads num_pes PE-local inputs laid out back to back (each PE's bytes a
 * sequence of NUL-terminated strings): pe_byte_begin[0..num_pes] are the byte
 * boundaries (the per-rank input of the reference's generators). */
int dss_set_create_pes(const uint8_t* chars, const uint64_t* pe_byte_begin, uint32_t num_pes,
                       dss_set* out);

/* Capped suffixes of a text as views into one arena (the suffix workload:
 * SuffixGenerator, src/util/random_string_generator.hpp:89-115, each suffix
 * capped at `cap` chars, without materialising them): string i =
 * text[i .. i + min(cap, text_len - i)), PEs split by the byte-slice rule
 * over the materialised strings.  The text must not contain NUL bytes. */
int dss_set_create_suffixes(const uint8_t* text, uint64_t text_len, uint32_t cap, uint32_t num_pes,
                            dss_set* out);
/* Same over a seeded order-3 Markov text over {A,C,G,T} (chain restarted
 * every 4096 chars) with `repeats` planted copies of rep_min..rep_max chars,
 * generated on the device. */
int dss_set_create_markov_suffixes(uint64_t text_len, uint32_t repeats, uint32_t rep_min,
                                   uint32_t rep_max, uint32_t cap, uint64_t seed, uint32_t num_pes,
                                   dss_set* out);

int dss_set_destroy(dss_set s);

/* The stream the set queues its work on: a caller's hipStream_t (passed as
 * void*), or NULL for the set's own stream (the default).  Work queued so far
 * finishes first.  The set never destroys a caller's stream; the caller keeps
 * it alive while the set uses it.  No counterpart in the reference (one MPI
 * rank is one thread); the explicit-stream form of the ABI, for callers that
 * order the set's work with their own. */
int dss_set_stream(dss_set s, void* stream);

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

/* Tuning of the local sort (no counterpart in the reference, whose tlx sort
 * has fixed thresholds): segments of more than large_min strings (16..1024,
 * default 512) go through the global MSD passes, smaller ones to the base
 * case; base-case classes up to 2^wave_class_max strings (0..6, default 6)
 * are sorted one segment per wave; skip_lanes (8 or 16, default 8) lanes
 * compare each string in the common-prefix skip, skip_strings (1, 2 or 4,
 * default 4; 0 selects the default) strings per lane group at a time;
 * merge_fan_in (2, 4, 8 or 16, default 2; 0 selects the default) runs are
 * merged per pass of dss_merge_lcp (the K of merge<K>, merging.hpp:23-63:
 * more runs than that take several passes; 2 is the pairwise merge tree, one
 * pass per level, larger values merge K runs per pass in LDS tiles);
 * pass_mode selects how the local sort's MSD passes are driven: 1 from the
 * host (each pass sized from counters read back after the previous one), 2
 * from the device (persistent kernels that read their work counts from
 * device counters, the launch sequence captured once per set in hipGraphs
 * and replayed: one read-back per sort), 0 (default) the device-driven form
 * for sets of at most 2^22 strings over at least 16 PEs (batches of jobs),
 * the host-driven one otherwise (the host-driven form is also used while
 * kernel timing is on, and on the null stream); merge_tile (512, 1024 or
 * 2048, default 1024; 0 selects the default) is the number of outputs per
 * tile of the pairwise merge (fan-in 2); digit_chars is the number of
 * characters one radix digit of the host-driven passes may cover: 1 one
 * (the byte), 2 or 3 up to that many for small alphabets (the sort scans the
 * arena's byte values and packs kc characters into a digit when (values + 1)
 * ^ kc <= 256), 0 (default) up to 3 for sets of at least 2^22 strings whose
 * arena holds at most 2 GiB, one otherwise.  Every setting produces the same
 * output.  NULL restores the defaults. */
typedef struct {
    uint32_t large_min;
    uint32_t wave_class_max;
    uint32_t skip_lanes;
    uint32_t skip_strings;
    uint32_t merge_fan_in;
    uint32_t pass_mode;
    uint32_t merge_tile;
    uint32_t digit_chars;
} dss_sort_params_t;

int dss_set_sort_params(dss_set s, const dss_sort_params_t* params);

/* Grows the set's device buffers for a local sort of num_strings strings
 * over num_pes PEs (for batch sets also their offset/length copies, keeping
 * their contents), so that later sorts up to that size allocate nothing (no
 * counterpart in the reference; for batch sets that are re-pointed at other
 * jobs).  Earlier results are dropped: sort the set again afterwards. */
int dss_set_reserve(dss_set s, uint64_t num_strings, uint32_t num_pes);
int dss_get_sort_params(dss_set s, dss_sort_params_t* params);
/* Sets a caller-computed local order instead of dss_sort_lcp (a StringLcpPtr
 * whose strings and LCPs the caller arranged, as the reference's exchange
 * tests do): perm permutes every PE's own string ids, lcp per position. */
int dss_set_local_order(dss_set s, const uint32_t* perm, const uint32_t* lcp);

/* ---- (a3) getAvgLcp (src/sorter/distributed/misc.hpp:67-87) ---- */
int dss_avg_lcp(dss_set s, uint64_t* avg);

/* ---- (a4, a5) computePartition (src/sorter/distributed/misc.hpp:473-648):
 *      splitter sampling (samplingStrategies.hpp; sampler 0 NumStrings,
 *      1 NumChars, 2 IndexedNumStrings, 3 IndexedNumChars), global sample
 *      order, splitter choice (getSplitters(Indexed), :89-189) and bucket
 *      sizes by binarySearch(Indexed) (:310-414).  Samples are truncated to
 *      max_len (the merge sort passes 200*(avgLcp+5)) or, with use_dist, to
 *      the prefix-doubling results.  interval (optional, p*p entries):
 *      interval[r*p+q] = strings PE r sends to PE q ---- */
int dss_partition(dss_set s, int sampler, uint64_t sampling_factor, uint64_t max_len,
                  int use_dist, uint64_t* interval);

/* Sets caller-computed send counts (the sendCounts argument of the exchange
 * policies' alltoallv, src/mpi/alltoall.hpp:642-1009) instead of
 * dss_partition; interval[r*p+q] as above, rows summing to the PE sizes. */
int dss_set_interval(dss_set s, const uint64_t* interval);
/* Sets caller-computed distinguishing prefixes (the distinguishingPrefix
 * argument of AllToAllStringImplPrefixDoubling::alltoallv) per position. */
int dss_set_dist(dss_set s, const uint32_t* dist);

/* ---- (a6, a7, a13) AllToAllStringImpl<..., EmptyByteEncoderMemCpy> (mode 0),
 *      <..., EmptyLcpByteEncoderMemCpy> (mode 1),
 *      AllToAllStringImplPrefixDoubling (mode 2) and
 *      AllToAllStringImpl<..., NoLcps> (mode 3: mode 0's bytes, no LCPs sent;
 *      the receiver derives its runs' LCPs from the received strings)
 *      (src/mpi/alltoall.hpp:567-804, 902-1009, encoders
 *      src/mpi/byte_encoder.hpp:81-172), for p virtual PEs as an in-HBM
 *      bucket scatter.  recv_begin / byte_begin (optional, p+1 entries) give
 *      each receiving PE's strings and arena bytes ---- */
int dss_exchange(dss_set s, int mode, uint64_t* recv_begin, uint64_t* byte_begin,
                 uint64_t* recv_nbytes);
/* Receive arena bytes, received LCPs (0 at every run start) and each
 * received string's origin (source PE, position in the source's sorted
 * order: StringIndexPEIndex); any pointer may be NULL. */
int dss_get_exchange(dss_set s, uint8_t* bytes, uint32_t* lcp, uint32_t* src_pe,
                     uint32_t* src_pos);

/* ---- (a8) choose_merge / merge<K> with LcpStringLoserTree_
 *      (src/sorter/distributed/merging.hpp:66-123,
 *      src/merge/bingmann-lcp_losertree.hpp:204-480): merges every PE's
 *      received runs with LCPs ---- */
int dss_merge_lcp(dss_set s);
/* Per receive-side output position: the received string (receive index),
 * the merged LCP and the saved LCP (oldLcps); any pointer may be NULL. */
int dss_get_merged(dss_set s, uint32_t* order, uint32_t* lcp, uint32_t* saved);

/* Sets a caller-built merged output instead of dss_merge_lcp: the input of
 * the reference's checker (is_sorted / is_complete_and_sorted,
 * src/mpi/is_sorted.hpp:27-198) as its tests build it
 * (src/tests/isSorted_tests.cpp:12-143).  Per receive-side output position
 * the received string (receive index < num_strings; a repeated index is
 * allowed and dss_check reports the output incomplete) and the merged LCP.
 * dss_check, dss_get_permutation and the distinguishing-prefix validators
 * then read this output; dss_extend_prefix refuses it (no saved LCPs). */
int dss_set_merged(dss_set s, const uint32_t* order, const uint32_t* lcp);

/* ---- (a9) StringLcpContainer::extendPrefix
 *      (src/strings/stringcontainer.hpp:247-319), called at
 *      src/executables/distributed_sorter.cpp:127-131 after a mode-1
 *      exchange ---- */
int dss_extend_prefix(dss_set s, uint64_t* out_nbytes);
/* Rebuilt arena (all PEs, merged order) and per-string byte offsets. */
int dss_get_extended(dss_set s, uint8_t* bytes, uint64_t* offsets);

/* ---- (a10-a12) computeDistinguishingPrefixes
 *      (src/sorter/distributed/prefix_doubling.hpp:156-198) with
 *      BloomFilter<...,XXHasher>::filter (bloomfilter.hpp:1415-1556); m is
 *      the hash range (bloomFilterSize, UINT64_MAX in the reference) ---- */
int dss_prefix_doubling(dss_set s, uint64_t m, uint32_t start_depth, uint32_t* num_rounds);
/* Per position: results (distinguishing-prefix estimate) and the number of
 * rounds the position was a candidate in; either pointer may be NULL. */
int dss_get_prefix_doubling(dss_set s, uint32_t* results, uint8_t* rounds);
/* XXHasher::hash (bloomfilter.hpp:1094-1102) of the first min(depth, len)
 * chars of the given strings (string ids), reduced % m. */
int dss_hash_prefixes(dss_set s, const uint32_t* ids, uint64_t count, uint32_t depth,
                      uint64_t m, uint64_t* out);

/* ---- drivers ---- */

/* DistributedMergeSort::sort (lcp_aware_number_string_splitting.hpp:
 * 189-289): compressed = 0 plain exchange (EmptyByteEncoderMemCpy) and LCP
 * merge; 1 LCP-compressed exchange and extendPrefix; 2 the NoLcps variant
 * (no LCPs sent, noLcpMerge, :139-182, 218-224, 276-281). */
int dss_merge_sort(dss_set s, int sampler, int compressed);
/* DistributedPrefixDoublingSort::sort (prefix_doubling.hpp:204-294). */
int dss_prefix_doubling_sort(dss_set s, int sampler, uint64_t m);
/* The permutation of DistributedPrefixDoublingSort (prefix_doubling.hpp:
 * 282-290): per receive-side output position, stringIndex and PEIndex. */
int dss_get_permutation(dss_set s, uint32_t* string_index, uint32_t* pe_index);

/* ---- (a14) is_complete_and_sorted + CheckerWithCompleteExchange
 *      (src/mpi/is_sorted.hpp:27-198) on the merged output of all PEs, and
 *      D = sum over the globally sorted output of max(lcp_i, lcp_i+1) + 1
 *      (computeD, src/tests/DNGenerator_tests.cpp:22-45) ---- */
typedef struct {
    uint64_t num_strings;
    uint64_t order_errors;      /* adjacent pairs out of order */
    uint64_t lcp_errors;        /* merged LCPs that differ from the recomputed LCP */
    uint64_t dist_prefix_size;  /* D */
    int32_t sorted;             /* 1 if globally sorted */
    int32_t complete;           /* 1 if every input string appears exactly once */
} dss_check_t;

int dss_check(dss_set s, dss_check_t* out);

/* ---- ExcatDistinguishingPrefix::filter_exact
 *      (src/sorter/distributed/bloomfilter.hpp:847-965) on the globally
 *      sorted output of dss_merge_sort / dss_prefix_doubling_sort ----
 * Per position (PE r's sorted index j at pe_begin[r] + j: the index space of
 * dss_get_prefix_doubling's results) the exact distinguishing-prefix length:
 * 1 + the larger full-string LCP with its neighbours in the global order, 0
 * when the set holds one string.  exact may be NULL (device only). */
int dss_exact_dist(dss_set s, uint32_t* exact);

typedef struct {
    uint64_t num_strings;
    uint64_t lower_violations;  /* results < DIST (< DIST - 1 where results = length) */
    uint64_t upper_violations;  /* results < length and results >= 2 * DIST + 8 */
    uint64_t exact_sum;         /* sum of DIST (= D of the output) */
    uint64_t results_sum;       /* sum of results */
} dss_dist_check_t;

/* Device validator of prefix doubling's results against the exact lengths:
 * a result is an upper bound of DIST, and without hash collisions the
 * doubling schedule 8, 24, 40, 72, ... keeps it below 2 * DIST + 8. */
int dss_check_dist(dss_set s, dss_dist_check_t* out);

/* ---- batching and serving: a device job pool ----
 *
 * No counterpart in the reference (its drivers sort one input per MPI job).
 * A dss_pool holds the character arenas and per-string offset/length arrays
 * of many independent sort jobs in HBM, allocated first-fit from fixed
 * capacities fixed at creation (the admission limit).  dss_pool_batch builds
 * a dss_set over any subset of live jobs without copying characters: job k
 * of the batch becomes virtual PE k, so dss_sort_lcp sorts every job of the
 * batch in the same kernel launches (src/sorter/RadixSortLCP call sites, per
 * job).  A pool may be used from several host threads; each batch set is
 * used by one thread at a time, and a job must not be freed while a batch
 * that holds it is alive. */
typedef struct dss_pool_s* dss_pool;

int dss_pool_create(uint64_t char_capacity, uint64_t string_capacity, dss_pool* out);
int dss_pool_destroy(dss_pool pool);

/* Adds a job from a host arena of NUL-terminated strings; *job = its id. */
int dss_pool_add_job(dss_pool pool, const uint8_t* chars, uint64_t nbytes, uint32_t* job);
/* Adds a job generated on the device, like dss_set_create_dton / _skewed. */
int dss_pool_add_job_dton(dss_pool pool, uint64_t n, uint32_t length, double dton, uint64_t seed,
                          uint32_t* job);
int dss_pool_add_job_skewed(dss_pool pool, uint64_t n, uint32_t long_len, uint32_t short_len,
                            double long_dn, double short_dn, uint32_t long_permille,
                            uint32_t dup_permille, uint64_t seed, uint32_t* job);
int dss_pool_free_job(dss_pool pool, uint32_t job);
int dss_pool_job_info(dss_pool pool, uint32_t job, uint64_t* num_strings, uint64_t* num_bytes);
int dss_pool_usage(dss_pool pool, uint64_t* bytes_used, uint64_t* strings_used,
                   uint64_t* bytes_capacity, uint64_t* strings_capacity);

/* A string set whose PE k is jobs[k] (characters borrowed from the pool). */
int dss_pool_batch(dss_pool pool, const uint32_t* jobs, uint32_t count, dss_set* out);
/* Re-points an existing batch set of the same pool at other jobs, keeping its
 * device buffers (no allocation once they are large enough; avoids the
 * device-wide synchronisation of freeing memory between batches). */
int dss_pool_rebatch(dss_pool pool, dss_set batch, const uint32_t* jobs, uint32_t count);

/* Per-PE results of the local sort: D of every PE's sorted strings alone
 * (computeD, src/tests/DNGenerator_tests.cpp:22-45; 0 below two strings),
 * and one PE's sorted order as PE-local string indices with its LCP array
 * (the sorted StringLcpContainer of one job). */
int dss_pe_dist_sizes(dss_set s, uint64_t* d_per_pe);
int dss_get_sorted_pe(dss_set s, uint32_t pe, uint32_t* perm, uint32_t* lcp);

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


/* ---- wire codecs of the exchange (multi-GPU wire formats; lossless) ----
 *
 * A dss_codec owns a HIP stream and scratch buffers.  The plain functions
 * take host buffers; the _dev functions take device pointers for the data
 * (values, bytes, words) and host pointers for counts and sizes.  An encoder
 * that returns DSS_ERANGE has stored the size it needs in its total. */
typedef struct dss_codec_s* dss_codec;

int dss_codec_create(dss_codec* out);
int dss_codec_destroy(dss_codec c);
int dss_codec_sync(dss_codec c);
int dss_codec_set_timing(dss_codec c, int enable);
int dss_codec_kernel_stats(dss_codec c, dss_kernel_stat_t* stats, uint32_t cap, uint32_t* count, int reset);

/* IntegerCompression::writeRanges (src/mpi/byte_encoder.hpp:31-53) over
 * PutVarint (src/encoding/integer_compression.hpp:67-150), as sendLcps sends
 * LCPs (src/mpi/alltoall.hpp:406-425): the values of num_ranges ranges
 * (counts[r] each) as varints back to back; range_bytes[r] = bytes of range r. */
int dss_varint_write_ranges(dss_codec c, const uint64_t* values, const uint64_t* counts, uint32_t num_ranges,
                            uint8_t* out, uint64_t out_cap, uint64_t* range_bytes, uint64_t* total_bytes);
int dss_varint_write_ranges_dev(dss_codec c, const uint64_t* values, const uint64_t* counts, uint32_t num_ranges,
                                uint8_t* out, uint64_t out_cap, uint64_t* range_bytes, uint64_t* total_bytes);
/* No counterpart in the reference: dss_varint_write_ranges_dev for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and output;
 * its scratch (per-value lengths and offsets, scan sums, range ends) lives in
 * one fixed device array of the library (about 34 MB, held on every GPU the
 * library runs on until the process ends), so a call allocates nothing.
 * Limits: at most DSS_NOALLOC_VARINT_MAX_VALUES values and
 * DSS_NOALLOC_VARINT_MAX_RANGES ranges per call (DSS_ERANGE beyond).  The
 * encoded bytes are complete once the codec's stream has finished
 * (dss_codec_sync).  All calls on one GPU share the array: calls that overlap
 * on one GPU (other codecs or host threads) race. */
#define DSS_NOALLOC_VARINT_MAX_VALUES (1u << 21)
#define DSS_NOALLOC_VARINT_MAX_RANGES 4096u
int dss_varint_write_ranges_noalloc(dss_codec c, const uint64_t* values, const uint64_t* counts,
                                    uint32_t num_ranges, uint8_t* out, uint64_t out_cap, uint64_t* range_bytes,
                                    uint64_t* total_bytes);
/* IntegerCompression::readRanges (src/mpi/byte_encoder.hpp:64-78): nbytes must
 * hold exactly sum(counts) varints; DSS_EINVAL where GetVarint
 * (integer_compression.hpp:209-236) throws or the input is short. */
int dss_varint_read_ranges(dss_codec c, const uint8_t* bytes, uint64_t nbytes, const uint64_t* counts,
                           uint32_t num_ranges, uint64_t* values);
int dss_varint_read_ranges_dev(dss_codec c, const uint8_t* bytes, uint64_t nbytes, const uint64_t* counts,
                               uint32_t num_ranges, uint64_t* values);

/* getB (src/encoding/golomb_encoding.hpp:49-58). */
int dss_golomb_b(uint64_t max_value, uint64_t num_values, uint64_t* b);

/* Delta + Golomb streams of ascending u64 values per range
 * (getDeltaEncoding, src/encoding/golomb_encoding.hpp:60-76), framed as
 *   DSS_GOLOMB_BLOCKS:    AllToAllHashesGolomb (bloomfilter.hpp:164-313):
 *                         [size][b][raw first value][codes] per non-empty
 *                         range, nothing for an empty one;
 *   DSS_GOLOMB_PIPELINED: AllToAllHashValuesPipeline::getEncoding
 *                         (bloomfilter.hpp:322-331): [size][raw][codes] per
 *                         range, [0] for an empty one.
 * b[r] is range r's Golomb parameter (the policies' getB choices are the
 * caller's); range_words[r] = words of range r's block. */
#define DSS_GOLOMB_BLOCKS 0
#define DSS_GOLOMB_PIPELINED 1
int dss_golomb_encode(dss_codec c, int layout, const uint64_t* values, const uint64_t* counts, const uint64_t* b,
                      uint32_t num_ranges, uint64_t* out, uint64_t out_cap, uint64_t* range_words,
                      uint64_t* total_words);
int dss_golomb_encode_dev(dss_codec c, int layout, const uint64_t* values, const uint64_t* counts, const uint64_t* b,
                          uint32_t num_ranges, uint64_t* out, uint64_t out_cap, uint64_t* range_words,
                          uint64_t* total_words);
/* getDeltaDecoding (golomb_encoding.hpp:93-109) of every block of nwords
 * words (the receive loops of bloomfilter.hpp:226-240 and 453-463): values
 * of all blocks back to back, block_counts[j] = values of block j,
 * *num_blocks = blocks found.  DSS_GOLOMB_BLOCKS reads b from each block and
 * accepts at most max_blocks blocks; DSS_GOLOMB_PIPELINED reads exactly
 * max_blocks blocks with parameter b_pipelined. */
int dss_golomb_decode(dss_codec c, int layout, const uint64_t* words, uint64_t nwords, uint32_t max_blocks,
                      uint64_t b_pipelined, uint64_t* values, uint64_t cap, uint64_t* block_counts,
                      uint32_t* num_blocks);
int dss_golomb_decode_dev(dss_codec c, int layout, const uint64_t* words, uint64_t nwords, uint32_t max_blocks,
                          uint64_t b_pipelined, uint64_t* values, uint64_t cap, uint64_t* block_counts,
                          uint32_t* num_blocks);

#ifdef __cplusplus
}
#endif

#endif /* DSS_GPU_H */
