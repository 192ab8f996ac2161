/*
 * dss_oracle.c -- CPU restatement of the per-PE string-sorting path of
 * bingmann/distributed-string-sorting.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the correctness oracle that the tests,
 * the smoke check and the CPU-baseline leg of the benchmark compare the HIP
 * implementation against.  Nothing in the product links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Paths are relative to the reference repository root; TLXV stands for the
 * vendored tlx copy at NeuKaDiS/KaDiS/external/RBC/external/tlx/tlx.
 *
 * Conventions shared with the device implementation:
 *   - a string set is one byte arena of NUL-terminated strings plus a u64
 *     start offset and a u32 length per string ("String = {ptr, length}",
 *     src/strings/stringset.hpp:700-845);
 *   - p virtual PEs own contiguous ranges of string ids, pe_begin[0..p];
 *   - a sorted PE is a permutation of its string ids (u32) plus a u32 LCP
 *     array, lcp[0] = 0 (the reference leaves it at its zero-initialised
 *     value, TLXV/sort/strings/radix_sort.hpp:585-604).
 *
 * Parity status of each function is recorded next to it ("pinned by").
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Seeded input generators.  The reference's generators draw their seed from  */
/* std::random_device (src/util/random_string_generator.hpp:290-298,501-504), */
/* so they are restated here with a counter-based RNG and a Feistel shuffle;  */
/* the string construction itself follows the reference.                     */
/* ------------------------------------------------------------------------- */

static inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

ORC_API uint64_t orc_rng(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
}

/* Bijection on [0, n): 4-round Feistel network on the smallest even bit width
 * covering n, with cycle walking. */
ORC_API uint64_t orc_shuffle_index(uint64_t x, uint64_t n, uint64_t seed) {
    if (n <= 1) return x;
    unsigned bits = 0;
    while ((1ull << bits) < n) ++bits;
    if (bits & 1) ++bits;
    if (bits < 2) bits = 2;
    const unsigned half = bits / 2;
    const uint64_t mask = (1ull << half) - 1;
    do {
        uint64_t L = x >> half, R = x & mask;
        for (unsigned r = 0; r < 4; ++r) {
            uint64_t F = mix64(seed ^ ((uint64_t)(r + 1) * 0xA24BAED4963EE407ull) ^ R) & mask;
            uint64_t t = R;
            R = L ^ F;
            L = t;
        }
        x = (L << half) | R;
    } while (x >= n);
    return x;
}

/* DNRatioGenerator::getRawStringsTimoStyle
 * (src/util/random_string_generator.hpp:188-235): string i is the base-26
 * representation of i over 'A'..'Z', left-padded with 'A' to k characters,
 * followed by stringLength-k copies of one random character; then the
 * strings are shuffled (:279-285). */
ORC_API void orc_dton_params(uint64_t n, uint32_t L, double dton,
                             uint32_t* k_out, uint32_t* slen_out) {
    double a = (double)L * dton;
    double b = ceil(log((double)n) / log(26.0));
    double m = a > b ? a : b; /* std::max of the two doubles (:197-198) */
    uint32_t k = (uint32_t)m;
    uint32_t slen = L > k ? L : k;
    *k_out = k;
    *slen_out = slen;
}

static void dton_string(uint64_t i, uint32_t k, uint32_t slen, uint8_t fill,
                        uint8_t* out) {
    for (uint32_t j = 0; j < k; ++j) out[j] = 'A';
    uint64_t cur = i;
    for (uint32_t j = 0; j < k && cur != 0; ++j) {
        out[k - 1 - j] = (uint8_t)('A' + cur % 26);
        cur /= 26;
    }
    for (uint32_t j = k; j < slen; ++j) out[j] = fill;
    out[slen] = 0;
}

ORC_API uint8_t orc_dton_fill(uint64_t seed) {
    return (uint8_t)('A' + orc_rng(seed, 1, 0) % 26);
}

/* out must hold n*(slen+1) bytes; arena position j holds string shuffle(j). */
ORC_API void orc_gen_dton(uint64_t n, uint32_t L, double dton, uint64_t seed,
                          uint8_t* out) {
    uint32_t k, slen;
    orc_dton_params(n, L, dton, &k, &slen);
    const uint8_t fill = orc_dton_fill(seed);
#pragma omp parallel for schedule(static)
    for (uint64_t j = 0; j < n; ++j) {
        uint64_t i = orc_shuffle_index(j, n, seed);
        dton_string(i, k, slen, fill, out + j * (uint64_t)(slen + 1));
    }
}

/* Skewed D/N generator (the skewed configuration of this project, modelled on
 * SkewedDNRatioGenerator, src/util/random_string_generator.hpp:431-509):
 * a fraction of long strings with a long distinguishing part, the rest short,
 * a fraction of exact duplicates, bytes 1..255.  String j of the arena is
 * built from id(i) with i = shuffle(j): duplicates reuse a random earlier id.
 * The distinguishing part is the base-255 representation of the id (bytes
 * 1..255, padded with byte 1) of length k = max(L*dn, ceil(log n/log 255)),
 * followed by per-string random filler bytes. */
typedef struct {
    uint64_t n;
    uint32_t long_len, short_len;
    double long_dn, short_dn;
    uint32_t long_permille, dup_permille;
} orc_skew_params;

static uint32_t skew_k(uint64_t n, uint32_t L, double dn) {
    double a = (double)L * dn;
    double b = ceil(log((double)(n > 1 ? n : 2)) / log(255.0));
    double m = a > b ? a : b;
    uint32_t k = (uint32_t)m;
    return k;
}

static void skew_id_len(const orc_skew_params* P, uint64_t seed, uint64_t i,
                        uint64_t* id, uint32_t* len, uint32_t* k) {
    uint64_t d = orc_rng(seed, 2, i) % 1000;
    uint64_t x = (d < P->dup_permille) ? orc_rng(seed, 3, i) % P->n : i;
    int is_long = (orc_rng(seed, 4, x) % 1000) < P->long_permille;
    uint32_t L = is_long ? P->long_len : P->short_len;
    uint32_t kk = skew_k(P->n, L, is_long ? P->long_dn : P->short_dn);
    if (L < kk) L = kk;
    *id = x;
    *len = L;
    *k = kk;
}

ORC_API uint64_t orc_gen_skewed_size(const orc_skew_params* P, uint64_t seed) {
    uint64_t total = 0;
    for (uint64_t i = 0; i < P->n; ++i) {
        uint64_t id;
        uint32_t len, k;
        skew_id_len(P, seed, i, &id, &len, &k);
        total += len + 1;
    }
    return total;
}

ORC_API void orc_gen_skewed(const orc_skew_params* P, uint64_t seed,
                            uint8_t* out) {
    uint64_t pos = 0;
    for (uint64_t j = 0; j < P->n; ++j) {
        uint64_t i = orc_shuffle_index(j, P->n, seed);
        uint64_t id;
        uint32_t len, k;
        skew_id_len(P, seed, i, &id, &len, &k);
        uint8_t* s = out + pos;
        for (uint32_t t = 0; t < k; ++t) s[t] = 1;
        uint64_t cur = id;
        for (uint32_t t = 0; t < k && cur != 0; ++t) {
            s[k - 1 - t] = (uint8_t)(1 + cur % 255);
            cur /= 255;
        }
        uint64_t h = orc_rng(seed, 5, id);
        for (uint32_t t = k; t < len; ++t) /* 8 filler bytes per mix64 */
            s[t] = (uint8_t)(1 + ((mix64(h + (t >> 3)) >> (8 * (t & 7))) & 0xFF) % 255);
        s[len] = 0;
        pos += len + 1;
    }
}

/* Suffix-style input (SuffixGenerator, src/util/random_string_generator.hpp:
 * 64-130, with suffixes capped): a text of text_len bytes over {A,C,G,T}
 * from an order-3 Markov chain whose transition table is drawn from the
 * seed, restarted every 4096 bytes, with `repeats` planted copies of earlier
 * substrings of length rep_min..rep_max; then every suffix capped at `cap`
 * characters, in text order.  Returns the arena size; writes only when out
 * is non-NULL. */
ORC_API void orc_gen_markov_text(uint64_t text_len, uint64_t seed,
                                 uint32_t repeats, uint32_t rep_min,
                                 uint32_t rep_max, uint8_t* text) {
    static const uint8_t ACGT[4] = {'A', 'C', 'G', 'T'};
    uint32_t table[64][4];
    for (int c = 0; c < 64; ++c) {
        uint32_t acc = 0;
        for (int s = 0; s < 4; ++s) {
            uint32_t w = 1 + (uint32_t)(orc_rng(seed, 7, (uint64_t)c * 4 + s) % 16);
            acc += w;
            table[c][s] = acc;
        }
    }
    for (uint64_t base = 0; base < text_len; base += 4096) {
        uint32_t ctx = (uint32_t)(orc_rng(seed, 8, base) % 64);
        uint64_t end = base + 4096 < text_len ? base + 4096 : text_len;
        for (uint64_t i = base; i < end; ++i) {
            uint32_t tot = table[ctx][3];
            uint32_t r = (uint32_t)(orc_rng(seed, 9, i) % tot);
            int s = 0;
            while (r >= table[ctx][s]) ++s;
            text[i] = ACGT[s];
            ctx = ((ctx << 2) | (uint32_t)s) & 63;
        }
    }
    for (uint32_t r = 0; r < repeats; ++r) {
        uint64_t span = rep_max > rep_min ? rep_max - rep_min + 1 : 1;
        uint64_t len = rep_min + orc_rng(seed, 10, r) % span;
        if (len * 2 >= text_len) continue;
        uint64_t src = orc_rng(seed, 11, r) % (text_len - len);
        uint64_t dst = orc_rng(seed, 12, r) % (text_len - len);
        memmove(text + dst, text + src, len);
    }
}

ORC_API uint64_t orc_suffix_arena(const uint8_t* text, uint64_t text_len,
                                  uint32_t cap, uint8_t* out) {
    uint64_t pos = 0;
    for (uint64_t i = 0; i < text_len; ++i) {
        uint64_t len = text_len - i;
        if (len > cap) len = cap;
        if (out) {
            memcpy(out + pos, text + i, len);
            out[pos + len] = 0;
        }
        pos += len + 1;
    }
    return pos;
}

/* ------------------------------------------------------------------------- */
/* String set basics                                                          */
/* ------------------------------------------------------------------------- */

ORC_API uint64_t orc_count_strings(const uint8_t* chars, uint64_t nbytes) {
    uint64_t n = 0;
    const uint8_t *p = chars, *e = chars + nbytes;
    while (p < e && (p = (const uint8_t*)memchr(p, 0, (size_t)(e - p))) != NULL) {
        ++n;
        ++p;
    }
    return n;
}

/* InitPolicy<GenericCharLengthStringSet>::init_strings
 * (src/strings/stringcontainer.hpp:49-68): one String per NUL-terminated run. */
ORC_API void orc_index(const uint8_t* chars, uint64_t nbytes, uint64_t* off,
                       uint32_t* len) {
    uint64_t s = 0, k = 0;
    const uint8_t* e = chars + nbytes;
    while (s < nbytes) {
        const uint8_t* z = (const uint8_t*)memchr(chars + s, 0, (size_t)(e - (chars + s)));
        if (!z) break;
        const uint64_t i = (uint64_t)(z - chars);
        off[k] = s;
        len[k] = (uint32_t)(i - s);
        ++k;
        s = i + 1;
    }
}

/* distribute_strings (src/mpi/readInput.hpp:40-104) on a NUL-separated arena:
 * PE r reads byte slice [o_r, o_{r+1}) (the first nbytes%p slices one byte
 * larger), drops its first partial string and hands it to PE r-1.  Net rule:
 * the string starting at byte s belongs to the PE whose slice holds byte s-1
 * (s = 0 belongs to PE 0).  Output: pe_begin[0..p] string-id boundaries. */
ORC_API void orc_pe_split(const uint64_t* off, uint64_t n, uint64_t nbytes,
                          uint32_t p, uint64_t* pe_begin) {
    uint64_t small = nbytes / p, larger = nbytes % p;
    pe_begin[0] = 0;
    uint64_t k = 0;
    for (uint32_t r = 1; r <= p; ++r) {
        /* slice r starts at o_r; strings with s-1 < o_r, i.e. s <= o_r, belong
         * to PEs before r */
        uint64_t o_r = r < larger ? r * (small + 1)
                                  : larger * (small + 1) + (r - larger) * small;
        if (r == p) o_r = nbytes;
        while (k < n && (off[k] == 0 || off[k] - 1 < o_r)) ++k;
        pe_begin[r] = (r == p) ? n : k;
    }
}

static inline int scmp0(const uint8_t* a, const uint8_t* b) {
    /* dss_schimek::scmp (src/strings/stringtools.hpp:108-115) */
    while (*a != 0 && *a == *b) ++a, ++b;
    return (int)*a - (int)*b;
}

static inline uint32_t calc_lcp(const uint8_t* a, const uint8_t* b) {
    uint32_t l = 0;
    while (a[l] != 0 && a[l] == b[l]) ++l;
    return l;
}

/* ------------------------------------------------------------------------- */
/* (a2) Local LCP sort: restatement of tlx radixsort_CI3                      */
/* (TLXV/sort/strings/radix_sort.hpp:539-880, insertion_sort.hpp): 16-bit     */
/* MSD radix steps for buckets of >= 65536 strings, 8-bit steps below, LCP    */
/* insertion sort below 32.  LCP rules: strings that end inside a step are    */
/* equal and get lcp = their length; the first string of every further non-   */
/* empty bucket gets lcp = depth (+1 when the 16-bit digits share the high    */
/* byte).  lcp[0] of the top-level call is left untouched (0).                */
/* Pinned by: textbook byte-lexicographic order + brute-force neighbour LCP   */
/* (tests/test_oracle.py); tie order among equal strings is not pinned.       */
/* ------------------------------------------------------------------------- */

typedef struct {
    const uint8_t* chars;
    const uint64_t* off;
    uint32_t* lcp; /* indexed by absolute position */
} sort_ctx;

static inline const uint8_t* S(const sort_ctx* c, uint32_t id) {
    return c->chars + c->off[id];
}

/* LCP insertion sort (TLXV/sort/strings/insertion_sort.hpp:81+). */
static void ins_sort(const sort_ctx* c, uint32_t* a, uint64_t base, uint64_t n,
                     uint32_t depth) {
    for (uint64_t i = 1; i < n; ++i) {
        uint32_t x = a[i];
        uint64_t j = i;
        while (j > 0 && scmp0(S(c, a[j - 1]) + depth, S(c, x) + depth) > 0) {
            a[j] = a[j - 1];
            --j;
        }
        a[j] = x;
    }
    for (uint64_t i = 1; i < n; ++i)
        c->lcp[base + i] = depth + calc_lcp(S(c, a[i - 1]) + depth, S(c, a[i]) + depth);
}

static void radix8(const sort_ctx* c, uint32_t* a, uint32_t* tmp, uint64_t base,
                   uint64_t n, uint32_t depth);

static void sort_rec(const sort_ctx* c, uint32_t* a, uint32_t* tmp,
                     uint64_t base, uint64_t n, uint32_t depth) {
    if (n < 2) return;
    if (n < 32) {
        ins_sort(c, a, base, n, depth);
        return;
    }
    if (n < 65536) {
        radix8(c, a, tmp, base, n, depth);
        return;
    }
    /* RadixStep_CI3: 16-bit digits = get_uint16 (big endian, 0 after end,
     * src/strings/stringset.hpp:157-168) */
    uint64_t* cnt = (uint64_t*)calloc(65536, sizeof(uint64_t));
    uint16_t* cc = (uint16_t*)malloc(n * sizeof(uint16_t));
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* s = S(c, a[i]) + depth;
        uint16_t v = 0;
        if (s[0] != 0) v = (uint16_t)((s[0] << 8) | s[1]);
        cc[i] = v;
        ++cnt[v];
    }
    uint64_t* pos = (uint64_t*)malloc(65536 * sizeof(uint64_t));
    uint64_t acc = 0;
    for (uint32_t d = 0; d < 65536; ++d) {
        pos[d] = acc;
        acc += cnt[d];
    }
    for (uint64_t i = 0; i < n; ++i) tmp[pos[cc[i]]++] = a[i];
    memcpy(a, tmp, n * sizeof(uint32_t));
    free(cc);
    /* lcps */
    uint64_t b = 0;
    int prev = -1;
    for (uint32_t d = 0; d < 65536; ++d) {
        uint64_t m = cnt[d];
        if (m == 0) continue;
        if (prev >= 0 && b > 0)
            c->lcp[base + b] = depth + (((uint32_t)prev >> 8) == (d >> 8) ? 1 : 0);
        if (d == 0) {
            for (uint64_t i = 1; i < m; ++i) c->lcp[base + b + i] = depth;
        } else if ((d & 0xFF) == 0) {
            for (uint64_t i = 1; i < m; ++i) c->lcp[base + b + i] = depth + 1;
        }
        prev = (int)d;
        b += m;
    }
    b = 0;
    for (uint32_t d = 0; d < 65536; ++d) {
        uint64_t m = cnt[d];
        if (m > 1 && d != 0 && (d & 0xFF) != 0)
            sort_rec(c, a + b, tmp, base + b, m, depth + 2);
        b += m;
    }
    free(pos);
    free(cnt);
}

/* RadixStep_CI2 (TLXV/sort/strings/radix_sort.hpp:539-612). */
static void radix8(const sort_ctx* c, uint32_t* a, uint32_t* tmp, uint64_t base,
                   uint64_t n, uint32_t depth) {
    uint64_t cnt[256];
    memset(cnt, 0, sizeof(cnt));
    uint8_t* cc = (uint8_t*)malloc(n);
    for (uint64_t i = 0; i < n; ++i) {
        cc[i] = S(c, a[i])[depth];
        ++cnt[cc[i]];
    }
    uint64_t pos[256], acc = 0;
    for (int d = 0; d < 256; ++d) {
        pos[d] = acc;
        acc += cnt[d];
    }
    for (uint64_t i = 0; i < n; ++i) tmp[pos[cc[i]]++] = a[i];
    memcpy(a, tmp, n * sizeof(uint32_t));
    free(cc);
    for (uint64_t i = 1; i < cnt[0]; ++i) c->lcp[base + i] = depth;
    uint64_t b = 0;
    for (int d = 0; d < 256; ++d) {
        if (cnt[d] == 0) continue;
        if (b > 0) c->lcp[base + b] = depth;
        b += cnt[d];
    }
    b = cnt[0];
    for (int d = 1; d < 256; ++d) {
        if (cnt[d] > 1) sort_rec(c, a + b, tmp, base + b, cnt[d], depth + 1);
        b += cnt[d];
    }
}

/* Sorts every PE's id range [pe_begin[r], pe_begin[r+1]) independently:
 * perm[i] receives the i-th id in sorted order, lcp[i] its LCP with
 * perm[i-1] inside the same PE (0 at each PE start). */
ORC_API void orc_local_sort(const uint8_t* chars, const uint64_t* off,
                            const uint64_t* pe_begin, uint32_t p, uint32_t* perm,
                            uint32_t* lcp) {
    uint64_t n = pe_begin[p];
    uint32_t* tmp = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    sort_ctx c = {chars, off, lcp};
    for (uint64_t i = 0; i < n; ++i) {
        perm[i] = (uint32_t)i;
        lcp[i] = 0;
    }
    for (uint32_t r = 0; r < p; ++r) {
        uint64_t b = pe_begin[r], m = pe_begin[r + 1] - b;
        sort_rec(&c, perm + b, tmp, b, m, 0);
    }
    free(tmp);
}

/* ------------------------------------------------------------------------- */
/* Distinguishing-prefix size D (src/tests/DNGenerator_tests.cpp:22-45):      */
/* D = sum_{i<n-1} (max(lcp_i, lcp_{i+1}) + 1) + lcp_{n-1} + 1, lcp_0 = 0,    */
/* n < 2 -> 0.  Pinned by the reference's own definition (test file).         */
/* ------------------------------------------------------------------------- */
ORC_API uint64_t orc_compute_d(const uint32_t* lcp, uint64_t n) {
    if (n < 2) return 0;
    uint64_t D = 0;
    for (uint64_t i = 0; i + 1 < n; ++i) {
        uint64_t a = i == 0 ? 0 : lcp[i], b = lcp[i + 1];
        D += (a > b ? a : b) + 1;
    }
    D += (uint64_t)lcp[n - 1] + 1;
    return D;
}

/* ------------------------------------------------------------------------- */
/* (a3) getAvgLcp (src/sorter/distributed/misc.hpp:67-87): integer mean of    */
/* all local LCP entries over all PEs.                                        */
/* ------------------------------------------------------------------------- */
ORC_API uint64_t orc_avg_lcp(const uint32_t* lcp, uint64_t n) {
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; ++i) s += lcp[i];
    return n ? s / n : 0;
}

/* ------------------------------------------------------------------------- */
/* (a4, a5) Partition: samplers (src/sorter/distributed/samplingStrategies.   */
/* hpp), global sample order (the RQuick result, misc.hpp:503-514), splitter  */
/* choice (getSplitters(Indexed), misc.hpp:89-189) and per-PE bucket sizes by */
/* binarySearch(Indexed) (misc.hpp:310-414).                                  */
/* sampler: 0 NumStrings, 1 NumChars, 2 IndexedNumStrings, 3 IndexedNumChars. */
/* max_len: truncation length (merge sort: 2*100*(avgLcp+5)); dist: per-      */
/* position distinguishing prefix (prefix doubling) or NULL.                   */
/* Output: interval[r*p + q] = strings PE r sends to PE q.                     */
/* Pinned by: counting (tests/test_oracle.py): with the indexed samplers the */
/* boundary before bucket q+1 equals the number of local (string, global      */
/* index) keys below splitter q's key; without an index it lies between the   */
/* counts of smaller and of not-greater strings; the splitters are the        */
/* samples at ranks i*floor(S/(nspl+1)) of the textbook-sorted sample list;   */
/* NumStrings samples sit at floor(i*n/(nr+1)), NumChars samples where the    */
/* character count since the previous sample first reaches num/(nr+1); the    */
/* buckets of all PEs respect the global order.  Also an independent Python  */
/* restatement of the rules, for all four samplers.                           */
/* ------------------------------------------------------------------------- */

typedef struct {
    const uint8_t* s; /* chars (NUL-terminated copy) */
    uint64_t index;   /* global index (indexed samplers) */
    uint32_t tlen;    /* truncated length */
} sample_t;

static int g_cmp_indexed;
static int sample_cmp(const void* x, const void* y) {
    const sample_t* a = (const sample_t*)x;
    const sample_t* b = (const sample_t*)y;
    int r = scmp0(a->s, b->s);
    if (r != 0) return r < 0 ? -1 : 1;
    if (g_cmp_indexed) return a->index < b->index ? -1 : (a->index > b->index ? 1 : 0);
    return 0;
}

/* indexStringCompare (misc.hpp:335-345) */
static int index_cmp(const uint8_t* a, uint64_t ia, const uint8_t* b, uint64_t ib) {
    while (*a == *b && *a != 0) ++a, ++b;
    if (*a != *b) return (int)*a - (int)*b;
    return ia < ib ? -1 : (ia > ib ? 1 : 0);
}

/* orc_partition_ex additionally reports, when the pointers are non-NULL,
 * every sample in emission order (PE by PE) as (global index, truncated
 * length) -- capacity p * sampling_factor * (p - 1) -- and the p - 1
 * splitters the same way (UINT64_MAX index for a missing splitter), so that
 * tests can check the bucket bounds against the splitters by counting. */
ORC_API int orc_partition_ex(const uint8_t* chars, const uint64_t* off,
                             const uint32_t* len, const uint64_t* pe_begin, uint32_t p,
                             const uint32_t* perm, int sampler, uint64_t sampling_factor,
                             uint64_t max_len, const uint32_t* dist, uint64_t* interval,
                             uint64_t* out_sample_gidx, uint32_t* out_sample_tlen,
                             uint64_t* out_num_samples, uint64_t* out_spl_gidx,
                             uint32_t* out_spl_tlen) {
    const int indexed = sampler >= 2;
    const int by_chars = sampler == 1 || sampler == 3;
    if (p == 0 || sampling_factor * (p - 1) > 8192) return -1;
    uint64_t nsamp_cap = 0;
    for (uint32_t r = 0; r < p; ++r) nsamp_cap += sampling_factor * (p - 1) + 1;
    sample_t* samp = (sample_t*)calloc(nsamp_cap + 1, sizeof(sample_t));
    uint8_t** bufs = (uint8_t**)calloc(nsamp_cap + 1, sizeof(uint8_t*));
    uint64_t ns = 0;
    uint64_t local_offset = 0;
    for (uint32_t r = 0; r < p; ++r) {
        const uint64_t b = pe_begin[r], m = pe_begin[r + 1] - b;
        /* getNumberSplitter (samplingStrategies.hpp:14-18) */
        uint64_t nr = sampling_factor * (p - 1);
        if (nr > m) nr = m;
        uint64_t idx_list[8192];
        uint64_t cnt = 0;
        if (!by_chars) {
            const double sd = (double)m / (double)(nr + 1);
            for (uint64_t i = 1; i <= nr; ++i) idx_list[cnt++] = (uint64_t)((double)i * sd);
        } else {
            /* NumChars policies (:199-236, :299-337).  With distinguishing
             * prefixes the indexed policy counts those instead of lengths
             * (:239-296); the plain one still counts lengths and only
             * truncates the splitter to dist (:339-387). */
            const int wdist = dist && sampler == 3;
            uint64_t num = 0;
            for (uint64_t i = 0; i < m; ++i) num += wdist ? dist[b + i] : len[perm[b + i]];
            const uint64_t sd = num / (nr + 1);
            uint64_t si = 0;
            for (uint64_t i = 1; i <= nr && si < m; ++i) {
                uint64_t seen = 0;
                while (seen < sd && si < m) {
                    seen += wdist ? dist[b + si] : len[perm[b + si]];
                    ++si;
                }
                idx_list[cnt++] = si ? si - 1 : 0; /* the reference reads index -1 here
                                                      when no characters were seen */
            }
        }
        for (uint64_t t = 0; t < cnt; ++t) {
            uint64_t idx = idx_list[t];
            uint32_t id = perm[b + idx];
            uint64_t ml = dist ? dist[b + idx] : max_len;
            uint64_t sl = len[id];
            uint64_t used = sl > ml ? ml : sl;
            uint8_t* copy = (uint8_t*)malloc(used + 1);
            memcpy(copy, chars + off[id], used);
            copy[used] = 0;
            bufs[ns] = copy;
            samp[ns].s = copy;
            samp[ns].index = local_offset + idx;
            samp[ns].tlen = (uint32_t)used;
            if (out_sample_gidx) out_sample_gidx[ns] = local_offset + idx;
            if (out_sample_tlen) out_sample_tlen[ns] = (uint32_t)used;
            ++ns;
        }
        local_offset += m;
    }
    if (out_num_samples) *out_num_samples = ns;
    g_cmp_indexed = indexed;
    qsort(samp, ns, sizeof(sample_t), sample_cmp);
    uint64_t nspl = (p - 1) < ns ? (p - 1) : ns;
    uint64_t sdist = ns / (nspl + 1);
    for (uint64_t q = 0; q + 1 < p; ++q) {
        const int have = q < nspl;
        if (out_spl_gidx) out_spl_gidx[q] = have ? samp[(q + 1) * sdist].index : UINT64_MAX;
        if (out_spl_tlen) out_spl_tlen[q] = have ? samp[(q + 1) * sdist].tlen : 0;
    }
    local_offset = 0;
    for (uint32_t r = 0; r < p; ++r) {
        const uint64_t b = pe_begin[r], m = pe_begin[r + 1] - b;
        uint64_t prev = 0;
        for (uint64_t q = 0; q + 1 < p; ++q) {
            uint64_t pos;
            if (q >= nspl) {
                pos = m; /* fewer splitters than PEs: trailing buckets empty */
            } else {
                const sample_t* sp = &samp[(q + 1) * sdist];
                uint64_t left = 0, right = m;
                pos = (uint64_t)-1;
                while (left != right) {
                    uint64_t d = (right - left) / 2;
                    const uint8_t* cur = chars + off[perm[b + left + d]];
                    int res = indexed ? index_cmp(cur, local_offset + left + d, sp->s, sp->index)
                                      : scmp0(cur, sp->s);
                    if (res < 0) left = left + d + 1;
                    else if (res == 0) { pos = left + d; break; }
                    else right = left + d;
                }
                if (pos == (uint64_t)-1) pos = left;
            }
            interval[(uint64_t)r * p + q] = pos - prev;
            prev = pos;
        }
        interval[(uint64_t)r * p + p - 1] = m - prev;
        local_offset += m;
    }
    for (uint64_t i = 0; i < ns; ++i) free(bufs[i]);
    free(bufs);
    free(samp);
    return 0;
}

ORC_API int orc_partition(const uint8_t* chars, const uint64_t* off,
                          const uint32_t* len, const uint64_t* pe_begin, uint32_t p,
                          const uint32_t* perm, int sampler, uint64_t sampling_factor,
                          uint64_t max_len, const uint32_t* dist, uint64_t* interval) {
    return orc_partition_ex(chars, off, len, pe_begin, p, perm, sampler, sampling_factor, max_len, dist,
                            interval, NULL, NULL, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------- */
/* (a6, a7, a13) String exchange between virtual PEs.                          */
/* mode 0: EmptyByteEncoderMemCpy (byte_encoder.hpp:81-107, whole strings);    */
/* mode 1: EmptyLcpByteEncoderMemCpy (:109-137, chars [lcp, len] incl. NUL);   */
/* mode 2: EmptyPrefixDoublingLcpByteEncoderMemCpy (:139-172, chars           */
/*         [lcp, dist) then NUL).                                             */
/* The sender zeroes the LCP at each non-empty bucket start                    */
/* (setLcpAtStartOfInterval, alltoall.hpp:396-404); receiver q's arena is the */
/* concatenation over source PEs r of bucket r->q (AllToAllvCombined order);  */
/* LCPs travel verbatim (sendLcps :406-425).  For each received string the    */
/* origin (source PE, position in the source's sorted order) is recorded     */
/* (InitPolicy<GenericCharIndexPEIndexStringSet>, stringcontainer.hpp:127-   */
/* 159 with recvOffsets, alltoall.hpp:981-987).                              */
/* Returns total received bytes; fills outputs when out_bytes != NULL.        */
/* Pinned by: alltoallStrings_tests.cpp:17-232 known answers                  */
/* (tests/test_oracle.py::test_exchange_reference_vectors).                   */
/* ------------------------------------------------------------------------- */
/* Receiver q's part of the exchange: its strings from every source PE in
 * source order, written from byte `total` and receive index `k` on (outputs
 * only when out_bytes != NULL); returns q's byte count. */
static uint64_t exch_receiver(int mode, uint32_t q, const uint8_t* chars, const uint64_t* off,
                              const uint32_t* len, const uint64_t* pe_begin, uint32_t p,
                              const uint32_t* perm, const uint32_t* lcp, const uint32_t* dist,
                              const uint64_t* interval, uint64_t total, uint64_t k,
                              uint8_t* out_bytes, uint32_t* out_lcp, uint32_t* out_src_pe,
                              uint32_t* out_src_pos) {
    const uint64_t total0 = total;
    for (uint32_t r = 0; r < p; ++r) {
        uint64_t start = 0;
        for (uint32_t t = 0; t < q; ++t) start += interval[(uint64_t)r * p + t];
        uint64_t cnt = interval[(uint64_t)r * p + q];
        for (uint64_t j = 0; j < cnt; ++j) {
            uint64_t pos = pe_begin[r] + start + j;
            uint32_t id = perm[pos];
            uint32_t l = j == 0 ? 0 : lcp[pos];
            uint64_t from, to; /* chars copied, then maybe NUL */
            int add_nul = 0;
            if (mode == 0) { from = 0; to = (uint64_t)len[id] + 1; }
            else if (mode == 1) { from = l; to = (uint64_t)len[id] + 1; }
            else { from = l; to = dist[pos]; add_nul = 1; }
            uint64_t m = to - from;
            if (out_bytes) {
                memcpy(out_bytes + total, chars + off[id] + from, m);
                if (add_nul) out_bytes[total + m] = 0;
                out_lcp[k] = l;
                out_src_pe[k] = r;
                out_src_pos[k] = (uint32_t)(start + j);
            }
            total += m + add_nul;
            ++k;
        }
    }
    return total - total0;
}

ORC_API uint64_t orc_exchange(int mode, const uint8_t* chars, const uint64_t* off,
                              const uint32_t* len, const uint64_t* pe_begin, uint32_t p,
                              const uint32_t* perm, const uint32_t* lcp,
                              const uint32_t* dist, const uint64_t* interval,
                              uint8_t* out_bytes, uint64_t* out_recv_begin,
                              uint64_t* out_byte_begin, uint32_t* out_lcp,
                              uint32_t* out_src_pe, uint32_t* out_src_pos) {
    uint64_t total = 0, k = 0;
    for (uint32_t q = 0; q < p; ++q) {
        if (out_bytes) {
            out_recv_begin[q] = k;
            out_byte_begin[q] = total;
        }
        total += exch_receiver(mode, q, chars, off, len, pe_begin, p, perm, lcp, dist, interval, total, k,
                               out_bytes, out_lcp, out_src_pe, out_src_pos);
        for (uint32_t r = 0; r < p; ++r) k += interval[(uint64_t)r * p + q];
    }
    if (out_bytes) {
        out_recv_begin[p] = k;
        out_byte_begin[p] = total;
    }
    return total;
}

/* ------------------------------------------------------------------------- */
/* (a8) LCP loser-tree K-way merge: dss_schimek::LcpStringLoserTree_          */
/* (src/merge/bingmann-lcp_losertree.hpp:204-480) behind choose_merge/merge<K> */
/* (src/sorter/distributed/merging.hpp:23-123): K = pow2roundup(p), missing   */
/* runs are empty; output[i] = winner, out_lcp[i] = LCP with the previous     */
/* output (nodes[1].lcp), saved[i] = the winner's own-run LCP (oldLcps,       */
/* :395) when prefix compressed.  Strings are given as char pointers; with    */
/* compression a string's stored chars start at logical position run_lcp.     */
/* Pinned by: oracle/_ref built from the reference header (bit-exact incl.    */
/* tie order) and its golden vectors, tests/test_oracle.py.                   */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t idx;
    uint32_t lcp;
} lt_node;

typedef struct {
    const uint8_t* const* str; /* per recv string: first stored char */
    const uint32_t* lcp;       /* per recv string: run LCP */
    uint64_t cur, end;
} lt_stream;

static inline int lt_empty(const lt_stream* s) { return s->cur >= s->end; }

static void lt_update(lt_stream* st, lt_node* contender, lt_node* defender, int compressed) {
    lt_stream* ds = &st[defender->idx];
    if (lt_empty(ds)) return;
    lt_stream* cs = &st[contender->idx];
    if (lt_empty(cs)) {
        lt_node t = *defender; *defender = *contender; *contender = t;
        return;
    }
    if (defender->lcp > contender->lcp) {
        lt_node t = *defender; *defender = *contender; *contender = t;
    } else if (defender->lcp == contender->lcp) {
        uint32_t l = defender->lcp;
        const uint8_t* s1 = ds->str[ds->cur] + (compressed ? l - ds->lcp[ds->cur] : l);
        const uint8_t* s2 = cs->str[cs->cur] + (compressed ? l - cs->lcp[cs->cur] : l);
        while (*s1 != 0 && *s1 == *s2) ++s1, ++s2, ++l;
        if (*s1 < *s2) {
            lt_node t = *defender; *defender = *contender; *contender = t;
        }
        defender->lcp = l;
    }
}

static uint64_t pow2roundup(uint64_t x) {
    if (x == 0) return 1;
    --x;
    for (int s = 1; s < 64; s <<= 1) x |= x >> s;
    return x + 1;
}

/* strs/lcp: all n strings of one receiver; run_begin[0..p]; outputs n. */
ORC_API void orc_merge_losertree(const uint8_t* const* strs, const uint32_t* lcp,
                                 const uint64_t* run_begin, uint32_t p, int compressed,
                                 uint64_t* out_order, uint32_t* out_lcp,
                                 uint32_t* out_saved) {
    uint64_t n = run_begin[p];
    if (n == 0) return;
    uint64_t K = pow2roundup(p);
    lt_stream* st = (lt_stream*)calloc(K + 1, sizeof(lt_stream));
    lt_node* nodes = (lt_node*)calloc(K + 1, sizeof(lt_node));
    for (uint64_t i = 1; i <= K; ++i) {
        st[i].str = strs;
        st[i].lcp = lcp;
        if (i - 1 < p) {
            st[i].cur = run_begin[i - 1];
            st[i].end = run_begin[i];
        }
    }
    /* initTree(0) (:335-359) uses the uncompressed updateNode */
    for (uint64_t k = 1; k <= K; ++k) {
        lt_node contender = {k, 0};
        uint64_t node = K + k;
        while (node % 2 == 0 && node > 2) {
            node >>= 1;
            lt_update(st, &contender, &nodes[node], 0);
        }
        node = (node + 1) / 2;
        nodes[node] = contender;
    }
    /* writeElementsToStream (:375-428 / :429-479) */
    for (uint64_t o = 0; o < n; ++o) {
        uint64_t w = nodes[1].idx;
        out_order[o] = st[w].cur;
        out_lcp[o] = nodes[1].lcp;
        if (out_saved) out_saved[o] = lcp[st[w].cur];
        ++st[w].cur;
        lt_node* contender = &nodes[1];
        if (!lt_empty(&st[w])) contender->lcp = lcp[st[w].cur];
        uint64_t node = w + K;
        while (node > 2) {
            node = (node + 1) / 2;
            lt_update(st, contender, &nodes[node], compressed);
        }
    }
    free(nodes);
    free(st);
}

/* ------------------------------------------------------------------------- */
/* (a9) StringLcpContainer::extendPrefix (src/strings/stringcontainer.hpp:    */
/* 247-319): rebuild full strings in merged order from prefix-compressed     */
/* strings and savedLcps; output arena = prefix(saved[i]) of the previous     */
/* rebuilt string + stored chars + NUL.  Returns output size.                 */
/* Pinned by: oracle/_ref built from the reference header; invariant          */
/* "output == uncompressed merge".                                            */
/* ------------------------------------------------------------------------- */
ORC_API uint64_t orc_extend_prefix(const uint8_t* const* strs, const uint64_t* order,
                                   const uint32_t* saved, uint64_t n, uint8_t* out,
                                   uint64_t* out_off) {
    uint64_t pos = 0;
    uint64_t prev_off = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* s = strs[order[i]];
        uint64_t pre = i == 0 ? 0 : saved[i];
        uint64_t sl = strlen((const char*)s);
        if (out) {
            if (pre) memcpy(out + pos, out + prev_off, pre);
            memcpy(out + pos + pre, s, sl + 1);
            out_off[i] = pos;
        }
        prev_off = pos;
        pos += pre + sl + 1;
    }
    return pos;
}

/* ------------------------------------------------------------------------- */
/* (a12) XXHasher::hash (src/sorter/distributed/bloomfilter.hpp:1094-1102):  */
/* xxh::hash_state_t<64> with seed 0, one update, digest, then % m.           */
/* The streaming digest (src/hash/xxhash.hpp:508-640) equals the published    */
/* XXH64 algorithm (Y. Collet) except that it takes the short-input branch    */
/* (v3 + PRIME5) when the total length is exactly 32 (test `total_len > 32`,  */
/* xxhash.hpp:573), with v3 already advanced by the one consumed stripe.      */
/* Pinned by: oracle/_ref built from the reference header (all lengths 0-300)*/
/* and the published XXH64 test vectors (tests/test_oracle.py).               */
/* ------------------------------------------------------------------------- */
#define P1 11400714785074694791ull
#define P2 14029467366897019727ull
#define P3 1609587929392839161ull
#define P4 9650029242287828579ull
#define P5 2870177450012600261ull

static inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static inline uint64_t rd64(const uint8_t* p) {
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}
static inline uint32_t rd32(const uint8_t* p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}
static inline uint64_t xround(uint64_t acc, uint64_t in) {
    acc += in * P2;
    acc = rotl64(acc, 31);
    return acc * P1;
}
static inline uint64_t xmerge(uint64_t acc, uint64_t v) {
    v = xround(0, v);
    acc ^= v;
    return acc * P1 + P4;
}

ORC_API uint64_t orc_xxh64_stream(const uint8_t* p, uint64_t len) {
    uint64_t v1 = P1 + P2, v2 = P2, v3 = 0, v4 = (uint64_t)0 - P1;
    const uint8_t* const end = p + len;
    if (len >= 32) {
        do {
            v1 = xround(v1, rd64(p));
            v2 = xround(v2, rd64(p + 8));
            v3 = xround(v3, rd64(p + 16));
            v4 = xround(v4, rd64(p + 24));
            p += 32;
        } while (p <= end - 32);
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
    while (p + 8 <= end) {
        h ^= xround(0, rd64(p));
        h = rotl64(h, 27) * P1 + P4;
        p += 8;
    }
    if (p + 4 <= end) {
        h ^= (uint64_t)rd32(p) * P1;
        h = rotl64(h, 23) * P2 + P3;
        p += 4;
    }
    while (p < end) {
        h ^= (*p) * P5;
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

ORC_API uint64_t orc_pd_hash(const uint8_t* p, uint64_t len, uint64_t m) {
    return orc_xxh64_stream(p, len) % m;
}

/* ------------------------------------------------------------------------- */
/* (a10, a11) Prefix doubling: computeDistinguishingPrefixes                  */
/* (src/sorter/distributed/prefix_doubling.hpp:156-198) with                  */
/* BloomFilter::filter (bloomfilter.hpp:1415-1556).  Per round with depth d   */
/* (8, then 8+16, 8+32, 8+64, ...) over the candidates of every PE, in local  */
/* sorted order:                                                              */
/*   d > len            -> EOS candidate, results = len (setDepth :1380-1411) */
/*   prev cand = c-1 and lcp[c] >= d -> local LCP duplicate; the last hashed  */
/*                          record, if it is c-1, becomes an LCP-local root   */
/*                          (generateHashStringIndices :1162-1227)            */
/*   otherwise          -> record (hash of chars[0,d) % m, c)                 */
/* records sorted by hash (ties: by position here; ips4o order is unpinned),  */
/* getIndicesOfLocalDuplicates (:1333-1378) flags equal-hash groups (first    */
/* member sent anyway) and lone LCP roots; the sent hashes of all PEs are     */
/* merged and every member of an equal-hash group is reported back            */
/* (FindDuplicates::findDuplicates :662-765); next candidates = local hash    */
/* duplicates + LCP duplicates + remote duplicates that were not sent anyway  */
/* (getSortedIndicesOfDuplicates :767-817), as a multiset in sorted order.    */
/* results[c] = d for every candidate of the round, len for EOS candidates.   */
/* The loop stops when no PE has candidates (allreduce_and, :183-189).        */
/* Outputs per position: results, rounds (number of rounds the position was a */
/* candidate in).  Returns the number of rounds run.                          */
/* Pinned by: the closed form "a candidate stays iff another candidate of any */
/* PE has the same hash of its d-prefix" checked against brute-force exact    */
/* prefix duplicates on small inputs (tests/test_oracle.py), and hash vectors */
/* from oracle/_ref.                                                          */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t hash;
    uint64_t pos; /* absolute position */
    uint32_t pe;
    uint8_t local_dup, send_anyway, lcp_root;
} pd_rec;

static int rec_cmp(const void* x, const void* y) {
    const pd_rec* a = (const pd_rec*)x;
    const pd_rec* b = (const pd_rec*)y;
    if (a->pe != b->pe) return a->pe < b->pe ? -1 : 1;
    if (a->hash != b->hash) return a->hash < b->hash ? -1 : 1;
    return a->pos < b->pos ? -1 : (a->pos > b->pos ? 1 : 0);
}

typedef struct {
    uint64_t hash;
    uint64_t rec; /* index into sent records */
} sent_t;

static int sent_cmp(const void* x, const void* y) {
    const sent_t* a = (const sent_t*)x;
    const sent_t* b = (const sent_t*)y;
    if (a->hash != b->hash) return a->hash < b->hash ? -1 : 1;
    return a->rec < b->rec ? -1 : (a->rec > b->rec ? 1 : 0);
}

static int u64_cmp(const void* x, const void* y) {
    uint64_t a = *(const uint64_t*)x, b = *(const uint64_t*)y;
    return a < b ? -1 : (a > b ? 1 : 0);
}

ORC_API uint32_t orc_prefix_doubling(const uint8_t* chars, const uint64_t* off,
                                     const uint32_t* len, const uint64_t* pe_begin,
                                     uint32_t p, const uint32_t* perm, const uint32_t* lcp,
                                     uint64_t m, uint32_t start_depth, uint32_t* results,
                                     uint8_t* rounds) {
    const uint64_t n = pe_begin[p];
    /* candidate multiset as a sorted list of positions (may repeat) */
    uint64_t* cand = (uint64_t*)malloc((n ? n : 1) * 2 * sizeof(uint64_t));
    uint64_t ncand = n;
    for (uint64_t i = 0; i < n; ++i) {
        cand[i] = i;
        results[i] = 0;
        rounds[i] = 0;
    }
    pd_rec* rec = (pd_rec*)malloc((n ? n : 1) * 2 * sizeof(pd_rec));
    uint64_t* next = (uint64_t*)malloc((n ? n : 1) * 4 * sizeof(uint64_t));
    sent_t* sent = (sent_t*)malloc((n ? n : 1) * 2 * sizeof(sent_t));
    uint32_t round = 0;
    uint64_t d = start_depth;
    uint64_t step = (uint64_t)start_depth * 2;
    while (ncand > 0) {
        /* PE id of each candidate */
        uint64_t nrec = 0, nnext = 0;
        uint64_t ci = 0;
        for (uint32_t r = 0; r < p; ++r) {
            uint64_t c0 = ci;
            while (ci < ncand && cand[ci] < pe_begin[r + 1]) ++ci;
            uint64_t rec0 = nrec;
            for (uint64_t t = c0; t < ci; ++t) {
                uint64_t c = cand[t];
                if ((t == c0 || cand[t - 1] != c) && rounds[c] < 255) rounds[c]++;
                results[c] = (uint32_t)d;
                uint32_t id = perm[c];
                if (d > len[id]) {
                    results[c] = len[id];
                    continue; /* EOS */
                }
                if (t > c0 && cand[t - 1] + 1 == c && lcp[c] >= d) {
                    next[nnext++] = c; /* local LCP duplicate */
                    if (nrec > rec0 && rec[nrec - 1].pos + 1 == c) rec[nrec - 1].lcp_root = 1;
                    continue;
                }
                pd_rec x;
                x.hash = orc_pd_hash(chars + off[id], d, m);
                x.pos = c;
                x.pe = r;
                x.local_dup = x.send_anyway = x.lcp_root = 0;
                rec[nrec++] = x;
            }
        }
        qsort(rec, nrec, sizeof(pd_rec), rec_cmp);
        /* getIndicesOfLocalDuplicates per PE */
        uint64_t nsent = 0;
        uint64_t a = 0;
        while (a < nrec) {
            uint64_t e = a;
            while (e < nrec && rec[e].pe == rec[a].pe) ++e;
            /* records [a, e) of one PE, sorted by hash */
            uint64_t i = a;
            while (e - a >= 1 && i + 1 < e) {
                uint64_t j = i + 1;
                if (rec[j].hash == rec[i].hash) {
                    rec[i].local_dup = rec[i].send_anyway = 1;
                    rec[j].local_dup = 1;
                    next[nnext++] = rec[i].pos;
                    next[nnext++] = rec[j].pos;
                    ++j;
                    while (j < e && rec[j].hash == rec[i].hash) {
                        rec[j].local_dup = 1;
                        next[nnext++] = rec[j].pos;
                        ++j;
                    }
                } else if (rec[i].lcp_root) {
                    rec[i].local_dup = rec[i].send_anyway = 1;
                    next[nnext++] = rec[i].pos;
                }
                i = j;
            }
            if (e > a && rec[e - 1].lcp_root) {
                rec[e - 1].local_dup = rec[e - 1].send_anyway = 1;
                next[nnext++] = rec[e - 1].pos;
            }
            for (uint64_t t = a; t < e; ++t)
                if (!rec[t].local_dup || rec[t].send_anyway) {
                    sent[nsent].hash = rec[t].hash;
                    sent[nsent].rec = t;
                    ++nsent;
                }
            a = e;
        }
        /* global duplicate detection over all sent hashes */
        qsort(sent, nsent, sizeof(sent_t), sent_cmp);
        for (uint64_t i = 0; i < nsent;) {
            uint64_t j = i + 1;
            while (j < nsent && sent[j].hash == sent[i].hash) ++j;
            if (j - i >= 2) {
                for (uint64_t t = i; t < j; ++t) {
                    const pd_rec* x = &rec[sent[t].rec];
                    if (!x->send_anyway) next[nnext++] = x->pos;
                }
            }
            i = j;
        }
        qsort(next, nnext, sizeof(uint64_t), u64_cmp);
        memcpy(cand, next, nnext * sizeof(uint64_t));
        ncand = nnext;
        ++round;
        d = start_depth + step;
        step *= 2;
        if (round > 60) break;
    }
    free(sent);
    free(next);
    free(rec);
    free(cand);
    return round;
}

/* ------------------------------------------------------------------------- */
/* (a14) Checker: is_complete_and_sorted / is_sorted                          */
/* (src/mpi/is_sorted.hpp:104-198): per-PE order, order across PE boundaries */
/* (skipping empty PEs), equal total char and string counts.                  */
/* strs: per output string pointer, pe_begin over outputs.                    */
/* Returns 1 if sorted and complete.                                          */
/* ------------------------------------------------------------------------- */
ORC_API int orc_is_sorted(const uint8_t* const* strs, const uint64_t* pe_begin,
                          uint32_t p) {
    const uint8_t* last = NULL;
    for (uint32_t r = 0; r < p; ++r) {
        for (uint64_t i = pe_begin[r]; i < pe_begin[r + 1]; ++i) {
            if (last && scmp0(last, strs[i]) > 0) return 0;
            last = strs[i];
        }
    }
    return 1;
}

/* ------------------------------------------------------------------------- */
/* (next) Wire codecs.                                                        */
/* ------------------------------------------------------------------------- */

/* PutVarint (src/encoding/integer_compression.hpp:67-150): 7 value bits per  */
/* byte, least significant group first, 0x80 on every byte but the last      */
/* (at most 10 bytes).  IntegerCompression::writeRanges                       */
/* (src/mpi/byte_encoder.hpp:31-53): the ranges are encoded back to back and  */
/* range_bytes[r] is the byte count of range r.  Returns the total bytes.     */
/* Pinned by: the reference's integerCompression_tests (round trip; the two  */
/* ranges of 8 values use 37 bytes each) and byte-exact vectors from the      */
/* reference header compiled in oracle/_ref.                                  */
ORC_API uint64_t orc_varint_write_ranges(const uint64_t* values, const uint64_t* counts,
                                         uint32_t num_ranges, uint8_t* out,
                                         uint64_t* range_bytes) {
    uint64_t o = 0, i = 0;
    for (uint32_t r = 0; r < num_ranges; ++r) {
        const uint64_t o0 = o;
        for (uint64_t k = 0; k < counts[r]; ++k, ++i) {
            uint64_t v = values[i];
            while (v >= 128) {
                out[o++] = (uint8_t)((v & 0x7F) | 0x80);
                v >>= 7;
            }
            out[o++] = (uint8_t)v;
        }
        range_bytes[r] = o - o0;
    }
    return o;
}

/* GetVarint (src/encoding/integer_compression.hpp:209-236) over             */
/* IntegerCompression::readRanges (src/mpi/byte_encoder.hpp:64-78): decodes   */
/* sum(counts) values in order.  Returns the bytes consumed, or -1 when the   */
/* input ends inside a value or a 10th byte carries bits beyond the 64th     */
/* (the reference throws std::overflow_error there).                          */
ORC_API int64_t orc_varint_read_ranges(const uint8_t* bytes, uint64_t nbytes,
                                       const uint64_t* counts, uint32_t num_ranges,
                                       uint64_t* values) {
    uint64_t o = 0, i = 0;
    for (uint32_t r = 0; r < num_ranges; ++r) {
        for (uint64_t k = 0; k < counts[r]; ++k, ++i) {
            uint64_t v = 0;
            for (unsigned s = 0;; s += 7) {
                if (o >= nbytes) return -1;
                const uint8_t b = bytes[o++];
                if (s == 63 && (b & 0xFE)) return -1;
                v |= (uint64_t)(b & 0x7F) << s;
                if (!(b & 0x80)) break;
                if (s == 63) return -1;
            }
            values[i] = v;
        }
    }
    return (int64_t)o;
}

/* getB (src/encoding/golomb_encoding.hpp:49-58): the power of two           */
/* 2^floor(log2(max_value / num_values)) (double log2, as the reference), 1   */
/* for no values or an average gap below 1.                                  */
ORC_API uint64_t orc_golomb_b(uint64_t max_value, uint64_t num_values) {
    if (num_values == 0) return 1;
    const uint64_t gap = max_value / num_values;
    if (gap < 1) return 1;
    const uint64_t k = (uint64_t)floor(log2((double)gap));
    return 1ull << k;
}

/* Bit writer of BitStreamWriter (src/encoding/bit_stream.hpp:40-110): bits   */
/* fill 64-bit words from the most significant end.                           */
typedef struct {
    uint64_t* out;
    uint64_t n;    /* words written */
    uint64_t buf;
    unsigned pos;  /* bits used in buf */
} orc_bits;

static void bits_put(orc_bits* w, uint64_t value, unsigned bits) {
    for (unsigned k = bits; k > 0; --k) {  /* one bit at a time, MSB first */
        const uint64_t bit = (value >> (k - 1)) & 1;
        w->buf |= bit << (63 - w->pos);
        if (++w->pos == 64) {
            w->out[w->n++] = w->buf;
            w->buf = 0;
            w->pos = 0;
        }
    }
}

static unsigned ceil_log2_u64(uint64_t b) {  /* tlx::integer_log2_ceil */
    unsigned k = 0;
    while (k < 64 && (1ull << k) < b) ++k;
    return k;
}

/* getDeltaEncoding (src/encoding/golomb_encoding.hpp:60-76) through         */
/* DeltaStreamWriter (src/encoding/delta_stream.hpp:24-44) and                */
/* GolombBitStreamWriter (src/encoding/golomb_bit_stream.hpp:27-136): the     */
/* first value as one raw word, then for every further value the gap g to its */
/* predecessor as q = g / b ones, a zero, and r = g - q*b in log2b - 1 bits   */
/* if r < 2^log2b - b, else r + 2^log2b - b in log2b bits (log2b =            */
/* ceil(log2 b)); the last word is filled up with ones.  Values ascending.    */
/* Returns the words written.                                                 */
/* Pinned by: word-exact vectors from the reference header compiled in       */
/* oracle/_ref (several b, including non-powers of two and long unary runs),  */
/* and the reference's pipelinedGolombEncoding_tests round trip.             */
ORC_API uint64_t orc_golomb_encode(const uint64_t* v, uint64_t n, uint64_t b, uint64_t* out) {
    if (n == 0) return 0;
    orc_bits w = {out, 0, 0, 0};
    out[w.n++] = v[0];
    const unsigned log2b = ceil_log2_u64(b);
    const uint64_t mlv = (log2b == 64 ? 0 : (1ull << log2b)) - b;
    for (uint64_t i = 1; i < n; ++i) {
        const uint64_t g = v[i] - v[i - 1];
        uint64_t q = g / b;
        const uint64_t r = g - q * b;
        for (; q > 0; --q) bits_put(&w, 1, 1);
        bits_put(&w, 0, 1);
        if (r >= mlv)
            bits_put(&w, r + mlv, log2b);
        else
            bits_put(&w, r, log2b - 1);
    }
    if (w.pos != 0) bits_put(&w, ~0ull >> w.pos, 64 - w.pos);
    return w.n;
}

/* getDeltaDecoding (src/encoding/golomb_encoding.hpp:93-109) through         */
/* GolombBitStreamReader (golomb_bit_stream.hpp:142-210) and BitStreamReader */
/* (bit_stream.hpp:113-230): a value follows while a zero bit remains.        */
/* Returns the number of values, or -1 if a code runs past the input or the   */
/* output capacity.  (For b = 2 the reference reader's GetBits(0) shifts a    */
/* 64-bit word by 64, which C++ leaves undefined; this restatement decodes    */
/* the code the writer produced.)                                             */
ORC_API int64_t orc_golomb_decode(const uint64_t* words, uint64_t nwords, uint64_t b,
                                  uint64_t* out, uint64_t cap) {
    if (nwords == 0) return 0;
    if (cap == 0) return -1;
    out[0] = words[0];
    uint64_t n = 1, acc = words[0];
    const uint64_t* bw = words + 1;
    const uint64_t nbits = (nwords - 1) * 64;
    const unsigned log2b = ceil_log2_u64(b);
    const uint64_t mlv = (log2b == 64 ? 0 : (1ull << log2b)) - b;
    uint64_t pos = 0;
#define ORC_BIT(p) ((bw[(p) >> 6] >> (63 - ((p) & 63))) & 1)
    for (;;) {
        uint64_t z = pos;
        while (z < nbits && ORC_BIT(z)) ++z;
        if (z >= nbits) break;  /* only ones remain: the padding */
        const uint64_t q = z - pos;
        uint64_t p = z + 1, r = 0;
        if (log2b > 0) {
            if (p + (log2b - 1) > nbits) return -1;
            for (unsigned k = 0; k + 1 < log2b; ++k) r = (r << 1) | ORC_BIT(p + k);
            p += log2b - 1;
            if (r >= mlv) {
                if (p + 1 > nbits) return -1;
                r = ((r << 1) | ORC_BIT(p)) - mlv;
                p += 1;
            }
        }
        if (n >= cap) return -1;
        acc += q * b + r;
        out[n++] = acc;
        pos = p;
    }
#undef ORC_BIT
    return (int64_t)n;
}

/* ------------------------------------------------------------------------- */
/* The whole DistributedPrefixDoublingSort::sort (prefix_doubling.hpp:204-294) */
/* with p PEs run PE-parallel on CPU threads, as p MPI ranks run it: the      */
/* local sort, the candidate classification and hashing of every round, the  */
/* local duplicate search, the exchange and the merge are per-PE work; the   */
/* global duplicate search of a round is split over hash ranges (the         */
/* reference's FindDuplicates owners, bloomfilter.hpp:662-765).  The CPU     */
/* baseline of the benchmark; the same steps as orc_local_sort,              */
/* orc_prefix_doubling, orc_partition, orc_exchange (mode 2) and             */
/* orc_merge_losertree, and bit-exact with them (tests/test_oracle.py).      */
/* Outputs per merged output position (receivers in order): stringIndex,     */
/* PEIndex and the merged LCP (the PE-boundary LCPs are 0); results per local */
/* sorted position; phase_s (6 entries, may be NULL): wall seconds of local  */
/* sort, prefix doubling, partition, exchange, merge and permutation.        */
/* Returns the number of prefix-doubling rounds.                              */
/* ------------------------------------------------------------------------- */
#include <omp.h>
#include <time.h>

static double wall_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static uint32_t pd_mt(const uint8_t* chars, const uint64_t* off, const uint32_t* len,
                      const uint64_t* pe_begin, uint32_t p, const uint32_t* perm,
                      const uint32_t* lcp, uint64_t m, uint32_t start_depth, uint32_t* results,
                      uint8_t* rounds) {
    const uint64_t n = pe_begin[p];
    const uint64_t nn = n ? n : 1;
    /* per-candidate buffers, grown when the candidate multiset (which may
     * hold a position twice) outgrows them */
    uint64_t cap = nn, ccap = nn;
    uint64_t* cand = (uint64_t*)malloc(ccap * sizeof(uint64_t));
    uint64_t ncand = n;
    pd_rec* rec = (pd_rec*)malloc(cap * sizeof(pd_rec));
    uint8_t* rflag = (uint8_t*)malloc(cap);
    uint64_t* next = (uint64_t*)malloc((cap * 3 + 2 * (uint64_t)p) * sizeof(uint64_t));
    sent_t* sent = (sent_t*)malloc(cap * sizeof(sent_t));
    uint64_t* cb = (uint64_t*)calloc(p + 1, 8);     /* candidate range per PE */
    uint64_t* nrec = (uint64_t*)calloc(p, 8);
    uint64_t* nsent = (uint64_t*)calloc(p, 8);
    uint64_t* nnext = (uint64_t*)calloc(p, 8);
    uint64_t* nb = (uint64_t*)calloc(p + 1, 8);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        cand[i] = i;
        results[i] = 0;
        rounds[i] = 0;
    }
    uint32_t round = 0;
    uint64_t d = start_depth, step = (uint64_t)start_depth * 2;
    while (ncand > 0) {
        if (ncand > cap) {
            cap = ncand;
            free(rec); free(rflag); free(next); free(sent);
            rec = (pd_rec*)malloc(cap * sizeof(pd_rec));
            rflag = (uint8_t*)malloc(cap);
            next = (uint64_t*)malloc((cap * 3 + 2 * (uint64_t)p) * sizeof(uint64_t));
            sent = (sent_t*)malloc(cap * sizeof(sent_t));
        }
        /* candidates of PE r: cand[cb[r], cb[r+1]) (sorted positions) */
        for (uint32_t r = 0, ci = 0; r <= p; ++r) {
            if (r == p) { cb[p] = ncand; break; }
            cb[r] = ci;
            while (ci < ncand && cand[ci] < pe_begin[r + 1]) ++ci;
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t r = 0; r < p; ++r) {
            const uint64_t c0 = cb[r], c1 = cb[r + 1];
            pd_rec* R = rec + c0;       /* at most one record per candidate */
            uint64_t* X = next + 3 * c0 + 2 * (uint64_t)r;
            uint64_t nr_ = 0, nx = 0;
            for (uint64_t t = c0; t < c1; ++t) {
                uint64_t c = cand[t];
                if ((t == c0 || cand[t - 1] != c) && rounds[c] < 255) rounds[c]++;
                results[c] = (uint32_t)d;
                uint32_t id = perm[c];
                if (d > len[id]) {
                    results[c] = len[id];
                    continue;
                }
                if (t > c0 && cand[t - 1] + 1 == c && lcp[c] >= d) {
                    X[nx++] = c;
                    if (nr_ > 0 && R[nr_ - 1].pos + 1 == c) R[nr_ - 1].lcp_root = 1;
                    continue;
                }
                pd_rec x;
                x.hash = orc_pd_hash(chars + off[id], d, m);
                x.pos = c;
                x.pe = r;
                x.local_dup = x.send_anyway = x.lcp_root = 0;
                R[nr_++] = x;
            }
            qsort(R, nr_, sizeof(pd_rec), rec_cmp);
            /* getIndicesOfLocalDuplicates (the single-thread loop, one PE) */
            uint64_t i = 0;
            while (i + 1 < nr_) {
                uint64_t j = i + 1;
                if (R[j].hash == R[i].hash) {
                    R[i].local_dup = R[i].send_anyway = 1;
                    R[j].local_dup = 1;
                    X[nx++] = R[i].pos;
                    X[nx++] = R[j].pos;
                    ++j;
                    while (j < nr_ && R[j].hash == R[i].hash) {
                        R[j].local_dup = 1;
                        X[nx++] = R[j].pos;
                        ++j;
                    }
                } else if (R[i].lcp_root) {
                    R[i].local_dup = R[i].send_anyway = 1;
                    X[nx++] = R[i].pos;
                }
                i = j;
            }
            if (nr_ > 0 && R[nr_ - 1].lcp_root) {
                R[nr_ - 1].local_dup = R[nr_ - 1].send_anyway = 1;
                X[nx++] = R[nr_ - 1].pos;
            }
            uint64_t ns = 0;
            sent_t* Sn = sent + c0;
            for (uint64_t t = 0; t < nr_; ++t) {
                rflag[c0 + t] = 0;
                if (!R[t].local_dup || R[t].send_anyway) {
                    Sn[ns].hash = R[t].hash;
                    Sn[ns].rec = c0 + t;
                    ++ns;
                }
            }
            nrec[r] = nr_;
            nsent[r] = ns;
            nnext[r] = nx;
        }
        /* global duplicates: hash range k owns hashes h with h / w == k;
         * each PE's sent list is sorted by hash, so its part of a range is
         * found by binary search */
        const uint64_t w = UINT64_MAX / p + 1;
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t k = 0; k < p; ++k) {
            uint64_t tot = 0;
            uint64_t* lo = (uint64_t*)malloc(2 * p * sizeof(uint64_t));
            uint64_t* hi = lo + p;
            for (uint32_t r = 0; r < p; ++r) {
                const sent_t* Sn = sent + cb[r];
                uint64_t a = 0, b = nsent[r];
                while (a < b) { uint64_t mid = (a + b) / 2; if (Sn[mid].hash / w < k) a = mid + 1; else b = mid; }
                lo[r] = a;
                b = nsent[r];
                while (a < b) { uint64_t mid = (a + b) / 2; if (Sn[mid].hash / w <= k) a = mid + 1; else b = mid; }
                hi[r] = a;
                tot += hi[r] - lo[r];
            }
            sent_t* G = (sent_t*)malloc((tot ? tot : 1) * sizeof(sent_t));
            uint64_t g = 0;
            for (uint32_t r = 0; r < p; ++r)
                for (uint64_t t = lo[r]; t < hi[r]; ++t) G[g++] = sent[cb[r] + t];
            qsort(G, tot, sizeof(sent_t), sent_cmp);
            for (uint64_t i = 0; i < tot;) {
                uint64_t j = i + 1;
                while (j < tot && G[j].hash == G[i].hash) ++j;
                if (j - i >= 2)
                    for (uint64_t t = i; t < j; ++t)
                        if (!rec[G[t].rec].send_anyway) rflag[G[t].rec] = 1;
                i = j;
            }
            free(G);
            free(lo);
        }
        /* next candidates per PE, sorted, concatenated in PE order */
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t r = 0; r < p; ++r) {
            const uint64_t c0 = cb[r];
            uint64_t* X = next + 3 * c0 + 2 * (uint64_t)r;
            uint64_t nx = nnext[r];
            for (uint64_t t = 0; t < nrec[r]; ++t)
                if (rflag[c0 + t]) X[nx++] = rec[c0 + t].pos;
            qsort(X, nx, sizeof(uint64_t), u64_cmp);
            nnext[r] = nx;
        }
        nb[0] = 0;
        for (uint32_t r = 0; r < p; ++r) nb[r + 1] = nb[r] + nnext[r];
        if (nb[p] > ccap) { /* the old candidates are no longer read */
            ccap = nb[p];
            free(cand);
            cand = (uint64_t*)malloc(ccap * sizeof(uint64_t));
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t r = 0; r < p; ++r)
            memcpy(cand + nb[r], next + 3 * cb[r] + 2 * (uint64_t)r, nnext[r] * sizeof(uint64_t));
        ncand = nb[p];
        ++round;
        d = start_depth + step;
        step *= 2;
        if (round > 60) break;
    }
    free(nb); free(nnext); free(nsent); free(nrec); free(cb);
    free(sent); free(next); free(rflag); free(rec); free(cand);
    return round;
}

ORC_API uint32_t orc_pd_sort_mt(const uint8_t* chars, const uint64_t* off, const uint32_t* len,
                                const uint64_t* pe_begin, uint32_t p, int sampler, uint64_t m,
                                int threads, uint32_t* out_si, uint32_t* out_pe, uint32_t* out_lcp,
                                uint32_t* out_results, double* phase_s, uint64_t* out_D) {
    if (threads > 0) omp_set_num_threads(threads);
    const uint64_t n = pe_begin[p];
    const uint64_t nn = n ? n : 1;
    double t0 = wall_s(), ph[6] = {0, 0, 0, 0, 0, 0};
    uint32_t* perm = (uint32_t*)malloc(nn * 4);
    uint32_t* lcp = (uint32_t*)malloc(nn * 4);
    /* (a2) local sort, one PE per thread */
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t r = 0; r < p; ++r) {
        const uint64_t b = pe_begin[r], mm = pe_begin[r + 1] - b;
        uint32_t* tmp = (uint32_t*)malloc((mm ? mm : 1) * 4);
        sort_ctx c = {chars, off, lcp};
        for (uint64_t i = b; i < b + mm; ++i) {
            perm[i] = (uint32_t)i;
            lcp[i] = 0;
        }
        sort_rec(&c, perm + b, tmp, b, mm, 0);
        free(tmp);
    }
    double t1 = wall_s();
    ph[0] = t1 - t0;
    uint32_t nrounds = 0;
    if (p > 1) {
        uint8_t* rounds = (uint8_t*)malloc(nn);
        nrounds = pd_mt(chars, off, len, pe_begin, p, perm, lcp, m, 8, out_results, rounds);
        free(rounds);
        double t2 = wall_s();
        ph[1] = t2 - t1;
        uint64_t* interval = (uint64_t*)malloc((uint64_t)p * p * 8);
        orc_partition_ex(chars, off, len, pe_begin, p, perm, sampler, 2, 0, out_results, interval, NULL, NULL,
                         NULL, NULL, NULL);
        double t3 = wall_s();
        ph[2] = t3 - t2;
        /* (a13) exchange, one receiver per thread after the size pass */
        uint64_t* rb = (uint64_t*)calloc(p + 1, 8);
        uint64_t* bb = (uint64_t*)calloc(p + 1, 8);
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t q = 0; q < p; ++q)
            bb[q + 1] = exch_receiver(2, q, chars, off, len, pe_begin, p, perm, lcp, out_results, interval, 0, 0,
                                      NULL, NULL, NULL, NULL);
        for (uint32_t q = 0; q < p; ++q) {
            uint64_t c = 0;
            for (uint32_t r = 0; r < p; ++r) c += interval[(uint64_t)r * p + q];
            rb[q + 1] = rb[q] + c;
            bb[q + 1] += bb[q];
        }
        uint8_t* bytes = (uint8_t*)malloc(bb[p] + 1);
        uint32_t* rlcp = (uint32_t*)malloc(nn * 4);
        uint32_t* rpe = (uint32_t*)malloc(nn * 4);
        uint32_t* rpos = (uint32_t*)malloc(nn * 4);
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t q = 0; q < p; ++q)
            exch_receiver(2, q, chars, off, len, pe_begin, p, perm, lcp, out_results, interval, bb[q], rb[q], bytes,
                          rlcp, rpe, rpos);
        double t4 = wall_s();
        ph[3] = t4 - t3;
        /* (a8) compressed LCP loser-tree merge, one receiver per thread */
        uint64_t* order = (uint64_t*)malloc(nn * 8);
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t q = 0; q < p; ++q) {
            const uint64_t r0 = rb[q], cnt = rb[q + 1] - r0;
            if (!cnt) continue;
            const uint8_t** strs = (const uint8_t**)malloc(cnt * sizeof(uint8_t*));
            const uint8_t* a = bytes + bb[q];
            for (uint64_t i = 0; i < cnt; ++i) {
                strs[i] = a;
                while (*a) ++a;
                ++a;
            }
            uint64_t* runb = (uint64_t*)malloc((p + 1) * 8);
            runb[0] = 0;
            for (uint32_t r = 0; r < p; ++r) runb[r + 1] = runb[r] + interval[(uint64_t)r * p + q];
            orc_merge_losertree(strs, rlcp + r0, runb, p, 1, order + r0, out_lcp + r0, NULL);
            free(runb);
            free((void*)strs);
        }
        double t5 = wall_s();
        ph[4] = t5 - t4;
        /* permutation writeback (prefix_doubling.hpp:282-290) */
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t q = 0; q < p; ++q)
            for (uint64_t k = rb[q]; k < rb[q + 1]; ++k) {
                const uint64_t a = rb[q] + order[k];
                out_si[k] = rpos[a];
                out_pe[k] = rpe[a];
            }
        ph[5] = wall_s() - t5;
        /* D of the global output (untimed): the merged LCPs with the LCP
         * across every receiver boundary recomputed from the strings */
        if (out_D) {
            uint32_t* glcp = (uint32_t*)malloc(nn * 4);
            memcpy(glcp, out_lcp, n * 4);
            int64_t prev = -1;
            for (uint32_t q = 0; q < p; ++q) {
                if (rb[q + 1] == rb[q]) continue;
                const uint64_t k = rb[q];
                if (prev >= 0) {
                    const uint8_t* x = chars + off[perm[pe_begin[out_pe[prev]] + out_si[prev]]];
                    const uint8_t* y = chars + off[perm[pe_begin[out_pe[k]] + out_si[k]]];
                    glcp[k] = calc_lcp(x, y);
                }
                prev = (int64_t)rb[q + 1] - 1;
            }
            *out_D = orc_compute_d(glcp, n);
            free(glcp);
        }
        free(order); free(rpos); free(rpe); free(rlcp); free(bytes); free(bb); free(rb); free(interval);
    }
    free(lcp);
    free(perm);
    if (phase_s) memcpy(phase_s, ph, sizeof(ph));
    return nrounds;
}
