/*
 * dss_gpu_adaptors.hpp -- header-only C++ adaptors for the reference side
 * (bingmann/distributed-string-sorting): drop-ins for its call sites over the
 * C ABI of dss_gpu.h.  Include after the reference's own headers; nothing
 * here depends on MPI.
 *
 *   dss_gpu::radixsort_CI3(strptr, depth, memory)
 *       for tlx::sort_strings_detail::radixsort_CI3
 *       (TLXV/sort/strings/radix_sort.hpp:860-880) at
 *       src/sorter/distributed/lcp_aware_number_string_splitting.hpp:218-224
 *       and src/sorter/distributed/prefix_doubling.hpp:228: same signature,
 *       strings of the StringLcpPtr permuted into lexicographic order and
 *       lcp[i] = LCP with the predecessor (lcp[0] untouched);
 *   dss_gpu::hash_round(set, ids, depth, out)
 *       XXHasher::hash (src/sorter/distributed/bloomfilter.hpp:1094-1102) of
 *       one prefix-doubling round's candidates in one call;
 *   dss_gpu::prefix_doubling_sort<Entry>(bytes, rank_byte_begin, sampler, m)
 *       DistributedPrefixDoublingSort::sort (prefix_doubling.hpp:204-294) for
 *       all ranks of one node as virtual PEs of one GPU; Entry is the
 *       reference's StringIndexPEIndex (src/util/structs.hpp:9-14) or any
 *       type constructible from (stringIndex, PEIndex).
 */
#ifndef DSS_GPU_ADAPTORS_HPP
#define DSS_GPU_ADAPTORS_HPP

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "dss_gpu.h"

namespace dss_gpu {

inline void check(int rc) {
    if (rc != DSS_OK) throw std::runtime_error(std::string("dss: ") + dss_last_error());
}

// RAII owner of a dss_set.
class Set {
public:
    explicit Set(dss_set h) : h_(h) {}
    Set(const Set&) = delete;
    Set& operator=(const Set&) = delete;
    ~Set() {
        if (h_) (void)dss_set_destroy(h_);
    }
    dss_set get() const { return h_; }

private:
    dss_set h_;
};

template <typename StringLcpPtr>
void radixsort_CI3(const StringLcpPtr& strptr, size_t depth, size_t /*memory*/) {
    if (depth != 0) throw std::invalid_argument("dss: radixsort_CI3 drop-in sorts from depth 0");
    auto ss = strptr.active();
    const size_t n = ss.size();
    if (n < 2) return;
    // the set's strings in their current order are the ids 0..n-1
    std::vector<typename StringLcpPtr::StringSet::String> strs(ss.begin(), ss.end());
    std::vector<unsigned char> chars;
    for (auto& s : strs) {
        const auto* c = ss.get_chars(s, 0);
        chars.insert(chars.end(), c, c + ss.get_length(s) + 1);  // with the NUL
    }
    const uint64_t pe_bytes[2] = {0, chars.size()};
    dss_set h = nullptr;
    check(dss_set_create_pes(chars.data(), pe_bytes, 1, &h));
    Set set(h);
    std::vector<uint32_t> perm(n), lcp(n);
    check(dss_sort_lcp(h));
    check(dss_get_sorted(h, perm.data(), lcp.data()));
    for (size_t i = 0; i < n; ++i) {
        ss[ss.begin() + i] = strs[perm[i]];
        if (i) strptr.set_lcp(i, lcp[i]);
    }
}

// ids: positions of the round's hashed candidates in the set.
inline void hash_round(dss_set h, const std::vector<uint32_t>& ids, uint32_t depth, std::vector<uint64_t>& out) {
    out.resize(ids.size());
    check(dss_hash_prefixes(h, ids.data(), ids.size(), depth, std::numeric_limits<uint64_t>::max(), out.data()));
}

// all_ranks: every rank's NUL-terminated strings back to back;
// rank_byte_begin: p + 1 byte boundaries; sampler: 0 NumStrings, 1 NumChars,
// 2 IndexedNumStrings, 3 IndexedNumChars; m: the hash range.
template <typename Entry>
std::vector<Entry> prefix_doubling_sort(const unsigned char* all_ranks, const std::vector<uint64_t>& rank_byte_begin,
                                        int sampler = 2, uint64_t m = std::numeric_limits<uint64_t>::max()) {
    if (rank_byte_begin.size() < 2) throw std::invalid_argument("dss: need at least one rank");
    const uint32_t p = (uint32_t)(rank_byte_begin.size() - 1);
    dss_set h = nullptr;
    check(dss_set_create_pes(all_ranks, rank_byte_begin.data(), p, &h));
    Set set(h);
    check(dss_prefix_doubling_sort(h, sampler, m));
    dss_set_info_t info;
    check(dss_set_info(h, &info));
    std::vector<Entry> perm;
    if (p == 1) return perm;  // one PE: no distributed sort (prefix_doubling.hpp:237)
    std::vector<uint32_t> si(info.num_strings), pi(info.num_strings);
    check(dss_get_permutation(h, si.data(), pi.data()));
    perm.reserve(si.size());
    for (size_t k = 0; k < si.size(); ++k) perm.emplace_back(si[k], pi[k]);
    return perm;  // concatenated over the receiving PEs, each PE's output in order
}

}  // namespace dss_gpu

#endif  // DSS_GPU_ADAPTORS_HPP
