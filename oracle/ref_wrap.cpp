// ref_wrap.cpp -- test-only C wrappers around reference headers, compiled
// in place from the reference sources by oracle/build_ref.sh into
// oracle/_ref/libdss_ref.so.  Used by the tests to pin the oracle's
// restatements bit-exactly:
//   - xxh::hash_state_t<64> (src/hash/xxhash.hpp:508-640) as called by
//     XXHasher::hash (src/sorter/distributed/bloomfilter.hpp:1094-1102);
//   - dss_schimek::LcpStringLoserTree_ (src/merge/bingmann-lcp_losertree.hpp:
//     204-480) driven like merge<K> (src/sorter/distributed/merging.hpp:23-63);
//   - StringLcpContainer::extendPrefix (src/strings/stringcontainer.hpp:247-319);
//   - IntegerCompression::writeRanges / readRanges (src/mpi/byte_encoder.hpp:
//     31-78) over PutVarint / GetVarint (src/encoding/integer_compression.hpp);
//   - getB, getDeltaEncoding, getDeltaDecoding (src/encoding/golomb_encoding.hpp:
//     49-109) over the Golomb and delta bit streams.
// The tlx radix sort (insertion_sort.hpp needs tlx/meta/enable_if.hpp, which
// the vendored tlx lacks) and the MPI-bound drivers are not built.
#include <cstdint>
#include <cstring>
#include <vector>

#include "hash/xxhash.hpp"
#include "merge/bingmann-lcp_losertree.hpp"
#include "strings/stringcontainer.hpp"
#include "encoding/golomb_encoding.hpp"
#include "mpi/byte_encoder.hpp"

using SSet = dss_schimek::UCharLengthStringSet;
using Cont = dss_schimek::StringLcpContainer<SSet>;

extern "C" uint64_t ref_xxh_stream(const uint8_t* p, uint64_t len) {
    xxh::hash_state_t<64> s;
    s.update(p, len);
    return s.digest();
}

template <size_t K>
static void run_merge(Cont& cont, std::vector<std::pair<size_t, size_t>>& ranges,
                      size_t n, int compressed, std::vector<SSet::String>& out,
                      std::vector<size_t>& out_lcp, std::vector<size_t>& saved) {
    SSet ss = cont.make_string_set();
    dss_schimek::StringLcpPtrMergeAdapter<SSet> in(ss, cont.lcp_array());
    dss_schimek::LcpStringLoserTree_<K, SSet> lt(in, ranges.data());
    SSet os(out.data(), out.data() + out.size());
    dss_schimek::StringLcpPtrMergeAdapter<SSet> o(os, out_lcp.data());
    if (compressed)
        lt.writeElementsToStream(o, n, saved);
    else
        lt.writeElementsToStream(o, n);
}

// bytes: receive arena of one PE (NUL-terminated strings, possibly prefix
// compressed), lcps: per received string; run_sizes[p].  Outputs per merged
// position: index of the received string, merged lcp, saved lcp.
extern "C" int ref_merge(const uint8_t* bytes, uint64_t nbytes, const uint64_t* lcps,
                         const uint64_t* run_sizes, uint32_t p, int compressed,
                         uint64_t* out_index, uint64_t* out_lcp, uint64_t* out_saved) {
    std::vector<unsigned char> raw(bytes, bytes + nbytes);
    size_t n = 0;
    for (uint32_t i = 0; i < p; ++i) n += run_sizes[i];
    std::vector<size_t> l(lcps, lcps + n);
    Cont cont(std::move(raw), std::move(l));
    std::vector<uint64_t> start(n);
    for (size_t i = 0; i < n; ++i) start[i] = cont[i].string - cont.raw_strings().data();
    std::vector<std::pair<size_t, size_t>> ranges;
    for (uint32_t i = 0, off = 0; i < p; ++i) {
        if (run_sizes[i] == 0) { ranges.emplace_back(0, 0); continue; }
        cont.lcp_array()[off] = 0;
        ranges.emplace_back(off, run_sizes[i]);
        off += run_sizes[i];
    }
    size_t K = 1;
    while (K < p) K *= 2;
    for (size_t i = p; i < K; ++i) ranges.emplace_back(0, 0);
    std::vector<SSet::String> out(n);
    std::vector<size_t> ol(n), saved;
    switch (K) {
        case 1: run_merge<1>(cont, ranges, n, compressed, out, ol, saved); break;
        case 2: run_merge<2>(cont, ranges, n, compressed, out, ol, saved); break;
        case 4: run_merge<4>(cont, ranges, n, compressed, out, ol, saved); break;
        case 8: run_merge<8>(cont, ranges, n, compressed, out, ol, saved); break;
        case 16: run_merge<16>(cont, ranges, n, compressed, out, ol, saved); break;
        case 32: run_merge<32>(cont, ranges, n, compressed, out, ol, saved); break;
        case 64: run_merge<64>(cont, ranges, n, compressed, out, ol, saved); break;
        default: return -1;
    }
    for (size_t i = 0; i < n; ++i) {
        uint64_t s = out[i].string - cont.raw_strings().data();
        size_t lo = 0, hi = n;
        while (lo + 1 < hi) {
            size_t mid = (lo + hi) / 2;
            if (start[mid] <= s) lo = mid; else hi = mid;
        }
        out_index[i] = lo;
        out_lcp[i] = ol[i];
        if (compressed) out_saved[i] = saved[i];
    }
    return 0;
}

// Rebuild full strings: bytes/order/saved as produced by a compressed merge.
// Writes the extended arena into out (size query when out == nullptr).
extern "C" uint64_t ref_extend_prefix(const uint8_t* bytes, uint64_t nbytes,
                                      const uint64_t* order, const uint64_t* saved,
                                      uint64_t n, uint8_t* out) {
    std::vector<unsigned char> raw(bytes, bytes + nbytes);
    Cont cont(std::move(raw));
    std::vector<SSet::String> sorted(n);
    for (uint64_t i = 0; i < n; ++i) sorted[i] = cont[order[i]];
    cont.set(std::move(sorted));
    std::vector<size_t> sv(saved, saved + n);
    cont.extendPrefix(cont.make_string_set(), sv);
    uint64_t total = cont.char_size();
    if (out) std::memcpy(out, cont.raw_strings().data(), total);
    return total;
}

// Varint ranges: writes the encoding into out (capacity 10 bytes per value)
// and the byte count of every range; returns the total bytes.
extern "C" uint64_t ref_varint_write_ranges(const uint64_t* values, const uint64_t* counts,
                                            uint32_t num_ranges, uint8_t* out,
                                            uint64_t* range_bytes) {
    std::vector<uint64_t> c(counts, counts + num_ranges);
    uint64_t n = 0;
    for (uint64_t x : c) n += x;
    std::vector<uint64_t> v(values, values + n);
    auto res = dss_schimek::IntegerCompression::writeRanges(c.begin(), c.end(), v.begin());
    uint64_t total = 0;
    for (uint32_t r = 0; r < num_ranges; ++r) {
        range_bytes[r] = res.counts[r];
        total += res.counts[r];
    }
    std::memcpy(out, res.integers.data(), total);
    return total;
}

extern "C" void ref_varint_read_ranges(const uint8_t* bytes, const uint64_t* counts,
                                       uint32_t num_ranges, uint64_t* values) {
    std::vector<uint64_t> c(counts, counts + num_ranges);
    auto out = dss_schimek::IntegerCompression::readRanges(c.begin(), c.end(), bytes);
    std::memcpy(values, out.data(), out.size() * sizeof(uint64_t));
}

extern "C" uint64_t ref_golomb_b(uint64_t max_value, uint64_t num_values) {
    return getB(max_value, num_values);
}

// Delta + Golomb encoding of ascending values; returns the words written.
extern "C" uint64_t ref_golomb_encode(const uint64_t* v, uint64_t n, uint64_t b, uint64_t* out) {
    std::vector<size_t> in(v, v + n), enc;
    getDeltaEncoding(in.begin(), in.end(), std::back_inserter(enc), b);
    std::memcpy(out, enc.data(), enc.size() * sizeof(uint64_t));
    return enc.size();
}

extern "C" uint64_t ref_golomb_decode(const uint64_t* words, uint64_t nwords, uint64_t b,
                                      uint64_t* out) {
    std::vector<size_t> in(words, words + nwords), dec;
    getDeltaDecoding(in.begin(), in.end(), std::back_inserter(dec), b);
    std::memcpy(out, dec.data(), dec.size() * sizeof(uint64_t));
    return dec.size();
}
