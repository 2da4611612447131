/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * Compiles the reference's own sequential Needleman-Wunsch translation unit
 * (src/pairwise/needleman/impl/sequential.cpp, with src/encoder.cpp) from
 * its sources under the reference tree, with the reference's production
 * flags (-std=c++14 -O3 -DPRODUCTION, makefile:36-39,69-71), and exposes
 * its file-local align_pair (sequential.cpp:29-68) through one C entry point.
 * The build recipe is oracle/Makefile (target `ref`); the output goes to
 * oracle/_ref/, which is never committed.  Tests use it to pin the
 * restatement in nw_oracle.c; bench.py may time it as the CPU baseline.
 *
 * Nothing from the reference is copied here: the translation unit is pulled
 * in by #include from the reference tree at build time.
 */
#include "pairwise/needleman/impl/sequential.cpp"

#include <cstdint>
#include <cstring>

namespace
{
    museqa::sequence make_sequence(const uint16_t* blocks, size_t n)
    {
        auto buf = museqa::encoder::buffer::make(n);
        for (size_t i = 0; i < n; ++i) buf[i] = blocks[i];
        return museqa::sequence {buf};
    }
}

extern "C" __attribute__((visibility("default")))
float ref_align_pair(const uint16_t* one, size_t one_blocks, const uint16_t* two, size_t two_blocks,
                     const float* table625, float penalty)
{
    using namespace museqa::pairwise;
    static thread_local scoring_table::raw_type raw;
    std::memcpy(&raw, table625, sizeof(raw));
    const scoring_table table {museqa::pointer<scoring_table::raw_type>::weak(&raw), penalty};
    return align_pair(make_sequence(one, one_blocks), make_sequence(two, two_blocks), table);
}

extern "C" __attribute__((visibility("default")))
size_t ref_encode(const char* text, size_t size, uint16_t* out, size_t capacity)
{
    const auto encoded = museqa::encoder::encode(text, size);
    for (size_t i = 0; i < encoded.size() && i < capacity; ++i) out[i] = encoded[i];
    return encoded.size();
}

#include <thread>
#include <vector>

/*
 * Scores count pairs of an explicit list with the reference's align_pair,
 * using `workers` threads on contiguous shares (utils::partition, src/utils.hpp:
 * 176-183), the way each MPI rank takes its range in the reference.  The
 * orientation is that of align() (sequential.cpp:87-91).  out[c] = score of
 * pairs[2c], pairs[2c+1].
 */
extern "C" __attribute__((visibility("default")))
int ref_align_list(const uint16_t* blocks, const int64_t* offset, const int64_t* nblocks,
                   const int32_t* pairs, int64_t count, const float* table625, float penalty,
                   float* out, int workers)
{
    if (workers < 1) workers = 1;
    std::vector<museqa::sequence> seqs;
    int64_t nseq = 0;
    for (int64_t c = 0; c < 2 * count; ++c) nseq = std::max<int64_t>(nseq, pairs[c] + 1);
    seqs.reserve(nseq);
    for (int64_t s = 0; s < nseq; ++s) seqs.push_back(make_sequence(blocks + offset[s], (size_t)nblocks[s]));
    auto work = [&](int64_t first, int64_t total) {
        using namespace museqa::pairwise;
        scoring_table::raw_type raw;
        std::memcpy(&raw, table625, sizeof(raw));
        const scoring_table table {museqa::pointer<scoring_table::raw_type>::weak(&raw), penalty};
        for (int64_t c = first; c < first + total; ++c) {
            const museqa::sequence& one = seqs[pairs[2 * c]];
            const museqa::sequence& two = seqs[pairs[2 * c + 1]];
            out[c] = one.size() > two.size() ? align_pair(one, two, table) : align_pair(two, one, table);
        }
    };
    std::vector<std::thread> threads;
    const int64_t quo = count / workers, rem = count % workers;
    for (int r = 0; r < workers; ++r) {
        const int64_t first = quo * r + std::min<int64_t>(rem, r);
        const int64_t total = quo + (rem > r);
        threads.emplace_back(work, first, total);
    }
    for (auto& t : threads) t.join();
    return 0;
}
