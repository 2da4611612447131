/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * Compiles the reference's own sequential Needleman-Wunsch translation unit
 * (src/pairwise/needleman/impl/sequential.cpp, with src/encoder.cpp and
 * src/database.cpp) from its sources under the reference tree, with the
 * reference's production flags (-std=c++14 -O3 -DPRODUCTION,
 * makefile:36-39,69-71), and exposes through C entry points:
 *   - its file-local align_pair (sequential.cpp:29-68), as given (no
 *     orientation), for the restatement's per-pair recurrence;
 *   - its file-local align (sequential.cpp:77-95), which picks the
 *     orientation itself, over a museqa::database and a buffer<pair> built
 *     from the caller's arrays, with the pairs cut into contiguous shares the
 *     way each MPI rank takes its range.
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
#include <string>
#include <thread>
#include <vector>

namespace
{
    museqa::sequence make_sequence(const uint16_t* blocks, size_t n)
    {
        auto buf = museqa::encoder::buffer::make(n);
        for (size_t i = 0; i < n; ++i) buf[i] = blocks[i];
        return museqa::sequence {buf};
    }

    museqa::pairwise::scoring_table make_table(museqa::pairwise::scoring_table::raw_type& raw,
                                               const float* table625, float penalty)
    {
        using namespace museqa::pairwise;
        std::memcpy(&raw, table625, sizeof(raw));
        return scoring_table {museqa::pointer<scoring_table::raw_type>::weak(&raw), penalty};
    }
}

extern "C" __attribute__((visibility("default")))
float ref_align_pair(const uint16_t* one, size_t one_blocks, const uint16_t* two, size_t two_blocks,
                     const float* table625, float penalty)
{
    static thread_local museqa::pairwise::scoring_table::raw_type raw;
    const auto table = make_table(raw, table625, penalty);
    return align_pair(make_sequence(one, one_blocks), make_sequence(two, two_blocks), table);
}

extern "C" __attribute__((visibility("default")))
size_t ref_encode(const char* text, size_t size, uint16_t* out, size_t capacity)
{
    const auto encoded = museqa::encoder::encode(text, size);
    for (size_t i = 0; i < encoded.size() && i < capacity; ++i) out[i] = encoded[i];
    return encoded.size();
}

/*
 * Scores count pairs of an explicit list with the reference's align()
 * (sequential.cpp:77-95) over a museqa::database of nseq sequences (sequence
 * s = blocks[offset[s], offset[s] + nblocks[s])), using `workers` threads on
 * contiguous shares (utils::partition, src/utils.hpp:176-183), the way each
 * MPI rank takes its range.  out[c] = score of pairs[2c], pairs[2c+1].
 * Every thread reads the database through const references only.
 */
extern "C" __attribute__((visibility("default")))
int ref_align_list(const uint16_t* blocks, const int64_t* offset, const int64_t* nblocks, int64_t nseq,
                   const int32_t* pairs, int64_t count, const float* table625, float penalty,
                   float* out, int workers)
{
    using namespace museqa::pairwise;
    if (workers < 1) workers = 1;
    for (int64_t c = 0; c < 2 * count; ++c)
        if (pairs[c] < 0 || pairs[c] >= nseq) return 1;
    museqa::database db((size_t)nseq);
    for (int64_t s = 0; s < nseq; ++s)
        db.add("seq" + std::to_string(s), make_sequence(blocks + offset[s], (size_t)nblocks[s]));
    auto work = [&](int64_t first, int64_t total) {
        scoring_table::raw_type raw;
        const auto table = make_table(raw, table625, penalty);
        auto share = museqa::buffer<pair>::make((size_t)total);
        for (int64_t k = 0; k < total; ++k) {
            share[k].first = pairs[2 * (first + k)];
            share[k].second = pairs[2 * (first + k) + 1];
        }
        const auto scores = align(share, db, table);
        for (int64_t k = 0; k < total; ++k) out[first + k] = scores[k];
    };
    std::vector<std::thread> threads;
    const int64_t quo = count / workers, rem = count % workers;
    for (int r = 0; r < workers; ++r) {
        const int64_t first = quo * r + std::min<int64_t>(rem, r);
        const int64_t total = quo + (rem > r);
        if (total > 0) threads.emplace_back(work, first, total);
    }
    for (auto& t : threads) t.join();
    return 0;
}
