// Host planner and the Needleman-Wunsch entry points of the C ABI.
//
// The planner replaces the reference's host batching (src/pairwise/needleman/
// impl/hybrid.cu:262-429: required_memory, needed_cache, load_input,
// make_input, align).  Where the reference cuts the pair list into batches
// that fit free device memory and launches one thread block per pair, a plan
// keeps the whole sequence store and every work list resident in HBM, orders
// the work longest first and issues one persistent launch per kernel family,
// optionally captured in a hipGraph.
//
// Routing (every score keeps the reference's sequential semantics,
// sequential.cpp:29-95):
//   - plain kernel: pairs whose column sequence (reference orientation,
//     :87-91) holds a padding code before its last residue, or whose row
//     sequence runs no row, or whose column sequence has no residue; and every
//     pair under MSA_KERNEL_PLAIN;
//   - long kernel: other pairs whose shorter side reaches the long threshold;
//   - batched kernel: the rest.  For these pairs the reference's rules reduce
//     to the global alignment of the two sequences cut at their first padding
//     code, so either one can be the query (rows; a transposed profile keeps
//     the reference's orientation): the planner takes the orientation the
//     batched kernel runs faster.  The shorter side is the query of a long
//     pair.
#include "nw_planner.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <new>
#include <thread>
#include <utility>

namespace msa {

namespace {

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Estimated lane instructions of the batched kernel for one pair, as
// query_weight(query) * target_steps(target): a query of lq rows under tile
// (G lanes, at most RMAX rows per lane) runs a first strip of R0 rows per lane
// and `full` strips of RMAX, each step costing 2R cell operations plus about
// 5 instructions (one more when the strip stores its bottom row, one more
// when it loads the one above), on G lanes; every strip walks the target's
// ceil((lt + 2G - 2) / G) * G steps.  A short query fills a lane's rows
// poorly; a short target pays each strip's G - 1 steps of skew.  A query that
// will be stacked (`stacks`: at least RMAX rows and at most half a strip)
// pays its share of a full strip: its rows and separator out of G * RMAX.
int64_t query_weight(int64_t lq, const TileConfig& tc, bool stacks)
{
    const int64_t g = tc.lanes, rmax = tc.rows;
    if (stacks && lq >= rmax && 2 * (lq + 1) <= g * rmax) return ((2 * rmax + 5) * (lq + 1) + rmax - 1) / rmax;
    const int64_t full = (lq - 1) / (g * rmax);
    const int64_t rem = lq - full * g * rmax;
    const int64_t step = (int64_t)kRowStep * g;
    const int64_t r0 = std::min(rmax, (rem + step - 1) / step * kRowStep);
    int64_t per_step = 2 * r0 + 5 + (full > 0 ? 1 : 0);
    if (full > 0) per_step += full * (2 * rmax + 6) + (full - 1);
    return per_step * g;
}

int64_t target_steps(int64_t lt, int64_t g) { return (lt + 2 * g - 2) / g * g; }

// Host threads for planning `work` pair visits: one below 2^20 (serving-size
// jobs plan on their submitting thread), else up to 16.
int plan_threads(int64_t work)
{
    if (work < (int64_t(1) << 20)) return 1;
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
}

// Runs f(t) for t = 0..T-1, thread 0 on the caller; false if any call threw
// (an allocation failure), after all have finished.
template <typename F>
bool run_parallel(int T, F&& f)
{
    std::atomic<bool> ok{true};
    auto guarded = [&](int t) {
        try {
            f(t);
        } catch (...) {
            ok = false;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) {
        try {
            pool.emplace_back(guarded, t);
        } catch (...) {
            guarded(t);      // no thread available: run it here
        }
    }
    guarded(0);
    for (auto& th : pool) th.join();
    return ok;
}

}  // namespace

int scan_sequences(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                   std::vector<HostSeq>& out, int64_t& total_blocks)
{
    out.resize(nseq);
    total_blocks = 0;
    for (int32_t s = 0; s < nseq; ++s) {
        if (size[s] < 0 || offset[s] < 0)
            return fail(MSA_ERR_INVALID_ARGUMENT, "sequence %d has a negative offset or size", s);
        total_blocks = std::max<int64_t>(total_blocks, offset[s] + size[s]);
        HostSeq h{offset[s], size[s], 3 * size[s], 0};
        const uint16_t* seq = blocks + offset[s];
        for (int32_t pos = 0; pos < 3 * size[s]; ++pos) {
            const int code = (seq[pos / 3] >> (1 + 5 * (pos % 3))) & 0x1F;
            if (code > kEndCode)
                return fail(MSA_ERR_GAP_CODE, "sequence %d holds code %d at position %d (the reference "
                            "scores only codes 0..24; 0x19 is the gap code)", s, code, pos);
            if (code == kEndCode) h.first_end = std::min(h.first_end, pos);
            else h.last_real = pos + 1;
        }
        out[s] = h;
    }
    if (total_blocks >= (int64_t(1) << 31))
        return fail(MSA_ERR_TOO_LARGE, "sequence store holds %lld blocks; at most 2^31-1 are supported",
                    (long long)total_blocks);
    return MSA_OK;
}

int make_sxx(const int32_t* table625, int32_t penalty, Sxx& sxx, bool& symmetric)
{
    std::memset(&sxx, 0, sizeof(sxx));
    symmetric = true;
    for (int a = 0; a < 25; ++a)
        for (int b = 0; b < 25; ++b) {
            const int64_t v = (int64_t)table625[a * 25 + b] + 2 * (int64_t)penalty;
            if (v < 0 || v > 255)
                return fail(MSA_ERR_TABLE, "table[%d][%d] + 2*penalty = %lld is outside 0..255, which the "
                            "byte profile of the kernels holds", a, b, (long long)v);
            sxx.v[a * 25 + b] = (uint8_t)v;
            if (table625[a * 25 + b] != table625[b * 25 + a]) symmetric = false;
        }
    return MSA_OK;
}

int device_shape(DeviceShape& out)
{
    static std::mutex lock;
    static DeviceShape cache[64];
    int dev = 0;
    MSA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(MSA_ERR_INVALID_ARGUMENT, "device %d out of range", dev);
    std::lock_guard<std::mutex> guard(lock);
    if (!cache[dev].compute_units) {
        hipDeviceProp_t p;
        MSA_HIP(hipGetDeviceProperties(&p, dev));
        cache[dev].compute_units = p.multiProcessorCount;
        cache[dev].lds_per_block = p.sharedMemPerBlock;
        cache[dev].lds_per_cu = p.maxSharedMemoryPerMultiProcessor;
    }
    out = cache[dev];
    return MSA_OK;
}

int resolve_tuning(const msa_nw_options& opt, const Tuning* tuning_override, Tuning& loaded, const Tuning*& out)
{
    out = &builtin_tuning();
    const char* tuning_file = tuning_override ? nullptr : opt.tuning_file;
    if (tuning_override) {
        out = tuning_override;
    } else if (!tuning_file) {
        const char* env = std::getenv("MSA_NW_TUNING");
        if (env && *env) tuning_file = env;
    }
    if (tuning_file) {
        MSA_TRY(load_tuning(tuning_file, loaded));
        out = &loaded;
    }
    return MSA_OK;
}

int long_threshold_of(const msa_nw_options& opt, int32_t* threshold)
{
    if (opt.kernel == MSA_KERNEL_LONG) return *threshold = 0, MSA_OK;
    if (opt.kernel == MSA_KERNEL_PLAIN || opt.kernel == MSA_KERNEL_BATCHED) return *threshold = INT32_MAX, MSA_OK;
    if (opt.long_threshold > 0) return *threshold = opt.long_threshold, MSA_OK;
    Tuning loaded;
    const Tuning* tuning = nullptr;
    MSA_TRY(resolve_tuning(opt, nullptr, loaded, tuning));
    *threshold = tuning->long_threshold;
    return MSA_OK;
}

int plan_host(HostPlan& hp, const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
              const int32_t* pairs, int64_t npairs, const int32_t* table625, int32_t penalty,
              const msa_nw_options& opt, const Tuning* tuning_override, bool copy_blocks)
{
    if (nseq < 0 || npairs < 0) return fail(MSA_ERR_INVALID_ARGUMENT, "negative sizes");
    if (nseq > 0 && (!offset || !size || !blocks)) return fail(MSA_ERR_INVALID_ARGUMENT, "null sequence database");
    if (!table625) return fail(MSA_ERR_INVALID_ARGUMENT, "null table");
    if (!pairs && npairs != (int64_t)nseq * (nseq - 1) / 2)
        return fail(MSA_ERR_INVALID_ARGUMENT, "all-pairs mode needs npairs = nseq*(nseq-1)/2 (%lld), got %lld",
                    (long long)nseq * (nseq - 1) / 2, (long long)npairs);
    if (npairs >= (int64_t(1) << 31))
        return fail(MSA_ERR_TOO_LARGE, "%lld pairs; at most 2^31-1 per plan", (long long)npairs);
    const int kernel = opt.kernel;
    if (kernel < MSA_KERNEL_AUTO || kernel > MSA_KERNEL_LONG)
        return fail(MSA_ERR_INVALID_ARGUMENT, "unknown kernel %d", kernel);
    // Tile shape: forced by the options, else per query-length bin from the
    // tuning table.
    int forced = -1;
    if (opt.lanes_per_pair || opt.rows_per_lane) {
        const int lanes = opt.lanes_per_pair ? opt.lanes_per_pair : 16;
        const int rows = opt.rows_per_lane ? opt.rows_per_lane : (lanes == 16 ? 32 : 16);
        forced = tile_config_index(lanes, rows);
        if (forced < 0)
            return fail(MSA_ERR_INVALID_ARGUMENT, "no batched tile of %d lanes x %d rows per lane", lanes, rows);
    }
    Tuning loaded;
    const Tuning* tuning = nullptr;
    MSA_TRY(resolve_tuning(opt, tuning_override, loaded, tuning));

    std::vector<HostSeq> seqs;
    int64_t total_blocks = 0;
    MSA_TRY(scan_sequences(blocks, offset, size, nseq, seqs, total_blocks));
    bool symmetric = true;
    MSA_TRY(make_sxx(table625, penalty, hp.sxx, symmetric));
    hp.penalty = penalty;

    // Classify pairs (host only), into parts: one for an explicit list, one
    // per planning thread (a contiguous range of queries) for all pairs.
    struct Fast { int32_t query, target, index; };
    struct Part {
        std::vector<Fast> fast, longs;
        std::vector<int32_t> plain_pairs, plain_index;
        int64_t cells = 0;
    };
    std::vector<Part> parts;
    const int64_t long_threshold = kernel == MSA_KERNEL_LONG ? 0
                                 : (opt.long_threshold > 0 ? opt.long_threshold : tuning->long_threshold);
    // Targets of one query are queued longest first: sequences ranked by
    // (rows descending, id descending).
    std::vector<int32_t> by_len(nseq), rank(nseq);
    for (int32_t q = 0; q < nseq; ++q) by_len[q] = q;
    std::sort(by_len.begin(), by_len.end(), [&](int32_t a, int32_t b) {
        return seqs[a].first_end != seqs[b].first_end ? seqs[a].first_end > seqs[b].first_end : a > b;
    });
    for (int32_t r = 0; r < nseq; ++r) rank[by_len[r]] = r;
    // Cost terms of each sequence for the orientation of batched pairs: as a
    // query (weight and lane-group width of its tile) and as a target (steps
    // under 16- and 32-lane groups).
    struct CostTerms { int64_t weight; int32_t lanes; int32_t steps16, steps32; };
    std::vector<CostTerms> cterm(nseq);
    for (int32_t s = 0; s < nseq; ++s) {
        const int cfg = forced >= 0 ? forced : tuning_config_for(*tuning, seqs[s].first_end);
        const int64_t len = seqs[s].first_end;
        cterm[s] = {query_weight(len, kTileConfig[cfg], opt.single_query_strips == 0), kTileConfig[cfg].lanes,
                    (int32_t)target_steps(len, 16),
                    (int32_t)target_steps(len, 32)};
    }
    auto pair_cost = [&](int32_t q, int32_t t) {
        const CostTerms& cq = cterm[q];
        return cq.weight * (cq.lanes == 16 ? cterm[t].steps16 : cterm[t].steps32);
    };
    // Routing of one pair (first, second) as the reference orders it.
    enum { kFast, kLong, kPlain };
    struct Route { int32_t kind, qkey, target; int64_t cells; };
    auto route = [&](int32_t first, int32_t second) -> Route {
        const int32_t rowref = seqs[first].blocks > seqs[second].blocks ? first : second;
        const int32_t colref = rowref == first ? second : first;
        const HostSeq& r = seqs[rowref];
        const HostSeq& col = seqs[colref];
        const int64_t c = (int64_t)r.first_end * col.last_real;
        const bool clean = col.first_end == col.last_real && r.first_end > 0 && col.last_real > 0;
        if (kernel == MSA_KERNEL_PLAIN || !clean) return {kPlain, 0, 0, c};
        const bool is_long = kernel != MSA_KERNEL_BATCHED && std::min(r.first_end, col.last_real) >= long_threshold;
        bool row_query = r.first_end < col.last_real || (r.first_end == col.last_real && rowref < colref);
        if (!is_long) {
            const int64_t as_row = pair_cost(rowref, colref);
            const int64_t as_col = pair_cost(colref, rowref);
            if (as_row != as_col) row_query = as_row < as_col;
        }
        const int32_t q = row_query ? rowref : colref;
        const int32_t t = row_query ? colref : rowref;
        const int32_t transpose = (!symmetric && !row_query) ? 1 : 0;
        return {is_long ? kLong : kFast, q * 2 + transpose, t, c};
    };
    auto add_plain = [](Part& P, int32_t first, int32_t second, int64_t idx) {
        P.plain_pairs.push_back(first);
        P.plain_pairs.push_back(second);
        P.plain_index.push_back((int32_t)idx);
    };
    if (pairs) {
        parts.resize(1);
        Part& P = parts[0];
        std::vector<Fast>& fast = P.fast;
        if (kernel != MSA_KERNEL_PLAIN) fast.reserve((size_t)npairs);
        for (int64_t c = 0; c < npairs; ++c) {
            const int32_t first = pairs[2 * c], second = pairs[2 * c + 1];
            if (first < 0 || first >= nseq || second < 0 || second >= nseq)
                return fail(MSA_ERR_INVALID_ARGUMENT, "pair %lld names sequence outside 0..%d", (long long)c,
                            nseq - 1);
            const Route rt = route(first, second);
            P.cells += rt.cells;
            if (rt.kind == kPlain) add_plain(P, first, second, c);
            else (rt.kind == kLong ? P.longs : fast).push_back({rt.qkey, rt.target, (int32_t)c});
        }
        // Group the batched pairs by query key, targets in rank order, list
        // order within a target: two stable counting passes.
        if (!fast.empty()) {
            std::vector<Fast> tmp(fast.size());
            std::vector<int64_t> start((size_t)2 * nseq + 1, 0);
            for (const Fast& f : fast) ++start[rank[f.target] + 1];
            for (int32_t r = 0; r < nseq; ++r) start[r + 1] += start[r];
            for (const Fast& f : fast) tmp[start[rank[f.target]]++] = f;
            std::fill(start.begin(), start.end(), 0);
            for (const Fast& f : tmp) ++start[f.query + 1];
            for (int32_t k = 0; k < 2 * nseq; ++k) start[k + 1] += start[k];
            for (const Fast& f : tmp) fast[start[f.query]++] = f;
        }
    } else {
        // All pairs, generated query by query with each query's partners in
        // rank order, so the batched list comes out grouped and ordered with
        // no sort.  Every pair is met from both of its sequences: batched and
        // long pairs are taken on their query's side, plain pairs on the side
        // of their larger index.  Threads take contiguous query ranges (every
        // query visits all nseq partners, so ranges of equal size balance).
        const int T = plan_threads((int64_t)nseq * nseq);
        parts.resize(T);
        const bool ok = run_parallel(T, [&](int t) {
            Part& P = parts[t];
            const int32_t q0 = (int32_t)((int64_t)nseq * t / T), q1 = (int32_t)((int64_t)nseq * (t + 1) / T);
            if (kernel != MSA_KERNEL_PLAIN) P.fast.reserve((size_t)npairs / T + 16);
            std::vector<Fast> transposed;
            for (int32_t q = q0; q < q1; ++q) {
                transposed.clear();
                for (int32_t k = 0; k < nseq; ++k) {
                    const int32_t tgt = by_len[k];
                    if (tgt == q) continue;
                    const int32_t first = std::max(q, tgt), second = std::min(q, tgt);
                    const int64_t idx = (int64_t)first * (first - 1) / 2 + second;
                    const Route rt = route(first, second);
                    if (rt.kind == kPlain) {
                        if (q == first) P.cells += rt.cells, add_plain(P, first, second, idx);
                        continue;
                    }
                    if (rt.qkey >> 1 != q) continue;
                    P.cells += rt.cells;
                    const Fast f{rt.qkey, rt.target, (int32_t)idx};
                    if (rt.kind == kLong) P.longs.push_back(f);
                    else if (rt.qkey & 1) transposed.push_back(f);
                    else P.fast.push_back(f);
                }
                P.fast.insert(P.fast.end(), transposed.begin(), transposed.end());
            }
        });
        if (!ok) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
    }
    // Plain and long pairs of every part, in part order.
    std::vector<Fast> longs;
    int64_t cells = 0, nfast = 0;
    std::vector<int64_t> fast_at(parts.size() + 1, 0);     // offset of each part in the target list
    for (size_t t = 0; t < parts.size(); ++t) {
        Part& P = parts[t];
        cells += P.cells;
        hp.plain_pairs.insert(hp.plain_pairs.end(), P.plain_pairs.begin(), P.plain_pairs.end());
        hp.plain_index.insert(hp.plain_index.end(), P.plain_index.begin(), P.plain_index.end());
        longs.insert(longs.end(), P.longs.begin(), P.longs.end());
        std::vector<int32_t>().swap(P.plain_pairs);
        std::vector<int32_t>().swap(P.plain_index);
        fast_at[t + 1] = fast_at[t] + (int64_t)P.fast.size();
    }
    nfast = fast_at.back();
    hp.info.pairs = npairs;
    hp.info.cells = cells;
    hp.info.plain_pairs = (int64_t)hp.plain_index.size();
    hp.info.batched_pairs = nfast;
    hp.info.long_pairs = (int64_t)longs.size();

    DeviceShape dev;
    MSA_TRY(device_shape(dev));

    // Sequence store, with the views rebased as the reference's to_device does
    // (src/pairwise/database.cu:56-72).
    if (copy_blocks) hp.blocks.assign(blocks, blocks + total_blocks);
    hp.seqinfo.resize(nseq);
    for (int32_t s = 0; s < nseq; ++s)
        hp.seqinfo[s] = make_int4((int)seqs[s].offset, seqs[s].blocks, seqs[s].first_end, seqs[s].last_real);

    // Plain part: one line per workgroup, as long as the longest column range.
    if (!hp.plain_index.empty()) {
        int32_t max_cols = 1;
        for (size_t k = 0; k < hp.plain_index.size(); ++k)
            max_cols = std::max({max_cols, seqs[hp.plain_pairs[2 * k]].last_real,
                                 seqs[hp.plain_pairs[2 * k + 1]].last_real});
        hp.plain_grid = plain_grid((int64_t)hp.plain_index.size(), dev.compute_units);
        hp.plain_stride = (max_cols + 63) & ~63;
    }

    // Batched part: group by (query, transpose), sort each query's targets by
    // length, cut into items of up to 16 targets, queue the items costliest
    // first.  Each item carries the tile shape of its query's length bin.
    // Short queries are stacked: several single-strip queries of one tile
    // shape, adjacent in length, share one strip (each below a neutral
    // separator row, its rows offset by its own constant K so that the rows
    // above cannot reach it) against the union of their target lists, with
    // one output index per (member, target) and -1 where a member does not
    // pair with a target.
    if (nfast > 0) {
        // Targets per item: 16 (four rounds of four) when the plan has enough
        // pairs to give every resident wave four items, else 8 or 4, so that
        // small plans (serving jobs) still spread over the whole GPU.
        int per_cu = 0;
        MSA_TRY(batched_occupancy(&per_cu));
        if (per_cu < 1) return fail(MSA_ERR_RUNTIME, "batched kernel does not fit a compute unit");
        const int64_t waves = (int64_t)per_cu * dev.compute_units * (batched_threads() / 64);
        int item_targets = 16;
        while (item_targets > 4 && nfast < (int64_t)item_targets * 4 * waves) item_targets /= 2;
        struct Item { int4 v; int64_t cost; };
        struct Group { int32_t part, key, cfg; int64_t begin, end; };
        struct PartItems {
            std::vector<Item> items;
            std::vector<Group> stackable;
            std::vector<Group> duoable;  // packed query groups that may share duo items
            int64_t kept = 0;           // pairs of the part's groups that are not stacked
            int32_t max_t = 1;
            bool need_lines = false;
            unsigned used = 0;
            int64_t packed = 0;         // pairs of the part's packed items
            int64_t packed_cells = 0;   // their unpadded cells
            bool packed_lines = false;  // some packed query is longer than a strip
        };
        std::vector<PartItems> pitems(parts.size());
        // The largest sxx entry bounds a member's scores: K grows by
        // lq * max_sxx + 1 from one member to the next.
        int32_t max_sxx = 0;
        for (int i = 0; i < 625; ++i) max_sxx = std::max<int32_t>(max_sxx, hp.sxx.v[i]);
        const bool stacking = opt.single_query_strips == 0;
        auto group_cfg = [&](int32_t lq) { return forced >= 0 ? forced : tuning_config_for(*tuning, lq); };
        // A query stacks when it fills a strip poorly on its own (at most half
        // of it) and holds at least a lane's rows, so that a lane sees at most
        // one member's last row and one separator.
        auto stackable = [&](int32_t lq, int cfg) {
            const TileConfig tc = kTileConfig[cfg];
            return stacking && lq >= tc.rows && 2 * ((int64_t)lq + 1) <= (int64_t)tc.lanes * tc.rows;
        };
        // A query takes the packed queue when it fills one 16 x 32 strip on
        // its own (225..512 rows: first-strip heights of 16..32 rows per lane)
        // and is not stacked, and every score of its pairs fits a packed
        // cell: in the shifted domain a strip's cells lie in [0, max(sxx) *
        // rows].
        // Packed items run in a launch of their own: a plan too small to give
        // every resident wave four items of 8 targets keeps its single launch.
        // With duo items (below) the test waits for the pair count.
        bool packing = opt.packed_cells > 0 ||
                       (opt.packed_cells == 0 && (opt.duo_queries >= 0 || item_targets >= 8));
        // A query longer than one strip is packed when its rebased strips'
        // cells fit (at most kPackedSpan x max(sxx) above their base).
        auto packable = [&](int32_t lq, int cfg) {
            const TileConfig tc = kTileConfig[cfg];
            if (!packing || cfg != 0 || lq < kPackedMinRows) return false;
            if (lq <= tc.lanes * tc.rows) return !stackable(lq, cfg) && (int64_t)max_sxx * lq <= kPackedMaxScore;
            return (int64_t)max_sxx * kPackedSpan <= kPackedMaxScore;
        };
        // Packed queries go in pairs to the duo kernel (two queries against
        // one target per lane group, about 1.1 VALU operations per cell
        // against 1.6 in the packed kernel) when their target lists agree,
        // from kDuoMinRows rows on (also queries that would stack).
        const bool duo = packing && opt.duo_queries >= 0;
        auto duo_ok = [&](int32_t lq, int cfg) {
            if (!packing || cfg != 0 || lq < kDuoMinRows) return false;
            if (lq <= 16 * 32) return (int64_t)max_sxx * lq <= kPackedMaxScore;
            return (int64_t)max_sxx * kPackedSpan <= kPackedMaxScore;
        };
        auto item_cost = [&](int64_t rows_in, const TileConfig& tc, const int2* tg, int32_t cnt) {
            const int groups = 64 / tc.lanes;
            const int64_t step = (int64_t)kRowStep * tc.lanes;
            const int64_t rows = (rows_in + step - 1) / step * step;
            int64_t cost = 0;
            // A round of `groups` targets costs as much as its longest (the first).
            for (int32_t u = 0; u < cnt; u += groups) {
                const int64_t chunks = (seqs[tg[u].x].first_end + 2 * tc.lanes - 2) / tc.lanes;
                cost += rows * chunks * tc.lanes;
            }
            return cost;
        };
        bool ok = run_parallel((int)parts.size(), [&](int t) {
            const std::vector<Fast>& fast = parts[t].fast;
            PartItems& out = pitems[t];
            size_t k = 0;
            while (k < fast.size()) {
                size_t e = k;
                while (e < fast.size() && fast[e].query == fast[k].query) ++e;
                const int32_t lq = seqs[fast[k].query >> 1].first_end;
                const int cfg = group_cfg(lq);
                if (duo && duo_ok(lq, cfg)) out.duoable.push_back({t, fast[k].query, cfg, (int64_t)k, (int64_t)e});
                else if (stackable(lq, cfg)) out.stackable.push_back({t, fast[k].query, cfg, (int64_t)k, (int64_t)e});
                else out.kept += (int64_t)(e - k);
                k = e;
            }
        });
        if (!ok) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
        // Duo items (32 pairs each) run a workgroup each: a plan with fewer
        // than 2 of them per resident workgroup scores its packable queries
        // in int32 cells unless packing is asked for (its few, long duo
        // items would set the time: uneven sets of 400 sequences, 1.6 duo
        // items per workgroup, ran 15% slower packed, of 480, 2.2 per
        // workgroup, 34% faster).
        if (duo && opt.packed_cells == 0) {
            int64_t cand_pairs = 0;
            for (const PartItems& pi : pitems)
                for (const Group& g : pi.duoable) cand_pairs += g.end - g.begin;
            if (cand_pairs < (int64_t)64 * per_cu * dev.compute_units) {
                packing = false;
                for (PartItems& pi : pitems) {
                    for (const Group& g : pi.duoable) {
                        if (stackable(seqs[g.key >> 1].first_end, g.cfg)) pi.stackable.push_back(g);
                        else pi.kept += g.end - g.begin;
                    }
                    pi.duoable.clear();
                }
            }
        }
        // Stacks: the stackable groups of each tile shape in rank order (length
        // descending), packed greedily while the strip holds them.
        std::vector<Group> cand;
        for (const PartItems& pi : pitems) cand.insert(cand.end(), pi.stackable.begin(), pi.stackable.end());
        std::stable_sort(cand.begin(), cand.end(), [&](const Group& x, const Group& y) {
            if (x.cfg != y.cfg) return x.cfg < y.cfg;
            const int32_t rx = rank[x.key >> 1], ry = rank[y.key >> 1];
            return rx != ry ? rx < ry : x.key < y.key;
        });
        struct Stack { size_t first; int32_t m, rows, cfg; int64_t uni, rec, at, souts; };
        std::vector<Stack> stacks;
        for (size_t g = 0; g < cand.size(); ++g) {
            const TileConfig tc = kTileConfig[cand[g].cfg];
            const int32_t lq = seqs[cand[g].key >> 1].first_end;
            if (!stacks.empty()) {
                Stack& st = stacks.back();
                if (st.cfg == cand[g].cfg && st.m < kMaxStackMembers &&
                    (int64_t)st.rows + 1 + lq <= (int64_t)tc.lanes * tc.rows) {
                    st.rows += 1 + lq;
                    ++st.m;
                    continue;
                }
            }
            stacks.push_back({g, 1, lq, cand[g].cfg, 0, 0, 0, 0});
        }
        // Union of each stack's target lists, in rank order; the j-th
        // occurrence of a target in a member's list (an explicit list can
        // name a pair twice) is its own union entry.
        auto merge = [&](const Stack& st, int2* tg_out, int32_t* so_out) -> int64_t {
            int64_t pos[kMaxStackMembers], occ[kMaxStackMembers];
            for (int i = 0; i < st.m; ++i) pos[i] = cand[st.first + i].begin, occ[i] = 0;
            int64_t u = 0;
            while (true) {
                int best = -1;
                int32_t brank = 0;
                int64_t bocc = 0;
                for (int i = 0; i < st.m; ++i) {
                    const Group& G = cand[st.first + i];
                    if (pos[i] >= G.end) continue;
                    const int32_t r = rank[parts[G.part].fast[pos[i]].target];
                    if (best < 0 || r < brank || (r == brank && occ[i] < bocc)) best = i, brank = r, bocc = occ[i];
                }
                if (best < 0) return u;
                const int32_t tgt = parts[cand[st.first + best].part].fast[pos[best]].target;
                if (tg_out) tg_out[u] = make_int2(tgt, (int)u);
                for (int i = 0; i < st.m; ++i) {
                    const Group& G = cand[st.first + i];
                    const std::vector<Fast>& f = parts[G.part].fast;
                    const bool hit = pos[i] < G.end && f[pos[i]].target == tgt && occ[i] == bocc;
                    if (so_out) so_out[(int64_t)i * st.uni + u] = hit ? f[pos[i]].index : -1;
                    if (hit) {
                        ++pos[i];
                        occ[i] = pos[i] < G.end && f[pos[i]].target == tgt ? occ[i] + 1 : 0;
                    }
                }
                ++u;
            }
        };
        ok = run_parallel(std::min<int>((int)parts.size(), std::max<int>(1, (int)stacks.size())), [&](int t) {
            const int T = std::min<int>((int)parts.size(), std::max<int>(1, (int)stacks.size()));
            for (size_t s = (size_t)t; s < stacks.size(); s += (size_t)T) stacks[s].uni = merge(stacks[s], nullptr, nullptr);
        });
        if (!ok) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
        // Duo pairs: the duo candidates by orientation (a query's transposed
        // and plain groups hold different targets), then in rank order
        // (length descending), each paired with the next when the pair's padded work (the union
        // of the two target lists, both queries at the longer one's rows)
        // stays within 1.3 times the work of their own pairs.  A query left
        // alone takes duo items of its own (their high halves idle) when such
        // queries hold at most a tenth of the pairs, else the packed kernel.
        std::vector<Group> dcand;
        for (const PartItems& pi : pitems) dcand.insert(dcand.end(), pi.duoable.begin(), pi.duoable.end());
        std::stable_sort(dcand.begin(), dcand.end(), [&](const Group& x, const Group& y) {
            if ((x.key & 1) != (y.key & 1)) return (x.key & 1) < (y.key & 1);
            const int32_t rx = rank[x.key >> 1], ry = rank[y.key >> 1];
            return rx != ry ? rx < ry : x.key < y.key;
        });
        // The union of two target lists in rank order (the j-th occurrence of
        // a target in one list meets its j-th occurrence in the other):
        // emit(target, output index in A or -1, output index in B or -1).
        auto duo_merge = [&](const Group& A, const Group* B, auto&& emit) -> int64_t {
            const std::vector<Fast>& fa = parts[A.part].fast;
            const std::vector<Fast>* fb = B ? &parts[B->part].fast : nullptr;
            int64_t pa = A.begin, pb = B ? B->begin : 0;
            const int64_t eb = B ? B->end : 0;
            int64_t u = 0;
            while (pa < A.end || pb < eb) {
                const int32_t ra = pa < A.end ? rank[fa[pa].target] : INT32_MAX;
                const int32_t rb = pb < eb ? rank[(*fb)[pb].target] : INT32_MAX;
                int32_t tgt = 0, ia = -1, ib = -1;
                if (ra <= rb) tgt = fa[pa].target, ia = fa[pa].index, ++pa;
                if (rb <= ra) tgt = (*fb)[pb].target, ib = (*fb)[pb].index, ++pb;
                emit(tgt, ia, ib);
                ++u;
            }
            return u;
        };
        std::vector<int64_t> next_uni(dcand.size(), 0);   // union of candidates i and i + 1
        const int TD = std::min<int>((int)parts.size(), std::max<int>(1, (int)dcand.size()));
        ok = run_parallel(TD, [&](int t) {
            for (size_t i = (size_t)t; i + 1 < dcand.size(); i += (size_t)TD)
                next_uni[i] = duo_merge(dcand[i], &dcand[i + 1], [](int32_t, int32_t, int32_t) {});
        });
        if (!ok) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
        struct Duo { size_t a, b; int64_t uni, rec; };    // b == a: a query alone
        std::vector<Duo> duos;
        std::vector<size_t> alone;
        int64_t paired = 0, alone_pairs = 0;
        for (size_t i = 0; i < dcand.size();) {
            const int64_t na = dcand[i].end - dcand[i].begin;
            if (i + 1 < dcand.size()) {
                const int64_t la = seqs[dcand[i].key >> 1].first_end, lb = seqs[dcand[i + 1].key >> 1].first_end;
                const int64_t nb = dcand[i + 1].end - dcand[i + 1].begin;
                if (20 * next_uni[i] * std::max(la, lb) <= 13 * (na * la + nb * lb)) {
                    duos.push_back({i, i + 1, next_uni[i], 0});
                    paired += na + nb;
                    i += 2;
                    continue;
                }
            }
            alone.push_back(i);
            alone_pairs += na;
            ++i;
        }
        const bool alone_in_duo = !duos.empty() && 10 * alone_pairs <= paired;
        std::vector<uint8_t> in_duo(dcand.empty() ? 0 : (size_t)2 * nseq, 0);
        // (A short query alone always takes a duo item: the stacks are made.)
        for (const size_t i : alone) {
            if (alone_in_duo || stackable(seqs[dcand[i].key >> 1].first_end, dcand[i].cfg))
                duos.push_back({i, i, dcand[i].end - dcand[i].begin, 0});
            else pitems[dcand[i].part].kept += dcand[i].end - dcand[i].begin;
        }
        for (const Duo& d : duos) {
            in_duo[dcand[d.a].key] = 1;
            in_duo[dcand[d.b].key] = 1;
        }
        constexpr int kDuoTargets = 16;    // targets per duo item: one per lane group of a workgroup
        constexpr int64_t kDuoItemInts2 = 1 + 2 * kDuoTargets;
        // Target list: the parts' unstacked groups in part order, then per
        // stack its record, its union of targets and its output table (one
        // row per member, in int32 units of the list), then per duo its items
        // (each a record and two entries per target).
        std::vector<int64_t> kept_at(parts.size() + 1, 0);
        for (size_t t = 0; t < parts.size(); ++t) kept_at[t + 1] = kept_at[t] + pitems[t].kept;
        int64_t ntargets = kept_at.back();
        for (Stack& st : stacks) {
            st.rec = ntargets;
            st.at = st.rec + 1 + 2 * (int64_t)st.m;
            st.souts = 2 * (st.at + st.uni);
            ntargets = st.at + st.uni + ((int64_t)st.m * st.uni + 1) / 2;
        }
        for (Duo& d : duos) {
            d.rec = ntargets;
            ntargets += (d.uni + kDuoTargets - 1) / kDuoTargets + 2 * d.uni;
        }
        if (2 * ntargets >= (int64_t(1) << 31))
            return fail(MSA_ERR_TOO_LARGE, "stacked target lists exceed 2^30 entries");
        hp.targets.resize((size_t)ntargets);
        // Every part holds whole queries, so parts build their items alone;
        // an item's start is its offset in the whole target list.
        ok = run_parallel((int)parts.size(), [&](int t) {
            const std::vector<Fast>& fast = parts[t].fast;
            PartItems& out = pitems[t];
            int64_t at = kept_at[t];
            size_t k = 0;
            while (k < fast.size()) {
                size_t e = k;
                while (e < fast.size() && fast[e].query == fast[k].query) ++e;
                const int32_t q = fast[k].query >> 1;
                const int32_t lq = seqs[q].first_end;
                const int cfg = group_cfg(lq);
                if (stackable(lq, cfg) || (!in_duo.empty() && in_duo[fast[k].query])) {
                    k = e;
                    continue;
                }
                const TileConfig tc = kTileConfig[cfg];
                out.used |= 1u << cfg;
                out.need_lines = out.need_lines || lq > tc.lanes * tc.rows;
                const bool pk = packable(lq, cfg);
                if (pk) out.packed += (int64_t)(e - k);
                if (pk && lq > tc.lanes * tc.rows) out.packed_lines = true;
                const int32_t flags = (fast[k].query & 1) | ((pk ? kPackedQueue : cfg) << 1);
                const int64_t first = at;
                for (size_t u = k; u < e; ++u) {
                    hp.targets[at++] = make_int2(fast[u].target, fast[u].index);
                    out.max_t = std::max(out.max_t, seqs[fast[u].target].first_end);
                    if (pk) out.packed_cells += (int64_t)lq * seqs[fast[u].target].first_end;
                }
                for (int64_t s = first; s < at; s += item_targets) {
                    const int32_t cnt = (int32_t)std::min<int64_t>(item_targets, at - s);
                    out.items.push_back({make_int4(q, (int)s, cnt, flags), item_cost(lq, tc, &hp.targets[s], cnt)});
                }
                k = e;
            }
        });
        if (!ok) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
        // Stacks: records, unions, output tables and items (stack s's items
        // go to the part s % T).
        const int TS = std::min<int>((int)parts.size(), std::max<int>(1, (int)stacks.size()));
        const bool pack_stacks = !duos.empty();
        std::vector<std::vector<Item>> sitems((size_t)TS);
        std::vector<int64_t> spairs((size_t)TS, 0), spacked((size_t)TS, 0), spacked_cells((size_t)TS, 0);
        std::vector<int32_t> smax_t((size_t)TS, 1);
        std::vector<unsigned> sused((size_t)TS, 0);
        ok = run_parallel(TS, [&](int t) {
            for (size_t s = (size_t)t; s < stacks.size(); s += (size_t)TS) {
                const Stack& st = stacks[s];
                const TileConfig tc = kTileConfig[st.cfg];
                int32_t* outs = reinterpret_cast<int32_t*>(hp.targets.data()) + st.souts;
                if ((int64_t)st.m * st.uni % 2) outs[(int64_t)st.m * st.uni] = -1;   // the last int2's padding
                merge(st, &hp.targets[st.at], outs);
                hp.targets[st.rec] = make_int2(st.m, st.rows);
                int64_t K = 0;
                for (int i = 0; i < st.m; ++i) {
                    const Group& G = cand[st.first + i];
                    const int32_t lq = seqs[G.key >> 1].first_end;
                    hp.targets[st.rec + 1 + 2 * i] = make_int2(G.key >> 1, (int)K);
                    hp.targets[st.rec + 2 + 2 * i] = make_int2((int)(st.souts + (int64_t)i * st.uni), lq * 2 + (G.key & 1));
                    K += (int64_t)lq * max_sxx + 1;
                    spairs[t] += G.end - G.begin;
                }
                // With duo items in the plan, a 16 x 32 stack whose values fit
                // 16 bits (K grows by lq * max_sxx + 1 a member, and a
                // member's cells by at most lq * max_sxx) takes packed cells,
                // two targets per lane group, in the duo launch.
                const bool pk = pack_stacks && st.cfg == 0 &&
                                (int64_t)max_sxx * st.rows + st.m <= kPackedMaxScore;
                if (pk)
                    for (int i = 0; i < st.m; ++i) {
                        const Group& G = cand[st.first + i];
                        const std::vector<Fast>& f = parts[G.part].fast;
                        const int64_t lq = seqs[G.key >> 1].first_end;
                        spacked[t] += G.end - G.begin;
                        for (int64_t u = G.begin; u < G.end; ++u) spacked_cells[t] += lq * seqs[f[u].target].first_end;
                    }
                sused[t] |= 1u << st.cfg;
                const int32_t flags = ((pk ? kPackedQueue : st.cfg) << 1) | kStackedFlag;
                for (int64_t u = 0; u < st.uni; u += item_targets) {
                    const int32_t cnt = (int32_t)std::min<int64_t>(item_targets, st.uni - u);
                    const int2* tg = &hp.targets[st.at + u];
                    for (int32_t v = 0; v < cnt; ++v) smax_t[t] = std::max(smax_t[t], seqs[tg[v].x].first_end);
                    sitems[t].push_back({make_int4((int)st.rec, (int)(st.at + u), cnt, flags),
                                         item_cost(st.rows, tc, tg, cnt)});
                }
            }
        });
        if (!ok) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
        for (int t = 0; t < TS; ++t) {
            PartItems& pi = pitems[(size_t)t];
            pi.items.insert(pi.items.end(), sitems[t].begin(), sitems[t].end());
            pi.max_t = std::max(pi.max_t, smax_t[t]);
            pi.used |= sused[t];
            pi.packed += spacked[t];
            pi.packed_cells += spacked_cells[t];
            hp.info.stacked_pairs += spairs[t];
        }
        // Duos: items (duo d's go to the part d % TD), the union entries of
        // both queries, and the duo item records.
        std::vector<PartItems> ditems((size_t)TD);
        ok = run_parallel(TD, [&](int t) {
            PartItems& out = ditems[(size_t)t];
            for (size_t s = (size_t)t; s < duos.size(); s += (size_t)TD) {
                const Duo& d = duos[s];
                const Group& A = dcand[d.a];
                const Group* B = d.b != d.a ? &dcand[d.b] : nullptr;
                const int32_t la = seqs[A.key >> 1].first_end, lb = B ? seqs[B->key >> 1].first_end : 0;
                const int64_t lmax = std::max(la, lb);
                int64_t u = 0;
                duo_merge(A, B, [&](int32_t tgt, int32_t ia, int32_t ib) {
                    const int64_t base = d.rec + (u / kDuoTargets) * kDuoItemInts2 + 2 * (u % kDuoTargets);
                    hp.targets[base + 1] = make_int2(tgt, ia);
                    hp.targets[base + 2] = make_int2(tgt, ib);
                    const int32_t lt = seqs[tgt].first_end;
                    out.max_t = std::max(out.max_t, lt);
                    if (ia >= 0) ++out.packed, out.packed_cells += (int64_t)la * lt;
                    if (ib >= 0) ++out.packed, out.packed_cells += (int64_t)lb * lt;
                    ++u;
                });
                const int32_t flags = (A.key & 1) | (kDuoQueue << 1);
                const int64_t rows = (lmax + 2 * 16 - 1) / (2 * 16) * (2 * 16);
                for (int64_t j = 0; j * kDuoTargets < d.uni; ++j) {
                    const int64_t rec = d.rec + j * kDuoItemInts2;
                    const int32_t cnt = (int32_t)std::min<int64_t>(kDuoTargets, d.uni - j * kDuoTargets);
                    hp.targets[rec] = B ? make_int2(B->key >> 1, B->key & 1) : make_int2(-1, 0);
                    const int64_t chunks = (seqs[hp.targets[rec + 1].x].first_end + 2 * 16 - 2) / 16;
                    out.items.push_back({make_int4(A.key >> 1, (int)rec, cnt, flags), rows * chunks * 16});
                }
                out.used |= 1u;
                // The duo kernel's waves own two lines per lane group (also
                // for the 16 x 32 items it takes on).
                out.packed_lines = true;
                if (lmax > 16 * 32) out.need_lines = true;
            }
        });
        if (!ok) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
        for (PartItems& di : ditems) {
            hp.info.duo_pairs += di.packed;
            hp.info.duo_cells += di.packed_cells;
            pitems[0].items.insert(pitems[0].items.end(), di.items.begin(), di.items.end());
            pitems[0].max_t = std::max(pitems[0].max_t, di.max_t);
            pitems[0].used |= di.used;
            pitems[0].need_lines = pitems[0].need_lines || di.need_lines;
            pitems[0].packed += di.packed;
            pitems[0].packed_cells += di.packed_cells;
            pitems[0].packed_lines = pitems[0].packed_lines || di.packed_lines;
        }
        for (Part& P : parts) std::vector<Fast>().swap(P.fast);
        std::vector<Item> items;
        int32_t max_t = 1;
        bool need_lines = false;
        unsigned used = 0;
        for (const PartItems& pi : pitems) {
            items.insert(items.end(), pi.items.begin(), pi.items.end());
            max_t = std::max(max_t, pi.max_t);
            need_lines = need_lines || pi.need_lines;
            used |= pi.used;
            hp.info.packed_pairs += pi.packed;
            hp.info.packed_cells += pi.packed_cells;
            hp.packed_lines = hp.packed_lines || pi.packed_lines;
        }
        // One queue per tile configuration and the packed queue (each has its
        // own kernel), every queue costliest first.  A stable sort, so the order is unique: the
        // planning threads sort contiguous ranges, then adjacent ranges are
        // merged (stably) level by level.
        auto before = [](const Item& x, const Item& y) {
            const int cx = (x.v.w >> 1) & kQueueMask, cy = (y.v.w >> 1) & kQueueMask;
            return cx != cy ? cx < cy : x.cost > y.cost;
        };
        const int T = (int)parts.size();
        if (T > 1) {
            std::vector<size_t> cut((size_t)T + 1);
            for (int t = 0; t <= T; ++t) cut[t] = items.size() * (size_t)t / (size_t)T;
            bool sorted = run_parallel(T, [&](int t) {
                std::stable_sort(items.begin() + cut[t], items.begin() + cut[t + 1], before);
            });
            for (int width = 1; sorted && width < T; width *= 2) {
                const int merges = (T + 2 * width - 1) / (2 * width);
                sorted = run_parallel(merges, [&](int m) {
                    const int a = 2 * width * m, b = std::min(a + width, T), e = std::min(a + 2 * width, T);
                    if (b < e) std::inplace_merge(items.begin() + cut[a], items.begin() + cut[b], items.begin() + cut[e], before);
                });
            }
            if (!sorted) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
        } else {
            std::stable_sort(items.begin(), items.end(), before);
        }
        hp.items.resize(items.size());
        for (size_t u = 0; u < items.size(); ++u) {
            hp.items[u] = items[u].v;
            ++hp.batch_first[((items[u].v.w >> 1) & kQueueMask) + 1];
        }
        for (int c = 0; c < kBatchQueues; ++c) {
            const int64_t n = hp.batch_first[c + 1];
            hp.batch_first[c + 1] += hp.batch_first[c];
            hp.batch_grid[c] = (int)std::min<int64_t>((int64_t)per_cu * dev.compute_units, (n + 3) / 4);
        }
        // Lines are touched only by queries of more than one strip.
        hp.batch_stride = need_lines ? batched_line_stride(max_t) : 0;
        hp.info.bins = __builtin_popcount(used);
    }

    // Long part: compact alphabet of the codes these pairs use, strip height
    // from the profile size, one pair per workgroup queued longest first.
    if (!longs.empty()) {
        bool present[32] = {};
        int32_t max_t = 1;
        for (const auto& f : longs) {
            for (const int32_t sidx : {f.query >> 1, f.target}) {
                const HostSeq& h = seqs[sidx];
                for (int32_t pos = 0; pos < h.first_end; ++pos)
                    present[(blocks[h.offset + pos / 3] >> (1 + 5 * (pos % 3))) & 0x1F] = true;
            }
            max_t = std::max(max_t, seqs[f.target].first_end);
        }
        int slots = 0;
        for (int code = 0; code < 32; ++code) hp.slot_code[code] = 0xFF;
        for (int code = 0; code < kEndCode; ++code)
            if (present[code]) {
                hp.remap[code] = (uint8_t)slots;
                hp.slot_code[slots] = (uint8_t)code;
                ++slots;
            }
        hp.long_slots = slots + 1;     // the neutral slot is last
        for (int code = 0; code < 32; ++code)
            if (code >= kEndCode || !present[code]) hp.remap[code] = (uint8_t)slots;
        hp.long_rows = long_rows_for_slots(hp.long_slots);
        const int strip = long_strip_rows(hp.long_rows);
        hp.long_stride = long_line_stride(max_t);
        // One buffer descriptor spans one line of 8-byte words: its byte size
        // must fit 31 bits.
        if ((int64_t)hp.long_stride * 8 >= (int64_t(1) << 31))
            return fail(MSA_ERR_TOO_LARGE, "a long pair's target of %d residues exceeds the boundary line's "
                        "limit of %d", max_t, (int)((int64_t(1) << 28) - kLongLineGuard - 64));
        hp.long_lds = long_lds_bytes(hp.long_rows, hp.long_slots);
        if (hp.long_lds > dev.lds_per_block)
            return fail(MSA_ERR_TOO_LARGE, "long pairs need %zu bytes of LDS per workgroup", hp.long_lds);
        int per_cu = 0;
        MSA_TRY(long_occupancy(hp.long_rows, hp.long_lds, &per_cu));
        if (per_cu < 1) return fail(MSA_ERR_RUNTIME, "long kernel does not fit a compute unit");
        // Resident one-wave workgroups per CU: a multiple of 4, so every SIMD
        // holds as many waves and the strips of a pair run at one speed; and
        // about one per long pair (at least 2 per SIMD), so that a strip's
        // producer is about a round of the grid ahead of it and the strips do
        // not run in lockstep, each waiting on the one above at every chunk
        // (measured on 64 x 30,000 DNA: 14 per CU 139 ms, 12 per CU 122 ms,
        // 8 per CU 110 ms).  Residency is capped by reserving LDS.
        if (per_cu >= 8 && dev.lds_per_cu > 0) {
            const int64_t per_pair = ((int64_t)longs.size() + dev.compute_units - 1) / dev.compute_units;
            const int want = (int)std::min<int64_t>(per_cu / 4 * 4, std::max<int64_t>(8, (per_pair + 3) / 4 * 4));
            const size_t capped = dev.lds_per_cu / (size_t)want / 512 * 512;
            if (want < per_cu && capped >= hp.long_lds && capped <= dev.lds_per_block) {
                int got = 0;
                MSA_TRY(long_occupancy(hp.long_rows, capped, &got));
                if (got == want) {
                    hp.long_lds = capped;
                    per_cu = got;
                }
            }
        }
        std::vector<std::pair<int64_t, int4>> order;
        for (const auto& f : longs)
            order.push_back({(int64_t)seqs[f.query >> 1].first_end * seqs[f.target].first_end,
                             make_int4(f.query >> 1, f.target, f.index, f.query & 1)});
        std::stable_sort(order.begin(), order.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        hp.long_pairs.resize(order.size());
        int64_t ntasks = 0;
        std::vector<int32_t> nst(order.size());
        for (size_t u = 0; u < order.size(); ++u) {
            hp.long_pairs[u] = order[u].second;
            nst[u] = (seqs[order[u].second.x].first_end + strip - 1) / strip;
            ntasks += nst[u];
        }
        if (ntasks >= (int64_t(1) << 31))
            return fail(MSA_ERR_TOO_LARGE, "long pairs make %lld strips; at most 2^31-1", (long long)ntasks);
        hp.long_grid = (int)std::min<int64_t>((int64_t)per_cu * dev.compute_units, ntasks);
        // Tasks: the pairs (costliest first) in groups as large as the grid,
        // each group level by level, so that the strip above a task comes
        // about a group earlier in the list (groups of a quarter of the grid
        // ran 5% slower).  Strips with a strip below them get the next line
        // number; a task records the line it reads (its producer's) and the
        // line it writes.
        const int64_t group = std::max<int64_t>(1, (int64_t)hp.long_grid);
        hp.long_tasks.reserve((size_t)ntasks);
        std::vector<int32_t> line_of(order.size(), -1);   // the last line each pair wrote
        std::vector<int64_t> task_of_line;                // task index of every line's writer
        int64_t max_gap = 0;                              // longest writer -> reader distance in the list
        int32_t nlines_used = 0;
        for (size_t g0 = 0; g0 < order.size(); g0 += (size_t)group) {
            const size_t g1 = std::min(order.size(), g0 + (size_t)group);
            int32_t levels = 0;
            for (size_t u = g0; u < g1; ++u) levels = std::max(levels, nst[u]);
            for (int32_t s = 0; s < levels; ++s)
                for (size_t u = g0; u < g1; ++u) {
                    if (s >= nst[u]) continue;
                    const int32_t in = s > 0 ? line_of[u] : -1;
                    const int32_t out = s + 1 < nst[u] ? nlines_used++ : -1;
                    const int64_t here = (int64_t)hp.long_tasks.size();
                    if (in >= 0) max_gap = std::max(max_gap, here - task_of_line[in]);
                    if (out >= 0) task_of_line.push_back(here);
                    line_of[u] = out;
                    hp.long_tasks.push_back(make_int4((int)u, s, in, out));
                }
        }
        // A writer waits for the reader of the line nlines numbers before its
        // own; that reader comes earlier in the list when the ring is longer
        // than every writer -> reader distance.
        hp.long_nlines = (int32_t)std::max<int64_t>(1, std::min<int64_t>(
            std::max<int64_t>((int64_t)kLongLinesPerWave * hp.long_grid + group, max_gap + 1), nlines_used));
        hp.long_occupants = nlines_used > 0 ? (nlines_used + hp.long_nlines - 1) / hp.long_nlines : 0;
    }
    int batched_launches = 0;
    for (int c = 0; c < kBatchQueues; ++c) batched_launches += hp.batch_grid[c] > 0 ? 2 : 0;
    if (hp.batch_grid[kDuoQueue] > 0) {   // the 16 x 32 and packed queues ride in the duo launch
        if (hp.batch_grid[0] > 0) batched_launches -= 2;
        if (hp.batch_grid[kPackedQueue] > 0) batched_launches -= 2;
    }
    hp.info.launches = (int32_t)(batched_launches + (hp.long_pairs.empty() ? 0 : 2) +
                                 (hp.plain_index.empty() ? 0 : 1));
    return MSA_OK;
}

DataLayout data_layout(const HostPlan& hp)
{
    DataLayout lay;
    size_t at = 0;
    auto place = [&](size_t& field, size_t bytes) { field = at; at = align256(at + std::max<size_t>(bytes, 1)); };
    place(lay.blocks, hp.blocks.size() * sizeof(uint16_t));
    place(lay.seqinfo, hp.seqinfo.size() * sizeof(SeqInfo));
    place(lay.plain_pairs, hp.plain_pairs.size() * sizeof(int32_t));
    place(lay.plain_index, hp.plain_index.size() * sizeof(int32_t));
    place(lay.items, hp.items.size() * sizeof(int4));
    place(lay.targets, hp.targets.size() * sizeof(int2));
    place(lay.long_pairs, hp.long_pairs.size() * sizeof(int4));
    place(lay.long_tasks, hp.long_tasks.size() * sizeof(int4));
    lay.bytes = at;
    return lay;
}

void pack_data(const HostPlan& hp, const DataLayout& lay, void* host_block)
{
    char* base = static_cast<char*>(host_block);
    // Copies of 8 MiB and more (the target list of a large plan) run on 8
    // threads.
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (bytes < (size_t(8) << 20)) {
            if (bytes) std::memcpy(base + off, src, bytes);
            return;
        }
        const char* from = static_cast<const char*>(src);
        const bool ok = run_parallel(8, [&](int t) {
            const size_t a = bytes * (size_t)t / 8, e = bytes * (size_t)(t + 1) / 8;
            std::memcpy(base + off + a, from + a, e - a);
        });
        if (!ok) std::memcpy(base + off, src, bytes);
    };
    put(lay.blocks, hp.blocks.data(), hp.blocks.size() * sizeof(uint16_t));
    put(lay.seqinfo, hp.seqinfo.data(), hp.seqinfo.size() * sizeof(SeqInfo));
    put(lay.plain_pairs, hp.plain_pairs.data(), hp.plain_pairs.size() * sizeof(int32_t));
    put(lay.plain_index, hp.plain_index.data(), hp.plain_index.size() * sizeof(int32_t));
    put(lay.items, hp.items.data(), hp.items.size() * sizeof(int4));
    put(lay.targets, hp.targets.data(), hp.targets.size() * sizeof(int2));
    put(lay.long_pairs, hp.long_pairs.data(), hp.long_pairs.size() * sizeof(int4));
    put(lay.long_tasks, hp.long_tasks.data(), hp.long_tasks.size() * sizeof(int4));
}

ScratchNeed scratch_need(const HostPlan& hp)
{
    ScratchNeed n;
    auto up64 = [](size_t v) { return (v + 63) & ~size_t(63); };
    if (!hp.plain_index.empty()) n.plain_lines = up64((size_t)hp.plain_grid * hp.plain_stride);
    if (!hp.items.empty()) {
        // The configurations' launches run one after another on the stream
        // and share the lines.
        const int grid = *std::max_element(hp.batch_grid, hp.batch_grid + kBatchQueues);
        n.batch_lines = up64((size_t)grid * batched_groups_per_block() * hp.batch_stride * (hp.packed_lines ? 2 : 1));
    }
    if (!hp.long_pairs.empty()) {
        n.long_lines = up64((size_t)2 * hp.long_nlines * hp.long_stride);
        n.long_slots = up64((size_t)hp.long_nlines);
    }
    return n;
}

Scratch carve_scratch(const ScratchNeed& need, int32_t* base)
{
    // Every region is a multiple of 64 ints, so each starts 256-byte aligned.
    Scratch s;
    s.batch_queue = base;
    s.long_queue = base + 64;
    s.long_error = base + 128;
    s.long_tag = reinterpret_cast<uint32_t*>(base + 192);
    int32_t* at = base + 256;
    s.long_turn = at;
    at += need.long_slots;
    s.plain_lines = at;
    at += need.plain_lines;
    s.batch_lines = at;
    at += need.batch_lines;
    s.long_lines = at;
    s.long_lines64 = reinterpret_cast<uint64_t*>(at);
    return s;
}

int launch_plan(const HostPlan& hp, const DataLayout& lay, const void* d_data, const Scratch& scr,
                int32_t* d_out, hipStream_t stream, const uint16_t* d_store)
{
    const char* base = static_cast<const char*>(d_data);
    const uint16_t* d_blocks = d_store ? d_store : reinterpret_cast<const uint16_t*>(base + lay.blocks);
    const SeqInfo* d_info = reinterpret_cast<const SeqInfo*>(base + lay.seqinfo);
    // With duo items, the packed and the 16 x 32 queues ride in the duo launch.
    const bool duo = hp.batch_first[kDuoQueue + 1] > hp.batch_first[kDuoQueue];
    auto count = [&](int c) { return (int32_t)(hp.batch_first[c + 1] - hp.batch_first[c]); };
    const int4* d_items = reinterpret_cast<const int4*>(base + lay.items);
    for (int c = 0; c < kBatchQueues; ++c) {
        const int32_t n = count(c);
        if (n == 0 || (duo && (c == 0 || c == kPackedQueue))) continue;
        const bool with = duo && c == kDuoQueue;
        const int grid = with ? std::max({hp.batch_grid[c], hp.batch_grid[0], hp.batch_grid[kPackedQueue]})
                              : hp.batch_grid[c];
        MSA_TRY(batched_launch(c, grid, d_blocks, d_info, d_items + hp.batch_first[c], n,
                               reinterpret_cast<const int2*>(base + lay.targets), scr.batch_queue + 8 * c, d_out,
                               scr.batch_lines, std::max(hp.batch_stride, 1), hp.sxx, hp.penalty, stream,
                               with ? d_items + hp.batch_first[0] : nullptr, with ? count(0) : 0,
                               with ? scr.batch_queue : nullptr,
                               with ? d_items + hp.batch_first[kPackedQueue] : nullptr, with ? count(kPackedQueue) : 0,
                               with ? scr.batch_queue + 8 * kPackedQueue : nullptr));
    }
    if (!hp.long_pairs.empty()) {
        LongLaunch l;
        l.rows = hp.long_rows;
        l.grid = hp.long_grid;
        l.slots = hp.long_slots;
        l.lds = hp.long_lds;
        l.blocks = d_blocks;
        l.info = d_info;
        l.pairs = reinterpret_cast<const int4*>(base + lay.long_pairs);
        l.tasks = reinterpret_cast<const int4*>(base + lay.long_tasks);
        l.ntasks = (int32_t)hp.long_tasks.size();
        l.queue = scr.long_queue;
        l.error = scr.long_error;
        l.out = d_out;
        l.lines = scr.long_lines64;
        l.turn = scr.long_turn;
        l.tag_word = scr.long_tag;
        l.line_stride = hp.long_stride;
        l.nlines = hp.long_nlines;
        l.occupants = hp.long_occupants;
        l.remap = hp.remap;
        l.slot_code = hp.slot_code;
        l.sxx = hp.sxx;
        l.penalty = hp.penalty;
        MSA_TRY(long_launch(l, stream));
    }
    if (!hp.plain_index.empty())
        MSA_TRY(plain_launch(d_blocks, d_info, reinterpret_cast<const int32_t*>(base + lay.plain_pairs),
                             reinterpret_cast<const int32_t*>(base + lay.plain_index),
                             (int64_t)hp.plain_index.size(), hp.sxx, hp.penalty, d_out, scr.plain_lines,
                             std::max(hp.plain_stride, 1), hp.plain_grid, stream));
    return MSA_OK;
}

}  // namespace msa

struct msa_nw_plan_opaque {
    msa::HostPlan hp;
    msa::DataLayout lay;
    msa::ScratchNeed need;
    msa::Scratch scratch;
    void* d_data = nullptr;
    int32_t* d_scratch = nullptr;
    bool use_graph = false;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t* graph_out = nullptr;
    // Every execution uses the plan's one scratch block (queue counters,
    // strip-boundary lines, the error flag), so executions are serialized:
    // each records `last_done` on its stream, and an execution on another
    // stream first waits for it.
    std::mutex lock;
    hipEvent_t last_done = nullptr;
    hipStream_t last_stream = nullptr;
    bool has_last = false;

    ~msa_nw_plan_opaque()
    {
        if (last_done) (void)hipEventDestroy(last_done);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        if (d_data) (void)hipFree(d_data);
        if (d_scratch) (void)hipFree(d_scratch);
    }
};

using msa::fail;

extern "C" {

int msa_nw_plan_create(msa_nw_plan* plan, const uint16_t* blocks, const int64_t* offset, const int32_t* size,
                       int32_t nseq, const int32_t* pairs, int64_t npairs, const int32_t* table625, int32_t penalty,
                       const msa_nw_options* options, msa_stream stream)
{
    msa_nw_options opt{};
    if (options) opt = *options;
    return msa::create_plan(plan, blocks, offset, size, nseq, pairs, npairs, table625, penalty, opt, nullptr,
                            (hipStream_t)stream);
}

}  // extern "C"

int msa::create_plan(msa_nw_plan* plan, const uint16_t* blocks, const int64_t* offset, const int32_t* size,
                     int32_t nseq, const int32_t* pairs, int64_t npairs, const int32_t* table625, int32_t penalty,
                     const msa_nw_options& opt, const Tuning* tuning, hipStream_t stream)
{
    if (!plan) return fail(MSA_ERR_INVALID_ARGUMENT, "plan is null");
    *plan = nullptr;
    auto* p = new (std::nothrow) msa_nw_plan_opaque();
    if (!p) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed");
    int rc = MSA_OK;
    try {
        rc = msa::plan_host(p->hp, blocks, offset, size, nseq, pairs, npairs, table625, penalty, opt, tuning);
        if (rc == MSA_OK) {
            p->lay = msa::data_layout(p->hp);
            p->need = msa::scratch_need(p->hp);
            // pack_data writes every array; the alignment gaps between them
            // are never read, so the block is not cleared first.
            std::unique_ptr<char[]> staging(new char[p->lay.bytes]);
            msa::pack_data(p->hp, p->lay, staging.get());
            rc = msa::hip_status(hipMalloc(&p->d_data, p->lay.bytes), "hipMalloc");
            if (rc == MSA_OK)
                rc = msa::hip_status(hipMalloc(&p->d_scratch, p->need.words() * sizeof(int32_t)), "hipMalloc");
            // The scratch starts zeroed (the long kernel's tag counter and
            // lines); this is plan creation, never inside a captured graph.
            if (rc == MSA_OK)
                rc = msa::hip_status(hipMemsetAsync(p->d_scratch, 0, p->need.words() * sizeof(int32_t), stream),
                                     "hipMemsetAsync");
            if (rc == MSA_OK)
                rc = msa::hip_status(hipMemcpyAsync(p->d_data, staging.get(), p->lay.bytes, hipMemcpyHostToDevice,
                                                    stream), "hipMemcpyAsync");
            if (rc == MSA_OK) rc = msa::hip_status(hipStreamSynchronize(stream), "hipStreamSynchronize");
            p->scratch = msa::carve_scratch(p->need, p->d_scratch);
            p->hp.info.device_bytes = (int64_t)(p->lay.bytes + p->need.words() * sizeof(int32_t));
            p->use_graph = opt.use_graph != 0;
        }
    } catch (const std::bad_alloc&) {
        rc = fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
    }
    if (rc != MSA_OK) {
        delete p;
        return rc;
    }
    *plan = p;
    return MSA_OK;
}

extern "C" {

int msa_nw_plan_execute(msa_nw_plan plan, int32_t* d_out, msa_stream stream)
{
    if (!plan || (!d_out && plan->hp.info.pairs)) return fail(MSA_ERR_INVALID_ARGUMENT, "null plan or output");
    if (!plan->hp.info.pairs) return MSA_OK;
    const hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> guard(plan->lock);
    // Under a caller's capture the order is the caller's graph's business (a
    // replay must not overlap another execution of this plan); otherwise wait
    // for the previous execution when it ran on another stream.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    MSA_HIP(hipStreamIsCapturing(s, &capture));
    const bool eager = capture == hipStreamCaptureStatusNone;
    if (eager && plan->has_last && plan->last_stream != s) MSA_HIP(hipStreamWaitEvent(s, plan->last_done, 0));
    MSA_TRY(msa::execute_locked(plan, d_out, s));
    if (eager) {
        if (!plan->last_done) MSA_HIP(hipEventCreateWithFlags(&plan->last_done, hipEventDisableTiming));
        MSA_HIP(hipEventRecord(plan->last_done, s));
        plan->last_stream = s;
        plan->has_last = true;
    }
    return MSA_OK;
}

}  // extern "C"

int msa::execute_locked(msa_nw_plan plan, int32_t* d_out, hipStream_t s)
{
    if (!plan->use_graph) return msa::launch_plan(plan->hp, plan->lay, plan->d_data, plan->scratch, d_out, s);
    if (plan->graph_out != d_out || !plan->exec) {
        if (plan->exec) { (void)hipGraphExecDestroy(plan->exec); plan->exec = nullptr; }
        if (plan->graph) { (void)hipGraphDestroy(plan->graph); plan->graph = nullptr; }
        if (!plan->capture_stream) MSA_HIP(hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking));
        MSA_HIP(hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal));
        const int rc = msa::launch_plan(plan->hp, plan->lay, plan->d_data, plan->scratch, d_out, plan->capture_stream);
        hipGraph_t g = nullptr;
        const hipError_t end = hipStreamEndCapture(plan->capture_stream, &g);
        if (rc != MSA_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        MSA_HIP(end);
        plan->graph = g;
        MSA_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
        plan->graph_out = d_out;
    }
    return msa::hip_status(hipGraphLaunch(plan->exec, s), "hipGraphLaunch");
}

extern "C" {

int msa_nw_plan_info_get(msa_nw_plan plan, msa_nw_plan_info* info)
{
    if (!plan || !info) return fail(MSA_ERR_INVALID_ARGUMENT, "null plan or info");
    *info = plan->hp.info;
    return MSA_OK;
}

int msa_nw_plan_status(msa_nw_plan plan, msa_stream stream)
{
    if (!plan) return fail(MSA_ERR_INVALID_ARGUMENT, "plan is null");
    MSA_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (!plan->hp.long_pairs.empty()) {
        int32_t err = 0;
        MSA_HIP(hipMemcpy(&err, plan->scratch.long_error, sizeof(err), hipMemcpyDeviceToHost));
        if (err) return fail(MSA_ERR_RUNTIME, "long kernel: a strip waited past its limit for the strip above");
    }
    return MSA_OK;
}

int msa_nw_plan_destroy(msa_nw_plan plan)
{
    delete plan;
    return MSA_OK;
}

int msa_nw_align(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                 const int32_t* pairs, int64_t npairs, const int32_t* table625, int32_t penalty,
                 const msa_nw_options* options, int32_t* out)
{
    if (npairs > 0 && !out) return fail(MSA_ERR_INVALID_ARGUMENT, "out is null");
    msa_nw_plan plan = nullptr;
    MSA_TRY(msa_nw_plan_create(&plan, blocks, offset, size, nseq, pairs, npairs, table625, penalty, options, nullptr));
    int32_t* d_out = nullptr;
    int rc = MSA_OK;
    if (npairs > 0) {
        rc = msa::hip_status(hipMalloc(&d_out, (size_t)npairs * sizeof(int32_t)), "hipMalloc");
        if (rc == MSA_OK) rc = msa_nw_plan_execute(plan, d_out, nullptr);
        if (rc == MSA_OK) rc = msa_nw_plan_status(plan, nullptr);
        if (rc == MSA_OK)
            rc = msa::hip_status(hipMemcpy(out, d_out, (size_t)npairs * sizeof(int32_t), hipMemcpyDeviceToHost),
                                 "hipMemcpy");
        if (d_out) (void)hipFree(d_out);
    }
    msa_nw_plan_destroy(plan);
    return rc;
}

}  // extern "C"

// ---- msa_nw_plain on caller-owned device memory -----------------------------

namespace {

// One workspace per (device, stream): calls on different streams never share
// memory, calls on one stream run in stream order.  A workspace grows on
// demand (hipFree waits for the device, so no earlier launch still reads the
// old buffers) and lives until the process ends.
struct PlainWorkspace {
    std::mutex lock;
    msa::SeqInfo* info = nullptr;
    size_t info_cap = 0;
    int32_t* lines = nullptr;
    size_t lines_cap = 0;
};

std::mutex g_workspaces_lock;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<PlainWorkspace>> g_workspaces;

PlainWorkspace& workspace_for(int dev, hipStream_t stream)
{
    std::lock_guard<std::mutex> guard(g_workspaces_lock);
    auto& slot = g_workspaces[{dev, stream}];
    if (!slot) slot.reset(new PlainWorkspace());
    return *slot;
}

template <typename T>
int grow(T*& ptr, size_t& cap, size_t need)
{
    if (need <= cap) return MSA_OK;
    if (ptr) MSA_HIP(hipFree(ptr));
    ptr = nullptr;
    cap = 0;
    void* p = nullptr;
    MSA_HIP(hipMalloc(&p, need * sizeof(T)));
    ptr = static_cast<T*>(p);
    cap = need;
    return MSA_OK;
}

}  // namespace

extern "C" int msa_nw_plain(const uint16_t* d_blocks, const int64_t* d_offset, const int32_t* d_size, int32_t nseq,
                            int32_t max_blocks, const int32_t* d_pairs, int64_t npairs, const int32_t* table625,
                            int32_t penalty, int32_t* d_out, msa_stream stream)
{
    if (npairs == 0) return MSA_OK;
    if (npairs < 0 || nseq < 1 || max_blocks < 0 || !d_blocks || !d_offset || !d_size || !d_out || !table625)
        return fail(MSA_ERR_INVALID_ARGUMENT, "invalid arguments to msa_nw_plain");
    if (!d_pairs && npairs != (int64_t)nseq * (nseq - 1) / 2)
        return fail(MSA_ERR_INVALID_ARGUMENT, "all-pairs mode needs npairs = nseq*(nseq-1)/2");
    msa::Sxx sxx;
    bool symmetric = true;
    MSA_TRY(msa::make_sxx(table625, penalty, sxx, symmetric));
    int dev = 0;
    MSA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(MSA_ERR_INVALID_ARGUMENT, "device %d out of range", dev);
    msa::DeviceShape shape;
    MSA_TRY(msa::device_shape(shape));
    const hipStream_t s = (hipStream_t)stream;
    // A captured graph would keep the workspace's pointers, which a later,
    // larger call frees.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    MSA_HIP(hipStreamIsCapturing(s, &capture));
    if (capture != hipStreamCaptureStatusNone)
        return fail(MSA_ERR_INVALID_ARGUMENT, "msa_nw_plain allocates its workspace and cannot be captured; "
                    "capture msa_nw_plan_execute instead");
    PlainWorkspace& ws = workspace_for(dev, s);
    std::lock_guard<std::mutex> guard(ws.lock);
    const int grid = msa::plain_grid(npairs, shape.compute_units);
    const int32_t stride = (int32_t)(((int64_t)3 * max_blocks + 63) & ~int64_t(63));
    MSA_TRY(grow(ws.info, ws.info_cap, (size_t)nseq));
    MSA_TRY(grow(ws.lines, ws.lines_cap, (size_t)grid * std::max(stride, 1)));
    MSA_TRY(msa::seq_info_launch(d_blocks, d_offset, d_size, nseq, ws.info, s));
    return msa::plain_launch(d_blocks, ws.info, d_pairs, nullptr, npairs, sxx, penalty, d_out, ws.lines,
                             std::max(stride, 1), grid, s);
}
