// Host planner and the Needleman-Wunsch entry points of the C ABI.
//
// The planner replaces the reference's host batching (src/pairwise/needleman/
// impl/hybrid.cu:262-429: required_memory, needed_cache, load_input,
// make_input, align).  Where the reference cuts the pair list into batches
// that fit free device memory and launches one thread block per pair, the
// plan keeps the whole sequence store and every work list resident in HBM,
// sorts pairs by shape and issues one persistent launch per length bin,
// optionally captured in a hipGraph.
//
// Routing (every score keeps the reference's sequential semantics):
//   - batched kernel: pairs whose column sequence (reference orientation,
//     sequential.cpp:87-91) has no padding code before its last residue and
//     whose row sequence has at least one row; their score is the global
//     alignment of the two sequences cut at their first padding code;
//   - plain kernel: every other pair (internal padding codes, empty sides),
//     and every pair when MSA_KERNEL_PLAIN is asked for.
#include "msa_internal.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <vector>

namespace msa {

int plain_launch(const uint16_t* d_blocks, const SeqInfo* d_info, const int32_t* d_pairs,
                 const int32_t* d_out_index, int64_t npairs, const Sxx& sxx, int32_t penalty,
                 int32_t* d_out, int32_t* d_lines, int32_t line_stride, int grid, hipStream_t stream);
int plain_grid(int64_t npairs, int compute_units);
int seq_info_launch(const uint16_t* d_blocks, const int64_t* d_offset, const int32_t* d_size, int32_t nseq,
                    SeqInfo* d_info, hipStream_t stream);
int batched_threads();
size_t batched_lds_bytes();
int batched_groups_per_block();
int batched_strip_rows();
int batched_occupancy(int* blocks_per_cu);
int batched_launch(int grid, const uint16_t* d_blocks, const SeqInfo* d_info, const int4* d_items, int32_t nitems,
                   const int2* d_targets, int32_t* d_queue, int32_t* d_out, int32_t* d_lines, int32_t line_stride,
                   const Sxx& sxx, int32_t penalty, hipStream_t stream);
int long_rows_for_slots(int slots);
int long_strip_rows(int rows);
size_t long_lds_bytes(int rows, int slots, int max_boundaries);
int long_occupancy(int rows, size_t lds, int* blocks_per_cu);
int long_launch(int rows, int grid, size_t lds, const uint16_t* d_blocks, const SeqInfo* d_info, const int4* d_pairs,
                int32_t npairs, int32_t* d_queue, int32_t* d_error, int32_t* d_out, int32_t* d_lines,
                int32_t line_stride, int32_t max_boundaries, int slots, const uint8_t* remap,
                const uint8_t* slot_code, const Sxx& sxx, int32_t penalty, hipStream_t stream);

namespace {

// Host view of one sequence (the same facts as SeqInfo).
struct HostSeq {
    int64_t offset;
    int32_t blocks;
    int32_t first_end;   // rows the reference runs for it as the row sequence
    int32_t last_real;   // columns up to its last non-padding code
};

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

struct DeviceState {
    int compute_units = 0;
    size_t lds_per_block = 0;
};

int device_state(DeviceState& st)
{
    int dev = 0;
    MSA_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    MSA_HIP(hipGetDeviceProperties(&p, dev));
    st.compute_units = p.multiProcessorCount;
    st.lds_per_block = p.sharedMemPerBlock;
    return MSA_OK;
}

struct LongPart {
    int rows = 32, grid = 0, slots = 0;
    size_t lds = 0;
    int32_t npairs = 0;
    int32_t line_stride = 1, max_boundaries = 1;
    uint8_t remap[32] = {};
    uint8_t slot_code[32] = {};
    int4* d_pairs = nullptr;
    int32_t* d_lines = nullptr;
    int32_t* d_queue = nullptr;
    int32_t* d_error = nullptr;
};

constexpr int kDefaultLongThreshold = 8192;

struct Batch {
    int grid = 0;
    int32_t nitems = 0;
    int32_t line_stride = 1;
    int4* d_items = nullptr;
    int2* d_targets = nullptr;
    int32_t* d_lines = nullptr;
    int32_t* d_queue = nullptr;
};

}  // namespace
}  // namespace msa

struct msa_nw_plan_opaque {
    int device = 0;
    int32_t penalty = 0;
    msa::Sxx sxx;
    msa_nw_plan_info info{};
    uint16_t* d_blocks = nullptr;
    msa::SeqInfo* d_info = nullptr;
    // plain part
    int64_t n_plain = 0;
    int32_t* d_plain_pairs = nullptr;
    int32_t* d_plain_index = nullptr;
    int32_t* d_plain_lines = nullptr;
    int32_t plain_line_stride = 1;
    int plain_grid = 0;
    // batched part
    msa::Batch batch;
    // long part
    msa::LongPart lng;
    // graph
    bool use_graph = false;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t* graph_out = nullptr;
    std::vector<void*> allocations;

    ~msa_nw_plan_opaque()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        for (void* p : allocations) (void)hipFree(p);
    }

    template <typename T>
    int alloc(T*& ptr, size_t count)
    {
        void* p = nullptr;
        MSA_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        allocations.push_back(p);
        info.device_bytes += (int64_t)(std::max<size_t>(count, 1) * sizeof(T));
        ptr = static_cast<T*>(p);
        return MSA_OK;
    }

    template <typename T>
    int upload(T*& ptr, const T* host, size_t count, hipStream_t stream)
    {
        MSA_TRY(alloc(ptr, count));
        // The host arrays are temporaries of the planner: copy synchronously.
        if (count) {
            MSA_HIP(hipMemcpyAsync(ptr, host, count * sizeof(T), hipMemcpyHostToDevice, stream));
            MSA_HIP(hipStreamSynchronize(stream));
        }
        return MSA_OK;
    }

    int launch_all(int32_t* d_out, hipStream_t stream)
    {
        if (batch.nitems)
            MSA_TRY(msa::batched_launch(batch.grid, d_blocks, d_info, batch.d_items, batch.nitems, batch.d_targets,
                                        batch.d_queue, d_out, batch.d_lines, batch.line_stride, sxx, penalty, stream));
        if (lng.npairs)
            MSA_TRY(msa::long_launch(lng.rows, lng.grid, lng.lds, d_blocks, d_info, lng.d_pairs, lng.npairs,
                                     lng.d_queue, lng.d_error, d_out, lng.d_lines, lng.line_stride,
                                     lng.max_boundaries, lng.slots, lng.remap, lng.slot_code, sxx, penalty, stream));
        if (n_plain)
            MSA_TRY(msa::plain_launch(d_blocks, d_info, d_plain_pairs, d_plain_index, n_plain, sxx, penalty, d_out,
                                      d_plain_lines, plain_line_stride, plain_grid, stream));
        return MSA_OK;
    }
};

namespace msa {
namespace {

int build_plan(msa_nw_plan_opaque& plan, const uint16_t* blocks, const int64_t* offset, const int32_t* size,
               int32_t nseq, const int32_t* pairs, int64_t npairs, const int32_t* table625, int32_t penalty,
               const msa_nw_options& opt, hipStream_t stream)
{
    if (nseq < 0 || npairs < 0) return fail(MSA_ERR_INVALID_ARGUMENT, "negative sizes");
    if (nseq > 0 && (!offset || !size)) return fail(MSA_ERR_INVALID_ARGUMENT, "null sequence views");
    if (!table625) return fail(MSA_ERR_INVALID_ARGUMENT, "null table");
    if (!pairs && npairs != (int64_t)nseq * (nseq - 1) / 2)
        return fail(MSA_ERR_INVALID_ARGUMENT, "all-pairs mode needs npairs = nseq*(nseq-1)/2 (%lld), got %lld",
                    (long long)nseq * (nseq - 1) / 2, (long long)npairs);
    if (npairs >= (int64_t(1) << 31))
        return fail(MSA_ERR_TOO_LARGE, "%lld pairs; at most 2^31-1 per plan", (long long)npairs);
    const int kernel = opt.kernel;
    if (kernel < MSA_KERNEL_AUTO || kernel > MSA_KERNEL_LONG)
        return fail(MSA_ERR_INVALID_ARGUMENT, "unknown kernel %d", kernel);

    std::vector<HostSeq> seqs;
    int64_t total_blocks = 0;
    MSA_TRY(scan_sequences(blocks, offset, size, nseq, seqs, total_blocks));
    bool symmetric = true;
    MSA_TRY(make_sxx(table625, penalty, plan.sxx, symmetric));
    plan.penalty = penalty;

    DeviceState dev;
    MSA_TRY(device_state(dev));
    MSA_HIP(hipGetDevice(&plan.device));

    // Classify pairs.
    std::vector<int32_t> plain_pairs, plain_index;
    struct Fast { int32_t query, target, index; };
    std::vector<Fast> fast, longs;
    const int64_t long_threshold = kernel == MSA_KERNEL_LONG ? 0
                                 : (opt.long_threshold > 0 ? opt.long_threshold : kDefaultLongThreshold);
    fast.reserve(kernel == MSA_KERNEL_PLAIN ? 0 : (size_t)npairs);
    int64_t cells = 0;
    int64_t c = 0;
    auto visit = [&](int32_t first, int32_t second, int64_t idx) -> int {
        if (first < 0 || first >= nseq || second < 0 || second >= nseq)
            return fail(MSA_ERR_INVALID_ARGUMENT, "pair %lld names sequence outside 0..%d", (long long)idx, nseq - 1);
        const int32_t rowref = seqs[first].blocks > seqs[second].blocks ? first : second;
        const int32_t colref = rowref == first ? second : first;
        const HostSeq& r = seqs[rowref];
        const HostSeq& col = seqs[colref];
        cells += (int64_t)r.first_end * col.last_real;
        const bool clean = col.first_end == col.last_real && r.first_end > 0 && col.last_real > 0;
        if (kernel == MSA_KERNEL_PLAIN || !clean) {
            plain_pairs.push_back(first);
            plain_pairs.push_back(second);
            plain_index.push_back((int32_t)idx);
            return MSA_OK;
        }
        const bool row_query = r.first_end < col.last_real || (r.first_end == col.last_real && rowref < colref);
        const int32_t q = row_query ? rowref : colref;
        const int32_t t = row_query ? colref : rowref;
        const int32_t transpose = (!symmetric && !row_query) ? 1 : 0;
        if (std::min(r.first_end, col.last_real) >= long_threshold)
            longs.push_back({q * 2 + transpose, t, (int32_t)idx});
        else
            fast.push_back({q * 2 + transpose, t, (int32_t)idx});
        return MSA_OK;
    };
    if (pairs) {
        for (c = 0; c < npairs; ++c) MSA_TRY(visit(pairs[2 * c], pairs[2 * c + 1], c));
    } else {
        for (int32_t i = 1; i < nseq; ++i)
            for (int32_t j = 0; j < i; ++j, ++c) MSA_TRY(visit(i, j, c));
    }

    plan.info.pairs = npairs;
    plan.info.cells = cells;
    plan.info.plain_pairs = (int64_t)plain_index.size();
    plan.info.batched_pairs = (int64_t)fast.size();
    plan.info.long_pairs = (int64_t)longs.size();

    // Upload the sequence store; per-sequence facts come from the device scan
    // so that msa_nw_plain and the plan agree on them.
    std::vector<SeqInfo> hinfo(nseq);
    for (int32_t s = 0; s < nseq; ++s)
        hinfo[s] = make_int4((int)seqs[s].offset, seqs[s].blocks, seqs[s].first_end, seqs[s].last_real);
    MSA_TRY(plan.upload(plan.d_blocks, blocks, (size_t)total_blocks, stream));
    MSA_TRY(plan.upload(plan.d_info, hinfo.data(), hinfo.size(), stream));

    // Plain part.
    if (!plain_index.empty()) {
        plan.n_plain = (int64_t)plain_index.size();
        MSA_TRY(plan.upload(plan.d_plain_pairs, plain_pairs.data(), plain_pairs.size(), stream));
        MSA_TRY(plan.upload(plan.d_plain_index, plain_index.data(), plain_index.size(), stream));
        int32_t max_cols = 1;
        for (size_t k = 0; k < plain_index.size(); ++k)
            max_cols = std::max({max_cols, seqs[plain_pairs[2 * k]].last_real, seqs[plain_pairs[2 * k + 1]].last_real});
        plan.plain_grid = plain_grid(plan.n_plain, dev.compute_units);
        plan.plain_line_stride = (max_cols + 63) & ~63;
        MSA_TRY(plan.alloc(plan.d_plain_lines, (size_t)plan.plain_grid * plan.plain_line_stride));
    }

    // Batched part: group by (query, transpose), sort each query's targets by
    // length, cut them into items of up to kItemTargets, and queue the items
    // longest first.
    if (!fast.empty()) {
        std::sort(fast.begin(), fast.end(), [&](const Fast& x, const Fast& y) {
            if (x.query != y.query) return x.query < y.query;
            const int32_t lx = seqs[x.target].first_end, ly = seqs[y.target].first_end;
            if (lx != ly) return lx > ly;
            return x.index < y.index;
        });
        constexpr int kItemTargets = 16;
        const int strip_rows = batched_strip_rows();
        struct Item { int4 v; int64_t cost; };
        std::vector<Item> items;
        std::vector<int2> targets(fast.size());
        int32_t max_t = 1;
        for (size_t k = 0; k < fast.size(); ++k) {
            targets[k] = make_int2(fast[k].target, fast[k].index);
            max_t = std::max(max_t, seqs[fast[k].target].first_end);
        }
        size_t k = 0;
        while (k < fast.size()) {
            size_t e = k;
            while (e < fast.size() && fast[e].query == fast[k].query) ++e;
            const int32_t q = fast[k].query >> 1;
            const int32_t flags = fast[k].query & 1;
            const int64_t lq = seqs[q].first_end;
            const int64_t rows = ((lq + 127) / 128) * 128;
            for (size_t s = k; s < e; s += kItemTargets) {
                const int32_t cnt = (int32_t)std::min<size_t>(kItemTargets, e - s);
                int64_t cost = 0;
                // A round of four targets costs as much as its longest.
                for (size_t u = s; u < s + cnt; u += 4)
                    cost += rows * (((seqs[fast[u].target].first_end + 30) / 16) * 16);
                items.push_back({make_int4(q, (int)s, cnt, flags), cost});
            }
            k = e;
        }
        std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
        std::vector<int4> hitems(items.size());
        for (size_t u = 0; u < items.size(); ++u) hitems[u] = items[u].v;
        Batch& b = plan.batch;
        b.nitems = (int32_t)items.size();
        int per_cu = 0;
        MSA_TRY(batched_occupancy(&per_cu));
        if (per_cu < 1) return fail(MSA_ERR_RUNTIME, "batched kernel does not fit a compute unit");
        b.grid = (int)std::min<int64_t>((int64_t)per_cu * dev.compute_units, (b.nitems + 3) / 4);
        b.line_stride = (max_t + 63) & ~63;
        (void)strip_rows;
        MSA_TRY(plan.upload(b.d_items, hitems.data(), hitems.size(), stream));
        MSA_TRY(plan.upload(b.d_targets, targets.data(), targets.size(), stream));
        MSA_TRY(plan.alloc(b.d_lines, (size_t)b.grid * batched_groups_per_block() * b.line_stride));
        MSA_TRY(plan.alloc(b.d_queue, 1));
        plan.info.bins = 1;
    }
    // Long part: compact alphabet of the codes these pairs use, strip height
    // from the profile size, one pair per workgroup queued longest first.
    if (!longs.empty()) {
        LongPart& L = plan.lng;
        bool present[32] = {};
        int32_t max_t = 1, max_q = 1;
        for (const auto& f : longs) {
            for (const int32_t sidx : {f.query >> 1, f.target}) {
                const HostSeq& h = seqs[sidx];
                for (int32_t pos = 0; pos < h.first_end; ++pos)
                    present[(blocks[h.offset + pos / 3] >> (1 + 5 * (pos % 3))) & 0x1F] = true;
            }
            max_t = std::max(max_t, seqs[f.target].first_end);
            max_q = std::max(max_q, seqs[f.query >> 1].first_end);
        }
        int slots = 0;
        for (int code = 0; code < 32; ++code) L.slot_code[code] = 0xFF;
        for (int code = 0; code < kEndCode; ++code)
            if (present[code]) { L.remap[code] = (uint8_t)slots; L.slot_code[slots] = (uint8_t)code; ++slots; }
        L.slots = slots + 1;                     // the neutral slot is last
        for (int code = 0; code < 32; ++code)
            if (code >= kEndCode || !present[code]) L.remap[code] = (uint8_t)slots;
        L.rows = long_rows_for_slots(L.slots);
        const int strip = long_strip_rows(L.rows);
        L.max_boundaries = std::max(1, (max_q + strip - 1) / strip - 1);
        L.line_stride = (max_t + 63) & ~63;
        L.lds = long_lds_bytes(L.rows, L.slots, L.max_boundaries);
        if (L.lds > dev.lds_per_block)
            return fail(MSA_ERR_TOO_LARGE, "long pairs need %zu bytes of LDS per workgroup", L.lds);
        int per_cu = 0;
        MSA_TRY(long_occupancy(L.rows, L.lds, &per_cu));
        if (per_cu < 1) return fail(MSA_ERR_RUNTIME, "long kernel does not fit a compute unit");
        std::vector<std::pair<int64_t, int4>> order;
        for (const auto& f : longs)
            order.push_back({(int64_t)seqs[f.query >> 1].first_end * seqs[f.target].first_end,
                             make_int4(f.query >> 1, f.target, f.index, f.query & 1)});
        std::stable_sort(order.begin(), order.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        std::vector<int4> hp(order.size());
        for (size_t u = 0; u < order.size(); ++u) hp[u] = order[u].second;
        L.npairs = (int32_t)hp.size();
        L.grid = (int)std::min<int64_t>((int64_t)per_cu * dev.compute_units, L.npairs);
        MSA_TRY(plan.upload(L.d_pairs, hp.data(), hp.size(), stream));
        MSA_TRY(plan.alloc(L.d_lines, (size_t)L.grid * L.max_boundaries * L.line_stride));
        MSA_TRY(plan.alloc(L.d_queue, 1));
        MSA_TRY(plan.alloc(L.d_error, 1));
        MSA_HIP(hipMemsetAsync(L.d_error, 0, sizeof(int32_t), stream));
    }
    plan.info.launches = (int32_t)((plan.batch.nitems ? 2 : 0) + (plan.lng.npairs ? 2 : 0) + (plan.n_plain ? 1 : 0));
    plan.use_graph = opt.use_graph != 0;
    MSA_HIP(hipStreamSynchronize(stream));
    return MSA_OK;
}

}  // namespace
}  // namespace msa

using msa::fail;

extern "C" {

int msa_nw_plan_create(msa_nw_plan* plan, const uint16_t* blocks, const int64_t* offset, const int32_t* size,
                       int32_t nseq, const int32_t* pairs, int64_t npairs, const int32_t* table625, int32_t penalty,
                       const msa_nw_options* options, msa_stream stream)
{
    if (!plan) return fail(MSA_ERR_INVALID_ARGUMENT, "plan is null");
    *plan = nullptr;
    if (nseq > 0 && !blocks) return fail(MSA_ERR_INVALID_ARGUMENT, "blocks is null");
    msa_nw_options opt{};
    if (options) opt = *options;
    auto* p = new (std::nothrow) msa_nw_plan_opaque();
    if (!p) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed");
    int rc = MSA_OK;
    try {
        rc = msa::build_plan(*p, blocks, offset, size, nseq, pairs, npairs, table625, penalty, opt,
                             (hipStream_t)stream);
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

int msa_nw_plan_execute(msa_nw_plan plan, int32_t* d_out, msa_stream stream)
{
    if (!plan || (!d_out && plan->info.pairs)) return fail(MSA_ERR_INVALID_ARGUMENT, "null plan or output");
    if (!plan->info.pairs) return MSA_OK;
    const hipStream_t s = (hipStream_t)stream;
    if (!plan->use_graph) return plan->launch_all(d_out, s);
    if (plan->graph_out != d_out || !plan->exec) {
        if (plan->exec) { (void)hipGraphExecDestroy(plan->exec); plan->exec = nullptr; }
        if (plan->graph) { (void)hipGraphDestroy(plan->graph); plan->graph = nullptr; }
        if (!plan->capture_stream) MSA_HIP(hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking));
        MSA_HIP(hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal));
        const int rc = plan->launch_all(d_out, plan->capture_stream);
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

int msa_nw_plan_info_get(msa_nw_plan plan, msa_nw_plan_info* info)
{
    if (!plan || !info) return fail(MSA_ERR_INVALID_ARGUMENT, "null plan or info");
    *info = plan->info;
    return MSA_OK;
}

int msa_nw_plan_status(msa_nw_plan plan, msa_stream stream)
{
    if (!plan) return fail(MSA_ERR_INVALID_ARGUMENT, "plan is null");
    MSA_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (plan->lng.d_error) {
        int32_t err = 0;
        MSA_HIP(hipMemcpy(&err, plan->lng.d_error, sizeof(err), hipMemcpyDeviceToHost));
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
            rc = msa::hip_status(hipMemcpy(out, d_out, (size_t)npairs * sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy");
        if (d_out) (void)hipFree(d_out);
    }
    msa_nw_plan_destroy(plan);
    return rc;
}

}  // extern "C"

// ---- msa_nw_plain on caller-owned device memory -----------------------------

namespace {

struct PlainWorkspace {
    std::mutex lock;
    msa::SeqInfo* info = nullptr;
    size_t info_cap = 0;
    int32_t* lines = nullptr;
    size_t lines_cap = 0;
};

PlainWorkspace g_workspace[64];

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
    msa::DeviceState st;
    MSA_TRY(msa::device_state(st));
    PlainWorkspace& ws = g_workspace[dev];
    std::lock_guard<std::mutex> guard(ws.lock);
    const int grid = msa::plain_grid(npairs, st.compute_units);
    const int32_t stride = (int32_t)(((int64_t)3 * max_blocks + 63) & ~int64_t(63));
    MSA_TRY(grow(ws.info, ws.info_cap, (size_t)nseq));
    MSA_TRY(grow(ws.lines, ws.lines_cap, (size_t)grid * std::max(stride, 1)));
    const hipStream_t s = (hipStream_t)stream;
    MSA_TRY(msa::seq_info_launch(d_blocks, d_offset, d_size, nseq, ws.info, s));
    return msa::plain_launch(d_blocks, ws.info, d_pairs, nullptr, npairs, sxx, penalty, d_out, ws.lines,
                             std::max(stride, 1), grid, s);
}
