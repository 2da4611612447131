// The autotuner and the tuning table it writes.
//
// The reference launches one fixed shape, 160 threads per pair
// (src/pairwise/needleman/impl/hybrid.cu:409-423).  Here the batched kernel
// has several tile shapes (lanes per pair x rows per lane, kTileConfig) and a
// second inter-tile scheme, the long kernel's pipelined strips.  Which one is
// fastest depends on the query length, so the tuner picks a shape per
// query-length bin and the length from which pairs take the long kernel, by
// timing whole executions of a real database (or a sample of its queries):
//
//   { "format": "msa-nw-tuning/1", "device": "...", "compute_units": 256,
//     "long_threshold": 8192,
//     "bins": [ {"min_query": 0, "lanes": 16, "rows": 32}, ... ],
//     "bin_cells": [...], "threshold_trials": [...], "measurements": [...] }
//
// Bins run from their min_query to the next bin's; the first starts at 0.
// Pairs whose shorter side reaches long_threshold take the long kernel.
#include "nw_planner.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

namespace msa {

int tile_config_index(int lanes, int rows)
{
    for (int c = 0; c < kTileConfigCount; ++c)
        if (kTileConfig[c].lanes == lanes && kTileConfig[c].rows == rows) return c;
    return -1;
}

const Tuning& builtin_tuning()
{
    static const Tuning t = [] {
        Tuning b;
        b.bins = {{0, 0}};
        b.long_threshold = kDefaultLongThreshold;
        return b;
    }();
    return t;
}

int tuning_config_for(const Tuning& t, int32_t query_len)
{
    int cfg = t.bins.empty() ? 0 : t.bins[0].config;
    for (const TuneBin& b : t.bins)
        if (query_len >= b.min_query) cfg = b.config;
    return cfg;
}

namespace {

// ---- a small JSON reader (objects, arrays, numbers, strings, literals) ----

struct JValue {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;

    const JValue* get(const char* key) const
    {
        if (kind != Obj) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class JReader {
public:
    JReader(const char* text, size_t n) : p_(text), end_(text + n) {}

    bool parse(JValue& v)
    {
        if (!value(v, 0)) return false;
        ws();
        return p_ == end_;
    }

private:
    const char* p_;
    const char* end_;

    void ws()
    {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }

    bool literal(const char* word)
    {
        const size_t n = std::strlen(word);
        if ((size_t)(end_ - p_) < n || std::strncmp(p_, word, n) != 0) return false;
        p_ += n;
        return true;
    }

    bool string(std::string& out)
    {
        if (p_ >= end_ || *p_ != '"') return false;
        ++p_;
        while (p_ < end_ && *p_ != '"') {
            if (*p_ == '\\') {
                if (++p_ >= end_) return false;
                const char c = *p_;
                out.push_back(c == 'n' ? '\n' : c == 't' ? '\t' : c);   // enough for names and notes
            } else {
                out.push_back(*p_);
            }
            ++p_;
        }
        if (p_ >= end_) return false;
        ++p_;
        return true;
    }

    bool value(JValue& v, int depth)
    {
        if (depth > 32) return false;
        ws();
        if (p_ >= end_) return false;
        const char c = *p_;
        if (c == '{') {
            v.kind = JValue::Obj;
            ++p_;
            ws();
            if (p_ < end_ && *p_ == '}') { ++p_; return true; }
            while (true) {
                ws();
                std::string key;
                if (!string(key)) return false;
                ws();
                if (p_ >= end_ || *p_ != ':') return false;
                ++p_;
                JValue item;
                if (!value(item, depth + 1)) return false;
                v.obj.emplace_back(std::move(key), std::move(item));
                ws();
                if (p_ < end_ && *p_ == ',') { ++p_; continue; }
                if (p_ < end_ && *p_ == '}') { ++p_; return true; }
                return false;
            }
        }
        if (c == '[') {
            v.kind = JValue::Arr;
            ++p_;
            ws();
            if (p_ < end_ && *p_ == ']') { ++p_; return true; }
            while (true) {
                JValue item;
                if (!value(item, depth + 1)) return false;
                v.arr.push_back(std::move(item));
                ws();
                if (p_ < end_ && *p_ == ',') { ++p_; continue; }
                if (p_ < end_ && *p_ == ']') { ++p_; return true; }
                return false;
            }
        }
        if (c == '"') {
            v.kind = JValue::Str;
            return string(v.str);
        }
        if (literal("true")) { v.kind = JValue::Bool; v.num = 1; return true; }
        if (literal("false")) { v.kind = JValue::Bool; v.num = 0; return true; }
        if (literal("null")) { v.kind = JValue::Null; return true; }
        std::string num;
        while (p_ < end_ && (std::strchr("+-.eE", *p_) || (*p_ >= '0' && *p_ <= '9'))) num.push_back(*p_++);
        if (num.empty()) return false;
        char* stop = nullptr;
        errno = 0;
        v.num = std::strtod(num.c_str(), &stop);
        if (errno || !stop || *stop) return false;
        v.kind = JValue::Num;
        return true;
    }
};

bool whole_number(const JValue* v, int64_t lo, int64_t hi, int64_t& out)
{
    if (!v || v->kind != JValue::Num || v->num != std::floor(v->num) || v->num < lo || v->num > hi) return false;
    out = (int64_t)v->num;
    return true;
}

}  // namespace

int load_tuning(const char* path, Tuning& out)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(MSA_ERR_INVALID_ARGUMENT, "cannot open tuning table '%s'", path);
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) {
        text.append(buf, n);
        if (text.size() > (size_t(1) << 24)) break;
    }
    std::fclose(f);
    JValue root;
    if (!JReader(text.data(), text.size()).parse(root) || root.kind != JValue::Obj)
        return fail(MSA_ERR_INVALID_ARGUMENT, "tuning table '%s' is not a JSON object", path);
    const JValue* fmt = root.get("format");
    if (!fmt || fmt->kind != JValue::Str || fmt->str != "msa-nw-tuning/1")
        return fail(MSA_ERR_INVALID_ARGUMENT, "tuning table '%s': format is not msa-nw-tuning/1", path);
    Tuning t;
    int64_t v = 0;
    if (!whole_number(root.get("long_threshold"), 1, INT32_MAX, v))
        return fail(MSA_ERR_INVALID_ARGUMENT, "tuning table '%s': long_threshold must be a positive integer", path);
    t.long_threshold = (int32_t)v;
    const JValue* bins = root.get("bins");
    if (!bins || bins->kind != JValue::Arr || bins->arr.empty())
        return fail(MSA_ERR_INVALID_ARGUMENT, "tuning table '%s': bins must be a non-empty array", path);
    for (size_t b = 0; b < bins->arr.size(); ++b) {
        const JValue& e = bins->arr[b];
        int64_t mq = 0, lanes = 0, rows = 0;
        if (!whole_number(e.get("min_query"), 0, INT32_MAX, mq) || !whole_number(e.get("lanes"), 1, 64, lanes) ||
            !whole_number(e.get("rows"), 1, 64, rows))
            return fail(MSA_ERR_INVALID_ARGUMENT, "tuning table '%s': bin %zu needs integer min_query, lanes, rows",
                        path, b);
        const int cfg = tile_config_index((int)lanes, (int)rows);
        if (cfg < 0)
            return fail(MSA_ERR_INVALID_ARGUMENT, "tuning table '%s': bin %zu asks for %lld lanes x %lld rows, "
                        "which the batched kernel does not have", path, b, (long long)lanes, (long long)rows);
        if ((b == 0 && mq != 0) || (b > 0 && mq <= t.bins.back().min_query))
            return fail(MSA_ERR_INVALID_ARGUMENT, "tuning table '%s': bins must start at 0 and increase", path);
        t.bins.push_back({(int32_t)mq, cfg});
    }
    out = std::move(t);
    return MSA_OK;
}

}  // namespace msa

namespace {

using msa::fail;

constexpr int32_t kBinEdges[] = {0, 128, 256, 512, 1024, 2048, 4096, 8192, 16384};
constexpr int kBins = sizeof(kBinEdges) / sizeof(kBinEdges[0]);
constexpr int32_t kLongCandidates[] = {1024, 2048, 4096, 8192, 16384, INT32_MAX};
constexpr int kReps = 3;
constexpr double kMinGain = 0.005;     // a change must save more than this fraction of the time

// One timed configuration: tile shape per bin and long threshold.
struct Trial {
    std::string step;
    int shape[kBins];
    int32_t long_threshold;
    double ms;
};

msa::Tuning make_tuning(const int (&shape)[kBins], int32_t long_threshold)
{
    msa::Tuning t;
    for (int b = 0; b < kBins; ++b)
        if (b == 0 || shape[b] != shape[b - 1]) t.bins.push_back({kBinEdges[b], shape[b]});
    t.long_threshold = long_threshold;
    return t;
}

// Median time of kReps executions (after one warm-up) of a plan over the
// sample under one tuning table.
int time_plan(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
              const std::vector<int32_t>& pairs, bool all_pairs, int64_t npairs, const int32_t* table625,
              int32_t penalty, const msa::Tuning& tuning, int32_t* d_out, double& ms)
{
    msa_nw_plan plan = nullptr;
    msa_nw_options opt {};
    opt.kernel = MSA_KERNEL_AUTO;
    MSA_TRY(msa::create_plan(&plan, blocks, offset, size, nseq, all_pairs ? nullptr : pairs.data(), npairs, table625,
                             penalty, opt, &tuning, nullptr));
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = msa::hip_status(hipEventCreate(&ev[0]), "hipEventCreate");
    if (rc == MSA_OK) rc = msa::hip_status(hipEventCreate(&ev[1]), "hipEventCreate");
    if (rc == MSA_OK) rc = msa_nw_plan_execute(plan, d_out, nullptr);
    if (rc == MSA_OK) rc = msa_nw_plan_status(plan, nullptr);
    float t[kReps] = {};
    for (int r = 0; r < kReps && rc == MSA_OK; ++r) {
        rc = msa::hip_status(hipEventRecord(ev[0], nullptr), "hipEventRecord");
        if (rc == MSA_OK) rc = msa_nw_plan_execute(plan, d_out, nullptr);
        if (rc == MSA_OK) rc = msa::hip_status(hipEventRecord(ev[1], nullptr), "hipEventRecord");
        if (rc == MSA_OK) rc = msa_nw_plan_status(plan, nullptr);
        if (rc == MSA_OK) rc = msa::hip_status(hipEventElapsedTime(&t[r], ev[0], ev[1]), "hipEventElapsedTime");
    }
    std::sort(t, t + kReps);
    ms = t[kReps / 2];
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    const int rc2 = msa_nw_plan_destroy(plan);
    return rc != MSA_OK ? rc : rc2;
}

void appendf(std::string& s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void appendf(std::string& s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    const int n = std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (n > 0) s.append(buf, std::min<size_t>((size_t)n, sizeof(buf) - 1));
}

// Coordinate descent over executions of the whole sample: each populated
// query-length bin in turn tries every tile shape with the others fixed,
// then every long threshold is tried with the chosen shapes.  A change is
// kept only when it saves more than kMinGain of the time, so that timing
// noise does not pick shapes.  Timing the
// whole mix (not each bin alone) keeps the GPU as full as in real use, so a
// small bin is judged by what it adds to a full launch, not by its own tail.
int tune(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq, const int32_t* table625,
         int32_t penalty, int64_t budget_cells, std::string& json)
{
    if (nseq < 2) return fail(MSA_ERR_INVALID_ARGUMENT, "tuning needs at least two sequences");
    if (!blocks || !offset || !size) return fail(MSA_ERR_INVALID_ARGUMENT, "null sequence database");
    if (!table625) return fail(MSA_ERR_INVALID_ARGUMENT, "null table");
    if (budget_cells <= 0) budget_cells = (int64_t)2e12;
    std::vector<msa::HostSeq> seqs;
    int64_t total_blocks = 0;
    MSA_TRY(msa::scan_sequences(blocks, offset, size, nseq, seqs, total_blocks));
    msa::Sxx sxx;
    bool symmetric = true;
    MSA_TRY(msa::make_sxx(table625, penalty, sxx, symmetric));

    // Query lengths: the shorter side of each pair (first padding code).  The
    // planner may run a batched pair with its longer side as the query; the
    // trials time whole executions, so that changes only which bin a pair's
    // time is charged to.
    auto len = [&](int32_t s) { return std::min(seqs[s].first_end, seqs[s].last_real); };
    std::vector<int32_t> order((size_t)nseq);
    for (int32_t s = 0; s < nseq; ++s) order[(size_t)s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return len(a) < len(b); });
    std::vector<int64_t> suffix((size_t)nseq + 1, 0);
    for (size_t i = (size_t)nseq; i-- > 0;) suffix[i] = suffix[i + 1] + len(order[i]);
    int64_t total_cells = 0;
    for (size_t i = 0; i < (size_t)nseq; ++i) total_cells += (int64_t)len(order[i]) * suffix[i + 1];

    // The sample: every pair, or every stride-th query (in length order)
    // with all its longer partners.
    const int64_t stride = std::max<int64_t>(1, (total_cells + budget_cells - 1) / budget_cells);
    const bool all_pairs = stride == 1;
    std::vector<int32_t> pairs;
    int64_t npairs = 0, cells = 0;
    int64_t bin_cells[kBins] = {};
    auto bin_of = [](int32_t q) {
        int b = 0;
        while (b + 1 < kBins && q >= kBinEdges[b + 1]) ++b;
        return b;
    };
    for (size_t i = 0; i + 1 < (size_t)nseq; ++i) {
        if (!all_pairs && (int64_t)i % stride != 0) continue;
        const int64_t c = (int64_t)len(order[i]) * suffix[i + 1];
        cells += c;
        bin_cells[bin_of(len(order[i]))] += c;
        npairs += (int64_t)nseq - 1 - (int64_t)i;
        if (!all_pairs)
            for (size_t j = i + 1; j < (size_t)nseq; ++j) {
                pairs.push_back(std::max(order[i], order[j]));
                pairs.push_back(std::min(order[i], order[j]));
            }
    }
    if (cells == 0) return fail(MSA_ERR_INVALID_ARGUMENT, "the database has no cell to tune on");
    if (npairs >= (int64_t(1) << 31)) return fail(MSA_ERR_TOO_LARGE, "tuning sample has too many pairs");

    msa::DeviceShape dev;
    MSA_TRY(msa::device_shape(dev));
    int device = 0;
    MSA_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    MSA_HIP(hipGetDeviceProperties(&prop, device));
    int32_t* d_out = nullptr;
    MSA_HIP(hipMalloc(&d_out, (size_t)npairs * sizeof(int32_t)));

    std::vector<Trial> trials;
    Trial cur;
    for (int b = 0; b < kBins; ++b) cur.shape[b] = 0;
    cur.long_threshold = INT32_MAX;
    auto run = [&](Trial& t) {
        return time_plan(blocks, offset, size, nseq, pairs, all_pairs, npairs, table625, penalty,
                         make_tuning(t.shape, t.long_threshold), d_out, t.ms);
    };
    cur.step = "start";
    int rc = run(cur);
    if (rc == MSA_OK) trials.push_back(cur);
    for (int b = 0; b < kBins && rc == MSA_OK; ++b) {
        if (bin_cells[b] == 0) continue;
        for (int c = 0; c < msa::kTileConfigCount && rc == MSA_OK; ++c) {
            if (c == cur.shape[b]) continue;
            Trial t = cur;
            t.shape[b] = c;
            char what[96];
            std::snprintf(what, sizeof(what), "bin %d: %d lanes x %d rows", kBinEdges[b], msa::kTileConfig[c].lanes,
                          msa::kTileConfig[c].rows);
            t.step = what;
            rc = run(t);
            if (rc != MSA_OK) break;
            trials.push_back(t);
            if (t.ms < cur.ms * (1.0 - kMinGain)) cur = t;
        }
    }
    // Long thresholds: only those that move some pair.
    const int32_t max_query = len(order[(size_t)nseq - 2]);   // the longest shorter side of a pair
    std::vector<std::pair<int32_t, double>> curve;
    curve.push_back({INT32_MAX, cur.ms});
    for (int32_t th : kLongCandidates) {
        if (rc != MSA_OK || th == INT32_MAX || th > max_query) continue;
        Trial t = cur;
        t.long_threshold = th;
        t.step = "long from " + std::to_string(th);
        rc = run(t);
        if (rc != MSA_OK) break;
        trials.push_back(t);
        curve.push_back({th, t.ms});
    }
    (void)hipFree(d_out);
    MSA_TRY(rc);
    Trial best = cur;
    for (const auto& pt : curve)
        if (pt.second < best.ms * (1.0 - kMinGain)) {
            best.long_threshold = pt.first;
            best.ms = pt.second;
        }

    char name[260];
    std::snprintf(name, sizeof(name), "%s %s", prop.name, prop.gcnArchName);
    for (char* c = name; *c; ++c)
        if (*c == '"' || *c == '\\' || (unsigned char)*c < 0x20) *c = ' ';
    const msa::Tuning table = make_tuning(best.shape, best.long_threshold);
    json.clear();
    appendf(json, "{\n  \"format\": \"msa-nw-tuning/1\",\n  \"device\": \"%s\",\n  \"compute_units\": %d,\n", name,
            dev.compute_units);
    appendf(json, "  \"method\": \"coordinate descent over executions of the sample\",\n");
    appendf(json, "  \"sample\": {\"pairs\": %lld, \"cells\": %lld, \"all_pairs\": %s},\n", (long long)npairs,
            (long long)cells, all_pairs ? "true" : "false");
    appendf(json, "  \"long_threshold\": %d,\n  \"bins\": [", table.long_threshold);
    for (size_t b = 0; b < table.bins.size(); ++b)
        appendf(json, "%s\n    {\"min_query\": %d, \"lanes\": %d, \"rows\": %d}", b ? "," : "", table.bins[b].min_query,
                msa::kTileConfig[table.bins[b].config].lanes, msa::kTileConfig[table.bins[b].config].rows);
    appendf(json, "\n  ],\n  \"bin_cells\": [");
    for (int b = 0; b < kBins; ++b)
        appendf(json, "%s\n    {\"min_query\": %d, \"cells\": %lld, \"lanes\": %d, \"rows\": %d}", b ? "," : "",
                kBinEdges[b], (long long)bin_cells[b], msa::kTileConfig[best.shape[b]].lanes,
                msa::kTileConfig[best.shape[b]].rows);
    appendf(json, "\n  ],\n  \"threshold_trials\": [");
    for (size_t i = 0; i < curve.size(); ++i)
        appendf(json, "%s\n    {\"long_threshold\": %d, \"ms\": %.4f}", i ? "," : "", curve[i].first, curve[i].second);
    appendf(json, "\n  ],\n  \"measurements\": [");
    for (size_t i = 0; i < trials.size(); ++i)
        appendf(json, "%s\n    {\"step\": \"%s\", \"long_threshold\": %d, \"ms\": %.4f, \"gcups\": %.1f}",
                i ? "," : "", trials[i].step.c_str(), trials[i].long_threshold, trials[i].ms,
                trials[i].ms > 0 ? (double)cells / (trials[i].ms * 1e6) : 0.0);
    appendf(json, "\n  ]\n}\n");
    return MSA_OK;
}

}  // namespace

extern "C" int msa_nw_tune(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                           const int32_t* table625, int32_t penalty, int64_t budget_cells, const char* path,
                           char* report, size_t report_bytes)
{
    std::string json;
    int rc = MSA_OK;
    try {
        rc = tune(blocks, offset, size, nseq, table625, penalty, budget_cells, json);
    } catch (const std::bad_alloc&) {
        rc = fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    if (rc != MSA_OK) return rc;
    if (path) {
        FILE* f = std::fopen(path, "wb");
        if (!f) return fail(MSA_ERR_INVALID_ARGUMENT, "cannot write tuning table '%s'", path);
        const bool ok = std::fwrite(json.data(), 1, json.size(), f) == json.size();
        if (std::fclose(f) != 0 || !ok) return fail(MSA_ERR_RUNTIME, "writing tuning table '%s' failed", path);
    }
    if (report && report_bytes) {
        const size_t n = std::min(json.size(), report_bytes - 1);
        std::memcpy(report, json.data(), n);
        report[n] = '\0';
    }
    return MSA_OK;
}
