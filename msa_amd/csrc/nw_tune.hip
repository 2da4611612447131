// The autotuner and the tuning table it writes.
//
// The reference launches one fixed shape, 160 threads per pair
// (src/pairwise/needleman/impl/hybrid.cu:409-423).  Here the batched kernel
// has several tile shapes (lanes per pair x rows per lane, kTileConfig) and a
// second inter-tile scheme, the long kernel's pipelined strips.  Which one is
// fastest depends on the query length, so the tuner times them on samples of
// each query-length bin of a real database and keeps the winners:
//
//   { "format": "msa-nw-tuning/1", "device": "...", "compute_units": 256,
//     "long_threshold": 8192,
//     "bins": [ {"min_query": 0, "lanes": 16, "rows": 32}, ... ],
//     "measurements": [ {...}, ... ] }
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
constexpr int64_t kMinSamplePairs = 16384;     // enough items to fill every wave of the GPU

struct Measurement {
    int bin;
    int64_t pairs, cells;
    bool long_scheme;
    int config;
    double ms, gcups;
};

// Times one execution of a plan over `pairs` (average of enough repeats to
// cover ~30 ms after one warm-up execution).
int time_plan(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
              const std::vector<int32_t>& pairs, const int32_t* table625, int32_t penalty, const msa_nw_options& opt,
              int32_t* d_out, double& ms)
{
    msa_nw_plan plan = nullptr;
    const int64_t npairs = (int64_t)pairs.size() / 2;
    MSA_TRY(msa_nw_plan_create(&plan, blocks, offset, size, nseq, pairs.data(), npairs, table625, penalty, &opt,
                               nullptr));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = msa::hip_status(hipEventCreate(&e0), "hipEventCreate");
    if (rc == MSA_OK) rc = msa::hip_status(hipEventCreate(&e1), "hipEventCreate");
    float t1 = 0;
    if (rc == MSA_OK) rc = msa_nw_plan_execute(plan, d_out, nullptr);
    if (rc == MSA_OK) rc = msa::hip_status(hipEventRecord(e0, nullptr), "hipEventRecord");
    if (rc == MSA_OK) rc = msa_nw_plan_execute(plan, d_out, nullptr);
    if (rc == MSA_OK) rc = msa::hip_status(hipEventRecord(e1, nullptr), "hipEventRecord");
    if (rc == MSA_OK) rc = msa_nw_plan_status(plan, nullptr);
    if (rc == MSA_OK) rc = msa::hip_status(hipEventElapsedTime(&t1, e0, e1), "hipEventElapsedTime");
    int reps = 1;
    if (rc == MSA_OK && t1 < 30.0f) {
        reps = std::min(20, std::max(2, (int)std::ceil(30.0 / std::max(t1, 0.01f))));
        rc = msa::hip_status(hipEventRecord(e0, nullptr), "hipEventRecord");
        for (int r = 0; r < reps && rc == MSA_OK; ++r) rc = msa_nw_plan_execute(plan, d_out, nullptr);
        if (rc == MSA_OK) rc = msa::hip_status(hipEventRecord(e1, nullptr), "hipEventRecord");
        if (rc == MSA_OK) rc = msa_nw_plan_status(plan, nullptr);
        if (rc == MSA_OK) rc = msa::hip_status(hipEventElapsedTime(&t1, e0, e1), "hipEventElapsedTime");
    }
    ms = (double)t1 / reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
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

int tune(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq, const int32_t* table625,
         int32_t penalty, int64_t budget_cells, std::string& json)
{
    if (nseq < 2) return fail(MSA_ERR_INVALID_ARGUMENT, "tuning needs at least two sequences");
    if (nseq > 0 && (!blocks || !offset || !size)) return fail(MSA_ERR_INVALID_ARGUMENT, "null sequence database");
    if (!table625) return fail(MSA_ERR_INVALID_ARGUMENT, "null table");
    if (budget_cells <= 0) budget_cells = (int64_t)4e10;
    std::vector<msa::HostSeq> seqs;
    int64_t total_blocks = 0;
    MSA_TRY(msa::scan_sequences(blocks, offset, size, nseq, seqs, total_blocks));
    msa::Sxx sxx;
    bool symmetric = true;
    MSA_TRY(msa::make_sxx(table625, penalty, sxx, symmetric));

    // Clean sequences in length order; a sequence is the query of the pairs
    // it makes with every later one (the shorter side of each pair).
    std::vector<int32_t> order;
    for (int32_t s = 0; s < nseq; ++s)
        if (seqs[s].first_end == seqs[s].last_real && seqs[s].first_end > 0) order.push_back(s);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return seqs[a].first_end < seqs[b].first_end; });
    const size_t m = order.size();
    std::vector<int64_t> suffix_cells(m + 1, 0);   // sum of lengths after position i
    for (size_t i = m; i-- > 0;) suffix_cells[i] = suffix_cells[i + 1] + seqs[order[i]].first_end;

    msa::DeviceShape dev;
    MSA_TRY(msa::device_shape(dev));
    int device = 0;
    MSA_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    MSA_HIP(hipGetDeviceProperties(&prop, device));

    std::vector<Measurement> meas;
    int best[kBins];
    double best_gcups[kBins], long_gcups[kBins];
    for (int b = 0; b < kBins; ++b) {
        best[b] = -1;
        best_gcups[b] = long_gcups[b] = 0;
    }
    int32_t* d_out = nullptr;
    size_t out_cap = 0;
    int rc = MSA_OK;
    for (int b = 0; b < kBins && rc == MSA_OK; ++b) {
        const int32_t lo = kBinEdges[b];
        const int32_t hi = b + 1 < kBins ? kBinEdges[b + 1] : INT32_MAX;
        std::vector<size_t> queries;
        int64_t bin_cells = 0, bin_pairs = 0;
        for (size_t i = 0; i < m; ++i) {
            const int32_t len = seqs[order[i]].first_end;
            if (len < lo || len >= hi || i + 1 == m) continue;
            queries.push_back(i);
            bin_cells += (int64_t)len * suffix_cells[i + 1];
            bin_pairs += (int64_t)(m - 1 - i);
        }
        if (queries.empty()) continue;
        // Every stride-th query (in length order) with all its targets.
        const int64_t stride = std::max<int64_t>(1, std::min(bin_cells / budget_cells, bin_pairs / kMinSamplePairs));
        std::vector<int32_t> pairs;
        int64_t cells = 0;
        for (size_t u = 0; u < queries.size(); u += (size_t)stride) {
            const size_t i = queries[u];
            const int32_t q = order[i];
            for (size_t j = i + 1; j < m; ++j) {
                const int32_t t = order[j];
                pairs.push_back(std::max(q, t));
                pairs.push_back(std::min(q, t));
            }
            cells += (int64_t)seqs[q].first_end * suffix_cells[i + 1];
        }
        const int64_t npairs = (int64_t)pairs.size() / 2;
        if ((size_t)npairs > out_cap) {
            if (d_out) (void)hipFree(d_out);
            d_out = nullptr;
            out_cap = (size_t)npairs;
            rc = msa::hip_status(hipMalloc(&d_out, out_cap * sizeof(int32_t)), "hipMalloc");
            if (rc != MSA_OK) break;
        }
        const int candidates = msa::kTileConfigCount + (lo >= 1024 ? 1 : 0);
        for (int c = 0; c < candidates && rc == MSA_OK; ++c) {
            msa_nw_options opt {};
            const bool long_scheme = c == msa::kTileConfigCount;
            if (long_scheme) {
                opt.kernel = MSA_KERNEL_LONG;
            } else {
                opt.kernel = MSA_KERNEL_BATCHED;
                opt.lanes_per_pair = msa::kTileConfig[c].lanes;
                opt.rows_per_lane = msa::kTileConfig[c].rows;
            }
            double ms = 0;
            rc = time_plan(blocks, offset, size, nseq, pairs, table625, penalty, opt, d_out, ms);
            if (rc != MSA_OK) break;
            const double gcups = ms > 0 ? (double)cells / (ms * 1e6) : 0;
            meas.push_back({b, npairs, cells, long_scheme, long_scheme ? -1 : c, ms, gcups});
            if (long_scheme) {
                long_gcups[b] = gcups;
            } else if (gcups > best_gcups[b]) {
                best_gcups[b] = gcups;
                best[b] = c;
            }
        }
    }
    if (d_out) (void)hipFree(d_out);
    MSA_TRY(rc);

    // Bins without samples take the shape of the nearest measured bin below
    // (or above, for the first ones).
    int first_measured = -1;
    for (int b = 0; b < kBins; ++b)
        if (best[b] >= 0 && first_measured < 0) first_measured = b;
    if (first_measured < 0) return fail(MSA_ERR_INVALID_ARGUMENT, "no clean pair to tune on");
    int shape[kBins];
    for (int b = 0, last = best[first_measured]; b < kBins; ++b) {
        if (best[b] >= 0) last = best[b];
        shape[b] = last;
    }
    // The long scheme starts at the lowest bin edge from which every measured
    // bin prefers it; with no such bin it starts past the longest measured bin.
    int32_t long_threshold = msa::kDefaultLongThreshold;
    int last_measured = -1;
    for (int b = 0; b < kBins; ++b)
        if (best[b] >= 0) last_measured = b;
    int from = -1;
    for (int b = last_measured; b >= 0; --b) {
        if (best[b] < 0) continue;
        if (long_gcups[b] > best_gcups[b]) from = b;
        else break;
    }
    if (from >= 0) long_threshold = std::max<int32_t>(1, kBinEdges[from]);
    else if (last_measured + 1 < kBins) long_threshold = std::max(long_threshold, kBinEdges[last_measured + 1]);
    else long_threshold = INT32_MAX;

    char name[260];
    std::snprintf(name, sizeof(name), "%s %s", prop.name, prop.gcnArchName);
    for (char* c = name; *c; ++c)
        if (*c == '"' || *c == '\\' || (unsigned char)*c < 0x20) *c = ' ';
    json.clear();
    appendf(json, "{\n  \"format\": \"msa-nw-tuning/1\",\n  \"device\": \"%s\",\n  \"compute_units\": %d,\n", name,
            dev.compute_units);
    appendf(json, "  \"budget_cells\": %lld,\n  \"long_threshold\": %d,\n  \"bins\": [", (long long)budget_cells,
            long_threshold);
    bool first = true;
    for (int b = 0; b < kBins; ++b) {
        if (b > 0 && shape[b] == shape[b - 1]) continue;
        appendf(json, "%s\n    {\"min_query\": %d, \"lanes\": %d, \"rows\": %d}", first ? "" : ",", kBinEdges[b],
                msa::kTileConfig[shape[b]].lanes, msa::kTileConfig[shape[b]].rows);
        first = false;
    }
    appendf(json, "\n  ],\n  \"measurements\": [");
    for (size_t i = 0; i < meas.size(); ++i) {
        const Measurement& x = meas[i];
        const int32_t hi = x.bin + 1 < kBins ? kBinEdges[x.bin + 1] : -1;
        appendf(json, "%s\n    {\"min_query\": %d, \"max_query\": %d, \"pairs\": %lld, \"cells\": %lld, ",
                i ? "," : "", kBinEdges[x.bin], hi < 0 ? -1 : hi - 1, (long long)x.pairs, (long long)x.cells);
        if (x.long_scheme)
            appendf(json, "\"scheme\": \"long\", \"lanes\": 64, \"rows\": 0, ");
        else
            appendf(json, "\"scheme\": \"batched\", \"lanes\": %d, \"rows\": %d, ", msa::kTileConfig[x.config].lanes,
                    msa::kTileConfig[x.config].rows);
        appendf(json, "\"ms\": %.4f, \"gcups\": %.1f}", x.ms, x.gcups);
    }
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
