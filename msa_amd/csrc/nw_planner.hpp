// Host planning shared by the plan API (nw_plan.hip) and the serving loop
// (nw_serve.hip).  A HostPlan is everything the host decides for one pair set:
// the sequence store, the routing of every pair to the plain, batched or long
// kernel, the work lists and the launch shapes.  Its arrays are laid out as
// one contiguous data block (one host->device copy); the per-launch scratch
// (strip-boundary lines, queue counters, the error flag) is sized separately
// so that a server can keep it per stream and reuse it across jobs.
#pragma once

#include "msa_internal.hpp"

#include <vector>

namespace msa {

// Kernel entry points (nw_plain.hip, nw_batched.hip, nw_long.hip).
int plain_launch(const uint16_t* d_blocks, const SeqInfo* d_info, const int32_t* d_pairs,
                 const int32_t* d_out_index, int64_t npairs, const Sxx& sxx, int32_t penalty,
                 int32_t* d_out, int32_t* d_lines, int32_t line_stride, int grid, hipStream_t stream);
int plain_grid(int64_t npairs, int compute_units);
int seq_info_launch(const uint16_t* d_blocks, const int64_t* d_offset, const int32_t* d_size, int32_t nseq,
                    SeqInfo* d_info, hipStream_t stream);
// Tile shapes of the batched kernel, chosen per item: `lanes` lanes score one
// pair, each lane owning up to `rows` rows of a strip (strips of lanes*rows
// query rows).  The index into this table travels in bits 1..3 of an item's
// flags (with the two packed queues below).
struct TileConfig {
    int lanes;
    int rows;
};
constexpr int kTileConfigCount = 3;
constexpr TileConfig kTileConfig[kTileConfigCount] = {{16, 32}, {16, 16}, {32, 16}};
// Batched queues: one per tile configuration, and two more for queries of the
// 16 x 32 shape whose cells fit 16-bit packed cells (those that fill one
// strip on their own and are not stacked, when max(sxx) x rows <= 0x7BFF;
// longer ones when max(sxx) x kPackedSpan <= 0x7BFF).  The packed kernel
// scores one query against two targets per lane group; the duo kernel two
// queries against one target per lane group, its workgroups taking whole
// items.  An item's queue travels in bits 1..3 of its flags.
constexpr int kPackedQueue = kTileConfigCount;
constexpr int kDuoQueue = kTileConfigCount + 1;
constexpr int kBatchQueues = kTileConfigCount + 2;
constexpr int kQueueMask = 7;                  // (flags >> 1) & kQueueMask: the item's queue
// Duo items: item.x is the first query, item.y the index in the target list
// of the item's record {second query or -1, its transpose flag}, followed by
// {target, first query's output index} and {target, second query's output
// index} for each of the item.z targets (-1: that query does not pair with
// the target).
constexpr int32_t kPackedMaxScore = 0x7BFF;   // largest packed cell value
constexpr int kPackedMinRows = 225;            // packed queries: first-strip heights of 16..32 rows per lane
// Duo queries: from 225 rows (shorter ones are stacked).  On uneven-2000
// moving the bound from 64 to 225 rows changed the time by under 2%
// (32.9 to 32.4 ms): short duo strips pay their padding rows and per-step
// work at 4-14 rows per lane, about what stacking saves in int32 cells.
constexpr int kDuoMinRows = 225;
constexpr int kPackedSpan = 592;               // a rebased strip's cells span at most kPackedSpan x max(sxx)
constexpr int kMaxTileBytes = 512;     // largest lanes*rows: one code's profile block
constexpr int kRowStep = 2;            // rows-per-lane granularity of a query's first strip
// Stacked items (bit 4 of an item's flags): item.x is the index, in the
// target list, of the stack's record: {members m, rows of the stack (queries
// plus separators)}, then per member {query, K} and {offset of its row of
// output indices (in int32 units of the target list), rows * 2 + transpose};
// the stack's targets are {sequence, index in the stack's target union}, and
// its output table (per member, an output index per union target, -1 where
// the member does not pair with it) sits in the target list too.
constexpr int kStackedFlag = 16;
constexpr int kMaxStackMembers = 16;
// A batched lane group's strip-boundary line: kLineGuard ints in front of
// column 0 and room for every column a strip's wavefront reaches (up to 2G
// past the longest target), so that the bottom row is stored without bounds
// checks; a region of batched_line_stride(longest target) ints per group.
constexpr int kLineGuard = 64;
constexpr int kLineMargin = 128;
inline int32_t batched_line_stride(int64_t max_len) { return (int32_t)((max_len + kLineMargin + 63) & ~int64_t(63)); }
// The long kernel's strip boundaries: kLongLineGuard ints before column 0 (a
// strip's last lane reaches 126 columns left of the target), none after it.
constexpr int kLongLineGuard = 128;
inline int32_t long_line_stride(int64_t max_len) { return (int32_t)((max_len + kLongLineGuard + 63) & ~int64_t(63)); }
int tile_config_index(int lanes, int rows);   // -1 when not in the table

int batched_threads();
size_t batched_lds_bytes();
int batched_groups_per_block();
int batched_occupancy(int* blocks_per_cu);
// One persistent launch of tile configuration `config` over d_items[0, nitems);
// the duo kernel's launch also takes the packed queue (d_items3[0, nitems3),
// counter d_queue3) and the 16 x 32 queue (d_items2[0, nitems2), counter
// d_queue2), whose items its workgroups score one per wave, in that order,
// once the duo items are gone, so that they fill the duo queue's tail.
int batched_launch(int config, int grid, const uint16_t* d_blocks, const SeqInfo* d_info, const int4* d_items,
                   int32_t nitems, const int2* d_targets, int32_t* d_queue, int32_t* d_out, int32_t* d_lines,
                   int32_t line_stride, const Sxx& sxx, int32_t penalty, hipStream_t stream,
                   const int4* d_items2 = nullptr, int32_t nitems2 = 0, int32_t* d_queue2 = nullptr,
                   const int4* d_items3 = nullptr, int32_t nitems3 = 0, int32_t* d_queue3 = nullptr);
int long_rows_for_slots(int slots);
int long_strip_rows(int rows);
size_t long_lds_bytes(int rows, int slots);
int long_occupancy(int rows, size_t lds, int* blocks_per_cu);
// One persistent launch of the long kernel (plus the reset of its queue,
// error flag, slot turns and tag counter): one-wave workgroups take the
// (pair, strip) tasks in list order; strip boundaries go through a ring of
// `nlines` lines of 64-bit (tag, score) words.
struct LongLaunch {
    int rows = 32, grid = 0, slots = 0;
    size_t lds = 0;
    const uint16_t* blocks = nullptr;
    const SeqInfo* info = nullptr;
    const int4* pairs = nullptr;      // {query, target, output index, transpose flag}
    const int4* tasks = nullptr;      // {pair, strip, line read or -1, line written or -1}
    int32_t ntasks = 0;
    int32_t* queue = nullptr;
    int32_t* error = nullptr;
    int32_t* out = nullptr;
    uint64_t* lines = nullptr;
    int32_t* turn = nullptr;
    uint32_t* tag_word = nullptr;     // grows by occupants + 1 per execution; starts at 0
    int32_t line_stride = 0, nlines = 0, occupants = 0;
    const uint8_t* remap = nullptr;
    const uint8_t* slot_code = nullptr;
    Sxx sxx;
    int32_t penalty = 0;
};
int long_launch(const LongLaunch& launch, hipStream_t stream);
// Most resident long workgroups on this device for any strip height.
int long_max_grid(int compute_units, int* grid);
// Line slots per resident long workgroup (plus one group of pairs): an
// occupant rarely waits for the slot's previous line to be read.
constexpr int kLongLinesPerWave = 2;
// Line slots a plan's long kernel uses at most, for a grid of `grid`: the
// planner's ring is kLongLinesPerWave per workgroup plus one group (as many
// pairs as the grid), and a group bounds the distance from a strip to the one
// below.
inline int64_t long_max_lines(int64_t grid) { return kLongLinesPerWave * grid + grid + 1; }

constexpr int kDefaultLongThreshold = 8192;

// Tuning table: the batched tile shape of each query-length bin and the query
// length from which pairs take the long kernel.  Bin b covers query lengths
// [bins[b].min_query, bins[b+1].min_query); bins[0].min_query is 0.
struct TuneBin {
    int32_t min_query;
    int config;                 // index into kTileConfig
};
struct Tuning {
    std::vector<TuneBin> bins;
    int32_t long_threshold = kDefaultLongThreshold;
};
const Tuning& builtin_tuning();
// Reads a table written by msa_nw_tune (nw_tune.hip).
int load_tuning(const char* path, Tuning& out);
int tuning_config_for(const Tuning& t, int32_t query_len);
// The table plans take under `opt`: `tuning_override` when not null, else the
// file the options or MSA_NW_TUNING name (read into `loaded`), else the
// built-in table.
int resolve_tuning(const msa_nw_options& opt, const Tuning* tuning_override, Tuning& loaded, const Tuning*& out);
// Shorter-side length from which plans under `opt` route clean pairs to the
// long kernel (0 under MSA_KERNEL_LONG, INT32_MAX when they never do).
int long_threshold_of(const msa_nw_options& opt, int32_t* threshold);

struct DeviceShape {
    int compute_units = 0;
    size_t lds_per_block = 0;
    size_t lds_per_cu = 0;
};
// Properties of the current device (cached per device).
int device_shape(DeviceShape& out);

int make_sxx(const int32_t* table625, int32_t penalty, Sxx& sxx, bool& symmetric);

struct HostSeq {
    int64_t offset;
    int32_t blocks;
    int32_t first_end;   // rows the reference runs for it as the row sequence
    int32_t last_real;   // columns up to its last non-padding code
};
// Validates a host sequence database and reads each sequence's padding facts.
int scan_sequences(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                   std::vector<HostSeq>& out, int64_t& total_blocks);

struct HostPlan {
    msa_nw_plan_info info{};
    Sxx sxx;
    int32_t penalty = 0;
    // sequence store
    std::vector<uint16_t> blocks;
    std::vector<SeqInfo> seqinfo;
    // plain kernel
    std::vector<int32_t> plain_pairs;     // 2 per pair
    std::vector<int32_t> plain_index;     // output index per pair
    int plain_grid = 0;
    int32_t plain_stride = 0;             // ints per boundary line
    // batched kernel
    std::vector<int4> items;              // grouped by tile configuration, costliest first in each
    std::vector<int2> targets;            // also the stack records and output tables of stacked items
    int64_t batch_first[kBatchQueues + 1] = {};   // items of queue c: [first[c], first[c+1])
    int batch_grid[kBatchQueues] = {};
    int32_t batch_stride = 0;
    bool packed_lines = false;            // the packed queue has queries longer than a strip (two lines per group)
    // long kernel
    std::vector<int4> long_pairs;
    std::vector<int4> long_tasks;         // {pair, strip, line read, line written} per strip
    int long_rows = 32, long_slots = 0, long_grid = 0;
    size_t long_lds = 0;
    int32_t long_stride = 0, long_nlines = 0, long_occupants = 0;   // stride in 64-bit words
    uint8_t remap[32] = {};
    uint8_t slot_code[32] = {};
};

// Builds a HostPlan for the current device.  `pairs` may be null (all pairs
// in reference order).  Input errors are reported before any device call.
// `tuning`, when not null, replaces the table the options name.
// `copy_blocks` false leaves hp.blocks empty: the sequences are already on the
// device at the same offsets (a resident store) and the data block carries
// only the work lists.
int plan_host(HostPlan& hp, const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
              const int32_t* pairs, int64_t npairs, const int32_t* table625, int32_t penalty,
              const msa_nw_options& opt, const Tuning* tuning = nullptr, bool copy_blocks = true);

// msa_nw_plan_create with an explicit tuning table (null: from the options).
// One execution of a plan on `stream` (graph or eager), with the plan's lock
// held by the caller.
int execute_locked(msa_nw_plan plan, int32_t* d_out, hipStream_t stream);

int create_plan(msa_nw_plan* plan, const uint16_t* blocks, const int64_t* offset, const int32_t* size,
                int32_t nseq, const int32_t* pairs, int64_t npairs, const int32_t* table625, int32_t penalty,
                const msa_nw_options& opt, const Tuning* tuning, hipStream_t stream);

// The HostPlan's arrays inside one block of memory (offsets 256-aligned).
struct DataLayout {
    size_t blocks = 0, seqinfo = 0, plain_pairs = 0, plain_index = 0, items = 0, targets = 0, long_pairs = 0,
           long_tasks = 0;
    size_t bytes = 0;
};
DataLayout data_layout(const HostPlan& hp);
void pack_data(const HostPlan& hp, const DataLayout& lay, void* host_block);

// Scratch one execution needs (int32 words), laid out the same way.
// The whole block is zeroed once when it is allocated (the long kernel's tag
// counter and lines must start at 0); executions clear what they need.
struct ScratchNeed {
    size_t plain_lines = 0, batch_lines = 0, long_lines = 0;   // long_lines in 32-bit words (2 per line word)
    size_t long_slots = 0;      // line slots of the long kernel: one turn each
    size_t words() const { return plain_lines + batch_lines + long_lines + long_slots + 4 * 64; }
};
ScratchNeed scratch_need(const HostPlan& hp);

struct Scratch {
    int32_t* plain_lines = nullptr;
    int32_t* batch_lines = nullptr;
    int32_t* long_lines = nullptr;
    int32_t* batch_queue = nullptr;
    int32_t* long_queue = nullptr;
    int32_t* long_error = nullptr;
    uint32_t* long_tag = nullptr;
    int32_t* long_turn = nullptr;
    uint64_t* long_lines64 = nullptr;
};
Scratch carve_scratch(const ScratchNeed& need, int32_t* base);

// Launches every kernel of one execution of `hp` (data at `d_data`; the
// sequence blocks at `d_store` instead when it is not null).
int launch_plan(const HostPlan& hp, const DataLayout& lay, const void* d_data, const Scratch& scratch,
                int32_t* d_out, hipStream_t stream, const uint16_t* d_store = nullptr);

}  // namespace msa
