// The batched Needleman-Wunsch kernel: one persistent launch per tile
// configuration over a queue of work items, each item one "query" sequence and
// a run of "targets" (the other sequence of each pair), sorted by cost by the
// planner (nw_plan.hip).
//
// Which pairs come here.  Only pairs whose score is the plain global
// alignment of the two sequences cut at their first padding code: the
// reference's padding rules (src/pairwise/needleman/impl/sequential.cpp:
// 40-66) reduce to that when the column sequence holds no padding code before
// its last residue.  Either sequence of a pair can be the query (the planner
// picks the cheaper orientation); with a non-symmetric table the profile is
// built from the transposed table for pairs whose query is the reference's
// column sequence, so every score keeps the reference's orientation (:87-91).
//
// Shape of the computation.  Every wave of a workgroup works alone: it takes
// an item from an atomic counter, splits its query into row strips, and its
// lane groups (64/G groups of G lanes) score 64/G targets at a time.  Inside a
// group, lane k owns R consecutive rows of the strip; the DP runs as an
// anti-diagonal wavefront -- at step s lane k updates column s-k -- with the
// bottom score of each lane moving one lane down per step by a DPP shift and
// each lane reading its column's residue code from a small ring in LDS.  The
// substitution scores of a lane's rows come from the strip's profile in the
// wave's LDS region: R bytes per (column code, lane) in 16-byte planes
// (conflict-free), read one step ahead of use.  A strip's bottom row goes to
// a per-group line in global memory for the next strip, one buffer store per
// step from the group's last lane.
//
// Tile configurations.  Each query's tile shape is picked per query length by
// the planner from the tuning table, and each shape has its own kernel and
// queue: G lanes per pair (16 or 32) and
// at most RMAX rows per lane (16 or 32), so a full strip is G*RMAX rows.  The
// first strip of a query takes R0 rows per lane, R0 a multiple of kRowStep (2),
// so at most 2*G - 1 padding rows are computed per query.
//
// Neutral padding.  The recurrence runs in the shifted domain G = H +
// (i+j)*penalty, where every boundary cell is 0 and every cell is
// max(G_diag + sxx, G_up, G_left) with sxx = T + 2*penalty >= 0.  Rows and
// columns whose profile entries are 0 and whose neighbours are all 0 stay 0,
// so the query is padded with such rows at the top of its first strip (the
// last real row is always the last lane's register R-1 of the last strip) and
// each target with such columns at its left (every group ends exactly at the
// end of a G-step chunk).  The inner loop therefore needs no per-step
// masking: results are captured at chunk ends.
//
// Packed cells.  Where a strip's values fit 15 bits, two cells share a
// 32-bit register as 16-bit halves: the packed queue scores one query (or
// one stack) against two targets per lane group, the duo queue two queries
// against one target per lane group, its workgroups taking whole items and
// sharing one profile; the duo launch then drains the packed and the 16 x 32
// queues, one item per wave (both sections below).
//
// Build layout.  The device code is this header; each tile configuration's
// kernel is instantiated in its own source file (nw_batched.hip for 16 x 32,
// nw_batched_16x16.hip, nw_batched_32x16.hip, nw_batched_packed.hip,
// nw_batched_duo.hip), so that they compile in parallel.
#pragma once

#include "msa_internal.hpp"
#include "nw_device.hpp"
#include "nw_planner.hpp"

#include <type_traits>

namespace msa {
namespace batched {

constexpr int kCodes = 25;           // codes 0..23 and the neutral code 24
constexpr int kNull = 24;
constexpr int kWaves = 4;            // waves per workgroup (independent)
constexpr int kMaxGroups = 4;        // lane groups per wave at 16 lanes
constexpr int kProfWords = kCodes * kMaxTileBytes / 4;   // one wave's profile region
// The sxx table in LDS for the profile builders: rows of 28 bytes (the row
// residue's code; columns: the column code), row 24 and columns 24..27 zero,
// so that a neutral row and the neutral column code both read 0; a second
// copy holds the transposed table.
constexpr int kTabStride = 28;
constexpr int kTabCopy = kCodes * kTabStride;                  // 700 bytes
constexpr int kTableBytes = (2 * kTabCopy + 15) / 16 * 16;     // both copies, padded to 16 bytes
// One wave's code rings: per lane group 4G bytes (one ring of 2G columns,
// twice), 256 bytes a wave; in the packed kernel 6G bytes (two rings of 3G
// bytes, one per target), 384 bytes a wave of four 16-lane groups (the duo
// kernel uses the first ring of each group).
constexpr int kRingBytes = 256;
constexpr int kPackedRingBytes = 384;
constexpr int kItemSlotBytes = 16;   // duo kernel: the workgroup's current item

constexpr bool packed_queue(int cfg) { return cfg == kPackedQueue || cfg == kDuoQueue; }

// Dynamic LDS of a workgroup: the waves' profile regions (in the duo kernel
// one shared profile), the table, the rings, the duo kernel's item slot.
constexpr size_t lds_bytes(int cfg = 0)
{
    return (size_t)kWaves * kProfWords * 4 + kTableBytes + kWaves * (packed_queue(cfg) ? kPackedRingBytes : kRingBytes) +
           (cfg == kDuoQueue ? kItemSlotBytes : 0);
}

static_assert(kTileConfigCount == 3, "one kernel source per tile configuration");
static_assert(kPackedQueue == 3 && kDuoQueue == 4 && kBatchQueues == 5, "the packed kernels are the last two queues");

struct BatchArgs {
    const uint16_t* blocks;
    const SeqInfo* info;
    const int4* items;          // {query or stack record, first target, target count, flags}
    const int2* targets;        // {sequence, output index}; stacked items: {sequence, index in the
                                // stack's union}, and the stack records and output tables
    int32_t nitems;
    int32_t* queue;
    int32_t* out;
    int32_t* lines;             // one line of line_stride ints per lane group
    int32_t line_stride;
    int32_t penalty;
    int32_t margin;             // packed kernel: 16 x the largest sxx (how far below the line a base is set)
    Sxx sxx;
    const int4* items2;         // duo kernel: the 16 x 32 queue, scored after the duo items
    int32_t* queue2;
    int32_t nitems2;
    int32_t nitems3;
    const int4* items3;         // duo kernel: the packed queue, scored before the 16 x 32 queue
    int32_t* queue3;
};

// Calls f(integral_constant<int, U>) for U = B..E-1, so that U can name DPP
// controls, which must be compile-time constants.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B> {});
        static_for<B + 1, E>(f);
    }
}

// Every lane of each 16-lane row reads lane U of its row.
template <int U>
__device__ __forceinline__ int row_bcast(int v)
{
    return __builtin_amdgcn_mov_dpp(v, dev::kRowNewBcast0 + U, 0xF, 0xF, false);
}

// Profile layout inside one code's block: NP = ceil(W/4) planes of 16 bytes
// per lane ([plane][lane][4 words]; W = ceil(R/4) words of four rows per
// lane), so that every read of a lane is the lane's 16-byte slot plus a
// constant, one address per step, and the block stride G*NP*16 bytes is a
// power of two.  Words of the last plane past W are never read.
template <int R>
constexpr int prof_planes = (dev::prof_words<R> + 3) / 4;
template <int G, int R>
constexpr int code_words = G * 4 * prof_planes<R>;

template <int G, int R>
__device__ __forceinline__ void load_prof(uint32_t (&w)[dev::prof_words<R>], const dev::lds_u32* code_base, int k)
{
    constexpr int W = dev::prof_words<R>;
    constexpr int F = W / 4, REM = W % 4;
    const dev::lds_u32* slot = code_base + k * 4;
#pragma unroll
    for (int j = 0; j < F; ++j) {
        const dev::u32x4 v = *(const dev::lds_u32x4*)(slot + j * G * 4);
        w[4 * j + 0] = v.x;
        w[4 * j + 1] = v.y;
        w[4 * j + 2] = v.z;
        w[4 * j + 3] = v.w;
    }
    if constexpr (REM >= 2) {
        const dev::u32x2 v = *(const dev::lds_u32x2*)(slot + F * G * 4);
        w[4 * F + 0] = v.x;
        w[4 * F + 1] = v.y;
    }
    if constexpr (REM % 2) w[4 * F + (REM & 2)] = slot[F * G * 4 + (REM & 2)];
}

// Word index of word `w` of lane `k` inside a code block (inverse of load_prof).
template <int G>
__device__ __forceinline__ int prof_word(int k, int w)
{
    return (w >> 2) * G * 4 + k * 4 + (w & 3);
}

// The profile words of one (lane, word) slot for every column code.  Byte t
// of the word for code b is the table entry at byte row[t] + b of `tab` (row[t]
// is the table row of the slot's row t: a row of either copy, or the zero row
// 24 for a neutral row).  Four 4-byte table reads give a 4 x 4 block of bytes
// (rows x codes 4j..4j+3), which eight byte permutes transpose into the words
// of four codes.
__device__ __forceinline__ void store_slot(dev::lds_u32* dst, int code_words, const dev::lds_u8* tab,
                                           const int (&row)[4])
{
#pragma unroll
    for (int j = 0; j < (kCodes + 3) / 4; ++j) {
        const uint32_t a0 = *(const dev::lds_u32*)(tab + row[0] + 4 * j);
        const uint32_t a1 = *(const dev::lds_u32*)(tab + row[1] + 4 * j);
        const uint32_t a2 = *(const dev::lds_u32*)(tab + row[2] + 4 * j);
        const uint32_t a3 = *(const dev::lds_u32*)(tab + row[3] + 4 * j);
        // v_perm_b32: selector bytes 0-3 pick bytes of the second operand,
        // 4-7 bytes of the first.
        const uint32_t lo01 = __builtin_amdgcn_perm(a1, a0, 0x05010400u);   // a0.0 a1.0 a0.1 a1.1
        const uint32_t hi01 = __builtin_amdgcn_perm(a1, a0, 0x07030602u);   // a0.2 a1.2 a0.3 a1.3
        const uint32_t lo23 = __builtin_amdgcn_perm(a3, a2, 0x05010400u);
        const uint32_t hi23 = __builtin_amdgcn_perm(a3, a2, 0x07030602u);
        const uint32_t o[4] = {__builtin_amdgcn_perm(lo23, lo01, 0x05040100u),
                               __builtin_amdgcn_perm(lo23, lo01, 0x07060302u),
                               __builtin_amdgcn_perm(hi23, hi01, 0x05040100u),
                               __builtin_amdgcn_perm(hi23, hi01, 0x07060302u)};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * j + i < kCodes) dst[(4 * j + i) * code_words] = o[i];
    }
}

// Builds the profile of one strip of R rows per lane (R a run-time value:
// one instance serves every strip height): rows [row0, row0 + G*R) of the
// padded query (rows below pad_top are neutral; bytes past a lane's R rows
// are 0).  Layout as load_prof<G, R> reads it.
template <int G>
__device__ __forceinline__ void build_profile(dev::lds_u32* prof, const dev::lds_u8* tab,
                                              const uint16_t* __restrict__ qseq, int row0, int pad_top,
                                              bool transpose, int R, int lane)
{
    const int W = (R + 3) / 4;
    const int code_words_r = G * 4 * ((W + 3) / 4);
    const int copy = transpose ? kTabCopy : 0;
    for (int slot = lane; slot < G * W; slot += 64) {
        const int w = slot / G;
        const int k = slot % G;
        int row[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = row0 + k * R + 4 * w + t - pad_top;
            const int code = (q >= 0 && 4 * w + t < R) ? dev::code_at(qseq, q) : kNull;
            row[t] = copy + min(code, kNull) * kTabStride;
        }
        store_slot(prof + prof_word<G>(k, w), code_words_r, tab, row);
    }
}

// Builds the profile of a stacked strip of R rows per lane: rows [0, pad_top)
// neutral, then the members in order, each after a neutral separator row but
// the first; a member's rows take its own orientation.  Same layout as
// build_profile<G>.
template <int G>
__device__ __forceinline__ void build_profile_stacked(dev::lds_u32* prof, const dev::lds_u8* tab,
                                                      const BatchArgs& a, int stack, int pad_top, int R, int lane)
{
    const int m = a.targets[stack].x;
    const int W = (R + 3) / 4;
    const int code_words_r = G * 4 * ((W + 3) / 4);
    for (int slot = lane; slot < G * W; slot += 64) {
        const int w = slot / G;
        const int k = slot % G;
        int row[4];
        for (int t = 0; t < 4; ++t) {
            row[t] = kNull * kTabStride;
            const int u = k * R + 4 * w + t - pad_top;
            if (u < 0 || 4 * w + t >= R) continue;
            int start = 0;
            for (int i = 0; i < m; ++i) {
                const int lw = a.targets[stack + 2 + 2 * i].y;
                const int lq = lw >> 1;
                if (u < start + lq) {
                    const int code = dev::code_at(a.blocks + a.info[a.targets[stack + 1 + 2 * i].x].x, u - start);
                    row[t] = ((lw & 1) ? kTabCopy : 0) + min(code, kNull) * kTabStride;
                    break;
                }
                if (u == start + lq) break;     // separator row: neutral
                start += lq + 1;
            }
        }
        store_slot(prof + prof_word<G>(k, w), code_words_r, tab, row);
    }
}

// What a lane's rows start from and which of them it reports.  Its rows begin
// at K = k_lo, from row kb on at k_hi (a stacked member's separator); after
// the pass it returns row sel (-1: none).  An unstacked strip starts every
// row at 0 (the shifted boundary) and reports its last row on the last lane.
struct LaneRows {
    int k_lo, k_hi, kb, sel;
};

// The bottom-row store of a strip: the wave's line regions as one raw buffer
// (dword 3 of the descriptor: 32-bit data format, as for any raw buffer
// access on gfx9), and this lane's byte offset into it -- out of range
// (the store is dropped) on every lane but a group's last.
constexpr int kRawBufferFlags = 0x00020000;
constexpr uint32_t kOutOfRange = 0x80000000u;
struct LineStore {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t lane_off;
};

// Per-group view of the target being scored.
struct TargetState {
    const uint16_t* seq;
    int len;        // columns
    int lpad;       // neutral columns in front, so that len + lpad + G-1 is a multiple of G
    int chunks;     // (lpad + len + G-1) / G, 0 for an idle group
};

template <int G>
__device__ __forceinline__ int col_code(const TargetState& t, int chunk, int k)
{
    const int idx = chunk * G + k - t.lpad;
    return (idx >= 0 && idx < t.len) ? min(dev::code_at(t.seq, idx), kNull) : kNull;
}

template <int G>
__device__ __forceinline__ int line_value(const TargetState& t, const int32_t* __restrict__ line, int chunk, int k)
{
    const int idx = chunk * G + k - t.lpad;
    return (idx >= 0 && idx < t.len) ? line[idx] : 0;
}

// A lane's column of the target, chunk after chunk (idx = chunk*G + k -
// lpad), with its block and bit shift kept up to date by additions instead
// of a division by 3 per chunk: a chunk moves G columns, i.e. G/3 blocks and
// G%3 slots of 5 bits.
template <int G>
struct ColumnCursor {
    int idx;
    int blk;
    int shift;      // 1 + 5 * (idx mod 3)

    __device__ __forceinline__ ColumnCursor(const TargetState& t, int chunk, int k)
    {
        idx = chunk * G + k - t.lpad;
        const int q = (idx >= 0 ? idx : idx - 2) / 3;   // floor, also left of the target
        blk = q;
        shift = 1 + 5 * (idx - 3 * q);
    }
    __device__ __forceinline__ void advance()
    {
        idx += G;
        blk += G / 3;
        shift += 5 * (G % 3);
        if (shift > 11) {
            shift -= 15;
            blk += 1;
        }
    }
    // The residue code (the neutral code outside the target), as col_code.
    __device__ __forceinline__ int code(const TargetState& t) const
    {
        return (idx >= 0 && idx < t.len) ? min((t.seq[blk] >> shift) & 0x1F, kNull) : kNull;
    }
    // The line value (0 outside the target), as line_value.
    __device__ __forceinline__ int line(const TargetState& t, const int32_t* __restrict__ l) const
    {
        return (idx >= 0 && idx < t.len) ? l[idx] : 0;
    }
};

// One strip of R rows per lane over the whole target.  LOAD: the strip above
// left its bottom row in `line`; STORE: this strip leaves its bottom row there.
// Returns (on the group's last lane) the last row's score at the last column.
//
// Per step, each lane needs the code of its column and the cell above its
// first row.  Codes come from the group's code ring in LDS (4G bytes): at step
// s lane k works virtual column s-k, and column i sits at byte i mod 2G and
// again at i mod 2G + 2G, so a lane reads the code for step u of a chunk at a
// per-chunk base plus u -- one LDS byte read two steps ahead, no VALU work;
// each chunk's codes go in one chunk ahead.  The reads are volatile so that
// they stay single-byte reads (merged into wider loads they would cost VALU
// extracts).  The cell above: lane 0 takes the line value, lane k lane k-1's
// bottom cell.  With 16-lane groups (one DPP row each) the line value is
// broadcast from the chunk's line vector with row_newbcast and the row shift
// takes it as its "old" value; with 32-lane groups the vector is shifted down
// the wave and the carry into the upper row comes from a row broadcast of
// lane 15.
template <int G, int R, bool LOAD, bool STORE>
__device__ __forceinline__ int run_strip(const dev::lds_u32* prof, const TargetState& t, int32_t* __restrict__ line,
                         volatile dev::lds_u8* cring, const LineStore& ls, int wchunks, int k, const LaneRows& lr)
{
    static_assert(G == 16 || G == 32, "lane groups of 16 or 32");
    static_assert(code_words<G, R> * 4 <= kMaxTileBytes, "profile block exceeds the wave's region");
    constexpr int W = dev::prof_words<R>;
    constexpr int CODE_WORDS = code_words<G, R>;
    constexpr int RING = 2 * G;
    int h[R];
#pragma unroll
    for (int r = 0; r < R; ++r) h[r] = r < lr.kb ? lr.k_lo : lr.k_hi;
    int dtop = 0;

    // Code ring: virtual column i at byte i mod 2G and again at i mod 2G + 2G;
    // columns -G..-1 are neutral, chunk 0 goes in before the first read.
    cring[G + k] = kNull;
    cring[3 * G + k] = kNull;
    {
        const int c0 = col_code<G>(t, 0, k);
        cring[k] = c0;
        cring[RING + k] = c0;
    }
    int cc_next = col_code<G>(t, 1, k);
    ColumnCursor<G> cur(t, 2, k);           // the column fetched at the end of chunk c is chunk c+2's
    int ln = 0, ln_next = 0;
    if constexpr (LOAD) {
        ln = line_value<G>(t, line, 0, k);
        ln_next = line_value<G>(t, line, 1, k);
    }

    // Codes of steps 0 and 1 (virtual columns -k and 1-k).
    const int base0 = (RING - k) & (RING - 1);
    int b = cring[base0];
    int bn = cring[base0 + 1];
    uint32_t w[W];
    load_prof<G, R>(w, prof + b * CODE_WORDS, k);

    for (int c = 0; c < wchunks; ++c) {
        // Byte of the code for step u + 2 of this chunk: base + u.
        const int base = (c * G + 2 - k) & (RING - 1);
        // Byte offset in the wave's line buffer of the bottom cell the
        // group's last lane makes at step 0 of this chunk (column c*G - (G-1)
        // - lpad); far out of range on every other lane.
        const uint32_t voff = ls.lane_off + (uint32_t)(c * G - (G - 1) - t.lpad) * 4u;
        static_for<0, G>([&](auto uc) {
            constexpr int u = decltype(uc)::value;
            // Chunk c+1 replaces chunk c-1, whose last reader was step G-4.
            if constexpr (u == G - 3) {
                const int slot = ((c + 1) * G + k) & (RING - 1);
                cring[slot] = cc_next;
                cring[RING + slot] = cc_next;
            }
            const int bnn = cring[base + u];
            uint32_t wn[W];
            load_prof<G, R>(wn, prof + bn * CODE_WORDS, k);

            int hup;
            if constexpr (LOAD) {
                if constexpr (G == 16) {
                    hup = dev::dpp<dev::kRowShr1>(row_bcast<u>(ln), h[R - 1]);
                } else {
                    hup = dev::group_shr<G>(ln, h[R - 1]);
                    if constexpr (u < G - 1) ln = dev::group_shl<G>(ln);
                }
            } else {
                if constexpr (G == 16) hup = dev::dpp_zero<dev::kRowShr1>(h[R - 1]);
                else hup = dev::group_shr<G>(0, h[R - 1]);
            }

            dev::cells<R>(h, hup, dtop, w);
            dtop = hup;
            // The bottom cell goes straight to the line: one buffer store
            // (the step's 4u in the scalar offset, no VALU work), whose lanes
            // other than the group's last are out of range and dropped;
            // columns outside the target land in the line's guard and tail,
            // which nothing reads.
            if constexpr (STORE)
                __builtin_amdgcn_raw_buffer_store_b32((uint32_t)h[R - 1], ls.rsrc, (int)voff, 4 * u, 0);
            b = bn;
            bn = bnn;
#pragma unroll
            for (int q = 0; q < W; ++q) w[q] = wn[q];
        });
        if constexpr (LOAD) ln = ln_next;
        cc_next = cur.code(t);
        if constexpr (LOAD) ln_next = cur.line(t, line);
        cur.advance();
    }
    // Every row keeps its value over the neutral columns past the target (the
    // code ring and the line give the neutral code and 0 there: a row's value
    // is then the max of its own and the row above's, which is no larger), so
    // the rows hold their final values after the wave's last chunk.
    // The row compares stay here: hoisted above the loop, their 64-bit lane
    // masks would hold R pairs of scalar registers across the strip.
    int sel = lr.sel;
    asm volatile("" : "+v"(sel));
    int result = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) result = r == sel ? h[r] : result;
    return result;
}

// First strips: R0 rows per lane, R0 in kRowStep..RMAX.
template <int G, int RMAX, bool STORE, int R = kRowStep>
__device__ __forceinline__ int run_first_strip(int r0, const dev::lds_u32* prof, const TargetState& t, int32_t* line,
                               volatile dev::lds_u8* cring, const LineStore& ls, int wchunks, int k,
                               const LaneRows& lr)
{
    if constexpr (R >= RMAX) {
        return run_strip<G, RMAX, false, STORE>(prof, t, line, cring, ls, wchunks, k, lr);
    } else {
        if (r0 == R) return run_strip<G, R, false, STORE>(prof, t, line, cring, ls, wchunks, k, lr);
        return run_first_strip<G, RMAX, STORE, R + kRowStep>(r0, prof, t, line, cring, ls, wchunks, k, lr);
    }
}


// LDS accesses of one wave to its own region are ordered; this keeps the
// compiler from moving them across the profile rebuild and waits for them.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
}

// Lane k's rows of a stacked strip (R0 rows per lane, pad_top neutral rows
// on top): member i > 0 starts its K at its separator row, and a lane holds at
// most one separator and one member's last row (every member has at least a
// lane's rows).  `member`: the member whose last row the lane reports, or -1.
struct StackedRows {
    LaneRows rows;
    int member;
};
__device__ __forceinline__ StackedRows stacked_rows(const BatchArgs& a, int stack, int pad_top, int R0, int k)
{
    const int m = a.targets[stack].x;
    StackedRows sr {{0, 0, R0, -1}, -1};
    const int lo = k * R0;
    int start = pad_top;
    for (int i = 0; i < m; ++i) {
        const int K = a.targets[stack + 1 + 2 * i].y;
        const int lq = a.targets[stack + 2 + 2 * i].y >> 1;
        if (i > 0) {
            const int sep = start - 1;
            if (sep <= lo) sr.rows.k_lo = K;
            else if (sep < lo + R0) sr.rows.kb = sep - lo, sr.rows.k_hi = K;
        }
        const int last = start + lq - 1;
        if (last >= lo && last < lo + R0) sr.rows.sel = last - lo, sr.member = i;
        start += lq + 1;
    }
    return sr;
}

// All strips of one item with tile shape (G, RMAX).
template <int G, int RMAX>
__device__ __forceinline__ void run_item(const BatchArgs& a, const int4 item, dev::lds_u32* prof, const dev::lds_u8* tab,
                         dev::lds_u8* rings, int32_t* lines, int lane)
{
    constexpr int GROUPS = 64 / G;
    constexpr int STRIP_MAX = G * RMAX;
    const int g = lane / G;
    const int k = lane % G;
    int32_t* line = lines + (int64_t)g * a.line_stride + kLineGuard;
    volatile dev::lds_u8* cring = rings + g * 4 * G;
    // The wave's line regions as one buffer (a wave-uniform descriptor).
    LineStore ls;
    {
        const uint64_t wb = (uint64_t)(uintptr_t)lines;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)wb);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(wb >> 32));
        ls.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)hi << 32) | lo), 0,
                                                    kMaxGroups * a.line_stride * 4, kRawBufferFlags);
        ls.lane_off = k == G - 1 ? (uint32_t)(g * a.line_stride + kLineGuard) * 4u : kOutOfRange;
    }
    const bool stacked = (item.w & kStackedFlag) != 0;
    const SeqInfo qi = a.info[stacked ? 0 : item.x];
    const int Lq = stacked ? a.targets[item.x].y : qi.z;  // rows: the query's, or the stack's with its separators
    const uint16_t* qseq = a.blocks + qi.x;
    const bool transpose = (item.w & 1) != 0;
    // Strips: a first strip of R0 rows per lane, then full strips of RMAX
    // (a stack is one strip).
    const int full = (Lq - 1) / STRIP_MAX;            // strips after the first
    const int rem = Lq - full * STRIP_MAX;            // rows of the first strip
    const int R0 = min(RMAX, ((rem + G * kRowStep - 1) / (G * kRowStep)) * kRowStep);
    const int pad_top = G * R0 - rem;
    const int nst = full + 1;
    const int rounds = (item.z + GROUPS - 1) / GROUPS;
    for (int rd = 0; rd < rounds; ++rd) {
        const int ti = rd * GROUPS + g;
        TargetState t {a.blocks, 0, 0, 0};
        int2 tg = make_int2(0, -1);
        if (ti < item.z) {
            tg = a.targets[item.y + ti];
            const SeqInfo si = a.info[tg.x];
            t.seq = a.blocks + si.x;
            t.len = si.z;
            t.chunks = (t.len + 2 * G - 2) / G;
            t.lpad = t.chunks * G - (t.len + G - 1);
        }
        int wchunks = t.chunks;
        if constexpr (GROUPS > 2) wchunks = max(wchunks, __shfl_xor(wchunks, 16));
        if constexpr (GROUPS > 1) wchunks = max(wchunks, __shfl_xor(wchunks, 32));

        int res = 0;
        for (int st = 0; st < nst; ++st) {
            const int R = st == 0 ? R0 : RMAX;
            const int row0 = st == 0 ? 0 : G * R0 + (st - 1) * STRIP_MAX;
            if (nst > 1 || rd == 0) {
                wave_lds_sync();
                if (stacked) build_profile_stacked<G>(prof, tab, a, item.x, pad_top, R, lane);
                else build_profile<G>(prof, tab, qseq, row0, pad_top, transpose, R, lane);
                wave_lds_sync();
            }
            const bool store = st + 1 < nst;
            // An unstacked query reports its last row: the last lane's row
            // R - 1 of the last strip.  A stacked lane's rows come from the
            // stack (worked out per pass, so that nothing of it stays live
            // across the strip).
            LaneRows lr {0, 0, R, (st + 1 == nst && k == G - 1) ? R - 1 : -1};
            if (stacked) lr = stacked_rows(a, item.x, pad_top, R0, k).rows;
            if (st == 0) {
                res = store ? run_first_strip<G, RMAX, true>(R0, prof, t, line, cring, ls, wchunks, k, lr)
                            : run_first_strip<G, RMAX, false>(R0, prof, t, line, cring, ls, wchunks, k, lr);
            } else {
                res = store ? run_strip<G, RMAX, true, true>(prof, t, line, cring, ls, wchunks, k, lr)
                            : run_strip<G, RMAX, true, false>(prof, t, line, cring, ls, wchunks, k, lr);
            }
            // The next strip reads this strip's bottom row from `line`.
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
        if (tg.y >= 0) {
            if (!stacked) {
                if (k == G - 1) a.out[tg.y] = res - (Lq + t.len) * a.penalty;
            } else {
                const int mine = stacked_rows(a, item.x, pad_top, R0, k).member;
                if (mine >= 0) {
                    const int2 m0 = a.targets[item.x + 1 + 2 * mine];     // {query, K}
                    const int2 m1 = a.targets[item.x + 2 + 2 * mine];     // {output row, rows * 2 + transpose}
                    const int o = reinterpret_cast<const int32_t*>(a.targets)[m1.x + tg.y];
                    if (o >= 0) a.out[o] = res - m0.y - ((m1.y >> 1) + t.len) * a.penalty;
                }
            }
        }
    }
}

// ---- Packed cells: two targets per lane group ---------------------------
//
// For a query that fills one 16 x 32 strip on its own, every cell of the
// shifted recurrence lies in [0, max(sxx) * rows]; the planner sends such a
// query here only when that bound is at most 0x7BFF.  A lane then holds a
// row of two targets in one register, target A in the low 16 bits and target
// B in the high 16 bits, and one step of a row is three instructions for two
// cells: v_perm_b32 takes the row's profile byte for A's column code and for
// B's into the two halves, v_pk_add_u16 adds them to the diagonal cells, and
// v_pk_maximum3_f16 takes the max3.  The f16 maximum of 16-bit patterns in
// 0x0000..0x7BFF (positive finite f16 values, denormals included) is the
// integer maximum of the patterns, so every cell is the exact integer of the
// int32 recurrence.

// Max of three packed cells (the f16 maximum of their bit patterns).
__device__ __forceinline__ uint32_t pk_max3(uint32_t a, uint32_t b, uint32_t c)
{
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 x = __builtin_bit_cast(h2, a), y = __builtin_bit_cast(h2, b), z = __builtin_bit_cast(h2, c);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_maximum(__builtin_elementwise_maximum(x, y), z));
}

__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b)
{
    typedef unsigned short u2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u2, a) + __builtin_bit_cast(u2, b));
}

// R packed rows of one lane at one column: wa / wb are the profile words of
// A's and B's column codes (row r in byte r%4 of word r/4).
template <int R>
__device__ __forceinline__ void cells_packed(uint32_t (&h)[R], uint32_t up, uint32_t diag,
                                             const uint32_t (&wa)[dev::prof_words<R>],
                                             const uint32_t (&wb)[dev::prof_words<R>])
{
    // bytes: A's byte r%4, 0, B's byte r%4, 0
    constexpr uint32_t kSel[4] = {0x0C040C00u, 0x0C050C01u, 0x0C060C02u, 0x0C070C03u};
    uint32_t u = up;
    // Row r+1's permute comes before row r's max3 and its add after: no
    // packed add then reads the instruction just before it, which would cost
    // a wait state before the max3 that reads the add.
    uint32_t t = pk_add(diag, __builtin_amdgcn_perm(wb[0], wa[0], kSel[0]));
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint32_t s = 0;
        if (r + 1 < R) s = __builtin_amdgcn_perm(wb[(r + 1) >> 2], wa[(r + 1) >> 2], kSel[(r + 1) & 3]);
        const uint32_t nh = pk_max3(t, u, h[r]);
        if (r + 1 < R) t = pk_add(h[r], s);
        h[r] = nh;
        u = nh;
    }
}

// Where a packed strip below another reads its two lines, and where a strip
// with one below it stores its bottom rows: each target's line holds int32
// scores of the shifted domain, as the int32 kernel's lines do.
struct PackedLines {
    const int32_t* la;          // target A's line, column 0
    const int32_t* lb;          // target B's line
    __amdgpu_buffer_rsrc_t rsrc;    // the wave's line regions
    uint32_t off_a, off_b;      // this lane's byte offsets of A's and B's lines (out of range but on the last lane)
};

// A lane's line value of `t` at its column of chunk `chunk` (-1 outside the
// target: its relative value is 0).
__device__ __forceinline__ int packed_line_abs(const TargetState& t, const int32_t* __restrict__ line, int chunk, int k)
{
    const int idx = chunk * 16 + k - t.lpad;
    return (idx >= 0 && idx < t.len) ? line[idx] : -1;
}

__device__ __forceinline__ uint32_t packed_rel(int abs_a, int base_a, int abs_b, int base_b)
{
    const uint32_t ra = abs_a < 0 ? 0u : (uint32_t)(abs_a - base_a);
    const uint32_t rb = abs_b < 0 ? 0u : (uint32_t)(abs_b - base_b);
    return ra | (rb << 16);
}

__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b)
{
    typedef unsigned short u2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u2, a) - __builtin_bit_cast(u2, b));
}

// One strip of R rows per lane against targets A and B.  LOAD: the strip
// above left each target's bottom row in its line; STORE: this strip leaves
// its bottom rows there (absolute int32 values, one buffer store per target
// and step from the group's last lane).  A packed cell is the cell's value
// minus its target's base.  A strip with no strip above never needs a base:
// its cells are at most max(sxx) x rows.  A strip below another starts at
// base 0 (its first columns' cells are at most max(sxx) x columns) and every
// fourth chunk raises each base to the line value at the first lane's next
// column minus `margin` (16 max(sxx)): every cell the group holds then is at
// least the base (a cell is at least the line value above it, and the line
// grows by at most max(sxx) a column), and until the next raise at most
// 592 max(sxx) above it (rows x max(sxx) over the line, and the line's growth
// over 4 chunks), which the planner keeps within 0x7BFF.  Each target has its
// own code ring of 3G bytes: virtual column i at byte i mod 2G, and columns
// with i mod 2G < G again at i mod 2G + 2G, so that a lane reads the code for
// step u of a chunk at a per-chunk base plus u, as in run_strip.  A stacked
// strip starts its rows at the lane's K values (`lr`, both halves alike: the
// two targets share the stack).  Returns both targets' cells of row lr.sel
// (row R-1 when lr.sel < 0) and their bases.
template <int G, int R, bool LOAD, bool STORE>
__device__ __forceinline__ uint32_t run_strip_packed(const dev::lds_u32* prof, const TargetState& ta,
                                                     const TargetState& tb, volatile dev::lds_u8* ra,
                                                     volatile dev::lds_u8* rb, const PackedLines& pl, int wchunks,
                                                     int k, int margin, int& base_a, int& base_b, const LaneRows& lr)
{
    static_assert(G == 16, "packed strips use 16-lane groups");
    constexpr int W = dev::prof_words<R>;
    constexpr int CODE_WORDS = code_words<G, R>;
    constexpr int RING = 2 * G;
    uint32_t h[R];
    {
        const uint32_t klo = (uint32_t)lr.k_lo | ((uint32_t)lr.k_lo << 16);
        const uint32_t khi = (uint32_t)lr.k_hi | ((uint32_t)lr.k_hi << 16);
#pragma unroll
        for (int r = 0; r < R; ++r) h[r] = r < lr.kb ? klo : khi;
    }
    uint32_t dtop = 0;
    base_a = 0;
    base_b = 0;

    // Columns -G..-1 are neutral; chunk 0 goes in before the first read.
    ra[G + k] = kNull;
    rb[G + k] = kNull;
    {
        const int ca = col_code<G>(ta, 0, k), cb = col_code<G>(tb, 0, k);
        ra[k] = ca;
        ra[RING + k] = ca;
        rb[k] = cb;
        rb[RING + k] = cb;
    }
    int na = col_code<G>(ta, 1, k), nb = col_code<G>(tb, 1, k);
    ColumnCursor<G> ka(ta, 2, k), kb(tb, 2, k);
    uint32_t ln = 0;                // this chunk's line values, packed relative (lane k: column chunk*G + k)
    int la_next = -1, lb_next = -1; // the next chunk's, absolute
    if constexpr (LOAD) {
        ln = packed_rel(packed_line_abs(ta, pl.la, 0, k), 0, packed_line_abs(tb, pl.lb, 0, k), 0);
        la_next = packed_line_abs(ta, pl.la, 1, k);
        lb_next = packed_line_abs(tb, pl.lb, 1, k);
    }

    const int base0 = (RING - k) & (RING - 1);
    int bna = ra[base0 + 1], bnb = rb[base0 + 1];
    uint32_t wa[W], wb[W];
    load_prof<G, R>(wa, prof + ra[base0] * CODE_WORDS, k);
    load_prof<G, R>(wb, prof + rb[base0] * CODE_WORDS, k);

    for (int c = 0; c < wchunks; ++c) {
        const int base = (c * G + 2 - k) & (RING - 1);
        // Byte offsets of the bottom cells the last lane makes at step 0 of
        // this chunk (column c*G - (G-1) - lpad of each target).
        const uint32_t va = pl.off_a + (uint32_t)(c * G - (G - 1) - ta.lpad) * 4u;
        const uint32_t vb = pl.off_b + (uint32_t)(c * G - (G - 1) - tb.lpad) * 4u;
        static_for<0, G>([&](auto uc) {
            constexpr int u = decltype(uc)::value;
            // Chunk c+1 replaces chunk c-1, whose last reader was step G-4.
            if constexpr (u == G - 3) {
                const int slot = ((c + 1) * G + k) & (RING - 1);
                ra[slot] = na;
                rb[slot] = nb;
                if (slot < G) {
                    ra[RING + slot] = na;
                    rb[RING + slot] = nb;
                }
            }
            const int bnna = ra[base + u], bnnb = rb[base + u];
            uint32_t wna[W], wnb[W];
            load_prof<G, R>(wna, prof + bna * CODE_WORDS, k);
            load_prof<G, R>(wnb, prof + bnb * CODE_WORDS, k);
            uint32_t hup;
            if constexpr (LOAD)
                hup = (uint32_t)dev::dpp<dev::kRowShr1>(row_bcast<u>((int)ln), (int)h[R - 1]);
            else
                hup = (uint32_t)dev::dpp_zero<dev::kRowShr1>((int)h[R - 1]);
            cells_packed<R>(h, hup, dtop, wa, wb);
            dtop = hup;
            if constexpr (STORE) {
                const int va_abs = (int)(h[R - 1] & 0xFFFFu) + base_a;
                const int vb_abs = (int)(h[R - 1] >> 16) + base_b;
                __builtin_amdgcn_raw_buffer_store_b32((uint32_t)va_abs, pl.rsrc, (int)va, 4 * u, 0);
                __builtin_amdgcn_raw_buffer_store_b32((uint32_t)vb_abs, pl.rsrc, (int)vb, 4 * u, 0);
            }
            bna = bnna;
            bnb = bnnb;
#pragma unroll
            for (int q = 0; q < W; ++q) {
                wa[q] = wna[q];
                wb[q] = wnb[q];
            }
        });
        if constexpr (LOAD) {
            if ((c & 3) == 3) {
                // Raise the bases from the first lane's line values of the
                // next chunk (its column (c+1)*G - lpad), where the target
                // still runs.
                const int fa = row_bcast<0>(la_next), fb = row_bcast<0>(lb_next);
                const int nba = fa < 0 ? base_a : max(base_a, fa - margin);
                const int nbb = fb < 0 ? base_b : max(base_b, fb - margin);
                const uint32_t delta = (uint32_t)(nba - base_a) | ((uint32_t)(nbb - base_b) << 16);
#pragma unroll
                for (int r = 0; r < R; ++r) h[r] = pk_sub(h[r], delta);
                dtop = pk_sub(dtop, delta);
                base_a = nba;
                base_b = nbb;
            }
            ln = packed_rel(la_next, base_a, lb_next, base_b);
            la_next = packed_line_abs(ta, pl.la, c + 2, k);
            lb_next = packed_line_abs(tb, pl.lb, c + 2, k);
        }
        na = ka.code(ta);
        nb = kb.code(tb);
        ka.advance();
        kb.advance();
    }
    // As in run_strip: the row select stays after the loop.
    int sel = lr.sel;
    asm volatile("" : "+v"(sel));
    uint32_t result = h[R - 1];
#pragma unroll
    for (int r = 0; r < R - 1; ++r) result = r == sel ? h[r] : result;
    return result;
}

template <bool STORE, int R = kRowStep>
__device__ __forceinline__ uint32_t run_first_strip_packed(int r0, const dev::lds_u32* prof, const TargetState& ta,
                                                           const TargetState& tb, volatile dev::lds_u8* ra,
                                                           volatile dev::lds_u8* rb, const PackedLines& pl,
                                                           int wchunks, int k, int margin, int& base_a, int& base_b,
                                                           const LaneRows& lr)
{
    if constexpr (R >= 32) {
        return run_strip_packed<16, 32, false, STORE>(prof, ta, tb, ra, rb, pl, wchunks, k, margin, base_a, base_b,
                                                      lr);
    } else {
        if (r0 == R)
            return run_strip_packed<16, R, false, STORE>(prof, ta, tb, ra, rb, pl, wchunks, k, margin, base_a,
                                                         base_b, lr);
        return run_first_strip_packed<STORE, R + kRowStep>(r0, prof, ta, tb, ra, rb, pl, wchunks, k, margin, base_a,
                                                           base_b, lr);
    }
}

__device__ __forceinline__ TargetState target_state(const BatchArgs& a, int2 tg, bool valid)
{
    TargetState t {a.blocks, 0, 0, 0};
    if (valid) {
        const SeqInfo si = a.info[tg.x];
        t.seq = a.blocks + si.x;
        t.len = si.z;
        t.chunks = (t.len + 2 * 16 - 2) / 16;
        t.lpad = t.chunks * 16 - (t.len + 16 - 1);
    }
    return t;
}

// One packed item: a query of the 16 x 32 shape and up to 16 targets, eight
// at a time (two per lane group), strip after strip; a query of one strip
// builds its profile once, a longer one per strip and round.  The wave's line
// regions hold two lines per lane group (A's, then B's).
__device__ __forceinline__ void run_item_packed(const BatchArgs& a, const int4 item, dev::lds_u32* prof,
                                                const dev::lds_u8* tab, dev::lds_u8* rings, int32_t* lines, int lane)
{
    constexpr int G = 16, RMAX = 32, GROUPS = 4, STRIP_MAX = G * RMAX;
    const int g = lane / G;
    const int k = lane % G;
    volatile dev::lds_u8* ra = rings + g * 6 * G;
    volatile dev::lds_u8* rb = ra + 3 * G;
    // A stack (one strip of short queries, as in run_item) or one query.
    const bool stacked = (item.w & kStackedFlag) != 0;
    const SeqInfo qi = a.info[stacked ? 0 : item.x];
    const int Lq = stacked ? a.targets[item.x].y : qi.z;
    const uint16_t* qseq = a.blocks + qi.x;
    const bool transpose = (item.w & 1) != 0;
    const int full = (Lq - 1) / STRIP_MAX;
    const int rem = Lq - full * STRIP_MAX;
    const int R0 = min(RMAX, ((rem + G * kRowStep - 1) / (G * kRowStep)) * kRowStep);
    const int pad_top = G * R0 - rem;
    const int nst = full + 1;
    PackedLines pl;
    {
        pl.la = lines + (int64_t)(2 * g) * a.line_stride + kLineGuard;
        pl.lb = pl.la + a.line_stride;
        const uint64_t wb = (uint64_t)(uintptr_t)lines;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)wb);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(wb >> 32));
        pl.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)hi << 32) | lo), 0,
                                                    2 * kMaxGroups * a.line_stride * 4, kRawBufferFlags);
        pl.off_a = k == G - 1 ? (uint32_t)(2 * g * a.line_stride + kLineGuard) * 4u : kOutOfRange;
        pl.off_b = k == G - 1 ? (uint32_t)((2 * g + 1) * a.line_stride + kLineGuard) * 4u : kOutOfRange;
    }
    const int rounds = (item.z + 2 * GROUPS - 1) / (2 * GROUPS);
    for (int rd = 0; rd < rounds; ++rd) {
        const int ia = rd * 2 * GROUPS + 2 * g, ib = ia + 1;
        const int2 tga = ia < item.z ? a.targets[item.y + ia] : make_int2(0, -1);
        const int2 tgb = ib < item.z ? a.targets[item.y + ib] : make_int2(0, -1);
        const TargetState ta = target_state(a, tga, ia < item.z), tb = target_state(a, tgb, ib < item.z);
        int wchunks = max(ta.chunks, tb.chunks);
        wchunks = max(wchunks, __shfl_xor(wchunks, 16));
        wchunks = max(wchunks, __shfl_xor(wchunks, 32));
        uint32_t res = 0;
        int base_a = 0, base_b = 0;
        for (int st = 0; st < nst; ++st) {
            if (nst > 1 || rd == 0) {
                wave_lds_sync();
                const int row0 = st == 0 ? 0 : G * R0 + (st - 1) * STRIP_MAX;
                if (stacked) build_profile_stacked<G>(prof, tab, a, item.x, pad_top, R0, lane);
                else build_profile<G>(prof, tab, qseq, row0, pad_top, transpose, st == 0 ? R0 : RMAX, lane);
                wave_lds_sync();
            }
            const bool store = st + 1 < nst;
            const int R = st == 0 ? R0 : RMAX;
            LaneRows lr {0, 0, R, -1};
            if (stacked) lr = stacked_rows(a, item.x, pad_top, R0, k).rows;
            if (st == 0)
                res = store ? run_first_strip_packed<true>(R0, prof, ta, tb, ra, rb, pl, wchunks, k, a.margin, base_a,
                                                           base_b, lr)
                            : run_first_strip_packed<false>(R0, prof, ta, tb, ra, rb, pl, wchunks, k, a.margin, base_a,
                                                            base_b, lr);
            else
                res = store ? run_strip_packed<G, RMAX, true, true>(prof, ta, tb, ra, rb, pl, wchunks, k, a.margin,
                                                                    base_a, base_b, lr)
                            : run_strip_packed<G, RMAX, true, false>(prof, ta, tb, ra, rb, pl, wchunks, k, a.margin,
                                                                     base_a, base_b, lr);
            // The next strip reads this strip's bottom rows from the lines.
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
        if (!stacked) {
            if (k == G - 1) {
                if (tga.y >= 0) a.out[tga.y] = (int)(res & 0xFFFFu) + base_a - (Lq + ta.len) * a.penalty;
                if (tgb.y >= 0) a.out[tgb.y] = (int)(res >> 16) + base_b - (Lq + tb.len) * a.penalty;
            }
        } else {
            // Each member's last row is one lane's row lr.sel; its scores go
            // through the stack's output table (-1: no such pair).
            const int mine = stacked_rows(a, item.x, pad_top, R0, k).member;
            if (mine >= 0) {
                const int2 m0 = a.targets[item.x + 1 + 2 * mine];     // {query, K}
                const int2 m1 = a.targets[item.x + 2 + 2 * mine];     // {output row, rows * 2 + transpose}
                const int32_t* table = reinterpret_cast<const int32_t*>(a.targets) + m1.x;
                const int oa = tga.y >= 0 ? table[tga.y] : -1;
                const int ob = tgb.y >= 0 ? table[tgb.y] : -1;
                if (oa >= 0) a.out[oa] = (int)(res & 0xFFFFu) - m0.y - ((m1.y >> 1) + ta.len) * a.penalty;
                if (ob >= 0) a.out[ob] = (int)(res >> 16) - m0.y - ((m1.y >> 1) + tb.len) * a.penalty;
            }
        }
        // The next round's rings are rewritten only after this round's reads.
        wave_lds_sync();
    }
}

// ---- Duo items: two queries per lane group ------------------------------
//
// A duo item pairs two queries of the packed kernel's kind whose target lists
// (almost) agree, and each lane group scores one target against both: the
// low half of every packed cell belongs to the first query, the high half to
// the second.  Both queries share the column, so one profile word per (column
// code, row) holds both queries' sxx -- the cell update is one packed add and
// one packed max3 per two cells, with no byte permute.  That profile takes
// four bytes a row, 51,200 bytes for a 16 x 32 strip: the four waves of a
// workgroup share it and take whole items together, each lane group one
// target of the item's round of 16.  The queries are aligned at their last
// rows (the shorter one gets neutral rows on top); lines, bases and results
// work per half as in the packed kernel.

// Profile layout of a duo strip of R rows per lane: per column code, planes of
// 16 bytes per lane ([plane][lane][4 words]), word r of a lane holding row r
// of both queries; a lane reads its words of a code at its slot plus constants.
template <int R>
constexpr int duo_planes = (R + 3) / 4;
template <int R>
constexpr int duo_code_words = 16 * 4 * duo_planes<R>;
static_assert(kCodes * duo_code_words<32> == kWaves * kProfWords, "a duo profile fills the waves' profile regions");

// Words [B, E) of a lane's profile words of one code (B a multiple of 4, E even).
template <int R, int B = 0, int E = R>
__device__ __forceinline__ void load_prof_duo(uint32_t (&w)[R], const dev::lds_u32* code_base, int k)
{
    static_assert(R % 2 == 0 && B % 4 == 0 && E % 2 == 0 && E <= R, "strip heights are even");
    const dev::lds_u32* slot = code_base + k * 4;
#pragma unroll
    for (int j = B / 4; j < E / 4; ++j) {
        const dev::u32x4 v = *(const dev::lds_u32x4*)(slot + j * 64);
        w[4 * j + 0] = v.x;
        w[4 * j + 1] = v.y;
        w[4 * j + 2] = v.z;
        w[4 * j + 3] = v.w;
    }
    if constexpr (E % 4 == 2) {
        const dev::u32x2 v = *(const dev::lds_u32x2*)(slot + (E / 4) * 64);
        w[E - 2] = v.x;
        w[E - 1] = v.y;
    }
}

// Builds a duo strip's profile with the whole workgroup: rows [row0, row0 +
// 16R) of the padded row space, where query q's row i sits at i + pad_q (rows
// of a query before its first are neutral; an absent second query, q2 ==
// nullptr, is neutral throughout).  Per (lane, row) slot, four-code runs of
// both queries' table rows become four words by byte permutes (word of code c:
// bytes q1's sxx, 0, q2's sxx, 0).
__device__ __forceinline__ void build_profile_duo(dev::lds_u32* prof, const dev::lds_u8* tab, const uint16_t* q1,
                                                  const uint16_t* q2, int row0, int pad1, int pad2, bool tr1,
                                                  bool tr2, int R, int tid)
{
    constexpr uint32_t kSel[4] = {0x0C040C00u, 0x0C050C01u, 0x0C060C02u, 0x0C070C03u};
    const int code_words_r = 16 * 4 * ((R + 3) / 4);
    for (int slot = tid; slot < 16 * R; slot += kWaves * 64) {
        const int k = slot % 16;
        const int r = slot / 16;
        const int p = row0 + k * R + r;
        const int i1 = p - pad1, i2 = p - pad2;
        const int c1 = i1 >= 0 ? min(dev::code_at(q1, i1), kNull) : kNull;
        const int c2 = (q2 != nullptr && i2 >= 0) ? min(dev::code_at(q2, i2), kNull) : kNull;
        const dev::lds_u8* t1 = tab + (tr1 ? kTabCopy : 0) + c1 * kTabStride;
        const dev::lds_u8* t2 = tab + (tr2 ? kTabCopy : 0) + c2 * kTabStride;
        dev::lds_u32* dst = prof + (r >> 2) * 64 + k * 4 + (r & 3);
#pragma unroll
        for (int j = 0; j < (kCodes + 3) / 4; ++j) {
            const uint32_t a = *(const dev::lds_u32*)(t1 + 4 * j);
            const uint32_t b = *(const dev::lds_u32*)(t2 + 4 * j);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (4 * j + i < kCodes) dst[(4 * j + i) * code_words_r] = __builtin_amdgcn_perm(b, a, kSel[i]);
        }
    }
}

// Rows [B, E) of one step of a lane's R rows, two queries at once (h packed;
// u: the cell above row B, new; t: row B's diagonal plus its sxx; w: the
// column code's profile words).  Row r+1's add comes after row r's max3.
template <int R, int B, int E>
__device__ __forceinline__ void cells_duo(uint32_t (&h)[R], uint32_t& u, uint32_t& t, const uint32_t (&w)[R])
{
    // Rows in runs of four: the run's adds first (row B's is `t`), then its
    // chain of max3s.
#pragma unroll
    for (int r0 = B; r0 < E; r0 += 4) {
        constexpr int N = 4;
        uint32_t s[N];
        s[0] = t;
#pragma unroll
        for (int i = 1; i < N; ++i)
            if (r0 + i < E) s[i] = pk_add(h[r0 + i - 1], w[r0 + i]);
        if (r0 + N < R && r0 + N <= E) t = pk_add(h[r0 + N - 1], w[r0 + N]);
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (r0 + i < E) {
                const uint32_t nh = pk_max3(s[i], u, h[r0 + i]);
                h[r0 + i] = nh;
                u = nh;
            }
    }
}

// One duo strip of R rows per lane over the target `t`: run_strip_packed with
// one target, one code ring (3G bytes) and one profile read per step; the
// lines `pl.la` and `pl.lb` are the first and the second query's.
template <int R, bool LOAD, bool STORE>
__device__ __forceinline__ uint32_t run_strip_duo(const dev::lds_u32* prof, const TargetState& t,
                                                  volatile dev::lds_u8* ring, const PackedLines& pl, int wchunks,
                                                  int k, int margin, int& base_a, int& base_b)
{
    constexpr int G = 16;
    constexpr int CODE_WORDS = duo_code_words<R>;
    constexpr int RING = 2 * G;
    uint32_t h[R];
#pragma unroll
    for (int r = 0; r < R; ++r) h[r] = 0;
    uint32_t dtop = 0;
    base_a = 0;
    base_b = 0;

    ring[G + k] = kNull;
    {
        const int c = col_code<G>(t, 0, k);
        ring[k] = c;
        ring[RING + k] = c;
    }
    int nc = col_code<G>(t, 1, k);
    ColumnCursor<G> cur(t, 2, k);
    uint32_t ln = 0;
    int la_next = -1, lb_next = -1;
    if constexpr (LOAD) {
        ln = packed_rel(packed_line_abs(t, pl.la, 0, k), 0, packed_line_abs(t, pl.lb, 0, k), 0);
        la_next = packed_line_abs(t, pl.la, 1, k);
        lb_next = packed_line_abs(t, pl.lb, 1, k);
    }

    const int base0 = (RING - k) & (RING - 1);
    int bn = ring[base0 + 1];
    uint32_t w[R];
    load_prof_duo<R>(w, prof + ring[base0] * CODE_WORDS, k);
    // Tall strips read the next column's words in two halves, each after the
    // current column's half is done (16 registers of read-ahead, not R).
    constexpr int H = R >= 20 ? (R / 2) / 4 * 4 : 0;

    for (int c = 0; c < wchunks; ++c) {
        const int base = (c * G + 2 - k) & (RING - 1);
        const uint32_t va = pl.off_a + (uint32_t)(c * G - (G - 1) - t.lpad) * 4u;
        const uint32_t vb = pl.off_b + (uint32_t)(c * G - (G - 1) - t.lpad) * 4u;
        static_for<0, G>([&](auto uc) {
            constexpr int u = decltype(uc)::value;
            if constexpr (u == G - 3) {
                const int slot = ((c + 1) * G + k) & (RING - 1);
                ring[slot] = nc;
                if (slot < G) ring[RING + slot] = nc;
            }
            const int bnn = ring[base + u];
            uint32_t hup;
            if constexpr (LOAD)
                hup = (uint32_t)dev::dpp<dev::kRowShr1>(row_bcast<u>((int)ln), (int)h[R - 1]);
            else
                hup = (uint32_t)dev::dpp_zero<dev::kRowShr1>((int)h[R - 1]);
            uint32_t up = hup;
            uint32_t tt = pk_add(dtop, w[0]);
            if constexpr (H > 0) {
                cells_duo<R, 0, H>(h, up, tt, w);
                load_prof_duo<R, 0, H>(w, prof + bn * CODE_WORDS, k);
                cells_duo<R, H, R>(h, up, tt, w);
                load_prof_duo<R, H, R>(w, prof + bn * CODE_WORDS, k);
            } else {
                uint32_t wn[R];
                load_prof_duo<R>(wn, prof + bn * CODE_WORDS, k);
                cells_duo<R, 0, R>(h, up, tt, w);
#pragma unroll
                for (int q = 0; q < R; ++q) w[q] = wn[q];
            }
            dtop = hup;
            if constexpr (STORE) {
                const int va_abs = (int)(h[R - 1] & 0xFFFFu) + base_a;
                const int vb_abs = (int)(h[R - 1] >> 16) + base_b;
                __builtin_amdgcn_raw_buffer_store_b32((uint32_t)va_abs, pl.rsrc, (int)va, 4 * u, 0);
                __builtin_amdgcn_raw_buffer_store_b32((uint32_t)vb_abs, pl.rsrc, (int)vb, 4 * u, 0);
            }
            bn = bnn;
        });
        if constexpr (LOAD) {
            if ((c & 3) == 3) {
                const int fa = row_bcast<0>(la_next), fb = row_bcast<0>(lb_next);
                const int nba = fa < 0 ? base_a : max(base_a, fa - margin);
                const int nbb = fb < 0 ? base_b : max(base_b, fb - margin);
                const uint32_t delta = (uint32_t)(nba - base_a) | ((uint32_t)(nbb - base_b) << 16);
#pragma unroll
                for (int r = 0; r < R; ++r) h[r] = pk_sub(h[r], delta);
                dtop = pk_sub(dtop, delta);
                base_a = nba;
                base_b = nbb;
            }
            ln = packed_rel(la_next, base_a, lb_next, base_b);
            la_next = packed_line_abs(t, pl.la, c + 2, k);
            lb_next = packed_line_abs(t, pl.lb, c + 2, k);
        }
        nc = cur.code(t);
        cur.advance();
    }
    return h[R - 1];
}

template <bool STORE, int R = kRowStep>
__device__ __forceinline__ uint32_t run_first_strip_duo(int r0, const dev::lds_u32* prof, const TargetState& t,
                                                        volatile dev::lds_u8* ring, const PackedLines& pl, int wchunks,
                                                        int k, int margin, int& base_a, int& base_b)
{
    if constexpr (R >= 32) {
        return run_strip_duo<32, false, STORE>(prof, t, ring, pl, wchunks, k, margin, base_a, base_b);
    } else {
        if (r0 == R) return run_strip_duo<R, false, STORE>(prof, t, ring, pl, wchunks, k, margin, base_a, base_b);
        return run_first_strip_duo<STORE, R + kRowStep>(r0, prof, t, ring, pl, wchunks, k, margin, base_a, base_b);
    }
}

// One duo item with the whole workgroup: rounds of 16 targets (one per lane
// group), strip after strip; the shared profile is built once for a pair of
// one-strip queries, else per strip and round, between workgroup barriers.
__device__ __forceinline__ void run_item_duo(const BatchArgs& a, const int4 item, dev::lds_u32* prof,
                                             const dev::lds_u8* tab, dev::lds_u8* rings, int32_t* lines, int wave,
                                             int lane)
{
    constexpr int G = 16, RMAX = 32, GROUPS = 4, STRIP_MAX = G * RMAX, SLOTS = kWaves * GROUPS;
    const int g = lane / G;
    const int k = lane % G;
    volatile dev::lds_u8* ring = rings + g * 3 * G;
    const int2 rec = a.targets[item.y];
    const SeqInfo q1i = a.info[item.x];
    const int lq1 = q1i.z;
    const uint16_t* q1 = a.blocks + q1i.x;
    int lq2 = 0;
    const uint16_t* q2 = nullptr;
    if (rec.x >= 0) {
        const SeqInfo q2i = a.info[rec.x];
        lq2 = q2i.z;
        q2 = a.blocks + q2i.x;
    }
    const int Lq = max(lq1, lq2);
    const int full = (Lq - 1) / STRIP_MAX;
    const int rem = Lq - full * STRIP_MAX;
    const int R0 = min(RMAX, ((rem + G * kRowStep - 1) / (G * kRowStep)) * kRowStep);
    const int pad_top = G * R0 - rem;
    const int pad1 = pad_top + Lq - lq1, pad2 = pad_top + Lq - lq2;
    const int nst = full + 1;
    PackedLines pl;
    {
        pl.la = lines + (int64_t)(2 * g) * a.line_stride + kLineGuard;
        pl.lb = pl.la + a.line_stride;
        const uint64_t wb = (uint64_t)(uintptr_t)lines;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)wb);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(wb >> 32));
        pl.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)hi << 32) | lo), 0,
                                                    2 * kMaxGroups * a.line_stride * 4, kRawBufferFlags);
        pl.off_a = k == G - 1 ? (uint32_t)(2 * g * a.line_stride + kLineGuard) * 4u : kOutOfRange;
        pl.off_b = k == G - 1 ? (uint32_t)((2 * g + 1) * a.line_stride + kLineGuard) * 4u : kOutOfRange;
    }
    const int rounds = (item.z + SLOTS - 1) / SLOTS;
    for (int rd = 0; rd < rounds; ++rd) {
        const int ti = rd * SLOTS + wave * GROUPS + g;
        const bool valid = ti < item.z;
        const int2 tg = valid ? a.targets[item.y + 1 + 2 * ti] : make_int2(0, -1);
        const int out2 = valid ? a.targets[item.y + 2 + 2 * ti].y : -1;
        const TargetState t = target_state(a, tg, valid);
        int wchunks = t.chunks;
        wchunks = max(wchunks, __shfl_xor(wchunks, 16));
        wchunks = max(wchunks, __shfl_xor(wchunks, 32));
        uint32_t res = 0;
        int base_a = 0, base_b = 0;
        for (int st = 0; st < nst; ++st) {
            if (nst > 1 || rd == 0) {
                __syncthreads();
                const int row0 = st == 0 ? 0 : G * R0 + (st - 1) * STRIP_MAX;
                build_profile_duo(prof, tab, q1, q2, row0, pad1, pad2, (item.w & 1) != 0, (rec.y & 1) != 0,
                                  st == 0 ? R0 : RMAX, wave * 64 + lane);
                __syncthreads();
            }
            const bool store = st + 1 < nst;
            if (st == 0)
                res = store ? run_first_strip_duo<true>(R0, prof, t, ring, pl, wchunks, k, a.margin, base_a, base_b)
                            : run_first_strip_duo<false>(R0, prof, t, ring, pl, wchunks, k, a.margin, base_a, base_b);
            else
                res = store ? run_strip_duo<RMAX, true, true>(prof, t, ring, pl, wchunks, k, a.margin, base_a, base_b)
                            : run_strip_duo<RMAX, true, false>(prof, t, ring, pl, wchunks, k, a.margin, base_a, base_b);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
        if (k == G - 1) {
            if (tg.y >= 0) a.out[tg.y] = (int)(res & 0xFFFFu) + base_a - (lq1 + t.len) * a.penalty;
            if (out2 >= 0) a.out[out2] = (int)(res >> 16) + base_b - (lq2 + t.len) * a.penalty;
        }
        wave_lds_sync();
    }
}

template <int CFG>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(3))) void nw_batched_kernel(BatchArgs a)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];   // 16-byte aligned: ds_read_b128 of profiles
    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    dev::lds_u32* lds_base = (dev::lds_u32*)lds;   // address-space cast: 32-bit LDS offsets
    dev::lds_u32* prof = lds_base + wave * kProfWords;
    dev::lds_u8* tab = (dev::lds_u8*)(lds_base + kWaves * kProfWords);
    dev::lds_u8* rings = tab + kTableBytes + wave * (packed_queue(CFG) ? kPackedRingBytes : kRingBytes);
    for (int i = threadIdx.x; i < 2 * kTabCopy; i += kWaves * 64) {
        const int flip = i >= kTabCopy;
        const int j = i - flip * kTabCopy;
        const int r = j / kTabStride, c = j - r * kTabStride;
        tab[i] = (r < kNull && c < kNull) ? a.sxx.v[flip ? c * 25 + r : r * 25 + c] : 0;
    }
    __syncthreads();

    // A wave's line regions: one line per lane group, two in the packed kernels.
    int32_t* lines = a.lines + (int64_t)(blockIdx.x * kWaves + wave) * (packed_queue(CFG) ? 2 : 1) * kMaxGroups *
                               a.line_stride;

    if constexpr (CFG == kDuoQueue) {
        // The workgroup takes items together (they share the profile).
        volatile dev::lds_u32* slot = (volatile dev::lds_u32*)(tab + kTableBytes + kWaves * kPackedRingBytes);
        while (true) {
            if (threadIdx.x == 0) *slot = (uint32_t)atomicAdd(a.queue, 1);
            __syncthreads();
            const int it = (int)*slot;
            __syncthreads();
            if (it >= a.nitems) break;
            run_item_duo(a, a.items[it], lds_base, tab, rings, lines, wave, lane);
        }
        // Then the packed and the 16 x 32 items, one per wave, each wave in
        // its own profile region and ring bytes: the duo items of other
        // workgroups may still run, and these fill their tail.
        while (a.nitems3 > 0) {
            int it = 0;
            if (lane == 0) it = atomicAdd(a.queue3, 1);
            it = __shfl(it, 0);
            if (it >= a.nitems3) break;
            run_item_packed(a, a.items3[it], prof, tab, rings, lines, lane);
        }
        while (a.nitems2 > 0) {
            int it = 0;
            if (lane == 0) it = atomicAdd(a.queue2, 1);
            it = __shfl(it, 0);
            if (it >= a.nitems2) break;
            run_item<kTileConfig[0].lanes, kTileConfig[0].rows>(a, a.items2[it], prof, tab, rings, lines, lane);
        }
    } else {
        while (true) {
            int it = 0;
            if (lane == 0) it = atomicAdd(a.queue, 1);
            it = __shfl(it, 0);
            if (it >= a.nitems) break;
            if constexpr (CFG == kPackedQueue) run_item_packed(a, a.items[it], prof, tab, rings, lines, lane);
            else run_item<kTileConfig[CFG].lanes, kTileConfig[CFG].rows>(a, a.items[it], prof, tab, rings, lines, lane);
        }
    }
}

// Host side of the kernel of tile configuration CFG.
template <int CFG>
int kernel_occupancy(int* blocks_per_cu)
{
    const void* fn = reinterpret_cast<const void*>(&nw_batched_kernel<CFG>);
    const size_t lds = lds_bytes(CFG);
    MSA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int n = 0;
    MSA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kWaves * 64, lds));
    *blocks_per_cu = n;
    return MSA_OK;
}

template <int CFG>
int kernel_launch(int grid, const BatchArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL(nw_batched_kernel<CFG>, dim3(grid), dim3(kWaves * 64), lds_bytes(CFG), stream, a);
    return hip_status(hipGetLastError(), "nw_batched_kernel launch");
}

// One pair of entry points per tile configuration, each defined in the
// source file that instantiates its kernel.
int occupancy_16x32(int* blocks_per_cu);
int launch_16x32(int grid, const BatchArgs& a, hipStream_t stream);
int occupancy_16x16(int* blocks_per_cu);
int launch_16x16(int grid, const BatchArgs& a, hipStream_t stream);
int occupancy_32x16(int* blocks_per_cu);
int launch_32x16(int grid, const BatchArgs& a, hipStream_t stream);
int occupancy_packed(int* blocks_per_cu);
int launch_packed(int grid, const BatchArgs& a, hipStream_t stream);
int occupancy_duo(int* blocks_per_cu);
int launch_duo(int grid, const BatchArgs& a, hipStream_t stream);

}  // namespace batched
}  // namespace msa
