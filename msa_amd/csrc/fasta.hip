// FASTA loading on the GPU: the reference's io::parser::fasta
// (src/io/loader/parser/fasta.cpp:26-70), encoder::encode (src/encoder.cpp:
// 50-98) and the pairwise::database packing (src/pairwise/database.cu:24-72),
// over text resident in device memory.
//
// Semantics, byte for byte those of the reference:
//   - lines end at '\n' ('\r' stays in the line); a final piece without '\n'
//     is a line, an empty final piece is not;
//   - a line whose first byte is '>' opens a record, its rest is the
//     description; the content lines that follow belong to it until the next
//     '>' line or the first empty line; lines after an empty line, and before
//     the first '>' line, belong to no record;
//   - every byte of a record's content lines is one residue: letters by the
//     encoder's table (case-folded), '-' the gap code, anything else the end
//     code 0x18; three codes per 16-bit block at bits 1, 6, 11, the last block
//     of a record padded with 0x18 and flagged in bit 0.
//
// Pipeline over N bytes in tiles of 16 KiB (256 threads x 64 bytes):
//   1. count_kernel       '\n' per tile (SWAR over 32-bit words)
//   2. scan               index of the first line each tile starts
//   3. index_kernel       the line table: start << 2 | class of its first byte
//   4. line scans         (reduce, partials, apply) of the record state over
//                         lines: headers so far, record open or not, residues
//                         so far -> each line's record and residue rank, each
//                         record's first residue and description range
//   5. records_kernel     blocks per record; scan -> block offsets
//   6. emit_desc_kernel   per line: block and slot of its first residue,
//                         owned / last-line-of-record flags
//   7. emit_kernel        codes into blocks staged in LDS per tile, written
//                         out coalesced.  A block belongs to the tile holding
//                         its first residue; when its residues run past the
//                         tile, that tile reads the few bytes ahead.
#include "msa_internal.hpp"

#include <algorithm>
#include <cstring>
#include <new>

namespace msa {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 64;                                  // text bytes per thread
constexpr int64_t kTile = (int64_t)kThreads * kChunk;       // text bytes per workgroup
constexpr int kLinesPerThread = 16;
constexpr int64_t kLineTile = (int64_t)kThreads * kLinesPerThread;
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = (int64_t)kThreads * kScanItems;
constexpr int kPartialThreads = 1024;

enum : uint32_t { kContent = 0, kHeader = 1, kEmpty = 2 };

// Residue code of a byte (src/encoder.cpp:50-61).
__host__ __device__ inline uint32_t encode_byte(uint32_t c)
{
    constexpr uint8_t letter[26] = {0x00, 0x14, 0x01, 0x06, 0x08, 0x0E, 0x03, 0x09, 0x0A, 0x15, 0x0C, 0x0B, 0x0D,
                                    0x05, 0x17, 0x0F, 0x07, 0x04, 0x10, 0x02, 0x17, 0x13, 0x11, 0x17, 0x12, 0x16};
    const uint32_t upper = (c >= 'a' && c <= 'z') ? c - 'a' + 'A' : c;
    if (upper >= 'A' && upper <= 'Z') return letter[upper - 'A'];
    return c == '-' ? (uint32_t)kGapCode : (uint32_t)kEndCode;
}

__device__ inline uint32_t line_class(uint32_t c) { return c == '>' ? kHeader : (c == '\n' ? kEmpty : kContent); }

// High bit of every byte of `w` that equals `b` (exact, no carries across bytes).
__device__ inline uint32_t bytes_equal(uint32_t w, uint32_t pattern)
{
    const uint32_t t = w ^ pattern;
    return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t | 0x7F7F7F7Fu);
}

// Those high bits packed into 4 bits, byte 0 lowest.
__device__ inline uint32_t pack4(uint32_t m)
{
    m >>= 7;
    return (m | (m >> 7) | (m >> 14) | (m >> 21)) & 0xFu;
}

// A thread's 64 bytes: 16 words (bytes past `valid` are 0), the '\n' mask
// and the number of valid bytes.
struct Chunk {
    uint32_t w[16];
    uint64_t nl;
    int valid;
};

__device__ inline void load_words(const uint8_t* text, int64_t n, int64_t c0, int vec, Chunk& ck)
{
    const int64_t left = n - c0;
    ck.valid = left <= 0 ? 0 : (left >= kChunk ? kChunk : (int)left);
    if (vec && ck.valid == kChunk) {
        const uint4* p = reinterpret_cast<const uint4*>(text + c0);
        uint4 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = p[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ck.w[4 * k] = v[k].x, ck.w[4 * k + 1] = v[k].y, ck.w[4 * k + 2] = v[k].z, ck.w[4 * k + 3] = v[k].w;
        }
    } else {
        // Static indices only, so the words stay in registers.
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (4 * k + b < ck.valid) v |= (uint32_t)text[c0 + 4 * k + b] << (8 * b);
            ck.w[k] = v;
        }
    }
}

__device__ inline void load_chunk(const uint8_t* text, int64_t n, int64_t c0, int vec, Chunk& ck)
{
    load_words(text, n, c0, vec, ck);
    ck.nl = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) ck.nl |= (uint64_t)pack4(bytes_equal(ck.w[k], 0x0A0A0A0Au)) << (4 * k);
    ck.nl &= ck.valid == 64 ? ~0ull : ((1ull << ck.valid) - 1);
}

// ---- block-wide helpers ------------------------------------------------------

// Exclusive scan of one int64 per thread; *total gets the sum.
__device__ int64_t block_exclusive_sum(int64_t v, int64_t* lds, int64_t* total)
{
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1) {
        const int64_t add = tid >= d ? lds[tid - d] : 0;
        __syncthreads();
        lds[tid] += add;
        __syncthreads();
    }
    const int64_t incl = lds[tid];
    if (total) *total = lds[kThreads - 1];
    __syncthreads();
    return incl - v;
}

// ---- line-state scan -----------------------------------------------------------

// A run of lines, for the record-state scan.  `marker` is 0 when the run has
// no header or empty line, else 1 (the last one is a header) or 2 (an empty
// line).  Residues of content lines before the run's first marker are owned
// only if the record open before the run stays open (`undecided`); those after
// it are settled inside the run (`decided`).
struct LineAgg {
    int64_t headers;
    int64_t undecided;
    int64_t decided;
    int32_t marker;
};

struct LineState {
    int64_t headers;     // records opened so far; the current one is headers - 1
    int64_t residues;    // owned residues so far
    int32_t open;        // the current record still takes content lines
};

__host__ __device__ inline LineAgg combine(const LineAgg& a, const LineAgg& b)
{
    LineAgg r;
    r.headers = a.headers + b.headers;
    r.marker = b.marker ? b.marker : a.marker;
    if (!a.marker) {
        r.undecided = a.undecided + b.undecided;
        r.decided = b.decided;
    } else {
        r.undecided = a.undecided;
        r.decided = a.decided + (a.marker == 1 ? b.undecided : 0) + b.decided;
    }
    return r;
}

__host__ __device__ inline LineState advance(const LineState& s, const LineAgg& a)
{
    LineState r;
    r.headers = s.headers + a.headers;
    r.residues = s.residues + (s.open ? a.undecided : 0) + a.decided;
    r.open = a.marker ? (a.marker == 1) : s.open;
    return r;
}

struct LineInfo {
    int64_t start, end;
    uint32_t cls;
};

__device__ inline LineAgg line_element(const LineInfo& li)
{
    if (li.cls == kHeader) return {1, 0, 0, 1};
    if (li.cls == kEmpty) return {0, 0, 0, 2};
    return {0, li.end - li.start, 0, 0};
}

// Ordered inclusive scan of one aggregate per thread (Hillis-Steele in LDS).
__device__ LineAgg block_scan_agg(LineAgg v, LineAgg* lds)
{
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1) {
        LineAgg mine = lds[tid];
        if (tid >= d) mine = combine(lds[tid - d], mine);
        __syncthreads();
        lds[tid] = mine;
        __syncthreads();
    }
    return lds[tid];
}

__device__ inline int64_t last_line_end(const uint8_t* text, int64_t n)
{
    return n - (text[n - 1] == '\n' ? 1 : 0);
}

// ---- kernels -------------------------------------------------------------------

__global__ __launch_bounds__(kThreads) void count_kernel(const uint8_t* text, int64_t n, int vec, int64_t* counts)
{
    __shared__ int64_t lds[kThreads];
    const int64_t c0 = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kChunk;
    int cnt = 0;
    const int64_t left = n - c0;
    if (vec && left >= kChunk) {
        const uint4* p = reinterpret_cast<const uint4*>(text + c0);
        uint4 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = p[k];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            cnt += __builtin_popcount(bytes_equal(v[k].x, 0x0A0A0A0Au)) +
                   __builtin_popcount(bytes_equal(v[k].y, 0x0A0A0A0Au)) +
                   __builtin_popcount(bytes_equal(v[k].z, 0x0A0A0A0Au)) +
                   __builtin_popcount(bytes_equal(v[k].w, 0x0A0A0A0Au));
    } else {
        for (int64_t i = 0; i < kChunk && i < left; ++i) cnt += text[c0 + i] == '\n';
    }
    int64_t total = 0;
    block_exclusive_sum(cnt, lds, &total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Line table: line 0 starts at 0; every '\n' at p < n-1 starts line
// (first_line[tile] + rank) at p + 1.
__global__ __launch_bounds__(kThreads) void index_kernel(const uint8_t* text, int64_t n, int vec,
                                                         const int64_t* first_line, uint64_t* line, int64_t m,
                                                         uint16_t* chunk_line, uint64_t* chunk_nl)
{
    __shared__ int64_t lds[kThreads];
    const int64_t c0 = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kChunk;
    Chunk ck;
    load_chunk(text, n, c0, vec, ck);
    const int64_t before = block_exclusive_sum(__builtin_popcountll(ck.nl), lds, nullptr);
    // '\n' before this chunk within its tile (at most kTile - kChunk).
    chunk_line[(int64_t)blockIdx.x * kThreads + threadIdx.x] = (uint16_t)before;
    chunk_nl[(int64_t)blockIdx.x * kThreads + threadIdx.x] = ck.nl;
    int64_t idx = first_line[blockIdx.x] + before;
    if (blockIdx.x == 0 && threadIdx.x == 0) line[0] = line_class(text[0]);
    uint64_t bits = ck.nl;
    while (bits) {
        const int j = __builtin_ctzll(bits);
        bits &= bits - 1;
        const int64_t p = c0 + j;
        if (p + 1 < n && idx < m) line[idx] = ((uint64_t)(p + 1) << 2) | line_class(text[p + 1]);
        ++idx;
    }
}

// A line tile in LDS: entries a = 0..kLineTile of the tile (the last one is
// the next tile's first line, for the end of the tile's last line), stored at
// a + a / 16 so that a thread's 16 consecutive entries fall in distinct banks.
constexpr int kLinePad = (int)(kLineTile + kLineTile / 16 + 2);

__device__ inline int line_slot(int a) { return a + (a >> 4); }

__device__ void load_line_tile(const uint64_t* line, int64_t m, int64_t base, uint64_t* sl)
{
    for (int a = threadIdx.x; a <= (int)kLineTile; a += kThreads) {
        const int64_t l = base + a;
        sl[line_slot(a)] = l < m ? line[l] : 0;
    }
    __syncthreads();
}

// Line a of the staged tile (global index base + a).
__device__ inline LineInfo staged_line(const uint64_t* sl, int a, int64_t l, int64_t m, int64_t last_end)
{
    LineInfo li;
    const uint64_t e = sl[line_slot(a)];
    li.start = (int64_t)(e >> 2);
    li.cls = (uint32_t)(e & 3);
    li.end = l + 1 < m ? (int64_t)(sl[line_slot(a + 1)] >> 2) - 1 : last_end;
    return li;
}

__global__ __launch_bounds__(kThreads) void line_reduce_kernel(const uint8_t* text, int64_t n, const uint64_t* line,
                                                               int64_t m, LineAgg* aggs)
{
    __shared__ uint64_t sl[kLinePad];
    __shared__ LineAgg lds[kThreads];
    const int64_t last_end = last_line_end(text, n);
    const int64_t base = (int64_t)blockIdx.x * kLineTile;
    load_line_tile(line, m, base, sl);
    const int a0 = threadIdx.x * kLinesPerThread;
    LineAgg acc{0, 0, 0, 0};
    for (int k = 0; k < kLinesPerThread; ++k) {
        const int64_t l = base + a0 + k;
        if (l >= m) break;
        acc = combine(acc, line_element(staged_line(sl, a0 + k, l, m, last_end)));
    }
    const LineAgg incl = block_scan_agg(acc, lds);
    if (threadIdx.x == kThreads - 1) aggs[blockIdx.x] = incl;
}

// One workgroup: the record state entering every line tile, and the totals.
__global__ __launch_bounds__(kPartialThreads) void line_partials_kernel(const LineAgg* aggs, int64_t parts,
                                                                        LineState* states, LineState* totals)
{
    __shared__ LineAgg lds[kPartialThreads];
    const int tid = threadIdx.x;
    const int64_t span = (parts + kPartialThreads - 1) / kPartialThreads;
    const int64_t b0 = tid * span;
    LineAgg acc{0, 0, 0, 0};
    for (int64_t b = b0; b < b0 + span && b < parts; ++b) acc = combine(acc, aggs[b]);
    lds[tid] = acc;
    __syncthreads();
    for (int d = 1; d < kPartialThreads; d <<= 1) {
        LineAgg mine = lds[tid];
        if (tid >= d) mine = combine(lds[tid - d], mine);
        __syncthreads();
        lds[tid] = mine;
        __syncthreads();
    }
    LineState s{0, 0, 0};
    if (tid > 0) s = advance(s, lds[tid - 1]);
    for (int64_t b = b0; b < b0 + span && b < parts; ++b) {
        states[b] = s;
        s = advance(s, aggs[b]);
    }
    if (tid == kPartialThreads - 1) *totals = advance(LineState{0, 0, 0}, lds[tid]);
}

// Per line: its record (-1 when it holds no residue of one) and residue rank;
// per record: its first residue rank and description range.  Lines are read
// and the per-line results written through LDS, so global accesses coalesce.
__global__ __launch_bounds__(kThreads) void line_apply_kernel(const uint8_t* text, int64_t n, const uint64_t* line,
                                                              int64_t m, const LineState* states, int32_t* rec,
                                                              uint64_t* rank, int64_t* rec_first, int64_t* desc_start,
                                                              int64_t* desc_end, int64_t records)
{
    __shared__ uint64_t sl[kLinePad];
    __shared__ int32_t sr[kLinePad];
    __shared__ LineAgg lds[kThreads];
    const int64_t last_end = last_line_end(text, n);
    const int64_t base = (int64_t)blockIdx.x * kLineTile;
    load_line_tile(line, m, base, sl);
    const int a0 = threadIdx.x * kLinesPerThread;
    LineAgg acc{0, 0, 0, 0};
    for (int k = 0; k < kLinesPerThread; ++k) {
        const int64_t l = base + a0 + k;
        if (l >= m) break;
        acc = combine(acc, line_element(staged_line(sl, a0 + k, l, m, last_end)));
    }
    block_scan_agg(acc, lds);
    LineState s = states[blockIdx.x];
    if (threadIdx.x > 0) s = advance(s, lds[threadIdx.x - 1]);
    uint64_t ranks[kLinesPerThread];
    int32_t recs[kLinesPerThread];
#pragma unroll
    for (int k = 0; k < kLinesPerThread; ++k) {
        const int64_t l = base + a0 + k;
        ranks[k] = 0;
        recs[k] = -1;
        if (l < m) {
            const LineInfo li = staged_line(sl, a0 + k, l, m, last_end);
            if (li.cls == kHeader && s.headers < records) {
                rec_first[s.headers] = s.residues;
                desc_start[s.headers] = li.start + 1;
                desc_end[s.headers] = li.end;
            } else if (li.cls == kContent && s.open) {
                recs[k] = (int32_t)(s.headers - 1);
            }
            ranks[k] = (uint64_t)s.residues;
            s = advance(s, line_element(li));
        }
    }
    __syncthreads();            // every thread is done reading sl
#pragma unroll
    for (int k = 0; k < kLinesPerThread; ++k) {
        sl[line_slot(a0 + k)] = ranks[k];
        sr[line_slot(a0 + k)] = recs[k];
    }
    __syncthreads();
    for (int a = threadIdx.x; a < (int)kLineTile; a += kThreads) {
        const int64_t l = base + a;
        if (l < m) {
            rank[l] = sl[line_slot(a)];
            rec[l] = sr[line_slot(a)];
        }
    }
}

// Blocks of every record: ceil(residues / 3).
__global__ __launch_bounds__(kThreads) void records_kernel(const int64_t* rec_first, int64_t records,
                                                           const LineState* totals, int64_t* blocks, int32_t* size)
{
    const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= records) return;
    const int64_t next = r + 1 < records ? rec_first[r + 1] : totals->residues;
    const int64_t b = (next - rec_first[r] + 2) / 3;
    blocks[r] = b;
    size[r] = (int32_t)b;
}

// Per line: (block of its first residue) << 4 | slot << 2 | last << 1 | owned.
__global__ __launch_bounds__(kThreads) void emit_desc_kernel(int64_t m, const int32_t* rec, uint64_t* rank,
                                                             const int64_t* rec_first, const int64_t* offset,
                                                             int64_t records)
{
    const int64_t l = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (l >= m) return;
    const int32_t r = rec[l];
    uint64_t e = 0;
    if (r >= 0 && r < records) {
        const int64_t k = (int64_t)rank[l] - rec_first[r];
        const int64_t q = offset[r] + k / 3;
        const bool last = l + 1 >= m || rec[l + 1] != r;
        e = ((uint64_t)q << 4) | ((uint64_t)(k % 3) << 2) | ((uint64_t)last << 1) | 1u;
    }
    rank[l] = e;
}

// Block and slot of residue position `pos` of a line that starts at `start`
// with descriptor `e`.
__device__ inline void locate(uint64_t e, int64_t start, int64_t pos, int64_t& blk, int& slot)
{
    const int64_t t = (int64_t)((e >> 2) & 3) + (pos - start);
    if (t < (int64_t(1) << 30)) {          // lines shorter than 1 GiB: 32-bit division
        const uint32_t t32 = (uint32_t)t, q = t32 / 3u;
        blk = (int64_t)(e >> 4) + q;
        slot = (int)(t32 - 3u * q);
    } else {
        blk = (int64_t)(e >> 4) + t / 3;
        slot = (int)(t % 3);
    }
}

// The tile's text in LDS, as residue codes (a '\n' as kNewlineCode, which no
// residue has): chunk t (64 bytes) at byte 68 t, so that the lanes of a
// wave, each reading in its own chunk, fall in different banks.  The 16 bytes
// after the tile follow the last chunk.
constexpr uint32_t kNewlineCode = 0x1F;
constexpr int kStageStride = kChunk + 4;
constexpr int kStageBytes = kThreads * kStageStride + 16;

__device__ inline int stage_at(int64_t p)          // tile-relative byte -> LDS byte
{
    return (int)(p + (p >> 6) * (kStageStride - kChunk));
}

// One owned line segment of a thread's chunk: residue bytes [b, e), the block
// and slot of the residue at b, whether the residue at e - 1 ends its record,
// and its line.
struct Segment {
    int b, e;
    int64_t blk;
    int slot;
    bool record_end;
    int64_t line;
};

// Segment s of the chunk (bytes [b, e) up to the s-th '\n'), or one with
// e == b when the line holds no residue of a record there.
__device__ inline Segment make_segment(uint64_t d, int s, int b, int e, int64_t l, int64_t c0, int64_t n,
                                       int valid, uint8_t after, int64_t start0)
{
    Segment g{b, b, 0, 0, false, l};
    if ((d & 1) && e > b) {
        g.e = e;
        const int64_t start = s == 0 ? start0 : c0 + b;
        locate(d, start, c0 + b, g.blk, g.slot);
        const bool line_ends = e < valid || c0 + e >= n || after == '\n';
        g.record_end = ((d >> 1) & 1) && line_ends;
    }
    return g;
}

// Blocks whose first residue lies in one 16 KiB tile: at most 16384 / 3 + 1,
// since two records' residues are at least 3 bytes apart ('\n', '>', '\n').
constexpr int kMaxTileBlocks = 5632;

// Blocks of a segment whose first residue lies in it.
__device__ inline int segment_blocks(const Segment& g)
{
    const int o = (3 - g.slot) % 3;
    const int len = g.e - g.b;
    return o < len ? (len - o + 2) / 3 : 0;
}

// Emit: every thread assembles the blocks whose first residue lies in its own
// chunk.  A block inside one line segment is three consecutive bytes of the
// staged text; a block that runs past its segment continues on the next lines
// of its record (at most two line breaks away), read from the stage or, past
// the tile, from the text; a record's last block is padded.  The tile's blocks
// are contiguous in the store: they are collected in LDS at positions from a
// scan of per-thread counts and written out coalesced.
__global__ __launch_bounds__(kThreads) void emit_kernel(const uint8_t* text, int64_t n, int vec, const uint64_t* line,
                                                        const uint64_t* desc, int64_t m, const int64_t* first_line,
                                                        const uint16_t* chunk_line, const uint64_t* chunk_nl,
                                                        uint16_t* out, int64_t nblocks)
{
    __shared__ uint32_t lut[256];                   // byte -> code, '\n' -> kNewlineCode
    __shared__ __attribute__((aligned(16))) uint8_t stage[kStageBytes];
    __shared__ uint16_t blocks_out[kMaxTileBlocks];
    __shared__ int64_t lds[kThreads];
    __shared__ int64_t tile_first_block;
    const int tid = threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.x * kTile;
    const int64_t c0 = t0 + (int64_t)tid * kChunk;
    lut[tid] = tid == '\n' ? kNewlineCode : encode_byte((uint32_t)tid);

    // Loads that do not depend on the text go first: the line holding the
    // chunk's first byte, its start and the next descriptors.
    const int64_t l0 = first_line[blockIdx.x] - 1 + chunk_line[(int64_t)blockIdx.x * kThreads + tid];
    uint64_t dpre[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) dpre[k] = l0 + k < m ? desc[l0 + k] : 0;
    const int64_t start0 = l0 < m ? (int64_t)(line[l0] >> 2) : 0;
    Chunk ck;
    ck.nl = chunk_nl[(int64_t)blockIdx.x * kThreads + tid];
    load_words(text, n, c0, vec, ck);
    const uint8_t after = c0 + kChunk < n ? text[c0 + kChunk] : (uint8_t)'\n';
    __syncthreads();                                // the code table is complete
    {
        // The chunk's codes, 64 independent table reads, packed back into words.
        uint32_t* dst = reinterpret_cast<uint32_t*>(stage + tid * kStageStride);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t w = ck.w[k];
            dst[k] = lut[w & 0xFF] | (lut[(w >> 8) & 0xFF] << 8) | (lut[(w >> 16) & 0xFF] << 16) |
                     (lut[w >> 24] << 24);
        }
        if (tid < 16) {
            const int64_t p = t0 + kTile + tid;
            stage[kThreads * kStageStride + tid] = (uint8_t)(p < n ? lut[text[p]] : kNewlineCode);
        }
    }
    __syncthreads();

    // The code at tile-relative position q, from the stage or past it.
    auto code_at = [&](int64_t q) -> uint32_t {
        if (q < kTile + 16) return stage[stage_at(q)];
        return t0 + q < n ? lut[text[t0 + q]] : kNewlineCode;
    };
    auto line_desc = [&](int64_t l) -> uint64_t {
        const int64_t k = l - l0;
        if (k == 0) return dpre[0];
        if (k == 1) return dpre[1];
        if (k == 2) return dpre[2];
        return l < m ? desc[l] : 0;
    };

    const int segs = ck.valid > 0 ? __builtin_popcountll(ck.nl) + 1 : 0;
    // Pass 1: this thread's block count and first block; the first three
    // segments are kept for pass 2.
    int mine = 0;
    int64_t my_first = -1;
    Segment g0{0, 0, 0, 0, false, 0}, g1 = g0, g2 = g0;
    {
        uint64_t rest = ck.nl;
        int b = 0;
        for (int s = 0; s < segs; ++s) {
            const int e = rest ? __builtin_ctzll(rest) : ck.valid;
            rest &= rest - 1;
            const int64_t l = l0 + s;
            const Segment g = make_segment(line_desc(l), s, b, e, l, c0, n, ck.valid, after, start0);
            b = e + 1;
            if (s == 0) g0 = g;
            else if (s == 1) g1 = g;
            else if (s == 2) g2 = g;
            const int nb = g.e > g.b ? segment_blocks(g) : 0;
            if (nb && my_first < 0) my_first = g.blk + (g.slot != 0);
            mine += nb;
        }
    }
    int64_t total = 0;
    const int base = (int)block_exclusive_sum(mine, lds, &total);
    if (mine > 0 && base == 0) tile_first_block = my_first;
    // Pass 2: the blocks, into LDS at base + (block - my_first).
    auto index_of = [&](const Segment& g) -> int {       // LDS index of the segment's first block
        return base + (int)(g.blk + (g.slot != 0) - my_first);
    };
    auto put = [&](int i, uint32_t v) {             // i is below kMaxTileBlocks (bound above)
        blocks_out[i < kMaxTileBlocks ? i : kMaxTileBlocks - 1] = (uint16_t)v;
    };
    auto whole = [](const Segment& g) -> int {            // blocks entirely inside the segment
        const int o = (3 - g.slot) % 3, len = g.e - g.b;
        return g.e > g.b && o + 2 < len ? (len - o) / 3 : 0;
    };
    // (a) Whole blocks of the first three segments in one loop, so the lanes
    // of a wave run about the same number of iterations wherever their line
    // breaks fall.
    {
        const int w0 = segs > 0 ? whole(g0) : 0, w1 = segs > 1 ? whole(g1) : 0, w2 = segs > 2 ? whole(g2) : 0;
        const int a0 = tid * kStageStride + g0.b + (3 - g0.slot) % 3, i0 = index_of(g0);
        const int a1 = tid * kStageStride + g1.b + (3 - g1.slot) % 3, i1 = index_of(g1);
        const int a2 = tid * kStageStride + g2.b + (3 - g2.slot) % 3, i2 = index_of(g2);
        const int wtot = w0 + w1 + w2;
        // Segment of the k-th whole block by sequential selects on values (a
        // select among named values may be lowered to a table in private
        // memory; products would use the quarter-rate integer multiplier).
        const int w01 = w0 + w1;
#define MSA_WHERE(k, addr, idx)                                         \
    do {                                                                \
        const bool past0_ = (k) >= w0, past1_ = (k) >= w01;             \
        int first_ = 0, a_ = a0, i_ = i0;                               \
        first_ = past0_ ? w0 : first_;                                  \
        a_ = past0_ ? a1 : a_;                                          \
        i_ = past0_ ? i1 : i_;                                          \
        first_ = past1_ ? w01 : first_;                                 \
        a_ = past1_ ? a2 : a_;                                          \
        i_ = past1_ ? i2 : i_;                                          \
        const int j_ = (k) - first_;                                    \
        addr = a_ + j_ + j_ + j_;                                       \
        idx = i_ + j_;                                                  \
    } while (0)
        int k = 0;
        for (; k + 1 < wtot; k += 2) {                // two blocks, six reads in flight
            int aa, ia, ab, ib;
            MSA_WHERE(k, aa, ia);
            MSA_WHERE(k + 1, ab, ib);
            const uint32_t p0 = stage[aa], p1 = stage[aa + 1], p2 = stage[aa + 2];
            const uint32_t q0 = stage[ab], q1 = stage[ab + 1], q2 = stage[ab + 2];
            put(ia, (p0 << 1) | (p1 << 6) | (p2 << 11));
            put(ib, (q0 << 1) | (q1 << 6) | (q2 << 11));
        }
        if (k < wtot) {
            int aa, ia;
            MSA_WHERE(k, aa, ia);
            put(ia, ((uint32_t)stage[aa] << 1) | ((uint32_t)stage[aa + 1] << 6) | ((uint32_t)stage[aa + 2] << 11));
        }
#undef MSA_WHERE
    }
    // (b) Whole blocks of later segments (chunks of very short lines), and
    // (c) every segment's block that runs past it: one or two residues here,
    // the rest on the record's next line(s), or padding.
    {
        uint64_t rest = ck.nl;
        int b = 0;
        for (int s = 0; s < segs; ++s) {
            const int e = rest ? __builtin_ctzll(rest) : ck.valid;
            rest &= rest - 1;
            const int64_t l = l0 + s;
            const Segment g = s == 0 ? g0 : s == 1 ? g1 : s == 2 ? g2
                                                              : make_segment(line_desc(l), s, b, e, l, c0, n,
                                                                             ck.valid, after, start0);
            b = e + 1;
            if (g.e <= g.b) continue;
            const int len = g.e - g.b;
            const int nw = whole(g);
            const uint8_t* sp = stage + tid * kStageStride + g.b;
            int o = (3 - g.slot) % 3;
            if (s >= 3)
                for (int j = 0; j < nw; ++j)
                    put(index_of(g) + j, ((uint32_t)sp[o + 3 * j] << 1) | ((uint32_t)sp[o + 3 * j + 1] << 6) |
                                             ((uint32_t)sp[o + 3 * j + 2] << 11));
            o += 3 * nw;
            if (o >= len) continue;
            uint32_t v = (uint32_t)sp[o] << 1;
            int got = 1;
            if (o + 1 < len) v |= (uint32_t)sp[o + 1] << 6, got = 2;
            if (!g.record_end) {
                // The record goes on: past the segment comes its line's '\n' (then
                // the next line) or, at the chunk's end, more of the same line.
                int64_t q = (int64_t)tid * kChunk + g.e;           // tile-relative
                int64_t ln = g.line;
                if (code_at(q) == kNewlineCode) ++q, ++ln;
                while (got < 3 && t0 + q < n) {
                    v |= code_at(q) << (1 + 5 * got);
                    ++got;
                    const uint32_t nx = code_at(q + 1);
                    const bool line_ends = nx == kNewlineCode || t0 + q + 1 >= n;
                    if (line_ends && ((line_desc(ln) >> 1) & 1)) break;     // record ends here
                    if (line_ends) q += 2, ++ln;
                    else q += 1;
                }
            }
            for (int k = got; k < 3; ++k) v |= (uint32_t)kEndCode << (1 + 5 * k);
            if (got < 3) v |= 1u;
            put(index_of(g) + nw, v);
        }
    }
    __syncthreads();
    if (total > 0) {
        const int64_t first = tile_first_block;
        int64_t count = total < kMaxTileBlocks ? total : kMaxTileBlocks;
        if (count > nblocks - first) count = nblocks - first;
        for (int i = tid; i < count; i += kThreads) out[first + i] = blocks_out[i];
    }
}

// ---- generic exclusive scan of int64 (reduce, partials, apply) -------------------

__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const int64_t* data, int64_t count, int64_t* partial)
{
    __shared__ int64_t lds[kThreads];
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t sum = 0;
    for (int k = 0; k < kScanItems; ++k)
        if (i0 + k < count) sum += data[i0 + k];
    int64_t total = 0;
    block_exclusive_sum(sum, lds, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(kPartialThreads) void scan_partials_kernel(int64_t* partial, int64_t parts,
                                                                        int64_t initial, int64_t* total)
{
    __shared__ int64_t lds[kPartialThreads];
    const int tid = threadIdx.x;
    const int64_t span = (parts + kPartialThreads - 1) / kPartialThreads;
    const int64_t b0 = tid * span;
    int64_t sum = 0;
    for (int64_t b = b0; b < b0 + span && b < parts; ++b) sum += partial[b];
    lds[tid] = sum;
    __syncthreads();
    for (int d = 1; d < kPartialThreads; d <<= 1) {
        const int64_t add = tid >= d ? lds[tid - d] : 0;
        __syncthreads();
        lds[tid] += add;
        __syncthreads();
    }
    int64_t run = initial + lds[tid] - sum;
    for (int64_t b = b0; b < b0 + span && b < parts; ++b) {
        const int64_t v = partial[b];
        partial[b] = run;
        run += v;
    }
    if (tid == kPartialThreads - 1) *total = initial + lds[tid];
}

__global__ __launch_bounds__(kThreads) void scan_apply_kernel(int64_t* data, int64_t count, const int64_t* partial)
{
    __shared__ int64_t lds[kThreads];
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = i0 + k < count ? data[i0 + k] : 0;
        sum += v[k];
    }
    int64_t run = partial[blockIdx.x] + block_exclusive_sum(sum, lds, nullptr);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (i0 + k < count) {
            data[i0 + k] = run;
            run += v[k];
        }
}

}  // namespace

// ---- host side -------------------------------------------------------------------

namespace {

size_t up256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

}  // namespace msa

struct msa_fasta_opaque {
    int device = 0;
    // results
    int64_t records = 0, blocks = 0, residues = 0, lines = 0, bytes = 0;
    uint16_t* d_blocks = nullptr;
    int64_t* d_offset = nullptr;
    int32_t* d_size = nullptr;
    int64_t* d_desc_start = nullptr;
    int64_t* d_desc_end = nullptr;
    // buffers, grown on demand and kept for the next parse
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    Buf text, tiles, chunks, partial, line, rank, rec, aggs, states, recs, out, small;
    int64_t* h_small = nullptr;       // pinned readback of the totals
    // Stage events of the last parse: start, line table done, records done,
    // emit started, emit done.
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool timed = false;

    ~msa_fasta_opaque()
    {
        for (Buf* b : {&text, &tiles, &chunks, &partial, &line, &rank, &rec, &aggs, &states, &recs, &out, &small})
            if (b->p) (void)hipFree(b->p);
        if (h_small) (void)hipHostFree(h_small);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

namespace msa {
namespace {

int grow(msa_fasta_opaque::Buf& b, size_t bytes)
{
    bytes = std::max<size_t>(up256(bytes), 256);
    if (b.bytes >= bytes) return MSA_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    MSA_HIP(hipMalloc(&b.p, bytes));
    b.bytes = bytes;
    return MSA_OK;
}

template <typename T>
T* as(const msa_fasta_opaque::Buf& b, size_t offset = 0)
{
    return reinterpret_cast<T*>(static_cast<char*>(b.p) + offset);
}

int64_t blocks_for(int64_t count, int64_t per) { return (count + per - 1) / per; }

// Exclusive scan of `count` int64 in place, `initial` first; the total lands
// in *d_total (device).
int scan_i64(msa_fasta_opaque* f, int64_t* data, int64_t count, int64_t initial, int64_t* d_total, hipStream_t st)
{
    const int64_t parts = std::max<int64_t>(1, blocks_for(count, kScanTile));
    MSA_TRY(grow(f->partial, (size_t)parts * sizeof(int64_t)));
    int64_t* partial = as<int64_t>(f->partial);
    if (count > 0) scan_reduce_kernel<<<(unsigned)parts, kThreads, 0, st>>>(data, count, partial);
    else MSA_HIP(hipMemsetAsync(partial, 0, sizeof(int64_t), st));
    scan_partials_kernel<<<1, kPartialThreads, 0, st>>>(partial, parts, initial, d_total);
    if (count > 0) scan_apply_kernel<<<(unsigned)parts, kThreads, 0, st>>>(data, count, partial);
    return hip_status(hipGetLastError(), "scan kernels");
}

// Reads `count` int64 totals from the small buffer.
int read_small(msa_fasta_opaque* f, int count, hipStream_t st)
{
    MSA_HIP(hipMemcpyAsync(f->h_small, f->small.p, count * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    MSA_HIP(hipStreamSynchronize(st));
    return MSA_OK;
}

int parse(msa_fasta_opaque* f, const uint8_t* d_text, int64_t n, hipStream_t st)
{
    f->records = f->blocks = f->residues = f->lines = 0;
    f->bytes = n;
    f->timed = false;
    if (n == 0) return MSA_OK;
    if (!f->ev[0])
        for (hipEvent_t& e : f->ev) MSA_HIP(hipEventCreate(&e));
    MSA_HIP(hipEventRecord(f->ev[0], st));
    if (!f->h_small) MSA_HIP(hipHostMalloc(reinterpret_cast<void**>(&f->h_small), 64 * sizeof(int64_t)));
    MSA_TRY(grow(f->small, 64 * sizeof(int64_t)));
    int64_t* small = as<int64_t>(f->small);
    const int vec = (reinterpret_cast<uintptr_t>(d_text) & 15) == 0;
    const int64_t tiles = blocks_for(n, kTile);
    if (tiles >= (int64_t(1) << 31)) return fail(MSA_ERR_TOO_LARGE, "text of %lld bytes is too large", (long long)n);

    // 1-2. '\n' per tile, then the first line each tile starts.
    MSA_TRY(grow(f->tiles, (size_t)tiles * sizeof(int64_t)));
    int64_t* first_line = as<int64_t>(f->tiles);
    count_kernel<<<(unsigned)tiles, kThreads, 0, st>>>(d_text, n, vec, first_line);
    MSA_TRY(scan_i64(f, first_line, tiles, 1, small + 0, st));
    MSA_TRY(read_small(f, 1, st));
    // Every '\n' but one at the very end starts a line.
    uint8_t last_byte = 0;
    MSA_HIP(hipMemcpyAsync(&last_byte, d_text + n - 1, 1, hipMemcpyDeviceToHost, st));
    MSA_HIP(hipStreamSynchronize(st));
    const int64_t m = f->h_small[0] - (last_byte == '\n' ? 1 : 0);
    f->lines = m;

    // 3. The line table.
    MSA_TRY(grow(f->line, (size_t)m * sizeof(uint64_t)));
    MSA_TRY(grow(f->rank, (size_t)m * sizeof(uint64_t)));
    MSA_TRY(grow(f->rec, (size_t)(m + 1) * sizeof(int32_t)));
    uint64_t* line = as<uint64_t>(f->line);
    const size_t nchunks = (size_t)tiles * kThreads;
    MSA_TRY(grow(f->chunks, nchunks * (sizeof(uint64_t) + sizeof(uint16_t))));
    uint64_t* chunk_nl = as<uint64_t>(f->chunks);
    uint16_t* chunk_line = as<uint16_t>(f->chunks, nchunks * sizeof(uint64_t));
    index_kernel<<<(unsigned)tiles, kThreads, 0, st>>>(d_text, n, vec, first_line, line, m, chunk_line, chunk_nl);
    MSA_HIP(hipEventRecord(f->ev[1], st));

    // 4. Record state over lines.
    const int64_t ltiles = blocks_for(m, kLineTile);
    MSA_TRY(grow(f->aggs, (size_t)ltiles * sizeof(LineAgg)));
    MSA_TRY(grow(f->states, (size_t)(ltiles + 1) * sizeof(LineState)));
    LineAgg* aggs = as<LineAgg>(f->aggs);
    LineState* states = as<LineState>(f->states);
    LineState* totals = states + ltiles;
    line_reduce_kernel<<<(unsigned)ltiles, kThreads, 0, st>>>(d_text, n, line, m, aggs);
    line_partials_kernel<<<1, kPartialThreads, 0, st>>>(aggs, ltiles, states, totals);
    MSA_HIP(hipGetLastError());
    MSA_HIP(hipMemcpyAsync(f->h_small, totals, sizeof(LineState), hipMemcpyDeviceToHost, st));
    MSA_HIP(hipStreamSynchronize(st));
    LineState tot;
    std::memcpy(&tot, f->h_small, sizeof(tot));
    const int64_t records = tot.headers;
    f->records = records;
    f->residues = tot.residues;
    if (records >= (int64_t(1) << 31))
        return fail(MSA_ERR_TOO_LARGE, "%lld records; at most 2^31-1 are supported", (long long)records);

    // Records: first residue, descriptions, blocks, offsets.
    const size_t r_al = up256((size_t)std::max<int64_t>(records, 1) * sizeof(int64_t));
    const size_t r_i32 = up256((size_t)std::max<int64_t>(records, 1) * sizeof(int32_t));
    MSA_TRY(grow(f->recs, 4 * r_al + r_i32));
    int64_t* rec_first = as<int64_t>(f->recs, 0);
    f->d_offset = as<int64_t>(f->recs, r_al);
    f->d_desc_start = as<int64_t>(f->recs, 2 * r_al);
    f->d_desc_end = as<int64_t>(f->recs, 3 * r_al);
    f->d_size = as<int32_t>(f->recs, 4 * r_al);
    int32_t* rec = as<int32_t>(f->rec);
    uint64_t* rank = as<uint64_t>(f->rank);
    line_apply_kernel<<<(unsigned)ltiles, kThreads, 0, st>>>(d_text, n, line, m, states, rec, rank, rec_first,
                                                              f->d_desc_start, f->d_desc_end, records);
    if (records > 0) {
        records_kernel<<<(unsigned)blocks_for(records, kThreads), kThreads, 0, st>>>(rec_first, records, totals,
                                                                                    f->d_offset, f->d_size);
        MSA_HIP(hipGetLastError());
    }
    MSA_TRY(scan_i64(f, f->d_offset, records, 0, small + 1, st));
    MSA_TRY(read_small(f, 2, st));
    f->blocks = f->h_small[1];

    // 6-7. Emit descriptors, then the blocks.
    MSA_HIP(hipEventRecord(f->ev[2], st));
    emit_desc_kernel<<<(unsigned)blocks_for(m, kThreads), kThreads, 0, st>>>(m, rec, rank, rec_first, f->d_offset,
                                                                                records);
    MSA_TRY(grow(f->out, (size_t)f->blocks * sizeof(uint16_t)));
    f->d_blocks = as<uint16_t>(f->out);
    MSA_HIP(hipEventRecord(f->ev[3], st));
    emit_kernel<<<(unsigned)tiles, kThreads, 0, st>>>(d_text, n, vec, line, rank, m, first_line, chunk_line,
                                                      chunk_nl, f->d_blocks, f->blocks);
    MSA_HIP(hipGetLastError());
    MSA_HIP(hipEventRecord(f->ev[4], st));
    MSA_HIP(hipStreamSynchronize(st));
    f->timed = true;
    return MSA_OK;
}

}  // namespace
}  // namespace msa

using msa::fail;

extern "C" {

int msa_fasta_create(msa_fasta* out)
{
    if (!out) return fail(MSA_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    auto* f = new (std::nothrow) msa_fasta_opaque();
    if (!f) return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed");
    int rc = msa::hip_status(hipGetDevice(&f->device), "hipGetDevice");
    if (rc != MSA_OK) {
        delete f;
        return rc;
    }
    *out = f;
    return MSA_OK;
}

int msa_fasta_destroy(msa_fasta f)
{
    delete f;
    return MSA_OK;
}

int msa_fasta_parse_device(msa_fasta f, const char* d_text, int64_t bytes, msa_stream stream)
{
    if (!f) return fail(MSA_ERR_INVALID_ARGUMENT, "null loader");
    if (bytes < 0) return fail(MSA_ERR_INVALID_ARGUMENT, "negative text size");
    if (bytes > 0 && !d_text) return fail(MSA_ERR_INVALID_ARGUMENT, "null text");
    return msa::parse(f, reinterpret_cast<const uint8_t*>(d_text), bytes, (hipStream_t)stream);
}

int msa_fasta_parse(msa_fasta f, const char* text, int64_t bytes)
{
    if (!f) return fail(MSA_ERR_INVALID_ARGUMENT, "null loader");
    if (bytes < 0) return fail(MSA_ERR_INVALID_ARGUMENT, "negative text size");
    if (bytes > 0 && !text) return fail(MSA_ERR_INVALID_ARGUMENT, "null text");
    if (bytes > 0) {
        MSA_TRY(msa::grow(f->text, (size_t)bytes));
        MSA_HIP(hipMemcpy(f->text.p, text, (size_t)bytes, hipMemcpyHostToDevice));
    }
    return msa::parse(f, static_cast<const uint8_t*>(f->text.p), bytes, nullptr);
}

int msa_fasta_info_get(msa_fasta f, msa_fasta_info* info)
{
    if (!f || !info) return fail(MSA_ERR_INVALID_ARGUMENT, "null argument");
    info->sequences = f->records;
    info->blocks = f->blocks;
    info->residues = f->residues;
    info->lines = f->lines;
    info->bytes = f->bytes;
    return MSA_OK;
}

int msa_fasta_timing_get(msa_fasta f, msa_fasta_timing* t)
{
    if (!f || !t) return fail(MSA_ERR_INVALID_ARGUMENT, "null argument");
    *t = msa_fasta_timing{0, 0, 0, 0};
    if (!f->timed) return MSA_OK;
    float ms[4];
    const int pairs[4][2] = {{0, 4}, {0, 1}, {1, 3}, {3, 4}};
    for (int k = 0; k < 4; ++k)
        MSA_HIP(hipEventElapsedTime(&ms[k], f->ev[pairs[k][0]], f->ev[pairs[k][1]]));
    t->total_ms = ms[0];
    t->lines_ms = ms[1];
    t->records_ms = ms[2];
    t->emit_ms = ms[3];
    return MSA_OK;
}

int msa_fasta_device_store(msa_fasta f, const uint16_t** d_blocks, const int64_t** d_offset, const int32_t** d_size,
                           const int64_t** d_desc_start, const int64_t** d_desc_end)
{
    if (!f) return fail(MSA_ERR_INVALID_ARGUMENT, "null loader");
    if (d_blocks) *d_blocks = f->d_blocks;
    if (d_offset) *d_offset = f->d_offset;
    if (d_size) *d_size = f->d_size;
    if (d_desc_start) *d_desc_start = f->d_desc_start;
    if (d_desc_end) *d_desc_end = f->d_desc_end;
    return MSA_OK;
}

int msa_fasta_copy(msa_fasta f, uint16_t* blocks, int64_t* offset, int32_t* size, int64_t* desc_start,
                   int64_t* desc_end)
{
    if (!f) return fail(MSA_ERR_INVALID_ARGUMENT, "null loader");
    const size_t r = (size_t)f->records;
    if (blocks && f->blocks)
        MSA_HIP(hipMemcpy(blocks, f->d_blocks, (size_t)f->blocks * sizeof(uint16_t), hipMemcpyDeviceToHost));
    if (offset && r) MSA_HIP(hipMemcpy(offset, f->d_offset, r * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (size && r) MSA_HIP(hipMemcpy(size, f->d_size, r * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (desc_start && r) MSA_HIP(hipMemcpy(desc_start, f->d_desc_start, r * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (desc_end && r) MSA_HIP(hipMemcpy(desc_end, f->d_desc_end, r * sizeof(int64_t), hipMemcpyDeviceToHost));
    return MSA_OK;
}

}  // extern "C"
