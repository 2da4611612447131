// The long-sequence Needleman-Wunsch kernel: every pair's rows cut into
// strips, and the strips of all pairs scored by one-wave workgroups that take
// them from a global queue, each strip an anti-diagonal wavefront that
// streams the bottom row of the strip above it from global memory.
//
// A pair whose shorter side runs to tens of thousands of residues is too big
// for one lane group (the batched kernel), and too few such pairs exist to
// fill the GPU one pair per group.  Here the query of each pair is cut into
// strips of 64*R rows (a first, partial strip at the top takes the neutral
// padding rows) and the work list holds one task per (pair, strip): the pairs
// (costliest first) in groups as large as the grid, each group level by
// level -- strip 0 of every pair of the group, then strip 1, ...  So the
// strip above a task was handed out about a group earlier and is well on its
// way when the task starts (the planner keeps about one resident wave per
// long pair for this, two to four per SIMD, as many on every SIMD).
// Persistent one-wave workgroups take tasks in
// list order from one counter: a pair's strips go to whichever waves are free,
// on any compute unit, and the last pairs spread over the whole GPU instead
// of leaving it idle at the end.
//
// A wave runs its strip as a 64-lane wavefront (lane k owns R rows, bottom
// scores move one lane per step by DPP wave shifts, each lane reads its
// column's slot from a ring in LDS) over 64-column chunks.  The last lane
// stores the strip's bottom cell at every step into a line in global memory;
// the strip below loads the line two chunks ahead and hands it to its first
// lane through a small LDS ring.  Every line word is 64
// bits, (tag << 32) | score, stored with the cache policy of an agent-scope
// atomic store (sc1) and loaded by relaxed agent-scope atomic loads, so the
// words carry their own synchronisation: a reader takes a word only when its
// tag is the producer's, and otherwise loads it again (the atomics of one
// location are coherent, so the producer's word arrives).  No fence is needed
// on the way, and a reader that keeps up with its producer just waits for it.
//
// Lines live in a ring of `nlines` slots; the strips that have a strip below
// are numbered in list order, and strip o writes slot o % nlines as its
// (o / nlines)-th occupant.  An occupant's tag is unique to it and to this
// execution (a per-execution base from a counter that only grows, plus the
// occupant number), so no earlier line is ever mistaken for it.  Before an
// occupant writes, it waits (agent-scope acquire) until the slot's previous
// line has been read in full: its reader advances the slot's turn with an
// agent-scope release when its strip is done.  Every wait -- for a word, for
// a turn -- is for a task handed out earlier, whose wave is running, so no
// wait deadlocks whatever the grid (the planner makes the ring longer than
// the distance between a strip and the strip below it in the list).  Waiting
// is bounded: a wave that waits too long, or sees the error flag, raises the
// flag and stops.
//
// Only pairs the batched kernel could take come here (no padding code
// before the last residue of the reference's column sequence); the planner
// (nw_plan.hip) routes them by length.  Scores follow the same shifted
// recurrence as the other kernels.  Profiles index a compact alphabet: the
// planner maps the codes present in the data to slots 0..S-2 and slot S-1 is
// the neutral code, so a DNA strip profile is 5 x 64 x R bytes.
#include "msa_internal.hpp"
#include "nw_device.hpp"
#include "nw_planner.hpp"

#include <algorithm>
#include <type_traits>

namespace msa {

namespace {

constexpr int kLanes = 64;
constexpr int kRingBytes = 4 * kLanes;   // the wave's slot ring
constexpr int kLineRingBytes = 2 * kLanes * 4;   // the wave's line ring: two chunks of ints
constexpr int kSpinLimit = 1 << 24;     // ~1e9 cycles of waiting before giving up
constexpr int kSc1 = 16;                // buffer-instruction cache policy: sc1 (agent scope)
constexpr int kFirstStep = 8;           // rows-per-lane granularity of a pair's first strip

struct LongArgs {
    const uint16_t* blocks;
    const SeqInfo* info;
    const int4* pairs;          // {query, target, output index, transpose flag}
    const int4* tasks;          // {pair, strip, line read (-1: none), line written (-1: none)}
    int32_t ntasks;
    int32_t* queue;
    int32_t* error;
    int32_t* out;
    uint64_t* lines;            // nlines lines of line_stride words: (tag << 32) | score
    int32_t* turn;              // per line slot: occupants whose line has been read
    const uint32_t* tag_word;   // this execution's tags end just below it
    int32_t line_stride;
    int32_t nlines;
    int32_t occupants;          // occupants per slot in one execution (at most)
    int32_t slots;              // profile slots, neutral slot last
    int32_t penalty;
    Sxx sxx;
    uint8_t remap[32];          // code -> slot (codes absent from the data -> neutral slot)
    uint8_t slot_code[32];      // slot -> code (the neutral slot -> 0xFF)
};

template <int R>
__device__ __forceinline__ void load_prof(uint32_t (&w)[R / 4], const dev::lds_u32* slot_base, int k)
{
    constexpr int FULL = R / 16;
#pragma unroll
    for (int j = 0; j < FULL; ++j) {
        const dev::u32x4 v = *(const dev::lds_u32x4*)(slot_base + j * kLanes * 4 + k * 4);
        w[4 * j + 0] = v.x;
        w[4 * j + 1] = v.y;
        w[4 * j + 2] = v.z;
        w[4 * j + 3] = v.w;
    }
    if constexpr (R % 16 == 8) {
        const dev::u32x2 v = *(const dev::lds_u32x2*)(slot_base + FULL * kLanes * 4 + k * 2);
        w[4 * FULL + 0] = v.x;
        w[4 * FULL + 1] = v.y;
    }
}

template <int R>
__device__ __forceinline__ int prof_word(int k, int w)
{
    const int j = w >> 2;
    if (j < R / 16) return j * kLanes * 4 + k * 4 + (w & 3);
    return (R / 16) * kLanes * 4 + k * 2 + (w & 3);
}

// Profile of rows [row0, row0 + 64R) of the padded query.
template <int R>
__device__ void build_profile(dev::lds_u32* prof, const dev::lds_u8* tab, const dev::lds_u8* slot_code, int slots,
                              const uint16_t* __restrict__ qseq, int row0, int pad_top, bool transpose, int lane)
{
    constexpr int W = R / 4;
    constexpr int SLOT_WORDS = kLanes * W;
    for (int idx = lane; idx < kLanes * W; idx += kLanes) {
        const int k = idx / W;
        const int w = idx - k * W;
        int codes[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = row0 + k * R + 4 * w + t - pad_top;
            const int code = q >= 0 ? dev::code_at(qseq, q) : -1;
            codes[t] = code < kEndCode ? code : -1;
        }
        dev::lds_u32* dst = prof + prof_word<R>(k, w);
        for (int s = 0; s < slots; ++s) {
            const int b = slot_code[s];
            uint32_t word = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                uint32_t v = 0;
                if (codes[t] >= 0 && b < 25) v = transpose ? tab[b * 25 + codes[t]] : tab[codes[t] * 25 + b];
                word |= v << (8 * t);
            }
            dst[s * SLOT_WORDS] = word;
        }
    }
}

struct Target {
    const uint16_t* seq;
    int len;
    int lpad;       // neutral columns in front: len + lpad + 63 is a multiple of 64
    int chunks;
};

__device__ __forceinline__ int col_slot(const Target& t, const dev::lds_u8* remap, int null_slot, int chunk, int k)
{
    const int idx = chunk * kLanes + k - t.lpad;
    return (idx >= 0 && idx < t.len) ? (int)remap[dev::code_at(t.seq, idx)] : null_slot;
}

// A lane's column of the target, chunk after chunk (idx = chunk*64 + k -
// lpad), as block index and bit shift advanced by additions (a chunk moves 64
// columns: 21 blocks and one 5-bit slot).  The block is loaded at the start of
// a chunk and decoded after its steps: a load used at once would make the
// wave wait for it and, the wait counter being shared with stores, for every
// line store the chunk has issued.
struct SlotCursor {
    int idx;
    int blk;
    int shift;      // 1 + 5 * (idx mod 3)

    __device__ __forceinline__ SlotCursor(const Target& t, int chunk, int k)
    {
        idx = chunk * kLanes + k - t.lpad;
        const int q = (idx >= 0 ? idx : idx - 2) / 3;   // floor, also left of the target
        blk = q;
        shift = 1 + 5 * (idx - 3 * q);
    }
    __device__ __forceinline__ bool inside(const Target& t) const { return idx >= 0 && idx < t.len; }
    __device__ __forceinline__ void advance()
    {
        idx += kLanes;
        blk += kLanes / 3;
        shift += 5 * (kLanes % 3);
        if (shift > 11) {
            shift -= 15;
            blk += 1;
        }
    }
};

__device__ __forceinline__ void raise_error(int32_t* error)
{
    __hip_atomic_store(error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool give_up(int spins, int32_t* error)
{
    if ((spins & 255) != 0) return false;
    if (spins > kSpinLimit || __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        raise_error(error);
        return true;
    }
    return false;
}

// Waits until the slot's turn reaches `want`, then acquires (agent scope);
// false when the wait gave up or the error flag is up.
__device__ __forceinline__ bool wait_turn(const int32_t* turn, int want, int32_t* error)
{
    int spins = 0;
    while (__hip_atomic_load(turn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
        __builtin_amdgcn_s_sleep(2);
        if (give_up(++spins, error)) return false;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// Lane k's word of the line above for chunk `chunk` (0 outside the target).
__device__ __forceinline__ uint64_t line_word(const Target& t, const uint64_t* line, int chunk, int k)
{
    const int idx = chunk * kLanes + k - t.lpad;
    return (idx >= 0 && idx < t.len) ? __hip_atomic_load(line + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                     : 0;
}

// The score in a word loaded by line_word, once its tag is the producer's:
// lanes holding an older word load it again until the producer's arrives.
__device__ __forceinline__ bool take_word(const Target& t, const uint64_t* line, int chunk, int k, uint64_t raw,
                                          uint32_t tag, int32_t* error, int& value)
{
    const int idx = chunk * kLanes + k - t.lpad;
    const bool inside = idx >= 0 && idx < t.len;
    bool stale = inside && (uint32_t)(raw >> 32) != tag;
    if (__any(stale)) {
        int spins = 0;
        while (true) {
            __builtin_amdgcn_s_sleep(1);
            if (stale) {
                raw = __hip_atomic_load(line + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                stale = (uint32_t)(raw >> 32) != tag;
            }
            if (!__any(stale)) break;
            if (give_up(++spins, error)) return false;
        }
    }
    value = inside ? (int)(uint32_t)raw : 0;
    return true;
}

// Column slots come from the wave's slot ring in LDS (256 bytes): at step s
// lane k works virtual column s-k, and column i sits at byte i mod 128 and
// again at i mod 128 + 128, so a lane reads the slot for step u of a chunk at
// a per-chunk base plus u -- one LDS byte read two steps ahead instead of two
// DPP moves; each chunk's slots go in before step 61 of the chunk before.
// The reads are volatile so that they stay single-byte reads.
// The bottom-row store of a strip: the strip's line as one raw buffer
// (wave-uniform descriptor; dword 3: 32-bit data format, as for any raw
// buffer access on gfx9) and this lane's byte offset in it -- out of range,
// so the store is dropped, on every lane but the last.  One descriptor spans
// one line, so every offset fits in 32 bits.
constexpr int kRawBufferFlags = 0x00020000;
constexpr uint32_t kOutOfRange = 0x80000000u;
typedef int v2i32 __attribute__((ext_vector_type(2)));

// Where a strip reads the line above and writes its own.
struct Exchange {
    const uint64_t* line_in;    // column 0 of the line above
    uint32_t tag_in;            // its producer's tag
    __amdgpu_buffer_rsrc_t out; // this strip's line
    uint32_t lane_off;          // byte offset of column -63-lpad on the last lane, else out of range
    int32_t tag_out;
};

// The line above, chunk by chunk, for the first lane: every lane writes its
// word of chunk c into an LDS ring of two chunks (64 ints each), and at step u
// the wave reads int u of the chunk (one address for all lanes, read one step
// ahead); the DPP shift that hands each lane the bottom cell of the lane
// before it writes into that register in place, so the first lane keeps the
// line value as its cell above.
typedef __attribute__((address_space(3))) int lds_i32;

template <int R, bool LOAD, bool STORE>
__device__ int run_strip(const dev::lds_u32* prof, const dev::lds_u8* remap, int null_slot, const Target& t,
                         const Exchange& x, int32_t* error, volatile dev::lds_u8* ring, volatile lds_i32* lline, int k,
                         bool& ok)
{
    constexpr int W = R / 4;
    constexpr int SLOT_WORDS = kLanes * W;
    constexpr int RING = 2 * kLanes;
    int h[R];
#pragma unroll
    for (int r = 0; r < R; ++r) h[r] = 0;
    int dtop = 0;

    // Columns -64..-1 are neutral; chunk 0 goes in before the first read.
    ring[kLanes + k] = null_slot;
    ring[3 * kLanes + k] = null_slot;
    {
        const int c0 = col_slot(t, remap, null_slot, 0, k);
        ring[k] = c0;
        ring[RING + k] = c0;
    }
    int cc_next = col_slot(t, remap, null_slot, 1, k);
    int lv = 0;     // the line value of the current step (first lane)
    uint64_t raw_next = 0;
    if constexpr (LOAD) {
        const uint64_t raw0 = line_word(t, x.line_in, 0, k);
        raw_next = line_word(t, x.line_in, 1, k);
        int l0 = 0;
        ok = take_word(t, x.line_in, 0, k, raw0, x.tag_in, error, l0);
        lline[k] = l0;
        lv = lline[0];
    }
    // The last lane stores its bottom cell and the tag at every step: one
    // 8-byte buffer store with the sc1 policy at an offset out of range (so
    // dropped) on every other lane, the step's 8u in the scalar offset.
    uint32_t voff = 0;
    // Slots of steps 0 and 1 (virtual columns -k and 1-k).
    const int base0 = (RING - k) & (RING - 1);
    int b = ring[base0];
    int bn = ring[base0 + 1];
    uint32_t w[W];
    load_prof<R>(w, prof + b * SLOT_WORDS, k);

    // Step u of a chunk whose slot for step u + 2 sits at byte base + u and
    // whose line values sit at lcur (the next chunk's at lnext); LAST: the
    // step that moves to the next chunk.
    auto step = [&](auto last_tag, int base, int u, volatile lds_i32* lcur, volatile lds_i32* lnext) {
        constexpr bool LAST = decltype(last_tag)::value;
        const int bnn = ring[base + u];
        uint32_t wn[W];
        load_prof<R>(wn, prof + bn * SLOT_WORDS, k);
        int hup;
        if constexpr (LOAD) {
            const int lvn = LAST ? lnext[0] : lcur[u + 1];
            hup = dev::dpp<dev::kWaveShr1>(lv, h[R - 1]);
            lv = lvn;
        } else {
            hup = dev::dpp_zero<dev::kWaveShr1>(h[R - 1]);
        }
        dev::cells<R>(h, hup, dtop, w);
        dtop = hup;
        if constexpr (STORE) {
            const v2i32 word = {h[R - 1], x.tag_out};
            __builtin_amdgcn_raw_buffer_store_b64(word, x.out, (int)voff, 8 * u, kSc1);
        }
        b = bn;
        bn = bnn;
#pragma unroll
        for (int q = 0; q < W; ++q) w[q] = wn[q];
    };

    SlotCursor cur(t, 2, k);    // the column decoded after chunk c's steps is chunk c+2's
    for (int c = 0; c < t.chunks && ok; ++c) {
        // Chunk c+2's block, loaded before this chunk's steps and decoded
        // after them.
        const bool in2 = cur.inside(t);
        const int raw2 = in2 ? (int)t.seq[cur.blk] : 0;
        const int base = (c * kLanes + 2 - k) & (RING - 1);
        volatile lds_i32* lcur = lline + (c & 1) * kLanes;
        volatile lds_i32* lnext = lline + ((c + 1) & 1) * kLanes;
        // Step u stores column c*64 + u - 63 - lpad (last lane).
        if constexpr (STORE) voff = x.lane_off + (uint32_t)(c * kLanes) * 8u;
#pragma unroll 9
        for (int u = 0; u < kLanes - 3; ++u) step(std::false_type {}, base, u, lcur, lnext);
        // Chunk c+1 replaces chunk c-1, whose last reader was step 60.
        {
            const int slot = ((c + 1) * kLanes + k) & (RING - 1);
            ring[slot] = cc_next;
            ring[RING + slot] = cc_next;
        }
        step(std::false_type {}, base, kLanes - 3, lcur, lnext);
        step(std::false_type {}, base, kLanes - 2, lcur, lnext);
        // The line above for chunk c+1, loaded a chunk ago, is taken and put
        // in the LDS ring (over chunk c-1, last read at step 62 of chunk c-1)
        // before the step that reads its first value.
        if constexpr (LOAD) {
            int lnv = 0;
            ok = take_word(t, x.line_in, c + 1, k, raw_next, x.tag_in, error, lnv);
            lnext[k] = lnv;
        }
        step(std::true_type {}, base, kLanes - 1, lcur, lnext);
        cc_next = in2 ? (int)remap[(raw2 >> cur.shift) & 0x1F] : null_slot;
        cur.advance();
        if constexpr (LOAD) raw_next = line_word(t, x.line_in, c + 2, k);
    }
    return h[R - 1];
}

// First strips: R0 rows per lane, R0 in kFirstStep..R (the first strip never
// loads a line; it stores one when a strip follows it).
template <int R, bool STORE, int R0 = kFirstStep>
__device__ __forceinline__ int run_first_strip(int r0, const dev::lds_u32* prof, const dev::lds_u8* remap,
                                               int null_slot, const Target& t, const Exchange& x, int32_t* error,
                                               volatile dev::lds_u8* ring, volatile lds_i32* lline, int k, bool& ok)
{
    if constexpr (R0 >= R) {
        return run_strip<R, false, STORE>(prof, remap, null_slot, t, x, error, ring, lline, k, ok);
    } else {
        if (r0 == R0) return run_strip<R0, false, STORE>(prof, remap, null_slot, t, x, error, ring, lline, k, ok);
        return run_first_strip<R, STORE, R0 + kFirstStep>(r0, prof, remap, null_slot, t, x, error, ring, lline, k, ok);
    }
}

template <int R, int R0 = kFirstStep>
__device__ __forceinline__ void build_first_profile(int r0, dev::lds_u32* prof, const dev::lds_u8* tab,
                                                    const dev::lds_u8* slot_code, int slots,
                                                    const uint16_t* __restrict__ qseq, int pad_top, bool transpose,
                                                    int lane)
{
    if constexpr (R0 >= R) {
        build_profile<R>(prof, tab, slot_code, slots, qseq, 0, pad_top, transpose, lane);
    } else {
        if (r0 == R0) build_profile<R0>(prof, tab, slot_code, slots, qseq, 0, pad_top, transpose, lane);
        else build_first_profile<R, R0 + kFirstStep>(r0, prof, tab, slot_code, slots, qseq, pad_top, transpose, lane);
    }
}

// Dynamic LDS of one workgroup (one wave): the strip profile, the slot ring,
// the line ring, the byte table, the code -> slot and slot -> code maps.
template <int R>
__global__ __launch_bounds__(kLanes) void nw_long_kernel(LongArgs a)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];   // 16-byte aligned: ds_read_b128 of profiles
    constexpr int STRIP = kLanes * R;
    const int lane = threadIdx.x;
    const int prof_words = a.slots * kLanes * R / 4;
    dev::lds_u32* prof = (dev::lds_u32*)lds;
    volatile dev::lds_u8* ring = (dev::lds_u8*)(prof + prof_words);
    volatile lds_i32* lline = (lds_i32*)(prof + prof_words + kRingBytes / 4);
    dev::lds_u8* tab = (dev::lds_u8*)(prof + prof_words) + kRingBytes + kLineRingBytes;
    dev::lds_u8* remap = tab + 640;
    dev::lds_u8* slot_code = remap + 32;

    for (int i = lane; i < 640; i += kLanes) tab[i] = a.sxx.v[i];
    if (lane < 32) {
        remap[lane] = a.remap[lane];
        slot_code[lane] = a.slot_code[lane];
    }
    const int null_slot = a.slots - 1;
    // Tags of this execution: tag_base + occupant, all above any earlier tag.
    const uint32_t tag_base = *a.tag_word - (uint32_t)a.occupants;

    while (true) {
        int t = 0;
        if (lane == 0) t = atomicAdd(a.queue, 1);
        t = __shfl(t, 0);
        if (t >= a.ntasks) break;
        const int4 task = a.tasks[t];
        const int4 pr = a.pairs[task.x];
        const int st = task.y;
        const SeqInfo qi = a.info[pr.x];
        const SeqInfo ti = a.info[pr.y];
        const int Lq = qi.z;
        const int nst = (Lq + STRIP - 1) / STRIP;
        // The first strip takes the rows left over by the full strips below
        // it, kFirstStep rows per lane at a time, with at most 64 kFirstStep
        // - 1 neutral padding rows on top.
        const int rem = Lq - (nst - 1) * STRIP;
        const int R0 = min(R, (rem + kLanes * kFirstStep - 1) / (kLanes * kFirstStep) * kFirstStep);
        const int pad_top = kLanes * R0 - rem;
        const int row0 = st == 0 ? 0 : kLanes * R0 + (st - 1) * STRIP;
        Target tg;
        tg.seq = a.blocks + ti.x;
        tg.len = ti.z;
        tg.chunks = (tg.len + 2 * kLanes - 2) / kLanes;
        tg.lpad = tg.chunks * kLanes - (tg.len + kLanes - 1);

        // The previous task's LDS reads (profile, ring) are done before the
        // profile is rebuilt.
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();
        if (st == 0)
            build_first_profile<R>(R0, prof, tab, slot_code, a.slots, a.blocks + qi.x, pad_top, pr.w != 0, lane);
        else
            build_profile<R>(prof, tab, slot_code, a.slots, a.blocks + qi.x, row0, pad_top, pr.w != 0, lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();

        const bool load = task.z >= 0, store = task.w >= 0;
        const int si = load ? task.z % a.nlines : 0, ki = load ? task.z / a.nlines : 0;
        const int so = store ? task.w % a.nlines : 0, ko = store ? task.w / a.nlines : 0;
        Exchange x;
        x.line_in = a.lines + (int64_t)si * a.line_stride + kLongLineGuard;
        x.tag_in = tag_base + (uint32_t)ki;
        x.tag_out = (int32_t)(tag_base + (uint32_t)ko);
        {
            const uint64_t wb = (uint64_t)(uintptr_t)(a.lines + (int64_t)so * a.line_stride);
            const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)wb);
            const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(wb >> 32));
            x.out = __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)hi << 32) | lo), 0, a.line_stride * 8,
                                                      kRawBufferFlags);
        }
        // Lane u of chunk c's row holds column c*64 + u - 63 - lpad, stored at
        // word guard + that column.
        x.lane_off = lane == kLanes - 1 ? (uint32_t)(kLongLineGuard - (kLanes - 1) - tg.lpad) * 8u : kOutOfRange;

        bool ok = true;
        // The slot's previous line must have been read in full before this
        // occupant overwrites it.
        if (store) ok = wait_turn(a.turn + so, ko, a.error);
        int res = 0;
        if (ok) {
            if (!load && !store)
                res = run_first_strip<R, false>(R0, prof, remap, null_slot, tg, x, a.error, ring, lline, lane, ok);
            else if (!load)
                res = run_first_strip<R, true>(R0, prof, remap, null_slot, tg, x, a.error, ring, lline, lane, ok);
            else if (store)
                res = run_strip<R, true, true>(prof, remap, null_slot, tg, x, a.error, ring, lline, lane, ok);
            else
                res = run_strip<R, true, false>(prof, remap, null_slot, tg, x, a.error, ring, lline, lane, ok);
        }
        if (!ok) break;     // the error flag is up; every wave stops
        if (load) {
            // The line above has been read: its slot goes to the next occupant.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (lane == 0) __hip_atomic_store(a.turn + si, ki + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // The last strip's last lane holds the last query row at the last column.
        if (st == nst - 1 && lane == kLanes - 1) a.out[pr.z] = res - (Lq + tg.len) * a.penalty;
    }
}

// Clears the queue counter, the error flag and every line slot's turn (a fill
// kernel: no memset node in a captured graph), and moves the tag counter past
// every tag this execution uses: occupants + 1 higher than before.
__global__ void long_reset_kernel(int32_t* queue, int32_t* error, int32_t* turn, uint32_t* tag_word, int32_t nlines,
                                  int32_t occupants)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        queue[0] = 0;
        error[0] = 0;
        tag_word[0] += (uint32_t)occupants + 1u;
    }
    for (int s = i; s < nlines; s += gridDim.x * blockDim.x) turn[s] = 0;
}

template <int R>
const void* long_fn() { return reinterpret_cast<const void*>(&nw_long_kernel<R>); }

}  // namespace

int long_rows_for_slots(int slots) { return slots <= 8 ? 32 : (slots <= 16 ? 16 : 8); }

int long_strip_rows(int rows) { return kLanes * rows; }

size_t long_lds_bytes(int rows, int slots)
{
    return (size_t)slots * kLanes * rows + kRingBytes + kLineRingBytes + 640 + 64;
}

static const void* long_kernel_fn(int rows)
{
    switch (rows) {
    case 8: return long_fn<8>();
    case 16: return long_fn<16>();
    default: return long_fn<32>();
    }
}

int long_occupancy(int rows, size_t lds, int* blocks_per_cu)
{
    const void* fn = long_kernel_fn(rows);
    MSA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int n = 0;
    MSA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kLanes, lds));
    *blocks_per_cu = n;
    return MSA_OK;
}

int long_max_grid(int compute_units, int* grid)
{
    // The fewest profile slots each strip height is used with (a compact
    // alphabet of 1, 8 or 16 codes plus the neutral slot) gives its least LDS
    // and so its most resident workgroups.
    int most = 0;
    for (const int slots : {2, 9, 17}) {
        const int rows = long_rows_for_slots(slots);
        int per_cu = 0;
        MSA_TRY(long_occupancy(rows, long_lds_bytes(rows, slots), &per_cu));
        most = std::max(most, per_cu);
    }
    *grid = most * compute_units;
    return MSA_OK;
}

int long_launch(const LongLaunch& l, hipStream_t stream)
{
    const int reset_grid = (int)std::min<int64_t>(1024, ((int64_t)l.nlines + 255) / 256);
    hipLaunchKernelGGL(long_reset_kernel, dim3(std::max(reset_grid, 1)), dim3(256), 0, stream, l.queue, l.error,
                       l.turn, l.tag_word, l.nlines, l.occupants);
    MSA_HIP(hipGetLastError());
    LongArgs a;
    a.blocks = l.blocks;
    a.info = l.info;
    a.pairs = l.pairs;
    a.tasks = l.tasks;
    a.ntasks = l.ntasks;
    a.queue = l.queue;
    a.error = l.error;
    a.out = l.out;
    a.lines = l.lines;
    a.turn = l.turn;
    a.tag_word = l.tag_word;
    a.line_stride = l.line_stride;
    a.nlines = l.nlines;
    a.occupants = l.occupants;
    a.slots = l.slots;
    a.penalty = l.penalty;
    a.sxx = l.sxx;
    for (int i = 0; i < 32; ++i) {
        a.remap[i] = l.remap[i];
        a.slot_code[i] = l.slot_code[i];
    }
    switch (l.rows) {
    case 8:
        hipLaunchKernelGGL((nw_long_kernel<8>), dim3(l.grid), dim3(kLanes), l.lds, stream, a);
        break;
    case 16:
        hipLaunchKernelGGL((nw_long_kernel<16>), dim3(l.grid), dim3(kLanes), l.lds, stream, a);
        break;
    default:
        hipLaunchKernelGGL((nw_long_kernel<32>), dim3(l.grid), dim3(kLanes), l.lds, stream, a);
        break;
    }
    return hip_status(hipGetLastError(), "nw_long_kernel launch");
}

}  // namespace msa
