// The long-sequence Needleman-Wunsch kernel: one pair per workgroup, its rows
// cut into strips that the workgroup's waves score as a pipelined
// anti-diagonal wavefront of tiles.
//
// A pair whose shorter side runs to tens of thousands of residues is too big
// for one lane group (the batched kernel) and too few pairs exist to fill the
// GPU one pair per group.  Here a workgroup of NW waves takes one pair at a
// time.  The query is cut into strips of 64*R rows (a first, partial strip at
// the top takes the neutral padding rows); strip s goes to wave s % NW.  Each
// wave runs its strip as a 64-lane wavefront (lane k owns R rows, bottom
// scores move one lane per step by DPP wave shifts, each lane reads its
// column's slot from a ring in LDS) over 64-column chunks, i.e. over a row of
// 64R x 64 tiles.  The strip below
// follows a few tiles behind: after every tile the upper wave stores the
// strip's bottom row for those 64 columns to a line in global memory (the
// int32 tile boundary) and publishes its progress in an LDS counter; the
// lower wave waits for that counter before loading the boundary.  All waves
// of a workgroup share one compute unit, so workgroup-scope release/acquire
// orders those stores and loads.  Waiting is bounded: a wave that waits too
// long raises the error flag and stops.
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

#include <type_traits>

namespace msa {

namespace {

constexpr int kLanes = 64;
constexpr int kRingBytes = 4 * kLanes;   // one wave's slot ring
constexpr int kSpinLimit = 1 << 24;     // ~1e9 cycles of waiting before giving up

struct LongArgs {
    const uint16_t* blocks;
    const SeqInfo* info;
    const int4* pairs;          // {query, target, output index, transpose flag}
    int32_t npairs;
    int32_t* queue;
    int32_t* error;
    int32_t* out;
    int32_t* lines;             // per workgroup: max_boundaries lines of line_stride ints
    int32_t line_stride;
    int32_t max_boundaries;
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

__device__ __forceinline__ int line_value(const Target& t, const int32_t* line, int chunk, int k)
{
    const int idx = chunk * kLanes + k - t.lpad;
    return (idx >= 0 && idx < t.len) ? line[idx] : 0;
}

// Waits until the strip above has published at least `need` columns.
__device__ __forceinline__ bool wait_for(int* progress, int need, int32_t* error)
{
    int spins = 0;
    while (__hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < need) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > kSpinLimit) {
            if (__lane_id() == 0) atomicExch(error, 1);
            return false;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}

// Column slots come from the wave's slot ring in LDS (256 bytes): at step s
// lane k works virtual column s-k, and column i sits at byte i mod 128 and
// again at i mod 128 + 128, so a lane reads the slot for step u of a chunk at
// a per-chunk base plus u -- one LDS byte read two steps ahead instead of two
// DPP moves; each chunk's slots go in before step 61 of the chunk before.
// The reads are volatile so that they stay single-byte reads.
// The bottom-row store of a strip: the workgroup's boundary lines as one raw
// buffer (wave-uniform descriptor; dword 3: 32-bit data format, as for any
// raw buffer access on gfx9) and this lane's byte offset of the strip's line
// -- out of range, so the store is dropped, on every lane but the last.
constexpr int kRawBufferFlags = 0x00020000;
constexpr uint32_t kOutOfRange = 0x80000000u;
struct LineStore {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t lane_off;
};

template <int R, bool LOAD, bool STORE>
__device__ int run_strip(const dev::lds_u32* prof, const dev::lds_u8* remap, int null_slot, const Target& t,
                         const int32_t* line_in, int* progress_in, const LineStore& out, int* progress_out,
                         int32_t* error, volatile dev::lds_u8* ring, int k)
{
    constexpr int W = R / 4;
    constexpr int SLOT_WORDS = kLanes * W;
    constexpr int RING = 2 * kLanes;
    int h[R];
#pragma unroll
    for (int r = 0; r < R; ++r) h[r] = 0;
    int dtop = 0;
    bool ok = true;

    if constexpr (LOAD) ok = wait_for(progress_in, min(t.len, 2 * kLanes - t.lpad), error);
    // Columns -64..-1 are neutral; chunk 0 goes in before the first read.
    ring[kLanes + k] = null_slot;
    ring[3 * kLanes + k] = null_slot;
    {
        const int c0 = col_slot(t, remap, null_slot, 0, k);
        ring[k] = c0;
        ring[RING + k] = c0;
    }
    int cc_next = col_slot(t, remap, null_slot, 1, k);
    int ln = 0, ln_next = 0;
    if constexpr (LOAD) {
        ln = line_value(t, line_in, 0, k);
        ln_next = line_value(t, line_in, 1, k);
    }
    uint32_t voff = 0;
    // Slots of steps 0 and 1 (virtual columns -k and 1-k).
    const int base0 = (RING - k) & (RING - 1);
    int b = ring[base0];
    int bn = ring[base0 + 1];
    uint32_t w[W];
    load_prof<R>(w, prof + b * SLOT_WORDS, k);

    // Step u of a chunk whose slot for step u + 2 sits at byte base + u; LAST:
    // the step that moves to the next chunk.
    auto step = [&](auto last_tag, int base, int u) {
        constexpr bool LAST = decltype(last_tag)::value;
        const int bnn = ring[base + u];
        uint32_t wn[W];
        load_prof<R>(wn, prof + bn * SLOT_WORDS, k);
        const int hup = LOAD ? dev::dpp<dev::kWaveShr1>(ln, h[R - 1]) : dev::dpp_zero<dev::kWaveShr1>(h[R - 1]);
        if constexpr (LOAD) ln = LAST ? ln_next : dev::dpp<dev::kWaveShl1>(ln, ln);
        dev::cells<R>(h, hup, dtop, w);
        dtop = hup;
        // The last lane's bottom cell goes straight to the line (column
        // c*64 + u - 63 - lpad): one buffer store, the step's 4u in the
        // scalar offset; columns left of the target land in the guard.
        if constexpr (STORE) __builtin_amdgcn_raw_buffer_store_b32((uint32_t)h[R - 1], out.rsrc, (int)voff, 4 * u, 0);
        b = bn;
        bn = bnn;
#pragma unroll
        for (int q = 0; q < W; ++q) w[q] = wn[q];
    };

    for (int c = 0; c < t.chunks && ok; ++c) {
        const int base = (c * kLanes + 2 - k) & (RING - 1);
        if constexpr (STORE) voff = out.lane_off + (uint32_t)(c * kLanes - (kLanes - 1) - t.lpad) * 4u;
#pragma unroll 9
        for (int u = 0; u < kLanes - 3; ++u) step(std::false_type {}, base, u);
        // Chunk c+1 replaces chunk c-1, whose last reader was step 60.
        {
            const int slot = ((c + 1) * kLanes + k) & (RING - 1);
            ring[slot] = cc_next;
            ring[RING + slot] = cc_next;
        }
        step(std::false_type {}, base, kLanes - 3);
        step(std::false_type {}, base, kLanes - 2);
        step(std::true_type {}, base, kLanes - 1);
        if constexpr (STORE) {
            // Columns up to c*64 - lpad are stored.
            const int done = min(t.len, c * kLanes + 1 - t.lpad);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (k == 0 && done > 0) __hip_atomic_store(progress_out, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        cc_next = col_slot(t, remap, null_slot, c + 2, k);
        if constexpr (LOAD) {
            const int need = min(t.len, (c + 3) * kLanes - t.lpad);
            if (need > 0) ok = wait_for(progress_in, need, error);
            ln_next = line_value(t, line_in, c + 2, k);
        }
    }
    if constexpr (STORE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (k == 0) __hip_atomic_store(progress_out, t.len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return h[R - 1];
}

template <int R, int NW>
__global__ __launch_bounds__(NW * 64) void nw_long_kernel(LongArgs a)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];   // 16-byte aligned: ds_read_b128 of profiles
    constexpr int STRIP = kLanes * R;
    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int prof_words = a.slots * kLanes * R / 4;
    dev::lds_u32* base = (dev::lds_u32*)lds;
    dev::lds_u32* prof = base + wave * prof_words;
    volatile dev::lds_u8* ring = (dev::lds_u8*)(base + NW * prof_words) + wave * kRingBytes;
    dev::lds_u8* tab = (dev::lds_u8*)(base + NW * prof_words) + NW * kRingBytes;
    dev::lds_u8* remap = tab + 640;
    dev::lds_u8* slot_code = remap + 32;
    int* progress = (int*)(slot_code + 32);      // max_boundaries counters
    __shared__ int s_pair;
    __shared__ int s_result;

    for (int i = threadIdx.x; i < 640; i += NW * 64) tab[i] = a.sxx.v[i];
    if (threadIdx.x < 32) {
        remap[threadIdx.x] = a.remap[threadIdx.x];
        slot_code[threadIdx.x] = a.slot_code[threadIdx.x];
    }
    const int null_slot = a.slots - 1;
    int32_t* lines = a.lines + (int64_t)blockIdx.x * a.max_boundaries * a.line_stride;
    __amdgpu_buffer_rsrc_t lines_rsrc;
    {
        const uint64_t wb = (uint64_t)(uintptr_t)lines;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)wb);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(wb >> 32));
        lines_rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)hi << 32) | lo), 0,
                                                       a.max_boundaries * a.line_stride * 4, kRawBufferFlags);
    }

    while (true) {
        __syncthreads();
        if (threadIdx.x == 0) s_pair = atomicAdd(a.queue, 1);
        for (int i = threadIdx.x; i < a.max_boundaries; i += NW * 64) progress[i] = 0;
        __syncthreads();
        const int it = s_pair;
        if (it >= a.npairs) break;
        const int4 pr = a.pairs[it];
        const SeqInfo qi = a.info[pr.x];
        const SeqInfo ti = a.info[pr.y];
        const int Lq = qi.z;
        const int nst = (Lq + STRIP - 1) / STRIP;
        const int pad_top = nst * STRIP - Lq;
        Target t;
        t.seq = a.blocks + ti.x;
        t.len = ti.z;
        t.chunks = (t.len + 2 * kLanes - 2) / kLanes;
        t.lpad = t.chunks * kLanes - (t.len + kLanes - 1);

        for (int st = wave; st < nst; st += NW) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_s_waitcnt(0xC07F);
            __builtin_amdgcn_wave_barrier();
            build_profile<R>(prof, tab, slot_code, a.slots, a.blocks + qi.x, st * STRIP, pad_top, pr.w != 0, lane);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_s_waitcnt(0xC07F);
            __builtin_amdgcn_wave_barrier();
            const bool load = st > 0, store = st + 1 < nst;
            const int32_t* lin = lines + (int64_t)(st - 1) * a.line_stride + kLongLineGuard;
            LineStore lout;
            lout.rsrc = lines_rsrc;
            lout.lane_off = lane == kLanes - 1 ? (uint32_t)(st * a.line_stride + kLongLineGuard) * 4u : kOutOfRange;
            int res;
            if (!load && !store)
                res = run_strip<R, false, false>(prof, remap, null_slot, t, lin, progress, lout, progress, a.error, ring, lane);
            else if (!load)
                res = run_strip<R, false, true>(prof, remap, null_slot, t, lin, progress, lout, progress + st, a.error, ring, lane);
            else if (store)
                res = run_strip<R, true, true>(prof, remap, null_slot, t, lin, progress + st - 1, lout, progress + st, a.error,
                                               ring, lane);
            else
                res = run_strip<R, true, false>(prof, remap, null_slot, t, lin, progress + st - 1, lout, progress, a.error,
                                                ring, lane);
            if (st == nst - 1 && lane == kLanes - 1) s_result = res;
        }
        __syncthreads();
        if (threadIdx.x == 0) a.out[pr.z] = s_result - (Lq + t.len) * a.penalty;
    }
}

__global__ void long_reset_kernel(int32_t* queue, int32_t* error)
{
    if (threadIdx.x == 0) {
        queue[0] = 0;
        error[0] = 0;
    }
}

template <int R, int NW>
const void* long_fn() { return reinterpret_cast<const void*>(&nw_long_kernel<R, NW>); }

constexpr int kLongWaves = 4;

}  // namespace

int long_rows_for_slots(int slots) { return slots <= 8 ? 32 : (slots <= 16 ? 16 : 8); }

int long_threads() { return kLongWaves * 64; }

int long_strip_rows(int rows) { return kLanes * rows; }

size_t long_lds_bytes(int rows, int slots, int max_boundaries)
{
    return (size_t)kLongWaves * (slots * kLanes * rows + kRingBytes) + 640 + 64 + 4 * (size_t)max_boundaries;
}

static const void* long_kernel_fn(int rows)
{
    switch (rows) {
    case 8: return long_fn<8, kLongWaves>();
    case 16: return long_fn<16, kLongWaves>();
    default: return long_fn<32, kLongWaves>();
    }
}

int long_occupancy(int rows, size_t lds, int* blocks_per_cu)
{
    const void* fn = long_kernel_fn(rows);
    MSA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int n = 0;
    MSA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kLongWaves * 64, lds));
    *blocks_per_cu = n;
    return MSA_OK;
}

int long_launch(int rows, int grid, size_t lds, const uint16_t* d_blocks, const SeqInfo* d_info, const int4* d_pairs,
                int32_t npairs, int32_t* d_queue, int32_t* d_error, int32_t* d_out, int32_t* d_lines,
                int32_t line_stride, int32_t max_boundaries, int slots, const uint8_t* remap,
                const uint8_t* slot_code, const Sxx& sxx, int32_t penalty, hipStream_t stream)
{
    hipLaunchKernelGGL(long_reset_kernel, dim3(1), dim3(64), 0, stream, d_queue, d_error);
    MSA_HIP(hipGetLastError());
    LongArgs a;
    a.blocks = d_blocks;
    a.info = d_info;
    a.pairs = d_pairs;
    a.npairs = npairs;
    a.queue = d_queue;
    a.error = d_error;
    a.out = d_out;
    a.lines = d_lines;
    a.line_stride = line_stride;
    a.max_boundaries = max_boundaries;
    a.slots = slots;
    a.penalty = penalty;
    a.sxx = sxx;
    for (int i = 0; i < 32; ++i) {
        a.remap[i] = remap[i];
        a.slot_code[i] = slot_code[i];
    }
    switch (rows) {
    case 8:
        hipLaunchKernelGGL((nw_long_kernel<8, kLongWaves>), dim3(grid), dim3(kLongWaves * 64), lds, stream, a);
        break;
    case 16:
        hipLaunchKernelGGL((nw_long_kernel<16, kLongWaves>), dim3(grid), dim3(kLongWaves * 64), lds, stream, a);
        break;
    default:
        hipLaunchKernelGGL((nw_long_kernel<32, kLongWaves>), dim3(grid), dim3(kLongWaves * 64), lds, stream, a);
        break;
    }
    return hip_status(hipGetLastError(), "nw_long_kernel launch");
}

}  // namespace msa
