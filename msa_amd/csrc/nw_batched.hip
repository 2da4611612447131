// The batched Needleman-Wunsch kernel: one persistent launch over a queue of
// work items, each item one "query" sequence and a run of "targets" (the
// other sequence of each pair), all sorted by length by the planner
// (nw_plan.hip).
//
// Which pairs come here.  Only pairs whose score is the plain global
// alignment of the two sequences cut at their first padding code: the
// reference's padding rules (src/pairwise/needleman/impl/sequential.cpp:
// 40-66) reduce to that when the column sequence holds no padding code before
// its last residue.  The shorter sequence of a pair is the query; with a
// non-symmetric table the profile is built from the transposed table for
// pairs whose query is the reference's column sequence, so every score keeps
// the reference's orientation (:87-91).
//
// Shape of the computation.  Every wave of a workgroup works alone: it takes
// an item from an atomic counter, splits its query into row strips, and its
// four 16-lane groups score four targets at a time.  Inside a group, lane k
// owns R consecutive rows of the strip; the DP runs as an anti-diagonal
// wavefront -- at step s lane k updates column s-k -- with the column residue
// and the bottom score of each lane moving one lane down per step by DPP row
// shifts.  The substitution scores of a lane's rows come from the strip's
// profile in the wave's LDS region: R bytes per (column code, lane), read with
// ds_read_b128 / b64 at a 16- or 8-byte lane stride (conflict-free), one step
// ahead of use.  A strip's bottom row goes to a per-group line in global
// memory for the next strip.
//
// Neutral padding.  The recurrence runs in the shifted domain G = H +
// (i+j)*penalty, where every boundary cell is 0 and every cell is
// max(G_diag + sxx, G_up, G_left) with sxx = T + 2*penalty >= 0.  Rows and
// columns whose profile entries are 0 and whose neighbours are all 0 stay 0,
// so the query is padded with such rows at the top of its first strip (the
// last real row is always lane 15, register R-1 of the last strip) and each
// target with such columns at its left (every group ends exactly at the end
// of a 16-step chunk).  The inner loop therefore needs no per-step masking:
// results are captured at chunk ends.
#include "msa_internal.hpp"
#include "nw_device.hpp"

namespace msa {

namespace {

constexpr int kG = 16;               // lanes per pair
constexpr int kGroups = 64 / kG;     // pairs per wave
constexpr int kRMax = 32;            // rows per lane in a full strip
constexpr int kStripMax = kG * kRMax;
constexpr int kCodes = 25;           // codes 0..23 and the neutral code 24
constexpr int kNull = 24;
constexpr int kWaves = 4;            // waves per workgroup (independent)
constexpr int kProfWords = kCodes * kG * kRMax / 4;   // one wave's profile region

struct BatchArgs {
    const uint16_t* blocks;
    const SeqInfo* info;
    const int4* items;          // {query, first target, target count, flags}
    const int2* targets;        // {sequence, output index}
    int32_t nitems;
    int32_t* queue;
    int32_t* out;
    int32_t* lines;             // one line of line_stride ints per lane group
    int32_t line_stride;
    int32_t penalty;
    Sxx sxx;
};

// Word offset, inside one code's block, of the first word a lane reads.  Full
// 16-byte planes come first ([plane][lane][4 words]), then, when R is not a
// multiple of 16, an 8-byte plane ([lane][2 words]).
template <int R>
__device__ __forceinline__ void load_prof(uint32_t (&w)[R / 4], const dev::lds_u32* code_base, int k)
{
    constexpr int FULL = R / 16;
#pragma unroll
    for (int j = 0; j < FULL; ++j) {
        const dev::u32x4 v = *(const dev::lds_u32x4*)(code_base + j * kG * 4 + k * 4);
        w[4 * j + 0] = v.x;
        w[4 * j + 1] = v.y;
        w[4 * j + 2] = v.z;
        w[4 * j + 3] = v.w;
    }
    if constexpr (R % 16 == 8) {
        const dev::u32x2 v = *(const dev::lds_u32x2*)(code_base + FULL * kG * 4 + k * 2);
        w[4 * FULL + 0] = v.x;
        w[4 * FULL + 1] = v.y;
    }
}

// Word index of word `w` of lane `k` inside a code block (inverse of load_prof).
template <int R>
__device__ __forceinline__ int prof_word(int k, int w)
{
    const int j = w >> 2;
    if (j < R / 16) return j * kG * 4 + k * 4 + (w & 3);
    return (R / 16) * kG * 4 + k * 2 + (w & 3);
}

// Builds the profile of one strip: rows [row0, row0 + 16R) of the padded
// query (rows below pad_top are neutral).
template <int R>
__device__ void build_profile(dev::lds_u32* prof, const dev::lds_u8* tab,
                              const uint16_t* __restrict__ qseq, int row0, int pad_top, bool transpose, int lane)
{
    constexpr int W = R / 4;
    constexpr int CODE_WORDS = kG * W;
    for (int slot = lane; slot < kG * W; slot += 64) {
        const int k = slot / W;
        const int w = slot - k * W;
        int codes[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = row0 + k * R + 4 * w + t - pad_top;
            const int code = q >= 0 ? dev::code_at(qseq, q) : -1;
            codes[t] = code < kNull ? code : -1;
        }
        dev::lds_u32* dst = prof + prof_word<R>(k, w);
        for (int b = 0; b < kCodes; ++b) {
            uint32_t word = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int code = codes[t];
                uint32_t v = 0;
                if (code >= 0 && b < kNull) v = transpose ? tab[b * 25 + code] : tab[code * 25 + b];
                word |= v << (8 * t);
            }
            dst[b * CODE_WORDS] = word;
        }
    }
}

// Per-group view of the target being scored.
struct TargetState {
    const uint16_t* seq;
    int len;        // columns
    int lpad;       // neutral columns in front, so that len + lpad + 15 is a multiple of 16
    int chunks;     // (lpad + len + 15) / 16, 0 for an idle group
};

__device__ __forceinline__ int col_code(const TargetState& t, int chunk, int k)
{
    const int idx = chunk * kG + k - t.lpad;
    return (idx >= 0 && idx < t.len) ? min(dev::code_at(t.seq, idx), kNull) : kNull;
}

__device__ __forceinline__ int line_value(const TargetState& t, const int32_t* __restrict__ line, int chunk, int k)
{
    const int idx = chunk * kG + k - t.lpad;
    return (idx >= 0 && idx < t.len) ? line[idx] : 0;
}

// One strip of R rows per lane over the whole target.  LOAD: the strip above
// left its bottom row in `line`; STORE: this strip leaves its bottom row there.
// Returns (on lane 15 of the group) the last row's score at the last column.
template <int R, bool LOAD, bool STORE>
__device__ int run_strip(const dev::lds_u32* prof, const TargetState& t, int32_t* __restrict__ line,
                         int wchunks, int k)
{
    constexpr int W = R / 4;
    constexpr int CODE_WORDS = kG * W;
    int h[R];
#pragma unroll
    for (int r = 0; r < R; ++r) h[r] = 0;
    int dtop = 0;
    int result = 0;

    int cc = col_code(t, 0, k);
    int cc_next = col_code(t, 1, k);
    int ln = 0, ln_next = 0;
    if constexpr (LOAD) {
        ln = line_value(t, line, 0, k);
        ln_next = line_value(t, line, 1, k);
    }
    int outv = 0;

    // b(0): lane 0 takes the first residue, the other lanes start neutral.
    int b = dev::group_shr<kG>(cc, kNull);
    uint32_t w[W];
    load_prof<R>(w, prof + b * CODE_WORDS, k);

    for (int c = 0; c < wchunks; ++c) {
#pragma unroll
        for (int u = 0; u < kG; ++u) {
            // Residue stream one step ahead: lane 0 takes the next residue.
            cc = (u < kG - 1) ? dev::group_shl<kG>(cc) : cc_next;
            const int bn = dev::group_shr<kG>(cc, b);
            uint32_t wn[W];
            load_prof<R>(wn, prof + bn * CODE_WORDS, k);

            const int hup = LOAD ? dev::group_shr<kG>(ln, h[R - 1]) : dev::dpp_zero<dev::kRowShr1>(h[R - 1]);
            if constexpr (LOAD) ln = (u < kG - 1) ? dev::group_shl<kG>(ln) : ln_next;

            dev::cells<R>(h, hup, dtop, w);
            dtop = hup;
            if constexpr (STORE) outv = dev::dpp<dev::kRowShl1>(h[R - 1], outv);

            b = bn;
#pragma unroll
            for (int q = 0; q < W; ++q) w[q] = wn[q];
        }
        if constexpr (STORE) {
            const int col = c * kG + k - (kG - 1) - t.lpad;
            if (col >= 0 && col < t.len) line[col] = outv;
        }
        result = (c == t.chunks - 1) ? h[R - 1] : result;
        cc_next = col_code(t, c + 2, k);
        if constexpr (LOAD) ln_next = line_value(t, line, c + 2, k);
    }
    return result;
}

template <bool LOAD, bool STORE>
__device__ int run_strip_r(int R, const dev::lds_u32* prof, const TargetState& t, int32_t* line, int wchunks, int k)
{
    switch (R) {
    case 8: return run_strip<8, LOAD, STORE>(prof, t, line, wchunks, k);
    case 16: return run_strip<16, LOAD, STORE>(prof, t, line, wchunks, k);
    case 24: return run_strip<24, LOAD, STORE>(prof, t, line, wchunks, k);
    default: return run_strip<32, LOAD, STORE>(prof, t, line, wchunks, k);
    }
}

__device__ void build_profile_r(int R, dev::lds_u32* prof, const dev::lds_u8* tab, const uint16_t* qseq, int row0,
                                int pad_top, bool transpose, int lane)
{
    switch (R) {
    case 8: build_profile<8>(prof, tab, qseq, row0, pad_top, transpose, lane); break;
    case 16: build_profile<16>(prof, tab, qseq, row0, pad_top, transpose, lane); break;
    case 24: build_profile<24>(prof, tab, qseq, row0, pad_top, transpose, lane); break;
    default: build_profile<32>(prof, tab, qseq, row0, pad_top, transpose, lane); break;
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

__global__ __launch_bounds__(kWaves * 64) void nw_batched_kernel(BatchArgs a)
{
    extern __shared__ uint32_t lds[];
    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int g = lane / kG;
    const int k = lane % kG;
    dev::lds_u32* lds_base = (dev::lds_u32*)lds;   // address-space cast: 32-bit LDS offsets
    dev::lds_u32* prof = lds_base + wave * kProfWords;
    dev::lds_u8* tab = (dev::lds_u8*)(lds_base + kWaves * kProfWords);
    for (int i = threadIdx.x; i < 640; i += kWaves * 64) tab[i] = a.sxx.v[i];
    __syncthreads();

    int32_t* line = a.lines + ((int64_t)(blockIdx.x * kWaves + wave) * kGroups + g) * a.line_stride;
    const int p = a.penalty;

    while (true) {
        int it = 0;
        if (lane == 0) it = atomicAdd(a.queue, 1);
        it = __shfl(it, 0);
        if (it >= a.nitems) break;

        const int4 item = a.items[it];
        const SeqInfo qi = a.info[item.x];
        const int Lq = qi.z;
        const uint16_t* qseq = a.blocks + qi.x;
        const bool transpose = (item.w & 1) != 0;
        // Strips: a first strip of R0 rows per lane (R0 in 8..32, step 8),
        // then full strips of 32 rows per lane.
        const int full = (Lq - 1) / kStripMax;            // strips after the first
        const int rem = Lq - full * kStripMax;              // rows of the first strip (1..512)
        const int R0 = min(kRMax, ((rem + kG * 8 - 1) / (kG * 8)) * 8);
        const int pad_top = kG * R0 - rem;
        const int nst = full + 1;

        const int rounds = (item.z + kGroups - 1) / kGroups;
        for (int rd = 0; rd < rounds; ++rd) {
            const int ti = rd * kGroups + g;
            TargetState t{a.blocks, 0, 0, 0};
            int2 tg = make_int2(0, -1);
            if (ti < item.z) {
                tg = a.targets[item.y + ti];
                const SeqInfo si = a.info[tg.x];
                t.seq = a.blocks + si.x;
                t.len = si.z;
                t.chunks = (t.len + 2 * kG - 2) / kG;
                t.lpad = t.chunks * kG - (t.len + kG - 1);
            }
            int wchunks = t.chunks;
            wchunks = max(wchunks, __shfl_xor(wchunks, 16));
            wchunks = max(wchunks, __shfl_xor(wchunks, 32));

            int res = 0;
            for (int st = 0; st < nst; ++st) {
                const int R = st == 0 ? R0 : kRMax;
                const int row0 = st == 0 ? 0 : kG * R0 + (st - 1) * kStripMax;
                if (nst > 1 || rd == 0) {
                    wave_lds_sync();
                    build_profile_r(R, prof, tab, qseq, row0, pad_top, transpose, lane);
                    wave_lds_sync();
                }
                const bool load = st > 0;
                const bool store = st + 1 < nst;
                if (!load && !store) res = run_strip_r<false, false>(R, prof, t, line, wchunks, k);
                else if (!load) res = run_strip_r<false, true>(R, prof, t, line, wchunks, k);
                else if (store) res = run_strip<kRMax, true, true>(prof, t, line, wchunks, k);
                else res = run_strip<kRMax, true, false>(prof, t, line, wchunks, k);
                // The next strip reads this strip's bottom row from `line`.
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            }
            if (tg.y >= 0 && k == kG - 1) a.out[tg.y] = res - (Lq + t.len) * p;
        }
    }
}

__global__ void queue_reset_kernel(int32_t* queue)
{
    if (threadIdx.x == 0) queue[0] = 0;
}

}  // namespace

int batched_threads() { return kWaves * 64; }

size_t batched_lds_bytes() { return (size_t)kWaves * kProfWords * 4 + 640; }

int batched_groups_per_block() { return kWaves * kGroups; }

int batched_strip_rows() { return kStripMax; }

int batched_occupancy(int* blocks_per_cu)
{
    const void* fn = reinterpret_cast<const void*>(&nw_batched_kernel);
    const size_t lds = batched_lds_bytes();
    MSA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int n = 0;
    MSA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kWaves * 64, lds));
    *blocks_per_cu = n;
    return MSA_OK;
}

int batched_launch(int grid, const uint16_t* d_blocks, const SeqInfo* d_info, const int4* d_items, int32_t nitems,
                   const int2* d_targets, int32_t* d_queue, int32_t* d_out, int32_t* d_lines, int32_t line_stride,
                   const Sxx& sxx, int32_t penalty, hipStream_t stream)
{
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, stream, d_queue);
    MSA_HIP(hipGetLastError());
    BatchArgs a;
    a.blocks = d_blocks;
    a.info = d_info;
    a.items = d_items;
    a.targets = d_targets;
    a.nitems = nitems;
    a.queue = d_queue;
    a.out = d_out;
    a.lines = d_lines;
    a.line_stride = line_stride;
    a.penalty = penalty;
    a.sxx = sxx;
    hipLaunchKernelGGL(nw_batched_kernel, dim3(grid), dim3(kWaves * 64), batched_lds_bytes(), stream, a);
    return hip_status(hipGetLastError(), "nw_batched_kernel launch");
}

}  // namespace msa
