// The batched Needleman-Wunsch kernel: a persistent, length-binned work queue.
//
// Work is organised around a "query" sequence whose substitution profile sits
// in LDS and is shared by every wave of the workgroup: a work item is one
// query plus a run of targets (the other sequence of each pair), sorted by
// length so that the lane groups of one wave stay in step.  The pair list
// comes from the planner (nw_plan.cpp), which sends here only pairs whose
// score is the textbook global alignment of the two unpadded sequences -- the
// pairs where the reference's padding rules (sequential.cpp:40-66) reduce to
// "stop at the first padding code" on both sides.  The shorter sequence of a
// pair is the query; when the table is not symmetric the profile is built
// from its transpose for pairs whose query is the reference's column
// sequence, so every score keeps the reference's orientation (:87-91).
//
// A workgroup is NW waves; each wave holds 64/G lane groups and each group
// scores one target at a time as an anti-diagonal wavefront: lane k owns R
// consecutive rows of a G*R-row strip of the query, column residues and
// bottom-row scores move one lane per step by DPP, and a strip's bottom row
// goes to a per-group line in global memory for the next strip.  The query
// is padded with neutral rows at the top of its first strip, so the last row
// is always lane G-1, register R-1.  Workgroups take items from an atomic
// counter until the queue is empty (every wave reaches that exit).
#include "msa_internal.hpp"
#include "nw_device.hpp"

namespace msa {

namespace {

constexpr int kBatchCodes = 25;   // codes 0..23 and the neutral code 24
constexpr int kBatchNull = 24;

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

template <int G, int R, int NW>
__global__ __launch_bounds__(NW * 64) void nw_batched_kernel(BatchArgs a)
{
    constexpr int NG = NW * 64 / G;
    constexpr int W = R / 4;
    constexpr int STRIP = G * R;
    constexpr int SLOT = G * W;                 // words per code per strip
    extern __shared__ uint32_t prof[];          // [strip][code][lane][W]
    __shared__ uint8_t tab[640];
    __shared__ int s_item;

    const int tid = threadIdx.x;
    const int g = tid / G;
    const int k = tid % G;
    for (int i = tid; i < 640; i += NW * 64) tab[i] = a.sxx.v[i];
    int32_t* line = a.lines + ((int64_t)blockIdx.x * NG + g) * a.line_stride;
    const int p = a.penalty;

    while (true) {
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(a.queue, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= a.nitems) break;

        const int4 item = a.items[it];
        const SeqInfo qi = a.info[item.x];
        const int Lq = qi.z;
        const int nst = (Lq + STRIP - 1) / STRIP;
        const int pad_top = nst * STRIP - Lq;
        const bool transpose = (item.w & 1) != 0;
        const uint16_t* qseq = a.blocks + qi.x;

        // Profile: prof[st][b][lane][r] = sxx[query code][b] (or the
        // transpose), 0 for the neutral top rows and the neutral code.
        for (int wi = tid; wi < nst * G * W; wi += NW * 64) {
            const int rho = wi * 4;
            const int st = rho / STRIP;
            const int rem = rho - st * STRIP;
            const int kk = rem / R;
            const int q = (rem - kk * R) >> 2;
            int codes[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int qpos = rho + t - pad_top;
                const int code = qpos >= 0 ? dev::code_at(qseq, qpos) : -1;
                codes[t] = code < kBatchNull ? code : -1;
            }
            uint32_t* dst = prof + st * kBatchCodes * SLOT + kk * W + q;
            for (int b = 0; b < kBatchCodes; ++b) {
                uint32_t word = 0;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int code = codes[t];
                    uint32_t v = 0;
                    if (code >= 0 && b < kBatchNull) v = transpose ? tab[b * 25 + code] : tab[code * 25 + b];
                    word |= v << (8 * t);
                }
                dst[b * SLOT] = word;
            }
        }
        __syncthreads();

        const int tcount = item.z;
        const int rounds = (tcount + NG - 1) / NG;
        for (int m = 0; m < rounds; ++m) {
            const int t = g + m * NG;
            const bool active = t < tcount;
            int2 tg = make_int2(0, 0);
            SeqInfo ti = make_int4(0, 0, 0, 0);
            if (active) {
                tg = a.targets[item.y + t];
                ti = a.info[tg.x];
            }
            const int Lt = ti.z;
            const uint16_t* tseq = a.blocks + ti.x;
            const int my_steps = active ? Lt + G - 1 : 0;
            int wmax = my_steps;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off));

            int h[R];
            for (int st = 0; st < nst; ++st) {
                const uint32_t* pbase = prof + st * kBatchCodes * SLOT + k * W;
                const bool first = st == 0;
                const bool last = st == nst - 1;
#pragma unroll
                for (int r = 0; r < R; ++r) h[r] = 0;
                int dtop = 0;
                int b = kBatchNull;
                int colchunk = kBatchNull;
                int linechunk = 0;
                int outchunk = 0;

                for (int s = 0; s < wmax; ++s) {
                    if (s < my_steps) {
                        if ((s & (G - 1)) == 0) {
                            const int cc = s + k;
                            colchunk = cc < Lt ? min(dev::code_at(tseq, cc), kBatchNull) : kBatchNull;
                            if (!first) linechunk = cc < Lt ? line[cc] : 0;
                        }
                        b = dev::group_shr<G>(colchunk, b);
                        colchunk = dev::group_shl<G>(colchunk);
                        const int hup = dev::group_shr<G>(first ? 0 : linechunk, h[R - 1]);
                        if (!first) linechunk = dev::group_shl<G>(linechunk);

                        uint32_t w[W];
                        dev::load_profile<R>(w, pbase + b * SLOT);
                        dev::cells<R>(h, hup, dtop, w);
                        dtop = hup;

                        if (!last) {
                            outchunk = dev::group_shl<G>(outchunk);
                            if (k == G - 1) outchunk = h[R - 1];
                            const int col = s - (G - 1);
                            if (col >= 0 && (((col + 1) & (G - 1)) == 0 || col == Lt - 1)) {
                                const int dcol = col - (G - 1) + k;
                                if (dcol >= (col & ~(G - 1))) line[dcol] = outchunk;
                            }
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            }
            if (active && k == G - 1) a.out[tg.y] = h[R - 1] - (Lq + Lt) * p;
        }
    }
}

__global__ void queue_reset_kernel(int32_t* queue, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) queue[i] = 0;
}

template <int G, int R, int NW>
struct BatchedKernel {
    static const void* fn() { return reinterpret_cast<const void*>(&nw_batched_kernel<G, R, NW>); }
    static void launch(dim3 grid, size_t lds, hipStream_t stream, const BatchArgs& a)
    {
        hipLaunchKernelGGL((nw_batched_kernel<G, R, NW>), grid, dim3(NW * 64), lds, stream, a);
    }
};

constexpr int kBatchWaves = 4;

using LaunchFn = void (*)(dim3, size_t, hipStream_t, const BatchArgs&);
using FnPtr = const void* (*)();

struct Variant {
    int lanes, rows;
    LaunchFn launch;
    FnPtr fn;
};

#define MSA_VARIANT(G, R) {G, R, &BatchedKernel<G, R, kBatchWaves>::launch, &BatchedKernel<G, R, kBatchWaves>::fn}
const Variant kVariants[] = {
    MSA_VARIANT(16, 8),  MSA_VARIANT(16, 16), MSA_VARIANT(16, 32),
    MSA_VARIANT(32, 8),  MSA_VARIANT(32, 16), MSA_VARIANT(32, 32),
    MSA_VARIANT(64, 8),  MSA_VARIANT(64, 16), MSA_VARIANT(64, 32),
};
#undef MSA_VARIANT

const Variant* find_variant(int lanes, int rows)
{
    for (const auto& v : kVariants)
        if (v.lanes == lanes && v.rows == rows) return &v;
    return nullptr;
}

}  // namespace

bool batched_supported(int lanes, int rows) { return find_variant(lanes, rows) != nullptr; }

int batched_threads() { return kBatchWaves * 64; }

size_t batched_lds_bytes(int lanes, int rows, int strips)
{
    return (size_t)strips * kBatchCodes * lanes * rows;
}

int batched_groups_per_block(int lanes) { return kBatchWaves * 64 / lanes; }

int batched_occupancy(int lanes, int rows, size_t lds, int* blocks_per_cu)
{
    const Variant* v = find_variant(lanes, rows);
    if (!v) return fail(MSA_ERR_INVALID_ARGUMENT, "no batched kernel for %d lanes x %d rows", lanes, rows);
    MSA_HIP(hipFuncSetAttribute(v->fn(), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int n = 0;
    MSA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, v->fn(), kBatchWaves * 64, lds));
    *blocks_per_cu = n;
    return MSA_OK;
}

int batched_launch(int lanes, int rows, int grid, size_t lds, const uint16_t* d_blocks, const SeqInfo* d_info,
                   const int4* d_items, int32_t nitems, const int2* d_targets, int32_t* d_queue, int32_t* d_out,
                   int32_t* d_lines, int32_t line_stride, const Sxx& sxx, int32_t penalty, hipStream_t stream)
{
    const Variant* v = find_variant(lanes, rows);
    if (!v) return fail(MSA_ERR_INVALID_ARGUMENT, "no batched kernel for %d lanes x %d rows", lanes, rows);
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, stream, d_queue, 1);
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
    v->launch(dim3(grid), lds, stream, a);
    return hip_status(hipGetLastError(), "nw_batched_kernel launch");
}

}  // namespace msa
