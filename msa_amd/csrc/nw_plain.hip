// The plain Needleman-Wunsch kernel: one pair per workgroup at a time.
//
// It is the GPU form of the reference's align_kernel (src/pairwise/needleman/
// impl/hybrid.cu:222-252) and keeps every rule of the sequential path
// (sequential.cpp:29-95) exactly, so it is also the kernel that scores the
// pairs the batched kernel does not take:
//   - orientation: rows = the first sequence of the pair when it has more
//     blocks, else the second (sequential.cpp:87-91);
//   - rows stop at the first padding code of the row sequence (:40-45);
//   - a padding code in the column sequence copies the left cell (:57-66);
//   - row 0 is j*-penalty over the padded width, the result is the cell at
//     the padded last column (:33-35, :68).
// A workgroup is one wave64.  Lane k owns R rows of a 64R-row strip; the
// column residues stream through the lanes (anti-diagonal wavefront) and the
// strip's bottom row goes to a per-workgroup line in global memory for the
// next strip, like the reference's column cache (hybrid.cu:292-296).
#include "msa_internal.hpp"
#include "nw_device.hpp"

namespace msa {

namespace {

constexpr int kPlainLanes = 64;
constexpr int kPlainRows = 8;
constexpr int kPlainCodes = 32;   // profile slots, one per 5-bit code
constexpr int kPlainNull = 31;    // code streamed outside the column range

struct PlainArgs {
    const uint16_t* blocks;
    const SeqInfo* info;
    const int32_t* pairs;       // 2 per pair, or null = all pairs in reference order
    const int32_t* out_index;   // destination of pair c, or null = c
    int64_t npairs;
    int32_t* out;
    int32_t* lines;             // gridDim.x lines of line_stride ints
    int32_t line_stride;
    int32_t penalty;
    Sxx sxx;
};

// Pair c of the reference order i = 1..n-1, j < i (pairwise.cpp:79-81).
__device__ __forceinline__ void pair_of(int64_t c, int& i, int& j)
{
    int64_t r = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)c)) * 0.5);
    while (r * (r - 1) / 2 > c) --r;
    while ((r + 1) * r / 2 <= c) ++r;
    i = (int)r;
    j = (int)(c - r * (r - 1) / 2);
}

template <int R>
__global__ __launch_bounds__(kPlainLanes) void nw_plain_kernel(PlainArgs a)
{
    constexpr int G = kPlainLanes;
    constexpr int W = R / 4;
    __shared__ __attribute__((aligned(16))) uint32_t prof[kPlainCodes * G * W];   // [code][lane][W words]
    __shared__ uint8_t tab[640];

    const int lane = threadIdx.x;
    for (int i = lane; i < 640; i += G) tab[i] = a.sxx.v[i];
    __syncthreads();

    int32_t* line = a.lines + (int64_t)blockIdx.x * a.line_stride;
    const int p = a.penalty;

    for (int64_t c = blockIdx.x; c < a.npairs; c += gridDim.x) {
        int first, second;
        if (a.pairs) {
            first = a.pairs[2 * c];
            second = a.pairs[2 * c + 1];
        } else {
            pair_of(c, first, second);
        }
        const SeqInfo fi = a.info[first];
        const SeqInfo se = a.info[second];
        const bool first_rows = fi.y > se.y;
        const SeqInfo ri = first_rows ? fi : se;
        const SeqInfo ci = first_rows ? se : fi;
        const uint16_t* rowseq = a.blocks + ri.x;
        const uint16_t* colseq = a.blocks + ci.x;
        const int La = ri.z;            // rows the reference runs
        const int Lb = ci.w;            // columns up to the last non-padding code
        const bool pads = ci.z < ci.w;  // a padding code before the last residue
        const int64_t dst = a.out_index ? a.out_index[c] : c;

        if (La == 0) {
            // No row runs: the result is row 0 at the padded last column.
            if (lane == 0) a.out[dst] = -p * 3 * ci.y;
            continue;
        }

        const int strip_rows = G * R;
        const int nstrips = (La + strip_rows - 1) / strip_rows;
        int g_final = 0;

        for (int st = 0; st < nstrips; ++st) {
            const int row0 = st * strip_rows;
            const int rows_here = min(strip_rows, La - row0);
            const int k_last = (rows_here - 1) / R;
            const int r_last = (rows_here - 1) % R;
            const bool first_strip = st == 0;
            const bool last_strip = st == nstrips - 1;

            // Profile of this strip: prof[b][lane][r] = sxx[row code][b].
            __syncthreads();
            {
                int codes[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int row = row0 + lane * R + r;
                    codes[r] = row < La ? dev::code_at(rowseq, row) : -1;
                }
                for (int b = 0; b < kPlainCodes; ++b) {
#pragma unroll
                    for (int q = 0; q < W; ++q) {
                        uint32_t word = 0;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int code = codes[4 * q + t];
                            const uint32_t v = (code >= 0 && code < 25 && b < 25) ? tab[code * 25 + b] : 0u;
                            word |= v << (8 * t);
                        }
                        prof[(b * G + lane) * W + q] = word;
                    }
                }
            }
            __syncthreads();

            int h[R];
#pragma unroll
            for (int r = 0; r < R; ++r) h[r] = 0;
            int dtop = 0;
            int b = kPlainNull;
            int colchunk = kPlainNull;
            int linechunk = 0;
            int outchunk = 0;
            const int steps = Lb + (last_strip ? k_last : G - 1);

            for (int s = 0; s < steps; ++s) {
                if ((s & (G - 1)) == 0) {
                    const int cc = s + lane;
                    colchunk = cc < Lb ? dev::code_at(colseq, cc) : kPlainNull;
                    if (!first_strip) linechunk = cc < Lb ? line[cc] : 0;
                }
                b = dev::group_shr<G>(colchunk, b);
                colchunk = dev::group_shl<G>(colchunk);
                const int hup = dev::group_shr<G>(first_strip ? 0 : linechunk, h[R - 1]);
                if (!first_strip) linechunk = dev::group_shl<G>(linechunk);

                uint32_t w[W];
                dev::load_profile<R>(w, &prof[(b * G + lane) * W]);
                if (pads)
                    dev::cells_pads<R>(h, hup, dtop, w, b == kEndCode, p);
                else
                    dev::cells<R>(h, hup, dtop, w);
                dtop = hup;

                if (!last_strip) {
                    // Lane G-1 finishes column s-(G-1); collect the strip's
                    // bottom row G columns at a time and store it coalesced.
                    outchunk = dev::group_shl<G>(outchunk);
                    if (lane == G - 1) outchunk = h[R - 1];
                    const int col = s - (G - 1);
                    if (col >= 0 && (((col + 1) & (G - 1)) == 0 || col == Lb - 1)) {
                        const int dcol = col - (G - 1) + lane;
                        if (dcol >= (col & ~(G - 1))) line[dcol] = outchunk;
                    }
                }
            }

            if (last_strip) {
                int g = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) g = (r == r_last) ? h[r] : g;
                g_final = __builtin_amdgcn_readlane(g, k_last);
            }
            // The next strip reads this strip's bottom row from `line`.
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }

        if (lane == 0) a.out[dst] = g_final - (La + Lb) * p;
    }
}

__global__ void seq_info_kernel(const uint16_t* __restrict__ blocks, const int64_t* __restrict__ offset,
                                const int32_t* __restrict__ size, int32_t nseq, SeqInfo* __restrict__ info)
{
    // One wave per sequence: first padding position and last non-padding
    // position over the padded length.
    const int s = blockIdx.x;
    if (s >= nseq) return;
    const int lane = threadIdx.x;
    const uint16_t* seq = blocks + offset[s];
    const int len = 3 * size[s];
    int first_end = len;
    int last_real = 0;
    for (int pos = lane; pos < len; pos += 64) {
        const int code = dev::code_at(seq, pos);
        if (code == kEndCode) first_end = min(first_end, pos);
        else last_real = max(last_real, pos + 1);
    }
    for (int off = 32; off >= 1; off >>= 1) {
        first_end = min(first_end, __shfl_xor(first_end, off));
        last_real = max(last_real, __shfl_xor(last_real, off));
    }
    if (lane == 0) info[s] = make_int4((int)offset[s], size[s], first_end, last_real);
}

}  // namespace

int plain_launch(const uint16_t* d_blocks, const SeqInfo* d_info, const int32_t* d_pairs,
                 const int32_t* d_out_index, int64_t npairs, const Sxx& sxx, int32_t penalty,
                 int32_t* d_out, int32_t* d_lines, int32_t line_stride, int grid, hipStream_t stream)
{
    if (npairs <= 0) return MSA_OK;
    PlainArgs a;
    a.blocks = d_blocks;
    a.info = d_info;
    a.pairs = d_pairs;
    a.out_index = d_out_index;
    a.npairs = npairs;
    a.out = d_out;
    a.lines = d_lines;
    a.line_stride = line_stride;
    a.penalty = penalty;
    a.sxx = sxx;
    hipLaunchKernelGGL(nw_plain_kernel<kPlainRows>, dim3(grid), dim3(kPlainLanes), 0, stream, a);
    return hip_status(hipGetLastError(), "nw_plain_kernel launch");
}

int plain_grid(int64_t npairs, int compute_units)
{
    const int64_t cap = (int64_t)compute_units * 16;
    return (int)(npairs < cap ? npairs : cap);
}

int seq_info_launch(const uint16_t* d_blocks, const int64_t* d_offset, const int32_t* d_size, int32_t nseq,
                    SeqInfo* d_info, hipStream_t stream)
{
    if (nseq <= 0) return MSA_OK;
    hipLaunchKernelGGL(seq_info_kernel, dim3(nseq), dim3(64), 0, stream, d_blocks, d_offset, d_size, nseq, d_info);
    return hip_status(hipGetLastError(), "seq_info_kernel launch");
}

}  // namespace msa
