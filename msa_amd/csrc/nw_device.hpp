// Device-side building blocks shared by the Needleman-Wunsch kernels (gfx950,
// wave64).  The DP runs as an anti-diagonal wavefront across the lanes of a
// lane group: lane k owns R consecutive rows, the column residues and the
// bottom-row scores travel one lane per step through DPP lane shifts, and the
// substitution scores of a lane's rows come from a per-row-strip profile in
// LDS, R bytes per (column code, lane).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msa {
namespace dev {

// encoder::access (src/encoder.hpp:54-70): code o of a sequence sits in block
// o/3 at bit shift 1 + 5*(o%3).
__device__ __forceinline__ int code_at(const uint16_t* __restrict__ seq, int pos)
{
    const int blk = pos / 3;
    const int slot = pos - 3 * blk;
    return (seq[blk] >> (1 + 5 * slot)) & 0x1F;
}

// DPP controls (GFX9 encoding).
constexpr int kRowShl1 = 0x101;
constexpr int kRowShr1 = 0x111;
constexpr int kWaveShl1 = 0x130;
constexpr int kWaveShr1 = 0x138;
constexpr int kRowBcast15 = 0x142;
constexpr int kRowNewBcast0 = 0x150;   // row_newbcast:n = 0x150 + n (gfx90a and later)

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, CTRL, ROW_MASK, 0xF, false);
}

// DPP move whose lanes without a source lane read 0.
template <int CTRL>
__device__ __forceinline__ int dpp_zero(int src)
{
    return __builtin_amdgcn_mov_dpp(src, CTRL, 0xF, 0xF, true);
}

// Words in LDS, addressed with 32-bit offsets.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x2 lds_u32x2;
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;

// Lane k of a G-lane group receives lane k-1's `src`; the first lane of each
// group receives its own `first`.
template <int G>
__device__ __forceinline__ int group_shr(int first, int src)
{
    if constexpr (G == 16) {
        return dpp<kRowShr1>(first, src);
    } else if constexpr (G == 64) {
        return dpp<kWaveShr1>(first, src);
    } else {
        static_assert(G == 32, "lane groups of 16, 32 or 64");
        // Rows 1 and 3 (lanes 16..31, 48..63) first take lane 15 / 47, so that
        // the in-row shift leaves lane 16 / 48 holding its upper neighbour.
        const int carry = dpp<kRowBcast15, 0xA>(first, src);
        return dpp<kRowShr1>(carry, src);
    }
}

// Lane k of a group receives lane k+1's `src`; what the last lane of a group
// receives is unspecified (callers never read it before a refill).
template <int G>
__device__ __forceinline__ int group_shl(int src)
{
    if constexpr (G == 16) return dpp<kRowShl1>(src, src);
    else return dpp<kWaveShl1>(src, src);
}

// Profile words holding R rows, four rows per word.
template <int R>
constexpr int prof_words = (R + 3) / 4;

// One DP column segment: R cells of one lane.  h[] holds the lane's rows at
// the previous column (left neighbours) and receives this column; `up` is the
// cell above the first row (this column), `diag` the cell above-left.  w[]
// holds the R profile bytes sxx[row][column code], row r in byte r%4 of word
// r/4.  In the shifted domain every cell is add + max3.
template <int R>
__device__ __forceinline__ void cells(int (&h)[R], int up, int diag, const uint32_t (&w)[prof_words<R>])
{
    int d = diag;
    int u = up;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int s = (int)((w[r >> 2] >> (8 * (r & 3))) & 0xFFu);
        const int nh = max(max(d + s, u), h[r]);
        d = h[r];
        h[r] = nh;
        u = nh;
    }
}

// As cells(), plus the reference's padding columns: a column whose code is
// the padding code copies its left neighbour (sequential.cpp:57-66), which in
// the shifted domain is left + penalty.
template <int R>
__device__ __forceinline__ void cells_pads(int (&h)[R], int up, int diag, const uint32_t (&w)[R / 4],
                                           bool pad, int penalty)
{
    int d = diag;
    int u = up;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int s = (int)((w[r >> 2] >> (8 * (r & 3))) & 0xFFu);
        int nh = max(max(d + s, u), h[r]);
        nh = pad ? h[r] + penalty : nh;
        d = h[r];
        h[r] = nh;
        u = nh;
    }
}

// Loads the R profile bytes of one (code, lane) slot from LDS.
template <int R>
__device__ __forceinline__ void load_profile(uint32_t (&w)[R / 4], const uint32_t* __restrict__ slot)
{
    static_assert(R % 4 == 0, "rows per lane must be a multiple of 4");
    if constexpr (R % 16 == 0) {
#pragma unroll
        for (int q = 0; q < R / 16; ++q) {
            const uint4 v = reinterpret_cast<const uint4*>(slot)[q];
            w[4 * q + 0] = v.x;
            w[4 * q + 1] = v.y;
            w[4 * q + 2] = v.z;
            w[4 * q + 3] = v.w;
        }
    } else if constexpr (R == 8) {
        const uint2 v = *reinterpret_cast<const uint2*>(slot);
        w[0] = v.x;
        w[1] = v.y;
    } else {
#pragma unroll
        for (int q = 0; q < R / 4; ++q) w[q] = slot[q];
    }
}

}  // namespace dev
}  // namespace msa
