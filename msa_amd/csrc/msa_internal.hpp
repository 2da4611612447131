// Internal declarations shared by the translation units of libmsa_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>

#include "msa_hip.h"

namespace msa {

// Records a printf-style detail for msa_last_error() and returns `status`.
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Maps a HIP error to a status code (recording the detail); MSA_OK on success.
int hip_status(hipError_t err, const char* what);

int table_lookup(const char* name, int32_t* table625, int32_t* penalty);

#define MSA_TRY(expr)                       \
    do {                                    \
        const int msa_rc_ = (expr);         \
        if (msa_rc_ != MSA_OK) return msa_rc_; \
    } while (0)

#define MSA_HIP(call) MSA_TRY(::msa::hip_status((call), #call))

// Residue codes (src/encoder.hpp:44-56).
constexpr int kEndCode = 0x18;   // encoder::end, the padding code
constexpr int kGapCode = 0x19;   // encoder::gap

// Per-sequence facts the kernels need, computed once per sequence store.
//   x: offset of the first block      y: number of blocks
//   z: first24 = position of the first padding code (3*y when none): the
//      number of rows the reference's DP runs when this sequence is the row
//      sequence (sequential.cpp:40-45)
//   w: lastreal = 1 + position of the last non-padding code (0 when none):
//      the column of the reference's result after the copy-left chain of
//      trailing padding columns (sequential.cpp:54-67)
// A sequence is "clean" when z == w: every padding code is trailing.
using SeqInfo = int4;

// Scoring table with the gap penalty folded in, as bytes: sxx[a*25+b] =
// T[a][b] + 2*penalty.  Scores are carried in the shifted domain
// G[i][j] = H[i][j] + (i+j)*penalty, where the recurrence of
// sequential.cpp:54-63 becomes G = max(G_diag + sxx, G_up, G_left) and every
// boundary cell is 0; H = G - (rows+cols)*penalty at the end.
struct Sxx {
    uint8_t v[640];
};

}  // namespace msa
