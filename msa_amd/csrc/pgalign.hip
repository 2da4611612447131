// The profile-alignment stage (pgalign) on the GPU: progressive alignment of
// a sequence database along the neighbor-joining guide tree.
//
// The reference's stage is a stub -- its "myers" algorithm returns an empty
// alignment (src/pgalign/myers/impl/sequential.cpp:29-32) -- so what is built
// here follows the objective its pipeline implies: every join of the guide
// tree, children before parents, aligns the two children's alignments
// (profiles) optimally under the sum-of-pairs score with the pairwise stage's
// table T and linear gap penalty p (sigma(x, y) = T[x][y] for residues,
// sigma(residue, gap) = sigma(gap, residue) = -p, sigma(gap, gap) = 0).  For two
// single sequences that is exactly the pairwise stage's Needleman-Wunsch.
// Rows hold the first unpadded() codes of each sequence (src/sequence.hpp:
// 110-119); the output gives the column of every residue, ascending, like the
// reference's gapped sequence (src/pgalign/sequence.cuh, m_columns).
//
// Device layout: the residues of all sequences in guide-tree leaf order (the
// members of every tree node are one contiguous range), one byte of code and
// one int32 column each.  Joins of one tree level are independent; a level is
// run in batches bounded by a cell budget, each batch as six launches:
//   pg_hist   column code counts of both profiles (atomics in global memory)
//   pg_prep   per column of A: V[i][y] = sum_x cA[i][x] sigma(x, y) and the
//             gap-column cost dA[i] = -p rA[i] nB; per column of B: its gap
//             count and dB[j] = -p rB[j] nA
//   pg_score  S[i][j] = sum_y V[i][y] cB[j][y] (int32; the planner refuses
//             profiles whose bound nA nB max|sigma| would not fit)
//   pg_dp     one wave per join: rows in order, columns across the lanes
//             (lane k, slice s: column 1 + 64 s + k); each row is
//             E = max(diag + S, up + dA) then H = P + prefix-max(E - P) with P
//             the prefix sums of dB (the left moves), a 64-bit max-scan across
//             the wave; one direction byte per cell goes to HBM
//   pg_trace  one lane per join walks the directions back from (LA, LB)
//   pg_remap  every member residue of A and B moves to its new column
// Ties resolve diagonal, then up, then left.  The traceback of a join is kept
// whole in HBM (one byte per cell: a 2,000 x 2,000 join holds 4 MB), where
// Myers-Miller's linear-space recursion would recompute half the matrix per
// level on the critical path of a tree that is often a chain.
#include "msa_internal.hpp"
#include "nw_planner.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace msa {

namespace {

constexpr int kCodes = 26;      // 25 residue codes and the gap
constexpr int kGap = 25;
constexpr int kLanes = 64;

struct Join {
    int64_t ra, rb;             // first residue of A and of B (leaf order)
    int64_t na_res, nb_res;     // residues of A and of B
    int32_t nA, nB;             // members
    int32_t LA, LB;             // columns
    int64_t ca, cb;             // offsets (ints) of cA [LA x 26] and cB [LB x 26]
    int64_t v;                  // offset (ints) of V [LA x 26]
    int64_t da, db;             // offsets (int64) of dA [LA] and dB [LB]
    int64_t s;                  // offset (ints) of S [LA x LB]
    int64_t dir;                // offset (bytes) of the directions [LA x LB]
    int64_t row;                // offset (int64) of the row and prefix buffers [2 (LB + 1)]
    int64_t map;                // offset (ints) of mapA [LA] then mapB [LB]
};

__global__ void pg_zero_kernel(int32_t* p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void pg_hist_kernel(const Join* joins, const uint8_t* codes, const int32_t* col, int32_t* ints)
{
    const Join j = joins[blockIdx.y];
    const int64_t total = j.na_res + j.nb_res;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < total; r += (int64_t)gridDim.x * blockDim.x) {
        const bool a = r < j.na_res;
        const int64_t at = a ? j.ra + r : j.rb + (r - j.na_res);
        int32_t* counts = ints + (a ? j.ca : j.cb);
        atomicAdd(counts + (int64_t)col[at] * kCodes + codes[at], 1);
    }
}

__global__ void pg_prep_kernel(const Join* joins, const int32_t* sigma, int32_t penalty, int32_t* ints, int64_t* wide)
{
    const Join j = joins[blockIdx.y];
    __shared__ int32_t sg[kCodes * kCodes];
    for (int i = threadIdx.x; i < kCodes * kCodes; i += blockDim.x) sg[i] = sigma[i];
    __syncthreads();
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < (int64_t)j.LA + j.LB;
         c += (int64_t)gridDim.x * blockDim.x) {
        if (c < j.LA) {
            int32_t* cnt = ints + j.ca + c * kCodes;
            int32_t r = 0;
            for (int x = 0; x < kGap; ++x) r += cnt[x];
            cnt[kGap] = j.nA - r;
            int32_t* v = ints + j.v + c * kCodes;
            for (int y = 0; y < kCodes; ++y) {
                int32_t acc = 0;
                for (int x = 0; x < kCodes; ++x) acc += cnt[x] * sg[x * kCodes + y];
                v[y] = acc;
            }
            wide[j.da + c] = -(int64_t)penalty * r * j.nB;
        } else {
            const int64_t jj = c - j.LA;
            int32_t* cnt = ints + j.cb + jj * kCodes;
            int32_t r = 0;
            for (int y = 0; y < kGap; ++y) r += cnt[y];
            cnt[kGap] = j.nB - r;
            wide[j.db + jj] = -(int64_t)penalty * r * j.nA;
        }
    }
}

__global__ void pg_score_kernel(const Join* joins, int32_t* ints)
{
    const Join j = joins[blockIdx.y];
    const int64_t cells = (int64_t)j.LA * j.LB;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < cells; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = c / j.LB, jj = c - i * j.LB;
        const int32_t* v = ints + j.v + i * kCodes;
        const int32_t* cb = ints + j.cb + jj * kCodes;
        int32_t acc = 0;
#pragma unroll
        for (int y = 0; y < kCodes; ++y) acc += v[y] * cb[y];
        ints[j.s + c] = acc;
    }
}

__device__ __forceinline__ int64_t shfl_up64(int64_t v, int d)
{
    const int lo = __shfl_up((int)(uint32_t)v, d);
    const int hi = __shfl_up((int)(uint32_t)((uint64_t)v >> 32), d);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ int64_t shfl64(int64_t v, int lane)
{
    const int lo = __shfl((int)(uint32_t)v, lane);
    const int hi = __shfl((int)(uint32_t)((uint64_t)v >> 32), lane);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Inclusive max-scan over the 64 lanes.
__device__ __forceinline__ int64_t wave_max_scan(int64_t v, int lane)
{
#pragma unroll
    for (int d = 1; d < kLanes; d <<= 1) {
        const int64_t o = shfl_up64(v, d);
        if (lane >= d) v = max(v, o);
    }
    return v;
}

// Inclusive sum-scan over the 64 lanes.
__device__ __forceinline__ int64_t wave_sum_scan(int64_t v, int lane)
{
#pragma unroll
    for (int d = 1; d < kLanes; d <<= 1) {
        const int64_t o = shfl_up64(v, d);
        if (lane >= d) v += o;
    }
    return v;
}

constexpr int64_t kFloor = INT64_MIN / 4;   // below every score, with room to add to it

// One wave per join.  row[0..LB] holds H of the previous row, row[LB+1 ..
// 2LB+1] the prefix sums P[0..LB] of dB.  Directions: 0 diagonal, 1 up, 2 left.
__global__ __launch_bounds__(kLanes) void pg_dp_kernel(const Join* joins, const int32_t* ints, const int64_t* wide_in,
                                                       int64_t* wide, uint8_t* dirs, int64_t* scores)
{
    const Join j = joins[blockIdx.x];
    const int lane = threadIdx.x;
    int64_t* row = wide + j.row;
    int64_t* P = row + j.LB + 1;
    const int64_t* dA = wide_in + j.da;
    const int64_t* dB = wide_in + j.db;
    const int32_t* S = ints + j.s;
    uint8_t* dir = dirs + j.dir;
    const int slices = (j.LB + kLanes - 1) / kLanes;
    // Row 0: H[0][j] = P[j].
    int64_t carry = 0;
    if (lane == 0) {
        row[0] = 0;
        P[0] = 0;
    }
    for (int s = 0; s < slices; ++s) {
        const int c = 1 + s * kLanes + lane;
        const int64_t d = c <= j.LB ? dB[c - 1] : 0;
        const int64_t p = carry + wave_sum_scan(d, lane);
        if (c <= j.LB) {
            P[c] = p;
            row[c] = p;
        }
        carry = shfl64(p, kLanes - 1);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    int64_t h0 = 0;     // H[i-1][0]
    for (int i = 1; i <= j.LA; ++i) {
        const int64_t da = dA[i - 1];
        const int64_t h0n = h0 + da;        // H[i][0]
        int64_t run = h0n;                  // max over k < current column of (E_k - P_k), with E_0 - P_0 = H[i][0]
        int64_t prev_last = h0;             // H[i-1][column before the slice's first]
        for (int s = 0; s < slices; ++s) {
            const int c = 1 + s * kLanes + lane;
            const bool in = c <= j.LB;
            const int64_t old = in ? row[c] : 0;                    // H[i-1][c]
            const int64_t up_lane = shfl_up64(old, 1);
            const int64_t diag = lane == 0 ? prev_last : up_lane;   // H[i-1][c-1]
            prev_last = shfl64(old, kLanes - 1);
            int64_t q = kFloor;
            uint8_t de = 0;
            int64_t pc = 0;
            if (in) {
                const int64_t ed = diag + S[(int64_t)(i - 1) * j.LB + (c - 1)];
                const int64_t eu = old + da;
                const int64_t e = ed >= eu ? ed : eu;
                de = ed >= eu ? 0 : 1;
                pc = P[c];
                q = e - pc;
            }
            const int64_t incl = max(wave_max_scan(q, lane), run);
            int64_t excl = shfl_up64(incl, 1);
            if (lane == 0) excl = run;
            if (in) {
                const int64_t h = pc + (q >= excl ? q : excl);
                dir[(int64_t)(i - 1) * j.LB + (c - 1)] = q >= excl ? de : 2;
                // Every lane read its H[i-1][c] above (and the next lane's
                // diagonal came through the shuffle), so the row is updated
                // in place.
                row[c] = h;
            }
            run = shfl64(incl, kLanes - 1);
        }
        h0 = h0n;
        if (lane == 0) row[0] = h0n;
    }
    if (lane == 0) scores[blockIdx.x] = j.LB > 0 ? row[j.LB] : (j.LA > 0 ? h0 : 0);
}

__global__ void pg_trace_kernel(const Join* joins, int32_t njoins, const uint8_t* dirs, int32_t* ints, int32_t* newlen)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= njoins) return;
    const Join j = joins[t];
    const uint8_t* dir = dirs + j.dir;
    int32_t* mapA = ints + j.map;
    int32_t* mapB = mapA + j.LA;
    int64_t i = j.LA, c = j.LB, steps = 0;
    while (i > 0 || c > 0) {
        const uint8_t d = i == 0 ? 2 : (c == 0 ? 1 : dir[(i - 1) * j.LB + (c - 1)]);
        if (d == 0) {
            mapA[--i] = (int32_t)steps;
            mapB[--c] = (int32_t)steps;
        } else if (d == 1) {
            mapA[--i] = (int32_t)steps;
        } else {
            mapB[--c] = (int32_t)steps;
        }
        ++steps;
    }
    newlen[t] = (int32_t)steps;
}

__global__ void pg_remap_kernel(const Join* joins, const int32_t* ints, const int32_t* newlen, int32_t* col)
{
    const Join j = joins[blockIdx.y];
    const int32_t n = newlen[blockIdx.y];
    const int32_t* mapA = ints + j.map;
    const int32_t* mapB = mapA + j.LA;
    const int64_t total = j.na_res + j.nb_res;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < total; r += (int64_t)gridDim.x * blockDim.x) {
        const bool a = r < j.na_res;
        const int64_t at = a ? j.ra + r : j.rb + (r - j.na_res);
        col[at] = n - 1 - (a ? mapA : mapB)[col[at]];
    }
}

// Device buffers of one run, grown on demand.
struct Buffer {
    void* p = nullptr;
    size_t bytes = 0;
    ~Buffer() { if (p) (void)hipFree(p); }
    int need(size_t b)
    {
        if (b <= bytes) return MSA_OK;
        if (p) MSA_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
        MSA_HIP(hipMalloc(&p, std::max<size_t>(b, 256)));
        bytes = std::max<size_t>(b, 256);
        return MSA_OK;
    }
};

int grid_for(int64_t items, int threads)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(4096, (items + threads - 1) / threads));
}

}  // namespace

// sequence::unpadded (src/sequence.hpp:110-119): the padded length minus the
// padding codes at positions 1 and 2 of the last block.
int64_t pg_unpadded(const uint16_t* seq, int32_t blocks)
{
    if (blocks <= 0) return 0;
    int64_t length = 3 * (int64_t)blocks;
    for (int k = 1; k < 3; ++k) length -= ((seq[blocks - 1] >> (1 + 5 * k)) & 0x1F) == kEndCode;
    return length;
}

}  // namespace msa

using msa::fail;

extern "C" int msa_pg_residues(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                               int64_t* residues)
{
    if (!residues || nseq < 0 || (nseq > 0 && (!offset || !size)))
        return fail(MSA_ERR_INVALID_ARGUMENT, "invalid arguments to msa_pg_residues");
    int64_t total = 0;
    for (int32_t s = 0; s < nseq; ++s) {
        if (size[s] < 0 || offset[s] < 0) return fail(MSA_ERR_INVALID_ARGUMENT, "sequence %d has a negative offset or size", s);
        if (size[s] > 0 && !blocks) return fail(MSA_ERR_INVALID_ARGUMENT, "null blocks");
        total += msa::pg_unpadded(blocks + offset[s], size[s]);
    }
    *residues = total;
    return MSA_OK;
}

extern "C" int msa_pg_align(const uint16_t* blocks, const int64_t* offset, const int32_t* size, int32_t nseq,
                            const msa_nj_node* nodes, const int32_t* table625, int32_t penalty, int32_t* columns,
                            int64_t* length, int64_t* join_scores)
{
    using namespace msa;
    if (!length || nseq < 0 || !table625 || (nseq > 0 && (!offset || !size || !nodes)))
        return fail(MSA_ERR_INVALID_ARGUMENT, "invalid arguments to msa_pg_align");
    for (int32_t s = 0; s < nseq; ++s)
        if (size[s] > 0 && !blocks) return fail(MSA_ERR_INVALID_ARGUMENT, "null blocks");
    *length = 0;
    if (nseq == 0) return MSA_OK;
    if (nseq > (1 << 24)) return fail(MSA_ERR_TOO_LARGE, "msa_pg_align takes at most 2^24 sequences");
    // Residues per sequence, in database order; gap codes are refused as in
    // the pairwise stage.
    std::vector<int64_t> len(nseq), out_at(nseq);
    int64_t total = 0;
    for (int32_t s = 0; s < nseq; ++s) {
        if (size[s] < 0 || offset[s] < 0) return fail(MSA_ERR_INVALID_ARGUMENT, "sequence %d has a negative offset or size", s);
        len[s] = pg_unpadded(blocks + offset[s], size[s]);
        out_at[s] = total;
        total += len[s];
        for (int64_t r = 0; r < len[s]; ++r) {
            const int code = (blocks[offset[s] + r / 3] >> (1 + 5 * (r % 3))) & 0x1F;
            if (code > kEndCode)
                return fail(MSA_ERR_GAP_CODE, "sequence %d holds code %d at position %lld", s, code, (long long)r);
        }
        if (len[s] >= (int64_t(1) << 30)) return fail(MSA_ERR_TOO_LARGE, "sequence %d is too long", s);
    }
    if (total > 0 && !columns) return fail(MSA_ERR_INVALID_ARGUMENT, "columns is null");
    // The tree: children smaller than their node; leaf order by a post-order
    // walk, so that every node's members are one contiguous range.
    const int32_t nn = 2 * nseq - 1;
    std::vector<int32_t> lo(nn), hi(nn), order, level(nn, 0);
    order.reserve(nseq);
    {
        std::vector<int32_t> stack;
        stack.push_back(nn - 1);
        std::vector<char> seen(nn, 0);
        while (!stack.empty()) {
            const int32_t v = stack.back();
            if (v < nseq) {
                stack.pop_back();
                if (seen[v]) return fail(MSA_ERR_INVALID_ARGUMENT, "guide tree reaches leaf %d twice", v);
                seen[v] = 1;
                lo[v] = (int32_t)order.size();
                order.push_back(v);
                hi[v] = (int32_t)order.size();
                continue;
            }
            const uint32_t c0 = nodes[v].child[0], c1 = nodes[v].child[1];
            if (c0 >= (uint32_t)v || c1 >= (uint32_t)v || c0 == c1)
                return fail(MSA_ERR_INVALID_ARGUMENT, "guide tree node %d has children %u and %u", v, c0, c1);
            if (!seen[v]) {
                seen[v] = 1;
                stack.push_back((int32_t)c1);
                stack.push_back((int32_t)c0);
            } else {
                stack.pop_back();
                lo[v] = lo[c0];
                hi[v] = hi[c1];
                level[v] = 1 + std::max(level[c0], level[c1]);
                if (hi[c0] != lo[c1]) return fail(MSA_ERR_INVALID_ARGUMENT, "guide tree is not a tree");
            }
            if (stack.size() > (size_t)4 * nn) return fail(MSA_ERR_INVALID_ARGUMENT, "guide tree has a cycle");
        }
        if ((int32_t)order.size() != nseq) return fail(MSA_ERR_INVALID_ARGUMENT, "guide tree does not hold every leaf");
    }
    std::vector<int64_t> res_at(nseq + 1, 0);
    for (int32_t k = 0; k < nseq; ++k) res_at[k + 1] = res_at[k] + len[order[k]];
    std::vector<uint8_t> h_codes((size_t)std::max<int64_t>(total, 1));
    std::vector<int32_t> h_col((size_t)std::max<int64_t>(total, 1));
    for (int32_t k = 0; k < nseq; ++k) {
        const int32_t s = order[k];
        for (int64_t r = 0; r < len[s]; ++r) {
            h_codes[res_at[k] + r] = (uint8_t)((blocks[offset[s] + r / 3] >> (1 + 5 * (r % 3))) & 0x1F);
            h_col[res_at[k] + r] = (int32_t)r;
        }
    }
    // sigma over 25 codes and the gap; the bound of S must fit int32.
    int32_t sigma[kCodes * kCodes];
    int64_t smax = 0;
    for (int x = 0; x < kCodes; ++x)
        for (int y = 0; y < kCodes; ++y) {
            const int32_t v = (x < kGap && y < kGap) ? table625[x * 25 + y] : ((x == kGap && y == kGap) ? 0 : -penalty);
            sigma[x * kCodes + y] = v;
            smax = std::max<int64_t>(smax, v < 0 ? -(int64_t)v : v);
        }
    std::vector<int64_t> plen(nn, 0);
    for (int32_t s = 0; s < nseq; ++s) plen[s] = len[s];
    int maxlevel = 0;
    for (int32_t v = nseq; v < nn; ++v) maxlevel = std::max(maxlevel, level[v]);
    std::vector<std::vector<int32_t>> by_level(maxlevel + 1);
    for (int32_t v = nseq; v < nn; ++v) by_level[level[v]].push_back(v);

    Buffer d_codes, d_col, d_sigma, d_joins, d_ints, d_wide, d_wide_in, d_dirs, d_scores, d_newlen;
    MSA_TRY(d_codes.need((size_t)std::max<int64_t>(total, 1)));
    MSA_TRY(d_col.need(sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1)));
    MSA_TRY(d_sigma.need(sizeof(sigma)));
    MSA_HIP(hipMemcpy(d_codes.p, h_codes.data(), (size_t)std::max<int64_t>(total, 1), hipMemcpyHostToDevice));
    MSA_HIP(hipMemcpy(d_col.p, h_col.data(), sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1),
                      hipMemcpyHostToDevice));
    MSA_HIP(hipMemcpy(d_sigma.p, sigma, sizeof(sigma), hipMemcpyHostToDevice));
    // Cells (S ints and direction bytes) per batch.
    const int64_t kBudget = int64_t(1) << 28;
    for (int lv = 1; lv <= maxlevel; ++lv) {
        const std::vector<int32_t>& js = by_level[lv];
        size_t first = 0;
        while (first < js.size()) {
            // One batch: joins until the cell budget (at least one join).
            std::vector<Join> batch;
            std::vector<int32_t> who;
            int64_t ints = 0, wide = 0, wide_in = 0, dbytes = 0, cells = 0;
            size_t k = first;
            for (; k < js.size(); ++k) {
                const int32_t v = js[k];
                const uint32_t a = nodes[v].child[0], b = nodes[v].child[1];
                Join jn{};
                jn.ra = res_at[lo[a]];
                jn.na_res = res_at[hi[a]] - jn.ra;
                jn.rb = res_at[lo[b]];
                jn.nb_res = res_at[hi[b]] - jn.rb;
                jn.nA = hi[a] - lo[a];
                jn.nB = hi[b] - lo[b];
                jn.LA = (int32_t)plen[a];
                jn.LB = (int32_t)plen[b];
                const int64_t c = (int64_t)jn.LA * jn.LB;
                // The join index travels in blockIdx.y (at most 65535).
                if (!batch.empty() && (cells + c > kBudget || batch.size() >= 65535)) break;
                if ((int64_t)jn.nA * jn.nB * smax * kCodes >= (int64_t(1) << 31))
                    return fail(MSA_ERR_TOO_LARGE, "join of %d and %d sequences: sum-of-pairs column scores exceed "
                                "32 bits", jn.nA, jn.nB);
                if (plen[a] + plen[b] >= (int64_t(1) << 31))
                    return fail(MSA_ERR_TOO_LARGE, "alignment would exceed 2^31 columns");
                auto up64 = [](int64_t x) { return (x + 63) & ~int64_t(63); };
                jn.ca = ints;
                ints += up64((int64_t)jn.LA * kCodes);
                jn.cb = ints;
                ints += up64((int64_t)jn.LB * kCodes);
                jn.v = ints;
                ints += up64((int64_t)jn.LA * kCodes);
                jn.s = ints;
                ints += up64(c);
                jn.map = ints;
                ints += up64((int64_t)jn.LA + jn.LB);
                jn.da = wide_in;
                wide_in += up64(jn.LA);
                jn.db = wide_in;
                wide_in += up64(jn.LB);
                jn.row = wide;
                wide += up64(2 * ((int64_t)jn.LB + 1));
                jn.dir = dbytes;
                dbytes += up64(c);
                cells += c;
                batch.push_back(jn);
                who.push_back(v);
            }
            first = k;
            const int nj = (int)batch.size();
            MSA_TRY(d_joins.need(sizeof(Join) * batch.size()));
            MSA_TRY(d_ints.need(sizeof(int32_t) * (size_t)ints));
            MSA_TRY(d_wide.need(sizeof(int64_t) * (size_t)wide));
            MSA_TRY(d_wide_in.need(sizeof(int64_t) * (size_t)wide_in));
            MSA_TRY(d_dirs.need((size_t)dbytes));
            MSA_TRY(d_scores.need(sizeof(int64_t) * batch.size()));
            MSA_TRY(d_newlen.need(sizeof(int32_t) * batch.size()));
            MSA_HIP(hipMemcpy(d_joins.p, batch.data(), sizeof(Join) * batch.size(), hipMemcpyHostToDevice));
            const Join* jd = static_cast<const Join*>(d_joins.p);
            int32_t* di = static_cast<int32_t*>(d_ints.p);
            int64_t max_res = 1, max_cols = 1, max_cells = 1;
            for (const Join& jn : batch) {
                max_res = std::max(max_res, jn.na_res + jn.nb_res);
                max_cols = std::max<int64_t>(max_cols, (int64_t)jn.LA + jn.LB);
                max_cells = std::max<int64_t>(max_cells, (int64_t)jn.LA * jn.LB);
            }
            hipLaunchKernelGGL(pg_zero_kernel, dim3(grid_for(ints, 256)), dim3(256), 0, nullptr, di, ints);
            hipLaunchKernelGGL(pg_hist_kernel, dim3(grid_for(max_res, 256), nj), dim3(256), 0, nullptr, jd,
                               static_cast<const uint8_t*>(d_codes.p), static_cast<const int32_t*>(d_col.p), di);
            hipLaunchKernelGGL(pg_prep_kernel, dim3(grid_for(max_cols, 256), nj), dim3(256), 0, nullptr, jd,
                               static_cast<const int32_t*>(d_sigma.p), penalty, di,
                               static_cast<int64_t*>(d_wide_in.p));
            hipLaunchKernelGGL(pg_score_kernel, dim3(grid_for(max_cells, 256), nj), dim3(256), 0, nullptr, jd, di);
            hipLaunchKernelGGL(pg_dp_kernel, dim3(nj), dim3(kLanes), 0, nullptr, jd, di,
                               static_cast<const int64_t*>(d_wide_in.p), static_cast<int64_t*>(d_wide.p),
                               static_cast<uint8_t*>(d_dirs.p), static_cast<int64_t*>(d_scores.p));
            hipLaunchKernelGGL(pg_trace_kernel, dim3((nj + 63) / 64), dim3(64), 0, nullptr, jd, nj,
                               static_cast<const uint8_t*>(d_dirs.p), di, static_cast<int32_t*>(d_newlen.p));
            hipLaunchKernelGGL(pg_remap_kernel, dim3(grid_for(max_res, 256), nj), dim3(256), 0, nullptr, jd, di,
                               static_cast<const int32_t*>(d_newlen.p), static_cast<int32_t*>(d_col.p));
            MSA_HIP(hipGetLastError());
            std::vector<int32_t> nl(batch.size());
            std::vector<int64_t> sc(batch.size());
            MSA_HIP(hipMemcpy(nl.data(), d_newlen.p, sizeof(int32_t) * batch.size(), hipMemcpyDeviceToHost));
            MSA_HIP(hipMemcpy(sc.data(), d_scores.p, sizeof(int64_t) * batch.size(), hipMemcpyDeviceToHost));
            for (int t = 0; t < nj; ++t) {
                plen[who[t]] = nl[t];
                if (join_scores) join_scores[who[t] - nseq] = sc[t];
            }
        }
    }
    if (total > 0) {
        MSA_HIP(hipMemcpy(h_col.data(), d_col.p, sizeof(int32_t) * (size_t)total, hipMemcpyDeviceToHost));
        for (int32_t k = 0; k < nseq; ++k) {
            const int32_t s = order[k];
            std::memcpy(columns + out_at[s], h_col.data() + res_at[k], sizeof(int32_t) * (size_t)len[s]);
        }
    }
    *length = plen[nn - 1];
    return MSA_OK;
}
