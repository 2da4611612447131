// Neighbor-joining on the GPU: the phylogeny stage that consumes the pairwise
// distance matrix (src/phylogeny/njoining/impl/sequential.cpp).
//
// Semantics are those of the reference's sequential algorithm, float for
// float, so the tree is the same node for node:
//   - the matrix is the pairwise scores (similarities: the pair with the
//     largest Q is joined), symmetric with a zero diagonal
//     (src/phylogeny/matrix.cuh:124-131);
//   - row sums S are accumulated in float, row k in index order (:79-94);
//   - Q(i, j) = (count-2) * d(i,j) - S_i - S_j in float, left to right, no
//     fused multiply-add (:123-127); the first pair (i > j, row-major) with
//     the strictly largest Q wins (:159-177);
//   - branch lengths dx = .5 * d + (S_x - S_y) / (2(count-2)) in double
//     rounded to float, dy = d - dx (:136-146), x = the larger index;
//   - the joined OTU takes x's place with d(i, new) = (d(i,x) + d(i,y) - d(x,y))
//     * .5 in double, S_i += updated - previous in double, and its own S the
//     float sum, in index order, of the updated values (:228-262);
//   - y leaves by a swap with index 0 and removal of index 0 (symmetric
//     variant, the reference's default) or a swap with the last index and its
//     removal (linear variant) (:188-221).
//
// Layout: the matrix is n x n float in HBM, kept in logical order so that the
// arg-max streams rows.  The reference's "swap y with index 0, remove index 0"
// (symmetric variant) is a copy of row and column 0 into y and a shift of the
// base of the logical window (matrix, node ids, row sums) by one row and one
// column; "swap y with the last, remove the last" (linear) is a copy of the
// last row and column into y.  Three launches per join, no atomics between
// workgroups: every workgroup of the arg-max leaves the best key of its rows
// (64-bit keys: order-preserving Q bits, then the complement of the pair
// index, so ties keep the first pair); the first join kernel reduces the keys
// and computes the joined OTU's distances and the row sums over all OTUs; the
// second writes rows and columns while one lane sums the joined OTU's row in
// the reference's order from LDS.
#include "msa_internal.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace msa {

namespace {

constexpr int kStepThreads = 256;
constexpr int kStage = 8192;             // floats staged in LDS for the sequential sum
constexpr int kMaxStepBlocks = 2048;     // partial arg-max slots
constexpr uint32_t kUndefined = 0xFFFFFFFFu;

struct NjState {
    float* m;                     // n x n; logical (i, j) at (base + i, base + j)
    uint32_t* map;                // logical -> node id (from base)
    float* cache;                 // logical row sums (from base)
    double* upd;                  // the joined OTU's distances, this iteration
    unsigned long long* partial;  // per workgroup: its arg-max key of this iteration
    int32_t* pick;                // the pair joined this iteration (x, y)
    int32_t* inexact;             // some distance of this join is not a float
    int32_t* error;               // set when no pair has a finite Q
    msa_nj_node* nodes;           // 2n-1
    int32_t n;
};

// Row i of pair index c = i(i-1)/2 + j, j < i.
__device__ __forceinline__ int64_t pair_row(int64_t c)
{
    int64_t i = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)c)) * 0.5);
    while (i * (i - 1) / 2 > c) --i;
    while ((i + 1) * i / 2 <= c) ++i;
    return i;
}

// Order-preserving key of a float: larger floats give larger keys.
__device__ __forceinline__ uint32_t float_key(float v)
{
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Maximum over the workgroup (every thread gets it).
__device__ unsigned long long block_max(unsigned long long v)
{
    __shared__ unsigned long long part[kStepThreads / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    __syncthreads();
    if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = v;
    __syncthreads();
    v = part[0];
    for (int w = 1; w < kStepThreads / 64; ++w) v = part[w] > v ? part[w] : v;
    return v;
}

template <typename T>
__global__ void expand_kernel(const T* __restrict__ packed, float* __restrict__ m, int32_t n)
{
    const int64_t total = (int64_t)n * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / n, j = e - i * n;
        float v = 0.0f;
        if (i != j) {
            const int64_t a = i > j ? i : j, b = i > j ? j : i;
            v = (float)packed[a * (a - 1) / 2 + b];
        }
        m[e] = v;
    }
}

__global__ void init_kernel(NjState s)
{
    const int32_t n = s.n;
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < 2 * n - 1; k += gridDim.x * blockDim.x) {
        if (k < n) {
            s.map[k] = (uint32_t)k;
            // Row sum in index order: sum_{j != k} d(j, k), read down column k
            // (the matrix is symmetric), coalesced across threads.
            float sum = 0.0f;
            for (int32_t j = 0; j < n; ++j)
                if (j != k) sum += s.m[(int64_t)j * n + k];
            s.cache[k] = sum;
        }
        msa_nj_node node;
        node.parent = kUndefined;
        node.child[0] = node.child[1] = kUndefined;
        node.distance = 3.40282347e+38f;
        node.level = 0;
        s.nodes[k] = node;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *s.error = 0;
        *s.inexact = 0;
    }
}

// Arg-max of Q over the count(count-1)/2 logical pairs: workgroups take rows
// i = 1.. round-robin, lanes run along j < i; each workgroup leaves its key.
__global__ __launch_bounds__(kStepThreads) void argmax_kernel(NjState s, int32_t count, int32_t base)
{
    const int64_t n = s.n;
    const float* mb = s.m + (int64_t)base * n + base;
    const float* cache = s.cache + base;
    const float factor = (float)(count - 2);
    unsigned long long best = 0;
    for (int32_t i = 1 + blockIdx.x; i < count; i += gridDim.x) {
        const float* row = mb + (int64_t)i * n;
        const float ci = cache[i];
        const uint32_t c0 = (uint32_t)((int64_t)i * (i - 1) / 2);
#pragma unroll 4
        for (int32_t j = threadIdx.x; j < i; j += kStepThreads) {
            float q = factor * row[j] - ci - cache[j];
            if (q > -INFINITY) {                     // NaN and -inf never win (the reference starts at -inf)
                q = q + 0.0f;                        // -0 and +0 compare equal: one key for both
                const unsigned long long key =
                    ((unsigned long long)float_key(q) << 32) | (unsigned long long)(0xFFFFFFFFu - (c0 + (uint32_t)j));
                best = key > best ? key : best;
            }
        }
    }
    best = block_max(best);
    if (threadIdx.x == 0) s.partial[blockIdx.x] = best;
}

// Reduces the arg-max partials, decodes the pair, and (every workgroup, on
// its share of the OTUs) computes the joined OTU's distances and updates the
// row sums; workgroup 0 also records the join in the tree.
__global__ __launch_bounds__(kStepThreads) void join_a_kernel(NjState s, int32_t count, int32_t base, uint32_t parent,
                                                              int32_t partials)
{
    __shared__ int32_t sx, sy;
    __shared__ float sdxy;
    const int64_t n = s.n;
    float* mb = s.m + (int64_t)base * n + base;
    uint32_t* map = s.map + base;
    float* cache = s.cache + base;
    unsigned long long key = 0;
    for (int32_t b = threadIdx.x; b < partials; b += kStepThreads) key = s.partial[b] > key ? s.partial[b] : key;
    key = block_max(key);
    if (threadIdx.x == 0) {
        // No finite Q (the reference's pick is undefined there): flag it and
        // join the first pair so that every access stays in bounds.
        const int64_t c = key ? (int64_t)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFu)) : 0;
        const int64_t i = pair_row(c);
        const int32_t x = (int32_t)i, y = (int32_t)(c - i * (i - 1) / 2);
        const float dxy = mb[(int64_t)x * n + y];
        sx = x, sy = y, sdxy = dxy;
        if (blockIdx.x == 0) {
            if (key == 0) *s.error = 1;
            const float pairsum = cache[x] - cache[y];
            const float dx = (float)(.5 * (double)dxy + (double)(pairsum / (float)(2 * (count - 2))));
            const float dy = dxy - dx;
            // star::join (src/phylogeny/njoining/star.cuh:63-80)
            const uint32_t a = map[x], b = map[y];
            const uint32_t c1 = a < b ? a : b, c2 = a < b ? b : a;
            s.nodes[c1].distance = a < b ? dx : dy;
            s.nodes[c1].parent = parent;
            s.nodes[c2].distance = a < b ? dy : dx;
            s.nodes[c2].parent = parent;
            const uint32_t l1 = s.nodes[c1].level, l2 = s.nodes[c2].level;
            s.nodes[parent].level = (l1 > l2 ? l1 : l2) + 1;
            s.nodes[parent].child[0] = c1;
            s.nodes[parent].child[1] = c2;
            s.pick[0] = x;
            s.pick[1] = y;
        }
    }
    __syncthreads();
    const int32_t x = sx, y = sy;
    const float dxy = sdxy;
    const float* rowx = mb + (int64_t)x * n;
    const float* rowy = mb + (int64_t)y * n;
    // The row sums of x and y are read above for the branch lengths; their
    // updates here are dead in the reference (x's is replaced by the new sum,
    // y leaves or is replaced by `other`'s), so they are skipped, which also
    // keeps workgroup 0's reads free of races.
    int32_t inexact = 0;
    for (int32_t i = blockIdx.x * kStepThreads + threadIdx.x; i < count; i += gridDim.x * kStepThreads) {
        const float previous = rowx[i] + rowy[i];
        const double updated = (double)(previous - dxy) * .5;
        s.upd[i] = updated;
        inexact |= (double)(float)updated != updated;
        if (i != x && i != y) cache[i] = (float)((double)cache[i] + (updated - (double)previous));
    }
    if (__any(inexact) && threadIdx.x % 64 == 0) atomicOr(s.inexact, 1);
}

// Writes the joined OTU's row and column, moves the OTU at `other` (index 0,
// or the last) into y's place (y leaves; `other` is dropped), and sums the
// joined OTU's row in index order on wave 0 of workgroup 0.
__global__ __launch_bounds__(kStepThreads) void join_b_kernel(NjState s, int32_t count, int32_t base, uint32_t parent,
                                                              int32_t variant)
{
    __shared__ __attribute__((aligned(16))) float stage[kStage];
    const int64_t n = s.n;
    float* mb = s.m + (int64_t)base * n + base;
    uint32_t* map = s.map + base;
    float* cache = s.cache + base;
    const int32_t x = s.pick[0], y = s.pick[1];
    const int32_t other = variant == MSA_NJ_SYMMETRIC ? 0 : count - 1;
    const bool move = y != other;
    // Row and column x; y's own row and column are rewritten below (or y is
    // dropped), so each element has one writer.
    for (int32_t i = blockIdx.x * kStepThreads + threadIdx.x; i < count; i += gridDim.x * kStepThreads) {
        if (i == y) continue;
        const float v = (float)s.upd[i];
        mb[(int64_t)x * n + i] = v;
        mb[(int64_t)i * n + x] = v;
    }
    if (move) {
        // The new row y is `other`'s row after this join: the joined OTU's
        // distance where `other` meets x, else unchanged.
        for (int32_t k = blockIdx.x * kStepThreads + threadIdx.x; k < count; k += gridDim.x * kStepThreads) {
            if (k == y || k == other) continue;
            const float v = other == x ? (float)s.upd[k] : (k == x ? (float)s.upd[other] : mb[(int64_t)other * n + k]);
            mb[(int64_t)y * n + k] = v;
            mb[(int64_t)k * n + y] = v;
        }
    }
    if (blockIdx.x != 0) return;
    // The joined OTU's row sum, in index order: float += double.  When every
    // term is a float (always, unless a halving underflowed), that is the
    // float sum, because rounding the exact double sum of two floats to
    // float is the float sum (53 >= 2 * 24 + 2 bits).  The terms are staged
    // in LDS and lane 0 reads 64 at a time into registers, so the dependent
    // chain is one add per term.
    const bool inexact = *s.inexact != 0;
    float new_sum = 0.0f;
    for (int32_t c0 = 0; c0 < count; c0 += kStage) {
        const int32_t len = count - c0 < kStage ? count - c0 : kStage;
        __syncthreads();
        if (!inexact)
            for (int32_t i = threadIdx.x; i < len; i += kStepThreads) stage[i] = (float)s.upd[c0 + i];
        __syncthreads();
        if (threadIdx.x != 0) continue;
        if (inexact) {
            for (int32_t i = 0; i < len; ++i) new_sum = (float)((double)new_sum + s.upd[c0 + i]);
            continue;
        }
        int32_t i = 0;
        for (; i + 64 <= len; i += 64) {
            float4 v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = *reinterpret_cast<const float4*>(&stage[i + 4 * q]);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                new_sum += v[q].x;
                new_sum += v[q].y;
                new_sum += v[q].z;
                new_sum += v[q].w;
            }
        }
        for (; i < len; ++i) new_sum += stage[i];
    }
    if (threadIdx.x == 0) {
        cache[x] = new_sum;
        map[x] = parent;
        if (move) {
            cache[y] = other == x ? new_sum : cache[other];
            map[y] = other == x ? parent : map[other];
        }
        *s.inexact = 0;          // for the next join (join_a only ever sets it)
    }
}

// Arg-max workgroups: one per row, up to 8 per compute unit.
int step_grid(int32_t count, int cus) { return std::max(1, std::min({count - 1, cus * 8, kMaxStepBlocks})); }

// Workgroups for the per-OTU phases of a join.
int otu_grid(int32_t count) { return std::max(1, (count + kStepThreads - 1) / kStepThreads); }

struct NjBuffers {
    void* block = nullptr;
    NjState s {};
    ~NjBuffers()
    {
        if (block) (void)hipFree(block);
    }
};

int nj_alloc(int32_t n, NjBuffers& b)
{
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t nn = (size_t)n * (size_t)n;
    const size_t o_m = 0;
    const size_t o_map = up(o_m + nn * sizeof(float));
    const size_t o_cache = up(o_map + (size_t)n * sizeof(uint32_t));
    const size_t o_upd = up(o_cache + (size_t)n * sizeof(float));
    const size_t o_partial = up(o_upd + (size_t)n * sizeof(double));
    const size_t o_pick = up(o_partial + kMaxStepBlocks * sizeof(unsigned long long));
    const size_t o_error = up(o_pick + 4 * sizeof(int32_t));
    const size_t bytes = up(o_error + sizeof(int32_t));
    MSA_HIP(hipMalloc(&b.block, bytes));
    char* base = static_cast<char*>(b.block);
    b.s.m = reinterpret_cast<float*>(base + o_m);
    b.s.map = reinterpret_cast<uint32_t*>(base + o_map);
    b.s.cache = reinterpret_cast<float*>(base + o_cache);
    b.s.upd = reinterpret_cast<double*>(base + o_upd);
    b.s.partial = reinterpret_cast<unsigned long long*>(base + o_partial);
    b.s.pick = reinterpret_cast<int32_t*>(base + o_pick);
    b.s.inexact = b.s.pick + 2;
    b.s.error = reinterpret_cast<int32_t*>(base + o_error);
    b.s.n = n;
    return MSA_OK;
}

template <typename T>
int nj_run(const T* d_packed, int32_t n, int32_t variant, msa_nj_node* d_nodes, hipStream_t stream)
{
    if (n < 2) return MSA_OK;
    int dev = 0;
    MSA_HIP(hipGetDevice(&dev));
    int cus = 0;
    MSA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    NjBuffers b;
    MSA_TRY(nj_alloc(n, b));
    b.s.nodes = d_nodes;
    const int64_t elems = (int64_t)n * n;
    const int eg = (int)std::min<int64_t>((elems + 255) / 256, (int64_t)cus * 64);
    hipLaunchKernelGGL(expand_kernel<T>, dim3(eg), dim3(256), 0, stream, d_packed, b.s.m, n);
    MSA_HIP(hipGetLastError());
    const int ig = (int)std::min<int64_t>((2 * (int64_t)n - 1 + 255) / 256, (int64_t)cus * 16);
    hipLaunchKernelGGL(init_kernel, dim3(ig), dim3(256), 0, stream, b.s);
    MSA_HIP(hipGetLastError());
    for (int32_t it = 0; it + 1 < n; ++it) {
        const int32_t count = n - it;
        const int32_t base = variant == MSA_NJ_SYMMETRIC ? it : 0;
        const int grid = step_grid(count, cus);
        hipLaunchKernelGGL(argmax_kernel, dim3(grid), dim3(kStepThreads), 0, stream, b.s, count, base);
        hipLaunchKernelGGL(join_a_kernel, dim3(otu_grid(count)), dim3(kStepThreads), 0, stream, b.s, count, base,
                           (uint32_t)(n + it), grid);
        hipLaunchKernelGGL(join_b_kernel, dim3(otu_grid(count)), dim3(kStepThreads), 0, stream, b.s, count, base,
                           (uint32_t)(n + it), variant);
    }
    MSA_HIP(hipGetLastError());
    // The buffers are freed when this returns: wait for the stream first.
    MSA_HIP(hipStreamSynchronize(stream));
    int32_t error = 0;
    MSA_HIP(hipMemcpy(&error, b.s.error, sizeof(error), hipMemcpyDeviceToHost));
    if (error) return fail(MSA_ERR_RUNTIME, "neighbor-joining found no pair with a finite Q value");
    return MSA_OK;
}

int check_nj_args(int32_t n, int32_t variant)
{
    if (n < 0) return fail(MSA_ERR_INVALID_ARGUMENT, "negative OTU count");
    if (n > 65535) return fail(MSA_ERR_TOO_LARGE, "%d OTUs; at most 65535 are supported", n);
    if (variant != MSA_NJ_SYMMETRIC && variant != MSA_NJ_LINEAR)
        return fail(MSA_ERR_INVALID_ARGUMENT, "unknown neighbor-joining variant %d", variant);
    return MSA_OK;
}

}  // namespace

}  // namespace msa

extern "C" {

int msa_nj_build_device(const int32_t* d_scores, int32_t n, int32_t variant, msa_nj_node* d_nodes, msa_stream stream)
{
    MSA_TRY(msa::check_nj_args(n, variant));
    if (n >= 2 && (!d_scores || !d_nodes)) return msa::fail(MSA_ERR_INVALID_ARGUMENT, "null pointer");
    return msa::nj_run<int32_t>(d_scores, n, variant, d_nodes, (hipStream_t)stream);
}

int msa_nj_build(const float* distances, int32_t n, int32_t variant, msa_nj_node* nodes)
{
    MSA_TRY(msa::check_nj_args(n, variant));
    if (n < 2) return MSA_OK;
    if (!distances || !nodes) return msa::fail(MSA_ERR_INVALID_ARGUMENT, "null pointer");
    const size_t npairs = (size_t)n * (size_t)(n - 1) / 2;
    float* d_in = nullptr;
    msa_nj_node* d_nodes = nullptr;
    int rc = msa::hip_status(hipMalloc(&d_in, npairs * sizeof(float)), "hipMalloc");
    if (rc == MSA_OK) rc = msa::hip_status(hipMalloc(&d_nodes, (2 * (size_t)n - 1) * sizeof(msa_nj_node)), "hipMalloc");
    if (rc == MSA_OK)
        rc = msa::hip_status(hipMemcpy(d_in, distances, npairs * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc == MSA_OK) rc = msa::nj_run<float>(d_in, n, variant, d_nodes, nullptr);
    if (rc == MSA_OK)
        rc = msa::hip_status(hipMemcpy(nodes, d_nodes, (2 * (size_t)n - 1) * sizeof(msa_nj_node),
                                       hipMemcpyDeviceToHost), "hipMemcpy");
    if (d_in) (void)hipFree(d_in);
    if (d_nodes) (void)hipFree(d_nodes);
    return rc;
}

}  // extern "C"
