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
// Layout: the matrix is n x n float in HBM and never moves.  The reference's
// swaps and removals are relabelings, so logical index k of an iteration maps
// to a physical row through `slot`; the symmetric variant's "remove index 0"
// is a shift of the base of the logical arrays (slot, node ids, row sums) by
// one per iteration.  Per iteration: one grid-wide arg-max of Q over the
// count(count-1)/2 logical pairs (64-bit keys: order-preserving Q bits, then
// the complement of the pair index, so ties keep the first pair), and one
// workgroup that joins the pair.  The row sums of the joined OTU are a
// sequential float sum (the reference's order), done by one lane from LDS.
#include "msa_internal.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace msa {

namespace {

constexpr int kJoinThreads = 1024;
constexpr int kSumChunk = 4096;          // doubles staged in LDS for the sequential sum
constexpr int kArgmaxThreads = 256;
constexpr uint32_t kUndefined = 0xFFFFFFFFu;

struct NjState {
    float* m;                     // n x n, physical rows
    int32_t* slot;                // logical -> physical row (from base)
    uint32_t* map;                // logical -> node id (from base)
    float* cache;                 // logical row sums (from base)
    double* upd;                  // the joined OTU's distances, this iteration
    unsigned long long* best;     // arg-max key of this iteration
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
            s.slot[k] = k;
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
        *s.best = 0;
        *s.error = 0;
    }
}

// Arg-max of Q over the count(count-1)/2 logical pairs of this iteration:
// workgroups take rows i = 1.. round-robin, lanes run along j < i.
__global__ __launch_bounds__(kArgmaxThreads) void argmax_kernel(NjState s, int32_t count, int32_t base)
{
    const int32_t* slot = s.slot + base;
    const float* cache = s.cache + base;
    const float factor = (float)(count - 2);
    const int64_t n = s.n;
    unsigned long long best = 0;
    for (int32_t i = 1 + blockIdx.x; i < count; i += gridDim.x) {
        const float* row = s.m + (int64_t)slot[i] * n;
        const float ci = cache[i];
        const uint32_t c0 = (uint32_t)((int64_t)i * (i - 1) / 2);
        for (int32_t j = threadIdx.x; j < i; j += kArgmaxThreads) {
            float q = factor * row[slot[j]] - ci - cache[j];
            if (q > -INFINITY) {                     // NaN and -inf never win (the reference starts at -inf)
                q = q + 0.0f;                        // -0 and +0 compare equal: one key for both
                const unsigned long long key =
                    ((unsigned long long)float_key(q) << 32) | (unsigned long long)(0xFFFFFFFFu - (c0 + (uint32_t)j));
                best = key > best ? key : best;
            }
        }
    }
    // Wave, then workgroup, then device maximum.
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off);
        best = o > best ? o : best;
    }
    __shared__ unsigned long long part[kArgmaxThreads / 64];
    if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kArgmaxThreads / 64; ++w) best = part[w] > best ? part[w] : best;
        if (best) atomicMax(s.best, best);
    }
}

// Joins the chosen pair (one workgroup).
__global__ __launch_bounds__(kJoinThreads) void join_kernel(NjState s, int32_t count, int32_t base, uint32_t parent,
                                                            int32_t variant)
{
    __shared__ int32_t sx, sy, px, py;
    __shared__ float sdxy;
    __shared__ double stage[kSumChunk];
    int32_t* slot = s.slot + base;
    uint32_t* map = s.map + base;
    float* cache = s.cache + base;
    const int64_t n = s.n;
    if (threadIdx.x == 0) {
        const unsigned long long key = *s.best;
        // No finite Q (the reference's pick is undefined there): flag it and
        // join the first pair so that every access stays in bounds.
        if (key == 0) *s.error = 1;
        const int64_t c = key ? (int64_t)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFu)) : 0;
        const int64_t i = pair_row(c);
        const int32_t x = (int32_t)i, y = (int32_t)(c - i * (i - 1) / 2);
        const float dxy = s.m[(int64_t)slot[x] * n + slot[y]];
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
        sx = x, sy = y, px = slot[x], py = slot[y], sdxy = dxy;
    }
    __syncthreads();
    const int32_t x = sx, y = sy;
    const int64_t rx = px, ry = py;
    const float dxy = sdxy;
    for (int32_t i = threadIdx.x; i < count; i += kJoinThreads) {
        const int64_t ri = slot[i];
        const float previous = s.m[ri * n + rx] + s.m[ri * n + ry];
        const double updated = (double)(previous - dxy) * .5;
        s.upd[i] = updated;
        cache[i] = (float)((double)cache[i] + (updated - (double)previous));
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < count; i += kJoinThreads) {
        const int64_t ri = slot[i];
        const float v = (float)s.upd[i];
        s.m[ri * n + rx] = v;
        s.m[rx * n + ri] = v;
    }
    // The joined OTU's row sum: float += double, in index order.
    float new_sum = 0.0f;
    for (int32_t c0 = 0; c0 < count; c0 += kSumChunk) {
        const int32_t len = count - c0 < kSumChunk ? count - c0 : kSumChunk;
        __syncthreads();
        for (int32_t i = threadIdx.x; i < len; i += kJoinThreads) stage[i] = s.upd[c0 + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int32_t i = 0; i < len; ++i) new_sum = (float)((double)new_sum + stage[i]);
    }
    if (threadIdx.x == 0) {
        cache[x] = new_sum;
        map[x] = parent;
        const int32_t other = variant == MSA_NJ_SYMMETRIC ? 0 : count - 1;
        if (y != other) {
            const float tc = cache[other];
            cache[other] = cache[y];
            cache[y] = tc;
            const int32_t ts = slot[other];
            slot[other] = slot[y];
            slot[y] = ts;
            const uint32_t tm = map[other];
            map[other] = map[y];
            map[y] = tm;
        }
        *s.best = 0;
    }
}

// One workgroup per row, up to 8 per compute unit (rows are taken round-robin).
int argmax_grid(int32_t count, int cus) { return std::max(1, std::min(count - 1, cus * 8)); }

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
    const size_t o_slot = up(o_m + nn * sizeof(float));
    const size_t o_map = up(o_slot + (size_t)n * sizeof(int32_t));
    const size_t o_cache = up(o_map + (size_t)n * sizeof(uint32_t));
    const size_t o_upd = up(o_cache + (size_t)n * sizeof(float));
    const size_t o_best = up(o_upd + (size_t)n * sizeof(double));
    const size_t o_error = up(o_best + sizeof(unsigned long long));
    const size_t bytes = up(o_error + sizeof(int32_t));
    MSA_HIP(hipMalloc(&b.block, bytes));
    char* base = static_cast<char*>(b.block);
    b.s.m = reinterpret_cast<float*>(base + o_m);
    b.s.slot = reinterpret_cast<int32_t*>(base + o_slot);
    b.s.map = reinterpret_cast<uint32_t*>(base + o_map);
    b.s.cache = reinterpret_cast<float*>(base + o_cache);
    b.s.upd = reinterpret_cast<double*>(base + o_upd);
    b.s.best = reinterpret_cast<unsigned long long*>(base + o_best);
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
        hipLaunchKernelGGL(argmax_kernel, dim3(argmax_grid(count, cus)), dim3(kArgmaxThreads), 0, stream, b.s, count,
                           base);
        hipLaunchKernelGGL(join_kernel, dim3(1), dim3(kJoinThreads), 0, stream, b.s, count, base,
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
