// The batched Needleman-Wunsch kernel's host entry points and its 16 x 32
// tile configuration (the device code is in nw_batched.hpp).
#include "nw_batched.hpp"

namespace msa {

namespace batched {

static_assert(kTileConfig[0].lanes == 16 && kTileConfig[0].rows == 32, "configuration 0 is 16 x 32");

int occupancy_16x32(int* blocks_per_cu) { return kernel_occupancy<0>(blocks_per_cu); }
int launch_16x32(int grid, const BatchArgs& a, hipStream_t stream) { return kernel_launch<0>(grid, a, stream); }

}  // namespace batched

namespace {

__global__ void queue_reset_kernel(int32_t* queue, int32_t* queue2, int32_t* queue3)
{
    if (threadIdx.x == 0) queue[0] = 0;
    if (threadIdx.x == 1 && queue2 != nullptr) queue2[0] = 0;
    if (threadIdx.x == 2 && queue3 != nullptr) queue3[0] = 0;
}

}  // namespace

int batched_threads() { return batched::kWaves * 64; }

size_t batched_lds_bytes() { return batched::lds_bytes(); }

int batched_groups_per_block() { return batched::kWaves * batched::kMaxGroups; }

int batched_occupancy(int* blocks_per_cu)
{
    // The smallest of the five kernels' occupancies: a plan sizes its grids
    // and line buffers once for whichever configurations it uses.
    int n[kBatchQueues] = {};
    MSA_TRY(batched::occupancy_16x32(&n[0]));
    MSA_TRY(batched::occupancy_16x16(&n[1]));
    MSA_TRY(batched::occupancy_32x16(&n[2]));
    MSA_TRY(batched::occupancy_packed(&n[3]));
    MSA_TRY(batched::occupancy_duo(&n[4]));
    *blocks_per_cu = std::min({n[0], n[1], n[2], n[3], n[4]});
    return MSA_OK;
}

int batched_launch(int config, int grid, const uint16_t* d_blocks, const SeqInfo* d_info, const int4* d_items,
                   int32_t nitems, const int2* d_targets, int32_t* d_queue, int32_t* d_out, int32_t* d_lines,
                   int32_t line_stride, const Sxx& sxx, int32_t penalty, hipStream_t stream,
                   const int4* d_items2, int32_t nitems2, int32_t* d_queue2, const int4* d_items3, int32_t nitems3,
                   int32_t* d_queue3)
{
    if (config < 0 || config >= kBatchQueues) return fail(MSA_ERR_INVALID_ARGUMENT, "no batched queue %d", config);
    if ((nitems2 > 0 && (config != kDuoQueue || d_items2 == nullptr || d_queue2 == nullptr)) ||
        (nitems3 > 0 && (config != kDuoQueue || d_items3 == nullptr || d_queue3 == nullptr)))
        return fail(MSA_ERR_INVALID_ARGUMENT, "only the duo launch takes more queues");
    if (nitems2 <= 0) d_queue2 = nullptr;
    if (nitems3 <= 0) d_queue3 = nullptr;
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, stream, d_queue, d_queue2, d_queue3);
    MSA_HIP(hipGetLastError());
    batched::BatchArgs a;
    a.blocks = d_blocks;
    a.info = d_info;
    a.items = d_items;
    a.targets = d_targets;
    a.nitems = nitems;
    a.queue = d_queue;
    a.items2 = nitems2 > 0 ? d_items2 : nullptr;
    a.nitems2 = std::max(nitems2, 0);
    a.queue2 = d_queue2;
    a.items3 = nitems3 > 0 ? d_items3 : nullptr;
    a.nitems3 = std::max(nitems3, 0);
    a.queue3 = d_queue3;
    a.out = d_out;
    a.lines = d_lines;
    a.line_stride = line_stride;
    a.penalty = penalty;
    int max_sxx = 0;
    for (int i = 0; i < 625; ++i) max_sxx = std::max<int>(max_sxx, sxx.v[i]);
    a.margin = 16 * max_sxx;
    a.sxx = sxx;
    switch (config) {
    case 0: return batched::launch_16x32(grid, a, stream);
    case 1: return batched::launch_16x16(grid, a, stream);
    case 2: return batched::launch_32x16(grid, a, stream);
    case 3: return batched::launch_packed(grid, a, stream);
    default: return batched::launch_duo(grid, a, stream);
    }
}

}  // namespace msa
