// The batched Needleman-Wunsch kernel of the duo queue: pairs of queries of
// the packed kind against one target per lane group in packed 16-bit cells,
// workgroups taking whole items (the device code is in nw_batched.hpp; its own
// source file so that it compiles in parallel with the others).
#include "nw_batched.hpp"

namespace msa {
namespace batched {

int occupancy_duo(int* blocks_per_cu) { return kernel_occupancy<kDuoQueue>(blocks_per_cu); }
int launch_duo(int grid, const BatchArgs& a, hipStream_t stream) { return kernel_launch<kDuoQueue>(grid, a, stream); }

}  // namespace batched
}  // namespace msa
