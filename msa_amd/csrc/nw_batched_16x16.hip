// The batched Needleman-Wunsch kernel for the 16x16 tile configuration (the
// device code is in nw_batched.hpp; its own source file so that it compiles
// in parallel with the others).
#include "nw_batched.hpp"

namespace msa {
namespace batched {

static_assert(kTileConfig[1].lanes == 16 && kTileConfig[1].rows == 16, "configuration 1 is 16x16");

int occupancy_16x16(int* blocks_per_cu) { return kernel_occupancy<1>(blocks_per_cu); }
int launch_16x16(int grid, const BatchArgs& a, hipStream_t stream) { return kernel_launch<1>(grid, a, stream); }

}  // namespace batched
}  // namespace msa
