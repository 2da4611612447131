// The batched Needleman-Wunsch kernel for the 32x16 tile configuration (the
// device code is in nw_batched.hpp; its own source file so that it compiles
// in parallel with the others).
#include "nw_batched.hpp"

namespace msa {
namespace batched {

static_assert(kTileConfig[2].lanes == 32 && kTileConfig[2].rows == 16, "configuration 2 is 32x16");

int occupancy_32x16(int* blocks_per_cu) { return kernel_occupancy<2>(blocks_per_cu); }
int launch_32x16(int grid, const BatchArgs& a, hipStream_t stream) { return kernel_launch<2>(grid, a, stream); }

}  // namespace batched
}  // namespace msa
