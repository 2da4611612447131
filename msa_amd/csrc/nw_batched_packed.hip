// The batched Needleman-Wunsch kernel of the packed queue: queries that fill
// one 16 x 32 strip on their own, two targets per lane group in packed 16-bit
// cells (the device code is in nw_batched.hpp; its own source file so that it
// compiles in parallel with the others).
#include "nw_batched.hpp"

namespace msa {
namespace batched {

int occupancy_packed(int* blocks_per_cu) { return kernel_occupancy<kPackedQueue>(blocks_per_cu); }
int launch_packed(int grid, const BatchArgs& a, hipStream_t stream) { return kernel_launch<kPackedQueue>(grid, a, stream); }

}  // namespace batched
}  // namespace msa
