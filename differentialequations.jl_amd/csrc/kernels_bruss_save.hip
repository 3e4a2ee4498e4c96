// kernels_bruss_save.hip -- the persistent team kernels of the uneven
// Brusselator batch built with saveat (bruss_team.hpp with SAVE): every
// accepted step writes the saveat points it passes from its dense output.
// A translation unit of its own so that the two sets of instantiations
// compile in parallel.
#include "bruss_team.hpp"

namespace mideq {

hipError_t bruss_occupancy_save(int lanes, int nk, int alg, int mix, int fast, int* per_cu) {
  return dispatch(lanes, nk, alg, mix, fast, OccupancyOp<true>{per_cu});
}

hipError_t bruss_launch_save(const BrussArgs& a, int lanes, int nk, int alg, int mix, int fast,
                             int grid, hipStream_t s) {
  return dispatch(lanes, nk, alg, mix, fast, LaunchOp<true>{&a, grid, s});
}

}  // namespace mideq
