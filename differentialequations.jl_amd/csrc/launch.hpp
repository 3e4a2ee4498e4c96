// launch.hpp -- host-side declarations shared by the kernel files and the
// C-ABI runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "lane_solver.hpp"

namespace mideq {

enum { MODEL_EXP = 0, MODEL_LORENZ = 1, MODEL_ROBERTSON = 2, MODEL_VDP = 3,
       MODEL_BRUSSELATOR = 4, MODEL_LINEAR = 5 };
enum { SCHED_AUTO = 0, SCHED_LANE = 1, SCHED_PERSISTENT = 2 };

// One lane per trajectory (n <= 3 models).
hipError_t launch_lane_solver(const LaneArgs& a, int model, int alg, int mix, int fast,
                              int sched, int block, int grid, hipStream_t s);
int lane_model_dims(int model, int* n, int* np);
// result[r] = fold over x[r][0..N) (op 1 sum, 2 mean, 3 min, 4 max) for
// r < nrow; partial holds nrow * ceil(N / 256) doubles.
hipError_t launch_fold(const double* x, int64_t N, int nrow, int op, double* partial,
                       double* result, hipStream_t s);
// *p = v by a one-thread kernel (no memset node inside a captured graph).
hipError_t launch_set_u64(unsigned long long* p, unsigned long long v, hipStream_t s);

}  // namespace mideq
