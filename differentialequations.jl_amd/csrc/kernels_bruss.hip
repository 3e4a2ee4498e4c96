// kernels_bruss.hip -- the persistent team kernels of the uneven
// Brusselator batch without saved points (bruss_team.hpp), the queue and
// sample kernels, and the host-side planning and launch entry points (the
// builds with saveat live in kernels_bruss_save.hip).
#include "bruss_team.hpp"

namespace mideq {

__global__ void k_bruss_reset(unsigned long long* q) { *q = 0ull; }

// out[k] = order[(k * ntraj) / S] (natural order when order is NULL): a
// strided sample of the launch order, for the autotuner.
__global__ void k_sample_order(int64_t* out, const int64_t* order, int64_t ntraj, int64_t S) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S) return;
  // floor(k ntraj / S) = k a + floor(k r / S) with ntraj = a S + r (no overflow)
  const int64_t a = ntraj / S, r = ntraj % S;
  int64_t q = k * a + (k * r) / S;
  out[k] = order ? order[q] : q;
}

hipError_t bruss_plan(int lanes, int nk, int alg, int mix, int fast, int ncu, int64_t ntraj,
                      int* grid, int* teams_per_block, bool save) {
  int per_cu = 0;
  hipError_t e = save ? bruss_occupancy_save(lanes, nk, alg, mix, fast, &per_cu)
                      : dispatch(lanes, nk, alg, mix, fast, OccupancyOp<false>{&per_cu});
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  int tpb = 256 / lanes;
  int64_t need = (ntraj + tpb - 1) / tpb;
  int64_t g = (int64_t)per_cu * ncu;
  *grid = (int)(g < need ? g : need);
  if (*grid < 1) *grid = 1;
  *teams_per_block = tpb;
  return hipSuccess;
}

hipError_t bruss_sample_order(int64_t* out, const int64_t* order, int64_t ntraj, int64_t S,
                              hipStream_t s) {
  if (S <= 0 || S > ntraj || S > (1 << 24)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_sample_order, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s, out,
                     order, ntraj, S);
  return hipGetLastError();
}

hipError_t launch_bruss_solver(const BrussArgs& a, int lanes, int nk, int alg, int mix, int fast,
                               int grid, hipStream_t s) {
  if (lanes != 32 && lanes != 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_bruss_reset, dim3(1), dim3(1), 0, s, a.queue);
  if (a.us && a.ns > 0) return bruss_launch_save(a, lanes, nk, alg, mix, fast, grid, s);
  return dispatch(lanes, nk, alg, mix, fast, LaunchOp<false>{&a, grid, s});
}

}  // namespace mideq

