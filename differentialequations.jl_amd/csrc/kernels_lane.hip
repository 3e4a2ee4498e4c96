// kernels_lane.hip -- ensemble kernels for small systems (n <= 3): one lane
// per trajectory, state in VGPRs, structure-of-arrays I/O in HBM.
//
// Two schedules:
//   lane        one thread per trajectory; the hardware dispatcher refills
//               CU slots wave by wave as waves retire.
// Both walk the trajectories in the caller's launch order when one is given
// (consecutive entries share a wavefront); results land at natural indices.
//   persistent  a fixed grid; each lane that finishes (detected with a
//               wavefront ballot) takes the next trajectory from a global
//               atomic counter, one atomic per wave per refill, so lanes
//               stay busy while the queue lasts.
#include "lane_solver.hpp"
#include "launch.hpp"

namespace mideq {

template <class M, bool MIX, int ALG, bool FAST>
__global__ __launch_bounds__(256) void k_lane(LaneArgs a) {
  lane_kernel_run<M, MIX, ALG, FAST>(a);
}

template <class M, bool MIX, int ALG, bool FAST>
__global__ __launch_bounds__(256) void k_persistent(LaneArgs a) {
  persistent_kernel_run<M, MIX, ALG, FAST>(a);
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

template <class M, bool MIX, int ALG, bool FAST>
static hipError_t launch_t(const LaneArgs& a, int sched, int block, int grid,
                           hipStream_t s) {
  const size_t lds = lane_lds_bytes(a.ns);
  if (sched == SCHED_PERSISTENT) {
    // grid <= 0: as many blocks as can be resident at once (a truly
    // persistent grid), from the occupancy of this instantiation.
    if (grid <= 0) {
      int per_cu = 0;
      hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, k_persistent<M, MIX, ALG, FAST>, block, lds);
      if (e != hipSuccess || per_cu < 1) per_cu = 1;
      int dev = 0, ncu = 256;
      if (hipGetDevice(&dev) == hipSuccess)
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) ncu = 256;
      int64_t need = (a.ntraj + block - 1) / block;
      int64_t g = (int64_t)per_cu * ncu;
      grid = (int)(g < need ? g : need);
    }
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, a.queue, 0ull);
    hipLaunchKernelGGL((k_persistent<M, MIX, ALG, FAST>), dim3(grid), dim3(block), lds, s, a);
  } else {
    hipLaunchKernelGGL((k_lane<M, MIX, ALG, FAST>), dim3(grid), dim3(block), lds, s, a);
  }
  return hipGetLastError();
}

template <class M, bool MIX, bool FAST>
static hipError_t launch_alg(const LaneArgs& a, int alg, int sched, int block, int grid,
                             hipStream_t s) {
  // saveat with a Rosenbrock method: the build with its dense output
  const bool dense = a.ns > 0 && a.us;
  switch (alg) {
    case ALG_TSIT5: return launch_t<M, MIX, ALG_TSIT5, FAST>(a, sched, block, grid, s);
    case ALG_RODAS5:
      return dense ? launch_t<M, MIX, ALG_RODAS5 | ALG_DENSE, FAST>(a, sched, block, grid, s)
                   : launch_t<M, MIX, ALG_RODAS5, FAST>(a, sched, block, grid, s);
    case ALG_RODAS5P:
      return dense ? launch_t<M, MIX, ALG_RODAS5P | ALG_DENSE, FAST>(a, sched, block, grid, s)
                   : launch_t<M, MIX, ALG_RODAS5P, FAST>(a, sched, block, grid, s);
    case ALG_AUTO5P:
      return dense ? launch_t<M, MIX, ALG_AUTO5P | ALG_DENSE, FAST>(a, sched, block, grid, s)
                   : launch_t<M, MIX, ALG_AUTO5P, FAST>(a, sched, block, grid, s);
    case ALG_AUTO_T: return launch_t<M, MIX, ALG_AUTO_T, FAST>(a, sched, block, grid, s);
    case ALG_TRUST: return launch_t<M, MIX, ALG_TRUST, FAST>(a, sched, block, grid, s);
    default:
      return dense ? launch_t<M, MIX, ALG_AUTO | ALG_DENSE, FAST>(a, sched, block, grid, s)
                   : launch_t<M, MIX, ALG_AUTO, FAST>(a, sched, block, grid, s);
  }
}

template <class M>
static hipError_t launch_model(const LaneArgs& a, int alg, int mix, int fast, int sched,
                               int block, int grid, hipStream_t s) {
  if (fast)
    return mix ? launch_alg<M, true, true>(a, alg, sched, block, grid, s)
               : launch_alg<M, false, true>(a, alg, sched, block, grid, s);
  return mix ? launch_alg<M, true, false>(a, alg, sched, block, grid, s)
             : launch_alg<M, false, false>(a, alg, sched, block, grid, s);
}

// Deterministic fold of x[row][0..N) per row: blocks of 256 by a halving
// tree (identity past the end), then the block results in block order.
__device__ __forceinline__ double fold_identity(int op) {
  return op == 3 ? __builtin_inf() : op == 4 ? -__builtin_inf() : 0.0;
}
__device__ __forceinline__ double fold2(int op, double a, double b) {
  return op == 3 ? fmin(a, b) : op == 4 ? fmax(a, b) : a + b;
}

__global__ __launch_bounds__(256) void k_fold_blocks(const double* x, int64_t N, int op,
                                                     int64_t nb, double* partial) {
  __shared__ double s[256];
  const int64_t b = blockIdx.x, row = blockIdx.y;
  const int tl = threadIdx.x;
  const int64_t i = b * 256 + tl;
  s[tl] = i < N ? x[row * N + i] : fold_identity(op);
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (tl < w) s[tl] = fold2(op, s[tl], s[tl + w]);
    __syncthreads();
  }
  if (tl == 0) partial[row * nb + b] = s[0];
}

__global__ void k_fold_final(const double* partial, int64_t nb, int op, int64_t N, int nrow,
                             double* result) {
  const int row = threadIdx.x;
  if (row >= nrow) return;
  double acc = partial[(int64_t)row * nb];
  for (int64_t b = 1; b < nb; b++) acc = fold2(op, acc, partial[(int64_t)row * nb + b]);
  if (op == 2) acc = acc / (double)N;
  result[row] = acc;
}

hipError_t launch_fold(const double* x, int64_t N, int nrow, int op, double* partial,
                       double* result, hipStream_t s) {
  if (N <= 0 || nrow <= 0 || nrow > 1024 || op < 1 || op > 4) return hipErrorInvalidValue;
  const int64_t nb = (N + 255) / 256;
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fold_blocks, dim3((unsigned)nb, (unsigned)nrow), dim3(256), 0, s, x, N,
                     op, nb, partial);
  hipLaunchKernelGGL(k_fold_final, dim3(1), dim3(((nrow + 63) / 64) * 64), 0, s, partial, nb,
                     op, N, nrow, result);
  return hipGetLastError();
}

hipError_t launch_set_u64(unsigned long long* p, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_lane_solver(const LaneArgs& a, int model, int alg, int mix, int fast,
                              int sched, int block, int grid, hipStream_t s) {
  switch (model) {
    case MODEL_EXP: return launch_model<ModelExp>(a, alg, mix, fast, sched, block, grid, s);
    case MODEL_LORENZ: return launch_model<ModelLorenz>(a, alg, mix, fast, sched, block, grid, s);
    case MODEL_ROBERTSON:
      return launch_model<ModelRobertson>(a, alg, mix, fast, sched, block, grid, s);
    case MODEL_VDP: return launch_model<ModelVdP>(a, alg, mix, fast, sched, block, grid, s);
    case MODEL_LINEAR: return launch_model<ModelLinear3>(a, alg, mix, fast, sched, block, grid, s);
    default: return hipErrorInvalidValue;
  }
}

int lane_model_dims(int model, int* n, int* np) {
  switch (model) {
    case MODEL_EXP: *n = ModelExp::N; *np = ModelExp::NP; return 0;
    case MODEL_LORENZ: *n = ModelLorenz::N; *np = ModelLorenz::NP; return 0;
    case MODEL_ROBERTSON: *n = ModelRobertson::N; *np = ModelRobertson::NP; return 0;
    case MODEL_VDP: *n = ModelVdP::N; *np = ModelVdP::NP; return 0;
    case MODEL_LINEAR: *n = ModelLinear3::N; *np = ModelLinear3::NP; return 0;
    default: return -1;
  }
}

}  // namespace mideq
