// bruss.hpp -- uneven batch of 1-D periodic Brusselator instances (state
// size n = 2*Ngrid, 16 <= n <= 128, per-instance tspan) solved by one
// persistent kernel: a team of L lanes per instance pulls instance indices
// from an atomic work queue.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mideq {

constexpr int BRUSS_NMAX = 128;
constexpr int BRUSS_NMIN = 16;

struct BrussArgs {
  int64_t ntraj;
  int32_t nmax;
  int32_t pad0;
  const int32_t* nvec;   // [ntraj] state sizes (NULL: all nmax)
  const double* u0;      // [nmax][ntraj]
  const double* p;       // [3][ntraj]: A, B, alpha
  const double* t0;
  const double* tf;
  double tspan0, tspan1;
  double abstol, reltol;
  double dt_fixed, dtmin, dtmax;
  int64_t maxiters;
  const double* dt0;
  const int8_t* alg_hint;
  int32_t qsteady_default;
  int32_t pad1;
  double* u_final;       // [nmax][ntraj]
  double* t_final;
  int8_t* retcode;
  int32_t* naccept;
  int32_t* nreject;
  int32_t* nf;
  int32_t* nstiff;
  double* dt_first;
  unsigned long long* queue;  // next position in `order`
  const int64_t* order;       // work order (NULL: natural order)
  int64_t nqueue;             // queue length (<= ntraj; a sample when autotuning)
  const double* saveat;       // [ns] ascending save times, or NULL
  void* us;                   // saved values (double or float), or NULL
  int32_t ns;
  int32_t save_f32;           // 1: us holds float
  int32_t save_layout;        // 0: [ns][nmax][ntraj]; 1: [ntraj][ns][nmax]
  int32_t pad2;
};

// Resident teams of the persistent grid for (team shape, alg, precision,
// controller): grid blocks and teams per block.  Returns hipSuccess or an
// error.  Team shapes (lanes, nk): (64, 128), (32, 128) and (32, 64), nk the
// largest state size a team holds.  fast: the division-free log2-domain PI
// controller.
hipError_t bruss_plan(int lanes, int nk, int alg, int mix, int fast, int ncu, int64_t ntraj,
                      int* grid, int* teams_per_block, bool save = false);
// out[k] = order[(k * ntraj) / S], k < S (natural order for a NULL order):
// the autotuner's sample, strided over the launch order.
hipError_t bruss_sample_order(int64_t* out, const int64_t* order, int64_t ntraj, int64_t S,
                              hipStream_t s);
// Launches the solver over the queue of `a`.  Every queued instance must
// have n <= nk (a larger one finishes with RC_INITIALFAILURE).
hipError_t launch_bruss_solver(const BrussArgs& a, int lanes, int nk, int alg, int mix, int fast,
                               int grid, hipStream_t s);
// The builds that write saveat points (a.us and a.ns > 0), compiled in their
// own translation unit: occupancy and launch of one configuration.
hipError_t bruss_occupancy_save(int lanes, int nk, int alg, int mix, int fast, int* per_cu);
hipError_t bruss_launch_save(const BrussArgs& a, int lanes, int nk, int alg, int mix, int fast,
                             int grid, hipStream_t s);

}  // namespace mideq
