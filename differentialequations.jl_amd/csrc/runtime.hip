// runtime.hip -- the C ABI of libmideq.so (include/mideq.h): contexts,
// argument validation, device buffers, schedule choice and launches.
//
// Boundary replaced: SciMLBase.__solve(::EnsembleProblem, alg,
// ::EnsembleThreads; trajectories, saveat, abstol, reltol) -- the ensemble
// fan-out of test/qa/qa.jl:50,56 -- for problems packed into SoA buffers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mideq.h"
#include "launch.hpp"

using namespace mideq;

struct DevBuf {
  void* ptr = nullptr;
  size_t size = 0;
};

struct mideq_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  mideq_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool ev_pending = false;
  int ncu = 256;
  unsigned long long* d_queue = nullptr;
  DevBuf bufs[16];
};

static int set_err(mideq_ctx* c, int code, const char* fmt, ...) {
  char msg[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                           \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return set_err(ctx, MIDEQ_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static int ensure_buf(mideq_ctx* c, int slot, size_t bytes, void** out) {
  DevBuf& b = c->bufs[slot];
  if (bytes == 0) {
    *out = nullptr;
    return 0;
  }
  if (b.size < bytes) {
    if (b.ptr) {
      hipError_t e = hipFree(b.ptr);
      (void)e;
      b.ptr = nullptr;
      b.size = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess)
      return set_err(c, MIDEQ_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    b.size = bytes;
  }
  *out = b.ptr;
  return 0;
}

extern "C" int mideq_version(void) { return MIDEQ_VERSION; }

extern "C" int mideq_ctx_create(mideq_ctx** out, int32_t device) {
  if (!out) return MIDEQ_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MIDEQ_E_NODEVICE;
  if (device < 0 || device >= ndev) return MIDEQ_E_ARG;
  mideq_ctx* c = new mideq_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return MIDEQ_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->ncu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipMalloc((void**)&c->d_queue, 256) != hipSuccess) {
    delete c;
    return MIDEQ_E_HIP;
  }
  *out = c;
  return MIDEQ_OK;
}

extern "C" void mideq_ctx_destroy(mideq_ctx* c) {
  if (!c) return;
  hipError_t e = hipSetDevice(c->device);
  if (c->stream) e = hipStreamSynchronize(c->stream);
  for (auto& b : c->bufs)
    if (b.ptr) e = hipFree(b.ptr);
  if (c->d_queue) e = hipFree(c->d_queue);
  if (c->ev0) e = hipEventDestroy(c->ev0);
  if (c->ev1) e = hipEventDestroy(c->ev1);
  if (c->stream) e = hipStreamDestroy(c->stream);
  (void)e;
  delete c;
}

extern "C" const char* mideq_last_error(const mideq_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

extern "C" int mideq_last_stats(const mideq_ctx* cc, mideq_stats* out) {
  mideq_ctx* c = const_cast<mideq_ctx*>(cc);
  if (!c || !out) return MIDEQ_E_ARG;
  if (c->ev_pending) {
    HIP_TRY(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.kernel_ms = ms;
    c->ev_pending = false;
  }
  *out = c->stats;
  return MIDEQ_OK;
}

// ---------------------------------------------------------------- validation
static int model_dims(int model, int* n, int* np) {
  if (model == MODEL_BRUSSELATOR) {
    *n = -1;
    *np = 3;
    return 0;
  }
  return lane_model_dims(model, n, np);
}

static int validate(mideq_ctx* c, const mideq_problem* P, const mideq_options* O) {
  if (!P || !O) return set_err(c, MIDEQ_E_ARG, "null problem or options");
  int n, np;
  if (model_dims(P->model, &n, &np) != 0)
    return set_err(c, MIDEQ_E_MODEL, "unknown model id %d", P->model);
  if (P->ntraj < 0) return set_err(c, MIDEQ_E_ARG, "ntraj < 0");
  if (P->model == MODEL_BRUSSELATOR) {
    return set_err(c, MIDEQ_E_UNSUPPORTED, "brusselator kernel not built");
  } else if (P->n != n) {
    return set_err(c, MIDEQ_E_ARG, "model %d has n=%d, got n=%d", P->model, n, P->n);
  }
  if (P->nvec && P->model != MODEL_BRUSSELATOR)
    return set_err(c, MIDEQ_E_UNSUPPORTED, "ragged nvec only for the brusselator model");
  if (P->np < np) return set_err(c, MIDEQ_E_ARG, "model %d needs np>=%d, got %d", P->model, np, P->np);
  if (P->ntraj > 0 && (!P->u0 || !P->p)) return set_err(c, MIDEQ_E_ARG, "null u0 or p");
  if (O->alg < 0 || O->alg > 2) return set_err(c, MIDEQ_E_ARG, "bad alg %d", O->alg);
  if (O->precision < 0 || O->precision > 1)
    return set_err(c, MIDEQ_E_ARG, "bad precision %d", O->precision);
  if (!(O->abstol > 0) || !(O->reltol > 0))
    return set_err(c, MIDEQ_E_ARG, "abstol and reltol must be > 0");
  if (O->ns < 0 || (O->ns > 0 && !O->saveat)) return set_err(c, MIDEQ_E_ARG, "bad saveat");
  if (O->ns > 0 && O->alg != MIDEQ_ALG_TSIT5)
    return set_err(c, MIDEQ_E_UNSUPPORTED, "saveat needs Tsit5 (Rodas5 dense output not provided)");
  if (O->save_dtype < 0 || O->save_dtype > 1) return set_err(c, MIDEQ_E_ARG, "bad save_dtype");
  if (O->schedule < 0 || O->schedule > 2) return set_err(c, MIDEQ_E_ARG, "bad schedule");
  if (O->block != 0 && (O->block < 64 || O->block > 256 || (O->block & 63)))
    return set_err(c, MIDEQ_E_ARG, "block must be 64..256, multiple of 64");
  return MIDEQ_OK;
}

static int check_saveat_host(mideq_ctx* c, const double* sa, int ns) {
  for (int k = 1; k < ns; k++)
    if (!(sa[k] >= sa[k - 1])) return set_err(c, MIDEQ_E_ARG, "saveat must be ascending");
  return MIDEQ_OK;
}

// ---------------------------------------------------------------- launch
static int run_solver(mideq_ctx* c, const mideq_problem* P, const mideq_options* O,
                      const mideq_outputs* out, hipStream_t s) {
  c->stats = mideq_stats{};
  c->ev_pending = false;
  if (P->ntraj == 0) return MIDEQ_OK;
  LaneArgs a{};
  a.ntraj = P->ntraj;
  a.u0 = P->u0;
  a.p = P->p;
  a.t0 = P->t0;
  a.tf = P->tf;
  a.tspan0 = P->tspan0;
  a.tspan1 = P->tspan1;
  a.abstol = O->abstol;
  a.reltol = O->reltol;
  a.dt_fixed = O->dt;
  a.dtmin = O->dtmin;
  a.dtmax = O->dtmax;
  a.maxiters = O->maxiters;
  a.saveat = O->saveat;
  a.ns = O->ns;
  a.save_f32 = O->save_dtype;
  a.dt0 = O->dt0;
  a.alg_hint = O->alg_hint;
  a.qsteady_default = O->qsteady_default;
  a.us = out->us;
  a.u_final = out->u_final;
  a.t_final = out->t_final;
  a.retcode = out->retcode;
  a.naccept = out->naccept;
  a.nreject = out->nreject;
  a.nf = out->nf;
  a.nstiff = out->nstiff;
  a.queue = c->d_queue;
  int sched = O->schedule == MIDEQ_SCHED_AUTO ? SCHED_LANE : O->schedule;
  int block = O->block > 0 ? O->block : 256;
  int grid;
  if (sched == SCHED_PERSISTENT) {
    // 0: sized from the kernel's occupancy by the launcher
    int64_t want = (int64_t)c->ncu * O->blocks_per_cu;
    int64_t need = (P->ntraj + block - 1) / block;
    grid = O->blocks_per_cu > 0 ? (int)(want < need ? want : need) : 0;
  } else {
    grid = (int)((P->ntraj + block - 1) / block);
  }
  HIP_TRY(c, hipEventRecord(c->ev0, s));
  hipError_t e = launch_lane_solver(a, P->model, O->alg, O->precision, sched, block, grid, s);
  if (e != hipSuccess) return set_err(c, MIDEQ_E_HIP, "kernel launch: %s", hipGetErrorString(e));
  HIP_TRY(c, hipEventRecord(c->ev1, s));
  c->ev_pending = true;
  c->stats.schedule = sched;
  c->stats.block = block;
  c->stats.grid = grid;
  c->stats.launches = sched == SCHED_PERSISTENT ? 2 : 1;
  return MIDEQ_OK;
}

extern "C" int mideq_ensemble_solve_device(mideq_ctx* c, const mideq_problem* P,
                                           const mideq_options* O, mideq_outputs* out,
                                           void* stream) {
  if (!c) return MIDEQ_E_ARG;
  int rc = validate(c, P, O);
  if (rc) return rc;
  if (!out) return set_err(c, MIDEQ_E_ARG, "null outputs");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return run_solver(c, P, O, out, s);
}

extern "C" int mideq_ensemble_solve(mideq_ctx* c, const mideq_problem* P,
                                    const mideq_options* O, mideq_outputs* out) {
  if (!c) return MIDEQ_E_ARG;
  int rc = validate(c, P, O);
  if (rc) return rc;
  if (!out) return set_err(c, MIDEQ_E_ARG, "null outputs");
  if (O->ns > 0 && (rc = check_saveat_host(c, O->saveat, O->ns))) return rc;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int64_t N = P->ntraj;
  const int n = P->n;
  if (N == 0) return MIDEQ_OK;
  hipEvent_t t0e, t1e;
  HIP_TRY(c, hipEventCreate(&t0e));
  HIP_TRY(c, hipEventCreate(&t1e));
  HIP_TRY(c, hipEventRecord(t0e, s));

  mideq_problem dP = *P;
  mideq_options dO = *O;
  mideq_outputs dOut{};
  void* d;
  // inputs
  if ((rc = ensure_buf(c, 0, sizeof(double) * (size_t)n * N, &d))) return rc;
  HIP_TRY(c, hipMemcpyAsync(d, P->u0, sizeof(double) * (size_t)n * N, hipMemcpyHostToDevice, s));
  dP.u0 = (const double*)d;
  if ((rc = ensure_buf(c, 1, sizeof(double) * (size_t)P->np * N, &d))) return rc;
  HIP_TRY(c, hipMemcpyAsync(d, P->p, sizeof(double) * (size_t)P->np * N, hipMemcpyHostToDevice, s));
  dP.p = (const double*)d;
  if (P->t0) {
    if ((rc = ensure_buf(c, 2, sizeof(double) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, P->t0, sizeof(double) * N, hipMemcpyHostToDevice, s));
    dP.t0 = (const double*)d;
  }
  if (P->tf) {
    if ((rc = ensure_buf(c, 3, sizeof(double) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, P->tf, sizeof(double) * N, hipMemcpyHostToDevice, s));
    dP.tf = (const double*)d;
  }
  if (P->nvec) {
    if ((rc = ensure_buf(c, 4, sizeof(int32_t) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, P->nvec, sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
    dP.nvec = (const int32_t*)d;
  }
  if (O->ns > 0) {
    if ((rc = ensure_buf(c, 5, sizeof(double) * O->ns, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->saveat, sizeof(double) * O->ns, hipMemcpyHostToDevice, s));
    dO.saveat = (const double*)d;
  }
  if (O->dt0) {
    if ((rc = ensure_buf(c, 6, sizeof(double) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->dt0, sizeof(double) * N, hipMemcpyHostToDevice, s));
    dO.dt0 = (const double*)d;
  }
  if (O->alg_hint) {
    if ((rc = ensure_buf(c, 7, N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->alg_hint, N, hipMemcpyHostToDevice, s));
    dO.alg_hint = (const int8_t*)d;
  }
  // outputs
  size_t us_bytes = (size_t)O->ns * n * N * (O->save_dtype ? 4 : 8);
  if (out->us && O->ns > 0) {
    if ((rc = ensure_buf(c, 8, us_bytes, &dOut.us))) return rc;
  }
  if (out->u_final && (rc = ensure_buf(c, 9, sizeof(double) * (size_t)n * N, (void**)&dOut.u_final))) return rc;
  if (out->t_final && (rc = ensure_buf(c, 10, sizeof(double) * N, (void**)&dOut.t_final))) return rc;
  if (out->retcode && (rc = ensure_buf(c, 11, N, (void**)&dOut.retcode))) return rc;
  if (out->naccept && (rc = ensure_buf(c, 12, 4 * N, (void**)&dOut.naccept))) return rc;
  if (out->nreject && (rc = ensure_buf(c, 13, 4 * N, (void**)&dOut.nreject))) return rc;
  if (out->nf && (rc = ensure_buf(c, 14, 4 * N, (void**)&dOut.nf))) return rc;
  if (out->nstiff && (rc = ensure_buf(c, 15, 4 * N, (void**)&dOut.nstiff))) return rc;

  rc = run_solver(c, &dP, &dO, &dOut, s);
  if (rc) return rc;
  if (out->us && O->ns > 0)
    HIP_TRY(c, hipMemcpyAsync(out->us, dOut.us, us_bytes, hipMemcpyDeviceToHost, s));
  if (out->u_final)
    HIP_TRY(c, hipMemcpyAsync(out->u_final, dOut.u_final, sizeof(double) * (size_t)n * N, hipMemcpyDeviceToHost, s));
  if (out->t_final)
    HIP_TRY(c, hipMemcpyAsync(out->t_final, dOut.t_final, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  if (out->retcode) HIP_TRY(c, hipMemcpyAsync(out->retcode, dOut.retcode, N, hipMemcpyDeviceToHost, s));
  if (out->naccept) HIP_TRY(c, hipMemcpyAsync(out->naccept, dOut.naccept, 4 * N, hipMemcpyDeviceToHost, s));
  if (out->nreject) HIP_TRY(c, hipMemcpyAsync(out->nreject, dOut.nreject, 4 * N, hipMemcpyDeviceToHost, s));
  if (out->nf) HIP_TRY(c, hipMemcpyAsync(out->nf, dOut.nf, 4 * N, hipMemcpyDeviceToHost, s));
  if (out->nstiff) HIP_TRY(c, hipMemcpyAsync(out->nstiff, dOut.nstiff, 4 * N, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipEventRecord(t1e, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_TRY(c, hipEventElapsedTime(&ms, t0e, t1e));
  HIP_TRY(c, hipEventDestroy(t0e));
  HIP_TRY(c, hipEventDestroy(t1e));
  mideq_stats st;
  if ((rc = mideq_last_stats(c, &st))) return rc;
  c->stats.total_ms = ms;
  return MIDEQ_OK;
}
