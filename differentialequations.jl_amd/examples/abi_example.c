/*
 * abi_example.c -- the C ABI of libmideq.so driven from C, the way the Julia
 * shim's ccalls drive it:
 *   1. a built-in Lorenz ensemble (test/core.jl:22-27) with Tsit5;
 *   2. the same right-hand side as a run-time compiled user model with a
 *      terminating ContinuousCallback at t = 5 and an on-device mean of u[0]
 *      over the trajectories (output_func + reduction), only the mean and
 *      the per-trajectory outputs asked for.
 * Prints "ok" and exits 0 when every check holds.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mideq.h"

#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      fprintf(stderr, __VA_ARGS__);       \
      fprintf(stderr, "\n");              \
      return 1;                           \
    }                                     \
  } while (0)

int main(void) {
  mideq_ctx *ctx = NULL;
  int rc = mideq_ctx_create(&ctx, 0);
  if (rc == MIDEQ_E_NODEVICE) {
    printf("no GPU\n");
    return 3;
  }
  CHECK(rc == MIDEQ_OK, "mideq_ctx_create: %d", rc);
  const int64_t N = 4096;
  double *u0 = calloc(3 * N, sizeof(double)), *p = malloc(3 * N * sizeof(double));
  double *uf = malloc(3 * N * sizeof(double)), *tfin = malloc(N * sizeof(double));
  int8_t *ret = malloc(N);
  for (int64_t i = 0; i < N; i++) {
    u0[i] = 1.0;
    p[i] = 10.0;
    p[N + i] = 28.0 * (1.0 + 0.1 * (double)(i % 7) / 7.0);
    p[2 * N + i] = 8.0 / 3.0;
  }
  mideq_problem P = {0};
  P.model = MIDEQ_MODEL_LORENZ;
  P.n = 3;
  P.np = 3;
  P.ntraj = N;
  P.u0 = u0;
  P.p = p;
  P.tspan0 = 0.0;
  P.tspan1 = 10.0;
  mideq_options O = {0};
  O.alg = MIDEQ_ALG_TSIT5;
  O.abstol = O.reltol = 1e-8;
  O.qsteady_default = 1;
  mideq_outputs out = {0};
  out.u_final = uf;
  out.t_final = tfin;
  out.retcode = ret;
  rc = mideq_ensemble_solve(ctx, &P, &O, &out);
  CHECK(rc == MIDEQ_OK, "solve: %s", mideq_last_error(ctx));
  for (int64_t i = 0; i < N; i++)
    CHECK(ret[i] == MIDEQ_RC_SUCCESS && tfin[i] == 10.0, "trajectory %lld: rc %d t %g",
          (long long)i, ret[i], tfin[i]);

  /* the same model compiled at run time, a callback and an on-device mean */
  mideq_callback cb = {MIDEQ_CB_CONTINUOUS, 0, 1, 10, "t - 5.0", NULL};
  mideq_user_model m = {0};
  m.n = 3;
  m.np = 3;
  m.rhs =
      "du[0] = p[0] * (u[1] - u[0]);\n"
      "du[1] = u[0] * (p[1] - u[2]) - u[1];\n"
      "du[2] = u[0] * u[1] - p[2] * u[2];";
  m.ncallbacks = 1;
  m.callbacks = &cb;
  m.nout = 1;
  m.output = "out[0] = u[0];";
  int32_t id = -1;
  rc = mideq_model_register(ctx, &m, &id);
  CHECK(rc == MIDEQ_OK, "register: %s", mideq_last_error(ctx));
  CHECK(id >= MIDEQ_USER_MODEL_BASE, "model id %d", id);
  P.model = id;
  O.reduction = MIDEQ_REDUCE_MEAN;
  double *vals = malloc(N * sizeof(double)), mean = 0.0;
  mideq_outputs out2 = {0};
  out2.t_final = tfin;
  out2.retcode = ret;
  out2.out = vals;
  out2.reduced = &mean;
  rc = mideq_ensemble_solve(ctx, &P, &O, &out2);
  CHECK(rc == MIDEQ_OK, "solve (user model): %s", mideq_last_error(ctx));
  double s = 0.0;
  for (int64_t i = 0; i < N; i++) {
    CHECK(ret[i] == MIDEQ_RC_TERMINATED, "trajectory %lld: rc %d", (long long)i, ret[i]);
    CHECK(fabs(tfin[i] - 5.0) < 1e-12, "trajectory %lld: t %.17g", (long long)i, tfin[i]);
    s += vals[i];
  }
  CHECK(fabs(mean - s / (double)N) <= 1e-12 * (fabs(mean) + 1.0), "mean %.17g vs %.17g", mean,
        s / (double)N);
  /* a model that does not compile: an error with the compiler's message */
  mideq_user_model bad = {0};
  bad.n = 1;
  bad.np = 1;
  bad.rhs = "du[0] = u[0] +;";
  rc = mideq_model_register(ctx, &bad, &id);
  CHECK(rc == MIDEQ_E_ARG, "bad model accepted (%d)", rc);
  printf("ok (mean u[0](5) = %.6f, error message: %.40s...)\n", mean, mideq_last_error(ctx));
  mideq_ctx_destroy(ctx);
  free(u0);
  free(p);
  free(uf);
  free(tfin);
  free(ret);
  free(vals);
  return 0;
}
