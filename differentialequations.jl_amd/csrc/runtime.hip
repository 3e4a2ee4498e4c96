// runtime.hip -- the C ABI of libmideq.so (include/mideq.h): contexts,
// argument validation, device buffers, schedule choice, the autotuner for
// the large-system kernel, hipGraph replay and launches.
//
// Boundary replaced: SciMLBase.__solve(::EnsembleProblem, alg,
// ::EnsembleThreads; trajectories, saveat, abstol, reltol) -- the ensemble
// fan-out of test/qa/qa.jl:50,56 -- for problems packed into SoA buffers.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <cctype>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/mideq.h"
#include "bruss.hpp"
#include "launch.hpp"

using namespace mideq;

struct DevBuf {
  void* ptr = nullptr;
  size_t size = 0;
};

enum {
  BUF_U0 = 0, BUF_P, BUF_T0, BUF_TF, BUF_NVEC, BUF_SAVEAT, BUF_DT0, BUF_HINT, BUF_US,
  BUF_UF, BUF_TFIN, BUF_RC, BUF_NA, BUF_NR, BUF_NF, BUF_NS, BUF_ORDER,
  BUF_DTF, BUF_NEV, BUF_TEV, BUF_OUT, BUF_RED, BUF_CAP, BUF_EVT, BUF_EVU, BUF_EVN, BUF_TST,
  BUF_COUNT
};

// Device memory a solve uses beyond its inputs and outputs: the work-queue
// counters, the two-pass handoff list, the reduction's block partials and
// the autotuner's sample order.
enum { SCR_HANDOFF = 0, SCR_FOLD, SCR_SAMPLE, SCR_COUNT };

// A run-time compiled right-hand side: its generated model source and the
// kernels built from it so far, one code object per (algorithm, precision,
// controller) configuration.
struct UserKernels {
  hipModule_t mod = nullptr;
  hipFunction_t lane = nullptr, persistent = nullptr;
};
struct UserModel {
  int n = 0, np = 0;
  int ncb = 0;   // callbacks compiled into the model
  int nout = 0;  // output_func values per trajectory
  bool everystep = false;  // save_everystep compiled into the model
  bool tstops = false;     // compiled to honour mideq_options.tstops
  std::string model_src;
  std::map<int, UserKernels> kernels;
};

struct GraphCache {
  std::vector<unsigned char> key;
  hipGraphExec_t exec = nullptr;
};

// The scratch of the solves enqueued on one stream: solves in flight on
// different streams of a context each use their own queue counters, scratch
// buffers and cached graph.
struct StreamSlot {
  hipStream_t stream = nullptr;
  unsigned long long* queue = nullptr;  // [0] persistent queue, [1] handoff count
  DevBuf bufs[SCR_COUNT];
  GraphCache graph;
  uint64_t last_use = 0;
};
constexpr int kMaxSlots = 8;

struct mideq_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  mideq_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t tune0 = nullptr, tune1 = nullptr;  // autotuner timing
  bool ev_pending = false;
  int ncu = 256;
  DevBuf bufs[BUF_COUNT];  // host entry point's copies of the caller's arrays
  std::vector<StreamSlot> slots;
  uint64_t slot_clock = 0;
  // autotuned team shape (lanes per instance) and resident blocks per CU (0:
  // as many as fit), keyed by (alg, precision, controller, nmax)
  std::map<std::tuple<int, int, int, int>, std::pair<int, int>> tuned_lanes;
  std::vector<UserModel> user_models;  // model id MIDEQ_USER_MODEL_BASE + index
  std::string arch;                    // device code target for hipRTC
  // host entry point: pinned staging buffers and streams of the copy
  // workers that move large outputs into pageable caller memory
  struct Stage {
    void* pinned = nullptr;
    hipStream_t s = nullptr;
  };
  std::vector<Stage> stages;
};

// What one solve call uses for its scratch: a stream slot's buffers (grown
// on demand), or a fixed region it may not grow (null slot).
struct Scratch {
  unsigned long long* queue = nullptr;
  StreamSlot* slot = nullptr;
  char* base = nullptr;  // fixed region past the counters
  size_t cap = 0, used = 0;
  GraphCache* graph = nullptr;  // null: no graph replay
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
      // the previous contents may still be in use by work queued on the
      // context's stream (the only stream these buffers are used on)
      hipError_t e = hipStreamSynchronize(c->stream);
      e = hipFree(b.ptr);
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

static void drop_graph(GraphCache* g) {
  if (g->exec) {
    hipError_t e = hipGraphExecDestroy(g->exec);
    (void)e;
  }
  g->exec = nullptr;
  g->key.clear();
}

static void free_slot(StreamSlot& sl) {
  hipError_t e = hipStreamSynchronize(sl.stream);
  drop_graph(&sl.graph);
  for (auto& b : sl.bufs)
    if (b.ptr) e = hipFree(b.ptr);
  if (sl.queue) e = hipFree(sl.queue);
  (void)e;
  sl = StreamSlot{};
}

// The slot of stream s: found, else a new one (the least recently used one
// is recycled past kMaxSlots streams, after its stream has drained).
static int slot_for(mideq_ctx* c, hipStream_t s, StreamSlot** out) {
  StreamSlot* lru = nullptr;
  for (auto& sl : c->slots) {
    if (sl.stream == s) {
      sl.last_use = ++c->slot_clock;
      *out = &sl;
      return MIDEQ_OK;
    }
    if (!lru || sl.last_use < lru->last_use) lru = &sl;
  }
  StreamSlot* sl;
  if ((int)c->slots.size() < kMaxSlots) {
    c->slots.emplace_back();
    sl = &c->slots.back();
  } else {
    free_slot(*lru);
    sl = lru;
  }
  hipError_t e = hipMalloc((void**)&sl->queue, 256);
  if (e != hipSuccess) {
    sl->queue = nullptr;
    return set_err(c, MIDEQ_E_NOMEM, "hipMalloc(256) failed: %s", hipGetErrorString(e));
  }
  sl->stream = s;
  sl->last_use = ++c->slot_clock;
  *out = sl;
  return MIDEQ_OK;
}

// A scratch buffer of `bytes` for this call: from the slot (grown after its
// stream has drained), or carved from the fixed region (256-byte aligned).
static int scratch_get(mideq_ctx* c, Scratch& S, int kind, size_t bytes, void** out) {
  *out = nullptr;
  if (bytes == 0) return MIDEQ_OK;
  if (!S.slot) {
    size_t at = (S.used + 255) & ~(size_t)255;
    if (at > S.cap || bytes > S.cap - at)
      return set_err(c, MIDEQ_E_ARG,
                     "the solve needs %zu bytes of scratch beyond the %zu-byte fixed region",
                     at + bytes, S.cap);
    *out = S.base + at;
    S.used = at + bytes;
    return MIDEQ_OK;
  }
  DevBuf& b = S.slot->bufs[kind];
  if (b.size < bytes) {
    if (b.ptr) {
      hipError_t e = hipStreamSynchronize(S.slot->stream);
      e = hipFree(b.ptr);
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
  return MIDEQ_OK;
}

// ---------------------------------------------------------------- fixed scratch
// The scratch of mideq_ensemble_solve_device_noalloc: a region of the
// library's own device data, one copy per GPU, held until the process ends.
// [0, 256): work-queue counters; then what the call needs of the two-pass
// handoff list and the reduction's block partials.
static __device__ __align__(256) unsigned char scratch[MIDEQ_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];       // device address of `scratch` per device
static hipError_t g_scratch_err[kMaxDevices];  // result of looking it up

// With `device` current: look up the device address of `scratch` on it (the
// array's name is a host-side stand-in; kernels get this address).
static void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess && ptr) g_scratch_ptr[device] = ptr;
}

static int scratch_for_stream(mideq_ctx* c, hipStream_t s, Scratch* S) {
  StreamSlot* sl = nullptr;
  int rc = slot_for(c, s, &sl);
  if (rc) return rc;
  *S = Scratch{};
  S->queue = sl->queue;
  S->slot = sl;
  S->graph = &sl->graph;
  return MIDEQ_OK;
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
  c->arch = "gfx950";
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->ncu = prop.multiProcessorCount;
    // "gfx950:sramecc+:xnack-" -> the processor name
    std::string a = prop.gcnArchName;
    if (!a.empty()) c->arch = a.substr(0, a.find(':'));
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreate(&c->tune0) != hipSuccess || hipEventCreate(&c->tune1) != hipSuccess) {
    mideq_ctx_destroy(c);
    return MIDEQ_E_HIP;
  }
  // the no-alloc entry point's scratch on this device (looked up once per
  // device; a failure is reported by that entry point)
  scratch_setup(device);
  *out = c;
  return MIDEQ_OK;
}

extern "C" void mideq_ctx_destroy(mideq_ctx* c) {
  if (!c) return;
  hipError_t e = hipSetDevice(c->device);
  // only the context's own work is waited for: its stream and the streams
  // its device-entry solves were enqueued on
  for (auto& sl : c->slots) free_slot(sl);
  if (c->stream) e = hipStreamSynchronize(c->stream);
  for (auto& b : c->bufs)
    if (b.ptr) e = hipFree(b.ptr);
  for (auto& um : c->user_models)
    for (auto& kv : um.kernels)
      if (kv.second.mod) e = hipModuleUnload(kv.second.mod);
  for (auto& st : c->stages) {
    if (st.s) e = hipStreamDestroy(st.s);
    if (st.pinned) e = hipHostFree(st.pinned);
  }
  if (c->ev0) e = hipEventDestroy(c->ev0);
  if (c->ev1) e = hipEventDestroy(c->ev1);
  if (c->tune0) e = hipEventDestroy(c->tune0);
  if (c->tune1) e = hipEventDestroy(c->tune1);
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

// ---------------------------------------------------------------- user models
// The solver headers as text (generated from the sources at build time).
static const char* const kRtcSources[] = {
#include "rtc_sources.inc"
};

static UserModel* user_model(mideq_ctx* c, int model) {
  int k = model - MIDEQ_USER_MODEL_BASE;
  if (!c || k < 0 || k >= (int)c->user_models.size()) return nullptr;
  return &c->user_models[k];
}

static int check_user_spec(const mideq_user_model* m, std::string* why) {
  if (!m || !m->rhs) {
    *why = "null model or rhs";
    return MIDEQ_E_ARG;
  }
  if (m->n < 1 || m->n > MIDEQ_USER_NMAX || m->np < 0 || m->np > MIDEQ_USER_NPMAX) {
    char b[128];
    snprintf(b, sizeof b, "user model n=%d np=%d outside 1..%d / 0..%d", m->n, m->np,
             MIDEQ_USER_NMAX, MIDEQ_USER_NPMAX);
    *why = b;
    return MIDEQ_E_ARG;
  }
  if (m->nout < 0 || m->nout > MIDEQ_USER_NOUTMAX || (m->nout > 0 && !m->output)) {
    *why = "bad nout or null output";
    return MIDEQ_E_ARG;
  }
  if (m->with_tstops != 0 && m->with_tstops != 1) {
    *why = "with_tstops must be 0 or 1";
    return MIDEQ_E_ARG;
  }
  if (m->save_everystep != 0 && m->save_everystep != 1) {
    *why = "save_everystep must be 0 or 1";
    return MIDEQ_E_ARG;
  }
  if (m->ncallbacks < 0 || m->ncallbacks > MIDEQ_USER_NCBMAX ||
      (m->ncallbacks > 0 && !m->callbacks)) {
    *why = "bad callback count or null callbacks";
    return MIDEQ_E_ARG;
  }
  for (int k = 0; k < m->ncallbacks; k++) {
    const mideq_callback& cb = m->callbacks[k];
    if (!cb.condition || cb.kind < 0 || cb.kind > 1 || cb.direction < -1 || cb.direction > 1 ||
        cb.interp_points < 0 || cb.interp_points > 1000 || cb.terminate < 0 ||
        cb.terminate > 1 || cb.save_positions < 0 || cb.save_positions > 3) {
      char b[112];
      snprintf(b, sizeof b,
               "callback %d: bad kind, direction, interp_points, save_positions or condition", k);
      *why = b;
      return MIDEQ_E_ARG;
    }
  }
  return MIDEQ_OK;
}

// struct ModelUser: the user's statements as the templated RHS (T = double,
// float, or Dual<N> for the Jacobian when none is given).
// Does the statement text use the identifier t (the time)?  Then the model
// is non-autonomous: f and jac take t, and df/dt comes from dual numbers.
// Comments (// and /* */) and character/string literals are skipped.  The
// statements sit in a block of their own, so a model that declares a local
// t shadows the time parameter (it is then merely treated as
// non-autonomous, with df/dt = 0).
extern "C" int mideq_uses_time(const char* text) {
  if (!text) return 0;
  auto ident = [](char ch) { return std::isalnum((unsigned char)ch) || ch == '_'; };
  for (size_t k = 0; text[k]; k++) {
    const char ch = text[k];
    if (ch == '/' && text[k + 1] == '/') {
      while (text[k + 1] && text[k + 1] != '\n') k++;
    } else if (ch == '/' && text[k + 1] == '*') {
      k += 2;
      while (text[k] && !(text[k] == '*' && text[k + 1] == '/')) k++;
      if (!text[k]) return 0;
      k++;
    } else if (ch == '"' || ch == '\'') {
      for (k++; text[k] && text[k] != ch; k++)
        if (text[k] == '\\' && text[k + 1]) k++;
      if (!text[k]) return 0;
    } else if (ident(ch)) {
      size_t b = k;
      while (ident(text[k + 1])) k++;
      if (k == b && ch == 't' && (b == 0 || text[b - 1] != '.')) return 1;
    }
  }
  return 0;
}
static bool uses_time(const char* text) { return mideq_uses_time(text) != 0; }

static std::string user_model_source(const mideq_user_model* m) {
  const int np = m->np > 0 ? m->np : 1;
  const bool tdep = uses_time(m->rhs) || uses_time(m->jac);
  std::string s = "namespace mideq {\nstruct ModelUser {\n  static constexpr int N = " +
                  std::to_string(m->n) + ", NP = " + std::to_string(np) + ";\n";
  s += std::string("  static constexpr bool TDEP = ") + (tdep ? "true" : "false") + ";\n";
  if (m->save_everystep) s += "  static constexpr bool EVERYSTEP = true;\n";
  if (m->with_tstops) s += "  static constexpr bool TSTOPS = true;\n";
  s += std::string("  template <class T>\n  __device__ static void f(T* du, const T* u, const T* p") +
       (tdep ? ", T t" : "") + ") {\n";
  // the statements in a block of their own: a local t shadows the time
  s += "  {\n#line 1 \"rhs\"\n";
  s += m->rhs;
  s += std::string("\n  }\n  }\n  __device__ static void jac(double (*J)[N], const double* u, "
                   "const double* p") + (tdep ? ", double t" : "") + ") {\n";
  if (m->jac) {
    s += "  {\n#line 1 \"jac\"\n";
    s += m->jac;
    s += "\n  }\n";
  } else if (m->n <= 8) {
    s += "    Dual<N> ud[N], pd[NP], fd[N];\n"
         "    for (int i = 0; i < N; i++) { ud[i].v = u[i]; ud[i].d[i] = 1.0; }\n"
         "    for (int j = 0; j < NP; j++) pd[j].v = p[j];\n";
    s += tdep ? "    f(fd, ud, pd, Dual<N>(t));\n" : "    f(fd, ud, pd);\n";
    s += "    for (int i = 0; i < N; i++)\n"
         "      for (int j = 0; j < N; j++) J[i][j] = fd[i].d[j];\n";
  } else {
    // n > 8: one direction per pass, column by column, in a loop that is
    // not unrolled.  Every partial follows the same operations as the n
    // directions of one Dual<N> pass (the rules act on each partial
    // alone), so J is the same bit for bit; the kernel holds 2n + 2 live
    // dual values instead of n (n + 1), and its Rodas5 kernel compiles in
    // ~30 s instead of ~150 s at n = 16.
    s += "#pragma unroll 1\n"
         "    for (int c = 0; c < N; c++) {\n"
         "      Dual<1> ud[N], pd[NP], fd[N];\n"
         "      for (int i = 0; i < N; i++) { ud[i].v = u[i]; ud[i].d[0] = i == c ? 1.0 : 0.0; }\n"
         "      for (int j = 0; j < NP; j++) pd[j].v = p[j];\n";
    s += tdep ? "      f(fd, ud, pd, Dual<1>(t));\n" : "      f(fd, ud, pd);\n";
    s += "      for (int i = 0; i < N; i++) J[i][c] = fd[i].d[0];\n"
         "    }\n";
  }
  s += "  }\n";
  if (tdep)  // df/dt with the time as the one dual direction
    s += "  __device__ static void tgrad(double* dT, const double* u, const double* p, "
         "double t) {\n"
         "    Dual<1> ud[N], pd[NP], fd[N], td(t);\n"
         "    td.d[0] = 1.0;\n"
         "    for (int i = 0; i < N; i++) ud[i].v = u[i];\n"
         "    for (int j = 0; j < NP; j++) pd[j].v = p[j];\n"
         "    f(fd, ud, pd, td);\n"
         "    for (int i = 0; i < N; i++) dT[i] = fd[i].d[0];\n"
         "  }\n";
  // callbacks: kind / direction / terminate / interp points as switches
  const int ncb = m->ncallbacks;
  if (ncb > 0) {
    s += "  static constexpr int NCB = " + std::to_string(ncb) + ";\n";
    auto table = [&](const char* name, auto field) {
      std::string t = "  __device__ static constexpr int " + std::string(name) +
                      "(int k) {\n    switch (k) {\n";
      for (int k = 0; k < ncb; k++)
        t += "      case " + std::to_string(k) + ": return " +
             std::to_string(field(m->callbacks[k])) + ";\n";
      t += "    }\n    return 0;\n  }\n";
      return t;
    };
    s += table("cb_kind", [](const mideq_callback& c) { return (int)c.kind; });
    s += table("cb_dir", [](const mideq_callback& c) { return (int)c.direction; });
    s += table("cb_term", [](const mideq_callback& c) { return (int)c.terminate; });
    // a continuous callback without an affect (and not terminating: then
    // terminate! is upstream's affect) leaves u unmodified: upstream's
    // apply_callback! returns before its save_positions[2] point
    s += table("cb_save", [](const mideq_callback& c) {
      return (int)(c.kind == 0 && !c.affect && !c.terminate ? (c.save_positions & 1)
                                                            : c.save_positions);
    });
    s += table("cb_interp", [](const mideq_callback& c) {
      return c.interp_points > 0 ? (int)c.interp_points : 1;  // 0: the step end only
    });
    s += "  __device__ static double cb_g(int k, const double* u, const double* p, double t) {\n"
         "    switch (k) {\n";
    for (int k = 0; k < ncb; k++) {
      s += "      case " + std::to_string(k) + ":\n#line 1 \"condition" + std::to_string(k) +
           "\"\n        return (double)(" + std::string(m->callbacks[k].condition) + ");\n";
    }
    s += "    }\n    return 0.0;\n  }\n";
    s += "  __device__ static void cb_affect(int k, double* u, const double* p, double t) {\n"
         "    switch (k) {\n";
    for (int k = 0; k < ncb; k++) {
      s += "      case " + std::to_string(k) + ": {\n";
      if (m->callbacks[k].affect) {
        s += "#line 1 \"affect" + std::to_string(k) + "\"\n";
        s += m->callbacks[k].affect;
        s += "\n";
      }
      s += "      } break;\n";
    }
    s += "    }\n  }\n";
  }
  if (m->nout > 0) {
    s += "  static constexpr int NOUT = " + std::to_string(m->nout) + ";\n";
    s += "  __device__ static void output(double* out, const double* u, const double* p, "
         "double t, int rc) {\n  {\n#line 1 \"output\"\n";
    s += m->output;
    s += "\n  }\n  }\n";
  }
  s += "};\n}  // namespace mideq\n";
  return s;
}

// Compiles the solver for one configuration of a user model: kernels
// mideq_user_lane / mideq_user_persistent.  Returns the code object.
static int rtc_compile(const std::string& model_src, const std::string& arch, int alg, int mix,
                       int fast, std::vector<char>* code, std::string* log) {
  std::string src =
      "typedef signed char int8_t; typedef short int16_t; typedef int int32_t;\n"
      "typedef long long int64_t; typedef unsigned char uint8_t;\n"
      "typedef unsigned int uint32_t; typedef unsigned long long uint64_t;\n";
  for (const char* part : kRtcSources) src += part;
  src += model_src;
  char k[512];
  snprintf(k, sizeof k,
           "extern \"C\" __global__ __launch_bounds__(256) void mideq_user_lane(mideq::LaneArgs a) "
           "{ mideq::lane_kernel_run<mideq::ModelUser, %s, %d, %s>(a); }\n"
           "extern \"C\" __global__ __launch_bounds__(256) void "
           "mideq_user_persistent(mideq::LaneArgs a) "
           "{ mideq::persistent_kernel_run<mideq::ModelUser, %s, %d, %s>(a); }\n",
           mix ? "true" : "false", alg, fast ? "true" : "false", mix ? "true" : "false", alg,
           fast ? "true" : "false");
  src += k;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "mideq_user.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS) {
    *log = "hiprtcCreateProgram failed";
    return MIDEQ_E_HIP;
  }
  // the library's own flags: FP64 results reproducible (only explicit fma
  // fuses), no packed FP32, no machine-level constant hoisting
  std::string arch_opt = "--offload-arch=" + arch;
  const char* opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-mllvm", "-disable-machine-licm"};
  hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  size_t ls = 0;
  if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
    log->assign(ls, '\0');
    if (hiprtcGetProgramLog(prog, &(*log)[0]) != HIPRTC_SUCCESS) log->clear();
  }
  int rc = MIDEQ_OK;
  if (r != HIPRTC_SUCCESS) {
    rc = MIDEQ_E_ARG;
  } else {
    size_t cs = 0;
    if (hiprtcGetCodeSize(prog, &cs) != HIPRTC_SUCCESS || cs == 0) {
      rc = MIDEQ_E_HIP;
    } else {
      code->resize(cs);
      if (hiprtcGetCode(prog, code->data()) != HIPRTC_SUCCESS) rc = MIDEQ_E_HIP;
    }
  }
  (void)hiprtcDestroyProgram(&prog);
  return rc;
}

static int user_kernels(mideq_ctx* c, UserModel* um, int alg, int mix, int fast,
                        UserKernels** out, bool may_build = true) {
  int key = (alg << 2) | (mix << 1) | fast;
  auto it = um->kernels.find(key);
  if (it == um->kernels.end()) {
    if (!may_build)
      return set_err(c, MIDEQ_E_ARG,
                     "the user model's kernels for this configuration are not built yet");
    std::vector<char> code;
    std::string log;
    int rc = rtc_compile(um->model_src, c->arch, alg, mix, fast, &code, &log);
    if (rc) return set_err(c, rc, "user model compile: %.400s", log.c_str());
    UserKernels uk;
    HIP_TRY(c, hipModuleLoadData(&uk.mod, code.data()));
    HIP_TRY(c, hipModuleGetFunction(&uk.lane, uk.mod, "mideq_user_lane"));
    HIP_TRY(c, hipModuleGetFunction(&uk.persistent, uk.mod, "mideq_user_persistent"));
    it = um->kernels.emplace(key, uk).first;
  }
  *out = &it->second;
  return MIDEQ_OK;
}

// Template ALG of a lane launch: Rosenbrock methods with saveat use the
// build with the dense output.
static int lane_alg_code(const LaneArgs& a, int alg, bool callbacks = false) {
  const bool dense = (a.ns > 0 && a.us) || callbacks;
  return (dense && alg != ALG_TSIT5 && alg != ALG_AUTO_T) ? (alg | ALG_DENSE) : alg;
}

// The lane-kernel launch of launch_lane_solver for a user model whose
// kernels for this configuration are built (user_kernels).
static hipError_t launch_user(const UserKernels* uk, const LaneArgs& a, int sched, int block,
                              int grid, int ncu, hipStream_t s) {
  const unsigned lds = (unsigned)lane_lds_bytes(a.ns);
  LaneArgs arg = a;
  void* params[] = {&arg};
  if (sched == SCHED_PERSISTENT) {
    if (grid <= 0) {
      int per_cu = 0;
      if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, uk->persistent, block,
                                                             lds) != hipSuccess ||
          per_cu < 1)
        per_cu = 1;
      int64_t need = (a.ntraj + block - 1) / block;
      int64_t g = (int64_t)per_cu * ncu;
      grid = (int)(g < need ? g : need);
    }
    hipError_t e = launch_set_u64(a.queue, 0ull, s);
    if (e != hipSuccess) return e;
    return hipModuleLaunchKernel(uk->persistent, grid, 1, 1, block, 1, 1, lds, s, params,
                                 nullptr);
  }
  return hipModuleLaunchKernel(uk->lane, grid, 1, 1, block, 1, 1, lds, s, params, nullptr);
}

extern "C" int mideq_model_check(const mideq_user_model* m, char* log, size_t log_size) {
  std::string why;
  int rc = check_user_spec(m, &why);
  if (rc == MIDEQ_OK) {
    std::vector<char> code;
    // AutoTsit5 with FP32 stages instantiates f for double, float and (for
    // the Jacobian) dual numbers
    rc = rtc_compile(user_model_source(m), "gfx950",
                     ALG_AUTO | (m->ncallbacks > 0 ? ALG_DENSE : 0), 1, 0, &code, &why);
  }
  if (log && log_size) {
    snprintf(log, log_size, "%s", why.c_str());
  }
  return rc;
}

extern "C" int mideq_model_register(mideq_ctx* c, const mideq_user_model* m,
                                    int32_t* model_id) {
  if (!c || !model_id) return MIDEQ_E_ARG;
  std::string why;
  int rc = check_user_spec(m, &why);
  if (rc) return set_err(c, rc, "%s", why.c_str());
  HIP_TRY(c, hipSetDevice(c->device));
  UserModel um;
  um.n = m->n;
  um.np = m->np > 0 ? m->np : 1;  // the kernel reads at least one parameter row
  um.ncb = m->ncallbacks;
  um.nout = m->nout;
  um.everystep = m->save_everystep != 0;
  um.tstops = m->with_tstops != 0;
  um.model_src = user_model_source(m);
  c->user_models.push_back(um);
  UserModel* p = &c->user_models.back();
  // compile one configuration now, so that errors surface here: AutoTsit5
  // with FP32 stages instantiates f for double, float and dual numbers
  UserKernels* uk = nullptr;
  rc = user_kernels(c, p, ALG_AUTO | (p->ncb > 0 ? ALG_DENSE : 0), 1, 0, &uk);
  if (rc) {
    for (auto& kv : p->kernels)
      if (kv.second.mod) (void)hipModuleUnload(kv.second.mod);
    c->user_models.pop_back();
    return rc;
  }
  *model_id = MIDEQ_USER_MODEL_BASE + (int32_t)(c->user_models.size() - 1);
  return MIDEQ_OK;
}

// ---------------------------------------------------------------- validation
static int model_dims(mideq_ctx* c, int model, int* n, int* np) {
  if (model == MODEL_BRUSSELATOR) {
    *n = -1;
    *np = 3;
    return 0;
  }
  if (UserModel* um = user_model(c, model)) {
    *n = um->n;
    *np = um->np;
    return 0;
  }
  return lane_model_dims(model, n, np);
}

static int validate(mideq_ctx* c, const mideq_problem* P, const mideq_options* O) {
  if (!P || !O) return set_err(c, MIDEQ_E_ARG, "null problem or options");
  if (O->ntstops < 0 || (O->ntstops > 0 && !O->tstops))
    return set_err(c, MIDEQ_E_ARG, "ntstops < 0, or tstops NULL with ntstops > 0");
  if (O->ntstops > 0) {
    const UserModel* um = user_model(c, P->model);
    if (!um || !um->tstops)
      return set_err(c, MIDEQ_E_UNSUPPORTED,
                     "tstops need a user model registered with with_tstops = 1");
  }
  int n, np;
  if (model_dims(c, P->model, &n, &np) != 0)
    return set_err(c, MIDEQ_E_MODEL, "unknown model id %d", P->model);
  if (P->ntraj < 0) return set_err(c, MIDEQ_E_ARG, "ntraj < 0");
  if (P->model == MODEL_BRUSSELATOR) {
    if (P->n < BRUSS_NMIN || P->n > BRUSS_NMAX || (P->n & 1))
      return set_err(c, MIDEQ_E_ARG, "brusselator n=%d must be even in [%d,%d]", P->n,
                     BRUSS_NMIN, BRUSS_NMAX);
    if (O->lanes != 0 && O->lanes != 32 && O->lanes != 64)
      return set_err(c, MIDEQ_E_ARG, "lanes must be 0, 32 or 64");
  } else if (P->n != n) {
    return set_err(c, MIDEQ_E_ARG, "model %d has n=%d, got n=%d", P->model, n, P->n);
  }
  if (P->nvec && P->model != MODEL_BRUSSELATOR)
    return set_err(c, MIDEQ_E_UNSUPPORTED, "ragged nvec only for the brusselator model");
  if (P->np < np) return set_err(c, MIDEQ_E_ARG, "model %d needs np>=%d, got %d", P->model, np, P->np);
  if (P->ntraj > 0 && (!P->u0 || !P->p)) return set_err(c, MIDEQ_E_ARG, "null u0 or p");
  if (O->alg < 0 || O->alg > 4) return set_err(c, MIDEQ_E_ARG, "bad alg %d", O->alg);
  if (O->precision < 0 || O->precision > 1)
    return set_err(c, MIDEQ_E_ARG, "bad precision %d", O->precision);
  if (!(O->abstol > 0) || !(O->reltol > 0))
    return set_err(c, MIDEQ_E_ARG, "abstol and reltol must be > 0");
  if (O->ns < 0 || (O->ns > 0 && !O->saveat)) return set_err(c, MIDEQ_E_ARG, "bad saveat");
  if (O->ns > 0 && (int64_t)P->n * P->ntraj >= ((int64_t)1 << 32))
    return set_err(c, MIDEQ_E_UNSUPPORTED, "saveat needs n * ntraj < 2^32 (got %lld); split the "
                   "ensemble", (long long)((int64_t)P->n * P->ntraj));
  if (O->save_dtype < 0 || O->save_dtype > 1) return set_err(c, MIDEQ_E_ARG, "bad save_dtype");
  if (O->save_layout < 0 || O->save_layout > 1) return set_err(c, MIDEQ_E_ARG, "bad save_layout");
  if (O->save_layout == 1 && O->ns > 0 && (int64_t)O->ns * P->n >= ((int64_t)1 << 32))
    return set_err(c, MIDEQ_E_UNSUPPORTED, "save_layout 1 needs ns * n < 2^32");
  if (O->schedule < 0 || O->schedule > 2) return set_err(c, MIDEQ_E_ARG, "bad schedule");
  if (O->controller < 0 || O->controller > 1) return set_err(c, MIDEQ_E_ARG, "bad controller");
  if (O->auto_passes < 0 || O->auto_passes > 3) return set_err(c, MIDEQ_E_ARG, "bad auto_passes");
  if (O->reduction < 0 || O->reduction > 4) return set_err(c, MIDEQ_E_ARG, "bad reduction");
  if (O->reduction != 0) {
    UserModel* um = user_model(c, P->model);
    if (!um || um->nout == 0)
      return set_err(c, MIDEQ_E_ARG, "a reduction needs a user model with an output");
  }
  if (O->block != 0 && (O->block < 64 || O->block > 256 || (O->block & 63)))
    return set_err(c, MIDEQ_E_ARG, "block must be 64..256, multiple of 64");
  return MIDEQ_OK;
}

static int check_saveat_host(mideq_ctx* c, const double* sa, int ns) {
  for (int k = 1; k < ns; k++)
    if (!(sa[k] >= sa[k - 1])) return set_err(c, MIDEQ_E_ARG, "saveat must be ascending");
  return MIDEQ_OK;
}

// Enqueue a launch sequence, optionally through a cached hipGraph whose key
// is the exact argument bytes of the sequence plus the stream.
template <class F>
static int enqueue(mideq_ctx* c, GraphCache* gc, bool use_graph, const void* key, size_t keylen,
                   hipStream_t s, F&& body) {
  c->stats.graph = 0;
  if (use_graph && !gc)
    return set_err(c, MIDEQ_E_ARG, "graph replay is not available on this entry point");
  if (!use_graph) {
    hipError_t e = body(s);
    if (e != hipSuccess) return set_err(c, MIDEQ_E_HIP, "kernel launch: %s", hipGetErrorString(e));
    return MIDEQ_OK;
  }
  const unsigned char* kb = (const unsigned char*)key;
  std::vector<unsigned char> k(kb, kb + keylen);
  const unsigned char* sb = (const unsigned char*)&s;
  k.insert(k.end(), sb, sb + sizeof(s));
  if (!gc->exec || gc->key != k) {
    drop_graph(gc);
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = body(s);
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return set_err(c, MIDEQ_E_HIP, "graph capture: %s / %s", hipGetErrorString(e),
                     hipGetErrorString(e2));
    }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_err(c, MIDEQ_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
    gc->exec = ex;
    gc->key = k;
  }
  HIP_TRY(c, hipGraphLaunch(gc->exec, s));
  c->stats.graph = 1;
  return MIDEQ_OK;
}

// ---------------------------------------------------------------- brusselator
// Team shape for a lanes choice: 32-lane teams hold (u, v) pairs for n <= 64
// in the registers of the 64-lane shape when the batch allows.
static int bruss_nk(int lanes, int nmax) {
  return (lanes == 32 && nmax <= BRUSS_NMAX / 2) ? BRUSS_NMAX / 2 : BRUSS_NMAX;
}

// Persistent blocks per CU the kernel's occupancy allows for this shape.
static int bruss_blocks_per_cu(mideq_ctx* c, const BrussArgs& b, int lanes, int alg, int mix,
                               int fast) {
  int grid = 0, tpb = 0;
  if (bruss_plan(lanes, bruss_nk(lanes, b.nmax), alg, mix, fast, c->ncu, INT64_MAX / 2, &grid,
                 &tpb, b.us && b.ns > 0) != hipSuccess)
    return 1;
  return std::max(1, grid / c->ncu);
}

// bpc > 0: at most bpc resident blocks per CU (fewer teams in flight);
// 0: as many as the occupancy allows.
static int bruss_launch_with(mideq_ctx* c, GraphCache* gc, BrussArgs b, int lanes, int alg,
                             int mix, int fast, bool use_graph, hipStream_t s, int* grid_used,
                             int bpc = 0) {
  int grid = 0, tpb = 0;
  const int nk = bruss_nk(lanes, b.nmax);
  const bool save = b.us && b.ns > 0;
  hipError_t e = bruss_plan(lanes, nk, alg, mix, fast, c->ncu, b.nqueue, &grid, &tpb, save);
  if (e != hipSuccess) return set_err(c, MIDEQ_E_HIP, "bruss_plan: %s", hipGetErrorString(e));
  if (bpc > 0) grid = std::max(1, std::min(grid, bpc * c->ncu));
  *grid_used = grid;
  struct Key {
    BrussArgs b;
    int lanes, alg, mix, fast, grid;
  } key{b, lanes, alg, mix, fast, grid};
  return enqueue(c, gc, use_graph, &key, sizeof key, s, [&](hipStream_t st) {
    return launch_bruss_solver(b, lanes, nk, alg, mix, fast, grid, st);
  });
}

static int run_bruss(mideq_ctx* c, const mideq_problem* P, const mideq_options* O,
                     const mideq_outputs* out, const int64_t* d_order, hipStream_t s,
                     Scratch& Sc) {
  BrussArgs b{};
  b.ntraj = P->ntraj;
  b.nqueue = P->ntraj;
  b.nmax = P->n;
  b.nvec = P->nvec;
  b.u0 = P->u0;
  b.p = P->p;
  b.t0 = P->t0;
  b.tf = P->tf;
  b.tspan0 = P->tspan0;
  b.tspan1 = P->tspan1;
  b.abstol = O->abstol;
  b.reltol = O->reltol;
  b.dt_fixed = O->dt;
  b.dtmin = O->dtmin;
  b.dtmax = O->dtmax;
  b.maxiters = O->maxiters;
  b.dt0 = O->dt0;
  b.alg_hint = O->alg_hint;
  b.qsteady_default = O->qsteady_default;
  b.u_final = out->u_final;
  b.t_final = out->t_final;
  b.retcode = out->retcode;
  b.naccept = out->naccept;
  b.nreject = out->nreject;
  b.nf = out->nf;
  b.nstiff = out->nstiff;
  b.dt_first = out->dt_first;
  b.queue = Sc.queue;
  b.order = d_order;
  if (O->ns > 0 && out->us) {
    b.saveat = O->saveat;
    b.us = out->us;
    b.ns = O->ns;
    b.save_f32 = O->save_dtype;
    b.save_layout = O->save_layout;
  }
  int lanes = O->lanes;
  int bpc = O->blocks_per_cu > 0 ? O->blocks_per_cu : 0;
  int grid = 0;
  if (lanes == 0) {
    // autotuner: time each candidate team shape (64- or 32-lane teams) and
    // residency (as many blocks per CU as fit, or one fewer: fewer teams
    // contending for the SIMDs) on a strided sample of the batch and keep
    // the fastest for this (alg, precision, nmax)
    auto key = std::make_tuple((int)O->alg, (int)O->precision, (int)O->controller, (int)P->n);
    auto it = c->tuned_lanes.find(key);
    if (it != c->tuned_lanes.end()) {
      lanes = it->second.first;
      if (!bpc) bpc = it->second.second;
    } else if (!Sc.slot) {
      // fixed scratch (no allocation, no tuning run): 64-lane teams
      lanes = 64;
    } else {
      // The sample is strided over the launch order and, for Tsit5, holds
      // ~8 instances per resident team: with fewer, the run is its tail
      // (the last instances' latency), which favours the shape with one
      // instance per wavefront over the one with the higher throughput.
      // Rosenbrock steps on n up to 128 cost ~500 Tsit5 steps: 1,024.
      int64_t S = 1024;
      if (O->alg == MIDEQ_ALG_TSIT5) {
        int64_t resident = 0;
        for (int cand : {64, 32}) {
          int g = 0, tpb = 0;
          if (bruss_plan(cand, bruss_nk(cand, P->n), O->alg, O->precision, O->controller,
                         c->ncu, INT64_MAX / 2, &g, &tpb, b.us != nullptr) == hipSuccess)
            resident = std::max<int64_t>(resident, (int64_t)g * tpb);
        }
        S = std::max<int64_t>(8192, 8 * resident);
      }
      S = std::min<int64_t>(S, P->ntraj);
      void* d;
      int rc = scratch_get(c, Sc, SCR_SAMPLE, sizeof(int64_t) * S, &d);
      if (rc) return rc;
      {
        hipError_t e = bruss_sample_order((int64_t*)d, d_order, P->ntraj, S, s);
        if (e != hipSuccess)
          return set_err(c, MIDEQ_E_HIP, "sample order: %s", hipGetErrorString(e));
      }
      BrussArgs bs = b;
      bs.order = (const int64_t*)d;
      bs.nqueue = S;
      float best = 1e30f;
      int best_l = 64, best_b = 0;
      bool warmed = false;
      for (int cand : {64, 32}) {
        const int full = bruss_blocks_per_cu(c, bs, cand, O->alg, O->precision, O->controller);
        // residencies: all that fit (0), one block per CU fewer; the
        // caller's blocks_per_cu, when given, is the only one.  Compared
        // only when the sample holds >= 4 instances per resident team at
        // full residency: with fewer, the full one idles teams in the
        // sample's tail and loses there what it gains on the batch.
        const int64_t teams = (int64_t)full * c->ncu * (256 / cand);
        const int cands[2] = {0, full - 1};
        const int ncand = (bpc > 0 || full < 2 || S < 4 * teams) ? 1 : 2;
        for (int ci = 0; ci < ncand; ci++) {
          const int use_b = bpc > 0 ? bpc : cands[ci];
          int g = 0;
          // one untimed run first (clocks up, caches warm); the candidates
          // after it are timed on their first run (repeats of a candidate
          // differ by ~1 %, the candidates by 10-40 %)
          if (!warmed) {
            if ((rc = bruss_launch_with(c, nullptr, bs, cand, O->alg, O->precision,
                                        O->controller, false, s, &g, use_b)))
              return rc;
            warmed = true;
          }
          HIP_TRY(c, hipEventRecord(c->tune0, s));
          if ((rc = bruss_launch_with(c, nullptr, bs, cand, O->alg, O->precision, O->controller,
                                      false, s, &g, use_b)))
            return rc;
          HIP_TRY(c, hipEventRecord(c->tune1, s));
          HIP_TRY(c, hipEventSynchronize(c->tune1));
          float ms = 0.f;
          HIP_TRY(c, hipEventElapsedTime(&ms, c->tune0, c->tune1));
          if (ms < best) {
            best = ms;
            best_l = cand;
            best_b = use_b;
          }
        }
      }
      lanes = best_l;
      if (!bpc) bpc = best_b;
      c->tuned_lanes[key] = {lanes, best_b};
    }
  }
  HIP_TRY(c, hipEventRecord(c->ev0, s));
  int rc = bruss_launch_with(c, Sc.graph, b, lanes, O->alg, O->precision, O->controller,
                             O->graph != 0, s, &grid, bpc);
  if (rc) return rc;
  HIP_TRY(c, hipEventRecord(c->ev1, s));
  c->ev_pending = true;
  c->stats.schedule = MIDEQ_SCHED_PERSISTENT;
  c->stats.block = 256;
  c->stats.grid = grid;
  c->stats.launches = 2;
  c->stats.lanes = lanes;
  return MIDEQ_OK;
}

// ---------------------------------------------------------------- launch
static int run_solver(mideq_ctx* c, const mideq_problem* P, const mideq_options* O,
                      const mideq_outputs* out, const int64_t* d_order, hipStream_t s,
                      Scratch& Sc) {
  c->stats = mideq_stats{};
  c->ev_pending = false;
  if (P->ntraj == 0) return MIDEQ_OK;
  if (P->model == MODEL_BRUSSELATOR) return run_bruss(c, P, O, out, d_order, s, Sc);
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
  a.save_layout = O->save_layout;
  a.us = out->us;
  a.u_final = out->u_final;
  a.t_final = out->t_final;
  a.retcode = out->retcode;
  a.naccept = out->naccept;
  a.nreject = out->nreject;
  a.nf = out->nf;
  a.nstiff = out->nstiff;
  a.dt_first = out->dt_first;
  a.nevent = out->nevent;
  a.t_event = out->t_event;
  a.ev_t = out->ev_t;
  a.ev_u = out->ev_u;
  a.ev_n = out->ev_n;
  a.ev_cap = out->ev_cap;
  a.tstops = O->ntstops > 0 ? O->tstops : nullptr;
  a.nts = O->ntstops > 0 ? O->ntstops : 0;
  a.out = out->out;
  a.queue = Sc.queue;
  a.order = d_order;
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
  // AutoSwitch in two passes: a Tsit5 pass with the stiffness test solves
  // every trajectory that never switches and lists the rest, which the full
  // AutoSwitch kernel then solves from the start.  The Tsit5-only kernel is
  // leaner (no Rosenbrock code or registers), and the second pass runs only
  // the switching trajectories side by side; results equal the single pass.
  // Opt-in: the Tsit5 pass re-solves the non-stiff start of every switching
  // trajectory, so it pays only when few trajectories switch.
  const bool two_pass = (O->alg == MIDEQ_ALG_AUTO || O->alg == MIDEQ_ALG_AUTO_RODAS5P) &&
                        O->auto_passes >= 2;
  // the first pass: Tsit5 with the stiffness test (2), or plain Tsit5 for the
  // trajectories hinted non-stiff, capped per trajectory (3)
  const int first_alg = O->auto_passes == 3 ? ALG_TRUST : ALG_AUTO_T;
  a.iter_cap = O->auto_passes == 3 ? O->iter_cap : nullptr;
  if (two_pass) {
    void* d;
    int rc0 = scratch_get(c, Sc, SCR_HANDOFF, sizeof(int64_t) * (size_t)P->ntraj, &d);
    if (rc0) return rc0;
    a.handoff = (int64_t*)d;
    a.handoff_count = Sc.queue + 1;  // queue[0] is the persistent queue
  }
  HIP_TRY(c, hipEventRecord(c->ev0, s));
  struct Key {
    LaneArgs a;
    int model, alg, prec, ctl, sched, block, grid, two_pass, first_alg;
  } key{a, P->model, O->alg, O->precision, O->controller, sched, block, grid, two_pass,
        first_alg};
  // user model: build its kernels for this configuration before any capture
  UserModel* um = user_model(c, P->model);
  UserKernels* uk_main = nullptr;
  UserKernels* uk_first = nullptr;
  if (um) {
    const bool may_build = Sc.slot != nullptr;
    int rc0 = user_kernels(c, um, lane_alg_code(a, O->alg, um->ncb > 0), O->precision,
                           O->controller, &uk_main, may_build);
    if (rc0 == MIDEQ_OK && two_pass)
      rc0 = user_kernels(c, um, first_alg, O->precision, O->controller, &uk_first, may_build);
    if (rc0) return rc0;
  }
  // on-device reduction of the output values: block partials in scratch
  double* fold_partial = nullptr;
  const bool fold = O->reduction != 0 && um && um->nout > 0;
  if (fold) {
    if (!out->out || !out->reduced)
      return set_err(c, MIDEQ_E_ARG, "a reduction needs the out and reduced buffers");
    void* d;
    int rc0 = scratch_get(c, Sc, SCR_FOLD,
                         sizeof(double) * (size_t)um->nout * (size_t)((P->ntraj + 255) / 256),
                         &d);
    if (rc0) return rc0;
    fold_partial = (double*)d;
  }
  auto launch = [&](const LaneArgs& la, int alg, hipStream_t st) -> hipError_t {
    if (um)
      return launch_user(alg == first_alg ? uk_first : uk_main, la, sched, block, grid, c->ncu,
                         st);
    return launch_lane_solver(la, P->model, alg, O->precision, O->controller, sched, block, grid,
                              st);
  };
  auto finish_fold = [&](hipError_t e, hipStream_t st) -> hipError_t {
    if (e != hipSuccess || !fold) return e;
    return launch_fold(out->out, P->ntraj, um->nout, O->reduction, fold_partial, out->reduced,
                       st);
  };
  int rc = enqueue(c, Sc.graph, O->graph != 0, &key, sizeof key, s, [&](hipStream_t st) {
    if (!two_pass) {
      LaneArgs a1 = a;
      a1.handoff = nullptr;
      a1.handoff_count = nullptr;
      return finish_fold(launch(a1, O->alg, st), st);
    }
    hipError_t e = launch_set_u64(a.handoff_count, 0ull, st);
    if (e != hipSuccess) return e;
    e = launch(a, first_alg, st);
    if (e != hipSuccess) return e;
    LaneArgs a2 = a;
    a2.order = a.handoff;
    a2.order_len = a.handoff_count;
    a2.handoff = nullptr;
    a2.handoff_count = nullptr;
    return finish_fold(launch(a2, O->alg, st), st);
  });
  if (rc) return rc;
  HIP_TRY(c, hipEventRecord(c->ev1, s));
  c->ev_pending = true;
  c->stats.schedule = sched;
  c->stats.block = block;
  c->stats.grid = grid;
  c->stats.launches = (sched == SCHED_PERSISTENT ? 2 : 1) * (two_pass ? 2 : 1) + (two_pass ? 1 : 0);
  c->stats.lanes = 1;
  return MIDEQ_OK;
}

// Event points (mideq_outputs.ev_*): user models with callbacks only (their
// kernels are the ones that apply affects), both arrays or neither.
static int validate_outputs(mideq_ctx* c, const mideq_problem* P, const mideq_outputs* out) {
  if (!out) return set_err(c, MIDEQ_E_ARG, "null outputs");
  if (!out->ev_t && !out->ev_u) return MIDEQ_OK;
  if (!out->ev_t || !out->ev_u)
    return set_err(c, MIDEQ_E_ARG, "ev_t and ev_u are given together or not at all");
  if (out->ev_cap < 1) return set_err(c, MIDEQ_E_ARG, "ev_cap must be >= 1 with ev_t");
  const UserModel* um = user_model(c, P->model);
  if (!um || (um->ncb == 0 && !um->everystep))
    return set_err(c, MIDEQ_E_UNSUPPORTED,
                   "extra points (ev_t) need a user model with callbacks or save_everystep");
  return MIDEQ_OK;
}

extern "C" int mideq_ensemble_solve_device(mideq_ctx* c, const mideq_problem* P,
                                           const mideq_options* O, mideq_outputs* out,
                                           void* stream) {
  if (!c) return MIDEQ_E_ARG;
  int rc = validate(c, P, O);
  if (rc) return rc;
  if ((rc = validate_outputs(c, P, out))) return rc;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  Scratch Sc;
  if ((rc = scratch_for_stream(c, s, &Sc))) return rc;
  return run_solver(c, P, O, out, O->order, s, Sc);
}

extern "C" int mideq_ensemble_solve_device_noalloc(mideq_ctx* c, const mideq_problem* P,
                                                   const mideq_options* O, mideq_outputs* out,
                                                   void* stream) {
  if (!c) return MIDEQ_E_ARG;
  int rc = validate(c, P, O);
  if (rc) return rc;
  if ((rc = validate_outputs(c, P, out))) return rc;
  if (O->graph)
    return set_err(c, MIDEQ_E_ARG,
                   "graph = 1 instantiates a graph, which allocates; capture this call in a "
                   "graph of the caller's instead");
  void* base = nullptr;
  hipError_t se = hipErrorInvalidDevice;
  if (c->device >= 0 && c->device < kMaxDevices) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[c->device];
    se = g_scratch_err[c->device];
  }
  if (!base)
    return set_err(c, MIDEQ_E_HIP, "the fixed scratch of device %d is unavailable: %s",
                   c->device, se == hipSuccess ? "not looked up" : hipGetErrorString(se));
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  Scratch Sc;
  Sc.queue = (unsigned long long*)base;
  Sc.base = (char*)base + 256;
  Sc.cap = MIDEQ_NOALLOC_SCRATCH_BYTES - 256;
  return run_solver(c, P, O, out, O->order, s, Sc);
}

// A launch order must be a permutation of 0..N-1.
static int check_order_host(mideq_ctx* c, const int64_t* order, int64_t N) {
  std::vector<unsigned char> seen((size_t)N, 0);
  for (int64_t k = 0; k < N; k++) {
    int64_t i = order[k];
    if (i < 0 || i >= N || seen[(size_t)i])
      return set_err(c, MIDEQ_E_ARG, "order is not a permutation of 0..ntraj-1 (entry %lld = %lld)",
                     (long long)k, (long long)i);
    seen[(size_t)i] = 1;
  }
  return MIDEQ_OK;
}

// Device -> host copies of large outputs into pageable caller memory (numpy,
// Julia arrays): a plain hipMemcpy from pageable memory moves ~15 GB/s.
// kStageWorkers host threads, each with a pinned buffer of kStageBytes and a
// stream of its own, take chunks in turn: a DMA into the pinned buffer, then
// a memcpy into the caller's array while the other workers' DMAs run.
// Pinned destinations and small copies go straight through on `s`.  Returns
// once the data is in `dst`.
constexpr int kStageWorkers = 8;
constexpr size_t kStageBytes = size_t(32) << 20;
constexpr size_t kStageMin = size_t(64) << 20;

static bool host_pinned(const void* p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory may be reported as an error
    return false;
  }
  return at.type == hipMemoryTypeHost;
}

static int copy_to_host(mideq_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < kStageMin || host_pinned(dst)) {
    HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    return MIDEQ_OK;
  }
  if (c->stages.empty()) {
    // all workers or none: a failure part way frees what was made
    std::vector<mideq_ctx::Stage> made(kStageWorkers);
    hipError_t e = hipSuccess;
    for (auto& st : made) {
      if ((e = hipHostMalloc(&st.pinned, kStageBytes, hipHostMallocDefault)) != hipSuccess) break;
      if ((e = hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking)) != hipSuccess) break;
    }
    if (e != hipSuccess) {
      for (auto& st : made) {
        if (st.s) (void)hipStreamDestroy(st.s);
        if (st.pinned) (void)hipHostFree(st.pinned);
      }
      return set_err(c, MIDEQ_E_HIP, "staging buffers: %s", hipGetErrorString(e));
    }
    c->stages = std::move(made);
  }
  // the workers' DMAs start once everything enqueued on s so far is done
  hipEvent_t ready;
  HIP_TRY(c, hipEventCreateWithFlags(&ready, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ready, s);
  const size_t nchunk = (bytes + kStageBytes - 1) / kStageBytes;
  std::atomic<size_t> next{0};
  std::atomic<int> bad{e == hipSuccess ? 0 : (int)e};
  auto work = [&](int w) {
    const mideq_ctx::Stage& st = c->stages[(size_t)w];
    if (hipSetDevice(c->device) != hipSuccess || hipStreamWaitEvent(st.s, ready, 0) != hipSuccess) {
      bad = 1;
      return;
    }
    for (size_t k; !bad && (k = next++) < nchunk;) {
      const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
      hipError_t r = hipMemcpyAsync(st.pinned, (const char*)src + off, len,
                                    hipMemcpyDeviceToHost, st.s);
      if (r == hipSuccess) r = hipStreamSynchronize(st.s);
      if (r != hipSuccess) {
        bad = (int)r;
        return;
      }
      memcpy((char*)dst + off, st.pinned, len);
    }
  };
  if (!bad) {
    std::vector<std::thread> th;
    const int nw = (int)std::min<size_t>(kStageWorkers, nchunk);
    for (int w = 1; w < nw; w++) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
  }
  const int b = bad;
  HIP_TRY(c, hipEventDestroy(ready));
  if (b) return set_err(c, MIDEQ_E_HIP, "staged copy to host memory failed: %s",
                        hipGetErrorString((hipError_t)b));
  return MIDEQ_OK;
}

extern "C" int mideq_ensemble_solve(mideq_ctx* c, const mideq_problem* P,
                                    const mideq_options* O, mideq_outputs* out) {
  if (!c) return MIDEQ_E_ARG;
  int rc = validate(c, P, O);
  if (rc) return rc;
  if ((rc = validate_outputs(c, P, out))) return rc;
  if (O->ns > 0 && (rc = check_saveat_host(c, O->saveat, O->ns))) return rc;
  for (int k = 1; k < O->ntstops; k++)
    if (!(O->tstops[k] >= O->tstops[k - 1]))
      return set_err(c, MIDEQ_E_ARG, "tstops must be ascending");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int64_t N = P->ntraj;
  const int n = P->n;
  if (N == 0) return MIDEQ_OK;
  if (P->nvec) {
    for (int64_t i = 0; i < N; i++)
      if (P->nvec[i] < BRUSS_NMIN || P->nvec[i] > n || (P->nvec[i] & 1))
        return set_err(c, MIDEQ_E_ARG, "nvec[%lld]=%d outside [%d,%d] or odd", (long long)i,
                       P->nvec[i], BRUSS_NMIN, n);
  }
  hipEvent_t t0e, t1e;
  HIP_TRY(c, hipEventCreate(&t0e));
  HIP_TRY(c, hipEventCreate(&t1e));
  HIP_TRY(c, hipEventRecord(t0e, s));

  mideq_problem dP = *P;
  mideq_options dO = *O;
  mideq_outputs dOut{};
  void* d;
  if ((rc = ensure_buf(c, BUF_U0, sizeof(double) * (size_t)n * N, &d))) return rc;
  HIP_TRY(c, hipMemcpyAsync(d, P->u0, sizeof(double) * (size_t)n * N, hipMemcpyHostToDevice, s));
  dP.u0 = (const double*)d;
  if ((rc = ensure_buf(c, BUF_P, sizeof(double) * (size_t)P->np * N, &d))) return rc;
  HIP_TRY(c, hipMemcpyAsync(d, P->p, sizeof(double) * (size_t)P->np * N, hipMemcpyHostToDevice, s));
  dP.p = (const double*)d;
  if (P->t0) {
    if ((rc = ensure_buf(c, BUF_T0, sizeof(double) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, P->t0, sizeof(double) * N, hipMemcpyHostToDevice, s));
    dP.t0 = (const double*)d;
  }
  if (P->tf) {
    if ((rc = ensure_buf(c, BUF_TF, sizeof(double) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, P->tf, sizeof(double) * N, hipMemcpyHostToDevice, s));
    dP.tf = (const double*)d;
  }
  if (P->nvec) {
    if ((rc = ensure_buf(c, BUF_NVEC, sizeof(int32_t) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, P->nvec, sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
    dP.nvec = (const int32_t*)d;
  }
  if (O->ns > 0) {
    if ((rc = ensure_buf(c, BUF_SAVEAT, sizeof(double) * O->ns, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->saveat, sizeof(double) * O->ns, hipMemcpyHostToDevice, s));
    dO.saveat = (const double*)d;
  }
  if (O->ntstops > 0) {
    if ((rc = ensure_buf(c, BUF_TST, sizeof(double) * O->ntstops, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->tstops, sizeof(double) * O->ntstops,
                              hipMemcpyHostToDevice, s));
    dO.tstops = (const double*)d;
  }
  if (O->dt0) {
    if ((rc = ensure_buf(c, BUF_DT0, sizeof(double) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->dt0, sizeof(double) * N, hipMemcpyHostToDevice, s));
    dO.dt0 = (const double*)d;
  }
  if (O->alg_hint) {
    if ((rc = ensure_buf(c, BUF_HINT, N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->alg_hint, N, hipMemcpyHostToDevice, s));
    dO.alg_hint = (const int8_t*)d;
  }
  if (O->iter_cap && O->auto_passes == 3) {
    if ((rc = ensure_buf(c, BUF_CAP, sizeof(int32_t) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->iter_cap, sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
    dO.iter_cap = (const int32_t*)d;
  }
  // launch order: the caller's, else for the large-system kernel most
  // expensive first (LPT), cost ~ n * Ngrid^2 * (tf - t0), the Tsit5 work
  // when stability-bound
  std::vector<int64_t> order;
  const int64_t* d_order = nullptr;
  if (O->order) {
    if ((rc = check_order_host(c, O->order, N))) return rc;
    if ((rc = ensure_buf(c, BUF_ORDER, sizeof(int64_t) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, O->order, sizeof(int64_t) * N, hipMemcpyHostToDevice, s));
    d_order = (const int64_t*)d;
  } else if (P->model == MODEL_BRUSSELATOR && O->schedule != MIDEQ_SCHED_LANE) {
    order.resize(N);
    std::iota(order.begin(), order.end(), 0);
    std::vector<double> cost(N);
    for (int64_t i = 0; i < N; i++) {
      double ni = P->nvec ? P->nvec[i] : n;
      double span = (P->tf ? P->tf[i] : P->tspan1) - (P->t0 ? P->t0[i] : P->tspan0);
      const double ci = ni * (ni * 0.5) * (ni * 0.5) * std::fabs(span);
      cost[i] = ci == ci ? ci : 0.0;  // a NaN span is not integrated (and would break the sort)
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return cost[x] > cost[y]; });
    if ((rc = ensure_buf(c, BUF_ORDER, sizeof(int64_t) * N, &d))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d, order.data(), sizeof(int64_t) * N, hipMemcpyHostToDevice, s));
    d_order = (const int64_t*)d;
  }
  size_t us_bytes = (size_t)O->ns * n * N * (O->save_dtype ? 4 : 8);
  if (out->us && O->ns > 0) {
    if ((rc = ensure_buf(c, BUF_US, us_bytes, &dOut.us))) return rc;
    // rows past a ragged instance's size are not written by the kernel
    if (P->nvec)
      HIP_TRY(c, hipMemcpyAsync(dOut.us, out->us, us_bytes, hipMemcpyHostToDevice, s));
  }
  if (out->u_final) {
    if ((rc = ensure_buf(c, BUF_UF, sizeof(double) * (size_t)n * N, (void**)&dOut.u_final))) return rc;
    // rows past a ragged instance's size are not written by the kernel
    if (P->nvec)
      HIP_TRY(c, hipMemcpyAsync(dOut.u_final, out->u_final, sizeof(double) * (size_t)n * N,
                                hipMemcpyHostToDevice, s));
  }
  if (out->t_final && (rc = ensure_buf(c, BUF_TFIN, sizeof(double) * N, (void**)&dOut.t_final))) return rc;
  if (out->retcode && (rc = ensure_buf(c, BUF_RC, N, (void**)&dOut.retcode))) return rc;
  if (out->naccept && (rc = ensure_buf(c, BUF_NA, 4 * N, (void**)&dOut.naccept))) return rc;
  if (out->nreject && (rc = ensure_buf(c, BUF_NR, 4 * N, (void**)&dOut.nreject))) return rc;
  if (out->nf && (rc = ensure_buf(c, BUF_NF, 4 * N, (void**)&dOut.nf))) return rc;
  if (out->nstiff && (rc = ensure_buf(c, BUF_NS, 4 * N, (void**)&dOut.nstiff))) return rc;
  if (out->dt_first && (rc = ensure_buf(c, BUF_DTF, 8 * N, (void**)&dOut.dt_first))) return rc;
  if (out->nevent && (rc = ensure_buf(c, BUF_NEV, 4 * N, (void**)&dOut.nevent))) return rc;
  if (out->t_event && (rc = ensure_buf(c, BUF_TEV, 8 * N, (void**)&dOut.t_event))) return rc;
  const size_t nev = out->ev_t ? (size_t)out->ev_cap * (size_t)N : 0;
  if (out->ev_t) {
    dOut.ev_cap = out->ev_cap;
    if ((rc = ensure_buf(c, BUF_EVT, 8 * nev, (void**)&dOut.ev_t))) return rc;
    if ((rc = ensure_buf(c, BUF_EVU, 8 * nev * (size_t)n, (void**)&dOut.ev_u))) return rc;
    if (out->ev_n && (rc = ensure_buf(c, BUF_EVN, 4 * N, (void**)&dOut.ev_n))) return rc;
  }
  // output_func values: kept on the device when only their reduction is asked for
  UserModel* um_out = user_model(c, P->model);
  const int nout = um_out ? um_out->nout : 0;
  if (nout > 0 && (out->out || O->reduction != 0)) {
    if ((rc = ensure_buf(c, BUF_OUT, sizeof(double) * (size_t)nout * N, (void**)&dOut.out)))
      return rc;
  }
  if (nout > 0 && O->reduction != 0 && out->reduced) {
    if ((rc = ensure_buf(c, BUF_RED, sizeof(double) * nout, (void**)&dOut.reduced))) return rc;
  }

  Scratch Sc;
  if ((rc = scratch_for_stream(c, s, &Sc))) return rc;
  rc = run_solver(c, &dP, &dO, &dOut, d_order, s, Sc);
  if (rc) return rc;
  if (out->us && O->ns > 0 && (rc = copy_to_host(c, out->us, dOut.us, us_bytes, s))) return rc;
  if (out->u_final)
    HIP_TRY(c, hipMemcpyAsync(out->u_final, dOut.u_final, sizeof(double) * (size_t)n * N, hipMemcpyDeviceToHost, s));
  if (out->t_final)
    HIP_TRY(c, hipMemcpyAsync(out->t_final, dOut.t_final, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  if (out->retcode) HIP_TRY(c, hipMemcpyAsync(out->retcode, dOut.retcode, N, hipMemcpyDeviceToHost, s));
  if (out->naccept) HIP_TRY(c, hipMemcpyAsync(out->naccept, dOut.naccept, 4 * N, hipMemcpyDeviceToHost, s));
  if (out->nreject) HIP_TRY(c, hipMemcpyAsync(out->nreject, dOut.nreject, 4 * N, hipMemcpyDeviceToHost, s));
  if (out->nf) HIP_TRY(c, hipMemcpyAsync(out->nf, dOut.nf, 4 * N, hipMemcpyDeviceToHost, s));
  if (out->nstiff) HIP_TRY(c, hipMemcpyAsync(out->nstiff, dOut.nstiff, 4 * N, hipMemcpyDeviceToHost, s));
  if (out->dt_first) HIP_TRY(c, hipMemcpyAsync(out->dt_first, dOut.dt_first, 8 * N, hipMemcpyDeviceToHost, s));
  if (out->nevent) HIP_TRY(c, hipMemcpyAsync(out->nevent, dOut.nevent, 4 * N, hipMemcpyDeviceToHost, s));
  if (out->t_event) HIP_TRY(c, hipMemcpyAsync(out->t_event, dOut.t_event, 8 * N, hipMemcpyDeviceToHost, s));
  if (out->ev_t) {
    HIP_TRY(c, hipMemcpyAsync(out->ev_t, dOut.ev_t, 8 * nev, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(out->ev_u, dOut.ev_u, 8 * nev * (size_t)n, hipMemcpyDeviceToHost, s));
    if (out->ev_n)
      HIP_TRY(c, hipMemcpyAsync(out->ev_n, dOut.ev_n, 4 * N, hipMemcpyDeviceToHost, s));
  }
  if (out->out && dOut.out)
    HIP_TRY(c, hipMemcpyAsync(out->out, dOut.out, sizeof(double) * (size_t)nout * N,
                              hipMemcpyDeviceToHost, s));
  if (out->reduced && dOut.reduced)
    HIP_TRY(c, hipMemcpyAsync(out->reduced, dOut.reduced, sizeof(double) * nout,
                              hipMemcpyDeviceToHost, s));
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
