"""Host build of a run-time compiled user model, for the oracle.

Test infrastructure.  The device compiles the user's statements into
`struct ModelUser` (f templated over double, float and the dual numbers of
differentialequations.jl_amd/csrc/user_dual.hpp; jac given or
differentiated); this module compiles the same text with g++ (no
contraction, IEEE float and double on x86-64 SSE) into a shared object and
hands its functions to the oracle (oracle_set_user_model)."""
import ctypes as C
import hashlib
import os
import re
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUAL = os.path.join(ROOT, "differentialequations.jl_amd", "csrc", "user_dual.hpp")
_CACHE = {}


def uses_time(text):
    """The device's rule: the statements name the identifier t, outside
    comments and character/string literals (and not as a member after '.')."""
    if not text:
        return False
    code = re.sub(r'//[^\n]*|/\*.*?\*/|"(?:\\.|[^"\\])*"|\'(?:\\.|[^\'\\])*\'', " ", text,
                  flags=re.S)
    return re.search(r"(?<![A-Za-z0-9_.])t(?![A-Za-z0-9_])", code) is not None


def model_source(n, np_, rhs, jac=None, callbacks=()):
    np_ = max(int(np_), 1)
    tdep = uses_time(rhs) or uses_time(jac)
    ta, tj = (", T t", ", double t") if tdep else ("", "")
    s = ["#include <math.h>", "#define MIDEQ_HD inline", f'#include "{DUAL}"',
         "namespace mideq {", "struct ModelUser {",
         f"  static constexpr int N = {n}, NP = {np_};",
         "  template <class T>", f"  static void f(T* du, const T* u, const T* p{ta}) {{",
         "  {", rhs, "  }", "  }",
         f"  static void jac(double (*J)[N], const double* u, const double* p{tj}) {{"]
    if jac is not None:
        s += ["  {", jac, "  }"]
    else:
        s += ["    Dual<N> ud[N], pd[NP], fd[N];",
              "    for (int i = 0; i < N; i++) { ud[i].v = u[i]; ud[i].d[i] = 1.0; }",
              "    for (int j = 0; j < NP; j++) pd[j].v = p[j];",
              "    f(fd, ud, pd, Dual<N>(t));" if tdep else "    f(fd, ud, pd);",
              "    for (int i = 0; i < N; i++)",
              "      for (int j = 0; j < N; j++) J[i][j] = fd[i].d[j];"]
    s.append("  }")
    if tdep:
        s += ["  static void tgrad(double* dT, const double* u, const double* p, double t) {",
              "    Dual<1> ud[N], pd[NP], fd[N], td(t);",
              "    td.d[0] = 1.0;",
              "    for (int i = 0; i < N; i++) ud[i].v = u[i];",
              "    for (int j = 0; j < NP; j++) pd[j].v = p[j];",
              "    f(fd, ud, pd, td);",
              "    for (int i = 0; i < N; i++) dT[i] = fd[i].d[0];",
              "  }"]
    targ = ", t" if tdep else ""
    s += ["};", "}  // namespace mideq",
          'extern "C" void user_f64(double* du, const double* u, const double* p, double t) '
          f"{{ (void)t; mideq::ModelUser::f<double>(du, u, p{targ}); }}",
          'extern "C" void user_f32(float* du, const float* u, const float* p, float t) '
          f"{{ (void)t; mideq::ModelUser::f<float>(du, u, p{targ}); }}",
          'extern "C" void user_jac(double* J, const double* u, const double* p, double t) '
          f"{{ (void)t; mideq::ModelUser::jac((double (*)[mideq::ModelUser::N])J, u, p{targ}); }}",
          'extern "C" void user_tgrad(double* dT, const double* u, const double* p, double t) '
          + ("{ mideq::ModelUser::tgrad(dT, u, p, t); }" if tdep else
             "{ (void)u; (void)p; (void)t; for (int i = 0; i < mideq::ModelUser::N; i++) dT[i] = 0.0; }"),
          f'extern "C" int user_tdep(void) {{ return {1 if tdep else 0}; }}']
    # callbacks, as the device compiles them (double)
    s.append('extern "C" double user_cb_g(int k, const double* u, const double* p, double t) {')
    s.append("  switch (k) {")
    for k, cb in enumerate(callbacks):
        s.append(f"    case {k}: return (double)({cb.condition});")
    s += ["  }", "  return 0.0;", "}"]
    s.append('extern "C" void user_cb_affect(int k, double* u, const double* p, double t) {')
    s.append("  switch (k) {")
    for k, cb in enumerate(callbacks):
        s.append(f"    case {k}: {{ {cb.affect or ''} }} break;")
    s += ["  }", "}"]
    return "\n".join(s) + "\n"


def build(n, np_, rhs, jac=None, callbacks=()):
    """ctypes handle of the host-compiled model (cached by source)."""
    src = model_source(n, np_, rhs, jac, callbacks)
    key = hashlib.sha1(src.encode()).hexdigest()[:16]
    if key in _CACHE:
        return _CACHE[key]
    d = tempfile.mkdtemp(prefix="mideq_user_")
    cpp = os.path.join(d, "user.cpp")
    so = os.path.join(d, "user.so")
    with open(cpp, "w") as f:
        f.write(src)
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-fno-fast-math", cpp, "-o", so], check=True, capture_output=True)
    lib = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    lib.user_f64.argtypes = [dp, dp, dp, C.c_double]
    lib.user_f32.argtypes = [fp, fp, fp, C.c_float]
    lib.user_jac.argtypes = [dp, dp, dp, C.c_double]
    lib.user_tgrad.argtypes = [dp, dp, dp, C.c_double]
    lib.user_cb_g.argtypes = [C.c_int32, dp, dp, C.c_double]
    lib.user_cb_g.restype = C.c_double
    _CACHE[key] = lib
    return lib


def install(oracle, n, np_, rhs, jac=None, callbacks=(), save_everystep=False):
    """Make the oracle's "user" model this one (callbacks: the mideq
    ContinuousCallback / DiscreteCallback objects)."""
    lib = build(n, np_, rhs, jac, callbacks)
    addr = lambda fn: C.cast(fn, C.c_void_p).value
    L = oracle.lib
    L.oracle_set_user_model.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int32]
    rc = L.oracle_set_user_model(n, max(int(np_), 1), addr(lib.user_f64), addr(lib.user_f32),
                                 addr(lib.user_jac), addr(lib.user_tgrad), lib.user_tdep())
    assert rc == 0
    k = len(callbacks)
    arr = lambda vals: (C.c_int32 * max(k, 1))(*vals)
    kind = [0 if type(cb).__name__ == "ContinuousCallback" else 1 for cb in callbacks]
    direction = [getattr(cb, "direction", 0) for cb in callbacks]
    term = [int(bool(cb.terminate)) for cb in callbacks]
    ip = [getattr(cb, "interp_points", 0) for cb in callbacks]
    L.oracle_set_user_callbacks.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p]
    rc = L.oracle_set_user_callbacks(k, arr(kind), arr(direction), arr(term), arr(ip),
                                     addr(lib.user_cb_g), addr(lib.user_cb_affect))
    assert rc == 0
    saves = [(1 if cb.save_positions[0] else 0) | (2 if cb.save_positions[1] else 0)
             for cb in callbacks]
    L.oracle_set_user_callback_saves.argtypes = [C.c_int32, C.c_void_p]
    rc = L.oracle_set_user_callback_saves(k, arr(saves))
    assert rc == 0
    L.oracle_set_user_callback_affects.argtypes = [C.c_int32, C.c_void_p]
    assert L.oracle_set_user_callback_affects(
        k, arr([int(getattr(cb, "affect", None) is not None) for cb in callbacks])) == 0
    L.oracle_set_user_everystep.argtypes = [C.c_int32]
    assert L.oracle_set_user_everystep(int(bool(save_everystep))) == 0
    return lib
