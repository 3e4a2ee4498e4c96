"""ctypes binding of libmideq.so (include/mideq.h).

This is the same binding the Julia shim makes with `ccall`; the structs
below mirror the C structs field for field.  The library must be built
(`make -C differentialequations.jl_amd/csrc` or `__graft_entry__.build()`);
there is no fallback: a missing library raises at import of the solver.
"""
import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
# MIDEQ_LIB selects another build of the library (the host-sanitizer build)
LIB_PATH = os.environ.get("MIDEQ_LIB") or os.path.join(PKG_ROOT, "_build", "libmideq.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "mideq.h")

MODEL = dict(exp=0, lorenz=1, robertson=2, vdp=3, brusselator=4, linear=5)
ALG = dict(tsit5=0, rodas5=1, auto=2, rodas5p=3, auto5p=4)
PRECISION = dict(fp64=0, mixed=1)
SCHEDULE = dict(auto=0, lane=1, persistent=2)
STATUS = {0: "OK", -1: "E_ARG", -2: "E_HIP", -3: "E_NOMEM", -4: "E_MODEL",
          -5: "E_UNSUPPORTED", -6: "E_NODEVICE"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_int8)


class Problem(C.Structure):
    _fields_ = [("model", C.c_int32), ("n", C.c_int32), ("np", C.c_int32),
                ("reserved0", C.c_int32), ("ntraj", C.c_int64), ("u0", C.c_void_p),
                ("p", C.c_void_p), ("t0", C.c_void_p), ("tf", C.c_void_p),
                ("tspan0", C.c_double), ("tspan1", C.c_double), ("nvec", C.c_void_p)]


class CallbackSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("direction", C.c_int32), ("terminate", C.c_int32),
                ("interp_points", C.c_int32), ("condition", C.c_char_p), ("affect", C.c_char_p),
                ("save_positions", C.c_int32), ("reserved", C.c_int32)]


class UserModelSpec(C.Structure):
    _fields_ = [("n", C.c_int32), ("np", C.c_int32), ("rhs", C.c_char_p), ("jac", C.c_char_p),
                ("ncallbacks", C.c_int32), ("save_everystep", C.c_int32),
                ("callbacks", C.POINTER(CallbackSpec)), ("nout", C.c_int32),
                ("with_tstops", C.c_int32), ("output", C.c_char_p)]


USER_MODEL_BASE = 1000
USER_NMAX, USER_NPMAX, USER_NCBMAX, USER_NOUTMAX = 16, 16, 8, 16
REDUCTION = dict(none=0, sum=1, mean=2, min=3, max=4)


class Options(C.Structure):
    _fields_ = [("alg", C.c_int32), ("precision", C.c_int32), ("abstol", C.c_double),
                ("reltol", C.c_double), ("dt", C.c_double), ("dtmin", C.c_double),
                ("dtmax", C.c_double), ("maxiters", C.c_int64), ("saveat", C.c_void_p),
                ("ns", C.c_int32), ("save_dtype", C.c_int32), ("dt0", C.c_void_p),
                ("alg_hint", C.c_void_p), ("qsteady_default", C.c_int32),
                ("schedule", C.c_int32), ("block", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("lanes", C.c_int32), ("graph", C.c_int32), ("controller", C.c_int32),
                ("auto_passes", C.c_int32), ("order", C.c_void_p),
                ("reduction", C.c_int32), ("save_layout", C.c_int32), ("iter_cap", C.c_void_p),
                ("tstops", C.c_void_p), ("ntstops", C.c_int32), ("reserved", C.c_int32)]


class Outputs(C.Structure):
    _fields_ = [("us", C.c_void_p), ("u_final", C.c_void_p), ("t_final", C.c_void_p),
                ("retcode", C.c_void_p), ("naccept", C.c_void_p), ("nreject", C.c_void_p),
                ("nf", C.c_void_p), ("nstiff", C.c_void_p), ("dt_first", C.c_void_p),
                ("nevent", C.c_void_p), ("t_event", C.c_void_p), ("out", C.c_void_p),
                ("reduced", C.c_void_p), ("ev_t", C.c_void_p), ("ev_u", C.c_void_p),
                ("ev_n", C.c_void_p), ("ev_cap", C.c_int32), ("reserved", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("schedule", C.c_int32),
                ("block", C.c_int32), ("grid", C.c_int32), ("launches", C.c_int32),
                ("lanes", C.c_int32), ("graph", C.c_int32)]


class MideqError(RuntimeError):
    pass


_LIB = None


def header_functions(path=HEADER):
    """Names of the functions include/mideq.h declares."""
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mideq_[a-z0-9_]+)\s*\(", txt)))


def load(path=LIB_PATH):
    """Load libmideq.so and declare its signatures.  Raises if missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise MideqError(f"libmideq.so not built: {path} (run __graft_entry__.build())")
    # PyTorch-ROCm ships its own HIP runtime under the same SONAME
    # (libamdhip64.so.7).  Loaded first, it is the one libmideq binds to, so
    # the process has a single HIP runtime and torch tensors and library
    # kernels share one device context.  Loaded after libmideq, torch would
    # pull in a second runtime that cannot open the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    L.mideq_version.restype = C.c_int
    L.mideq_ctx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32]
    L.mideq_ctx_destroy.argtypes = [C.c_void_p]
    L.mideq_last_error.argtypes = [C.c_void_p]
    L.mideq_last_error.restype = C.c_char_p
    L.mideq_last_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.mideq_ensemble_solve.argtypes = [C.c_void_p, C.POINTER(Problem), C.POINTER(Options),
                                       C.POINTER(Outputs)]
    L.mideq_model_register.argtypes = [C.c_void_p, C.POINTER(UserModelSpec),
                                       C.POINTER(C.c_int32)]
    L.mideq_model_check.argtypes = [C.POINTER(UserModelSpec), C.c_char_p, C.c_size_t]
    L.mideq_ensemble_solve_device.argtypes = [C.c_void_p, C.POINTER(Problem),
                                              C.POINTER(Options), C.POINTER(Outputs),
                                              C.c_void_p]
    if hasattr(L, "mideq_ensemble_solve_device_noalloc"):
        L.mideq_ensemble_solve_device_noalloc.argtypes = [C.c_void_p, C.POINTER(Problem),
                                                          C.POINTER(Options),
                                                          C.POINTER(Outputs), C.c_void_p]
    _LIB = L
    return L


class Context:
    """One mideq_ctx: a HIP stream and cached device buffers on one GPU."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.mideq_ctx_create(C.byref(h), device)
        if rc != 0:
            raise MideqError(f"mideq_ctx_create failed: {STATUS.get(rc, rc)}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.mideq_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, what):
        if rc != 0:
            msg = self.lib.mideq_last_error(self.h).decode()
            raise MideqError(f"{what}: {STATUS.get(rc, rc)}: {msg}")

    def solve(self, prob, opt, out):
        self.check(self.lib.mideq_ensemble_solve(self.h, C.byref(prob), C.byref(opt),
                                                 C.byref(out)), "mideq_ensemble_solve")

    def solve_device(self, prob, opt, out, stream=None):
        self.check(self.lib.mideq_ensemble_solve_device(self.h, C.byref(prob), C.byref(opt),
                                                        C.byref(out), stream),
                   "mideq_ensemble_solve_device")

    def solve_device_noalloc(self, prob, opt, out, stream=None):
        self.check(self.lib.mideq_ensemble_solve_device_noalloc(
            self.h, C.byref(prob), C.byref(opt), C.byref(out), stream),
            "mideq_ensemble_solve_device_noalloc")

    def register_model(self, spec):
        mid = C.c_int32()
        self.check(self.lib.mideq_model_register(self.h, C.byref(spec), C.byref(mid)),
                   "mideq_model_register")
        return mid.value

    def stats(self):
        s = Stats()
        self.check(self.lib.mideq_last_stats(self.h, C.byref(s)), "mideq_last_stats")
        return s


_CTX = {}


def context(device=0):
    if device not in _CTX:
        _CTX[device] = Context(device)
    return _CTX[device]
