"""Learned initial step, method class and cost for ensemble solves.

A 16 -> 3x64 (GELU) -> 4 MLP, run by the BF16 matrix-core kernels of
libmideq.so (`mideq_mlp_*`), predicts for every trajectory
  * log10 of a step size the controller accepts first time (the plain
    solver's first accepted step), used as `dt0`;
  * a stiffness logit, used as the initial method of AutoTsit5(Rodas5())
    (`alg_hint`);
  * log2 of the solve's cost (attempted steps, a Rodas5 step counted twice),
    which with the class sets the launch `order`: trajectories of one class
    and similar cost share a wavefront, so lanes neither wait for a much
    longer neighbour nor mix Tsit5 and Rodas5 steps;
  * a never-stiff logit: the solve takes no Rosenbrock step at all, which
    is when plain Tsit5 (no stiffness test) gives AutoTsit5's result and
    can be trusted (`auto_passes = 3`, `iter_cap`).
The plain initdt heuristic and AutoTsit5 switching stay in place: the
prediction only sets where each trajectory starts and when it is scheduled,
and the solver's own controller and stiffness test still decide every step.

Training data are the plain solver's own logs (first accepted step, accepted
Rodas5 steps and attempted steps per trajectory).  Training runs on the GPU with the
library's forward/backward kernels and Adam; tensors live in torch device
memory, passed to the C ABI as raw pointers.
"""
import ctypes as C

import numpy as np

from . import _abi

NIN, NOUT = 16, 4
NPARAM = 9668
RODAS_STEP_COST = 2.0  # a Rodas5 step in Tsit5 steps, for the cost target
MODEL_SLOT = {"lorenz": 0, "robertson": 1, "vdp": 2}


def _lib():
    L = _abi.load()
    if not getattr(L, "_mlp_declared", False):
        vp = C.c_void_p
        L.mideq_mlp_create.argtypes = [C.POINTER(vp), C.c_int32]
        L.mideq_mlp_destroy.argtypes = [vp]
        L.mideq_mlp_last_error.argtypes = [vp]
        L.mideq_mlp_last_error.restype = C.c_char_p
        L.mideq_mlp_set_params.argtypes = [vp, vp]
        L.mideq_mlp_get_params.argtypes = [vp, vp]
        L.mideq_mlp_get_grads.argtypes = [vp, vp]
        L.mideq_mlp_set_norm.argtypes = [vp, vp, vp]
        L.mideq_mlp_infer.argtypes = [vp, vp, C.c_int64, vp, vp]
        L.mideq_mlp_grad.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int64, C.c_float, C.c_float,
                                     C.c_float, C.c_float, vp, vp]
        L.mideq_mlp_adam_step.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float, vp]
        L.mideq_mlp_launch_order.argtypes = [vp, vp, C.c_int64, vp, vp]
        L._mlp_declared = True
    return L


def features(model, p, u0, t0, tf, abstol, reltol):
    """[N, 16] float32 features: model one-hot (3), log10|p| (3), u0 (3),
    log10(tf - t0), log10 abstol, log10 reltol, zero padding (4)."""
    p = np.atleast_2d(np.asarray(p, dtype=np.float64))
    u0 = np.atleast_2d(np.asarray(u0, dtype=np.float64))
    N = u0.shape[1]
    X = np.zeros((N, NIN), dtype=np.float64)
    X[:, MODEL_SLOT[model]] = 1.0
    for j in range(min(3, p.shape[0])):
        X[:, 3 + j] = np.log10(np.abs(p[j]) + 1e-300)
    for j in range(min(3, u0.shape[0])):
        X[:, 6 + j] = u0[j]
    span = np.broadcast_to(np.asarray(tf, dtype=np.float64) - np.asarray(t0, dtype=np.float64), (N,))
    X[:, 9] = np.log10(np.abs(span) + 1e-300)
    X[:, 10] = np.log10(np.broadcast_to(abstol, (N,)))
    X[:, 11] = np.log10(np.broadcast_to(reltol, (N,)))
    return X.astype(np.float32)


class MLP:
    """Device-resident parameters, Adam state and normalisation of the net."""

    def __init__(self, device=0, params=None, seed=0):
        self.L = _lib()
        h = C.c_void_p()
        rc = self.L.mideq_mlp_create(C.byref(h), device)
        if rc != 0:
            raise _abi.MideqError(f"mideq_mlp_create failed: {_abi.STATUS.get(rc, rc)}")
        self.h = h
        self.device = device
        if params is None:
            params = init_params(seed)
        self.set_params(params)

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.mideq_mlp_last_error(self.h).decode()
            raise _abi.MideqError(f"{what}: {_abi.STATUS.get(rc, rc)}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mideq_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        assert a.size == NPARAM
        self._check(self.L.mideq_mlp_set_params(self.h, a.ctypes.data), "set_params")

    def get_params(self):
        a = np.empty(NPARAM, dtype=np.float32)
        self._check(self.L.mideq_mlp_get_params(self.h, a.ctypes.data), "get_params")
        return a

    def get_grads(self):
        a = np.empty(NPARAM, dtype=np.float32)
        self._check(self.L.mideq_mlp_get_grads(self.h, a.ctypes.data), "get_grads")
        return a

    def set_norm(self, mean, std):
        m = np.ascontiguousarray(mean, dtype=np.float32)
        s = np.ascontiguousarray(std, dtype=np.float32)
        self.mean, self.std = m.copy(), s.copy()
        self._check(self.L.mideq_mlp_set_norm(self.h, m.ctypes.data, s.ctypes.data), "set_norm")

    # device-pointer entry points (torch tensors)
    def infer(self, X, Y=None, stream=None):
        import torch
        B = X.shape[0]
        if Y is None:
            Y = torch.empty((B, NOUT), dtype=torch.float32, device=X.device)
        s = None if stream is None else stream.cuda_stream
        self._check(self.L.mideq_mlp_infer(self.h, X.data_ptr(), B, Y.data_ptr(), s), "infer")
        return Y

    def launch_order(self, Y, order=None, stream=None):
        """Launch order (device int64 [B]) from infer()'s output Y [B, 4]: the
        library's counting sort by (predicted class, decreasing predicted
        cost in 1/64-octave buckets); see order_keys()."""
        import torch
        B = Y.shape[0]
        if order is None:
            order = torch.empty(B, dtype=torch.int64, device=Y.device)
        s = None if stream is None else stream.cuda_stream
        self._check(self.L.mideq_mlp_launch_order(self.h, Y.data_ptr(), B, order.data_ptr(), s),
                    "launch_order")
        return order

    def grad(self, X, target, label, cost=None, w_dt=1.0, w_cls=1.0, w_cost=1.0,
             want_loss=True, stream=None, never=None, w_never=1.0):
        loss = C.c_float(0.0)
        s = None if stream is None else stream.cuda_stream
        cp = None if cost is None else cost.data_ptr()
        npt = None if never is None else never.data_ptr()
        self._check(self.L.mideq_mlp_grad(self.h, X.data_ptr(), target.data_ptr(), label.data_ptr(),
                                          cp, npt, X.shape[0], w_dt, w_cls, w_cost, w_never,
                                          C.byref(loss) if want_loss else None, s), "grad")
        return loss.value if want_loss else None

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        s = None if stream is None else stream.cuda_stream
        self._check(self.L.mideq_mlp_adam_step(self.h, lr, beta1, beta2, eps, s), "adam")

    def fit(self, X, target, label, cost=None, epochs=4, batch=8192, lr=2e-3, seed=0,
            w_cls=1.0, w_cost=0.25, init_bias=True, log=None, never=None, w_never=1.0):
        """Minibatch Adam over device tensors X [N,16], target [N], label [N]
        and (optional) cost [N] and never-stiff label [N].  init_bias: start
        the output biases at the target means and the labels' log-odds, so
        that no head's offset dominates the loss at the start (pass False to
        continue training)."""
        import torch
        N = X.shape[0]
        if init_bias:
            theta = self.get_params()
            pl = float(label.float().mean().clamp(1e-3, 1 - 1e-3))
            theta[NPARAM - 4] = float(target.float().mean())
            theta[NPARAM - 3] = float(np.log(pl / (1.0 - pl)))
            if cost is not None:
                theta[NPARAM - 2] = float(cost.float().mean())
            if never is not None:
                pn = float(never.float().mean().clamp(1e-3, 1 - 1e-3))
                theta[NPARAM - 1] = float(np.log(pn / (1.0 - pn)))
            self.set_params(theta)
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(N, generator=g).to(X.device)
            tot, nb = 0.0, 0
            for k in range(0, N - batch + 1, batch):
                idx = perm[k:k + batch]
                xb = X.index_select(0, idx).contiguous()
                tb = target.index_select(0, idx).contiguous()
                lb = label.index_select(0, idx).contiguous()
                cb = None if cost is None else cost.index_select(0, idx).contiguous()
                nvb = None if never is None else never.index_select(0, idx).contiguous()
                want = (nb % 16 == 0)
                lv = self.grad(xb, tb, lb, cb, 1.0, w_cls, w_cost, want_loss=want, never=nvb,
                               w_never=w_never)
                self.adam_step(lr * (0.5 ** ep))
                if want:
                    tot += lv
                    losses.append(lv)
                nb += 1
            if log:
                log(f"epoch {ep}: loss {losses[-1] if losses else float('nan'):.4f}")
        torch.cuda.synchronize()
        return losses


def init_params(seed=0):
    """Xavier-uniform weights, zero biases (same layout as the C ABI)."""
    rng = np.random.default_rng(seed)
    parts = []
    for (fan_in, fan_out) in ((NIN, 64), (64, 64), (64, 64), (64, NOUT)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-lim, lim, fan_in * fan_out))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def training_targets(dt_first, naccept, nstiff, nreject=None):
    """(log10 first accepted dt, stiff label, log2 cost) from plain-solver
    logs; cost = attempted steps with an accepted Rodas5 step counted
    RODAS_STEP_COST times."""
    tgt = np.log10(np.maximum(dt_first, 1e-300)).astype(np.float32)
    lab = (nstiff > 0.5 * np.maximum(naccept, 1)).astype(np.float32)
    rej = 0 if nreject is None else nreject
    cost = naccept + rej + (RODAS_STEP_COST - 1.0) * nstiff
    return tgt, lab, np.log2(np.maximum(cost, 1.0)).astype(np.float32)


def never_stiff_label(nstiff):
    """1 where the plain AutoTsit5 solve took no Rosenbrock step: plain
    Tsit5 then gives the same result."""
    return (np.asarray(nstiff) == 0).astype(np.float32)


ORDER_BUCKETS = 2048  # log2-cost buckets per class, 1/64 octave each over [0, 32)


def order_keys(stiff_logit, log2_cost):
    """Sort keys of the launch order (numpy): predicted non-stiff class
    first, each class by decreasing predicted log2 cost in 1/64-octave
    buckets -- the keys mideq_mlp_launch_order sorts by."""
    c = np.asarray(log2_cost, dtype=np.float32)
    b = np.where(c >= 0, np.minimum(np.floor(np.where(c >= 0, c, 0) * np.float32(64)),
                                    ORDER_BUCKETS - 1), 0).astype(np.int64)
    cls = (np.asarray(stiff_logit) > 0).astype(np.int64)
    return cls * ORDER_BUCKETS + (ORDER_BUCKETS - 1 - b)


def launch_order(stiff_logit, log2_cost):
    """Launch order from predictions on the host (numpy): a stable sort by
    order_keys (longest first inside each class, so the last wavefronts to
    start are the short ones).  MLP.launch_order does the same on the GPU."""
    return np.argsort(order_keys(stiff_logit, log2_cost), kind="stable")
