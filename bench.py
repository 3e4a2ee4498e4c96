#!/usr/bin/env python3
"""Benchmark: trajectories/s of the seeded Lorenz ensemble (1,048,576
trajectories, Tsit5, saveat 100 uniform points on [0, 10], SoA output
100 x 3 x N) on one MI355X, with the final-state error against an FP64
solve at reltol = abstol = 1e-12.

One step = one ensemble solve with inputs already resident in HBM.  Prints
one JSON line (metric, value, roofline, cpu_baseline, per-form details).

    python bench.py [--steps K] [--warmup W] [--ntraj N] [--forms plain,mixed,tuned]
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "differentialequations.jl_amd"))

import numpy as np  # noqa: E402

# Algorithmic work per unit (counted from the method definitions; FMA = 2,
# add/mul/div/sqrt/pow = 1): Tsit5 attempted step = 6 RHS + n*(48 stage
# sums + 14 embedded error + 6 norm) + 10 controller; Lorenz RHS = 8 flops.
LORENZ_F = 8
N_STATE = 3
FLOPS_PER_STEP = 6 * LORENZ_F + N_STATE * (48 + 14 + 6) + 10        # 262
FLOPS_PER_SAVE = 56 + 14 * N_STATE                                 # 98
PEAK_FP64 = 78.6   # TFLOP/s, MI355X FP64 vector
PEAK_HBM = 8000.0  # GB/s
# HBM bytes per launch of the solver kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a 1 GiB copy kernel (factors 2.0 and 1.0);
# profiles/tsit5_lorenz_1M_saveat100_hbm_counters.md.  Re-measured when
# the kernel changes; None where no counter pass exists for that form.
MEASURED_TRAFFIC = {"plain": 11549978186}
ALGORITHMIC_BYTES_PER_TRAJ = 48 + 100 * 3 * 8 + 24 + 8 + 1 + 12

FORMS = {
    # name: (precision, abstol=reltol, schedule, save dtype)
    "plain": ("fp64", 1e-8, "lane", "float64"),
    "mixed": ("mixed", 1e-6, "lane", "float32"),
    "tuned": ("fp64", 1e-8, "persistent", "float64"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ntraj", type=int, default=1 << 20)
    ap.add_argument("--forms", default="plain,mixed,tuned")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=20000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration(torch):
    """A plain copy of exactly 1 GiB (read 1 GiB, write 1 GiB): the
    reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters."""
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 3}))


class DeviceEnsemble:
    """The seeded Lorenz ensemble resident in HBM plus output buffers."""

    def __init__(self, torch, N, ns=100):
        from mideq import workloads
        self.torch = torch
        w = workloads.lorenz(N, seed=1234)
        dev = torch.device("cuda:0")
        self.N = N
        self.ns = ns
        self.u0 = torch.from_numpy(w["u0"]).to(dev).contiguous()
        self.p = torch.from_numpy(w["p"]).to(dev).contiguous()
        self.saveat = torch.from_numpy(workloads.saveat_uniform(0.0, 10.0, ns)).to(dev)
        self.u_final = torch.empty((3, N), dtype=torch.float64, device=dev)
        self.t_final = torch.empty(N, dtype=torch.float64, device=dev)
        self.retcode = torch.empty(N, dtype=torch.int8, device=dev)
        self.naccept = torch.empty(N, dtype=torch.int32, device=dev)
        self.nreject = torch.empty(N, dtype=torch.int32, device=dev)
        self.nf = torch.empty(N, dtype=torch.int32, device=dev)
        self.us = {}
        self.stream = torch.cuda.Stream()

    def us_buf(self, dtype):
        if dtype not in self.us:
            tdt = self.torch.float64 if dtype == "float64" else self.torch.float32
            self.us[dtype] = self.torch.empty((self.ns, 3, self.N), dtype=tdt, device="cuda:0")
        return self.us[dtype]

    def structs(self, precision, tol, schedule, save_dtype, with_save=True):
        from mideq import _abi
        P = _abi.Problem()
        P.model = _abi.MODEL["lorenz"]
        P.n = 3
        P.np = 3
        P.ntraj = self.N
        P.u0 = self.u0.data_ptr()
        P.p = self.p.data_ptr()
        P.tspan0 = 0.0
        P.tspan1 = 10.0
        O = _abi.Options()
        O.alg = _abi.ALG["tsit5"]
        O.precision = _abi.PRECISION[precision]
        O.abstol = tol
        O.reltol = tol
        O.qsteady_default = 1
        O.schedule = _abi.SCHEDULE[schedule]
        if with_save:
            O.saveat = self.saveat.data_ptr()
            O.ns = self.ns
            O.save_dtype = 0 if save_dtype == "float64" else 1
        out = _abi.Outputs()
        out.us = self.us_buf(save_dtype).data_ptr() if with_save else None
        out.u_final = self.u_final.data_ptr()
        out.t_final = self.t_final.data_ptr()
        out.retcode = self.retcode.data_ptr()
        out.naccept = self.naccept.data_ptr()
        out.nreject = self.nreject.data_ptr()
        out.nf = self.nf.data_ptr()
        return P, O, out


def run_form(torch, ctx, ens, form, steps, warmup):
    precision, tol, schedule, save_dtype = FORMS[form]
    P, O, out = ens.structs(precision, tol, schedule, save_dtype)
    # a dedicated stream: the kernels and the timing events share it
    stream = ens.stream
    sp = C.c_void_p(stream.cuda_stream)
    for _ in range(warmup):
        ctx.solve_device(P, O, out, sp)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        ctx.solve_device(P, O, out, sp)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [a.elapsed_time(b) for a, b in ev]
    na = ens.naccept.cpu().numpy()
    nr = ens.nreject.cpu().numpy()
    rc = ens.retcode.cpu().numpy()
    uf = ens.u_final.cpu().numpy().copy()
    attempted = int(na.sum()) + int(nr.sum())
    return dict(form=form, precision=precision, tol=tol, schedule=schedule,
                save_dtype=save_dtype, wall_s=wall, kernel_ms=kms,
                kernel_ms_mean=statistics.fmean(kms), attempted_steps=attempted,
                naccept_mean=float(na.mean()), nreject_mean=float(nr.mean()),
                all_success=bool((rc == 1).all()), u_final=uf)


def reference_solution(torch, ctx, ens):
    """FP64 Tsit5 at reltol = abstol = 1e-12 on the device (the kernel whose
    FP64 results match the CPU oracle bit for bit)."""
    P, O, out = ens.structs("fp64", 1e-12, "lane", "float64", with_save=False)
    ctx.solve_device(P, O, out, C.c_void_p(ens.stream.cuda_stream))
    torch.cuda.synchronize()
    rc = ens.retcode.cpu().numpy()
    assert (rc == 1).all(), "reference solve failed"
    return ens.u_final.cpu().numpy().copy()


def errors(u, ref):
    e = np.linalg.norm(u - ref, axis=0) / np.linalg.norm(ref, axis=0)
    return dict(median=float(np.median(e)), p99=float(np.quantile(e, 0.99)),
                max=float(e.max()), frac_le_1e6=float((e <= 1e-6).mean()))


def cpu_baseline(sample, ns=100):
    """The oracle's EnsembleThreads restatement (C, pthreads) on this host's
    cores, on the first `sample` trajectories of the same workload."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib
    from mideq import workloads
    O = oracle_lib.load()
    w = workloads.lorenz(sample, seed=1234)
    sa = workloads.saveat_uniform(0.0, 10.0, ns)
    threads = min(os.cpu_count() or 1, 64)
    t0 = time.perf_counter()
    O.ensemble("lorenz", w["u0"], w["p"], 0.0, 10.0, nthreads=threads, saveat=sa,
               abstol=1e-8, reltol=1e-8)
    dt = time.perf_counter() - t0
    return dict(value=sample / dt, unit="trajectories/s", cores=threads, kind="port",
                sample=f"first {sample} of the 1,048,576 seeded Lorenz trajectories, Tsit5 FP64 "
                       f"abstol=reltol=1e-8, saveat 100 points, oracle EnsembleThreads "
                       f"restatement on {threads} host threads, {dt:.2f} s")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from mideq import _abi
    if not torch.cuda.is_available():
        raise SystemExit("no GPU visible")
    if a.copy_calibration:
        copy_calibration(torch)
        return
    ctx = _abi.context(0)
    ens = DeviceEnsemble(torch, a.ntraj)
    forms = [f for f in a.forms.split(",") if f]
    results = {}
    for f in forms:
        print(f"[bench] form {f}: {FORMS[f]}", file=sys.stderr, flush=True)
        results[f] = run_form(torch, ctx, ens, f, a.steps, a.warmup)
    print("[bench] reference solve 1e-12", file=sys.stderr, flush=True)
    ref = reference_solution(torch, ctx, ens)
    detail = {}
    for f, r in results.items():
        err = errors(r.pop("u_final"), ref)
        r["error"] = err
        r["traj_per_s"] = a.ntraj * a.steps / r["wall_s"]
        flops = r["attempted_steps"] * FLOPS_PER_STEP + a.ntraj * 100 * FLOPS_PER_SAVE
        r["tflops"] = flops / (r["kernel_ms_mean"] * 1e-3) / 1e12
        r["gate_ok"] = err["median"] <= 1e-6 and r["all_success"]
        detail[f] = r
    ok = [f for f in forms if detail[f]["gate_ok"]]
    head = max(ok, key=lambda f: detail[f]["traj_per_s"]) if ok else forms[0]
    h = detail[head]
    ms = h["wall_s"] / a.steps * 1e3
    out_bytes = a.ntraj * (100 * 3 * (8 if h["save_dtype"] == "float64" else 4) + 3 * 8 + 8 + 13)
    roof = dict(bound="valu", achieved=h["tflops"], peak=PEAK_FP64, unit="TFLOP/s",
                bound_note="compute-bound on the FP64 vector ALUs; peak = MI355X FP64 vector "
                           "78.6 TFLOP/s (no MFMA on this path)",
                frac=h["tflops"] / PEAK_FP64, traffic=MEASURED_TRAFFIC.get(head),
                algorithmic_bytes=a.ntraj * ALGORITHMIC_BYTES_PER_TRAJ,
                kernel=f"k_lane/k_persistent<ModelLorenz,{h['precision']},Tsit5>",
                flops_per_attempted_step=FLOPS_PER_STEP, flops_per_saved_point=FLOPS_PER_SAVE,
                attempted_steps=h["attempted_steps"], kernel_ms=h["kernel_ms_mean"],
                output_gbps=out_bytes / (h["kernel_ms_mean"] * 1e-3) / 1e9)
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_sample)
    line = {
        "metric": "trajectories/s at <=1e-6 final-state err vs FP64 oracle, plain vs tuned form",
        "value": h["traj_per_s"], "unit": "trajectories/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if h["precision"] == "fp64" else "f32+f64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": "Lorenz ensemble, 1,048,576 trajectories, sigma/rho/beta +-10% "
                               "seeded, u0=[1,0,0], tspan (0,10), Tsit5, saveat 100 uniform "
                               "points, SoA output 100x3xN",
                   "ntraj": a.ntraj, "headline_form": head,
                   "gate": "median final-state rel err vs FP64 tol-1e-12 solve <= 1e-6"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": {f: {k: v for k, v in d.items() if k != "kernel_ms"} for f, d in detail.items()},
    }
    if cpu:
        line["gpu_vs_cpu"] = h["traj_per_s"] / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
