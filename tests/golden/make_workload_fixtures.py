"""Writes tests/golden/workloads.json: the first 8 rows of every benchmark
input table and a SHA-256 of each full table, so that the seeded generator
(mideq.workloads) is pinned at every size the benchmark uses.

Tables (seed 1234 unless stated; arrays are the SoA layout the solver reads,
row = trajectory):
  lorenz_10000       u0, p            (CPU-runnable Lorenz case, tspan (0, 10))
  lorenz_1048576     u0, p            (the 1,048,576-trajectory saveat workload)
  robertson_262144   u0, p
  brusselator_131072 u0 (rows 0..nvec-1 filled), p, tf, nvec
  mix_1000000        lorenz / robertson / vdp thirds of 333,333 each: u0, p, tf
  mix_train_999999   the training mix (seed 777), same layout

Run from the repository root:  python tests/golden/make_workload_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "differentialequations.jl_amd"))

from mideq import workloads  # noqa: E402

HEAD = 8


def digest(a):
    """SHA-256 of the array's bytes in C order, little-endian, with its dtype
    and shape, so that a layout change is a different digest."""
    a = np.ascontiguousarray(a)
    h = hashlib.sha256()
    h.update(f"{a.dtype.str}{a.shape}".encode())
    h.update(a.tobytes())
    return h.hexdigest()


def table(w, keys):
    out = {}
    for k in keys:
        a = np.asarray(w[k])
        if a.ndim == 0:
            out[k] = {"scalar": float(a)}
            continue
        rows = a[..., :HEAD] if a.ndim == 2 else a[:HEAD]
        out[k] = {"shape": list(a.shape), "dtype": a.dtype.str,
                  "head": rows.tolist(), "sha256": digest(a)}
    return out


def tables():
    t = {}
    t["lorenz_10000"] = table(workloads.lorenz(10_000, seed=1234), ["u0", "p"])
    t["lorenz_1048576"] = table(workloads.lorenz(1_048_576, seed=1234), ["u0", "p"])
    t["robertson_262144"] = table(workloads.robertson(262_144, seed=1234), ["u0", "p"])
    t["brusselator_131072"] = table(workloads.brusselator(131_072, seed=1234),
                                    ["u0", "p", "tf", "nvec"])
    for name, seed, n_each in (("mix_1000000", 1234, 333_333), ("mix_train_999999", 777, 333_333)):
        sub = {}
        for m in ("lorenz", "robertson", "vdp"):
            w = workloads.GENERATORS[m](n_each, seed=seed)
            w["tf"] = np.broadcast_to(np.asarray(w["tf"], dtype=np.float64), (n_each,)).copy()
            sub[m] = table(w, ["u0", "p", "tf"])
        t[name] = sub
    return t


def main():
    path = os.path.join(HERE, "workloads.json")
    with open(path, "w") as f:
        json.dump(tables(), f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
