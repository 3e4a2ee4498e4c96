"""Seeded ensemble inputs for the benchmark configurations.

Random numbers come from a counter-based generator: the k-th draw of
trajectory i under seed s is the SplitMix64 finalizer of
    s * 0xD1B54A32D192ED03 + (16 * i + k + 1) * 0x9E3779B97F4A7C15   (mod 2^64)
mapped to [0, 1) with 53 bits.  Every draw is independent of N and of the
order of evaluation, so any subset of an ensemble can be regenerated.

Problem definitions (the reference's problems, test/core.jl:22-27 Lorenz,
test/core.jl:39-45 Robertson; src/DifferentialEquations.jl:14-20):
  lorenz      sigma, rho, beta = (10, 28, 8/3) * (1 + U(-0.1, 0.1)); u0 = [1,0,0];
              tspan (0, 10)
  robertson   k = (0.04, 3e7, 1e4) * 10^U(-0.5, 0.5); u0 = [1,0,0]; tspan (0, 1e5)
  vdp         classic x' = y, y' = mu (1 - x^2) y - x; mu = 10^U(0, 4); u0 = [2, 0];
              tspan (0, 2 mu)
  brusselator Ngrid ~ U{8..64}, A = 1, B ~ U[2, 4], alpha = 0.02, periodic,
              u_j(0) = 1 + sin(2 pi j / Ngrid), v_j(0) = 3, tspan (0, U[1, 20])
"""
import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)
_S = np.uint64(0xD1B54A32D192ED03)


def uniform(seed, idx, k):
    """U[0,1) draws for trajectory indices `idx` (array), draw number k < 16."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * _S + (np.uint64(16) * idx + np.uint64(k + 1)) * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def lorenz(ntraj, seed=1234, start=0):
    i = np.arange(start, start + ntraj)
    base = np.array([10.0, 28.0, 8.0 / 3.0])
    p = np.empty((3, ntraj))
    for j in range(3):
        p[j] = base[j] * (1.0 + 0.1 * (2.0 * uniform(seed, i, j) - 1.0))
    u0 = np.zeros((3, ntraj))
    u0[0] = 1.0
    return dict(model="lorenz", u0=u0, p=p, t0=0.0, tf=10.0)


def robertson(ntraj, seed=1234, start=0):
    i = np.arange(start, start + ntraj)
    base = np.array([0.04, 3.0e7, 1.0e4])
    p = np.empty((3, ntraj))
    for j in range(3):
        p[j] = base[j] * 10.0 ** (uniform(seed, i, j) - 0.5)
    u0 = np.zeros((3, ntraj))
    u0[0] = 1.0
    return dict(model="robertson", u0=u0, p=p, t0=0.0, tf=1.0e5)


def vdp(ntraj, seed=1234, start=0):
    i = np.arange(start, start + ntraj)
    mu = 10.0 ** (4.0 * uniform(seed, i, 0))
    u0 = np.zeros((2, ntraj))
    u0[0] = 2.0
    return dict(model="vdp", u0=u0, p=mu[None, :].copy(), t0=0.0, tf=2.0 * mu)


def brusselator(ntraj, seed=1234, start=0, nmax=128):
    i = np.arange(start, start + ntraj)
    ngrid = 8 + np.floor(57.0 * uniform(seed, i, 0)).astype(np.int32)
    ngrid = np.minimum(ngrid, 64)
    B = 2.0 + 2.0 * uniform(seed, i, 1)
    tf = 1.0 + 19.0 * uniform(seed, i, 2)
    p = np.stack([np.ones(ntraj), B, np.full(ntraj, 0.02)])
    u0 = np.zeros((nmax, ntraj))
    for g in np.unique(ngrid):
        sel = ngrid == g
        x = np.arange(g) / g
        u0[0:2 * g:2, sel] = (1.0 + np.sin(2.0 * np.pi * x))[:, None]
        u0[1:2 * g:2, sel] = 3.0
    return dict(model="brusselator", u0=u0, p=p, t0=0.0, tf=tf, nvec=(2 * ngrid).astype(np.int32))


def saveat_uniform(t0, t1, npts):
    """npts points from t0 to t1 inclusive: t0 + ((t1 - t0) * k) / (npts - 1)."""
    k = np.arange(npts, dtype=np.float64)
    return t0 + ((t1 - t0) * k) / (npts - 1)


GENERATORS = dict(lorenz=lorenz, robertson=robertson, vdp=vdp, brusselator=brusselator)
