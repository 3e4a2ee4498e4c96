"""mideq -- MI355X-native ensemble ODE solver (Tsit5 / Rodas5 / Rodas5P /
AutoTsit5(Rodas5()) / AutoTsit5(Rodas5P()) over many independent
trajectories), exposed with the
SciMLBase surface of the reference:

    from mideq import *
    prob = ODEProblem(lorenz, [1.0, 0.0, 0.0], (0.0, 10.0), [10.0, 28.0, 8 / 3])
    ens = EnsembleProblem(prob, prob_func=lambda prob, i, repeat: remake(prob, p=...))
    sol = solve(ens, Tsit5(), EnsembleMI355X(), trajectories=10_000,
                abstol=1e-8, reltol=1e-8)

The compute path is libmideq.so (HIP kernels for gfx950 behind the C ABI in
include/mideq.h); there is no CPU fallback.
"""
from . import workloads
from ._abi import MideqError, header_functions
from .scimlbase import (AutoTsit5, CallbackSet, ContinuousCallback, DeviceOutputFunc,
                        DeviceReduction, DiscreteCallback, EnsembleArrays, EnsembleMI355X, EnsembleProblem,
                        EnsembleSolution, ODEFunction, ODEProblem, ODESolution, ReturnCode,
                        Rodas5, Rodas5P, Stats, Tsit5, UserModel, brusselator, device_solve,
                        exp_growth, linear3, lorenz, remake, robertson, solve,
                        successful_retcode, vanderpol)

__all__ = ["AutoTsit5", "CallbackSet", "ContinuousCallback", "DeviceOutputFunc",
           "DeviceReduction", "DiscreteCallback", "EnsembleArrays", "EnsembleMI355X", "EnsembleProblem",
           "EnsembleSolution", "MideqError", "ODEFunction", "ODEProblem", "ODESolution",
           "ReturnCode", "Rodas5", "Rodas5P", "Stats", "Tsit5", "UserModel", "brusselator",
           "device_solve",
           "exp_growth", "header_functions", "linear3", "lorenz", "remake", "robertson",
           "solve", "successful_retcode", "vanderpol", "workloads"]
