"""
    MI355XEnsembles

`EnsembleMI355X`: an `SciMLBase.EnsembleAlgorithm` that solves every
trajectory of an `EnsembleProblem` on one AMD MI355X through libmideq.so
(C ABI in include/mideq.h).  Drop-in for `EnsembleThreads()`:

    using DifferentialEquations, MI355XEnsembles
    sol = solve(EnsembleProblem(prob; prob_func, output_func), Tsit5(),
                EnsembleMI355X(); trajectories = 10_000, saveat = 0.1,
                abstol = 1e-8, reltol = 1e-8)

Reference plug point: `SciMLBase.__solve(::AbstractEnsembleProblem, alg,
::EnsembleAlgorithm; trajectories, ...)`, the dispatch EnsembleThreads uses
(names re-exported by DifferentialEquations.jl, src/DifferentialEquations.jl:8-9;
test/qa/qa.jl:50,52,56).  The right-hand side must be registered
(`MODEL_FOR` below): as a built-in device model, or with its body as C++
statements that the library compiles for the GPU at run time
(`register_model!(f, n, np, rhs)`); any other `f` falls back to
`EnsembleThreads` with a warning, as the backend cannot run arbitrary Julia
code on the device.

This file is source only in this repository (no Julia toolchain is present
where it is built and tested); tests/test_gpu_api.py drives the same C ABI
through Python ctypes with the same packing, retcode mapping and
prob_func / output_func / rerun / reduction order.
"""
module MI355XEnsembles

using SciMLBase
import SciMLBase: __solve, EnsembleSolution, ReturnCode
# method types re-exported by DifferentialEquations (test/qa/qa.jl:15,97,119)
using OrdinaryDiffEq: Tsit5, Rodas5, Rodas5P, CompositeAlgorithm

export EnsembleMI355X, register_model!, DeviceCallback, solve_resident!

const LIB = Ref{String}(get(ENV, "MIDEQ_LIB",
    joinpath(@__DIR__, "..", "_build", "libmideq.so")))

Base.@kwdef struct EnsembleMI355X <: SciMLBase.EnsembleAlgorithm
    precision::Symbol = :fp64          # :fp64 or :mixed
    schedule::Symbol = :auto           # :auto, :lane, :persistent
    device::Int = 0
    save_dtype::Type = Float64
end

# ---- C structs (field order and types mirror include/mideq.h) -------------
struct MideqProblem
    model::Int32; n::Int32; np::Int32; reserved0::Int32
    ntraj::Int64
    u0::Ptr{Float64}; p::Ptr{Float64}; t0::Ptr{Float64}; tf::Ptr{Float64}
    tspan0::Float64; tspan1::Float64
    nvec::Ptr{Int32}
end

struct MideqOptions
    alg::Int32; precision::Int32
    abstol::Float64; reltol::Float64
    dt::Float64; dtmin::Float64; dtmax::Float64
    maxiters::Int64
    saveat::Ptr{Float64}; ns::Int32; save_dtype::Int32
    dt0::Ptr{Float64}; alg_hint::Ptr{Int8}
    qsteady_default::Int32; schedule::Int32; block::Int32; blocks_per_cu::Int32
    lanes::Int32; graph::Int32; controller::Int32; auto_passes::Int32
    order::Ptr{Int64}
    reduction::Int32; save_layout::Int32
    iter_cap::Ptr{Int32}
    tstops::Ptr{Float64}; ntstops::Int32; reserved::Int32
end

struct MideqOutputs
    us::Ptr{Cvoid}; u_final::Ptr{Float64}; t_final::Ptr{Float64}; retcode::Ptr{Int8}
    naccept::Ptr{Int32}; nreject::Ptr{Int32}; nf::Ptr{Int32}; nstiff::Ptr{Int32}
    dt_first::Ptr{Float64}
    nevent::Ptr{Int32}; t_event::Ptr{Float64}
    out::Ptr{Float64}; reduced::Ptr{Float64}
    ev_t::Ptr{Float64}; ev_u::Ptr{Float64}; ev_n::Ptr{Int32}; ev_cap::Int32; reserved::Int32
end

# ---- registered right-hand sides ------------------------------------------
# f (the user's Julia function object) => device model id.  Users register
# their in-place f! once: register_model!(lorenz!, :lorenz).
const MODEL_IDS = Dict(:exp => 0, :lorenz => 1, :robertson => 2, :vdp => 3,
                       :brusselator => 4, :linear => 5)
const MODEL_FOR = IdDict{Any, Int32}()
register_model!(f, name::Symbol) = (MODEL_FOR[f] = MODEL_IDS[name]; f)

struct MideqCallback
    kind::Int32; direction::Int32; terminate::Int32; interp_points::Int32
    condition::Ptr{UInt8}; affect::Ptr{UInt8}
    save_positions::Int32; reserved::Int32
end

# registered models whose callbacks save points at events (save_positions)
const EVENT_SAVES = IdDict{Any, Bool}()
# f => compiled with save_everystep (every accepted step's end point kept)
const EVERYSTEP = IdDict{Any, Bool}()

struct MideqUserModel
    n::Int32; np::Int32; rhs::Ptr{UInt8}; jac::Ptr{UInt8}
    ncallbacks::Int32; save_everystep::Int32; callbacks::Ptr{MideqCallback}
    nout::Int32; with_tstops::Int32; output::Ptr{UInt8}
end

"""
    DeviceCallback(kind, condition; affect = nothing, direction = 0,
                   terminate = false, interp_points = 10, save_positions = (true, true))

The device form of `ContinuousCallback(condition, affect!)` (`kind = :continuous`)
or `DiscreteCallback(condition, affect!)` (`kind = :discrete`) (test/qa/qa.jl:26,36):
`condition` a C++ expression of `u[]`, `p[]`, `t`; `affect` statements changing
`u[]`; compiled into a model by `register_model!(...; callbacks)`.  `save_positions`
as upstream: the state just before / just after each affect joins the solution.
"""
Base.@kwdef struct DeviceCallback
    kind::Symbol
    condition::String
    affect::Union{Nothing, String} = nothing
    direction::Int = 0
    terminate::Bool = false
    interp_points::Int = 10
    save_positions::Tuple{Bool, Bool} = (true, true)
end
save_bits(cb::DeviceCallback) = Int32(cb.save_positions[1]) | (Int32(cb.save_positions[2]) << 1)
DeviceCallback(kind::Symbol, condition::AbstractString; kw...) =
    DeviceCallback(; kind, condition = String(condition), kw...)

"""
    register_model!(f!, n, np, rhs; jac = nothing, callbacks = [], output = nothing,
                    nout = 0, save_everystep = false, with_tstops = false, device = 0)

Registers `f!` with a device right-hand side written as C++ statements for
u' = f(u, p, t): assign `du[0..n-1]` from `u[i]`, `p[j]` (0-based) and the
time `t` (naming `t` makes the model non-autonomous); the number type is `T`, constants `2.0` or `T(2)`, functions
`sqrt exp log sin cos tanh fabs pow(x, c)`.  `jac` assigns `J[i][j]`;
without it the Rosenbrock methods differentiate `rhs` in forward mode.
`with_tstops = true` lets `solve(...; tstops)` land steps on those times.
`save_everystep = true` compiles in upstream's save_everystep: each
solution then holds every accepted step's end (up to `event_cap` per
trajectory).
The model id belongs to `device`'s context.  Example (test/core.jl:22-27):

    register_model!(lorenz!, 3, 3, """
        du[0] = p[0] * (u[1] - u[0]);
        du[1] = u[0] * (p[1] - u[2]) - u[1];
        du[2] = u[0] * u[1] - p[2] * u[2];""")
"""
function register_model!(f, n::Integer, np::Integer, rhs::AbstractString;
                         jac::Union{Nothing, AbstractString} = nothing,
                         callbacks::AbstractVector{DeviceCallback} = DeviceCallback[],
                         output::Union{Nothing, AbstractString} = nothing, nout::Integer = 0,
                         save_everystep::Bool = false, with_tstops::Bool = false,
                         device::Int = 0)
    ctx = context(device)
    id = Ref{Int32}(0)
    r = String(rhs)
    j = jac === nothing ? nothing : String(jac)
    o = output === nothing ? nothing : String(output)
    conds = [cb.condition for cb in callbacks]
    affs = [cb.affect === nothing ? "" : cb.affect for cb in callbacks]
    rc = GC.@preserve r j o conds affs begin
        cbs = [MideqCallback(cb.kind === :discrete ? 1 : 0, cb.direction, cb.terminate,
                             cb.interp_points, pointer(conds[k]),
                             cb.affect === nothing ? C_NULL : pointer(affs[k]),
                             save_bits(cb), 0)
               for (k, cb) in enumerate(callbacks)]
        GC.@preserve cbs begin
            m = MideqUserModel(n, np, pointer(r), j === nothing ? C_NULL : pointer(j),
                               length(cbs), Int32(save_everystep),
                               isempty(cbs) ? C_NULL : pointer(cbs),
                               o === nothing ? 0 : nout, Int32(with_tstops),
                               o === nothing ? C_NULL : pointer(o))
            ccall((:mideq_model_register, LIB[]), Cint,
                  (Ptr{Cvoid}, Ref{MideqUserModel}, Ref{Int32}), ctx.h, m, id)
        end
    end
    check(ctx, rc, "mideq_model_register")
    MODEL_FOR[f] = id[]
    # points beside saveat (events, every step) come back in the ev arrays
    EVENT_SAVES[f] = save_everystep || any(cb -> save_bits(cb) != 0, callbacks)
    EVERYSTEP[f] = save_everystep
    return f
end

alg_code(::Tsit5) = Int32(0)
alg_code(::Rodas5) = Int32(1)
alg_code(::Rodas5P) = Int32(3)
# AutoTsit5(Rodas5()) / AutoTsit5(Rodas5P())
alg_code(alg::CompositeAlgorithm) = alg.algs[2] isa Rodas5P ? Int32(4) :
                                    alg.algs[2] isa Rodas5 ? Int32(2) :
                                    error("EnsembleMI355X: AutoTsit5 needs Rodas5() or Rodas5P()")
alg_code(alg) = error("EnsembleMI355X supports Tsit5, Rodas5, Rodas5P and AutoTsit5 of those")

const RETCODES = (ReturnCode.Default, ReturnCode.Success, ReturnCode.MaxIters,
                  ReturnCode.DtLessThanMin, ReturnCode.Unstable, ReturnCode.DtNaN,
                  ReturnCode.InitialFailure, ReturnCode.Terminated)

mutable struct Ctx
    h::Ptr{Cvoid}
end
const CTX = Dict{Int, Ctx}()

function context(dev::Int)
    get!(CTX, dev) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:mideq_ctx_create, LIB[]), Cint, (Ref{Ptr{Cvoid}}, Int32), h, dev)
        rc == 0 || error("mideq_ctx_create failed ($rc)")
        c = Ctx(h[])
        finalizer(x -> ccall((:mideq_ctx_destroy, LIB[]), Cvoid, (Ptr{Cvoid},), x.h), c)
        c
    end
end

function check(ctx, rc, what)
    rc == 0 && return
    msg = unsafe_string(ccall((:mideq_last_error, LIB[]), Cstring, (Ptr{Cvoid},), ctx.h))
    error("$what failed ($rc): $msg")
end

# save_start / save_end (nothing: upstream's defaults -- true for a scalar
# saveat, else whether t0 / tf is among the points): OrdinaryDiffEq keeps the
# points strictly inside the span and saves the ends by these flags.
function _saveat_vector(saveat, tspan, save_start = nothing, save_end = nothing)
    saveat === nothing && return Float64[]
    if saveat isa Number
        # in the direction of integration, handed over ascending
        dir = tspan[2] >= tspan[1] ? 1 : -1
        pts = Float64.(collect(tspan[1]:(dir * saveat):tspan[2]))
        (last(pts) - tspan[2]) * dir < 0 && push!(pts, tspan[2])
    else
        pts = Float64.(collect(saveat))
    end
    for (e, flag) in ((Float64(tspan[1]), save_start), (Float64(tspan[2]), save_end))
        flag === nothing && continue
        filter!(x -> x != e, pts)
        flag && push!(pts, e)
    end
    return sort(pts)
end

function __solve(ens::SciMLBase.AbstractEnsembleProblem, alg, ealg::EnsembleMI355X;
                 trajectories, batch_size = trajectories, saveat = nothing, abstol = 1e-6,
                 reltol = 1e-3, dt = 0.0, dtmin = 0.0, dtmax = 0.0, maxiters = 100_000,
                 dense = saveat === nothing, save_start = nothing, save_end = nothing,
                 tstops = nothing, kwargs...)
    f = ens.prob.f.f
    if !haskey(MODEL_FOR, f)
        @warn "EnsembleMI355X: unregistered right-hand side, falling back to EnsembleThreads"
        return __solve(ens, alg, EnsembleThreads(); trajectories, batch_size, saveat, abstol,
                       reltol, kwargs...)
    end
    t_start = time()
    N = trajectories
    trajectories ÷ batch_size < 1 && error("trajectories ÷ batch_size cannot be less than 1")
    # batches of batch_size trajectories, the reduction after each, stopping
    # once it reports convergence
    u = ens.u_init === nothing ? [] : ens.u_init
    converged = false
    for first in 1:batch_size:N
        II = first:min(first + batch_size - 1, N)
        batch = _solve_batch(ens, alg, ealg, f, II; saveat, abstol, reltol, dt, dtmin,
                             dtmax, maxiters, dense, save_start, save_end, tstops)
        u, converged = ens.reduction(u, batch, II)
        converged && break
    end
    # one batch with the default reduction: the ensemble is complete
    length(1:batch_size:N) == 1 && ens.reduction === SciMLBase.DEFAULT_REDUCTION &&
        (converged = true)
    return EnsembleSolution(u, time() - t_start, converged)
end

# One mideq_ensemble_solve call over the problems `probs` (one per trajectory,
# all of one registered f) with the ascending saveat points `sa`; `lanes` = 0
# lets the library pick the team shape.  Returns the output arrays and the
# team shape used.
struct MideqStats
    kernel_ms::Float64; total_ms::Float64
    schedule::Int32; block::Int32; grid::Int32; launches::Int32; lanes::Int32; graph::Int32
end

function _device_solve(probs, alg, ealg::EnsembleMI355X, f, sa; abstol, reltol, dt, dtmin,
                       dtmax, maxiters, lanes = 0, event_cap = 0, tstops = Float64[])
    M = length(probs)
    n = maximum(length(pr.u0) for pr in probs)
    np = max(length(probs[1].p), 1)   # a user model without parameters reads one row
    # column-major (M, n): trajectory index fastest, the SoA layout of the ABI
    U0 = zeros(Float64, M, n); P = zeros(Float64, M, np)
    T0 = Float64[pr.tspan[1] for pr in probs]; TF = Float64[pr.tspan[2] for pr in probs]
    nvec = Int32[length(pr.u0) for pr in probs]
    for (k, pr) in enumerate(probs)
        U0[k, 1:length(pr.u0)] .= pr.u0
        P[k, 1:length(pr.p)] .= collect(pr.p)
    end
    ns = length(sa)
    US = ealg.save_dtype === Float32 ? zeros(Float32, M, n, ns) : zeros(Float64, M, n, ns)
    UF = zeros(Float64, M, n); TFo = zeros(Float64, M)
    RC = zeros(Int8, M); NA = zeros(Int32, M); NR = zeros(Int32, M)
    NF = zeros(Int32, M); NS = zeros(Int32, M)
    # points saved at callback events: the ABI's [ntraj][cap] times and
    # [ntraj][cap][n] states are (cap, M) and (n, cap, M) column-major
    ev = get(EVENT_SAVES, f, false)
    # points per trajectory beside saveat: 16 for events; with save_everystep
    # what ~1 GiB holds over the batch (16..65,536); event_cap > 0 sets it
    event_cap = event_cap > 0 ? event_cap :
                get(EVERYSTEP, f, false) ? clamp((1 << 27) ÷ (M * (n + 1)), 16, 65536) : 16
    EVT = ev ? fill(NaN, event_cap, M) : zeros(Float64, 0)
    EVU = ev ? fill(NaN, n, event_cap, M) : zeros(Float64, 0)
    EVN = ev ? zeros(Int32, M) : zeros(Int32, 0)
    st = Ref(MideqStats(0, 0, 0, 0, 0, 0, 0, 0))
    ctx = context(ealg.device)
    GC.@preserve U0 P T0 TF nvec sa US UF TFo RC NA NR NF NS EVT EVU EVN tstops begin
        prob_c = MideqProblem(MODEL_FOR[f], n, np, 0, M, pointer(U0), pointer(P),
                              pointer(T0), pointer(TF), T0[1], TF[1],
                              MODEL_FOR[f] == 4 ? pointer(nvec) : C_NULL)
        opt_c = MideqOptions(alg_code(alg), ealg.precision === :mixed ? 1 : 0,
                             abstol, reltol, dt, dtmin, dtmax, maxiters,
                             ns > 0 ? pointer(sa) : C_NULL, ns,
                             ealg.save_dtype === Float32 ? 1 : 0, C_NULL, C_NULL, 1,
                             ealg.schedule === :lane ? 1 : ealg.schedule === :persistent ? 2 : 0,
                             0, 0, lanes, 0, 0, 0, C_NULL, 0, 0, C_NULL,
                             isempty(tstops) ? C_NULL : pointer(tstops), length(tstops), 0)
        out_c = MideqOutputs(ns > 0 ? pointer(US) : C_NULL, pointer(UF), pointer(TFo),
                             pointer(RC), pointer(NA), pointer(NR), pointer(NF), pointer(NS),
                             C_NULL, C_NULL, C_NULL, C_NULL, C_NULL,
                             ev ? pointer(EVT) : C_NULL, ev ? pointer(EVU) : C_NULL,
                             ev ? pointer(EVN) : C_NULL, ev ? event_cap : 0, 0)
        rc = ccall((:mideq_ensemble_solve, LIB[]), Cint,
                   (Ptr{Cvoid}, Ref{MideqProblem}, Ref{MideqOptions}, Ref{MideqOutputs}),
                   ctx.h, prob_c, opt_c, out_c)
        check(ctx, rc, "mideq_ensemble_solve")
        rc = ccall((:mideq_last_stats, LIB[]), Cint, (Ptr{Cvoid}, Ref{MideqStats}), ctx.h, st)
        check(ctx, rc, "mideq_last_stats")
    end
    return (; US, UF, TFo, RC, NA, NR, NF, NS, EVT, EVU, EVN, ev, lanes = st[].lanes)
end

# Device-resident solve (the zero-copy EnsembleSolution view): the inputs and
# outputs are arrays already in HBM (AMDGPU.jl ROCArrays, or any array whose
# `pointer` is a device address) in the ABI's layout -- U0 (N, n) and P
# (N, np) column-major, so the trajectory index runs fastest; T0, TF (N);
# outputs UF (N, n), TFo, RC (Int8), NA, NR, NF, NS (Int32) of length N, and
# US (N, n, ns) with saveat.  Calls mideq_ensemble_solve_device, enqueued on
# `stream` (a hipStream_t; C_NULL = the context's stream) without
# synchronizing: nothing is copied to or from the host, and the caller keeps
# every array alive until the stream has passed the solve.
function solve_resident!(f, alg, ealg::EnsembleMI355X, U0, P, T0, TF, out;
                         saveat = nothing, nvec = nothing, abstol = 1e-6, reltol = 1e-3,
                         dt = 0.0, dtmin = 0.0, dtmax = 0.0, maxiters = 0,
                         stream::Ptr{Cvoid} = C_NULL)
    haskey(MODEL_FOR, f) || error("solve_resident!: unregistered right-hand side")
    N, n = size(U0, 1), size(U0, 2)
    ns = saveat === nothing ? 0 : length(saveat)
    ctx = context(ealg.device)
    ptr(x) = x === nothing ? C_NULL : reinterpret(Ptr{Cvoid}, pointer(x))
    prob_c = MideqProblem(MODEL_FOR[f], n, max(size(P, 2), 1), 0, N, ptr(U0), ptr(P),
                          ptr(T0), ptr(TF), 0.0, 0.0, ptr(nvec))
    opt_c = MideqOptions(alg_code(alg), ealg.precision === :mixed ? 1 : 0,
                         abstol, reltol, dt, dtmin, dtmax, maxiters, ptr(saveat), ns,
                         ealg.save_dtype === Float32 ? 1 : 0, C_NULL, C_NULL, 1,
                         ealg.schedule === :lane ? 1 : ealg.schedule === :persistent ? 2 : 0,
                         0, 0, 0, 0, 0, 0, C_NULL, 0, 0, C_NULL, C_NULL, 0, 0)
    out_c = MideqOutputs(ns > 0 ? ptr(out.US) : C_NULL, ptr(out.UF), ptr(out.TFo),
                         ptr(out.RC), ptr(out.NA), ptr(out.NR), ptr(out.NF), ptr(out.NS),
                         C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, 0, 0)
    rc = ccall((:mideq_ensemble_solve_device, LIB[]), Cint,
               (Ptr{Cvoid}, Ref{MideqProblem}, Ref{MideqOptions}, Ref{MideqOutputs},
                Ptr{Cvoid}), ctx.h, prob_c, opt_c, out_c, stream)
    check(ctx, rc, "mideq_ensemble_solve_device")
    return out
end

# Dense output sol(t) (test/core.jl:51-58): the trajectory solved again with
# the requested times as saveat points.  Saved points do not change the steps
# taken, so the values are the interpolants of the steps of the solve that
# made the solution; NaN outside the integrated span.  Values only.
struct MideqDense{F} <: SciMLBase.AbstractDiffEqInterpolation
    resolve::F   # ascending Vector{Float64} of times -> states at them
end

function _dense_values(id::MideqDense, tvals)
    ts = collect(Float64, tvals)
    perm = sortperm(ts)
    vals = id.resolve(ts[perm])
    out = similar(vals)
    out[perm] = vals
    return out
end

function (id::MideqDense)(tvals, idxs, deriv, p, continuity::Symbol = :left)
    deriv === Val{0} || error("EnsembleMI355X dense output gives values only")
    vals = _dense_values(id, tvals)
    vals = idxs === nothing ? vals : [v[idxs] for v in vals]
    return SciMLBase.DiffEqArray(vals, collect(tvals))
end

function (id::MideqDense)(t::Number, idxs, deriv, p, continuity::Symbol = :left)
    deriv === Val{0} || error("EnsembleMI355X dense output gives values only")
    v = _dense_values(id, (t,))[1]
    return idxs === nothing ? v : v[idxs]
end

# The trajectories II of one batch: prob_func, the device solve, output_func
# with its rerun loop; returns their outputs in index order.  Without saveat
# every solution has dense output (sol(t)), as upstream's default.
function _solve_batch(ens, alg, ealg::EnsembleMI355X, f, II; saveat, abstol, reltol, dt,
                      dtmin, dtmax, maxiters, dense = saveat === nothing,
                      save_start = nothing, save_end = nothing, tstops = nothing)
    outputs = Dict{Int,Any}()
    pending = collect(II)
    repeat = 1
    sa = _saveat_vector(saveat, ens.prob.tspan, save_start, save_end)
    ns = length(sa)
    ts = tstops === nothing ? Float64[] : sort(Float64.(collect(tstops)))
    tol = (; abstol, reltol, dt, dtmin, dtmax, maxiters, tstops = ts)
    while !isempty(pending)
        probs = [ens.prob_func(deepcopy(ens.prob), i, repeat) for i in pending]
        r = _device_solve(probs, alg, ealg, f, sa; tol...)
        rerun = Int[]
        for (k, i) in enumerate(pending)
            pr = probs[k]
            nk = length(pr.u0)
            # a scalar problem (u0 isa Number) gets scalar states, as upstream
            val(row) = pr.u0 isa Number ? row[1] : Vector{Float64}(row)
            back = r.TFo[k] < pr.tspan[1]
            # points saved at events, in the order of integration
            m = r.ev ? min(Int(r.EVN[k]), size(r.EVT, 1)) : 0
            et = r.ev ? r.EVT[1:m, k] : Float64[]
            eu = [val(r.EVU[1:nk, j, k]) for j in 1:m]
            if ns > 0
                keep = [j for j in 1:ns if !isnan(r.US[k, 1, j])]
                back && reverse!(keep)   # backward: order of integration
                ts = sa[keep]; us = [val(r.US[k, 1:nk, j]) for j in keep]
                if m > 0   # merged; a saveat point first at an equal time
                    allt = vcat(ts, et); allu = vcat(us, eu)
                    perm = sortperm(back ? -allt : allt; alg = Base.Sort.DEFAULT_STABLE)
                    ts = allt[perm]; us = allu[perm]
                end
                # save_everystep saves the start (save_start) when saveat lacks it
                if get(EVERYSTEP, f, false) && (isempty(ts) || ts[1] != pr.tspan[1])
                    ts = vcat([pr.tspan[1]], ts); us = vcat([val([pr.u0;])], us)
                end
            else
                ts = vcat([pr.tspan[1]], et); us = vcat([val([pr.u0;])], eu)
                # the postamble's final point, unless the last point saved has its time
                cut = r.ev && Int(r.EVN[k]) > size(r.EVT, 1)
                if (m == 0 || et[end] != r.TFo[k] || cut) && save_end !== false
                    ts = vcat(ts, [r.TFo[k]]); us = vcat(us, [val(r.UF[k, 1:nk])])
                end
                if save_start === false
                    ts = ts[2:end]; us = us[2:end]
                end
            end
            if dense
                lanes = r.lanes
                resolve = function (tq)
                    d = _device_solve([pr], alg, ealg, f, tq; tol..., lanes)
                    [val(d.US[1, 1:nk, j]) for j in eachindex(tq)]
                end
                sol = SciMLBase.build_solution(pr, alg, ts, us; dense = true,
                                               interp = MideqDense(resolve),
                                               retcode = RETCODES[r.RC[k] + 1])
            else
                sol = SciMLBase.build_solution(pr, alg, ts, us;
                                               retcode = RETCODES[r.RC[k] + 1])
            end
            out, again = ens.output_func(sol, i)
            outputs[i] = out
            again && push!(rerun, i)
        end
        pending = rerun
        repeat += 1
    end
    return [outputs[i] for i in II]
end

end # module
