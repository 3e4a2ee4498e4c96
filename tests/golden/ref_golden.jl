# Reference vectors for the ensemble path, from the reference itself
# (DifferentialEquations.jl 8.0.3: OrdinaryDiffEq 7, SciMLBase 3), for a
# machine with Julia and those packages installed:
#
#     julia -t auto ref_golden.jl ref_vectors.json
#
# No Julia exists where this repository is built and tested; the script is
# the recipe for pinning what only the reference can pin (accept/reject
# counts of upstream Tsit5/Rodas5/AutoTsit5, its interpolant's values).  The
# inputs are the seeded tables of mideq/workloads.py (the same SplitMix64
# counter generator, restated below), so every vector compares trajectory
# by trajectory with this library's output; tests/test_ref_golden.py reads
# the file when it is present.
#
# Call sites followed: solve(prob, Tsit5()) and solve(prob, Rodas5P())
# (src/DifferentialEquations.jl:27-28, test/core.jl:14,47), EnsembleProblem /
# EnsembleThreads (test/qa/qa.jl:50,56), saveat (test/core.jl:90-96).
using DifferentialEquations

const M1 = 0xBF58476D1CE4E5B9
const M2 = 0x94D049BB133111EB
const G = 0x9E3779B97F4A7C15
const S = 0xD1B54A32D192ED03

# k-th U[0,1) draw of 0-based trajectory i under seed
function uniform(seed::Integer, i::Integer, k::Integer)
    z = UInt64(seed) * S + (UInt64(16) * UInt64(i) + UInt64(k + 1)) * G
    z = (z ⊻ (z >> 30)) * M1
    z = (z ⊻ (z >> 27)) * M2
    z = z ⊻ (z >> 31)
    return Float64(z >> 11) * (1.0 / 9007199254740992.0)
end

lorenz_p(i; seed = 1234) =
    [b * (1.0 + 0.1 * (2.0 * uniform(seed, i, j - 1) - 1.0))
     for (j, b) in enumerate((10.0, 28.0, 8.0 / 3.0))]
robertson_p(i; seed = 1234) =
    [b * 10.0^(uniform(seed, i, j - 1) - 0.5) for (j, b) in enumerate((0.04, 3.0e7, 1.0e4))]
vdp_mu(i; seed = 1234) = 10.0^(4.0 * uniform(seed, i, 0))

function lorenz!(du, u, p, t)
    du[1] = p[1] * (u[2] - u[1])
    du[2] = u[1] * (p[2] - u[3]) - u[2]
    du[3] = u[1] * u[2] - p[3] * u[3]
end
function robertson!(du, u, p, t)
    du[1] = -p[1] * u[1] + p[3] * u[2] * u[3]
    du[2] = p[1] * u[1] - p[2] * u[2]^2 - p[3] * u[2] * u[3]
    du[3] = p[2] * u[2]^2
end
function robertson_jac!(J, u, p, t)
    J[1, 1] = -p[1];  J[1, 2] = p[3] * u[3];                      J[1, 3] = p[3] * u[2]
    J[2, 1] = p[1];   J[2, 2] = -2p[2] * u[2] - p[3] * u[3];      J[2, 3] = -p[3] * u[2]
    J[3, 1] = 0.0;    J[3, 2] = 2p[2] * u[2];                     J[3, 3] = 0.0
end
function vdp!(du, u, p, t)
    du[1] = u[2]
    du[2] = p[1] * (1 - u[1]^2) * u[2] - u[1]
end

function run(name, f, u0, tspan_of, p_of, N, alg; saveat = nothing, kw...)
    prob = ODEProblem(f, u0, tspan_of(0), p_of(0))
    pf = (prob, i, repeat) -> remake(prob; p = p_of(i - 1), tspan = tspan_of(i - 1))
    ens = EnsembleProblem(prob; prob_func = pf)
    skw = saveat === nothing ? (save_everystep = false, save_start = false) : (saveat = saveat,)
    sol = solve(ens, alg, EnsembleThreads(); trajectories = N, skw..., kw...)
    out = Dict{String, Any}()
    out["u_final"] = [collect(s.u[end]) for s in sol.u]
    out["naccept"] = [s.stats.naccept for s in sol.u]
    out["nreject"] = [s.stats.nreject for s in sol.u]
    out["nf"] = [s.stats.nf for s in sol.u]
    out["retcode"] = [string(s.retcode) for s in sol.u]
    if saveat !== nothing
        out["t"] = collect(sol.u[1].t)
        out["saved"] = [[collect(v) for v in s.u] for s in sol.u]
    end
    return name => out
end

# minimal JSON writer (numbers at full precision)
js(x::AbstractString) = "\"" * x * "\""
js(x::Integer) = string(x)
js(x::AbstractFloat) = isfinite(x) ? repr(Float64(x)) : "null"
js(x::AbstractVector) = "[" * join(js.(x), ",") * "]"
js(x::AbstractDict) = "{" * join([js(string(k)) * ":" * js(v) for (k, v) in x], ",") * "}"

function main(path)
    res = Dict{String, Any}()
    push!(res, run("lorenz_tsit5_1e-8_endpoint", lorenz!, [1.0, 0.0, 0.0], i -> (0.0, 10.0),
                   lorenz_p, 10_000, Tsit5(); abstol = 1e-8, reltol = 1e-8))
    push!(res, run("lorenz_tsit5_1e-8_saveat100", lorenz!, [1.0, 0.0, 0.0], i -> (0.0, 10.0),
                   lorenz_p, 1_000, Tsit5(); saveat = range(0.0, 10.0; length = 100),
                   abstol = 1e-8, reltol = 1e-8))
    rf = ODEFunction(robertson!; jac = robertson_jac!)
    push!(res, run("robertson_rodas5_1e-8_1e-6", rf, [1.0, 0.0, 0.0], i -> (0.0, 1e5),
                   robertson_p, 2_000, Rodas5(); abstol = 1e-8, reltol = 1e-6))
    push!(res, run("robertson_rodas5p_1e-8_1e-6", rf, [1.0, 0.0, 0.0], i -> (0.0, 1e5),
                   robertson_p, 2_000, Rodas5P(); abstol = 1e-8, reltol = 1e-6))
    push!(res, run("vdp_autotsit5_1e-8_1e-6", vdp!, [2.0, 0.0], i -> (0.0, 2vdp_mu(i)),
                   i -> [vdp_mu(i)], 2_000, AutoTsit5(Rodas5()); abstol = 1e-8, reltol = 1e-6))
    open(path, "w") do io
        write(io, js(res))
    end
end

main(length(ARGS) > 0 ? ARGS[1] : "ref_vectors.json")
