"""
    DiffEqFluxHIP

Routes DiffEqFlux's `NeuralODE` (src/neural_de.jl:35-61 of the reference) to
the MI355X HIP library `libnode_hip.so` (C ABI: include/node_hip.h) for the
supported case, and falls back to the reference method otherwise:

  * model: a Lux `Chain` of `Dense` layers (identity / tanh / softplus / relu /
    sigmoid), optionally preceded by `x -> x.^3`, with equal input/output width;
  * alg: `Tsit5()`;
  * sensealg: the default `InterpolatingAdjoint(autojacvec = ZygoteVJP())`, with
    or without `checkpointing` (the forward keeps checkpoint states only; the
    adjoint re-solves each interval), or `BacksolveAdjoint(; checkpointing)`
    (autojacvec unset or `ZygoteVJP()`);
  * GEMM precision: the keyword `hip_precision` (`:fp32`, `:bf16`, `:bf16x3`;
    no counterpart in the reference, removed before any reference call) or
    the module default `DiffEqFluxHIP.PRECISION[]` (`:fp32`);
  * kwargs: saveat (range, vector or number), save_everystep, save_start,
    save_end, abstol, reltol, dt, adaptive, dtmax, dtmin, maxiters;
  * u0: `Vector{Float32}` (D) or `Matrix{Float32}` (D x B), t1 > t0.

`ChainRulesCore.rrule(::typeof(hip_neuralode), ...)` is the pullback: it
calls `node_backward` (the HIP InterpolatingAdjoint, or BacksolveAdjoint as
selected by `NODE_OPT_SENSEALG`) on the forward tape kept in the handle.  A
handle keeps the tape of its latest forward only: a pullback whose forward was
superseded (the same layer called again at the same batch size before the
first pullback ran) raises instead of differentiating the wrong solve.  The
call returns a SciMLBase `ODESolution` built from the saved states
(`build_solution`: `.u`, `.t`, `.retcode`, `.stats`, `Array(sol)`, indexing,
`sol(t)` by linear interpolation between the saved states as for any solve
without dense output).
`FFJORD` (src/ffjord.jl:55-178) with a `Chain` of `Dense` layers, `Tsit5()`
and the standard-normal base goes to `node_cnf_forward` (`monte_carlo = false`:
the exact trace, `NODE_MAP_CNF_EXACT_TRACE`; `regularize = true`: the two
regulariser rows, `NODE_MAP_CNF_REGULARIZE`, returned as λ₁, λ₂ -- the rows,
where the reference's `selectdim(pred, N, ...)` picks sample columns), its
pullback to `node_cnf_backward` (the regularisers carry no gradient there: a
non-zero cotangent for λ₁ / λ₂ raises).  `FFJORDDistribution` of such an
`FFJORD` (src/ffjord.jl:216-246) samples through the same solve over the
reversed time span (`rand`, as `__backward_ffjord`, :183-214) and evaluates
`logpdf` through the forward solve.  Julia was not available where this
package was written; the call contract it binds is exercised from Python
through the same C ABI.
"""
module DiffEqFluxHIP

using ChainRulesCore: ChainRulesCore, NoTangent, ZeroTangent, unthunk
using DiffEqFlux: DiffEqFlux, NeuralODE, FFJORD, FFJORDDistribution
using Distributions: Distributions
using Random: AbstractRNG
using Lux: Lux, Chain, Dense, WrappedFunction
using OrdinaryDiffEq: Tsit5
using SciMLBase: SciMLBase, ODEProblem, ODEFunction, ReturnCode
using SciMLSensitivity: InterpolatingAdjoint, BacksolveAdjoint, ZygoteVJP

const CRC = ChainRulesCore

libnode() = get(ENV, "DIFFEQFLUX_HIP_LIB",
                joinpath(@__DIR__, "..", "..", "..", "diffeqflux_hip", "libnode_hip.so"))

const MAX_LAYERS = 8

# ---- C structs (layout of include/node_hip.h)
struct NodeDesc
    n_layers::Int32
    widths::NTuple{MAX_LAYERS + 1, Int32}
    acts::NTuple{MAX_LAYERS, Int32}
    input_map::Int32
    precision::Int32
    batch::Int64
end

struct SolveOpts
    t0::Float64; t1::Float64; abstol::Float64; reltol::Float64
    adaptive::Int32; save_start::Int32; save_end::Int32; save_everystep::Int32
    dt::Float64; dtmax::Float64; dtmin::Float64; maxiters::Int64
end

mutable struct Stats
    nf::Int64; naccept::Int64; nreject::Int64; nsaved::Int64
    retcode::Int32; reserved::Int32; t_final::Float64
end
Stats() = Stats(0, 0, 0, 0, 0, 0, 0.0)

const ACT = Dict(identity => Int32(0), tanh => Int32(1), Lux.NNlib.softplus => Int32(2),
                 Lux.NNlib.relu => Int32(3), Lux.NNlib.sigmoid => Int32(4))
const RETCODES = (:Default, :Success, :Terminated, :DtNaN, :MaxIters, :DtLessThanMin,
                  :Unstable)

# ---- handle
mutable struct Handle
    ptr::Ptr{Cvoid}
    D::Int
    B::Int
    P::Int
    saved_t::Vector{Float64}
    gen::Int            # forward solves run on this handle (its tape is the latest's)
    function Handle(desc::NodeDesc)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:node_create, libnode()), Int32, (Ref{NodeDesc}, Ref{Ptr{Cvoid}}), desc, r)
        rc == 0 || error("node_create failed ($rc)")
        P = ccall((:node_num_params, libnode()), Int64, (Ref{NodeDesc},), desc)
        h = new(r[], Int(desc.widths[1]), Int(desc.batch), Int(P), Float64[], 0)
        finalizer(x -> ccall((:node_destroy, libnode()), Int32, (Ptr{Cvoid},), x.ptr), h)
        return h
    end
end

function check_current(h::Handle, gen::Int, what)
    h.gen == gen && return nothing
    error("$what: the layer ran another forward solve at this batch size since this one " *
          "(a handle keeps the tape of its latest forward); take the gradient before calling " *
          "the layer again")
end

function check(h::Handle, rc, what)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:node_last_error, libnode()), Cstring, (Ptr{Cvoid},), h.ptr))
    error("$what failed ($rc): $msg")
end

# ---- support check (anything else falls back to the reference path)
function chain_spec(model)
    model isa Chain || return nothing
    layers = collect(values(model.layers))
    imap = Int32(0)
    if !isempty(layers) && layers[1] isa WrappedFunction
        imap = Int32(1)          # x -> x.^3 is the only wrapped function on the path
        layers = layers[2:end]
    end
    (isempty(layers) || length(layers) > MAX_LAYERS) && return nothing
    all(l -> l isa Dense && haskey(ACT, l.activation) && Lux.has_bias(l), layers) ||
        return nothing
    widths = Int32[layers[1].in_dims; [l.out_dims for l in layers]...]
    widths[1] == widths[end] || return nothing
    acts = Int32[ACT[l.activation] for l in layers]
    return widths, acts, imap
end

# NODE_OPT_SENSEALG value of the layer's sensealg (nothing: not on the HIP path)
const NODE_OPT_SENSEALG = Int32(5)
const NODE_OPT_CHECKPOINTING = Int32(6)
# (NODE_SENSE_*, checkpointing flag) of the layer's sensealg; nothing: not on the HIP path
function sense_mode(n::NeuralODE)
    haskey(n.kwargs, :sensealg) || return (Int64(0), Int64(0))
    s = n.kwargs[:sensealg]
    s isa InterpolatingAdjoint && s.autojacvec isa ZygoteVJP &&
        return (Int64(0), Int64(s.checkpointing ? 1 : 0))
    if s isa BacksolveAdjoint && (s.autojacvec === nothing || s.autojacvec isa ZygoteVJP)
        return (s.checkpointing ? Int64(1) : Int64(2), Int64(0))
    end
    return nothing
end

# GEMM precision (NODE_PREC_*): the hip_precision keyword, else the module default
const PRECISION = Ref{Symbol}(:fp32)
const PREC = Dict(:fp32 => Int32(0), :bf16 => Int32(1), :bf16x3 => Int32(2))
hip_precision(n) = PREC[get(n.kwargs, :hip_precision, PRECISION[])]
# the layer without the keywords that only this package reads (for the reference path)
function reference_layer(n::NeuralODE)
    haskey(n.kwargs, :hip_precision) || return n
    kw = Base.structdiff(NamedTuple(n.kwargs), NamedTuple{(:hip_precision,)})
    return NeuralODE(n.model, n.tspan, n.args...; kw...)
end

supported(n::NeuralODE, x) = chain_spec(n.model) !== nothing && length(n.args) >= 1 &&
    n.args[1] isa Tsit5 && eltype(x) === Float32 && ndims(x) in (1, 2) &&
    n.tspan[2] > n.tspan[1] && sense_mode(n) !== nothing

# OrdinaryDiffEq's saving keywords -> interior save times + flags
function save_plan(t0, t1, kw)
    saveat = get(kw, :saveat, Float64[])
    isnum = saveat isa Number
    seq = isnum ? Float64[] : Float64.(collect(saveat))
    empty = !isnum && isempty(seq)
    sev = get(kw, :save_everystep, empty)
    interior = isnum ? [t for t in collect(t0:saveat:t1)[2:end] if t0 < t < t1] :
                       sort(unique([t for t in seq if t0 < t < t1]))
    ss = get(kw, :save_start, sev || empty || isnum || t0 in seq)
    se = get(kw, :save_end, sev || empty || isnum || t1 in seq)
    return Float64.(interior), ss, se, sev
end

function make_opts(n::NeuralODE)
    kw = n.kwargs
    t0, t1 = Float64.(n.tspan)
    interior, ss, se, sev = save_plan(t0, t1, kw)
    o = SolveOpts(t0, t1, get(kw, :abstol, 1e-6), get(kw, :reltol, 1e-3),
                  Int32(get(kw, :adaptive, true)), Int32(ss), Int32(se), Int32(sev),
                  Float64(get(kw, :dt, 0.0)), Float64(get(kw, :dtmax, 0.0)),
                  Float64(get(kw, :dtmin, 0.0)), Int64(get(kw, :maxiters, 0)))
    return o, interior
end

const HANDLES = Dict{Tuple{UInt, Int, Int32}, Handle}()

function handle_for(n::NeuralODE, B::Int; precision::Int32 = Int32(0))
    key = (objectid(n), B, precision)
    get!(HANDLES, key) do
        widths, acts, imap = chain_spec(n.model)
        w = ntuple(i -> i <= length(widths) ? widths[i] : Int32(0), MAX_LAYERS + 1)
        a = ntuple(i -> i <= length(acts) ? acts[i] : Int32(0), MAX_LAYERS)
        Handle(NodeDesc(Int32(length(acts)), w, a, imap, precision, Int64(B)))
    end
end

"""
    hip_neuralode(x, p_flat, h, opts, interior) -> Array (D x B x Nt) and times

Forward Tsit5 solve on the GPU (node_forward).  Differentiable through the
`rrule` below (node_backward = InterpolatingAdjoint).
"""
function hip_neuralode(x::AbstractArray{Float32}, p::AbstractVector{Float32}, h::Handle,
                       o::SolveOpts, interior::Vector{Float64})
    check(h, ccall((:node_set_params, libnode()), Int32, (Ptr{Cvoid}, Ptr{Float32}, Int64),
                   h.ptr, p, length(p)), "node_set_params")
    st = Stats()
    cap = o.save_everystep != 0 ? 0 : length(interior) + o.save_start + o.save_end
    out = Array{Float32}(undef, h.D, h.B, max(cap, 1))
    u0 = Matrix{Float32}(reshape(x, h.D, h.B))
    rc = GC.@preserve u0 interior out ccall((:node_forward, libnode()), Int32,
        (Ptr{Cvoid}, Ref{SolveOpts}, Ptr{Float32}, Ptr{Float64}, Int64, Ptr{Float32}, Int64,
         Ref{Stats}), h.ptr, o, u0, interior, length(interior),
        cap > 0 ? pointer(out) : C_NULL, cap, st)
    check(h, rc, "node_forward")
    h.gen += 1
    if o.save_everystep != 0
        out = Array{Float32}(undef, h.D, h.B, st.nsaved)
        check(h, ccall((:node_copy_saved, libnode()), Int32, (Ptr{Cvoid}, Ptr{Float32}, Int64),
                       h.ptr, out, st.nsaved), "node_copy_saved")
    end
    ts = Vector{Float64}(undef, st.nsaved)
    check(h, ccall((:node_saved_times, libnode()), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int64),
                   h.ptr, ts, length(ts)), "node_saved_times")
    return out[:, :, 1:st.nsaved], ts, RETCODES[st.retcode + 1], st
end

function CRC.rrule(::typeof(hip_neuralode), x, p, h::Handle, o::SolveOpts, interior)
    out, ts, ret, fst = hip_neuralode(x, p, h, o, interior)
    gen = h.gen
    function hip_neuralode_pullback(Δ)
        check_current(h, gen, "NeuralODE pullback")
        Δout = unthunk(Δ[1])
        Δout isa Union{NoTangent, ZeroTangent} && (Δout = zero(out))
        g = Array{Float32}(Δout)
        du0 = Matrix{Float32}(undef, h.D, h.B)
        dp = Vector{Float64}(undef, h.P)
        st = Stats()
        rc = ccall((:node_backward, libnode()), Int32,
                   (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float64}, Ref{Stats}),
                   h.ptr, g, du0, dp, st)
        check(h, rc, "node_backward")
        return NoTangent(), reshape(du0, size(x)), Float32.(dp), NoTangent(), NoTangent(),
               NoTangent()
    end
    return (out, ts, ret, fst), hip_neuralode_pullback
end

const RETCODE_ENUM = (ReturnCode.Default, ReturnCode.Success, ReturnCode.Terminated,
                      ReturnCode.DtNaN, ReturnCode.MaxIters, ReturnCode.DtLessThanMin,
                      ReturnCode.Unstable)

# the drop-in: a more specific call method for supported NeuralODEs
function (n::NeuralODE{<:Chain})(x::AbstractArray{Float32}, p, st)
    modes = sense_mode(n)
    (supported(n, x) && modes !== nothing) ||
        return invoke(reference_layer(n), Tuple{Any, Any, Any}, x, p, st)  # reference path
    B = ndims(x) == 1 ? 1 : size(x, 2)
    h = handle_for(n, B; precision = hip_precision(n))
    check(h, ccall((:node_set_option, libnode()), Int32, (Ptr{Cvoid}, Int32, Int64),
                   h.ptr, NODE_OPT_SENSEALG, modes[1]), "node_set_option")
    check(h, ccall((:node_set_option, libnode()), Int32, (Ptr{Cvoid}, Int32, Int64),
                   h.ptr, NODE_OPT_CHECKPOINTING, modes[2]), "node_set_option")
    o, interior = make_opts(n)
    # flat Lux/ComponentArray order (test/BasicNeuralDE/neural_de_tests.jl:45); for a
    # ComponentArray `Vector(p)` is already that vector and Zygote restructures dp
    pflat = p isa AbstractVector ? Vector{Float32}(p) :
        reduce(vcat, [vcat(vec(l.weight), l.bias) for l in values(p)
                      if l isa NamedTuple && haskey(l, :weight)])
    out, ts, ret, fst = hip_neuralode(x, pflat, h, o, interior)
    # the ODESolution of src/neural_de.jl:54-60, built from the saved states
    # (no dense output: sol(t) interpolates linearly between them)
    u = ndims(x) == 1 ? [out[:, 1, i] for i in axes(out, 3)] :
                        [out[:, :, i] for i in axes(out, 3)]
    dudt(u, p, t) = first(n.model(u, p, st))
    prob = ODEProblem{false}(ODEFunction{false}(dudt), x, n.tspan, p)
    stats = SciMLBase.DEStats(; nf = fst.nf, naccept = fst.naccept, nreject = fst.nreject)
    sol = SciMLBase.build_solution(prob, n.args[1], ts, u; retcode = RETCODE_ENUM[ret_index(ret)],
                                   stats = stats, dense = false)
    return sol, st
end
ret_index(ret::Symbol) = findfirst(==(ret), RETCODES)

# ---- FFJORD (src/ffjord.jl:55-178): standard-normal base distribution,
# model(z) (NODE_MAP_CNF); regularize / monte_carlo = false as input_map flags
const NODE_MAP_CNF = Int32(2)
const NODE_MAP_CNF_REGULARIZE = Int32(0x10)
const NODE_MAP_CNF_EXACT_TRACE = Int32(0x20)

function ffjord_spec(model)
    spec = chain_spec(model)
    spec === nothing && return nothing
    widths, acts, imap = spec
    (imap == 0 && acts[end] == 0 && length(acts) >= 2 && widths[end] + 3 <= 64) || return nothing
    return widths, acts
end

supported(n::FFJORD, x, st) = ffjord_spec(n.model) !== nothing && length(n.args) >= 1 &&
    n.args[1] isa Tsit5 && n.basedist === nothing && eltype(x) === Float32 && ndims(x) == 2 &&
    n.tspan[2] > n.tspan[1]

function ffjord_handle_for(n::FFJORD, B::Int, regularize::Bool, monte_carlo::Bool;
                           precision::Int32 = Int32(0))
    imap = NODE_MAP_CNF | (regularize ? NODE_MAP_CNF_REGULARIZE : Int32(0)) |
           (monte_carlo ? Int32(0) : NODE_MAP_CNF_EXACT_TRACE)
    key = (objectid(n), B, precision | (imap << 8))
    get!(HANDLES, key) do
        widths, acts = ffjord_spec(n.model)
        w = ntuple(i -> i <= length(widths) ? widths[i] : Int32(0), MAX_LAYERS + 1)
        a = ntuple(i -> i <= length(acts) ? acts[i] : Int32(0), MAX_LAYERS)
        Handle(NodeDesc(Int32(length(acts)), w, a, imap, precision, Int64(B)))
    end
end

"""
    hip_ffjord(x, p_flat, e, h, opts, regularize) -> (logpx (1 x B), z(T) (D x B),
                                                       delta_logp (1 x B), λ₁ (1 x B), λ₂ (1 x B))

The FFJORD log-likelihood solve on the GPU (node_cnf_forward) with the
Hutchinson probe `e` (D x B) held for the whole solve.  Differentiable through
the `rrule` below (node_cnf_backward: the layer's InterpolatingAdjoint with the
second-order VJP of the trace term, same probe).
"""
function hip_ffjord(x::Matrix{Float32}, p::AbstractVector{Float32}, e::Matrix{Float32}, h::Handle,
                    o::SolveOpts, regularize::Bool)
    pv = Vector{Float32}(p)
    check(h, ccall((:node_set_params, libnode()), Int32, (Ptr{Cvoid}, Ptr{Float32}, Int64),
                   h.ptr, pv, length(pv)), "node_set_params")
    D, B = size(x)
    z = Matrix{Float32}(undef, D, B)
    dl = Vector{Float32}(undef, B)
    lp = Vector{Float32}(undef, B)
    st = Stats()
    rc = ccall((:node_cnf_forward, libnode()), Int32,
               (Ptr{Cvoid}, Ref{SolveOpts}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                Ptr{Float32}, Ptr{Float32}, Ref{Stats}), h.ptr, o, x, e, z, dl, lp, st)
    check(h, rc, "node_cnf_forward")
    h.gen += 1
    st.retcode == 1 || error("FFJORD solve failed: $(RETCODES[st.retcode + 1])")
    l1, l2 = dl, dl                 # λ₁ = λ₂ = Δlogp without regularisation (:166-167)
    if regularize
        l1, l2 = Vector{Float32}(undef, B), Vector{Float32}(undef, B)
        check(h, ccall((:node_cnf_regularizers, libnode()), Int32,
                       (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}), h.ptr, l1, l2),
              "node_cnf_regularizers")
    end
    return reshape(lp, 1, B), z, reshape(dl, 1, B), reshape(l1, 1, B), reshape(l2, 1, B)
end

function CRC.rrule(::typeof(hip_ffjord), x, p, e, h::Handle, o::SolveOpts, regularize::Bool)
    lp, z, dl, l1, l2 = hip_ffjord(x, p, e, h, o, regularize)
    gen = h.gen
    function hip_ffjord_pullback(Δ)
        check_current(h, gen, "FFJORD pullback")
        tang(t, like) = (t = unthunk(t); t isa Union{NoTangent, ZeroTangent} ? zero(like) : t)
        Δlp, Δz, Δdl = tang(Δ[1], lp), tang(Δ[2], z), tang(Δ[3], dl)
        if regularize
            iszero(tang(Δ[4], l1)) && iszero(tang(Δ[5], l2)) ||
                error("FFJORD: gradients through the regularisers (λ₁, λ₂) are not on the HIP path")
        else  # λ₁ = λ₂ = Δlogp
            Δdl = Δdl .+ tang(Δ[4], l1) .+ tang(Δ[5], l2)
        end
        # logpx = -|z|^2 / 2 - D/2 log(2π) - delta_logp
        gz = Matrix{Float32}(Δz .- Δlp .* z)
        gd = Vector{Float32}(vec(Δdl .- Δlp))
        dx = Matrix{Float32}(undef, size(x))
        dp = Vector{Float64}(undef, h.P)
        st = Stats()
        rc = ccall((:node_cnf_backward, libnode()), Int32,
                   (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float64}, Ref{Stats}),
                   h.ptr, gz, gd, dx, dp, st)
        check(h, rc, "node_cnf_backward")
        return NoTangent(), dx, Float32.(dp), NoTangent(), NoTangent(), NoTangent(), NoTangent()
    end
    return (lp, z, dl, l1, l2), hip_ffjord_pullback
end

function make_opts(n::FFJORD)
    kw = n.kwargs
    t0, t1 = Float64.(n.tspan)
    return SolveOpts(t0, t1, get(kw, :abstol, 1e-6), get(kw, :reltol, 1e-3),
                     Int32(get(kw, :adaptive, true)), Int32(0), Int32(1), Int32(0),
                     Float64(get(kw, :dt, 0.0)), Float64(get(kw, :dtmax, 0.0)),
                     Float64(get(kw, :dtmin, 0.0)), Int64(get(kw, :maxiters, 0)))
end

flat_params(ps) = ps isa AbstractVector ? Vector{Float32}(ps) :
    reduce(vcat, [vcat(vec(l.weight), l.bias) for l in values(ps)
                  if l isa NamedTuple && haskey(l, :weight)])

# the drop-in for FFJORD: one Gaussian probe per call, held for the whole solve
# (the reference redraws randn! inside every right-hand-side call, :96)
function (n::FFJORD{<:Chain})(x::AbstractMatrix{Float32}, ps, st)
    supported(n, x, st) || return invoke(n, Tuple{Any, Any, Any}, x, ps, st)  # reference path
    D, B = size(x)
    h = ffjord_handle_for(n, B, st.regularize, st.monte_carlo)
    pflat = flat_params(ps)
    e = CRC.@ignore_derivatives randn(Float32, D, B)
    lp, _, dl, l1, l2 = hip_ffjord(Matrix{Float32}(x), pflat, e, h, make_opts(n), st.regularize)
    return (lp, l1, l2), (; model = st.model, regularize = st.regularize,
                          monte_carlo = st.monte_carlo)
end

# ---- FFJORDDistribution (src/ffjord.jl:216-246)
const FFJORDDistHIP = FFJORDDistribution{<:FFJORD{<:Chain}}

# base samples x (D x n) through the flow backwards in time: the augmented
# solve of `__backward_ffjord` (:183-214) over reverse(tspan), z returned
function hip_ffjord_sample(n::FFJORD, x::Matrix{Float32}, pflat, st)
    D, B = size(x)
    h = ffjord_handle_for(n, B, st.regularize, st.monte_carlo)
    o = make_opts(n)
    rev = SolveOpts(o.t1, o.t0, o.abstol, o.reltol, o.adaptive, o.save_start, o.save_end,
                    o.save_everystep, o.dt, o.dtmax, o.dtmin, o.maxiters)
    _, z, _, _, _ = hip_ffjord(x, pflat, randn(Float32, D, B), h, rev, st.regularize)
    return z
end

function Distributions._rand!(rng::AbstractRNG, d::FFJORDDistHIP, x::AbstractArray{<:Real})
    n = d.model
    nsamp = size(x, ndims(x))
    base = randn(rng, Float32, n.input_dims..., nsamp)             # :186-191
    supported(n, base, d.st) ||                                    # reference path
        return invoke(Distributions._rand!, Tuple{AbstractRNG, FFJORDDistribution,
                                                  AbstractArray{<:Real}}, rng, d, x)
    copyto!(x, hip_ffjord_sample(n, base, flat_params(d.ps), d.st))
    return x
end

# first(first(...)) as the reference (:235-240): the first sample's logpx
function Distributions._logpdf(d::FFJORDDistHIP, x::AbstractVector)
    return first(first(d.model(reshape(Vector{Float32}(x), :, 1), d.ps, d.st)))
end
function Distributions._logpdf(d::FFJORDDistHIP, x::AbstractArray)
    return first(first(d.model(Matrix{Float32}(reshape(x, size(x, 1), :)), d.ps, d.st)))
end

end # module
