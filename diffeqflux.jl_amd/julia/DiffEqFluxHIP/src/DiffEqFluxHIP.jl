"""
    DiffEqFluxHIP

Routes DiffEqFlux's `NeuralODE` (src/neural_de.jl:35-61 of the reference) to
the MI355X HIP library `libnode_hip.so` (C ABI: include/node_hip.h) for the
supported case, and falls back to the reference method otherwise:

  * model: a Lux `Chain` of `Dense` layers (identity / tanh / softplus / relu /
    sigmoid), optionally preceded by `x -> x.^3`, with equal input/output width;
  * alg: `Tsit5()`;
  * sensealg: the default `InterpolatingAdjoint(autojacvec = ZygoteVJP())`, or
    `BacksolveAdjoint(; checkpointing)` (autojacvec unset or `ZygoteVJP()`);
  * kwargs: saveat (range, vector or number), save_everystep, save_start,
    save_end, abstol, reltol, dt, adaptive, dtmax, dtmin, maxiters;
  * u0: `Vector{Float32}` (D) or `Matrix{Float32}` (D x B), t1 > t0.

`ChainRulesCore.rrule(::typeof(hip_neuralode), ...)` is the pullback: it
calls `node_backward` (the HIP InterpolatingAdjoint, or BacksolveAdjoint as
selected by `NODE_OPT_SENSEALG`) on the forward tape kept in the handle.
`FFJORD` (src/ffjord.jl:55-178) with a `Chain` of `Dense` layers, `Tsit5()`,
`monte_carlo = true`, no regularisation and the standard-normal base goes to
`node_cnf_forward`, its pullback to `node_cnf_backward`.  Julia was not available where this package was written; the
call contract it binds is exercised from Python through the same C ABI.
"""
module DiffEqFluxHIP

using ChainRulesCore: ChainRulesCore, NoTangent, ZeroTangent, unthunk
using DiffEqFlux: DiffEqFlux, NeuralODE, FFJORD
using Lux: Lux, Chain, Dense, WrappedFunction
using OrdinaryDiffEq: Tsit5
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
    function Handle(desc::NodeDesc)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:node_create, libnode()), Int32, (Ref{NodeDesc}, Ref{Ptr{Cvoid}}), desc, r)
        rc == 0 || error("node_create failed ($rc)")
        P = ccall((:node_num_params, libnode()), Int64, (Ref{NodeDesc},), desc)
        h = new(r[], Int(desc.widths[1]), Int(desc.batch), Int(P), Float64[])
        finalizer(x -> ccall((:node_destroy, libnode()), Int32, (Ptr{Cvoid},), x.ptr), h)
        return h
    end
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
function sense_mode(n::NeuralODE)
    haskey(n.kwargs, :sensealg) || return Int64(0)
    s = n.kwargs[:sensealg]
    s isa InterpolatingAdjoint && s.autojacvec isa ZygoteVJP && return Int64(0)
    if s isa BacksolveAdjoint && (s.autojacvec === nothing || s.autojacvec isa ZygoteVJP)
        return s.checkpointing ? Int64(1) : Int64(2)
    end
    return nothing
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

const HANDLES = Dict{Tuple{UInt, Int}, Handle}()

function handle_for(n::NeuralODE, B::Int; precision::Int32 = Int32(0))
    key = (objectid(n), B)
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
    if o.save_everystep != 0
        out = Array{Float32}(undef, h.D, h.B, st.nsaved)
        check(h, ccall((:node_copy_saved, libnode()), Int32, (Ptr{Cvoid}, Ptr{Float32}, Int64),
                       h.ptr, out, st.nsaved), "node_copy_saved")
    end
    ts = Vector{Float64}(undef, st.nsaved)
    check(h, ccall((:node_saved_times, libnode()), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int64),
                   h.ptr, ts, length(ts)), "node_saved_times")
    return out[:, :, 1:st.nsaved], ts, RETCODES[st.retcode + 1]
end

function CRC.rrule(::typeof(hip_neuralode), x, p, h::Handle, o::SolveOpts, interior)
    out, ts, ret = hip_neuralode(x, p, h, o, interior)
    function hip_neuralode_pullback(Δ)
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
    return (out, ts, ret), hip_neuralode_pullback
end

"""
Solution returned on the HIP path: `.u`, `.t`, `.retcode`, `Array(sol)`,
`last`, `sum`, indexing -- what downstream code uses of an `ODESolution`
(docs/src/examples/neural_ode.md:37, GPUs.md:107, cuda_tests.jl:52).
"""
struct HIPSolution{A, T}
    array::A
    t::T
    retcode::Symbol
end
Base.Array(s::HIPSolution) = s.array
Base.getproperty(s::HIPSolution, f::Symbol) =
    f === :u ? [selectdim(getfield(s, :array), ndims(getfield(s, :array)), i)
                for i in 1:size(getfield(s, :array), ndims(getfield(s, :array)))] :
    getfield(s, f)
Base.last(s::HIPSolution) = selectdim(s.array, ndims(s.array), size(s.array, ndims(s.array)))
Base.sum(s::HIPSolution) = sum(s.array)
Base.length(s::HIPSolution) = length(s.t)

# the drop-in: a more specific call method for supported NeuralODEs
function (n::NeuralODE{<:Chain})(x::AbstractArray{Float32}, p, st)
    supported(n, x) || return invoke(n, Tuple{Any, Any, Any}, x, p, st)  # reference path
    B = ndims(x) == 1 ? 1 : size(x, 2)
    h = handle_for(n, B)
    check(h, ccall((:node_set_option, libnode()), Int32, (Ptr{Cvoid}, Int32, Int64),
                   h.ptr, NODE_OPT_SENSEALG, sense_mode(n)), "node_set_option")
    o, interior = make_opts(n)
    # flat Lux/ComponentArray order (test/BasicNeuralDE/neural_de_tests.jl:45); for a
    # ComponentArray `Vector(p)` is already that vector and Zygote restructures dp
    pflat = p isa AbstractVector ? Vector{Float32}(p) :
        reduce(vcat, [vcat(vec(l.weight), l.bias) for l in values(p)
                      if l isa NamedTuple && haskey(l, :weight)])
    out, ts, ret = hip_neuralode(x, pflat, h, o, interior)
    arr = ndims(x) == 1 ? reshape(out, size(out, 1), size(out, 3)) : out
    return HIPSolution(arr, ts, ret), st
end

# ---- FFJORD (src/ffjord.jl:55-178): monte_carlo = true, no regularisation,
# standard-normal base distribution, model(z) (NODE_MAP_CNF)
const NODE_MAP_CNF = Int32(2)

function ffjord_spec(model)
    spec = chain_spec(model)
    spec === nothing && return nothing
    widths, acts, imap = spec
    (imap == 0 && acts[end] == 0 && length(acts) >= 2 && widths[end] + 1 <= 64) || return nothing
    return widths, acts
end

supported(n::FFJORD, x, st) = ffjord_spec(n.model) !== nothing && length(n.args) >= 1 &&
    n.args[1] isa Tsit5 && n.basedist === nothing && eltype(x) === Float32 && ndims(x) == 2 &&
    !st.regularize && st.monte_carlo && n.tspan[2] > n.tspan[1]

function ffjord_handle_for(n::FFJORD, B::Int; precision::Int32 = Int32(0))
    key = (objectid(n), B)
    get!(HANDLES, key) do
        widths, acts = ffjord_spec(n.model)
        w = ntuple(i -> i <= length(widths) ? widths[i] : Int32(0), MAX_LAYERS + 1)
        a = ntuple(i -> i <= length(acts) ? acts[i] : Int32(0), MAX_LAYERS)
        Handle(NodeDesc(Int32(length(acts)), w, a, NODE_MAP_CNF, precision, Int64(B)))
    end
end

"""
    hip_ffjord(x, p_flat, e, h, opts) -> (logpx (1 x B), z(T) (D x B), delta_logp (1 x B))

The FFJORD log-likelihood solve on the GPU (node_cnf_forward) with the
Hutchinson probe `e` (D x B) held for the whole solve.  Differentiable through
the `rrule` below (node_cnf_backward: the layer's InterpolatingAdjoint with the
second-order VJP of the trace term, same probe).
"""
function hip_ffjord(x::Matrix{Float32}, p::AbstractVector{Float32}, e::Matrix{Float32}, h::Handle,
                    o::SolveOpts)
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
    st.retcode == 1 || error("FFJORD solve failed: $(RETCODES[st.retcode + 1])")
    return reshape(lp, 1, B), z, reshape(dl, 1, B)
end

function CRC.rrule(::typeof(hip_ffjord), x, p, e, h::Handle, o::SolveOpts)
    lp, z, dl = hip_ffjord(x, p, e, h, o)
    function hip_ffjord_pullback(Δ)
        tang(t, like) = (t = unthunk(t); t isa Union{NoTangent, ZeroTangent} ? zero(like) : t)
        Δlp, Δz, Δdl = tang(Δ[1], lp), tang(Δ[2], z), tang(Δ[3], dl)
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
        return NoTangent(), dx, Float32.(dp), NoTangent(), NoTangent(), NoTangent()
    end
    return (lp, z, dl), hip_ffjord_pullback
end

function make_opts(n::FFJORD)
    kw = n.kwargs
    t0, t1 = Float64.(n.tspan)
    return SolveOpts(t0, t1, get(kw, :abstol, 1e-6), get(kw, :reltol, 1e-3),
                     Int32(get(kw, :adaptive, true)), Int32(0), Int32(1), Int32(0),
                     Float64(get(kw, :dt, 0.0)), Float64(get(kw, :dtmax, 0.0)),
                     Float64(get(kw, :dtmin, 0.0)), Int64(get(kw, :maxiters, 0)))
end

# the drop-in for FFJORD: one Gaussian probe per call, held for the whole solve
# (the reference redraws randn! inside every right-hand-side call, :96)
function (n::FFJORD{<:Chain})(x::AbstractMatrix{Float32}, ps, st)
    supported(n, x, st) || return invoke(n, Tuple{Any, Any, Any}, x, ps, st)  # reference path
    D, B = size(x)
    h = ffjord_handle_for(n, B)
    pflat = ps isa AbstractVector ? Vector{Float32}(ps) :
        reduce(vcat, [vcat(vec(l.weight), l.bias) for l in values(ps)
                      if l isa NamedTuple && haskey(l, :weight)])
    e = CRC.@ignore_derivatives randn(Float32, D, B)
    lp, _, dl = hip_ffjord(Matrix{Float32}(x), pflat, e, h, make_opts(n))
    return (lp, dl, dl), (; model = st.model, regularize = st.regularize,
                          monte_carlo = st.monte_carlo)
end

end # module
