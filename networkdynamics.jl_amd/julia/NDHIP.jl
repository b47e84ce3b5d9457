# NDHIP.jl -- binds NetworkDynamics.jl's `Network` to libnd_hip.so through `ccall`.
#
#     using NetworkDynamics, AMDGPU, NDHIP
#     nw = Network(g, vertexm, edgem; execution=HIPExecution{true}(), aggregator=SegmentedAggregator(+))
#     nw(du, u, p, t)            # du, u, p :: ROCVector{Float64}
#
# The reference constructor (src/construction.jl:31-236) runs unchanged: it builds
# its IndexManager and batches and calls the aggregator constructor
# (src/construction.jl:198).  `SegmentedAggregator(im, batches, +)` hands the graph and
# component kinds to `nd_network_create`, checks that the library's layout equals
# `im` bit for bit, and creates the device plan.  The specialised call method
# replaces src/coreloop.jl:1-102 for `Network{<:HIPExecution}`.
#
# Only the component functions of the library's table below run on the GPU; any
# other `compf`/`compg` is an ArgumentError at construction.
module NDHIP

using NetworkDynamics
using NetworkDynamics: ExecutionStyle, Aggregator, IndexManager, ComponentBatch, compf, compg, fftype,
                       dim, pdim, outdim, extdim, PureFeedForward, PureStateMap, StateMask, AntiSymmetric,
                       Symmetric, Directed, Fiducial, ExtMap, StateBufIdx, has_external_input
using AMDGPU: ROCVector, AMDGPU
using ForwardDiff: Dual

export HIPExecution, SegmentedAggregator, register_model!, rk4!, tsit5!

const libnd = normpath(joinpath(@__DIR__, "..", "libnd_hip.so"))

struct HIPExecution{buffered} <: ExecutionStyle{buffered} end
NetworkDynamics.iscudacompatible(::Type{<:HIPExecution}) = false

# mirrors nd_kind in include/nd_hip.h
struct NDKind
    is_edge::Int32
    model::Int32
    wrapper::Int32
    ff::Int32
    dim::Int32
    pdim::Int32
    outdim_src::Int32
    outdim_dst::Int32
    mask_src::NTuple{4,Int32}
    mask_dst::NTuple{4,Int32}
    extdim::Int32          # external inputs (extdim(c), external_inputs.jl)
end

# component function -> ND_VM_* / ND_EM_* id (fill from the user's component library,
# e.g. register_model!(Lib.kuramoto_inertia!, 3) for ND_VM_SWING)
const MODEL_IDS = IdDict{Any,Int32}()
register_model!(f, id::Integer) = (MODEL_IDS[f] = Int32(id); nothing)
register_model!(NetworkDynamics.LOOPBACK_G, 10)      # ND_EM_LOOPBACK (post_utils.jl:105-108)

_check(rc) = rc == 0 || throw(ArgumentError(unsafe_string(ccall((:nd_last_error, libnd), Cstring, ()))))

_mask(s::StateMask) = ntuple(i -> i <= length(s.idxs) ? Int32(s.idxs[i]) : Int32(0), 4)
_nomask() = ntuple(_ -> Int32(0), 4)

function _kind(c::VertexModel)
    g = compg(c)
    if !(g isa StateMask)
        # a computed output (a source's NoFeedForward g, an injector's PureFeedForward g):
        # the registered g names the model, the trait goes along as ND_FF_*
        ffid = fftype(c) isa PureFeedForward ? 1 : fftype(c) isa NetworkDynamics.FeedForward ? 2 : 3
        return NDKind(0, get(MODEL_IDS, g, -1), 0, ffid, dim(c), pdim(c), 0, outdim(c), _nomask(), _nomask(),
                      extdim(c))
    end
    NDKind(0, get(MODEL_IDS, compf(c), -1), 0, 0, dim(c), pdim(c), 0, outdim(c), _nomask(), _mask(g),
           extdim(c))
end

function _kind(c::EdgeModel)
    g = compg(c)
    od = outdim(c)
    if g isa Fiducial && g.src isa StateMask
        return NDKind(1, get(MODEL_IDS, compf(c), -1), 4, 0, dim(c), pdim(c), od.src, od.dst,
                      _mask(g.src), _mask(g.dst), extdim(c))
    end
    wrap = g isa AntiSymmetric ? 1 : g isa Symmetric ? 2 : g isa Directed ? 3 : 0
    if wrap != 0 && g.g isa StateMask
        # an ODE edge whose output is a StateMask inside the wrapper (component_functions.jl:
        # 117-175): the model is its f; the mask serves both sides, Directed has no src side
        m = _mask(g.g)
        return NDKind(1, get(MODEL_IDS, compf(c), -1), wrap, 0, dim(c), pdim(c), od.src, od.dst,
                      wrap == 3 ? _nomask() : m, m, extdim(c))
    end
    inner = wrap == 0 ? g : g.g
    NDKind(1, get(MODEL_IDS, inner, -1), wrap, 1, dim(c), pdim(c), od.src, od.dst, _nomask(), _nomask(),
           extdim(c))
end

mutable struct SegmentedAggregator{F} <: Aggregator
    f::F
    net::Ptr{Cvoid}
    plan::Ptr{Cvoid}
end
SegmentedAggregator(f) = (im, batches) -> SegmentedAggregator(im, batches, f)

# the copy constructor Network(nw; ...) / copy(nw) (src/construction.jl:475-536) rebuilds
# the aggregator through this (src/aggregators.jl:287-292): each copy gets its own device
# plan and buffers, so copies can run concurrently (one per thread or stream)
NetworkDynamics.get_aggr_constructor(a::SegmentedAggregator) = SegmentedAggregator(a.f)

function SegmentedAggregator(im::IndexManager, batches, f)
    f === (+) || throw(ArgumentError("SegmentedAggregator only supports +"))
    src = Int64[e.src for e in im.edgevec]
    dst = Int64[e.dst for e in im.edgevec]
    kinds = NDKind[]
    keys = Dict{NDKind,Int32}()
    kid(k) = get!(() -> (push!(kinds, k); Int32(length(kinds) - 1)), keys, k)
    vk = Int32[kid(_kind(m)) for m in im.vertexm]
    ek = Int32[kid(_kind(m)) for m in im.edgem]
    net = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:nd_network_create, libnd), Cint,
                 (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int32, Ptr{NDKind}, Ptr{Int32}, Ptr{Int32}, Ptr{Ptr{Cvoid}}),
                 length(im.vertexm), length(im.edgem), src, dst, length(kinds), kinds, vk, ek, net))
    _check_layout(net[], im)
    if has_external_input(im)
        # the reference's own ExtMap (external_inputs.jl:10-31), as +u / -o indices
        em = [x isa StateBufIdx ? Int64(x.idx) : -Int64(x.idx) for x in ExtMap(im).map]
        _check(ccall((:nd_network_set_extmap, libnd), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}), net[], length(em), em))
    end
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:nd_plan_create, libnd), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), net[], plan))
    a = SegmentedAggregator(f, net[], plan[])
    finalizer(a) do x
        ccall((:nd_plan_destroy, libnd), Cvoid, (Ptr{Cvoid},), x.plan)
        ccall((:nd_network_destroy, libnd), Cvoid, (Ptr{Cvoid},), x.net)
    end
    a
end

# bit-exact check of the library's flat layout against the reference IndexManager
function _check_layout(net, im)
    nv, ne = length(im.v_data), length(im.e_data)
    vr = zeros(Int64, 8, nv)
    er = zeros(Int64, 12, max(ne, 1))
    _check(ccall((:nd_network_export_ranges, libnd), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}), net, vr, er))
    for i in 1:nv
        for (j, r) in enumerate((im.v_data[i], im.v_out[i], im.v_para[i], im.v_aggr[i]))
            (vr[2j-1, i], vr[2j, i]) == (first(r), last(r)) || error("layout mismatch at vertex $i")
        end
    end
    for i in 1:ne
        rs = (im.e_data[i], im.e_out[i].src, im.e_out[i].dst, im.e_para[i], im.e_gbufr[i].src, im.e_gbufr[i].dst)
        for (j, r) in enumerate(rs)
            (er[2j-1, i], er[2j, i]) == (first(r), last(r)) || error("layout mismatch at edge $i")
        end
    end
    nothing
end

# aggregate!(::SegmentedAggregator, aggbuf, o): additive, deterministic (src/aggregators.jl:140)
function NetworkDynamics.aggregate!(a::SegmentedAggregator, aggbuf::ROCVector{Float64}, o::ROCVector{Float64})
    _check(ccall((:nd_aggregate, libnd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                 a.plan, pointer(aggbuf), pointer(o), AMDGPU.stream().stream))
    AMDGPU.synchronize()
    nothing
end

_flatmap(m) = isempty(m) ? Int64[] : collect(Int64, Iterators.flatten(m))

# nw(du, u, p, t; perturb, perturb_maps) for HIPExecution (replaces src/coreloop.jl:1-102).
# With perturb: the buffered loop with the perturbations of coreloop.jl:49-87; the maps are
# the (buffer index, channel) tuples of _classify_perturbation_channels.
function (nw::Network{<:HIPExecution})(du::ROCVector{Float64}, u::ROCVector{Float64}, p::ROCVector{Float64}, t;
                                       perturb=nothing, perturb_maps=nothing)
    a = nw.layer.aggregator::SegmentedAggregator
    if !isnothing(perturb)
        vo, vi = _flatmap(perturb_maps.vo_map), _flatmap(perturb_maps.vi_map)
        eo, ei = _flatmap(perturb_maps.eo_map), _flatmap(perturb_maps.ei_map)
        _check(ccall((:nd_plan_set_perturbation, libnd), Cint,
                     (Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Int64, Ptr{Int64}, Int64, Ptr{Int64}, Int64, Ptr{Int64}),
                     a.plan, length(perturb), length(vo) ÷ 2, vo, length(vi) ÷ 2, vi, length(eo) ÷ 2, eo,
                     length(ei) ÷ 2, ei))
        δ = ROCVector{Float64}(perturb)
        _check(ccall((:nd_rhs_perturbed_jvp, libnd), Cint,
                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                      Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                      Ptr{Cvoid}),
                     a.plan, pointer(du), C_NULL, pointer(u), C_NULL, pointer(p), C_NULL, Float64(t), pointer(δ),
                     C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, AMDGPU.stream().stream))
        AMDGPU.synchronize()
        return nothing
    end
    _check(ccall((:nd_rhs, libnd), Cint,
                 (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Int64, Int64, Int64, Ptr{Cvoid}),
                 a.plan, pointer(du), pointer(u), pointer(p), Float64(t), length(du), length(u), length(p),
                 AMDGPU.stream().stream))
    AMDGPU.synchronize()   # the reference call returns after its work is done (coreloop.jl:100)
    nothing
end

# nw(du, u, p, t) on first-order dual numbers (one partial, Float64), as ForwardDiff calls
# it through the DiffCache buffers (src/network_structure.jl:152-155) for Jacobian-vector
# products (linearize_network, implicit solvers): du = f(u) and its partial = J(u) v.
# The Duals are split into value and partial vectors for nd_rhs_jvp and joined back.
function (nw::Network{<:HIPExecution})(du::ROCVector{Dual{T,Float64,1}}, u::ROCVector{Dual{T,Float64,1}},
                                       p::ROCVector{Float64}, t) where {T}
    a = nw.layer.aggregator::SegmentedAggregator
    n = length(u)
    length(du) == n || throw(ArgumentError("du or u does not have expected size $(n)"))
    uf = reinterpret(Float64, u)           # (value, partial) pairs, contiguous
    uval, upart = uf[1:2:end], uf[2:2:end]
    dval, dpart = similar(uval), similar(upart)
    _check(ccall((:nd_rhs_jvp, libnd), Cint,
                 (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Int64, Int64,
                  Int64, Ptr{Cvoid}),
                 a.plan, pointer(dval), pointer(dpart), pointer(uval), pointer(upart), pointer(p), Float64(t), n, n,
                 length(p), AMDGPU.stream().stream))
    df = reinterpret(Float64, du)
    df[1:2:end] .= dval
    df[2:2:end] .= dpart
    AMDGPU.synchronize()
    nothing
end

# Chunked duals, N = 2..8 partials (ForwardDiff's chunk, the size the reference gives its
# DiffCache buffers, src/construction.jl:205-208): the value and the N partial columns go
# to nd_rhs_jvp_multi in one launch.  Networks on the library's generic kernels have no
# K-partial form there (ND_ERR_UNSUPPORTED): they take one nd_rhs_jvp per partial.
function (nw::Network{<:HIPExecution})(du::ROCVector{Dual{T,Float64,N}}, u::ROCVector{Dual{T,Float64,N}},
                                       p::ROCVector{Float64}, t) where {T,N}
    2 <= N <= 8 || throw(ArgumentError("HIPExecution carries at most 8 partials per dual (got $(N))"))
    a = nw.layer.aggregator::SegmentedAggregator
    n = length(u)
    length(du) == n || throw(ArgumentError("du or u does not have expected size $(n)"))
    um = reshape(reinterpret(Float64, u), N + 1, n)     # per state: value, then the partials
    uval = um[1, :]
    upart = permutedims(um[2:end, :])                  # n x N: partial j contiguous, column j
    dval, dpart = similar(uval), similar(upart)
    rc = ccall((:nd_rhs_jvp_multi, libnd), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int32, Ptr{Float64}, Float64,
                Int64, Int64, Int64, Ptr{Cvoid}),
               a.plan, pointer(dval), pointer(dpart), pointer(uval), pointer(upart), Int32(N), pointer(p), Float64(t),
               n, n, length(p), AMDGPU.stream().stream)
    if rc == 2                                          # ND_ERR_UNSUPPORTED
        for j in 1:N
            off = (j - 1) * n + 1
            _check(ccall((:nd_rhs_jvp, libnd), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64,
                          Int64, Int64, Int64, Ptr{Cvoid}),
                         a.plan, pointer(dval), pointer(dpart, off), pointer(uval), pointer(upart, off), pointer(p),
                         Float64(t), n, n, length(p), AMDGPU.stream().stream))
        end
    else
        _check(rc)
    end
    dm = reshape(reinterpret(Float64, du), N + 1, n)
    dm[1, :] .= dval
    dm[2:end, :] .= permutedims(dpart)
    AMDGPU.synchronize()
    nothing
end

# Everything else the HIP path cannot run: host arrays, other element types, duals with
# more than 8 partials.  The reference's own argument checks throw ArgumentError
# (src/coreloop.jl:2-7, src/network_structure.jl:146-151); so does this, before any
# device work, instead of falling through to the generic coreloop.
function (nw::Network{<:HIPExecution})(du, u, p, t; kwargs...)
    throw(ArgumentError("HIPExecution runs nw(du, u, p, t) with du, u, p as ROCVector{Float64} on the HIP " *
                        "device (du and u may be ROCVectors of ForwardDiff.Dual{T,Float64,N}, N <= 8, for " *
                        "Jacobian-vector products); got $(typeof(du)), $(typeof(u)), $(typeof(p))"))
end

# Solves of ODEProblem(nw, u0, tspan, p) with RK4 / Tsit5 (src/post_utils.jl:43-82) in the
# library's steppers, in place on u.  tsit5! returns `nothing` when the network does not fit
# one workgroup (ND_ERR_UNSUPPORTED): solve it with OrdinaryDiffEq on `nw` instead.
function rk4!(nw::Network{<:HIPExecution}, u::ROCVector{Float64}, p::ROCVector{Float64}, t0, dt, nsteps::Integer)
    a = nw.layer.aggregator::SegmentedAggregator
    _check(ccall((:nd_rk4, libnd), Cint,
                 (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Int64, Ptr{Float64}, Int64, Int64, Ptr{Cvoid}),
                 a.plan, pointer(u), pointer(p), Float64(t0), Float64(dt), Int64(nsteps), C_NULL, length(u), length(p),
                 AMDGPU.stream().stream))
    AMDGPU.synchronize()
    u
end

# mirrors nd_ensemble_stat in include/nd_hip.h
struct NDStat
    t::Float64
    dt::Float64
    naccept::Int64
    nreject::Int64
    nf::Int64
    retcode::Int32
    ntrip::Int32
end

function tsit5!(nw::Network{<:HIPExecution}, u::ROCVector{Float64}, p::ROCVector{Float64}, tspan;
                reltol=1e-8, abstol=1e-8, dt0=0.0, maxiters=10^6)
    a = nw.layer.aggregator::SegmentedAggregator
    st = ROCVector{NDStat}(undef, 1)
    rc = ccall((:nd_tsit5, libnd), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Float64, Float64, Int64, Ptr{NDStat},
                Int64, Int64, Ptr{Cvoid}),
               a.plan, pointer(u), pointer(p), Float64(tspan[1]), Float64(tspan[2]), Float64(reltol), Float64(abstol),
               Float64(dt0), Int64(maxiters), pointer(st), length(u), length(p), AMDGPU.stream().stream)
    rc == 2 && return nothing                              # ND_ERR_UNSUPPORTED
    _check(rc)
    AMDGPU.synchronize()
    Array(st)[1]
end

end # module
