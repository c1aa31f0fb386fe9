# DDDHip.jl -- routes DDD.jl's force path to libdddhip.so when
# dlnParams.parGPU is true.  Include after `using DDD`; the methods below take
# the same arguments as the reference's (src/Processing/DislocationForces.jl,
# src/Processing/DislocationFEM.jl) and pass the network's arrays to the C ABI
# declared in include/ddd_hip.h without copying them.
#
# This file cannot be executed in the build container (no Julia); the same
# ABI is exercised from Python ctypes with identical memory layouts.
module DDDHip

using DDD
using SparseArrays
using StaticArrays

const libddd = joinpath(@__DIR__, "..", "libdddhip.so")

const DDD_ACCUMULATE = Cint(1)
const DDD_OUT_BY_POSITION = Cint(2)
const DDD_PK_HALF_TO_NODES = Cint(4)

struct Consts
    mu4pi::Float64
    mu8pi::Float64
    mu4pinu::Float64
    nuMu4pinu::Float64
    omNuInv::Float64
    nuOmNuInv::Float64
    coreRad::Float64
    coreRadSq::Float64
    coreEnergy::Float64
    sigmaPN::Float64
end

struct Network
    numNode::Int64
    numSeg::Int64
    segIdx::Ptr{Int64}
    segIdxLd::Int64
    coord::Ptr{Float64}
    coordCols::Int64
    bVec::Ptr{Float64}
    bVecCols::Int64
    extSeg::Ptr{UInt8}
    links::Ptr{Int64}
    linksCols::Int64
    connectivity::Ptr{Int64}
    connLd::Int64
    connCols::Int64
end

struct Mesh
    C::Ptr{Float64}
    coord::Ptr{Float64}
    numNode::Int64
    connectivity::Ptr{Int64}
    numElem::Int64
    mx::Int64
    my::Int64
    mz::Int64
    w::Float64
    h::Float64
    d::Float64
    uHat::Ptr{Float64}
end

consts(d::DislocationParameters, m::MaterialParameters) = Consts(
    m.μ4π, m.μ8π, m.μ4πν, m.νμ4πν, m.omνInv, m.νomνInv,
    d.coreRad, d.coreRadSq, d.coreEnergy, m.σPN)

# Vector{Bool} is one byte per element, as the ABI expects.
netstruct(n::DislocationNetwork) = Network(
    n.numNode[1], n.numSeg[1],
    pointer(n.segIdx), size(n.segIdx, 1),
    pointer(n.coord), size(n.coord, 2),
    pointer(n.bVec), size(n.bVec, 2),
    Ptr{UInt8}(pointer(n.extSeg)),
    pointer(n.links), size(n.links, 2),
    pointer(n.connectivity), size(n.connectivity, 1), size(n.connectivity, 2))

# The caller keeps C (Matrix(mesh.C)) and uHat (the densified uHat) alive with
# GC.@preserve for as long as the struct is used.
meshstruct(mesh, C::Matrix{Float64}, uHat::Vector{Float64}) = Mesh(
    pointer(C), pointer(mesh.coord), size(mesh.coord, 2),
    pointer(mesh.connectivity), size(mesh.connectivity, 2),
    mesh.mx, mesh.my, mesh.mz, mesh.w, mesh.h, mesh.d, pointer(uHat))

mutable struct Ctx
    h::Ptr{Cvoid}
end
const CTX = Ref{Ctx}()
function ctx()
    if !isassigned(CTX)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:ddd_ctx_create, libddd), Cint, (Ref{Ptr{Cvoid}}, Cint), h, 0)
        rc == 0 || error("ddd_ctx_create failed ($rc): no gfx950 device")
        CTX[] = Ctx(h[])
    end
    return CTX[].h
end

check(rc) = rc == 0 || error(unsafe_string(ccall((:ddd_last_error, libddd), Cstring, (Ptr{Cvoid},), ctx())))

idxvec(idx) = isnothing(idx) ? Int64[] : collect(Int64, idx isa Integer ? (idx,) : idx)

function calcSegSegForce!(dlnParams, matParams, network::DislocationNetwork, idx = nothing)
    dlnParams.parGPU || return DDD.calcSegSegForce!(dlnParams, matParams, network, idx)
    k = Ref(consts(dlnParams, matParams))
    ix = idxvec(idx)
    GC.@preserve network ix begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_segseg_force, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ptr{Int64}, Int64, Ptr{Float64}, Int64, Cint),
            ctx(), k, nw, isnothing(idx) ? C_NULL : pointer(ix), length(ix),
            pointer(network.segForce), size(network.segForce, 3), DDD_ACCUMULATE))
    end
    return nothing
end

function calcSegSegForce(dlnParams, matParams, network::DislocationNetwork, idx = nothing)
    dlnParams.parGPU || return DDD.calcSegSegForce(dlnParams, matParams, network, idx)
    k = Ref(consts(dlnParams, matParams))
    ix = idxvec(idx)
    m = isnothing(idx) ? network.numSeg[1] : length(ix)
    out = zeros(3, 2, m)
    GC.@preserve network ix out begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_segseg_force, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ptr{Int64}, Int64, Ptr{Float64}, Int64, Cint),
            ctx(), k, nw, isnothing(idx) ? C_NULL : pointer(ix), length(ix),
            pointer(out), m, DDD_OUT_BY_POSITION))
    end
    return out
end

function calcSegForce!(dlnParams, matParams, mesh, forceDisplacement, network::DislocationNetwork,
                       idx = nothing)
    dlnParams.parGPU ||
        return DDD.calcSegForce!(dlnParams, matParams, mesh, forceDisplacement, network, idx)
    k = Ref(consts(dlnParams, matParams))
    ix = idxvec(idx)
    range = isnothing(idx) ? (1:network.numSeg[1]) : ix
    network.segForce[:, :, range] .= 0
    uHat = Vector{Float64}(forceDisplacement.uHat)  # the ABI takes uHat dense
    C = Matrix(mesh.C)
    GC.@preserve network ix uHat C mesh begin
        nw = Ref(netstruct(network))
        ms = Ref(meshstruct(mesh, C, uHat))
        check(ccall((:ddd_seg_force, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ref{Mesh}, Ptr{Int64}, Int64, Ptr{Float64},
             Int64, Cint),
            ctx(), k, nw, ms, isnothing(idx) ? C_NULL : pointer(ix), length(ix),
            pointer(network.segForce), size(network.segForce, 3), DDD_ACCUMULATE))
    end
    return nothing
end

function calcSegForce(dlnParams, matParams, mesh, forceDisplacement, network::DislocationNetwork,
                      idx = nothing)
    dlnParams.parGPU ||
        return DDD.calcSegForce(dlnParams, matParams, mesh, forceDisplacement, network, idx)
    k = Ref(consts(dlnParams, matParams))
    ix = idxvec(idx)
    m = isnothing(idx) ? network.numSeg[1] : length(ix)
    out = zeros(3, 2, m)
    uHat = Vector{Float64}(forceDisplacement.uHat)
    C = Matrix(mesh.C)
    GC.@preserve network ix uHat C mesh out begin
        nw = Ref(netstruct(network))
        ms = Ref(meshstruct(mesh, C, uHat))
        check(ccall((:ddd_seg_force, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ref{Mesh}, Ptr{Int64}, Int64, Ptr{Float64},
             Int64, Cint),
            ctx(), k, nw, ms, isnothing(idx) ? C_NULL : pointer(ix), length(ix),
            pointer(out), m, DDD_OUT_BY_POSITION))
    end
    return out
end

function self_call(dlnParams, matParams, network, idx, out, flags)
    k = Ref(consts(dlnParams, matParams))
    ix = idxvec(idx)
    GC.@preserve network ix out begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_self_force, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ptr{Int64}, Int64, Ptr{Float64}, Int64, Cint),
            ctx(), k, nw, isnothing(idx) ? C_NULL : pointer(ix), length(ix),
            pointer(out), size(out, 3), flags))
    end
end

# calcSelfForce: (selfForceNode1, selfForceNode2), each 3 x numSeg (or
# 3 x length(idx)), node 1 = -node 2.
function calcSelfForce(dlnParams, matParams, network::DislocationNetwork, idx = nothing)
    dlnParams.parGPU || return DDD.calcSelfForce(dlnParams, matParams, network, idx)
    m = isnothing(idx) ? network.numSeg[1] : length(idxvec(idx))
    out = zeros(3, 2, m)
    self_call(dlnParams, matParams, network, idx, out, DDD_OUT_BY_POSITION)
    return out[:, 1, :], out[:, 2, :]
end

function calcSelfForce!(dlnParams, matParams, network::DislocationNetwork, idx = nothing)
    dlnParams.parGPU || return DDD.calcSelfForce!(dlnParams, matParams, network, idx)
    self_call(dlnParams, matParams, network, idx, network.segForce, DDD_ACCUMULATE)
    return nothing
end

# calcPKForce / calcPKForce! / calc_σHat carry no DislocationParameters in the
# reference, so there is no parGPU switch: these methods always run on the GPU.
function pk_call(mesh, forceDisplacement, network, idx, out, cols, flags)
    ix = idxvec(idx)
    uHat = Vector{Float64}(forceDisplacement.uHat)
    C = Matrix(mesh.C)
    GC.@preserve network ix uHat C mesh out begin
        nw = Ref(netstruct(network))
        ms = Ref(meshstruct(mesh, C, uHat))
        check(ccall((:ddd_pk_force, libddd), Cint,
            (Ptr{Cvoid}, Ref{Network}, Ref{Mesh}, Ptr{Int64}, Int64, Ptr{Float64}, Int64, Cint),
            ctx(), nw, ms, isnothing(idx) ? C_NULL : pointer(ix), length(ix), pointer(out), cols,
            flags))
    end
end

function calcPKForce(mesh, forceDisplacement, network::DislocationNetwork, idx = nothing)
    m = isnothing(idx) ? network.numSeg[1] : length(idxvec(idx))
    out = zeros(3, m)
    pk_call(mesh, forceDisplacement, network, idx, out, m, DDD_OUT_BY_POSITION)
    return out
end

function calcPKForce!(mesh, forceDisplacement, network::DislocationNetwork, idx = nothing)
    pk_call(mesh, forceDisplacement, network, idx, network.segForce, size(network.segForce, 3),
            DDD_ACCUMULATE | DDD_PK_HALF_TO_NODES)
    return nothing
end

function calc_σHat(mesh, forceDisplacement, x0)
    x = Vector{Float64}(collect(x0))
    σ = zeros(3, 3)
    uHat = Vector{Float64}(forceDisplacement.uHat)
    C = Matrix(mesh.C)
    GC.@preserve x σ uHat C mesh begin
        ms = Ref(meshstruct(mesh, C, uHat))
        check(ccall((:ddd_sigma_hat, libddd), Cint,
            (Ptr{Cvoid}, Ref{Mesh}, Ptr{Float64}, Int64, Ptr{Float64}),
            ctx(), ms, pointer(x), Int64(1), pointer(σ)))
    end
    return SMatrix{3, 3, Float64}(σ)
end

function calc_σTilde!(σ, x0, dlnParams, matParams, network::DislocationNetwork, idx = nothing)
    dlnParams.parGPU || return DDD.calc_σTilde!(σ, x0, dlnParams, matParams, network, idx)
    pts = x0 isa Base.Iterators.ProductIterator ?
          reduce(hcat, [collect(p) for p in x0]) : Matrix{Float64}(x0)  # x fastest
    @assert size(σ, 1) == 6 && size(σ, 2) == size(pts, 2)
    k = Ref(consts(dlnParams, matParams))
    ix = idxvec(idx)
    GC.@preserve network ix pts σ begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_sigma_tilde, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ptr{Int64}, Int64, Ptr{Float64}, Int64,
             Ptr{Float64}, Cint),
            ctx(), k, nw, isnothing(idx) ? C_NULL : pointer(ix), length(ix),
            pointer(pts), size(pts, 2), pointer(σ), Cint(0)))
    end
    return σ
end

# dlnMobility / dlnMobility! for mobBCC (src/Processing/DislocationMobility.jl:
# 16-171, 186-341): node forces gathered from network.segForce and velocities
# from the BCC drag tensor, on the device.
struct Drag
    edge::Float64
    screw::Float64
    climb::Float64
    line::Float64
end
drag(d::DislocationParameters) = Drag(d.dragCoeffs[:edge], d.dragCoeffs[:screw],
                                      d.dragCoeffs[:climb], d.dragCoeffs[:line])

function mobility_call(dlnParams, matParams, network, idx, nodeForce, nodeVel, flags)
    k = Ref(consts(dlnParams, matParams))
    dr = Ref(drag(dlnParams))
    ix = idxvec(idx)
    cols = min(size(nodeForce, 2), size(nodeVel, 2))
    GC.@preserve network ix nodeForce nodeVel begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_mobility_bcc, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Drag}, Ref{Network}, Ptr{Float64}, Int64, Ptr{Int64},
             Int64, Ptr{Float64}, Ptr{Float64}, Int64, Cint),
            ctx(), k, dr, nw, pointer(network.segForce), size(network.segForce, 3),
            isnothing(idx) ? C_NULL : pointer(ix), length(ix), pointer(nodeForce),
            pointer(nodeVel), cols, flags))
    end
end

function dlnMobility(dlnParams::DislocationParameters{T1} where {T1 <: mobBCC}, matParams,
                     network::DislocationNetwork, idx = nothing)
    dlnParams.parGPU || return DDD.dlnMobility(dlnParams, matParams, network, idx)
    n = isnothing(idx) ? network.numNode[1] : length(idx)
    nodeForce = zeros(3, n)
    nodeVel = zeros(3, n)
    mobility_call(dlnParams, matParams, network, idx, nodeForce, nodeVel, DDD_OUT_BY_POSITION)
    return nodeForce, nodeVel
end

function dlnMobility!(dlnParams::DislocationParameters{T1} where {T1 <: mobBCC}, matParams,
                      network::DislocationNetwork, idx = nothing)
    dlnParams.parGPU || return DDD.dlnMobility!(dlnParams, matParams, network, idx)
    mobility_call(dlnParams, matParams, network, idx, network.nodeForce, network.nodeVel, Cint(0))
    return network
end

calc_σTilde(x0, dlnParams, matParams, network, idx = nothing) =
    calc_σTilde!(zeros(6, x0 isa Base.Iterators.ProductIterator ? prod(size(x0)) :
                         div(length(x0), 3)), x0, dlnParams, matParams, network, idx)

# ---- precision modes, device-resident evaluation, learned far field ------
# (no counterpart in the reference; include/ddd_hip.h documents each call)

const DDD_MODE_FP64 = Cint(0)
const DDD_MODE_MIXED = Cint(1)
const DDD_MODE_HYBRID = Cint(2)

struct Opts
    mode::Cint
    farDist::Float64     # 0: 5 core radii
    sinMin::Float64      # 0: 1e-3
    farfield::Ptr{Cvoid}
end

struct Stats
    pairs::UInt64
    pairsFP64::UInt64
    deviceBytes::UInt64
    chunks::UInt64
    tileMs::Float64
end

struct TrainOpts
    samples::Int64
    valSamples::Int64
    steps::Cint
    batch::Cint
    lr::Float32
    lrFinal::Float32
    seed::UInt64
    denFloor::Float32
    nearFrac::Float32
    nearShells::Cint
    plainOffsets::Cint
end

struct TrainReport
    lossFirst::Float64
    lossLast::Float64
    relErrVal::Float64
    relErrValNear::Float64
    seconds::Float64
    relErrShell::NTuple{8,Float64}
end

"""Whole-network seg-seg forces in a precision mode (3 x 2 x numSeg) and the
pair statistics (pairs evaluated exactly, of which in FP64)."""
function calcSegSegForceMode(dlnParams, matParams, network::DislocationNetwork, opts::Opts)
    k = Ref(consts(dlnParams, matParams))
    m = network.numSeg[1]
    out = zeros(3, 2, m)
    st = Ref(Stats(0, 0, 0, 0, 0.0))
    GC.@preserve network out begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_segseg_force_opts, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ref{Opts}, Ptr{Float64}, Int64, Cint, Ref{Stats}),
            ctx(), k, nw, Ref(opts), pointer(out), m, Cint(0), st))
    end
    return out, st[]
end

"""A far field on a G^3 grid of cells of edge h from corner lo, pairs of
cells within Chebyshev distance nearRadius exact."""
function farfield(G::Integer, lo::AbstractVector{Float64}, h::Float64, nearRadius::Integer)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    l = Vector{Float64}(lo)
    GC.@preserve l check(ccall((:ddd_farfield_create, libddd), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Float64}, Float64, Cint, Ref{Ptr{Cvoid}}),
        ctx(), Cint(G), pointer(l), h, Cint(nearRadius), out))
    return out[]
end

function farfield_mlp(seed::Integer = 11)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ddd_mlp_create, libddd), Cint, (Ptr{Cvoid}, UInt64, Ref{Ptr{Cvoid}}),
                ctx(), UInt64(seed), out))
    return out[]
end

"""Train the far-field network on exact local expansions of `trainNetwork`
and install its operator into `ff`."""
function train_farfield!(ff, mlp, dlnParams, matParams, trainNetwork::DislocationNetwork;
                         samples = 1 << 24, valSamples = 1 << 16, steps = 60000, batch = 4096,
                         lr = 2f-3, lrFinal = 5f-6, seed = 7, denFloor = 0.1f0)
    k = Ref(consts(dlnParams, matParams))
    o = Ref(TrainOpts(samples, valSamples, steps, batch, lr, lrFinal, seed, denFloor, 0f0, 0, 0))
    rep = Ref{TrainReport}()
    GC.@preserve trainNetwork begin
        nw = Ref(netstruct(trainNetwork))
        check(ccall((:ddd_farfield_train, libddd), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ref{TrainOpts},
             Ref{TrainReport}), ctx(), ff, mlp, k, nw, o, rep))
    end
    return rep[]
end

"""Keep `network` on the GPU for repeated evaluations (time stepping):
upload once, set a mode, then `resident_seg_force!` fills network.segForce."""
function resident_upload(dlnParams, matParams, network::DislocationNetwork)
    k = Ref(consts(dlnParams, matParams))
    GC.@preserve network begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_net_upload, libddd), Cint, (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ptr{Cvoid}),
                    ctx(), k, nw, C_NULL))
    end
    return nothing
end

resident_mode!(opts::Opts) =
    check(ccall((:ddd_net_set_opts, libddd), Cint, (Ptr{Cvoid}, Ref{Opts}), ctx(), Ref(opts)))

function resident_seg_force!(network::DislocationNetwork)
    check(ccall((:ddd_net_seg_force, libddd), Cint, (Ptr{Cvoid}, Cint), ctx(), Cint(0)))
    GC.@preserve network check(ccall((:ddd_net_download, libddd), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64), ctx(), pointer(network.segForce),
        size(network.segForce, 3)))
    return nothing
end

# calc_uTilde / calc_uTilde! (src/Processing/DislocationFEM.jl:355-707) on the
# device.  The reference's signature carries no DislocationParameters, so there
# is no parGPU switch: these methods always run on the GPU.
function u_tilde_points(mesh, matParams, network, points::Matrix{Float64})
    P = size(points, 2)
    u = zeros(3, P)
    label = Int32.(network.label)
    vertices = reduce(hcat, mesh.vertices.vertices)  # 3 x 8
    faceNorm = Matrix{Float64}(mesh.faceNorm)
    faceMidPt = Matrix{Float64}(mesh.faceMidPt)
    GC.@preserve network label vertices faceNorm faceMidPt points u begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_u_tilde, libddd), Cint,
            (Ptr{Cvoid}, Float64, Float64, Ref{Network}, Ptr{Int32}, Ptr{Float64}, Ptr{Float64},
             Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Cint),
            ctx(), matParams.omνInv8π, matParams.om2νomνInv8π, nw, pointer(label),
            pointer(vertices), pointer(faceNorm), pointer(faceMidPt), pointer(points), P,
            pointer(u), Cint(0)))
    end
    return u
end

calc_uTilde(mesh, boundaries, matParams, network::DislocationNetwork) =
    vec(u_tilde_points(mesh, matParams, network, Matrix{Float64}(mesh.coord[:, boundaries.uGammaDln])))

function calc_uTilde!(forceDisplacement, mesh, boundaries, matParams, network::DislocationNetwork)
    uTilde = forceDisplacement.uTilde
    uTilde .= 0
    v = calc_uTilde(mesh, boundaries, matParams, network)
    uDofs = boundaries.uDofsDln
    for k in eachindex(v)           # the reference's dof mapping, value k -> uDofs[k]
        uTilde[uDofs[k]] += v[k]
    end
    return nothing
end

# detectCollision (src/Processing/DislocationTopologyCollision.jl:1-222) on the
# device; same arguments and the same 10-tuple result.
struct Collision
    collision::Int32
    type::Int32
    n1s1::Int64
    n2s1::Int64
    n1s2::Int64
    n2s2::Int64
    s1::Int64
    s2::Int64
    L1::Float64
    L2::Float64
end

function detectCollision(dlnParams::DislocationParameters, network::DislocationNetwork,
                         skipSegs = Vector{Tuple{Int, Int}}())
    dlnParams.parGPU || return DDD.detectCollision(dlnParams, network, skipSegs)
    label = Int32.(network.label)
    skip = Int64[x for p in skipSegs for x in p]          # 2 x nSkip
    res = Ref(Collision(0, 0, 0, 0, 0, 0, 0, 0, 0.0, 0.0))
    GC.@preserve network label skip begin
        nw = Ref(netstruct(network))
        check(ccall((:ddd_detect_collision, libddd), Cint,
            (Ptr{Cvoid}, Float64, Ref{Network}, Ptr{Int32}, Ptr{Float64}, Ptr{Int64}, Int64,
             Ref{Collision}), ctx(), dlnParams.collisionDistSq, nw, pointer(label),
            pointer(network.nodeVel), isempty(skip) ? C_NULL : pointer(skip),
            length(skipSegs), res))
    end
    r = res[]
    type = (:null, :hinge, :twoLine)[r.type + 1]
    return r.collision == 1, type, r.n1s1, r.n2s1, r.n1s2, r.n2s2, r.s1, r.s2, r.L1, r.L2
end

# deriv! / integrate! (src/Processing/IntegrationProcessing.jl:13-168) with the
# network on the device: coordinates, velocities and forces stay in HBM across
# the adaptive step's trial evaluations; results are copied back at the end.
struct Integ
    tmin::Float64
    tmax::Float64
    dtmin::Float64
    dtmax::Float64
    abstol::Float64
    reltol::Float64
    maxchange::Float64
    exponent::Float64
    maxiter::Int64
end
struct ITime
    dt::Float64
    time::Float64
    step::Int64
end

function dynamics_upload(dlnParams, matParams, mesh, forceDisplacement, network)
    k = Ref(consts(dlnParams, matParams))
    dr = Ref(drag(dlnParams))
    label = Int32.(network.label)  # nodeTypeDln is an Int32 enum
    uHat = Vector{Float64}(forceDisplacement.uHat)
    C = Matrix(mesh.C)
    GC.@preserve network label uHat C mesh begin
        nw = Ref(netstruct(network))
        ms = Ref(meshstruct(mesh, C, uHat))
        check(ccall((:ddd_net_upload, libddd), Cint,
            (Ptr{Cvoid}, Ref{Consts}, Ref{Network}, Ref{Mesh}), ctx(), k, nw, ms))
        check(ccall((:ddd_net_upload_dynamics, libddd), Cint,
            (Ptr{Cvoid}, Ref{Network}, Ptr{Int32}, Ref{Drag}), ctx(), nw, pointer(label), dr))
    end
end

function dynamics_download!(network)
    n = network.numNode[1]
    GC.@preserve network begin
        check(ccall((:ddd_net_download_nodes, libddd), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64), ctx(),
            pointer(network.coord), pointer(network.nodeVel), pointer(network.nodeForce), n))
        check(ccall((:ddd_net_download, libddd), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), ctx(),
            pointer(network.segForce), size(network.segForce, 3)))
    end
end

function deriv!(dlnParams::DislocationParameters{T1} where {T1 <: mobBCC}, matParams, mesh,
                forceDisplacement, network::DislocationNetwork)
    dlnParams.parGPU || return DDD.deriv!(dlnParams, matParams, mesh, forceDisplacement, network)
    dynamics_upload(dlnParams, matParams, mesh, forceDisplacement, network)
    check(ccall((:ddd_net_deriv, libddd), Cint, (Ptr{Cvoid}, Cint), ctx(), Cint(1)))
    dynamics_download!(network)
    return nothing
end

function integrate!(intParams::IntegrationParameters{T1} where {T1 <: AdaptiveEulerTrapezoid},
                    intVars::IntegrationTime,
                    dlnParams::DislocationParameters{T2} where {T2 <: mobBCC}, matParams, mesh,
                    forceDisplacement, network::DislocationNetwork)
    dlnParams.parGPU || return DDD.integrate!(intParams, intVars, dlnParams, matParams, mesh,
                                              forceDisplacement, network)
    dynamics_upload(dlnParams, matParams, mesh, forceDisplacement, network)
    ip = Ref(Integ(intParams.tmin, intParams.tmax, intParams.dtmin, intParams.dtmax,
                   intParams.abstol, intParams.reltol, intParams.maxchange, intParams.exponent,
                   intParams.maxiter))
    t = Ref(ITime(intVars.dt, intVars.time, intVars.step))
    check(ccall((:ddd_net_integrate, libddd), Cint, (Ptr{Cvoid}, Ref{Integ}, Ref{ITime}, Cint),
                ctx(), ip, t, Cint(1)))
    dynamics_download!(network)
    return IntegrationTime(t[].dt, t[].time, t[].step)
end

end # module
