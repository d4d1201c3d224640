"""
    TopOptHIP

Julia side of `libtopopt_hip`: a linear-solver plugin for TopOpt.jl v0.14's
`GenericFEASolver` whose displacement solve, compliance sensitivity, density
filter and (optionally) the whole SIMP loop run on an MI355X GPU.

    using TopOpt, TopOptHIP
    problem = PointLoadCantilever((120, 40, 40), (1.0, 1.0, 1.0), 1.0, 0.3, 1.0)
    solver  = FEASolver(HIPMatrixFreeSolver, problem; xmin=1e-9,
                        penalty=TopOpt.PowerPenaltyFun(3.0))
    comp    = TopOpt.Compliance(solver)          # unchanged TopOpt API
    filter  = DensityFilterFun(solver; rmin=1.5) # J applied on the device
    volfrac = VolumeFun(solver)                   # device reduction
    obj = x -> comp(filter(PseudoDensities(x)))   # Zygote: the rrules run on the device

Plugin points used (TopOpt.jl source):
  * `AbstractLinearSolver`                      src/FEA/solvers_api.jl:44
  * `solve_system!(::Type{S}, solver, K, f, lhs)` src/FEA/solvers_api.jl:131-282
  * the solver call operator                    src/FEA/solvers_api.jl:285-374
    (specialised here so the host `assemble!` of K is skipped)
  * `ComplianceFun` call                         src/Functions/compliance.jl:58-76
    (energies on the displacement the solve left on the device)
  * `DensityFilterFun(solver, rmin)` constructor  src/CheqFilters/density_filter.jl:25-33
    (its J is a device operator, so the call :35-40 and rrule :41-47 run on the GPU)
  * `VolumeFun` call                             src/Functions/volume.jl:59-62
  * `FEASolver(S, problem)` fallback             src/FEA/solvers_api.jl:599-603

Every `ccall` passes plain pointers and sizes; host arrays are kept alive
with `GC.@preserve`.  Return codes: 0 ok, 1 = PCG hit maxiter (a warning,
the reference's CG is silent there), < 0 = error with a message.
"""
module TopOptHIP

using LinearAlgebra
using TopOpt
using TopOpt: PseudoDensities, PENALTY_BEFORE_INTERPOLATION
using TopOpt.FEA: AbstractLinearSolver, GenericFEASolver, LinearElasticity, HeatTransfer
using TopOpt.TopOptProblems: StiffnessTopOptProblem, getcloaddict
import TopOpt.FEA: solve_system!
import TopOpt.Functions: ComplianceFun

export HIPMatrixFreeSolver, HIPCGAssemblySolver, hip_filter, hip_projected_filter, hip_simp!,
       hip_beso, hip_heat_context, MmaOpts

const libtopopt_hip = get(ENV, "TOPOPT_HIP_LIB", "libtopopt_hip.so")

const TOHIP_FP64 = Cint(0)
const TOHIP_MIXED = Cint(1)
const TOHIP_MIXED_MG = Cint(2)
const TOHIP_FP64_MG = Cint(3)

# precision keyword -> tohip precision code (:fp64, :mixed, :mixed_mg, :fp64_mg)
function _precision(p::Symbol)
    p === :fp64 && return TOHIP_FP64
    p === :mixed && return TOHIP_MIXED
    p === :mixed_mg && return TOHIP_MIXED_MG
    p === :fp64_mg && return TOHIP_FP64_MG
    throw(ArgumentError("precision must be :fp64, :mixed, :mixed_mg or :fp64_mg, got :$p"))
end

"""Linear-solver tag: Jacobi PCG on the GPU, matrix-free (hex8 / quad4)."""
struct HIPMatrixFreeSolver <: AbstractLinearSolver end

"""Linear-solver tag: CGAssemblySolver on the GPU — K(vars) assembled in CSR
with apply!'s boundary rule, then CG (identity, or Jacobi with
`preconditioner=:jacobi` at the solve)."""
struct HIPCGAssemblySolver <: AbstractLinearSolver end

mutable struct HIPContext
    ptr::Ptr{Cvoid}
    pen::Tuple{Cint,Float64,Cint,Float64}  # penalty (kind, p, proj, beta) last sent
    f::Vector{Float64}  # point-load vector, Ferrite DOF order
    rmin::Float64       # radius of the device filter (0: none yet)
end

const CONTEXTS = WeakKeyDict{Any,HIPContext}()

function _err(ptr)
    msg = ccall((:tohip_last_error, libtopopt_hip), Cstring, (Ptr{Cvoid},), ptr)
    return msg == C_NULL ? "" : unsafe_string(msg)
end

function _check(rc::Integer, ptr, what)
    # TOHIP_ERR_DOMAIN: EnergyCriteria's DomainError (src/FEA/convergence_criteria.jl:35-41)
    rc == -4 && throw(DomainError(rc, "$what: $(_err(ptr))"))
    rc < 0 && error("$what failed ($rc): $(_err(ptr))")
    # TOHIP_BREAKDOWN: u is NaN, as the reference's direct solver leaves it
    # with a warning for a matrix that is not positive definite (solvers_api.jl:152-158)
    rc == 2 && @warn "$what: $(_err(ptr))"
    rc == 1 && @warn "$what: PCG stopped at maxiter before reaching its tolerance"
    return rc
end

# make_cload (src/TopOptProblems/matrices_and_vectors.jl:346-365) as a dense vector
function _point_load(problem, ndofs)
    dim = TopOpt.TopOptProblems.getdim(problem)
    node_dofs = problem.metadata.node_dofs
    f = zeros(Float64, ndofs)
    for (node, load) in getcloaddict(problem), (d, v) in enumerate(load)
        v != 0 && (f[node_dofs[(node - 1) * dim + d]] += v)
    end
    f[problem.ch.prescribed_dofs] .= 0
    return f
end

# Penalty families (src/Utilities/penalties.jl:30-101) as tohip_set_penalty_fun
# codes: TOHIP_PEN_POWER / RATIONAL / SINH and TOHIP_PROJ_NONE / HEAVISIDE / SIGMOID.
_pen_kind(::TopOpt.PowerPenaltyFun) = Cint(0)
_pen_kind(::TopOpt.RationalPenaltyFun) = Cint(1)
_pen_kind(::TopOpt.SinhPenaltyFun) = Cint(2)
_proj_code(::TopOpt.HeavisideProjectionFun) = Cint(1)
_proj_code(::TopOpt.SigmoidProjectionFun) = Cint(2)
_penalty_spec(P) = (_pen_kind(P), Float64(P.p), Cint(0), 0.0)
function _penalty_spec(P::TopOpt.ProjectedPenaltyFun)
    return (_pen_kind(P.penalty), Float64(P.penalty.p), _proj_code(P.proj), Float64(P.proj.β))
end

function _set_penalty!(c::HIPContext, P)
    spec = _penalty_spec(P)
    _check(ccall((:tohip_set_penalty_fun, libtopopt_hip), Cint,
                 (Ptr{Cvoid}, Cint, Cdouble, Cint, Cdouble), c.ptr, spec...),
           c.ptr, "tohip_set_penalty_fun")
    c.pen = spec
    return c
end

"""Device context of a solver (created on first use, freed with the solver)."""
function hip_context(s::GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver}) where {T}
    return _hip_context_impl(s)
end

function _hip_context_impl(s)
    return get!(CONTEXTS, s) do
        problem = s.problem
        grid = problem.rect_grid
        dim = length(grid.nels)
        nels = Int64[grid.nels...]
        h = Float64[grid.sizes...]
        node_dofs = Vector{Int64}(vec(problem.metadata.node_dofs))  # dim x nnodes, 1-based
        ndofs = length(s.u)
        fixed = zeros(UInt8, ndofs)
        fixed[problem.ch.prescribed_dofs] .= 0x01
        ref = Ref{Ptr{Cvoid}}(C_NULL)
        rc = GC.@preserve nels h node_dofs fixed ccall(
            (:tohip_create, libtopopt_hip), Cint,
            (Ptr{Ptr{Cvoid}}, Cint, Ptr{Int64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble,
             Cdouble, Cint, Ptr{Int64}, Ptr{UInt8}),
            ref, dim, nels, h, Float64(problem.E), Float64(problem.ν), Float64(s.xmin),
            Float64(s.penalty.p), Cint(PENALTY_BEFORE_INTERPOLATION), node_dofs, fixed)
        _check(rc, C_NULL, "tohip_create")
        c = HIPContext(ref[], (Cint(0), Float64(s.penalty.p), Cint(0), 0.0),
                       _point_load(problem, ndofs), 0.0)
        finalizer(c) do c
            c.ptr == C_NULL || ccall((:tohip_destroy, libtopopt_hip), Cvoid, (Ptr{Cvoid},), c.ptr)
            c.ptr = C_NULL
        end
        c
    end
end

# CG stopping rule, FEASolver's `conv` keyword (src/FEA/solvers_api.jl:482):
# uploaded before every device solve; EnergyCriteria.energy, which every
# check updates (src/FEA/convergence_criteria.jl:43), is read back after it.
function _sync_conv!(c::HIPContext, conv)
    energy = conv isa TopOpt.EnergyCriteria
    _check(ccall((:tohip_set_convergence, libtopopt_hip), Cint, (Ptr{Cvoid}, Cint, Cdouble),
                 c.ptr, Cint(energy ? 1 : 0), energy ? Float64(conv.energy) : 0.0),
           c.ptr, "tohip_set_convergence")
    return c
end

function _pull_conv!(c::HIPContext, conv)
    if conv isa TopOpt.EnergyCriteria
        e = Ref{Float64}(0)
        _check(ccall((:tohip_get_convergence_energy, libtopopt_hip), Cint,
                     (Ptr{Cvoid}, Ptr{Float64}), c.ptr, e), c.ptr, "tohip_get_convergence_energy")
        conv.energy = e[]
    end
    return conv
end

function _sync!(c::HIPContext, s)
    # setpenalty! (src/FEA/solvers_api.jl:419-429) may have changed p or the family
    _penalty_spec(s.penalty) != c.pen && _set_penalty!(c, s.penalty)
    _sync_conv!(c, s.conv)
    vars = Vector{Float64}(s.vars)
    GC.@preserve vars _check(
        ccall((:tohip_set_densities, libtopopt_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}),
              c.ptr, vars), c.ptr, "tohip_set_densities")
    return c
end

"""
    solve_system!(HIPMatrixFreeSolver, s, K, rhs, lhs; initially_zero=true,
                  reltol=sqrt(eps()), precision=:fp64)

Jacobi PCG on the GPU, stopping at ||r|| <= max(reltol ||r0||, s.abstol) or
after s.cg_max_iter iterations (the reference's `cg!` rule).  `K` is unused:
the operator is applied matrix-free from the densities in `s.vars`.
`precision=:mixed` runs FP32 PCG with FP64 residual refreshes; `:mixed_mg` the same
with a geometric multigrid V-cycle as the FP32 preconditioner (even cell counts).
"""
function solve_system!(::Type{HIPMatrixFreeSolver},
                       s::GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver}, K, rhs, lhs;
                       initially_zero=true, reltol=sqrt(eps(Float64)), precision=:fp64,
                       kwargs...) where {T}
    c = _sync!(hip_context(s), s)
    f = Vector{Float64}(rhs)
    u = Vector{Float64}(lhs)
    iters = Ref{Int64}(0)
    resid = Ref{Float64}(0)
    rc = GC.@preserve f u ccall(
        (:tohip_solve, libtopopt_hip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Int64, Cint, Cint,
         Ptr{Int64}, Ptr{Float64}),
        c.ptr, f, u, Float64(s.abstol), Float64(reltol), Int64(s.cg_max_iter),
        Cint(initially_zero), _precision(precision), iters, resid)
    _pull_conv!(c, s.conv)
    _check(rc, c.ptr, "tohip_solve")
    lhs .= u
    return nothing
end

# The generic call operator runs the host assemble! of K before every solve
# (src/FEA/solvers_api.jl:353-361); this one goes straight to the device.
function (s::GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver})(
    reuse_fact::Bool=false, (::Type{Val{safe}})=Val{false};
    assemble_f=true, rhs=nothing, lhs=assemble_f ? s.u : s.lhs, kwargs...,
) where {T,safe}
    c = hip_context(s)
    if rhs === nothing
        rhs = assemble_f ? c.f : s.rhs
    else  # apply_zero! on the prescribed DOFs (solvers_api.jl:363-367)
        rhs = copy(rhs)
        rhs[s.problem.ch.prescribed_dofs] .= 0
    end
    solve_system!(HIPMatrixFreeSolver, s, nothing, rhs, lhs; kwargs...)
    return nothing
end

# ComplianceFun on the device: solve, then obj / cell_comp / grad in one call
# (compute_element_energy, src/Functions/compute_element_energy.jl:18-38) on
# the displacement the solve left on the device (u = C_NULL: no re-upload;
# solver.u still receives its copy, as the reference's solver() fills it).
# The reference rrule (compliance.jl:72-76) reads o.grad and is unchanged.
function (o::ComplianceFun{T,<:GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver}})(
    x::PseudoDensities) where {T}
    s = o.solver
    s.vars .= x.x
    s()
    c = hip_context(s)
    obj = Ref{Float64}(0)
    cc = Vector{Float64}(undef, length(o.cell_comp))
    g = Vector{Float64}(undef, length(o.grad))
    GC.@preserve cc g _check(
        ccall((:tohip_compliance, libtopopt_hip), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
              c.ptr, C_NULL, obj, cc, g), c.ptr, "tohip_compliance")
    o.cell_comp .= cc
    o.grad .= g
    return T(obj[])
end

# ---- DensityFilterFun with a device Jacobian --------------------------------
# DensityFilterFun stores J and applies it as mul!(out, jacobian, x.x) in its
# call and jacobian' * Δ in its rrule (src/CheqFilters/density_filter.jl
# :35-47).  For a GPU solver the constructor below builds the same struct
# with a Jacobian that is an operator on the device: both reference methods
# then run the stencil kernels unchanged, and the host neighbour search and
# sparse J (getJacobian, :49-108) are never built.
struct HIPFilterJacobian <: AbstractMatrix{Float64}
    ctx::HIPContext
    rmin::Float64
    n::Int
    transposed::Bool
end
Base.size(J::HIPFilterJacobian) = (J.n, J.n)
Base.getindex(::HIPFilterJacobian, ::Int, ::Int) =
    error("HIPFilterJacobian: the filter is a device operator; its entries are not stored")
Base.adjoint(J::HIPFilterJacobian) = HIPFilterJacobian(J.ctx, J.rmin, J.n, !J.transposed)
Base.transpose(J::HIPFilterJacobian) = adjoint(J)

function _filter_setup!(c::HIPContext, rmin)
    if c.rmin != rmin
        _check(ccall((:tohip_filter_setup, libtopopt_hip), Cint, (Ptr{Cvoid}, Cdouble),
                     c.ptr, Float64(rmin)), c.ptr, "tohip_filter_setup")
        c.rmin = rmin
    end
    return c
end

function LinearAlgebra.mul!(y::AbstractVector, J::HIPFilterJacobian, x::AbstractVector)
    c = _filter_setup!(J.ctx, J.rmin)
    xin = Vector{Float64}(x)
    out = Vector{Float64}(undef, J.n)
    GC.@preserve xin out _check(
        ccall((:tohip_filter_apply, libtopopt_hip), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint), c.ptr, xin, out,
              Cint(J.transposed)), c.ptr, "tohip_filter_apply")
    y .= out
    return y
end
Base.:*(J::HIPFilterJacobian, x::AbstractVector) = mul!(Vector{Float64}(undef, J.n), J, x)

function TopOpt.CheqFilters.DensityFilterFun(
    solver::GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver}, rmin::T,
    (::Type{TI})=Int) where {T,TI<:Integer}
    c = _filter_setup!(hip_context(solver), rmin)
    J = HIPFilterJacobian(c, Float64(rmin), length(solver.vars), false)
    return TopOpt.CheqFilters.DensityFilterFun(TopOpt.CheqFilters.FilterMetadata(T, TI), rmin, J)
end

# VolumeFun call (src/Functions/volume.jl:59-62) as a device reduction; the
# rrule (:64-67) returns the stored constant gradient and is unchanged.
function (v::TopOpt.Functions.VolumeFun{T,<:GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver}})(
    x::PseudoDensities) where {T}
    c = hip_context(v.solver)
    xin = Vector{Float64}(x.x)
    vol = Ref{Float64}(0)
    GC.@preserve xin _check(
        ccall((:tohip_volume, libtopopt_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
              c.ptr, xin, vol), c.ptr, "tohip_volume")
    # tohip_volume is the fraction sum_e x_e V_e / V_total (equal cells)
    return T(v.fraction ? vol[] : vol[] * v.total_volume)
end

"""
    hip_filter(s, rmin, x; transpose=false) -> J*x (or J'*x)

The DensityFilterFun product (src/CheqFilters/density_filter.jl:35-47) as
a stencil on the GPU; identical J (closed form of the neighbour search,
src/CheqFilters/CheqFilters.jl:66-118).
"""
function hip_filter(s::GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver}, rmin, x;
                    transpose=false) where {T}
    c = hip_context(s)
    if c.rmin != rmin
        _check(ccall((:tohip_filter_setup, libtopopt_hip), Cint, (Ptr{Cvoid}, Cdouble),
                     c.ptr, Float64(rmin)), c.ptr, "tohip_filter_setup")
        c.rmin = rmin
    end
    xin = Vector{Float64}(x)
    y = similar(xin)
    GC.@preserve xin y _check(
        ccall((:tohip_filter_apply, libtopopt_hip), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint), c.ptr, xin, y, Cint(transpose)),
        c.ptr, "tohip_filter_apply")
    return y
end

_proj_spec(::Nothing) = (Cint(0), 0.0)
_proj_spec(P) = (_proj_code(P), Float64(P.β))

"""
    hip_projected_filter(s, f::TopOpt.ProjectedDensityFilterFun, x; cotangent=nothing)

`f(x)` of a ProjectedDensityFilterFun (src/CheqFilters/density_filter.jl:130-169) on
the GPU, or with `cotangent = Δ` its pullback at x.  Pre/post projections must be
`nothing`, `HeavisideProjectionFun` or `SigmoidProjectionFun`.
"""
function hip_projected_filter(s::GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver}, f, x;
                              cotangent=nothing) where {T}
    c = hip_context(s)
    rmin = f.filter.rmin
    _check(ccall((:tohip_filter_setup, libtopopt_hip), Cint, (Ptr{Cvoid}, Cdouble), c.ptr,
                 Float64(rmin)), c.ptr, "tohip_filter_setup")
    c.rmin = rmin
    pre, post = _proj_spec(f.preproj), _proj_spec(f.postproj)
    xin = Vector{Float64}(x)
    y = similar(xin)
    if cotangent === nothing
        GC.@preserve xin y _check(
            ccall((:tohip_filter_apply_projected, libtopopt_hip), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Cdouble, Cint, Cdouble),
                  c.ptr, xin, y, pre..., post...), c.ptr, "tohip_filter_apply_projected")
    else
        d = Vector{Float64}(cotangent)
        GC.@preserve xin d y _check(
            ccall((:tohip_filter_pullback_projected, libtopopt_hip), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Cdouble, Cint,
                   Cdouble), c.ptr, xin, d, y, pre..., post...),
            c.ptr, "tohip_filter_pullback_projected")
    end
    return y
end

# struct tohip_beso_opts / tohip_beso_result (include/topopt_hip.h)
struct BesoOpts
    maxiter::Int64
    vol_limit::Cdouble
    rmin::Cdouble
    er::Cdouble
    tol::Cdouble
    sens_tol::Cdouble
    k::Cint
    penalty_p::Cdouble
    rel_tol::Cdouble
    pcg_maxiter::Int64
    precision::Cint
end
mutable struct BesoResultC
    objval::Cdouble
    change::Cdouble
    converged::Cint
    fevals::Int64
    volume::Cdouble
end

"""
    hip_beso(b::TopOpt.BESO, x0; black=BitVector(), white=BitVector(), precision=:fp64)

`b(x0; black, white)` (src/Algorithms/beso.jl:93-188) with the whole loop on the GPU;
returns a `BESOResult` and leaves the topology in `b.topology`.
"""
function hip_beso(b, x0::AbstractVector; black=BitVector(), white=BitVector(), precision=:fp64,
                  rel_tol=1e-8)
    s = b.comp.solver
    c = _sync!(hip_context(s), s)
    GC.@preserve c _check(ccall((:tohip_set_load, libtopopt_hip), Cint,
                                (Ptr{Cvoid}, Ptr{Float64}), c.ptr, c.f), c.ptr, "tohip_set_load")
    n = length(x0)
    passive = zeros(UInt8, n)
    isempty(black) || (passive[black] .= 0x01)
    isempty(white) || (passive[white] .= 0x02)
    opts = Ref(BesoOpts(b.maxiter, b.vol_limit, b.filter.rmin, b.er, b.tol, b.sens_tol,
                        length(b.obj_trace), b.penalty.p, rel_tol, s.cg_max_iter,
                        _precision(precision)))
    x = Vector{Float64}(x0)
    topo = zeros(Float64, n)
    res = Ref(BesoResultC(0, 0, 0, 0, 0))
    rc = GC.@preserve x passive topo ccall(
        (:tohip_beso_run, libtopopt_hip), Cint,
        (Ptr{Cvoid}, Ref{BesoOpts}, Ptr{Float64}, Ptr{UInt8}, Ptr{Float64}, Ref{BesoResultC},
         Ptr{Float64}),
        c.ptr, opts, x, passive, topo, res, C_NULL)
    _pull_conv!(c, s.conv)
    _check(rc, c.ptr, "tohip_beso_run")
    b.topology .= topo
    r = res[]
    return TopOpt.Algorithms.BESOResult(topo, r.objval, r.change, r.converged != 0, Int(r.fevals))
end

# struct tohip_simp_opts (include/topopt_hip.h)
struct SimpOpts
    iterations::Int64
    volfrac::Cdouble
    rmin::Cdouble
    move::Cdouble
    oc_tol::Cdouble
    rel_tol::Cdouble
    maxiter::Int64
    precision::Cint
    warm_start::Cint
    inner_rel_tol::Cdouble
end

# struct tohip_mma_opts (include/topopt_hip.h)
struct MmaOpts
    s_init::Cdouble
    s_incr::Cdouble
    s_decr::Cdouble
end

"""
    hip_simp!(s, x; volfrac, rmin, iterations, move=0.2, rel_tol=1e-8,
              precision=:fp64, warm_start=true, optimizer=:oc) -> (obj, pcg_iters, change)

`iterations` SIMP iterations (filter, solve, compliance, pullback, design
update) with every vector resident on the GPU; x is the design, updated in
place.  `optimizer=:oc` is the OC update; `:mma` (or an `MmaOpts`) is MMA87
on min c(x) s.t. V(J x) <= volfrac, the optimizer of the reference tutorials
(docs/tutorials/simp.qmd:103-114) run on the device.
"""
function hip_simp!(s::GenericFEASolver{T,LinearElasticity,HIPMatrixFreeSolver},
                   x::Vector{Float64}; volfrac, rmin, iterations, move=0.2, oc_tol=1e-12,
                   rel_tol=1e-8, maxiter=200_000, precision=:fp64, warm_start=true,
                   inner_rel_tol=(precision === :mixed_mg ? 1e-3 : 0.1),
                   optimizer=:oc) where {T}
    c = hip_context(s)
    _set_penalty!(c, s.penalty)
    _sync_conv!(c, s.conv)
    GC.@preserve c _check(ccall((:tohip_set_load, libtopopt_hip), Cint,
                                (Ptr{Cvoid}, Ptr{Float64}), c.ptr, c.f), c.ptr, "tohip_set_load")
    opts = Ref(SimpOpts(iterations, volfrac, rmin, move, oc_tol, rel_tol, maxiter,
                        _precision(precision),
                        warm_start === :learned ? Cint(2) : Cint(warm_start ? 1 : 0),
                        inner_rel_tol))
    obj = zeros(Float64, iterations)
    its = zeros(Int64, iterations)
    chg = zeros(Float64, iterations)
    mma = optimizer === :mma ? MmaOpts(0.5, 1.2, 0.7) :
          optimizer isa MmaOpts ? optimizer : nothing
    (mma === nothing && optimizer !== :oc) &&
        throw(ArgumentError("optimizer must be :oc, :mma or MmaOpts"))
    rc = if mma === nothing
        GC.@preserve x obj its chg ccall(
            (:tohip_simp_run, libtopopt_hip), Cint,
            (Ptr{Cvoid}, Ref{SimpOpts}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64}),
            c.ptr, opts, x, obj, its, chg)
    else
        GC.@preserve x obj its chg ccall(
            (:tohip_simp_run_mma, libtopopt_hip), Cint,
            (Ptr{Cvoid}, Ref{SimpOpts}, Ref{MmaOpts}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64},
             Ptr{Float64}),
            c.ptr, opts, Ref(mma), x, obj, its, chg)
    end
    _pull_conv!(c, s.conv)
    _check(rc, c.ptr, "tohip_simp_run")
    return obj, its, chg
end

# ------------------------------------------------------------------ heat --
# HeatTransfer physics (HeatConductionProblem, linear cells) on the device
# through tohip_heat_*.  The forward solve applies the prescribed temperatures
# as apply! does (T_p prescribed, K_ff T_f = Q_f - K_fp T_p), so nonzero
# Dirichlet values are supported (the reference's CGMatrixFreeSolver rejects
# them, src/FEA/solvers_api.jl:514-523).

mutable struct HIPHeatContext
    ptr::Ptr{Cvoid}
    pen::Tuple{Cint,Float64,Cint,Float64}
end

const HEAT_CONTEXTS = WeakKeyDict{Any,HIPHeatContext}()

function _heat_check(rc::Integer, ptr, what)
    if rc < 0
        msg = ccall((:tohip_heat_last_error, libtopopt_hip), Cstring, (Ptr{Cvoid},), ptr)
        error("$what failed ($rc): $(msg == C_NULL ? "" : unsafe_string(msg))")
    end
    # TOHIP_BREAKDOWN: u is NaN, as the reference's direct solver leaves it
    # with a warning for a matrix that is not positive definite (solvers_api.jl:152-158)
    rc == 2 && @warn "$what: $(_err(ptr))"
    rc == 1 && @warn "$what: PCG stopped at maxiter before reaching its tolerance"
    return rc
end

"""Device context of a heat-conduction solver (freed with the solver)."""
function hip_heat_context(s::GenericFEASolver{T,HeatTransfer,HIPMatrixFreeSolver}) where {T}
    return get!(HEAT_CONTEXTS, s) do
        problem = s.problem
        grid = problem.rect_grid
        dim = length(grid.nels)
        nels = Int64[grid.nels...]
        h = Float64[grid.sizes...]
        node_dofs = Vector{Int64}(vec(problem.metadata.node_dofs))  # 1 x nnodes, 1-based
        n = length(s.u)
        ch = problem.ch
        fixed = zeros(UInt8, n)
        fixed[ch.prescribed_dofs] .= 0x01
        vals = zeros(Float64, n)
        vals[ch.prescribed_dofs] .= ch.inhomogeneities
        ref = Ref{Ptr{Cvoid}}(C_NULL)
        rc = GC.@preserve nels h node_dofs fixed vals ccall(
            (:tohip_heat_create, libtopopt_hip), Cint,
            (Ptr{Ptr{Cvoid}}, Cint, Ptr{Int64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble, Cint,
             Ptr{Int64}, Ptr{UInt8}, Ptr{Float64}),
            ref, dim, nels, h, Float64(problem.k), Float64(s.xmin), Float64(s.penalty.p),
            Cint(PENALTY_BEFORE_INTERPOLATION), node_dofs, fixed, vals)
        _heat_check(rc, C_NULL, "tohip_heat_create")
        c = HIPHeatContext(ref[], (Cint(0), Float64(s.penalty.p), Cint(0), 0.0))
        finalizer(c) do c
            c.ptr == C_NULL ||
                ccall((:tohip_heat_destroy, libtopopt_hip), Cvoid, (Ptr{Cvoid},), c.ptr)
            c.ptr = C_NULL
        end
        c
    end
end

function _heat_sync!(c::HIPHeatContext, s)
    spec = _penalty_spec(s.penalty)
    if spec != c.pen
        _heat_check(ccall((:tohip_heat_set_penalty_fun, libtopopt_hip), Cint,
                          (Ptr{Cvoid}, Cint, Cdouble, Cint, Cdouble), c.ptr, spec...),
                    c.ptr, "tohip_heat_set_penalty_fun")
        c.pen = spec
    end
    vars = Vector{Float64}(s.vars)
    GC.@preserve vars _heat_check(
        ccall((:tohip_heat_set_densities, libtopopt_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}),
              c.ptr, vars), c.ptr, "tohip_heat_set_densities")
    return c
end

"""
    solve_system!(HIPMatrixFreeSolver, s::GenericFEASolver{T,HeatTransfer}, K, rhs, lhs;
                  initially_zero=true, reltol=sqrt(eps()), apply_dirichlet=false)

Jacobi PCG for the temperature on the GPU.  `apply_dirichlet=true` takes
`rhs` = the unconstrained heat load Q and prescribes the Dirichlet values;
`false` solves with homogeneous conditions (the adjoint / pullback solves).
"""
function solve_system!(::Type{HIPMatrixFreeSolver},
                       s::GenericFEASolver{T,HeatTransfer,HIPMatrixFreeSolver}, K, rhs, lhs;
                       initially_zero=true, reltol=sqrt(eps(Float64)), apply_dirichlet=false,
                       kwargs...) where {T}
    c = _heat_sync!(hip_heat_context(s), s)
    f = Vector{Float64}(rhs)
    u = Vector{Float64}(lhs)
    iters = Ref{Int64}(0)
    resid = Ref{Float64}(0)
    rc = GC.@preserve f u ccall(
        (:tohip_heat_solve, libtopopt_hip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Int64, Cint, Cint,
         Ptr{Int64}, Ptr{Float64}),
        c.ptr, f, u, Float64(s.abstol), Float64(reltol), Int64(s.cg_max_iter),
        Cint(initially_zero), Cint(apply_dirichlet), iters, resid)
    _heat_check(rc, c.ptr, "tohip_heat_solve")
    lhs .= u
    return nothing
end

# Call operator without the host assemble!: the forward solve uses the
# non-penalized load Q (elementinfo.fixedload) with the prescribed values; an
# explicit rhs (assemble_f=false, the pullback solves) is zeroed on the
# prescribed DOFs (apply_zero!, solvers_api.jl:363-367).
function (s::GenericFEASolver{T,HeatTransfer,HIPMatrixFreeSolver})(
    reuse_fact::Bool=false, (::Type{Val{safe}})=Val{false};
    assemble_f=true, rhs=nothing, lhs=assemble_f ? s.u : s.lhs, kwargs...,
) where {T,safe}
    if rhs === nothing && assemble_f
        solve_system!(HIPMatrixFreeSolver, s, nothing, s.elementinfo.fixedload, lhs;
                      apply_dirichlet=true, kwargs...)
    else
        r = Vector{Float64}(rhs === nothing ? s.rhs : rhs)
        r[s.problem.ch.prescribed_dofs] .= 0
        solve_system!(HIPMatrixFreeSolver, s, nothing, r, lhs; apply_dirichlet=false, kwargs...)
    end
    return nothing
end

# solve_adjoint! (src/Functions/thermal_compliance.jl) for this solver.
function TopOpt.Functions.solve_adjoint!(s::GenericFEASolver{T,HeatTransfer,HIPMatrixFreeSolver},
                                         lhs, rhs) where {T}
    s(; assemble_f=false, rhs=rhs, lhs=lhs)
    return nothing
end

# ThermalComplianceFun with the element loop on the device: J = Q'T, the
# adjoint K_cond lam = -Q_cond, cell_e = lam_e' Ke T_e, grad_e = drho_e cell_e
# (compute_thermal_compliance!, src/Functions/thermal_compliance.jl).  The
# reference rrule reads tc.grad and is unchanged.
function (tc::TopOpt.Functions.ThermalComplianceFun{
    T,<:GenericFEASolver{T,HeatTransfer,HIPMatrixFreeSolver}})(x::PseudoDensities) where {T}
    s = tc.solver
    s.vars .= x.x
    s()
    Q = Vector{Float64}(s.elementinfo.fixedload)
    obj = dot(Q, s.u)
    lam = zeros(Float64, length(s.u))
    s(; assemble_f=false, rhs=-Q, lhs=lam)
    c = hip_heat_context(s)
    u = Vector{Float64}(s.u)
    cc = Vector{Float64}(undef, length(tc.cell_comp))
    g = Vector{Float64}(undef, length(tc.grad))
    GC.@preserve lam u cc g _heat_check(
        ccall((:tohip_heat_cell_products, libtopopt_hip), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Ptr{Float64}, Ptr{Float64}),
              c.ptr, lam, u, 1.0, cc, g), c.ptr, "tohip_heat_cell_products")
    tc.adjoint_sol .= lam
    tc.cell_comp .= cc
    tc.grad .= g
    return T(obj)
end

# ------------------------------------------------------- assembled K --
# solve_system!(CGAssemblySolver, ...) (src/FEA/solvers_api.jl:223-250) on the
# device: the context of the matrix-free solver (same grid, densities and
# prescribed DOFs), K assembled there in CSR on every solve as the
# reference's call operator re-assembles it.
function hip_context(s::GenericFEASolver{T,LinearElasticity,HIPCGAssemblySolver}) where {T}
    return _hip_context_impl(s)
end

function solve_system!(::Type{HIPCGAssemblySolver},
                       s::GenericFEASolver{T,LinearElasticity,HIPCGAssemblySolver}, K, rhs, lhs;
                       safe=false, initially_zero=true, reltol=sqrt(eps(Float64)),
                       preconditioner=:identity, kwargs...) where {T}
    c = _sync!(hip_context(s), s)
    f = Vector{Float64}(rhs)
    u = Vector{Float64}(lhs)
    iters = Ref{Int64}(0)
    resid = Ref{Float64}(0)
    rc = GC.@preserve f u ccall(
        (:tohip_csr_solve, libtopopt_hip), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Int64, Cint, Cint, Cint,
         Ptr{Int64}, Ptr{Float64}),
        c.ptr, f, u, Float64(s.abstol), Float64(reltol), Int64(s.cg_max_iter),
        Cint(initially_zero), Cint(preconditioner === :jacobi), Cint(safe), iters, resid)
    _check(rc, c.ptr, "tohip_csr_solve")
    lhs .= u
    return nothing
end

function (s::GenericFEASolver{T,LinearElasticity,HIPCGAssemblySolver})(
    reuse_fact::Bool=false, (::Type{Val{safe}})=Val{false};
    assemble_f=true, rhs=nothing, lhs=assemble_f ? s.u : s.lhs, kwargs...,
) where {T,safe}
    c = hip_context(s)
    if rhs === nothing
        rhs = assemble_f ? c.f : s.rhs
    else
        rhs = copy(rhs)
        rhs[s.problem.ch.prescribed_dofs] .= 0
    end
    solve_system!(HIPCGAssemblySolver, s, nothing, rhs, lhs; safe=safe, kwargs...)
    return nothing
end

end # module
