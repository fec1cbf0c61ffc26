# GridapHybridMI355X.jl -- Julia binding of libhdg (include/hdg.h).
#
# Drop-in for the per-cell hybrid path of GridapHybrid.jl:
#   HybridAffineFEOperator(weakform, X, Y, [1,2], [3])
#   (src/HybridAffineFEOperators.jl:10-50) -> hdg_mesh_create + hdg_condense_assemble
#   when `weakform` is one of the typed weak forms below (DarcyHDGWeakForm,
#   ElasticityHDGWeakForm); any other weak form (a plain Function) keeps
#   GridapHybrid's own method, selected by Julia's dispatch.
#   solve(op) (:67-100) -> host sparse solve of the returned CSC + hdg_recover.
# Written against the C ABI and Gridap 0.18's public API; this image has no
# Julia, so this file is not executed by the test suite (tests/test_abi.py
# checks that every symbol it binds is exported by the library).
module GridapHybridMI355X

using SparseArrays
using LinearAlgebra
using Gridap
using Gridap.Geometry
using Gridap.ReferenceFEs
using Gridap.FESpaces
using Gridap.MultiField
import GridapHybrid

const libhdg = get(ENV, "LIBHDG", joinpath(@__DIR__, "..", "gridaphybrid", "libhdg.so"))

const HDG_OK = Cint(0)
const HDG_SINGULAR = Cint(1)
const HDG_FP64 = Cint(0)
const HDG_MIXED = Cint(1)
const HDG_DARCY = Cint(0)
const HDG_ELASTICITY = Cint(1)
const HDG_SPACE_Q = Cint(0)
const HDG_SPACE_P = Cint(1)

struct HdgProblem            # hdg_problem
  physics::Cint
  dim::Cint
  space::Cint
  k_flux::Cint
  k_pot::Cint
  k_trace::Cint
  quad_degree::Cint
  tau::Cdouble
  young::Cdouble
  source::Cint
  projection::Cint         # elasticity: bulk-to-skeleton projection P_m (0 / 1)
end

struct HdgMeshDesc           # hdg_mesh_desc
  ncells::Int64
  nfacets::Int64
  x0::Ptr{Float64}
  jac::Ptr{Float64}
  cell_facets::Ptr{Int32}
  facet_dirichlet::Ptr{Int8}
  coef::Ptr{Float64}
  facet_perm::Ptr{Int8}
end

struct HdgMeshInfo           # hdg_mesh_info
  ni::Cint; nb::Cint; n_flux::Cint; n_pot::Cint; n_trace_facet::Cint; nq_cell::Cint
  ncells::Int64; nfacets::Int64; nfree::Int64; ndir::Int64; nnz::Int64; nfree_total::Int64
end

mutable struct Context
  h::Ptr{Cvoid}
end

function check(ctx, rc, what)
  rc == HDG_OK && return
  msg = ctx === nothing ? "" : unsafe_string(ccall((:hdg_last_error, libhdg), Cstring, (Ptr{Cvoid},), ctx.h))
  # mirrors the reference's @check info==0 (src/StaticCondensationMap.jl:180)
  rc == HDG_SINGULAR && error("$what: singular A_ii: $msg")
  error("$what: libhdg error $rc: $msg")
end

function Context(device::Integer=0)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(nothing, ccall((:hdg_ctx_create, libhdg), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, h), "hdg_ctx_create")
  ctx = Context(h[])
  finalizer(c -> ccall((:hdg_ctx_destroy, libhdg), Cint, (Ptr{Cvoid},), c.h), ctx)
  ctx
end

"""
    MI355XHybridOperator(ctx, model, physics; order, tau, kappa_or_nu, E, source, precision)

GPU replacement of `HybridAffineFEOperator(weakform, X, Y, [1,2], [3])` for the
Darcy HDG weak form of test/DarcyHDGTests.jl:125-135 and the elasticity HDG
weak form of test/LinearElasticityHDGTests.jl:140-150 on a
`CartesianDiscreteModel(domain, partition)`.
"""
mutable struct MI355XHybridOperator
  ctx::Context
  mesh::Ptr{Cvoid}
  info::HdgMeshInfo
end

function MI355XHybridOperator(ctx::Context, domain::NTuple{N,Float64}, partition::NTuple{D,Int};
                              physics=:darcy, order=1, tau=1.0, coef=1.0, E=1.0, source=2,
                              precision=HDG_FP64) where {N,D}
  @assert N == 2D
  el = physics == :elasticity
  prob = HdgProblem(el ? HDG_ELASTICITY : HDG_DARCY, D, el ? HDG_SPACE_P : HDG_SPACE_Q,
                    order, el ? order + 1 : order, order, 2 * (order + 1), tau, E, source, 0)
  ncells = prod(partition)
  c = coef isa AbstractVector ? Vector{Float64}(coef) : fill(Float64(coef), ncells)
  cells = Int64[partition...; ones(Int64, 3 - D)]
  dom = Float64[domain...; zeros(6 - 2D)]
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ctx, ccall((:hdg_mesh_create_cartesian, libhdg), Cint,
                   (Ptr{Cvoid}, Ref{HdgProblem}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
                   ctx.h, prob, cells, dom, c, h), "hdg_mesh_create_cartesian")
  info = Ref{HdgMeshInfo}()
  check(ctx, ccall((:hdg_mesh_get_info, libhdg), Cint, (Ptr{Cvoid}, Ref{HdgMeshInfo}), h[], info),
        "hdg_mesh_get_info")
  check(ctx, ccall((:hdg_mesh_set_dirichlet_mms, libhdg), Cint, (Ptr{Cvoid},), h[]), "dirichlet")
  nbad = Ref{Int64}(0)
  check(ctx, ccall((:hdg_condense_assemble, libhdg), Cint, (Ptr{Cvoid}, Cint, Ref{Int64}),
                   h[], precision, nbad), "HybridAffineFEOperator")
  op = MI355XHybridOperator(ctx, h[], info[])
  finalizer(o -> ccall((:hdg_mesh_destroy, libhdg), Cint, (Ptr{Cvoid},), o.mesh), op)
  op
end

"The skeleton system as a host SparseMatrixCSC (1-based) and rhs."
function skeleton_system(op::MI355XHybridOperator)
  cp, rv, nz, rh = Ref{Ptr{Int64}}(), Ref{Ptr{Int32}}(), Ref{Ptr{Float64}}(), Ref{Ptr{Float64}}()
  check(op.ctx, ccall((:hdg_skeleton_system, libhdg), Cint,
                      (Ptr{Cvoid}, Ref{Ptr{Int64}}, Ref{Ptr{Int32}}, Ref{Ptr{Float64}}, Ref{Ptr{Float64}}),
                      op.mesh, cp, rv, nz, rh), "hdg_skeleton_system")
  n, nnz = op.info.nfree, op.info.nnz
  colptr = Vector{Int64}(undef, n + 1); rowval = Vector{Int32}(undef, nnz)
  nzval = Vector{Float64}(undef, nnz); rhs = Vector{Float64}(undef, n)
  for (dst, src) in ((colptr, cp[]), (rowval, rv[]), (nzval, nz[]), (rhs, rh[]))
    check(op.ctx, ccall((:hdg_memcpy_d2h, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                        op.ctx.h, dst, src, sizeof(dst)), "d2h")
  end
  SparseMatrixCSC(n, n, colptr .+ 1, Int64.(rowval) .+ 1, nzval), rhs
end

"""
    solve(op) -> free values [q of all cells; u of all cells; lambda free]

src/HybridAffineFEOperators.jl:67-100: host sparse direct solve of the
skeleton system, then GPU recovery (BackwardStaticCondensationMap).
"""
function solve(op::MI355XHybridOperator)
  K, rhs = skeleton_system(op)
  _recover(op, K \ rhs)
end

function _recover(op::MI355XHybridOperator, lam::Vector{Float64})
  nt = op.info.nfree_total
  dlam, dout = Ref{Ptr{Cvoid}}(), Ref{Ptr{Cvoid}}()
  # device buffers through HIP's allocator (libamdhip64)
  @assert ccall((:hipMalloc, "libamdhip64"), Cint, (Ref{Ptr{Cvoid}}, Csize_t), dlam, max(sizeof(lam), 8)) == 0
  @assert ccall((:hipMalloc, "libamdhip64"), Cint, (Ref{Ptr{Cvoid}}, Csize_t), dout, 8nt) == 0
  try
    check(op.ctx, ccall((:hdg_memcpy_h2d, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                        op.ctx.h, dlam[], lam, sizeof(lam)), "h2d")
    check(op.ctx, ccall((:hdg_recover, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                        op.mesh, dlam[], dout[]), "hdg_recover")
    out = Vector{Float64}(undef, nt)
    check(op.ctx, ccall((:hdg_memcpy_d2h, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                        op.ctx.h, out, dout[], 8nt), "d2h")
    return out
  finally
    ccall((:hipFree, "libamdhip64"), Cint, (Ptr{Cvoid},), dlam[])
    ccall((:hipFree, "libamdhip64"), Cint, (Ptr{Cvoid},), dout[])
  end
end

"""
    static_condensation(ctx, A, b, field_sizes, interior_fields, boundary_fields)

Map-level StaticCondensationMap over a batch of cells (src/StaticCondensationMap.jl:152-196):
A is (n, n, ncells) column-major per cell, fields 1-based as in the reference.
Returns (S, g) with S (nb, nb, ncells).
"""
function static_condensation(ctx::Context, A::Array{Float64,3}, b::Matrix{Float64},
                             field_sizes::Vector{Int}, interior_fields::Vector{Int},
                             boundary_fields::Vector{Int})
  n = sum(field_sizes); ncells = size(A, 3)
  nb = sum(field_sizes[boundary_fields])
  fs = Int32.(field_sizes); fi = Int32.(interior_fields .- 1); fb = Int32.(boundary_fields .- 1)
  fields = (Cint(length(fs)), pointer(fs), Cint(length(fi)), pointer(fi), Cint(length(fb)), pointer(fb))
  dev(x) = (p = Ref{Ptr{Cvoid}}(); @assert ccall((:hipMalloc, "libamdhip64"), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, max(x, 8)) == 0; p[])
  dA, db, dS, dg, di = dev(sizeof(A)), dev(sizeof(b)), dev(8nb * nb * ncells), dev(8nb * ncells), dev(4ncells)
  try
    GC.@preserve fs fi fb begin
      hf = Ref(fields)
      ccall((:hdg_memcpy_h2d, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64), ctx.h, dA, A, sizeof(A))
      ccall((:hdg_memcpy_h2d, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64), ctx.h, db, b, sizeof(b))
      check(ctx, ccall((:hdg_static_condensation, libhdg), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                       ctx.h, hf, ncells, dA, db, dS, dg, di), "StaticCondensationMap")
    end
    S = Array{Float64,3}(undef, nb, nb, ncells); g = Matrix{Float64}(undef, nb, ncells)
    ccall((:hdg_memcpy_d2h, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64), ctx.h, S, dS, sizeof(S))
    ccall((:hdg_memcpy_d2h, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64), ctx.h, g, dg, sizeof(g))
    return S, g
  finally
    for p in (dA, db, dS, dg, di)
      ccall((:hipFree, "libamdhip64"), Cint, (Ptr{Cvoid},), p)
    end
  end
end

# ---------------------------------------------------------------------------
# HybridAffineFEOperator(weakform, X, Y, bulk_fields, skeleton_fields) on the GPU
# ---------------------------------------------------------------------------
"""
Weak forms the GPU path integrates itself.  Each is callable as
`wf(x, y) -> (a(x, y), l(y))`, exactly the reference's weak form, so Gridap's
and GridapHybrid's own code accept it; `HybridAffineFEOperator` dispatches it
to libhdg instead.
"""
abstract type HDGWeakForm <: Function end

"""
    DarcyHDGWeakForm(dΩ, d∂K, n, nₒ; τ, κ, f, g, order, space=:Q, k_pot=order)

Darcy HDG of test/DarcyHDGTests.jl:125-135 with κ⁻¹ on the mass term (κ a
number or one value per cell).  `f` (vector) and `g` (scalar) are the sources.
`order`, `k_pot` and `space` must be those of the reffes of X (flux order
`order`, potential `k_pot`, trace `order`).
"""
struct DarcyHDGWeakForm{A,B,C,D,F,G} <: HDGWeakForm
  dΩ::A; d∂K::B; n::C; nₒ::D
  τ::Float64
  κ::Union{Float64,Vector{Float64}}
  f::F; g::G
  order::Int; k_pot::Int; space::Symbol
end
DarcyHDGWeakForm(dΩ, d∂K, n, nₒ; τ=1.0, κ=1.0, f, g, order, space=:Q, k_pot=order) =
  DarcyHDGWeakForm(dΩ, d∂K, n, nₒ, Float64(τ), κ isa Number ? Float64(κ) : Vector{Float64}(κ),
                   f, g, order, k_pot, space)

function (wf::DarcyHDGWeakForm)((uh, ph, lh), (vh, qh, mh))
  κinv = wf.κ isa Float64 ? 1.0 / wf.κ : CellField(1.0 ./ wf.κ, get_triangulation(wf.dΩ.quad))
  n, nₒ, τ, dΩ, d∂K = wf.n, wf.nₒ, wf.τ, wf.dΩ, wf.d∂K
  a = ∫(vh ⋅ (κinv * uh) - (∇ ⋅ vh) * ph - ∇(qh) ⋅ uh)dΩ + ∫((vh ⋅ n) * lh)d∂K +
      ∫(qh * (uh ⋅ n))d∂K + ∫(τ * qh * ph * (n ⋅ nₒ))d∂K - ∫(τ * qh * lh * (n ⋅ nₒ))d∂K +
      ∫(mh * (uh ⋅ n))d∂K + ∫(τ * mh * ph * (n ⋅ nₒ))d∂K - ∫(τ * mh * lh * (n ⋅ nₒ))d∂K
  l = ∫(vh ⋅ wf.f + qh * wf.g)dΩ
  a, l
end

"""
    ElasticityHDGWeakForm(dΩ, d∂K, n; τ, ν, E=1.0, f, order, bulk_to_skeleton_projection=false)

Linear-elasticity HDG of test/LinearElasticityHDGTests.jl:140-150, compliance
A(σ) of :15-17 with ν a number or one value per cell; P_order stress,
P_{order+1} displacement, P_order trace.  With bulk_to_skeleton_projection the
displacement in the τ terms is the reference's Pₘ projection (test/P_m.jl);
the callable form below then needs GridapHybrid's Pₘ, the GPU path builds it
into its tables.
"""
struct ElasticityHDGWeakForm{A,B,C,F} <: HDGWeakForm
  dΩ::A; d∂K::B; n::C
  τ::Float64
  ν::Union{Float64,Vector{Float64}}
  E::Float64
  f::F
  order::Int
  projection::Bool   # bulk_to_skeleton_projection (P_m of test/P_m.jl) in the tau terms
end
ElasticityHDGWeakForm(dΩ, d∂K, n; τ, ν, E=1.0, f, order, bulk_to_skeleton_projection=false) =
  ElasticityHDGWeakForm(dΩ, d∂K, n, Float64(τ), ν isa Number ? Float64(ν) : Vector{Float64}(ν),
                        Float64(E), f, order, bulk_to_skeleton_projection)

function (wf::ElasticityHDGWeakForm)((σh, uh, uhΓ), (v, ω, μ))
  wf.projection && error("the lazy form with P_m needs the Pm of GridapHybrid's test/P_m.jl; " *
                         "HybridAffineFEOperator(wf, ...) builds P_m into the GPU tables")
  ν = wf.ν isa Float64 ? wf.ν : CellField(wf.ν, get_triangulation(wf.dΩ.quad))
  E, τ, n, dΩ, d∂K = wf.E, wf.τ, wf.n, wf.dΩ, wf.d∂K
  Aσ = (1.0 + ν) / E * σh - (ν / E) * (tr ∘ σh) * one(SymTensorValue{num_cell_dims(get_triangulation(dΩ.quad)),Float64})
  a = ∫(v ⊙ Aσ + (∇ ⋅ v) ⋅ uh - ω ⋅ (∇ ⋅ σh))dΩ - ∫((v ⋅ n) ⋅ uhΓ)d∂K +
      ∫(τ * (ω ⋅ uh))d∂K - ∫(τ * (ω ⋅ uhΓ))d∂K + ∫(μ ⋅ (σh ⋅ n))d∂K -
      ∫(τ * μ ⋅ uh)d∂K + ∫(τ * μ ⋅ uhΓ)d∂K
  l = ∫(-ω ⋅ wf.f)dΩ
  a, l
end

"""The GPU-resident HybridAffineFEOperator: condensed skeleton system on the
device, trial/test spaces of the reference for the result."""
mutable struct MI355XHybridAffineFEOperator
  ctx::Context
  mesh::Ptr{Cvoid}
  info::HdgMeshInfo
  trial::MultiFieldFESpace
  test::MultiFieldFESpace
  src::Ptr{Cvoid}           # device source values at quadrature points (HIP allocation)
end

const _CTX = Ref{Union{Nothing,Context}}(nothing)
default_context() = (_CTX[] === nothing && (_CTX[] = Context(0)); _CTX[])

# Gridap's lexicographic vertex order of an n-cube: vertex 1 is the origin and
# vertex 1 + 2^(e-1) the end of reference edge e.
function _affine_cell(coords, D)
  x0 = coords[1]
  J = [coords[1 + 2^(e - 1)][a] - x0[a] for a in 1:D, e in 1:D]
  for (v, c) in enumerate(coords)   # every vertex must be x0 + J * its reference point
    ξ = [((v - 1) >> (e - 1)) & 1 for e in 1:D]
    maximum(abs.(Vector(c) .- (Vector(x0) .+ J * ξ))) <= 1e-12 * (1 + maximum(abs.(J))) ||
      return nothing
  end
  Vector(x0), J
end

# hdg_mesh_desc orientation code of a facet vertex permutation of Gridap
# (get_vertex_permutations of the facet polytope): the code whose coordinate
# map sends every cell-local facet vertex to the facet's own vertex it is
# identified with.
function _orientation_codes(facet_polytope, D)
  verts = [Vector(v) for v in get_vertex_coordinates(facet_polytope)]
  codes = Int8[]
  for perm in get_vertex_permutations(facet_polytope)
    code = -1
    for p in 0:(D == 2 ? 1 : 7)
      ok = all(1:length(verts)) do i
        y = verts[i]
        yo = if D == 2
          [(p & 1) != 0 ? 1 - y[1] : y[1]]
        else
          a, b = (p & 1) != 0 ? (y[2], y[1]) : (y[1], y[2])
          [(p & 2) != 0 ? 1 - a : a, (p & 4) != 0 ? 1 - b : b]
        end
        maximum(abs.(yo .- verts[perm[i]])) < 1e-12
      end
      ok && (code = p; break)
    end
    code < 0 && error("facet vertex permutation $perm has no orientation code")
    push!(codes, Int8(code))
  end
  codes
end

function GridapHybrid.HybridAffineFEOperator(wf::HDGWeakForm, trial::MultiFieldFESpace,
                                             test::MultiFieldFESpace,
                                             bulk_fields::Vector{<:Integer},
                                             skeleton_fields::Vector{<:Integer};
                                             precision=HDG_FP64, ctx::Context=default_context())
  model = get_background_model(get_triangulation(trial.spaces[1]))
  D = num_cell_dims(model)
  el = wf isa ElasticityHDGWeakForm
  # elasticity: the displacement trace as one vector field [3] or as one
  # scalar field per component [3, ..., 2+D] (src/HybridAffineFEOperators.jl:343-356)
  ntrace = el && skeleton_fields == collect(3:2 + D) ? D : 1
  (bulk_fields == [1, 2] && (skeleton_fields == [3] || ntrace > 1)) ||
    error("the GPU path hybridizes fields [1,2] onto [3] (elasticity: or onto [3, ..., 2+D]); " *
          "pass the weak form as a plain Function to use GridapHybrid's own operator")
  is_n_cube(first(get_polytopes(model))) ||
    error("the GPU path integrates parallelogram/parallelepiped cells only")
  grid = get_grid(model)
  topo = get_grid_topology(model)
  cc = get_cell_coordinates(grid)
  nc = length(cc)
  x0 = zeros(D, nc); jac = zeros(D, D, nc)
  for c in 1:nc
    r = _affine_cell(cc[c], D)
    r === nothing && error("cell $c is not affine; use GridapHybrid's own operator")
    x0[:, c] = r[1]; jac[:, :, c] = r[2]
  end
  c2f = get_faces(topo, D, D - 1)
  nf = num_faces(topo, D - 1)
  cf = Int32[c2f[c][lf] - 1 for lf in 1:2D, c in 1:nc]
  codes = _orientation_codes(get_facet_polytope(first(get_polytopes(model))), D)
  cperm = Gridap.Geometry.get_cell_permutations(topo, D - 1)
  fperm = Int8[codes[cperm[c][lf]] for lf in 1:2D, c in 1:nc]
  # Dirichlet facets and values from the trial trace space (negative ids)
  L = trial.spaces[3]
  fids = get_cell_dof_ids(L)
  fdir = Int8[any(<(0), fids[f]) for f in 1:nf]
  coef = el ? (wf.ν isa Float64 ? fill(wf.ν, nc) : wf.ν) : (wf.κ isa Float64 ? fill(wf.κ, nc) : wf.κ)
  length(coef) == nc || error("one coefficient per cell expected")
  k = wf.order
  prob = HdgProblem(el ? HDG_ELASTICITY : HDG_DARCY, D,
                    el ? HDG_SPACE_P : (wf.space == :P ? HDG_SPACE_P : HDG_SPACE_Q),
                    k, el ? k + 1 : wf.k_pot, k, 2 * (k + 1), wf.τ, el ? wf.E : 1.0, Cint(100),
                    el && wf.projection ? 1 : 0)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  GC.@preserve x0 jac cf fdir coef fperm begin
    desc = HdgMeshDesc(nc, nf, pointer(x0), pointer(jac), pointer(cf), pointer(fdir),
                       pointer(coef), any(!=(0), fperm) ? pointer(fperm) : Ptr{Int8}(C_NULL))
    check(ctx, ccall((:hdg_mesh_create, libhdg), Cint,
                     (Ptr{Cvoid}, Ref{HdgProblem}, Ref{HdgMeshDesc}, Ptr{Ptr{Cvoid}}),
                     ctx.h, prob, desc, h), "hdg_mesh_create")
  end
  ntrace > 1 && check(ctx, ccall((:hdg_mesh_set_trace_fields, libhdg), Cint, (Ptr{Cvoid}, Cint),
                                 h[], ntrace), "hdg_mesh_set_trace_fields")
  info = Ref{HdgMeshInfo}()
  check(ctx, ccall((:hdg_mesh_get_info, libhdg), Cint, (Ptr{Cvoid}, Ref{HdgMeshInfo}), h[], info),
        "hdg_mesh_get_info")
  num_free_dofs(trial) == info[].nfree_total ||
    error("the spaces of X do not match the weak form's orders ($(num_free_dofs(trial)) free " *
          "dofs, the GPU path has $(info[].nfree_total))")
  # Dirichlet values of the trial trace space(s), in Gridap's Dirichlet numbering
  lam_dir = ntrace > 1 ?
    reduce(vcat, [Vector{Float64}(get_dirichlet_dof_values(trial.spaces[2 + i])) for i in 1:ntrace]) :
    Vector{Float64}(get_dirichlet_dof_values(L))
  check(ctx, ccall((:hdg_mesh_set_dirichlet, libhdg), Cint, (Ptr{Cvoid}, Ptr{Float64}), h[],
                   lam_dir), "hdg_mesh_set_dirichlet")
  # sources at the cell quadrature points, x = x0 + J xi
  nq = info[].nq_cell
  pts = zeros(D, nq); wts = zeros(nq)
  check(ctx, ccall((:hdg_mesh_cell_quadrature, libhdg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                   h[], pts, wts), "hdg_mesh_cell_quadrature")
  ncomp = el ? D : D + 1
  src = zeros(ncomp, nq, nc)
  for c in 1:nc, q in 1:nq
    x = Point(Tuple(x0[:, c] .+ jac[:, :, c] * pts[:, q]))
    fv = wf.f(x)
    for a in 1:D
      src[a, q, c] = fv[a]
    end
    el || (src[D + 1, q, c] = wf.g(x))
  end
  dsrc = Ref{Ptr{Cvoid}}()
  @assert ccall((:hipMalloc, "libamdhip64"), Cint, (Ref{Ptr{Cvoid}}, Csize_t), dsrc, sizeof(src)) == 0
  check(ctx, ccall((:hdg_memcpy_h2d, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                   ctx.h, dsrc[], src, sizeof(src)), "h2d")
  check(ctx, ccall((:hdg_mesh_set_source, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), h[], dsrc[]),
        "hdg_mesh_set_source")
  nbad = Ref{Int64}(0)
  check(ctx, ccall((:hdg_condense_assemble, libhdg), Cint, (Ptr{Cvoid}, Cint, Ref{Int64}),
                   h[], precision, nbad), "HybridAffineFEOperator")
  op = MI355XHybridAffineFEOperator(ctx, h[], info[], trial, test, dsrc[])
  finalizer(op) do o
    ccall((:hdg_mesh_destroy, libhdg), Cint, (Ptr{Cvoid},), o.mesh)
    ccall((:hipFree, "libamdhip64"), Cint, (Ptr{Cvoid},), o.src)
  end
  op
end

"""
    solve(op::MI355XHybridAffineFEOperator) -> MultiFieldFEFunction

src/HybridAffineFEOperators.jl:67-100: the skeleton system is solved on the
host (`K \\ rhs`), the interior fields are recovered on the GPU, and the free
values [flux; potential; trace] become an FEFunction of the trial space.
"""
function Gridap.solve(op::MI355XHybridAffineFEOperator; skeleton_solver::Symbol=:host,
                      rtol::Float64=1e-12)
  skeleton_solver === :device &&
    return FEFunction(op.trial, _recover(op, skeleton_solve_device(op; rtol=rtol)))
  K, rhs = skeleton_system(op)
  FEFunction(op.trial, _recover(op, K \ rhs))
end

"""
    skeleton_solve_device(op; rtol=1e-12, maxit=0) -> lambda (host vector)

The skeleton solve on the GPU (hdg_skeleton_solve: MINRES with a facet
block-Jacobi preconditioner on the resident system) instead of the host
`A \\ b` of src/HybridAffineFEOperators.jl:74; errors when rtol is not met.
"""
function skeleton_solve_device(op::MI355XHybridAffineFEOperator; rtol::Float64=1e-12,
                               maxit::Integer=0)
  n = op.info.nfree
  n == 0 && return Float64[]
  dlam = Ref{Ptr{Cvoid}}()
  @assert ccall((:hipMalloc, "libamdhip64"), Cint, (Ref{Ptr{Cvoid}}, Csize_t), dlam, 8n) == 0
  try
    iters, relres = Ref{Cint}(0), Ref{Float64}(0.0)
    check(op.ctx, ccall((:hdg_skeleton_solve, libhdg), Cint,
                        (Ptr{Cvoid}, Float64, Cint, Ptr{Cvoid}, Ref{Cint}, Ref{Float64}),
                        op.mesh, rtol, maxit, dlam[], iters, relres), "hdg_skeleton_solve")
    lam = Vector{Float64}(undef, n)
    check(op.ctx, ccall((:hdg_memcpy_d2h, libhdg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                        op.ctx.h, lam, dlam[], 8n), "d2h")
    return lam
  finally
    ccall((:hipFree, "libamdhip64"), Cint, (Ptr{Cvoid},), dlam[])
  end
end

skeleton_system(op::MI355XHybridAffineFEOperator) =
  skeleton_system(MI355XHybridOperator(op.ctx, op.mesh, op.info))

_recover(op::MI355XHybridAffineFEOperator, lam) =
  _recover(MI355XHybridOperator(op.ctx, op.mesh, op.info), lam)

end # module
