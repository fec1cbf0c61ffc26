# GridapHybridMI355X.jl -- Julia binding of libhdg (include/hdg.h).
#
# Drop-in for the per-cell hybrid path of GridapHybrid.jl:
#   HybridAffineFEOperator construction (src/HybridAffineFEOperators.jl:10-50)
#   -> hdg_condense_assemble; solve(op) (:67-100) -> host sparse solve of the
#   returned CSC + hdg_recover.
# Written against the C ABI only; this image has no Julia, so this file is not
# executed by the test suite (tests/test_abi.py checks that every symbol it
# binds is exported by the library).
module GridapHybridMI355X

using SparseArrays
using LinearAlgebra

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
                    order, el ? order + 1 : order, order, 2 * (order + 1), tau, E, source)
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
  lam = K \ rhs
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

end # module
