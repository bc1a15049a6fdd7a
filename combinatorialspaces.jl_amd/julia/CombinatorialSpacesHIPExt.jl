"""
Package extension adding `Val(:HIP)` methods to CombinatorialSpaces.jl, backed by
libdec_hip.so (HIP, gfx950) through `ccall`.  It mirrors the method list of
ext/CombinatorialSpacesCUDAExt.jl: the same positional signatures, with the
backend value selecting this library.  Forms are AMDGPU.jl `ROCArray`s; the
library never copies them.  Register it in Project.toml as

    [weakdeps]   AMDGPU = "21141c5a-9bdb-4563-92ae-f87d6854732e"
    [extensions] CombinatorialSpacesHIPExt = "AMDGPU"

This file is not executed by this repository's tests (no Julia in the build
image); tests drive the same C ABI from Python.
"""
module CombinatorialSpacesHIPExt

using CombinatorialSpaces
using CombinatorialSpaces.DiscreteExteriorCalculus: DiscreteHodge, Circumcenter, Barycenter,
  geometric_center, triangle_center, tetrahedron_center
using CombinatorialSpaces.Multigrid: BinarySubdivision, AbstractSubdivisionScheme,
  AbstractMultigridMode, DirectMode, GalerkinMode, MultigridData, MGData, PrimalGeometricMapSeries
using AMDGPU
using Krylov
using LinearAlgebra
using SparseArrays
import CombinatorialSpaces: cache_wedge, dec_wedge_product, dec_wedge_product_dd,
  dec_c_wedge_product!, dec_c_wedge_product, dec_Δ⁻¹, ∇²,
  dec_wedge_product_pd, dec_wedge_product_dp, ♭♯_mat, interior_product_dd, ℒ_dd, Δᵈ,
  dec_boundary, dec_differential, dec_dual_derivative,
  dec_hodge_star, dec_inv_hodge_star, avg₀₁_mat,
  d0_p0_interpolation, p0_d0_interpolation, p2_d2_interpolation, p3_d3_interpolation,
  multigrid_vcycles, multigrid_wcycles, full_multigrid

const libdec_hip = get(ENV, "DEC_HIP_LIB", "libdec_hip")

# C ABI constants (include/dec_hip.h)
const DEC_F64, DEC_F32 = Int32(0), Int32(1)
const DEC_CIRCUMCENTER, DEC_BARYCENTER = Int32(0), Int32(1)
const DEC_GEOMETRIC_HODGE, DEC_DIAGONAL_HODGE = Int32(0), Int32(1)
const DEC_OP_WEDGE, DEC_OP_WEDGE_DD, DEC_OP_DIFFERENTIAL, DEC_OP_BOUNDARY,
      DEC_OP_DUAL_DERIVATIVE, DEC_OP_HODGE, DEC_OP_INV_HODGE, DEC_OP_LAPLACIAN0,
      DEC_OP_AVG01, DEC_OP_WEDGE_PD, DEC_OP_FLAT_SHARP, DEC_OP_WEDGE_DP,
      DEC_OP_INTERIOR_PRODUCT_DD, DEC_OP_LIE_DD, DEC_OP_DUAL_LAPLACIAN,
      DEC_OP_P2_D2_INTERPOLATION, DEC_OP_P3_D3_INTERPOLATION, DEC_OP_D0_P0_INTERPOLATION,
      DEC_OP_P0_D0_INTERPOLATION = Int32.(1:19)
const DEC_SMOOTH_CG, DEC_SMOOTH_GS, DEC_SMOOTH_JACOBI, DEC_SMOOTH_BLOCK_GS = Int32.(0:3)

function check(status::Cint)
  status == 0 && return nothing
  msg = unsafe_string(ccall((:dec_last_error, libdec_hip), Cstring, ()))
  status == 2 ? error(msg) : throw(ErrorException("dec_hip status $status: $msg"))
end

dtype_code(::Type{Float64}) = DEC_F64
dtype_code(::Type{Float32}) = DEC_F32
hodge_code(::GeometricHodge) = DEC_GEOMETRIC_HODGE
hodge_code(::DiagonalHodge) = DEC_DIAGONAL_HODGE
center_code(::Circumcenter) = DEC_CIRCUMCENTER
center_code(::Barycenter) = DEC_BARYCENTER

# The FastDEC methods below that CombinatorialSpaces defines for every
# HasDeltaSet2D also serve tetrahedral complexes (FastDEC.jl:42-133, 433-462, 606-636).
const DualComplex = Union{EmbeddedDeltaDualComplex2D, EmbeddedDeltaDualComplex3D}
stream_ptr() = AMDGPU.stream().stream

# ---------------------------------------------------------------- meshes

mutable struct HIPMesh
  handle::Ptr{Cvoid}
end

const MESHES = WeakKeyDict{Any, HIPMesh}()

function pack_points(sd)
  pts = Vector{Float64}(undef, 3 * nv(sd))
  for (i, p) in enumerate(sd[:point])
    pts[3i-2] = p[1]; pts[3i-1] = p[2]; pts[3i] = length(p) > 2 ? p[3] : 0.0
  end
  pts
end
orientation_codes(xs) = Int8[isnothing(o) ? -1 : Int8(o) for o in xs]

function register_mesh(sd, h, center)
  check(ccall((:dec_mesh_dualize, libdec_hip), Cint, (Ptr{Cvoid}, Int32), h, center_code(center)))
  m = HIPMesh(h)
  finalizer(x -> ccall((:dec_mesh_destroy, libdec_hip), Cint, (Ptr{Cvoid},), x.handle), m)
  m
end

# The center `dualize(s, center)` used for sd, read from sd's own dual points:
# the dual point of a cell is compared with the barycenter and the
# circumcenter of its vertices (geometric_center, DiscreteExteriorCalculus.jl:
# 100-135), on the first cell where those two differ.  A complex whose cells
# all have coinciding centers gets the same geometry either way.
function dual_center(sd, cells, center_of, verts_of)
  for x in cells
    ps = sd[verts_of(sd, x), :point]
    bary, circ = geometric_center(ps, Barycenter()), geometric_center(ps, Circumcenter())
    scale = maximum(norm(p - bary) for p in ps)
    norm(bary - circ) > 1e-6 * scale || continue
    dp = sd[center_of(sd, x), :dual_point]
    norm(dp - bary) <= 1e-9 * scale && return Barycenter()
    norm(dp - circ) <= 1e-9 * scale && return Circumcenter()
    error("HIP backend: dual points are neither barycenters nor circumcenters")
  end
  Circumcenter()
end
dual_center(sd::EmbeddedDeltaDualComplex2D) =
  dual_center(sd, triangles(sd), triangle_center, triangle_vertices)
dual_center(sd::EmbeddedDeltaDualComplex3D) =
  dual_center(sd, tetrahedra(sd), tetrahedron_center, tetrahedron_vertices)

"""Upload (once per dual complex) the primal boundary maps, orientations and
points; the dual geometry is then computed on the GPU in Float64 with the
center that sd itself was dualized with (`dual_center`)."""
function hip_mesh(sd::EmbeddedDeltaDualComplex2D)
  get!(MESHES, sd) do
    center = dual_center(sd)
    nv_, ne_, nt_ = nv(sd), ne(sd), ntriangles(sd)
    v0 = Int32.(sd[:∂v0]); v1 = Int32.(sd[:∂v1])
    e0 = Int32.(sd[:∂e0]); e1 = Int32.(sd[:∂e1]); e2 = Int32.(sd[:∂e2])
    eo = orientation_codes(sd[:edge_orientation])
    to = orientation_codes(sd[:tri_orientation])
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:dec_mesh_create, libdec_hip), Cint,
      (Int32, Int32, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
       Ptr{Int8}, Ptr{Int8}, Ptr{Float64}, Int32, Ptr{Ptr{Cvoid}}),
      nv_, ne_, nt_, v0, v1, e0, e1, e2, eo, to, pack_points(sd), Int32(1), h))
    register_mesh(sd, h[], center)
  end
end

# A tetrahedral complex: dec_mesh3_create with ∂t0..∂t3 and the tetrahedron
# orientation as well (SimplicialSets.jl:418-606).
function hip_mesh(sd::EmbeddedDeltaDualComplex3D)
  get!(MESHES, sd) do
    center = dual_center(sd)
    v0 = Int32.(sd[:∂v0]); v1 = Int32.(sd[:∂v1])
    e0 = Int32.(sd[:∂e0]); e1 = Int32.(sd[:∂e1]); e2 = Int32.(sd[:∂e2])
    t0 = Int32.(sd[:∂t0]); t1 = Int32.(sd[:∂t1]); t2 = Int32.(sd[:∂t2]); t3 = Int32.(sd[:∂t3])
    eo = orientation_codes(sd[:edge_orientation])
    to = orientation_codes(sd[:tri_orientation])
    teo = orientation_codes(sd[:tet_orientation])
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:dec_mesh3_create, libdec_hip), Cint,
      (Int32, Int32, Int32, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
       Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int8}, Ptr{Int8}, Ptr{Int8},
       Ptr{Float64}, Int32, Ptr{Ptr{Cvoid}}),
      nv(sd), ne(sd), ntriangles(sd), ntetrahedra(sd), v0, v1, e0, e1, e2, t0, t1, t2, t3,
      eo, to, teo, pack_points(sd), Int32(1), h))
    register_mesh(sd, h[], center)
  end
end

# ---------------------------------------------------------------- operators

mutable struct HIPOperator{T}
  handle::Ptr{Cvoid}
  rows::Int
  cols::Int
  cols2::Int
end

function HIPOperator{T}(sd, kind, m, n, hodge) where T
  mesh = hip_mesh(sd)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:dec_op_create, libdec_hip), Cint,
    (Ptr{Cvoid}, Int32, Int32, Int32, Int32, Int32, Ptr{Ptr{Cvoid}}),
    mesh.handle, Int32(kind), Int32(m), Int32(n), hodge, dtype_code(T), h))
  r, c, c2 = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
  check(ccall((:dec_op_shape, libdec_hip), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}),
    h[], r, c, c2))
  op = HIPOperator{T}(h[], r[], c[], c2[])
  finalizer(x -> ccall((:dec_op_destroy, libdec_hip), Cint, (Ptr{Cvoid},), x.handle), op)
  op
end

function apply!(res::ROCVector{T}, op::HIPOperator{T}, x::ROCVector{T}, y=nothing) where T
  length(x) == op.cols || throw(DimensionMismatch("first form has length $(length(x)), expected $(op.cols)"))
  op.cols2 > 0 && length(y) != op.cols2 && throw(DimensionMismatch("second form has wrong length"))
  check(ccall((:dec_op_apply, libdec_hip), Cint,
    (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{Cvoid}),
    op.handle, pointer(x), isnothing(y) ? Ptr{T}(C_NULL) : pointer(y), pointer(res), stream_ptr()))
  res
end

Base.:*(op::HIPOperator{T}, x::ROCVector{T}) where T = apply!(ROCVector{T}(undef, op.rows), op, x)
Base.size(op::HIPOperator) = (op.rows, op.cols)

float_of(sd::EmbeddedDeltaDualComplex2D{Bool, F, P}) where {F, P} = F
float_of(sd::EmbeddedDeltaDualComplex3D{Bool, F, P}) where {F, P} = F

# Wedge products.  cache_wedge returns the reference's (tables, coefficients,
# count) triple (src/FastDEC.jl:136-155): here the first slot is the device
# operator holding the library's tables, the second its coefficient store
# (the same operator), the last the result length.  dec_c_wedge_product! and
# dec_c_wedge_product (:217-238) then run the library's kernel when handed
# that triple, so caller code that caches tables keeps working.
struct HIPWedgeTable{T}
  op::HIPOperator{T}
end
struct HIPWedgeCache{T}
  table::HIPWedgeTable{T}
  n::Int
end
Base.getindex(c::HIPWedgeCache, i::Int) = i == 1 || i == 2 ? c.table :
  i == 3 ? c.n : throw(BoundsError(c, i))
Base.first(c::HIPWedgeCache) = c.table
Base.last(c::HIPWedgeCache) = c.n
Base.length(::HIPWedgeCache) = 3

function cache_wedge(::Val{m}, ::Val{n}, sd::DualComplex, ::Val{:HIP},
                     arr_cons=ROCArray, cast_float=nothing) where {m, n}
  T = isnothing(cast_float) ? float_of(sd) : cast_float
  op = HIPOperator{T}(sd, DEC_OP_WEDGE, m, n, DEC_GEOMETRIC_HODGE)
  HIPWedgeCache{T}(HIPWedgeTable{T}(op), op.rows)
end

function dec_c_wedge_product!(::Val{j}, ::Val{k}, res::ROCVector{T}, α, β, p::HIPWedgeTable{T}, c) where {j, k, T}
  apply!(res, p.op, α, β)
end

function dec_c_wedge_product(::Val{m}, ::Val{n}, α, β, wc::HIPWedgeCache{T}) where {m, n, T}
  α_data = α isa SimplexForm ? α.data : α
  β_data = β isa SimplexForm ? β.data : β
  apply!(ROCVector{T}(undef, last(wc)), wc.table.op, α_data, β_data)
end

function dec_wedge_product(::Val{m}, ::Val{n}, sd::DualComplex, ::Val{:HIP},
                           arr_cons=ROCArray, cast_float=nothing) where {m, n}
  (m, n) == (0, 0) && return (f, g) -> f .* g
  wc = cache_wedge(Val(m), Val(n), sd, Val(:HIP), arr_cons, cast_float)
  (α, β) -> dec_c_wedge_product(Val(m), Val(n), α, β, wc)
end

function dec_wedge_product_dd(::Val{m}, ::Val{n}, sd::EmbeddedDeltaDualComplex2D, ::Val{:HIP},
                              cast_float=nothing) where {m, n}
  T = isnothing(cast_float) ? float_of(sd) : cast_float
  op = HIPOperator{T}(sd, DEC_OP_WEDGE_DD, m, n, DEC_GEOMETRIC_HODGE)
  (f, g) -> apply!(similar(f, op.rows), op, f, g)
end

# Sparse operators (ext/CombinatorialSpacesCUDAExt.jl:30-61).
dec_boundary(n::Int, sd::DualComplex, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_BOUNDARY, n, 0, DEC_GEOMETRIC_HODGE)
dec_dual_derivative(n::Int, sd::DualComplex, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_DUAL_DERIVATIVE, n, 0, DEC_GEOMETRIC_HODGE)
dec_differential(n::Int, sd::DualComplex, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_DIFFERENTIAL, n, 0, DEC_GEOMETRIC_HODGE)
avg₀₁_mat(sd::EmbeddedDeltaDualComplex2D, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_AVG01, 0, 0, DEC_GEOMETRIC_HODGE)

# Hodge stars (ext/CombinatorialSpacesCUDAExt.jl:70-96): diagonal stars stay
# vectors on the device (no dense n×n matrix), geometric ⋆1 is a sparse operator.
dec_hodge_star(n::Int, sd::DualComplex, h::DiscreteHodge, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_HODGE, n, 0, hodge_code(h))
# Geometric ⋆1⁻¹ (CUDA ext :93-96 runs GMRES on -⋆1): the library's operator is
# the solve itself (CG or BiCGSTAB on the device), so it is returned as a closure.
function dec_inv_hodge_star(n::Int, sd::DualComplex, h::DiscreteHodge, ::Val{:HIP})
  op = HIPOperator{float_of(sd)}(sd, DEC_OP_INV_HODGE, n, 0, hodge_code(h))
  n == 1 && h isa GeometricHodge ? (x -> op * x) : op
end

# Primal-dual / dual-primal wedges (src/FastDEC.jl:374-400): closures as in FastDEC.
function dec_wedge_product_pd(::Val{m}, ::Val{n}, sd::EmbeddedDeltaDualComplex2D, ::Val{:HIP}) where {m, n}
  op = HIPOperator{float_of(sd)}(sd, DEC_OP_WEDGE_PD, m, n, DEC_GEOMETRIC_HODGE)
  (a, b) -> apply!(similar(a, op.rows), op, a, b)
end
function dec_wedge_product_dp(::Val{m}, ::Val{n}, sd::EmbeddedDeltaDualComplex2D, ::Val{:HIP}) where {m, n}
  op = HIPOperator{float_of(sd)}(sd, DEC_OP_WEDGE_DP, m, n, DEC_GEOMETRIC_HODGE)
  (a, b) -> apply!(similar(a, op.rows), op, a, b)
end

# ♭♯_mat (DiscreteExteriorCalculus.jl:1992) as a device operator.
♭♯_mat(sd::EmbeddedDeltaDualComplex2D, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_FLAT_SHARP, 0, 0, DEC_GEOMETRIC_HODGE)

# Interior products, Lie derivative and dual Laplacians (src/FastDEC.jl:758-860).
function interior_product_dd(::Val{m}, ::Val{n}, sd::EmbeddedDeltaDualComplex2D, ::Val{:HIP}) where {m, n}
  op = HIPOperator{float_of(sd)}(sd, DEC_OP_INTERIOR_PRODUCT_DD, m, n, DEC_GEOMETRIC_HODGE)
  (f, g) -> apply!(similar(f, op.rows), op, f, g)
end
function ℒ_dd(::Val{1}, ::Val{1}, sd::EmbeddedDeltaDualComplex2D, ::Val{:HIP})
  op = HIPOperator{float_of(sd)}(sd, DEC_OP_LIE_DD, 1, 1, DEC_GEOMETRIC_HODGE)
  (f, g) -> apply!(similar(f, op.rows), op, f, g)
end
function Δᵈ(::Val{n}, sd::DualComplex, ::Val{:HIP}) where n
  op = HIPOperator{float_of(sd)}(sd, DEC_OP_DUAL_LAPLACIAN, n, 0, DEC_GEOMETRIC_HODGE)
  x -> op * x
end

# Interpolation matrices (FastDEC.jl:915-941, DiscreteExteriorCalculus.jl:2016-2053)
# as device operators.
p2_d2_interpolation(sd::EmbeddedDeltaDualComplex2D, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_P2_D2_INTERPOLATION, 0, 0, DEC_GEOMETRIC_HODGE)
p3_d3_interpolation(sd::EmbeddedDeltaDualComplex3D, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_P3_D3_INTERPOLATION, 0, 0, DEC_GEOMETRIC_HODGE)
d0_p0_interpolation(sd::DualComplex, ::Val{:HIP}; hodge=GeometricHodge()) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_D0_P0_INTERPOLATION, 0, 0, hodge_code(hodge))
p0_d0_interpolation(sd::DualComplex, ::Val{:HIP}) =
  HIPOperator{float_of(sd)}(sd, DEC_OP_P0_D0_INTERPOLATION, 0, 0, DEC_GEOMETRIC_HODGE)

# ---------------------------------------------------------------- multigrid

mutable struct HIPMultigridData
  handle::Ptr{Cvoid}
  n::Int
  steps::Int
end

const DEC_MG_DIRECT, DEC_MG_GALERKIN = Int32(0), Int32(1)
mode_code(::DirectMode) = DEC_MG_DIRECT
mode_code(::GalerkinMode) = DEC_MG_GALERKIN

"""HIP counterpart of `MultigridData(s, BinarySubdivision(), levels, op, steps;
mode)` (src/Multigrid.jl:698-702) with op = fast_laplace_beltrami using `hodge`
for ⋆1; `mode` is DirectMode() or GalerkinMode()."""
function HIPMultigridData(s::EmbeddedDeltaSet2D, ::BinarySubdivision, levels::Int, steps::Int;
                          hodge::DiscreteHodge=GeometricHodge(), project::Bool=false,
                          mode::AbstractMultigridMode=DirectMode())
  base = hip_mesh(dualize(s, Circumcenter()))   # the library dualizes every level itself
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:dec_mg_create_mode, libdec_hip), Cint,
    (Ptr{Cvoid}, Int32, Int32, Int32, Int32, Int32, Int32, Ptr{Ptr{Cvoid}}),
    base.handle, Int32(levels), Int32(project), hodge_code(hodge), DEC_CIRCUMCENTER, Int32(steps),
    mode_code(mode), h))
  n = Ref{Int32}(0)
  check(ccall((:dec_mg_level_size, libdec_hip), Cint, (Ptr{Cvoid}, Int32, Ref{Int32}), h[], 0, n))
  md = HIPMultigridData(h[], n[], steps)
  finalizer(x -> ccall((:dec_mg_destroy, libdec_hip), Cint, (Ptr{Cvoid},), x.handle), md)
  md
end

# CSR arrays (1-based, Int32) of a sparse matrix: the CSC arrays of its transpose.
function csr_arrays(M::AbstractMatrix)
  T = SparseMatrixCSC{Float64, Int32}(sparse(transpose(sparse(M))))
  (Vector{Int32}(T.colptr), Vector{Int32}(T.rowval), Vector{Float64}(T.nzval), size(M, 1))
end

"""HIP counterpart of `MultigridData(ops, restrictions, prolongations, steps)`
(src/Multigrid.jl:545-546): the hierarchy of a CPU `MultigridData` (operators,
restrictions and prolongations finest first) uploaded as is.  The reference's
other constructors go through it: `MultigridData(s, scheme, levels, op, steps)`
with any `op`, and `MGData(series, op, steps)`."""
function HIPMultigridData(md::MultigridData)
  steps = unique(md.steps)
  length(steps) == 1 || error("HIP backend: one smoothing step count for all levels")
  A = map(csr_arrays, md.operators)
  R = map(csr_arrays, md.restrictions)
  P = map(csr_arrays, md.prolongations)
  nl = length(A)
  sizes = Int32[a[4] for a in A]
  ptrs(ms, j) = Ptr{Cvoid}[pointer(m[j]) for m in ms]
  h = Ref{Ptr{Cvoid}}(C_NULL)
  GC.@preserve A R P begin
    check(ccall((:dec_mg_create_from_csr, libdec_hip), Cint,
      (Int32, Ptr{Int32}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
       Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
       Int32, Int32, Ptr{Ptr{Cvoid}}),
      Int32(nl), sizes, ptrs(A, 1), ptrs(A, 2), ptrs(A, 3), ptrs(R, 1), ptrs(R, 2), ptrs(R, 3),
      ptrs(P, 1), ptrs(P, 2), ptrs(P, 3), Int32(1), Int32(only(steps)), h))
  end
  out = HIPMultigridData(h[], Int(sizes[1]), Int(only(steps)))
  finalizer(x -> ccall((:dec_mg_destroy, libdec_hip), Cint, (Ptr{Cvoid},), x.handle), out)
  out
end

"""`MultigridData(s, scheme, levels, op, steps; alg, mode)` (src/Multigrid.jl:
698-702) with any operator function `op`: the reference builds the operators,
restrictions and prolongations on the host, the V-cycles run on the GPU."""
HIPMultigridData(s::EmbeddedDeltaSet2D, scheme::AbstractSubdivisionScheme, levels::Int, op::Function, steps;
                 alg=Circumcenter(), mode::AbstractMultigridMode=DirectMode()) =
  HIPMultigridData(MultigridData(s, scheme, levels, op, steps; alg=alg, mode=mode))

"""dec_Δ⁻¹(Val(0), series; steps, cycles, alg, μ) (src/FastDEC.jl:866-869) on
the GPU: MGData(series, sd -> ∇²(0, sd), steps) uploaded once, then full
multigrid with μ-cycles per call."""
function dec_Δ⁻¹(::Val{0}, s::PrimalGeometricMapSeries, ::Val{:HIP};
                 scheme::AbstractSubdivisionScheme=BinarySubdivision(), steps=3, cycles=5, alg=Krylov.cg, μ=2)
  md = HIPMultigridData(MGData(s, sd -> ∇²(0, sd), steps, scheme))
  b -> full_multigrid(b, md, cycles, alg, μ)
end

smoother_code(alg) = alg === Krylov.cg ? DEC_SMOOTH_CG :
  alg === :gauss_seidel ? DEC_SMOOTH_GS : alg === :jacobi ? DEC_SMOOTH_JACOBI :
  alg === :block_gauss_seidel ? DEC_SMOOTH_BLOCK_GS :
  error("unsupported smoother $alg")

function set_smoother!(md::HIPMultigridData, alg; omega=2/3, coarse_direct=false)
  check(ccall((:dec_mg_set_smoother, libdec_hip), Cint, (Ptr{Cvoid}, Int32, Int32, Float64, Int32),
    md.handle, smoother_code(alg), Int32(md.steps), omega, Int32(coarse_direct)))
end

function multigrid_μ_cycles(u::ROCVector{Float64}, b::ROCVector{Float64}, md::HIPMultigridData,
                            cycles, alg, μ)
  set_smoother!(md, alg)
  u = copy(u)
  check(ccall((:dec_mg_cycles, libdec_hip), Cint,
    (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
    md.handle, Int32(μ), Int32(cycles), pointer(u), pointer(b), stream_ptr()))
  u
end

multigrid_vcycles(u::ROCVector{Float64}, b::ROCVector{Float64}, md::HIPMultigridData, cycles, alg=Krylov.cg) =
  multigrid_μ_cycles(u, b, md, cycles, alg, 1)
multigrid_wcycles(u::ROCVector{Float64}, b::ROCVector{Float64}, md::HIPMultigridData, cycles, alg=Krylov.cg) =
  multigrid_μ_cycles(u, b, md, cycles, alg, 2)

# A u = b to ‖b - A u‖₂/‖b‖₂ ≤ tol (checked in float64), by dec_mg_solve.  No
# counterpart in the reference (its solves run a fixed number of cycles,
# src/Multigrid.jl:814-855).  variant: :plain (float64 V-cycles), :mixed
# (float32 V-cycles in float64 iterative refinement), :pcg / :pcg_mixed
# (conjugate gradients preconditioned by one float64 / float32 V-cycle); the
# smoother is the hybrid block Gauss-Seidel.  Returns (u, cycles or CG
# iterations, relative residual).
const SOLVE_VARIANTS = Dict(:plain => Int32(0), :mixed => Int32(1), :pcg => Int32(2), :pcg_mixed => Int32(3))

function multigrid_solve(b::ROCVector{Float64}, md::HIPMultigridData; u=nothing, tol=1e-8, max_cycles=200,
                         variant=:plain)
  set_smoother!(md, :block_gauss_seidel; coarse_direct=true)
  v = get(SOLVE_VARIANTS, variant) do
    error("unsupported solve variant $variant")
  end
  x = u === nothing ? fill!(similar(b), 0) : copy(u)
  done = Ref{Int32}(0)
  rr = Ref{Float64}(0)
  check(ccall((:dec_mg_solve, libdec_hip), Cint,
    (Ptr{Cvoid}, Int32, Float64, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Float64}, Ptr{Cvoid}),
    md.handle, v, tol, Int32(max_cycles), pointer(x), pointer(b), done, rr, stream_ptr()))
  x, Int(done[]), rr[]
end

function full_multigrid(b::ROCVector{Float64}, md::HIPMultigridData, cycles, alg=Krylov.cg, μ=1)
  set_smoother!(md, alg)
  u = similar(b)
  check(ccall((:dec_mg_full_multigrid, libdec_hip), Cint,
    (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
    md.handle, Int32(μ), Int32(cycles), pointer(b), pointer(u), stream_ptr()))
  u
end

end
