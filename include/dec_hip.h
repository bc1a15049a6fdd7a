/*
 * dec_hip.h -- C ABI of libdec_hip.so, the MI355X (gfx950) implementation of
 * CombinatorialSpaces.jl's DEC hot path: FastDEC wedge products, exterior
 * derivative / boundary / dual derivative, Hodge stars and inverses, the
 * 0-form Laplace-Beltrami operator, and the geometric multigrid V-cycle.
 *
 * Conventions
 *  - Status: every entry point returns int32 status (DEC_OK == 0); on error
 *    dec_last_error() returns a thread-local message.  Nothing throws.
 *  - Indices are int32.  Host index arrays passed in are 0-based unless the
 *    index_base argument says 1 (Julia callers pass 1).
 *  - Forms (x, y, out) are caller-owned DEVICE pointers (ROCArray, torch,
 *    hipMalloc).  Meshes, operators and multigrid hierarchies own their device
 *    tables and are freed by the matching *_destroy.
 *  - stream is a hipStream_t (NULL = default stream).  Operators are
 *    immutable after creation: concurrent applies on different streams are
 *    safe.  A multigrid handle owns mutable work vectors: one stream at a time.
 *  - dtype selects the arithmetic type of form values: DEC_F64 or DEC_F32.
 */
#ifndef DEC_HIP_H
#define DEC_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dec_mesh_s *dec_mesh;
typedef struct dec_op_s *dec_op;
typedef struct dec_mg_s *dec_mg;
typedef struct dec_batch_s *dec_batch;
typedef struct dec_nn_s *dec_nn;

enum dec_status {
  DEC_OK = 0,
  DEC_ERR_ARG = 1,          /* bad argument / shape mismatch */
  DEC_ERR_UNSUPPORTED = 2,  /* "Unsupported combination of degrees ..." (FastDEC.jl:229, 256) */
  DEC_ERR_OOM = 3,
  DEC_ERR_HIP = 4,
  DEC_ERR_NOCONV = 5,       /* solver did not reach the tolerance */
  DEC_ERR_NONORIENTABLE = 6 /* orient! returned false (DiscreteExteriorCalculus.jl:576) */
};
enum dec_dtype { DEC_F64 = 0, DEC_F32 = 1 };
enum dec_center { DEC_CIRCUMCENTER = 0, DEC_BARYCENTER = 1 };
enum dec_hodge { DEC_GEOMETRIC_HODGE = 0, DEC_DIAGONAL_HODGE = 1 };
enum dec_op_kind {
  DEC_OP_WEDGE = 1,           /* primal-primal wedge, degrees (m,n) */
  DEC_OP_WEDGE_DD = 2,        /* dual-dual wedge (0,1) / (1,0) */
  DEC_OP_DIFFERENTIAL = 3,    /* d_n */
  DEC_OP_BOUNDARY = 4,        /* ∂_n */
  DEC_OP_DUAL_DERIVATIVE = 5, /* d̃_n */
  DEC_OP_HODGE = 6,           /* ⋆_n */
  DEC_OP_INV_HODGE = 7,       /* ⋆_n⁻¹ (diagonal) */
  DEC_OP_LAPLACIAN0 = 8,      /* -⋆0⁻¹ d̃1 ⋆1 d0 */
  DEC_OP_AVG01 = 9,           /* avg₀₁ */
  DEC_OP_WEDGE_PD = 10,       /* primal-dual wedge 0-1, 1-1 */
  DEC_OP_FLAT_SHARP = 11,     /* ♭♯_mat: dual 1-form -> primal 1-form */
  DEC_OP_WEDGE_DP = 12,       /* dual-primal wedge 1-0, 1-1 */
  DEC_OP_INTERIOR_PRODUCT_DD = 13, /* interior_product_dd (1,1), (1,2) */
  DEC_OP_LIE_DD = 14,         /* ℒ_dd (1,1) */
  DEC_OP_DUAL_LAPLACIAN = 15, /* Δᵈ n = 0, 1 */
  DEC_OP_P2_D2_INTERPOLATION = 16, /* primal 2-forms -> dual 2-forms, nv x nt */
  DEC_OP_P3_D3_INTERPOLATION = 17, /* primal 3-forms -> dual 3-forms, nv x ntet */
  DEC_OP_D0_P0_INTERPOLATION = 18, /* dual 0-forms -> primal 0-forms */
  DEC_OP_P0_D0_INTERPOLATION = 19  /* primal 0-forms -> dual 0-forms, nt x nv */
};
enum dec_smoother { DEC_SMOOTH_CG = 0, DEC_SMOOTH_GS = 1, DEC_SMOOTH_JACOBI = 2, DEC_SMOOTH_BLOCK_GS = 3 };
enum dec_solve_variant {
  DEC_SOLVE_PLAIN_F64 = 0,
  DEC_SOLVE_MIXED_IR = 1,
  DEC_SOLVE_PCG_F64 = 2,
  DEC_SOLVE_PCG_MIXED = 3
};

/* ---------------------------------------------------------------- library */
const char *dec_last_error(void);
const char *dec_version(void);
/* Fill a device vector with the seeded uniform [-1,1) form
 * x[i] = 2 u - 1, u = (splitmix64(seed * 2^32 + i) >> 11) * 2^-53. */
int dec_fill_uniform(void *x, int64_t n, uint64_t seed, int dtype, void *stream);

/* ------------------------------------------------------------------ meshes */
/* EmbeddedDeltaSet2D from boundary maps (Multigrid.jl:27-64 MeshTopology /
 * topo_to_mesh; SimplicialSets.jl:85-411).  edge_orient / tri_orient may be
 * NULL or hold -1 for "unset": unset edge orientations become false and unset
 * triangle orientations are set by orient! when the dual complex is built
 * (DiscreteExteriorCalculus.jl:311-321, 573-587). points: nv x 3 float64. */
int dec_mesh_create(int32_t nv, int32_t ne, int32_t nt, const int32_t *v0,
                    const int32_t *v1, const int32_t *e0, const int32_t *e1,
                    const int32_t *e2, const int8_t *edge_orient,
                    const int8_t *tri_orient, const double *points,
                    int32_t index_base, dec_mesh *out);
/* Pole-aligned icosahedron refined k times by binary subdivision with the new
 * points projected to the unit sphere (replaces loadmesh(Icosphere(n)),
 * CombMeshes.jl:31-36, whose artifact is a downloaded file). */
int dec_mesh_icosphere(int32_t k, dec_mesh *out);
/* EmbeddedDeltaSet2D(mesh) from a triangle list (MeshInterop.jl:116-131):
 * glue_sorted_triangle! (sorted = 1) or glue_triangle! (sorted = 0) for each
 * triangle in order (SimplicialSets.jl:343-364), then orient!.  points: nv x 3
 * float64; tris: nt x 3 vertex ids counted from index_base. */
int dec_mesh_from_triangles(int32_t nv, int32_t nt, const double *points,
                            const int32_t *tris, int32_t sorted, int32_t index_base,
                            dec_mesh *out);
/* triangulated_grid(lx, ly, dx, dy, Point3d, compress) (CombMeshes.jl:79-130). */
int dec_mesh_triangulated_grid(double lx, double ly, double dx, double dy,
                               int32_t compress, dec_mesh *out);
/* subdivision(s, BinarySubdivision()) (Multigrid.jl:447-452): refine +
 * propagate_points; project=1 also rescales new points to |p| = 1. */
int dec_mesh_subdivide(dec_mesh coarse, int32_t project, dec_mesh *fine);
int dec_mesh_counts(dec_mesh m, int32_t *nv, int32_t *ne, int32_t *nt);
/* Host copies of the primal complex (after orientation). Any pointer may be NULL. */
int dec_mesh_export(dec_mesh m, int32_t *v0, int32_t *v1, int32_t *e0, int32_t *e1,
                    int32_t *e2, int8_t *edge_orient, int8_t *tri_orient,
                    double *points);
/* dualize(s, center) geometry on the GPU in float64 (subdivide_duals_2d! +
 * precompute_volumes_2d!, DiscreteExteriorCalculus.jl:941-990). */
int dec_mesh_dualize(dec_mesh m, int32_t center);
/* Host copies of the dual geometry in the reference's DualV / DualE /
 * DualTri numbering: dual_point (nv+ne+nt) x 3, length ne, area nt,
 * dual_length 2ne+6nt, dual_area 6nt.  Any pointer may be NULL. */
int dec_mesh_export_dual(dec_mesh m, double *dual_point, double *length,
                         double *area, double *dual_length, double *dual_area);
int dec_mesh_destroy(dec_mesh m);

/* ------------------------------------------------------- 3-D (tetrahedra) */
/* EmbeddedDeltaSet3D from boundary maps (SimplicialSets.jl:418-606): edges
 * (∂v0, ∂v1), triangles (∂e0, ∂e1, ∂e2), tetrahedra (∂t0..∂t3).  Orientation
 * arrays may be NULL or hold -1 for "unset": when the dual complex is built,
 * unset edges and triangles become false and unset tetrahedra are set by
 * orient!(s, Val(3)) (DiscreteExteriorCalculus.jl:311-321, 573-584, 1198-1207).
 * The handle is a dec_mesh: dec_mesh_dualize, dec_mesh_export (the triangle
 * skeleton), dec_op_create and dec_mesh_destroy take it. */
int dec_mesh3_create(int32_t nv, int32_t ne, int32_t nt, int32_t ntet,
                     const int32_t *v0, const int32_t *v1, const int32_t *e0,
                     const int32_t *e1, const int32_t *e2, const int32_t *t0,
                     const int32_t *t1, const int32_t *t2, const int32_t *t3,
                     const int8_t *edge_orient, const int8_t *tri_orient,
                     const int8_t *tet_orient, const double *points,
                     int32_t index_base, dec_mesh *out);
/* EmbeddedDeltaSet3D(mesh) from a tetrahedron list (MeshInterop.jl:141-153):
 * glue_sorted_tetrahedron! (sorted = 1) or glue_tetrahedron! (sorted = 0) of
 * each tetrahedron in order (SimplicialSets.jl:511-534), then orient!.
 * points: nv x 3 float64; tets: ntet x 4 vertex ids counted from index_base. */
int dec_mesh3_from_tets(int32_t nv, int32_t ntet, const double *points,
                        const int32_t *tets, int32_t sorted, int32_t index_base,
                        dec_mesh *out);
/* 2 for a triangle complex, 3 for a tetrahedral one. */
int dec_mesh_dim(dec_mesh m, int32_t *dim);
int dec_mesh3_counts(dec_mesh m, int32_t *nv, int32_t *ne, int32_t *nt, int32_t *ntet);
/* Host copies of ∂t0..∂t3 and the tetrahedron orientation (after orient!). */
int dec_mesh3_export(dec_mesh m, int32_t *t0, int32_t *t1, int32_t *t2, int32_t *t3,
                     int8_t *tet_orient);
/* Host copies of the 3-D dual geometry (subdivide_duals_3d! +
 * precompute_volumes_3d!, DiscreteExteriorCalculus.jl:1303-1335):
 * dual_point (nv+ne+nt+ntet) x 3, length ne, area nt, vol ntet, dual_length
 * and dual_area of the first 2ne+6nt dual edges and 6nt dual triangles (the
 * 1-D and 2-D blocks, numbered as on a 2-D complex), and dual_vol 24 ntet
 * (DualTet 24x + q, D_tetrahedron_schemas order).  Any pointer may be NULL. */
int dec_mesh3_export_dual(dec_mesh m, double *dual_point, double *length, double *area,
                          double *vol, double *dual_length, double *dual_area,
                          double *dual_vol);

/* --------------------------------------------------------------- operators */
/* Create an operator on a dualized mesh.
 *  DEC_OP_WEDGE, (m,n) in {(0,1),(1,0),(0,2),(2,0),(1,1)}:
 *     dec_wedge_product(m, n, sd, Val(:HIP), ROCArray, T) (FastDEC.jl:255-289,
 *     CUDA counterpart ext/CombinatorialSpacesCUDAExt.jl:16-21); apply(x=α, y=β).
 *  DEC_OP_WEDGE_DD, (0,1) / (1,0): dec_wedge_product_dd (FastDEC.jl:313-325).
 *  DEC_OP_DIFFERENTIAL n=0,1: dec_differential (FastDEC.jl:468-469, CUDA ext :55-61).
 *  DEC_OP_BOUNDARY n=1,2: dec_boundary (FastDEC.jl:433-437, CUDA ext :30-36).
 *  DEC_OP_DUAL_DERIVATIVE n=0,1: dec_dual_derivative (FastDEC.jl:443-453, CUDA ext :42-48).
 *  DEC_OP_HODGE n=0,1,2 with hodge: dec_hodge_star (FastDEC.jl:641-716, CUDA ext :70-77).
 *  DEC_OP_INV_HODGE n=0,1,2: dec_inv_hodge_star (FastDEC.jl:721-743, CUDA ext :83-91).
 *     Geometric ⋆1⁻¹ (FastDEC.jl:744-748, CUDA ext :93-96) is a solve: apply
 *     returns (-⋆1)⁻¹ x from a Jacobi-scaled Krylov iteration on the device
 *     (CG where ⋆1 is symmetric, BiCGSTAB otherwise) stopped at
 *     ‖x - ⋆1 z‖ ≤ 1e-14 ‖x‖ (float64; 1e-6 in float32).  Apply synchronises
 *     the stream to read the stop flag after 2, 6, 14 and then every 8
 *     iterations, except while the stream is captured into a graph (then the
 *     whole budget of 200 (float64) or 100 (float32) iterations is recorded;
 *     those past convergence return at once).  The operator owns work
 *     vectors: one apply at a time.
 *  DEC_OP_LAPLACIAN0 (hodge selects ⋆1): fast_laplace_beltrami
 *     (test/Multigrid.jl:46-48) = ∇²(0, sd) form (DiscreteExteriorCalculus.jl:1887-1890).
 *  DEC_OP_AVG01: avg₀₁_mat (FastDEC.jl:874-892).
 *  DEC_OP_WEDGE_PD (0,1): dec_wedge_product_pd(0,1), apply(x=g primal 0-form,
 *     y=f dual 1-form) = (M g) .* f, M = wedge_pd_01_mat (FastDEC.jl:327-346, 379-382);
 *     (1,1): apply(x=f primal, y=g dual 1-form) = ∧11(f, ♭♯ g) (FastDEC.jl:390-394).
 *  DEC_OP_WEDGE_DP (1,0): dec_wedge_product_dp(1,0), apply(x=f dual 1-form,
 *     y=g primal 0-form) = f .* (M g) (FastDEC.jl:374-377); (1,1): apply(x=f
 *     dual, y=g primal 1-form) = ∧11(♭♯ f, g) (FastDEC.jl:396-400).
 *  DEC_OP_FLAT_SHARP: ♭♯_mat (DiscreteExteriorCalculus.jl:1985-1992), ne x ne;
 *     ♯ (LLSDDSharp) is the pseudo-inverse in each triangle's plane.
 *  DEC_OP_INTERIOR_PRODUCT_DD (1,1): interior_product_dd, apply(f, g) dual
 *     1-forms -> dual 0-form; (1,2): apply(f dual 1-form, g dual 2-form) ->
 *     dual 1-form (FastDEC.jl:758-786).
 *  DEC_OP_LIE_DD (1,1): ℒ_dd, apply(f, g) dual 1-forms (FastDEC.jl:788-803).
 *  DEC_OP_DUAL_LAPLACIAN n=0,1: Δᵈ(Val(n)) on a 2-D complex (FastDEC.jl:808-860).
 *  DEC_OP_P2_D2_INTERPOLATION: p2_d2_interpolation (DiscreteExteriorCalculus.jl:2016-2031).
 *  DEC_OP_P0_D0_INTERPOLATION: p0_d0_interpolation (FastDEC.jl:928-941).
 *  DEC_OP_D0_P0_INTERPOLATION: d0_p0_interpolation = ⋆0⁻¹ p2_d2 ⋆2⁻¹
 *     (FastDEC.jl:915-917); hodge is accepted, both stars are diagonal.
 *  The last three and the 1-1 wedges are compositions of other operators
 *  (and of the geometric ⋆1⁻¹ solve); they own work vectors.
 * On a 3-D complex (FastDEC.jl methods for HasDeltaSet3D):
 *  DEC_OP_WEDGE (0,1),(1,0),(0,2),(2,0),(1,1) as above on the triangle
 *     skeleton; (0,3),(3,0) (FastDEC.jl:64-76, 173-183) and (2,1),(1,2)
 *     (FastDEC.jl:96-133, 194-208, 281-289).
 *  DEC_OP_DIFFERENTIAL n=0,1,2; DEC_OP_BOUNDARY n=1,2,3;
 *  DEC_OP_DUAL_DERIVATIVE n=0,1,2: d̃0 = -∂3, d̃1 = ∂2, d̃2 = -∂1 (FastDEC.jl:455-462).
 *  DEC_OP_HODGE n=0..3 with DEC_DIAGONAL_HODGE (FastDEC.jl:606-636);
 *  DEC_OP_INV_HODGE n=0..3 diagonal, and geometric for n=0,3 (FastDEC.jl:750-757).
 *  DEC_OP_LAPLACIAN0 with DEC_DIAGONAL_HODGE: -⋆0⁻¹ d̃2 ⋆1 d0.
 *  DEC_OP_DUAL_LAPLACIAN n=0: ⋆3 d2 ⋆2⁻¹ d̃0 as one matrix (FastDEC.jl:832-839).
 *  DEC_OP_P3_D3_INTERPOLATION: p3_d3_interpolation (DiscreteExteriorCalculus.jl:2038-2053);
 *  DEC_OP_D0_P0_INTERPOLATION: ⋆0⁻¹ p3_d3 ⋆3⁻¹ (FastDEC.jl:924-926);
 *  DEC_OP_P0_D0_INTERPOLATION on the triangle skeleton as above.
 *  Other kinds return DEC_ERR_UNSUPPORTED.
 * deg_n is ignored for single-degree kinds. */
int dec_op_create(dec_mesh m, int32_t kind, int32_t deg_m, int32_t deg_n,
                  int32_t hodge, int32_t dtype, dec_op *out);
/* rows = output length; cols = length of the first input, cols2 of the
 * second input (0 if unary). */
int dec_op_shape(dec_op op, int64_t *rows, int64_t *cols, int64_t *cols2);
/* out = op(x[, y]).  Unary operators ignore y.  Same dtype as the operator. */
int dec_op_apply(dec_op op, const void *x, const void *y, void *out, void *stream);
/* For the geometric ⋆1⁻¹ solve: iterations and ‖x - ⋆1 z‖/‖x‖ of the last
 * apply on `stream` (synchronises it).  DEC_ERR_ARG for other operators.
 * No counterpart in the reference, whose solve is a direct factorisation. */
int dec_op_solve_info(dec_op op, int32_t *iterations, double *relres, void *stream);
/* Several applications in one launch: out[i] = ops[i](x[i], y[i]) for
 * i < count (y[i] ignored by one-form operators).  All operators must share
 * a dtype.  The items run concurrently, so no output may overlap another
 * item's input or output (DEC_ERR_ARG otherwise; a chain such as ⋆ then d
 * takes two calls).  With that, results are bit-identical to count
 * dec_op_apply calls; the single grid removes the per-launch ramp-up and
 * drain.  The geometric ⋆1⁻¹ solve and composite operators are refused
 * (DEC_ERR_UNSUPPORTED).  No counterpart in the
 * reference (each FastDEC operator is its own kernel, FastDEC.jl:255-289). */
int dec_op_apply_many(int32_t count, const dec_op *ops, const void *const *x, const void *const *y,
                      void *const *out, void *stream);
/* Sparse-matrix operators (d, ∂, d̃, ⋆ (geometric or diagonal), ⋆⁻¹, ∇²,
 * avg01, wedge_dd's averaging matrix): CSR export for parity checks, columns
 * ascending.  Call with NULL arrays to read nnz first. */
int dec_op_export_csr(dec_op op, int64_t *nnz, int32_t *rowptr, int32_t *col,
                      double *val);
/* Wedge coefficient tables as the kernels read them (SoA, float64 copies):
 * wedge 0-1: idx 2 x ne; wedge 0-2: idx 3 x nt, coeff 3 x nt (pairs of the
 * reference's six coefficients pre-summed); wedge 1-1: idx 3 x nt, coeff 3 x nt;
 * wedge 0-3: idx 4 x ntet (tetrahedron_vertices), coeff 4 x ntet (the
 * reference's 24 coefficients summed per vertex); wedge 2-1: idx 10 x ntet
 * (tetrahedron_edges, then tetrahedron_triangles), coeff 12 x ntet (the
 * reference's signed coefficients). */
int dec_op_export_tables(dec_op op, int32_t *idx, double *coeff);
int dec_op_destroy(dec_op op);

/* --------------------------------------------------------------- multigrid */
/* MultigridData(s, BinarySubdivision(), levels, op, steps; alg, mode=DirectMode())
 * (Multigrid.jl:698-702, 752-797) with op = fast_laplace_beltrami using the
 * given Hodge for ⋆1 and dual center.  Levels are stored finest first.
 * project=1 rescales every new point to the unit sphere (icosphere family). */
int dec_mg_create(dec_mesh base, int32_t levels, int32_t project, int32_t hodge,
                  int32_t center, int32_t steps, dec_mg *out);
/* The same with mode = DEC_MG_DIRECT (DirectMode(): every level dualized and
 * discretized) or DEC_MG_GALERKIN (GalerkinMode(): only the finest level is
 * dualized; A_{k+1} = R_k A_k P_k, Multigrid.jl:558-569, 740-748). */
enum { DEC_MG_DIRECT = 0, DEC_MG_GALERKIN = 1 };
int dec_mg_create_mode(dec_mesh base, int32_t levels, int32_t project, int32_t hodge,
                       int32_t center, int32_t steps, int32_t mode, dec_mg *out);
/* MultigridData(ops, restrictions, prolongations, steps) (Multigrid.jl:545-546,
 * MGData at :548-551): a hierarchy from the caller's host matrices, finest
 * first, as CSR arrays per level (index_base 0, or 1 for Julia's ids).
 * sizes[l] = n_l; A_l is n_l x n_l with a nonzero diagonal; R_l (level l ->
 * l+1) is n_{l+1} x n_l and P_l is n_l x n_{l+1} (l < nlevels - 1).  Row sums
 * follow each row's stored column order.  This also carries the reference's
 * generic `op` (MultigridData(s, scheme, levels, op, steps), :698-702): the
 * caller evaluates op on each level's dual complex and passes the matrices.
 * Gauss-Seidel colours are greedy; dec_mg_finest_mesh is unavailable. */
int dec_mg_create_from_csr(int32_t nlevels, const int32_t *sizes,
                           const int32_t *const *a_rowptr, const int32_t *const *a_col,
                           const double *const *a_val, const int32_t *const *r_rowptr,
                           const int32_t *const *r_col, const double *const *r_val,
                           const int32_t *const *p_rowptr, const int32_t *const *p_col,
                           const double *const *p_val, int32_t index_base, int32_t steps,
                           dec_mg *out);
int dec_mg_nlevels(dec_mg mg, int32_t *nlevels);
int dec_mg_level_size(dec_mg mg, int32_t level, int32_t *n);
/* The fine mesh of the hierarchy (owned by the hierarchy; do not destroy). */
int dec_mg_finest_mesh(dec_mg mg, dec_mesh *m);
/* CSR exports of A_l, R_l (level l -> l+1) and P_l = transpose(S_l). */
int dec_mg_export_level(dec_mg mg, int32_t level, int32_t which, int64_t *nnz,
                        int32_t *rowptr, int32_t *col, double *val);
/* The Gauss-Seidel colour of each vertex of level l (n_l entries) and the
 * number of colours.  No counterpart in the reference, whose smoother is
 * Krylov cg: the multicolour Gauss-Seidel smoother is this library's, and its
 * colouring is exported so that callers and tests can check it and reproduce
 * the sweep order.  Four colours where the base mesh admits them (found by
 * search), else greedy colours. */
int dec_mg_export_colors(dec_mg mg, int32_t level, int32_t *color, int32_t *ncolors);
/* The block of each vertex of level l (n_l entries, caller's numbering) for
 * the hybrid block Gauss-Seidel smoother, and the number of blocks.  No
 * counterpart in the reference (the smoother is this library's); exported so
 * that tests can reproduce the sweep.  Levels of at most 3000 rows are one
 * block. */
int dec_mg_export_blocks(dec_mg mg, int32_t level, int32_t *block, int32_t *nblocks);
/* Smoother per level: alg (cg = Krylov.jl's cg as in Multigrid.jl:860,
 * gauss_seidel = multicolour GS, jacobi = damped Jacobi, block_gs = hybrid
 * block Gauss-Seidel: multicolour GS inside blocks of up to 512 vertices
 * grown over the mesh, neighbours in other blocks taken from the start of
 * the sweep, one launch per sweep; blocks are exported by
 * dec_mg_export_blocks), steps (0 = skip,
 * Multigrid.jl:860, 871), omega (Jacobi damping).  coarse_direct=1 solves the
 * coarsest level with a dense float64 operator instead of smoothing it. */
int dec_mg_set_smoother(dec_mg mg, int32_t alg, int32_t steps, double omega,
                        int32_t coarse_direct);
/* multigrid_vcycles / multigrid_wcycles (Multigrid.jl:814-824): mu = 1 (V)
 * or 2 (W); u (in/out) and b are float64 device vectors of the finest size. */
int dec_mg_cycles(dec_mg mg, int32_t mu, int32_t cycles, void *u, const void *b,
                  void *stream);
/* full_multigrid (Multigrid.jl:830-855); dec_Δ⁻¹ (FastDEC.jl:866-869) is
 * full_multigrid with mu = 2, cycles = 5, steps = 3, alg = cg. */
int dec_mg_full_multigrid(dec_mg mg, int32_t mu, int32_t cycles, const void *b,
                          void *u, void *stream);
/* Solve A u = b to ||b - A u||_2 / ||b||_2 <= tol (checked in float64).
 * variant DEC_SOLVE_PLAIN_F64: float64 V-cycles; DEC_SOLVE_MIXED_IR: float32
 * V-cycles inside float64 iterative refinement; DEC_SOLVE_PCG_F64 /
 * DEC_SOLVE_PCG_MIXED: conjugate gradients in float64 with one V-cycle
 * (float64 / float32) from a zero start as the preconditioner, for operators
 * that are symmetric up to a positive diagonal scaling (the 0-form Laplacian
 * is); an iteration that cuts the residual by less than 0.9x hands the rest
 * to the V-cycles of the matching first two variants.  The CG variants do not
 * use the learned smoother.  No counterpart in the reference (its solves are
 * multigrid_vcycles / full_multigrid, Multigrid.jl:814-855).  Returns the
 * cycle (iteration) count and the final float64 relative residual of b - A u. */
int dec_mg_solve(dec_mg mg, int32_t variant, double tol, int32_t max_cycles,
                 void *u, const void *b, int32_t *cycles_done, double *relres,
                 void *stream);
int dec_mg_destroy(dec_mg mg);

/* ------------------------------------------------------- learned smoother */
/* A 3-layer edge message-passing network (hidden 32) used as the multigrid
 * smoother: c = N(r) for a residual r on a level of a hierarchy, with dense
 * layers on BF16 MFMA (FP32 accumulate).  No counterpart in the reference,
 * whose smoother is Krylov.jl's cg (Multigrid.jl:860).  Parameters are a flat
 * float vector (dec_nn_nparams entries); the untrained network (seeded
 * initialisation) is damped Jacobi.  r, c are float32 device arrays of
 * nsamples x n (samples back to back, vertex order).  precise=1 runs the
 * dense layers in FP32 on the VALU with FP32 activations (for checking). */
int dec_nn_create(dec_mg mg, uint64_t seed, dec_nn *out);
int dec_nn_nparams(dec_nn nn, int32_t *n);
int dec_nn_get_params(dec_nn nn, float *host);
int dec_nn_set_params(dec_nn nn, const float *host);
int dec_nn_forward(dec_nn nn, int32_t level, int32_t nsamples, const float *r, float *c,
                   int32_t precise, void *stream);
/* One smoothing application u += c(r) for a single residual r (n floats,
 * device) with the inference kernels the V-cycles use: BF16 MFMA dense
 * layers, hidden layers 1 and 3 kept in registers.  Bit-identical to
 * dec_nn_forward (precise = 0) followed by u += c. */
int dec_nn_smooth(dec_nn nn, int32_t level, const float *r, float *u, void *stream);
/* Loss and its gradient (host copies).  e == NULL: the residual loss
 * mean((r - A c)² / s²); e given (the errors A⁻¹r paired with r, same
 * layout): the error loss mean((κ H (e - c))² / s²), κ = mean(diag A),
 * H = I - Pr R with R the
 * restriction to the next coarser level and Pr the prolongation back (the
 * error left that the coarse grid cannot represent; not on the coarsest
 * level). */
int dec_nn_loss_grad(dec_nn nn, int32_t level, int32_t nsamples, const float *r, const float *e,
                     int32_t precise, double *loss, float *grad, void *stream);
/* Adam on minibatches drawn (seeded) from the npairs residuals in R
 * (npairs x n, device) and, if E is given, their errors (error loss, else
 * residual loss); loss_hist (host, steps entries) may be NULL. */
int dec_nn_train(dec_nn nn, int32_t level, int32_t npairs, const float *R, const float *E,
                 int32_t steps, int32_t batch, float lr, uint64_t seed, double *loss_hist,
                 void *stream);
int dec_nn_destroy(dec_nn nn);
/* Use the network as the smoother of the float32 V-cycles (DEC_SOLVE_MIXED_IR)
 * on the levels of min_rows to max_rows vertices (max_rows <= 0: no upper
 * bound; Gauss-Seidel elsewhere), `applications` times before and after the
 * coarse correction.  During dec_mg_solve, a cycle whose residual drops by
 * less than 0.5x switches the rest of the solve back to Gauss-Seidel;
 * dec_mg_learned_stats reports those fallbacks and the cycles run with the
 * network.  The network must come from dec_nn_create on this hierarchy. */
int dec_mg_set_learned(dec_mg mg, dec_nn nn, int32_t enable, int32_t min_rows, int32_t max_rows,
                       int32_t applications);
int dec_mg_learned_stats(dec_mg mg, int32_t *fallbacks, int32_t *learned_cycles);

/* ------------------------------------------------------------------ batches */
/* Many explicit heat-equation instances u_t = Δu on meshes of uneven size,
 * stepped together: u ← u - dt L u with L the 0-form Laplace-Beltrami
 * operator (fast_laplace_beltrami with the diagonal Hodge, test/Multigrid.jl:46-48)
 * and dt = 0.1 h², h the mesh's mean edge length.  No counterpart in the
 * reference, which applies its operators one mesh at a time
 * (FastDEC.jl:208-238).  meshes[i] is instance i's mesh (repeats share one
 * operator).  The state is one device vector holding the instances back to
 * back; dec_batch_layout returns the offsets (ninst + 1 entries) and steps. */
int dec_batch_create(int32_t ninst, const dec_mesh *meshes, int32_t dtype, dec_batch *out);
int dec_batch_layout(dec_batch b, int64_t *total, int64_t *offsets, double *dt);
/* Autotune the launch shape (block size, rows per thread, workgroups per CU)
 * on the batch itself, starting from the state u (left unchanged); the winner
 * is cached in the text file cache_path (NULL: no file) keyed by the batch shape. */
int dec_batch_tune(dec_batch b, void *u, const char *cache_path, int32_t *block,
                   int32_t *rows_per_thread, int32_t *grid_mult);
/* nsteps explicit Euler steps of every instance, one persistent launch per
 * step; use_graph=1 captures the nsteps launches in a hipGraph (kept for the
 * same state pointer and step count) and replays it. */
int dec_batch_heat_steps(dec_batch b, void *u, int32_t nsteps, int32_t use_graph, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the steps of dec_batch_heat_steps, same kernels and
 * results, with the ping-pong partner of u in a fixed 32 MiB device array of
 * the library instead of a library allocation, and no graph of its own
 * (callers may capture the call).  Calls no hipMalloc/hipFree.  States above
 * 32 MiB (b's total rows times 8 bytes in float64, 4 in float32) are rejected
 * with DEC_ERR_ARG.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other threads) race. */
int dec_batch_heat_steps_noalloc(dec_batch b, void *u, int32_t nsteps, void *stream);
int dec_batch_destroy(dec_batch b);

#ifdef __cplusplus
}
#endif
#endif
