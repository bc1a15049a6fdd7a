/*
 * dec_oracle.h -- CPU restatement of CombinatorialSpaces.jl's DEC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library is tested against this oracle:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * it.  The product path (libdec_hip.so) never links or calls it.
 *
 * Conventions: every index is 0-based int32 (the reference is 1-based; the
 * tests shift by one when comparing with the reference's known answers).
 * Orientations are signed char: 1 = true, 0 = false, -1 = unset.
 * Sparse matrices are CSC, as SparseArrays stores them, so that the oracle's
 * accumulation orders are the reference's.
 */
#ifndef DEC_ORACLE_H
#define DEC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int m, n, nnz;
  int *colptr;   /* n+1, 0-based offsets */
  int *rowval;   /* nnz */
  double *nzval; /* nnz */
} om_csc;

void om_csc_free(om_csc *A);

/* ---- primal meshes (src/SimplicialSets.jl, src/CombMeshes.jl) ---- */
int om_glue_triangles(int nv, int nt, const int *tri_verts, int sorted,
                      int *v0, int *v1, int *e0, int *e1, int *e2);
int om_orient2(int ne, int nt, const int *e0, const int *e1, const int *e2,
               signed char *tri_orient);
void om_icosahedron_points(double *p);
void om_icosahedron_faces(int *tv);
int om_grid_counts(double lx, double ly, double dx, double dy, int *nx, int *ny);
int om_triangulated_grid(double lx, double ly, double dx, double dy, int compress,
                         double *p, int *v0, int *v1, int *e0, int *e1, int *e2,
                         signed char *edge_orient, signed char *tri_orient);

/* ---- binary subdivision (src/Multigrid.jl) ---- */
void om_refine(int nv, int ne, int nt, const int *v0, const int *v1,
               const int *e0, const int *e1, const int *e2,
               int *fv0, int *fv1, int *fe0, int *fe1, int *fe2);
void om_propagate_points(int nv, int ne, const int *v0, const int *v1,
                         const double *p, double *pf, int project);
void om_subdivision_matrix(int nv, int ne, const int *v0, const int *v1, om_csc *S);
void om_make_restriction(const om_csc *S, om_csc *R);

/* ---- dual complex (src/DiscreteExteriorCalculus.jl) ---- */
void om_dual_complex(int nv, int ne, int nt, const int *v0, const int *v1,
                     const int *e0, const int *e1, const int *e2,
                     signed char *edge_orient, signed char *tri_orient,
                     int *D_v0, int *D_v1, int *D_e0, int *D_e1, int *D_e2,
                     signed char *D_edge_orient, signed char *D_tri_orient);
void om_dual_geometry(int nv, int ne, int nt, const int *v0, const int *v1,
                      const int *e0, const int *e1, const int *e2,
                      const int *D_v0, const int *D_v1,
                      const int *D_e0, const int *D_e1, const int *D_e2,
                      const double *p, int circumcenter,
                      double *dual_point, double *length, double *area,
                      double *dual_length, double *dual_area);

/* ---- FastDEC (src/FastDEC.jl) ---- */
void om_wedge02_coeffs(int nv, int ne, int nt, const int *D_v1, const int *D_e2,
                       const double *dual_area, const double *area,
                       int *verts6, double *coeffs6);
void om_wedge11_coeffs(int nt, const int *e0, const int *e1, const int *e2,
                       const signed char *eo, const signed char *to,
                       const double *dual_area, const double *area,
                       int *edges3, double *coeffs3);
void om_wedge01(int ne, const int *v0, const int *v1, const double *f,
                const double *a, double *res);
void om_wedge02(int nt, const int *verts6, const double *coeffs6,
                const double *f, const double *a, double *res);
void om_wedge11(int nt, const int *edges3, const double *coeffs3,
                const double *a, const double *b, double *res);
void om_wedge_pd01_mat(int nv, int ne, int nt, const int *v0, const int *v1, const int *e1,
                       const int *e2, const int *D_v0, const int *D_v1, const double *dl,
                       om_csc *M);
void om_wedge_dd01_mat(int nv, int ne, int nt, const int *D_v0, const int *D_v1,
                       const double *dual_length, om_csc *M);
void om_derivbound(int k, int nv, int ne, int nt, const int *v0, const int *v1,
                   const int *e0, const int *e1, const int *e2,
                   const signed char *eo, const signed char *to,
                   int transpose, int negate, om_csc *A);
void om_hodge_diag(int k, int nv, int ne, int nt, const int *D_v0, const int *D_v1,
                   const int *D_e1, const double *dual_length,
                   const double *dual_area, const double *length,
                   const double *area, double *h);
void om_inv_hodge_diag(int k, int n, const double *h, double *out);
void om_geo_hodge1(int nv, int ne, int nt, const int *v0, const int *v1,
                   const int *e0, const int *e1, const int *e2,
                   const double *dual_point, om_csc *H);

/* ---- generic CSC kernels (SparseArrays semantics) ---- */
void om_sparse(int m, int n, int nnz, const int *I, const int *J, const double *V,
               om_csc *A);
void om_spmv(const om_csc *A, const double *x, double *y);
void om_spmv_t(const om_csc *A, const double *x, double *y);
void om_spgemm(const om_csc *A, const om_csc *B, om_csc *C);
void om_diag_left(const double *d, const om_csc *A, om_csc *C);
void om_diag_right(const om_csc *A, const double *d, om_csc *C);
void om_scale(double a, const om_csc *A, om_csc *C);
void om_transpose(const om_csc *A, om_csc *C);
void om_laplacian0(const om_csc *d0, const double *h1, const om_csc *h1_geo,
                   const double *inv_h0, om_csc *L);

/* ---- multigrid (src/Multigrid.jl + project smoothers) ---- */
enum { OM_CG = 0, OM_GS = 1, OM_JACOBI = 2, OM_BLOCK_GS = 3 };
int om_greedy_colors(const om_csc *A, int *color);
void om_cg(const om_csc *A, const double *b, double *x, int itmax);
void om_coarse_inverse(const om_csc *A, double *Z);
void om_mu_cycles(int nlev, const om_csc *A, const om_csc *R, const om_csc *S,
                  const int *steps, int alg, int mu, int cycles, int coarse_direct,
                  double omega, double *u, const double *b, const int *colors,
                  const int *blocks);
void om_full_multigrid(int nlev, const om_csc *A, const om_csc *R, const om_csc *S,
                       const int *steps, int alg, int mu, int cycles,
                       int coarse_direct, double omega, const double *b, double *u,
                       const int *colors, const int *blocks);
/* colors: NULL, or every level's Gauss-Seidel colour array (finest first)
 * concatenated; NULL colours each level with om_greedy_colors.
 * blocks: for OM_BLOCK_GS, every level's block index per vertex (finest
 * first, concatenated; NULL: one block per level, which is multicolour GS). */

/* ---- inputs ---- */
void om_uniform_form(unsigned long long seed, int n, double *x);
double om_splitmix_u01(unsigned long long seed, unsigned long long i);

#ifdef __cplusplus
}
#endif
#endif
