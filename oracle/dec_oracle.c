/*
 * dec_oracle.c -- CPU restatement of CombinatorialSpaces.jl's DEC hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the restatement the HIP library is compared
 * against.  It is never linked into, called by, or shipped with the product.
 *
 * Every function restates one reference routine and cites it (paths are inside
 * the CombinatorialSpaces.jl v0.10.1 source tree).  Arithmetic is written in the
 * reference's evaluation order (no FMA contraction: build with
 * -ffp-contract=off) so that orders of summation are the reference's wherever
 * the reference fixes them (SparseArrays CSC SpMV/SpGEMM, left-to-right sums).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *  - primal/dual numbering and orientation: reference known-answer tests
 *    (test/SimplicialSets.jl:93-145, test/DiscreteExteriorCalculus.jl:183-233,
 *    test/Multigrid.jl:14-41, test/CombMeshes.jl:14-84);
 *  - geometry/Hodge: single-triangle closed forms
 *    (test/DiscreteExteriorCalculus.jl:288-371) and invariants;
 *  - wedges: avg01 identity, 1^a = a, antisymmetry, brute-force per-simplex
 *    formulas on small meshes;
 *  - Laplacian: null space, symmetry of d0' *1 d0, 1-D/2-D closed forms;
 *  - multigrid: residual thresholds of test/Multigrid.jl.
 *  Parity unpinned: Krylov.jl `cg` iterates (third-party, BLAS dot order) --
 *  om_cg follows Krylov.jl's published CG recurrence and stopping rule but is
 *  checked only through residual thresholds.
 */
#include "dec_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define XALLOC(T, n) ((T *)calloc((size_t)((n) > 0 ? (n) : 1), sizeof(T)))

void om_csc_free(om_csc *A) {
  if (!A) return;
  free(A->colptr);
  free(A->rowval);
  free(A->nzval);
  A->colptr = NULL;
  A->rowval = NULL;
  A->nzval = NULL;
  A->nnz = 0;
}

static void csc_alloc(om_csc *A, int m, int n, int nnz) {
  A->m = m;
  A->n = n;
  A->nnz = nnz;
  A->colptr = XALLOC(int, n + 1);
  A->rowval = XALLOC(int, nnz);
  A->nzval = XALLOC(double, nnz);
}

/* ------------------------------------------------------------------------ */
/* Inputs: counter-based splitmix64 stream, x = 2u - 1, u in [0,1).          */
/* ------------------------------------------------------------------------ */
static unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

double om_splitmix_u01(unsigned long long seed, unsigned long long i) {
  return (double)(splitmix64((seed << 32) + i) >> 11) * 0x1.0p-53;
}

void om_uniform_form(unsigned long long seed, int n, double *x) {
  for (int i = 0; i < n; i++) x[i] = 2.0 * om_splitmix_u01(seed, (unsigned long long)i) - 1.0;
}

/* ------------------------------------------------------------------------ */
/* Primal meshes                                                             */
/* ------------------------------------------------------------------------ */

/* glue_triangle!/get_edge!/add_triangle!/glue_sorted_triangle!
 * (src/SimplicialSets.jl:323-364).  get_edge!(src,tgt) reuses the first edge
 * (ascending id) with ∂v1 == src and ∂v0 == tgt, else appends one; edges are
 * created in the order (v0,v1), (v1,v2), (v0,v2); the triangle stores
 * ∂e0 = e(v1,v2), ∂e1 = e(v0,v2), ∂e2 = e(v0,v1).  Returns ne. */
int om_glue_triangles(int nv, int nt, const int *tv, int sorted, int *v0, int *v1,
                      int *e0, int *e1, int *e2) {
  /* per-source adjacency: head[src] -> edge, next[edge] */
  int *head = XALLOC(int, nv);
  int *tail = XALLOC(int, nv);
  int *next = XALLOC(int, 3 * nt);
  for (int i = 0; i < nv; i++) head[i] = tail[i] = -1;
  int ne = 0;
  for (int t = 0; t < nt; t++) {
    int a = tv[3 * t], b = tv[3 * t + 1], c = tv[3 * t + 2];
    if (sorted) { /* sort(SVector(v0,v1,v2)) */
      int x;
      if (a > b) { x = a; a = b; b = x; }
      if (b > c) { x = b; b = c; c = x; }
      if (a > b) { x = a; a = b; b = x; }
    }
    int src[3] = {a, b, a}, tgt[3] = {b, c, c}, eid[3];
    for (int k = 0; k < 3; k++) {
      int found = -1;
      for (int e = head[src[k]]; e >= 0; e = next[e])
        if (v0[e] == tgt[k]) { found = e; break; }
      if (found < 0) {
        found = ne++;
        v1[found] = src[k];
        v0[found] = tgt[k];
        next[found] = -1;
        if (tail[src[k]] < 0) head[src[k]] = found; else next[tail[src[k]]] = found;
        tail[src[k]] = found;
      }
      eid[k] = found;
    }
    e0[t] = eid[1];
    e1[t] = eid[2];
    e2[t] = eid[0];
  }
  free(head);
  free(tail);
  free(next);
  return ne;
}

/* orient!(s, Val(2)) (src/SimplicialSets.jl:812-852): depth-first search from
 * the lowest unvisited triangle (seed orientation +1 == true); the neighbour y
 * across face i of x and face j of y gets -o_x when i+j is even, else o_x.
 * cofaces are visited in ascending triangle order.  Returns 1 on success,
 * 0 (orientations untouched) if the complex is not orientable. */
int om_orient2(int ne, int nt, const int *e0, const int *e1, const int *e2,
               signed char *to) {
  if (nt == 0) return 1;
  const int *E[3] = {e0, e1, e2};
  /* coface lists per (face index j, edge): CSR over edges, ascending t */
  int *cnt[3], *ptr[3], *lst[3];
  for (int j = 0; j < 3; j++) {
    cnt[j] = XALLOC(int, ne + 1);
    ptr[j] = XALLOC(int, ne + 1);
    lst[j] = XALLOC(int, nt);
    for (int t = 0; t < nt; t++) cnt[j][E[j][t] + 1]++;
    for (int e = 0; e < ne; e++) cnt[j][e + 1] += cnt[j][e];
    memcpy(ptr[j], cnt[j], sizeof(int) * (ne + 1));
    for (int t = 0; t < nt; t++) lst[j][ptr[j][E[j][t]]++] = t;
  }
  signed char *ors = XALLOC(signed char, nt);
  int *stack = XALLOC(int, nt + 1);
  int ok = 1;
  for (int seed = 0; seed < nt && ok; seed++) {
    if (ors[seed] != 0) continue;
    int sp = 0;
    stack[sp++] = seed;
    ors[seed] = 1;
    while (sp > 0 && ok) {
      int x = stack[--sp];
      int ox = ors[x], nox = -ox;
      for (int i = 0; i < 3 && ok; i++) {
        int face = E[i][x];
        for (int j = 0; j < 3 && ok; j++) {
          int same = ((i + j) % 2) == 0;
          for (int k = cnt[j][face]; k < cnt[j][face + 1]; k++) {
            int y = lst[j][k];
            if (y == x) continue;
            int oy = ors[y];
            if (oy == 0) {
              ors[y] = (signed char)(same ? nox : ox);
              stack[sp++] = y;
            } else if (same && oy == ox) {
              ok = 0;
              break;
            }
          }
        }
      }
    }
  }
  if (ok)
    for (int t = 0; t < nt; t++) to[t] = ors[t] == 1 ? 1 : 0;
  for (int j = 0; j < 3; j++) {
    free(cnt[j]);
    free(ptr[j]);
    free(lst[j]);
  }
  free(ors);
  free(stack);
  return ok;
}

/* Pole-aligned icosahedron (vertex on each pole, rings at z = ±1/sqrt 5 offset
 * by 36 degrees), the base of the icosphere family.  Coordinates are fixed
 * literals so every implementation starts from identical bits. */
static const double ICO_P[12][3] = {
    {0.0, 0.0, 1.0},
    {0.8944271909999159, 0.0, 0.4472135954999579},
    {0.27639320225002106, 0.8506508083520399, 0.4472135954999579},
    {-0.7236067977499788, 0.5257311121191337, 0.4472135954999579},
    {-0.723606797749979, -0.5257311121191335, 0.4472135954999579},
    {0.27639320225002084, -0.85065080835204, 0.4472135954999579},
    {0.7236067977499789, 0.5257311121191336, -0.4472135954999579},
    {-0.27639320225002095, 0.85065080835204, -0.4472135954999579},
    {-0.8944271909999159, 0.0, -0.4472135954999579},
    {-0.2763932022500211, -0.8506508083520399, -0.4472135954999579},
    {0.7236067977499788, -0.5257311121191338, -0.4472135954999579},
    {0.0, 0.0, -1.0},
};

void om_icosahedron_points(double *p) { memcpy(p, ICO_P, sizeof(ICO_P)); }

/* 20 faces, glued in this order with glue_sorted_triangle!: top cap, upper
 * band, lower band, bottom cap. */
void om_icosahedron_faces(int *tv) {
  int t = 0;
  for (int k = 0; k < 5; k++) { tv[3*t] = 0; tv[3*t+1] = 1 + k; tv[3*t+2] = 1 + (k + 1) % 5; t++; }
  for (int k = 0; k < 5; k++) { tv[3*t] = 1 + k; tv[3*t+1] = 6 + k; tv[3*t+2] = 1 + (k + 1) % 5; t++; }
  for (int k = 0; k < 5; k++) { tv[3*t] = 1 + (k + 1) % 5; tv[3*t+1] = 6 + k; tv[3*t+2] = 6 + (k + 1) % 5; t++; }
  for (int k = 0; k < 5; k++) { tv[3*t] = 11; tv[3*t+1] = 6 + k; tv[3*t+2] = 6 + (k + 1) % 5; t++; }
}

/* length(0:d:l) for a Float64 range: floor(l/d) + 1 (with the range's
 * tolerance for a stop that lands on a grid point). */
static int range_len(double l, double d) {
  double q = l / d;
  double r = floor(q + 0.5);
  if (fabs(q - r) <= 1e-10 * (r > 1 ? r : 1)) return (int)r + 1;
  return (int)floor(q) + 1;
}

int om_grid_counts(double lx, double ly, double dx, double dy, int *nx, int *ny) {
  *nx = range_len(lx, dx);
  *ny = range_len(ly, dy);
  return (*nx) * (*ny);
}

/* triangulated_grid(lx, ly, dx, dy, Point3d, compress) (src/CombMeshes.jl:79-130).
 * Returns ne.  p must hold nx*ny*3, edges 3*ntri, ntri = 2(nx-1)(ny-1). */
int om_triangulated_grid(double lx, double ly, double dx, double dy, int compress,
                         double *p, int *v0, int *v1, int *e0, int *e1, int *e2,
                         signed char *eo, signed char *to) {
  int nx, ny;
  om_grid_counts(lx, ly, dx, dy, &nx, &ny);
  double scale = lx / (lx + dx / 2);
  double shift = dx / 2;
  for (int yi = 1; yi <= ny; yi++) {
    double raw_y = (double)(yi - 1) * dy;
    for (int xi = 1; xi <= nx; xi++) {
      double raw_x = (double)(xi - 1) * dx;
      double xc = raw_x + shift * (double)((yi % 2) == 0);
      if (compress) xc *= scale;
      int i = (xi - 1) + nx * (yi - 1);
      p[3 * i] = xc;
      p[3 * i + 1] = raw_y;
      p[3 * i + 2] = 0.0;
    }
  }
  int ntri = 2 * (nx - 1) * (ny - 1);
  int *tv = XALLOC(int, 3 * ntri);
  int t = 0;
  for (int y = 1; y <= ny - 1; y++)
    for (int x = 1; x <= nx - 1; x++) {
      int i = x + nx * (y - 1); /* 1-based */
      if (y % 2 == 0) {
        tv[3*t] = i - 1; tv[3*t+1] = i + nx - 1; tv[3*t+2] = i + nx; t++;
        tv[3*t] = i - 1; tv[3*t+1] = i; tv[3*t+2] = i + nx; t++;
      } else {
        tv[3*t] = i - 1; tv[3*t+1] = i; tv[3*t+2] = i + nx - 1; t++;
        tv[3*t] = i; tv[3*t+1] = i + nx - 1; tv[3*t+2] = i + nx; t++;
      }
    }
  int ne = om_glue_triangles(nx * ny, ntri, tv, 0, v0, v1, e0, e1, e2);
  free(tv);
  for (int e = 0; e < ne; e++) eo[e] = 0;
  int nxtri = 2 * (nx - 1), nytri = ny - 1;
  for (int y = 1; y <= nytri; y++) {
    int o = (y % 2) != 0;
    for (int x = 1; x <= nxtri; x += 2) {
      int i = x + nxtri * (y - 1) - 1;
      to[i] = (signed char)o;
      to[i + 1] = (signed char)!o;
    }
  }
  return ne;
}

/* ------------------------------------------------------------------------ */
/* Binary subdivision (src/Multigrid.jl:105-219, 595-612)                    */
/* ------------------------------------------------------------------------ */
void om_refine(int nv, int ne, int nt, const int *v0, const int *v1,
               const int *e0, const int *e1, const int *e2,
               int *fv0, int *fv1, int *fe0, int *fe1, int *fe2) {
  /* 0-based restatement of refine(::BinarySubdivision, topo), lines 130-183.
   * Coarse edge e (0-based) -> fine edges 2e (2e-1 in 1-based: ∂v1=∂v0_c)
   * and 2e+1 (∂v1 = ∂v1_c), both with ∂v0 = nv+e. */
  for (int e = 0; e < ne; e++) {
    int m = nv + e;
    fv0[2 * e] = m;     fv1[2 * e] = v0[e];
    fv0[2 * e + 1] = m; fv1[2 * e + 1] = v1[e];
  }
  for (int t = 0; t < nt; t++) {
    int m0 = e0[t] + nv, m1 = e1[t] + nv, m2 = e2[t] + nv;
    int m0_m1 = 2 * ne + 3 * t, m1_m2 = m0_m1 + 1, m0_m2 = m0_m1 + 2;
    fv0[m0_m1] = m1; fv1[m0_m1] = m0;
    fv0[m1_m2] = m2; fv1[m1_m2] = m1;
    fv0[m0_m2] = m2; fv1[m0_m2] = m0;
    int m0_v1 = 2 * e0[t] + 1, v2_m0 = 2 * e0[t];
    int m1_v0 = 2 * e1[t] + 1, v2_m1 = 2 * e1[t];
    int m2_v0 = 2 * e2[t] + 1, v1_m2 = 2 * e2[t];
    int b = 4 * t;
    fe0[b] = m1_m2;     fe1[b] = m0_m2;     fe2[b] = m0_m1;
    fe0[b + 1] = m1_m2; fe1[b + 1] = m2_v0; fe2[b + 1] = m1_v0;
    fe0[b + 2] = m0_m2; fe1[b + 2] = v1_m2; fe2[b + 2] = m0_v1;
    fe0[b + 3] = m0_m1; fe1[b + 3] = v2_m1; fe2[b + 3] = v2_m0;
  }
}

/* propagate_points(::BinarySubdivision) (src/Multigrid.jl:111-122): midpoints
 * (p[∂v0] + p[∂v1]) / 2.  project=1 additionally rescales each new point to
 * the unit sphere (icosphere family; not part of the reference routine). */
void om_propagate_points(int nv, int ne, const int *v0, const int *v1,
                         const double *p, double *pf, int project) {
  memcpy(pf, p, sizeof(double) * 3 * (size_t)nv);
  for (int e = 0; e < ne; e++) {
    double *q = pf + 3 * (size_t)(nv + e);
    for (int c = 0; c < 3; c++) q[c] = (p[3 * v0[e] + c] + p[3 * v1[e] + c]) / 2;
    if (project) {
      double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
      q[0] = q[0] / n; q[1] = q[1] / n; q[2] = q[2] / n;
    }
  }
}

/* subdivision_matrix(::BinarySubdivision) (src/Multigrid.jl:191-219):
 * nv_c x nv_f CSC; identity block, then column nv_c+e holds rows
 * (min, max) of (∂v0, ∂v1)[e] with 0.5 each. */
void om_subdivision_matrix(int nv, int ne, const int *v0, const int *v1, om_csc *S) {
  int nvf = nv + ne, nnz = nv + 2 * ne;
  csc_alloc(S, nv, nvf, nnz);
  for (int i = 0; i < nv; i++) {
    S->colptr[i] = i;
    S->rowval[i] = i;
    S->nzval[i] = 1.0;
  }
  for (int e = 0; e < ne; e++) {
    int k = nv + 2 * e;
    S->colptr[nv + e] = k;
    int r0 = v0[e], r1 = v1[e];
    if (r0 > r1) { int x = r0; r0 = r1; r1 = x; }
    S->rowval[k] = r0; S->nzval[k] = 0.5;
    S->rowval[k + 1] = r1; S->nzval[k + 1] = 0.5;
  }
  S->colptr[nvf] = nnz;
}

/* make_restriction (src/Multigrid.jl:595-612): same pattern as S, each value
 * divided by its row sum (row sums accumulated in storage order). */
void om_make_restriction(const om_csc *S, om_csc *R) {
  csc_alloc(R, S->m, S->n, S->nnz);
  memcpy(R->colptr, S->colptr, sizeof(int) * (S->n + 1));
  memcpy(R->rowval, S->rowval, sizeof(int) * S->nnz);
  double *rs = XALLOC(double, S->m);
  for (int k = 0; k < S->nnz; k++) rs[S->rowval[k]] += S->nzval[k];
  for (int k = 0; k < S->nnz; k++) R->nzval[k] = S->nzval[k] / rs[S->rowval[k]];
  free(rs);
}

/* ------------------------------------------------------------------------ */
/* Dual complex (src/DiscreteExteriorCalculus.jl:301-328, 545-611)           */
/* ------------------------------------------------------------------------ */
/* DualV: vertex centers 0..nv-1, edge centers nv..nv+ne-1, triangle centers
 * nv+ne..  DualE: A0 [0,ne), A1 [ne,2ne), B_j [2ne + j nt, ...),
 * C_j [2ne + 3nt + j nt, ...).  DualTri k (0-based) at k*nt + t with schema
 * (v,e,ev) = (0,1,1),(0,2,1),(1,2,0),(1,0,1),(2,0,0),(2,1,0).
 * Unset edge orientations become false; unset triangle orientations are set
 * by orient! (the arrays passed in are updated in place). */
static const int DT_SCHEMA[6][3] = {{0, 1, 1}, {0, 2, 1}, {1, 2, 0}, {1, 0, 1}, {2, 0, 0}, {2, 1, 0}};

void om_dual_complex(int nv, int ne, int nt, const int *v0, const int *v1,
                     const int *e0, const int *e1, const int *e2,
                     signed char *eo, signed char *to,
                     int *D_v0, int *D_v1, int *D_e0, int *D_e1, int *D_e2,
                     signed char *Deo, signed char *Dto) {
  const int *E[3] = {e0, e1, e2};
  const int *Vv[2] = {v0, v1};
  int ecen = nv, tcen = nv + ne;
  for (int e = 0; e < ne; e++) {
    D_v0[e] = ecen + e;      D_v1[e] = v0[e];
    D_v0[ne + e] = ecen + e; D_v1[ne + e] = v1[e];
  }
  for (int j = 0; j < 3; j++)
    for (int t = 0; t < nt; t++) {
      int de = 2 * ne + j * nt + t;
      D_v0[de] = tcen + t;
      D_v1[de] = ecen + E[j][t];
    }
  for (int j = 0; j < 3; j++)
    for (int t = 0; t < nt; t++) {
      int tvs[3] = {v1[e1[t]], v0[e2[t]], v0[e1[t]]};
      int de = 2 * ne + 3 * nt + j * nt + t;
      D_v0[de] = tcen + t;
      D_v1[de] = tvs[j];
    }
  for (int k = 0; k < 6; k++) {
    int v = DT_SCHEMA[k][0], e = DT_SCHEMA[k][1], ev = DT_SCHEMA[k][2];
    for (int t = 0; t < nt; t++) {
      int dt = k * nt + t;
      D_e0[dt] = 2 * ne + e * nt + t;
      D_e1[dt] = 2 * ne + 3 * nt + v * nt + t;
      D_e2[dt] = ev * ne + E[e][t];
    }
  }
  (void)Vv;
  /* orientation: make_dual_simplices_1d! (311-324), make_dual_simplices_2d! (573-604) */
  for (int e = 0; e < ne; e++)
    if (eo[e] < 0) eo[e] = 0;
  int unset = 0;
  for (int t = 0; t < nt; t++)
    if (to[t] < 0) unset = 1;
  if (unset) om_orient2(ne, nt, e0, e1, e2, to);
  for (int e = 0; e < ne; e++) {
    Deo[e] = (signed char)!eo[e];
    Deo[ne + e] = eo[e];
  }
  for (int k = 0; k < 6; k++)
    for (int t = 0; t < nt; t++)
      Dto[k * nt + t] = (signed char)((k % 2 == 0) ? !to[t] : to[t]);
  for (int j = 0; j < 3; j++)
    for (int t = 0; t < nt; t++) {
      int x = eo[E[j][t]], y = (j % 2) ? !to[t] : to[t];
      Deo[2 * ne + j * nt + t] = (signed char)((x && y) || (!x && !y)); /* relative_sign(::Bool,::Bool), :611 */
    }
  for (int j = 0; j < 3; j++)
    for (int t = 0; t < nt; t++) Deo[2 * ne + 3 * nt + j * nt + t] = 1;
}

/* sqdistance (src/SimplicialSets.jl:945): sum((x-y).^2), left to right. */
static double sqdist(const double *x, const double *y) {
  double a = x[0] - y[0], b = x[1] - y[1], c = x[2] - y[2];
  return (a * a + b * b) + c * c;
}

/* volume(points) (src/SimplicialSets.jl:896-903) for n = 1, 2. */
static double vol1(const double *a, const double *b) { return sqrt(sqdist(a, b)); }
static double vol2(const double *p1, const double *p2, const double *p3) {
  double u[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  double w[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
  double uu = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
  double ww = (w[0] * w[0] + w[1] * w[1]) + w[2] * w[2];
  double uw = (u[0] * w[0] + u[1] * w[1]) + u[2] * w[2];
  return sqrt(fabs(uu * ww - uw * uw)) / 2;
}

/* geometric_center(points, Circumcenter()) (src/DiscreteExteriorCalculus.jl:124-135):
 * barycentric coordinates = first row of inv(cayley_menger(points...)),
 * columns 2..end; the inverse is taken here by Gauss-Jordan elimination with
 * partial pivoting (an independent route to the closed form the HIP library
 * uses). */
static void circumcenter(int np, const double *const *pts, double *out) {
  int n = np + 1;
  double M[4][8];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) M[i][j] = 0.0;
  for (int i = 1; i < n; i++) { M[0][i] = 1.0; M[i][0] = 1.0; }
  for (int i = 1; i < n; i++)
    for (int j = 1; j < n; j++) M[i][j] = (i == j) ? 0.0 : sqdist(pts[i - 1], pts[j - 1]);
  for (int i = 0; i < n; i++) M[i][n + i] = 1.0;
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
    if (piv != c)
      for (int j = 0; j < 2 * n; j++) { double x = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = x; }
    double d = M[c][c];
    for (int j = 0; j < 2 * n; j++) M[c][j] /= d;
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      double f = M[r][c];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * n; j++) M[r][j] -= f * M[c][j];
    }
  }
  out[0] = out[1] = out[2] = 0.0;
  for (int k = 0; k < np; k++) {
    double lam = M[0][n + 1 + k];
    for (int c = 0; c < 3; c++) out[c] = out[c] + lam * pts[k][c];
  }
}

/* geometric_center(points, Barycenter()) (:100-102): sum(points)/length. */
static void barycenter(int np, const double *const *pts, double *out) {
  for (int c = 0; c < 3; c++) {
    double s = pts[0][c];
    for (int k = 1; k < np; k++) s = s + pts[k][c];
    out[c] = s / np;
  }
}

/* subdivide_duals_2d! + precompute_volumes_2d! (src/DiscreteExteriorCalculus.jl:941-990). */
void om_dual_geometry(int nv, int ne, int nt, const int *v0, const int *v1,
                      const int *e0, const int *e1, const int *e2,
                      const int *D_v0, const int *D_v1,
                      const int *D_e0, const int *D_e1, const int *D_e2,
                      const double *p, int circ, double *dp, double *length,
                      double *area, double *dual_length, double *dual_area) {
  (void)e0;
  (void)D_e2;
  int nde = 2 * ne + 6 * nt, ndt = 6 * nt;
  memcpy(dp, p, sizeof(double) * 3 * (size_t)nv);
  for (int e = 0; e < ne; e++) {
    const double *pts[2] = {p + 3 * v0[e], p + 3 * v1[e]}; /* edge_vertices = (∂v0, ∂v1) */
    if (circ) {
      /* For two points the first row of inv(cayley_menger) is (-d/2, d/(2d),
       * d/(2d)) = (-d/2, 1/2, 1/2) exactly (cofactor form of the 3x3 inverse),
       * so the center is 0.5 p0 + 0.5 p1. */
      double *q = dp + 3 * (size_t)(nv + e);
      for (int c = 0; c < 3; c++) q[c] = 0.5 * pts[0][c] + 0.5 * pts[1][c];
    } else {
      barycenter(2, pts, dp + 3 * (size_t)(nv + e));
    }
  }
  for (int t = 0; t < nt; t++) {
    const double *pts[3] = {p + 3 * v1[e1[t]], p + 3 * v0[e2[t]], p + 3 * v0[e1[t]]};
    if (circ) circumcenter(3, pts, dp + 3 * (size_t)(nv + ne + t));
    else barycenter(3, pts, dp + 3 * (size_t)(nv + ne + t));
  }
  for (int e = 0; e < ne; e++) length[e] = vol1(p + 3 * v0[e], p + 3 * v1[e]);
  for (int de = 0; de < nde; de++) dual_length[de] = vol1(dp + 3 * D_v0[de], dp + 3 * D_v1[de]);
  for (int t = 0; t < nt; t++)
    area[t] = vol2(p + 3 * v1[e1[t]], p + 3 * v0[e2[t]], p + 3 * v0[e1[t]]);
  for (int dt = 0; dt < ndt; dt++) {
    /* dual_triangle_vertices (:224-228): (D_∂v1[D_∂e1], D_∂v1[D_∂e0], D_∂v0[D_∂e0]) */
    int a = D_v1[D_e1[dt]], b = D_v1[D_e0[dt]], c = D_v0[D_e0[dt]];
    dual_area[dt] = vol2(dp + 3 * a, dp + 3 * b, dp + 3 * c);
  }
}

/* ------------------------------------------------------------------------ */
/* FastDEC (src/FastDEC.jl)                                                  */
/* ------------------------------------------------------------------------ */
static double nsign(signed char o) { return o ? 1.0 : -1.0; } /* numeric_sign, SimplicialSets.jl:763 */

/* wedge_kernel_coeffs(Val(0), Val(2)) (FastDEC.jl:50-62). */
void om_wedge02_coeffs(int nv, int ne, int nt, const int *D_v1, const int *D_e2,
                       const double *dual_area, const double *area, int *verts,
                       double *coeffs) {
  (void)nv;
  (void)ne;
  for (int t = 0; t < nt; t++)
    for (int k = 0; k < 6; k++) {
      int dtr = t + k * nt;
      verts[6 * t + k] = D_v1[D_e2[dtr]];
      coeffs[6 * t + k] = dual_area[dtr] / area[t];
    }
}

/* wedge_kernel_coeffs(Val(1), Val(1)) (FastDEC.jl:78-93). */
void om_wedge11_coeffs(int nt, const int *e0, const int *e1, const int *e2,
                       const signed char *eo, const signed char *to,
                       const double *da, const double *area, int *edges3,
                       double *c) {
  for (int i = 0; i < nt; i++) {
    double ar = (nsign(to[i]) * 2) * area[i];
    int a = e0[i], b = e1[i], d = e2[i];
    edges3[3 * i] = a; edges3[3 * i + 1] = b; edges3[3 * i + 2] = d;
    c[3 * i]     = ((nsign(eo[b]) * nsign(eo[d])) * (da[i] + da[i + nt])) / ar;
    c[3 * i + 1] = ((nsign(eo[a]) * nsign(eo[d])) * (da[i + 2 * nt] + da[i + 3 * nt])) / ar;
    c[3 * i + 2] = ((nsign(eo[a]) * nsign(eo[b])) * (da[i + 4 * nt] + da[i + 5 * nt])) / ar;
  }
}

/* wedge_kernel_01! (FastDEC.jl:159-163): res = half*α*(f[p1]+f[p2]); f is the
 * 0-form, a the 1-form. */
void om_wedge01(int ne, const int *v0, const int *v1, const double *f,
                const double *a, double *res) {
  for (int e = 0; e < ne; e++) res[e] = (0.5 * a[e]) * (f[v0[e]] + f[v1[e]]);
}

/* wedge_kernel_02! (FastDEC.jl:165-170). */
void om_wedge02(int nt, const int *p, const double *c, const double *f,
                const double *a, double *res) {
  for (int i = 0; i < nt; i++) {
    const int *q = p + 6 * i;
    const double *k = c + 6 * i;
    double s = k[0] * f[q[0]] + k[1] * f[q[1]];
    s = s + k[2] * f[q[2]];
    s = s + k[3] * f[q[3]];
    s = s + k[4] * f[q[4]];
    s = s + k[5] * f[q[5]];
    res[i] = a[i] * s;
  }
}

/* wedge_kernel_11! (FastDEC.jl:184-191). */
void om_wedge11(int nt, const int *e, const double *c, const double *a,
                const double *b, double *res) {
  for (int i = 0; i < nt; i++) {
    int i0 = e[3 * i], i1 = e[3 * i + 1], i2 = e[3 * i + 2];
    double ae0 = a[i0], ae1 = a[i1], ae2 = a[i2];
    double be0 = b[i0], be1 = b[i1], be2 = b[i2];
    double x = c[3 * i] * (ae2 * be1 - ae1 * be2);
    double y = c[3 * i + 1] * (ae2 * be0 - ae0 * be2);
    double z = c[3 * i + 2] * (ae1 * be0 - ae0 * be1);
    res[i] = (x + y) + z;
  }
}

/* ------------------------------------------------------------------------ */
/* Generic CSC (SparseArrays semantics)                                      */
/* ------------------------------------------------------------------------ */

/* sparse(I, J, V, m, n): duplicates summed in input order; rows sorted within
 * each column (stable counting sorts, as SparseArrays.sparse! does). */
void om_sparse(int m, int n, int nnz, const int *I, const int *J, const double *V,
               om_csc *A) {
  /* pass 1: sort by row (stable), pass 2: by column (stable) => (col, row, input order) */
  int *rc = XALLOC(int, m + 1), *ord1 = XALLOC(int, nnz);
  for (int k = 0; k < nnz; k++) rc[I[k] + 1]++;
  for (int i = 0; i < m; i++) rc[i + 1] += rc[i];
  for (int k = 0; k < nnz; k++) ord1[rc[I[k]]++] = k;
  int *cc = XALLOC(int, n + 1), *ord2 = XALLOC(int, nnz);
  for (int k = 0; k < nnz; k++) cc[J[k] + 1]++;
  for (int j = 0; j < n; j++) cc[j + 1] += cc[j];
  int *cptr = XALLOC(int, n + 1);
  memcpy(cptr, cc, sizeof(int) * (n + 1));
  for (int q = 0; q < nnz; q++) { int k = ord1[q]; ord2[cptr[J[k]]++] = k; }
  /* within equal (col,row), the entries appear in input order because both
   * sorts are stable and pass 1 preserved input order within a row. */
  A->m = m; A->n = n;
  A->colptr = XALLOC(int, n + 1);
  A->rowval = XALLOC(int, nnz);
  A->nzval = XALLOC(double, nnz);
  int w = 0;
  for (int j = 0; j < n; j++) {
    A->colptr[j] = w;
    int last_row = -1;
    for (int q = cc[j]; q < cc[j + 1]; q++) {
      int k = ord2[q];
      if (I[k] == last_row) {
        A->nzval[w - 1] = A->nzval[w - 1] + V[k];
      } else {
        A->rowval[w] = I[k];
        A->nzval[w] = V[k];
        last_row = I[k];
        w++;
      }
    }
  }
  A->colptr[n] = w;
  A->nnz = w;
  free(rc); free(ord1); free(cc); free(ord2); free(cptr);
}

/* A*x for CSC A (SparseArrays spmatmul): y = 0; for col asc: y[row] += a*x[col]. */
void om_spmv(const om_csc *A, const double *x, double *y) {
  for (int i = 0; i < A->m; i++) y[i] = 0.0;
  for (int j = 0; j < A->n; j++) {
    double xj = x[j];
    for (int k = A->colptr[j]; k < A->colptr[j + 1]; k++) y[A->rowval[k]] += A->nzval[k] * xj;
  }
}

/* transpose(A)*x: y[col] = sum over the column's entries in storage order. */
void om_spmv_t(const om_csc *A, const double *x, double *y) {
  for (int j = 0; j < A->n; j++) {
    double s = 0.0;
    for (int k = A->colptr[j]; k < A->colptr[j + 1]; k++) s += A->nzval[k] * x[A->rowval[k]];
    y[j] = s;
  }
}

/* C = A*B, Gustavson over the columns of B: C[:,j] = sum_k A[:,k] B[k,j] with
 * k in B's storage order; rows of each output column sorted ascending. */
void om_spgemm(const om_csc *A, const om_csc *B, om_csc *C) {
  int m = A->m, n = B->n;
  double *acc = XALLOC(double, m);
  int *mark = XALLOC(int, m), *rows = XALLOC(int, m);
  for (int i = 0; i < m; i++) mark[i] = -1;
  int cap = A->nnz + B->nnz + 16, nnz = 0;
  int *rv = XALLOC(int, cap);
  double *nz = XALLOC(double, cap);
  int *cp = XALLOC(int, n + 1);
  for (int j = 0; j < n; j++) {
    cp[j] = nnz;
    int nr = 0;
    for (int kb = B->colptr[j]; kb < B->colptr[j + 1]; kb++) {
      int k = B->rowval[kb];
      double b = B->nzval[kb];
      for (int ka = A->colptr[k]; ka < A->colptr[k + 1]; ka++) {
        int i = A->rowval[ka];
        if (mark[i] != j) { mark[i] = j; rows[nr++] = i; acc[i] = A->nzval[ka] * b; }
        else acc[i] += A->nzval[ka] * b;
      }
    }
    /* insertion sort of rows (short columns) */
    for (int a = 1; a < nr; a++) {
      int r = rows[a], q = a - 1;
      while (q >= 0 && rows[q] > r) { rows[q + 1] = rows[q]; q--; }
      rows[q + 1] = r;
    }
    if (nnz + nr > cap) {
      cap = 2 * (nnz + nr) + 16;
      rv = (int *)realloc(rv, sizeof(int) * cap);
      nz = (double *)realloc(nz, sizeof(double) * cap);
    }
    for (int a = 0; a < nr; a++) { rv[nnz] = rows[a]; nz[nnz] = acc[rows[a]]; nnz++; }
  }
  cp[n] = nnz;
  C->m = m; C->n = n; C->nnz = nnz; C->colptr = cp; C->rowval = rv; C->nzval = nz;
  free(acc); free(mark); free(rows);
}

void om_diag_left(const double *d, const om_csc *A, om_csc *C) {
  csc_alloc(C, A->m, A->n, A->nnz);
  memcpy(C->colptr, A->colptr, sizeof(int) * (A->n + 1));
  memcpy(C->rowval, A->rowval, sizeof(int) * A->nnz);
  for (int k = 0; k < A->nnz; k++) C->nzval[k] = d[A->rowval[k]] * A->nzval[k];
}

void om_diag_right(const om_csc *A, const double *d, om_csc *C) {
  csc_alloc(C, A->m, A->n, A->nnz);
  memcpy(C->colptr, A->colptr, sizeof(int) * (A->n + 1));
  memcpy(C->rowval, A->rowval, sizeof(int) * A->nnz);
  for (int j = 0; j < A->n; j++)
    for (int k = A->colptr[j]; k < A->colptr[j + 1]; k++) C->nzval[k] = A->nzval[k] * d[j];
}

void om_scale(double a, const om_csc *A, om_csc *C) {
  csc_alloc(C, A->m, A->n, A->nnz);
  memcpy(C->colptr, A->colptr, sizeof(int) * (A->n + 1));
  memcpy(C->rowval, A->rowval, sizeof(int) * A->nnz);
  for (int k = 0; k < A->nnz; k++) C->nzval[k] = a * A->nzval[k];
}

void om_transpose(const om_csc *A, om_csc *C) {
  int m = A->m, n = A->n;
  csc_alloc(C, n, m, A->nnz);
  for (int k = 0; k < A->nnz; k++) C->colptr[A->rowval[k] + 1]++;
  for (int i = 0; i < m; i++) C->colptr[i + 1] += C->colptr[i];
  int *w = XALLOC(int, m + 1);
  memcpy(w, C->colptr, sizeof(int) * (m + 1));
  for (int j = 0; j < n; j++)
    for (int k = A->colptr[j]; k < A->colptr[j + 1]; k++) {
      int q = w[A->rowval[k]]++;
      C->rowval[q] = j;
      C->nzval[q] = A->nzval[k];
    }
  free(w);
}

/* dec_p_derivbound(Val(0|1)) (FastDEC.jl:471-524) -> sparse(I,J,V).
 * k = 0: d0 (ne x nv); k = 1: d1 (nt x ne).  transpose/negate as the
 * keyword arguments of the reference (boundary = transpose, dual derivative
 * = transpose [+ negate]). */
void om_derivbound(int k, int nv, int ne, int nt, const int *v0, const int *v1,
                   const int *e0, const int *e1, const int *e2,
                   const signed char *eo, const signed char *to, int transpose,
                   int negate, om_csc *A) {
  int nnz = k == 0 ? 2 * ne : 3 * nt;
  int *I = XALLOC(int, nnz), *J = XALLOC(int, nnz);
  double *V = XALLOC(double, nnz);
  int m, n;
  if (k == 0) {
    m = ne; n = nv;
    for (int i = 0; i < ne; i++) {
      int j = 2 * i;
      I[j] = I[j + 1] = i;
      J[j] = v0[i]; J[j + 1] = v1[i];
      double s = nsign(eo[i]);
      V[j] = s; V[j + 1] = -1 * s;
    }
  } else {
    m = nt; n = ne;
    for (int i = 0; i < nt; i++) {
      int j = 3 * i;
      I[j] = I[j + 1] = I[j + 2] = i;
      J[j] = e0[i]; J[j + 1] = e1[i]; J[j + 2] = e2[i];
      double ts = nsign(to[i]);
      V[j] = nsign(eo[e0[i]]) * ts;
      V[j + 1] = -1 * nsign(eo[e1[i]]) * ts;
      V[j + 2] = nsign(eo[e2[i]]) * ts;
    }
  }
  if (negate)
    for (int q = 0; q < nnz; q++) V[q] = -1 * V[q];
  if (transpose) om_sparse(n, m, nnz, J, I, V, A);
  else om_sparse(m, n, nnz, I, J, V, A);
  free(I); free(J); free(V);
}

/* dec_p_hodge_diag(Val(k)) for 2-D complexes (FastDEC.jl:580-604). */
void om_hodge_diag(int k, int nv, int ne, int nt, const int *D_v0, const int *D_v1,
                   const int *D_e1, const double *dual_length,
                   const double *dual_area, const double *length,
                   const double *area, double *h) {
  (void)D_v0;
  if (k == 0) {
    for (int v = 0; v < nv; v++) h[v] = 0.0;
    for (int dt = 0; dt < 6 * nt; dt++) h[D_v1[D_e1[dt]]] += dual_area[dt];
  } else if (k == 1) {
    for (int e = 0; e < ne; e++) h[e] = 0.0;
    for (int de = 0; de < 2 * ne + 6 * nt; de++) {
      int s = D_v1[de] - nv;
      if (s >= 0 && s < ne) h[s] += dual_length[de] / length[s];
    }
  } else {
    for (int t = 0; t < nt; t++) h[t] = 1 / area[t];
  }
}

/* dec_inv_hodge_star(Val(k), sd, DiagonalHodge()) (FastDEC.jl:725-730), 2-D:
 * (1 ./ hdg) .* (-1)^(k(2-k)). */
void om_inv_hodge_diag(int k, int n, const double *h, double *out) {
  double mult = ((k * (2 - k)) % 2 == 0) ? 1.0 : -1.0;
  for (int i = 0; i < n; i++) out[i] = (1 / h[i]) * mult;
}

/* wedge_dd_01_mat (FastDEC.jl:293-305): m[e,t] = dual_length / sum over the
 * elementary duals of e (B-block dual edges whose D_∂v1 is e's edge center,
 * ascending dual-edge id). */
void om_wedge_dd01_mat(int nv, int ne, int nt, const int *D_v0, const int *D_v1,
                       const double *dl, om_csc *M) {
  int nde = 2 * ne + 6 * nt;
  int *cnt = XALLOC(int, ne + 1);
  for (int de = 0; de < nde; de++) {
    int s = D_v1[de] - nv;
    if (s >= 0 && s < ne) cnt[s + 1]++;
  }
  for (int e = 0; e < ne; e++) cnt[e + 1] += cnt[e];
  int tot = cnt[ne];
  int *pos = XALLOC(int, ne + 1), *lst = XALLOC(int, tot);
  memcpy(pos, cnt, sizeof(int) * (ne + 1));
  for (int de = 0; de < nde; de++) {
    int s = D_v1[de] - nv;
    if (s >= 0 && s < ne) lst[pos[s]++] = de;
  }
  int *I = XALLOC(int, tot), *J = XALLOC(int, tot);
  double *V = XALLOC(double, tot);
  int q = 0;
  for (int e = 0; e < ne; e++) {
    double sum = 0.0;
    for (int a = cnt[e]; a < cnt[e + 1]; a++) sum = sum + dl[lst[a]];
    for (int a = cnt[e]; a < cnt[e + 1]; a++) {
      I[q] = e;
      J[q] = D_v0[lst[a]] - nv - ne;
      V[q] = dl[lst[a]] / sum;
      q++;
    }
  }
  /* m[e,t] = w assigns (the last assignment wins on a repeated (e,t)); the
   * dual complex never repeats one, so sparse() of the list is identical. */
  om_sparse(ne, nt, tot, I, J, V, M);
  free(cnt); free(pos); free(lst); free(I); free(J); free(V);
}

/* wedge_pd_01_mat (src/FastDEC.jl:329-346): edge e = (α, β) = (∂v0, ∂v1);
 * for each elementary dual edge de of e (dual edges ending at e's center,
 * ascending) with w = dual_length[de] / Σ dual_length, the triangle t whose
 * center starts de and its third vertex l: m[e,α] += w*5/12,
 * m[e,β] += w*5/12, m[e,l] += w*2/12 (the += chain is sparse()'s duplicate
 * summation in input order). */
void om_wedge_pd01_mat(int nv, int ne, int nt, const int *v0, const int *v1, const int *e1,
                       const int *e2, const int *D_v0, const int *D_v1, const double *dl,
                       om_csc *M) {
  int nde = 2 * ne + 6 * nt;
  int *cnt = XALLOC(int, ne + 1);
  for (int de = 0; de < nde; de++) {
    int s = D_v1[de] - nv;
    if (s >= 0 && s < ne) cnt[s + 1]++;
  }
  for (int e = 0; e < ne; e++) cnt[e + 1] += cnt[e];
  int tot = cnt[ne];
  int *pos = XALLOC(int, ne + 1), *lst = XALLOC(int, tot);
  memcpy(pos, cnt, sizeof(int) * (ne + 1));
  for (int de = 0; de < nde; de++) {
    int s = D_v1[de] - nv;
    if (s >= 0 && s < ne) lst[pos[s]++] = de;
  }
  int *I = XALLOC(int, 3 * tot), *J = XALLOC(int, 3 * tot);
  double *V = XALLOC(double, 3 * tot);
  int q = 0;
  for (int e = 0; e < ne; e++) {
    int a = v0[e], b = v1[e];
    double sum = 0.0;
    for (int k = cnt[e]; k < cnt[e + 1]; k++) sum = sum + dl[lst[k]];
    for (int k = cnt[e]; k < cnt[e + 1]; k++) {
      int t = D_v0[lst[k]] - nv - ne;
      /* triangle_vertices(t) = (∂v1[∂e1], ∂v0[∂e2], ∂v0[∂e1]) (src/SimplicialSets.jl:305-309) */
      int tv[3] = {v1[e1[t]], v0[e2[t]], v0[e1[t]]};
      int l = -1;
      for (int i = 0; i < 3; i++)
        if (tv[i] != a && tv[i] != b) l = tv[i];
      double w = dl[lst[k]] / sum;
      I[q] = e; J[q] = a; V[q] = w * 5 / 12; q++;
      I[q] = e; J[q] = b; V[q] = w * 5 / 12; q++;
      I[q] = e; J[q] = l; V[q] = w * 2 / 12; q++;
    }
  }
  om_sparse(ne, nv, q, I, J, V, M);
  free(cnt); free(pos); free(lst); free(I); free(J); free(V);
}

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static double norm3(const double *a) { return sqrt(dot3(a, a)); }
static double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }
/* crossdot (src/DiscreteExteriorCalculus.jl:1667-1670) */
static double crossdot(const double *a, const double *b) {
  double c[3];
  cross3(a, b, c);
  return norm3(c) * (c[2] == 0 ? 1.0 : sgn(c[2]));
}

/* dec_hodge_star(Val(1), sd::EmbeddedDeltaDualComplex2D, GeometricHodge())
 * (FastDEC.jl:660-716). */
void om_geo_hodge1(int nv, int ne, int nt, const int *v0, const int *v1,
                   const int *e0, const int *e1, const int *e2,
                   const double *dp, om_csc *H) {
  int *I = XALLOC(int, 9 * nt), *J = XALLOC(int, 9 * nt);
  double *V = XALLOC(double, 9 * nt);
  const int *te[3] = {e2, e1, e0}; /* reversed by construction */
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int idx = 0;
  for (int t = 0; t < nt; t++) {
    const double *tc = dp + 3 * (size_t)(nv + ne + t);
    double evt[3][3], dvt[3][3];
    for (int i = 0; i < 3; i++) {
      int e = te[i][t];
      const double *a = dp + 3 * v0[e], *b = dp + 3 * v1[e], *ec = dp + 3 * (size_t)(nv + e);
      for (int c = 0; c < 3; c++) { evt[i][c] = a[c] - b[c]; dvt[i][c] = tc[c] - ec[c]; }
    }
    for (int c = 0; c < 3; c++) dvt[1][c] *= -1;
    double cr[3];
    cross3(evt[0], dvt[0], cr);
    double rel = cr[2] == 0 ? 1.0 : sgn(cr[2]);
    for (int i = 0; i < 3; i++) {
      double dc = crossdot(evt[i], dvt[i]);
      if (dc != 0.0) {
        I[idx] = te[i][t]; J[idx] = te[i][t];
        V[idx] = rel * dc / dot3(evt[i], evt[i]);
        idx++;
      }
    }
    for (int q = 0; q < 6; q++) {
      int a = perm[q][0], b = perm[q][1], c = perm[q][2];
      double ddot = dot3(evt[a], dvt[a]) / dot3(evt[a], evt[a]);
      double val = ddot * dot3(evt[a], evt[c]);
      if (val != 0.0) {
        I[idx] = te[a][t]; J[idx] = te[b][t];
        V[idx] = rel * val / crossdot(evt[b], evt[c]);
        idx++;
      }
    }
  }
  om_sparse(ne, ne, idx, I, J, V, H);
  free(I); free(J); free(V);
}

/* fast_laplace_beltrami (test/Multigrid.jl:46-48):
 *   -1 * ⋆0⁻¹ * d̃1 * ⋆1 * d0, with d̃1 = dec_dual_derivative(1) = -d0ᵀ,
 * evaluated left to right as Julia does: ((((-1*D(inv_h0)) * d̃1) * H1) * d0).
 * H1 is Diagonal(h1) when h1_geo is NULL, else the sparse geometric ⋆1. */
void om_laplacian0(const om_csc *d0, const double *h1, const om_csc *h1_geo,
                   const double *inv_h0, om_csc *L) {
  int nv = d0->n;
  double *mih = XALLOC(double, nv);
  for (int v = 0; v < nv; v++) mih[v] = -1 * inv_h0[v];
  om_csc dd1t, tmp, X;
  om_transpose(d0, &tmp);          /* d0ᵀ (nv x ne) */
  om_scale(-1.0, &tmp, &dd1t);     /* d̃1 = -d0ᵀ (Int8 values, exact) */
  om_csc_free(&tmp);
  om_diag_left(mih, &dd1t, &tmp);  /* Diagonal(-inv_h0) * d̃1 */
  om_csc_free(&dd1t);
  if (h1_geo) om_spgemm(&tmp, h1_geo, &X);
  else om_diag_right(&tmp, h1, &X);
  om_csc_free(&tmp);
  om_spgemm(&X, d0, L);
  om_csc_free(&X);
  free(mih);
}

/* ------------------------------------------------------------------------ */
/* Multigrid                                                                  */
/* ------------------------------------------------------------------------ */

/* Krylov.jl `cg(A, b, x0; itmax)` as called by _multigrid_μ_cycle
 * (src/Multigrid.jl:860, 871): warm start solves A Δ = b - A x0 from zero with
 * atol = rtol = sqrt(eps), zero-curvature guard, and returns Δ + x0.
 * Third-party (Krylov.jl 0.9.6/0.10, not vendored): parity unpinned beyond
 * the reference's residual thresholds. */
void om_cg(const om_csc *A, const double *b, double *x0, int itmax) {
  int n = A->m;
  double *x = XALLOC(double, n), *r = XALLOC(double, n), *p = XALLOC(double, n),
         *Ap = XALLOC(double, n);
  om_spmv(A, x0, r);
  for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
  memcpy(p, r, sizeof(double) * n);
  double gamma = 0.0;
  for (int i = 0; i < n; i++) gamma += r[i] * r[i];
  double rnorm = sqrt(gamma);
  if (gamma == 0) goto done;
  if (itmax == 0) itmax = 2 * n;
  const double eps = 2.220446049250313e-16, sq = 1.4901161193847656e-08;
  double eps_stop = sq + sq * rnorm, pnorm2 = gamma;
  int iter = 0, solved = rnorm <= eps_stop;
  while (!solved && iter < itmax) {
    om_spmv(A, p, Ap);
    double pAp = 0.0;
    for (int i = 0; i < n; i++) pAp += p[i] * Ap[i];
    if (pAp <= eps * pnorm2 && fabs(pAp) <= eps * pnorm2) break; /* zero curvature */
    double alpha = gamma / pAp;
    for (int i = 0; i < n; i++) x[i] += alpha * p[i];
    for (int i = 0; i < n; i++) r[i] -= alpha * Ap[i];
    double gnext = 0.0;
    for (int i = 0; i < n; i++) gnext += r[i] * r[i];
    rnorm = sqrt(gnext);
    solved = (rnorm <= eps_stop) || (rnorm + 1.0 <= 1.0);
    if (!solved) {
      double beta = gnext / gamma;
      pnorm2 = gnext + beta * beta * pnorm2;
      gamma = gnext;
      for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
    }
    iter++;
  }
done:
  for (int i = 0; i < n; i++) x0[i] = x[i] + x0[i];
  free(x); free(r); free(p); free(Ap);
}

/* Greedy colouring in vertex order over the symmetric pattern of A: each
 * vertex takes the smallest colour unused by its already-coloured neighbours.
 * Returns the number of colours. */
int om_greedy_colors(const om_csc *A, int *color) {
  int n = A->n;
  om_csc At;
  om_transpose(A, &At);
  int *used = XALLOC(int, n + 2), nc = 0;
  for (int i = 0; i < n + 2; i++) used[i] = -1;
  for (int v = 0; v < n; v++) color[v] = -1;
  for (int v = 0; v < n; v++) {
    for (int k = A->colptr[v]; k < A->colptr[v + 1]; k++) {
      int w = A->rowval[k];
      if (w != v && color[w] >= 0) used[color[w]] = v;
    }
    for (int k = At.colptr[v]; k < At.colptr[v + 1]; k++) {
      int w = At.rowval[k];
      if (w != v && color[w] >= 0) used[color[w]] = v;
    }
    int c = 0;
    while (used[c] == v) c++;
    color[v] = c;
    if (c + 1 > nc) nc = c + 1;
  }
  free(used);
  om_csc_free(&At);
  return nc;
}

/* Dense coarse-grid operator Z (row-major n x n) with z = Z r solving A z = r.
 * When A annihilates constants (closed or Neumann surfaces) the bordered
 * system [A 1; 1ᵀ 0] is inverted and Z is its leading block: z has zero mean
 * and absorbs any inconsistent component of r into the border multiplier. */
void om_coarse_inverse(const om_csc *A, double *Z) {
  int n = A->m;
  double amax = 0.0, rmax = 0.0;
  double *rs = XALLOC(double, n);
  for (int j = 0; j < n; j++)
    for (int k = A->colptr[j]; k < A->colptr[j + 1]; k++) {
      rs[A->rowval[k]] += A->nzval[k];
      if (fabs(A->nzval[k]) > amax) amax = fabs(A->nzval[k]);
    }
  for (int i = 0; i < n; i++) if (fabs(rs[i]) > rmax) rmax = fabs(rs[i]);
  int sing = rmax <= 1e-10 * amax;
  int N = n + (sing ? 1 : 0);
  double *M = XALLOC(double, (size_t)N * 2 * N);
  int W = 2 * N;
  for (int j = 0; j < n; j++)
    for (int k = A->colptr[j]; k < A->colptr[j + 1]; k++) M[(size_t)A->rowval[k] * W + j] += A->nzval[k];
  if (sing) {
    for (int i = 0; i < n; i++) { M[(size_t)i * W + n] = 1.0; M[(size_t)n * W + i] = 1.0; }
  }
  for (int i = 0; i < N; i++) M[(size_t)i * W + N + i] = 1.0;
  for (int c = 0; c < N; c++) {
    int piv = c;
    for (int r = c + 1; r < N; r++)
      if (fabs(M[(size_t)r * W + c]) > fabs(M[(size_t)piv * W + c])) piv = r;
    if (piv != c)
      for (int j = 0; j < W; j++) { double x = M[(size_t)c * W + j]; M[(size_t)c * W + j] = M[(size_t)piv * W + j]; M[(size_t)piv * W + j] = x; }
    double d = M[(size_t)c * W + c];
    for (int j = 0; j < W; j++) M[(size_t)c * W + j] /= d;
    for (int r = 0; r < N; r++) {
      if (r == c) continue;
      double f = M[(size_t)r * W + c];
      if (f == 0.0) continue;
      for (int j = 0; j < W; j++) M[(size_t)r * W + j] -= f * M[(size_t)c * W + j];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Z[(size_t)i * n + j] = M[(size_t)i * W + N + j];
  free(M);
  free(rs);
}

typedef struct {
  const om_csc *A, *R, *S;
  om_csc At;      /* rows of A */
  int *color, ncolors;
  int *block;     /* OM_BLOCK_GS: block index of each vertex */
  double *Z;      /* coarsest, direct */
} om_level;

/* One Gauss-Seidel sweep over the colours in the given direction: within a
 * colour every row is updated from the current iterate,
 * u[v] = (b[v] - sum_{w != v, ascending} A[v,w] u[w]) / A[v,v]. */
static void gs_sweep(om_level *L, const double *b, double *u, int reverse) {
  const om_csc *At = &L->At;
  int n = At->n;
  for (int ci = 0; ci < L->ncolors; ci++) {
    int c = reverse ? L->ncolors - 1 - ci : ci;
    for (int v = 0; v < n; v++) {
      if (L->color[v] != c) continue;
      double s = 0.0, d = 0.0;
      for (int k = At->colptr[v]; k < At->colptr[v + 1]; k++) {
        int w = At->rowval[k];
        if (w == v) d = At->nzval[k];
        else s += At->nzval[k] * u[w];
      }
      u[v] = (b[v] - s) / d;
    }
  }
}

/* One hybrid block Gauss-Seidel sweep (the build's GPU smoother, no
 * counterpart in the reference): the vertices are split into blocks; within a
 * block the sweep is the multicolour sweep above, and a neighbour in another
 * block enters with its value from the start of the sweep,
 * u[v] = (b[v] - sum_{w != v, ascending} A[v,w] x[w]) / A[v,v] with
 * x[w] = u[w] when block[w] == block[v], else u_start[w].
 * With a single block this is gs_sweep. */
static void block_gs_sweep(om_level *L, const double *b, double *u, int reverse) {
  const om_csc *At = &L->At;
  int n = At->n;
  double *u0 = XALLOC(double, n);
  memcpy(u0, u, sizeof(double) * n);
  for (int ci = 0; ci < L->ncolors; ci++) {
    int c = reverse ? L->ncolors - 1 - ci : ci;
    for (int v = 0; v < n; v++) {
      if (L->color[v] != c) continue;
      double s = 0.0, d = 0.0;
      for (int k = At->colptr[v]; k < At->colptr[v + 1]; k++) {
        int w = At->rowval[k];
        if (w == v) d = At->nzval[k];
        else s += At->nzval[k] * (L->block[w] == L->block[v] ? u[w] : u0[w]);
      }
      u[v] = (b[v] - s) / d;
    }
  }
  free(u0);
}

/* Damped Jacobi: u += omega * (b - A u) ./ diag(A). */
static void jacobi_sweep(om_level *L, const double *b, double *u, double omega) {
  const om_csc *At = &L->At;
  int n = At->n;
  double *un = XALLOC(double, n);
  for (int v = 0; v < n; v++) {
    double s = 0.0, d = 0.0;
    for (int k = At->colptr[v]; k < At->colptr[v + 1]; k++) {
      int w = At->rowval[k];
      if (w == v) d = At->nzval[k];
      s += At->nzval[k] * u[w];
    }
    un[v] = u[v] + omega * ((b[v] - s) / d);
  }
  memcpy(u, un, sizeof(double) * n);
  free(un);
}

static void smooth(om_level *L, int alg, int steps, double omega, const double *b,
                   double *u, int post) {
  if (steps == 0) return;
  if (alg == OM_CG) om_cg(L->A, b, u, steps);
  else if (alg == OM_GS) for (int s = 0; s < steps; s++) gs_sweep(L, b, u, post);
  else if (alg == OM_BLOCK_GS) for (int s = 0; s < steps; s++) block_gs_sweep(L, b, u, post);
  else for (int s = 0; s < steps; s++) jacobi_sweep(L, b, u, omega);
}

/* _multigrid_μ_cycle (src/Multigrid.jl:857-872).  With coarse_direct the
 * coarsest level is solved with the dense operator Z (u += Z (b - A u))
 * instead of running the smoother there. */
static void mu_cycle(om_level *lv, int nlev, int lvl, const int *steps, int alg,
                     int mu, int coarse_direct, double omega, double *u,
                     const double *b) {
  om_level *L = lv + lvl;
  int n = L->A->m;
  if (lvl == nlev - 1 && coarse_direct) {
    double *r = XALLOC(double, n);
    om_spmv(L->A, u, r);
    for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = 0; j < n; j++) s += L->Z[(size_t)i * n + j] * r[j];
      u[i] = u[i] + s;
    }
    free(r);
    return;
  }
  smooth(L, alg, steps[lvl], omega, b, u, 0);
  if (lvl >= nlev - 1) return;
  int nc = L->R->m;
  double *r = XALLOC(double, n), *rc = XALLOC(double, nc), *z = XALLOC(double, nc),
         *pz = XALLOC(double, n);
  om_spmv(L->A, u, r);
  for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
  om_spmv(L->R, r, rc);
  mu_cycle(lv, nlev, lvl + 1, steps, alg, mu, coarse_direct, omega, z, rc);
  if (mu > 1) mu_cycle(lv, nlev, lvl + 1, steps, alg, mu - 1, coarse_direct, omega, z, rc);
  om_spmv_t(L->S, z, pz); /* P = transpose(S) */
  for (int i = 0; i < n; i++) u[i] = u[i] + pz[i];
  smooth(L, alg, steps[lvl], omega, b, u, 1);
  free(r); free(rc); free(z); free(pz);
}

/* Per-level Gauss-Seidel colours: `colors` holds every level's colour array,
 * finest first, concatenated (NULL: greedy colours, om_greedy_colors). */
static om_level *levels_setup(int nlev, const om_csc *A, const om_csc *R,
                              const om_csc *S, int alg, int coarse_direct,
                              const int *colors, const int *blocks) {
  om_level *lv = XALLOC(om_level, nlev);
  size_t off = 0;
  int gs = alg == OM_GS || alg == OM_BLOCK_GS;
  for (int l = 0; l < nlev; l++) {
    lv[l].A = A + l;
    lv[l].R = l < nlev - 1 ? R + l : NULL;
    lv[l].S = l < nlev - 1 ? S + l : NULL;
    om_transpose(A + l, &lv[l].At);
    lv[l].color = XALLOC(int, A[l].n);
    lv[l].block = XALLOC(int, A[l].n);
    for (int v = 0; v < A[l].n; v++) lv[l].block[v] = blocks ? blocks[off + v] : 0;
    if (gs && colors) {
      int nc = 0;
      for (int v = 0; v < A[l].n; v++) {
        lv[l].color[v] = colors[off + v];
        if (colors[off + v] + 1 > nc) nc = colors[off + v] + 1;
      }
      lv[l].ncolors = nc;
    } else {
      lv[l].ncolors = gs ? om_greedy_colors(A + l, lv[l].color) : 1;
    }
    off += (size_t)A[l].n;
    lv[l].Z = NULL;
    if (l == nlev - 1 && coarse_direct) {
      lv[l].Z = XALLOC(double, (size_t)A[l].m * A[l].m);
      om_coarse_inverse(A + l, lv[l].Z);
    }
  }
  return lv;
}

static void levels_free(om_level *lv, int nlev) {
  for (int l = 0; l < nlev; l++) {
    om_csc_free(&lv[l].At);
    free(lv[l].color);
    free(lv[l].block);
    free(lv[l].Z);
  }
  free(lv);
}

/* multigrid_μ_cycles (src/Multigrid.jl:819-824). */
void om_mu_cycles(int nlev, const om_csc *A, const om_csc *R, const om_csc *S,
                  const int *steps, int alg, int mu, int cycles, int coarse_direct,
                  double omega, double *u, const double *b, const int *colors,
                  const int *blocks) {
  om_level *lv = levels_setup(nlev, A, R, S, alg, coarse_direct, colors, blocks);
  for (int c = 0; c < cycles; c++) mu_cycle(lv, nlev, 0, steps, alg, mu, coarse_direct, omega, u, b);
  levels_free(lv, nlev);
}

/* full_multigrid / _full_multigrid (src/Multigrid.jl:830-855). */
static void fmg_rec(om_level *lv, int nlev, int lvl, const int *steps, int alg, int mu,
                    int cycles, int coarse_direct, double omega, const double *b,
                    double *z) {
  int n = lv[lvl].A->m;
  for (int i = 0; i < n; i++) z[i] = 0.0;
  if (lvl < nlev - 1) {
    int nc = lv[lvl].R->m;
    double *bc = XALLOC(double, nc), *zc = XALLOC(double, nc);
    om_spmv(lv[lvl].R, b, bc);
    fmg_rec(lv, nlev, lvl + 1, steps, alg, mu, cycles, coarse_direct, omega, bc, zc);
    om_spmv_t(lv[lvl].S, zc, z);
    free(bc); free(zc);
  }
  if (lvl > 0)
    for (int c = 0; c < cycles; c++) mu_cycle(lv, nlev, lvl, steps, alg, mu, coarse_direct, omega, z, b);
}

void om_full_multigrid(int nlev, const om_csc *A, const om_csc *R, const om_csc *S,
                       const int *steps, int alg, int mu, int cycles,
                       int coarse_direct, double omega, const double *b, double *u,
                       const int *colors, const int *blocks) {
  om_level *lv = levels_setup(nlev, A, R, S, alg, coarse_direct, colors, blocks);
  fmg_rec(lv, nlev, 0, steps, alg, mu, cycles, coarse_direct, omega, b, u);
  for (int c = 0; c < cycles; c++) mu_cycle(lv, nlev, 0, steps, alg, mu, coarse_direct, omega, u, b);
  levels_free(lv, nlev);
}
