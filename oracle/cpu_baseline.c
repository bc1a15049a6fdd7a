/* CPU baseline of the operator set (test and benchmark infrastructure only):
 * the oracle's loops for wedge 0-1, 1-1, 0-2, d0, d1, the diagonal Hodge
 * stars and the 0-form Laplacian, built with -O3 and OpenMP so that the
 * benchmark can time the restated reference loops at one thread and at every
 * host core.  Each loop is the oracle's (dec_oracle.c: om_wedge01, om_wedge11,
 * om_wedge02, om_spmv), row-parallel: the sparse operators run on CSR copies
 * whose row sums follow the same ascending column order as the oracle's CSC
 * product, so the outputs equal the oracle's bit for bit.
 *
 * Reference loops restated: wedge_kernel_01!/02!/11! (src/FastDEC.jl:159-191),
 * the SparseArrays products of dec_differential (src/FastDEC.jl:468-469),
 * dec_hodge_star's diagonal (src/FastDEC.jl:641-648) and the 0-form
 * Laplacian (test/Multigrid.jl:46-48). */
#include <omp.h>
#include <stddef.h>

typedef struct {
  int nv, ne, nt;
  const int *v0, *v1;                /* wedge 0-1: edge endpoints */
  const int *e11; const double *c11; /* wedge 1-1: 3 edges + 3 coefficients per triangle */
  const int *v02; const double *c02; /* wedge 0-2: 6 vertices + 6 coefficients per triangle */
  const int *d0p, *d0c; const double *d0v;  /* CSR of d0 (ne rows) */
  const int *d1p, *d1c; const double *d1v;  /* CSR of d1 (nt rows) */
  const int *Lp, *Lc; const double *Lv;     /* CSR of the Laplacian (nv rows) */
  const double *h0, *h1, *h2;               /* diagonal ⋆0, ⋆1, ⋆2 */
  const double *V, *E, *E2, *T;             /* input forms */
  double *oV, *oE, *oT;                     /* outputs (overwritten by each operator) */
} cb_opset;

static void csr_rows(int n, const int *p, const int *c, const double *v, const double *x, double *y) {
#pragma omp for schedule(static)
  for (int r = 0; r < n; r++) {
    double s = 0.0;
    for (int k = p[r]; k < p[r + 1]; k++) s += v[k] * x[c[k]];
    y[r] = s;
  }
}

static void diag(int n, const double *h, const double *x, double *y) {
#pragma omp for schedule(static)
  for (int i = 0; i < n; i++) y[i] = h[i] * x[i];
}

/* Operator `which` of the set, called inside a parallel region:
 * 0 wedge 0-1 -> oE, 1 wedge 1-1 -> oT, 2 wedge 0-2 -> oT, 3 d0 -> oE,
 * 4 d1 -> oT, 5 ⋆0 -> oV, 6 ⋆1 -> oE, 7 ⋆2 -> oT, 8 Laplacian -> oV. */
static void op(const cb_opset *o, int which) {
  const int ne = o->ne, nt = o->nt, nv = o->nv;
  switch (which) {
    case 0:
#pragma omp for schedule(static)
      for (int i = 0; i < ne; i++) o->oE[i] = 0.5 * o->E[i] * (o->V[o->v0[i]] + o->V[o->v1[i]]);
      break;
    case 1:
#pragma omp for schedule(static)
      for (int i = 0; i < nt; i++) {
        const int *e = o->e11 + 3 * (size_t)i;
        const double *c = o->c11 + 3 * (size_t)i;
        double a0 = o->E[e[0]], a1 = o->E[e[1]], a2 = o->E[e[2]];
        double b0 = o->E2[e[0]], b1 = o->E2[e[1]], b2 = o->E2[e[2]];
        o->oT[i] = c[0] * (a2 * b1 - a1 * b2) + c[1] * (a2 * b0 - a0 * b2) + c[2] * (a1 * b0 - a0 * b1);
      }
      break;
    case 2:
#pragma omp for schedule(static)
      for (int i = 0; i < nt; i++) {
        const int *q = o->v02 + 6 * (size_t)i;
        const double *k = o->c02 + 6 * (size_t)i;
        double s = k[0] * o->V[q[0]] + k[1] * o->V[q[1]];
        s = s + k[2] * o->V[q[2]];
        s = s + k[3] * o->V[q[3]];
        s = s + k[4] * o->V[q[4]];
        s = s + k[5] * o->V[q[5]];
        o->oT[i] = o->T[i] * s;
      }
      break;
    case 3: csr_rows(ne, o->d0p, o->d0c, o->d0v, o->V, o->oE); break;
    case 4: csr_rows(nt, o->d1p, o->d1c, o->d1v, o->E, o->oT); break;
    case 5: diag(nv, o->h0, o->V, o->oV); break;
    case 6: diag(ne, o->h1, o->E, o->oE); break;
    case 7: diag(nt, o->h2, o->T, o->oT); break;
    default: csr_rows(nv, o->Lp, o->Lc, o->Lv, o->V, o->oV); break;
  }
}

/* Runs `passes` passes of the nine operators on `threads` threads and
 * returns the elapsed wall seconds (omp_get_wtime around the passes). */
double cb_run(const cb_opset *o, int threads, int passes) {
  double t0 = omp_get_wtime();
#pragma omp parallel num_threads(threads)
  for (int p = 0; p < passes; p++)
    for (int w = 0; w < 9; w++) op(o, w);
  return omp_get_wtime() - t0;
}

/* Operator `which` alone, once (for checking its output). */
void cb_apply(const cb_opset *o, int threads, int which) {
#pragma omp parallel num_threads(threads)
  op(o, which);
}
