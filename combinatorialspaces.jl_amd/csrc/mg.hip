// Geometric multigrid of libdec_hip.so (src/Multigrid.jl): the binary-
// subdivision hierarchy in DirectMode with op = fast_laplace_beltrami, and the
// μ-cycles / full multigrid / tolerance solves on the GPU.
//
// Per level the operator is held in slot-major ELL with rows sorted by
// (colour, vertex): a multicolour Gauss-Seidel pass over one colour is then a
// contiguous range of rows.  The same slots with rows in vertex order serve
// residuals and the CG smoother.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>

#include "dec_internal.h"
#include "ell_rows.h"
#include "kernels_common.h"
#include "learned.h"

using dec::DBuf;
using dec::fail;

namespace {

constexpr int kBlock = 256;
constexpr int kDotBlocks = 256;
constexpr int kNormSlots = 4;  // residual norms of a solve in flight (at most 2 are)

// ---------------------------------------------------------------- kernels

template <class T>
__global__ void k_zero(int n, T *__restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = static_cast<T>(0);
}

// r[v] = b[v] - (A u)[v], row sums in ascending column order (b - A*u); the
// rows at ELL positions p are rowid[p] (rows in vertex order: rowid null).
template <int W, class T>
__global__ void k_residual(int n, int width, const int *__restrict__ col, const T *__restrict__ val,
                           const int *__restrict__ rowid, const T *__restrict__ b, const T *__restrict__ u,
                           T *__restrict__ r) {
  int p = dec_ell::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n) return;
  T acc = dec_ell::row<W, T, T>(n, width, p, col, val, u);
  int v = rowid ? rowid[p] : p;
  r[v] = b[v] - acc;
}

// The residual r = b - A u of k_residual and, per workgroup, the sum of its
// rows' r² (a fixed tree) into part[blockIdx.x]; k_sum_rows adds the
// partial sums.  (A ticket letting the last workgroup add them needs a
// device-scope fence in every workgroup, which writes back the XCD's L2:
// measured six times slower.)
template <int W>
__global__ void __launch_bounds__(256)
    k_residual_norm(int n, int width, const int *__restrict__ col, const double *__restrict__ val,
                    const int *__restrict__ rowid, const double *__restrict__ b, const double *__restrict__ u,
                    double *__restrict__ r, double *__restrict__ part) {
  __shared__ double sh[256];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0;
  if (p < n) {
    double acc = dec_ell::row<W, double, double>(n, width, p, col, val, u);
    int v = rowid ? rowid[p] : p;
    double ri = b[v] - acc;
    r[v] = ri;
    rr = ri * ri;
  }
  sh[threadIdx.x] = rr;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// *out = Σ part[0..nb) in a fixed order: 256 consecutive runs, then a tree.
__global__ void __launch_bounds__(256) k_sum_rows(int nb, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double sh[256];
  int per = (nb + 255) / 256;
  double acc = 0.0;
  for (int j = threadIdx.x * per; j < min(nb, (static_cast<int>(threadIdx.x) + 1) * per); j++) acc = acc + part[j];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// y[v] = (A x)[v]
template <int W, class T>
__global__ void k_spmv_rows(int n, int width, const int *__restrict__ col, const T *__restrict__ val,
                            const int *__restrict__ rowid, const T *__restrict__ x, T *__restrict__ y,
                            const int *__restrict__ done) {
  if (done && *done) return;
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  y[rowid[p]] = dec_ell::row<W, T, T>(n, width, p, col, val, x);
}

// One colour of Gauss-Seidel: u[v] = (b[v] - Σ_{w≠v} A[v,w] u[w]) / A[v,v].
template <int W, class T>
__global__ void k_gs_color(int p0, int p1, int n, int width, const int *__restrict__ col,
                           const T *__restrict__ val, const int *__restrict__ rowid, const T *__restrict__ b,
                           T *__restrict__ u) {
  int p = p0 + dec_ell::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  int v = rowid[p];
  T s, d;
  dec_ell::gs_row<W, T>(n, width, p, v, col, val, u, s, d);
  u[v] = (b[v] - s) / d;
}

// Damped Jacobi: un[v] = u[v] + ω (b[v] - (A u)[v]) / A[v,v].
template <int W, class T>
__global__ void k_jacobi(int n, int width, const int *__restrict__ col, const T *__restrict__ val,
                         const int *__restrict__ rowid, const T *__restrict__ b, const T *__restrict__ u,
                         T *__restrict__ un, T omega) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int v = rowid[p];
  T s, d;
  dec_ell::gs_row<W, T>(n, width, p, v, col, val, u, s, d);
  // (A u)[v] in ascending column order: the diagonal term enters at its place
  T full = dec_ell::row<W, T, T>(n, width, p, col, val, u);
  (void)s;
  un[v] = u[v] + omega * ((b[v] - full) / d);
}

template <class T>
__global__ void k_copy(int n, const T *__restrict__ x, T *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// b_c = R r_f; rows are coarse vertices in natural order (CSC order of R:
// the identity column first, then the midpoint columns ascending).
// uc (when given) is the coarse iterate, zeroed here for the coarse cycle.
template <int W, class T>
__global__ void k_restrict(int nc, int width, const int *__restrict__ col, const T *__restrict__ val,
                           const T *__restrict__ rf, T *__restrict__ bc, T *__restrict__ uc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  bc[i] = dec_ell::row<W, T, T>(nc, width, i, col, val, rf);
  if (uc) uc[i] = static_cast<T>(0);
}

// k_residual and k_restrict in one launch for levels where launches, not
// bytes, bound the cycle.  Eight lanes per coarse row i: lane s evaluates
// r_j = b_j - (A u)_j for the row's slot s (j = R's column; the fine rows in
// vertex order, summed as k_residual does) and multiplies it by R's
// entry; lane 0 then adds the eight products in slot order, so b_c is
// bit-identical to the two launches.  Fine rows shared by two coarse rows are
// evaluated twice.  R's width is at most 8.
template <int W, class T>
__global__ void __launch_bounds__(256)
    k_residual_restrict(int nc, int rwidth, const int *__restrict__ rcol, const T *__restrict__ rval, int n,
                        int width, const int *__restrict__ col, const T *__restrict__ val, const T *__restrict__ b,
                        const T *__restrict__ u, T *__restrict__ bc, T *__restrict__ uc) {
  __shared__ T prod[256];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int i = t >> 3, sl = t & 7;
  T pr = static_cast<T>(0);
  bool has = false;
  if (i < nc && sl < rwidth) {
    size_t q = static_cast<size_t>(sl) * nc + i;
    int j = rcol[q];
    if (j >= 0) {
      T rj = b[j] - dec_ell::row<W, T, T>(n, width, j, col, val, u);
      pr = rval[q] * rj;
      has = true;
    }
  }
  prod[threadIdx.x] = pr;
  unsigned long long m = __ballot(has);
  __syncthreads();
  if (sl == 0 && i < nc) {
    T acc = static_cast<T>(0);
    int base = threadIdx.x;
    unsigned lanebits = static_cast<unsigned>((m >> (threadIdx.x & 63)) & 0xffu);
    for (int k = 0; k < 8; k++)
      if (lanebits & (1u << k)) acc = acc + prod[base + k];
    bc[i] = acc;
    if (uc) uc[i] = static_cast<T>(0);
  }
}

// u_f += P z_c, P in ELL rows (for a binary subdivision P = transpose(S):
// a coarse vertex copies its value with weight 1, a midpoint takes 0.5 of its
// edge's endpoints in the order (min, max) of S's column).
template <int W, class T>
__global__ void k_prolong_add(int nf, int width, const int *__restrict__ pcol, const T *__restrict__ pval,
                              const T *__restrict__ zc, T *__restrict__ uf) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nf) return;
  uf[j] = uf[j] + dec_ell::row<W, T, T>(nf, width, j, pcol, pval, zc);
}

// u[j], or with zc given u[j] + (P zc)[j] (P an ELL of width 2: the sum of
// k_prolong_add), for sweeps that apply the prolongation themselves.
template <class T>
__device__ __forceinline__ T prolonged(int n, int j, const T *__restrict__ u, const int *__restrict__ pcol,
                                       const T *__restrict__ pval, const T *__restrict__ zc) {
  return zc ? u[j] + dec_ell::row<2, T, T>(n, 2, j, pcol, pval, zc) : u[j];
}

// One hybrid block Gauss-Seidel sweep.  The level's vertices are numbered so
// that every block is a contiguous range [bstart[k], bstart[k+1]); the ELL rows
// are sorted by (colour, vertex), so the rows of colour c in block k are the
// positions [seg[c*(nb+1)+k], seg[c*(nb+1)+k+1]).  One workgroup per block
// stages the block's iterate in LDS and runs the colours in order (reversed
// for the post-smoother) with a barrier between colours; a neighbour in
// another block enters with its value in uin (the iterate at the start of the
// sweep).  The result goes to uout, so no workgroup reads a value another one
// writes in the same launch.  Per row: the sum over the off-diagonal entries
// in slot (ascending column) order, then (b - sum) / diagonal.
template <int W, class T, bool FOLD>
__global__ void __launch_bounds__(256)
    k_block_gs(int n, int width, int ncolors, int nb, const int *__restrict__ col, const T *__restrict__ val,
               const int *__restrict__ rowid, const int *__restrict__ bstart, const int *__restrict__ seg,
               const T *__restrict__ b, const T *__restrict__ uin, T *__restrict__ uout, int reverse,
               const int *__restrict__ pcol, const T *__restrict__ pval, const T *__restrict__ zc_arg) {
  const T *zc = FOLD ? zc_arg : nullptr;  // folded prolongation only in the FOLD instance
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *x = reinterpret_cast<T *>(smem_raw);
  int k = blockIdx.x;
  int r0 = bstart[k], nr = bstart[k + 1] - r0;
  for (int i = threadIdx.x; i < nr; i += blockDim.x) x[i] = prolonged(n, r0 + i, uin, pcol, pval, zc);
  __syncthreads();
  for (int ci = 0; ci < ncolors; ci++) {
    int c = reverse ? ncolors - 1 - ci : ci;
    int p0 = seg[c * (nb + 1) + k], p1 = seg[c * (nb + 1) + k + 1];
    for (int p = p0 + static_cast<int>(threadIdx.x); p < p1; p += blockDim.x) {
      int v = rowid[p];
      T s = static_cast<T>(0), d = static_cast<T>(0);
      if constexpr (W > 0) {
        int cc[W];
        T vv[W], xv[W];
#pragma unroll
        for (int q = 0; q < W; q++) {
          size_t o = static_cast<size_t>(q) * n + p;
          cc[q] = col[o];
          vv[q] = val[o];
        }
#pragma unroll
        for (int q = 0; q < W; q++) {
          unsigned loc = static_cast<unsigned>(cc[q] - r0);
          xv[q] = (cc[q] < 0 || loc < static_cast<unsigned>(nr)) ? static_cast<T>(0)
                                                                 : prolonged(n, cc[q], uin, pcol, pval, zc);
        }
#pragma unroll
        for (int q = 0; q < W; q++) {
          if (cc[q] < 0) continue;
          unsigned loc = static_cast<unsigned>(cc[q] - r0);
          if (cc[q] == v) d = vv[q];
          else s = s + vv[q] * (loc < static_cast<unsigned>(nr) ? x[loc] : xv[q]);
        }
      } else {
        for (int q = 0; q < width; q++) {
          size_t o = static_cast<size_t>(q) * n + p;
          int cq = col[o];
          if (cq < 0) break;
          unsigned loc = static_cast<unsigned>(cq - r0);
          if (cq == v) d = val[o];
          else s = s + val[o] * (loc < static_cast<unsigned>(nr) ? x[loc] : prolonged(n, cq, uin, pcol, pval, zc));
        }
      }
      x[v - r0] = (b[v] - s) / d;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < nr; i += blockDim.x) uout[r0 + i] = x[i];
}

// The streaming sweep with the next colour's loads in flight while the
// current colour is computed, for blocks whose (colour, block) segments hold
// at most one row per thread: thread t takes the t-th row of each segment.
// Its slots, values, b and the values of its neighbours in other blocks are
// loaded one colour ahead, unconditionally (from the segment's last row when
// the thread has none, and uin itself where a neighbour is inside the block),
// so no load waits on a branch; the barrier between colours waits for LDS
// only.  Per-row arithmetic is k_block_gs's.
template <int W, class T>
struct SweepRow {
  int v;
  bool has;
  int c[W];
  T val[W], halo[W], b;
};

template <int W, class T>
__device__ __forceinline__ void load_sweep_row(SweepRow<W, T> &o, int c, int k, int nb, int r0, int nr, int n,
                                               const int *__restrict__ col, const T *__restrict__ val,
                                               const int *__restrict__ rowid, const int *__restrict__ seg,
                                               const T *__restrict__ b, const T *__restrict__ uin) {
  int p0 = seg[c * (nb + 1) + k], p1 = seg[c * (nb + 1) + k + 1];
  int p = p0 + static_cast<int>(threadIdx.x);
  o.has = p < p1;
  int pc = p1 > p0 ? min(p, p1 - 1) : 0;
  o.v = rowid[pc];
#pragma unroll
  for (int q = 0; q < W; q++) {
    size_t off = static_cast<size_t>(q) * n + pc;
    o.c[q] = col[off];
    o.val[q] = val[off];
  }
  o.b = b[o.v];
#pragma unroll
  for (int q = 0; q < W; q++) o.halo[q] = uin[o.c[q] < 0 ? o.v : o.c[q]];
}

template <int W, class T>
__global__ void __launch_bounds__(256)
    k_block_gs_pl(int n, int ncolors, int nb, const int *__restrict__ col, const T *__restrict__ val,
                  const int *__restrict__ rowid, const int *__restrict__ bstart, const int *__restrict__ seg,
                  const T *__restrict__ b, const T *__restrict__ uin, T *__restrict__ uout, int reverse) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *x = reinterpret_cast<T *>(smem_raw);
  int k = blockIdx.x;
  int r0 = bstart[k], nr = bstart[k + 1] - r0;
  SweepRow<W, T> cur, nxt;
  load_sweep_row<W, T>(cur, reverse ? ncolors - 1 : 0, k, nb, r0, nr, n, col, val, rowid, seg, b, uin);
  for (int i = threadIdx.x; i < nr; i += blockDim.x) x[i] = uin[r0 + i];
  __syncthreads();
  for (int ci = 0; ci < ncolors; ci++) {
    if (ci + 1 < ncolors)
      load_sweep_row<W, T>(nxt, reverse ? ncolors - 2 - ci : ci + 1, k, nb, r0, nr, n, col, val, rowid, seg, b, uin);
    if (cur.has) {
      T sum = static_cast<T>(0), d = static_cast<T>(0);
#pragma unroll
      for (int q = 0; q < W; q++) {
        int cq = cur.c[q];
        if (cq < 0) continue;
        unsigned loc = static_cast<unsigned>(cq - r0);
        if (cq == cur.v) d = cur.val[q];
        else sum = sum + cur.val[q] * (loc < static_cast<unsigned>(nr) ? x[loc] : cur.halo[q]);
      }
      x[cur.v - r0] = (cur.b - sum) / d;
    }
    __syncthreads();
    cur = nxt;
  }
  for (int i = threadIdx.x; i < nr; i += blockDim.x) uout[r0 + i] = x[i];
}

// The same sweep with every memory access issued before the colour loop, so
// that the colour phases touch only LDS and registers.  Thread t owns the
// block's rows r0 + t + i·blockDim.x (i < R) and holds their slots, values,
// b and the values of their neighbours in other blocks.  With zc given, the
// sweep starts from u + P zc (the prolongation of the coarse correction,
// P an ELL of width 2, the sums of k_prolong_add), both for the block's rows
// and for those neighbours.  The per-row arithmetic is k_block_gs's.
template <int W, class T, int R>
struct OwnedRows {
  int v[R], c[R];
  int col[R][W];
  T val[R][W], halo[R][W], b[R];
};

constexpr int kPfColors = 8;  // colours the prefetching sweep supports

// Thread t owns the block's rows r0 + t + i·blockDim.x (i < R) of the
// vertex-ordered ELL, so a row's number, colour, slots, values and b need no
// lookup first; the neighbours in other blocks are gathered next.  Every load
// is unconditional (rows past the block's end load its last row, and a
// neighbour inside the block loads the row's own entry of uin), so none
// waits on a branch.
template <int W, class T, int R>
__device__ __forceinline__ void load_owned(OwnedRows<W, T, R> &o, int r0, int nr, int n,
                                           const int *__restrict__ vcol, const T *__restrict__ vval,
                                           const signed char *__restrict__ vcolor, const T *__restrict__ b,
                                           const T *__restrict__ uin, const int *__restrict__ pcol,
                                           const T *__restrict__ pval, const T *__restrict__ zc) {
#pragma unroll
  for (int i = 0; i < R; i++) {
    int q = static_cast<int>(threadIdx.x) + i * static_cast<int>(blockDim.x);
    int v = r0 + min(q, nr - 1);
    o.v[i] = q < nr ? v : -1;
    int c = vcolor[v];
    o.c[i] = q < nr ? c : -1;
#pragma unroll
    for (int s = 0; s < W; s++) {
      size_t off = static_cast<size_t>(s) * n + v;
      o.col[i][s] = vcol[off];
      o.val[i][s] = vval[off];
    }
    o.b[i] = b[v];
  }
#pragma unroll
  for (int i = 0; i < R; i++) {
    int v = r0 + min(static_cast<int>(threadIdx.x) + i * static_cast<int>(blockDim.x), nr - 1);
#pragma unroll
    for (int s = 0; s < W; s++) {
      int cq = o.col[i][s];
      bool inside = cq < 0 || static_cast<unsigned>(cq - r0) < static_cast<unsigned>(nr);
      T h = prolonged(n, inside ? v : cq, uin, pcol, pval, zc);
      o.halo[i][s] = inside ? static_cast<T>(0) : h;
    }
  }
}

// The colour phases of one sweep over the owned rows (x: the block's iterate in LDS).
template <int W, class T, int R>
__device__ __forceinline__ void sweep_owned(const OwnedRows<W, T, R> &o, T *x, int r0, int nr, int ncolors,
                                            bool reverse) {
  for (int ci = 0; ci < ncolors; ci++) {
    int c = reverse ? ncolors - 1 - ci : ci;
#pragma unroll
    for (int i = 0; i < R; i++) {
      if (o.c[i] != c) continue;
      int v = o.v[i];
      T s = static_cast<T>(0), d = static_cast<T>(0);
#pragma unroll
      for (int q = 0; q < W; q++) {
        int cq = o.col[i][q];
        if (cq < 0) continue;
        unsigned loc = static_cast<unsigned>(cq - r0);
        if (cq == v) d = o.val[i][q];
        else s = s + o.val[i][q] * (loc < static_cast<unsigned>(nr) ? x[loc] : o.halo[i][q]);
      }
      x[v - r0] = (o.b[i] - s) / d;
    }
    __syncthreads();
  }
}

template <int W, class T, int R>
__global__ void __launch_bounds__(256)
    k_block_gs_pf(int n, int ncolors, const int *__restrict__ vcol, const T *__restrict__ vval,
                  const signed char *__restrict__ vcolor, const int *__restrict__ bstart, const T *__restrict__ b,
                  const T *__restrict__ uin, T *__restrict__ uout, int reverse, const int *__restrict__ pcol,
                  const T *__restrict__ pval, const T *__restrict__ zc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *x = reinterpret_cast<T *>(smem_raw);
  int k = blockIdx.x;
  int r0 = bstart[k], nr = bstart[k + 1] - r0;
  OwnedRows<W, T, R> o;
  load_owned<W, T, R>(o, r0, nr, n, vcol, vval, vcolor, b, uin, pcol, pval, zc);
  for (int i = threadIdx.x; i < nr; i += blockDim.x) x[i] = prolonged(n, r0 + i, uin, pcol, pval, zc);
  __syncthreads();
  sweep_owned<W, T, R>(o, x, r0, nr, ncolors, reverse != 0);
  for (int i = threadIdx.x; i < nr; i += blockDim.x) uout[r0 + i] = x[i];
}

// Renumbering between the caller's vertex order and the hierarchy's own:
// perm[i] is the caller's index of internal vertex i.
template <class T>
__global__ void k_perm_gather(int n, const int *__restrict__ perm, const T *__restrict__ src, T *__restrict__ dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

template <class T>
__global__ void k_perm_scatter(int n, const int *__restrict__ perm, const T *__restrict__ src, T *__restrict__ dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}

// u += Z r with Z stored column-major (ZT[j * n + i] = Z[i, j]), in two
// launches so that the whole chip streams Z: the columns are cut into chunks of
// kCoarseChunk, one workgroup of 64 rows per (row tile, chunk) writes
// part[c * n + i] = Σ_{j in chunk c, ascending} Z[i, j] r[j], and the second
// pass adds Σ_c part[c * n + i] (c ascending) to u.  The fused coarse tail
// uses the same two sums in the same order.
constexpr int kCoarseChunk = 64;

template <class T>
__global__ void __launch_bounds__(64)
    k_coarse_partial(int n, const T *__restrict__ ZT, const T *__restrict__ r, T *__restrict__ part) {
  int i = blockIdx.x * 64 + threadIdx.x;
  int j0 = blockIdx.y * kCoarseChunk;
  int j1 = min(n, j0 + kCoarseChunk);
  if (i >= n) return;
  T acc = static_cast<T>(0);
#pragma unroll 16
  for (int j = j0; j < j1; j++) acc = acc + ZT[static_cast<size_t>(j) * n + i] * r[j];
  part[static_cast<size_t>(blockIdx.y) * n + i] = acc;
}

template <class T>
__device__ __forceinline__ T coarse_sum(int n, int nchunks, int i, const T *__restrict__ part) {
  T acc = static_cast<T>(0);
  for (int c = 0; c < nchunks; c++) acc = acc + part[static_cast<size_t>(c) * n + i];
  return acc;
}

template <class T>
__global__ void k_coarse_finish(int n, int nchunks, const T *__restrict__ part, T *__restrict__ u) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] = u[i] + coarse_sum(n, nchunks, i, part);
}

inline int coarse_chunks(int n) { return (n + kCoarseChunk - 1) / kCoarseChunk; }

template <class T>
__device__ double block_sum(double v) {
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// Deterministic dot product: fixed grid of partial sums, then one block.
template <class T>
__global__ void k_dot_partial(int n, const T *__restrict__ x, const T *__restrict__ y, double *__restrict__ part,
                              const int *__restrict__ done) {
  if (done && *done) return;
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kDotBlocks * kBlock)
    acc = acc + static_cast<double>(x[i]) * static_cast<double>(y[i]);
  double s = block_sum<T>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_sum_partials(const double *__restrict__ part, double *__restrict__ out,
                               const int *__restrict__ done) {
  if (done && *done) return;
  double s = block_sum<double>(threadIdx.x < kDotBlocks ? part[threadIdx.x] : 0.0);
  if (threadIdx.x == 0) *out = s;
}

// Krylov.jl cg recurrence with device-resident scalars (warm start,
// atol = rtol = sqrt(eps), zero-curvature exit).  γ and ‖p‖² are kept in two
// slots: iteration `it` reads slot it & 1 and writes the other, so that the
// blocks of one launch never see a value written in that launch.
struct CgState {
  double gamma[2], pnorm2[2], rnorm, eps_stop, alpha, beta;
  int done, iter;
};

// Levels above kCgSmallRows rows: two launches per iteration.  The scalar
// steps run inside the vector kernels: every block reduces all partial sums of
// the previous launch itself, in one fixed order, so all blocks hold the same
// γ, β and α, and block 0 alone writes the scalars back.  p is double-
// buffered: the direction launch of iteration it > 0 forms p = r + β p_old for
// its own rows into the other buffer, and its product evaluates r + β p_old at
// every column, the expression the column's own row stores, so A p is the
// product of the stored p without a launch in between.  The last step launch
// adds the correction to u.  Against the four-launch sequence (product, dot,
// α step, β step) the per-row values are the same expressions; only the dot
// products add in another partition, so the iterates agree to rounding.
constexpr int kCgBlocksMax = 1024;  // partial sums per reduction

inline int cg_blocks(int n) { return std::min(kCgBlocksMax, dec::nblocks(n, kBlock)); }

// Σ part[0..np) in a fixed order, the same in every workgroup: thread t adds
// a run of ceil(np / kBlock) consecutive partials, then block_sum's tree
// (every thread has read the result before a later block_sum of the kernel
// writes the same LDS array)
__device__ double sum_partials(const double *__restrict__ part, int np) {
  constexpr int kRun = 16;  // runs up to this long load all their partials at once
  int per = (np + kBlock - 1) / kBlock;
  int j0 = min(np, static_cast<int>(threadIdx.x) * per), j1 = min(np, j0 + per);
  double acc = 0.0;
  if (per <= kRun) {
    double v[kRun];
#pragma unroll
    for (int k = 0; k < kRun; k++) v[k] = part[min(j0 + k, np - 1)];
#pragma unroll
    for (int k = 0; k < kRun; k++)
      if (j0 + k < j1) acc = acc + v[k];
  } else {
    for (int j = j0; j < j1; j++) acc = acc + part[j];
  }
  double s = block_sum<double>(acc);
  __syncthreads();
  return s;
}

// (A p)_i with p = r + β q evaluated at each column, summed in slot order.
template <int W>
__device__ __forceinline__ double row_dir(int n, int width, int i, const int *__restrict__ col,
                                          const double *__restrict__ val, const double *__restrict__ r,
                                          const double *__restrict__ q, double beta) {
  double acc = 0.0;
  if constexpr (W > 0) {
    int c[W];
    double v[W], rv[W], qv[W];
#pragma unroll
    for (int k = 0; k < W; k++) {
      size_t o = static_cast<size_t>(k) * n + i;
      c[k] = col[o];
      v[k] = val[o];
    }
#pragma unroll
    for (int k = 0; k < W; k++) {
      int cc = c[k] < 0 ? 0 : c[k];
      rv[k] = r[cc];
      qv[k] = q[cc];
    }
#pragma unroll
    for (int k = 0; k < W; k++)
      if (c[k] >= 0) acc = acc + v[k] * (rv[k] + beta * qv[k]);
  } else {
    for (int k = 0; k < width; k++) {
      size_t o = static_cast<size_t>(k) * n + i;
      int cc = col[o];
      if (cc < 0) break;
      acc = acc + val[o] * (r[cc] + beta * q[cc]);
    }
  }
  return acc;
}

// *f as read by thread 0, the same for every thread of the block.  Block 0's
// thread 0 may write *f later in the same launch; a block that reads that
// value takes the exit its own scalars would take, and reading once keeps the
// waves of a block on one path through its barriers.
__device__ int block_flag(const int *f) {
  __shared__ int v;
  if (threadIdx.x == 0) v = *f;
  __syncthreads();
  return v;
}

// r = b - A u, p = r, x = 0, and the partial sums of r·r.
template <int W>
__global__ void __launch_bounds__(kBlock)
    k_cg2_start(int n, int width, const int *__restrict__ vcol, const double *__restrict__ vval,
                const double *__restrict__ b, const double *__restrict__ u, double *__restrict__ r,
                double *__restrict__ p, double *__restrict__ x, double *__restrict__ part) {
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double ri = b[i] - dec_ell::row<W, double, double>(n, width, i, vcol, vval, u);
    r[i] = ri;
    p[i] = ri;
    x[i] = 0.0;
    acc = acc + ri * ri;
  }
  double sum = block_sum<double>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

// Iteration it's direction: at it = 0, γ = r·r, ‖r‖, the stopping threshold
// and the immediate exit; at it > 0, ‖r‖ and the solved exit, β = γ_new / γ,
// ‖p‖² and γ to slot it & 1, and p = r + β p_old.  Then A p and the partial
// sums of p·Ap.
template <int W>
__global__ void __launch_bounds__(kBlock)
    k_cg2_dir(CgState *st, int it, int n, int width, const int *__restrict__ vcol, const double *__restrict__ vval,
              const double *__restrict__ r, const double *__restrict__ pold, double *__restrict__ p,
              double *__restrict__ Ap, const double *__restrict__ part_rr, int nb, double *__restrict__ part_pap) {
  if (it > 0 && block_flag(&st->done)) return;
  const int stride = gridDim.x * kBlock, i0 = blockIdx.x * kBlock + threadIdx.x;
  // The thread's first row is loaded before the scalar step, which does not
  // need it, so its loads are in flight while the partial sums are reduced.
  constexpr int WP = W > 0 ? W : 1;
  int c0[WP];
  double v0[WP], rc0[WP], qc0[WP], ri0 = 0.0, qi0 = 0.0;
  if constexpr (W > 0) {
    int ia = min(i0, n - 1);
#pragma unroll
    for (int k = 0; k < W; k++) {
      size_t o = static_cast<size_t>(k) * n + ia;
      c0[k] = vcol[o];
      v0[k] = vval[o];
    }
#pragma unroll
    for (int k = 0; k < W; k++) {
      int cc = c0[k] < 0 ? 0 : c0[k];
      rc0[k] = r[cc];
      qc0[k] = pold[cc];
    }
    ri0 = r[ia];
    qi0 = pold[ia];
  }
  double g = sum_partials(part_rr, nb);
  double rn = sqrt(g);
  bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  double beta = 0.0;
  if (it == 0) {
    const double sq = 1.4901161193847656e-08;
    double es = sq + sq * rn;
    bool stop = (g == 0.0) || (rn <= es);
    if (lead) {
      st->gamma[0] = g;
      st->rnorm = rn;
      st->eps_stop = es;
      st->pnorm2[0] = g;
      st->iter = 0;
      st->done = stop;
    }
    if (stop) return;
  } else {
    int c = (it - 1) & 1;
    if (rn <= st->eps_stop || rn + 1.0 <= 1.0) {
      if (lead) {
        st->rnorm = rn;
        st->iter += 1;
        st->done = 2;  // solved: p is left as is
      }
      return;
    }
    beta = g / st->gamma[c];
    if (lead) {
      st->rnorm = rn;
      st->iter += 1;
      st->beta = beta;
      st->pnorm2[c ^ 1] = g + beta * beta * st->pnorm2[c];
      st->gamma[c ^ 1] = g;
    }
  }
  double acc = 0.0;
  int i = i0;
  if constexpr (W > 0) {
    // the first row from the registers, with the sums of dec_ell::row and
    // row_dir (slot order, the same products)
    if (i0 < n) {
      double pi, api = 0.0;
      if (it == 0) {
        pi = qi0;
#pragma unroll
        for (int k = 0; k < W; k++)
          if (c0[k] >= 0) api = api + v0[k] * qc0[k];
      } else {
        pi = ri0 + beta * qi0;
        p[i0] = pi;
#pragma unroll
        for (int k = 0; k < W; k++)
          if (c0[k] >= 0) api = api + v0[k] * (rc0[k] + beta * qc0[k]);
      }
      Ap[i0] = api;
      acc = acc + pi * api;
    }
    i += stride;
  }
  for (; i < n; i += stride) {
    double pi, api;
    if (it == 0) {
      pi = pold[i];
      api = dec_ell::row<W, double, double>(n, width, i, vcol, vval, pold);
    } else {
      pi = r[i] + beta * pold[i];
      p[i] = pi;
      api = row_dir<W>(n, width, i, vcol, vval, r, pold, beta);
    }
    Ap[i] = api;
    acc = acc + pi * api;
  }
  double sum = block_sum<double>(acc);
  if (threadIdx.x == 0) part_pap[blockIdx.x] = sum;
}

// Iteration it's step: α = γ / p·Ap (zero-curvature exit), x += α p,
// r -= α Ap and the partial sums of r·r; on the last iteration (LAST) u += x
// instead (x as it stands when an exit came earlier).  The thread's first
// element is loaded before the scalar step, so those loads are in flight
// while the partial sums are reduced.
template <bool LAST>
__global__ void __launch_bounds__(kBlock)
    k_cg2_step(CgState *st, int it, int n, double *__restrict__ x, const double *__restrict__ p,
               double *__restrict__ r, const double *__restrict__ Ap, const double *__restrict__ part_pap, int nb,
               double *__restrict__ part_rr, double *__restrict__ u) {
  const double eps = 2.220446049250313e-16;
  const int stride = gridDim.x * kBlock, i0 = blockIdx.x * kBlock + threadIdx.x, ia = min(i0, n - 1);
  double x0 = x[ia], p0 = p[ia], r0 = 0.0, a0 = 0.0, u0 = 0.0;
  if constexpr (LAST) {
    u0 = u[ia];
  } else {
    r0 = r[ia];
    a0 = Ap[ia];
  }
  bool stop = block_flag(&st->done) != 0;
  double a = 0.0;
  if (!stop) {
    double pAp = sum_partials(part_pap, nb);
    int c = it & 1;
    double pn2 = st->pnorm2[c];
    if (pAp <= eps * pn2 && fabs(pAp) <= eps * pn2) {
      if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
      stop = true;
    } else {
      a = st->gamma[c] / pAp;
      if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = a;
    }
  }
  if constexpr (LAST) {
    if (i0 < n) {
      double xi = stop ? x0 : x0 + a * p0;
      u[i0] = xi + u0;
    }
    for (int i = i0 + stride; i < n; i += stride) {
      double xi = stop ? x[i] : x[i] + a * p[i];
      u[i] = xi + u[i];
    }
    return;
  }
  if (stop) return;
  double acc = 0.0;
  if (i0 < n) {
    x[i0] = x0 + a * p0;
    double ri = r0 - a * a0;
    r[i0] = ri;
    acc = acc + ri * ri;
  }
  for (int i = i0 + stride; i < n; i += stride) {
    x[i] = x[i] + a * p[i];
    double ri = r[i] - a * Ap[i];
    r[i] = ri;
    acc = acc + ri * ri;
  }
  double sum = block_sum<double>(acc);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = sum;
}

// The CG smoother (the recurrence of k_cg2_*) on a level of at most
// kCgSmallRows rows, as one workgroup of kCgThreads threads: thread t owns
// rows t + kCgThreads k (k < R) and holds their slots, values, x, r and A p
// in registers for the whole smoothing; p lives in LDS, so after the first
// residual every product is LDS gathers only.  Each dot product is one
// workgroup reduction in a fixed order (waves, then the waves' sums).  One
// launch per smoothing in place of 1 + 2 steps launches.  Its dot products
// add in another order than the multi-launch partition, so its iterates
// agree with those of the multi-launch smoother to rounding.
constexpr int kCgThreads = 1024;
constexpr int kCgRowsPerThread = 3;  // slots and values of 3 rows of up to 7 fit in 128 VGPRs
constexpr int kCgSmallRows = kCgThreads * kCgRowsPerThread;

// The workgroup's sum of v: each wave's shuffle tree, then the 16 wave sums
// in order, read by every thread.  Successive calls alternate between the two
// halves of red[2][16]: a thread writes a half again only two calls later,
// after the barrier of the call between, by which time every thread has read
// it, so one barrier per call suffices.
__device__ double wg_sum(double v, double (*red)[kCgThreads / 64], int &half) {
  for (int off = 32; off > 0; off >>= 1) v = v + __shfl_down(v, off, 64);
  double *h = red[half];
  half ^= 1;
  if ((threadIdx.x & 63) == 0) h[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < kCgThreads / 64; k++) t = t + h[k];
  return t;
}

// Σ_s v[s] x[c[s]] over the slots with c[s] ≥ 0, in slot order (the sum of
// dec_ell::row).
template <int W>
__device__ __forceinline__ double row_regs(const int (&c)[W], const double (&v)[W], const double *x) {
  double xv[W];
#pragma unroll
  for (int s = 0; s < W; s++) xv[s] = x[c[s] < 0 ? 0 : c[s]];
  double acc = 0.0;
#pragma unroll
  for (int s = 0; s < W; s++)
    if (c[s] >= 0) acc = acc + v[s] * xv[s];
  return acc;
}

template <int W, int R>
__global__ void __launch_bounds__(kCgThreads)
    k_cg_smooth_small(int n, const int *__restrict__ vcol, const double *__restrict__ vval,
                      const double *__restrict__ b, double *__restrict__ u, int steps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *p = reinterpret_cast<double *>(smem_raw);
  __shared__ double red[2][kCgThreads / 64];
  int half = 0;
  int c[R][W];
  double v[R][W], x[R], r[R], Ap[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    int i = static_cast<int>(threadIdx.x) + k * kCgThreads;
#pragma unroll
    for (int s = 0; s < W; s++) {
      size_t o = static_cast<size_t>(s) * n + (i < n ? i : 0);
      c[k][s] = i < n ? vcol[o] : -1;
      v[k][s] = vval[o];
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < R; k++) {
    int i = static_cast<int>(threadIdx.x) + k * kCgThreads;
    x[k] = 0.0;
    r[k] = 0.0;
    Ap[k] = 0.0;
    if (i < n) {
      r[k] = b[i] - row_regs<W>(c[k], v[k], u);
      p[i] = r[k];
      acc = acc + r[k] * r[k];
    }
  }
  double gamma = wg_sum(acc, red, half);
  double rnorm = sqrt(gamma);
  const double sq = 1.4901161193847656e-08, eps = 2.220446049250313e-16;
  const double eps_stop = sq + sq * rnorm;
  double pnorm2 = gamma;
  bool done = (gamma == 0.0) || (rnorm <= eps_stop);
  for (int it = 0; it < steps && !done; it++) {
    __syncthreads();  // p complete
    double pap = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      int i = static_cast<int>(threadIdx.x) + k * kCgThreads;
      if (i < n) {
        Ap[k] = row_regs<W>(c[k], v[k], p);
        pap = pap + p[i] * Ap[k];
      }
    }
    double pAp = wg_sum(pap, red, half);
    if (pAp <= eps * pnorm2 && fabs(pAp) <= eps * pnorm2) break;
    double a = gamma / pAp;
    double rr = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) {
      int i = static_cast<int>(threadIdx.x) + k * kCgThreads;
      if (i < n) {
        x[k] = x[k] + a * p[i];
        r[k] = r[k] - a * Ap[k];
        rr = rr + r[k] * r[k];
      }
    }
    double gn = wg_sum(rr, red, half);
    double rn = sqrt(gn);
    if (rn <= eps_stop || rn + 1.0 <= 1.0) break;
    double beta = gn / gamma;
    pnorm2 = gn + beta * beta * pnorm2;
    gamma = gn;
#pragma unroll
    for (int k = 0; k < R; k++) {
      int i = static_cast<int>(threadIdx.x) + k * kCgThreads;
      if (i < n) p[i] = r[k] + beta * p[i];
    }
  }
#pragma unroll
  for (int k = 0; k < R; k++) {
    int i = static_cast<int>(threadIdx.x) + k * kCgThreads;
    if (i < n) u[i] = x[k] + u[i];
  }
}

__global__ void k_d2f(int n, const double *__restrict__ x, float *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = static_cast<float>(x[i]);
}

__global__ void k_add_f2d(int n, const float *__restrict__ c, double *__restrict__ u) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] = u[i] + static_cast<double>(c[i]);
}

__global__ void k_f2d(int n, const float *__restrict__ c, double *__restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = static_cast<double>(c[i]);
}

// ---------------------------------------------- V-cycle-preconditioned CG
// Conjugate gradients on the finest level with one V-cycle (from a zero
// start) as the preconditioner z = M r:
//   β = (r·z) / (r·z)_prev,  p = z + β p,  q = A p,  α = (r·z) / (p·q),
//   x += α p,  r -= α q.
// Three launches per iteration after the V-cycle, one element per thread,
// with the scalars on the device: every workgroup adds the previous launch's
// per-workgroup partial sums itself, in one fixed order (sum_partials), so all
// hold the same β or α.  Vectors are indexed by the hierarchy's vertex
// numbers; the product runs over the level's rows in vertex order.
// sc[0] = (r·z)_prev (1 before the first iteration, when p = 0),
// sc[1] = r·z, sc[2] = α: the direction launch reads sc[0] and its block 0
// writes sc[1]; a one-workgroup launch forms α and copies sc[1] to sc[0];
// the update launch reads α.  No launch with several workgroups reads a slot
// it writes.  The update launch also prepares the next V-cycle's input: z = 0
// (float64 preconditioner), or the float32 r and a zero float32 iterate
// (float32 preconditioner).


// kPcgDotBlocks partial sums of r·z over a grid-stride loop (z32 non-null:
// z = the float32 preconditioner's result, stored in float64 first)
constexpr int kPcgDotBlocks = kBlock;

__global__ void __launch_bounds__(kBlock) k_pcg_dot(int n, const double *__restrict__ r, const float *__restrict__ z32,
                                                    double *__restrict__ z, double *__restrict__ part) {
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kPcgDotBlocks * kBlock) {
    double zi;
    if (z32) {
      zi = static_cast<double>(z32[i]);
      z[i] = zi;
    } else {
      zi = z[i];
    }
    acc = acc + r[i] * zi;
  }
  double s = block_sum<double>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// α = (r·z) / (p·q) into sc[2] (zero when p·q is not positive), from the
// direction launch's per-workgroup partial sums of p·q; sc[0] = sc[1] for
// the next iteration's β
__global__ void __launch_bounds__(kBlock) k_pcg_alpha(const double *__restrict__ part_pq, int np,
                                                      double *__restrict__ sc) {
  double pq = sum_partials(part_pq, np);
  if (threadIdx.x != 0) return;
  double rz = sc[1];
  double alpha = pq > 0.0 ? rz / pq : 0.0;
  if (!(alpha == alpha) || fabs(alpha) > 1e300) alpha = 0.0;
  sc[2] = alpha;
  sc[0] = rz;
}

// β, then per row p = z + β p and q = A z + β q (= A p, the product of the
// previous direction carried over, so one vector is gathered; slots in
// ascending column order), in place, and the partial sums of p·q
template <int W>
__global__ void __launch_bounds__(kBlock)
    k_pcg_dir(int n, int width, const int *__restrict__ col, const double *__restrict__ val,
              const int *__restrict__ rowid, const double *__restrict__ z, double *__restrict__ p,
              double *__restrict__ q, const double *__restrict__ part_rz, int np, double *__restrict__ sc,
              double *__restrict__ part_pq) {
  double rz = sum_partials(part_rz, np);
  double beta = rz / sc[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[1] = rz;
  int pos = dec_ell::xcd_block() * kBlock + threadIdx.x;
  double acc = 0.0;
  if (pos < n) {
    double s = dec_ell::row<W, double, double>(n, width, pos, col, val, z);
    int v = rowid ? rowid[pos] : pos;
    double pv = z[v] + beta * p[v];
    double qv = s + beta * q[v];
    p[v] = pv;
    q[v] = qv;
    acc = pv * qv;
  }
  double t = block_sum<double>(acc);
  if (threadIdx.x == 0) part_pq[blockIdx.x] = t;
}

// x += α p, r -= α q, the partial sums of r·r, and the next V-cycle's input
__global__ void __launch_bounds__(kBlock)
    k_pcg_update(int n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                 const double *__restrict__ q, const double *__restrict__ sc, double *__restrict__ part_rr,
                 double *__restrict__ z, float *__restrict__ b32, float *__restrict__ u32) {
  const double alpha = sc[2];
  int i = blockIdx.x * kBlock + threadIdx.x;
  double acc = 0.0;
  if (i < n) {
    x[i] = x[i] + alpha * p[i];
    double ri = r[i] - alpha * q[i];
    r[i] = ri;
    acc = ri * ri;
    if (z) z[i] = 0.0;
    if (b32) {
      b32[i] = static_cast<float>(ri);
      u32[i] = 0.0f;
    }
  }
  double t = block_sum<double>(acc);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = t;
}

// (re)start: r = src, p = q = 0, sc[0] = 1 (the first direction is z), and
// the first V-cycle's input as k_pcg_update prepares it
__global__ void k_pcg_start(int n, const double *__restrict__ src, double *__restrict__ r, double *__restrict__ p,
                            double *__restrict__ q, double *__restrict__ sc, double *__restrict__ z,
                            float *__restrict__ b32, float *__restrict__ u32) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double ri = src[i];
    r[i] = ri;
    p[i] = 0.0;
    q[i] = 0.0;
    if (z) z[i] = 0.0;
    if (b32) {
      b32[i] = static_cast<float>(ri);
      u32[i] = 0.0f;
    }
  }
  if (i == 0) {
    sc[0] = 1.0;
    sc[1] = 0.0;
    sc[2] = 0.0;
  }
}


// ------------------------------------------------------------- fused coarse tail
// The levels whose size is at most fuse_rows (DEC_MG_FUSE_ROWS, default
// kFuseRowsDefault, at most kFuseRowsMax) run as one workgroup of 1024
// threads: every phase of the V-cycle below that level (colour sweeps,
// residual, restriction, prolongation) is a strided loop over the level's
// rows followed by a workgroup barrier, in place of one kernel launch per
// phase.  Per-row arithmetic is the same as the multi-kernel path (same
// expressions, same order).  One workgroup streams at a small fraction of
// the chip's bandwidth, so this pays only for levels of a few thousand rows.
constexpr int kFuseRowsDefault = 3000;
constexpr int kFuseRowsMax = 4096;  // levels up to this size are one block (LDS)
constexpr int kFuseThreads = 1024;
constexpr int kMaxColors = 16;

template <class T>
struct LevelDesc {
  int n, width, ncolors, nc, rwidth, pwidth;
  int cstart[kMaxColors + 1];
  const int *col, *rowid, *rcol, *pcol;
  const T *val, *rval, *pval;
  T *u, *b, *r;
};

template <class T>
__device__ __forceinline__ T fused_row(const LevelDesc<T> &L, int p, const T *x) {
  return L.width == 7 ? dec_ell::row<7, T, T>(L.n, L.width, p, L.col, L.val, x)
                      : dec_ell::row<0, T, T>(L.n, L.width, p, L.col, L.val, x);
}

template <class T>
__device__ void fused_smooth(const LevelDesc<T> &L, const T *b, T *u, int steps, int alg, T omega, bool post) {
  for (int it = 0; it < steps; it++) {
    if (alg == DEC_SMOOTH_GS || alg == DEC_SMOOTH_BLOCK_GS) {
      // a level in the fused tail is one block: the hybrid sweep is the
      // multicolour sweep
      for (int ci = 0; ci < L.ncolors; ci++) {
        int c = post ? L.ncolors - 1 - ci : ci;
        for (int p = L.cstart[c] + threadIdx.x; p < L.cstart[c + 1]; p += blockDim.x) {
          int v = L.rowid[p];
          T sum, d;
          if (L.width == 7) dec_ell::gs_row<7, T>(L.n, L.width, p, v, L.col, L.val, u, sum, d);
          else dec_ell::gs_row<0, T>(L.n, L.width, p, v, L.col, L.val, u, sum, d);
          u[v] = (b[v] - sum) / d;
        }
        __syncthreads();
      }
    } else {
      for (int p = threadIdx.x; p < L.n; p += blockDim.x) {
        int v = L.rowid[p];
        T sum, d;
        if (L.width == 7) dec_ell::gs_row<7, T>(L.n, L.width, p, v, L.col, L.val, u, sum, d);
        else dec_ell::gs_row<0, T>(L.n, L.width, p, v, L.col, L.val, u, sum, d);
        T full = fused_row(L, p, u);
        L.r[v] = u[v] + omega * ((b[v] - full) / d);
      }
      __syncthreads();
      for (int i = threadIdx.x; i < L.n; i += blockDim.x) u[i] = L.r[i];
      __syncthreads();
    }
  }
}

// phase 0: the whole V-cycle below l0, the coarsest level smoothed.
// phase 1: the down sweep; at the coarsest level only its residual is formed
//          (into Ls[nl-1].r) for the dense direct solve, which runs as
//          k_coarse_partial on the whole chip.
// phase 2: the coarsest correction u += Σ_c part[c] (k_coarse_finish's sum),
//          then the up sweep.
template <class T>
__global__ void __launch_bounds__(kFuseThreads)
    k_fused_vcycle(const LevelDesc<T> *__restrict__ Ls, int l0, int nl, T *u0, const T *b0, int steps, int alg,
                   T omega, int phase, const T *__restrict__ part) {
  if (phase != 2) {
    for (int l = l0; l < nl; l++) {
      const LevelDesc<T> &L = Ls[l];
      T *u = l == l0 ? u0 : L.u;
      const T *b = l == l0 ? b0 : L.b;
      if (l > l0) {
        for (int i = threadIdx.x; i < L.n; i += blockDim.x) u[i] = static_cast<T>(0);
        __syncthreads();
      }
      if (l == nl - 1 && phase == 1) {
        for (int p = threadIdx.x; p < L.n; p += blockDim.x) {
          int v = L.rowid[p];
          L.r[v] = b[v] - fused_row(L, p, u);
        }
        return;
      }
      fused_smooth(L, b, u, steps, alg, omega, false);
      if (l == nl - 1) break;
      for (int p = threadIdx.x; p < L.n; p += blockDim.x) {
        int v = L.rowid[p];
        L.r[v] = b[v] - fused_row(L, p, u);
      }
      __syncthreads();
      const LevelDesc<T> &C = Ls[l + 1];
      for (int i = threadIdx.x; i < L.nc; i += blockDim.x) {
        T acc = L.rwidth == 7 ? dec_ell::row<7, T, T>(L.nc, L.rwidth, i, L.rcol, L.rval, L.r)
                              : dec_ell::row<0, T, T>(L.nc, L.rwidth, i, L.rcol, L.rval, L.r);
        C.b[i] = acc;
      }
      __syncthreads();
    }
  } else {
    const LevelDesc<T> &C = Ls[nl - 1];
    T *uc = nl - 1 == l0 ? u0 : C.u;
    int nch = (C.n + kCoarseChunk - 1) / kCoarseChunk;
    for (int i = threadIdx.x; i < C.n; i += blockDim.x) uc[i] = uc[i] + coarse_sum(C.n, nch, i, part);
    __syncthreads();
  }
  for (int l = nl - 2; l >= l0; l--) {
    const LevelDesc<T> &L = Ls[l];
    const LevelDesc<T> &C = Ls[l + 1];
    T *u = l == l0 ? u0 : L.u;
    const T *b = l == l0 ? b0 : L.b;
    for (int j = threadIdx.x; j < L.n; j += blockDim.x)
      u[j] = u[j] + (L.pwidth == 2 ? dec_ell::row<2, T, T>(L.n, L.pwidth, j, L.pcol, L.pval, C.u)
                                   : dec_ell::row<0, T, T>(L.n, L.pwidth, j, L.pcol, L.pval, C.u));
    __syncthreads();
    fused_smooth(L, b, u, steps, alg, omega, true);
  }
}

// The coarse tail when it is one smoothed level A (at most 3 x 1024 rows,
// ELL width 1-8) above the dense coarsest level C, in Gauss-Seidel V-cycles:
// the same phases and per-row arithmetic as k_fused_vcycle with A's rows held
// in registers (thread t owns positions t + 1024 i) and its vectors in LDS, so
// the colour phases wait on no memory access.
//   phase 1: A's pre-smoothing, residual and restriction; C's b and its
//            residual (C's iterate starts at zero, so that residual is b).
//   phase 2: C's correction from k_coarse_partial's chunks, the prolongation
//            to A and A's post-smoothing.
template <int W, class T, int R>
__global__ void __launch_bounds__(kFuseThreads)
    k_tail2(const LevelDesc<T> A, const LevelDesc<T> C, T *__restrict__ uA, const T *__restrict__ bA, int steps,
            int phase, const T *__restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int n = A.n;
  T *x = reinterpret_cast<T *>(smem_raw);
  T *w = x + n;  // phase 1: A's residual; phase 2: C's correction
  OwnedRows<W, T, R> o;
#pragma unroll
  for (int i = 0; i < R; i++) {
    int q = static_cast<int>(threadIdx.x) + i * kFuseThreads;
    o.v[i] = -1;
    o.c[i] = -1;
    if (q >= n) continue;
    int c = 0;
    while (c + 1 < A.ncolors && q >= A.cstart[c + 1]) c++;
    o.c[i] = c;
    o.v[i] = A.rowid[q];
#pragma unroll
    for (int sl = 0; sl < W; sl++) {
      size_t off = static_cast<size_t>(sl) * n + q;
      o.col[i][sl] = A.col[off];
      o.val[i][sl] = A.val[off];
    }
    o.b[i] = bA[o.v[i]];
  }
  if (phase == 1) {
    // the restriction row this thread forms at the end, loaded up front
    // (one coarse row per thread when R is at most 8 wide and C has at most
    // 1024 rows; else the rows are read at the end)
    bool rpre = A.rwidth <= 8 && A.nc <= kFuseThreads;
    int rcl[8];
    T rvl[8];
    int ri = threadIdx.x;
#pragma unroll
    for (int sl = 0; sl < 8; sl++) {
      rcl[sl] = -1;
      rvl[sl] = static_cast<T>(0);
      if (rpre && ri < A.nc && sl < A.rwidth) {
        size_t off = static_cast<size_t>(sl) * A.nc + ri;
        rcl[sl] = A.rcol[off];
        rvl[sl] = A.rval[off];
      }
    }
    for (int i = threadIdx.x; i < n; i += kFuseThreads) x[i] = uA[i];
    __syncthreads();
    for (int it = 0; it < steps; it++) sweep_owned<W, T, R>(o, x, 0, n, A.ncolors, false);
#pragma unroll
    for (int i = 0; i < R; i++) {
      if (o.v[i] < 0) continue;
      T acc = static_cast<T>(0);
#pragma unroll
      for (int sl = 0; sl < W; sl++)
        if (o.col[i][sl] >= 0) acc = acc + o.val[i][sl] * x[o.col[i][sl]];
      w[o.v[i]] = o.b[i] - acc;
    }
    __syncthreads();
    if (rpre) {
      if (ri < A.nc) {
        T bc = static_cast<T>(0);
#pragma unroll
        for (int sl = 0; sl < 8; sl++)
          if (rcl[sl] >= 0) bc = bc + rvl[sl] * w[rcl[sl]];
        C.b[ri] = bc;
        C.r[ri] = bc;
      }
    } else {
      for (int i = threadIdx.x; i < A.nc; i += kFuseThreads) {
        T bc = A.rwidth == 7 ? dec_ell::row<7, T, T>(A.nc, A.rwidth, i, A.rcol, A.rval, w)
                             : dec_ell::row<0, T, T>(A.nc, A.rwidth, i, A.rcol, A.rval, w);
        C.b[i] = bc;
        C.r[i] = bc;
      }
    }
    for (int i = threadIdx.x; i < n; i += kFuseThreads) uA[i] = x[i];
    return;
  }
  int nch = (C.n + kCoarseChunk - 1) / kCoarseChunk;
  for (int i = threadIdx.x; i < C.n; i += kFuseThreads) w[i] = static_cast<T>(0) + coarse_sum(C.n, nch, i, part);
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += kFuseThreads)
    x[j] = uA[j] + (A.pwidth == 2 ? dec_ell::row<2, T, T>(n, A.pwidth, j, A.pcol, A.pval, w)
                                  : dec_ell::row<0, T, T>(n, A.pwidth, j, A.pcol, A.pval, w));
  __syncthreads();
  for (int it = 0; it < steps; it++) sweep_owned<W, T, R>(o, x, 0, n, A.ncolors, true);
  for (int i = threadIdx.x; i < n; i += kFuseThreads) uA[i] = x[i];
}

// ---------------------------------------------------------------- storage

template <class T>
struct LevelVals {
  DBuf<T> val;   // ELL values of A (colour-sorted rows)
  DBuf<T> vval;  // the same rows in vertex order (residuals and the CG product)
  DBuf<T> rval;  // ELL values of R (this level -> next coarser)
  DBuf<T> pval;  // ELL values of P (next coarser -> this level)
  DBuf<T> Z;     // dense coarse operator, column-major (coarsest level, direct solve)
  DBuf<T> zpart; // chunk partial products of the direct solve (coarsest level)
  DBuf<T> u, b, r, t;  // work vectors
};

// Every level keeps its own vertex numbering (internal numbers): vertices are
// grouped into blocks grown breadth-first over the operator's graph, each
// block a contiguous range, so that a block is a compact patch of the mesh
// (few neighbours outside it) and its vectors are contiguous.  The caller's
// order is restored at the boundary (solve, cycles, full multigrid); host
// copies and exports stay in the caller's order.  A level of at most
// fuse_rows rows is one block and keeps the caller's order.
struct Level {
  int n = 0, width = 0;
  DBuf<int32_t> col, rowid;  // ELL of A over internal numbers, rows sorted by (colour, vertex)
  DBuf<int32_t> vcol;        // the same ELL with rows in vertex order (same slots)
  std::vector<int> cstart;  // colour ranges of positions
  int ncolors = 0;
  // restriction to the next coarser level (rows: coarse internal numbers)
  int nc = 0, rwidth = 0;
  DBuf<int32_t> rcol;
  // prolongation from the next coarser level (rows: this level's internal numbers)
  int pwidth = 0;
  DBuf<int32_t> pcol;
  LevelVals<double> d;
  LevelVals<float> f;
  dec::HostCsr A, R, P;  // host copies for export (caller's numbering)
  std::vector<int32_t> color;  // Gauss-Seidel colour of each vertex (caller's numbering)
  std::vector<double> Zh;
  // numbering: perm[i] = caller's index of internal vertex i, iperm its inverse
  std::vector<int32_t> perm, iperm;
  bool identity = true;
  DBuf<int32_t> dperm;
  // hybrid block Gauss-Seidel: block ranges and (colour, block) ELL segments
  int nblk = 0, bmax = 0, seg_threads = 64, seg_max = 0;  // seg_max: rows of the largest (colour, block) segment
  DBuf<int32_t> bstart, seg;
  DBuf<signed char> vcolor;  // colour of each internal vertex (prefetching sweep)
  std::vector<int32_t> hbstart;
};

}  // namespace

struct dec_mg_s {
  std::vector<std::unique_ptr<dec_mesh_s>> meshes;  // finest first
  std::vector<std::unique_ptr<Level>> lv;            // finest first
  int alg = DEC_SMOOTH_CG, steps = 3, coarse_direct = 0;
  double omega = 2.0 / 3.0;
  DBuf<double> part, part2, scal;  // reduction partials, scalar outputs
  DBuf<CgState> cg;
  DBuf<double> cgx, cgp, cgp2, cgAp;  // CG vectors (finest size; p double-buffered)
  int fuse_from = -1;           // first level handled by the fused coarse kernel
  DBuf<LevelDesc<double>> desc64;
  DBuf<LevelDesc<float>> desc32;
  std::vector<LevelDesc<double>> hdesc64;  // host copies (kernel arguments)
  std::vector<LevelDesc<float>> hdesc32;
  bool use_fused = true;
  int fuse_rows = kFuseRowsDefault;
  // graph of one solve cycle or preconditioned-CG iteration (per variant)
  // for fixed u, b
  static constexpr int kGraphs = 4;
  hipGraphExec_t gexec[kGraphs] = {};
  hipGraph_t graph[kGraphs] = {};
  const void *gu[kGraphs] = {}, *gb[kGraphs] = {};
  int gkey[kGraphs] = {-1, -1, -1, -1};
  // preconditioned CG on the finest level: residual, preconditioned residual,
  // direction, A times the direction, the partial sums
  // of r·z, p·q and r·r (at most one per workgroup of the finest level's
  // element launches) and the scalars (r·z)_prev, r·z, α
  DBuf<double> pcg_r, pcg_z, pcg_p, pcg_q, pcg_part, pcg_sc;
  hipStream_t cap = nullptr;
  bool use_graph = true;
  // residual norms of a solve on their way to the host (pinned ring + events)
  DBuf<double> rpart;  // per-workgroup partial sums of the finest residual's r²
  // (slot kNormSlots: ||b||^2 of the solve)
  double *hnorm = nullptr;
  hipEvent_t nev[kNormSlots + 1] = {};
  // learned smoother (float32 cycles, levels of learned_min_rows to
  // learned_max_rows rows; learned_max_rows <= 0: no upper bound)
  dec_nn_s *nn = nullptr;
  int learned = 0, learned_min_rows = 16385, learned_max_rows = 0, learned_apps = 1;
  bool learned_active = false;
  int fallbacks = 0, learned_cycles = 0;
  ~dec_mg_s() {
    for (int i = 0; i < kGraphs; i++) {
      if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
      if (graph[i]) (void)hipGraphDestroy(graph[i]);
    }
    if (cap) (void)hipStreamDestroy(cap);
    for (auto &e : nev)
      if (e) (void)hipEventDestroy(e);
    if (hnorm) (void)hipHostFree(hnorm);
  }
};

namespace {

inline dim3 G(long long n) { return dim3(dec::nblocks(n, kBlock)); }

template <class T>
const std::vector<LevelDesc<T>> &host_desc(const dec_mg_s &mg);
template <>
const std::vector<LevelDesc<double>> &host_desc<double>(const dec_mg_s &mg) { return mg.hdesc64; }
template <>
const std::vector<LevelDesc<float>> &host_desc<float>(const dec_mg_s &mg) { return mg.hdesc32; }

template <class T>
LevelVals<T> &vals(Level &L);
template <>
LevelVals<double> &vals<double>(Level &L) { return L.d; }
template <>
LevelVals<float> &vals<float>(Level &L) { return L.f; }

// Symmetric adjacency (CSR, no diagonal) of the sparsity pattern of A.
void adjacency(const dec::HostCsr &A, std::vector<int32_t> &deg, std::vector<int32_t> &adj) {
  int n = A.m;
  deg.assign(n + 1, 0);
  for (int r = 0; r < n; r++)
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
      if (A.col[k] != r) {
        deg[r + 1]++;
        deg[A.col[k] + 1]++;
      }
  for (int r = 0; r < n; r++) deg[r + 1] += deg[r];
  adj.assign(deg[n], 0);
  std::vector<int32_t> w(deg.begin(), deg.end() - 1);
  for (int r = 0; r < n; r++)
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
      if (A.col[k] != r) {
        adj[w[r]++] = A.col[k];
        adj[w[A.col[k]]++] = r;
      }
}

// Blocks of about B vertices grown breadth-first over the symmetric graph
// (deg, adj) of a level's operator: a block grows from its seed until it
// holds B vertices, and the next block is seeded next to it (from the
// unassigned neighbours of the last block's BFS frontier, else the lowest
// unassigned vertex), so consecutive blocks are neighbours.  perm lists the
// vertices block by block in BFS order; bstart holds the block ranges.
void grow_blocks(int n, const std::vector<int32_t> &deg, const std::vector<int32_t> &adj, int B,
                 std::vector<int32_t> &perm, std::vector<int32_t> &bstart) {
  std::vector<int32_t> blk(n, -1);
  perm.clear();
  perm.reserve(n);
  bstart.assign(1, 0);
  std::vector<int32_t> seeds, queue;
  size_t sh = 0;
  int next = 0, nb = 0;
  while (static_cast<int>(perm.size()) < n) {
    int seed = -1;
    while (sh < seeds.size() && seed < 0) {
      int c = seeds[sh++];
      if (blk[c] < 0) seed = c;
    }
    if (seed < 0) {
      while (next < n && blk[next] >= 0) next++;
      seed = next;
    }
    queue.clear();
    queue.push_back(seed);
    blk[seed] = nb;
    size_t qh = 0;
    int cnt = 1;
    while (qh < queue.size() && cnt < B) {
      int v = queue[qh++];
      for (int k = deg[v]; k < deg[v + 1] && cnt < B; k++) {
        int w = adj[k];
        if (blk[w] < 0) {
          blk[w] = nb;
          cnt++;
          queue.push_back(w);
        }
      }
    }
    seeds.clear();
    sh = 0;
    for (size_t q = qh; q < queue.size(); q++)
      for (int k = deg[queue[q]]; k < deg[queue[q] + 1]; k++)
        if (blk[adj[k]] < 0) seeds.push_back(adj[k]);
    perm.insert(perm.end(), queue.begin(), queue.end());
    bstart.push_back(static_cast<int32_t>(perm.size()));
    nb++;
  }
}

// Block size of the hybrid Gauss-Seidel: the largest power of two from 64 to
// 512 rows that still gives at least 640 blocks (2.5 per CU).  The
// icosphere levels hold 10 * 4^k + 2 rows, just above a power of two: at
// 163,842 rows the rule picks 256-row blocks (640 blocks), where a threshold
// of "b * 640 < n" picked 512 (320 blocks), and the 3->8 solve takes 2.01
// against 2.14 ms.
int block_rows(int n) {
  int b = 64;
  while (b < 512 && n / (2 * b) >= 640) b *= 2;
  return b;
}

// Colour-sorted ELL of a host CSR (caller's numbering) over the level's
// internal numbers.  `color` is the proposed Gauss-Seidel colouring (empty:
// none); a proposal that is not a proper colouring of A's pattern is replaced
// by greedy colours.  On return `color` holds the colouring used.  A level of
// more than `block_min` rows is renumbered into BFS-grown blocks; slots keep
// the caller's ascending column order, so row sums are those of A.
void build_level_matrix(Level &L, const dec::HostCsr &A, std::vector<int32_t> &color, int block_min) {
  int n = A.m;
  L.n = n;
  std::vector<int32_t> deg, adj;
  adjacency(A, deg, adj);
  int nc = 0;
  for (int c : color) nc = std::max(nc, c + 1);
  if (color.empty() || !dec::proper_coloring(n, deg, adj, color, nc)) nc = dec::greedy_colors(n, deg, adj, color);
  L.ncolors = nc;
  L.color = color;
  std::vector<int32_t> bstart;
  if (n > block_min) {
    grow_blocks(n, deg, adj, block_rows(n), L.perm, bstart);
  } else {
    L.perm.resize(n);
    for (int v = 0; v < n; v++) L.perm[v] = v;
    bstart = {0, n};
  }
  L.iperm.assign(n, 0);
  L.identity = true;
  for (int i = 0; i < n; i++) {
    L.iperm[L.perm[i]] = i;
    L.identity = L.identity && L.perm[i] == i;
  }
  L.dperm.upload(L.perm);
  // positions: rows sorted by (colour, internal number)
  std::vector<int32_t> order(n);
  L.cstart.assign(L.ncolors + 1, 0);
  for (int v = 0; v < n; v++) L.cstart[color[v] + 1]++;
  for (int c = 0; c < L.ncolors; c++) L.cstart[c + 1] += L.cstart[c];
  std::vector<int32_t> pos(L.cstart.begin(), L.cstart.end() - 1);
  for (int i = 0; i < n; i++) order[pos[color[L.perm[i]]]++] = i;
  // (colour, block) segments of the positions
  int nb = static_cast<int>(bstart.size()) - 1;
  L.nblk = nb;
  L.bmax = 0;
  for (int k = 0; k < nb; k++) L.bmax = std::max(L.bmax, bstart[k + 1] - bstart[k]);
  std::vector<int32_t> seg(static_cast<size_t>(L.ncolors) * (nb + 1));
  {
    std::vector<int32_t> blk_of(n);
    for (int k = 0; k < nb; k++)
      for (int i = bstart[k]; i < bstart[k + 1]; i++) blk_of[i] = k;
    std::vector<int32_t> cnt(static_cast<size_t>(L.ncolors) * nb, 0);
    for (int i = 0; i < n; i++) cnt[static_cast<size_t>(color[L.perm[i]]) * nb + blk_of[i]]++;
    for (int c = 0; c < L.ncolors; c++) {
      int p = L.cstart[c];
      for (int k = 0; k < nb; k++) {
        seg[static_cast<size_t>(c) * (nb + 1) + k] = p;
        p += cnt[static_cast<size_t>(c) * nb + k];
      }
      seg[static_cast<size_t>(c) * (nb + 1) + nb] = p;
    }
    // threads per block of the sweep kernel: one per row of the largest
    // (colour, block) segment, in whole wavefronts, at most 256
    int smax = 0;
    for (int c : cnt) smax = std::max(smax, c);
    L.seg_threads = std::min(256, std::max(64, (smax + 63) / 64 * 64));
    L.seg_max = smax;
  }
  L.bstart.upload(bstart);
  L.seg.upload(seg);
  L.hbstart = bstart;
  int width = 0;
  for (int r = 0; r < n; r++) width = std::max(width, A.ptr[r + 1] - A.ptr[r]);
  L.width = width;
  size_t sz = static_cast<size_t>(n) * std::max(width, 1);
  std::vector<int32_t> col(sz, -1);
  std::vector<double> vd(sz, 0.0);
  std::vector<float> vf(sz, 0.0f);
  for (int p = 0; p < n; p++) {
    int r = L.perm[order[p]];
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
      size_t q = static_cast<size_t>(k - A.ptr[r]) * n + p;
      col[q] = L.iperm[A.col[k]];
      vd[q] = A.val[k];
      vf[q] = static_cast<float>(A.val[k]);
    }
  }
  L.col.upload(col);
  L.rowid.upload(order);
  {
    std::vector<int32_t> where(n);
    for (int p = 0; p < n; p++) where[order[p]] = p;
    // rows in vertex order, for the kernels that sweep no colours: their
    // vector accesses are then contiguous, where colour-sorted rows touch
    // every cache line of a vector once per colour
    std::vector<int32_t> vc(sz, -1);
    std::vector<double> vvd(sz, 0.0);
    std::vector<float> vvf(sz, 0.0f);
    int w = std::max(width, 1);
    for (int i = 0; i < n; i++)
      for (int k = 0; k < w; k++) {
        size_t a = static_cast<size_t>(k) * n + i, b = static_cast<size_t>(k) * n + where[i];
        vc[a] = col[b];
        vvd[a] = vd[b];
        vvf[a] = vf[b];
      }
    L.vcol.upload(vc);
    L.d.vval.upload(vvd);
    L.f.vval.upload(vvf);
    std::vector<signed char> vcl(n);
    for (int i = 0; i < n; i++) vcl[i] = static_cast<signed char>(color[L.perm[i]]);
    L.vcolor.upload(vcl);
  }
  L.d.val.upload(vd);
  L.f.val.upload(vf);
  L.A = A;
}

// make_restriction(transpose(S)) (Multigrid.jl:595-612) as rows of R, and the
// endpoints of P = transpose(S) = subdivision_matrix(BinarySubdivision) (:191-219).
// R = row-normalised S and P = Sᵀ of one binary refinement of `coarse`
// (Multigrid.jl:191-219, 595-612); pa, pb are the endpoints of each midpoint.
void transfer_csr(const dec::Topo &coarse, dec::HostCsr &R, dec::HostCsr &P, std::vector<int32_t> &pa,
                  std::vector<int32_t> &pb) {
  int nc = coarse.nv, ne = coarse.ne, nf = nc + ne;
  std::vector<std::vector<int32_t>> mids(nc);
  pa.assign(ne, 0);
  pb.assign(ne, 0);
  for (int e = 0; e < ne; e++) {
    int a = std::min(coarse.v0[e], coarse.v1[e]), b = std::max(coarse.v0[e], coarse.v1[e]);
    pa[e] = a;
    pb[e] = b;
    mids[a].push_back(nc + e);
    if (b != a) mids[b].push_back(nc + e);
  }
  R = dec::HostCsr();
  R.m = nc;
  R.n = nf;
  R.ptr.assign(nc + 1, 0);
  for (int i = 0; i < nc; i++) {
    double rs = 1.0;  // row sum in storage order: 1.0, then 0.5 per midpoint
    for (size_t k = 0; k < mids[i].size(); k++) rs += 0.5;
    R.col.push_back(i);
    R.val.push_back(1.0 / rs);
    for (int c : mids[i]) {
      R.col.push_back(c);
      R.val.push_back(0.5 / rs);
    }
    R.ptr[i + 1] = static_cast<int32_t>(R.col.size());
  }
  P = dec::HostCsr();
  P.m = nf;
  P.n = nc;
  P.ptr.assign(nf + 1, 0);
  for (int j = 0; j < nf; j++) {
    if (j < nc) {
      P.col.push_back(j);
      P.val.push_back(1.0);
    } else {
      P.col.push_back(pa[j - nc]);
      P.val.push_back(0.5);
      P.col.push_back(pb[j - nc]);
      P.val.push_back(0.5);
    }
    P.ptr[j + 1] = static_cast<int32_t>(P.col.size());
  }
}

// Slot-major ELL of a host CSR with rows and columns renumbered: row i is
// CSR row rperm[i], columns map through cimap; slots keep the CSR order.
template <class T>
void ell_renumbered(const dec::HostCsr &M, const std::vector<int32_t> &rperm, const std::vector<int32_t> &cimap,
                    int &width, std::vector<int32_t> &col, std::vector<T> &val) {
  int m = M.m;
  width = 0;
  for (int r = 0; r < m; r++) width = std::max(width, M.ptr[r + 1] - M.ptr[r]);
  size_t sz = static_cast<size_t>(m) * std::max(width, 1);
  col.assign(sz, -1);
  val.assign(sz, T(0));
  for (int i = 0; i < m; i++) {
    int r = rperm[i];
    for (int k = M.ptr[r]; k < M.ptr[r + 1]; k++) {
      size_t q = static_cast<size_t>(k - M.ptr[r]) * m + i;
      col[q] = cimap[M.col[k]];
      val[q] = static_cast<T>(M.val[k]);
    }
  }
}

// R (fine -> coarse) and P (coarse -> fine) of one level pair as ELL over
// the two levels' internal numbers; R and P are host CSRs in the caller's
// numbering, kept for export.
void set_transfer(Level &fine, const Level &coarse, dec::HostCsr R, dec::HostCsr P) {
  if (R.m != coarse.n || R.n != fine.n || P.m != fine.n || P.n != coarse.n)
    fail(DEC_ERR_ARG, "multigrid: transfer operator shapes do not match the levels");
  fine.nc = coarse.n;
  std::vector<int32_t> col;
  std::vector<double> vd;
  std::vector<float> vf;
  ell_renumbered(R, coarse.perm, fine.iperm, fine.rwidth, col, vd);
  ell_renumbered(R, coarse.perm, fine.iperm, fine.rwidth, col, vf);
  fine.rcol.upload(col);
  fine.d.rval.upload(vd);
  fine.f.rval.upload(vf);
  ell_renumbered(P, fine.perm, coarse.iperm, fine.pwidth, col, vd);
  ell_renumbered(P, fine.perm, coarse.iperm, fine.pwidth, col, vf);
  fine.pcol.upload(col);
  fine.d.pval.upload(vd);
  fine.f.pval.upload(vf);
  fine.R = std::move(R);
  fine.P = std::move(P);
}

void build_transfer(Level &fine, const Level &coarse, const dec::Topo &coarse_topo) {
  dec::HostCsr R, P;
  std::vector<int32_t> pa, pb;
  transfer_csr(coarse_topo, R, P, pa, pb);
  set_transfer(fine, coarse, std::move(R), std::move(P));
}

void alloc_work(Level &L) {
  for (auto *v : {&L.d.u, &L.d.b, &L.d.r, &L.d.t}) v->alloc(L.n);
  for (auto *v : {&L.f.u, &L.f.b, &L.f.r, &L.f.t}) v->alloc(L.n);
}

dec_mesh_s *copy_mesh(const dec_mesh_s &m) {
  auto *c = new dec_mesh_s();
  c->h = m.h;
  c->hp = m.hp;
  c->to_dfs = m.to_dfs;
  c->v0.upload(c->h.v0);
  c->v1.upload(c->h.v1);
  c->e0.upload(c->h.e0);
  c->e1.upload(c->h.e1);
  c->e2.upload(c->h.e2);
  c->p.upload(c->hp);
  return c;
}

// Gauss-Seidel colours of every level (finest first), four per level where
// possible.  The base level is 4-coloured by search.  A proper colouring c of
// a level with values in {0,1,2,3} colours the next finer level properly:
// coarse vertices take 0 and the midpoint of coarse edge (a, b) takes
// c(a) ^ c(b), which is nonzero because a and b are adjacent; two midpoints
// are adjacent only when their edges (a, b), (b, z) bound a common coarse
// triangle, and c(a) ^ c(b) = c(b) ^ c(z) would force c(a) = c(z) for the
// adjacent a, z; coarse vertices are never adjacent on the finer level.
// A level without a proposal (or whose proposal fails the check in
// build_level_matrix) is coloured greedily.
std::vector<std::vector<int32_t>> level_colors(const dec_mg_s &mg, const std::vector<dec::HostCsr> &As) {
  int nl = static_cast<int>(As.size());
  std::vector<std::vector<int32_t>> colors(nl);
  const dec::HostCsr &Ab = As[nl - 1];
  if (Ab.m <= 65536) {
    std::vector<int32_t> deg, adj;
    adjacency(Ab, deg, adj);
    if (!dec::four_colors(Ab.m, deg, adj, colors[nl - 1], 64L * Ab.m + 4096)) colors[nl - 1].clear();
  }
  for (int l = nl - 2; l >= 0; l--) {
    const std::vector<int32_t> &cc = colors[l + 1];
    if (cc.empty()) break;
    const dec::Topo &coarse = mg.meshes[l + 1]->h;
    int nc = coarse.nv;
    std::vector<int32_t> &cf = colors[l];
    cf.assign(static_cast<size_t>(nc) + coarse.ne, 0);
    for (int e = 0; e < coarse.ne; e++) cf[nc + e] = cc[coarse.v0[e]] ^ cc[coarse.v1[e]];
  }
  return colors;
}

// Levels, transfers, the dense coarse operator and the fused-tail
// descriptors of a hierarchy, from its operators (caller's numbering, finest
// first), R_l / P_l between levels l and l+1, and proposed colourings.
void finish_hierarchy(dec_mg_s &mg, const std::vector<dec::HostCsr> &As, std::vector<dec::HostCsr> &Rs,
                      std::vector<dec::HostCsr> &Ps, std::vector<std::vector<int32_t>> &colors) {
  int nl = static_cast<int>(As.size());
  for (int l = 0; l < nl; l++) {
    auto L = std::make_unique<Level>();
    build_level_matrix(*L, As[l], colors[l], mg.fuse_rows);
    alloc_work(*L);
    mg.lv.push_back(std::move(L));
  }
  for (int l = 0; l + 1 < nl; l++) set_transfer(*mg.lv[l], *mg.lv[l + 1], std::move(Rs[l]), std::move(Ps[l]));
  // dense coarse operator for the direct solve (over the level's internal numbers)
  Level &C = *mg.lv.back();
  int n = C.n;
  if (static_cast<long long>(n) * n <= 64LL * 1024 * 1024) {
    std::vector<double> Ad(static_cast<size_t>(n) * n, 0.0);
    for (int r = 0; r < n; r++)
      for (int k = C.A.ptr[r]; k < C.A.ptr[r + 1]; k++) Ad[static_cast<size_t>(r) * n + C.A.col[k]] += C.A.val[k];
    dec::coarse_inverse(n, Ad, C.Zh);
    std::vector<double> zt(C.Zh.size());
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        zt[static_cast<size_t>(j) * n + i] = C.Zh[static_cast<size_t>(C.perm[i]) * n + C.perm[j]];
    C.d.Z.upload(zt);
    std::vector<float> zf(zt.begin(), zt.end());
    C.f.Z.upload(zf);
    C.d.zpart.alloc(static_cast<size_t>(coarse_chunks(n)) * n);
    C.f.zpart.alloc(static_cast<size_t>(coarse_chunks(n)) * n);
  }
  // descriptors of the coarse tail for the fused kernel
  for (int l = 0; l < nl; l++)
    if (mg.lv[l]->n <= mg.fuse_rows && mg.lv[l]->ncolors <= kMaxColors) {
      bool ok = true;
      for (int q = l; q < nl; q++) ok = ok && mg.lv[q]->n <= mg.fuse_rows && mg.lv[q]->ncolors <= kMaxColors;
      if (ok) {
        mg.fuse_from = l;
        break;
      }
    }
  {
    std::vector<LevelDesc<double>> d64(nl);
    std::vector<LevelDesc<float>> d32(nl);
    for (int l = 0; l < nl; l++) {
      Level &L = *mg.lv[l];
      auto fill = [&](auto &D, auto &V) {
        D.n = L.n;
        D.width = L.width;
        D.ncolors = L.ncolors;
        D.nc = L.nc;
        D.rwidth = L.rwidth;
        D.pwidth = L.pwidth;
        for (int c = 0; c <= kMaxColors; c++) D.cstart[c] = c < static_cast<int>(L.cstart.size()) ? L.cstart[c] : L.n;
        D.col = L.col.p;
        D.rowid = L.rowid.p;
        D.rcol = L.rcol.p;
        D.pcol = L.pcol.p;
        D.val = V.val.p;
        D.rval = V.rval.p;
        D.pval = V.pval.p;
        D.u = V.u.p;
        D.b = V.b.p;
        D.r = V.r.p;
      };
      fill(d64[l], L.d);
      fill(d32[l], L.f);
    }
    mg.desc64.upload(d64);
    mg.desc32.upload(d32);
    mg.hdesc64 = d64;
    mg.hdesc32 = d32;
  }
  mg.part.alloc(kCgBlocksMax);  // also the dot partials (kDotBlocks of them)
  mg.part2.alloc(kCgBlocksMax);
  mg.scal.alloc(8);
  mg.rpart.alloc(static_cast<size_t>(dec::nblocks(mg.lv[0]->n, kBlock)));
  if (hipHostMalloc(reinterpret_cast<void **>(&mg.hnorm), sizeof(double) * (kNormSlots + 1), hipHostMallocDefault) !=
      hipSuccess)
    fail(DEC_ERR_OOM, "multigrid: pinned host buffer");
  for (auto &e : mg.nev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) fail(DEC_ERR_HIP, "hipEventCreate failed");
  mg.cg.alloc(1);
  int nf = mg.lv[0]->n;
  mg.cgx.alloc(nf);
  mg.cgp.alloc(nf);
  mg.cgp2.alloc(nf);
  mg.cgAp.alloc(nf);
}

void env_switches(dec_mg_s &mg) {
  // debugging switches: run every phase as its own launch / launch cycles eagerly
  mg.use_fused = std::getenv("DEC_MG_NO_FUSE") == nullptr;
  if (const char *e = std::getenv("DEC_MG_FUSE_ROWS")) mg.fuse_rows = std::min(std::max(std::atoi(e), 0), kFuseRowsMax);
  mg.use_graph = std::getenv("DEC_MG_NO_GRAPH") == nullptr;
}

// A caller's m x n CSR (index_base 0 or 1) checked and copied 0-based.
dec::HostCsr csr_in(int m, int n, const int32_t *ptr, const int32_t *col, const double *val, int base,
                    const char *what) {
  if (!ptr || (!col && ptr[m] > ptr[0]) || (!val && ptr[m] > ptr[0]))
    fail(DEC_ERR_ARG, std::string("dec_mg_create_from_csr: NULL array for ") + what);
  dec::HostCsr A;
  A.m = m;
  A.n = n;
  A.ptr.resize(m + 1);
  for (int r = 0; r <= m; r++) {
    A.ptr[r] = ptr[r] - ptr[0];
    if (r > 0 && A.ptr[r] < A.ptr[r - 1])
      fail(DEC_ERR_ARG, std::string("dec_mg_create_from_csr: row pointers of ") + what + " decrease");
  }
  int nnz = A.ptr[m];
  A.col.resize(nnz);
  A.val.assign(val ? val : static_cast<const double *>(nullptr), val ? val + nnz : nullptr);
  for (int k = 0; k < nnz; k++) {
    int c = col[k] - base;
    if (c < 0 || c >= n) fail(DEC_ERR_ARG, std::string("dec_mg_create_from_csr: column out of range in ") + what);
    A.col[k] = c;
  }
  for (int r = 0; r < m; r++)
    for (int k = A.ptr[r] + 1; k < A.ptr[r + 1]; k++)
      if (A.col[k] == A.col[k - 1])
        fail(DEC_ERR_ARG, std::string("dec_mg_create_from_csr: repeated column in a row of ") + what);
  return A;
}

// MultigridData(ops, restrictions, prolongations, steps) (Multigrid.jl:545-546):
// a hierarchy from the caller's matrices, finest first.
dec_mg_s *mg_from_csr(int nl, const int32_t *sizes, const int32_t *const *ap, const int32_t *const *ac,
                      const double *const *av, const int32_t *const *rp, const int32_t *const *rc,
                      const double *const *rv, const int32_t *const *pp, const int32_t *const *pc,
                      const double *const *pv, int base, int steps) {
  if (nl < 1 || nl > 16) fail(DEC_ERR_ARG, "dec_mg_create_from_csr: 1 to 16 levels");
  if (!sizes || !ap || !ac || !av || (nl > 1 && (!rp || !rc || !rv || !pp || !pc || !pv)))
    fail(DEC_ERR_ARG, "dec_mg_create_from_csr: NULL argument");
  if (base != 0 && base != 1) fail(DEC_ERR_ARG, "dec_mg_create_from_csr: index_base must be 0 or 1");
  if (steps < 0) fail(DEC_ERR_ARG, "dec_mg_create_from_csr: steps must be >= 0");
  for (int l = 0; l < nl; l++)
    if (sizes[l] < 1) fail(DEC_ERR_ARG, "dec_mg_create_from_csr: empty level");
  auto mg = std::make_unique<dec_mg_s>();
  mg->steps = steps;
  env_switches(*mg);
  std::vector<dec::HostCsr> As(nl), Rs(nl - 1), Ps(nl - 1);
  for (int l = 0; l < nl; l++) {
    As[l] = csr_in(sizes[l], sizes[l], ap[l], ac[l], av[l], base, "an operator");
    bool diag = true;
    for (int r = 0; r < sizes[l] && diag; r++) {
      bool has = false;
      for (int k = As[l].ptr[r]; k < As[l].ptr[r + 1]; k++) has = has || (As[l].col[k] == r && As[l].val[k] != 0.0);
      diag = has;
    }
    if (!diag) fail(DEC_ERR_ARG, "dec_mg_create_from_csr: an operator has a zero diagonal entry");
  }
  for (int l = 0; l + 1 < nl; l++) {
    Rs[l] = csr_in(sizes[l + 1], sizes[l], rp[l], rc[l], rv[l], base, "a restriction");
    Ps[l] = csr_in(sizes[l], sizes[l + 1], pp[l], pc[l], pv[l], base, "a prolongation");
  }
  std::vector<std::vector<int32_t>> colors(nl);  // greedy colours per level
  finish_hierarchy(*mg, As, Rs, Ps, colors);
  return mg.release();
}

dec_mg_s *mg_create(const dec_mesh_s &base, int levels, bool project, int hodge, int center, int steps,
                    bool galerkin) {
  if (base.dim != 2) fail(DEC_ERR_UNSUPPORTED, "dec_mg_create: the multigrid hierarchy refines 2-D complexes");
  if (levels < 0 || levels > 11) fail(DEC_ERR_ARG, "dec_mg_create: levels must be in 0..11");
  if (steps < 0) fail(DEC_ERR_ARG, "dec_mg_create: steps must be >= 0");
  auto mg = std::make_unique<dec_mg_s>();
  mg->steps = steps;
  env_switches(*mg);
  // coarse -> fine walk (Multigrid.jl:752-797), stored finest first
  std::vector<std::unique_ptr<dec_mesh_s>> walk;
  walk.emplace_back(copy_mesh(base));
  for (int i = 0; i < levels; i++) walk.emplace_back(dec::subdivide(*walk.back(), project));
  for (int i = levels; i >= 0; i--) mg->meshes.push_back(std::move(walk[i]));
  int nl = levels + 1;
  std::vector<dec::HostCsr> As(nl);
  for (int l = 0; l < nl; l++) {
    dec_mesh_s &m = *mg->meshes[l];
    m.owned_by_mg = true;
    if (galerkin && l > 0) {
      // GalerkinMode: only the finest mesh is dualized; Aₖ₊₁ = (Rₖ Aₖ) Pₖ (Multigrid.jl:740-748)
      dec::HostCsr R, P, RA;
      std::vector<int32_t> pa, pb;
      transfer_csr(m.h, R, P, pa, pb);
      dec::spgemm(R, As[l - 1], RA);
      dec::spgemm(RA, P, As[l]);
      continue;
    }
    dec::dualize(m, center);
    std::vector<double> h0 = m.h0.download(), h1 = m.h1.download(), ih0(m.h.nv);
    for (int v = 0; v < m.h.nv; v++) ih0[v] = 1 / h0[v];
    if (hodge == DEC_GEOMETRIC_HODGE) {
      dec::HostCsr H;
      dec::geo_hodge1(m, H);
      dec::assemble_laplacian0(m.h, ih0, h1, &H, As[l]);
    } else {
      dec::assemble_laplacian0(m.h, ih0, h1, nullptr, As[l]);
    }
  }
  std::vector<std::vector<int32_t>> colors = level_colors(*mg, As);
  std::vector<dec::HostCsr> Rs(nl > 0 ? nl - 1 : 0), Ps(nl > 0 ? nl - 1 : 0);
  for (int l = 0; l + 1 < nl; l++) {
    std::vector<int32_t> pa, pb;
    transfer_csr(mg->meshes[l + 1]->h, Rs[l], Ps[l], pa, pb);
  }
  finish_hierarchy(*mg, As, Rs, Ps, colors);
  dec::check_launch("dec_mg_create");
  dec::dsync();
  return mg.release();
}



// ---------------------------------------------------------------- cycles

template <class T>
void residual(Level &L, const T *b, const T *u, T *r, hipStream_t s) {
  LevelVals<T> &V = vals<T>(L);
  DEC_ELL_WIDTH_SWITCH(L.width, hipLaunchKernelGGL((k_residual<W, T>), dim3(dec::xcd_grid(L.n)), dim3(kBlock), 0, s,
                                                   L.n, L.width, L.vcol.p, V.vval.p, static_cast<const int *>(nullptr),
                                                   b, u, r));
}

void dot(dec_mg_s &mg, int n, const double *x, const double *y, double *out, const int *done, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partial<double>, dim3(kDotBlocks), dim3(kBlock), 0, s, n, x, y, mg.part.p, done);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, mg.part.p, out, done);
}

// cg(A, b, u; itmax = steps) as in Multigrid.jl:860 (float64 only), on a
// level above kCgSmallRows rows: 1 + 2 steps launches (the residual with the
// start, then a direction and a step launch per iteration).
void cg_smooth(dec_mg_s &mg, Level &L, const double *b, double *u, int steps, hipStream_t s) {
  if (steps <= 0) return;
  int n = L.n, nb = cg_blocks(n);
  CgState *st = mg.cg.p;
  double *x = mg.cgx.p, *Ap = mg.cgAp.p, *r = L.d.t.p;
  double *pb[2] = {mg.cgp.p, mg.cgp2.p};
  DEC_ELL_WIDTH_SWITCH(L.width, hipLaunchKernelGGL((k_cg2_start<W>), dim3(nb), dim3(kBlock), 0, s, n, L.width,
                                                   L.vcol.p, L.d.vval.p, b, u, r, pb[0], x, mg.part.p));
  for (int it = 0; it < steps; it++) {
    // iteration it's p is pb[it & 1]; the direction launch forms it from
    // pb[(it - 1) & 1] (at it = 0, pb[0] already holds p = r and is only read)
    double *pold = it == 0 ? pb[0] : pb[(it + 1) & 1], *pnew = it == 0 ? pb[1] : pb[it & 1];
    DEC_ELL_WIDTH_SWITCH(L.width, hipLaunchKernelGGL((k_cg2_dir<W>), dim3(nb), dim3(kBlock), 0, s, st, it, n,
                                                     L.width, L.vcol.p, L.d.vval.p, r, pold, pnew, Ap, mg.part.p,
                                                     nb, mg.part2.p));
    if (it + 1 == steps)
      hipLaunchKernelGGL(k_cg2_step<true>, dim3(nb), dim3(kBlock), 0, s, st, it, n, x, pb[it & 1], r, Ap,
                         mg.part2.p, nb, mg.part.p, u);
    else
      hipLaunchKernelGGL(k_cg2_step<false>, dim3(nb), dim3(kBlock), 0, s, st, it, n, x, pb[it & 1], r, Ap,
                         mg.part2.p, nb, mg.part.p, u);
  }
}

inline bool learned_on(const dec_mg_s &mg) { return mg.nn && mg.learned && mg.learned_active; }
inline bool learned_rows(const dec_mg_s &mg, int n) {
  return n >= mg.learned_min_rows && (mg.learned_max_rows <= 0 || n <= mg.learned_max_rows);
}

// u_zero: u is known to be zero on entry (the residual is then b itself).
// The prefetching sweep serves levels up to this many rows, where a sweep is
// bound by memory latency; on larger levels its register footprint lowers
// occupancy and the streaming sweep is faster (measured at 655,362 rows:
// 25.6 against 18.2 µs per sweep).
constexpr int kPrefetchMaxRows = 200000;
// Residual and restriction run as one launch on levels up to this many rows
// (measured: 6.8 against 9.0 µs at 40,962 rows, 18.5 against 10.6 µs at
// 163,842, where evaluating shared fine rows twice costs more than a launch).
constexpr int kFusedRestrictMaxRows = 50000;

// Rows per thread of the prefetching sweep for this level's blocks (1 or 2
// with up to 256 threads), or 0 when it does not apply (wide rows, or a
// single block larger than 512 rows).
int prefetch_rows(const Level &L) {
  if (L.width < 1 || L.width > 8 || L.ncolors > kPfColors) return 0;
  if (L.bmax <= kBlock) return 1;
  if (L.bmax <= 2 * kBlock) return 2;
  return 0;
}

template <class T>
void launch_block_gs_pf(Level &L, LevelVals<T> &V, int R, int nthr, size_t lds, const T *b, const T *src, T *dst,
                        bool post, const T *zc, hipStream_t s) {
  const T *pv = zc ? V.pval.p : nullptr;
  const int *pc = zc ? L.pcol.p : nullptr;
#define DEC_BGS_PF(WW, RR)                                                                                            \
  hipLaunchKernelGGL((k_block_gs_pf<WW, T, RR>), dim3(L.nblk), dim3(nthr), lds, s, L.n, L.ncolors, L.vcol.p,      \
                     V.vval.p, L.vcolor.p, L.bstart.p, b, src, dst, post ? 1 : 0, pc, pv, zc)
#define DEC_BGS_PF_R(WW) \
  if (R == 1) DEC_BGS_PF(WW, 1); \
  else DEC_BGS_PF(WW, 2);
  switch (L.width) {
    case 1: DEC_BGS_PF_R(1) break;
    case 2: DEC_BGS_PF_R(2) break;
    case 3: DEC_BGS_PF_R(3) break;
    case 4: DEC_BGS_PF_R(4) break;
    case 5: DEC_BGS_PF_R(5) break;
    case 6: DEC_BGS_PF_R(6) break;
    case 7: DEC_BGS_PF_R(7) break;
    case 8: DEC_BGS_PF_R(8) break;
    default: fail(DEC_ERR_ARG, "multigrid: prefetching sweep width");
  }
#undef DEC_BGS_PF_R
#undef DEC_BGS_PF
}

template <class T>
void smooth(dec_mg_s &mg, int lvl, Level &L, const T *b, T *u, bool post, bool u_zero, hipStream_t s,
            const T *zc = nullptr) {
  int steps = mg.steps;
  if (steps == 0) return;
  LevelVals<T> &V = vals<T>(L);
  if constexpr (sizeof(T) == 4) {
    if (learned_on(mg) && learned_rows(mg, L.n)) {
      for (int a = 0; a < mg.learned_apps; a++) {
        if (a == 0 && u_zero) {
          dec::nn_smooth(*mg.nn, lvl, b, u, s);
          continue;
        }
        residual<float>(L, b, u, V.r.p, s);
        dec::nn_smooth(*mg.nn, lvl, V.r.p, u, s);
      }
      return;
    }
  }
  if (mg.alg == DEC_SMOOTH_CG) {
    if constexpr (sizeof(T) == 8) {
      if (L.n <= kCgSmallRows && L.width >= 1 && L.width <= 7) {
        size_t lds = static_cast<size_t>(L.n) * sizeof(double);  // p
        DEC_ELL_WIDTH_SWITCH(L.width, if constexpr (W > 0) {
          hipLaunchKernelGGL((k_cg_smooth_small<W, kCgRowsPerThread>), dim3(1), dim3(kCgThreads), lds, s, L.n,
                             L.vcol.p, L.d.vval.p, b, u, steps);
        });
      } else {
        cg_smooth(mg, L, b, u, steps, s);
      }
    } else {
      fail(DEC_ERR_UNSUPPORTED, "the CG smoother runs in float64 only");
    }
    return;
  }
  if (mg.alg == DEC_SMOOTH_BLOCK_GS) {
    // one launch per sweep, ping-ponging between u and the level's t
    // (an odd step count ends with a copy back into u); zc (post-smoothing
    // only) folds the prolongation u += P zc into the first sweep
    T *bufs[2] = {u, V.t.p};
    size_t lds = static_cast<size_t>(L.bmax) * sizeof(T);
    int R = prefetch_rows(L);
    for (int it = 0; it < steps; it++) {
      const T *src = bufs[it & 1];
      T *dst = bufs[(it + 1) & 1];
      const T *z = it == 0 ? zc : nullptr;
      if (R > 0 && L.n <= kPrefetchMaxRows) {
        int nthr = R == 1 ? (L.bmax + 63) / 64 * 64 : kBlock;
        launch_block_gs_pf<T>(L, V, R, nthr, lds, b, src, dst, post, z, s);
      } else {
        const T *pv = z ? V.pval.p : nullptr;
        const int *pc = z ? L.pcol.p : nullptr;
        if (z) {
          DEC_ELL_WIDTH_SWITCH(L.width, hipLaunchKernelGGL((k_block_gs<W, T, true>), dim3(L.nblk),
                                                           dim3(L.seg_threads), lds, s, L.n, L.width, L.ncolors,
                                                           L.nblk, L.col.p, V.val.p, L.rowid.p, L.bstart.p, L.seg.p,
                                                           b, src, dst, post ? 1 : 0, pc, pv, z));
        } else {
          // one row per thread and colour: the float32 sweep with the next
          // colour's loads in flight (measured 1.93 against 1.99 ms for the
          // mixed 3->8 solve; the float64 sweep measured 1% slower with it,
          // at 85 against 54 registers)
          bool pipelined = sizeof(T) == 4 && L.seg_max <= L.seg_threads;
          DEC_ELL_WIDTH_SWITCH(L.width, {
            if constexpr (W > 0) {
              if (pipelined) {
                hipLaunchKernelGGL((k_block_gs_pl<W, T>), dim3(L.nblk), dim3(L.seg_threads), lds, s, L.n,
                                   L.ncolors, L.nblk, L.col.p, V.val.p, L.rowid.p, L.bstart.p, L.seg.p, b, src, dst,
                                   post ? 1 : 0);
                break;
              }
            }
            hipLaunchKernelGGL((k_block_gs<W, T, false>), dim3(L.nblk), dim3(L.seg_threads), lds, s, L.n, L.width,
                               L.ncolors, L.nblk, L.col.p, V.val.p, L.rowid.p, L.bstart.p, L.seg.p, b, src, dst,
                               post ? 1 : 0, pc, pv, z);
          });
        }
      }
    }
    if (steps & 1) hipLaunchKernelGGL(k_copy<T>, G(L.n), dim3(kBlock), 0, s, L.n, V.t.p, u);
    return;
  }
  for (int it = 0; it < steps; it++) {
    if (mg.alg == DEC_SMOOTH_GS) {
      for (int ci = 0; ci < L.ncolors; ci++) {
        int c = post ? L.ncolors - 1 - ci : ci;
        int p0 = L.cstart[c], p1 = L.cstart[c + 1];
        if (p1 > p0)
          DEC_ELL_WIDTH_SWITCH(L.width, hipLaunchKernelGGL((k_gs_color<W, T>), dim3(dec::xcd_grid(p1 - p0)), dim3(kBlock),
                                                           0, s, p0, p1,
                                                           L.n, L.width, L.col.p, V.val.p, L.rowid.p, b, u));
      }
    } else {
      DEC_ELL_WIDTH_SWITCH(L.width, hipLaunchKernelGGL((k_jacobi<W, T>), G(L.n), dim3(kBlock), 0, s, L.n, L.width,
                                                     L.col.p, V.val.p, L.rowid.p, b, u, V.t.p,
                                                     static_cast<T>(mg.omega)));
      hipLaunchKernelGGL(k_copy<T>, G(L.n), dim3(kBlock), 0, s, L.n, V.t.p, u);
    }
  }
}

// u += Z r (r = the level's residual vector) by the chunked two-pass product.
template <class T>
void coarse_partials(Level &L, hipStream_t s) {
  LevelVals<T> &V = vals<T>(L);
  dim3 grid(static_cast<unsigned>((L.n + 63) / 64), static_cast<unsigned>(coarse_chunks(L.n)));
  hipLaunchKernelGGL(k_coarse_partial<T>, grid, dim3(64), 0, s, L.n, V.Z.p, V.r.p, V.zpart.p);
}

template <class T>
void coarse_gemv(Level &L, T *u, hipStream_t s) {
  coarse_partials<T>(L, s);
  hipLaunchKernelGGL(k_coarse_finish<T>, G(L.n), dim3(kBlock), 0, s, L.n, coarse_chunks(L.n), vals<T>(L).zpart.p, u);
}

template <class T>
void coarse_solve(dec_mg_s &mg, Level &L, const T *b, T *u, hipStream_t s) {
  LevelVals<T> &V = vals<T>(L);
  if (!V.Z.p) fail(DEC_ERR_UNSUPPORTED, "coarsest level too large for the dense direct solve");
  residual<T>(L, b, u, V.r.p, s);
  coarse_gemv<T>(L, u, s);
  (void)mg;
}

// _multigrid_μ_cycle (Multigrid.jl:857-872).
template <class T>
void mu_cycle(dec_mg_s &mg, int lvl, T *u, const T *b, int mu, bool u_zero, hipStream_t s) {
  int nl = static_cast<int>(mg.lv.size());
  Level &L = *mg.lv[lvl];
  // the fused tail has no learned smoother: a level that uses it runs unfused
  bool learned_here = sizeof(T) == 4 && learned_on(mg) && learned_rows(mg, L.n);
  if (mg.use_fused && mu == 1 && mg.alg != DEC_SMOOTH_CG && mg.fuse_from >= 0 && lvl >= mg.fuse_from &&
      !learned_here && (mg.coarse_direct == 0 || vals<T>(*mg.lv.back()).Z.p)) {
    const LevelDesc<T> *D;
    if constexpr (sizeof(T) == 8) D = mg.desc64.p;
    else D = mg.desc32.p;
    T om = static_cast<T>(mg.omega);
    if (!mg.coarse_direct) {
      hipLaunchKernelGGL(k_fused_vcycle<T>, dim3(1), dim3(kFuseThreads), 0, s, D, lvl, nl, u, b, mg.steps, mg.alg, om,
                         0, static_cast<const T *>(nullptr));
      return;
    }
    Level &C = *mg.lv.back();
    // one Gauss-Seidel level above the dense coarsest level: the register/LDS tail
    if (lvl == nl - 2 && mg.alg != DEC_SMOOTH_JACOBI && L.n <= 3 * kFuseThreads && L.width >= 1 && L.width <= 8 &&
        static_cast<size_t>(L.n + std::max(L.n, C.n)) * sizeof(T) <= 64 * 1024) {
      size_t lds = static_cast<size_t>(L.n + std::max(L.n, C.n)) * sizeof(T);
      const T *zp = vals<T>(C).zpart.p;
      const std::vector<LevelDesc<T>> &HD = host_desc<T>(mg);
      auto tail = [&](int phase) {
        switch (L.width) {
#define DEC_TAIL2(WW)                                                                                                 \
  case WW:                                                                                                           \
    hipLaunchKernelGGL((k_tail2<WW, T, 3>), dim3(1), dim3(kFuseThreads), lds, s, HD[lvl], HD[lvl + 1], u, b,       \
                       mg.steps, phase, zp);                                                                          \
    break;
          DEC_TAIL2(1) DEC_TAIL2(2) DEC_TAIL2(3) DEC_TAIL2(4) DEC_TAIL2(5) DEC_TAIL2(6) DEC_TAIL2(7) DEC_TAIL2(8)
#undef DEC_TAIL2
          default: break;
        }
      };
      tail(1);
      coarse_partials<T>(C, s);
      tail(2);
      return;
    }
    hipLaunchKernelGGL(k_fused_vcycle<T>, dim3(1), dim3(kFuseThreads), 0, s, D, lvl, nl, u, b, mg.steps, mg.alg, om,
                       1, static_cast<const T *>(nullptr));
    coarse_partials<T>(C, s);
    hipLaunchKernelGGL(k_fused_vcycle<T>, dim3(1), dim3(kFuseThreads), 0, s, D, lvl, nl, u, b, mg.steps, mg.alg, om,
                       2, static_cast<const T *>(vals<T>(C).zpart.p));
    return;
  }
  if (lvl == nl - 1 && mg.coarse_direct) {
    coarse_solve<T>(mg, L, b, u, s);
    return;
  }
  smooth<T>(mg, lvl, L, b, u, false, u_zero, s);
  if (lvl >= nl - 1) return;
  Level &C = *mg.lv[lvl + 1];
  LevelVals<T> &V = vals<T>(L);
  LevelVals<T> &CV = vals<T>(C);
  if (L.n <= kFusedRestrictMaxRows && L.rwidth <= 8) {
    DEC_ELL_WIDTH_SWITCH(L.width, hipLaunchKernelGGL((k_residual_restrict<W, T>), G(8LL * L.nc), dim3(kBlock), 0, s,
                                                     L.nc, L.rwidth, L.rcol.p, V.rval.p, L.n, L.width, L.vcol.p,
                                                     V.vval.p, b, u, CV.b.p, CV.u.p));
  } else {
    residual<T>(L, b, u, V.r.p, s);
    DEC_ELL_WIDTH_SWITCH(L.rwidth, hipLaunchKernelGGL((k_restrict<W, T>), G(L.nc), dim3(kBlock), 0, s, L.nc,
                                                      L.rwidth, L.rcol.p, V.rval.p, V.r.p, CV.b.p, CV.u.p));
  }
  mu_cycle<T>(mg, lvl + 1, CV.u.p, CV.b.p, mu, true, s);
  if (mu > 1) mu_cycle<T>(mg, lvl + 1, CV.u.p, CV.b.p, mu - 1, false, s);
  // the hybrid smoother's first post-sweep applies the prolongation itself
  // (on the prefetching sweep only: there the extra loads join the load
  // phase; on the streaming sweep of the largest level they cost more than
  // the separate launch, measured 37 against 18 + 6.5 µs at 655,362 rows)
  bool fold = mg.alg == DEC_SMOOTH_BLOCK_GS && mg.steps > 0 && L.pwidth == 2 && prefetch_rows(L) > 0 &&
              L.n <= kPrefetchMaxRows && !(sizeof(T) == 4 && learned_on(mg) && learned_rows(mg, L.n));
  if (!fold)
    DEC_ELL_WIDTH_SWITCH(L.pwidth, hipLaunchKernelGGL((k_prolong_add<W, T>), G(L.n), dim3(kBlock), 0, s, L.n,
                                                      L.pwidth, L.pcol.p, V.pval.p, CV.u.p, u));
  smooth<T>(mg, lvl, L, b, u, true, false, s, fold ? static_cast<const T *>(CV.u.p) : nullptr);
}

// _full_multigrid (Multigrid.jl:839-855).
void fmg_rec(dec_mg_s &mg, int lvl, const double *b, double *z, int mu, int cycles, hipStream_t s) {
  int nl = static_cast<int>(mg.lv.size());
  Level &L = *mg.lv[lvl];
  hipLaunchKernelGGL(k_zero<double>, G(L.n), dim3(kBlock), 0, s, L.n, z);
  if (lvl < nl - 1) {
    Level &C = *mg.lv[lvl + 1];
    DEC_ELL_WIDTH_SWITCH(L.rwidth, hipLaunchKernelGGL((k_restrict<W, double>), G(L.nc), dim3(kBlock), 0, s, L.nc,
                                                      L.rwidth, L.rcol.p, L.d.rval.p, b, C.d.b.p,
                                                      static_cast<double *>(nullptr)));
    fmg_rec(mg, lvl + 1, C.d.b.p, C.d.u.p, mu, cycles, s);
    DEC_ELL_WIDTH_SWITCH(L.pwidth, hipLaunchKernelGGL((k_prolong_add<W, double>), G(L.n), dim3(kBlock), 0, s, L.n,
                                                      L.pwidth, L.pcol.p, L.d.pval.p, C.d.u.p, z));
  }
  if (lvl > 0)
    for (int c = 0; c < cycles; c++) mu_cycle<double>(mg, lvl, z, b, mu, false, s);
}

// The caller's finest-level vectors in the hierarchy's numbering: b into
// the finest level's b, u (when load_u) into its u; the caller's own pointers
// when the finest level keeps the caller's order.
struct FineIO {
  double *u;
  const double *b;
};

FineIO fine_in(dec_mg_s &mg, double *u, const double *b, bool load_u, hipStream_t s) {
  Level &F = *mg.lv[0];
  if (F.identity) return FineIO{u, b};
  hipLaunchKernelGGL(k_perm_gather<double>, G(F.n), dim3(kBlock), 0, s, F.n, F.dperm.p, b, F.d.b.p);
  if (load_u) hipLaunchKernelGGL(k_perm_gather<double>, G(F.n), dim3(kBlock), 0, s, F.n, F.dperm.p, u, F.d.u.p);
  return FineIO{F.d.u.p, F.d.b.p};
}

void fine_out(dec_mg_s &mg, double *u, hipStream_t s) {
  Level &F = *mg.lv[0];
  if (F.identity) return;
  hipLaunchKernelGGL(k_perm_scatter<double>, G(F.n), dim3(kBlock), 0, s, F.n, F.dperm.p, F.d.u.p, u);
}

double norm2(dec_mg_s &mg, int n, const double *x, hipStream_t s) {
  dot(mg, n, x, x, mg.scal.p, nullptr, s);
  double v = 0.0;
  if (hipMemcpyAsync(&v, mg.scal.p, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    fail(DEC_ERR_HIP, "norm: copy failed");
  return std::sqrt(v);
}

void cycle_body(dec_mg_s &mg, int variant, double *u, const double *b, hipStream_t s) {
  Level &F = *mg.lv[0];
  int n = F.n;
  if (variant == DEC_SOLVE_PLAIN_F64) {
    mu_cycle<double>(mg, 0, u, b, 1, false, s);
  } else {
    // float32 V-cycle on the float64 residual, correction added in float64
    hipLaunchKernelGGL(k_d2f, G(n), dim3(kBlock), 0, s, n, F.d.r.p, F.f.b.p);
    hipLaunchKernelGGL(k_zero<float>, G(n), dim3(kBlock), 0, s, n, F.f.u.p);
    mu_cycle<float>(mg, 0, F.f.u.p, F.f.b.p, 1, true, s);
    hipLaunchKernelGGL(k_add_f2d, G(n), dim3(kBlock), 0, s, n, F.f.u.p, u);
  }
}

void gcheck(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(DEC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ||b||^2 on its way to the host (pinned slot kNormSlots with its event), and
// its square root once it has arrived
void send_bnorm(dec_mg_s &mg, int n, const double *b, hipStream_t s) {
  dot(mg, n, b, b, mg.scal.p + kNormSlots, nullptr, s);
  gcheck(hipMemcpyAsync(mg.hnorm + kNormSlots, mg.scal.p + kNormSlots, sizeof(double), hipMemcpyDeviceToHost, s),
         "hipMemcpyAsync");
  gcheck(hipEventRecord(mg.nev[kNormSlots], s), "hipEventRecord");
}
double read_bnorm(dec_mg_s &mg) {
  gcheck(hipEventSynchronize(mg.nev[kNormSlots]), "hipEventSynchronize");
  return std::sqrt(mg.hnorm[kNormSlots]);
}

// body(stream) replayed from the hipGraph in slot i, captured for (u, b) and
// the current smoother settings (key); captured again when any of them changes.
template <class Body>
void replay(dec_mg_s &mg, int i, const void *u, const void *b, Body body, hipStream_t s) {
  int key = mg.alg * 1000003 + mg.steps * 1009 + mg.coarse_direct * 7 + static_cast<int>(mg.omega * 1e6) +
            (mg.learned && mg.learned_active ? 1 << 29 : 0) + mg.learned_apps * 31 + mg.learned_min_rows +
            mg.learned_max_rows * 7919;
  if (!mg.gexec[i] || mg.gu[i] != u || mg.gb[i] != b || mg.gkey[i] != key) {
    if (mg.gexec[i]) (void)hipGraphExecDestroy(mg.gexec[i]);
    if (mg.graph[i]) (void)hipGraphDestroy(mg.graph[i]);
    mg.gexec[i] = nullptr;
    mg.graph[i] = nullptr;
    if (!mg.cap) gcheck(hipStreamCreateWithFlags(&mg.cap, hipStreamNonBlocking), "hipStreamCreate");
    gcheck(hipStreamBeginCapture(mg.cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      body(mg.cap);
    } catch (...) {
      // close the capture, or every later allocation on this thread fails
      hipGraph_t partial = nullptr;
      (void)hipStreamEndCapture(mg.cap, &partial);
      if (partial) (void)hipGraphDestroy(partial);
      throw;
    }
    gcheck(hipStreamEndCapture(mg.cap, &mg.graph[i]), "hipStreamEndCapture");
    gcheck(hipGraphInstantiate(&mg.gexec[i], mg.graph[i], nullptr, nullptr, 0), "hipGraphInstantiate");
    mg.gu[i] = u;
    mg.gb[i] = b;
    mg.gkey[i] = key;
  }
  gcheck(hipGraphLaunch(mg.gexec[i], s), "hipGraphLaunch");
}

// One cycle of the solve, replayed from a hipGraph captured for (u, b) and
// the current smoother settings.
void run_cycle(dec_mg_s &mg, int variant, double *u, const double *b, hipStream_t s) {
  if (!mg.use_graph) {
    cycle_body(mg, variant, u, b, s);
    return;
  }
  replay(mg, variant == DEC_SOLVE_PLAIN_F64 ? 0 : 1, u, b,
         [&](hipStream_t cs) { cycle_body(mg, variant, u, b, cs); }, s);
}

void solve(dec_mg_s &mg, int variant, double tol, int max_cycles, double *u_caller, const double *b_caller,
           int *done_cycles, double *relres, hipStream_t s) {
  Level &F = *mg.lv[0];
  int n = F.n;
  // With the iterate in the hierarchy's own numbering, the caller's u stays
  // untouched until the end, so the first cycles are queued before the host
  // has read ||b|| and the starting residual (one wait instead of two, with
  // the GPU busy during it); a b of zero or a start that already meets the
  // tolerance then returns without them.
  const bool ahead = !F.identity;
  double bn = 0.0;
  if (ahead) {
    send_bnorm(mg, n, b_caller, s);
  } else {
    bn = norm2(mg, n, b_caller, s);
    if (bn == 0.0) {
      hipLaunchKernelGGL(k_zero<double>, G(n), dim3(kBlock), 0, s, n, u_caller);
      *done_cycles = 0;
      *relres = 0.0;
      return;
    }
  }
  FineIO io = fine_in(mg, u_caller, b_caller, true, s);
  double *u = io.u;
  const double *b = io.b;
  // the network smooths the float32 cycles on the levels of at least
  // learned_min_rows .. learned_max_rows rows; with no such level the solve
  // is a plain mixed one
  bool any = false;
  for (auto &Lv : mg.lv) any = any || learned_rows(mg, Lv->n);
  mg.learned_active = mg.learned && mg.nn && variant == DEC_SOLVE_MIXED_IR && any;
  mg.fallbacks = 0;
  mg.learned_cycles = 0;
  // A residual norm is copied to pinned host memory with an event per slot,
  // and the host reads it when it must decide whether to stop.
  bool plain = variant == DEC_SOLVE_PLAIN_F64;
  // after cycle k: the float64 residual (which the mixed variant's next cycle
  // refines) and, when `check`, its norm on its way to the host
  auto post = [&](int k, bool check) {
    if (plain && !check) return;
    int nb = dec::nblocks(n, kBlock);
    if (mg.rpart.n < static_cast<size_t>(nb)) fail(DEC_ERR_ARG, "multigrid: residual partials");
    DEC_ELL_WIDTH_SWITCH(F.width, hipLaunchKernelGGL((k_residual_norm<W>), dim3(nb), dim3(kBlock), 0, s, n, F.width,
                                                     F.vcol.p, F.d.vval.p, static_cast<const int *>(nullptr), b, u,
                                                     F.d.r.p, mg.rpart.p));
    if (!check) return;
    double *slot = mg.scal.p + (k % kNormSlots);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(kBlock), 0, s, nb, mg.rpart.p, slot);
    gcheck(hipMemcpyAsync(mg.hnorm + (k % kNormSlots), slot, sizeof(double), hipMemcpyDeviceToHost, s),
           "hipMemcpyAsync");
    gcheck(hipEventRecord(mg.nev[k % kNormSlots], s), "hipEventRecord");
  };
  auto read_norm = [&](int k) {
    gcheck(hipEventSynchronize(mg.nev[k % kNormSlots]), "hipEventSynchronize");
    return std::sqrt(mg.hnorm[k % kNormSlots]) / bn;
  };
  int enq = 0;               // cycles queued
  bool enq_checked = true;   // whether the norm after cycle `enq` was queued
  auto enqueue_cycle = [&](bool check) {
    enq++;
    if (mg.learned_active) mg.learned_cycles++;
    run_cycle(mg, variant, u, b, s);
    post(enq, check);
    enq_checked = check;
  };
  // The norm is read only where the decision needs it: after cycle 1 (the
  // convergence rate), then at the first cycle that could meet the
  // tolerance, then after every cycle (every cycle while the learned
  // smoother is on: its fallback test needs each drop).  Cycles before a
  // check run without waiting for the host; while the last known residual
  // stays above 1000 tol after the queued cycles, the cycle after the check
  // is queued too (no V-cycle cuts the residual 1000x).
  post(0, true);
  if (ahead) {
    // what the loop below queues first for a start far from the solution
    if (max_cycles >= 1) enqueue_cycle(true);
    if (max_cycles >= 2) enqueue_cycle(false);
    bn = read_bnorm(mg);
    if (bn == 0.0) {
      hipLaunchKernelGGL(k_zero<double>, G(n), dim3(kBlock), 0, s, n, u_caller);
      mg.learned_cycles = 0;
      *done_cycles = 0;
      *relres = 0.0;
      return;
    }
  }
  double rr = read_norm(0), prev = rr, rho = -1.0;
  int c = 0;
  if (ahead && rr <= tol) {
    mg.learned_cycles = 0;
    *done_cycles = 0;
    *relres = rr;
    return;
  }
  while (rr > tol && c < max_cycles) {
    int kc = c + 1;
    // (the prediction takes each cycle to cut twice as much as the last ones
    // did, so the check comes no later than the first cycle that can meet
    // the tolerance)
    if (rho > 0.0 && !mg.learned_active)
      while (kc < max_cycles && rr * std::pow(0.5 * rho, kc - c) > tol) kc++;
    while (enq < kc) enqueue_cycle(enq + 1 == kc);
    if (enq == kc && !enq_checked) {
      // the check cycle was queued ahead without its norm; nothing after it
      // is queued, so the iterate is still the one after cycle kc
      post(kc, true);
      enq_checked = true;
    }
    if (enq < kc + 1 && kc + 1 <= max_cycles && rr * std::pow(1e-3, kc + 1 - c) > tol) enqueue_cycle(false);
    double rk = read_norm(kc);
    rho = std::min(0.95, std::max(1e-4, std::pow(rk / rr, 1.0 / (kc - c))));
    // a learned cycle that cut the residual by less than 0.5x hands the rest
    // of the solve to Gauss-Seidel (from the next cycle not yet queued)
    if (mg.learned_active && rk > 0.5 * prev) {
      mg.learned_active = false;
      mg.fallbacks++;
    }
    prev = rr = rk;
    c = kc;
  }
  if (enq > c) {  // a queued cycle ran past the stop (a drop above 1000x): report it
    if (!enq_checked) post(enq, true);
    c = enq;
    rr = read_norm(c);
  }
  fine_out(mg, u_caller, s);
  *done_cycles = c;
  *relres = rr;
  if (rr > tol) fail(DEC_ERR_NOCONV, "dec_mg_solve: relative residual " + std::to_string(rr) + " above tolerance after " +
                                         std::to_string(c) + " cycles");
}

// One iteration of the preconditioned CG: z = one V-cycle on r from a zero
// start (float32 cycles on the rounded residual for the mixed variant, z
// their float64 value), then the three CG launches.  The V-cycle's input (a
// zero z, or the float32 r and iterate) was prepared by the previous
// iteration's update launch or by the start.
void pcg_iteration(dec_mg_s &mg, bool mixed, double *x, hipStream_t s) {
  Level &F = *mg.lv[0];
  int n = F.n, nb = dec::nblocks(n, kBlock), ng = dec::xcd_grid(n);
  // partial sums: r·z in [0, kPcgDotBlocks), then p·q (one per workgroup of
  // the direction launch, ng) and r·r (one per workgroup of the update, nb)
  double *r = mg.pcg_r.p, *z = mg.pcg_z.p, *part = mg.pcg_part.p;
  double *part_pq = part + kPcgDotBlocks, *part_rr = part_pq + ng;
  if (!mixed) mu_cycle<double>(mg, 0, z, r, 1, true, s);
  else mu_cycle<float>(mg, 0, F.f.u.p, F.f.b.p, 1, true, s);
  hipLaunchKernelGGL(k_pcg_dot, dim3(kPcgDotBlocks), dim3(kBlock), 0, s, n, r, mixed ? F.f.u.p : nullptr, z, part);
  DEC_ELL_WIDTH_SWITCH(F.width, hipLaunchKernelGGL((k_pcg_dir<W>), dim3(ng), dim3(kBlock), 0, s, n,
                                                   F.width, F.vcol.p, F.d.vval.p, static_cast<const int *>(nullptr),
                                                   z, mg.pcg_p.p,
                                                   mg.pcg_q.p, part, kPcgDotBlocks, mg.pcg_sc.p, part_pq));
  hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(kBlock), 0, s, part_pq, ng, mg.pcg_sc.p);
  hipLaunchKernelGGL(k_pcg_update, dim3(nb), dim3(kBlock), 0, s, n, x, r, mg.pcg_p.p, mg.pcg_q.p, mg.pcg_sc.p,
                     part_rr, mixed ? nullptr : z, mixed ? F.f.b.p : nullptr, mixed ? F.f.u.p : nullptr);
}

// Solve with the preconditioned CG (variants DEC_SOLVE_PCG_F64 / _MIXED).
// The stop test reads the recursively updated residual's norm where the
// decision needs it, as solve() does; once it meets the tolerance the true
// residual b - A x is formed and must meet it too (otherwise CG restarts
// from the true residual).  An iteration rate of 0.9 or worse hands the rest
// of the solve to plain V-cycles from the current iterate.  The learned
// smoother is not used in these variants.
void pcg_solve(dec_mg_s &mg, bool mixed, double tol, int max_iters, double *u_caller, const double *b_caller,
               int *done_iters, double *relres, hipStream_t s) {
  Level &F = *mg.lv[0];
  int n = F.n;
  // the first iterations are queued before the host reads ||b|| and the
  // starting residual, as in solve()
  const bool ahead = !F.identity;
  double bn = 0.0;
  if (ahead) {
    send_bnorm(mg, n, b_caller, s);
  } else {
    bn = norm2(mg, n, b_caller, s);
    if (bn == 0.0) {
      hipLaunchKernelGGL(k_zero<double>, G(n), dim3(kBlock), 0, s, n, u_caller);
      *done_iters = 0;
      *relres = 0.0;
      return;
    }
  }
  int nb = dec::nblocks(n, kBlock);
  if (mg.pcg_r.n < static_cast<size_t>(n)) {
    for (auto *v : {&mg.pcg_r, &mg.pcg_z, &mg.pcg_p, &mg.pcg_q}) v->alloc(n);
    mg.pcg_part.alloc(kPcgDotBlocks + static_cast<size_t>(dec::xcd_grid(n)) + nb);
    mg.pcg_sc.alloc(3);
  }
  FineIO io = fine_in(mg, u_caller, b_caller, true, s);
  double *x = io.u;
  const double *b = io.b;
  mg.learned_active = false;
  if (mg.rpart.n < static_cast<size_t>(nb)) fail(DEC_ERR_ARG, "multigrid: residual partials");
  // a norm on its way to the host: the partial sums part[0..np) added into
  // ring slot k with an event
  auto send = [&](int k, const double *part, int np) {
    double *slot = mg.scal.p + (k % kNormSlots);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(kBlock), 0, s, np, part, slot);
    gcheck(hipMemcpyAsync(mg.hnorm + (k % kNormSlots), slot, sizeof(double), hipMemcpyDeviceToHost, s),
           "hipMemcpyAsync");
    gcheck(hipEventRecord(mg.nev[k % kNormSlots], s), "hipEventRecord");
  };
  auto read_norm = [&](int k) {
    gcheck(hipEventSynchronize(mg.nev[k % kNormSlots]), "hipEventSynchronize");
    return std::sqrt(mg.hnorm[k % kNormSlots]) / bn;
  };
  // the true residual b - A x into F.d.r, its norm to slot k
  auto true_residual = [&](int k) {
    DEC_ELL_WIDTH_SWITCH(F.width, hipLaunchKernelGGL((k_residual_norm<W>), dim3(nb), dim3(kBlock), 0, s, n, F.width,
                                                     F.vcol.p, F.d.vval.p, static_cast<const int *>(nullptr), b, x,
                                                     F.d.r.p, mg.rpart.p));
    send(k, mg.rpart.p, nb);
  };
  // CG from the true residual in F.d.r
  auto restart = [&]() {
    hipLaunchKernelGGL(k_pcg_start, G(n), dim3(kBlock), 0, s, n, F.d.r.p, mg.pcg_r.p, mg.pcg_p.p, mg.pcg_q.p,
                       mg.pcg_sc.p, mixed ? nullptr : mg.pcg_z.p, mixed ? F.f.b.p : nullptr,
                       mixed ? F.f.u.p : nullptr);
  };
  int enq = 0;
  bool enq_checked = true;
  auto enqueue = [&](bool check) {
    enq++;
    if (mg.use_graph)
      replay(mg, mixed ? 3 : 2, x, b, [&](hipStream_t cs) { pcg_iteration(mg, mixed, x, cs); }, s);
    else
      pcg_iteration(mg, mixed, x, s);
    if (check) send(enq, mg.pcg_part.p + kPcgDotBlocks + dec::xcd_grid(n), nb);
    enq_checked = check;
  };
  true_residual(0);
  restart();
  if (ahead) {
    if (max_iters >= 1) enqueue(true);
    if (max_iters >= 2) enqueue(false);
    bn = read_bnorm(mg);
    if (bn == 0.0) {
      hipLaunchKernelGGL(k_zero<double>, G(n), dim3(kBlock), 0, s, n, u_caller);
      *done_iters = 0;
      *relres = 0.0;
      return;
    }
  }
  double rr = read_norm(0), rho = -1.0;
  int c = 0;
  bool stalled = false, verified = rr <= tol;
  if (ahead && verified) {
    *done_iters = 0;
    *relres = rr;
    return;
  }
  while (c < max_iters && !verified) {
    if (rr <= tol) {
      // the recursive residual meets the tolerance: confirm it on the
      // iterate after every queued iteration
      c = enq;
      true_residual(c);
      rr = read_norm(c);
      verified = rr <= tol;
      if (!verified) {
        restart();
        rho = -1.0;
      }
      continue;
    }
    // the check comes no later than the first iteration that can meet the
    // tolerance (each predicted to cut twice as much as the last ones did)
    int kc = c + 1;
    if (rho > 0.0)
      while (kc < max_iters && rr * std::pow(0.5 * rho, kc - c) > tol) kc++;
    while (enq < kc) enqueue(enq + 1 == kc);
    if (enq == kc && !enq_checked) {
      // queued ahead without its norm; nothing after it is queued, so its
      // partial sums are still in place
      send(kc, mg.pcg_part.p + kPcgDotBlocks + dec::xcd_grid(n), nb);
      enq_checked = true;
    }
    if (enq < kc + 1 && kc + 1 <= max_iters && rr * std::pow(1e-3, kc + 1 - c) > tol) enqueue(false);
    double rk = read_norm(kc);
    double rate = std::pow(rk / rr, 1.0 / (kc - c));
    rr = rk;
    c = kc;
    if (!(rate < 0.9)) {
      stalled = true;
      break;
    }
    rho = std::min(0.95, std::max(1e-4, rate));
  }
  if (!verified) {
    // the true residual after every queued iteration; when CG stalled,
    // plain V-cycles from there
    c = enq;
    true_residual(c);
    rr = read_norm(c);
    int variant = mixed ? DEC_SOLVE_MIXED_IR : DEC_SOLVE_PLAIN_F64;
    while (stalled && rr > tol && c < max_iters) {
      run_cycle(mg, variant, x, b, s);
      c++;
      true_residual(c);
      rr = read_norm(c);
    }
  }
  fine_out(mg, u_caller, s);
  *done_iters = c;
  *relres = rr;
  if (rr > tol) fail(DEC_ERR_NOCONV, "dec_mg_solve: relative residual " + std::to_string(rr) + " above tolerance after " +
                                         std::to_string(c) + " iterations");
}

}  // namespace

extern "C" {

int dec_mg_create(dec_mesh base, int32_t levels, int32_t project, int32_t hodge, int32_t center, int32_t steps,
                  dec_mg *out) {
  DEC_TRY
  if (!base || !out) fail(DEC_ERR_ARG, "dec_mg_create: NULL handle");
  *out = nullptr;
  *out = mg_create(*base, levels, project != 0, hodge, center, steps, false);
  DEC_CATCH
}

int dec_mg_create_mode(dec_mesh base, int32_t levels, int32_t project, int32_t hodge, int32_t center,
                       int32_t steps, int32_t mode, dec_mg *out) {
  DEC_TRY
  if (!base || !out) fail(DEC_ERR_ARG, "dec_mg_create_mode: NULL handle");
  if (mode != DEC_MG_DIRECT && mode != DEC_MG_GALERKIN) fail(DEC_ERR_ARG, "dec_mg_create_mode: unknown mode");
  *out = nullptr;
  *out = mg_create(*base, levels, project != 0, hodge, center, steps, mode == DEC_MG_GALERKIN);
  DEC_CATCH
}

int dec_mg_create_from_csr(int32_t nlevels, const int32_t *sizes, const int32_t *const *a_rowptr,
                           const int32_t *const *a_col, const double *const *a_val,
                           const int32_t *const *r_rowptr, const int32_t *const *r_col,
                           const double *const *r_val, const int32_t *const *p_rowptr,
                           const int32_t *const *p_col, const double *const *p_val, int32_t index_base,
                           int32_t steps, dec_mg *out) {
  DEC_TRY
  if (!out) fail(DEC_ERR_ARG, "dec_mg_create_from_csr: out is NULL");
  *out = nullptr;
  dec_mg_s *mg = mg_from_csr(nlevels, sizes, a_rowptr, a_col, a_val, r_rowptr, r_col, r_val, p_rowptr, p_col,
                             p_val, index_base, steps);
  dec::check_launch("dec_mg_create_from_csr");
  dec::dsync();
  *out = mg;
  DEC_CATCH
}

int dec_mg_nlevels(dec_mg mg, int32_t *nlevels) {
  DEC_TRY
  if (!mg || !nlevels) fail(DEC_ERR_ARG, "dec_mg_nlevels: NULL argument");
  *nlevels = static_cast<int32_t>(mg->lv.size());
  DEC_CATCH
}

int dec_mg_level_size(dec_mg mg, int32_t level, int32_t *n) {
  DEC_TRY
  if (!mg || !n) fail(DEC_ERR_ARG, "dec_mg_level_size: NULL argument");
  if (level < 0 || level >= static_cast<int>(mg->lv.size())) fail(DEC_ERR_ARG, "dec_mg_level_size: bad level");
  *n = mg->lv[level]->n;
  DEC_CATCH
}

int dec_mg_finest_mesh(dec_mg mg, dec_mesh *m) {
  DEC_TRY
  if (!mg || !m) fail(DEC_ERR_ARG, "dec_mg_finest_mesh: NULL argument");
  if (mg->meshes.empty()) fail(DEC_ERR_UNSUPPORTED, "dec_mg_finest_mesh: the hierarchy was built from matrices");
  *m = mg->meshes[0].get();
  DEC_CATCH
}

int dec_mg_export_level(dec_mg mg, int32_t level, int32_t which, int64_t *nnz, int32_t *rowptr, int32_t *col,
                        double *val) {
  DEC_TRY
  if (!mg) fail(DEC_ERR_ARG, "dec_mg_export_level: NULL handle");
  int nl = static_cast<int>(mg->lv.size());
  if (level < 0 || level >= nl) fail(DEC_ERR_ARG, "dec_mg_export_level: bad level");
  if (which != 0 && level == nl - 1) fail(DEC_ERR_ARG, "dec_mg_export_level: coarsest level has no transfer");
  const dec::HostCsr &A = which == 0 ? mg->lv[level]->A : (which == 1 ? mg->lv[level]->R : mg->lv[level]->P);
  if (which < 0 || which > 2) fail(DEC_ERR_ARG, "dec_mg_export_level: which must be 0 (A), 1 (R) or 2 (P)");
  if (nnz) *nnz = static_cast<int64_t>(A.col.size());
  if (rowptr) std::copy(A.ptr.begin(), A.ptr.end(), rowptr);
  if (col) std::copy(A.col.begin(), A.col.end(), col);
  if (val) std::copy(A.val.begin(), A.val.end(), val);
  DEC_CATCH
}

int dec_mg_export_colors(dec_mg mg, int32_t level, int32_t *color, int32_t *ncolors) {
  DEC_TRY
  if (!mg) fail(DEC_ERR_ARG, "dec_mg_export_colors: NULL handle");
  if (level < 0 || level >= static_cast<int>(mg->lv.size())) fail(DEC_ERR_ARG, "dec_mg_export_colors: bad level");
  const Level &L = *mg->lv[level];
  if (color) std::copy(L.color.begin(), L.color.end(), color);
  if (ncolors) *ncolors = L.ncolors;
  DEC_CATCH
}

int dec_mg_export_blocks(dec_mg mg, int32_t level, int32_t *block, int32_t *nblocks) {
  DEC_TRY
  if (!mg) fail(DEC_ERR_ARG, "dec_mg_export_blocks: NULL handle");
  if (level < 0 || level >= static_cast<int>(mg->lv.size())) fail(DEC_ERR_ARG, "dec_mg_export_blocks: bad level");
  const Level &L = *mg->lv[level];
  if (block)
    for (int k = 0; k < L.nblk; k++)
      for (int i = L.hbstart[k]; i < L.hbstart[k + 1]; i++) block[L.perm[i]] = k;
  if (nblocks) *nblocks = L.nblk;
  DEC_CATCH
}

int dec_mg_set_smoother(dec_mg mg, int32_t alg, int32_t steps, double omega, int32_t coarse_direct) {
  DEC_TRY
  if (!mg) fail(DEC_ERR_ARG, "dec_mg_set_smoother: NULL handle");
  if (alg < DEC_SMOOTH_CG || alg > DEC_SMOOTH_BLOCK_GS) fail(DEC_ERR_ARG, "dec_mg_set_smoother: unknown smoother");
  if (steps < 0) fail(DEC_ERR_ARG, "dec_mg_set_smoother: steps must be >= 0");
  if (coarse_direct && !mg->lv.back()->d.Z.p)
    fail(DEC_ERR_UNSUPPORTED, "dec_mg_set_smoother: coarsest level too large for the dense direct solve");
  mg->alg = alg;
  mg->steps = steps;
  mg->omega = omega;
  mg->coarse_direct = coarse_direct != 0;
  DEC_CATCH
}

int dec_mg_cycles(dec_mg mg, int32_t mu, int32_t cycles, void *u, const void *b, void *stream) {
  DEC_TRY
  if (!mg || !u || !b) fail(DEC_ERR_ARG, "dec_mg_cycles: NULL argument");
  if (mu < 1 || cycles < 0) fail(DEC_ERR_ARG, "dec_mg_cycles: mu >= 1 and cycles >= 0 required");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FineIO io = fine_in(*mg, static_cast<double *>(u), static_cast<const double *>(b), true, s);
  for (int c = 0; c < cycles; c++) mu_cycle<double>(*mg, 0, io.u, io.b, mu, false, s);
  fine_out(*mg, static_cast<double *>(u), s);
  dec::check_launch("dec_mg_cycles");
  DEC_CATCH
}

int dec_mg_full_multigrid(dec_mg mg, int32_t mu, int32_t cycles, const void *b, void *u, void *stream) {
  DEC_TRY
  if (!mg || !u || !b) fail(DEC_ERR_ARG, "dec_mg_full_multigrid: NULL argument");
  if (mu < 1 || cycles < 0) fail(DEC_ERR_ARG, "dec_mg_full_multigrid: mu >= 1 and cycles >= 0 required");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FineIO io = fine_in(*mg, static_cast<double *>(u), static_cast<const double *>(b), false, s);
  fmg_rec(*mg, 0, io.b, io.u, mu, cycles, s);
  for (int c = 0; c < cycles; c++) mu_cycle<double>(*mg, 0, io.u, io.b, mu, false, s);
  fine_out(*mg, static_cast<double *>(u), s);
  dec::check_launch("dec_mg_full_multigrid");
  DEC_CATCH
}

int dec_mg_solve(dec_mg mg, int32_t variant, double tol, int32_t max_cycles, void *u, const void *b,
                 int32_t *cycles_done, double *relres, void *stream) {
  DEC_TRY
  if (!mg || !u || !b) fail(DEC_ERR_ARG, "dec_mg_solve: NULL argument");
  if (variant < DEC_SOLVE_PLAIN_F64 || variant > DEC_SOLVE_PCG_MIXED) fail(DEC_ERR_ARG, "dec_mg_solve: bad variant");
  if ((variant == DEC_SOLVE_MIXED_IR || variant == DEC_SOLVE_PCG_MIXED) && mg->alg == DEC_SMOOTH_CG)
    fail(DEC_ERR_UNSUPPORTED, "dec_mg_solve: the mixed variants need a Gauss-Seidel or Jacobi smoother");
  if ((variant == DEC_SOLVE_PCG_F64 || variant == DEC_SOLVE_PCG_MIXED) && mg->alg == DEC_SMOOTH_CG)
    fail(DEC_ERR_UNSUPPORTED, "dec_mg_solve: the CG variants take the V-cycle of a Gauss-Seidel or Jacobi smoother");
  int c = 0;
  double rr = 0.0;
  try {
    if (variant == DEC_SOLVE_PCG_F64 || variant == DEC_SOLVE_PCG_MIXED)
      pcg_solve(*mg, variant == DEC_SOLVE_PCG_MIXED, tol, max_cycles, static_cast<double *>(u),
                static_cast<const double *>(b), &c, &rr, static_cast<hipStream_t>(stream));
    else
      solve(*mg, variant, tol, max_cycles, static_cast<double *>(u), static_cast<const double *>(b), &c, &rr,
            static_cast<hipStream_t>(stream));
  } catch (...) {
    if (cycles_done) *cycles_done = c;
    if (relres) *relres = rr;
    throw;
  }
  if (cycles_done) *cycles_done = c;
  if (relres) *relres = rr;
  DEC_CATCH
}

int dec_nn_create(dec_mg mg, uint64_t seed, dec_nn *out) {
  DEC_TRY
  if (!mg || !out) fail(DEC_ERR_ARG, "dec_nn_create: NULL argument");
  std::vector<dec::NnOps> ops;
  int nl = static_cast<int>(mg->lv.size());
  for (int l = 0; l < nl; l++) {
    Level &L = *mg->lv[l];
    bool has = l + 1 < nl;
    ops.push_back({&L.A, has ? &L.R : nullptr, has ? &L.P : nullptr, &L.perm});
  }
  *out = dec::nn_create(ops, seed);
  DEC_CATCH
}

int dec_mg_set_learned(dec_mg mg, dec_nn nn, int32_t enable, int32_t min_rows, int32_t max_rows,
                       int32_t applications) {
  DEC_TRY
  if (!mg) fail(DEC_ERR_ARG, "dec_mg_set_learned: NULL handle");
  if (enable && !nn) fail(DEC_ERR_ARG, "dec_mg_set_learned: no network");
  if (applications < 1) fail(DEC_ERR_ARG, "dec_mg_set_learned: applications must be >= 1");
  if (nn) {
    std::vector<int> sizes;
    for (auto &L : mg->lv) sizes.push_back(L->n);
    // the network must carry this hierarchy's operators; its work buffers
    // are sized here, outside any graph capture of the cycles
    dec::nn_bind(*nn, sizes);
  }
  mg->nn = nn;
  mg->learned = enable != 0;
  mg->learned_min_rows = min_rows;
  mg->learned_max_rows = max_rows;
  mg->learned_apps = applications;
  DEC_CATCH
}

int dec_mg_learned_stats(dec_mg mg, int32_t *fallbacks, int32_t *learned_cycles) {
  DEC_TRY
  if (!mg) fail(DEC_ERR_ARG, "dec_mg_learned_stats: NULL handle");
  if (fallbacks) *fallbacks = mg->fallbacks;
  if (learned_cycles) *learned_cycles = mg->learned_cycles;
  DEC_CATCH
}

int dec_mg_destroy(dec_mg mg) {
  DEC_TRY
  if (mg)
    for (auto &m : mg->meshes) m->owned_by_mg = false;
  delete mg;
  DEC_CATCH
}

}  // extern "C"
