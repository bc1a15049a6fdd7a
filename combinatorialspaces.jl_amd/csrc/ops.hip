// FastDEC operators of libdec_hip.so (src/FastDEC.jl): wedge products, exterior
// derivative / boundary / dual derivative, Hodge stars and inverses, the 0-form
// Laplace-Beltrami operator and avg₀₁, as gather kernels over int32 SoA
// incidence tables.  Coefficient tables are computed on the GPU in float64
// and cast once to the operator's dtype (cache_wedge's cast_float).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <utility>

#include "dec_internal.h"
#include "ell_rows.h"
#include "kernels_common.h"

using dec::DBuf;
using dec::fail;

namespace {

enum Impl {
  I_W01,   // wedge 0-1
  I_W02,   // wedge 0-2
  I_W11,   // wedge 1-1
  I_WDD,   // dual-dual wedge 0-1 / 1-0
  I_D0,    // d0
  I_D1,    // d1
  I_SELL,  // ±1 ELL: ∂1 = d0ᵀ, d̃1 = -d0ᵀ, ∂2 = d̃0 = d1ᵀ
  I_DIAG,  // diagonal ⋆, ⋆⁻¹
  I_ELL,   // general sparse: geometric ⋆1, ∇²
  I_AVG,   // avg₀₁
  I_WPD,   // primal-dual wedge 0-1 / dual-primal 1-0: (M g) .* f, M = wedge_pd_01_mat
  I_W03,   // wedge 0-3 (tetrahedra)
  I_W21,   // wedge 2-1 (tetrahedra)
  I_GINV,  // geometric ⋆1⁻¹: a solve with ⋆1 (Jacobi-preconditioned CG on the device)
  I_COMP   // a composition of other operators (see Comp)
};

// Composite operators, each a fixed sequence of sub-operator applications.
enum Comp {
  C_WPD11,  // primal-dual 1-1: ∧11(f, ♭♯ g); dual-primal 1-1: ∧11(♭♯ f, g)
  C_I11,    // interior_product_dd(1,1): ⋆2 ∧11(⋆1⁻¹ g, ♭♯ f)
  C_I12,    // interior_product_dd(1,2): ⋆1 ♭♯ ((Λ01 ⋆0⁻¹ g) .* f)
  C_LIE,    // ℒ_dd(1,1): -(d̃0 i11(f, g)) - i12(f, d̃1 g)
  C_DL0,    // Δᵈ(0): (⋆2 d1) ⋆1⁻¹(d̃0 x)
  C_DL1     // Δᵈ(1): M x + N ⋆1⁻¹(x), M = ⋆1 d0 ⋆0⁻¹ d̃1, N = d̃0 ⋆2 d1
};

}  // namespace

struct dec_op_s {
  int kind = 0, m = 0, n = 0, hodge = 0, dtype = 0;
  int impl = 0;
  bool swap = false;  // apply(x, y) evaluates the kernel on (y, x)
  int64_t rows = 0, cols = 0, cols2 = 0;
  // index tables
  DBuf<int32_t> i0, i1, i2;
  DBuf<int8_t> s8;
  // coefficient tables, stored as raw bytes of the op dtype
  DBuf<unsigned char> c0, c1, c2;
  // tetrahedron tables (slot-major: slot k of row x at k * rows + x): vertex
  // or edge/triangle ids, per-vertex weights, and sign bits of the 2-1 terms
  DBuf<int32_t> ix;
  DBuf<unsigned char> cx;
  DBuf<uint16_t> m16;
  // sparse forms
  dec::Ell<int8_t> sell;
  dec::Ell<double> ell64;
  dec::Ell<float> ell32;
  dec::HostCsr csr;  // host copy for export (sparse-matrix operators)
  // geometric ⋆1⁻¹: the Jacobi scaling 1/diag(⋆1), CG vectors, reduction
  // partials and device-resident scalars (mutable: one apply at a time)
  DBuf<unsigned char> dinv, cr, cp, cq, cz, crh, ct;
  DBuf<double> cpart, cscal;
  int max_iter = 0;
  double rtol = 0.0;
  bool symmetric = true;  // CG when ⋆1 is symmetric, BiCGSTAB otherwise
  // spatial tile schedule of the gather kernels (8 x tiles-per-XCD entries)
  DBuf<int32_t> sched;
  int nsched = 0;
  // composite operators: sub-operators and work vectors (one apply at a time)
  int comp = 0;
  std::vector<std::unique_ptr<dec_op_s>> sub;
  std::vector<DBuf<unsigned char>> tmp;
};

namespace {

// The tile (workgroup-sized row range) a workgroup processes: its entry in
// the operator's spatial schedule when it has one, else the default order.
// A schedule deals tiles to the XCDs (workgroup i runs on XCD i % 8) by
// position on the mesh, so the neighbours one XCD gathers stay in its L2;
// -1 pads the XCDs with fewer tiles.
__device__ __forceinline__ int tile_index(const int *__restrict__ sched, int dflt) {
  return sched ? sched[blockIdx.x] : dflt;
}

// Apply kernels: each thread handles the pair of simplices (2i, 2i+1), loading
// index and coefficient tables as int2 / T2 vectors; a trailing odd element is
// handled by a scalar path.  Caller forms must be 2·sizeof(T)-aligned for the
// vector path (checked at apply time; otherwise the scalar kernels run).
template <class T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};
template <class T>
using V2 = typename Vec2<T>::type;

template <class T>
__device__ __forceinline__ V2<T> ld2(const T *p, int i) {
  return *reinterpret_cast<const V2<T> *>(p + i);
}
// Operator outputs are written with nontemporal stores: a result is not
// read again by the same launch, and streaming it out measured 8% faster on
// the wedge 1-1 kernel at k = 8 (as fast as a plain copy of the same bytes).
template <class T>
__device__ __forceinline__ void stn(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}
template <class T>
__device__ __forceinline__ void st2(T *p, int i, T a, T b) {
  stn(p + i, a);
  stn(p + i + 1, b);
}
__device__ __forceinline__ int2 ldi2(const int *p, int i) { return *reinterpret_cast<const int2 *>(p + i); }

// wedge_kernel_01! (FastDEC.jl:159-163): half * α[i] * (f[p1] + f[p2])
template <class T>
__device__ __forceinline__ T w01(T a, T f0, T f1) {
  return (static_cast<T>(0.5) * a) * (f0 + f1);
}
template <class T>
__device__ __forceinline__ void b_wedge01(int blk, int ne, int vec, const int *__restrict__ v0, const int *__restrict__ v1,
                          const T *__restrict__ f, const T *__restrict__ a, T *__restrict__ res) {
  int e = 2 * (blk * blockDim.x + threadIdx.x);
  if (vec && e + 1 < ne) {
    int2 p = ldi2(v0, e), q = ldi2(v1, e);
    V2<T> av = ld2(a, e);
    T f0 = f[p.x], f1 = f[q.x], f2 = f[p.y], f3 = f[q.y];
    st2(res, e, w01(av.x, f0, f1), w01(av.y, f2, f3));
  } else {
    for (int j = e; j < e + 2 && j < ne; j++) {
      stn(res + j, w01(a[j], f[v0[j]], f[v1[j]]));
    }
  }
}
template <class T>
__global__ void k_wedge01(int ne, int vec, const int *__restrict__ v0, const int *__restrict__ v1,
                          const T *__restrict__ f, const T *__restrict__ a, T *__restrict__ res,
                          const int *__restrict__ sched) {
  int blk = tile_index(sched, dec_ell::xcd_block());
  if (blk >= 0) b_wedge01<T>(blk, ne, vec, v0, v1, f, a, res);
}

// wedge_kernel_02! (FastDEC.jl:165-170) with the six coefficients summed
// pairwise: the six vertex slots are (v0, v0, v1, v1, v2, v2).
template <class T>
__device__ __forceinline__ T w02(T a, T c0, T c1, T c2, T f0, T f1, T f2) {
  T s = c0 * f0 + c1 * f1;
  s = s + c2 * f2;
  return a * s;
}
template <class T>
__device__ __forceinline__ void b_wedge02(int blk, int nt, int vec, const int *__restrict__ i0, const int *__restrict__ i1,
                          const int *__restrict__ i2, const T *__restrict__ c0, const T *__restrict__ c1,
                          const T *__restrict__ c2, const T *__restrict__ f, const T *__restrict__ a,
                          T *__restrict__ res) {
  int t = 2 * (blk * blockDim.x + threadIdx.x);
  if (vec && t + 1 < nt) {
    int2 p = ldi2(i0, t), q = ldi2(i1, t), u = ldi2(i2, t);
    V2<T> k0 = ld2(c0, t), k1 = ld2(c1, t), k2 = ld2(c2, t), av = ld2(a, t);
    T fa = f[p.x], fb = f[q.x], fc = f[u.x], fd = f[p.y], fe = f[q.y], ff = f[u.y];
    st2(res, t, w02(av.x, k0.x, k1.x, k2.x, fa, fb, fc), w02(av.y, k0.y, k1.y, k2.y, fd, fe, ff));
  } else {
    for (int j = t; j < t + 2 && j < nt; j++) {
      stn(res + j, w02(a[j], c0[j], c1[j], c2[j], f[i0[j]], f[i1[j]], f[i2[j]]));
    }
  }
}
template <class T>
__global__ void k_wedge02(int nt, int vec, const int *__restrict__ i0, const int *__restrict__ i1,
                          const int *__restrict__ i2, const T *__restrict__ c0, const T *__restrict__ c1,
                          const T *__restrict__ c2, const T *__restrict__ f, const T *__restrict__ a,
                          T *__restrict__ res, const int *__restrict__ sched) {
  int blk = tile_index(sched, dec_ell::xcd_block());
  if (blk >= 0) b_wedge02<T>(blk, nt, vec, i0, i1, i2, c0, c1, c2, f, a, res);
}

// wedge_kernel_11! (FastDEC.jl:184-191)
template <class T>
__device__ __forceinline__ T w11(T c0, T c1, T c2, T ae0, T ae1, T ae2, T be0, T be1, T be2) {
  T x = c0 * (ae2 * be1 - ae1 * be2);
  T y = c1 * (ae2 * be0 - ae0 * be2);
  T z = c2 * (ae1 * be0 - ae0 * be1);
  return (x + y) + z;
}
template <class T>
__device__ __forceinline__ void b_wedge11(int blk, int nt, int vec, const int *__restrict__ i0, const int *__restrict__ i1,
                          const int *__restrict__ i2, const T *__restrict__ c0, const T *__restrict__ c1,
                          const T *__restrict__ c2, const T *__restrict__ a, const T *__restrict__ b,
                          T *__restrict__ res) {
  int t = 2 * (blk * blockDim.x + threadIdx.x);
  if (vec && t + 1 < nt) {
    int2 p = ldi2(i0, t), q = ldi2(i1, t), u = ldi2(i2, t);
    V2<T> k0 = ld2(c0, t), k1 = ld2(c1, t), k2 = ld2(c2, t);
    T a0 = a[p.x], a1 = a[q.x], a2 = a[u.x], b0 = b[p.x], b1 = b[q.x], b2 = b[u.x];
    T a3 = a[p.y], a4 = a[q.y], a5 = a[u.y], b3 = b[p.y], b4 = b[q.y], b5 = b[u.y];
    st2(res, t, w11(k0.x, k1.x, k2.x, a0, a1, a2, b0, b1, b2), w11(k0.y, k1.y, k2.y, a3, a4, a5, b3, b4, b5));
  } else {
    for (int j = t; j < t + 2 && j < nt; j++) {
      int e0 = i0[j], e1 = i1[j], e2 = i2[j];
      stn(res + j, w11(c0[j], c1[j], c2[j], a[e0], a[e1], a[e2], b[e0], b[e1], b[e2]));
    }
  }
}
template <class T>
__global__ void k_wedge11(int nt, int vec, const int *__restrict__ i0, const int *__restrict__ i1,
                          const int *__restrict__ i2, const T *__restrict__ c0, const T *__restrict__ c1,
                          const T *__restrict__ c2, const T *__restrict__ a, const T *__restrict__ b,
                          T *__restrict__ res, const int *__restrict__ sched) {
  int blk = tile_index(sched, static_cast<int>(blockIdx.x));
  if (blk >= 0) b_wedge11<T>(blk, nt, vec, i0, i1, i2, c0, c1, c2, a, b, res);
}

// (m * f)[e] .* g[e], m = wedge_dd_01_mat (FastDEC.jl:293-305); the row's
// triangles are ascending (ta < tb, tb < 0 when e has one coface).
template <class T>
__device__ __forceinline__ T wdd(int a, int b, T wa, T wb, const T *__restrict__ f, T g) {
  T s = static_cast<T>(0);
  if (a >= 0) s += wa * f[a];
  if (b >= 0) s += wb * f[b];
  return s * g;
}
template <class T>
__device__ __forceinline__ void b_wedge_dd(int blk, int ne, int vec, const int *__restrict__ ta, const int *__restrict__ tb,
                           const T *__restrict__ wa, const T *__restrict__ wb, const T *__restrict__ f,
                           const T *__restrict__ g, T *__restrict__ res) {
  int e = 2 * (blk * blockDim.x + threadIdx.x);
  if (vec && e + 1 < ne) {
    int2 p = ldi2(ta, e), q = ldi2(tb, e);
    V2<T> x = ld2(wa, e), y = ld2(wb, e), gg = ld2(g, e);
    st2(res, e, wdd(p.x, q.x, x.x, y.x, f, gg.x), wdd(p.y, q.y, x.y, y.y, f, gg.y));
  } else {
    for (int j = e; j < e + 2 && j < ne; j++) {
      stn(res + j, wdd(ta[j], tb[j], wa[j], wb[j], f, g[j]));
    }
  }
}
template <class T>
__global__ void k_wedge_dd(int ne, int vec, const int *__restrict__ ta, const int *__restrict__ tb,
                           const T *__restrict__ wa, const T *__restrict__ wb, const T *__restrict__ f,
                           const T *__restrict__ g, T *__restrict__ res) {
  b_wedge_dd<T>(dec_ell::xcd_block(), ne, vec, ta, tb, wa, wb, f, g, res);
}

// row e of d0 (FastDEC.jl:471-492): σ at ∂v0, -σ at ∂v1
template <class T>
__device__ __forceinline__ T d0v(int8_t s, T x0, T x1) {
  T d = x0 - x1;
  return s > 0 ? d : -d;
}
template <class T>
__device__ __forceinline__ void b_d0(int blk, int ne, int vec, const int *__restrict__ v0, const int *__restrict__ v1,
                     const int8_t *__restrict__ sgn, const T *__restrict__ x, T *__restrict__ res) {
  int e = 2 * (blk * blockDim.x + threadIdx.x);
  if (vec && e + 1 < ne) {
    int2 p = ldi2(v0, e), q = ldi2(v1, e);
    char2 s = *reinterpret_cast<const char2 *>(sgn + e);
    T a = x[p.x], b = x[q.x], c = x[p.y], d = x[q.y];
    st2(res, e, d0v<T>(s.x, a, b), d0v<T>(s.y, c, d));
  } else {
    for (int j = e; j < e + 2 && j < ne; j++) {
      stn(res + j, d0v<T>(sgn[j], x[v0[j]], x[v1[j]]));
    }
  }
}
template <class T>
__global__ void k_d0(int ne, int vec, const int *__restrict__ v0, const int *__restrict__ v1,
                     const int8_t *__restrict__ sgn, const T *__restrict__ x, T *__restrict__ res,
                     const int *__restrict__ sched) {
  int blk = tile_index(sched, dec_ell::xcd_block());
  if (blk >= 0) b_d0<T>(blk, ne, vec, v0, v1, sgn, x, res);
}

// row t of d1 (FastDEC.jl:496-524), its three edges in ascending order with
// signs packed in bits 0..2 (set = +1)
template <class T>
__device__ __forceinline__ T d1v(int s, T xa, T xb, T xc) {
  T r = (s & 1) ? xa : -xa;
  r = r + ((s & 2) ? xb : -xb);
  r = r + ((s & 4) ? xc : -xc);
  return r;
}
template <class T>
__device__ __forceinline__ void b_d1(int blk, int nt, int vec, const int *__restrict__ ia, const int *__restrict__ ib, const int *__restrict__ ic,
                     const int8_t *__restrict__ sg, const T *__restrict__ x, T *__restrict__ res) {
  int t = 2 * (blk * blockDim.x + threadIdx.x);
  if (vec && t + 1 < nt) {
    int2 p = ldi2(ia, t), q = ldi2(ib, t), u = ldi2(ic, t);
    char2 s = *reinterpret_cast<const char2 *>(sg + t);
    T a = x[p.x], b = x[q.x], c = x[u.x], d = x[p.y], e = x[q.y], g = x[u.y];
    st2(res, t, d1v<T>(s.x, a, b, c), d1v<T>(s.y, d, e, g));
  } else {
    for (int j = t; j < t + 2 && j < nt; j++) {
      stn(res + j, d1v<T>(sg[j], x[ia[j]], x[ib[j]], x[ic[j]]));
    }
  }
}
template <class T>
__global__ void k_d1(int nt, int vec, const int *__restrict__ ia, const int *__restrict__ ib, const int *__restrict__ ic,
                     const int8_t *__restrict__ sg, const T *__restrict__ x, T *__restrict__ res,
                     const int *__restrict__ sched) {
  int blk = tile_index(sched, dec_ell::xcd_block());
  if (blk >= 0) b_d1<T>(blk, nt, vec, ia, ib, ic, sg, x, res);
}

// wedge_kernel_03! (FastDEC.jl:173-183): res = Σ_j (α c_j) f[p_j] over the 24
// dual tetrahedra.  The six dual tetrahedra at a vertex share p_j, so their
// coefficients are stored summed: one weight per vertex (Σ dual_vol / vol),
// four terms per tetrahedron in tetrahedron_vertices order.
template <class T>
__device__ __forceinline__ void b_wedge03(int blk, int n, const int *__restrict__ vi, const T *__restrict__ c,
                                          const T *__restrict__ f, const T *__restrict__ a, T *__restrict__ res) {
  int x = blk * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const size_t N = static_cast<size_t>(n);
  int p0 = vi[x], p1 = vi[N + x], p2 = vi[2 * N + x], p3 = vi[3 * N + x];
  T f0 = f[p0], f1 = f[p1], f2 = f[p2], f3 = f[p3];
  T c0 = c[x], c1 = c[N + x], c2 = c[2 * N + x], c3 = c[3 * N + x];
  T av = a[x];
  T s = (av * c0) * f0;
  s = s + (av * c1) * f1;
  s = s + (av * c2) * f2;
  s = s + (av * c3) * f3;
  stn(res + x, s);
}
template <class T>
__global__ void k_wedge03(int n, const int *__restrict__ vi, const T *__restrict__ c, const T *__restrict__ f,
                          const T *__restrict__ a, T *__restrict__ res) {
  b_wedge03<T>(dec_ell::xcd_block(), n, vi, c, f, a, res);
}

// wedge_kernel_21! (FastDEC.jl:194-208): α a primal 2-form on the triangles
// t0..t3, β a primal 1-form on the edges e0..e5 of each tetrahedron.  Each of
// the 12 coefficients (FastDEC.jl:96-133) is ± the weight ws_v = (dual volume
// at vertex v / vol) / 3 of its vertex; bit k of sg is set when coefficient
// k + 1 is negative.  Terms and their order are the reference kernel's.
template <class T>
__device__ __forceinline__ void b_wedge21(int blk, int n, const int *__restrict__ et, const T *__restrict__ ws,
                                          const uint16_t *__restrict__ sg, const T *__restrict__ a,
                                          const T *__restrict__ b, T *__restrict__ res) {
  int x = blk * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const size_t N = static_cast<size_t>(n);
  int e[6], t[4];
#pragma unroll
  for (int k = 0; k < 6; k++) e[k] = et[k * N + x];
#pragma unroll
  for (int k = 0; k < 4; k++) t[k] = et[(6 + k) * N + x];
  T be[6], at[4];
#pragma unroll
  for (int k = 0; k < 6; k++) be[k] = b[e[k]];
#pragma unroll
  for (int k = 0; k < 4; k++) at[k] = a[t[k]];
  T w0 = ws[x], w1 = ws[N + x], w2 = ws[2 * N + x], w3 = ws[3 * N + x];
  unsigned m = sg[x];
  auto C = [m](int k, T w) { return ((m >> k) & 1u) ? -w : w; };
  T r = C(0, w0) * be[3] * at[3];
  r = r - C(1, w0) * be[4] * at[2];
  r = r + C(2, w0) * be[5] * at[1];
  r = r + C(3, w1) * be[1] * at[3];
  r = r - C(4, w1) * be[2] * at[2];
  r = r + C(5, w1) * be[5] * at[0];
  r = r + C(6, w2) * be[0] * at[3];
  r = r - C(7, w2) * be[2] * at[1];
  r = r + C(8, w2) * be[4] * at[0];
  r = r + C(9, w3) * be[0] * at[2];
  r = r - C(10, w3) * be[1] * at[1];
  r = r + C(11, w3) * be[3] * at[0];
  stn(res + x, r);
}
template <class T>
__global__ void k_wedge21(int n, const int *__restrict__ et, const T *__restrict__ ws, const uint16_t *__restrict__ sg,
                          const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ res) {
  b_wedge21<T>(dec_ell::xcd_block(), n, et, ws, sg, a, b, res);
}

// Per-vertex weights of the 0-3 (div = 1) and 2-1 (div = 3) wedges from the
// dual volume at each vertex of each tetrahedron (slot-major 4 x n).
template <class T>
__global__ void k_coeff_tet(int n, const double *__restrict__ tw, const double *__restrict__ vol, int div3,
                            T *__restrict__ out) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const size_t N = static_cast<size_t>(n);
  for (int i = 0; i < 4; i++) {
    double w = tw[i * N + x] / vol[x];
    out[i * N + x] = static_cast<T>(div3 ? w / 3 : w);
  }
}

// One row per thread (two rows per thread with int2 / T2 slot loads measured
// 9% slower on the 0-form Laplacian at k = 8).
template <int W, class T, class V>
__device__ __forceinline__ void b_ell(int blk, int n, int width, const int *__restrict__ col, const V *__restrict__ val,
                                      const T *__restrict__ x, T *__restrict__ y) {
  int r = blk * blockDim.x + threadIdx.x;
  if (r < n) stn(y + r, dec_ell::row<W, T, V>(n, width, r, col, val, x));
}
template <int W, class T, class V>
__global__ void k_ell(int n, int width, const int *__restrict__ col, const V *__restrict__ val,
                      const T *__restrict__ x, T *__restrict__ y, const int *__restrict__ sched) {
  int blk = tile_index(sched, static_cast<int>(blockIdx.x));
  if (blk >= 0) b_ell<W, T, V>(blk, n, width, col, val, x, y);
}
template <int WMAX, class T, class V>
__global__ void k_ell_upto(int n, int width, const int *__restrict__ col, const V *__restrict__ val,
                           const T *__restrict__ x, T *__restrict__ y) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) stn(y + r, dec_ell::row_upto<WMAX, T, V>(n, width, r, col, val, x));
}

// y = h ⊙ x, two entries per thread (T2 loads when the forms are aligned)
template <class T>
__device__ __forceinline__ void b_diag(int blk, int n, int vec, const T *__restrict__ h, const T *__restrict__ x,
                                       T *__restrict__ y) {
  int i = 2 * (blk * blockDim.x + threadIdx.x);
  if (vec && i + 1 < n) {
    V2<T> hv = ld2(h, i), xv = ld2(x, i);
    st2(y, i, hv.x * xv.x, hv.y * xv.y);
  } else {
    for (int j = i; j < i + 2 && j < n; j++) stn(y + j, h[j] * x[j]);
  }
}
template <class T>
__global__ void k_diag(int n, int vec, const T *__restrict__ h, const T *__restrict__ x, T *__restrict__ y) {
  b_diag<T>(static_cast<int>(blockIdx.x), n, vec, h, x, y);
}

template <class T>
__device__ __forceinline__ void b_avg01(int blk, int ne, int vec, const int *__restrict__ v0, const int *__restrict__ v1, const T *__restrict__ x,
                        T *__restrict__ y) {
  int e = 2 * (blk * blockDim.x + threadIdx.x);
  const T h = static_cast<T>(0.5);
  if (vec && e + 1 < ne) {
    int2 p = ldi2(v0, e), q = ldi2(v1, e);
    T a = x[p.x], b = x[q.x], c = x[p.y], d = x[q.y];
    st2(y, e, h * a + h * b, h * c + h * d);
  } else {
    for (int j = e; j < e + 2 && j < ne; j++) {
      stn(y + j, h * x[v0[j]] + h * x[v1[j]]);
    }
  }
}
template <class T>
__global__ void k_avg01(int ne, int vec, const int *__restrict__ v0, const int *__restrict__ v1, const T *__restrict__ x,
                        T *__restrict__ y) {
  b_avg01<T>(dec_ell::xcd_block(), ne, vec, v0, v1, x, y);
}

// (M g)[e] f[e], M = wedge_pd_01_mat (FastDEC.jl:329-346) in ELL with
// ascending columns (the CSC SpMV order), one row per thread.
template <int W, class T>
__device__ __forceinline__ void b_wedge_pd(int blk, int ne, int width, const int *__restrict__ col,
                                           const T *__restrict__ val, const T *__restrict__ g, const T *__restrict__ f,
                                           T *__restrict__ res) {
  int e = blk * blockDim.x + threadIdx.x;
  if (e < ne) stn(res + e, dec_ell::row<W, T, T>(ne, width, e, col, val, g) * f[e]);
}
template <int W, class T>
__global__ void k_wedge_pd(int ne, int width, const int *__restrict__ col, const T *__restrict__ val,
                           const T *__restrict__ g, const T *__restrict__ f, T *__restrict__ res) {
  b_wedge_pd<W, T>(static_cast<int>(blockIdx.x), ne, width, col, val, g, f, res);
}

// ---- geometric ⋆1⁻¹: a Krylov solve with Jacobi scaling, scalars on the device ----
// The reference factorizes -⋆1 (FastDEC.jl:745-748; GMRES on CUDA,
// ext/CombinatorialSpacesCUDAExt.jl:93-96).  ⋆1 is symmetric positive definite
// on a circumcentric dual (a mass matrix) and is solved by conjugate
// gradients; on a barycentric dual it is not symmetric and BiCGSTAB
// (van der Vorst 1992, right Jacobi preconditioning) is used.  All scalars
// stay in device memory; every kernel of an iteration returns at once when
// the stop flag is set, so a fixed launch sequence needs no host round trip.
constexpr int kCgChunk = 8;  // most iterations between host reads of the stop flag
enum { CS_BB = 0, CS_RR, CS_DONE, CS_ITER, CS_ALPHA, CS_OMEGA, CS_RZ0, CS_RZ1, CS_COUNT };

// The stop flag as read by thread 0, the same for every thread of the block.
// Block 0's thread 0 may set the flag later in the same launch; a block that
// reads it set takes the exit its own scalars would take, and reading once
// keeps the waves of a block on one path through its barriers.
__device__ bool block_stopped(const double *sc) {
  __shared__ int v;
  if (threadIdx.x == 0) v = sc[CS_DONE] != 0.0;
  __syncthreads();
  return v != 0;
}

// ---- fused conjugate gradients: two launches per iteration ----
// Every block of a launch reduces the previous launch's kFcBlocks partial sums
// itself, in one fixed order, so all blocks hold the same scalars and no
// scalar kernel runs between the vector kernels.  The search direction is
// double-buffered: the SpMV launch writes p_new = z + β p_old for its rows
// and forms the same values for its neighbours from z and p_old.
constexpr int kFcBlocks = 1024;

__device__ __forceinline__ void fc_tree(double *lds) {
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      lds[threadIdx.x] += lds[threadIdx.x + w];
      lds[256 + threadIdx.x] += lds[256 + threadIdx.x + w];
    }
    __syncthreads();
  }
}

// (Σ part[0..kFcBlocks), Σ part[kFcBlocks..2 kFcBlocks)), the same in every block
__device__ __forceinline__ void fc_sum(const double *__restrict__ part, double &s1, double &s2, double *lds) {
  double a = 0.0, b = 0.0;
  for (int j = threadIdx.x; j < kFcBlocks; j += 256) {
    a += part[j];
    b += part[kFcBlocks + j];
  }
  lds[threadIdx.x] = a;
  lds[256 + threadIdx.x] = b;
  __syncthreads();
  fc_tree(lds);
  s1 = lds[0];
  s2 = lds[256];
  __syncthreads();
}

__device__ __forceinline__ void fc_partial(double a, double b, double *__restrict__ part, double *lds) {
  lds[threadIdx.x] = a;
  lds[256 + threadIdx.x] = b;
  __syncthreads();
  fc_tree(lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = lds[0];
    part[kFcBlocks + blockIdx.x] = lds[256];
  }
}

// x = 0, r = b, z = D⁻¹ b, p_old = 0, scalars cleared; partials r·z, r·r
template <class T>
__global__ __launch_bounds__(256) void k_fcg_init(int n, const T *__restrict__ b, const T *__restrict__ dinv,
                                                  T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
                                                  T *__restrict__ p0, double *__restrict__ partB,
                                                  double *__restrict__ sc) {
  __shared__ double lds[512];
  if (blockIdx.x == 0 && threadIdx.x < CS_COUNT) sc[threadIdx.x] = 0.0;  // clears the stop flag
  double a1 = 0.0, a2 = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kFcBlocks * 256) {
    T bi = b[i];
    T zi = dinv[i] * bi;
    x[i] = static_cast<T>(0);
    r[i] = bi;
    z[i] = zi;
    p0[i] = static_cast<T>(0);
    a1 += static_cast<double>(bi) * static_cast<double>(zi);
    a2 += static_cast<double>(bi) * static_cast<double>(bi);
  }
  fc_partial(a1, a2, partB, lds);
}

// Iteration `it`: the stopping test on the residual after `it` updates, then
// p_new = z + β p_old, q = A p_new and partials p_new·q.  `last` only tests.
template <int W, class T>
__global__ __launch_bounds__(256) void k_fcg_spmv(int n, int width, const int *__restrict__ col,
                                                  const T *__restrict__ val, const T *__restrict__ z,
                                                  const T *__restrict__ pold, T *__restrict__ pnew,
                                                  T *__restrict__ q, const double *__restrict__ partB,
                                                  double *__restrict__ partA, double *__restrict__ sc, int it,
                                                  int last, double rtol) {
  __shared__ double lds[512];
  if (block_stopped(sc)) return;
  double rz, rr;
  fc_sum(partB, rz, rr, lds);
  double bb = it == 0 ? rr : sc[CS_BB];
  bool conv = rr <= rtol * rtol * bb;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it == 0) sc[CS_BB] = rr;
    sc[CS_RR] = rr;
    sc[CS_ITER] = it;
    sc[CS_RZ0 + (it & 1)] = rz;
    if (conv) sc[CS_DONE] = 1.0;
  }
  if (conv || last) return;
  double rzp = it == 0 ? 0.0 : sc[CS_RZ0 + ((it - 1) & 1)];
  T beta = static_cast<T>(rzp != 0.0 ? rz / rzp : 0.0);
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kFcBlocks * 256) {
    T pi = it == 0 ? z[i] : z[i] + beta * pold[i];
    pnew[i] = pi;
    T qi = static_cast<T>(0);
    if constexpr (W > 0) {
      int c[W];
      T v[W];
#pragma unroll
      for (int k = 0; k < W; k++) {
        size_t o = static_cast<size_t>(k) * n + i;
        c[k] = col[o];
        v[k] = val[o];
      }
#pragma unroll
      for (int k = 0; k < W; k++)
        if (c[k] >= 0) {
          T pc = it == 0 ? z[c[k]] : z[c[k]] + beta * pold[c[k]];
          qi = qi + v[k] * pc;
        }
    } else {
      for (int k = 0; k < width; k++) {
        size_t o = static_cast<size_t>(k) * n + i;
        int c = col[o];
        if (c < 0) break;
        T pc = it == 0 ? z[c] : z[c] + beta * pold[c];
        qi = qi + val[o] * pc;
      }
    }
    q[i] = qi;
    acc += static_cast<double>(pi) * static_cast<double>(qi);
  }
  fc_partial(acc, 0.0, partA, lds);
}

// x += α p, r -= α q, z = D⁻¹ r; partials r·z, r·r
template <class T>
__global__ __launch_bounds__(256) void k_fcg_update(int n, const T *__restrict__ p, const T *__restrict__ q,
                                                    const T *__restrict__ dinv, T *__restrict__ x,
                                                    T *__restrict__ r, T *__restrict__ z,
                                                    const double *__restrict__ partA, double *__restrict__ partB,
                                                    double *__restrict__ sc, int it) {
  __shared__ double lds[512];
  if (block_stopped(sc)) return;
  double pq, unused;
  fc_sum(partA, pq, unused, lds);
  if (pq == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[CS_DONE] = 1.0;
    return;
  }
  T alpha = static_cast<T>(sc[CS_RZ0 + (it & 1)] / pq);
  double a1 = 0.0, a2 = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kFcBlocks * 256) {
    x[i] = x[i] + alpha * p[i];
    T ri = r[i] - alpha * q[i];
    T zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    a1 += static_cast<double>(ri) * static_cast<double>(zi);
    a2 += static_cast<double>(ri) * static_cast<double>(ri);
  }
  fc_partial(a1, a2, partB, lds);
}

// ---- fused BiCGSTAB: five launches per iteration, same scalar scheme ----
// ρ_it = r̂·r sits in CS_RZ0 + (it & 1); α and ω of an iteration are written
// by one block of the launch that forms them and read by later launches.

// x = 0, r = r̂ = b, scalars cleared; partials r̂·r, r·r
template <class T>
__global__ __launch_bounds__(256) void k_fbi_init(int n, const T *__restrict__ b, T *__restrict__ x,
                                                  T *__restrict__ r, T *__restrict__ rh, double *__restrict__ partR,
                                                  double *__restrict__ sc) {
  __shared__ double lds[512];
  if (blockIdx.x == 0 && threadIdx.x < CS_COUNT) sc[threadIdx.x] = 0.0;
  double a = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kFcBlocks * 256) {
    T bi = b[i];
    x[i] = static_cast<T>(0);
    r[i] = bi;
    rh[i] = bi;
    a += static_cast<double>(bi) * static_cast<double>(bi);
  }
  fc_partial(a, a, partR, lds);
}

// Iteration `it`: the stopping test on the residual after `it` updates, then
// p = r + β (p - ω v) and y = D⁻¹ p.  `last` only tests.
template <class T>
__global__ __launch_bounds__(256) void k_fbi_dir(int n, const T *__restrict__ r, const T *__restrict__ v,
                                                 const T *__restrict__ dinv, T *__restrict__ p, T *__restrict__ y,
                                                 const double *__restrict__ partR, double *__restrict__ sc, int it,
                                                 int last, double rtol) {
  __shared__ double lds[512];
  if (block_stopped(sc)) return;
  double rho, rr;
  fc_sum(partR, rho, rr, lds);
  double bb = it == 0 ? rr : sc[CS_BB];
  bool stop = rr <= rtol * rtol * bb || rho == 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it == 0) sc[CS_BB] = rr;
    sc[CS_RR] = rr;
    sc[CS_ITER] = it;
    sc[CS_RZ0 + (it & 1)] = rho;
    if (stop) sc[CS_DONE] = 1.0;
  }
  if (stop || last) return;
  T beta = static_cast<T>(0), omega = static_cast<T>(0);
  if (it > 0) {
    double rhop = sc[CS_RZ0 + ((it - 1) & 1)];
    beta = static_cast<T>(rho / rhop * (sc[CS_ALPHA] / sc[CS_OMEGA]));
    omega = static_cast<T>(sc[CS_OMEGA]);
  }
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kFcBlocks * 256) {
    T pi = it == 0 ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    y[i] = dinv[i] * pi;
  }
}

// q = A x_in and partials a·q, q·q (b may be NULL: the second sum is then q·q)
template <int W, class T>
__global__ __launch_bounds__(256) void k_fbi_spmv(int n, int width, const int *__restrict__ col,
                                                  const T *__restrict__ val, const T *__restrict__ xin,
                                                  T *__restrict__ q, const T *__restrict__ a, int with_qq,
                                                  double *__restrict__ part, const double *__restrict__ sc) {
  __shared__ double lds[512];
  if (block_stopped(sc)) return;
  double s1 = 0.0, s2 = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kFcBlocks * 256) {
    T qi = dec_ell::row<W, T, T>(n, width, i, col, val, xin);
    q[i] = qi;
    s1 += static_cast<double>(a[i]) * static_cast<double>(qi);
    if (with_qq) s2 += static_cast<double>(qi) * static_cast<double>(qi);
  }
  fc_partial(s1, s2, part, lds);
}

// α = ρ / r̂·v; x += α y; s = r - α v (kept in r); yz = D⁻¹ s in place of y
template <class T>
__global__ __launch_bounds__(256) void k_fbi_half(int n, T *__restrict__ yz, const T *__restrict__ v,
                                                  const T *__restrict__ dinv, T *__restrict__ x, T *__restrict__ r,
                                                  const double *__restrict__ partV, double *__restrict__ sc, int it) {
  __shared__ double lds[512];
  if (block_stopped(sc)) return;
  double rv, unused;
  fc_sum(partV, rv, unused, lds);
  if (rv == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[CS_DONE] = 1.0;
    return;
  }
  double ad = sc[CS_RZ0 + (it & 1)] / rv;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[CS_ALPHA] = ad;
  T alpha = static_cast<T>(ad);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kFcBlocks * 256) {
    x[i] = x[i] + alpha * yz[i];
    T si = r[i] - alpha * v[i];
    r[i] = si;
    yz[i] = dinv[i] * si;
  }
}

// ω = t·s / t·t; x += ω z; r = s - ω t; partials r̂·r, r·r.  t = 0 means s = 0:
// x already solves the system.
template <class T>
__global__ __launch_bounds__(256) void k_fbi_full(int n, const T *__restrict__ z, const T *__restrict__ t,
                                                  const T *__restrict__ rh, T *__restrict__ x, T *__restrict__ r,
                                                  const double *__restrict__ partT, double *__restrict__ partR,
                                                  double *__restrict__ sc, int it) {
  __shared__ double lds[512];
  if (block_stopped(sc)) return;
  double ts, tt;
  fc_sum(partT, ts, tt, lds);
  if (tt == 0.0 || ts == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc[CS_DONE] = 1.0;
      if (tt == 0.0) {
        sc[CS_RR] = 0.0;
        sc[CS_ITER] = it + 1;
      }
    }
    return;
  }
  double od = ts / tt;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[CS_OMEGA] = od;
  T omega = static_cast<T>(od);
  double a1 = 0.0, a2 = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kFcBlocks * 256) {
    x[i] = x[i] + omega * z[i];
    T ri = r[i] - omega * t[i];
    r[i] = ri;
    a1 += static_cast<double>(rh[i]) * static_cast<double>(ri);
    a2 += static_cast<double>(ri) * static_cast<double>(ri);
  }
  fc_partial(a1, a2, partR, lds);
}

// out = sa a + sb b with sa, sb = ±1 (exact scalings)
template <class T>
__global__ void k_pm(int n, T sa, const T *__restrict__ a, T sb, const T *__restrict__ b, T *__restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sa * a[i] + sb * b[i];
}

template <class T>
__global__ void k_negate(int n, T *__restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = -x[i];
}

// ---- coefficient precomputation (float64) ----

// wedge_kernel_coeffs(Val(0), Val(2)) (FastDEC.jl:50-62): dual_area / area for
// the six dual triangles of t, summed pairwise.
template <class T>
__global__ void k_coeff02(int nt, const double *__restrict__ da, const double *__restrict__ area,
                          T *__restrict__ c0, T *__restrict__ c1, T *__restrict__ c2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double A = area[t];
  double k[6];
#pragma unroll
  for (int j = 0; j < 6; j++) k[j] = da[static_cast<size_t>(j) * nt + t] / A;
  c0[t] = static_cast<T>(k[0] + k[1]);
  c1[t] = static_cast<T>(k[2] + k[3]);
  c2[t] = static_cast<T>(k[4] + k[5]);
}

// wedge_kernel_coeffs(Val(1), Val(1)) (FastDEC.jl:78-93).
template <class T>
__global__ void k_coeff11(int nt, const int *__restrict__ e0, const int *__restrict__ e1,
                          const int *__restrict__ e2, const int8_t *__restrict__ eo,
                          const int8_t *__restrict__ to, const double *__restrict__ da,
                          const double *__restrict__ area, T *__restrict__ c0, T *__restrict__ c1,
                          T *__restrict__ c2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double ar = ((to[t] ? 1.0 : -1.0) * 2) * area[t];
  double s0 = eo[e0[t]] ? 1.0 : -1.0, s1 = eo[e1[t]] ? 1.0 : -1.0, s2 = eo[e2[t]] ? 1.0 : -1.0;
  size_t n = static_cast<size_t>(nt);
  c0[t] = static_cast<T>(((s1 * s2) * (da[t] + da[t + n])) / ar);
  c1[t] = static_cast<T>(((s0 * s2) * (da[t + 2 * n] + da[t + 3 * n])) / ar);
  c2[t] = static_cast<T>(((s0 * s1) * (da[t + 4 * n] + da[t + 5 * n])) / ar);
}

// wedge_dd_01_mat weights: dual_length / Σ dual_length over e's elementary duals.
template <class T>
__global__ void k_coeff_dd(int ne, const int *__restrict__ ptr, const int *__restrict__ idx,
                           const double *__restrict__ dlB, T *__restrict__ wa, T *__restrict__ wb) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int a = ptr[e], cnt = ptr[e + 1] - a;
  double la = cnt > 0 ? dlB[idx[a]] : 0.0, lb = cnt > 1 ? dlB[idx[a + 1]] : 0.0;
  double s = cnt > 1 ? la + lb : la;
  // weights are listed in ascending triangle order (the CSC row order)
  wa[e] = static_cast<T>(cnt > 0 ? la / s : 0.0);
  wb[e] = static_cast<T>(cnt > 1 ? lb / s : 0.0);
}

template <class T>
__global__ void k_cast(size_t n, const double *__restrict__ x, T *__restrict__ y) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = static_cast<T>(x[i]);
}

template <class T>
__global__ void k_inv_diag(int n, const double *__restrict__ h, double mult, T *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  // dec_inv_hodge_star: (1 ./ hdg) .* (-1)^(k(2-k)) (FastDEC.jl:725-730)
  if (i < n) y[i] = static_cast<T>((1 / h[i]) * mult);
}

// Geometric ⋆1, per triangle (FastDEC.jl:660-716): nine candidate entries
// (three diagonal, six off-diagonal in the reference's permutation order);
// an entry equal to 0.0 is not stored by the reference and is dropped here.
struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 sub3(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 cross3(D3 a, D3 b) {
  return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dot3(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ double sgnd(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }
// crossdot (DiscreteExteriorCalculus.jl:1667-1670)
__device__ __forceinline__ double crossdot(D3 a, D3 b) {
  D3 c = cross3(a, b);
  return sqrt(dot3(c, c)) * (c.z == 0 ? 1.0 : sgnd(c.z));
}
__device__ __forceinline__ D3 ld(const double *p, size_t i) { return D3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

__global__ void k_geo_hodge1_tri(int nv, int ne, int nt, const int *__restrict__ v0, const int *__restrict__ v1,
                                 const int *__restrict__ e0, const int *__restrict__ e1,
                                 const int *__restrict__ e2, const double *__restrict__ dpt,
                                 double *__restrict__ out9) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int te[3] = {e2[t], e1[t], e0[t]};  // reversed by construction
  D3 tc = ld(dpt, static_cast<size_t>(nv) + ne + t);
  D3 evt[3], dvt[3];
  for (int i = 0; i < 3; i++) {
    int e = te[i];
    evt[i] = sub3(ld(dpt, v0[e]), ld(dpt, v1[e]));
    dvt[i] = sub3(tc, ld(dpt, static_cast<size_t>(nv) + e));
  }
  dvt[1] = D3{-1 * dvt[1].x, -1 * dvt[1].y, -1 * dvt[1].z};
  D3 cr = cross3(evt[0], dvt[0]);
  double rel = cr.z == 0 ? 1.0 : sgnd(cr.z);
  double *o = out9 + static_cast<size_t>(9) * t;
  for (int i = 0; i < 3; i++) {
    double dc = crossdot(evt[i], dvt[i]);
    o[i] = dc != 0.0 ? rel * dc / dot3(evt[i], evt[i]) : 0.0;
  }
  const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int q = 0; q < 6; q++) {
    int a = P[q][0], b = P[q][1], c = P[q][2];
    double ddot = dot3(evt[a], dvt[a]) / dot3(evt[a], evt[a]);
    double val = ddot * dot3(evt[a], evt[c]);
    o[3 + q] = val != 0.0 ? rel * val / crossdot(evt[b], evt[c]) : 0.0;
  }
}

template <class T>
void launch_cast(const DBuf<double> &src, DBuf<unsigned char> &dst) {
  dst.alloc(sizeof(T) * src.n);
  if (src.n)
    hipLaunchKernelGGL(k_cast<T>, dim3(dec::nblocks(static_cast<long long>(src.n))), dim3(256), 0, 0, src.n, src.p,
                       reinterpret_cast<T *>(dst.p));
}

}  // namespace

namespace dec {

// Per-row restatement of the SpGEMM chain ((((-1 * D(inv_h0)) * d̃1) * H) * d0)
// in Gustavson order:
//   A[v,e] = inv_h0[v] * d0[e,v]              (exact: d0 = ±1)
//   X[v,f] = Σ_{e ∋ v, ascending} A[v,e] * H[e,f]
//   L[v,w] = Σ_{f ∋ w, ascending} X[v,f] * d0[f,w]
void assemble_laplacian0(const Topo &t, const std::vector<double> &inv_h0, const std::vector<double> &h1,
                         const HostCsr *H, HostCsr &L) {
  Csr ve;
  vertex_edges(t, ve);
  int nv = t.nv;
  L.m = L.n = nv;
  L.ptr.assign(nv + 1, 0);
  L.col.clear();
  L.val.clear();
  std::vector<std::pair<int, double>> X;   // (f, value), f in first-touch order
  std::vector<std::pair<int, double>> C;   // (w, value)
  std::vector<int> xpos;
  for (int v = 0; v < nv; v++) {
    X.clear();
    C.clear();
    double iv = inv_h0[v];
    for (int k = ve.ptr[v]; k < ve.ptr[v + 1]; k++) {
      int e = ve.idx[k];
      double a = iv * static_cast<double>(ve.val[k]);
      auto add = [&](int f, double hval) {
        double prod = a * hval;
        for (auto &x : X)
          if (x.first == f) {
            x.second += prod;
            return;
          }
        X.emplace_back(f, prod);
      };
      if (H) {
        for (int q = H->ptr[e]; q < H->ptr[e + 1]; q++) add(H->col[q], H->val[q]);
      } else {
        add(e, h1[e]);
      }
    }
    std::sort(X.begin(), X.end(), [](const auto &p, const auto &q) { return p.first < q.first; });
    for (auto &x : X) {
      int f = x.first;
      double s = t.eo[f] == 1 ? 1.0 : -1.0;
      int ws[2] = {t.v0[f], t.v1[f]};
      double ds[2] = {s, -s};
      for (int j = 0; j < 2; j++) {
        int w = ws[j];
        double prod = x.second * ds[j];
        bool found = false;
        for (auto &c : C)
          if (c.first == w) {
            c.second += prod;
            found = true;
            break;
          }
        if (!found) C.emplace_back(w, prod);
      }
    }
    std::sort(C.begin(), C.end(), [](const auto &p, const auto &q) { return p.first < q.first; });
    for (auto &c : C) {
      L.col.push_back(c.first);
      L.val.push_back(c.second);
    }
    L.ptr[v + 1] = static_cast<int32_t>(L.col.size());
  }
}

template <class V>
void upload_ell(const HostCsr &A, Ell<V> &E) {
  int n = A.m, w = 0;
  for (int r = 0; r < n; r++) w = std::max(w, A.ptr[r + 1] - A.ptr[r]);
  E.nrows = n;
  E.width = w;
  std::vector<int32_t> col(static_cast<size_t>(n) * std::max(w, 1), -1);
  std::vector<V> val(static_cast<size_t>(n) * std::max(w, 1), static_cast<V>(0));
  for (int r = 0; r < n; r++)
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
      size_t q = static_cast<size_t>(k - A.ptr[r]) * n + r;
      col[q] = A.col[k];
      val[q] = static_cast<V>(A.val[k]);
    }
  E.col.upload(col);
  E.val.upload(val);
}
template void upload_ell<double>(const HostCsr &, Ell<double> &);
template void upload_ell<float>(const HostCsr &, Ell<float> &);

template <class T>
void ell_spmv(const Ell<T> &E, const T *x, T *y, hipStream_t s, const int *sched, int nsched) {
  if (E.nrows == 0) return;
  // rows of 9..16 slots (tetrahedral ∇²₀, the dual Laplacians' products) issue
  // all their loads at once; the grouped launch keeps the generic loop so that
  // its register budget stays that of width 8
  if (E.width > 8 && E.width <= 12) {
    hipLaunchKernelGGL((k_ell_upto<12, T, T>), dim3(nblocks(E.nrows)), dim3(256), 0, s, E.nrows, E.width, E.col.p,
                       E.val.p, x, y);
    return;
  }
  if (E.width > 12 && E.width <= 16) {
    hipLaunchKernelGGL((k_ell_upto<16, T, T>), dim3(nblocks(E.nrows)), dim3(256), 0, s, E.nrows, E.width, E.col.p,
                       E.val.p, x, y);
    return;
  }
  DEC_ELL_WIDTH_SWITCH(E.width, hipLaunchKernelGGL((k_ell<W, T, T>), dim3(sched ? nsched : nblocks(E.nrows)),
                                                    dim3(256), 0, s, E.nrows, E.width, E.col.p, E.val.p, x, y, sched));
}
template void ell_spmv<double>(const Ell<double> &, const double *, double *, hipStream_t, const int *, int);
template void ell_spmv<float>(const Ell<float> &, const float *, float *, hipStream_t, const int *, int);

// Host CSR of the geometric ⋆1 from the per-triangle entries: rows ascending
// columns, duplicates (diagonals shared by two triangles) summed in the
// reference's input order (triangle, then slot), zero entries dropped.
void geo_hodge1_csr(const Topo &t, const std::vector<double> &v9, HostCsr &H) {
  int ne = t.ne, nt = t.nt;
  struct Ent {
    int row, col;
    long long order;
    double val;
  };
  std::vector<Ent> ents;
  ents.reserve(static_cast<size_t>(9) * nt);
  const int P[6][2] = {{0, 1}, {0, 2}, {1, 0}, {1, 2}, {2, 0}, {2, 1}};
  for (int i = 0; i < nt; i++) {
    int te[3] = {t.e2[i], t.e1[i], t.e0[i]};
    for (int a = 0; a < 3; a++) {
      double v = v9[static_cast<size_t>(9) * i + a];
      if (v != 0.0) ents.push_back({te[a], te[a], 9LL * i + a, v});
    }
    for (int q = 0; q < 6; q++) {
      double v = v9[static_cast<size_t>(9) * i + 3 + q];
      if (v != 0.0) ents.push_back({te[P[q][0]], te[P[q][1]], 9LL * i + 3 + q, v});
    }
  }
  std::sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
    if (a.row != b.row) return a.row < b.row;
    if (a.col != b.col) return a.col < b.col;
    return a.order < b.order;
  });
  H.m = H.n = ne;
  H.ptr.assign(ne + 1, 0);
  H.col.clear();
  H.val.clear();
  for (size_t k = 0; k < ents.size(); k++) {
    if (k > 0 && ents[k].row == ents[k - 1].row && ents[k].col == ents[k - 1].col) {
      H.val.back() += ents[k].val;
      continue;
    }
    H.col.push_back(ents[k].col);
    H.val.push_back(ents[k].val);
    H.ptr[ents[k].row + 1] = static_cast<int32_t>(H.col.size());
  }
  for (int r = 0; r < ne; r++) H.ptr[r + 1] = std::max(H.ptr[r + 1], H.ptr[r]);
}

void geo_hodge1(dec_mesh_s &m, HostCsr &H) {
  const Topo &t = m.h;
  DBuf<double> v9;
  v9.alloc(static_cast<size_t>(9) * t.nt);
  if (t.nt > 0)
    hipLaunchKernelGGL(k_geo_hodge1_tri, dim3(nblocks(t.nt)), dim3(256), 0, 0, t.nv, t.ne, t.nt, m.v0.p, m.v1.p,
                       m.e0.p, m.e1.p, m.e2.p, m.dpt.p, v9.p);
  check_launch("geometric hodge");
  geo_hodge1_csr(t, v9.download(), H);
}

// C = A B over host CSR, each entry accumulated over k ascending (the order
// of SparseArrays' column-by-column product), columns ascending.
void spgemm(const HostCsr &A, const HostCsr &B, HostCsr &C) {
  if (A.n != B.m) fail(DEC_ERR_ARG, "spgemm: shape mismatch");
  C.m = A.m;
  C.n = B.n;
  C.ptr.assign(A.m + 1, 0);
  C.col.clear();
  C.val.clear();
  std::vector<double> acc(B.n, 0.0);
  std::vector<int> mark(B.n, -1), cols;
  for (int i = 0; i < A.m; i++) {
    cols.clear();
    for (int ka = A.ptr[i]; ka < A.ptr[i + 1]; ka++) {
      int k = A.col[ka];
      double av = A.val[ka];
      for (int kb = B.ptr[k]; kb < B.ptr[k + 1]; kb++) {
        int j = B.col[kb];
        if (mark[j] != i) {
          mark[j] = i;
          acc[j] = av * B.val[kb];
          cols.push_back(j);
        } else {
          acc[j] = acc[j] + av * B.val[kb];
        }
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int j : cols) {
      C.col.push_back(j);
      C.val.push_back(acc[j]);
    }
    C.ptr[i + 1] = static_cast<int32_t>(C.col.size());
  }
}

}  // namespace dec

namespace {

void to_sign_ell(const dec::Csr &c, dec::Ell<int8_t> &E, bool negate) {
  int n = c.n, w = 0;
  for (int r = 0; r < n; r++) w = std::max(w, c.ptr[r + 1] - c.ptr[r]);
  E.nrows = n;
  E.width = w;
  std::vector<int32_t> col(static_cast<size_t>(n) * std::max(w, 1), -1);
  std::vector<int8_t> val(static_cast<size_t>(n) * std::max(w, 1), 0);
  for (int r = 0; r < n; r++)
    for (int k = c.ptr[r]; k < c.ptr[r + 1]; k++) {
      size_t q = static_cast<size_t>(k - c.ptr[r]) * n + r;
      col[q] = c.idx[k];
      val[q] = static_cast<int8_t>(negate ? -c.val[k] : c.val[k]);
    }
  E.col.upload(col);
  E.val.upload(val);
}

void csr_from_signs(const dec::Csr &c, int ncols, bool negate, dec::HostCsr &h) {
  h.m = c.n;
  h.n = ncols;
  h.ptr = c.ptr;
  h.col = c.idx;
  h.val.resize(c.val.size());
  for (size_t k = 0; k < c.val.size(); k++) h.val[k] = negate ? -c.val[k] : c.val[k];
}

void csr_d0(const dec::Topo &t, dec::HostCsr &h) {
  h.m = t.ne;
  h.n = t.nv;
  h.ptr.assign(t.ne + 1, 0);
  h.col.clear();
  h.val.clear();
  for (int e = 0; e < t.ne; e++) {
    double s = t.eo[e] == 1 ? 1.0 : -1.0;
    int a = t.v0[e], b = t.v1[e];
    if (a < b) {
      h.col.insert(h.col.end(), {a, b});
      h.val.insert(h.val.end(), {s, -s});
    } else {
      h.col.insert(h.col.end(), {b, a});
      h.val.insert(h.val.end(), {-s, s});
    }
    h.ptr[e + 1] = static_cast<int32_t>(h.col.size());
  }
}

void csr_d1(const dec::Topo &t, dec::HostCsr &h) {
  h.m = t.nt;
  h.n = t.ne;
  h.ptr.assign(t.nt + 1, 0);
  h.col.clear();
  h.val.clear();
  for (int i = 0; i < t.nt; i++) {
    double ts = t.to[i] == 1 ? 1.0 : -1.0;
    std::pair<int, double> ent[3] = {{t.e0[i], (t.eo[t.e0[i]] == 1 ? 1.0 : -1.0) * ts},
                                     {t.e1[i], -1 * (t.eo[t.e1[i]] == 1 ? 1.0 : -1.0) * ts},
                                     {t.e2[i], (t.eo[t.e2[i]] == 1 ? 1.0 : -1.0) * ts}};
    std::sort(ent, ent + 3, [](auto &a, auto &b) { return a.first < b.first; });
    for (auto &x : ent) {
      h.col.push_back(x.first);
      h.val.push_back(x.second);
    }
    h.ptr[i + 1] = static_cast<int32_t>(h.col.size());
  }
}

void csr_diag(const std::vector<double> &d, dec::HostCsr &h) {
  int n = static_cast<int>(d.size());
  h.m = h.n = n;
  h.ptr.resize(n + 1);
  h.col.resize(n);
  h.val = d;
  for (int i = 0; i <= n; i++) h.ptr[i] = i;
  for (int i = 0; i < n; i++) h.col[i] = i;
}

inline void hip_status(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(DEC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// out = -⋆1⁻¹ x, the solve of ⋆1 z = x negated as dec_inv_hodge_star's
// (-⋆1)⁻¹ x.  Iterations go out in chunks; between chunks the host reads the
// device stop flag (one 8-byte copy and a stream synchronisation).  While the
// stream is being captured into a graph the whole iteration budget is
// recorded instead, and iterations past convergence return at once.
template <class T>
void ginv_apply(dec_op_s &op, const T *bvec, T *x, hipStream_t s) {
  int n = static_cast<int>(op.rows);
  const dec::Ell<T> *E;
  if constexpr (sizeof(T) == 8) E = &op.ell64;
  else E = &op.ell32;
  auto P = [](DBuf<unsigned char> &v) { return reinterpret_cast<T *>(v.p); };
  T *r = P(op.cr), *p = P(op.cp), *q = P(op.cq), *z = P(op.cz), *rh = P(op.crh), *t = P(op.ct);
  const T *dinv = reinterpret_cast<const T *>(op.dinv.p);
  double *sc = op.cscal.p;
  const bool bicg = !op.symmetric;
  dim3 g(dec::nblocks(n)), blk(256), gf(kFcBlocks);
  // three partial-sum pairs: A and B for CG; R, V and T for BiCGSTAB
  double *partA = op.cpart.p, *partB = partA + 2 * kFcBlocks, *partC = partB + 2 * kFcBlocks;
  T *pbuf[2] = {p, rh};  // CG: the direction, double-buffered
  if (bicg) hipLaunchKernelGGL(k_fbi_init<T>, gf, blk, 0, s, n, bvec, x, r, rh, partA, sc);
  else hipLaunchKernelGGL(k_fcg_init<T>, gf, blk, 0, s, n, bvec, dinv, x, r, z, pbuf[0], partB, sc);
  int it = 0;
  // BiCGSTAB buffers: p, y = z (then D⁻¹ s), v = q, t; r holds s between the halves
  auto spmv = [&](const T *xin, T *out, const T *a, int with_qq, double *part) {
    DEC_ELL_WIDTH_SWITCH(E->width, hipLaunchKernelGGL((k_fbi_spmv<W, T>), gf, blk, 0, s, n, E->width, E->col.p,
                                                      E->val.p, xin, out, a, with_qq, part, sc));
  };
  auto test = [&](int last) {
    if (!bicg) {
      DEC_ELL_WIDTH_SWITCH(E->width, hipLaunchKernelGGL((k_fcg_spmv<W, T>), gf, blk, 0, s, n, E->width, E->col.p,
                                                        E->val.p, z, pbuf[it & 1], pbuf[(it + 1) & 1], q, partB,
                                                        partA, sc, it, last, op.rtol));
    } else {
      hipLaunchKernelGGL(k_fbi_dir<T>, gf, blk, 0, s, n, r, q, dinv, p, z, partA, sc, it, last, op.rtol);
    }
  };
  auto iteration = [&]() {
    if (!bicg) {
      test(0);
      hipLaunchKernelGGL(k_fcg_update<T>, gf, blk, 0, s, n, pbuf[(it + 1) & 1], q, dinv, x, r, z, partA, partB, sc,
                         it);
    } else {
      test(0);                                          // ρ, β, p, y
      spmv(z, q, rh, 0, partB);                         // v = A y, r̂·v
      hipLaunchKernelGGL(k_fbi_half<T>, gf, blk, 0, s, n, z, q, dinv, x, r, partB, sc, it);
      spmv(z, t, r, 1, partC);                          // t = A z, t·s, t·t
      hipLaunchKernelGGL(k_fbi_full<T>, gf, blk, 0, s, n, z, t, rh, x, r, partC, partA, sc, it);
    }
    it++;
  };
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  hip_status(hipStreamIsCapturing(s, &cap), "hipStreamIsCapturing");
  // checks after 2, 6, 14, 22, ... iterations: the diagonal-dominant ⋆1
  // converges in a few, so the first checks come early
  int chunk = cap == hipStreamCaptureStatusNone ? 2 : op.max_iter;
  for (int done_it = 0; done_it < op.max_iter; done_it += chunk, chunk = std::min(2 * chunk, kCgChunk)) {
    for (int k = 0; k < chunk && done_it + k < op.max_iter; k++) iteration();
    // test the latest residual now, so that the flag read below sees it
    if (done_it + chunk >= op.max_iter || cap == hipStreamCaptureStatusNone) test(1);
    if (cap != hipStreamCaptureStatusNone) break;
    double flag = 0.0;
    hip_status(hipMemcpyAsync(&flag, sc + CS_DONE, sizeof(double), hipMemcpyDeviceToHost, s), "stop flag copy");
    hip_status(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (flag != 0.0) break;
  }
  hipLaunchKernelGGL(k_negate<T>, g, blk, 0, s, n, x);
}

template <class T>
void comp_apply(dec_op_s &op, const T *x, const T *y, T *out, hipStream_t s);

template <class T>
void apply_t(dec_op_s &op, const void *xv, const void *yv, void *outv, hipStream_t s) {
  const T *x = static_cast<const T *>(op.swap ? yv : xv);
  const T *y = static_cast<const T *>(op.swap ? xv : yv);
  T *out = static_cast<T *>(outv);
  int rows = static_cast<int>(op.rows);
  if (rows == 0) return;
  dim3 g(dec::xcd_grid(rows)), b(256);
  dim3 g2(dec::xcd_grid((static_cast<long long>(rows) + 1) / 2));
  // the pair kernels load the caller's forms as 2-vectors: require alignment
  const uintptr_t A = 2 * sizeof(T);
  int vec = (reinterpret_cast<uintptr_t>(out) % A == 0) && (!y || reinterpret_cast<uintptr_t>(y) % A == 0) &&
            (!x || reinterpret_cast<uintptr_t>(x) % A == 0);
  auto C = [](DBuf<unsigned char> &c) { return reinterpret_cast<const T *>(c.p); };
  // grid of a gather kernel: one workgroup per schedule entry when it has one
  auto gs = [&](dim3 dflt) { return op.sched.p ? dim3(op.nsched) : dflt; };
  switch (op.impl) {
    case I_W01:
      hipLaunchKernelGGL(k_wedge01<T>, gs(g2), b, 0, s, rows, vec, op.i0.p, op.i1.p, x, y, out, op.sched.p);
      break;
    case I_W02:
      hipLaunchKernelGGL(k_wedge02<T>, gs(g2), b, 0, s, rows, vec, op.i0.p, op.i1.p, op.i2.p, C(op.c0), C(op.c1),
                         C(op.c2), x, y, out, op.sched.p);
      break;
    case I_W11:
      hipLaunchKernelGGL(k_wedge11<T>, gs(g2), b, 0, s, rows, vec, op.i0.p, op.i1.p, op.i2.p, C(op.c0), C(op.c1),
                         C(op.c2), x, y, out, op.sched.p);
      break;
    case I_WDD:
      if (op.m == 0)  // (m*f) .* g : f = x (dual 0-form), g = y
        hipLaunchKernelGGL(k_wedge_dd<T>, g2, b, 0, s, rows, vec, op.i0.p, op.i1.p, C(op.c0), C(op.c1), x, y, out);
      else  // f .* (m*g)
        hipLaunchKernelGGL(k_wedge_dd<T>, g2, b, 0, s, rows, vec, op.i0.p, op.i1.p, C(op.c0), C(op.c1), y, x, out);
      break;
    case I_D0:
      hipLaunchKernelGGL(k_d0<T>, gs(g2), b, 0, s, rows, vec, op.i0.p, op.i1.p, op.s8.p, x, out, op.sched.p);
      break;
    case I_D1:
      hipLaunchKernelGGL(k_d1<T>, gs(g2), b, 0, s, rows, vec, op.i0.p, op.i1.p, op.i2.p, op.s8.p, x, out,
                         op.sched.p);
      break;
    case I_SELL:
      DEC_ELL_WIDTH_SWITCH(op.sell.width, hipLaunchKernelGGL((k_ell<W, T, int8_t>), g, b, 0, s, rows, op.sell.width,
                                                             op.sell.col.p, op.sell.val.p, x, out,
                                                             static_cast<const int *>(nullptr)));
      break;
    case I_DIAG:
      hipLaunchKernelGGL(k_diag<T>, g2, b, 0, s, rows, vec, C(op.c0), x, out);
      break;
    case I_ELL:
      if constexpr (sizeof(T) == 8) dec::ell_spmv(op.ell64, x, out, s, op.sched.p, op.nsched);
      else dec::ell_spmv(op.ell32, x, out, s, op.sched.p, op.nsched);
      break;
    case I_AVG:
      hipLaunchKernelGGL(k_avg01<T>, g2, b, 0, s, rows, vec, op.i0.p, op.i1.p, x, out);
      break;
    case I_W03:
      hipLaunchKernelGGL(k_wedge03<T>, g, b, 0, s, rows, op.ix.p, C(op.cx), x, y, out);
      break;
    case I_W21:
      hipLaunchKernelGGL(k_wedge21<T>, g, b, 0, s, rows, op.ix.p, C(op.cx), op.m16.p, x, y, out);
      break;
    case I_GINV:
      ginv_apply<T>(op, x, out, s);
      break;
    case I_COMP:
      comp_apply<T>(op, x, y, out, s);
      break;
    case I_WPD: {
      const dec::Ell<T> *E;
      if constexpr (sizeof(T) == 8) E = &op.ell64;
      else E = &op.ell32;
      DEC_ELL_WIDTH_SWITCH(E->width, hipLaunchKernelGGL((k_wedge_pd<W, T>), g, b, 0, s, rows, E->width, E->col.p,
                                                        E->val.p, x, y, out));
      break;
    }
    default:
      fail(DEC_ERR_ARG, "dec_op_apply: corrupt operator");
  }
  dec::check_launch("dec_op_apply");
}

// Composite operators: the sub-operator applications of the reference's
// closures, in the same order, through device work vectors.
template <class T>
void comp_apply(dec_op_s &op, const T *x, const T *y, T *out, hipStream_t s) {
  auto W = [&](int i) { return reinterpret_cast<T *>(op.tmp[i].p); };
  auto A = [&](int i, const T *a, const T *b, T *o) { apply_t<T>(*op.sub[i], a, b, o, s); };
  auto pm = [&](int64_t n, T sa, const T *a, T sb, const T *b, T *o) {
    if (n) hipLaunchKernelGGL(k_pm<T>, dim3(dec::nblocks(n)), dim3(256), 0, s, static_cast<int>(n), sa, a, sb, b, o);
  };
  switch (op.comp) {
    case C_WPD11:  // FastDEC.jl:390-400
      if (op.kind == DEC_OP_WEDGE_PD) {  // (f, g) -> Λ11(f, ♭♯ g)
        A(0, y, nullptr, W(0));
        A(1, x, W(0), out);
      } else {  // (f, g) -> Λ11(♭♯ f, g)
        A(0, x, nullptr, W(0));
        A(1, W(0), y, out);
      }
      break;
    case C_I11:  // (f, g) -> ⋆2 Λpd11(⋆1⁻¹ g, f) (FastDEC.jl:768-773)
      A(0, y, nullptr, W(0));
      A(1, x, nullptr, W(1));
      A(2, W(0), W(1), W(2));
      A(3, W(2), nullptr, out);
      break;
    case C_I12:  // (f, g) -> ⋆1 ♭♯ ((Λ01 ⋆0⁻¹ g) .* f) (FastDEC.jl:779-786)
      A(0, y, nullptr, W(0));
      A(1, W(0), x, W(1));
      A(2, W(1), nullptr, W(2));
      A(3, W(2), nullptr, out);
      break;
    case C_LIE:  // (f, g) -> -(d̃0 i11(f, g)) - i12(f, d̃1 g) (FastDEC.jl:796-803)
      A(0, x, y, W(0));
      A(1, W(0), nullptr, W(1));
      A(2, y, nullptr, W(2));
      A(3, x, W(2), W(3));
      pm(op.rows, static_cast<T>(-1), W(1), static_cast<T>(-1), W(3), out);
      break;
    case C_DL0:  // x -> (⋆2 d1) ⋆1⁻¹(d̃0 x) (FastDEC.jl:821-829)
      A(0, x, nullptr, W(0));
      A(1, W(0), nullptr, W(1));
      A(2, W(1), nullptr, out);
      break;
    case C_DL1:  // x -> M x + N ⋆1⁻¹(x) (FastDEC.jl:845-860)
      A(0, x, nullptr, W(0));
      A(1, x, nullptr, W(1));
      A(2, W(0), nullptr, W(2));
      pm(op.rows, static_cast<T>(1), W(1), static_cast<T>(1), W(2), out);
      break;
    default:
      fail(DEC_ERR_ARG, "dec_op_apply: corrupt composite operator");
  }
}

// ---- grouped launch: several operators in one kernel -----------------------
// Workgroups are dealt to the operators in order (start[i] .. start[i+1]-1
// run operator i, with that operator's own grid); each workgroup runs the
// same device body as the single-operator kernel, so results are identical.
// One launch replaces `count` launches and their ramp-up and drain.
constexpr int kMaxGroup = 16;

struct GroupOp {
  int impl, rows, vec, width, swap_wdd;
  const int *i0, *i1, *i2;
  const int8_t *s8;
  const void *c0, *c1, *c2;
  const int *ecol;
  const void *eval;
  const int *ix;
  const void *cx;
  const uint16_t *m16;
  const int *sched;  // spatial tile schedule (null: tiles in order)
  const void *x, *y;
  void *out;
};

struct GroupArgs {
  int count;
  int start[kMaxGroup + 1];
  GroupOp op[kMaxGroup];
};

template <class T>
__global__ void __launch_bounds__(256) k_apply_group(GroupArgs g) {
  int b = blockIdx.x, i = 0;
  while (i + 1 < g.count && b >= g.start[i + 1]) i++;
  const GroupOp &o = g.op[i];
  int blk = b - g.start[i];
  // every operator's range starts at a multiple of 8, so b and its local
  // index run on the same XCD (b % 8), as the schedule assumes
  if (o.sched) {
    blk = o.sched[blk];
    if (blk < 0) return;
  }
  const T *x = static_cast<const T *>(o.x), *y = static_cast<const T *>(o.y);
  const T *c0 = static_cast<const T *>(o.c0), *c1 = static_cast<const T *>(o.c1), *c2 = static_cast<const T *>(o.c2);
  T *out = static_cast<T *>(o.out);
  switch (o.impl) {
    case I_W01: b_wedge01<T>(blk, o.rows, o.vec, o.i0, o.i1, x, y, out); break;
    case I_W02: b_wedge02<T>(blk, o.rows, o.vec, o.i0, o.i1, o.i2, c0, c1, c2, x, y, out); break;
    case I_W11: b_wedge11<T>(blk, o.rows, o.vec, o.i0, o.i1, o.i2, c0, c1, c2, x, y, out); break;
    case I_WDD:
      if (o.swap_wdd) b_wedge_dd<T>(blk, o.rows, o.vec, o.i0, o.i1, c0, c1, y, x, out);
      else b_wedge_dd<T>(blk, o.rows, o.vec, o.i0, o.i1, c0, c1, x, y, out);
      break;
    case I_D0: b_d0<T>(blk, o.rows, o.vec, o.i0, o.i1, o.s8, x, out); break;
    case I_D1: b_d1<T>(blk, o.rows, o.vec, o.i0, o.i1, o.i2, o.s8, x, out); break;
    case I_SELL:
      DEC_ELL_WIDTH_SWITCH(o.width, (b_ell<W, T, int8_t>(blk, o.rows, o.width, o.ecol,
                                                          static_cast<const int8_t *>(o.eval), x, out)));
      break;
    case I_DIAG: b_diag<T>(blk, o.rows, o.vec, c0, x, out); break;
    case I_ELL:
      DEC_ELL_WIDTH_SWITCH(o.width, (b_ell<W, T, T>(blk, o.rows, o.width, o.ecol,
                                                     static_cast<const T *>(o.eval), x, out)));
      break;
    case I_AVG: b_avg01<T>(blk, o.rows, o.vec, o.i0, o.i1, x, out); break;
    case I_W03: b_wedge03<T>(blk, o.rows, o.ix, static_cast<const T *>(o.cx), x, y, out); break;
    case I_W21: b_wedge21<T>(blk, o.rows, o.ix, static_cast<const T *>(o.cx), o.m16, x, y, out); break;
    case I_WPD:
      DEC_ELL_WIDTH_SWITCH(o.width, (b_wedge_pd<W, T>(blk, o.rows, o.width, o.ecol, static_cast<const T *>(o.eval), x, y,
                                                      out)));
      break;
    default: break;
  }
}

// Workgroups of one operator in the grouped launch (its own grid, without the
// XCD rounding, and without the grid-stride cap of the diagonal kernel).
// Rounded up to a multiple of 8 (the XCD count), so that every operator's
// range of workgroups starts on XCD 0; surplus workgroups find no rows.
int group_blocks(const dec_op_s &op) {
  long long rows = op.rows;
  int nb;
  if (op.sched.p) {
    nb = op.nsched;
  } else {
    switch (op.impl) {
      case I_SELL:
      case I_ELL:
      case I_WPD:
      case I_W03:
      case I_W21: nb = dec::nblocks(rows, 256); break;
      default: nb = dec::nblocks((rows + 1) / 2, 256); break;
    }
  }
  return (nb + 7) / 8 * 8;
}

template <class T>
void apply_group_t(int count, dec_op_s *const *ops, const void *const *xs, const void *const *ys, void *const *outs,
                   hipStream_t s) {
  for (int base = 0; base < count; base += kMaxGroup) {
    GroupArgs g{};
    int k = 0, blocks = 0;
    // Launch order: sparse-matrix operators first (their gathers make the
    // slowest workgroups), the streaming diagonal ones last to fill the
    // drain; measured 3% faster than the caller's order on the k = 8 set.
    std::vector<int> order;
    for (int q = base; q < count && q < base + kMaxGroup; q++) order.push_back(q);
    auto rank = [&](int q) {
      int m = ops[q]->impl;
      return m == I_ELL || m == I_SELL || m == I_WPD ? 0 : (m == I_DIAG ? 2 : 1);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rank(a) < rank(b); });
    for (int q : order) {
      dec_op_s &op = *ops[q];
      if (op.rows == 0) continue;
      const T *x = static_cast<const T *>(op.swap ? ys[q] : xs[q]);
      const T *y = static_cast<const T *>(op.swap ? xs[q] : ys[q]);
      const uintptr_t A = 2 * sizeof(T);
      GroupOp &o = g.op[k];
      o.impl = op.impl;
      o.rows = static_cast<int>(op.rows);
      o.vec = (reinterpret_cast<uintptr_t>(outs[q]) % A == 0) && (!y || reinterpret_cast<uintptr_t>(y) % A == 0) &&
              (!x || reinterpret_cast<uintptr_t>(x) % A == 0);
      o.swap_wdd = op.impl == I_WDD && op.m != 0;
      o.i0 = op.i0.p;
      o.i1 = op.i1.p;
      o.i2 = op.i2.p;
      o.s8 = op.s8.p;
      o.c0 = op.c0.p;
      o.c1 = op.c1.p;
      o.c2 = op.c2.p;
      o.ix = op.ix.p;
      o.cx = op.cx.p;
      o.m16 = op.m16.p;
      o.sched = op.sched.p;
      o.x = x;
      o.y = y;
      o.out = outs[q];
      if (op.impl == I_SELL) {
        o.width = op.sell.width;
        o.ecol = op.sell.col.p;
        o.eval = op.sell.val.p;
      } else if (op.impl == I_ELL || op.impl == I_WPD) {
        if constexpr (sizeof(T) == 8) {
          o.width = op.ell64.width;
          o.ecol = op.ell64.col.p;
          o.eval = op.ell64.val.p;
        } else {
          o.width = op.ell32.width;
          o.ecol = op.ell32.col.p;
          o.eval = op.ell32.val.p;
        }
      }
      g.start[k] = blocks;
      blocks += group_blocks(op);
      k++;
    }
    if (k == 0) continue;
    g.count = k;
    g.start[k] = blocks;
    hipLaunchKernelGGL(k_apply_group<T>, dim3(blocks), dim3(256), 0, s, g);
  }
  dec::check_launch("dec_op_apply_many");
}

void unsupported(int m, int n) {
  fail(DEC_ERR_UNSUPPORTED, "Unsupported combination of degrees " + std::to_string(m) + " and " + std::to_string(n) +
                                ". Ensure that their sum is not greater than the degree of the complex.");
}

template <class T>
void make_diag(dec_op_s &op, const DBuf<double> &h, bool inverse, double mult) {
  if (inverse) {
    op.c0.alloc(sizeof(T) * h.n);
    if (h.n)
      hipLaunchKernelGGL(k_inv_diag<T>, dim3(dec::nblocks(static_cast<long long>(h.n))), dim3(256), 0, 0,
                         static_cast<int>(h.n), h.p, mult, reinterpret_cast<T *>(op.c0.p));
  } else {
    launch_cast<T>(h, op.c0);
  }
}

// wedge_pd_01_mat (FastDEC.jl:329-346): for edge e = (α, β) = (∂v0, ∂v1) and
// each elementary dual edge of e (B-block dual edges, ascending) with weight
// w = dual_length / Σ dual_length, and l the third vertex of its triangle:
// M[e,α] += w·5/12, M[e,β] += w·5/12, M[e,l] += w·2/12.  Rows in ascending
// column order, repeated (row, column) entries summed in that input order.
void pd01_matrix(const dec_mesh_s &m, dec::HostCsr &M) {
  const dec::Topo &t = m.h;
  int nv = t.nv, ne = t.ne, nt = t.nt;
  dec::Csr eb;
  dec::edge_bduals(t, eb);
  std::vector<double> dlB(3 * static_cast<size_t>(nt));
  if (nt) dec::d2h(dlB.data(), m.dl.p + 2 * static_cast<size_t>(ne), sizeof(double) * dlB.size());
  M.m = ne;
  M.n = nv;
  M.ptr.assign(ne + 1, 0);
  M.col.clear();
  M.val.clear();
  std::vector<std::pair<int, double>> ent;
  for (int e = 0; e < ne; e++) {
    int a = t.v0[e], b = t.v1[e];
    double tot = 0.0;
    for (int k = eb.ptr[e]; k < eb.ptr[e + 1]; k++) tot += dlB[eb.idx[k]];
    ent.clear();
    for (int k = eb.ptr[e]; k < eb.ptr[e + 1]; k++) {
      int tri = eb.idx[k] % nt;
      int tv[3] = {t.v1[t.e1[tri]], t.v0[t.e2[tri]], t.v0[t.e1[tri]]};
      int l = -1;
      for (int q = 0; q < 3; q++)
        if (tv[q] != a && tv[q] != b) l = tv[q];
      if (l < 0) fail(DEC_ERR_ARG, "wedge_pd_01_mat: degenerate triangle");
      double w = dlB[eb.idx[k]] / tot;
      ent.push_back({a, w * 5 / 12});
      ent.push_back({b, w * 5 / 12});
      ent.push_back({l, w * 2 / 12});
    }
    // stable sort by column, then sum runs in input order
    std::stable_sort(ent.begin(), ent.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (size_t i = 0; i < ent.size();) {
      int c = ent[i].first;
      double v = ent[i].second;
      size_t j = i + 1;
      for (; j < ent.size() && ent[j].first == c; j++) v = v + ent[j].second;
      M.col.push_back(c);
      M.val.push_back(v);
      i = j;
    }
    M.ptr[e + 1] = static_cast<int32_t>(M.col.size());
  }
}

// ♭♯_mat = only.(♭_mat(DPPFlat) * ♯_mat(LLSDDSharp)) (DiscreteExteriorCalculus.jl:684-714,
// 876-909, 1992).  Per triangle t with edges e_j (slot j) and B-block dual
// edges b_j from the triangle center to the edge centers:
//   ♯[t, e_j] = pinv(X)[:, j] · w_j,  X rows = sign(b_j) (center_t - center_{e_j}),
//   ♭[e, t]   = w_e,t · (p[∂v0 e] - p[∂v1 e]) sign(e),
// with w_e,t = |b_{e,t}| / Σ_{t' ∋ e} |b_{e,t'}|.  pinv(X) is taken in the
// plane of the triangle (X has rank 2: its rows are tangent), which is the
// reference's SVD pseudo-inverse on flat meshes and stays well defined where
// rounding lifts the rows out of the plane on a curved surface.
// p2_d2_interpolation (DiscreteExteriorCalculus.jl:2016-2031), nv x nt: for each
// triangle t ascending and each of its six dual triangles k·nt + t in order,
// mat[v, t] += dual_area / area with v the dual triangle's primal vertex
// (slots v0, v0, v1, v1, v2, v2 of triangle_vertices).  transposed = 1 gives
// p0_d0_interpolation (FastDEC.jl:928-941), nt x nv, the same sums by rows of t.
void p2_d2_matrix(const dec_mesh_s &m, bool transposed, dec::HostCsr &M) {
  const dec::Topo &t = m.h;
  int nv = t.nv, nt = t.nt;
  std::vector<double> da = m.da.download(), area = m.area.download();
  const int sv[6] = {0, 0, 1, 1, 2, 2};
  // per triangle: three (vertex, weight) pairs, each the sum of two dual triangles
  std::vector<int32_t> tv(3 * static_cast<size_t>(nt));
  std::vector<double> w(3 * static_cast<size_t>(nt));
  for (int i = 0; i < nt; i++) {
    int vs[3] = {t.v1[t.e1[i]], t.v0[t.e2[i]], t.v0[t.e1[i]]};
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 6; k++) acc[sv[k]] = acc[sv[k]] + da[static_cast<size_t>(k) * nt + i] / area[i];
    for (int j = 0; j < 3; j++) {
      tv[3 * static_cast<size_t>(i) + j] = vs[j];
      w[3 * static_cast<size_t>(i) + j] = acc[j];
    }
  }
  M = dec::HostCsr();
  if (transposed) {
    M.m = nt;
    M.n = nv;
    M.ptr.assign(nt + 1, 0);
    for (int i = 0; i < nt; i++) {
      std::pair<int, double> e[3];
      for (int j = 0; j < 3; j++) e[j] = {tv[3 * static_cast<size_t>(i) + j], w[3 * static_cast<size_t>(i) + j]};
      std::sort(e, e + 3, [](const auto &a, const auto &b) { return a.first < b.first; });
      for (int j = 0; j < 3; j++) {
        if (j && e[j].first == e[j - 1].first) fail(DEC_ERR_UNSUPPORTED, "interpolation: triangle with a repeated vertex");
        M.col.push_back(e[j].first);
        M.val.push_back(e[j].second);
      }
      M.ptr[i + 1] = static_cast<int32_t>(M.col.size());
    }
    return;
  }
  M.m = nv;
  M.n = nt;
  M.ptr.assign(nv + 1, 0);
  for (size_t q = 0; q < tv.size(); q++) M.ptr[tv[q] + 1]++;
  for (int v = 0; v < nv; v++) M.ptr[v + 1] += M.ptr[v];
  M.col.assign(M.ptr[nv], 0);
  M.val.assign(M.ptr[nv], 0.0);
  std::vector<int32_t> pos(M.ptr.begin(), M.ptr.end() - 1);
  for (int i = 0; i < nt; i++)  // ascending t => each row sorted
    for (int j = 0; j < 3; j++) {
      int v = tv[3 * static_cast<size_t>(i) + j];
      M.col[pos[v]] = i;
      M.val[pos[v]++] = w[3 * static_cast<size_t>(i) + j];
    }
}

// p3_d3_interpolation (DiscreteExteriorCalculus.jl:2038-2053), nv x ntet: for
// each tetrahedron and its 24 dual tetrahedra in order, mat[v, x] += dual_vol /
// vol, v the dual tetrahedron's primal vertex (six consecutive per vertex).
void p3_d3_matrix(dec_mesh_s &m, dec::HostCsr &M) {
  const dec::Topo &t = m.h;
  int nv = t.nv, n = m.ht.n;
  std::vector<double> dvol, vol = m.vol.download();
  dec::tet_dual_volumes(m, dvol);
  M = dec::HostCsr();
  M.m = nv;
  M.n = n;
  M.ptr.assign(nv + 1, 0);
  std::vector<int32_t> tvs(4 * static_cast<size_t>(n));
  for (int x = 0; x < n; x++) {
    dec::tet_vertices(t, m.ht, x, &tvs[4 * static_cast<size_t>(x)]);
    for (int i = 0; i < 4; i++) M.ptr[tvs[4 * static_cast<size_t>(x) + i] + 1]++;
  }
  for (int v = 0; v < nv; v++) M.ptr[v + 1] += M.ptr[v];
  M.col.assign(M.ptr[nv], 0);
  M.val.assign(M.ptr[nv], 0.0);
  std::vector<int32_t> pos(M.ptr.begin(), M.ptr.end() - 1);
  for (int x = 0; x < n; x++)  // ascending x => each row sorted
    for (int i = 0; i < 4; i++) {
      double acc = 0.0;
      for (int q = 6 * i; q < 6 * i + 6; q++) acc = acc + dvol[24 * static_cast<size_t>(x) + q] / vol[x];
      int v = tvs[4 * static_cast<size_t>(x) + i];
      M.col[pos[v]] = x;
      M.val[pos[v]++] = acc;
    }
}

// SparseMatrixCSC(⋆0⁻¹) * P * SparseMatrixCSC(⋆k⁻¹): entry (v, c) = (ih0[v] P[v, c]) ihk[c]
// (d0_p0_interpolation, FastDEC.jl:915-926).
void scale_rows_cols(dec::HostCsr &P, const std::vector<double> &rs, const std::vector<double> &cs) {
  for (int r = 0; r < P.m; r++)
    for (int k = P.ptr[r]; k < P.ptr[r + 1]; k++) P.val[k] = (rs[r] * P.val[k]) * cs[P.col[k]];
}

std::vector<double> inv_diag(const DBuf<double> &h) {
  std::vector<double> v = h.download();
  for (auto &x : v) x = (1 / x) * 1.0;  // dec_inv_hodge_star with (-1)^(k(n-k)) = 1 for these k
  return v;
}

void flat_sharp_matrix(const dec_mesh_s &m, dec::HostCsr &M) {
  const dec::Topo &t = m.h;
  int nv = t.nv, ne = t.ne, nt = t.nt;
  std::vector<double> dpt = m.dpt.download();
  std::vector<double> dlB(3 * static_cast<size_t>(nt));
  if (nt) dec::d2h(dlB.data(), m.dl.p + 2 * static_cast<size_t>(ne), sizeof(double) * dlB.size());
  dec::Csr eb;
  dec::edge_bduals(t, eb);
  std::vector<double> tot(ne, 0.0);
  for (int e = 0; e < ne; e++)
    for (int k = eb.ptr[e]; k < eb.ptr[e + 1]; k++) tot[e] = tot[e] + dlB[eb.idx[k]];
  const std::vector<int32_t> *E[3] = {&t.e0, &t.e1, &t.e2};
  const double *P = m.hp.data();
  struct Ent {
    int row, col, tri;
    double val;
  };
  std::vector<Ent> ents;
  ents.reserve(9 * static_cast<size_t>(nt));
  for (int tr = 0; tr < nt; tr++) {
    const double *c = &dpt[3 * static_cast<size_t>(nv + ne + tr)];
    double X[3][3], fl[3][3], sh[3][3];
    int ed[3];
    for (int j = 0; j < 3; j++) {
      int e = (*E[j])[tr];
      ed[j] = e;
      const double *ce = &dpt[3 * static_cast<size_t>(nv + e)];
      // D_edge_orientation of b_j: relative_sign(eo[e_j], j odd ? !to[t] : to[t]) (DiscreteExteriorCalculus.jl:595-599)
      bool tj = (j % 2) ? !t.to[tr] : static_cast<bool>(t.to[tr]);
      double sg = (static_cast<bool>(t.eo[e]) == tj) ? 1.0 : -1.0;
      for (int k = 0; k < 3; k++) X[j][k] = sg * (c[k] - ce[k]);
    }
    // tangent basis of the triangle (triangle_vertices, SimplicialSets.jl:305-309)
    int tv[3] = {t.v1[t.e1[tr]], t.v0[t.e2[tr]], t.v0[t.e1[tr]]};
    double a[3], b[3], n[3], u[3], w[3];
    for (int k = 0; k < 3; k++) {
      a[k] = P[3 * tv[1] + k] - P[3 * tv[0] + k];
      b[k] = P[3 * tv[2] + k] - P[3 * tv[0] + k];
    }
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int k = 0; k < 3; k++) u[k] = a[k] / la;
    w[0] = n[1] * u[2] - n[2] * u[1];
    w[1] = n[2] * u[0] - n[0] * u[2];
    w[2] = n[0] * u[1] - n[1] * u[0];
    double lw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (!(la > 0.0) || !(lw > 0.0)) fail(DEC_ERR_ARG, "♭♯_mat: degenerate triangle");
    for (int k = 0; k < 3; k++) w[k] /= lw;
    double Y[3][2], g00 = 0, g01 = 0, g11 = 0;
    for (int j = 0; j < 3; j++) {
      Y[j][0] = X[j][0] * u[0] + X[j][1] * u[1] + X[j][2] * u[2];
      Y[j][1] = X[j][0] * w[0] + X[j][1] * w[1] + X[j][2] * w[2];
      g00 += Y[j][0] * Y[j][0];
      g01 += Y[j][0] * Y[j][1];
      g11 += Y[j][1] * Y[j][1];
    }
    double det = g00 * g11 - g01 * g01;
    if (!(det > 0.0)) fail(DEC_ERR_ARG, "♭♯_mat: dual edges of a triangle are collinear");
    for (int j = 0; j < 3; j++) {
      // column j of pinv(X) = [u w] G⁻¹ Y_jᵀ, times the weight of b_j
      double q0 = (g11 * Y[j][0] - g01 * Y[j][1]) / det;
      double q1 = (g00 * Y[j][1] - g01 * Y[j][0]) / det;
      double wt = dlB[static_cast<size_t>(j) * nt + tr] / tot[ed[j]];
      for (int k = 0; k < 3; k++) sh[j][k] = (u[k] * q0 + w[k] * q1) * wt;
      // ♭[e_j, t] = del * e_vec / dels_sum (DiscreteExteriorCalculus.jl:705-710)
      int e = ed[j];
      double sgn = t.eo[e] ? 1.0 : -1.0;
      for (int k = 0; k < 3; k++) {
        double ev = (P[3 * t.v0[e] + k] - P[3 * t.v1[e] + k]) * sgn;
        fl[j][k] = dlB[static_cast<size_t>(j) * nt + tr] * ev / tot[e];
      }
    }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        ents.push_back({ed[i], ed[j], tr, fl[i][0] * sh[j][0] + fl[i][1] * sh[j][1] + fl[i][2] * sh[j][2]});
  }
  // entries summed over triangles in ascending order (the SpGEMM order)
  std::sort(ents.begin(), ents.end(), [](const Ent &x, const Ent &y) {
    return x.row != y.row ? x.row < y.row : (x.col != y.col ? x.col < y.col : x.tri < y.tri);
  });
  M.m = ne;
  M.n = ne;
  M.ptr.assign(ne + 1, 0);
  M.col.clear();
  M.val.clear();
  for (size_t i = 0; i < ents.size();) {
    size_t j = i + 1;
    double v = ents[i].val;
    for (; j < ents.size() && ents[j].row == ents[i].row && ents[j].col == ents[i].col; j++) v = v + ents[j].val;
    M.col.push_back(ents[i].col);
    M.val.push_back(v);
    M.ptr[ents[i].row + 1]++;
    i = j;
  }
  for (int r = 0; r < ne; r++) M.ptr[r + 1] += M.ptr[r];
}

// A sparse-matrix operator from a host CSR (products assembled at create).
std::unique_ptr<dec_op_s> ell_op(dec::HostCsr &&A, int dtype) {
  auto o = std::make_unique<dec_op_s>();
  o->impl = I_ELL;
  o->dtype = dtype;
  o->rows = A.m;
  o->cols = A.n;
  o->csr = std::move(A);
  if (dtype == DEC_F64) dec::upload_ell(o->csr, o->ell64);
  else dec::upload_ell(o->csr, o->ell32);
  return o;
}

// ±1 matrix of d2 (FastDEC.jl:527-548) by rows (tetrahedra, triangles
// ascending) or, transposed, by triangles (tetrahedra ascending).
void csr_d2(const dec::Topo &t, const dec::Tets &T, bool transposed, dec::Csr &c) {
  int n = T.n, nt = t.nt;
  auto sgn = [&](int x, int k) {
    int st = t.to[T.t[k][x]] == 1 ? 1 : -1, sx = T.o[x] == 1 ? 1 : -1;
    return static_cast<int8_t>((k % 2 ? -1 : 1) * st * sx);
  };
  if (!transposed) {
    c.n = n;
    c.ptr.assign(n + 1, 0);
    c.idx.clear();
    c.val.clear();
    for (int x = 0; x < n; x++) {
      std::pair<int, int8_t> ent[4];
      for (int k = 0; k < 4; k++) ent[k] = {T.t[k][x], sgn(x, k)};
      std::sort(ent, ent + 4, [](const auto &a, const auto &b) { return a.first < b.first; });
      for (int k = 0; k < 4; k++) {
        if (k && ent[k].first == ent[k - 1].first) fail(DEC_ERR_UNSUPPORTED, "d2: tetrahedron with a repeated triangle");
        c.idx.push_back(ent[k].first);
        c.val.push_back(ent[k].second);
      }
      c.ptr[x + 1] = static_cast<int32_t>(c.idx.size());
    }
    return;
  }
  c.n = nt;
  c.ptr.assign(nt + 1, 0);
  for (int x = 0; x < n; x++)
    for (int k = 0; k < 4; k++) c.ptr[T.t[k][x] + 1]++;
  for (int f = 0; f < nt; f++) c.ptr[f + 1] += c.ptr[f];
  c.idx.assign(c.ptr[nt], 0);
  c.val.assign(c.ptr[nt], 0);
  std::vector<int32_t> w(c.ptr.begin(), c.ptr.end() - 1);
  for (int x = 0; x < n; x++)  // ascending x => each row sorted
    for (int k = 0; k < 4; k++) {
      int f = T.t[k][x];
      c.idx[w[f]] = x;
      c.val[w[f]++] = sgn(x, k);
    }
}

// The operators of a 3-D complex that differ from those of a 2-D one.  Returns
// null for the operators the triangle skeleton serves unchanged (wedges 0-1,
// 0-2, 1-1 with the 2-D dual triangles of each triangle; d0, d1; ∂1, ∂2),
// which create() then builds exactly as on a 2-D complex (FastDEC.jl methods
// on HasDeltaSet2D apply to a HasDeltaSet3D).
std::unique_ptr<dec_op_s> create3(dec_mesh_s &m, int kind, int dm, int dn, int hodge, int dtype) {
  const dec::Topo &t = m.h;
  const dec::Tets &T = m.ht;
  int nv = t.nv, ne = t.ne, nt = t.nt, n = T.n;
  const size_t N = static_cast<size_t>(n);
  bool f64 = dtype == DEC_F64;
  auto op = std::make_unique<dec_op_s>();
  op->kind = kind;
  op->m = dm;
  op->n = dn;
  op->hodge = hodge;
  op->dtype = dtype;
  auto sell = [&](dec::Csr &&c, int ncols, bool neg) {
    op->impl = I_SELL;
    op->rows = c.n;
    op->cols = ncols;
    to_sign_ell(c, op->sell, neg);
    csr_from_signs(c, ncols, neg, op->csr);
  };
  auto weights = [&](int div3) {
    op->cx.alloc((f64 ? 8 : 4) * 4 * std::max<size_t>(N, 1));
    if (n) {
      if (f64)
        hipLaunchKernelGGL(k_coeff_tet<double>, dim3(dec::nblocks(n)), dim3(256), 0, 0, n, m.tw.p, m.vol.p, div3,
                           reinterpret_cast<double *>(op->cx.p));
      else
        hipLaunchKernelGGL(k_coeff_tet<float>, dim3(dec::nblocks(n)), dim3(256), 0, 0, n, m.tw.p, m.vol.p, div3,
                           reinterpret_cast<float *>(op->cx.p));
    }
  };
  switch (kind) {
    case DEC_OP_WEDGE: {
      int lo = std::min(dm, dn), hi = std::max(dm, dn);
      if (lo < 0 || lo + hi > 3) unsupported(dm, dn);
      if (!((lo == 0 && hi == 3) || (lo == 1 && hi == 2))) return nullptr;
      op->rows = n;
      std::vector<int32_t> ix;
      if (lo == 0) {
        // dec_wedge_product(0,3) / (3,0): the (0,3) kernel on (α, β) / (β, α) (FastDEC.jl:266-274)
        op->impl = I_W03;
        op->swap = dm > dn;
        op->cols = dm == 0 ? nv : n;
        op->cols2 = dm == 0 ? n : nv;
        ix.resize(4 * N);
        for (int x = 0; x < n; x++) {
          int v[4];
          dec::tet_vertices(t, T, x, v);
          for (int i = 0; i < 4; i++) ix[i * N + x] = v[i];
        }
        op->ix.upload(ix);
        weights(0);
      } else {
        // dec_wedge_product(2,1) on (α, β); (1,2) runs it on (β, α) (FastDEC.jl:281-289)
        op->impl = I_W21;
        op->swap = dm < dn;
        op->cols = dm == 2 ? nt : ne;
        op->cols2 = dm == 2 ? ne : nt;
        ix.resize(10 * N);
        std::vector<uint16_t> sg(N);
        // (edge slot, triangle slot) of coefficients 1..12 (FastDEC.jl:117-130)
        const int pe[12] = {3, 4, 5, 1, 2, 5, 0, 2, 4, 0, 1, 3};
        const int pt[12] = {3, 2, 1, 3, 2, 0, 3, 1, 0, 2, 1, 0};
        for (int x = 0; x < n; x++) {
          int e[6];
          dec::tet_edges(t, T, x, e);
          for (int k = 0; k < 6; k++) ix[k * N + x] = e[k];
          for (int k = 0; k < 4; k++) ix[(6 + k) * N + x] = T.t[k][x];
          int sx = T.o[x] == 1 ? 1 : -1;
          uint16_t bits = 0;
          for (int k = 0; k < 12; k++) {
            int s = sx * (t.eo[e[pe[k]]] == 1 ? 1 : -1) * (t.to[T.t[pt[k]][x]] == 1 ? 1 : -1);
            if (s < 0) bits = static_cast<uint16_t>(bits | (1u << k));
          }
          sg[x] = bits;
        }
        op->ix.upload(ix);
        op->m16.upload(sg);
        weights(1);
      }
      break;
    }
    case DEC_OP_DIFFERENTIAL:
      if (dm == 0 || dm == 1) return nullptr;
      if (dm != 2) fail(DEC_ERR_UNSUPPORTED, "dec_differential: n must be 0, 1 or 2 on a 3-D complex");
      {
        dec::Csr c;
        csr_d2(t, T, false, c);
        sell(std::move(c), nt, false);
      }
      break;
    case DEC_OP_BOUNDARY:
      if (dm == 1 || dm == 2) return nullptr;
      if (dm != 3) fail(DEC_ERR_UNSUPPORTED, "dec_boundary: n must be 1, 2 or 3 on a 3-D complex");
      {
        dec::Csr c;
        csr_d2(t, T, true, c);
        sell(std::move(c), n, false);
      }
      break;
    case DEC_OP_DUAL_DERIVATIVE: {
      // d̃0 = -∂3, d̃1 = ∂2, d̃2 = -∂1 (FastDEC.jl:455-462)
      dec::Csr c;
      if (dm == 0) {
        csr_d2(t, T, true, c);
        sell(std::move(c), n, true);
      } else if (dm == 1) {
        dec::edge_triangles(t, c);
        sell(std::move(c), nt, false);
      } else if (dm == 2) {
        dec::vertex_edges(t, c);
        sell(std::move(c), ne, true);
      } else {
        fail(DEC_ERR_UNSUPPORTED, "dec_dual_derivative: n must be 0, 1 or 2 on a 3-D complex");
      }
      break;
    }
    case DEC_OP_HODGE:
    case DEC_OP_INV_HODGE: {
      if (dm < 0 || dm > 3) fail(DEC_ERR_UNSUPPORTED, "dec_hodge_star: n must be 0, 1, 2 or 3 on a 3-D complex");
      bool inv = kind == DEC_OP_INV_HODGE;
      if (hodge == DEC_GEOMETRIC_HODGE) {
        // dec_hodge_star has no geometric method on a 3-D complex; the inverse
        // falls back to the diagonal one for n = 0, 3 (FastDEC.jl:750-757)
        if (!inv)
          fail(DEC_ERR_UNSUPPORTED,
               "dec_hodge_star: the geometric Hodge star is not defined on a 3-D complex; use the diagonal one");
        if (dm == 1 || dm == 2)
          fail(DEC_ERR_UNSUPPORTED,
               "The Geometric Hodge star in 3D for 1-forms and 2-forms has not yet been implemented. Please use the "
               "Diagonal Hodge star instead.");
      }
      op->impl = I_DIAG;
      const DBuf<double> &h = dm == 0 ? m.h0 : (dm == 1 ? m.h1 : (dm == 2 ? m.h2 : m.h3));
      op->rows = op->cols = static_cast<int64_t>(h.n);
      // (-1)^(k(3-k)) = 1 for every k (FastDEC.jl:721-727)
      if (f64) make_diag<double>(*op, h, inv, 1.0);
      else make_diag<float>(*op, h, inv, 1.0);
      std::vector<double> hv = h.download();
      if (inv)
        for (auto &x : hv) x = (1 / x) * 1.0;
      csr_diag(hv, op->csr);
      break;
    }
    case DEC_OP_LAPLACIAN0: {
      // ∇²(0, s) = -⋆0⁻¹ d̃2 ⋆1 d0 (DiscreteExteriorCalculus.jl:1836-1890); d̃2 = -d0ᵀ as d̃1 in 2-D
      if (hodge != DEC_DIAGONAL_HODGE)
        fail(DEC_ERR_UNSUPPORTED, "∇²: a 3-D complex has only the diagonal Hodge star; pass DEC_DIAGONAL_HODGE");
      op->impl = I_ELL;
      op->rows = op->cols = nv;
      std::vector<double> h0 = m.h0.download(), h1 = m.h1.download(), ih0(nv);
      for (int v = 0; v < nv; v++) ih0[v] = 1 / h0[v];
      dec::assemble_laplacian0(t, ih0, h1, nullptr, op->csr);
      if (f64) dec::upload_ell(op->csr, op->ell64);
      else dec::upload_ell(op->csr, op->ell32);
      break;
    }
    case DEC_OP_DUAL_LAPLACIAN: {
      // Δᵈ(Val(0)) on a 3-D complex: m = ⋆3 d2 ⋆2⁻¹ d̃0 as one matrix (FastDEC.jl:832-839)
      if (dm != 0) fail(DEC_ERR_UNSUPPORTED, "Δᵈ: n must be 0 on a 3-D complex");
      std::vector<double> h3 = m.h3.download(), h2 = m.h2.download();
      dec::Csr rows, cols;
      csr_d2(t, T, false, rows);
      csr_d2(t, T, true, cols);
      dec::HostCsr A, B, M;
      A.m = n;
      A.n = nt;
      A.ptr = rows.ptr;
      A.col = rows.idx;
      A.val.resize(rows.idx.size());
      for (int x = 0; x < n; x++)
        for (int k = rows.ptr[x]; k < rows.ptr[x + 1]; k++)
          A.val[k] = (h3[x] * rows.val[k]) * ((1 / h2[rows.idx[k]]) * 1.0);
      csr_from_signs(cols, n, true, B);
      dec::spgemm(A, B, M);
      auto e = ell_op(std::move(M), dtype);
      e->kind = kind;
      e->m = dm;
      e->hodge = hodge;
      return e;
    }
    case DEC_OP_P0_D0_INTERPOLATION:
      return nullptr;  // the triangle skeleton's p0_d0 (FastDEC.jl:928 takes any HasDeltaSet2D)
    case DEC_OP_P3_D3_INTERPOLATION:
    case DEC_OP_D0_P0_INTERPOLATION: {
      // p3_d3_interpolation; d0_p0_interpolation = ⋆0⁻¹ p3_d3 ⋆3⁻¹ (FastDEC.jl:924-926)
      dec::HostCsr P;
      p3_d3_matrix(m, P);
      if (kind == DEC_OP_D0_P0_INTERPOLATION) scale_rows_cols(P, inv_diag(m.h0), inv_diag(m.h3));
      auto e = ell_op(std::move(P), dtype);
      e->kind = kind;
      e->hodge = hodge;
      return e;
    }
    default:
      fail(DEC_ERR_UNSUPPORTED, "dec_op_create: operator kind " + std::to_string(kind) +
                                    " is defined on 2-D complexes only");
  }
  return op;
}

// Spatial tile schedule: the rows are cut into tiles of `per` rows (one
// workgroup each), the tiles are sorted by the Morton code of their mean row
// position in the mesh's bounding box, and the sorted list is cut into 8
// contiguous groups, group x going to XCD x: sched[8 j + x] is the j-th tile
// of group x (-1 past its end).  Neighbouring rows, whose gathers overlap,
// then run on the same XCD and hit its L2.
void make_schedule(dec_op_s &op, const std::vector<double> &pos, int per) {
  long long rows = op.rows;
  int ntile = static_cast<int>((rows + per - 1) / per);
  if (ntile < 16) return;
  std::vector<double> c(3 * static_cast<size_t>(ntile), 0.0);
  for (long long r = 0; r < rows; r++)
    for (int k = 0; k < 3; k++) c[3 * (r / per) + k] += pos[3 * r + k];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int t = 0; t < ntile; t++) {
    long long cnt = std::min<long long>(per, rows - static_cast<long long>(t) * per);
    for (int k = 0; k < 3; k++) {
      c[3 * t + k] /= static_cast<double>(cnt);
      lo[k] = std::min(lo[k], c[3 * t + k]);
      hi[k] = std::max(hi[k], c[3 * t + k]);
    }
  }
  auto spread = [](uint64_t v) {
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x30000ff;
    v = (v | (v << 8)) & 0x300f00f;
    v = (v | (v << 4)) & 0x30c30c3;
    v = (v | (v << 2)) & 0x9249249;
    return v;
  };
  std::vector<std::pair<uint64_t, int>> key(ntile);
  for (int t = 0; t < ntile; t++) {
    uint64_t code = 0;
    for (int k = 0; k < 3; k++) {
      double span = hi[k] - lo[k];
      uint64_t q = span > 0 ? static_cast<uint64_t>((c[3 * t + k] - lo[k]) / span * 1023.0) : 0;
      code |= spread(q) << k;
    }
    key[t] = {code, t};
  }
  std::stable_sort(key.begin(), key.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  int per_xcd = (ntile + 7) / 8;
  std::vector<int32_t> sched(8 * static_cast<size_t>(per_xcd), -1);
  for (int x = 0; x < 8; x++) {
    // each XCD walks its tiles in row order (tables stream in address order)
    int i0 = std::min(ntile, x * per_xcd), i1 = std::min(ntile, (x + 1) * per_xcd);
    std::vector<int> mine;
    for (int i = i0; i < i1; i++) mine.push_back(key[i].second);
    std::sort(mine.begin(), mine.end());
    for (size_t j = 0; j < mine.size(); j++) sched[8 * j + x] = mine[j];
  }
  op.sched.upload(sched);
  op.nsched = static_cast<int>(sched.size());
}

// Schedules for the gather operators of a complex: rows at edge midpoints
// (wedge 0-1, d0), triangle centroids (wedges 0-2, 1-1, d1) or vertices
// (the 0-form Laplacian, rows of at most 8 slots).
void schedule_gathers(dec_op_s &op, const dec_mesh_s &m) {
  const dec::Topo &t = m.h;
  const std::vector<double> &p = m.hp;
  if (p.size() < 3 * static_cast<size_t>(t.nv)) return;
  std::vector<double> pos;
  int per = 512;
  switch (op.impl) {
    case I_W01:
    case I_D0:
      if (op.rows != t.ne) return;
      pos.resize(3 * static_cast<size_t>(t.ne));
      for (int e = 0; e < t.ne; e++)
        for (int k = 0; k < 3; k++) pos[3 * e + k] = 0.5 * p[3 * t.v0[e] + k] + 0.5 * p[3 * t.v1[e] + k];
      break;
    case I_W02:
    case I_W11:
    case I_D1:
      if (op.rows != t.nt) return;
      pos.resize(3 * static_cast<size_t>(t.nt));
      for (int i = 0; i < t.nt; i++) {
        int a = t.v1[t.e1[i]], b = t.v0[t.e2[i]], c = t.v0[t.e1[i]];
        for (int k = 0; k < 3; k++) pos[3 * i + k] = (p[3 * a + k] + p[3 * b + k] + p[3 * c + k]) / 3.0;
      }
      break;
    case I_ELL: {
      int width = op.dtype == DEC_F64 ? op.ell64.width : op.ell32.width;
      if (op.kind != DEC_OP_LAPLACIAN0 || op.rows != t.nv || width > 8) return;
      pos.assign(p.begin(), p.begin() + 3 * static_cast<size_t>(t.nv));
      per = 256;
      break;
    }
    default: return;
  }
  make_schedule(op, pos, per);
}

dec_op_s *create(dec_mesh_s &m, int kind, int dm, int dn, int hodge, int dtype) {
  if (dtype != DEC_F64 && dtype != DEC_F32) fail(DEC_ERR_ARG, "dec_op_create: dtype must be DEC_F64 or DEC_F32");
  if (hodge != DEC_GEOMETRIC_HODGE && hodge != DEC_DIAGONAL_HODGE) fail(DEC_ERR_ARG, "dec_op_create: unknown hodge");
  dec::ensure_dual(m);
  if (m.dim == 3) {
    if (auto o3 = create3(m, kind, dm, dn, hodge, dtype)) {
      dec::check_launch("dec_op_create");
      dec::dsync();
      return o3.release();
    }
  }
  const dec::Topo &t = m.h;
  int nv = t.nv, ne = t.ne, nt = t.nt;
  bool f64 = dtype == DEC_F64;
  auto op = std::make_unique<dec_op_s>();
  op->kind = kind;
  op->m = dm;
  op->n = dn;
  op->hodge = hodge;
  op->dtype = dtype;
  const size_t es = f64 ? sizeof(double) : sizeof(float);
  auto add_tmp = [&](int64_t n) {
    op->tmp.emplace_back();
    op->tmp.back().alloc(es * static_cast<size_t>(std::max<int64_t>(n, 1)));
  };
  switch (kind) {
    case DEC_OP_WEDGE: {
      int lo = std::min(dm, dn), hi = std::max(dm, dn);
      op->swap = dm > dn;  // (k,0) evaluates the (0,k) kernel on (β, α) (FastDEC.jl:266-269)
      if (lo == 0 && hi == 1) {
        op->impl = I_W01;
        op->rows = ne;
        op->i0.upload(t.v0);
        op->i1.upload(t.v1);
        op->cols = dm == 0 ? nv : ne;
        op->cols2 = dm == 0 ? ne : nv;
      } else if (lo == 0 && hi == 2) {
        op->impl = I_W02;
        op->rows = nt;
        // slots (v0,v0,v1,v1,v2,v2) = ∂v_ev[∂e_e] for schema (v,e,ev) (FastDEC.jl:56)
        std::vector<int32_t> a(nt), b(nt), c(nt);
        for (int i = 0; i < nt; i++) {
          int s0 = t.v1[t.e1[i]], s1 = t.v1[t.e2[i]], s2 = t.v0[t.e2[i]], s3 = t.v1[t.e0[i]],
              s4 = t.v0[t.e0[i]], s5 = t.v0[t.e1[i]];
          if (s0 != s1 || s2 != s3 || s4 != s5)
            fail(DEC_ERR_UNSUPPORTED, "wedge 0-2: triangle violates the simplicial identities");
          a[i] = s0;
          b[i] = s2;
          c[i] = s4;
        }
        op->i0.upload(a);
        op->i1.upload(b);
        op->i2.upload(c);
        size_t bytes = (f64 ? 8 : 4) * static_cast<size_t>(nt);
        op->c0.alloc(bytes);
        op->c1.alloc(bytes);
        op->c2.alloc(bytes);
        if (nt) {
          if (f64)
            hipLaunchKernelGGL(k_coeff02<double>, dim3(dec::nblocks(nt)), dim3(256), 0, 0, nt, m.da.p, m.area.p,
                               reinterpret_cast<double *>(op->c0.p), reinterpret_cast<double *>(op->c1.p),
                               reinterpret_cast<double *>(op->c2.p));
          else
            hipLaunchKernelGGL(k_coeff02<float>, dim3(dec::nblocks(nt)), dim3(256), 0, 0, nt, m.da.p, m.area.p,
                               reinterpret_cast<float *>(op->c0.p), reinterpret_cast<float *>(op->c1.p),
                               reinterpret_cast<float *>(op->c2.p));
        }
        op->cols = dm == 0 ? nv : nt;
        op->cols2 = dm == 0 ? nt : nv;
      } else if (lo == 1 && hi == 1) {
        op->impl = I_W11;
        op->rows = nt;
        op->i0.upload(t.e0);
        op->i1.upload(t.e1);
        op->i2.upload(t.e2);
        size_t bytes = (f64 ? 8 : 4) * static_cast<size_t>(nt);
        op->c0.alloc(bytes);
        op->c1.alloc(bytes);
        op->c2.alloc(bytes);
        if (nt) {
          if (f64)
            hipLaunchKernelGGL(k_coeff11<double>, dim3(dec::nblocks(nt)), dim3(256), 0, 0, nt, m.e0.p, m.e1.p, m.e2.p,
                               m.eo.p, m.to.p, m.da.p, m.area.p, reinterpret_cast<double *>(op->c0.p),
                               reinterpret_cast<double *>(op->c1.p), reinterpret_cast<double *>(op->c2.p));
          else
            hipLaunchKernelGGL(k_coeff11<float>, dim3(dec::nblocks(nt)), dim3(256), 0, 0, nt, m.e0.p, m.e1.p, m.e2.p,
                               m.eo.p, m.to.p, m.da.p, m.area.p, reinterpret_cast<float *>(op->c0.p),
                               reinterpret_cast<float *>(op->c1.p), reinterpret_cast<float *>(op->c2.p));
        }
        op->cols = ne;
        op->cols2 = ne;
      } else {
        unsupported(dm, dn);
      }
      break;
    }
    case DEC_OP_WEDGE_DD: {
      if (!((dm == 0 && dn == 1) || (dm == 1 && dn == 0)))
        fail(DEC_ERR_UNSUPPORTED, "Unsupported combination of degrees " + std::to_string(dm) + " and " +
                                      std::to_string(dn) +
                                      ". The currently supported dual-dual wedges are 0-1 and 1-0.");
      op->impl = I_WDD;
      op->rows = ne;
      dec::Csr eb;
      dec::edge_bduals(t, eb);
      std::vector<int32_t> ta(ne, -1), tb(ne, -1);
      std::vector<int32_t> order_ptr(ne + 1, 0), order_idx;
      for (int e = 0; e < ne; e++) {
        int cnt = eb.ptr[e + 1] - eb.ptr[e];
        if (cnt > 2) fail(DEC_ERR_UNSUPPORTED, "dual-dual wedge: edge with more than two cofaces");
        // weights in ascending triangle order
        std::vector<int32_t> ids(eb.idx.begin() + eb.ptr[e], eb.idx.begin() + eb.ptr[e + 1]);
        std::sort(ids.begin(), ids.end(), [&](int a, int b) { return a % nt < b % nt; });
        if (cnt > 0) ta[e] = ids[0] % nt;
        if (cnt > 1) tb[e] = ids[1] % nt;
        for (int id : ids) order_idx.push_back(id);
        order_ptr[e + 1] = static_cast<int32_t>(order_idx.size());
      }
      op->i0.upload(ta);
      op->i1.upload(tb);
      DBuf<int32_t> dp, di;
      dp.upload(order_ptr);
      di.upload(order_idx);
      size_t bytes = (f64 ? 8 : 4) * static_cast<size_t>(ne);
      op->c0.alloc(bytes);
      op->c1.alloc(bytes);
      const double *dlB = m.dl.p + 2 * static_cast<size_t>(ne);
      if (ne) {
        if (f64)
          hipLaunchKernelGGL(k_coeff_dd<double>, dim3(dec::nblocks(ne)), dim3(256), 0, 0, ne, dp.p, di.p, dlB,
                             reinterpret_cast<double *>(op->c0.p), reinterpret_cast<double *>(op->c1.p));
        else
          hipLaunchKernelGGL(k_coeff_dd<float>, dim3(dec::nblocks(ne)), dim3(256), 0, 0, ne, dp.p, di.p, dlB,
                             reinterpret_cast<float *>(op->c0.p), reinterpret_cast<float *>(op->c1.p));
      }
      dec::check_launch("wedge_dd coefficients");
      dec::dsync();
      op->cols = dm == 0 ? nt : ne;
      op->cols2 = dm == 0 ? ne : nt;
      break;
    }
    case DEC_OP_FLAT_SHARP: {
      op->impl = I_ELL;
      op->rows = op->cols = ne;
      flat_sharp_matrix(m, op->csr);
      if (f64) dec::upload_ell(op->csr, op->ell64);
      else dec::upload_ell(op->csr, op->ell32);
      break;
    }
    case DEC_OP_WEDGE_PD:
    case DEC_OP_WEDGE_DP: {
      if (dm == 1 && dn == 1) {
        // dec_wedge_product_pd(1,1) / dec_wedge_product_dp(1,1) (FastDEC.jl:390-400)
        op->impl = I_COMP;
        op->comp = C_WPD11;
        op->rows = nt;
        op->cols = op->cols2 = ne;
        op->sub.emplace_back(create(m, DEC_OP_FLAT_SHARP, 0, 0, hodge, dtype));
        op->sub.emplace_back(create(m, DEC_OP_WEDGE, 1, 1, hodge, dtype));
        add_tmp(ne);
        break;
      }
      if (kind == DEC_OP_WEDGE_DP && !(dm == 1 && dn == 0))
        fail(DEC_ERR_UNSUPPORTED, "Unsupported combination of degrees " + std::to_string(dm) + " and " +
                                      std::to_string(dn) +
                                      ". Ensure that their sum is not greater than the degree of the complex. The "
                                      "currently supported dual-primal wedges are 0-1, 1-0, and 1-1.");
      // dec_wedge_product_pd(0,1) / dec_wedge_product_dp(1,0) (FastDEC.jl:327-346, 379-387)
      if (kind == DEC_OP_WEDGE_PD && !(dm == 0 && dn == 1))
        fail(DEC_ERR_UNSUPPORTED, "Unsupported combination of degrees " + std::to_string(dm) + " and " +
                                      std::to_string(dn) +
                                      ". Ensure that their sum is not greater than the degree of the complex. The "
                                      "currently supported primal-dual wedges are 0-1, 1-0, and 1-1.");
      op->impl = I_WPD;
      op->rows = ne;
      op->swap = dm == 1;  // dp(1,0) takes (f, g): the kernel runs on (g, f)
      op->cols = dm == 0 ? nv : ne;
      op->cols2 = dm == 0 ? ne : nv;
      pd01_matrix(m, op->csr);
      if (f64) dec::upload_ell(op->csr, op->ell64);
      else dec::upload_ell(op->csr, op->ell32);
      break;
    }
    case DEC_OP_DIFFERENTIAL:
    case DEC_OP_BOUNDARY:
    case DEC_OP_DUAL_DERIVATIVE: {
      // ∂_n = d_{n-1}ᵀ; d̃0 = d1ᵀ; d̃1 = -d0ᵀ (FastDEC.jl:433-453)
      int which;        // 0: d0, 1: d1, 2: d0ᵀ, 3: d1ᵀ
      bool neg = false;
      if (kind == DEC_OP_DIFFERENTIAL) {
        if (dm != 0 && dm != 1) fail(DEC_ERR_UNSUPPORTED, "dec_differential: n must be 0 or 1 on a 2-D complex");
        which = dm;
      } else if (kind == DEC_OP_BOUNDARY) {
        if (dm != 1 && dm != 2) fail(DEC_ERR_UNSUPPORTED, "dec_boundary: n must be 1 or 2 on a 2-D complex");
        which = dm == 1 ? 2 : 3;
      } else {
        if (dm != 0 && dm != 1) fail(DEC_ERR_UNSUPPORTED, "dec_dual_derivative: n must be 0 or 1 on a 2-D complex");
        which = dm == 0 ? 3 : 2;
        neg = dm == 1;
      }
      if (which == 0) {
        op->impl = I_D0;
        op->rows = ne;
        op->cols = nv;
        op->i0.upload(t.v0);
        op->i1.upload(t.v1);
        std::vector<int8_t> sg(ne);
        for (int e = 0; e < ne; e++) sg[e] = t.eo[e] == 1 ? 1 : -1;
        op->s8.upload(sg);
        csr_d0(t, op->csr);
      } else if (which == 1) {
        op->impl = I_D1;
        op->rows = nt;
        op->cols = ne;
        csr_d1(t, op->csr);
        std::vector<int32_t> a(nt), b(nt), c(nt);
        std::vector<int8_t> sg(nt);
        for (int i = 0; i < nt; i++) {
          int p = op->csr.ptr[i];
          if (op->csr.ptr[i + 1] - p != 3) fail(DEC_ERR_UNSUPPORTED, "d1: triangle with a repeated edge");
          a[i] = op->csr.col[p];
          b[i] = op->csr.col[p + 1];
          c[i] = op->csr.col[p + 2];
          sg[i] = static_cast<int8_t>((op->csr.val[p] > 0 ? 1 : 0) | (op->csr.val[p + 1] > 0 ? 2 : 0) |
                                      (op->csr.val[p + 2] > 0 ? 4 : 0));
        }
        op->i0.upload(a);
        op->i1.upload(b);
        op->i2.upload(c);
        op->s8.upload(sg);
      } else if (which == 2) {
        op->impl = I_SELL;
        op->rows = nv;
        op->cols = ne;
        dec::Csr ve;
        dec::vertex_edges(t, ve);
        to_sign_ell(ve, op->sell, neg);
        csr_from_signs(ve, ne, neg, op->csr);
      } else {
        op->impl = I_SELL;
        op->rows = ne;
        op->cols = nt;
        dec::Csr et;
        dec::edge_triangles(t, et);
        to_sign_ell(et, op->sell, false);
        csr_from_signs(et, nt, false, op->csr);
      }
      break;
    }
    case DEC_OP_HODGE:
    case DEC_OP_INV_HODGE: {
      if (dm < 0 || dm > 2) fail(DEC_ERR_UNSUPPORTED, "dec_hodge_star: n must be 0, 1 or 2 on a 2-D complex");
      bool inv = kind == DEC_OP_INV_HODGE;
      if (hodge == DEC_GEOMETRIC_HODGE && dm == 1) {
        op->impl = inv ? I_GINV : I_ELL;
        op->rows = op->cols = ne;
        dec::geo_hodge1(m, op->csr);
        if (f64) dec::upload_ell(op->csr, op->ell64);
        else dec::upload_ell(op->csr, op->ell32);
        if (inv) {
          // Jacobi scaling from the diagonal of ⋆1; CG work vectors
          std::vector<double> dg(ne, 0.0);
          for (int r = 0; r < ne; r++)
            for (int k = op->csr.ptr[r]; k < op->csr.ptr[r + 1]; k++)
              if (op->csr.col[k] == r) dg[r] = op->csr.val[k];
          for (int r = 0; r < ne; r++)
            if (!(dg[r] > 0.0)) fail(DEC_ERR_UNSUPPORTED, "geometric ⋆1⁻¹: ⋆1 has a non-positive diagonal entry");
          size_t es = f64 ? 8 : 4;
          if (f64) {
            std::vector<double> di(ne);
            for (int r = 0; r < ne; r++) di[r] = 1.0 / dg[r];
            op->dinv.alloc(es * ne);
            if (ne) dec::h2d(op->dinv.p, di.data(), es * ne);
          } else {
            std::vector<float> di(ne);
            for (int r = 0; r < ne; r++) di[r] = static_cast<float>(1.0 / dg[r]);
            op->dinv.alloc(es * ne);
            if (ne) dec::h2d(op->dinv.p, di.data(), es * ne);
          }
          // symmetric to rounding (circumcentric duals): CG; else BiCGSTAB
          double amax = 0.0, asym = 0.0;
          {
            std::vector<std::unordered_map<int, double>> rowmap(ne);
            for (int r = 0; r < ne; r++)
              for (int k = op->csr.ptr[r]; k < op->csr.ptr[r + 1]; k++) {
                rowmap[r][op->csr.col[k]] += op->csr.val[k];
                amax = std::max(amax, std::fabs(op->csr.val[k]));
              }
            for (int r = 0; r < ne; r++)
              for (auto &kv : rowmap[r]) {
                auto it = rowmap[kv.first].find(r);
                double other = it == rowmap[kv.first].end() ? 0.0 : it->second;
                asym = std::max(asym, std::fabs(kv.second - other));
              }
          }
          op->symmetric = asym <= 1e-12 * amax;
          for (auto *v : {&op->cr, &op->cp, &op->cq, &op->cz, &op->crh, &op->ct}) v->alloc(es * ne);
          op->cpart.alloc(6 * kFcBlocks);
          op->cscal.alloc(CS_COUNT);
          {
            std::vector<double> init(CS_COUNT, 0.0);
            dec::h2d(op->cscal.p, init.data(), sizeof(double) * CS_COUNT);
          }
          op->max_iter = f64 ? 200 : 100;
          op->rtol = f64 ? 1e-14 : 1e-6;
          op->csr = dec::HostCsr();  // the inverse has no sparse matrix to export
        }
        break;
      }
      op->impl = I_DIAG;
      const DBuf<double> &h = dm == 0 ? m.h0 : (dm == 1 ? m.h1 : m.h2);
      op->rows = op->cols = static_cast<int64_t>(h.n);
      // dec_inv_hodge_star: (1 ./ hdg) .* (-1)^(k(n-k)) (FastDEC.jl:721-727)
      double mult = ((dm * (2 - dm)) % 2 == 0) ? 1.0 : -1.0;
      if (f64) make_diag<double>(*op, h, inv, mult);
      else make_diag<float>(*op, h, inv, mult);
      std::vector<double> hv = h.download();
      if (inv)
        for (auto &x : hv) x = (1 / x) * mult;
      csr_diag(hv, op->csr);
      break;
    }
    case DEC_OP_LAPLACIAN0: {
      op->impl = I_ELL;
      op->rows = op->cols = nv;
      std::vector<double> h0 = m.h0.download(), h1 = m.h1.download(), ih0(nv);
      for (int v = 0; v < nv; v++) ih0[v] = 1 / h0[v];
      if (hodge == DEC_GEOMETRIC_HODGE) {
        dec::HostCsr H;
        dec::geo_hodge1(m, H);
        dec::assemble_laplacian0(t, ih0, h1, &H, op->csr);
      } else {
        dec::assemble_laplacian0(t, ih0, h1, nullptr, op->csr);
      }
      if (f64) dec::upload_ell(op->csr, op->ell64);
      else dec::upload_ell(op->csr, op->ell32);
      break;
    }
    case DEC_OP_P2_D2_INTERPOLATION:
    case DEC_OP_P0_D0_INTERPOLATION:
    case DEC_OP_D0_P0_INTERPOLATION: {
      dec::HostCsr P;
      p2_d2_matrix(m, kind == DEC_OP_P0_D0_INTERPOLATION, P);
      if (kind == DEC_OP_D0_P0_INTERPOLATION) scale_rows_cols(P, inv_diag(m.h0), inv_diag(m.h2));
      auto e = ell_op(std::move(P), dtype);
      e->kind = kind;
      e->hodge = hodge;
      dec::check_launch("dec_op_create");
      dec::dsync();
      return e.release();
    }
    case DEC_OP_P3_D3_INTERPOLATION:
      fail(DEC_ERR_UNSUPPORTED, "p3_d3_interpolation is defined on 3-D complexes");
    case DEC_OP_AVG01: {
      op->impl = I_AVG;
      op->rows = ne;
      op->cols = nv;
      op->i0.upload(t.v0);
      op->i1.upload(t.v1);
      csr_d0(t, op->csr);
      for (auto &x : op->csr.val) x = 0.5;
      break;
    }
    case DEC_OP_INTERIOR_PRODUCT_DD: {
      op->impl = I_COMP;
      if (dm == 1 && dn == 1) {
        // interior_product_dd(1,1) (FastDEC.jl:762-773): dual 1-forms f, g -> dual 0-form
        op->comp = C_I11;
        op->rows = nt;
        op->cols = op->cols2 = ne;
        op->sub.emplace_back(create(m, DEC_OP_INV_HODGE, 1, 0, DEC_GEOMETRIC_HODGE, dtype));
        op->sub.emplace_back(create(m, DEC_OP_FLAT_SHARP, 0, 0, hodge, dtype));
        op->sub.emplace_back(create(m, DEC_OP_WEDGE, 1, 1, hodge, dtype));
        op->sub.emplace_back(create(m, DEC_OP_HODGE, 2, 0, DEC_GEOMETRIC_HODGE, dtype));
        add_tmp(ne);
        add_tmp(ne);
        add_tmp(nt);
      } else if (dm == 1 && dn == 2) {
        // interior_product_dd(1,2) (FastDEC.jl:775-786): dual 1-form f, dual 2-form g -> dual 1-form
        op->comp = C_I12;
        op->rows = ne;
        op->cols = ne;
        op->cols2 = nv;
        op->sub.emplace_back(create(m, DEC_OP_INV_HODGE, 0, 0, DEC_GEOMETRIC_HODGE, dtype));
        op->sub.emplace_back(create(m, DEC_OP_WEDGE_PD, 0, 1, hodge, dtype));
        op->sub.emplace_back(create(m, DEC_OP_FLAT_SHARP, 0, 0, hodge, dtype));
        op->sub.emplace_back(create(m, DEC_OP_HODGE, 1, 0, DEC_GEOMETRIC_HODGE, dtype));
        add_tmp(nv);
        add_tmp(ne);
        add_tmp(ne);
      } else {
        fail(DEC_ERR_UNSUPPORTED, "interior_product_dd: supported degrees are (1,1) and (1,2)");
      }
      break;
    }
    case DEC_OP_LIE_DD: {
      // ℒ_dd(1,1) (FastDEC.jl:791-803)
      if (!(dm == 1 && dn == 1)) fail(DEC_ERR_UNSUPPORTED, "ℒ_dd: supported degrees are (1,1)");
      op->impl = I_COMP;
      op->comp = C_LIE;
      op->rows = ne;
      op->cols = op->cols2 = ne;
      op->sub.emplace_back(create(m, DEC_OP_INTERIOR_PRODUCT_DD, 1, 1, hodge, dtype));
      op->sub.emplace_back(create(m, DEC_OP_DUAL_DERIVATIVE, 0, 0, hodge, dtype));
      op->sub.emplace_back(create(m, DEC_OP_DUAL_DERIVATIVE, 1, 0, hodge, dtype));
      op->sub.emplace_back(create(m, DEC_OP_INTERIOR_PRODUCT_DD, 1, 2, hodge, dtype));
      add_tmp(nt);
      add_tmp(ne);
      add_tmp(nv);
      add_tmp(ne);
      break;
    }
    case DEC_OP_DUAL_LAPLACIAN: {
      op->impl = I_COMP;
      auto csr_of = [&](int k, int a, int h) {
        std::unique_ptr<dec_op_s> o(create(m, k, a, 0, h, DEC_F64));
        return std::move(o->csr);
      };
      if (dm == 0) {
        // Δᵈ(Val(0)) on a 2-D complex (FastDEC.jl:821-829): m = ⋆2 d1; x -> m ⋆1⁻¹(d̃0 x)
        op->comp = C_DL0;
        op->rows = op->cols = nt;
        op->sub.emplace_back(create(m, DEC_OP_DUAL_DERIVATIVE, 0, 0, hodge, dtype));
        op->sub.emplace_back(create(m, DEC_OP_INV_HODGE, 1, 0, DEC_GEOMETRIC_HODGE, dtype));
        dec::HostCsr M;
        dec::spgemm(csr_of(DEC_OP_HODGE, 2, DEC_GEOMETRIC_HODGE), csr_of(DEC_OP_DIFFERENTIAL, 1, hodge), M);
        op->sub.emplace_back(ell_op(std::move(M), dtype));
        add_tmp(ne);
        add_tmp(ne);
      } else if (dm == 1) {
        // Δᵈ(Val(1)) (FastDEC.jl:835-860): M = ⋆1 d0 ⋆0⁻¹ d̃1, N = d̃0 ⋆2 d1; x -> M x + N ⋆1⁻¹(x)
        op->comp = C_DL1;
        op->rows = op->cols = ne;
        op->sub.emplace_back(create(m, DEC_OP_INV_HODGE, 1, 0, DEC_GEOMETRIC_HODGE, dtype));
        dec::HostCsr t1, t2, M, N;
        dec::spgemm(csr_of(DEC_OP_HODGE, 1, DEC_GEOMETRIC_HODGE), csr_of(DEC_OP_DIFFERENTIAL, 0, hodge), t1);
        dec::spgemm(t1, csr_of(DEC_OP_INV_HODGE, 0, DEC_GEOMETRIC_HODGE), t2);
        dec::spgemm(t2, csr_of(DEC_OP_DUAL_DERIVATIVE, 1, hodge), M);
        dec::spgemm(csr_of(DEC_OP_DUAL_DERIVATIVE, 0, hodge), csr_of(DEC_OP_HODGE, 2, DEC_GEOMETRIC_HODGE), t1);
        dec::spgemm(t1, csr_of(DEC_OP_DIFFERENTIAL, 1, hodge), N);
        op->sub.emplace_back(ell_op(std::move(M), dtype));
        op->sub.emplace_back(ell_op(std::move(N), dtype));
        add_tmp(ne);
        add_tmp(ne);
        add_tmp(ne);
      } else {
        fail(DEC_ERR_UNSUPPORTED, "Δᵈ: n must be 0 or 1 on a 2-D complex");
      }
      break;
    }
    default:
      fail(DEC_ERR_ARG, "dec_op_create: unknown operator kind");
  }
  schedule_gathers(*op, m);
  dec::check_launch("dec_op_create");
  dec::dsync();
  return op.release();
}

}  // namespace

extern "C" {

int dec_op_create(dec_mesh m, int32_t kind, int32_t deg_m, int32_t deg_n, int32_t hodge, int32_t dtype, dec_op *out) {
  DEC_TRY
  if (!m || !out) fail(DEC_ERR_ARG, "dec_op_create: NULL handle");
  *out = nullptr;
  *out = create(*m, kind, deg_m, deg_n, hodge, dtype);
  DEC_CATCH
}

int dec_op_shape(dec_op op, int64_t *rows, int64_t *cols, int64_t *cols2) {
  DEC_TRY
  if (!op) fail(DEC_ERR_ARG, "dec_op_shape: NULL handle");
  if (rows) *rows = op->rows;
  if (cols) *cols = op->cols;
  if (cols2) *cols2 = op->cols2;
  DEC_CATCH
}

int dec_op_solve_info(dec_op op, int32_t *iterations, double *relres, void *stream) {
  DEC_TRY
  if (!op) fail(DEC_ERR_ARG, "dec_op_solve_info: NULL handle");
  if (op->impl != I_GINV) fail(DEC_ERR_ARG, "dec_op_solve_info: the operator is not a solve");
  double sc[CS_COUNT];
  hipStream_t s = static_cast<hipStream_t>(stream);
  hip_status(hipMemcpyAsync(sc, op->cscal.p, sizeof(sc), hipMemcpyDeviceToHost, s), "solve info copy");
  hip_status(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (iterations) *iterations = static_cast<int32_t>(sc[CS_ITER]);
  if (relres) *relres = sc[CS_BB] > 0.0 ? std::sqrt(sc[CS_RR] / sc[CS_BB]) : 0.0;
  DEC_CATCH
}

int dec_op_apply(dec_op op, const void *x, const void *y, void *out, void *stream) {
  DEC_TRY
  if (!op) fail(DEC_ERR_ARG, "dec_op_apply: NULL handle");
  if (op->rows > 0 && (!x || !out)) fail(DEC_ERR_ARG, "dec_op_apply: NULL form");
  if (op->cols2 > 0 && !y) fail(DEC_ERR_ARG, "dec_op_apply: this operator takes two forms");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (op->dtype == DEC_F64) apply_t<double>(*op, x, y, out, s);
  else apply_t<float>(*op, x, y, out, s);
  DEC_CATCH
}

int dec_op_export_csr(dec_op op, int64_t *nnz, int32_t *rowptr, int32_t *col, double *val) {
  DEC_TRY
  if (!op) fail(DEC_ERR_ARG, "dec_op_export_csr: NULL handle");
  if (op->csr.ptr.empty()) fail(DEC_ERR_UNSUPPORTED, "dec_op_export_csr: operator has no matrix form");
  if (nnz) *nnz = static_cast<int64_t>(op->csr.col.size());
  if (rowptr) std::copy(op->csr.ptr.begin(), op->csr.ptr.end(), rowptr);
  if (col) std::copy(op->csr.col.begin(), op->csr.col.end(), col);
  if (val) std::copy(op->csr.val.begin(), op->csr.val.end(), val);
  DEC_CATCH
}

}  // extern "C"

namespace {
// slots x rows weights of a tetrahedron wedge as float64
void export_weights(const dec_op_s &op, int slots, double *out) {
  size_t cnt = static_cast<size_t>(slots) * static_cast<size_t>(op.rows);
  if (op.dtype == DEC_F64) {
    dec::d2h(out, op.cx.p, sizeof(double) * cnt);
  } else {
    std::vector<float> tmp(cnt);
    dec::d2h(tmp.data(), op.cx.p, sizeof(float) * cnt);
    for (size_t i = 0; i < cnt; i++) out[i] = tmp[i];
  }
}
}  // namespace

extern "C" {

int dec_op_export_tables(dec_op op, int32_t *idx, double *coeff) {
  DEC_TRY
  if (!op) fail(DEC_ERR_ARG, "dec_op_export_tables: NULL handle");
  size_t n = static_cast<size_t>(op->rows);
  auto dl_i = [&](DBuf<int32_t> &b, size_t slot) {
    if (idx && b.p && n) dec::d2h(idx + slot * n, b.p, sizeof(int32_t) * n);
  };
  auto dl_c = [&](DBuf<unsigned char> &b, size_t slot) {
    if (!coeff || !b.p || !n) return;
    if (op->dtype == DEC_F64) {
      dec::d2h(coeff + slot * n, b.p, sizeof(double) * n);
    } else {
      std::vector<float> tmp(n);
      dec::d2h(tmp.data(), b.p, sizeof(float) * n);
      for (size_t i = 0; i < n; i++) coeff[slot * n + i] = tmp[i];
    }
  };
  switch (op->impl) {
    case I_W01: dl_i(op->i0, 0); dl_i(op->i1, 1); break;
    case I_W02:
    case I_W11:
      dl_i(op->i0, 0); dl_i(op->i1, 1); dl_i(op->i2, 2);
      dl_c(op->c0, 0); dl_c(op->c1, 1); dl_c(op->c2, 2);
      break;
    case I_WDD:
      dl_i(op->i0, 0); dl_i(op->i1, 1); dl_c(op->c0, 0); dl_c(op->c1, 1);
      break;
    case I_W03:  // vertex slots 0..3 and their weights
      if (idx && n) dec::d2h(idx, op->ix.p, sizeof(int32_t) * 4 * n);
      if (coeff && n) export_weights(*op, 4, coeff);
      break;
    case I_W21: {  // edges e0..e5 and triangles t0..t3; the 12 signed coefficients
      if (idx && n) dec::d2h(idx, op->ix.p, sizeof(int32_t) * 10 * n);
      if (coeff && n) {
        std::vector<double> ws(4 * n);
        export_weights(*op, 4, ws.data());
        std::vector<uint16_t> sg(n);
        dec::d2h(sg.data(), op->m16.p, sizeof(uint16_t) * n);
        const int vert[12] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3};
        for (int k = 0; k < 12; k++)
          for (size_t x = 0; x < n; x++) {
            double w = ws[vert[k] * n + x];
            coeff[k * n + x] = ((sg[x] >> k) & 1u) ? -w : w;
          }
      }
      break;
    }
    default:
      fail(DEC_ERR_UNSUPPORTED, "dec_op_export_tables: operator has no wedge tables");
  }
  DEC_CATCH
}

int dec_op_apply_many(int32_t count, const dec_op *ops, const void *const *x, const void *const *y, void *const *out,
                      void *stream) {
  DEC_TRY
  if (count < 0 || (count > 0 && (!ops || !x || !y || !out))) fail(DEC_ERR_ARG, "dec_op_apply_many: bad argument");
  if (count == 0) return DEC_OK;
  int dtype = -1;
  for (int q = 0; q < count; q++) {
    dec_op_s *op = ops[q];
    if (!op) fail(DEC_ERR_ARG, "dec_op_apply_many: NULL operator");
    if (op->impl == I_GINV || op->impl == I_COMP)
      fail(DEC_ERR_UNSUPPORTED,
           "dec_op_apply_many: solves and composite operators run several launches; apply them with dec_op_apply");
    if (dtype < 0) dtype = op->dtype;
    if (op->dtype != dtype) fail(DEC_ERR_ARG, "dec_op_apply_many: operators of different dtypes");
    if (op->rows > 0 && !out[q]) fail(DEC_ERR_ARG, "dec_op_apply_many: NULL output");
    if (op->rows > 0 && op->cols > 0 && !x[q]) fail(DEC_ERR_ARG, "dec_op_apply_many: NULL input");
    if (op->rows > 0 && op->cols2 > 0 && !y[q]) fail(DEC_ERR_ARG, "dec_op_apply_many: NULL second input");
  }
  // The items of one call run concurrently: an output that overlaps another
  // item's input or output would race, so such lists are refused.
  const size_t es = dtype == DEC_F64 ? sizeof(double) : sizeof(float);
  auto span = [es](const void *p, int64_t n) {
    uintptr_t a = reinterpret_cast<uintptr_t>(p);
    return std::make_pair(a, a + static_cast<uintptr_t>(n > 0 ? n : 0) * es);
  };
  auto overlap = [](std::pair<uintptr_t, uintptr_t> a, std::pair<uintptr_t, uintptr_t> b) {
    return a.first < a.second && b.first < b.second && a.first < b.second && b.first < a.second;
  };
  for (int q = 0; q < count; q++) {
    if (ops[q]->rows <= 0) continue;
    auto o = span(out[q], ops[q]->rows);
    for (int r = 0; r < count; r++) {
      if (r == q) continue;
      const dec_op_s *opr = ops[r];
      if (opr->rows <= 0) continue;
      if (overlap(o, span(out[r], opr->rows)) || (opr->cols > 0 && overlap(o, span(x[r], opr->cols))) ||
          (opr->cols2 > 0 && overlap(o, span(y[r], opr->cols2))))
        fail(DEC_ERR_ARG, "dec_op_apply_many: the output of item " + std::to_string(q) +
                              " overlaps an input or output of item " + std::to_string(r) +
                              "; items of one call run concurrently");
    }
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == DEC_F64) apply_group_t<double>(count, ops, x, y, out, s);
  else apply_group_t<float>(count, ops, x, y, out, s);
  DEC_CATCH
}

int dec_op_destroy(dec_op op) {
  DEC_TRY
  delete op;
  DEC_CATCH
}

}  // extern "C"
