// Learned multigrid smoother of libdec_hip.so: a 3-layer edge message-passing
// network over a level's operator graph, with hand-written forward, backward
// and Adam kernels.  No counterpart in the reference (its smoother is Krylov
// CG, src/Multigrid.jl:860).
//
// For a residual r on a level with operator A (D = diag A, P = D⁻¹(D - A)):
//   s = sqrt(mean r²), x = r / s, f0 = [x, P x]
//   h1 = relu(f0 W1 + b1)             2 -> 32   (VALU)
//   h2 = relu([h1, P h1] W2 + b2)     64 -> 32  (BF16 MFMA, FP32 accumulate)
//   h3 = relu([h2, P h2] W3 + b3)     64 -> 32  (BF16 MFMA, FP32 accumulate)
//   y  = alpha x + h3 wo + bo, c = (s / D) y
// trained on mean((r - A c)² / s²).  Activations are stored in BF16 (FP32 in
// the "precise" mode, whose dense layers run on the VALU in FP32 and which
// exists to check the kernels against the float64 oracle).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>

#include "dec_internal.h"
#include "kernels_common.h"
#include "ell_rows.h"
#include "learned.h"

using dec::DBuf;
using dec::fail;

namespace {

constexpr int H = 32;          // hidden width
constexpr int NP = 4290;       // parameter count
// parameter offsets (flat float vector)
constexpr int OW1 = 0, OB1 = 64, OW2 = 96, OB2 = 2144, OW3 = 2176, OB3 = 4224, OWO = 4256, OBO = 4288, OAL = 4289;
constexpr int TB = 256;

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// float -> BF16, round to nearest even, by gfx950's v_cvt_pk_bf16_f32 (one
// instruction for two values).  For every finite input and ±inf it gives the
// bits of the integer rounding u + 0x7FFF + ((u >> 16) & 1), u = bits of f.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned int f2bf2(float a, float b) {
  f32x2_t f = {a, b};
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(f, bf16x2_t));
}
__device__ __forceinline__ unsigned short f2bf(float f) {
  return __builtin_bit_cast(unsigned short, static_cast<__bf16>(f));
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(static_cast<unsigned int>(h) << 16); }

// activation storage: BF16 (as raw bits) or FP32
template <class AT>
__device__ __forceinline__ float ald(const AT *p, size_t i);
template <>
__device__ __forceinline__ float ald<unsigned short>(const unsigned short *p, size_t i) { return bf2f(p[i]); }
template <>
__device__ __forceinline__ float ald<float>(const float *p, size_t i) { return p[i]; }
template <class AT>
__device__ __forceinline__ void ast(AT *p, size_t i, float v);
template <>
__device__ __forceinline__ void ast<unsigned short>(unsigned short *p, size_t i, float v) { p[i] = f2bf(v); }
template <>
__device__ __forceinline__ void ast<float>(float *p, size_t i, float v) { p[i] = v; }

__device__ __forceinline__ float ell_row(int n, int width, int r, const int *__restrict__ col,
                                         const float *__restrict__ val, const float *__restrict__ x) {
  float acc = 0.0f;
  for (int k = 0; k < width; k++) {
    size_t q = static_cast<size_t>(k) * n + r;
    int c = col[q];
    if (c < 0) break;
    acc = acc + val[q] * x[c];
  }
  return acc;
}

// ---- forward ------------------------------------------------------------

// Per-sample scale s_b = sqrt(mean r²): kScaleBlocks partial sums per sample
// (block j sums i ≡ j·TB + t mod kScaleBlocks·TB in double, then a fixed
// tree), summed in block order by one thread per sample.
constexpr int kScaleBlocks = 128;

__global__ void k_scale_partial(int n, const float *__restrict__ r, double *__restrict__ part) {
  __shared__ double sh[TB];
  const float *rb = r + static_cast<size_t>(blockIdx.y) * n;
  double acc = 0.0;
  constexpr int stride = kScaleBlocks * TB;
  // 32 loads in flight (one round trip up to 1,048,576 rows), added in the
  // order of i; past the end a clamped load adds +0, which leaves the sum of
  // squares unchanged
  for (int i = blockIdx.x * TB + threadIdx.x; i < n; i += 32 * stride) {
    float a[32];
#pragma unroll
    for (int u = 0; u < 32; u++) a[u] = rb[min(i + u * stride, n - 1)];
#pragma unroll
    for (int u = 0; u < 32; u++) acc += i + u * stride < n ? static_cast<double>(a[u]) * a[u] : 0.0;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[static_cast<size_t>(blockIdx.y) * kScaleBlocks + blockIdx.x] = sh[0];
}

__global__ void k_scale_final(int n, int B, const double *__restrict__ part, float *__restrict__ s) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double acc = 0.0;
  for (int j = 0; j < kScaleBlocks; j++) acc += part[static_cast<size_t>(b) * kScaleBlocks + j];
  double m = acc / n;
  s[b] = m > 0 ? static_cast<float>(sqrt(m)) : 1.0f;
}

// x = r / s, f0 = [x, P x], h1 = relu(f0 W1 + b1)
template <class AT>
__global__ void k_input(int n, int pw, const int *__restrict__ pcol, const float *__restrict__ pval,
                        const float *__restrict__ r, const float *__restrict__ s, const float *__restrict__ th,
                        float *__restrict__ x, float *__restrict__ px, AT *__restrict__ h1) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (v >= n) return;
  size_t o = static_cast<size_t>(b) * n;
  float inv = 1.0f / s[b];
  float xv = r[o + v] * inv;
  float acc = 0.0f;
  for (int k = 0; k < pw; k++) {
    size_t q = static_cast<size_t>(k) * n + v;
    int c = pcol[q];
    if (c < 0) break;
    acc = acc + pval[q] * (r[o + c] * inv);
  }
  x[o + v] = xv;
  px[o + v] = acc;
  for (int j = 0; j < H; j++) {
    float z = xv * th[OW1 + j] + acc * th[OW1 + H + j] + th[OB1 + j];
    ast(h1, (o + v) * H + j, z > 0.0f ? z : 0.0f);
  }
}

// m = P h (32 features): 32 lanes per vertex, one feature per lane.
template <class AT>
__global__ void k_spmm(int n, int pw, const int *__restrict__ pcol, const float *__restrict__ pval,
                       const AT *__restrict__ h, AT *__restrict__ m) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int v = t >> 5, f = t & 31;
  int b = blockIdx.y;
  if (v >= n) return;
  size_t o = static_cast<size_t>(b) * n;
  float acc = 0.0f;
  for (int k = 0; k < pw; k++) {
    size_t q = static_cast<size_t>(k) * n + v;
    int c = pcol[q];
    if (c < 0) break;
    acc = acc + pval[q] * ald(h, (o + c) * H + f);
  }
  ast(m, (o + v) * H + f, acc);
}

// Dense 64 -> 32 with bias and relu on BF16 MFMA: each 64-lane wavefront
// computes 16 rows x 32 columns as two 16x16 tiles, K = 64 in four steps of
// v_mfma_f32_16x16x16bf16_1k.  Operand layout: lane l holds A[l%16][4(l/16)+i]
// and B[4(l/16)+i][l%16]; the result D[4(l/16)+i][l%16].
__global__ void k_dense_bf16(int rows, const unsigned short *__restrict__ h, const unsigned short *__restrict__ m,
                             const float *__restrict__ W, const float *__restrict__ bias,
                             unsigned short *__restrict__ out) {
  int lane = threadIdx.x & 63;
  int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int r0 = wave * 16;
  if (r0 >= rows) return;
  int li = lane & 15, lg = lane >> 4;
  int row = r0 + li;
  short4_t A[4];
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    int k = ks * 16 + 4 * lg;  // 4 consecutive k of the concatenated [h, m]
    const unsigned short *src = k < H ? h + static_cast<size_t>(row) * H + k : m + static_cast<size_t>(row) * H + (k - H);
    short4_t a;
    if (row < rows) {
      a[0] = static_cast<short>(src[0]);
      a[1] = static_cast<short>(src[1]);
      a[2] = static_cast<short>(src[2]);
      a[3] = static_cast<short>(src[3]);
    } else {
      a[0] = a[1] = a[2] = a[3] = 0;
    }
    A[ks] = a;
  }
  float4_t acc[2];
#pragma unroll
  for (int jt = 0; jt < 2; jt++) {
    acc[jt] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      short4_t bv;
#pragma unroll
      for (int i = 0; i < 4; i++) bv[i] = static_cast<short>(f2bf(W[(ks * 16 + 4 * lg + i) * H + jt * 16 + li]));
      acc[jt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[ks], bv, acc[jt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int jt = 0; jt < 2; jt++) {
    int col = jt * 16 + li;
    float bc = bias[col];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int rr = r0 + 4 * lg + i;
      if (rr < rows) {
        float z = acc[jt][i] + bc;
        out[static_cast<size_t>(rr) * H + col] = f2bf(z > 0.0f ? z : 0.0f);
      }
    }
  }
}

// Dense 64 -> 32 in FP32 on the VALU (precise mode), thread per (row, col).
__global__ void k_dense_f32(int rows, const float *__restrict__ h, const float *__restrict__ m,
                            const float *__restrict__ W, const float *__restrict__ bias, float *__restrict__ out) {
  size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t row = t >> 5;
  int col = static_cast<int>(t & 31);
  if (row >= static_cast<size_t>(rows)) return;
  float z = bias[col];
  for (int k = 0; k < H; k++) z = z + h[row * H + k] * W[k * H + col];
  for (int k = 0; k < H; k++) z = z + m[row * H + k] * W[(H + k) * H + col];
  out[row * H + col] = z > 0.0f ? z : 0.0f;
}

// y = alpha x + h3 wo + bo; c = (s / D) y, optionally u += c.
template <class AT>
__global__ void k_output(int n, const float *__restrict__ x, const AT *__restrict__ h3, const float *__restrict__ th,
                         const float *__restrict__ s, const float *__restrict__ D, float *__restrict__ y,
                         float *__restrict__ c, float *__restrict__ u_add) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (v >= n) return;
  size_t o = static_cast<size_t>(b) * n + v;
  float acc = th[OAL] * x[o];
  float hw = 0.0f;
  for (int j = 0; j < H; j++) hw = hw + ald(h3, o * H + j) * th[OWO + j];
  float yy = acc + hw + th[OBO];
  if (y) y[o] = yy;
  float cc = (s[b] / D[v]) * yy;
  if (c) c[o] = cc;
  if (u_add) u_add[o] = u_add[o] + cc;
}

// ---- inference: the smoother inside V-cycles -------------------------------
// Three kernels for one residual (B = 1) after the scale's partial sums.
// Levels up to kInfSmallRows rows: x and P x; layer 1 recomputed for every
// neighbour from (x, P x) with layer 2 on MFMA; layer 3 on MFMA with the
// output head, adding c to u (16 rows per wavefront).  Larger levels: x, P x
// and the layer-1 rows h1 (stored); layer 2 gathering h1 and storing h2;
// layer 3 gathering h2 and adding c to u (64 rows per wavefront, four lanes
// to a row for the gathers).  The rows [h, P h] pass through LDS into the A
// operands of v_mfma_f32_16x16x16bf16 (four 16-row tiles; lane (li, lg) of
// tile q takes row 16q + li, K entries 4lg..4lg+3 of each K step).
// Expressions, summation orders and BF16 rounding points are those of
// k_input / k_spmm / k_dense_bf16 / k_output, so that u gains exactly the c
// of the BF16 forward.

constexpr int kInfWaves = TB / 64;
// levels up to this many rows use the 16-rows-per-wavefront layer kernels
constexpr int kInfSmallRows = 50000;
constexpr int kRowPitch = 2 * H + 8;  // bf16 per LDS row: [h | P h] plus padding
// tile-local indices of the halo kernels are bytes: 64 own rows, at most
// kHaloMax halo rows, 255 for a pad
constexpr size_t kHaloMax = 191;

__device__ __forceinline__ void load_wB(const float *__restrict__ W, int li, int lg, short4_t B[2][4]) {
#pragma unroll
  for (int jt = 0; jt < 2; jt++)
#pragma unroll
    for (int ks = 0; ks < 4; ks++)
#pragma unroll
      for (int i = 0; i < 4; i++) B[jt][ks][i] = static_cast<short>(f2bf(W[(ks * 16 + 4 * lg + i) * H + jt * 16 + li]));
}

__device__ __forceinline__ short4_t lds4(const unsigned short *p) {
  uint2 w = *reinterpret_cast<const uint2 *>(p);
  short4_t a;
  a[0] = static_cast<short>(w.x & 0xffffu);
  a[1] = static_cast<short>(w.x >> 16);
  a[2] = static_cast<short>(w.y & 0xffffu);
  a[3] = static_cast<short>(w.y >> 16);
  return a;
}

__device__ __forceinline__ void unpack8(uint4 w, float *f) {
  f[0] = bf2f(static_cast<unsigned short>(w.x & 0xffffu));
  f[1] = bf2f(static_cast<unsigned short>(w.x >> 16));
  f[2] = bf2f(static_cast<unsigned short>(w.y & 0xffffu));
  f[3] = bf2f(static_cast<unsigned short>(w.y >> 16));
  f[4] = bf2f(static_cast<unsigned short>(w.z & 0xffffu));
  f[5] = bf2f(static_cast<unsigned short>(w.z >> 16));
  f[6] = bf2f(static_cast<unsigned short>(w.w & 0xffffu));
  f[7] = bf2f(static_cast<unsigned short>(w.w >> 16));
}

// The four 16-row tiles of a wavefront's 64 LDS rows times W (64 x 32):
// acc[q][jt] holds D[4lg+i][16jt+li] of tile q.
__device__ __forceinline__ void mfma_rows(const unsigned short *rows, int li, int lg, const short4_t Bw[2][4],
                                          float4_t acc[4][2]) {
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const unsigned short *r = rows + (16 * q + li) * kRowPitch + 4 * lg;
    short4_t A0 = lds4(r), A1 = lds4(r + 16), A2 = lds4(r + 32), A3 = lds4(r + 48);
#pragma unroll
    for (int jt = 0; jt < 2; jt++) {
      float4_t c = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A0, Bw[jt][0], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A1, Bw[jt][1], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A2, Bw[jt][2], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A3, Bw[jt][3], c, 0, 0, 0);
      acc[q][jt] = c;
    }
  }
}

__device__ __forceinline__ uint4 pack8(const float *f) {
  uint4 w;
  w.x = f2bf2(f[0], f[1]);
  w.y = f2bf2(f[2], f[3]);
  w.z = f2bf2(f[4], f[5]);
  w.w = f2bf2(f[6], f[7]);
  return w;
}

// xp[v] = (x, P x) with x = r / s, and (h1 non-null) the row's layer-1
// features h1 = relu([x, P x] W1 + b1) in BF16, the values k_input stores.
// W > 0 is the width of P's rows: all slots' loads are issued before the
// sums, which run in slot order.
template <int W>
__global__ void k_inf_input(int n, int pw, const int *__restrict__ pcol, const float *__restrict__ pval,
                            const float *__restrict__ r, const double *__restrict__ part, float *__restrict__ s,
                            float2 *__restrict__ xp, const float *__restrict__ th, unsigned short *__restrict__ h1) {
  // s = sqrt(mean r²) from k_scale_partial's sums, added in k_scale_final's
  // order by one thread of every workgroup; workgroup 0 stores it for the
  // layer kernels
  __shared__ float s_sh;
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int j = 0; j < kScaleBlocks; j++) acc += part[j];
    double m = acc / n;
    float sv = m > 0 ? static_cast<float>(sqrt(m)) : 1.0f;
    s_sh = sv;
    if (blockIdx.x == 0) s[0] = sv;
  }
  __syncthreads();
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  float inv = 1.0f / s_sh;
  float xv = r[v] * inv;
  float acc = 0.0f;
  if constexpr (W > 0) {
    int c[W];
    float pv[W], rc[W];
#pragma unroll
    for (int k = 0; k < W; k++) {
      size_t q = static_cast<size_t>(k) * n + v;
      c[k] = pcol[q];
      pv[k] = pval[q];
    }
#pragma unroll
    for (int k = 0; k < W; k++) rc[k] = r[c[k] < 0 ? v : c[k]];
#pragma unroll
    for (int k = 0; k < W; k++)
      if (c[k] >= 0) acc = acc + pv[k] * (rc[k] * inv);
  } else {
    for (int k = 0; k < pw; k++) {
      size_t q = static_cast<size_t>(k) * n + v;
      int c = pcol[q];
      if (c < 0) break;
      acc = acc + pval[q] * (r[c] * inv);
    }
  }
  xp[v] = make_float2(xv, acc);
  if (h1) {
    float hh[H];
#pragma unroll
    for (int j = 0; j < H; j++) {
      float z = xv * th[OW1 + j] + acc * th[OW1 + H + j] + th[OB1 + j];
      hh[j] = z > 0.0f ? z : 0.0f;
    }
    uint4 *dst = reinterpret_cast<uint4 *>(h1 + static_cast<size_t>(v) * H);
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++) dst[k4] = pack8(hh + 8 * k4);
  }
}

// m += pv * (the 32 BF16 values of a row)
__device__ __forceinline__ void add_row(const uint4 *hc, float pv, float *m) {
#pragma unroll
  for (int k4 = 0; k4 < 4; k4++) {
    float f[8];
    unpack8(hc[k4], f);
#pragma unroll
    for (int e = 0; e < 8; e++) m[8 * k4 + e] = m[8 * k4 + e] + pv * f[e];
  }
}

// uint4 of BF16 values, or zeros where the slot is a pad (c < 0): the
// product pv * 0 is ±0, and adding ±0 leaves a sum unchanged (the sums start
// at +0 and never become -0), so a pad adds nothing, without a branch
__device__ __forceinline__ uint4 keep_if(uint4 a, bool keep) {
  unsigned int mk = keep ? 0xffffffffu : 0u;
  return make_uint4(a.x & mk, a.y & mk, a.z & mk, a.w & mk);
}

// k_inf_input for the large levels, with the layer-1 rows stored four lanes
// to a row (a store instruction writes 16 whole rows).  A lane first forms
// (x, P x) of its own row; lane l then computes features 8 (l & 3) .. +7 of
// row 16 g + (l >> 2) for g = 0..3 (the same operations as k_inf_input).
// rcol/rperm: r's index of each slot's column and of each row (r in another
// numbering than the rows; NULL: pcol and the row itself).  Wavefronts run
// independently after the scale; tiles are dealt to XCDs in contiguous ranges.
template <int W>
__global__ void __launch_bounds__(TB) k_inf_input_c(int n, const int *__restrict__ pcol, const float *__restrict__ pval,
                                                    const int *__restrict__ rcol, const int *__restrict__ rperm,
                                                    const float *__restrict__ r, const double *__restrict__ part,
                                                    float *__restrict__ s, float2 *__restrict__ xp,
                                                    const float *__restrict__ th, unsigned short *__restrict__ h1) {
  __shared__ float s_sh;
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int j = 0; j < kScaleBlocks; j++) acc += part[j];
    double m = acc / n;
    float sv = m > 0 ? static_cast<float>(sqrt(m)) : 1.0f;
    s_sh = sv;
    if (blockIdx.x == 0) s[0] = sv;
  }
  __syncthreads();
  int lane = threadIdx.x & 63;
  int t = dec_ell::xcd_block() * kInfWaves + (threadIdx.x >> 6);
  if (t * 64 >= n) return;
  int v = t * 64 + lane;
  size_t vv = static_cast<size_t>(min(v, n - 1));
  const int *cols = rcol ? rcol : pcol;
  float inv = 1.0f / s_sh;
  int rv = rperm ? rperm[vv] : static_cast<int>(vv);
  int c[W];
  float pv[W], rc[W];
#pragma unroll
  for (int k = 0; k < W; k++) {
    c[k] = cols[k * static_cast<size_t>(n) + vv];
    pv[k] = pval[k * static_cast<size_t>(n) + vv];
  }
  float rown = r[rv];
#pragma unroll
  for (int k = 0; k < W; k++) rc[k] = r[c[k] < 0 ? rv : c[k]];
  float xv = rown * inv;
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < W; k++) {
    float tk = pv[k] * (rc[k] * inv);
    acc = c[k] >= 0 ? acc + tk : acc;
  }
  if (v < n) xp[v] = make_float2(xv, acc);
  const int sub = lane >> 2, q = lane & 3;
  float w0[8], w1[8], b1[8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    w0[e] = th[OW1 + 8 * q + e];
    w1[e] = th[OW1 + H + 8 * q + e];
    b1[e] = th[OB1 + 8 * q + e];
  }
#pragma unroll
  for (int g = 0; g < 4; g++) {
    int src = 16 * g + sub;
    float xg = __shfl(xv, src), ag = __shfl(acc, src);
    float hh[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      float z = xg * w0[e] + ag * w1[e] + b1[e];
      hh[e] = z > 0.0f ? z : 0.0f;
    }
    int row = t * 64 + src;
    if (row < n) *reinterpret_cast<uint4 *>(h1 + static_cast<size_t>(row) * H + 8 * q) = pack8(hh);
  }
}

// One row per lane, for rows of any width: the row's own h (BF16) and
// m = P h (FP32 sums over the neighbours' rows, ascending ELL order, rounded
// to BF16) into its LDS row; rows past the end (live false) get a zero row.
__device__ __forceinline__ void gather_rows(int n, int pw, int v, bool live, const int *__restrict__ pcol,
                                            const float *__restrict__ pval, const unsigned short *__restrict__ h,
                                            unsigned short *my) {
  uint4 *dst = reinterpret_cast<uint4 *>(my);
  if (!live) {
#pragma unroll
    for (int k = 0; k < 8; k++) dst[k] = make_uint4(0u, 0u, 0u, 0u);
    return;
  }
  const uint4 *hv = reinterpret_cast<const uint4 *>(h + static_cast<size_t>(v) * H);
#pragma unroll
  for (int k = 0; k < 4; k++) dst[k] = hv[k];
  float m[H];
#pragma unroll
  for (int j = 0; j < H; j++) m[j] = 0.0f;
  for (int k = 0; k < pw; k++) {
    size_t qq = static_cast<size_t>(k) * n + v;
    int c = pcol[qq];
    if (c < 0) break;
    add_row(reinterpret_cast<const uint4 *>(h + static_cast<size_t>(c) * H), pval[qq], m);
  }
#pragma unroll
  for (int k4 = 0; k4 < 4; k4++) dst[4 + k4] = pack8(m + 8 * k4);
}

// h2 = relu([h1, P h1] W2 + b2) from the stored layer-1 rows, for rows wider
// than the wave-granular kernels take: each lane gathers its row's
// neighbours' h1 (the sums of k_spmm, in slot order).
__global__ void __launch_bounds__(TB) k_inf_layer2s(int n, int pw, const int *__restrict__ pcol,
                                                    const float *__restrict__ pval,
                                                    const unsigned short *__restrict__ h1,
                                                    const float *__restrict__ th, unsigned short *__restrict__ h2) {
  __shared__ __align__(16) unsigned short rows_s[kInfWaves][64 * kRowPitch];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int li = lane & 15, lg = lane >> 4;
  unsigned short *rows = rows_s[w];
  short4_t Bw[2][4];
  load_wB(th + OW2, li, lg, Bw);
  float bias0 = th[OB2 + li], bias1 = th[OB2 + 16 + li];
  int nsteps = (n + 63) / 64;
  for (int base = blockIdx.x * kInfWaves; base < nsteps; base += gridDim.x * kInfWaves) {
    int t = base + w;
    int v = t * 64 + lane;
    bool live = t < nsteps && v < n;
    unsigned short *my = rows + lane * kRowPitch;
    gather_rows(n, pw, v, live, pcol, pval, h1, my);
    __syncthreads();
    float4_t acc[4][2];
    mfma_rows(rows, li, lg, Bw, acc);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        unsigned short *o = rows + (16 * q + 4 * lg + i) * kRowPitch;
        float z0 = acc[q][0][i] + bias0, z1 = acc[q][1][i] + bias1;
        o[li] = f2bf(z0 > 0.0f ? z0 : 0.0f);
        o[16 + li] = f2bf(z1 > 0.0f ? z1 : 0.0f);
      }
    __syncthreads();
    if (live) {
      const uint4 *src = reinterpret_cast<const uint4 *>(my);
      uint4 *dst = reinterpret_cast<uint4 *>(h2 + static_cast<size_t>(v) * H);
#pragma unroll
      for (int k = 0; k < 4; k++) dst[k] = src[k];
    }
    __syncthreads();
  }
}

// h3 = relu([h2, P h2] W3 + b3); y = alpha x + h3 wo + bo (ascending sum over
// the 32 features); u += (s / D) y.  Rows of any width, one per lane.
__global__ void __launch_bounds__(TB) k_inf_layer3(int n, int pw, const int *__restrict__ pcol,
                                                   const float *__restrict__ pval, const unsigned short *__restrict__ h2,
                                                   const float2 *__restrict__ xo, const float *__restrict__ th,
                                                   const float *__restrict__ s, const float *__restrict__ D,
                                                   float *__restrict__ u) {
  __shared__ __align__(16) unsigned short rows_s[kInfWaves][64 * kRowPitch];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int li = lane & 15, lg = lane >> 4;
  unsigned short *rows = rows_s[w];
  float *prod = reinterpret_cast<float *>(rows);  // 64 x (H + 1) floats, after the MFMA reads
  short4_t Bw[2][4];
  load_wB(th + OW3, li, lg, Bw);
  float bias0 = th[OB3 + li], bias1 = th[OB3 + 16 + li];
  float wo0 = th[OWO + li], wo1 = th[OWO + 16 + li];
  float alpha = th[OAL], bo = th[OBO], sc = s[0];
  int nsteps = (n + 63) / 64;
  for (int base = blockIdx.x * kInfWaves; base < nsteps; base += gridDim.x * kInfWaves) {
    int t = base + w;
    int v = t * 64 + lane;
    bool live = t < nsteps && v < n;
    unsigned short *my = rows + lane * kRowPitch;
    float xv = 0.0f, dv = 1.0f, uv = 0.0f;
    if (live) {
      xv = xo[v].x;
      dv = D[v];
      uv = u[v];
    }
    gather_rows(n, pw, v, live, pcol, pval, h2, my);
    __syncthreads();
    float4_t acc[4][2];
    mfma_rows(rows, li, lg, Bw, acc);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        float *o = prod + (16 * q + 4 * lg + i) * (H + 1);
        float z0 = acc[q][0][i] + bias0, z1 = acc[q][1][i] + bias1;
        o[li] = bf2f(f2bf(z0 > 0.0f ? z0 : 0.0f)) * wo0;
        o[16 + li] = bf2f(f2bf(z1 > 0.0f ? z1 : 0.0f)) * wo1;
      }
    __syncthreads();
    if (live) {
      const float *o = prod + lane * (H + 1);
      float hw = 0.0f;
      for (int j = 0; j < H; j++) hw = hw + o[j];
      float yy = alpha * xv + hw + bo;
      u[v] = uv + (sc / dv) * yy;
    }
    __syncthreads();
  }
}

// The wavefront's rows t*64 .. t*64+63 into its LDS rows: the own row h into
// columns 0..31 and m = P h (BF16 of the FP32 sums) into columns 32..63, four
// lanes to a row.  A lane-per-row gather of 64-byte rows makes every load
// instruction touch 64 cache lines (16 bytes of each); here the rows are
// taken 16 at a time, lane l serving row 16g + (l >> 2) and its features
// 8 (l & 3) .. 8 (l & 3) + 7, so one instruction reads 16 whole rows.  A
// group's slots are gathered together (W loads in flight) and summed into 8
// registers.  A row's columns and weights are loaded by the lane that owns
// the row (lane l owns row l; c = -1 pads, and every slot of a row past the
// end) and passed on by cross-lane reads.  Each feature's sum runs over the
// slots in ascending order with the operations of k_spmm.
template <int W>
__device__ __forceinline__ void coop_rows_g(int n, int t, int lane, const int (&c)[W], const float (&pv)[W],
                                            const unsigned short *__restrict__ h, unsigned short *rows) {
  const int sub = lane >> 2, part = lane & 3;
#pragma unroll
  for (int g = 0; g < 4; g++) {
    int r = t * 64 + 16 * g + sub;
    int cg[W];
    float vg[W];
#pragma unroll
    for (int k = 0; k < W; k++) {
      cg[k] = __shfl(c[k], 16 * g + sub);
      vg[k] = __shfl(pv[k], 16 * g + sub);
    }
    uint4 own = *reinterpret_cast<const uint4 *>(h + static_cast<size_t>(min(r, n - 1)) * H + 8 * part);
    uint4 a[W];
#pragma unroll
    for (int k = 0; k < W; k++)
      a[k] = keep_if(*reinterpret_cast<const uint4 *>(h + static_cast<size_t>(cg[k] < 0 ? 0 : cg[k]) * H + 8 * part),
                     cg[k] >= 0);
    float m[8];
#pragma unroll
    for (int e = 0; e < 8; e++) m[e] = 0.0f;
#pragma unroll
    for (int k = 0; k < W; k++) {
      float f[8];
      unpack8(a[k], f);
#pragma unroll
      for (int e = 0; e < 8; e++) m[e] = m[e] + vg[k] * f[e];
    }
    *reinterpret_cast<uint4 *>(rows + (16 * g + sub) * kRowPitch + 8 * part) = keep_if(own, r < n);
    *reinterpret_cast<uint4 *>(rows + (16 * g + sub) * kRowPitch + H + 8 * part) = pack8(m);
  }
}

// this lane's row's slots (v = its row; rows past the end get only pads)
template <int W>
__device__ __forceinline__ void own_slots(int n, int v, const int *__restrict__ pcol, const float *__restrict__ pval,
                                          int (&c)[W], float (&pv)[W]) {
  size_t vv = static_cast<size_t>(min(v, n - 1));
#pragma unroll
  for (int k = 0; k < W; k++) {
    int cq = pcol[k * static_cast<size_t>(n) + vv];
    c[k] = v < n ? cq : -1;
    pv[k] = pval[k * static_cast<size_t>(n) + vv];
  }
}

// Wave-granular forms of the two layers: one 64-row tile per one-wavefront
// workgroup.  The LDS rows are the wavefront's own, so the steps are ordered
// by a wavefront-scope fence (LDS operations of one wavefront complete in
// order) instead of a workgroup barrier, and each 16-row MFMA tile's output
// goes to LDS as soon as it is formed.  Tiles are dealt to XCDs in
// contiguous ranges.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acc[jt] = rows 16q .. 16q+15 of the tile times W (64 x 32)
__device__ __forceinline__ void mfma_tile(const unsigned short *rows, int q, int li, int lg, const short4_t Bw[2][4],
                                          float4_t acc[2]) {
  const unsigned short *r = rows + (16 * q + li) * kRowPitch + 4 * lg;
  short4_t A0 = lds4(r), A1 = lds4(r + 16), A2 = lds4(r + 32), A3 = lds4(r + 48);
#pragma unroll
  for (int jt = 0; jt < 2; jt++) {
    float4_t c = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A0, Bw[jt][0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A1, Bw[jt][1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A2, Bw[jt][2], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A3, Bw[jt][3], c, 0, 0, 0);
    acc[jt] = c;
  }
}

template <int W>
__global__ void __launch_bounds__(64) k_inf_layer2w(int n, const int *__restrict__ pcol, const float *__restrict__ pval,
                                                    const unsigned short *__restrict__ h1,
                                                    const float *__restrict__ th, unsigned short *__restrict__ h2) {
  __shared__ __align__(16) unsigned short rows[64 * kRowPitch];
  int lane = threadIdx.x;
  int li = lane & 15, lg = lane >> 4;
  const int sub = lane >> 2, part = lane & 3;
  int t = dec_ell::xcd_block();
  if (t >= (n + 63) / 64) return;
  short4_t Bw[2][4];
  load_wB(th + OW2, li, lg, Bw);
  float bias0 = th[OB2 + li], bias1 = th[OB2 + 16 + li];
  int c[W];
  float pv[W];
  own_slots<W>(n, t * 64 + lane, pcol, pval, c, pv);
  coop_rows_g<W>(n, t, lane, c, pv, h1, rows);
  wave_sync();
#pragma unroll
  for (int q = 0; q < 4; q++) {
    float4_t acc[2];
    mfma_tile(rows, q, li, lg, Bw, acc);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      unsigned short *o = rows + (16 * q + 4 * lg + i) * kRowPitch;
      float z0 = acc[0][i] + bias0, z1 = acc[1][i] + bias1;
      o[li] = f2bf(z0 > 0.0f ? z0 : 0.0f);
      o[16 + li] = f2bf(z1 > 0.0f ? z1 : 0.0f);
    }
  }
  wave_sync();
#pragma unroll
  for (int g = 0; g < 4; g++) {
    int r = t * 64 + 16 * g + sub;
    if (r < n)
      *reinterpret_cast<uint4 *>(h2 + static_cast<size_t>(r) * H + 8 * part) =
          *reinterpret_cast<const uint4 *>(rows + (16 * g + sub) * kRowPitch + 8 * part);
  }
}

// x of (x, P x) or of x alone
__device__ __forceinline__ float x_of(float2 a) { return a.x; }
__device__ __forceinline__ float x_of(float a) { return a; }

// prod (64 x (H + 1) floats) shares the LDS of rows: tile q's products cover
// bytes below the start of tile q + 1's rows, and are written after tile q's
// rows were read.
template <int W, class XT>
__global__ void __launch_bounds__(64) k_inf_layer3w(int n, const int *__restrict__ pcol, const float *__restrict__ pval,
                                                    const unsigned short *__restrict__ h2,
                                                    const XT *__restrict__ xo, const float *__restrict__ th,
                                                    const float *__restrict__ s, const float *__restrict__ D,
                                                    const int *__restrict__ uperm, float *__restrict__ u) {
  __shared__ __align__(16) unsigned short rows[64 * kRowPitch];
  static_assert(16 * (H + 1) * 4 <= 16 * kRowPitch * 2, "products of a tile overlap the next tile's rows");
  int lane = threadIdx.x;
  int li = lane & 15, lg = lane >> 4;
  float *prod = reinterpret_cast<float *>(rows);
  int t = dec_ell::xcd_block();
  if (t >= (n + 63) / 64) return;
  short4_t Bw[2][4];
  load_wB(th + OW3, li, lg, Bw);
  float bias0 = th[OB3 + li], bias1 = th[OB3 + 16 + li];
  float wo0 = th[OWO + li], wo1 = th[OWO + 16 + li];
  float alpha = th[OAL], bo = th[OBO], sc = s[0];
  int v = t * 64 + lane;
  int vv = min(v, n - 1);
  int c[W];
  float pv[W];
  own_slots<W>(n, v, pcol, pval, c, pv);
  int ui = uperm ? uperm[vv] : vv;
  float xv = x_of(xo[vv]), dv = D[vv], uv = u[ui];
  coop_rows_g<W>(n, t, lane, c, pv, h2, rows);
  wave_sync();
#pragma unroll
  for (int q = 0; q < 4; q++) {
    float4_t acc[2];
    mfma_tile(rows, q, li, lg, Bw, acc);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float *o = prod + (16 * q + 4 * lg + i) * (H + 1);
      float z0 = acc[0][i] + bias0, z1 = acc[1][i] + bias1;
      o[li] = bf2f(f2bf(z0 > 0.0f ? z0 : 0.0f)) * wo0;
      o[16 + li] = bf2f(f2bf(z1 > 0.0f ? z1 : 0.0f)) * wo1;
    }
  }
  wave_sync();
  if (v < n) {
    const float *o = prod + lane * (H + 1);
    float hw = 0.0f;
    for (int j = 0; j < H; j++) hw = hw + o[j];
    float yy = alpha * xv + hw + bo;
    u[ui] = uv + (sc / dv) * yy;
  }
}

// ---- layer 1 of large levels with a staged halo -----------------------------
// Each 64-row tile (one wavefront) copies r of its own rows and of its halo
// rows (hrow, at most kHaloMax) into LDS and reads every neighbour's r from
// there; a slot costs 5 bytes (its weight and a one-byte tile-local index)
// instead of 8.  The sums and their order are those of k_inf_input_c.

// The halo's values of x into LDS: at most kHaloMax entries, 64 a time.
// Lanes past the end load and store the last entry again (the same value to
// the same place), so no load or store sits under a branch: a load under a
// branch makes the wavefront wait for it before going on.
__device__ __forceinline__ void stage_halo_f(int lane, int h0, int nh, const int *__restrict__ hrow,
                                             const float *__restrict__ x, float *dst) {
#pragma unroll
  for (int j = 0; j < 3; j++) {
    int i = min(lane + 64 * j, max(nh - 1, 0));
    dst[i] = x[hrow[h0 + i]];
  }
}

// this lane's row's slots as tile-local indices (255: pad, and every slot of
// a row past the end) and weights
template <int W>
__device__ __forceinline__ void own_lslots(int n, int v, const uint8_t *__restrict__ lidx,
                                           const float *__restrict__ pval, int (&l)[W], float (&pv)[W]) {
  size_t vv = static_cast<size_t>(min(v, n - 1));
#pragma unroll
  for (int k = 0; k < W; k++) {
    int lq = lidx[k * static_cast<size_t>(n) + vv];
    l[k] = v < n ? lq : 255;
    pv[k] = pval[k * static_cast<size_t>(n) + vv];
  }
}

// x = r / s and the layer-1 rows h1 (the values of k_inf_input_c), reading r
// of the tile's own rows and halo from LDS.  rperm / hrow_r: r's index of
// each own row (NULL: the row itself) and of each halo row.
template <int W>
__global__ void __launch_bounds__(TB) k_inf_input_h(int n, const uint8_t *__restrict__ lidx,
                                                    const float *__restrict__ pval, const int *__restrict__ hoff,
                                                    const int *__restrict__ hrow_r, const int *__restrict__ rperm,
                                                    const float *__restrict__ r, const double *__restrict__ part,
                                                    float *__restrict__ s, float *__restrict__ xo,
                                                    const float *__restrict__ th, unsigned short *__restrict__ h1) {
  __shared__ float s_sh;
  __shared__ float rs_s[kInfWaves][64 + kHaloMax];
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int j = 0; j < kScaleBlocks; j++) acc += part[j];
    double m = acc / n;
    float sv = m > 0 ? static_cast<float>(sqrt(m)) : 1.0f;
    s_sh = sv;
    if (blockIdx.x == 0) s[0] = sv;
  }
  __syncthreads();
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int t = dec_ell::xcd_block() * kInfWaves + w;
  if (t * 64 >= n) return;
  float *rs = rs_s[w];
  int v = t * 64 + lane;
  int vv = min(v, n - 1);
  int h0 = hoff[t], nh = hoff[t + 1] - h0;
  float rown = r[rperm ? rperm[vv] : vv];
  int l[W];
  float pv[W];
  own_lslots<W>(n, v, lidx, pval, l, pv);
  stage_halo_f(lane, h0, nh, hrow_r, r, rs + 64);
  rs[lane] = rown;
  wave_sync();
  float inv = 1.0f / s_sh;
  float xv = rown * inv;
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < W; k++) {
    float tk = pv[k] * (rs[l[k] < 255 ? l[k] : lane] * inv);
    acc = l[k] < 255 ? acc + tk : acc;
  }
  if (v < n) xo[v] = xv;
  const int sub = lane >> 2, q = lane & 3;
  float w0[8], w1[8], b1[8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    w0[e] = th[OW1 + 8 * q + e];
    w1[e] = th[OW1 + H + 8 * q + e];
    b1[e] = th[OB1 + 8 * q + e];
  }
#pragma unroll
  for (int g = 0; g < 4; g++) {
    int src = 16 * g + sub;
    float xg = __shfl(xv, src), ag = __shfl(acc, src);
    float hh[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      float z = xg * w0[e] + ag * w1[e] + b1[e];
      hh[e] = z > 0.0f ? z : 0.0f;
    }
    int row = t * 64 + src;
    if (row < n) *reinterpret_cast<uint4 *>(h1 + static_cast<size_t>(row) * H + 8 * q) = pack8(hh);
  }
}

__device__ __forceinline__ int feat(int lg, int q) { return q < 4 ? 4 * lg + q : 16 + 4 * lg + (q - 4); }

__device__ __forceinline__ short4_t pack4(const unsigned short *h) {
  short4_t a;
  a[0] = static_cast<short>(h[0]);
  a[1] = static_cast<short>(h[1]);
  a[2] = static_cast<short>(h[2]);
  a[3] = static_cast<short>(h[3]);
  return a;
}

__device__ __forceinline__ void unpack4(uint2 w, unsigned short *h) {
  h[0] = static_cast<unsigned short>(w.x & 0xffffu);
  h[1] = static_cast<unsigned short>(w.x >> 16);
  h[2] = static_cast<unsigned short>(w.y & 0xffffu);
  h[3] = static_cast<unsigned short>(w.y >> 16);
}

// The same two layers for small levels, 16 rows per wavefront: lane (li, lg)
// owns row r0 + li and the features F(lg) (its MFMA A-operand entries), so a
// row's gathers and layer-1 recomputation are split over four lanes, which
// shortens each wavefront's dependent chain where there are few rows per CU.
// h2 = relu([h1, P h1] W2 + b2), h1 = relu([x, P x] W1 + b1) evaluated per
// neighbour.  All waves of a block run the same number of tile steps.
__global__ void __launch_bounds__(TB) k_inf_layer2_t(int n, int pw, const int *__restrict__ pcol,
                                                   const float *__restrict__ pval, const float2 *__restrict__ xp,
                                                   const float *__restrict__ th, unsigned short *__restrict__ h2) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int li = lane & 15, lg = lane >> 4;
  float w0[8], w1[8], b1[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    int f = feat(lg, q);
    w0[q] = th[OW1 + f];
    w1[q] = th[OW1 + H + f];
    b1[q] = th[OB1 + f];
  }
  short4_t Bw[2][4];
  load_wB(th + OW2, li, lg, Bw);
  float bias0 = th[OB2 + li], bias1 = th[OB2 + 16 + li];
  int ntiles = (n + 15) / 16;
  constexpr int WPB = TB / 64;
  for (int base = blockIdx.x * WPB; base < ntiles; base += gridDim.x * WPB) {
    int t = base + w;
    if (t >= ntiles) continue;
    int r0 = t * 16, v = r0 + li;
    short4_t A[4];
    if (v < n) {
      float2 f = xp[v];
      unsigned short hb[8];
      float m[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        float z = f.x * w0[q] + f.y * w1[q] + b1[q];
        hb[q] = f2bf(z > 0.0f ? z : 0.0f);
        m[q] = 0.0f;
      }
      for (int k = 0; k < pw; k++) {
        size_t qq = static_cast<size_t>(k) * n + v;
        int c = pcol[qq];
        if (c < 0) break;
        float pv = pval[qq];
        float2 fc = xp[c];
#pragma unroll
        for (int q = 0; q < 8; q++) {
          float z = fc.x * w0[q] + fc.y * w1[q] + b1[q];
          m[q] = m[q] + pv * bf2f(f2bf(z > 0.0f ? z : 0.0f));
        }
      }
      unsigned short mb[8];
#pragma unroll
      for (int q = 0; q < 8; q++) mb[q] = f2bf(m[q]);
      A[0] = pack4(hb);
      A[1] = pack4(hb + 4);
      A[2] = pack4(mb);
      A[3] = pack4(mb + 4);
    } else {
      A[0] = A[1] = A[2] = A[3] = short4_t{0, 0, 0, 0};
    }
    float4_t acc0 = float4_t{0.0f, 0.0f, 0.0f, 0.0f}, acc1 = acc0;
#pragma unroll
    for (int ks = 0; ks < 4; ks++) acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[ks], Bw[0][ks], acc0, 0, 0, 0);
#pragma unroll
    for (int ks = 0; ks < 4; ks++) acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[ks], Bw[1][ks], acc1, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int rr = r0 + 4 * lg + i;
      if (rr < n) {
        float z0 = acc0[i] + bias0, z1 = acc1[i] + bias1;
        h2[static_cast<size_t>(rr) * H + li] = f2bf(z0 > 0.0f ? z0 : 0.0f);
        h2[static_cast<size_t>(rr) * H + 16 + li] = f2bf(z1 > 0.0f ? z1 : 0.0f);
      }
    }
  }
}

// h3 = relu([h2, P h2] W3 + b3); y = alpha x + h3 wo + bo (ascending sum over
// the 32 features, staged through LDS); u += (s / D) y.
__global__ void __launch_bounds__(TB) k_inf_layer3_t(int n, int pw, const int *__restrict__ pcol,
                                                   const float *__restrict__ pval, const unsigned short *__restrict__ h2,
                                                   const float2 *__restrict__ xp, const float *__restrict__ th,
                                                   const float *__restrict__ s, const float *__restrict__ D,
                                                   float *__restrict__ u) {
  constexpr int WPB = TB / 64;
  __shared__ float prod[WPB][16][H + 1];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int li = lane & 15, lg = lane >> 4;
  short4_t Bw[2][4];
  load_wB(th + OW3, li, lg, Bw);
  float bias0 = th[OB3 + li], bias1 = th[OB3 + 16 + li];
  float wo0 = th[OWO + li], wo1 = th[OWO + 16 + li];
  float alpha = th[OAL], bo = th[OBO], sc = s[0];
  int ntiles = (n + 15) / 16;
  for (int base = blockIdx.x * WPB; base < ntiles; base += gridDim.x * WPB) {
    int t = base + w;
    bool active = t < ntiles;
    int r0 = t * 16, v = r0 + li;
    if (active) {
      short4_t A[4];
      if (v < n) {
        const unsigned short *hv = h2 + static_cast<size_t>(v) * H;
        unsigned short hb[8];
        unpack4(*reinterpret_cast<const uint2 *>(hv + 4 * lg), hb);
        unpack4(*reinterpret_cast<const uint2 *>(hv + 16 + 4 * lg), hb + 4);
        float m[8];
#pragma unroll
        for (int q = 0; q < 8; q++) m[q] = 0.0f;
        for (int k = 0; k < pw; k++) {
          size_t qq = static_cast<size_t>(k) * n + v;
          int c = pcol[qq];
          if (c < 0) break;
          float pv = pval[qq];
          const unsigned short *hc = h2 + static_cast<size_t>(c) * H;
          unsigned short nb[8];
          unpack4(*reinterpret_cast<const uint2 *>(hc + 4 * lg), nb);
          unpack4(*reinterpret_cast<const uint2 *>(hc + 16 + 4 * lg), nb + 4);
#pragma unroll
          for (int q = 0; q < 8; q++) m[q] = m[q] + pv * bf2f(nb[q]);
        }
        unsigned short mb[8];
#pragma unroll
        for (int q = 0; q < 8; q++) mb[q] = f2bf(m[q]);
        A[0] = pack4(hb);
        A[1] = pack4(hb + 4);
        A[2] = pack4(mb);
        A[3] = pack4(mb + 4);
      } else {
        A[0] = A[1] = A[2] = A[3] = short4_t{0, 0, 0, 0};
      }
      float4_t acc0 = float4_t{0.0f, 0.0f, 0.0f, 0.0f}, acc1 = acc0;
#pragma unroll
      for (int ks = 0; ks < 4; ks++) acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[ks], Bw[0][ks], acc0, 0, 0, 0);
#pragma unroll
      for (int ks = 0; ks < 4; ks++) acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[ks], Bw[1][ks], acc1, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        float z0 = acc0[i] + bias0, z1 = acc1[i] + bias1;
        prod[w][4 * lg + i][li] = bf2f(f2bf(z0 > 0.0f ? z0 : 0.0f)) * wo0;
        prod[w][4 * lg + i][16 + li] = bf2f(f2bf(z1 > 0.0f ? z1 : 0.0f)) * wo1;
      }
    }
    __syncthreads();
    if (active && lane < 16) {
      int row = r0 + lane;
      if (row < n) {
        float hw = 0.0f;
        for (int j = 0; j < H; j++) hw = hw + prod[w][lane][j];
        float yy = alpha * xp[row].x + hw + bo;
        u[row] = u[row] + (sc / D[row]) * yy;
      }
    }
    __syncthreads();
  }
}

// ---- loss and backward ------------------------------------------------------

// rho = x - A (y / D) per sample; dρ = 2 rho / (n B); block partial of Σ rho².
__global__ void k_loss(int n, int B, int aw, const int *__restrict__ acol, const float *__restrict__ aval,
                       const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ D,
                       float *__restrict__ t, float *__restrict__ drho, double *__restrict__ part) {
  __shared__ double sh[TB];
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  size_t o = static_cast<size_t>(b) * n;
  double sq = 0.0;
  if (v < n) {
    // t = y / D is formed by the neighbours' threads; compute A t on the fly
    float acc = 0.0f;
    for (int k = 0; k < aw; k++) {
      size_t q = static_cast<size_t>(k) * n + v;
      int c = acol[q];
      if (c < 0) break;
      acc = acc + aval[q] * (y[o + c] / D[c]);
    }
    float rho = x[o + v] - acc;
    drho[o + v] = 2.0f * rho / (static_cast<float>(n) * B);
    sq = static_cast<double>(rho) * rho;
  }
  sh[threadIdx.x] = sq;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
  (void)t;
}

// ---- error loss: mean (H ε)², ε = e / s - y / D, H = I - Pr R ---------------

// y[b][i] = Σ_k val[k][i] x[b][col[k][i]] for a batch of vectors (rows ascending ELL)
__global__ void k_bspmv(int nrows, int ncols, int width, const int *__restrict__ col, const float *__restrict__ val,
                        const float *__restrict__ x, float *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (i >= nrows) return;
  const float *xb = x + static_cast<size_t>(b) * ncols;
  float acc = 0.0f;
  for (int k = 0; k < width; k++) {
    size_t q = static_cast<size_t>(k) * nrows + i;
    int c = col[q];
    if (c < 0) break;
    acc = acc + val[q] * xb[c];
  }
  y[static_cast<size_t>(b) * nrows + i] = acc;
}

// eps = κ (e / s - y / D), κ = mean(D): the error in residual units
__global__ void k_err(int n, float kap, const float *__restrict__ e, const float *__restrict__ s,
                      const float *__restrict__ y, const float *__restrict__ D, float *__restrict__ eps) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (v >= n) return;
  size_t o = static_cast<size_t>(b) * n + v;
  eps[o] = kap * (e[o] / s[b] - y[o] / D[v]);
}

// eh = eps - t (t = Pr R eps) in place of eps; block partials of Σ eh²
__global__ void k_hp(int n, const float *__restrict__ t, float *__restrict__ eps, double *__restrict__ part) {
  __shared__ double sh[TB];
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  double sq = 0.0;
  if (v < n) {
    size_t o = static_cast<size_t>(b) * n + v;
    float eh = eps[o] - t[o];
    eps[o] = eh;
    sq = static_cast<double>(eh) * eh;
  }
  sh[threadIdx.x] = sq;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

// g = scale (eh - t), t = Rᵀ Prᵀ eh: the gradient Hᵀ H ε scaled by 2 κ / (n B)
__global__ void k_hp_grad(int n, float scale, const float *__restrict__ eh, const float *__restrict__ t,
                          float *__restrict__ g) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (v >= n) return;
  size_t o = static_cast<size_t>(b) * n + v;
  g[o] = scale * (eh[o] - t[o]);
}

// dy = -(Aᵀ dρ) / D (residual loss) or -g / D for a given g (error loss);
// dz3 = (dy wo) ⊙ [h3 > 0]; block partials of
// (Σ dy x, Σ dy, Σ dy h3) for alpha, bo, wo.
template <class AT>
__global__ void k_out_back(int n, int atw, const int *__restrict__ atcol, const float *__restrict__ atval,
                           const float *__restrict__ drho, const float *__restrict__ D, const float *__restrict__ x,
                           const AT *__restrict__ h3, const float *__restrict__ th, AT *__restrict__ dz3,
                           float *__restrict__ part, int direct) {
  __shared__ float sh[TB][2 * H + 2];
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  size_t o = static_cast<size_t>(b) * n;
  float dy = 0.0f, xv = 0.0f;
  float hv[H];
  if (v < n) {
    float acc = 0.0f;
    if (direct) {
      acc = drho[o + v];  // the loss gradient with respect to y / D is given
    } else {
      for (int k = 0; k < atw; k++) {
        size_t q = static_cast<size_t>(k) * n + v;
        int c = atcol[q];
        if (c < 0) break;
        acc = acc + atval[q] * drho[o + c];
      }
    }
    dy = -acc / D[v];
    xv = x[o + v];
    for (int j = 0; j < H; j++) {
      hv[j] = ald(h3, (o + v) * H + j);
      float g = hv[j] > 0.0f ? dy * th[OWO + j] : 0.0f;
      ast(dz3, (o + v) * H + j, g);
      sh[threadIdx.x][H + 2 + j] = ald(dz3, (o + v) * H + j);
    }
  } else {
    for (int j = 0; j < H; j++) {
      hv[j] = 0.0f;
      sh[threadIdx.x][H + 2 + j] = 0.0f;
    }
  }
  for (int j = 0; j < H; j++) sh[threadIdx.x][j] = dy * hv[j];
  sh[threadIdx.x][H] = dy;
  sh[threadIdx.x][H + 1] = dy * xv;
  __syncthreads();
  // column sums in fixed order: [wo (32), bo, alpha, b3 (32)]
  if (threadIdx.x < 2 * H + 2) {
    float acc = 0.0f;
    for (int i = 0; i < TB; i++) acc += sh[i][threadIdx.x];
    part[(static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x) * (2 * H + 2) + threadIdx.x] = acc;
  }
}

// dg = dz Wᵀ (rows x 64) on BF16 MFMA: 16 rows per wavefront, four 16-column
// tiles, K = 32 in two steps.
__global__ void k_dgrad_bf16(int rows, const unsigned short *__restrict__ dz, const float *__restrict__ W,
                             float *__restrict__ dg) {
  int lane = threadIdx.x & 63;
  int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int r0 = wave * 16;
  if (r0 >= rows) return;
  int li = lane & 15, lg = lane >> 4;
  int row = r0 + li;
  short4_t A[2];
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    int k = ks * 16 + 4 * lg;
    short4_t a;
#pragma unroll
    for (int i = 0; i < 4; i++) a[i] = row < rows ? static_cast<short>(dz[static_cast<size_t>(row) * H + k + i]) : 0;
    A[ks] = a;
  }
#pragma unroll
  for (int jt = 0; jt < 4; jt++) {
    float4_t acc = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      short4_t bv;
      // B[k][j] = Wᵀ[k][j] = W[j][k], j = input feature (64), k = output (32)
#pragma unroll
      for (int i = 0; i < 4; i++) bv[i] = static_cast<short>(f2bf(W[(jt * 16 + li) * H + ks * 16 + 4 * lg + i]));
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[ks], bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int rr = r0 + 4 * lg + i;
      if (rr < rows) dg[static_cast<size_t>(rr) * 2 * H + jt * 16 + li] = acc[i];
    }
  }
}

__global__ void k_dgrad_f32(int rows, const float *__restrict__ dz, const float *__restrict__ W, float *__restrict__ dg) {
  size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t row = t >> 6;
  int j = static_cast<int>(t & 63);
  if (row >= static_cast<size_t>(rows)) return;
  float acc = 0.0f;
  for (int k = 0; k < H; k++) acc = acc + dz[row * H + k] * W[j * H + k];
  dg[row * 2 * H + j] = acc;
}

// Weight gradient partials dW = gᵀ dz (64 x 32) over a chunk of 256 rows per
// wavefront on BF16 MFMA: A = gᵀ (16 features x 16 rows), B = dz (16 rows x
// 16 cols); eight accumulator tiles.  Partials [wave][64][32].
__global__ void k_wgrad_bf16(int rows, const unsigned short *__restrict__ h, const unsigned short *__restrict__ m,
                             const unsigned short *__restrict__ dz, float *__restrict__ part) {
  int lane = threadIdx.x & 63;
  int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int base = wave * 256;
  int li = lane & 15, lg = lane >> 4;
  float4_t acc[4][2];
#pragma unroll
  for (int ti = 0; ti < 4; ti++)
#pragma unroll
    for (int tj = 0; tj < 2; tj++) acc[ti][tj] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
  for (int kc = 0; kc < 256 && base + kc < rows; kc += 16) {
    short4_t Bv[2];
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int rr = base + kc + 4 * lg + i;
        Bv[tj][i] = rr < rows ? static_cast<short>(dz[static_cast<size_t>(rr) * H + tj * 16 + li]) : 0;
      }
#pragma unroll
    for (int ti = 0; ti < 4; ti++) {
      short4_t Av;
      int f = ti * 16 + li;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int rr = base + kc + 4 * lg + i;
        unsigned short gv = 0;
        if (rr < rows) gv = f < H ? h[static_cast<size_t>(rr) * H + f] : m[static_cast<size_t>(rr) * H + f - H];
        Av[i] = static_cast<short>(gv);
      }
#pragma unroll
      for (int tj = 0; tj < 2; tj++) acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Av, Bv[tj], acc[ti][tj], 0, 0, 0);
    }
  }
  float *p = part + static_cast<size_t>(wave) * 2 * H * H;
#pragma unroll
  for (int ti = 0; ti < 4; ti++)
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
      for (int i = 0; i < 4; i++) p[(ti * 16 + 4 * lg + i) * H + tj * 16 + li] = acc[ti][tj][i];
}

// FP32 weight-gradient partials, thread per (feature, col) over 256 rows.
__global__ void k_wgrad_f32(int rows, const float *__restrict__ h, const float *__restrict__ m,
                            const float *__restrict__ dz, float *__restrict__ part) {
  int chunk = blockIdx.x;
  for (int e = threadIdx.x; e < 2 * H * H; e += blockDim.x) {
    int f = e / H, j = e % H;
    float acc = 0.0f;
    for (int rr = chunk * 256; rr < chunk * 256 + 256 && rr < rows; rr++) {
      float g = f < H ? h[static_cast<size_t>(rr) * H + f] : m[static_cast<size_t>(rr) * H + f - H];
      acc = acc + g * dz[static_cast<size_t>(rr) * H + j];
    }
    part[static_cast<size_t>(chunk) * 2 * H * H + e] = acc;
  }
}

// dh = dg[:, :32] + Pᵀ dg[:, 32:], dz = dh ⊙ [h > 0]; block partials of Σ dz (bias).
template <class AT>
__global__ void k_back_mix(int n, int ptw, const int *__restrict__ ptcol, const float *__restrict__ ptval,
                           const float *__restrict__ dg, const AT *__restrict__ h, AT *__restrict__ dz,
                           float *__restrict__ part) {
  __shared__ float sh[TB];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int v = t >> 5, f = t & 31;
  int b = blockIdx.y;
  size_t o = static_cast<size_t>(b) * n;
  float g = 0.0f;
  if (v < n) {
    float acc = dg[(o + v) * 2 * H + f];
    for (int k = 0; k < ptw; k++) {
      size_t q = static_cast<size_t>(k) * n + v;
      int c = ptcol[q];
      if (c < 0) break;
      acc = acc + ptval[q] * dg[(o + c) * 2 * H + H + f];
    }
    g = ald(h, (o + v) * H + f) > 0.0f ? acc : 0.0f;
    ast(dz, (o + v) * H + f, g);
  }
  sh[threadIdx.x] = g;
  __syncthreads();
  if (threadIdx.x < H) {
    float acc = 0.0f;
    for (int i = threadIdx.x; i < TB; i += H) acc += sh[i];
    part[(static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x) * H + threadIdx.x] = acc;
  }
}

// dW1 = f0ᵀ dz1 (2 x 32) block partials.
template <class AT>
__global__ void k_w1_back(int n, const float *__restrict__ x, const float *__restrict__ px, const AT *__restrict__ dz1,
                          float *__restrict__ part) {
  __shared__ float sh[TB][2];
  int b = blockIdx.y;
  size_t o = static_cast<size_t>(b) * n;
  int v0 = blockIdx.x * (TB / H);
  // thread (vl, j): vertex v0 + vl, column j
  int vl = threadIdx.x / H, j = threadIdx.x % H;
  float a0 = 0.0f, a1 = 0.0f;
  int v = v0 + vl;
  if (v < n) {
    float d = ald(dz1, (o + v) * H + j);
    a0 = x[o + v] * d;
    a1 = px[o + v] * d;
  }
  sh[threadIdx.x][0] = a0;
  sh[threadIdx.x][1] = a1;
  __syncthreads();
  if (threadIdx.x < 2 * H) {
    int jj = threadIdx.x % H, which = threadIdx.x / H;
    float acc = 0.0f;
    for (int q = 0; q < TB / H; q++) acc += sh[q * H + jj][which];
    part[(static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x) * 2 * H + which * H + jj] = acc;
  }
}

// out[j] = Σ_p part[p * stride + off + j] for j < width, sequential over p.
// Column sums out[j] = scale · Σ_p part[p·stride + off + j] (j < width) of
// per-block partials, in two deterministic passes: workgroup (x, c) sums the
// partials p of chunk c for 64 columns (coalesced over the columns, double
// accumulators), then one thread per column adds the chunk sums in order.
constexpr int kRedChunks = 4096;

template <class PT>
__global__ void k_reduce_chunks(const PT *__restrict__ part, int nparts, int stride, int off, int width, int len,
                                double *__restrict__ tmp) {
  int j = blockIdx.x * 64 + threadIdx.x;
  int c = blockIdx.y;
  if (j >= width) return;
  int p0 = c * len, p1 = min(nparts, p0 + len);
  double acc = 0.0;
  for (int p = p0; p < p1; p++) acc += static_cast<double>(part[static_cast<size_t>(p) * stride + off + j]);
  tmp[static_cast<size_t>(c) * width + j] = acc;
}

template <class OT>
__global__ void k_reduce_final(const double *__restrict__ tmp, int chunks, int width, double scale,
                               OT *__restrict__ out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= width) return;
  double acc = 0.0;
  for (int c = 0; c < chunks; c++) acc += tmp[static_cast<size_t>(c) * width + j];
  out[j] = static_cast<OT>(acc * scale);
}

__global__ void k_adam(int np, float *__restrict__ th, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float gi = g[i];
  float a = b1 * m1[i] + (1 - b1) * gi;
  float v = b2 * m2[i] + (1 - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  th[i] = th[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_gather_rows(int n, const float *__restrict__ src, const int *__restrict__ idx, float *__restrict__ dst) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (v < n) dst[static_cast<size_t>(b) * n + v] = src[static_cast<size_t>(idx[b]) * n + v];
}

}  // namespace

// ---------------------------------------------------------------- host side

namespace dec {

static void to_ell(const HostCsr &A, NnLevel::Op &E) {
  Ell<float> e;
  upload_ell(A, e);
  E.n = e.nrows;
  E.width = e.width;
  E.col = std::move(e.col);
  E.val = std::move(e.val);
}

static HostCsr transpose(const HostCsr &A) {
  HostCsr T;
  T.m = A.n;
  T.n = A.m;
  T.ptr.assign(A.n + 1, 0);
  for (int c : A.col) T.ptr[c + 1]++;
  for (int i = 0; i < A.n; i++) T.ptr[i + 1] += T.ptr[i];
  T.col.resize(A.col.size());
  T.val.resize(A.val.size());
  std::vector<int32_t> w(T.ptr.begin(), T.ptr.end() - 1);
  for (int r = 0; r < A.m; r++)
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
      int q = w[A.col[k]]++;
      T.col[q] = r;
      T.val[q] = A.val[k];
    }
  return T;
}

// The halo lists of the 64-row tiles of Q (rows and columns in one
// numbering; perm maps it to vertex numbers, NULL: the identity).
static void build_halos(const HostCsr &Q, const std::vector<int32_t> *perm, NnLevel &L) {
  int n = Q.m, w = 0;
  for (int i = 0; i < n; i++) w = std::max(w, Q.ptr[i + 1] - Q.ptr[i]);
  if (w < 1 || w > 8) return;
  int tiles = (n + 63) / 64;
  std::vector<int32_t> hoff(tiles + 1, 0), hrow, hrow_v, loc(n, -1), h;
  std::vector<uint8_t> lid(static_cast<size_t>(w) * n, 255);
  int hmax = 0;
  for (int t = 0; t < tiles; t++) {
    int r0 = 64 * t, r1 = std::min(n, r0 + 64);
    h.clear();
    for (int i = r0; i < r1; i++)
      for (int k = Q.ptr[i]; k < Q.ptr[i + 1]; k++)
        if (Q.col[k] < r0 || Q.col[k] >= r1) h.push_back(Q.col[k]);
    std::sort(h.begin(), h.end());
    h.erase(std::unique(h.begin(), h.end()), h.end());
    if (h.size() > kHaloMax) return;
    hmax = std::max(hmax, static_cast<int>(h.size()));
    for (size_t j = 0; j < h.size(); j++) loc[h[j]] = 64 + static_cast<int>(j);
    for (int i = r0; i < r1; i++)
      for (int k = Q.ptr[i]; k < Q.ptr[i + 1]; k++) {
        int c = Q.col[k];
        lid[static_cast<size_t>(k - Q.ptr[i]) * n + i] = static_cast<uint8_t>(c >= r0 && c < r1 ? c - r0 : loc[c]);
      }
    for (int c : h) {
      loc[c] = -1;
      hrow.push_back(c);
      hrow_v.push_back(perm ? (*perm)[c] : c);
    }
    hoff[t + 1] = static_cast<int32_t>(hrow.size());
  }
  // one more entry: the last tile's (clamped) loads stay in bounds
  hrow.push_back(0);
  hrow_v.push_back(0);
  L.lidx.upload(lid);
  L.hoff.upload(hoff);
  L.hrow.upload(hrow);
  L.hrow_v.upload(hrow_v);
  L.hmax = hmax;
}

NnLevel::NnLevel(const HostCsr &A, const HostCsr *R, const HostCsr *Pr, const std::vector<int32_t> *perm) {
  n = A.m;
  HostCsr P;
  P.m = P.n = n;
  P.ptr.assign(n + 1, 0);
  std::vector<float> d(n, 0.0f);
  for (int r = 0; r < n; r++) {
    double dd = 0.0;
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
      if (A.col[k] == r) dd = A.val[k];
    d[r] = static_cast<float>(dd);
    for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
      if (A.col[k] != r) {
        P.col.push_back(A.col[k]);
        P.val.push_back(-A.val[k] / dd);
      }
    P.ptr[r + 1] = static_cast<int32_t>(P.col.size());
  }
  D.upload(d);
  double dsum = 0.0;
  for (int r = 0; r < n; r++) dsum += d[r];
  dmean = n ? static_cast<float>(dsum / n) : 1.0f;
  to_ell(A, a);
  to_ell(transpose(A), at);
  to_ell(P, p);
  {
    // P's rows renumbered (rows perm[i], columns through the inverse), slots
    // in P's order, so each row's sum is the vertex-order row's
    std::vector<int32_t> iperm(n);
    for (int i = 0; i < n; i++) iperm[perm ? (*perm)[i] : i] = i;
    HostCsr Q, Qv;
    Q.m = Q.n = n;
    Q.ptr.assign(n + 1, 0);
    std::vector<float> di(n);
    for (int i = 0; i < n; i++) {
      int v = perm ? (*perm)[i] : i;
      di[i] = d[v];
      for (int k = P.ptr[v]; k < P.ptr[v + 1]; k++) {
        Q.col.push_back(iperm[P.col[k]]);
        Q.val.push_back(P.val[k]);
      }
      Q.ptr[i + 1] = static_cast<int32_t>(Q.col.size());
    }
    to_ell(Q, pi);
    Di.upload(di);
    if (perm) {
      perm_d.upload(*perm);
      Qv = Q;
      for (int i = 0; i < n; i++)
        for (int k = Q.ptr[i]; k < Q.ptr[i + 1]; k++) Qv.col[k] = (*perm)[Q.col[k]];
      Op tmp;
      to_ell(Qv, tmp);
      pi_vcol = std::move(tmp.col);
    }
    if (n > kInfSmallRows) build_halos(Q, perm, *this);
  }
  to_ell(transpose(P), pt);
  if (R && Pr) {
    nc = R->m;
    to_ell(*R, r);
    to_ell(transpose(*R), rt);
    to_ell(*Pr, pr);
    to_ell(transpose(*Pr), prt);
  }
}

}  // namespace dec

struct dec_nn_s {
  std::vector<std::unique_ptr<dec::NnLevel>> levels;  // one per hierarchy level (finest first)
  DBuf<float> th, grad, m1, m2;
  int adam_t = 0;
  // work buffers (grown on demand)
  DBuf<float> x, px, y, c, s, drho, dg, part, r, eps, tmp, crs;
  DBuf<double> dpart, dloss, spart, rtmp;
  DBuf<float2> ix;                                 // inference: (x, P x)
  DBuf<unsigned short> ih1, ih2;                   // inference: layer-1 and layer-2 rows
  DBuf<unsigned char> h1, hm1, h2, hm2, h3, dz;  // activation storage (bf16 or f32)
  DBuf<int32_t> idx;
  size_t cap_rows = 0;
  int cap_prec = -1;
};

namespace {

void ensure(dec_nn_s &nn, size_t rows, int precise) {
  if (rows <= nn.cap_rows && precise == nn.cap_prec) return;
  size_t as = precise ? 4 : 2;
  nn.x.alloc(rows);
  nn.px.alloc(rows);
  nn.y.alloc(rows);
  nn.c.alloc(rows);
  nn.drho.alloc(rows);
  nn.eps.alloc(rows);
  nn.tmp.alloc(rows);
  nn.crs.alloc(rows);
  nn.dg.alloc(rows * 2 * H);
  for (auto *b : {&nn.h1, &nn.hm1, &nn.h2, &nn.hm2, &nn.h3, &nn.dz}) b->alloc(rows * H * as);
  size_t parts = (rows + 255) / 256 + 64;
  nn.part.alloc(parts * 2 * H * H);
  nn.dpart.alloc(parts + 64);

  nn.cap_rows = rows;
  nn.cap_prec = precise;
}

// out[j] = scale · Σ_p part[p·stride + off + j], j < width (kernels above)
template <class PT, class OT>
void reduce_parts(dec_nn_s &nn, const PT *part, int nparts, int stride, int off, int width, double scale, OT *out,
                  hipStream_t st) {
  // about sqrt(nparts) partials per thread in each of the two passes
  int chunks = std::max(1, std::min(kRedChunks, static_cast<int>(std::ceil(std::sqrt(static_cast<double>(nparts))))));
  int len = (nparts + chunks - 1) / chunks;
  chunks = (nparts + len - 1) / len;
  size_t need = static_cast<size_t>(chunks) * width;
  if (nn.rtmp.n < need) nn.rtmp.alloc(need);
  hipLaunchKernelGGL((k_reduce_chunks<PT>), dim3((width + 63) / 64, chunks), dim3(64), 0, st, part, nparts, stride,
                     off, width, len, nn.rtmp.p);
  hipLaunchKernelGGL((k_reduce_final<OT>), dim3((width + 63) / 64), dim3(64), 0, st, nn.rtmp.p, chunks, width, scale,
                     out);
}

template <class AT>
struct Acts {
  AT *h1, *m1, *h2, *m2, *h3, *dz;
};
template <class AT>
Acts<AT> acts(dec_nn_s &nn) {
  return Acts<AT>{reinterpret_cast<AT *>(nn.h1.p), reinterpret_cast<AT *>(nn.hm1.p), reinterpret_cast<AT *>(nn.h2.p),
                  reinterpret_cast<AT *>(nn.hm2.p), reinterpret_cast<AT *>(nn.h3.p), reinterpret_cast<AT *>(nn.dz.p)};
}

template <class AT>
void dense(int R, const AT *h, const AT *m, const float *W, const float *b, AT *out, hipStream_t st) {
  if constexpr (sizeof(AT) == 2) {
    int waves = (R + 15) / 16;
    hipLaunchKernelGGL(k_dense_bf16, dim3((waves + 3) / 4), dim3(TB), 0, st, R, h, m, W, b, out);
  } else {
    hipLaunchKernelGGL(k_dense_f32, dim3(dec::nblocks(static_cast<long long>(R) * H, TB)), dim3(TB), 0, st, R, h, m,
                       W, b, out);
  }
}

void scale(dec_nn_s &nn, int n, int B, const float *r, hipStream_t st) {
  if (nn.s.n < static_cast<size_t>(B)) nn.s.alloc(B);
  if (nn.spart.n < static_cast<size_t>(B) * kScaleBlocks) nn.spart.alloc(static_cast<size_t>(B) * kScaleBlocks);
  hipLaunchKernelGGL(k_scale_partial, dim3(kScaleBlocks, B), dim3(TB), 0, st, n, r, nn.spart.p);
  hipLaunchKernelGGL(k_scale_final, dim3(dec::nblocks(B, 64)), dim3(64), 0, st, n, B, nn.spart.p, nn.s.p);
}

template <class AT>
void forward_t(dec_nn_s &nn, const dec::NnLevel &L, int B, const float *r, float *c_out, float *u_add,
               hipStream_t st) {
  int n = L.n;
  size_t rows = static_cast<size_t>(n) * B;
  ensure(nn, rows, sizeof(AT) == 4);
  Acts<AT> A = acts<AT>(nn);
  const float *th = nn.th.p;
  scale(nn, n, B, r, st);
  dim3 gv(dec::nblocks(n, TB), B), gf(dec::nblocks(static_cast<long long>(n) * H, TB), B);
  hipLaunchKernelGGL(k_input<AT>, gv, dim3(TB), 0, st, n, L.p.width, L.p.col.p, L.p.val.p, r, nn.s.p, th, nn.x.p,
                     nn.px.p, A.h1);
  int R = static_cast<int>(rows);
  hipLaunchKernelGGL(k_spmm<AT>, gf, dim3(TB), 0, st, n, L.p.width, L.p.col.p, L.p.val.p, A.h1, A.m1);
  dense<AT>(R, A.h1, A.m1, th + OW2, th + OB2, A.h2, st);
  hipLaunchKernelGGL(k_spmm<AT>, gf, dim3(TB), 0, st, n, L.p.width, L.p.col.p, L.p.val.p, A.h2, A.m2);
  dense<AT>(R, A.h2, A.m2, th + OW3, th + OB3, A.h3, st);
  hipLaunchKernelGGL(k_output<AT>, gv, dim3(TB), 0, st, n, nn.x.p, A.h3, th, nn.s.p, L.D.p, nn.y.p, c_out, u_add);
}

// Loss (into *loss_out, device) and gradient (into nn.grad) of a batch of B
// residuals stacked in r.
template <class AT>
// e: the B errors paired with the residuals (error loss), or NULL (residual loss).
void loss_grad_t(dec_nn_s &nn, const dec::NnLevel &L, int B, const float *r, const float *e, double *loss_out,
                 hipStream_t st) {
  forward_t<AT>(nn, L, B, r, nullptr, nullptr, st);
  int n = L.n;
  int R = n * B;
  Acts<AT> A = acts<AT>(nn);
  const float *th = nn.th.p;
  float *g = nn.grad.p;
  dim3 gv(dec::nblocks(n, TB), B), gf(dec::nblocks(static_cast<long long>(n) * H, TB), B);
  int nbv = gv.x * B, nbf = gf.x * B;
  if (!e) {
    hipLaunchKernelGGL(k_loss, gv, dim3(TB), 0, st, n, B, L.a.width, L.a.col.p, L.a.val.p, nn.x.p, nn.y.p, L.D.p,
                       nullptr, nn.drho.p, nn.dpart.p);
  } else {
    if (L.nc == 0) fail(DEC_ERR_ARG, "learned smoother: the error loss needs a coarser level below this one");
    dim3 gc(dec::nblocks(L.nc, TB), B);
    hipLaunchKernelGGL(k_err, gv, dim3(TB), 0, st, n, L.dmean, e, nn.s.p, nn.y.p, L.D.p, nn.eps.p);
    hipLaunchKernelGGL(k_bspmv, gc, dim3(TB), 0, st, L.nc, n, L.r.width, L.r.col.p, L.r.val.p, nn.eps.p, nn.crs.p);
    hipLaunchKernelGGL(k_bspmv, gv, dim3(TB), 0, st, n, L.nc, L.pr.width, L.pr.col.p, L.pr.val.p, nn.crs.p,
                       nn.tmp.p);
    hipLaunchKernelGGL(k_hp, gv, dim3(TB), 0, st, n, nn.tmp.p, nn.eps.p, nn.dpart.p);
    hipLaunchKernelGGL(k_bspmv, gc, dim3(TB), 0, st, L.nc, n, L.prt.width, L.prt.col.p, L.prt.val.p, nn.eps.p,
                       nn.crs.p);
    hipLaunchKernelGGL(k_bspmv, gv, dim3(TB), 0, st, n, L.nc, L.rt.width, L.rt.col.p, L.rt.val.p, nn.crs.p, nn.tmp.p);
    hipLaunchKernelGGL(k_hp_grad, gv, dim3(TB), 0, st, n, 2.0f * L.dmean / (static_cast<float>(n) * B), nn.eps.p,
                       nn.tmp.p, nn.drho.p);
  }
  reduce_parts(nn, nn.dpart.p, nbv, 1, 0, 1, 1.0 / (static_cast<double>(n) * B), loss_out, st);
  // output head: dz3, and the (wo, bo, alpha, b3) partials
  hipLaunchKernelGGL(k_out_back<AT>, gv, dim3(TB), 0, st, n, L.at.width, L.at.col.p, L.at.val.p, nn.drho.p, L.D.p,
                     nn.x.p, A.h3, th, A.dz, nn.part.p, e ? 1 : 0);
  reduce_parts(nn, nn.part.p, nbv, 2 * H + 2, 0, H + 2, 1.0, g + OWO, st);
  reduce_parts(nn, nn.part.p, nbv, 2 * H + 2, H + 2, H, 1.0, g + OB3, st);
  // dense layers 3 and 2: weight gradient, input gradient, mix through Pᵀ
  auto layer = [&](const AT *gh, const AT *gm, int ow, const AT *h_below, int ob_below) {
    if constexpr (sizeof(AT) == 2) {
      int waves = (R + 255) / 256;
      hipLaunchKernelGGL(k_wgrad_bf16, dim3((waves + 3) / 4), dim3(TB), 0, st, R, gh, gm, A.dz, nn.part.p);
      reduce_parts(nn, nn.part.p, waves, 2 * H * H, 0, 2 * H * H, 1.0, g + ow, st);
      int dw = (R + 15) / 16;
      hipLaunchKernelGGL(k_dgrad_bf16, dim3((dw + 3) / 4), dim3(TB), 0, st, R, A.dz, th + ow, nn.dg.p);
    } else {
      int chunks = (R + 255) / 256;
      hipLaunchKernelGGL(k_wgrad_f32, dim3(chunks), dim3(TB), 0, st, R, gh, gm, A.dz, nn.part.p);
      reduce_parts(nn, nn.part.p, chunks, 2 * H * H, 0, 2 * H * H, 1.0, g + ow, st);
      hipLaunchKernelGGL(k_dgrad_f32, dim3(dec::nblocks(static_cast<long long>(R) * 2 * H, TB)), dim3(TB), 0, st, R,
                         A.dz, th + ow, nn.dg.p);
    }
    // dz of the layer below (overwrites dz) and its bias partials
    hipLaunchKernelGGL(k_back_mix<AT>, gf, dim3(TB), 0, st, n, L.pt.width, L.pt.col.p, L.pt.val.p, nn.dg.p, h_below,
                       A.dz, nn.part.p);
    reduce_parts(nn, nn.part.p, nbf, H, 0, H, 1.0, g + ob_below, st);
  };
  layer(A.h2, A.m2, OW3, A.h2, OB2);
  layer(A.h1, A.m1, OW2, A.h1, OB1);
  // first layer: dW1 = f0ᵀ dz1
  dim3 g1(dec::nblocks(n, TB / H), B);
  hipLaunchKernelGGL(k_w1_back<AT>, g1, dim3(TB), 0, st, n, nn.x.p, nn.px.p, A.dz, nn.part.p);
  reduce_parts(nn, nn.part.p, static_cast<int>(g1.x) * B, 2 * H, 0, 2 * H, 1.0, g + OW1, st);
}

void lcheck(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(DEC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

const dec::NnLevel &level_of(dec_nn_s &nn, int lvl) {
  if (lvl < 0 || lvl >= static_cast<int>(nn.levels.size())) fail(DEC_ERR_ARG, "learned smoother: bad level");
  return *nn.levels[lvl];
}

}  // namespace

namespace dec {

void reserve_inference(dec_nn_s &nn, int n) {
  if (nn.ix.n < static_cast<size_t>(n)) nn.ix.alloc(n);
  if (nn.ih1.n < static_cast<size_t>(n) * H) nn.ih1.alloc(static_cast<size_t>(n) * H);
  if (nn.ih2.n < static_cast<size_t>(n) * H) nn.ih2.alloc(static_cast<size_t>(n) * H);
  if (nn.s.n < 1) nn.s.alloc(1);
  if (nn.spart.n < static_cast<size_t>(kScaleBlocks)) nn.spart.alloc(kScaleBlocks);
}

void nn_bind(dec_nn_s &nn, const std::vector<int> &sizes) {
  if (sizes.size() != nn.levels.size())
    fail(DEC_ERR_ARG, "learned smoother: the network was created for another hierarchy (level count)");
  int nmax = 0;
  for (size_t l = 0; l < sizes.size(); l++) {
    if (sizes[l] != nn.levels[l]->n)
      fail(DEC_ERR_ARG, "learned smoother: the network was created for another hierarchy (level sizes)");
    nmax = std::max(nmax, sizes[l]);
  }
  reserve_inference(nn, nmax);
}

void nn_smooth(dec_nn_s &nn, int lvl, const float *r, float *u, hipStream_t s, bool internal) {
  const NnLevel &NL = level_of(nn, lvl);
  int n = NL.n;
  reserve_inference(nn, n);
  // the scale's partial sums (over r in the caller's order); the next kernel
  // adds them (a last-workgroup ticket would need a device-scope fence per
  // workgroup, which writes back the XCD's L2)
  hipLaunchKernelGGL(k_scale_partial, dim3(kScaleBlocks, 1), dim3(TB), 0, s, n, r, nn.spart.p);
  if (n > kInfSmallRows && NL.pi.width >= 1 && NL.pi.width <= 8) {
    // Large levels run over the multigrid level's own numbering, whose BFS
    // blocks keep a row's neighbours close together (the vertex order of a
    // subdivided mesh puts a third of them a level's size away).  A caller in
    // vertex order has r read and u updated through perm; every row's sums
    // are those of vertex order (same slots, same order).
    bool vorder = !internal && NL.perm_d.n;
    const int *rcol = vorder ? NL.pi_vcol.p : nullptr;
    const int *rperm = vorder ? NL.perm_d.p : nullptr;
    int tiles = (n + 63) / 64;
    dim3 tg((tiles + 7) / 8 * 8);
    if (NL.hmax >= 0) {
      float *x = reinterpret_cast<float *>(nn.ix.p);
      DEC_ELL_WIDTH_SWITCH(NL.pi.width, if constexpr (W > 0) {
        hipLaunchKernelGGL((k_inf_input_h<W>), dim3(dec::xcd_grid(n, TB)), dim3(TB), 0, s, n, NL.lidx.p,
                           NL.pi.val.p, NL.hoff.p, vorder ? NL.hrow_v.p : NL.hrow.p, rperm, r, nn.spart.p, nn.s.p,
                           x, nn.th.p, nn.ih1.p);
        hipLaunchKernelGGL((k_inf_layer2w<W>), tg, dim3(64), 0, s, n, NL.pi.col.p, NL.pi.val.p, nn.ih1.p, nn.th.p,
                           nn.ih2.p);
        hipLaunchKernelGGL((k_inf_layer3w<W, float>), tg, dim3(64), 0, s, n, NL.pi.col.p, NL.pi.val.p, nn.ih2.p, x,
                           nn.th.p, nn.s.p, NL.Di.p, rperm, u);
      });
      return;
    }
    DEC_ELL_WIDTH_SWITCH(NL.pi.width, if constexpr (W > 0) {
      hipLaunchKernelGGL((k_inf_input_c<W>), dim3(dec::xcd_grid(n, TB)), dim3(TB), 0, s, n, NL.pi.col.p,
                         NL.pi.val.p, rcol, rperm, r, nn.spart.p, nn.s.p, nn.ix.p, nn.th.p, nn.ih1.p);
      hipLaunchKernelGGL((k_inf_layer2w<W>), tg, dim3(64), 0, s, n, NL.pi.col.p, NL.pi.val.p, nn.ih1.p, nn.th.p,
                         nn.ih2.p);
      hipLaunchKernelGGL((k_inf_layer3w<W, float2>), tg, dim3(64), 0, s, n, NL.pi.col.p, NL.pi.val.p, nn.ih2.p, nn.ix.p,
                         nn.th.p, nn.s.p, NL.Di.p, rperm, u);
    });
    return;
  }
  struct {
    const NnLevel::Op &p;
    const DBuf<float> &D;
  } L{internal ? NL.pi : NL.p, internal ? NL.Di : NL.D};
  bool big = n > kInfSmallRows;
  DEC_ELL_WIDTH_SWITCH(L.p.width, hipLaunchKernelGGL((k_inf_input<W>), dim3(dec::nblocks(n, TB)), dim3(TB), 0, s, n,
                                                      L.p.width, L.p.col.p, L.p.val.p, r, nn.spart.p, nn.s.p,
                                                      nn.ix.p, nn.th.p, big ? nn.ih1.p : nullptr));
  if (!big) {
    int tiles = (n + 15) / 16;
    int blocks = std::max(1, (tiles + kInfWaves - 1) / kInfWaves);
    hipLaunchKernelGGL(k_inf_layer2_t, dim3(blocks), dim3(TB), 0, s, n, L.p.width, L.p.col.p, L.p.val.p, nn.ix.p,
                       nn.th.p, nn.ih2.p);
    hipLaunchKernelGGL(k_inf_layer3_t, dim3(blocks), dim3(TB), 0, s, n, L.p.width, L.p.col.p, L.p.val.p, nn.ih2.p,
                       nn.ix.p, nn.th.p, nn.s.p, L.D.p, u);
  } else {
    // rows wider than 8 slots: one lane per row
    int steps = (n + 63) / 64;
    int blocks = std::max(1, (steps + kInfWaves - 1) / kInfWaves);
    hipLaunchKernelGGL(k_inf_layer2s, dim3(blocks), dim3(TB), 0, s, n, L.p.width, L.p.col.p, L.p.val.p, nn.ih1.p,
                       nn.th.p, nn.ih2.p);
    hipLaunchKernelGGL(k_inf_layer3, dim3(blocks), dim3(TB), 0, s, n, L.p.width, L.p.col.p, L.p.val.p, nn.ih2.p,
                       nn.ix.p, nn.th.p, nn.s.p, L.D.p, u);
  }
}

dec_nn_s *nn_create(const std::vector<NnOps> &ops, uint64_t seed) {
  auto nn = std::make_unique<dec_nn_s>();
  for (const NnOps &o : ops) nn->levels.push_back(std::make_unique<NnLevel>(*o.A, o.R, o.P, o.perm));
  // initialisation: W ~ N(0, 2/fan_in) (seeded), biases and output head zero,
  // alpha = 2/3: the untrained network is damped Jacobi
  std::vector<float> th(NP, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int i = 0; i < 2 * H; i++) th[OW1 + i] = static_cast<float>(nd(rng) * std::sqrt(2.0 / 2));
  for (int i = 0; i < 2 * H * H; i++) th[OW2 + i] = static_cast<float>(nd(rng) * std::sqrt(2.0 / (2 * H)));
  for (int i = 0; i < 2 * H * H; i++) th[OW3 + i] = static_cast<float>(nd(rng) * std::sqrt(2.0 / (2 * H)));
  th[OAL] = 2.0f / 3.0f;
  nn->th.upload(th);
  std::vector<float> z(NP, 0.0f);
  nn->grad.upload(z);
  nn->m1.upload(z);
  nn->m2.upload(z);
  nn->dloss.alloc(1);
  return nn.release();
}

}  // namespace dec

extern "C" {

int dec_nn_nparams(dec_nn nn, int32_t *n) {
  DEC_TRY
  if (!nn || !n) fail(DEC_ERR_ARG, "dec_nn_nparams: NULL argument");
  *n = NP;
  DEC_CATCH
}

int dec_nn_get_params(dec_nn nn, float *host) {
  DEC_TRY
  if (!nn || !host) fail(DEC_ERR_ARG, "dec_nn_get_params: NULL argument");
  dec::d2h(host, nn->th.p, sizeof(float) * NP);
  DEC_CATCH
}

int dec_nn_set_params(dec_nn nn, const float *host) {
  DEC_TRY
  if (!nn || !host) fail(DEC_ERR_ARG, "dec_nn_set_params: NULL argument");
  dec::h2d(nn->th.p, host, sizeof(float) * NP);
  nn->adam_t = 0;
  std::vector<float> z(NP, 0.0f);
  dec::h2d(nn->m1.p, z.data(), sizeof(float) * NP);
  dec::h2d(nn->m2.p, z.data(), sizeof(float) * NP);
  DEC_CATCH
}

int dec_nn_forward(dec_nn nn, int32_t level, int32_t nsamples, const float *r, float *c, int32_t precise,
                   void *stream) {
  DEC_TRY
  if (!nn || !r || !c || nsamples <= 0) fail(DEC_ERR_ARG, "dec_nn_forward: bad argument");
  const dec::NnLevel &L = level_of(*nn, level);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precise) forward_t<float>(*nn, L, nsamples, r, c, nullptr, s);
  else forward_t<unsigned short>(*nn, L, nsamples, r, c, nullptr, s);
  dec::check_launch("dec_nn_forward");
  DEC_CATCH
}

int dec_nn_smooth(dec_nn nn, int32_t level, const float *r, float *u, void *stream) {
  DEC_TRY
  if (!nn || !r || !u) fail(DEC_ERR_ARG, "dec_nn_smooth: bad argument");
  dec::nn_smooth(*nn, level, r, u, static_cast<hipStream_t>(stream), false);
  dec::check_launch("dec_nn_smooth");
  DEC_CATCH
}

int dec_nn_loss_grad(dec_nn nn, int32_t level, int32_t nsamples, const float *r, const float *e, int32_t precise,
                     double *loss, float *grad, void *stream) {
  DEC_TRY
  if (!nn || !r || nsamples <= 0) fail(DEC_ERR_ARG, "dec_nn_loss_grad: bad argument");
  const dec::NnLevel &L = level_of(*nn, level);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precise) loss_grad_t<float>(*nn, L, nsamples, r, e, nn->dloss.p, s);
  else loss_grad_t<unsigned short>(*nn, L, nsamples, r, e, nn->dloss.p, s);
  dec::check_launch("dec_nn_loss_grad");
  lcheck(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (loss) dec::d2h(loss, nn->dloss.p, sizeof(double));
  if (grad) dec::d2h(grad, nn->grad.p, sizeof(float) * NP);
  DEC_CATCH
}

int dec_nn_train(dec_nn nn, int32_t level, int32_t npairs, const float *R, const float *E, int32_t steps,
                 int32_t batch, float lr, uint64_t seed, double *loss_hist, void *stream) {
  DEC_TRY
  if (!nn || !R || npairs <= 0 || steps < 0 || batch <= 0) fail(DEC_ERR_ARG, "dec_nn_train: bad argument");
  const dec::NnLevel &L = level_of(*nn, level);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int n = L.n;
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int> pick(0, npairs - 1);
  std::vector<int32_t> ids(static_cast<size_t>(steps) * batch);
  for (auto &i : ids) i = pick(rng);
  DBuf<int32_t> dids;
  dids.upload(ids);
  DBuf<float> rb, eb;
  rb.alloc(static_cast<size_t>(n) * batch);
  if (E) eb.alloc(static_cast<size_t>(n) * batch);
  DBuf<double> hist;
  hist.alloc(std::max(steps, 1));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int k = 0; k < steps; k++) {
    hipLaunchKernelGGL(k_gather_rows, dim3(dec::nblocks(n, TB), batch), dim3(TB), 0, s, n, R,
                       dids.p + static_cast<size_t>(k) * batch, rb.p);
    if (E)
      hipLaunchKernelGGL(k_gather_rows, dim3(dec::nblocks(n, TB), batch), dim3(TB), 0, s, n, E,
                         dids.p + static_cast<size_t>(k) * batch, eb.p);
    loss_grad_t<unsigned short>(*nn, L, batch, rb.p, E ? eb.p : nullptr, hist.p + k, s);
    nn->adam_t++;
    float c1 = 1.0f - std::pow(b1, static_cast<float>(nn->adam_t));
    float c2 = 1.0f - std::pow(b2, static_cast<float>(nn->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(dec::nblocks(NP, TB)), dim3(TB), 0, s, NP, nn->th.p, nn->grad.p, nn->m1.p,
                       nn->m2.p, lr, b1, b2, eps, c1, c2);
  }
  dec::check_launch("dec_nn_train");
  lcheck(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (loss_hist && steps > 0) dec::d2h(loss_hist, hist.p, sizeof(double) * steps);
  DEC_CATCH
}

int dec_nn_destroy(dec_nn nn) {
  DEC_TRY
  delete nn;
  DEC_CATCH
}

}  // extern "C"

