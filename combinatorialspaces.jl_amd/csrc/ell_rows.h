// Row sums over a slot-major ELL (entry k of row r at [k * n + r], col < 0
// pads), in ascending column order (the order of a CSC SpMV in SparseArrays).
// The width is a template parameter when it is small: all loads of a row are
// issued before the sum, which keeps many gathers in flight per wavefront; the
// padding is predicated, not branched on.  W == 0 selects the runtime-width loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dec_ell {

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i % 8), and each XCD has its own L2.  Remapping the workgroup index so that
// XCD k processes the k-th contiguous eighth of the rows keeps the gathers of
// neighbouring rows in the same L2.  The grid must be a multiple of 8 (see
// xcd_grid); workgroups past the end find no rows.
__device__ __forceinline__ int xcd_block() {
  int nb = gridDim.x;
  return static_cast<int>((blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3));
}

template <class V, class T>
__device__ __forceinline__ T mul(V v, T x) {
  return static_cast<T>(v) * x;
}
template <>
__device__ __forceinline__ double mul<int8_t, double>(int8_t v, double x) {
  return v > 0 ? x : -x;
}
template <>
__device__ __forceinline__ float mul<int8_t, float>(int8_t v, float x) {
  return v > 0 ? x : -x;
}

template <int W, class T, class V>
__device__ __forceinline__ T row(int n, int width, int r, const int *__restrict__ col, const V *__restrict__ val,
                                 const T *__restrict__ x) {
  T acc = static_cast<T>(0);
  if constexpr (W > 0) {
    int c[W];
    V v[W];
    T xv[W];
#pragma unroll
    for (int k = 0; k < W; k++) {
      size_t q = static_cast<size_t>(k) * n + r;
      c[k] = col[q];
      v[k] = val[q];
    }
#pragma unroll
    for (int k = 0; k < W; k++) xv[k] = x[c[k] < 0 ? 0 : c[k]];
#pragma unroll
    for (int k = 0; k < W; k++)
      if (c[k] >= 0) acc = acc + mul<V, T>(v[k], xv[k]);
  } else {
    // wide rows: four slots' loads in flight at a time, summed in slot order
    int k = 0;
    for (; k + 4 <= width; k += 4) {
      int c[4];
      V v[4];
      T xv[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        size_t q = static_cast<size_t>(k + j) * n + r;
        c[j] = col[q];
        v[j] = val[q];
      }
      if (c[0] < 0) return acc;
#pragma unroll
      for (int j = 0; j < 4; j++) xv[j] = x[c[j] < 0 ? 0 : c[j]];
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (c[j] >= 0) acc = acc + mul<V, T>(v[j], xv[j]);
    }
    for (; k < width; k++) {
      size_t q = static_cast<size_t>(k) * n + r;
      int c = col[q];
      if (c < 0) break;
      acc = acc + mul<V, T>(val[q], x[c]);
    }
  }
  return acc;
}

// A row of at most WMAX slots (width ≤ WMAX, not a compile-time constant):
// every slot's loads are issued before the sum, as for an exact W, and the
// sum runs in slot order over the row's columns, so the result equals row<0>'s.
template <int WMAX, class T, class V>
__device__ __forceinline__ T row_upto(int n, int width, int r, const int *__restrict__ col,
                                      const V *__restrict__ val, const T *__restrict__ x) {
  int c[WMAX];
  V v[WMAX];
  T xv[WMAX];
#pragma unroll
  for (int k = 0; k < WMAX; k++) {
    c[k] = -1;
    v[k] = V(0);
    if (k < width) {
      size_t q = static_cast<size_t>(k) * n + r;
      c[k] = col[q];
      v[k] = val[q];
    }
  }
#pragma unroll
  for (int k = 0; k < WMAX; k++) xv[k] = c[k] >= 0 ? x[c[k]] : static_cast<T>(0);
  T acc = static_cast<T>(0);
#pragma unroll
  for (int k = 0; k < WMAX; k++)
    if (c[k] >= 0) acc = acc + mul<V, T>(v[k], xv[k]);
  return acc;
}

// Gauss-Seidel row: s = Σ_{c ≠ self} a x[c] in ascending order, d = diagonal.
template <int W, class T>
__device__ __forceinline__ void gs_row(int n, int width, int p, int self, const int *__restrict__ col,
                                       const T *__restrict__ val, const T *__restrict__ x, T &s, T &d) {
  s = static_cast<T>(0);
  d = static_cast<T>(0);
  if constexpr (W > 0) {
    int c[W];
    T v[W], xv[W];
#pragma unroll
    for (int k = 0; k < W; k++) {
      size_t q = static_cast<size_t>(k) * n + p;
      c[k] = col[q];
      v[k] = val[q];
    }
#pragma unroll
    for (int k = 0; k < W; k++) xv[k] = x[c[k] < 0 ? 0 : c[k]];
#pragma unroll
    for (int k = 0; k < W; k++) {
      if (c[k] == self) d = v[k];
      else if (c[k] >= 0) s = s + v[k] * xv[k];
    }
  } else {
    for (int k = 0; k < width; k++) {
      size_t q = static_cast<size_t>(k) * n + p;
      int c = col[q];
      if (c < 0) break;
      if (c == self) d = val[q];
      else s = s + val[q] * x[c];
    }
  }
}

}  // namespace dec_ell

// Instantiate `LAUNCH` (a statement using the constexpr int W) for the ELL width.
#define DEC_ELL_WIDTH_SWITCH(width, LAUNCH)   \
  switch (width) {                            \
    case 1: { constexpr int W = 1; LAUNCH; } break; \
    case 2: { constexpr int W = 2; LAUNCH; } break; \
    case 3: { constexpr int W = 3; LAUNCH; } break; \
    case 4: { constexpr int W = 4; LAUNCH; } break; \
    case 5: { constexpr int W = 5; LAUNCH; } break; \
    case 6: { constexpr int W = 6; LAUNCH; } break; \
    case 7: { constexpr int W = 7; LAUNCH; } break; \
    case 8: { constexpr int W = 8; LAUNCH; } break; \
    default: { constexpr int W = 0; LAUNCH; } break; \
  }
