// STREAM-style bandwidth ceiling of the device (libdec_stream.so): copy
// (b = a) and triad (a = b + s c) over float64 arrays, timed with HIP events
// on the launch stream.  It is the measured reference point for the
// operators' GB/s, not part of the operator library.
//
// Each thread moves 16-byte vectors (double2); the grid covers the arrays
// once (no grid-stride loop), so a launch is the same shape as an operator
// launch over the same bytes.  Stores may be nontemporal, as the operator
// kernels' output stores are.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kBlock = 256;

template <bool NT>
__device__ __forceinline__ void store2(double2 *p, double2 v) {
  if constexpr (NT) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
  } else {
    *p = v;
  }
}

template <bool NT>
__global__ __launch_bounds__(kBlock) void k_copy2(int64_t n2, const double2 *__restrict__ a, double2 *__restrict__ b) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n2) store2<NT>(b + i, a[i]);
}

template <bool NT>
__global__ __launch_bounds__(kBlock) void k_triad2(int64_t n2, double2 *__restrict__ a, const double2 *__restrict__ b,
                                                   const double2 *__restrict__ c, double s) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n2) {
    double2 x = b[i], y = c[i], r;
    r.x = x.x + s * y.x;
    r.y = x.y + s * y.y;
    store2<NT>(a + i, r);
  }
}

}  // namespace

extern "C" {

// kind 0: copy b = a over `sets` rotating pairs; kind 1: triad a = b + s c
// over `sets` rotating triples.  bufs holds 2*sets (copy) or 3*sets (triad)
// device pointers to arrays of n doubles (n even, 16-byte aligned).  Runs
// `warmup` untimed and `reps` timed launches, cycling through the sets, and
// writes the mean milliseconds per launch.  Returns 0, or a hipError_t code.
int dec_stream_run(int32_t kind, int32_t nontemporal, int64_t n, int32_t sets, void *const *bufs, int32_t warmup,
                   int32_t reps, void *stream, float *ms_per_launch) {
  if (n <= 0 || (n & 1) || sets <= 0 || reps <= 0 || !bufs || !ms_per_launch || (kind != 0 && kind != 1))
    return static_cast<int>(hipErrorInvalidValue);
  for (int i = 0; i < (kind == 0 ? 2 : 3) * sets; i++)
    if (!bufs[i] || reinterpret_cast<uintptr_t>(bufs[i]) % 16) return static_cast<int>(hipErrorInvalidValue);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int64_t n2 = n / 2;
  dim3 grid(static_cast<unsigned>((n2 + kBlock - 1) / kBlock)), blk(kBlock);
  auto launch = [&](int it) {
    int q = it % sets;
    if (kind == 0) {
      auto a = static_cast<const double2 *>(bufs[2 * q]);
      auto b = static_cast<double2 *>(bufs[2 * q + 1]);
      if (nontemporal) hipLaunchKernelGGL(k_copy2<true>, grid, blk, 0, s, n2, a, b);
      else hipLaunchKernelGGL(k_copy2<false>, grid, blk, 0, s, n2, a, b);
    } else {
      auto a = static_cast<double2 *>(bufs[3 * q]);
      auto b = static_cast<const double2 *>(bufs[3 * q + 1]);
      auto c = static_cast<const double2 *>(bufs[3 * q + 2]);
      if (nontemporal) hipLaunchKernelGGL(k_triad2<true>, grid, blk, 0, s, n2, a, b, c, 3.0);
      else hipLaunchKernelGGL(k_triad2<false>, grid, blk, 0, s, n2, a, b, c, 3.0);
    }
  };
  for (int it = 0; it < warmup; it++) launch(it);
  hipEvent_t e0, e1;
  hipError_t err = hipEventCreate(&e0);
  if (err != hipSuccess) return static_cast<int>(err);
  err = hipEventCreate(&e1);
  if (err != hipSuccess) {
    (void)hipEventDestroy(e0);
    return static_cast<int>(err);
  }
  (void)hipEventRecord(e0, s);
  for (int it = 0; it < reps; it++) launch(warmup + it);
  (void)hipEventRecord(e1, s);
  err = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  if (err == hipSuccess) err = hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_per_launch = ms / static_cast<float>(reps);
  return static_cast<int>(err);
}

}  // extern "C"
