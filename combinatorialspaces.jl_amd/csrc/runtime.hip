// Runtime pieces of libdec_hip.so: errors, device memory, seeded inputs.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "dec_internal.h"

namespace dec {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

void fail(int code, const std::string &msg) { throw Error{code, msg}; }

static void check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    if (e == hipErrorOutOfMemory) fail(DEC_ERR_OOM, std::string(what) + ": out of device memory");
    fail(DEC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
}

void *dmalloc(size_t bytes) {
  void *p = nullptr;
  check(hipMalloc(&p, bytes), "hipMalloc");
  return p;
}

void dfree(void *p) { (void)hipFree(p); }

void h2d(void *dst, const void *src, size_t bytes) {
  check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

void d2h(void *dst, const void *src, size_t bytes) {
  check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

void dsync() { check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

}  // namespace dec

namespace {

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

template <class T>
__global__ void k_fill_uniform(T *x, long long n, unsigned long long seed) {
  long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) {
    double u = static_cast<double>(splitmix64((seed << 32) + static_cast<unsigned long long>(i)) >> 11) *
               0x1.0p-53;
    x[i] = static_cast<T>(2.0 * u - 1.0);
  }
}

}  // namespace

extern "C" {

const char *dec_last_error(void) { return dec::g_last_error.c_str(); }

const char *dec_version(void) { return "dec_hip 0.1.0 (gfx950)"; }

int dec_fill_uniform(void *x, int64_t n, uint64_t seed, int dtype, void *stream) {
  try {
    if (n < 0 || (n > 0 && !x)) dec::fail(DEC_ERR_ARG, "dec_fill_uniform: bad vector");
    if (n == 0) return DEC_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int blocks = static_cast<int>((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    if (dtype == DEC_F64)
      hipLaunchKernelGGL(k_fill_uniform<double>, dim3(blocks), dim3(256), 0, s, static_cast<double *>(x), n, seed);
    else if (dtype == DEC_F32)
      hipLaunchKernelGGL(k_fill_uniform<float>, dim3(blocks), dim3(256), 0, s, static_cast<float *>(x), n, seed);
    else
      dec::fail(DEC_ERR_ARG, "dec_fill_uniform: dtype must be DEC_F64 or DEC_F32");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) dec::fail(DEC_ERR_HIP, std::string("dec_fill_uniform: ") + hipGetErrorString(e));
    return DEC_OK;
  } catch (const dec::Error &e) {
    dec::set_last_error(e.msg);
    return e.code;
  }
}

}  // extern "C"
