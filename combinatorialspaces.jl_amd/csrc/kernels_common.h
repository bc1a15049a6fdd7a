// Shared device helpers and C-ABI error wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "dec_internal.h"

namespace dec {

inline void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail(DEC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

inline int nblocks(long long n, int b = 256) { return static_cast<int>((n + b - 1) / b); }
// grid for kernels using xcd_block(): rounded up to a multiple of the 8 XCDs
inline int xcd_grid(long long n, int b = 256) { return (nblocks(n, b) + 7) / 8 * 8; }

// mesh.hip
void dualize(dec_mesh_s &m, int center);
void ensure_dual(dec_mesh_s &m);
dec_mesh_s *subdivide(const dec_mesh_s &c, bool project);
// tet.hip: orientations of a 3-D complex before dualizing, and the tetrahedron
// part of its dual geometry with the 3-D diagonal Hodge stars
void resolve_orientation3(dec_mesh_s &m);
void dualize_tets(dec_mesh_s &m, int center);
// the 24 dual volumes of each tetrahedron (DualTet 24x + q), host copy
void tet_dual_volumes(dec_mesh_s &m, std::vector<double> &dvol);

// Host CSR with float64 values, columns ascending within each row.
struct HostCsr {
  int m = 0, n = 0;
  std::vector<int32_t> ptr, col;
  std::vector<double> val;
};
// The 0-form Laplace-Beltrami -⋆0⁻¹ d̃1 ⋆1 d0 assembled in the reference's
// SpGEMM order (see ops.hip).  H (geometric ⋆1) may be null: then ⋆1 = diag(h1).
void assemble_laplacian0(const Topo &t, const std::vector<double> &inv_h0, const std::vector<double> &h1,
                         const HostCsr *H, HostCsr &L);
// C = A B, each entry accumulated over k ascending (SparseArrays' order).
void spgemm(const HostCsr &A, const HostCsr &B, HostCsr &C);
// Geometric ⋆1 (FastDEC.jl:660-716) as a host CSR, per-triangle entries on the GPU.
void geo_hodge1(dec_mesh_s &m, HostCsr &H);
// Upload a CSR as slot-major ELL with values cast to V.
template <class V>
void upload_ell(const HostCsr &A, Ell<V> &E);
// y = A x over a slot-major ELL, rows in storage order, sequential sums.
template <class T>
void ell_spmv(const Ell<T> &E, const T *x, T *y, hipStream_t s, const int *sched = nullptr, int nsched = 0);

}  // namespace dec

#define DEC_TRY try {
#define DEC_CATCH                                \
  }                                              \
  catch (const dec::Error &e) {                  \
    dec::set_last_error(e.msg);                  \
    return e.code;                               \
  }                                              \
  catch (const std::bad_alloc &) {               \
    dec::set_last_error("host out of memory");   \
    return DEC_ERR_OOM;                          \
  }                                              \
  return DEC_OK;
