// Learned smoother: per-level operator data (shared by mg.hip and learned.hip).
#pragma once

#include "dec_internal.h"
#include "kernels_common.h"

namespace dec {

// A level's operator A, its transpose, P = D⁻¹(D - A) and Pᵀ as float ELL,
// and D = diag(A), all with rows in vertex order.
struct NnLevel {
  int n = 0;
  struct Op {
    int n = 0, width = 0;
    DBuf<int32_t> col;
    DBuf<float> val;
  } a, at, p, pt;
  DBuf<float> D;
  explicit NnLevel(const HostCsr &A);
};

// u += N(r) for one residual r (both float, vertex order) on level `lvl` of
// the network's hierarchy.
void nn_smooth(dec_nn_s &nn, int lvl, const float *r, float *u, hipStream_t s);
// Check that the network was built for a hierarchy with these level sizes
// (finest first) and size its inference buffers for the largest level.
void nn_bind(dec_nn_s &nn, const std::vector<int> &sizes);
dec_nn_s *nn_create(const std::vector<const HostCsr *> &ops, uint64_t seed);

}  // namespace dec
