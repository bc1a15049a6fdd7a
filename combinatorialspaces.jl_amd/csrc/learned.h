// Learned smoother: per-level operator data (shared by mg.hip and learned.hip).
#pragma once

#include "dec_internal.h"
#include "kernels_common.h"

namespace dec {

// A level's operator A, its transpose, P = D⁻¹(D - A) and Pᵀ as float ELL,
// and D = diag(A), all with rows in vertex order; for the error loss, the
// restriction R to the next coarser level (nc rows), the prolongation Pr
// back (n rows) and their transposes (nc = 0 on the coarsest level).
struct NnLevel {
  int n = 0, nc = 0;
  float dmean = 1.0f;  // mean of D: the error loss's scale
  struct Op {
    int n = 0, width = 0;
    DBuf<int32_t> col;
    DBuf<float> val;
  } a, at, p, pt, r, rt, pr, prt;
  DBuf<float> D;
  // P and D over the multigrid level's own vertex numbering (perm[i] = vertex
  // of internal number i), for the smoother inside the V-cycles
  Op pi;
  DBuf<float> Di;
  // perm on the device and pi's columns as vertex numbers (both empty when
  // the numbering is vertex order)
  DBuf<int32_t> perm_d, pi_vcol;
  // Halo lists of pi's 64-row tiles (large levels): tile t's rows gather
  // from its own rows and from the rows hrow[hoff[t] .. hoff[t+1]) (hrow_v:
  // their vertex numbers, with perm); lidx (slot-major like pi) holds each
  // slot's column as a tile-local index, 0..63 the tile's own rows, 64 + j
  // its j-th halo row, 255 a pad.  hmax: the largest halo, -1 when the lists
  // were not built (a halo above 191 rows, or rows wider than 8 slots).
  DBuf<uint8_t> lidx;
  DBuf<int32_t> hoff, hrow, hrow_v;
  int hmax = -1;
  NnLevel(const HostCsr &A, const HostCsr *R, const HostCsr *Pr, const std::vector<int32_t> *perm);
};

// u += N(r) for one residual r (both float) on level `lvl` of the network's
// hierarchy: in the multigrid level's own numbering (internal), else in
// vertex order.
void nn_smooth(dec_nn_s &nn, int lvl, const float *r, float *u, hipStream_t s, bool internal = true);
// Check that the network was built for a hierarchy with these level sizes
// (finest first) and size its inference buffers for the largest level.
void nn_bind(dec_nn_s &nn, const std::vector<int> &sizes);
// ops[l] = (A_l, R_l, P_l), R and P NULL on the coarsest level.
struct NnOps {
  const HostCsr *A, *R, *P;
  const std::vector<int32_t> *perm;  // the level's numbering (NULL: vertex order)
};
dec_nn_s *nn_create(const std::vector<NnOps> &ops, uint64_t seed);

}  // namespace dec
