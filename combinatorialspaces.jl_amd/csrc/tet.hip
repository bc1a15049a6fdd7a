// Tetrahedral complexes of libdec_hip.so: EmbeddedDeltaSet3D from boundary
// maps or a tetrahedron list, and the float64 geometry of the 3-D dual complex
// (subdivide_duals_3d! + precompute_volumes_3d!, DiscreteExteriorCalculus.jl:
// 1303-1335) on the GPU.
//
// The reference subdivides each tetrahedron into 24 dual tetrahedra
// (vertex center, edge center, triangle center, tetrahedron center) in the
// order of D_tetrahedron_schemas (DiscreteExteriorCalculus.jl:1184-1188), so
// DualTet 24x + q belongs to tetrahedron x.  The diagonal Hodge stars only
// need sums over elementary duals (FastDEC.jl:606-636):
//   ⋆0[v] = Σ dual_vol of the dual tetrahedra at v,
//   ⋆1[e] = Σ dual_area of the dual triangles (center(e), center(t), center(x)) / |e|,
//   ⋆2[t] = Σ |center(t) - center(x)| / |t|,   ⋆3[x] = 1 / |x|,
// so one thread per tetrahedron computes its 24 dual volumes, 12 dual areas and
// 4 dual lengths, and gathers over host-built incidence lists sum them per
// vertex, edge and triangle, in ascending tetrahedron order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "dec_internal.h"
#include "kernels_common.h"

namespace {

using dec::DBuf;
using dec::fail;

struct Q3 {
  double x, y, z;
};
__device__ __forceinline__ Q3 ldq(const double *p, size_t i) { return Q3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void stq(double *p, size_t i, Q3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ __forceinline__ Q3 qsub(Q3 a, Q3 b) { return Q3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double qdot(Q3 a, Q3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ Q3 qcross(Q3 a, Q3 b) {
  return Q3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// volume(points) for n = 1, 2, 3 (SimplicialSets.jl:896-910)
__device__ __forceinline__ double qvol1(Q3 a, Q3 b) {
  Q3 d = qsub(a, b);
  return sqrt(qdot(d, d));
}
__device__ __forceinline__ double qvol2(Q3 a, Q3 b, Q3 c) {
  Q3 u = qsub(b, a), w = qsub(c, a);
  double uu = qdot(u, u), ww = qdot(w, w), uw = qdot(u, w);
  return sqrt(fabs(uu * ww - uw * uw)) / 2;
}
__device__ __forceinline__ double qvol3(Q3 a, Q3 b, Q3 c, Q3 d) {
  return fabs(qdot(qsub(b, a), qcross(qsub(c, a), qsub(d, a)))) / 6;
}
// Circumcenter of a tetrahedron: p0 + (|a|² b×c + |b|² c×a + |c|² a×b) / (2 a·(b×c))
// with a, b, c the edges from p0 -- the point the reference finds from the
// first row of the inverse Cayley-Menger matrix (DiscreteExteriorCalculus.jl:114-118).
__device__ __forceinline__ Q3 circum4(Q3 p0, Q3 p1, Q3 p2, Q3 p3) {
  Q3 a = qsub(p1, p0), b = qsub(p2, p0), c = qsub(p3, p0);
  Q3 bc = qcross(b, c), ca = qcross(c, a), ab = qcross(a, b);
  double aa = qdot(a, a), bb = qdot(b, b), cc = qdot(c, c);
  double den = 2 * qdot(a, bc);
  return Q3{p0.x + ((aa * bc.x + bb * ca.x) + cc * ab.x) / den, p0.y + ((aa * bc.y + bb * ca.y) + cc * ab.y) / den,
            p0.z + ((aa * bc.z + bb * ca.z) + cc * ab.z) / den};
}
// sum(points) / 4, summed left to right (DiscreteExteriorCalculus.jl:100-102)
__device__ __forceinline__ Q3 bary4(Q3 p0, Q3 p1, Q3 p2, Q3 p3) {
  return Q3{(((p0.x + p1.x) + p2.x) + p3.x) / 4, (((p0.y + p1.y) + p2.y) + p3.y) / 4,
            (((p0.z + p1.z) + p2.z) + p3.z) / 4};
}

// D_tetrahedron_schemas: (vertex slot, edge slot, triangle slot) of dual tetrahedron q.
__constant__ int8_t kSv[24] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3};
__constant__ int8_t kSe[24] = {4, 4, 3, 3, 5, 5, 5, 5, 1, 1, 2, 2, 2, 2, 0, 0, 4, 4, 3, 3, 0, 0, 1, 1};
__constant__ int8_t kSt[24] = {3, 1, 1, 2, 2, 3, 3, 2, 2, 0, 0, 3, 3, 0, 0, 1, 1, 3, 2, 1, 1, 0, 0, 2};
// For edge slot j, the two triangle slots whose dual triangles
// (center(e_j), center(t_k), center(x)) the dual complex creates first and second.
__constant__ int8_t kKa[6] = {0, 2, 0, 1, 3, 2};
__constant__ int8_t kKb[6] = {1, 0, 3, 2, 1, 3};

// One thread per tetrahedron x: its center (dual vertex nv+ne+nt+x), volume,
// the dual volume at each of its vertices (tw, slot-major), the two dual areas
// per edge slot (ta, slot 2j and 2j+1) and the dual length per triangle slot
// (tl); dvol, when given, receives the 24 dual volumes (DualTet 24x + q).
__global__ void k_tet_geometry(int nv, int ne, int nt, int n, const int *__restrict__ v0, const int *__restrict__ v1,
                               const int *__restrict__ e0, const int *__restrict__ e1, const int *__restrict__ e2,
                               const int *__restrict__ t0, const int *__restrict__ t1, const int *__restrict__ t2,
                               const int *__restrict__ t3, const double *__restrict__ p, int circumcenter,
                               double *__restrict__ dpt, double *__restrict__ vol, double *__restrict__ tw,
                               double *__restrict__ ta, double *__restrict__ tl, double *__restrict__ dvol) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const size_t N = static_cast<size_t>(n);
  int T[4] = {t0[x], t1[x], t2[x], t3[x]};
  // tetrahedron_vertices and tetrahedron_edges (SimplicialSets.jl:474-492)
  int V[4] = {v1[e2[T[2]]], v0[e2[T[2]]], v1[e0[T[0]]], v0[e0[T[0]]]};
  int E[6] = {e0[T[0]], e1[T[0]], e2[T[0]], e1[T[1]], e2[T[1]], e2[T[2]]};
  Q3 P[4], EC[6], TC[4];
  for (int i = 0; i < 4; i++) P[i] = ldq(p, V[i]);
  for (int j = 0; j < 6; j++) EC[j] = ldq(dpt, static_cast<size_t>(nv) + E[j]);
  for (int k = 0; k < 4; k++) TC[k] = ldq(dpt, static_cast<size_t>(nv) + ne + T[k]);
  Q3 c = circumcenter ? circum4(P[0], P[1], P[2], P[3]) : bary4(P[0], P[1], P[2], P[3]);
  stq(dpt, static_cast<size_t>(nv) + ne + nt + x, c);
  vol[x] = qvol3(P[0], P[1], P[2], P[3]);
  double w[4] = {0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < 24; q++) {
    double dv = qvol3(P[kSv[q]], EC[kSe[q]], TC[kSt[q]], c);
    w[kSv[q]] += dv;
    if (dvol) dvol[24 * static_cast<size_t>(x) + q] = dv;
  }
  for (int i = 0; i < 4; i++) tw[i * N + x] = w[i];
  for (int j = 0; j < 6; j++) {
    ta[(2 * j) * N + x] = qvol2(EC[j], TC[kKa[j]], c);
    ta[(2 * j + 1) * N + x] = qvol2(EC[j], TC[kKb[j]], c);
  }
  for (int k = 0; k < 4; k++) tl[k * N + x] = qvol1(c, TC[k]);
}

// out[r] = Σ vals[idx[k]] over the row's list (in list order), divided by den[r] if given.
__global__ void k_gather_sum(int n, const int64_t *__restrict__ ptr, const int64_t *__restrict__ idx,
                             const double *__restrict__ vals, const double *__restrict__ den,
                             double *__restrict__ out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int64_t k = ptr[r]; k < ptr[r + 1]; k++) s += vals[idx[k]];
  out[r] = den ? s / den[r] : s;
}

__global__ void k_recip(int n, const double *__restrict__ a, double *__restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = 1 / a[i];
}

// Incidence lists of flat indices into the per-tetrahedron tables, rows in
// ascending tetrahedron order.
struct Lists {
  std::vector<int64_t> ptr, idx;
};
void build_lists(int rows, int n, int per, const std::vector<int32_t> &key, const std::vector<int32_t> &slot,
                 int stride_mult, int copies, Lists &L) {
  // key[x*per + i] = row of (x, slot i); entries stride_mult*slot*N + x (+ N per copy)
  L.ptr.assign(rows + 1, 0);
  for (size_t q = 0; q < key.size(); q++) L.ptr[key[q] + 1] += copies;
  for (int r = 0; r < rows; r++) L.ptr[r + 1] += L.ptr[r];
  L.idx.assign(L.ptr[rows], 0);
  std::vector<int64_t> w(L.ptr.begin(), L.ptr.end() - 1);
  const int64_t N = n;
  for (int x = 0; x < n; x++)
    for (int i = 0; i < per; i++) {
      int r = key[static_cast<size_t>(x) * per + i];
      int64_t base = static_cast<int64_t>(stride_mult) * slot[static_cast<size_t>(x) * per + i] * N + x;
      for (int c = 0; c < copies; c++) L.idx[w[r]++] = base + c * N;
    }
}

void upload_lists(const Lists &L, DBuf<int64_t> &p, DBuf<int64_t> &i) {
  p.upload(L.ptr);
  i.upload(L.idx);
}

void launch_geometry(dec_mesh_s &m, int circ, double *vol, double *tw, double *ta, double *tl, double *dvol) {
  int n = m.ht.n;
  if (n == 0) return;
  const dec::Topo &t = m.h;
  hipLaunchKernelGGL(k_tet_geometry, dim3(dec::nblocks(n)), dim3(256), 0, 0, t.nv, t.ne, t.nt, n, m.v0.p, m.v1.p,
                     m.e0.p, m.e1.p, m.e2.p, m.t0.p, m.t1.p, m.t2.p, m.t3.p, m.p.p, circ, m.dpt.p, vol, tw, ta, tl,
                     dvol);
  dec::check_launch("tetrahedron geometry");
}

dec_mesh_s *new_mesh3(dec::Topo &&t, dec::Tets &&T, const std::vector<double> &pts) {
  auto *m = new dec_mesh_s();
  m->dim = 3;
  m->h = std::move(t);
  m->ht = std::move(T);
  m->hp = pts;
  try {
    m->v0.upload(m->h.v0);
    m->v1.upload(m->h.v1);
    m->e0.upload(m->h.e0);
    m->e1.upload(m->h.e1);
    m->e2.upload(m->h.e2);
    m->t0.upload(m->ht.t[0]);
    m->t1.upload(m->ht.t[1]);
    m->t2.upload(m->ht.t[2]);
    m->t3.upload(m->ht.t[3]);
    m->p.upload(pts);
  } catch (...) {
    delete m;
    throw;
  }
  return m;
}

dec_mesh_s &mesh3(dec_mesh m, const char *what) {
  if (!m) fail(DEC_ERR_ARG, std::string(what) + ": NULL handle");
  if (m->dim != 3) fail(DEC_ERR_ARG, std::string(what) + ": not a 3-D complex");
  return *m;
}

}  // namespace

namespace dec {

// Orientations of a 3-D complex before its dual is built: unset edges and
// triangles become false (zero(Bool), DiscreteExteriorCalculus.jl:311-321,
// 573-584) and unset tetrahedra are set by orient! (:1201-1207).
void resolve_orientation3(dec_mesh_s &m) {
  for (auto &o : m.h.eo)
    if (o < 0) o = 0;
  for (auto &o : m.h.to)
    if (o < 0) o = 0;
  if (std::any_of(m.ht.o.begin(), m.ht.o.end(), [](int8_t o) { return o < 0; })) {
    std::vector<int8_t> to;
    if (!orient3_dfs(m.h, m.ht, to))
      fail(DEC_ERR_NONORIENTABLE, "The 3-simplices of the given 3D delta set are non-orientable.");
    m.ht.o = to;
  }
}

// The tetrahedron part of the dual geometry and the 3-D diagonal Hodge stars;
// the edge and triangle parts (centers, lengths, areas, the 2-D dual
// triangles of each triangle) are already in m.
void dualize_tets(dec_mesh_s &m, int center) {
  const Topo &t = m.h;
  const Tets &T = m.ht;
  int nv = t.nv, ne = t.ne, nt = t.nt, n = T.n;
  const size_t N = static_cast<size_t>(n);
  m.teo.upload(T.o);
  m.vol.alloc(N);
  m.tw.alloc(4 * N);
  DBuf<double> ta, tl;
  ta.alloc(12 * N);
  tl.alloc(4 * N);
  launch_geometry(m, center == DEC_CIRCUMCENTER ? 1 : 0, m.vol.p, m.tw.p, ta.p, tl.p, nullptr);
  // incidence lists: vertex -> (x, vertex slot), edge -> (x, edge slot) twice, triangle -> (x, slot)
  std::vector<int32_t> kv(4 * N), sv(4 * N), ke(6 * N), se(6 * N), kt(4 * N), st(4 * N);
  for (int x = 0; x < n; x++) {
    int v[4], e[6];
    tet_vertices(t, T, x, v);
    tet_edges(t, T, x, e);
    const size_t X = static_cast<size_t>(x);
    for (int i = 0; i < 4; i++) {
      kv[4 * X + i] = v[i];
      sv[4 * X + i] = i;
      kt[4 * X + i] = T.t[i][x];
      st[4 * X + i] = i;
    }
    for (int j = 0; j < 6; j++) {
      ke[6 * X + j] = e[j];
      se[6 * X + j] = j;
    }
  }
  Lists Lv, Le, Lt;
  build_lists(nv, n, 4, kv, sv, 1, 1, Lv);
  build_lists(ne, n, 6, ke, se, 2, 2, Le);
  build_lists(nt, n, 4, kt, st, 1, 1, Lt);
  DBuf<int64_t> pv, iv, pe, ie, pt, it;
  upload_lists(Lv, pv, iv);
  upload_lists(Le, pe, ie);
  upload_lists(Lt, pt, it);
  m.h0.alloc(nv);
  m.h1.alloc(ne);
  m.h2.alloc(nt);
  m.h3.alloc(N);
  if (nv) hipLaunchKernelGGL(k_gather_sum, dim3(nblocks(nv)), dim3(256), 0, 0, nv, pv.p, iv.p, m.tw.p, nullptr, m.h0.p);
  if (ne) hipLaunchKernelGGL(k_gather_sum, dim3(nblocks(ne)), dim3(256), 0, 0, ne, pe.p, ie.p, ta.p, m.length.p, m.h1.p);
  if (nt) hipLaunchKernelGGL(k_gather_sum, dim3(nblocks(nt)), dim3(256), 0, 0, nt, pt.p, it.p, tl.p, m.area.p, m.h2.p);
  if (n) hipLaunchKernelGGL(k_recip, dim3(nblocks(n)), dim3(256), 0, 0, n, m.vol.p, m.h3.p);
  check_launch("3-D Hodge stars");
  dsync();
}

void tet_dual_volumes(dec_mesh_s &m, std::vector<double> &dvol) {
  const size_t N = static_cast<size_t>(m.ht.n);
  dvol.assign(24 * N, 0.0);
  if (!N) return;
  DBuf<double> v, w, a, l, dv;
  v.alloc(N);
  w.alloc(4 * N);
  a.alloc(12 * N);
  l.alloc(4 * N);
  dv.alloc(24 * N);
  launch_geometry(m, m.center == DEC_CIRCUMCENTER ? 1 : 0, v.p, w.p, a.p, l.p, dv.p);
  d2h(dvol.data(), dv.p, sizeof(double) * 24 * N);
}

}  // namespace dec

extern "C" {

int dec_mesh3_create(int32_t nv, int32_t ne, int32_t nt, int32_t ntet, const int32_t *v0, const int32_t *v1,
                     const int32_t *e0, const int32_t *e1, const int32_t *e2, const int32_t *t0, const int32_t *t1,
                     const int32_t *t2, const int32_t *t3, const int8_t *edge_orient, const int8_t *tri_orient,
                     const int8_t *tet_orient, const double *points, int32_t index_base, dec_mesh *out) {
  DEC_TRY
  if (!out) fail(DEC_ERR_ARG, "dec_mesh3_create: out is NULL");
  *out = nullptr;
  if (nv < 0 || ne < 0 || nt < 0 || ntet < 0) fail(DEC_ERR_ARG, "dec_mesh3_create: negative count");
  if ((ne && (!v0 || !v1)) || (nt && (!e0 || !e1 || !e2)) || (ntet && (!t0 || !t1 || !t2 || !t3)) ||
      (nv && !points))
    fail(DEC_ERR_ARG, "dec_mesh3_create: NULL array");
  if (index_base != 0 && index_base != 1) fail(DEC_ERR_ARG, "dec_mesh3_create: index_base must be 0 or 1");
  dec::Topo t;
  t.nv = nv;
  t.ne = ne;
  t.nt = nt;
  auto take = [&](const int32_t *src, int count, int limit, std::vector<int32_t> &dst, const char *what) {
    dst.resize(count);
    for (int i = 0; i < count; i++) {
      dst[i] = src[i] - index_base;
      if (dst[i] < 0 || dst[i] >= limit) fail(DEC_ERR_ARG, std::string("dec_mesh3_create: ") + what + " out of range");
    }
  };
  take(v0, ne, nv, t.v0, "edge vertex");
  take(v1, ne, nv, t.v1, "edge vertex");
  take(e0, nt, ne, t.e0, "triangle edge");
  take(e1, nt, ne, t.e1, "triangle edge");
  take(e2, nt, ne, t.e2, "triangle edge");
  dec::Tets T;
  T.n = ntet;
  const int32_t *src[4] = {t0, t1, t2, t3};
  for (int k = 0; k < 4; k++) take(src[k], ntet, nt, T.t[k], "tetrahedron triangle");
  auto orient = [](const int8_t *o, int count, std::vector<int8_t> &dst) {
    dst.assign(count, -1);
    if (o)
      for (int i = 0; i < count; i++) dst[i] = o[i] < 0 ? -1 : (o[i] ? 1 : 0);
  };
  orient(edge_orient, ne, t.eo);
  orient(tri_orient, nt, t.to);
  orient(tet_orient, ntet, T.o);
  std::vector<double> pts(points, points + 3 * static_cast<size_t>(nv));
  *out = new_mesh3(std::move(t), std::move(T), pts);
  DEC_CATCH
}

int dec_mesh3_from_tets(int32_t nv, int32_t ntet, const double *points, const int32_t *tets, int32_t sorted,
                        int32_t index_base, dec_mesh *out) {
  DEC_TRY
  if (!out) fail(DEC_ERR_ARG, "dec_mesh3_from_tets: out is NULL");
  *out = nullptr;
  if (nv < 0 || ntet < 0) fail(DEC_ERR_ARG, "dec_mesh3_from_tets: negative count");
  if ((nv > 0 && !points) || (ntet > 0 && !tets)) fail(DEC_ERR_ARG, "dec_mesh3_from_tets: NULL array");
  if (index_base != 0 && index_base != 1) fail(DEC_ERR_ARG, "dec_mesh3_from_tets: index_base must be 0 or 1");
  std::vector<int32_t> tv(4 * static_cast<size_t>(ntet));
  for (size_t i = 0; i < tv.size(); i++) tv[i] = tets[i] - index_base;
  dec::Topo t;
  dec::Tets T;
  dec::glue_tetrahedra(nv, tv, sorted != 0, t, T);
  // orient!(s) right after gluing (MeshInterop.jl:150); a non-orientable
  // complex keeps its tetrahedra unset and is refused when dualized
  std::vector<int8_t> to;
  if (dec::orient3_dfs(t, T, to)) T.o = to;
  std::vector<double> p(points, points + 3 * static_cast<size_t>(nv));
  *out = new_mesh3(std::move(t), std::move(T), p);
  DEC_CATCH
}

int dec_mesh_dim(dec_mesh m, int32_t *dim) {
  DEC_TRY
  if (!m || !dim) fail(DEC_ERR_ARG, "dec_mesh_dim: NULL argument");
  *dim = m->dim;
  DEC_CATCH
}

int dec_mesh3_counts(dec_mesh m, int32_t *nv, int32_t *ne, int32_t *nt, int32_t *ntet) {
  DEC_TRY
  dec_mesh_s &s = mesh3(m, "dec_mesh3_counts");
  if (nv) *nv = s.h.nv;
  if (ne) *ne = s.h.ne;
  if (nt) *nt = s.h.nt;
  if (ntet) *ntet = s.ht.n;
  DEC_CATCH
}

int dec_mesh3_export(dec_mesh m, int32_t *t0, int32_t *t1, int32_t *t2, int32_t *t3, int8_t *tet_orient) {
  DEC_TRY
  dec_mesh_s &s = mesh3(m, "dec_mesh3_export");
  int32_t *dst[4] = {t0, t1, t2, t3};
  for (int k = 0; k < 4; k++)
    if (dst[k]) std::copy(s.ht.t[k].begin(), s.ht.t[k].end(), dst[k]);
  if (tet_orient) std::copy(s.ht.o.begin(), s.ht.o.end(), tet_orient);
  DEC_CATCH
}

int dec_mesh3_export_dual(dec_mesh m, double *dual_point, double *length, double *area, double *vol,
                          double *dual_length, double *dual_area, double *dual_vol) {
  DEC_TRY
  dec_mesh_s &s = mesh3(m, "dec_mesh3_export_dual");
  if (!s.dualized) fail(DEC_ERR_ARG, "dec_mesh3_export_dual: mesh is not dualized");
  if (dual_point) dec::d2h(dual_point, s.dpt.p, sizeof(double) * s.dpt.n);
  if (length) dec::d2h(length, s.length.p, sizeof(double) * s.length.n);
  if (area) dec::d2h(area, s.area.p, sizeof(double) * s.area.n);
  if (vol) dec::d2h(vol, s.vol.p, sizeof(double) * s.vol.n);
  if (dual_length) dec::d2h(dual_length, s.dl.p, sizeof(double) * s.dl.n);
  if (dual_area) dec::d2h(dual_area, s.da.p, sizeof(double) * s.da.n);
  if (dual_vol && s.ht.n) {
    std::vector<double> dv;
    dec::tet_dual_volumes(s, dv);
    std::copy(dv.begin(), dv.end(), dual_vol);
  }
  DEC_CATCH
}

}  // extern "C"
