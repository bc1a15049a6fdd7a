// Meshes of libdec_hip.so: primal complexes, binary subdivision and the
// float64 dual geometry (circumcentric or barycentric), computed on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "dec_internal.h"
#include "kernels_common.h"

namespace {

using dec::fail;

// propagate_points(::BinarySubdivision) (Multigrid.jl:111-122): new vertex
// nv + e at (p[∂v0] + p[∂v1]) / 2, optionally rescaled to the unit sphere.
__global__ void k_propagate(int nv, int ne, const int *__restrict__ v0, const int *__restrict__ v1,
                            const double *__restrict__ p, double *__restrict__ pf, int project) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) {
    pf[3 * i] = p[3 * i];
    pf[3 * i + 1] = p[3 * i + 1];
    pf[3 * i + 2] = p[3 * i + 2];
  }
  if (i < ne) {
    const double *a = p + 3 * v0[i], *b = p + 3 * v1[i];
    double q0 = (a[0] + b[0]) / 2, q1 = (a[1] + b[1]) / 2, q2 = (a[2] + b[2]) / 2;
    if (project) {
      double n = sqrt((q0 * q0 + q1 * q1) + q2 * q2);
      q0 = q0 / n;
      q1 = q1 / n;
      q2 = q2 / n;
    }
    double *q = pf + 3 * static_cast<size_t>(nv + i);
    q[0] = q0;
    q[1] = q1;
    q[2] = q2;
  }
}

struct P3 {
  double x, y, z;
};
__device__ __forceinline__ P3 ld3(const double *p, int i) {
  return P3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
__device__ __forceinline__ void st3(double *p, size_t i, P3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
// sqdistance: sum((x - y).^2) left to right (SimplicialSets.jl:945)
__device__ __forceinline__ double sqdist(P3 a, P3 b) {
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return (dx * dx + dy * dy) + dz * dz;
}
// volume(points), n = 1 and n = 2 (SimplicialSets.jl:896-903)
__device__ __forceinline__ double vol1(P3 a, P3 b) { return sqrt(sqdist(a, b)); }
__device__ __forceinline__ double vol2(P3 a, P3 b, P3 c) {
  double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
  double wx = c.x - a.x, wy = c.y - a.y, wz = c.z - a.z;
  double uu = (ux * ux + uy * uy) + uz * uz;
  double ww = (wx * wx + wy * wy) + wz * wz;
  double uw = (ux * wx + uy * wy) + uz * wz;
  return sqrt(fabs(uu * ww - uw * uw)) / 2;
}
// Center of an edge: the circumcenter of two points is their midpoint
// 0.5 p0 + 0.5 p1 (first row of the inverse Cayley-Menger matrix is
// (-d/2, 1/2, 1/2)); the barycenter (p0 + p1)/2 has the same bits.
__device__ __forceinline__ P3 mid(P3 a, P3 b) {
  return P3{0.5 * a.x + 0.5 * b.x, 0.5 * a.y + 0.5 * b.y, 0.5 * a.z + 0.5 * b.z};
}
// Circumcenter of a triangle from its barycentric coordinates
// λ_i ∝ a_i (a_j + a_k - a_i), a_i the squared length of the side opposite p_i
// (the first row of inv(cayley_menger), DiscreteExteriorCalculus.jl:124-128),
// summed as mapreduce(*, +, λ, points).
__device__ __forceinline__ P3 circum(P3 p0, P3 p1, P3 p2) {
  double A = sqdist(p1, p2), B = sqdist(p0, p2), C = sqdist(p0, p1);
  double w0 = A * ((B + C) - A), w1 = B * ((C + A) - B), w2 = C * ((A + B) - C);
  double s = (w0 + w1) + w2;
  double l0 = w0 / s, l1 = w1 / s, l2 = w2 / s;
  return P3{(l0 * p0.x + l1 * p1.x) + l2 * p2.x, (l0 * p0.y + l1 * p1.y) + l2 * p2.y,
            (l0 * p0.z + l1 * p1.z) + l2 * p2.z};
}
__device__ __forceinline__ P3 bary3(P3 a, P3 b, P3 c) {
  return P3{((a.x + b.x) + c.x) / 3, ((a.y + b.y) + c.y) / 3, ((a.z + b.z) + c.z) / 3};
}

// Per edge e: length, edge center (dual vertex nv + e) and the two A-block
// dual lengths |center - p[∂v0]| (dual edge e) and |center - p[∂v1]| (ne + e).
__global__ void k_edge_geometry(int nv, int ne, const int *__restrict__ v0, const int *__restrict__ v1,
                                const double *__restrict__ p, double *__restrict__ length,
                                double *__restrict__ dl, double *__restrict__ dpt) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  P3 a = ld3(p, v0[e]), b = ld3(p, v1[e]);
  length[e] = vol1(a, b);
  P3 c = mid(a, b);
  st3(dpt, static_cast<size_t>(nv) + e, c);
  dl[e] = vol1(c, a);
  dl[ne + e] = vol1(c, b);
}

// Per triangle t: area, center (dual vertex nv + ne + t), the B_j and C_j dual
// lengths and the six dual areas in the reference numbering
// (DiscreteExteriorCalculus.jl:545-608, 941-990).
__global__ void k_tri_geometry(int nv, int ne, int nt, const int *__restrict__ v0,
                               const int *__restrict__ v1, const int *__restrict__ e0,
                               const int *__restrict__ e1, const int *__restrict__ e2,
                               const double *__restrict__ p, int circumcenter,
                               double *__restrict__ area, double *__restrict__ dl,
                               double *__restrict__ da, double *__restrict__ dpt) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int E[3] = {e0[t], e1[t], e2[t]};
  // triangle_vertices (SimplicialSets.jl:305-309)
  P3 P[3] = {ld3(p, v1[E[1]]), ld3(p, v0[E[2]]), ld3(p, v0[E[1]])};
  area[t] = vol2(P[0], P[1], P[2]);
  P3 tc = circumcenter ? circum(P[0], P[1], P[2]) : bary3(P[0], P[1], P[2]);
  st3(dpt, static_cast<size_t>(nv) + ne + t, tc);
  P3 ec[3];
  for (int j = 0; j < 3; j++) ec[j] = mid(ld3(p, v0[E[j]]), ld3(p, v1[E[j]]));
  for (int j = 0; j < 3; j++) {
    dl[2 * static_cast<size_t>(ne) + static_cast<size_t>(j) * nt + t] = vol1(tc, ec[j]);
    dl[2 * static_cast<size_t>(ne) + static_cast<size_t>(3 + j) * nt + t] = vol1(tc, P[j]);
  }
  // dual triangle k = (v, e, ev): vertices (P_v, center of ∂e_e, tc)
  const int sv[6] = {0, 0, 1, 1, 2, 2}, se[6] = {1, 2, 2, 0, 0, 1};
#pragma unroll
  for (int k = 0; k < 6; k++) da[static_cast<size_t>(k) * nt + t] = vol2(P[sv[k]], ec[se[k]], tc);
}

// ⋆0[v] = Σ dual_area over the dual triangles of v, ascending dual id (FastDEC.jl:580-587).
__global__ void k_hodge0(int nv, const int *__restrict__ ptr, const int *__restrict__ idx,
                         const double *__restrict__ da, double *__restrict__ h0) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  double s = 0.0;
  for (int k = ptr[v]; k < ptr[v + 1]; k++) s += da[idx[k]];
  h0[v] = s;
}

// ⋆1[e] = Σ dual_length / length over the B-block dual edges of e (FastDEC.jl:589-599).
__global__ void k_hodge1(int ne, const int *__restrict__ ptr, const int *__restrict__ idx,
                         const double *__restrict__ dlB, const double *__restrict__ length,
                         double *__restrict__ h1) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double s = 0.0, L = length[e];
  for (int k = ptr[e]; k < ptr[e + 1]; k++) s += dlB[idx[k]] / L;
  h1[e] = s;
}

// ⋆2[t] = 1 / area (FastDEC.jl:602-604).
__global__ void k_hodge2(int nt, const double *__restrict__ area, double *__restrict__ h2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nt) h2[t] = 1 / area[t];
}

void upload_topology(dec_mesh_s &m) {
  m.v0.upload(m.h.v0);
  m.v1.upload(m.h.v1);
  m.e0.upload(m.h.e0);
  m.e1.upload(m.h.e1);
  m.e2.upload(m.h.e2);
}

dec_mesh_s *new_mesh_from_topo(dec::Topo &&t, const std::vector<double> &pts) {
  auto *m = new dec_mesh_s();
  m->h = std::move(t);
  m->hp = pts;
  upload_topology(*m);
  m->p.upload(pts);
  return m;
}

}  // namespace

namespace dec {

// Resolve orientations (unset edge -> false; unset triangles -> orient!),
// upload them and compute the dual geometry and diagonal Hodge stars.
void dualize(dec_mesh_s &m, int center) {
  if (center != DEC_CIRCUMCENTER && center != DEC_BARYCENTER) fail(DEC_ERR_ARG, "dualize: unknown center");
  Topo &t = m.h;
  if (m.dim == 3) {
    resolve_orientation3(m);
  } else {
    for (auto &o : t.eo)
      if (o < 0) o = 0;
    bool unset = std::any_of(t.to.begin(), t.to.end(), [](int8_t o) { return o < 0; });
    if (unset) {
      std::vector<int8_t> to;
      if (!orient_dfs(t, to))
        fail(DEC_ERR_NONORIENTABLE, "The 2-simplices of the given 2D delta set are non-orientable.");
      t.to = to;
    }
  }
  m.eo.upload(t.eo);
  m.to.upload(t.to);
  int nv = t.nv, ne = t.ne, nt = t.nt;
  size_t ndv = static_cast<size_t>(nv) + ne + nt + (m.dim == 3 ? m.ht.n : 0);
  m.length.alloc(ne);
  m.area.alloc(nt);
  m.dl.alloc(2 * static_cast<size_t>(ne) + 6 * static_cast<size_t>(nt));
  m.da.alloc(6 * static_cast<size_t>(nt));
  m.dpt.alloc(3 * ndv);
  h2d(m.dpt.p, m.hp.data(), sizeof(double) * 3 * nv);
  if (ne > 0)
    hipLaunchKernelGGL(k_edge_geometry, dim3(dec::nblocks(ne)), dim3(256), 0, 0, nv, ne, m.v0.p, m.v1.p, m.p.p,
                       m.length.p, m.dl.p, m.dpt.p);
  if (nt > 0)
    hipLaunchKernelGGL(k_tri_geometry, dim3(dec::nblocks(nt)), dim3(256), 0, 0, nv, ne, nt, m.v0.p, m.v1.p, m.e0.p,
                       m.e1.p, m.e2.p, m.p.p, center == DEC_CIRCUMCENTER ? 1 : 0, m.area.p, m.dl.p, m.da.p,
                       m.dpt.p);
  dec::check_launch("dualize");
  if (m.dim == 3) {
    // the 1-D and 2-D steps above are shared with a 3-D complex
    // (DiscreteExteriorCalculus.jl:1308-1327); the tetrahedra follow
    dualize_tets(m, center);
    m.dualized = true;
    m.center = center;
    return;
  }
  Csr vdt, eb;
  vertex_dual_tris(t, vdt);
  edge_bduals(t, eb);
  DBuf<int32_t> vptr, vidx, eptr, eidx;
  vptr.upload(vdt.ptr);
  vidx.upload(vdt.idx);
  eptr.upload(eb.ptr);
  eidx.upload(eb.idx);
  m.h0.alloc(nv);
  m.h1.alloc(ne);
  m.h2.alloc(nt);
  if (nv > 0)
    hipLaunchKernelGGL(k_hodge0, dim3(dec::nblocks(nv)), dim3(256), 0, 0, nv, vptr.p, vidx.p, m.da.p, m.h0.p);
  if (ne > 0)
    hipLaunchKernelGGL(k_hodge1, dim3(dec::nblocks(ne)), dim3(256), 0, 0, ne, eptr.p, eidx.p,
                       m.dl.p + 2 * static_cast<size_t>(ne), m.length.p, m.h1.p);
  if (nt > 0) hipLaunchKernelGGL(k_hodge2, dim3(dec::nblocks(nt)), dim3(256), 0, 0, nt, m.area.p, m.h2.p);
  dec::check_launch("dualize");
  dsync();
  m.dualized = true;
  m.center = center;
}

void ensure_dual(dec_mesh_s &m) {
  if (!m.dualized) dualize(m, DEC_CIRCUMCENTER);
}

// Binary subdivision with points propagated on the GPU.  The fine mesh's
// triangle orientation is derived from the coarse one found by orient!.
dec_mesh_s *subdivide(const dec_mesh_s &c, bool project) {
  Topo ct = c.h;
  if (!c.to_dfs) {
    std::vector<int8_t> dfs;
    if (!orient_dfs(ct, dfs)) fail(DEC_ERR_NONORIENTABLE, "The 2-simplices of the given 2D delta set are non-orientable.");
    ct.to = dfs;
  }
  Topo ft;
  refine(ct, ft, true);
  // the derived orientation is what dualize's orient! would assign; keep it
  // only as the resolved value (edge orientations stay unset -> false).
  auto *f = new dec_mesh_s();
  f->h = std::move(ft);
  f->to_dfs = true;
  upload_topology(*f);
  size_t nvf = static_cast<size_t>(c.h.nv) + c.h.ne;
  f->p.alloc(3 * nvf);
  int n = std::max(c.h.nv, c.h.ne);
  if (n > 0)
    hipLaunchKernelGGL(k_propagate, dim3(dec::nblocks(n)), dim3(256), 0, 0, c.h.nv, c.h.ne, c.v0.p, c.v1.p, c.p.p, f->p.p,
                       project ? 1 : 0);
  dec::check_launch("subdivide");
  f->hp.resize(3 * nvf);
  d2h(f->hp.data(), f->p.p, sizeof(double) * 3 * nvf);
  return f;
}

}  // namespace dec

extern "C" {

int dec_mesh_create(int32_t nv, int32_t ne, int32_t nt, const int32_t *v0, const int32_t *v1,
                    const int32_t *e0, const int32_t *e1, const int32_t *e2, const int8_t *edge_orient,
                    const int8_t *tri_orient, const double *points, int32_t index_base, dec_mesh *out) {
  DEC_TRY
  if (!out) fail(DEC_ERR_ARG, "dec_mesh_create: out is NULL");
  *out = nullptr;
  if (nv < 0 || ne < 0 || nt < 0) fail(DEC_ERR_ARG, "dec_mesh_create: negative count");
  if ((ne && (!v0 || !v1)) || (nt && (!e0 || !e1 || !e2)) || (nv && !points))
    fail(DEC_ERR_ARG, "dec_mesh_create: NULL array");
  if (index_base != 0 && index_base != 1) fail(DEC_ERR_ARG, "dec_mesh_create: index_base must be 0 or 1");
  dec::Topo t;
  t.nv = nv;
  t.ne = ne;
  t.nt = nt;
  t.v0.resize(ne);
  t.v1.resize(ne);
  for (int e = 0; e < ne; e++) {
    t.v0[e] = v0[e] - index_base;
    t.v1[e] = v1[e] - index_base;
    if (t.v0[e] < 0 || t.v0[e] >= nv || t.v1[e] < 0 || t.v1[e] >= nv)
      fail(DEC_ERR_ARG, "dec_mesh_create: edge vertex out of range");
  }
  t.e0.resize(nt);
  t.e1.resize(nt);
  t.e2.resize(nt);
  for (int i = 0; i < nt; i++) {
    t.e0[i] = e0[i] - index_base;
    t.e1[i] = e1[i] - index_base;
    t.e2[i] = e2[i] - index_base;
    for (int x : {t.e0[i], t.e1[i], t.e2[i]})
      if (x < 0 || x >= ne) fail(DEC_ERR_ARG, "dec_mesh_create: triangle edge out of range");
  }
  t.eo.assign(ne, -1);
  t.to.assign(nt, -1);
  if (edge_orient)
    for (int e = 0; e < ne; e++) t.eo[e] = edge_orient[e] < 0 ? -1 : (edge_orient[e] ? 1 : 0);
  if (tri_orient)
    for (int i = 0; i < nt; i++) t.to[i] = tri_orient[i] < 0 ? -1 : (tri_orient[i] ? 1 : 0);
  std::vector<double> pts(points, points + 3 * static_cast<size_t>(nv));
  *out = new_mesh_from_topo(std::move(t), pts);
  DEC_CATCH
}

int dec_mesh_icosphere(int32_t k, dec_mesh *out) {
  DEC_TRY
  if (!out) fail(DEC_ERR_ARG, "dec_mesh_icosphere: out is NULL");
  *out = nullptr;
  if (k < 0 || k > 11) fail(DEC_ERR_ARG, "dec_mesh_icosphere: k must be in 0..11");
  std::vector<double> p;
  std::vector<int32_t> tv;
  dec::icosahedron(p, tv);
  dec::Topo t;
  dec::glue_triangles(12, tv, true, t);
  dec_mesh_s *m = new_mesh_from_topo(std::move(t), p);
  for (int i = 0; i < k; i++) {
    dec_mesh_s *f = nullptr;
    try {
      f = dec::subdivide(*m, true);
    } catch (...) {
      delete m;
      throw;
    }
    delete m;
    m = f;
  }
  *out = m;
  DEC_CATCH
}

int dec_mesh_from_triangles(int32_t nv, int32_t nt, const double *points, const int32_t *tris, int32_t sorted,
                            int32_t index_base, dec_mesh *out) {
  DEC_TRY
  if (!out) fail(DEC_ERR_ARG, "dec_mesh_from_triangles: out is NULL");
  *out = nullptr;
  if (nv < 0 || nt < 0) fail(DEC_ERR_ARG, "dec_mesh_from_triangles: negative count");
  if ((nv > 0 && !points) || (nt > 0 && !tris)) fail(DEC_ERR_ARG, "dec_mesh_from_triangles: NULL array");
  if (index_base != 0 && index_base != 1) fail(DEC_ERR_ARG, "dec_mesh_from_triangles: index_base must be 0 or 1");
  std::vector<int32_t> tv(3 * static_cast<size_t>(nt));
  for (size_t i = 0; i < tv.size(); i++) tv[i] = tris[i] - index_base;
  dec::Topo t;
  dec::glue_triangles(nv, tv, sorted != 0, t);
  // orient!(s) right after gluing (MeshInterop.jl:129); a non-orientable
  // complex keeps its triangles unset and is refused when dualized
  std::vector<int8_t> to;
  bool oriented = dec::orient_dfs(t, to);
  if (oriented) t.to = to;
  std::vector<double> p(points, points + 3 * static_cast<size_t>(nv));
  dec_mesh_s *m = new_mesh_from_topo(std::move(t), p);
  m->to_dfs = oriented;
  *out = m;
  DEC_CATCH
}

int dec_mesh_triangulated_grid(double lx, double ly, double dx, double dy, int32_t compress, dec_mesh *out) {
  DEC_TRY
  if (!out) fail(DEC_ERR_ARG, "dec_mesh_triangulated_grid: out is NULL");
  *out = nullptr;
  std::vector<double> p;
  dec::Topo t;
  dec::triangulated_grid(lx, ly, dx, dy, compress != 0, p, t);
  *out = new_mesh_from_topo(std::move(t), p);
  DEC_CATCH
}

int dec_mesh_subdivide(dec_mesh coarse, int32_t project, dec_mesh *fine) {
  DEC_TRY
  if (!coarse || !fine) fail(DEC_ERR_ARG, "dec_mesh_subdivide: NULL handle");
  if (coarse->dim != 2) fail(DEC_ERR_UNSUPPORTED, "dec_mesh_subdivide: binary subdivision is defined for 2-D complexes");
  *fine = dec::subdivide(*coarse, project != 0);
  DEC_CATCH
}

int dec_mesh_counts(dec_mesh m, int32_t *nv, int32_t *ne, int32_t *nt) {
  DEC_TRY
  if (!m) fail(DEC_ERR_ARG, "dec_mesh_counts: NULL handle");
  if (nv) *nv = m->h.nv;
  if (ne) *ne = m->h.ne;
  if (nt) *nt = m->h.nt;
  DEC_CATCH
}

int dec_mesh_export(dec_mesh m, int32_t *v0, int32_t *v1, int32_t *e0, int32_t *e1, int32_t *e2,
                    int8_t *edge_orient, int8_t *tri_orient, double *points) {
  DEC_TRY
  if (!m) fail(DEC_ERR_ARG, "dec_mesh_export: NULL handle");
  const dec::Topo &t = m->h;
  if (v0) std::copy(t.v0.begin(), t.v0.end(), v0);
  if (v1) std::copy(t.v1.begin(), t.v1.end(), v1);
  if (e0) std::copy(t.e0.begin(), t.e0.end(), e0);
  if (e1) std::copy(t.e1.begin(), t.e1.end(), e1);
  if (e2) std::copy(t.e2.begin(), t.e2.end(), e2);
  if (edge_orient) std::copy(t.eo.begin(), t.eo.end(), edge_orient);
  if (tri_orient) std::copy(t.to.begin(), t.to.end(), tri_orient);
  if (points) std::copy(m->hp.begin(), m->hp.end(), points);
  DEC_CATCH
}

int dec_mesh_dualize(dec_mesh m, int32_t center) {
  DEC_TRY
  if (!m) fail(DEC_ERR_ARG, "dec_mesh_dualize: NULL handle");
  dec::dualize(*m, center);
  DEC_CATCH
}

int dec_mesh_export_dual(dec_mesh m, double *dual_point, double *length, double *area, double *dual_length,
                         double *dual_area) {
  DEC_TRY
  if (!m) fail(DEC_ERR_ARG, "dec_mesh_export_dual: NULL handle");
  if (!m->dualized) fail(DEC_ERR_ARG, "dec_mesh_export_dual: mesh is not dualized");
  if (m->dim != 2) fail(DEC_ERR_ARG, "dec_mesh_export_dual: a 3-D complex exports with dec_mesh3_export_dual");
  if (dual_point) dec::d2h(dual_point, m->dpt.p, sizeof(double) * m->dpt.n);
  if (length) dec::d2h(length, m->length.p, sizeof(double) * m->length.n);
  if (area) dec::d2h(area, m->area.p, sizeof(double) * m->area.n);
  if (dual_length) dec::d2h(dual_length, m->dl.p, sizeof(double) * m->dl.n);
  if (dual_area) dec::d2h(dual_area, m->da.p, sizeof(double) * m->da.n);
  DEC_CATCH
}

int dec_mesh_destroy(dec_mesh m) {
  DEC_TRY
  if (m && m->owned_by_mg) fail(DEC_ERR_ARG, "dec_mesh_destroy: mesh belongs to a multigrid hierarchy");
  delete m;
  DEC_CATCH
}

}  // extern "C"
