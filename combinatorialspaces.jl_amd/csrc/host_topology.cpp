// Host-side integer topology of libdec_hip.so: primal complex assembly,
// orientation, binary refinement, incidence lists and colouring.  All
// floating-point work (points, geometry, coefficients) runs on the GPU.
#include <algorithm>
#include <queue>
#include <tuple>
#include <cmath>
#include <cstring>

#include "dec_internal.h"

namespace dec {

// glue_triangle!(s, v0, v1, v2) = add_triangle!(get_edge!(v0,v1), get_edge!(v1,v2),
// get_edge!(v0,v2)) with ∂e0 = e(v1,v2), ∂e1 = e(v0,v2), ∂e2 = e(v0,v1)
// (SimplicialSets.jl:323-351); get_edge! reuses the first edge (ascending id)
// with ∂v1 == src and ∂v0 == tgt; glue_sorted_triangle! sorts first (:361-364).
void glue_triangles(int nv, const std::vector<int32_t> &tv, bool sorted, Topo &t) {
  int nt = static_cast<int>(tv.size() / 3);
  t.nv = nv;
  t.nt = nt;
  t.v0.clear();
  t.v1.clear();
  t.e0.assign(nt, 0);
  t.e1.assign(nt, 0);
  t.e2.assign(nt, 0);
  std::vector<int32_t> head(nv, -1), tail(nv, -1), next;
  next.reserve(3 * nt);
  auto get_edge = [&](int src, int tgt) {
    if (src < 0 || src >= nv || tgt < 0 || tgt >= nv) fail(DEC_ERR_ARG, "triangle vertex out of range");
    for (int e = head[src]; e >= 0; e = next[e])
      if (t.v0[e] == tgt) return e;
    int e = static_cast<int>(t.v0.size());
    t.v0.push_back(tgt);
    t.v1.push_back(src);
    next.push_back(-1);
    if (tail[src] < 0) head[src] = e;
    else next[tail[src]] = e;
    tail[src] = e;
    return e;
  };
  for (int i = 0; i < nt; i++) {
    int a = tv[3 * i], b = tv[3 * i + 1], c = tv[3 * i + 2];
    if (sorted) {
      if (a > b) std::swap(a, b);
      if (b > c) std::swap(b, c);
      if (a > b) std::swap(a, b);
    }
    int eab = get_edge(a, b), ebc = get_edge(b, c), eac = get_edge(a, c);
    t.e0[i] = ebc;
    t.e1[i] = eac;
    t.e2[i] = eab;
  }
  t.ne = static_cast<int>(t.v0.size());
  t.eo.assign(t.ne, -1);
  t.to.assign(nt, -1);
}

// Depth-first orientation.  The traversal visits cofaces (face index j, then
// ascending triangle id) and pushes on a LIFO stack as orient! does; the
// orientation it returns on a connected orientable complex is unique given the
// seed, and each component's seed is its lowest triangle.
bool orient_dfs(const Topo &t, std::vector<int8_t> &to) {
  int ne = t.ne, nt = t.nt;
  to.assign(nt, 1);
  if (nt == 0) return true;
  const int32_t *E[3] = {t.e0.data(), t.e1.data(), t.e2.data()};
  std::vector<int32_t> ptr[3], lst[3];
  for (int j = 0; j < 3; j++) {
    ptr[j].assign(ne + 1, 0);
    lst[j].assign(nt, 0);
    for (int i = 0; i < nt; i++) ptr[j][E[j][i] + 1]++;
    for (int e = 0; e < ne; e++) ptr[j][e + 1] += ptr[j][e];
    std::vector<int32_t> w(ptr[j].begin(), ptr[j].end() - 1);
    for (int i = 0; i < nt; i++) lst[j][w[E[j][i]]++] = i;
  }
  std::vector<int8_t> ors(nt, 0);
  std::vector<int32_t> stack;
  stack.reserve(nt);
  for (int seed = 0; seed < nt; seed++) {
    if (ors[seed]) continue;
    stack.clear();
    stack.push_back(seed);
    ors[seed] = 1;
    while (!stack.empty()) {
      int x = stack.back();
      stack.pop_back();
      int ox = ors[x];
      for (int i = 0; i < 3; i++) {
        int face = E[i][x];
        for (int j = 0; j < 3; j++) {
          bool same = ((i + j) % 2) == 0;
          for (int k = ptr[j][face]; k < ptr[j][face + 1]; k++) {
            int y = lst[j][k];
            if (y == x) continue;
            if (ors[y] == 0) {
              ors[y] = static_cast<int8_t>(same ? -ox : ox);
              stack.push_back(y);
            } else if (same && ors[y] == ox) {
              return false;
            }
          }
        }
      }
    }
  }
  for (int i = 0; i < nt; i++) to[i] = ors[i] == 1 ? 1 : 0;
  return true;
}

// glue_tetrahedron!(s, v0, v1, v2, v3) = add_tetrahedron!(get_triangle!(v1,v2,v3),
// get_triangle!(v0,v2,v3), get_triangle!(v0,v1,v3), get_triangle!(v0,v1,v2))
// (SimplicialSets.jl:511-527).  get_triangle! finds the triangle whose edges
// are v1→v2, v0→v2, v0→v1 as ∂e0, ∂e1, ∂e2; since every triangle is glued
// from its vertex triple, that is the triangle glued from (v0, v1, v2), kept
// here in a list per first vertex.  Otherwise glue_triangle! adds it.
void glue_tetrahedra(int nv, const std::vector<int32_t> &tv, bool sorted, Topo &t, Tets &T) {
  int n = static_cast<int>(tv.size() / 4);
  t = Topo();
  t.nv = nv;
  T = Tets();
  T.n = n;
  for (auto &a : T.t) a.assign(n, 0);
  std::vector<int32_t> ehead(nv, -1), etail(nv, -1), enext;
  std::vector<int32_t> thead(nv, -1), ttail(nv, -1), tnext, tb, tc;
  auto get_edge = [&](int src, int tgt) {
    for (int e = ehead[src]; e >= 0; e = enext[e])
      if (t.v0[e] == tgt) return e;
    int e = static_cast<int>(t.v0.size());
    t.v0.push_back(tgt);
    t.v1.push_back(src);
    enext.push_back(-1);
    if (etail[src] < 0) ehead[src] = e;
    else enext[etail[src]] = e;
    etail[src] = e;
    return e;
  };
  auto get_tri = [&](int a, int b, int c) {
    for (int x = thead[a]; x >= 0; x = tnext[x])
      if (tb[x] == b && tc[x] == c) return x;
    int eab = get_edge(a, b), ebc = get_edge(b, c), eac = get_edge(a, c);
    int x = static_cast<int>(t.e0.size());
    t.e0.push_back(ebc);
    t.e1.push_back(eac);
    t.e2.push_back(eab);
    tb.push_back(b);
    tc.push_back(c);
    tnext.push_back(-1);
    if (ttail[a] < 0) thead[a] = x;
    else tnext[ttail[a]] = x;
    ttail[a] = x;
    return x;
  };
  for (int i = 0; i < n; i++) {
    int v[4] = {tv[4 * i], tv[4 * i + 1], tv[4 * i + 2], tv[4 * i + 3]};
    for (int k = 0; k < 4; k++)
      if (v[k] < 0 || v[k] >= nv) fail(DEC_ERR_ARG, "tetrahedron vertex out of range");
    if (sorted) std::sort(v, v + 4);
    for (int a = 0; a < 4; a++)
      for (int b = a + 1; b < 4; b++)
        if (v[a] == v[b]) fail(DEC_ERR_ARG, "tetrahedron with a repeated vertex");
    T.t[0][i] = get_tri(v[1], v[2], v[3]);
    T.t[1][i] = get_tri(v[0], v[2], v[3]);
    T.t[2][i] = get_tri(v[0], v[1], v[3]);
    T.t[3][i] = get_tri(v[0], v[1], v[2]);
  }
  t.ne = static_cast<int>(t.v0.size());
  t.nt = static_cast<int>(t.e0.size());
  t.eo.assign(t.ne, -1);
  t.to.assign(t.nt, -1);
  T.o.assign(n, -1);
}

// orient!(s, Val(3)): the depth-first pass of orient_dfs over tetrahedra and
// their four faces.
bool orient3_dfs(const Topo &t, const Tets &T, std::vector<int8_t> &to) {
  int nt = t.nt, n = T.n;
  to.assign(n, 1);
  if (n == 0) return true;
  std::vector<int32_t> ptr[4], lst[4];
  for (int j = 0; j < 4; j++) {
    ptr[j].assign(nt + 1, 0);
    lst[j].assign(n, 0);
    for (int i = 0; i < n; i++) ptr[j][T.t[j][i] + 1]++;
    for (int f = 0; f < nt; f++) ptr[j][f + 1] += ptr[j][f];
    std::vector<int32_t> w(ptr[j].begin(), ptr[j].end() - 1);
    for (int i = 0; i < n; i++) lst[j][w[T.t[j][i]]++] = i;
  }
  std::vector<int8_t> ors(n, 0);
  std::vector<int32_t> stack;
  stack.reserve(n);
  for (int seed = 0; seed < n; seed++) {
    if (ors[seed]) continue;
    stack.clear();
    stack.push_back(seed);
    ors[seed] = 1;
    while (!stack.empty()) {
      int x = stack.back();
      stack.pop_back();
      int ox = ors[x];
      for (int i = 0; i < 4; i++) {
        int face = T.t[i][x];
        for (int j = 0; j < 4; j++) {
          bool same = ((i + j) % 2) == 0;
          for (int k = ptr[j][face]; k < ptr[j][face + 1]; k++) {
            int y = lst[j][k];
            if (y == x) continue;
            if (ors[y] == 0) {
              ors[y] = static_cast<int8_t>(same ? -ox : ox);
              stack.push_back(y);
            } else if (same && ors[y] == ox) {
              return false;
            }
          }
        }
      }
    }
  }
  for (int i = 0; i < n; i++) to[i] = ors[i] == 1 ? 1 : 0;
  return true;
}

void tet_vertices(const Topo &t, const Tets &T, int x, int v[4]) {
  int a = T.t[2][x], b = T.t[0][x];
  v[0] = t.v1[t.e2[a]];
  v[1] = t.v0[t.e2[a]];
  v[2] = t.v1[t.e0[b]];
  v[3] = t.v0[t.e0[b]];
}

void tet_edges(const Topo &t, const Tets &T, int x, int e[6]) {
  int a = T.t[0][x], b = T.t[1][x], c = T.t[2][x];
  e[0] = t.e0[a];
  e[1] = t.e1[a];
  e[2] = t.e2[a];
  e[3] = t.e1[b];
  e[4] = t.e2[b];
  e[5] = t.e2[c];
}

// refine(::BinarySubdivision, topo), 0-based: coarse edge e -> fine edges 2e
// (∂v1 = ∂v0_c[e]) and 2e+1 (∂v1 = ∂v1_c[e]), both with ∂v0 = nv + e; interior
// edges 2ne + 3t + {0,1,2}; children 4t + {0,1,2,3}.
void refine(const Topo &c, Topo &f, bool derive_orientation) {
  int nv = c.nv, ne = c.ne, nt = c.nt;
  f.nv = nv + ne;
  f.ne = 2 * ne + 3 * nt;
  f.nt = 4 * nt;
  f.v0.assign(f.ne, 0);
  f.v1.assign(f.ne, 0);
  f.e0.assign(f.nt, 0);
  f.e1.assign(f.nt, 0);
  f.e2.assign(f.nt, 0);
  for (int e = 0; e < ne; e++) {
    int m = nv + e;
    f.v0[2 * e] = m;
    f.v1[2 * e] = c.v0[e];
    f.v0[2 * e + 1] = m;
    f.v1[2 * e + 1] = c.v1[e];
  }
  for (int t = 0; t < nt; t++) {
    int m0 = c.e0[t] + nv, m1 = c.e1[t] + nv, m2 = c.e2[t] + nv;
    int m0_m1 = 2 * ne + 3 * t, m1_m2 = m0_m1 + 1, m0_m2 = m0_m1 + 2;
    f.v0[m0_m1] = m1; f.v1[m0_m1] = m0;
    f.v0[m1_m2] = m2; f.v1[m1_m2] = m1;
    f.v0[m0_m2] = m2; f.v1[m0_m2] = m0;
    int m0_v1 = 2 * c.e0[t] + 1, v2_m0 = 2 * c.e0[t];
    int m1_v0 = 2 * c.e1[t] + 1, v2_m1 = 2 * c.e1[t];
    int m2_v0 = 2 * c.e2[t] + 1, v1_m2 = 2 * c.e2[t];
    int b = 4 * t;
    f.e0[b] = m1_m2;     f.e1[b] = m0_m2;     f.e2[b] = m0_m1;
    f.e0[b + 1] = m1_m2; f.e1[b + 1] = m2_v0; f.e2[b + 1] = m1_v0;
    f.e0[b + 2] = m0_m2; f.e1[b + 2] = v1_m2; f.e2[b + 2] = m0_v1;
    f.e0[b + 3] = m0_m1; f.e1[b + 3] = v2_m1; f.e2[b + 3] = v2_m0;
  }
  f.eo.assign(f.ne, -1);
  f.to.assign(f.nt, -1);
  if (derive_orientation) {
    // Child k of a parent with vertices (v0, v1, v2) has vertex order
    // (m0,m1,m2), (v0,m1,m2), (v1,m0,m2), (v2,m0,m1): children 0 and 2 turn
    // the same way as the parent, 1 and 3 the other way.
    for (int t = 0; t < nt; t++) {
      int8_t o = c.to[t];
      f.to[4 * t] = o;
      f.to[4 * t + 1] = static_cast<int8_t>(!o);
      f.to[4 * t + 2] = o;
      f.to[4 * t + 3] = static_cast<int8_t>(!o);
    }
  }
}

void icosahedron(std::vector<double> &p, std::vector<int32_t> &tv) {
  static const double P[12][3] = {
      {0.0, 0.0, 1.0},
      {0.8944271909999159, 0.0, 0.4472135954999579},
      {0.27639320225002106, 0.8506508083520399, 0.4472135954999579},
      {-0.7236067977499788, 0.5257311121191337, 0.4472135954999579},
      {-0.723606797749979, -0.5257311121191335, 0.4472135954999579},
      {0.27639320225002084, -0.85065080835204, 0.4472135954999579},
      {0.7236067977499789, 0.5257311121191336, -0.4472135954999579},
      {-0.27639320225002095, 0.85065080835204, -0.4472135954999579},
      {-0.8944271909999159, 0.0, -0.4472135954999579},
      {-0.2763932022500211, -0.8506508083520399, -0.4472135954999579},
      {0.7236067977499788, -0.5257311121191338, -0.4472135954999579},
      {0.0, 0.0, -1.0},
  };
  p.assign(&P[0][0], &P[0][0] + 36);
  tv.clear();
  for (int k = 0; k < 5; k++) tv.insert(tv.end(), {0, 1 + k, 1 + (k + 1) % 5});
  for (int k = 0; k < 5; k++) tv.insert(tv.end(), {1 + k, 6 + k, 1 + (k + 1) % 5});
  for (int k = 0; k < 5; k++) tv.insert(tv.end(), {1 + (k + 1) % 5, 6 + k, 6 + (k + 1) % 5});
  for (int k = 0; k < 5; k++) tv.insert(tv.end(), {11, 6 + k, 6 + (k + 1) % 5});
}

static int range_len(double l, double d) {
  double q = l / d;
  double r = std::floor(q + 0.5);
  if (std::fabs(q - r) <= 1e-10 * (r > 1 ? r : 1)) return static_cast<int>(r) + 1;
  return static_cast<int>(std::floor(q)) + 1;
}

// triangulated_grid(lx, ly, dx, dy, Point3d, compress) (CombMeshes.jl:79-130).
void triangulated_grid(double lx, double ly, double dx, double dy, bool compress,
                       std::vector<double> &p, Topo &t) {
  if (!(lx > 0 && ly > 0 && dx > 0 && dy > 0)) fail(DEC_ERR_ARG, "triangulated_grid: sizes must be positive");
  int nx = range_len(lx, dx), ny = range_len(ly, dy);
  if (nx < 2 || ny < 2) fail(DEC_ERR_ARG, "triangulated_grid: fewer than 2 points per side");
  double scale = lx / (lx + dx / 2), shift = dx / 2;
  p.assign(3 * static_cast<size_t>(nx) * ny, 0.0);
  for (int yi = 1; yi <= ny; yi++) {
    double raw_y = static_cast<double>(yi - 1) * dy;
    for (int xi = 1; xi <= nx; xi++) {
      double xc = static_cast<double>(xi - 1) * dx + shift * static_cast<double>((yi % 2) == 0);
      if (compress) xc *= scale;
      size_t i = static_cast<size_t>(xi - 1) + static_cast<size_t>(nx) * (yi - 1);
      p[3 * i] = xc;
      p[3 * i + 1] = raw_y;
      p[3 * i + 2] = 0.0;
    }
  }
  std::vector<int32_t> tv;
  for (int y = 1; y <= ny - 1; y++)
    for (int x = 1; x <= nx - 1; x++) {
      int i = x + nx * (y - 1) - 1;  // 0-based
      if (y % 2 == 0) {
        tv.insert(tv.end(), {i, i + nx, i + nx + 1});
        tv.insert(tv.end(), {i, i + 1, i + nx + 1});
      } else {
        tv.insert(tv.end(), {i, i + 1, i + nx});
        tv.insert(tv.end(), {i + 1, i + nx, i + nx + 1});
      }
    }
  glue_triangles(nx * ny, tv, false, t);
  t.eo.assign(t.ne, 0);
  int nxtri = 2 * (nx - 1), nytri = ny - 1;
  for (int y = 1; y <= nytri; y++) {
    int8_t o = (y % 2) != 0;
    for (int x = 1; x <= nxtri; x += 2) {
      int i = x + nxtri * (y - 1) - 1;
      t.to[i] = o;
      t.to[i + 1] = static_cast<int8_t>(!o);
    }
  }
}

void vertex_edges(const Topo &t, Csr &out) {
  out.n = t.nv;
  out.ptr.assign(t.nv + 1, 0);
  for (int e = 0; e < t.ne; e++) {
    out.ptr[t.v0[e] + 1]++;
    if (t.v1[e] != t.v0[e]) out.ptr[t.v1[e] + 1]++;
  }
  for (int v = 0; v < t.nv; v++) out.ptr[v + 1] += out.ptr[v];
  out.idx.assign(out.ptr[t.nv], 0);
  out.val.assign(out.ptr[t.nv], 0);
  std::vector<int32_t> w(out.ptr.begin(), out.ptr.end() - 1);
  for (int e = 0; e < t.ne; e++) {  // ascending e => each row sorted
    int8_t s = t.eo[e] == 1 ? 1 : -1;
    int a = t.v0[e], b = t.v1[e];
    if (a == b) continue;
    out.idx[w[a]] = e; out.val[w[a]++] = s;
    out.idx[w[b]] = e; out.val[w[b]++] = static_cast<int8_t>(-s);
  }
}

void edge_triangles(const Topo &t, Csr &out) {
  out.n = t.ne;
  out.ptr.assign(t.ne + 1, 0);
  const std::vector<int32_t> *E[3] = {&t.e0, &t.e1, &t.e2};
  for (int i = 0; i < t.nt; i++)
    for (int j = 0; j < 3; j++) out.ptr[(*E[j])[i] + 1]++;
  for (int e = 0; e < t.ne; e++) out.ptr[e + 1] += out.ptr[e];
  out.idx.assign(out.ptr[t.ne], 0);
  out.val.assign(out.ptr[t.ne], 0);
  std::vector<int32_t> w(out.ptr.begin(), out.ptr.end() - 1);
  for (int i = 0; i < t.nt; i++) {  // ascending t
    int8_t ts = t.to[i] == 1 ? 1 : -1;
    for (int j = 0; j < 3; j++) {
      int e = (*E[j])[i];
      int8_t es = t.eo[e] == 1 ? 1 : -1;
      int8_t v = static_cast<int8_t>((j == 1 ? -1 : 1) * es * ts);
      out.idx[w[e]] = i;
      out.val[w[e]++] = v;
    }
  }
}

void edge_bduals(const Topo &t, Csr &out) {
  out.n = t.ne;
  out.ptr.assign(t.ne + 1, 0);
  const std::vector<int32_t> *E[3] = {&t.e0, &t.e1, &t.e2};
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < t.nt; i++) out.ptr[(*E[j])[i] + 1]++;
  for (int e = 0; e < t.ne; e++) out.ptr[e + 1] += out.ptr[e];
  out.idx.assign(out.ptr[t.ne], 0);
  out.val.assign(out.ptr[t.ne], 0);
  std::vector<int32_t> w(out.ptr.begin(), out.ptr.end() - 1);
  for (int j = 0; j < 3; j++)  // ascending j * nt + i
    for (int i = 0; i < t.nt; i++) {
      int e = (*E[j])[i];
      out.idx[w[e]] = j * t.nt + i;
      out.val[w[e]++] = static_cast<int8_t>(j);
    }
}

// Primal vertex of dual triangle block k: schema v of (v,e,ev) =
// (0,1,1),(0,2,1),(1,2,0),(1,0,1),(2,0,0),(2,1,0); the vertex is ∂v_ev[∂e_e[t]]
// (the D_∂v1 of D_∂e2, FastDEC.jl:56) and, equivalently, triangle_vertices[v].
static const int DT_SCHEMA[6][3] = {{0, 1, 1}, {0, 2, 1}, {1, 2, 0}, {1, 0, 1}, {2, 0, 0}, {2, 1, 0}};

void vertex_dual_tris(const Topo &t, Csr &out) {
  out.n = t.nv;
  out.ptr.assign(t.nv + 1, 0);
  auto vert = [&](int k, int i) {
    // ⋆0 sums dual_area over dual triangles with D_∂v1[D_∂e1[dt]] == v
    // (FastDEC.jl:580-587); D_∂e1 is C_v(t) whose D_∂v1 is triangle_vertices[v].
    int v = DT_SCHEMA[k][0];
    return v == 0 ? t.v1[t.e1[i]] : (v == 1 ? t.v0[t.e2[i]] : t.v0[t.e1[i]]);
  };
  for (int k = 0; k < 6; k++)
    for (int i = 0; i < t.nt; i++) out.ptr[vert(k, i) + 1]++;
  for (int v = 0; v < t.nv; v++) out.ptr[v + 1] += out.ptr[v];
  out.idx.assign(out.ptr[t.nv], 0);
  std::vector<int32_t> w(out.ptr.begin(), out.ptr.end() - 1);
  for (int k = 0; k < 6; k++)
    for (int i = 0; i < t.nt; i++) out.idx[w[vert(k, i)]++] = k * t.nt + i;
}

int greedy_colors(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &adj,
                  std::vector<int32_t> &color) {
  color.assign(n, -1);
  std::vector<int32_t> used(n + 2, -1);
  int nc = 0;
  for (int v = 0; v < n; v++) {
    for (int k = ptr[v]; k < ptr[v + 1]; k++) {
      int w = adj[k];
      if (w != v && color[w] >= 0) used[color[w]] = v;
    }
    int c = 0;
    while (used[c] == v) c++;
    color[v] = c;
    nc = std::max(nc, c + 1);
  }
  return nc;
}

namespace {

bool four_colors_once(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &adj,
                      std::vector<int32_t> &color, long budget, uint32_t salt) {
  color.assign(n, -1);
  // ties of the removal order are broken by a hash of (vertex, salt)
  auto key = [salt](int v) {
    uint32_t h = static_cast<uint32_t>(v) * 2654435761u ^ salt * 0x9e3779b9u;
    h ^= h >> 15;
    return static_cast<int>(h & 0x7fffffff);
  };
  // smallest-last order: repeatedly remove a vertex of least remaining degree
  std::vector<int32_t> deg(n), order;
  order.reserve(n);
  std::vector<char> gone(n, 0);
  using Item = std::tuple<int, int, int>;  // (remaining degree, tie key, vertex)
  std::priority_queue<Item, std::vector<Item>, std::greater<>> pq;
  for (int v = 0; v < n; v++) {
    for (int k = ptr[v]; k < ptr[v + 1]; k++) deg[v] += adj[k] != v;
    pq.push({deg[v], salt ? key(v) : v, v});
  }
  while (!pq.empty()) {
    auto [d, t, v] = pq.top();
    pq.pop();
    if (gone[v] || d != deg[v]) continue;
    gone[v] = 1;
    order.push_back(v);
    for (int k = ptr[v]; k < ptr[v + 1]; k++) {
      int w = adj[k];
      if (w != v && !gone[w]) {
        --deg[w];
        pq.push({deg[w], salt ? key(w) : w, w});
      }
    }
  }
  // colour in reverse removal order; a vertex that sees all four colours is
  // freed by swapping the two colours a, b on the Kempe chains through its
  // a-coloured neighbours, when those chains reach none of its b-coloured ones
  std::vector<int32_t> stamp(n, 0), queue;
  int gen = 0;
  long work = 0;
  for (int i = n - 1; i >= 0; i--) {
    int v = order[i];
    int used = 0;
    for (int k = ptr[v]; k < ptr[v + 1]; k++)
      if (adj[k] != v && color[adj[k]] >= 0) used |= 1 << color[adj[k]];
    int c = 0;
    while (c < 4 && (used >> c & 1)) c++;
    if (c < 4) {
      color[v] = c;
      continue;
    }
    bool freed = false;
    for (int a = 0; a < 4 && !freed; a++)
      for (int b = 0; b < 4 && !freed; b++) {
        if (a == b) continue;
        ++gen;
        queue.clear();
        for (int k = ptr[v]; k < ptr[v + 1]; k++) {
          int w = adj[k];
          if (w != v && color[w] == a && stamp[w] != gen) {
            stamp[w] = gen;
            queue.push_back(w);
          }
        }
        bool blocked = false;
        for (size_t q = 0; q < queue.size() && !blocked; q++) {
          int x = queue[q];
          for (int k = ptr[x]; k < ptr[x + 1]; k++) {
            int y = adj[k];
            if (y == x || stamp[y] == gen || (color[y] != a && color[y] != b)) continue;
            stamp[y] = gen;
            queue.push_back(y);
          }
        }
        work += static_cast<long>(queue.size());
        if (work > budget) return false;
        for (int k = ptr[v]; k < ptr[v + 1] && !blocked; k++)
          if (adj[k] != v && color[adj[k]] == b && stamp[adj[k]] == gen) blocked = true;
        if (blocked) continue;
        for (int x : queue) color[x] = color[x] == a ? b : a;
        color[v] = a;
        freed = true;
      }
    if (!freed) return false;
  }
  return true;
}

}  // namespace

bool four_colors(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &adj,
                 std::vector<int32_t> &color, long budget) {
  for (uint32_t salt = 0; salt < 8; salt++)
    if (four_colors_once(n, ptr, adj, color, budget, salt)) return true;
  color.clear();
  return false;
}

bool proper_coloring(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &adj,
                     const std::vector<int32_t> &color, int ncolors) {
  if (static_cast<int>(color.size()) != n) return false;
  for (int v = 0; v < n; v++) {
    if (color[v] < 0 || color[v] >= ncolors) return false;
    for (int k = ptr[v]; k < ptr[v + 1]; k++)
      if (adj[k] != v && color[adj[k]] == color[v]) return false;
  }
  return true;
}

void coarse_inverse(int n, const std::vector<double> &A, std::vector<double> &Z) {
  double amax = 0.0, rmax = 0.0;
  for (int i = 0; i < n; i++) {
    double rs = 0.0;
    for (int j = 0; j < n; j++) {
      rs += A[static_cast<size_t>(i) * n + j];
      amax = std::max(amax, std::fabs(A[static_cast<size_t>(i) * n + j]));
    }
    rmax = std::max(rmax, std::fabs(rs));
  }
  bool sing = rmax <= 1e-10 * amax;
  int N = n + (sing ? 1 : 0), W = 2 * N;
  std::vector<double> M(static_cast<size_t>(N) * W, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) M[static_cast<size_t>(i) * W + j] = A[static_cast<size_t>(i) * n + j];
  if (sing)
    for (int i = 0; i < n; i++) {
      M[static_cast<size_t>(i) * W + n] = 1.0;
      M[static_cast<size_t>(n) * W + i] = 1.0;
    }
  for (int i = 0; i < N; i++) M[static_cast<size_t>(i) * W + N + i] = 1.0;
  for (int c = 0; c < N; c++) {
    int piv = c;
    for (int r = c + 1; r < N; r++)
      if (std::fabs(M[static_cast<size_t>(r) * W + c]) > std::fabs(M[static_cast<size_t>(piv) * W + c])) piv = r;
    if (M[static_cast<size_t>(piv) * W + c] == 0.0) fail(DEC_ERR_ARG, "coarse operator is singular");
    if (piv != c)
      for (int j = 0; j < W; j++) std::swap(M[static_cast<size_t>(c) * W + j], M[static_cast<size_t>(piv) * W + j]);
    double d = M[static_cast<size_t>(c) * W + c];
    for (int j = 0; j < W; j++) M[static_cast<size_t>(c) * W + j] /= d;
    for (int r = 0; r < N; r++) {
      if (r == c) continue;
      double f = M[static_cast<size_t>(r) * W + c];
      if (f == 0.0) continue;
      for (int j = 0; j < W; j++) M[static_cast<size_t>(r) * W + j] -= f * M[static_cast<size_t>(c) * W + j];
    }
  }
  Z.assign(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Z[static_cast<size_t>(i) * n + j] = M[static_cast<size_t>(i) * W + N + j];
}

}  // namespace dec
