// Internal declarations of libdec_hip.so (not part of the C ABI).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dec_hip.h"

namespace dec {

// Errors never cross the C ABI: entry points catch Error and return its code.
struct Error {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const std::string &msg);
void set_last_error(const std::string &msg);

// ---------------------------------------------------------------------------
// Host topology (integer work: built and kept on the host, uploaded once).
// ---------------------------------------------------------------------------
struct Topo {
  int nv = 0, ne = 0, nt = 0;
  std::vector<int32_t> v0, v1;      // ∂v0 (target), ∂v1 (source)
  std::vector<int32_t> e0, e1, e2;  // ∂e0, ∂e1, ∂e2
  std::vector<int8_t> eo, to;       // orientation: 1 true, 0 false, -1 unset
};

// Tetrahedra of a 3-D complex: ∂t0..∂t3 and the orientation (1, 0, -1 unset).
struct Tets {
  int n = 0;
  std::vector<int32_t> t[4];
  std::vector<int8_t> o;
};

// glue_triangle! / glue_sorted_triangle! over a triangle list (SimplicialSets.jl:323-364).
void glue_triangles(int nv, const std::vector<int32_t> &tv, bool sorted, Topo &t);
// glue_tetrahedron! / glue_sorted_tetrahedron! over a tetrahedron list
// (SimplicialSets.jl:511-534): the triangle skeleton goes to t.
void glue_tetrahedra(int nv, const std::vector<int32_t> &tv, bool sorted, Topo &t, Tets &T);
// orient!(s, Val(3)) (SimplicialSets.jl:812-852); false if not orientable.
bool orient3_dfs(const Topo &t, const Tets &T, std::vector<int8_t> &to);
// tetrahedron_vertices / tetrahedron_edges of tetrahedron x (SimplicialSets.jl:474-492).
void tet_vertices(const Topo &t, const Tets &T, int x, int v[4]);
void tet_edges(const Topo &t, const Tets &T, int x, int e[6]);
// orient!(s, Val(2)) (SimplicialSets.jl:812-852); false if not orientable.
bool orient_dfs(const Topo &t, std::vector<int8_t> &to);
// refine(::BinarySubdivision, topo) (Multigrid.jl:130-183).  Triangle
// orientations of the children follow from the parent's: the children 4t,
// 4t+2 keep the parent's orientation and 4t+1, 4t+3 take the opposite, which
// is the unique consistent orientation orient! finds when the parent
// orientation is itself the one orient! found.
void refine(const Topo &c, Topo &f, bool derive_orientation);
void icosahedron(std::vector<double> &p, std::vector<int32_t> &tv);
void triangulated_grid(double lx, double ly, double dx, double dy, bool compress,
                       std::vector<double> &p, Topo &t);

// CSR list with a payload per entry.
struct Csr {
  int n = 0;
  std::vector<int32_t> ptr, idx;
  std::vector<int8_t> val;
};
// vertex -> incident edges ascending, payload d0[e, v] (σ_e or -σ_e)
void vertex_edges(const Topo &t, Csr &out);
// edge -> incident triangles ascending, payload d1[t, e]
void edge_triangles(const Topo &t, Csr &out);
// edge -> B-block dual edges j*nt + t ascending (elementary duals of e)
void edge_bduals(const Topo &t, Csr &out);
// vertex -> dual triangles k*nt + t ascending whose primal vertex is v
void vertex_dual_tris(const Topo &t, Csr &out);
// greedy colouring of a symmetric sparsity pattern given as CSR adjacency
int greedy_colors(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &adj,
                  std::vector<int32_t> &color);
// Four colours for the vertices of a planar graph (CSR adjacency, symmetric):
// greedy in smallest-last order, freeing a colour by a Kempe-chain swap where
// a vertex sees all four; false when no swap frees one or the chains visited
// exceed `budget` vertices in total.
bool four_colors(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &adj,
                 std::vector<int32_t> &color, long budget);
// No two adjacent vertices share a colour and every colour is in [0, ncolors).
bool proper_coloring(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &adj,
                     const std::vector<int32_t> &color, int ncolors);
// Dense coarse operator Z (row-major n x n): z = Z r solves A z = r, bordered
// with constants when A annihilates them.  A given as dense row-major.
void coarse_inverse(int n, const std::vector<double> &A, std::vector<double> &Z);

// ---------------------------------------------------------------------------
// Device buffers
// ---------------------------------------------------------------------------
void *dmalloc(size_t bytes);
void dfree(void *p);
void h2d(void *dst, const void *src, size_t bytes);
void d2h(void *dst, const void *src, size_t bytes);
void dsync();

template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { reset(); }
  void alloc(size_t count) {
    reset();
    n = count;
    p = static_cast<T *>(dmalloc(sizeof(T) * (count ? count : 1)));
  }
  void upload(const std::vector<T> &v) {
    alloc(v.size());
    if (!v.empty()) h2d(p, v.data(), sizeof(T) * v.size());
  }
  std::vector<T> download() const {
    std::vector<T> v(n);
    if (n) d2h(v.data(), p, sizeof(T) * n);
    return v;
  }
  void reset() {
    if (p) dfree(p);
    p = nullptr;
    n = 0;
  }
};

// Slot-major ELL: entry k of row r at [k * nrows + r]; col < 0 pads.
template <class V>
struct Ell {
  int nrows = 0, width = 0;
  DBuf<int32_t> col;
  DBuf<V> val;
};

}  // namespace dec

// ---------------------------------------------------------------------------
// Handles
// ---------------------------------------------------------------------------
struct dec_mesh_s {
  int dim = 2;                  // 2: triangle complex; 3: tetrahedral complex
  dec::Topo h;                  // host topology (orientations resolved at dualize)
  dec::Tets ht;                 // host tetrahedra (dim 3)
  std::vector<double> hp;       // host points, nv x 3
  dec::DBuf<int32_t> v0, v1, e0, e1, e2;
  dec::DBuf<int8_t> eo, to;
  dec::DBuf<double> p;          // nv x 3
  bool dualized = false;
  int center = DEC_CIRCUMCENTER;
  // dual geometry, float64, reference numbering of DualE / DualTri blocks
  dec::DBuf<double> length, area;   // ne, nt
  dec::DBuf<double> dl;             // 2ne + 6nt
  dec::DBuf<double> da;             // 6nt
  dec::DBuf<double> dpt;            // (nv+ne+nt) x 3
  dec::DBuf<double> h0, h1, h2;     // diagonal Hodge stars
  // 3-D: tetrahedra, volumes, ⋆3, and per tetrahedron the dual volume at each
  // of its vertices (slot-major 4 x ntet, tetrahedron_vertices order)
  dec::DBuf<int32_t> t0, t1, t2, t3;
  dec::DBuf<int8_t> teo;
  dec::DBuf<double> vol, h3, tw;
  bool owned_by_mg = false;
  bool to_dfs = false;          // h.to is the orientation orient! assigns
};
