// Batched explicit heat steps u ← u - dt L u over many meshes of uneven size
// (icosphere levels mixed per instance) in one persistent launch per step.
//
// Instances are concatenated: instance i owns rows [off_i, off_i + n_i) of the
// state vector.  Instances on the same mesh share one Laplacian (slot-major
// ELL of -⋆0⁻¹ d̃1 ⋆1 d0 with the diagonal Hodge).  The work list holds one
// item per (instance, tile of rows); each workgroup walks the list with a
// fixed stride (no atomics: the launch is deterministic and can be replayed
// from a hipGraph).  Block size and rows per thread come from an autotuner
// that times the candidate shapes on the batch itself and caches the winner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "dec_internal.h"
#include "ell_rows.h"
#include "kernels_common.h"

using dec::DBuf;
using dec::fail;

namespace {

struct Item {
  int inst;   // instance
  int row0;   // first local row of the tile
  int rows;   // rows in the tile
};

struct OpRef {  // a Laplacian shared by the instances on one mesh
  int n = 0, width = 0;
  DBuf<int32_t> col;
  DBuf<double> v64;
  DBuf<float> v32;
};

struct InstDev {
  long long off;   // row offset in the concatenated state
  int op;          // index of the shared Laplacian
  int n;           // rows
  double dt;       // explicit Euler step
};

template <class T>
struct OpPtrs {
  const int *col;
  const T *val;
  int n, width;
};

// One explicit Euler step for every instance: dst = src - dt * (L src).
template <class T, int RPT>
__global__ void __launch_bounds__(1024) k_heat_step(int nitems, const Item *__restrict__ items,
                                                    const InstDev *__restrict__ inst,
                                                    const OpPtrs<T> *__restrict__ ops, const T *__restrict__ src,
                                                    T *__restrict__ dst) {
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    Item w = items[it];
    InstDev I = inst[w.inst];
    OpPtrs<T> O = ops[I.op];
    const T *u = src + I.off;
    T *v = dst + I.off;
    T dt = static_cast<T>(I.dt);
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      int r = w.row0 + j * blockDim.x + threadIdx.x;
      if (r >= w.row0 + w.rows) break;
      T acc = O.width == 7 ? dec_ell::row<7, T, T>(O.n, O.width, r, O.col, O.val, u)
                           : dec_ell::row<0, T, T>(O.n, O.width, r, O.col, O.val, u);
      v[r] = u[r] - dt * acc;
    }
  }
}

template <class T>
__global__ void k_copy_rows(long long n, const T *__restrict__ x, T *__restrict__ y) {
  long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// Partner buffer of dec_batch_heat_steps_noalloc: a fixed device array, so
// that entry point allocates nothing.  Each GPU the library runs on holds its
// own copy until the process ends.  Host code never passes `scratch` itself
// (in host code the name is a stand-in, not a device address): it passes the
// address hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct Shape {
  int block = 256;
  int rpt = 1;
  int grid_mult = 4;  // workgroups = grid_mult * 256 CUs (capped by items)
};

}  // namespace

struct dec_batch_s {
  int ninst = 0, dtype = DEC_F32;
  long long total = 0;
  std::vector<std::unique_ptr<OpRef>> ops;
  std::vector<InstDev> inst;
  DBuf<InstDev> dinst;
  DBuf<OpPtrs<double>> dops64;
  DBuf<OpPtrs<float>> dops32;
  DBuf<unsigned char> tmp;  // ping-pong partner of the caller's state
  Shape shape;
  bool tuned = false;
  DBuf<Item> items;
  int nitems = 0;
  // cached graph of `gsteps` steps for state pointer `gu`
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  void *gu = nullptr;
  int gsteps = -1;
  hipStream_t cap = nullptr;
  ~dec_batch_s() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {

// The work list, laid out for the XCDs.  Workgroup w runs on XCD w mod 8,
// and the grid is a multiple of 8, so item i is always served by XCD i mod 8.
// Whole instances are dealt to the 8 XCDs, largest first onto the least
// loaded, and item 8j + x is the j-th tile of XCD x's instances (an empty
// item where x has fewer tiles).  An instance's gathers of its state then
// stay in one XCD's L2 instead of being fetched by all eight.
constexpr int kXcds = 8;

void build_items(dec_batch_s &b) {
  int tile = b.shape.block * b.shape.rpt;
  std::vector<int> order(b.ninst);
  for (int i = 0; i < b.ninst; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b.inst[a].n > b.inst[c].n; });
  std::vector<long long> load(kXcds, 0);
  std::vector<std::vector<int>> group(kXcds);
  for (int i : order) {
    int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
    group[x].push_back(i);
    load[x] += b.inst[i].n;
  }
  std::vector<std::vector<Item>> per(kXcds);
  size_t longest = 0;
  for (int x = 0; x < kXcds; x++) {
    std::sort(group[x].begin(), group[x].end());
    for (int i : group[x])
      for (int r = 0; r < b.inst[i].n; r += tile) per[x].push_back(Item{i, r, std::min(tile, b.inst[i].n - r)});
    longest = std::max(longest, per[x].size());
  }
  std::vector<Item> items(longest * kXcds, Item{0, 0, 0});
  for (int x = 0; x < kXcds; x++)
    for (size_t j = 0; j < per[x].size(); j++) items[j * kXcds + x] = per[x][j];
  b.items.upload(items);
  b.nitems = static_cast<int>(items.size());
}

template <class T>
void launch_step(dec_batch_s &b, const T *src, T *dst, hipStream_t s) {
  int grid = std::min(b.nitems, b.shape.grid_mult * 256);
  if (grid <= 0) return;
  const OpPtrs<T> *ops;
  if constexpr (sizeof(T) == 8) ops = b.dops64.p;
  else ops = b.dops32.p;
  dim3 g(grid), blk(b.shape.block);
  switch (b.shape.rpt) {
    case 1: hipLaunchKernelGGL((k_heat_step<T, 1>), g, blk, 0, s, b.nitems, b.items.p, b.dinst.p, ops, src, dst); break;
    case 2: hipLaunchKernelGGL((k_heat_step<T, 2>), g, blk, 0, s, b.nitems, b.items.p, b.dinst.p, ops, src, dst); break;
    case 4: hipLaunchKernelGGL((k_heat_step<T, 4>), g, blk, 0, s, b.nitems, b.items.p, b.dinst.p, ops, src, dst); break;
    default: fail(DEC_ERR_ARG, "batch: bad rows per thread");
  }
}

// nsteps steps ping-ponging between the caller's state u and the partner
// buffer t (at least b.total entries); an odd count ends with a copy back.
template <class T>
void run_steps(dec_batch_s &b, T *u, T *t, int nsteps, hipStream_t s) {
  for (int k = 0; k < nsteps; k++) {
    if (k % 2 == 0) launch_step<T>(b, u, t, s);
    else launch_step<T>(b, t, u, s);
  }
  if (nsteps % 2 == 1)
    hipLaunchKernelGGL(k_copy_rows<T>, dim3(dec::nblocks(b.total)), dim3(256), 0, s, b.total, t, u);
}

void run_with(dec_batch_s &b, void *u, void *partner, int nsteps, hipStream_t s) {
  if (b.dtype == DEC_F64) run_steps<double>(b, static_cast<double *>(u), static_cast<double *>(partner), nsteps, s);
  else run_steps<float>(b, static_cast<float *>(u), static_cast<float *>(partner), nsteps, s);
}

size_t state_bytes(const dec_batch_s &b) { return static_cast<size_t>(b.total) * (b.dtype == DEC_F64 ? 8 : 4); }

// The library-owned partner buffer, allocated on first use (never under a
// capture: callers reach this before hipStreamBeginCapture).
void ensure_tmp(dec_batch_s &b) {
  if (!b.tmp.p) b.tmp.alloc(state_bytes(b));
}

// The device address of `scratch` on the current device, looked up on the
// first call made with that device current and kept per device.
void *scratch_ptr() {
  static std::mutex mu;
  static void *ptrs[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) fail(DEC_ERR_HIP, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) fail(DEC_ERR_UNSUPPORTED, "device index above the scratch table");
  std::lock_guard<std::mutex> g(mu);
  if (!ptrs[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      fail(DEC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptrs[dev] = p;
  }
  return ptrs[dev];
}

void run(dec_batch_s &b, void *u, int nsteps, hipStream_t s) {
  ensure_tmp(b);
  run_with(b, u, b.tmp.p, nsteps, s);
}

void drop_graph(dec_batch_s &b) {
  if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.gexec = nullptr;
  b.graph = nullptr;
  b.gsteps = -1;
  b.gu = nullptr;
}

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(DEC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Time candidate shapes on the batch's own work and keep the fastest.
// The autotuner's candidate launch shapes; a cached shape outside these sets
// is not trusted.
constexpr int kBlocks[] = {64, 128, 256, 512};
constexpr int kRpts[] = {1, 2, 4};
constexpr int kMults[] = {2, 4, 8};

template <size_t N>
bool one_of(const int (&set)[N], int v) {
  for (int x : set)
    if (x == v) return true;
  return false;
}

bool valid_shape(const Shape &s) {
  return one_of(kBlocks, s.block) && one_of(kRpts, s.rpt) && one_of(kMults, s.grid_mult);
}

void tune(dec_batch_s &b, void *u) {
  const auto &blocks = kBlocks;
  const auto &rpts = kRpts;
  const auto &mults = kMults;
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  Shape best = b.shape;
  float best_ms = 1e30f;
  for (int bl : blocks)
    for (int rp : rpts)
      for (int gm : mults) {
        b.shape = Shape{bl, rp, gm};
        build_items(b);
        run(b, u, 2, nullptr);  // warm
        check(hipEventRecord(e0, nullptr), "hipEventRecord");
        run(b, u, 8, nullptr);
        check(hipEventRecord(e1, nullptr), "hipEventRecord");
        check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (ms < best_ms) {
          best_ms = ms;
          best = b.shape;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b.shape = best;
  build_items(b);
  b.tuned = true;
  drop_graph(b);
}

std::string cache_key(const dec_batch_s &b) {
  // the device's architecture, the sizes of the instances and the dtype
  // determine the best shape
  unsigned long long h = 1469598103934665603ULL;
  for (auto &i : b.inst) {
    h ^= static_cast<unsigned long long>(i.n);
    h *= 1099511628211ULL;
  }
  int dev = 0;
  hipDeviceProp_t prop;
  check(hipGetDevice(&dev), "hipGetDevice");
  check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  std::string arch(prop.gcnArchName);
  arch = arch.substr(0, arch.find(':'));  // drop the feature suffix (":sramecc+:xnack-")
  char buf[160];
  snprintf(buf, sizeof buf, "heat:%s:%d:%lld:%d:%016llx", arch.c_str(), b.ninst, b.total, b.dtype, h);
  return buf;
}

bool cache_lookup(const char *path, const std::string &key, Shape &s) {
  if (!path || !*path) return false;
  FILE *f = fopen(path, "r");
  if (!f) return false;
  char k[128];
  int bl, rp, gm;
  bool found = false;
  while (fscanf(f, "%127s %d %d %d", k, &bl, &rp, &gm) == 4)
    if (key == k && valid_shape(Shape{bl, rp, gm})) {  // a later valid line wins
      s = Shape{bl, rp, gm};
      found = true;
    }
  fclose(f);
  return found;
}

void cache_store(const char *path, const std::string &key, const Shape &s) {
  if (!path || !*path) return;
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %d\n", key.c_str(), s.block, s.rpt, s.grid_mult);
  fclose(f);
}

dec_batch_s *create(int ninst, const dec_mesh *meshes, int dtype) {
  if (ninst <= 0 || !meshes) fail(DEC_ERR_ARG, "dec_batch_create: no instances");
  if (dtype != DEC_F64 && dtype != DEC_F32) fail(DEC_ERR_ARG, "dec_batch_create: bad dtype");
  auto b = std::make_unique<dec_batch_s>();
  b->ninst = ninst;
  b->dtype = dtype;
  std::map<dec_mesh, int> seen;
  std::vector<double> dts;
  for (int i = 0; i < ninst; i++) {
    dec_mesh m = meshes[i];
    if (!m) fail(DEC_ERR_ARG, "dec_batch_create: NULL mesh");
    if (m->dim != 2) fail(DEC_ERR_UNSUPPORTED, "dec_batch_create: heat steps run on 2-D complexes");
    auto it = seen.find(m);
    int oi;
    double dt;
    if (it == seen.end()) {
      dec::ensure_dual(*m);
      const dec::Topo &t = m->h;
      std::vector<double> h0 = m->h0.download(), h1 = m->h1.download(), ih0(t.nv);
      for (int v = 0; v < t.nv; v++) ih0[v] = 1 / h0[v];
      dec::HostCsr L;
      dec::assemble_laplacian0(t, ih0, h1, nullptr, L);
      auto op = std::make_unique<OpRef>();
      op->n = t.nv;
      dec::Ell<double> e64;
      dec::Ell<float> e32;
      dec::upload_ell(L, e64);
      dec::upload_ell(L, e32);
      op->width = e64.width;
      op->col = std::move(e64.col);
      op->v64 = std::move(e64.val);
      op->v32 = std::move(e32.val);
      std::vector<double> len = m->length.download();
      double hsum = 0.0;
      for (double x : len) hsum += x;
      double h = hsum / static_cast<double>(len.size());
      dt = 0.1 * h * h;  // dt = 0.1 h², h the mean edge length of the mesh
      oi = static_cast<int>(b->ops.size());
      b->ops.push_back(std::move(op));
      dts.push_back(dt);
      seen[m] = oi;
    } else {
      oi = it->second;
      dt = dts[oi];
    }
    b->inst.push_back(InstDev{b->total, oi, b->ops[oi]->n, dt});
    b->total += b->ops[oi]->n;
  }
  std::vector<OpPtrs<double>> p64;
  std::vector<OpPtrs<float>> p32;
  for (auto &o : b->ops) {
    p64.push_back(OpPtrs<double>{o->col.p, o->v64.p, o->n, o->width});
    p32.push_back(OpPtrs<float>{o->col.p, o->v32.p, o->n, o->width});
  }
  b->dops64.upload(p64);
  b->dops32.upload(p32);
  b->dinst.upload(b->inst);
  build_items(*b);
  dec::check_launch("dec_batch_create");
  dec::dsync();
  return b.release();
}

}  // namespace

extern "C" {

int dec_batch_create(int32_t ninst, const dec_mesh *meshes, int32_t dtype, dec_batch *out) {
  DEC_TRY
  if (!out) fail(DEC_ERR_ARG, "dec_batch_create: out is NULL");
  *out = nullptr;
  *out = create(ninst, meshes, dtype);
  DEC_CATCH
}

int dec_batch_layout(dec_batch b, int64_t *total, int64_t *offsets, double *dt) {
  DEC_TRY
  if (!b) fail(DEC_ERR_ARG, "dec_batch_layout: NULL handle");
  if (total) *total = b->total;
  for (int i = 0; i < b->ninst; i++) {
    if (offsets) offsets[i] = b->inst[i].off;
    if (dt) dt[i] = b->inst[i].dt;
  }
  if (offsets) offsets[b->ninst] = b->total;
  DEC_CATCH
}

int dec_batch_tune(dec_batch b, void *u, const char *cache_path, int32_t *block, int32_t *rows_per_thread,
                   int32_t *grid_mult) {
  DEC_TRY
  if (!b || !u) fail(DEC_ERR_ARG, "dec_batch_tune: NULL argument");
  std::string key = cache_key(*b);
  Shape s;
  if (cache_lookup(cache_path, key, s)) {
    b->shape = s;
    build_items(*b);
    b->tuned = true;
    drop_graph(*b);
  } else {
    // tuning runs steps on u: work on a copy so the caller's state is kept
    size_t bytes = state_bytes(*b);
    DBuf<unsigned char> work;
    work.alloc(bytes);
    check(hipMemcpy(work.p, u, bytes, hipMemcpyDeviceToDevice), "hipMemcpy");
    tune(*b, work.p);
    dec::dsync();
    cache_store(cache_path, key, b->shape);
  }
  if (block) *block = b->shape.block;
  if (rows_per_thread) *rows_per_thread = b->shape.rpt;
  if (grid_mult) *grid_mult = b->shape.grid_mult;
  DEC_CATCH
}

int dec_batch_heat_steps(dec_batch b, void *u, int32_t nsteps, int32_t use_graph, void *stream) {
  DEC_TRY
  if (!b || !u) fail(DEC_ERR_ARG, "dec_batch_heat_steps: NULL argument");
  if (nsteps < 0) fail(DEC_ERR_ARG, "dec_batch_heat_steps: nsteps must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  ensure_tmp(*b);
  if (!use_graph) {
    run(*b, u, nsteps, s);
    dec::check_launch("dec_batch_heat_steps");
    return DEC_OK;
  }
  if (!b->gexec || b->gu != u || b->gsteps != nsteps) {
    drop_graph(*b);
    if (!b->cap) check(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      run(*b, u, nsteps, b->cap);
    } catch (...) {
      hipGraph_t partial = nullptr;
      (void)hipStreamEndCapture(b->cap, &partial);
      if (partial) (void)hipGraphDestroy(partial);
      throw;
    }
    check(hipStreamEndCapture(b->cap, &b->graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    b->gu = u;
    b->gsteps = nsteps;
  }
  check(hipGraphLaunch(b->gexec, s), "hipGraphLaunch");
  DEC_CATCH
}

int dec_batch_heat_steps_noalloc(dec_batch b, void *u, int32_t nsteps, void *stream) {
  DEC_TRY
  if (!b || !u) fail(DEC_ERR_ARG, "dec_batch_heat_steps_noalloc: NULL argument");
  if (nsteps < 0) fail(DEC_ERR_ARG, "dec_batch_heat_steps_noalloc: nsteps must be >= 0");
  if (state_bytes(*b) > kScratchBytes)
    fail(DEC_ERR_ARG, "dec_batch_heat_steps_noalloc: the state takes " + std::to_string(state_bytes(*b)) +
                          " bytes, above the " + std::to_string(kScratchBytes) + "-byte scratch buffer");
  void *t = scratch_ptr();
  run_with(*b, u, t, nsteps, static_cast<hipStream_t>(stream));
  dec::check_launch("dec_batch_heat_steps_noalloc");
  DEC_CATCH
}

int dec_batch_destroy(dec_batch b) {
  DEC_TRY
  delete b;
  DEC_CATCH
}

}  // extern "C"
