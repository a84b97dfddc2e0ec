// Batched lid-driven cavities: many independent cubic D3Q19 BGK cavities of
// different edges and relaxation times advanced by ONE launch.
//
// Each cavity of edge L is the reference's cavity step chain on an L^3 grid
// (full-way bounce-back walls on x = 0, x = L-1, y = 0, z = 0, z = L-1 and the
// Zou-He moving lid of boundaries3d.py:324-404 on the fluid cells of y = L-1;
// pull stream, then lid, then wall swap, then BGK), i.e. exactly what tlbm_step
// computes for that cavity alone with the wall mask and TLBM_BC_LID_Y.  (Zou-He
// on the wall cells of the lid plane, whose values only bounce back, makes this
// scheme diverge within ~100 steps at u_lid = 0.1.)
// Walls are decided from (x, y, z) in registers: no mask is read.
//
// Memory: one buffer holds every cavity's populations, cavity i at element
// offset off_i in a [z][q][y][x] layout whose (y, x) plane is padded to a
// multiple of 64 elements.  Work: the flat list of 256-cell tiles (tile =
// 256 consecutive cells of one z-plane of one cavity, so small and large
// cavities fill whole wavefronts alike).  A launch either takes one workgroup
// per tile or a persistent grid that strides over the tile list; the cavity
// of a tile is found by a binary search over the tile prefix sums held in the
// kernel arguments (uniform, scalar loads).  A hipGraph of an even number of
// ping-pong steps replays the loop without per-step launch overhead.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <new>

#include "../../include/tlbm_hip.h"
#include "collide.hpp"
#include "store.hpp"

namespace tlbm {

int set_error(int code, const char* msg);  // tlbm.hip: shared tlbm_last_error() channel
// tlbm.hip: two steps of every cavity per launch (k_cav2), table already validated
int cav2_step(int store, const void* fin, void* fout, const tlbm_cavity* cavs, int n, double u_lid, int zc,
              hipStream_t s);

constexpr int CAV_TILE = 256;
constexpr int CAV_MAX = TLBM_CAVITIES_MAX;

struct CavDev {
  int64_t offset;         // element offset of the cavity
  uint32_t tiles_before;  // tiles of the cavities before it
  int32_t edge;
  double omega;  // 1 / tau
};

struct CavArgs {
  const void* fin;
  void* fout;
  double u_lid;
  uint32_t total_tiles;
  int32_t n;
  CavDev c[CAV_MAX];
};
static_assert(sizeof(CavArgs) < 4000, "cavity table travels in the kernel arguments");

__host__ __device__ inline int64_t cav_plane(int L) { return (int64_t(L) * L + 63) / 64 * 64; }

template <int S>
__global__ __launch_bounds__(CAV_TILE) void k_cavities(const CavArgs a) {
  using ST = Store<S>;
  using T = typename ST::T;
  using C = typename ST::C;
  using L = D3Q19;
  constexpr int Q = 19;
  for (uint32_t t = blockIdx.x; t < a.total_tiles; t += gridDim.x) {
    // cavity of tile t: last i with tiles_before <= t
    int lo = 0, hi = a.n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.c[mid].tiles_before <= t) lo = mid; else hi = mid - 1;
    }
    const int E = a.c[lo].edge;
    const uint32_t cells = uint32_t(E) * uint32_t(E);
    const uint32_t tpp = (cells + CAV_TILE - 1) / CAV_TILE;
    const uint32_t local = t - a.c[lo].tiles_before;
    const int z = int(local / tpp);
    const uint32_t c = (local - uint32_t(z) * tpp) * CAV_TILE + threadIdx.x;
    if (c >= cells) continue;
    // y = c / E exactly: (c + 0.5) / E sits at least 0.5 / E from an integer
    const int y = int((float(c) + 0.5f) * __builtin_amdgcn_rcpf(float(E)));
    const int x = int(c) - y * E;
    const int64_t plane = cav_plane(E);
    const int64_t zs = int64_t(Q) * plane;
    const T* fin = static_cast<const T*>(a.fin) + a.c[lo].offset;
    T* fout = static_cast<T*>(a.fout) + a.c[lo].offset;

    const int xd = x == 0 ? E - 1 : x - 1, xu = x == E - 1 ? 0 : x + 1;
    const int yd = y == 0 ? E - 1 : y - 1, yu = y == E - 1 ? 0 : y + 1;
    const int zd = z == 0 ? E - 1 : z - 1, zu = z == E - 1 ? 0 : z + 1;
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int sx = L::cx[q] > 0 ? xd : (L::cx[q] < 0 ? xu : x);
      const int sy = L::cy[q] > 0 ? yd : (L::cy[q] < 0 ? yu : y);
      const int sz = L::cz[q] > 0 ? zd : (L::cz[q] < 0 ? zu : z);
      const T* base = fin + int64_t(q) * plane + int64_t(sz) * zs;
      f[q] = ST::template get<L>(base[uint32_t(sy) * uint32_t(E) + uint32_t(sx)], q);
    }
    const bool wall = x == 0 || x == E - 1 || y == 0 || z == 0 || z == E - 1;
    if (y == E - 1 && !wall) zou_he_lid(f, C(a.u_lid));
    if (wall) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (q < L::opp[q]) {
          const C tmp = f[q];
          f[q] = f[L::opp[q]];
          f[L::opp[q]] = tmp;
        }
      }
    }
    const Macro<C> mc = macro<L>(f);
    bgk_relax<L>(f, mc.rho_safe, mc.ux, mc.uy, mc.uz, C(a.c[lo].omega));
#pragma unroll
    for (int q = 0; q < Q; ++q)
      __builtin_nontemporal_store(ST::template put<L>(f[q], q), fout + int64_t(q) * plane + int64_t(z) * zs + c);
  }
}

}  // namespace tlbm

using namespace tlbm;

namespace {

int cfail(int code, const char* fmt, long long v = 0) {
  char buf[256];
  snprintf(buf, sizeof(buf), fmt, v);
  return tlbm::set_error(code, buf);
}

int build_args(CavArgs& a, int store, const void* f_in, void* f_out, const tlbm_cavity* cavs, int n,
               int64_t capacity, double u_lid) {
  if (!f_in || !f_out || !cavs) return cfail(TLBM_ERR_INVALID, "null pointer");
  if (f_in == f_out) return cfail(TLBM_ERR_ALIAS, "input and output alias");
  if (n < 1 || n > CAV_MAX) return cfail(TLBM_ERR_SHAPE, "cavity count must be 1..%lld", CAV_MAX);
  if (store != TLBM_STORE_F32 && store != TLBM_STORE_F64 && store != TLBM_STORE_F16_DEV &&
      store != TLBM_STORE_BF16_DEV)
    return cfail(TLBM_ERR_DTYPE, "cavities take fp32, fp64 or 16-bit deviation storage (store %lld)", store);
  a.fin = f_in;
  a.fout = f_out;
  a.u_lid = u_lid;
  a.n = n;
  uint64_t tiles = 0;
  int64_t end = 0;
  for (int i = 0; i < n; ++i) {
    const int E = cavs[i].edge;
    if (E < 3 || E > 1024) return cfail(TLBM_ERR_SHAPE, "cavity edge must be 3..1024 (got %lld)", E);
    if (!(cavs[i].tau > 0.5)) return cfail(TLBM_ERR_INVALID, "cavity %lld: tau must be > 0.5", i);
    const int64_t size = int64_t(E) * 19 * cav_plane(E);
    if (cavs[i].offset < end || cavs[i].offset % 64 != 0)
      return cfail(TLBM_ERR_SHAPE, "cavity %lld: offsets must be 64-aligned, increasing and non-overlapping", i);
    end = cavs[i].offset + size;
    if (end > capacity) return cfail(TLBM_ERR_SHAPE, "cavity %lld lies beyond the buffer capacity", i);
    if (cavs[i].tiles_before != int64_t(tiles))
      return cfail(TLBM_ERR_INVALID, "cavity %lld: tiles_before does not match the edges (use tlbm_cavities_layout)", i);
    a.c[i].offset = cavs[i].offset;
    a.c[i].tiles_before = uint32_t(tiles);
    a.c[i].edge = E;
    a.c[i].omega = 1.0 / cavs[i].tau;
    tiles += uint64_t(E) * ((uint64_t(E) * E + CAV_TILE - 1) / CAV_TILE);
    if (tiles >= (uint64_t(1) << 31)) return cfail(TLBM_ERR_SHAPE, "too many tiles");
  }
  a.total_tiles = uint32_t(tiles);
  return TLBM_OK;
}

int launch(const CavArgs& a, int store, int grid, hipStream_t s) {
  const uint32_t g = grid > 0 ? uint32_t(grid) : a.total_tiles;
  switch (store) {
    case TLBM_STORE_F32: hipLaunchKernelGGL(k_cavities<ST_F32>, dim3(g), dim3(CAV_TILE), 0, s, a); break;
    case TLBM_STORE_F64: hipLaunchKernelGGL(k_cavities<ST_F64>, dim3(g), dim3(CAV_TILE), 0, s, a); break;
    case TLBM_STORE_F16_DEV: hipLaunchKernelGGL(k_cavities<ST_F16_DEV>, dim3(g), dim3(CAV_TILE), 0, s, a); break;
    default: hipLaunchKernelGGL(k_cavities<ST_BF16_DEV>, dim3(g), dim3(CAV_TILE), 0, s, a); break;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cfail(TLBM_ERR_HIP, "k_cavities launch failed (hip error %lld)", int(e));
  return TLBM_OK;
}

}  // namespace

struct tlbm_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipStream_t side;             // launch stream used when the caller passes the null stream
  hipEvent_t before, after;     // orders the side stream against the null stream
};

namespace {

// Capture `count` launches (launch(k, stream) for k = 0 .. count-1) into an
// instantiated graph with its own fence stream and events.
template <class Launch>
int capture(int count, Launch launch, tlbm_graph** out) {
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return cfail(TLBM_ERR_HIP, "hipStreamCreate failed");
  tlbm_graph* g = new (std::nothrow) tlbm_graph{nullptr, nullptr, nullptr, nullptr, nullptr};
  if (!g) {
    (void)hipStreamDestroy(s);
    return cfail(TLBM_ERR_HIP, "out of host memory");
  }
  int rc = TLBM_OK;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = cfail(TLBM_ERR_HIP, "hipStreamBeginCapture failed");
  } else {
    for (int k = 0; k < count && rc == TLBM_OK; ++k) rc = launch(k, s);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc == TLBM_OK && e != hipSuccess) rc = cfail(TLBM_ERR_HIP, "hipStreamEndCapture failed (%lld)", int(e));
    g->graph = graph;
    if (rc == TLBM_OK && hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess)
      rc = cfail(TLBM_ERR_HIP, "hipGraphInstantiate failed");
  }
  (void)hipStreamDestroy(s);
  if (rc == TLBM_OK && (hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking) != hipSuccess ||
                        hipEventCreateWithFlags(&g->before, hipEventDisableTiming) != hipSuccess ||
                        hipEventCreateWithFlags(&g->after, hipEventDisableTiming) != hipSuccess))
    rc = cfail(TLBM_ERR_HIP, "stream / event creation failed");
  if (rc != TLBM_OK) {
    tlbm_graph_destroy(g);
    return rc;
  }
  *out = g;
  return TLBM_OK;
}

}  // namespace

extern "C" {

int tlbm_cavities_layout(tlbm_cavity* cavs, int n, int64_t* total_elements, int64_t* total_tiles) {
  if (!cavs || n < 1 || n > CAV_MAX) return cfail(TLBM_ERR_SHAPE, "cavity count must be 1..%lld", CAV_MAX);
  int64_t off = 0, tiles = 0;
  for (int i = 0; i < n; ++i) {
    const int E = cavs[i].edge;
    if (E < 3 || E > 1024) return cfail(TLBM_ERR_SHAPE, "cavity edge must be 3..1024 (got %lld)", E);
    cavs[i].offset = off;
    cavs[i].tiles_before = tiles;
    off += int64_t(E) * 19 * cav_plane(E);
    tiles += int64_t(E) * ((int64_t(E) * E + CAV_TILE - 1) / CAV_TILE);
  }
  if (total_elements) *total_elements = off;
  if (total_tiles) *total_tiles = tiles;
  return TLBM_OK;
}

int tlbm_cavities_step(int store, const void* f_in, void* f_out, int64_t capacity, const tlbm_cavity* cavs, int n,
                       double u_lid, int grid, tlbm_stream_t stream) {
  CavArgs a;
  const int rc = build_args(a, store, f_in, f_out, cavs, n, capacity, u_lid);
  if (rc != TLBM_OK) return rc;
  if (grid < 0) return cfail(TLBM_ERR_INVALID, "grid must be >= 0 (0: one workgroup per tile)");
  return launch(a, store, grid, (hipStream_t)stream);
}

int tlbm_cavities_graph(int store, void* f_a, void* f_b, int64_t capacity, const tlbm_cavity* cavs, int n,
                        double u_lid, int grid, int steps, tlbm_graph** out) {
  if (!out) return cfail(TLBM_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (steps < 2 || steps % 2) return cfail(TLBM_ERR_INVALID, "a graph holds an even number of steps >= 2 (%lld)", steps);
  if (grid < 0) return cfail(TLBM_ERR_INVALID, "grid must be >= 0 (0: one workgroup per tile)");
  CavArgs ab, ba;
  int rc = build_args(ab, store, f_a, f_b, cavs, n, capacity, u_lid);
  if (rc != TLBM_OK) return rc;
  build_args(ba, store, f_b, f_a, cavs, n, capacity, u_lid);
  return capture(steps, [&](int k, hipStream_t s) { return launch(k % 2 ? ba : ab, store, grid, s); }, out);
}

int tlbm_cavities_step2(int store, const void* f_in, void* f_out, int64_t capacity, const tlbm_cavity* cavs, int n,
                        double u_lid, int zc, tlbm_stream_t stream) {
  CavArgs a;
  const int rc = build_args(a, store, f_in, f_out, cavs, n, capacity, u_lid);
  if (rc != TLBM_OK) return rc;
  if (zc < 0) return cfail(TLBM_ERR_INVALID, "zc must be >= 0 (0: the launch-shape model picks it)");
  return tlbm::cav2_step(store, f_in, f_out, cavs, n, u_lid, zc, (hipStream_t)stream);
}

int tlbm_cavities_graph2(int store, void* f_a, void* f_b, int64_t capacity, const tlbm_cavity* cavs, int n,
                         double u_lid, int zc, int launches, tlbm_graph** out) {
  if (!out) return cfail(TLBM_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (launches < 2 || launches % 2)
    return cfail(TLBM_ERR_INVALID, "a graph holds an even number of launches >= 2 (%lld)", launches);
  if (zc < 0) return cfail(TLBM_ERR_INVALID, "zc must be >= 0 (0: the launch-shape model picks it)");
  CavArgs check;
  const int rc = build_args(check, store, f_a, f_b, cavs, n, capacity, u_lid);
  if (rc != TLBM_OK) return rc;
  return capture(launches, [&](int k, hipStream_t s) {
    return tlbm::cav2_step(store, k % 2 ? f_b : f_a, k % 2 ? f_a : f_b, cavs, n, u_lid, zc, s);
  }, out);
}

int tlbm_graph_launch(tlbm_graph* g, tlbm_stream_t stream) {
  if (!g || !g->exec) return cfail(TLBM_ERR_INVALID, "null graph");
  hipStream_t s = (hipStream_t)stream;
  if (s != nullptr) {
    const hipError_t e = hipGraphLaunch(g->exec, s);
    if (e != hipSuccess) return cfail(TLBM_ERR_HIP, "hipGraphLaunch failed (%lld)", int(e));
    return TLBM_OK;
  }
  // null stream: run the graph on a private stream fenced by events on both sides,
  // so it starts after and finishes before the null stream's neighbouring work
  if (hipEventRecord(g->before, nullptr) != hipSuccess || hipStreamWaitEvent(g->side, g->before, 0) != hipSuccess)
    return cfail(TLBM_ERR_HIP, "event fence before the graph failed");
  const hipError_t e = hipGraphLaunch(g->exec, g->side);
  if (e != hipSuccess) return cfail(TLBM_ERR_HIP, "hipGraphLaunch failed (%lld)", int(e));
  if (hipEventRecord(g->after, g->side) != hipSuccess || hipStreamWaitEvent(nullptr, g->after, 0) != hipSuccess)
    return cfail(TLBM_ERR_HIP, "event fence after the graph failed");
  return TLBM_OK;
}

int tlbm_graph_destroy(tlbm_graph* g) {
  if (!g) return TLBM_OK;
  if (g->side) (void)hipStreamSynchronize(g->side);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->before) (void)hipEventDestroy(g->before);
  if (g->after) (void)hipEventDestroy(g->after);
  if (g->side) (void)hipStreamDestroy(g->side);
  delete g;
  return TLBM_OK;
}

}  // extern "C"
