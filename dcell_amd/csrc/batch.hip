// batch.hip -- many independent level sets (DCells) in one launch.
//
// The reference advects the DCells of a rank one after another
// (DWorld/DCellsArray.c:57-60), each over its own small window.  Here the
// windows are packed back to back in one buffer per field and cut into tiles;
// a persistent kernel claims the tiles of the packed tile list (work queue
// below) and runs the SL step on every tile, then a second persistent kernel
// does the curvature.  One step can be recorded as a hipGraph of these two
// launches.  3D cells: the curvature tile shape, the SL kernel (tile gathers
// or LDS-ring columns) and the resident blocks per CU come from a table of
// configurations, and the ring columns' z-chunk is chosen per cell-size
// class; dcl_batch_autotune times them on the caller's buffers and keeps the
// fastest.  2D cells: 64 x 4 node tiles, one node per thread.  Band mode
// (dcl_batch_band_step): only the listed nodes of every cell, as the
// reference's band loop (LevelSetMethod/AdvectSL.c:23-89), one thread per
// node over the packed band list.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "dcl_host.hpp"
#include "dcl_ring.hpp"
#include "dcl_tile.hpp"

namespace dcl {

struct BCell {
  long long phi_off;  // element offset of the cell in the packed phi / kappa buffers
  long long vel_off;  // element offset in the packed velocity buffer (dim per node)
  int px, py, pz, nx, ny, nz;
  int zc;              // z-chunk of the cell's ring columns (3D)
  int pad;
  double c[3], ch[3];  // dt / dh, dt / (2 dh)
  double dx, dy, dz;
};

struct TileShape {
  int tx, ty, tz;
};
// tile shapes of the configuration table (threads = tx * ty).  Shapes
// NBASE .. 2 NBASE - 1 use the same curvature tiles as shape s - NBASE with
// the SL step over 32 x 8 x zc columns staged in an LDS ring.
constexpr int NBASE = 6;
constexpr int NSHAPES = 2 * NBASE;
constexpr int RING_LIST = NBASE;  // index of the ring column list
__host__ __device__ constexpr TileShape shape_of(int s) {
  s = s % NBASE;
  return s == 0 ? TileShape{32, 8, 4}
       : s == 1 ? TileShape{32, 8, 8}
       : s == 2 ? TileShape{64, 4, 4}
       : s == 3 ? TileShape{16, 16, 4}
       : s == 4 ? TileShape{32, 4, 8}
                : TileShape{64, 8, 2};
}
constexpr int NOCC = 3;
constexpr int occ_of(int o) { return o == 0 ? 2 : o == 1 ? 4 : 8; }  // resident blocks per CU
// 2D tiles: 64 x 4 nodes
constexpr int T2X = 64, T2Y = 4;
// cell-size classes (largest axis up to 48, up to 80, beyond) and the ring
// z-chunks tried for each
constexpr int NCLASS = DCL_BATCH_CLASSES;
constexpr int CLASS_EDGE[NCLASS] = {48, 80, 1 << 30};
constexpr int NZC = DCL_BATCH_ZCHUNKS;
constexpr int ZC_OPT[NZC] = {16, 32, 64, 128};
// tiles per claim tried by the autotuner (0: static stride, no atomics)
constexpr int NCLAIM = DCL_BATCH_CLAIMS;
constexpr int CLAIM_OPT[NCLAIM] = {0, 1, 4, 16};

// Work distribution of one persistent launch: the tile list is cut into 8
// contiguous chunks of about equal cost, one per XCD (blocks are dealt to the
// XCDs round-robin, so block b runs on XCD b mod 8), so that neighbouring
// columns (which share halo rows) run in the same L2.  With k = 0 the blocks
// of an XCD take its chunk's tiles in a fixed stride (no atomics); with k > 0
// they claim k tiles at a time with an atomic counter per chunk and, when
// their chunk is exhausted, take tiles from the other chunks.  The last block
// to finish resets the counters, so the next launch (or graph replay) starts
// from zero.
struct WorkQ {
  unsigned claim[8];
  unsigned done;
  unsigned pad[7];
};

template <typename F>
__device__ __forceinline__ void claim_tiles(WorkQ* __restrict__ q, const int* __restrict__ bounds, int k, F body) {
  const int xcd = blockIdx.x & 7;
  if (k <= 0) {
    const int nslot = gridDim.x >> 3, slot = blockIdx.x >> 3;
    for (int t = bounds[xcd] + slot; t < bounds[xcd + 1]; t += nslot) body(t);
    return;
  }
  __shared__ int s_tile;
  const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
  for (int s = 0; s < 8; ++s) {
    const int x = (xcd + s) & 7;
    const int lo = bounds[x], hi = bounds[x + 1];
    if (lo >= hi) continue;
    for (;;) {
      if (lead) s_tile = lo + k * (int)atomicAdd(&q->claim[x], 1u);
      __syncthreads();
      const int t0 = s_tile;
      __syncthreads();
      if (t0 >= hi) break;
      const int t1 = min(t0 + k, hi);
      for (int t = t0; t < t1; ++t) body(t);
    }
  }
  if (lead && atomicAdd(&q->done, 1u) == gridDim.x - 1) {
    for (int x = 0; x < 8; ++x) atomicExch(&q->claim[x], 0u);
    atomicExch(&q->done, 0u);
  }
}

template <int DIM>
__device__ __forceinline__ SLP cell_slp(const BCell& c) {
  SLP P;
  P.w.px = c.px; P.w.py = c.py; P.w.pz = c.pz;
  P.w.nx = c.nx; P.w.ny = c.ny; P.w.nz = c.nz;
  P.v.w = P.w;
  P.v.dof = DIM;
  P.v.udof = 0;
  P.v.cs = 1;
  for (int a = 0; a < 3; ++a) {
    P.c[a] = c.c[a];
    P.ch[a] = c.ch[a];
  }
  return P;
}

__device__ __forceinline__ KWin cell_win(const BCell& c) {
  KWin w;
  w.px = c.px; w.py = c.py; w.pz = c.pz;
  w.nx = c.nx; w.ny = c.ny; w.nz = c.nz;
  return w;
}

template <typename T, int INTERP, int BT, int S>
__global__ __launch_bounds__(shape_of(S).tx * shape_of(S).ty) void batch_sl_k(
    const T* __restrict__ phi0, T* __restrict__ phi, const T* __restrict__ vel, const BCell* __restrict__ cells,
    const int4* __restrict__ tiles, const int* __restrict__ bounds, WorkQ* __restrict__ q, int k) {
  constexpr TileShape sh = shape_of(S);
  claim_tiles(q, bounds, k, [&](int t) {
    const int4 tl = tiles[t];
    const BCell c = cells[tl.x];
    const SLP P = cell_slp<3>(c);
    const unsigned nodes = (unsigned)c.nx * c.ny * c.nz;
    const rsrc_t rp = make_rsrc(phi0 + c.phi_off, nodes * (unsigned)sizeof(T));
    const rsrc_t rv = make_rsrc(vel + c.vel_off, 3u * nodes * (unsigned)sizeof(T));
    sl_tile<T, INTERP, BT, sh.tx, sh.ty, sh.tz>(rp, rv, phi0 + c.phi_off, phi + c.phi_off, P, tl.y, tl.z, tl.w);
  });
}

// ring columns of the batch: 32 x 8 nodes, the cell's z-chunk, velocity staged in FP32
constexpr int RING_TX = 32, RING_TY = 8;

template <typename T, int INTERP, int BT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ring_waves<T, sizeof(T) == 4>(), 8))) void
batch_ring_k(const T* __restrict__ phi0, T* __restrict__ phi,
                                                      const T* __restrict__ vel, const BCell* __restrict__ cells,
                                                      const int4* __restrict__ tiles,
                                                      const int* __restrict__ bounds, WorkQ* __restrict__ q, int k) {
  typedef typename RingFor<T, RING_TX, RING_TY>::type G;
  typedef VelGeom<RING_TX, RING_TY, G::VRP, G::VPP> V;
  constexpr bool VL = sizeof(T) == 4;
  __shared__ T ring[RG_R * G::PP];
  __shared__ T vr[VL ? VR_R * 3 * V::PP : 1];
  claim_tiles(q, bounds, k, [&](int t) {
    const int4 tl = tiles[t];
    const BCell c = cells[tl.x];
    const SLP P = cell_slp<3>(c);
    const unsigned nodes = (unsigned)c.nx * c.ny * c.nz;
    const rsrc_t rp = make_rsrc(phi0 + c.phi_off, nodes * (unsigned)sizeof(T));
    const rsrc_t rv = make_rsrc(vel + c.vel_off, 3u * nodes * (unsigned)sizeof(T));
    ring_column<T, INTERP, BT, G, VL>(phi0 + c.phi_off, phi + c.phi_off, rp, rv, P, tl.y, tl.z, tl.w, c.zc, ring,
                                      vr);
  });
}

template <typename T, int S>
__global__ __launch_bounds__(shape_of(S).tx * shape_of(S).ty) void batch_curv_k(
    const T* __restrict__ phi, T* __restrict__ kappa, const BCell* __restrict__ cells,
    const int4* __restrict__ tiles, const int* __restrict__ bounds, WorkQ* __restrict__ q, int k) {
  constexpr TileShape sh = shape_of(S);
  constexpr int PF = ((sh.tx + 2) * (sh.ty + 2) + sh.tx * sh.ty - 1) / (sh.tx * sh.ty);
  __shared__ T pl[4 * PF * sh.tx * sh.ty];
  claim_tiles(q, bounds, k, [&](int t) {
    const int4 tl = tiles[t];
    const BCell c = cells[tl.x];
    const unsigned nodes = (unsigned)c.nx * c.ny * c.nz;
    curv_column2<T, sh.tx, sh.ty, 1, sh.tz>(make_rsrc(phi + c.phi_off, nodes * (unsigned)sizeof(T)),
                                            kappa + c.phi_off, cell_win(c), curv_recips<T>(c.dx, c.dy, c.dz), tl.y,
                                            tl.z, tl.w, pl);
  });
}

// 2D cells: one node per thread over 64 x 4 tiles; nodes outside [2, n-3]
// copy phi0 (as the dense 2D step)
template <typename T, int INTERP, int BT>
__global__ __launch_bounds__(T2X * T2Y) void batch2d_sl_k(const T* __restrict__ phi0, T* __restrict__ phi,
                                                         const T* __restrict__ vel, const BCell* __restrict__ cells,
                                                         const int4* __restrict__ tiles,
                                                         const int* __restrict__ bounds, WorkQ* __restrict__ q, int k) {
  claim_tiles(q, bounds, k, [&](int t) {
    const int4 tl = tiles[t];
    const BCell c = cells[tl.x];
    const int i = tl.y + threadIdx.x, j = tl.z + threadIdx.y;
    if (i < c.nx && j < c.ny) {
      const SLP P = cell_slp<2>(c);
      const int id = j * c.nx + i;
      const T* p0 = phi0 + c.phi_off;
      T v;
      if (i >= 2 && i <= c.nx - 3 && j >= 2 && j <= c.ny - 3)
        v = sl_point<T, 2, INTERP, BT>(p0, vel + c.vel_off, P, c.px + i, c.py + j, c.pz);
      else
        v = p0[id];
      phi[c.phi_off + id] = v;
    }
  });
}

// GridFunction2D_Curv (Common/Grid.c:231-254) at a local node; 0 where the
// stencil leaves the window
template <typename T>
__device__ __forceinline__ T curv2d_node(const T* __restrict__ p, const BCell& c, int i, int j) {
  if (i < 1 || i > c.nx - 2 || j < 1 || j > c.ny - 2) return T(0);
  return curv2d_at<T>(p + j * c.nx + i, c.nx, (T)c.dx, (T)c.dy);
}

template <typename T>
__global__ __launch_bounds__(T2X * T2Y) void batch2d_curv_k(const T* __restrict__ phi, T* __restrict__ kappa,
                                                           const BCell* __restrict__ cells,
                                                           const int4* __restrict__ tiles,
                                                           const int* __restrict__ bounds, WorkQ* __restrict__ q, int k) {
  claim_tiles(q, bounds, k, [&](int t) {
    const int4 tl = tiles[t];
    const BCell c = cells[tl.x];
    const int i = tl.y + threadIdx.x, j = tl.z + threadIdx.y;
    if (i < c.nx && j < c.ny) kappa[c.phi_off + j * c.nx + i] = curv2d_node<T>(phi + c.phi_off, c, i, j);
  });
}

// Band mode: entry b of the packed list is {x, y, z, cell} in the cell's
// global coordinates.  The new value is written to phi at the node when the
// node lies in the cell's window.
template <typename T, int DIM, int INTERP, int BT>
__global__ __launch_bounds__(256) void batch_band_sl_k(const T* __restrict__ phi0, T* __restrict__ phi,
                                                       const T* __restrict__ vel, const BCell* __restrict__ cells,
                                                       const int4* __restrict__ band, int64_t n) {
  for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < n; b += (int64_t)gridDim.x * 256) {
    const int4 e = band[b];
    const BCell c = cells[e.w];
    const int x = e.x - c.px, y = e.y - c.py, z = DIM == 3 ? e.z - c.pz : 0;
    if (x < 0 || x >= c.nx || y < 0 || y >= c.ny || z < 0 || z >= (DIM == 3 ? c.nz : 1)) continue;
    const SLP P = cell_slp<DIM>(c);
    const T v = sl_point<T, DIM, INTERP, BT>(phi0 + c.phi_off, vel + c.vel_off, P, e.x, e.y, DIM == 3 ? e.z : c.pz);
    phi[c.phi_off + ((long long)z * c.ny + y) * c.nx + x] = v;
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void batch_band_curv_k(const T* __restrict__ phi, T* __restrict__ kappa,
                                                         const BCell* __restrict__ cells,
                                                         const int4* __restrict__ band, int64_t n) {
  for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < n; b += (int64_t)gridDim.x * 256) {
    const int4 e = band[b];
    const BCell c = cells[e.w];
    const int x = e.x - c.px, y = e.y - c.py, z = DIM == 3 ? e.z - c.pz : 0;
    if (x < 0 || x >= c.nx || y < 0 || y >= c.ny || z < 0 || z >= (DIM == 3 ? c.nz : 1)) continue;
    const T* p = phi + c.phi_off;
    T k;
    if (DIM == 3) {
      const bool in = x >= 1 && x <= c.nx - 2 && y >= 1 && y <= c.ny - 2 && z >= 1 && z <= c.nz - 2;
      k = in ? curv3d_at<T>(p + ((long long)z * c.ny + y) * c.nx + x, c.nx, c.nx * c.ny,
                            curv_recips<T>(c.dx, c.dy, c.dz))
             : T(0);
    } else {
      k = curv2d_node<T>(p, c, x, y);
    }
    kappa[c.phi_off + ((long long)z * c.ny + y) * c.nx + x] = k;
  }
}

// the kernel pointers of one precision / scheme
template <typename T, int INTERP, int BT>
struct KTable {
  static const void* sl(int s) {
    switch (s) {
      case 0: return (const void*)&batch_sl_k<T, INTERP, BT, 0>;
      case 1: return (const void*)&batch_sl_k<T, INTERP, BT, 1>;
      case 2: return (const void*)&batch_sl_k<T, INTERP, BT, 2>;
      case 3: return (const void*)&batch_sl_k<T, INTERP, BT, 3>;
      case 4: return (const void*)&batch_sl_k<T, INTERP, BT, 4>;
      default: return (const void*)&batch_sl_k<T, INTERP, BT, 5>;
    }
  }
  static const void* ring() { return (const void*)&batch_ring_k<T, INTERP, BT>; }
  static const void* sl2d() { return (const void*)&batch2d_sl_k<T, INTERP, BT>; }
  static const void* band(int dim) {
    return dim == 3 ? (const void*)&batch_band_sl_k<T, 3, INTERP, BT>
                    : (const void*)&batch_band_sl_k<T, 2, INTERP, BT>;
  }
};

template <typename T>
static const void* curv_kernel(int s) {
  switch (s) {
    case 0: return (const void*)&batch_curv_k<T, 0>;
    case 1: return (const void*)&batch_curv_k<T, 1>;
    case 2: return (const void*)&batch_curv_k<T, 2>;
    case 3: return (const void*)&batch_curv_k<T, 3>;
    case 4: return (const void*)&batch_curv_k<T, 4>;
    default: return (const void*)&batch_curv_k<T, 5>;
  }
}

// SL kernels by kind: 0 tile gathers of shape s, 1 ring columns, 2 2D
// tiles, 3 band mode of dimension s
template <typename K>
static const void* sl_pick(int kind, int s) {
  return kind == 0 ? K::sl(s) : kind == 1 ? K::ring() : kind == 2 ? K::sl2d() : K::band(s);
}

template <typename T>
static const void* sl_kernel(int interp, int bt, int kind, int s) {
  if (interp == DCL_CUBIC && bt == DCL_RK2) return sl_pick<KTable<T, DCL_CUBIC, DCL_RK2>>(kind, s);
  if (interp == DCL_CUBIC) return sl_pick<KTable<T, DCL_CUBIC, DCL_EULER>>(kind, s);
  if (bt == DCL_RK2) return sl_pick<KTable<T, DCL_LINEAR, DCL_RK2>>(kind, s);
  return sl_pick<KTable<T, DCL_LINEAR, DCL_EULER>>(kind, s);
}

}  // namespace dcl

using namespace dcl;

struct dcl_batch_s {
  int ncells, precision, interp, backtrace, dim;
  std::vector<dcl_window> wins;
  std::vector<BCell> host_cells;
  std::vector<long long> phi_off, vel_off;
  std::vector<int> cls;                 // size class of every cell
  BCell* cells = nullptr;
  int4* tiles[NBASE + 1] = {nullptr};   // per tile shape, then the ring column list (2D: [0])
  int* bounds[NBASE + 1] = {nullptr};   // 9 chunk bounds (one chunk per XCD) of each tile list
  int ntiles[NBASE + 1] = {0};
  WorkQ* queues = nullptr;              // [0] the SL launch, [1] the curvature launch
  int shape = 1, occ = 1;
  int claim = 0;                        // tiles per claim of both launches (0: static stride)
  int zc[NCLASS] = {32, 32, 32};        // ring z-chunk per size class
  int cus = 256;
  bool dt_set = false;
};

static int cell_class(const dcl_window& w) {
  const int e = std::max(w.n.x, std::max(w.n.y, w.n.z));
  int c = 0;
  while (c < NCLASS - 1 && e > CLASS_EDGE[c]) ++c;
  return c;
}

static int upload_cells(dcl_batch b) {
  for (int c = 0; c < b->ncells; ++c) b->host_cells[c].zc = b->zc[b->cls[c]];
  DCL_HIP(hipMemcpy(b->cells, b->host_cells.data(), sizeof(BCell) * b->ncells, hipMemcpyHostToDevice));
  return 0;
}

// tile list s < NBASE: tiles of shape s (2D batches: s = 0, 64 x 4 tiles);
// s == RING_LIST: 32 x 8 columns of each cell's z-chunk.  Cells come largest
// first (a block that claims late gets a small tile), each cell's tiles in
// raster order; the list is cut into 8 contiguous chunks of about equal cost,
// a tile costing its lanes times the planes it walks plus a halo's worth of
// planes.
static int build_tiles(dcl_batch b, int s) {
  if (b->tiles[s]) return 0;
  const bool two = b->dim == 2;
  std::vector<int> order(b->ncells);
  for (int c = 0; c < b->ncells; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return win_nodes(b->wins[a]) > win_nodes(b->wins[c]); });
  std::vector<int4> t;
  std::vector<double> cost;
  for (int c : order) {
    const dcl_window& w = b->wins[c];
    const TileShape sh = two                ? TileShape{T2X, T2Y, 1}
                         : s == RING_LIST ? TileShape{RING_TX, RING_TY, b->zc[b->cls[c]]}
                                          : shape_of(s);
    const int halo = two ? 0 : s == RING_LIST ? 2 * RG_H : 2;
    const int nz = two ? 1 : w.n.z;
    for (int z = 0; z < nz; z += sh.tz)
      for (int y = 0; y < w.n.y; y += sh.ty)
        for (int x = 0; x < w.n.x; x += sh.tx) {
          t.push_back(make_int4(c, x, y, z));
          cost.push_back((double)sh.tx * sh.ty * (std::min(sh.tz, nz - z) + halo));
        }
  }
  if (t.size() >= (size_t)1 << 31) return set_error(DCL_ERR_ARG_SIZ, "too many tiles");
  double total = 0;
  for (double c : cost) total += c;
  int bnd[9];
  bnd[0] = 0;
  double acc = 0;
  size_t i = 0;
  for (int x = 1; x < 8; ++x) {
    while (i < t.size() && acc + 0.5 * cost[i] < total * x / 8.0) acc += cost[i++];
    bnd[x] = (int)i;
  }
  bnd[8] = (int)t.size();
  void* p = nullptr;
  DCL_TRY(dcl_malloc(&p, sizeof(int4) * t.size() + sizeof(bnd)));
  hipError_t e = hipMemcpy(p, t.data(), sizeof(int4) * t.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(static_cast<char*>(p) + sizeof(int4) * t.size(), bnd, sizeof(bnd), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dcl_free(p);
    return hip_fail(e, "tile list upload", __FILE__, __LINE__);
  }
  b->tiles[s] = static_cast<int4*>(p);
  b->bounds[s] = reinterpret_cast<int*>(static_cast<char*>(p) + sizeof(int4) * t.size());
  b->ntiles[s] = (int)t.size();
  return 0;
}

static int drop_tiles(dcl_batch b, int s) {
  if (b->tiles[s]) {
    DCL_HIP(hipDeviceSynchronize());
    dcl_free(b->tiles[s]);
  }
  b->tiles[s] = nullptr;
  b->bounds[s] = nullptr;
  b->ntiles[s] = 0;
  return 0;
}

static int grid_blocks(dcl_batch b, int occ) {
  int g = b->cus * occ_of(occ);
  return g < 8 ? 8 : g - g % 8;
}

// the two launches of one step with configuration (s, occ)
struct Plan {
  const void* ksl;
  const void* kcv;
  dim3 grid, bsl, bcv;
  const int4* tsl;
  const int4* tcv;
  const int* bsl_bounds;
  const int* bcv_bounds;
  WorkQ* qsl;
  WorkQ* qcv;
  int k;
};

static int make_plan(dcl_batch b, int s, int occ, Plan& p) {
  const bool f64 = b->precision == DCL_F64;
  p.grid = dim3(grid_blocks(b, occ));
  p.k = b->claim;
  p.qsl = b->queues;
  p.qcv = b->queues + 1;
  if (b->dim == 2) {
    DCL_TRY(build_tiles(b, 0));
    p.bsl = p.bcv = dim3(T2X, T2Y);
    p.tsl = p.tcv = b->tiles[0];
    p.bsl_bounds = p.bcv_bounds = b->bounds[0];
    p.ksl = f64 ? sl_kernel<double>(b->interp, b->backtrace, 2, 0) : sl_kernel<float>(b->interp, b->backtrace, 2, 0);
    p.kcv = f64 ? (const void*)&batch2d_curv_k<double> : (const void*)&batch2d_curv_k<float>;
    return 0;
  }
  const int cs = s % NBASE;
  const bool ring = s >= NBASE;
  DCL_TRY(build_tiles(b, cs));
  if (ring) DCL_TRY(build_tiles(b, RING_LIST));
  const TileShape sh = shape_of(cs);
  p.bcv = dim3(sh.tx, sh.ty);
  p.tcv = b->tiles[cs];
  p.bcv_bounds = b->bounds[cs];
  p.kcv = f64 ? curv_kernel<double>(cs) : curv_kernel<float>(cs);
  if (ring) {
    p.ksl = f64 ? sl_kernel<double>(b->interp, b->backtrace, 1, 0) : sl_kernel<float>(b->interp, b->backtrace, 1, 0);
    p.bsl = dim3(RING_TX, RING_TY);
    p.tsl = b->tiles[RING_LIST];
    p.bsl_bounds = b->bounds[RING_LIST];
  } else {
    p.ksl = f64 ? sl_kernel<double>(b->interp, b->backtrace, 0, cs)
                : sl_kernel<float>(b->interp, b->backtrace, 0, cs);
    p.bsl = p.bcv;
    p.tsl = p.tcv;
    p.bsl_bounds = p.bcv_bounds;
  }
  return 0;
}

// enqueue one step with a given configuration
static int batch_launch(dcl_batch b, int s, int occ, const void* phi0, void* phi, void* kappa, const void* vel,
                        hipStream_t st) {
  Plan p;
  DCL_TRY(make_plan(b, s, occ, p));
  const BCell* cells = b->cells;
  void* a1[] = {(void*)&phi0, (void*)&phi, (void*)&vel, (void*)&cells, (void*)&p.tsl, (void*)&p.bsl_bounds,
                (void*)&p.qsl, (void*)&p.k};
  DCL_HIP(hipLaunchKernel(p.ksl, p.grid, p.bsl, a1, 0, st));
  const void* cphi = phi;
  void* a2[] = {(void*)&cphi, (void*)&kappa, (void*)&cells, (void*)&p.tcv, (void*)&p.bcv_bounds, (void*)&p.qcv,
                (void*)&p.k};
  DCL_HIP(hipLaunchKernel(p.kcv, p.grid, p.bcv, a2, 0, st));
  return 0;
}

extern "C" int dcl_batch_create(int32_t ncells, const dcl_window* wins, int precision, int interp,
                                int backtrace, dcl_batch* out) {
  DCL_TRY(check_precision(precision));
  if (!out) return set_error(DCL_ERR_ARG_NULL, "out is NULL");
  *out = nullptr;
  if (ncells <= 0) return set_error(DCL_ERR_ARG_SIZ, "ncells %d <= 0", ncells);
  if (!wins) return set_error(DCL_ERR_ARG_NULL, "wins is NULL");
  if (interp != DCL_LINEAR && interp != DCL_CUBIC) return set_error(DCL_ERR_ARG_WRONG, "bad interp");
  if (backtrace != DCL_EULER && backtrace != DCL_RK2) return set_error(DCL_ERR_ARG_WRONG, "bad backtrace");
  const int dim = wins[0].n.z < 2 ? 2 : 3;
  for (int c = 0; c < ncells; ++c) {
    DCL_TRY(check_window(&wins[c], "cell", 5, dim));
    if ((wins[c].n.z < 2 ? 2 : 3) != dim)
      return set_error(DCL_ERR_SUP, "cell %d is %dD, cell 0 %dD: a batch holds cells of one dimension", c,
                       wins[c].n.z < 2 ? 2 : 3, dim);
    // the 3D kernels address a cell's velocity through a buffer descriptor
    // with signed 32-bit byte offsets
    const int64_t vbytes = dim * win_nodes(wins[c]) * (precision == DCL_F64 ? 8 : 4);
    if (vbytes >= ((int64_t)1 << 31))
      return set_error(DCL_ERR_ARG_SIZ, "cell %d: its velocity takes %lld bytes; the limit is 2^31 - 1", c,
                       (long long)vbytes);
  }
  dcl_batch b = new (std::nothrow) dcl_batch_s();
  if (!b) return set_error(DCL_ERR_MEM, "host allocation failed");
  b->ncells = ncells;
  b->precision = precision;
  b->interp = interp;
  b->backtrace = backtrace;
  b->dim = dim;
  b->wins.assign(wins, wins + ncells);
  long long po = 0, vo = 0;
  for (int c = 0; c < ncells; ++c) {
    const dcl_window& w = wins[c];
    BCell cell;
    memset(&cell, 0, sizeof(cell));
    cell.phi_off = po;
    cell.vel_off = vo;
    cell.px = w.p.x; cell.py = w.p.y; cell.pz = w.p.z;
    cell.nx = w.n.x; cell.ny = w.n.y; cell.nz = dim == 2 ? 1 : w.n.z;
    cell.dx = w.d.x; cell.dy = w.d.y; cell.dz = w.d.z;
    b->phi_off.push_back(po);
    b->vel_off.push_back(vo);
    b->host_cells.push_back(cell);
    b->cls.push_back(cell_class(w));
    po += win_nodes(w);
    vo += dim * win_nodes(w);
  }
  b->phi_off.push_back(po);
  b->vel_off.push_back(vo);
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      b->cus = cus;
  }
  void* p = nullptr;
  int rc = dcl_malloc(&p, sizeof(BCell) * ncells + 2 * sizeof(WorkQ));
  if (rc) {
    delete b;
    return rc;
  }
  b->cells = static_cast<BCell*>(p);
  b->queues = reinterpret_cast<WorkQ*>(static_cast<char*>(p) + sizeof(BCell) * ncells);
  // the queues start at zero; each launch's last block leaves them at zero
  const hipError_t e = hipMemset(b->queues, 0, 2 * sizeof(WorkQ));
  if (e != hipSuccess) {
    dcl_free(p);
    delete b;
    return hip_fail(e, "work queue reset", __FILE__, __LINE__);
  }
  *out = b;
  return 0;
}

extern "C" int dcl_batch_destroy(dcl_batch b) {
  if (!b) return 0;
  DCL_HIP(hipDeviceSynchronize());
  dcl_free(b->cells);
  for (int s = 0; s <= NBASE; ++s) dcl_free(b->tiles[s]);
  delete b;
  return 0;
}

extern "C" int dcl_batch_layout(dcl_batch b, int64_t* phi_offsets, int64_t* vel_offsets) {
  if (!b) return set_error(DCL_ERR_ARG_NULL, "batch is NULL");
  for (int c = 0; c <= b->ncells; ++c) {
    if (phi_offsets) phi_offsets[c] = b->phi_off[c];
    if (vel_offsets) vel_offsets[c] = b->vel_off[c];
  }
  return 0;
}

extern "C" int dcl_batch_set_dt(dcl_batch b, const double* dt) {
  if (!b || !dt) return set_error(DCL_ERR_ARG_NULL, "batch or dt is NULL");
  for (int c = 0; c < b->ncells; ++c) {
    BCell& cell = b->host_cells[c];
    const double d[3] = {cell.dx, cell.dy, cell.dz};
    for (int a = 0; a < 3; ++a) {
      cell.c[a] = dt[c] / d[a];
      cell.ch[a] = dt[c] / (2.0 * d[a]);
    }
  }
  DCL_TRY(upload_cells(b));
  b->dt_set = true;
  return 0;
}

extern "C" int dcl_batch_set_config(dcl_batch b, int shape, int occupancy) {
  if (!b) return set_error(DCL_ERR_ARG_NULL, "batch is NULL");
  if (shape < 0 || shape >= NSHAPES || occupancy < 0 || occupancy >= NOCC)
    return set_error(DCL_ERR_ARG_WRONG, "configuration (%d, %d) out of range", shape, occupancy);
  b->shape = shape;
  b->occ = occupancy;
  return 0;
}

extern "C" int dcl_batch_get_config(dcl_batch b, int* shape, int* occupancy, int32_t tile[3],
                                    int* blocks) {
  if (!b) return set_error(DCL_ERR_ARG_NULL, "batch is NULL");
  if (shape) *shape = b->shape;
  if (occupancy) *occupancy = b->occ;
  const TileShape sh = b->dim == 2 ? TileShape{T2X, T2Y, 1} : shape_of(b->shape);
  if (tile) {
    tile[0] = sh.tx;
    tile[1] = sh.ty;
    tile[2] = sh.tz;
  }
  if (blocks) *blocks = grid_blocks(b, b->occ);
  return 0;
}

extern "C" int dcl_batch_set_claim(dcl_batch b, int tiles_per_claim) {
  if (!b) return set_error(DCL_ERR_ARG_NULL, "batch is NULL");
  if (tiles_per_claim < 0 || tiles_per_claim > 4096)
    return set_error(DCL_ERR_ARG_WRONG, "tiles per claim %d out of range", tiles_per_claim);
  b->claim = tiles_per_claim;
  return 0;
}

extern "C" int dcl_batch_get_claim(dcl_batch b, int* tiles_per_claim) {
  if (!b || !tiles_per_claim) return set_error(DCL_ERR_ARG_NULL, "batch or tiles_per_claim is NULL");
  *tiles_per_claim = b->claim;
  return 0;
}

static int set_zc(dcl_batch b, int cls, int zc) {
  if (b->zc[cls] == zc) return 0;
  b->zc[cls] = zc;
  DCL_TRY(drop_tiles(b, RING_LIST));
  if (b->dt_set) DCL_TRY(upload_cells(b));
  return 0;
}

extern "C" int dcl_batch_set_class_zchunk(dcl_batch b, int cls, int zchunk) {
  if (!b) return set_error(DCL_ERR_ARG_NULL, "batch is NULL");
  if (cls < 0 || cls >= NCLASS) return set_error(DCL_ERR_ARG_WRONG, "class %d out of range", cls);
  if (zchunk < 1 || zchunk > 4096) return set_error(DCL_ERR_ARG_WRONG, "z-chunk %d out of range", zchunk);
  return set_zc(b, cls, zchunk);
}

extern "C" int dcl_batch_get_classes(dcl_batch b, int32_t max_edge[DCL_BATCH_CLASSES],
                                     int32_t cells[DCL_BATCH_CLASSES], int32_t zchunk[DCL_BATCH_CLASSES]) {
  if (!b) return set_error(DCL_ERR_ARG_NULL, "batch is NULL");
  for (int k = 0; k < NCLASS; ++k) {
    if (max_edge) max_edge[k] = CLASS_EDGE[k] < (1 << 30) ? CLASS_EDGE[k] : -1;
    if (cells) cells[k] = (int32_t)std::count(b->cls.begin(), b->cls.end(), k);
    if (zchunk) zchunk[k] = b->zc[k];
  }
  return 0;
}

static int check_step(dcl_batch b, const void* phi0, void* phi, void* kappa, const void* vel) {
  if (!b) return set_error(DCL_ERR_ARG_NULL, "batch is NULL");
  if (!phi0 || !phi || !kappa || !vel) return set_error(DCL_ERR_ARG_NULL, "a buffer is NULL");
  if (phi0 == phi || kappa == phi || kappa == phi0)
    return set_error(DCL_ERR_ARG_WRONG, "phi0, phi and kappa must be distinct buffers");
  if (!b->dt_set) return set_error(DCL_ERR_ARG_WRONG, "dcl_batch_set_dt was not called");
  return 0;
}

extern "C" int dcl_batch_step(dcl_batch b, const void* phi0, void* phi, void* kappa, const void* vel,
                              void* stream) {
  DCL_TRY(check_step(b, phi0, phi, kappa, vel));
  return batch_launch(b, b->shape, b->occ, phi0, phi, kappa, vel, as_stream(stream));
}

extern "C" int dcl_batch_band_step(dcl_batch b, const void* phi0, void* phi, void* kappa, const void* vel,
                                   const int32_t* band, int64_t count, void* stream) {
  DCL_TRY(check_step(b, phi0, phi, kappa, vel));
  if (count < 0) return set_error(DCL_ERR_ARG_SIZ, "count < 0");
  if (count == 0) return 0;
  if (!band) return set_error(DCL_ERR_ARG_NULL, "band is NULL");
  const bool f64 = b->precision == DCL_F64;
  hipStream_t st = as_stream(stream);
  const int64_t nb64 = (count + 255) / 256;
  const unsigned nb = (unsigned)std::min<int64_t>(nb64, (int64_t)b->cus * 16);
  const void* ksl = f64 ? sl_kernel<double>(b->interp, b->backtrace, 3, b->dim)
                        : sl_kernel<float>(b->interp, b->backtrace, 3, b->dim);
  const void* kcv = f64 ? (b->dim == 3 ? (const void*)&batch_band_curv_k<double, 3>
                                       : (const void*)&batch_band_curv_k<double, 2>)
                        : (b->dim == 3 ? (const void*)&batch_band_curv_k<float, 3>
                                       : (const void*)&batch_band_curv_k<float, 2>);
  const BCell* cells = b->cells;
  const int4* bd = reinterpret_cast<const int4*>(band);
  void* a1[] = {(void*)&phi0, (void*)&phi, (void*)&vel, (void*)&cells, (void*)&bd, (void*)&count};
  DCL_HIP(hipLaunchKernel(ksl, dim3(nb), dim3(256), a1, 0, st));
  const void* cphi = phi;
  void* a2[] = {(void*)&cphi, (void*)&kappa, (void*)&cells, (void*)&bd, (void*)&count};
  DCL_HIP(hipLaunchKernel(kcv, dim3(nb), dim3(256), a2, 0, st));
  return 0;
}

// per-step time of configuration (s, occ) over reps steps after one warm-up step
static int time_config(dcl_batch b, int s, int o, const void* phi0, void* phi, void* kappa, const void* vel,
                       int reps, hipStream_t st, hipEvent_t e0, hipEvent_t e1, double* per) {
  DCL_TRY(batch_launch(b, s, o, phi0, phi, kappa, vel, st));  // warm-up (and tile list build)
  DCL_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) DCL_TRY(batch_launch(b, s, o, phi0, phi, kappa, vel, st));
  DCL_HIP(hipEventRecord(e1, st));
  DCL_HIP(hipEventSynchronize(e1));
  float ms = 0;
  DCL_HIP(hipEventElapsedTime(&ms, e0, e1));
  *per = ms / reps;
  return 0;
}

static int autotune(dcl_batch b, const void* phi0, void* phi, void* kappa, const void* vel, int reps,
                    double* ms_table, double* zc_table, double* claim_table, double* best_ms, hipStream_t st,
                    hipEvent_t e0, hipEvent_t e1) {
  double best = 1e300;
  b->claim = 0;
  int bs = b->shape, bo = b->occ;
  const int nshapes = b->dim == 2 ? 1 : NSHAPES;
  for (int s = 0; s < NSHAPES; ++s)
    for (int o = 0; o < NOCC; ++o) {
      double per = -1;
      if (s < nshapes) DCL_TRY(time_config(b, s, o, phi0, phi, kappa, vel, reps, st, e0, e1, &per));
      if (ms_table) ms_table[s * NOCC + o] = per;
      if (s < nshapes && per < best) {
        best = per;
        bs = s;
        bo = o;
      }
    }
  b->shape = bs;
  b->occ = bo;
  // with ring columns: the z-chunk of each size class, one class at a time
  // with the others at their current choice
  for (int k = 0; k < NCLASS; ++k)
    for (int z = 0; z < NZC; ++z) {
      if (zc_table) zc_table[k * NZC + z] = -1;
      if (b->dim != 3 || bs < NBASE || std::count(b->cls.begin(), b->cls.end(), k) == 0) continue;
      const int keep = b->zc[k];
      DCL_TRY(set_zc(b, k, ZC_OPT[z]));
      double per = 0;
      DCL_TRY(time_config(b, bs, bo, phi0, phi, kappa, vel, reps, st, e0, e1, &per));
      if (zc_table) zc_table[k * NZC + z] = per;
      if (per < best) best = per;
      else DCL_TRY(set_zc(b, k, keep));
    }
  // the work distribution: static stride or claims of 1, 4, 16 tiles
  int bc = 0;
  for (int c = 0; c < NCLAIM; ++c) {
    b->claim = CLAIM_OPT[c];
    double per = 0;
    DCL_TRY(time_config(b, bs, bo, phi0, phi, kappa, vel, reps, st, e0, e1, &per));
    if (claim_table) claim_table[c] = per;
    if (per < best) {
      best = per;
      bc = CLAIM_OPT[c];
    }
  }
  b->claim = bc;
  if (best_ms) *best_ms = best;
  return 0;
}

extern "C" int dcl_batch_autotune(dcl_batch b, const void* phi0, void* phi, void* kappa, const void* vel,
                                  int reps, double* ms_table, double* zc_table, double* claim_table, double* best_ms,
                                  void* stream) {
  DCL_TRY(check_step(b, phi0, phi, kappa, vel));
  if (reps < 1) return set_error(DCL_ERR_ARG_SIZ, "reps < 1");
  hipEvent_t e0, e1;
  DCL_HIP(hipEventCreate(&e0));
  DCL_HIP(hipEventCreate(&e1));
  const int rc =
      autotune(b, phi0, phi, kappa, vel, reps, ms_table, zc_table, claim_table, best_ms, as_stream(stream), e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

extern "C" int dcl_batch_graph_create(dcl_batch b, const void* phi0, void* phi, void* kappa, const void* vel,
                                      void** graph_exec) {
  DCL_TRY(check_step(b, phi0, phi, kappa, vel));
  if (!graph_exec) return set_error(DCL_ERR_ARG_NULL, "graph_exec is NULL");
  Plan p;
  DCL_TRY(make_plan(b, b->shape, b->occ, p));
  const BCell* cells = b->cells;
  const void* cphi = phi;
  void* a1[] = {(void*)&phi0, (void*)&phi, (void*)&vel, (void*)&cells, (void*)&p.tsl, (void*)&p.bsl_bounds,
                (void*)&p.qsl, (void*)&p.k};
  void* a2[] = {(void*)&cphi, (void*)&kappa, (void*)&cells, (void*)&p.tcv, (void*)&p.bcv_bounds, (void*)&p.qcv,
                (void*)&p.k};
  hipGraph_t g;
  DCL_HIP(hipGraphCreate(&g, 0));
  hipKernelNodeParams p1;
  memset(&p1, 0, sizeof(p1));
  p1.func = const_cast<void*>(p.ksl);
  p1.gridDim = p.grid;
  p1.blockDim = p.bsl;
  p1.kernelParams = a1;
  hipKernelNodeParams p2 = p1;
  p2.func = const_cast<void*>(p.kcv);
  p2.blockDim = p.bcv;
  p2.kernelParams = a2;
  hipGraphNode_t n1, n2;
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphAddKernelNode(&n1, g, nullptr, 0, &p1);
  if (e == hipSuccess) e = hipGraphAddKernelNode(&n2, g, &n1, 1, &p2);
  if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "batch graph build", __FILE__, __LINE__);
  *graph_exec = ex;
  return 0;
}

extern "C" int dcl_batch_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return set_error(DCL_ERR_ARG_NULL, "graph_exec is NULL");
  DCL_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), as_stream(stream)));
  return 0;
}

extern "C" int dcl_batch_graph_destroy(void* graph_exec) {
  if (graph_exec) DCL_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
  return 0;
}
