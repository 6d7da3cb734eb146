// batch.hip -- many independent level sets (DCells) in one launch.
//
// The reference advects the DCells of a rank one after another
// (DWorld/DCellsArray.c:57-60), each over its own small window.  Here the
// windows are packed back to back in one buffer per field and cut into tiles;
// a persistent kernel walks the packed tile list (each XCD a contiguous run
// of it, for L2 locality) and runs the SL step of dcl_tile.hpp on every tile,
// then a second persistent kernel does the curvature.  One step
// can be recorded as a hipGraph of these two launches.  Tile shape and
// resident blocks per CU come from a table of configurations; dcl_batch_autotune
// times them all on the caller's buffers and keeps the fastest.
#include <math.h>
#include <string.h>

#include <new>
#include <vector>

#include "dcl_host.hpp"
#include "dcl_ring.hpp"
#include "dcl_tile.hpp"

namespace dcl {

struct BCell {
  long long phi_off;  // element offset of the cell in the packed phi / kappa buffers
  long long vel_off;  // element offset in the packed velocity buffer (3 per node)
  int px, py, pz, nx, ny, nz;
  double c[3], ch[3];  // dt / dh, dt / (2 dh)
  double dx, dy, dz;
};

struct TileShape {
  int tx, ty, tz;
};
// tile shapes of the configuration table (threads = tx * ty).  Shapes
// NBASE .. 2 NBASE - 1 use the same curvature tiles as shape s - NBASE with
// the SL step over 32 x 8 x RING_ZC columns staged in an LDS ring.
constexpr int NBASE = 6;
constexpr int NSHAPES = 2 * NBASE;
constexpr int RING_LIST = NBASE;  // index of the ring column list
constexpr int RING_ZC = 32;
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

__device__ __forceinline__ void tile_range(int ntiles, int& t0, int& t1, int& slot, int& nslot) {
  const int xcd = blockIdx.x % 8;
  nslot = gridDim.x / 8;
  slot = blockIdx.x / 8;
  const int chunk = (ntiles + 7) / 8;
  t0 = xcd * chunk;
  t1 = min(t0 + chunk, ntiles);
}

template <typename T, int INTERP, int BT, int S>
__global__ __launch_bounds__(shape_of(S).tx * shape_of(S).ty) void batch_sl_k(
    const T* __restrict__ phi0, T* __restrict__ phi, const T* __restrict__ vel, const BCell* __restrict__ cells,
    const int4* __restrict__ tiles, int ntiles) {
  constexpr TileShape sh = shape_of(S);
  int t0, t1, slot, nslot;
  tile_range(ntiles, t0, t1, slot, nslot);
  for (int t = t0 + slot; t < t1; t += nslot) {
    const int4 tl = tiles[t];
    const BCell c = cells[tl.x];
    SLP P;
    P.w.px = c.px; P.w.py = c.py; P.w.pz = c.pz;
    P.w.nx = c.nx; P.w.ny = c.ny; P.w.nz = c.nz;
    P.v.w = P.w;
    P.v.dof = 3;
    P.v.udof = 0;
    for (int a = 0; a < 3; ++a) {
      P.c[a] = c.c[a];
      P.ch[a] = c.ch[a];
    }
    const unsigned nodes = (unsigned)c.nx * c.ny * c.nz;
    const rsrc_t rp = make_rsrc(phi0 + c.phi_off, nodes * (unsigned)sizeof(T));
    const rsrc_t rv = make_rsrc(vel + c.vel_off, 3u * nodes * (unsigned)sizeof(T));
    sl_tile<T, INTERP, BT, sh.tx, sh.ty, sh.tz>(rp, rv, phi0 + c.phi_off, phi + c.phi_off, P, tl.y, tl.z, tl.w);
  }
}

// ring columns of the batch: 32 x 8 nodes, velocity staged in FP32
constexpr int RING_TX = 32, RING_TY = 8;

template <typename T, int INTERP, int BT>
__global__ __launch_bounds__(256) void batch_ring_k(const T* __restrict__ phi0, T* __restrict__ phi,
                                                      const T* __restrict__ vel, const BCell* __restrict__ cells,
                                                      const int4* __restrict__ tiles, int ntiles) {
  typedef RingGeom<RING_TX, RING_TY> G;
  typedef VelGeom<RING_TX, RING_TY> V;
  constexpr bool VL = sizeof(T) == 4;
  __shared__ T ring[RG_R * G::PLANE];
  __shared__ T vr[VL ? VR_R * 3 * V::PLANE : 1];
  int t0, t1, slot, nslot;
  tile_range(ntiles, t0, t1, slot, nslot);
  for (int t = t0 + slot; t < t1; t += nslot) {
    const int4 tl = tiles[t];
    const BCell c = cells[tl.x];
    SLP P;
    P.w.px = c.px; P.w.py = c.py; P.w.pz = c.pz;
    P.w.nx = c.nx; P.w.ny = c.ny; P.w.nz = c.nz;
    P.v.w = P.w;
    P.v.dof = 3;
    P.v.udof = 0;
    for (int a = 0; a < 3; ++a) {
      P.c[a] = c.c[a];
      P.ch[a] = c.ch[a];
    }
    const unsigned nodes = (unsigned)c.nx * c.ny * c.nz;
    const rsrc_t rp = make_rsrc(phi0 + c.phi_off, nodes * (unsigned)sizeof(T));
    const rsrc_t rv = make_rsrc(vel + c.vel_off, 3u * nodes * (unsigned)sizeof(T));
    ring_column<T, INTERP, BT, RING_ZC, G, VL>(phi0 + c.phi_off, phi + c.phi_off, rp, rv, P, tl.y, tl.z, tl.w,
                                               ring, vr);
  }
}

template <typename T, int S>
__global__ __launch_bounds__(shape_of(S).tx * shape_of(S).ty) void batch_curv_k(
    const T* __restrict__ phi, T* __restrict__ kappa, const BCell* __restrict__ cells,
    const int4* __restrict__ tiles, int ntiles) {
  constexpr TileShape sh = shape_of(S);
  __shared__ T pl[4 * (sh.tx + 2) * (sh.ty + 2)];
  int t0, t1, slot, nslot;
  tile_range(ntiles, t0, t1, slot, nslot);
  for (int t = t0 + slot; t < t1; t += nslot) {
    const int4 tl = tiles[t];
    const BCell c = cells[tl.x];
    KWin w;
    w.px = c.px; w.py = c.py; w.pz = c.pz;
    w.nx = c.nx; w.ny = c.ny; w.nz = c.nz;
    const unsigned nodes = (unsigned)c.nx * c.ny * c.nz;
    curv_column<T, sh.tx, sh.ty, sh.tz>(make_rsrc(phi + c.phi_off, nodes * (unsigned)sizeof(T)), kappa + c.phi_off,
                                        w, curv_recips<T>(c.dx, c.dy, c.dz), tl.y, tl.z, tl.w, pl);
  }
}

typedef void (*SLKernel)(const void*, void*, const void*, const BCell*, const int4*, int);

// the kernel pointers of one precision / scheme for every shape
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

template <typename T>
static const void* ring_kernel(int interp, int bt) {
  if (interp == DCL_CUBIC && bt == DCL_RK2) return (const void*)&batch_ring_k<T, DCL_CUBIC, DCL_RK2>;
  if (interp == DCL_CUBIC) return (const void*)&batch_ring_k<T, DCL_CUBIC, DCL_EULER>;
  if (bt == DCL_RK2) return (const void*)&batch_ring_k<T, DCL_LINEAR, DCL_RK2>;
  return (const void*)&batch_ring_k<T, DCL_LINEAR, DCL_EULER>;
}

template <typename T>
static const void* sl_kernel(int interp, int bt, int s) {
  if (interp == DCL_CUBIC && bt == DCL_RK2) return KTable<T, DCL_CUBIC, DCL_RK2>::sl(s);
  if (interp == DCL_CUBIC) return KTable<T, DCL_CUBIC, DCL_EULER>::sl(s);
  if (bt == DCL_RK2) return KTable<T, DCL_LINEAR, DCL_RK2>::sl(s);
  return KTable<T, DCL_LINEAR, DCL_EULER>::sl(s);
}

}  // namespace dcl

using namespace dcl;

struct dcl_batch_s {
  int ncells, precision, interp, backtrace;
  std::vector<dcl_window> wins;
  std::vector<BCell> host_cells;
  std::vector<long long> phi_off, vel_off;
  BCell* cells = nullptr;
  int4* tiles[NBASE + 1] = {nullptr};  // per tile shape, then the ring column list
  int ntiles[NBASE + 1] = {0};
  int shape = 1, occ = 1;
  int cus = 256;
  bool dt_set = false;
};

// tile list s < NBASE: tiles of shape s; s == RING_LIST: 32 x 8 x RING_ZC columns
static int build_tiles(dcl_batch b, int s) {
  if (b->tiles[s]) return 0;
  const TileShape sh = s == RING_LIST ? TileShape{RING_TX, RING_TY, RING_ZC} : shape_of(s);
  std::vector<int4> t;
  for (int c = 0; c < b->ncells; ++c) {
    const dcl_window& w = b->wins[c];
    for (int z = 0; z < w.n.z; z += sh.tz)
      for (int y = 0; y < w.n.y; y += sh.ty)
        for (int x = 0; x < w.n.x; x += sh.tx) t.push_back(make_int4(c, x, y, z));
  }
  if (t.size() >= (size_t)1 << 31) return set_error(DCL_ERR_ARG_SIZ, "too many tiles");
  void* p = nullptr;
  DCL_TRY(dcl_malloc(&p, sizeof(int4) * t.size()));
  hipError_t e = hipMemcpy(p, t.data(), sizeof(int4) * t.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dcl_free(p);
    return hip_fail(e, "tile list upload", __FILE__, __LINE__);
  }
  b->tiles[s] = static_cast<int4*>(p);
  b->ntiles[s] = (int)t.size();
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
  int nsl, ncv;
};

static int make_plan(dcl_batch b, int s, int occ, Plan& p) {
  const int cs = s % NBASE;
  const bool ring = s >= NBASE;
  DCL_TRY(build_tiles(b, cs));
  if (ring) DCL_TRY(build_tiles(b, RING_LIST));
  const TileShape sh = shape_of(cs);
  const bool f64 = b->precision == DCL_F64;
  p.grid = dim3(grid_blocks(b, occ));
  p.bcv = dim3(sh.tx, sh.ty);
  p.tcv = b->tiles[cs];
  p.ncv = b->ntiles[cs];
  p.kcv = f64 ? curv_kernel<double>(cs) : curv_kernel<float>(cs);
  if (ring) {
    p.ksl = f64 ? ring_kernel<double>(b->interp, b->backtrace) : ring_kernel<float>(b->interp, b->backtrace);
    p.bsl = dim3(RING_TX, RING_TY);
    p.tsl = b->tiles[RING_LIST];
    p.nsl = b->ntiles[RING_LIST];
  } else {
    p.ksl = f64 ? sl_kernel<double>(b->interp, b->backtrace, cs)
                : sl_kernel<float>(b->interp, b->backtrace, cs);
    p.bsl = p.bcv;
    p.tsl = p.tcv;
    p.nsl = p.ncv;
  }
  return 0;
}

// enqueue one step with a given configuration
static int batch_launch(dcl_batch b, int s, int occ, const void* phi0, void* phi, void* kappa, const void* vel,
                        hipStream_t st) {
  Plan p;
  DCL_TRY(make_plan(b, s, occ, p));
  const BCell* cells = b->cells;
  void* a1[] = {(void*)&phi0, (void*)&phi, (void*)&vel, (void*)&cells, (void*)&p.tsl, (void*)&p.nsl};
  DCL_HIP(hipLaunchKernel(p.ksl, p.grid, p.bsl, a1, 0, st));
  const void* cphi = phi;
  void* a2[] = {(void*)&cphi, (void*)&kappa, (void*)&cells, (void*)&p.tcv, (void*)&p.ncv};
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
  for (int c = 0; c < ncells; ++c) {
    DCL_TRY(check_window(&wins[c], "cell", 5, 3));
    if (wins[c].n.z < 2) return set_error(DCL_ERR_SUP, "cell %d is 2D; batches are 3D", c);
    // the kernels address a cell's velocity through a buffer descriptor whose
    // byte size is 32-bit
    const int64_t vbytes = 3 * win_nodes(wins[c]) * (precision == DCL_F64 ? 8 : 4);
    if (vbytes >= ((int64_t)1 << 32))
      return set_error(DCL_ERR_ARG_SIZ, "cell %d: its velocity takes %lld bytes; the limit is 2^32 - 1", c,
                       (long long)vbytes);
  }
  dcl_batch b = new (std::nothrow) dcl_batch_s();
  if (!b) return set_error(DCL_ERR_MEM, "host allocation failed");
  b->ncells = ncells;
  b->precision = precision;
  b->interp = interp;
  b->backtrace = backtrace;
  b->wins.assign(wins, wins + ncells);
  long long po = 0, vo = 0;
  for (int c = 0; c < ncells; ++c) {
    const dcl_window& w = wins[c];
    BCell cell;
    memset(&cell, 0, sizeof(cell));
    cell.phi_off = po;
    cell.vel_off = vo;
    cell.px = w.p.x; cell.py = w.p.y; cell.pz = w.p.z;
    cell.nx = w.n.x; cell.ny = w.n.y; cell.nz = w.n.z;
    cell.dx = w.d.x; cell.dy = w.d.y; cell.dz = w.d.z;
    b->phi_off.push_back(po);
    b->vel_off.push_back(vo);
    b->host_cells.push_back(cell);
    po += win_nodes(w);
    vo += 3 * win_nodes(w);
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
  int rc = dcl_malloc(&p, sizeof(BCell) * ncells);
  if (rc) {
    delete b;
    return rc;
  }
  b->cells = static_cast<BCell*>(p);
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
  DCL_HIP(hipMemcpy(b->cells, b->host_cells.data(), sizeof(BCell) * b->ncells, hipMemcpyHostToDevice));
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
  const TileShape sh = shape_of(b->shape);
  if (tile) {
    tile[0] = sh.tx;
    tile[1] = sh.ty;
    tile[2] = sh.tz;
  }
  if (blocks) *blocks = grid_blocks(b, b->occ);
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

extern "C" int dcl_batch_autotune(dcl_batch b, const void* phi0, void* phi, void* kappa, const void* vel,
                                  int reps, double* ms_table, double* best_ms, void* stream) {
  DCL_TRY(check_step(b, phi0, phi, kappa, vel));
  if (reps < 1) return set_error(DCL_ERR_ARG_SIZ, "reps < 1");
  hipStream_t st = as_stream(stream);
  hipEvent_t e0, e1;
  DCL_HIP(hipEventCreate(&e0));
  DCL_HIP(hipEventCreate(&e1));
  double best = 1e300;
  int bs = b->shape, bo = b->occ;
  int rc = 0;
  for (int s = 0; s < NSHAPES && !rc; ++s)
    for (int o = 0; o < NOCC && !rc; ++o) {
      rc = batch_launch(b, s, o, phi0, phi, kappa, vel, st);  // warm-up (and tile list build)
      if (rc) break;
      if (hipEventRecord(e0, st) != hipSuccess) { rc = set_error(DCL_ERR_LIB, "event record"); break; }
      for (int r = 0; r < reps && !rc; ++r) rc = batch_launch(b, s, o, phi0, phi, kappa, vel, st);
      if (rc) break;
      float ms = 0;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = set_error(DCL_ERR_LIB, "event timing failed");
        break;
      }
      const double per = ms / reps;
      if (ms_table) ms_table[s * NOCC + o] = per;
      if (per < best) {
        best = per;
        bs = s;
        bo = o;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  b->shape = bs;
  b->occ = bo;
  if (best_ms) *best_ms = best;
  return 0;
}

extern "C" int dcl_batch_graph_create(dcl_batch b, const void* phi0, void* phi, void* kappa, const void* vel,
                                      void** graph_exec) {
  DCL_TRY(check_step(b, phi0, phi, kappa, vel));
  if (!graph_exec) return set_error(DCL_ERR_ARG_NULL, "graph_exec is NULL");
  Plan p;
  DCL_TRY(make_plan(b, b->shape, b->occ, p));
  const BCell* cells = b->cells;
  const void* cphi = phi;
  void* a1[] = {(void*)&phi0, (void*)&phi, (void*)&vel, (void*)&cells, (void*)&p.tsl, (void*)&p.nsl};
  void* a2[] = {(void*)&cphi, (void*)&kappa, (void*)&cells, (void*)&p.tcv, (void*)&p.ncv};
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
