// sl_advect.hip -- semi-Lagrangian advection of phi on gfx950.
//
// Replaces the band loop of LevelSetAdvectSL_3D / _2D
// (LevelSetMethod/AdvectSL.c:23-89): per node, the MAC velocity at the node
// (InterpolateVelocity3D, Common/InterpolateVelocity.c:35-67), the departure
// point (Euler, AdvectSL.c:80-82, or the RK2 midpoint rule of
// ParticleLS3D.c:25-32 run backwards) and a gather of phi0 there (Bilinear3D,
// Common/Grid3D.c:4-31, or the Catmull-Rom tricubic of TricubicInverse.h).
//
// Dense windows: one thread per node, 64x4 node tiles so each wavefront reads
// a contiguous x-run of phi0 and velocity; blocks are numbered so that the
// eight XCDs each take a contiguous run of z-slabs (their L2s then hold the
// slabs their gathers revisit).  Band lists: one thread per band node.
#include <stdlib.h>
#include <string.h>

#include "dcl_ring.hpp"
#include "dcl_host.hpp"

namespace dcl {

// Map the hardware block id so that consecutive ids on one XCD (ids are dealt
// round-robin over the 8 XCDs) walk consecutive tiles.
__device__ __forceinline__ unsigned xcd_remap(unsigned hw, unsigned nblocks) {
  const unsigned per = nblocks / 8u;
  if (per == 0 || hw >= per * 8u) return hw;
  return (hw % 8u) * per + hw / 8u;
}

template <typename T, int DIM, int INTERP, int BT, bool BUF>
__global__ __launch_bounds__(256) void sl_dense(const T* __restrict__ phi0, T* __restrict__ phi,
                                                const T* __restrict__ vel, SLP P, unsigned tx,
                                                unsigned ty, unsigned pbytes, unsigned vbytes) {
  const unsigned nb = tx * ty * gridDim.y;
  const unsigned lb = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, nb);
  const unsigned bx = lb % tx, by = (lb / tx) % ty, bz = lb / (tx * ty);
  const int i = bx * 64 + threadIdx.x;
  const int j = by * 4 + threadIdx.y;
  const int k = bz;
  if (i >= P.w.nx || j >= P.w.ny) return;
  const int id = (k * P.w.ny + j) * P.w.nx + i;
  const bool in = i >= 2 && i <= P.w.nx - 3 && j >= 2 && j <= P.w.ny - 3 &&
                  (DIM == 2 || (k >= 2 && k <= P.w.nz - 3));
  if (!in) {
    phi[id] = phi0[id];
    return;
  }
  if (BUF && DIM == 3)
    phi[id] = sl_point3_buf<T, INTERP, BT>(make_rsrc(phi0, pbytes), make_rsrc(vel, vbytes), P, P.w.px + i,
                                           P.w.py + j, P.w.pz + k);
  else
    phi[id] = sl_point<T, DIM, INTERP, BT>(phi0, vel, P, P.w.px + i, P.w.py + j, P.w.pz + k);
}

template <typename T, int DIM, int INTERP, int BT, bool COMPACT, bool BUF>
__global__ __launch_bounds__(256) void sl_band(const T* __restrict__ phi0, T* __restrict__ out,
                                               const T* __restrict__ vel, SLP P,
                                               const dcl_icoor* __restrict__ band, int64_t n,
                                               unsigned pbytes, unsigned vbytes) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  const dcl_icoor q = band[b];
  T v;
  if (BUF && DIM == 3)
    v = sl_point3_buf<T, INTERP, BT>(make_rsrc(phi0, pbytes), make_rsrc(vel, vbytes), P, q.x, q.y, q.z);
  else
    v = sl_point<T, DIM, INTERP, BT>(phi0, vel, P, q.x, q.y, DIM == 3 ? q.z : P.w.pz);
  if (COMPACT) {
    out[b] = v;
  } else {
    const int x = q.x - P.w.px, y = q.y - P.w.py, z = DIM == 3 ? q.z - P.w.pz : 0;
    if (x < 0 || x >= P.w.nx || y < 0 || y >= P.w.ny || z < 0 || z >= P.w.nz) return;
    out[(z * P.w.ny + y) * P.w.nx + x] = v;
  }
}

// Dense 3D windows: phi0 staged in an LDS ring (sl_ring3d) or gathered from
// global memory per node (sl_dense).  Measured on 256^3 / 384^3 windows, the
// ring is faster for the 64-value tricubic stencil and, with the velocity
// per node, for the trilinear RK2 midpoint; the gather for trilinear Euler,
// and for trilinear RK2 with component planes (FP64 256^3: 0.25 ms against
// the ring's 0.29).  DCL_SL_KERNEL=ring|gather overrides the
// choice (read at every call); DCL_SL_KERNEL=plain selects the pointer-based
// gather that windows of 2 GiB and more take.
static const char* sl_kernel_env() { return getenv("DCL_SL_KERNEL"); }

static bool sl_ring_selected(int interp, int bt, bool planes) {
  const char* e = sl_kernel_env();
  if (e && (strcmp(e, "gather") == 0 || strcmp(e, "plain") == 0)) return false;
  if (e && strcmp(e, "ring") == 0) return true;
  return interp == DCL_CUBIC || (bt == DCL_RK2 && !planes);
}

template <typename T, int DIM, int INTERP, int BT, bool BUF>
static int launch_sl_b(const void* phi0, void* phi, const void* vel, const SLP& P,
                       const dcl_icoor* band, int64_t band_len, void* band_values, unsigned pb,
                       unsigned vb, hipStream_t s) {
  const T* p0 = static_cast<const T*>(phi0);
  const T* v = static_cast<const T*>(vel);
  if (!band && BUF && DIM == 3 && sl_ring_selected(INTERP, BT, P.v.dof == 1)) {
    // 32 x 8 columns (DCL_RING_TILE=0|1|2 selects 16 x 16, 32 x 8, 64 x 4);
    // velocity staged too where it can be (same window, 3 components) and
    // measured faster: FP32, or FP64 with the trilinear stencil (the FP64
    // tricubic kernel is register-bound with it).  DCL_RING_VEL=0|1 overrides.
    // z-chunks of 32 planes, halved (down to 8) while the grid has fewer
    // than 4 blocks per CU: small windows (128^3 fills 256 blocks at 32)
    // need the blocks more than the shorter chunks' extra halo planes
    const char* ge = getenv("DCL_RING_TILE");
    const int g = ge ? atoi(ge) : 1;
    const char* ve = getenv("DCL_RING_VEL");
    const bool same = P.v.w.px == P.w.px && P.v.w.py == P.w.py && P.v.w.pz == P.w.pz && P.v.w.nx == P.w.nx &&
                      P.v.w.ny == P.w.ny && P.v.w.nz == P.w.nz && P.v.dof == 3 && P.v.udof == 0;
    const bool vl = same && (ve ? atoi(ve) != 0 : (sizeof(T) == 4 || INTERP == DCL_LINEAR));
#define DCL_RING(GI, TXV, TYV, VLV)                                                                      \
  if (g == GI && vl == VLV) {                                                                            \
    const unsigned tx = (P.w.nx + TXV - 1) / TXV, ty = (P.w.ny + TYV - 1) / TYV;                         \
    int zc = 32;                                                                                         \
    while (zc > 8 && (int64_t)tx * ty * ((P.w.nz + zc - 1) / zc) < 4 * (int64_t)device_cus()) zc /= 2;   \
    const unsigned nzc = (P.w.nz + zc - 1) / zc;                                                         \
    hipLaunchKernelGGL((sl_ring3d<T, INTERP, BT, TXV, TYV, VLV>), dim3(tx * ty * nzc), dim3(TXV, TYV), 0, \
                       s, p0, static_cast<T*>(phi), v, P, tx, ty, nzc, zc, pb, vb);                      \
  }
    DCL_RING(0, 16, 16, false)
    DCL_RING(1, 32, 8, false)
    DCL_RING(2, 64, 4, false)
    DCL_RING(0, 16, 16, true)
    DCL_RING(1, 32, 8, true)
    DCL_RING(2, 64, 4, true)
#undef DCL_RING
    DCL_LAUNCHED("sl_ring3d");
    return 0;
  }
  if (!band) {
    const unsigned tx = (P.w.nx + 63) / 64, ty = (P.w.ny + 3) / 4;
    dim3 grid(tx * ty, DIM == 3 ? P.w.nz : 1);
    hipLaunchKernelGGL((sl_dense<T, DIM, INTERP, BT, BUF>), grid, dim3(64, 4), 0, s, p0,
                       static_cast<T*>(phi), v, P, tx, ty, pb, vb);
    DCL_LAUNCHED("sl_dense");
    return 0;
  }
  if (band_len == 0) return 0;
  const unsigned g = grid1d(band_len, 256);
  if (band_values)
    hipLaunchKernelGGL((sl_band<T, DIM, INTERP, BT, true, BUF>), dim3(g), dim3(256), 0, s, p0,
                       static_cast<T*>(band_values), v, P, band, band_len, pb, vb);
  else
    hipLaunchKernelGGL((sl_band<T, DIM, INTERP, BT, false, BUF>), dim3(g), dim3(256), 0, s, p0,
                       static_cast<T*>(phi), v, P, band, band_len, pb, vb);
  DCL_LAUNCHED("sl_band");
  return 0;
}

// 3D windows whose bytes fit a buffer descriptor take the buffer-load path
template <typename T, int DIM, int INTERP, int BT>
static int launch_sl(const void* phi0, void* phi, const void* vel, const SLP& P,
                     const dcl_icoor* band, int64_t band_len, void* band_values,
                     hipStream_t s) {
  const long long pb = (long long)P.w.nx * P.w.ny * P.w.nz * (long long)sizeof(T);
  const long long vb =
      (long long)P.v.w.nx * P.v.w.ny * P.v.w.nz * (P.v.dof == 1 ? DIM : P.v.dof) * (long long)sizeof(T);
  const char* e = sl_kernel_env();
  if (DIM == 3 && buf_ok(pb, vb) && !(e && strcmp(e, "plain") == 0))
    return launch_sl_b<T, DIM, INTERP, BT, true>(phi0, phi, vel, P, band, band_len, band_values,
                                                  (unsigned)pb, (unsigned)vb, s);
  return launch_sl_b<T, DIM, INTERP, BT, false>(phi0, phi, vel, P, band, band_len, band_values, 0, 0, s);
}

template <typename T>
static int dispatch_sl(int dim, int interp, int bt, const void* phi0, void* phi, const void* vel,
                       const SLP& P, const dcl_icoor* band, int64_t n, void* bv, hipStream_t s) {
#define DCL_SL(D, I, B) \
  if (dim == D && interp == I && bt == B) return launch_sl<T, D, I, B>(phi0, phi, vel, P, band, n, bv, s);
  DCL_SL(3, DCL_LINEAR, DCL_EULER)
  DCL_SL(3, DCL_LINEAR, DCL_RK2)
  DCL_SL(3, DCL_CUBIC, DCL_EULER)
  DCL_SL(3, DCL_CUBIC, DCL_RK2)
  DCL_SL(2, DCL_LINEAR, DCL_EULER)
  DCL_SL(2, DCL_LINEAR, DCL_RK2)
  DCL_SL(2, DCL_CUBIC, DCL_EULER)
  DCL_SL(2, DCL_CUBIC, DCL_RK2)
#undef DCL_SL
  return set_error(DCL_ERR_ARG_WRONG, "interp %d / backtrace %d not supported", interp, bt);
}

int sl_params(const dcl_window* win, const dcl_velocity* vel, double dt, SLP& P) {
  P.w = kwin(*win);
  P.v = kvel(*vel);
  const double d[3] = {win->d.x, win->d.y, win->d.z};
  for (int a = 0; a < 3; ++a) {
    P.c[a] = d[a] > 0 ? dt / d[a] : 0.0;
    P.ch[a] = d[a] > 0 ? dt / (2.0 * d[a]) : 0.0;
  }
  return 0;
}

int sl_launch(const void* phi0, void* phi, const dcl_window* win, const dcl_velocity* vel,
              const dcl_icoor* band, int64_t band_len, void* band_values, double dt,
              int precision, int interp, int backtrace, hipStream_t s) {
  const int dim = win->n.z < 2 ? 2 : 3;
  SLP P;
  sl_params(win, vel, dt, P);
  if (precision == DCL_F64)
    return dispatch_sl<double>(dim, interp, backtrace, phi0, phi, vel->data, P, band, band_len,
                               band_values, s);
  return dispatch_sl<float>(dim, interp, backtrace, phi0, phi, vel->data, P, band, band_len,
                            band_values, s);
}

int check_sl_args(const void* phi0, void* phi, const dcl_window* win, const dcl_velocity* vel,
                  const dcl_icoor* band, int64_t band_len, void* band_values, int precision,
                  int interp, int backtrace) {
  DCL_TRY(check_precision(precision));
  if (!phi0) return set_error(DCL_ERR_ARG_NULL, "phi0 is NULL");
  const int min_n = band ? 2 : 5;
  DCL_TRY(check_window(win, "phi", min_n));
  const int dim = win->n.z < 2 ? 2 : 3;
  DCL_TRY(check_velocity(vel, dim, true));
  if (interp != DCL_LINEAR && interp != DCL_CUBIC)
    return set_error(DCL_ERR_ARG_WRONG, "interp %d is neither DCL_LINEAR nor DCL_CUBIC", interp);
  if (backtrace != DCL_EULER && backtrace != DCL_RK2)
    return set_error(DCL_ERR_ARG_WRONG, "backtrace %d is neither DCL_EULER nor DCL_RK2", backtrace);
  if (interp == DCL_CUBIC && (win->n.x < 4 || win->n.y < 4 || (dim == 3 && win->n.z < 4)))
    return set_error(DCL_ERR_ARG_SIZ, "cubic interpolation needs 4 nodes per axis");
  if (band && band_len < 0) return set_error(DCL_ERR_ARG_SIZ, "band_len %lld < 0", (long long)band_len);
  if (!band || !band_values) {
    if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
    if (phi == phi0) return set_error(DCL_ERR_ARG_WRONG, "phi and phi0 must be distinct buffers");
  }
  return 0;
}

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_sl_advect(const void* phi0, void* phi, const dcl_window* win,
                             const dcl_velocity* vel, const dcl_icoor* band, int64_t band_len,
                             void* band_values, double dt, int precision, int interp,
                             int backtrace, void* stream) {
  DCL_TRY(check_sl_args(phi0, phi, win, vel, band, band_len, band_values, precision, interp,
                        backtrace));
  return sl_launch(phi0, phi, win, vel, band, band_len, band_values, dt, precision, interp,
                   backtrace, as_stream(stream));
}
