// surface.hip -- surface quantities at the irregular nodes on gfx950:
// IIMUpdateSurfaceQuantities_3D / _2D (IIM/SurfaceQuantities.c:31-118).  At
// each record's X: the unit normal from Bilinear3D (Common/Grid3D.c:4-31) /
// Bilinear2D (Common/Grid.c:190-212) of the central-difference functionals
// (Common/Grid3D.c:38-108, Common/Grid.c:219-254), the tangents of
// LocalCoor3DTangential (IIM/LocalCoor.c:38-65), the principal curvatures
// from the Hessian rotated by IIMGlobalToLocal_1st / _2nd
// (IIM/Rotations.c:45-63), and the curvature functional.
//
// One thread per record.  The functionals at the 8 (4) corners of X's cell
// read a 4x4x4 (4x4) neighbourhood, held in registers (compile-time indices).
// All arithmetic is double, one IEEE operation at a time in the reference's
// operand order, so the results are the same on every run and device.  Reads
// clamp the index into the window.
#include <stddef.h>

#include "dcl_host.hpp"

namespace dcl {

constexpr int SQB = 128;

// neighbourhood value at offset (a, b, c) from the cell's corner minus one
#define NB3(a, b, c) nb[((c) * 4 + (b)) * 4 + (a)]
#define NB2(a, b) nb[(b) * 4 + (a)]

enum { SQ_DX, SQ_DY, SQ_DZ, SQ_DXX, SQ_DYY, SQ_DZZ, SQ_DXY, SQ_DXZ, SQ_DYZ, SQ_CURV };

// GridFunction3D_* (Common/Grid3D.c:38-108) at neighbourhood node (a, b, c)
template <int KIND>
__device__ __forceinline__ double sq_gf3(const double* nb, int a, int b, int c, double dx, double dy, double dz) {
#pragma clang fp contract(off)
#define P(u, v, w) NB3(u, v, w)
  if (KIND == SQ_DX) return (P(a + 1, b, c) - P(a - 1, b, c)) / (2.0 * dx);
  if (KIND == SQ_DY) return (P(a, b + 1, c) - P(a, b - 1, c)) / (2.0 * dy);
  if (KIND == SQ_DZ) return (P(a, b, c + 1) - P(a, b, c - 1)) / (2.0 * dz);
  if (KIND == SQ_DXX) return (P(a - 1, b, c) - 2. * P(a, b, c) + P(a + 1, b, c)) / (dx * dx);
  if (KIND == SQ_DYY) return (P(a, b - 1, c) - 2. * P(a, b, c) + P(a, b + 1, c)) / (dy * dy);
  if (KIND == SQ_DZZ) return (P(a, b, c - 1) - 2. * P(a, b, c) + P(a, b, c + 1)) / (dz * dz);
  if (KIND == SQ_DXY)
    return (P(a - 1, b - 1, c) + P(a + 1, b + 1, c) - P(a - 1, b + 1, c) - P(a + 1, b - 1, c)) / (4. * dx * dy);
  if (KIND == SQ_DXZ)
    return (P(a - 1, b, c - 1) + P(a + 1, b, c + 1) - P(a - 1, b, c + 1) - P(a + 1, b, c - 1)) / (4. * dx * dz);
  if (KIND == SQ_DYZ)
    return (P(a, b - 1, c - 1) + P(a, b + 1, c + 1) - P(a, b - 1, c + 1) - P(a, b + 1, c - 1)) / (4. * dy * dz);
  const double px = (P(a + 1, b, c) - P(a - 1, b, c)) / (2.0 * dx);
  const double py = (P(a, b + 1, c) - P(a, b - 1, c)) / (2.0 * dy);
  const double pz = (P(a, b, c + 1) - P(a, b, c - 1)) / (2.0 * dz);
  const double px2 = px * px, py2 = py * py, pz2 = pz * pz;
  const double ppp = px2 + py2 + pz2;
  const double pxx = (P(a - 1, b, c) - 2. * P(a, b, c) + P(a + 1, b, c)) / (dx * dx);
  const double pyy = (P(a, b - 1, c) - 2. * P(a, b, c) + P(a, b + 1, c)) / (dy * dy);
  const double pzz = (P(a, b, c - 1) - 2. * P(a, b, c) + P(a, b, c + 1)) / (dz * dz);
  const double pxy =
      (P(a - 1, b - 1, c) + P(a + 1, b + 1, c) - P(a - 1, b + 1, c) - P(a + 1, b - 1, c)) / (4. * dx * dy);
  const double pxz =
      (P(a - 1, b, c - 1) + P(a + 1, b, c + 1) - P(a - 1, b, c + 1) - P(a + 1, b, c - 1)) / (4. * dx * dz);
  const double pyz =
      (P(a, b - 1, c - 1) + P(a, b + 1, c + 1) - P(a, b - 1, c + 1) - P(a, b + 1, c - 1)) / (4. * dy * dz);
  double k = px2 * (pyy + pzz) + py2 * (pxx + pzz) + pz2 * (pxx + pyy);
  k -= 2. * (px * py * pxy + px * pz * pxz + py * pz * pyz);
  k /= sqrt(ppp * ppp * ppp);
  if (k != k) k = 0;
  return k;
#undef P
}

// Bilinear3D (Common/Grid3D.c:4-31) of functional KIND: loop order i, j, k
template <int KIND>
__device__ __forceinline__ double sq_bilinear3(const double* nb, double sx, double sy, double sz, double dx,
                                               double dy, double dz) {
#pragma clang fp contract(off)
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double xb = i * sx + (1 - i) * (1 - sx);
        const double yb = j * sy + (1 - j) * (1 - sy);
        const double zb = k * sz + (1 - k) * (1 - sz);
        sum += sq_gf3<KIND>(nb, 1 + i, 1 + j, 1 + k, dx, dy, dz) * xb * yb * zb;
      }
  return sum;
}

// GridFunction2D_* (Common/Grid.c:219-254) at neighbourhood node (a, b)
template <int KIND>
__device__ __forceinline__ double sq_gf2(const double* nb, int a, int b, double dx, double dy) {
#pragma clang fp contract(off)
#define P(u, v) NB2(u, v)
  if (KIND == SQ_DX) return (P(a + 1, b) - P(a - 1, b)) / (2. * dx);
  if (KIND == SQ_DY) return (P(a, b + 1) - P(a, b - 1)) / (2. * dy);
  const double px = (P(a + 1, b) - P(a - 1, b)) / (2. * dx);
  const double py = (P(a, b + 1) - P(a, b - 1)) / (2. * dy);
  const double px2 = px * px, py2 = py * py;
  const double pxx = (P(a - 1, b) - 2. * P(a, b) + P(a + 1, b)) / (dx * dx);
  const double pyy = (P(a, b - 1) - 2. * P(a, b) + P(a, b + 1)) / (dy * dy);
  const double pxy = (P(a - 1, b - 1) + P(a + 1, b + 1) - P(a - 1, b + 1) - P(a + 1, b - 1)) / (4. * dx * dy);
  double k = (pxx * py2 - 2. * px * py * pxy + pyy * px2) / sqrt((px2 + py2) * (px2 + py2) * (px2 + py2));
  if (k != k) k = 0;
  return k;
#undef P
}

// Bilinear2D (Common/Grid.c:190-212): the (2x - 1) form, loop order x, y
template <int KIND>
__device__ __forceinline__ double sq_bilinear2(const double* nb, double xf, double yf, double dx, double dy) {
#pragma clang fp contract(off)
  double sum = 0.;
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const double xb = 1. + (2. * k - 1.) * xf;
      const double yb = 1. + (2. * l - 1.) * yf;
      sum += sq_gf2<KIND>(nb, 1 + k, 1 + l, dx, dy) * xb * yb;
    }
  sum /= 4.;
  return sum;
}

template <typename T>
__device__ __forceinline__ double sq_at(const T* __restrict__ phi, const KWin& w, int x, int y, int z) {
  const int lx = clampi(x - w.px, 0, w.nx - 1), ly = clampi(y - w.py, 0, w.ny - 1);
  const int lz = w.nz < 2 ? 0 : clampi(z - w.pz, 0, w.nz - 1);
  return (double)phi[((size_t)lz * w.ny + ly) * w.nx + lx];
}

template <typename T>
__global__ __launch_bounds__(SQB) void surface3_k(const T* __restrict__ phi, KWin w, double dx, double dy, double dz,
                                                  const dcl_irregular_node* __restrict__ nodes, int64_t count,
                                                  dcl_surface* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * SQB + threadIdx.x;
  if (t >= count) return;
  const double X0 = nodes[t].X.x, X1 = nodes[t].X.y, X2 = nodes[t].X.z;
  const int xs = (int)floor(X0), ys = (int)floor(X1), zs = (int)floor(X2);
  const double sx = X0 - xs, sy = X1 - ys, sz = X2 - zs;
  double nb[64];
  const int lx = xs - 1 - w.px, ly = ys - 1 - w.py, lz = zs - 1 - w.pz;
  if (lx >= 0 && lx + 3 < w.nx && ly >= 0 && ly + 3 < w.ny && lz >= 0 && lz + 3 < w.nz) {
    const T* c0 = phi + ((size_t)lz * w.ny + ly) * w.nx + lx;
    const int64_t py = w.nx, pz = (int64_t)w.nx * w.ny;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int a = 0; a < 4; ++a) NB3(a, b, c) = (double)c0[c * pz + b * py + a];
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int a = 0; a < 4; ++a) NB3(a, b, c) = sq_at(phi, w, xs - 1 + a, ys - 1 + b, zs - 1 + c);
  }
  // Surface normal (SurfaceQuantities.c:86-93)
  double nx = sq_bilinear3<SQ_DX>(nb, sx, sy, sz, dx, dy, dz);
  double ny = sq_bilinear3<SQ_DY>(nb, sx, sy, sz, dx, dy, dz);
  double nz = sq_bilinear3<SQ_DZ>(nb, sx, sy, sz, dx, dy, dz);
  const double h = sqrt(nx * nx + ny * ny + nz * nz);
  nx /= h;
  ny /= h;
  nz /= h;
  // LocalCoor3DTangential (IIM/LocalCoor.c:38-65)
  double s0, s1, s2, r0, r1, r2, g;
  if (nx * nx + ny * ny >= nx * nx + nz * nz) {
    g = sqrt(nx * nx + ny * ny);
    s0 = ny / g;
    s1 = -nx / g;
    s2 = 0;
    r0 = nx * nz;
    r1 = ny * nz;
    r2 = -nx * nx - ny * ny;
  } else {
    g = sqrt(nx * nx + nz * nz);
    s0 = nz / g;
    s1 = 0;
    s2 = -nx / g;
    r0 = -nx * ny;
    r1 = nx * nx + nz * nz;
    r2 = -ny * nz;
  }
  g = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
  r0 /= g;
  r1 /= g;
  r2 /= g;
  // IIMGlobalToLocal_1st of the gradient (IIM/Rotations.c:45-50): e
  const double jx = nx * h, jy = ny * h, jz = nz * h;
  const double je = jx * nx + jy * ny + jz * nz;
  const double xx = sq_bilinear3<SQ_DXX>(nb, sx, sy, sz, dx, dy, dz);
  const double xy = sq_bilinear3<SQ_DXY>(nb, sx, sy, sz, dx, dy, dz);
  const double xz = sq_bilinear3<SQ_DXZ>(nb, sx, sy, sz, dx, dy, dz);
  const double yy = sq_bilinear3<SQ_DYY>(nb, sx, sy, sz, dx, dy, dz);
  const double yz = sq_bilinear3<SQ_DYZ>(nb, sx, sy, sz, dx, dy, dz);
  const double zz = sq_bilinear3<SQ_DZZ>(nb, sx, sy, sz, dx, dy, dz);
  // IIMGlobalToLocal_2nd (IIM/Rotations.c:51-63): the nn, nt and tt entries
  const double nn = s0 * (xx * s0 + xy * s1 + xz * s2) + s1 * (xy * s0 + yy * s1 + yz * s2) +
                    s2 * (xz * s0 + yz * s1 + zz * s2);
  const double nt = r0 * (xx * s0 + xy * s1 + xz * s2) + r1 * (xy * s0 + yy * s1 + yz * s2) +
                    r2 * (xz * s0 + yz * s1 + zz * s2);
  const double tt = r0 * (xx * r0 + xy * r1 + xz * r2) + r1 * (xy * r0 + yy * r1 + yz * r2) +
                    r2 * (xz * r0 + yz * r1 + zz * r2);
  dcl_surface o;
  o.n.x = nx, o.n.y = ny, o.n.z = nz;
  o.s.x = s0, o.s.y = s1, o.s.z = s2;
  o.r.x = r0, o.r.y = r1, o.r.z = r2;
  o.k_nn = -nn / je;
  o.k_tt = -tt / je;
  o.k_nt = -nt / je;
  o.k = sq_bilinear3<SQ_CURV>(nb, sx, sy, sz, dx, dy, dz);
  out[t] = o;
}

template <typename T>
__global__ __launch_bounds__(SQB) void surface2_k(const T* __restrict__ phi, KWin w, double dx, double dy,
                                                  const dcl_irregular_node* __restrict__ nodes, int64_t count,
                                                  dcl_surface* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * SQB + threadIdx.x;
  if (t >= count) return;
  const double x = nodes[t].X.x, y = nodes[t].X.y;
  const int xi = (int)floor(x), yi = (int)floor(y);
  const double xf = (2. * (x - xi) - 1.), yf = (2. * (y - yi) - 1.);
  double nb[16];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int a = 0; a < 4; ++a) NB2(a, b) = sq_at(phi, w, xi - 1 + a, yi - 1 + b, w.pz);
  // SurfaceQuantities.c:46-56
  dcl_surface o;
  o.k = sq_bilinear2<SQ_CURV>(nb, xf, yf, dx, dy);
  o.k_nn = o.k;
  o.k_tt = 0;
  o.k_nt = 0;
  double nx = sq_bilinear2<SQ_DX>(nb, xf, yf, dx, dy);
  double ny = sq_bilinear2<SQ_DY>(nb, xf, yf, dx, dy);
  const double h = sqrt(nx * nx + ny * ny);
  nx = nx / h;
  ny = ny / h;
  o.n.x = nx, o.n.y = ny, o.n.z = 0;
  o.s.x = ny, o.s.y = -nx, o.s.z = 0;
  o.r.x = o.r.y = o.r.z = 0;
  out[t] = o;
}

#undef NB3
#undef NB2

template <typename T>
static int surface_t(const void* phi, const dcl_window* win, const dcl_irregular_node* nodes, int64_t count,
                     dcl_surface* out, hipStream_t s) {
  const T* p = static_cast<const T*>(phi);
  const KWin w = kwin(*win);
  const unsigned nblk = grid1d(count, SQB);
  if (win->n.z >= 2)
    hipLaunchKernelGGL((surface3_k<T>), dim3(nblk), dim3(SQB), 0, s, p, w, win->d.x, win->d.y, win->d.z, nodes, count,
                       out);
  else
    hipLaunchKernelGGL((surface2_k<T>), dim3(nblk), dim3(SQB), 0, s, p, w, win->d.x, win->d.y, nodes, count, out);
  DCL_LAUNCHED("surface quantities");
  return 0;
}

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_surface_quantities(const void* phi, const dcl_window* win, const dcl_irregular_node* nodes,
                                      int64_t count, int precision, dcl_surface* out, void* stream) {
  DCL_TRY(check_precision(precision));
  if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
  if (count < 0) return set_error(DCL_ERR_ARG_SIZ, "count < 0");
  if (count > 0 && (!nodes || !out)) return set_error(DCL_ERR_ARG_NULL, "nodes or out is NULL");
  if (count > INT32_MAX) return set_error(DCL_ERR_ARG_SIZ, "count too large (at most 2^31 - 1 records)");
  DCL_TRY(check_window(win, "phi", 1));
  if (count == 0) return 0;
  hipStream_t s = as_stream(stream);
  if (precision == DCL_F64) return surface_t<double>(phi, win, nodes, count, out, s);
  return surface_t<float>(phi, win, nodes, count, out, s);
}
