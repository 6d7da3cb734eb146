// curvature.hip -- finite-difference curvature of phi on gfx950.
//
//   3D: GridFunction3D_Curv, 19 points (Common/Grid3D.c:83-108)
//   2D: GridFunction2D_Curv, 9 points (Common/Grid.c:231-254)
//   2D least squares: GridFunction2D_CurvLSS, 81 points
//       (LevelSetMethod/Curvature.c:3-186)
// kappa = 0 where the stencil leaves the window (and, for the first two, where
// the formula gives NaN, as in the reference).
#include "dcl_host.hpp"

namespace dcl {

struct CurvP {
  KWin w;
  double dx, dy, dz;
};

// GridFunction3D_Curv at local node (i, j, k) of a window with row stride Y and
// plane stride Z; p points at the node.
template <typename T>
__device__ __forceinline__ T curv3d_at(const T* __restrict__ p, int Y, int Z, T dx, T dy, T dz) {
  const T c0 = p[0];
  const T px = (p[1] - p[-1]) / (T(2.0) * dx);
  const T py = (p[Y] - p[-Y]) / (T(2.0) * dy);
  const T pz = (p[Z] - p[-Z]) / (T(2.0) * dz);
  const T px2 = px * px, py2 = py * py, pz2 = pz * pz;
  const T ppp = px2 + py2 + pz2;
  const T pxx = (p[-1] - T(2.) * c0 + p[1]) / (dx * dx);
  const T pyy = (p[-Y] - T(2.) * c0 + p[Y]) / (dy * dy);
  const T pzz = (p[-Z] - T(2.) * c0 + p[Z]) / (dz * dz);
  const T pxy = (p[-Y - 1] + p[Y + 1] - p[Y - 1] - p[-Y + 1]) / (T(4.) * dx * dy);
  const T pxz = (p[-Z - 1] + p[Z + 1] - p[Z - 1] - p[-Z + 1]) / (T(4.) * dx * dz);
  const T pyz = (p[-Z - Y] + p[Z + Y] - p[Z - Y] - p[-Z + Y]) / (T(4.) * dy * dz);
  T c = px2 * (pyy + pzz) + py2 * (pxx + pzz) + pz2 * (pxx + pyy);
  c -= T(2.) * (px * py * pxy + px * pz * pxz + py * pz * pyz);
  c /= sqrt(ppp * ppp * ppp);
  if (c != c) c = T(0);
  return c;
}

template <typename T>
__device__ __forceinline__ T curv2d_at(const T* __restrict__ p, int Y, T dx, T dy) {
  const T px = (p[1] - p[-1]) / (T(2.) * dx);
  const T py = (p[Y] - p[-Y]) / (T(2.) * dy);
  const T px2 = px * px, py2 = py * py;
  const T pxx = (p[-1] - T(2.) * p[0] + p[1]) / (dx * dx);
  const T pyy = (p[-Y] - T(2.) * p[0] + p[Y]) / (dy * dy);
  const T pxy = (p[-Y - 1] + p[Y + 1] - p[Y - 1] - p[-Y + 1]) / (T(4.) * dx * dy);
  T k = (pxx * py2 - T(2.) * px * py * pxy + pyy * px2) / sqrt((px2 + py2) * (px2 + py2) * (px2 + py2));
  if (k != k) k = T(0);
  return k;
}

// Least-squares fit c0 + c1 a + c2 b + c3 a^2 + c4 ab + c5 b^2 over the 9x9
// offsets; the reference's generated sums are these coefficients, with
// pxx = c3 and pyy = c5 (half the second derivatives).  Decoupled normal
// equations over the symmetric design: c1 = sum a p / 540, c4 = sum ab p /
// 3600, c3 = sum (a^2 - 20/3) p / 2772.
template <typename T>
__device__ __forceinline__ T curv2d_lss_at(const T* __restrict__ p, int Y, T dx, T dy) {
  T s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
  for (int b = -4; b <= 4; ++b) {
    const T* r = p + b * Y;
    const T bb = T(b * b) - T(20.0 / 3.0);
#pragma unroll
    for (int a = -4; a <= 4; ++a) {
      const T v = r[a];
      s1 += T(a) * v;
      s2 += T(b) * v;
      s3 += (T(a * a) - T(20.0 / 3.0)) * v;
      s4 += T(a * b) * v;
      s5 += bb * v;
    }
  }
  const T px = s1 / T(540.0) / dx;
  const T py = s2 / T(540.0) / dy;
  const T pxx = s3 / T(2772.0) / (dx * dx);
  const T pyy = s5 / T(2772.0) / (dy * dy);
  const T pxy = s4 / T(3600.0) / (dx * dy);
  const T px2 = px * px, py2 = py * py;
  return (pxx * py2 - T(2.) * px * py * pxy + pyy * px2) / sqrt((px2 + py2) * (px2 + py2) * (px2 + py2));
}

template <typename T, int DIM, int ST>
__device__ __forceinline__ T curv_node(const T* __restrict__ phi, const CurvP& P, int i, int j, int k) {
  const int m = ST == DCL_CURV_LSS ? 4 : 1;
  if (i < m || i > P.w.nx - 1 - m || j < m || j > P.w.ny - 1 - m) return T(0);
  if (DIM == 3 && (k < 1 || k > P.w.nz - 2)) return T(0);
  const int Y = P.w.nx, Z = P.w.nx * P.w.ny;
  const T* p = phi + (k * P.w.ny + j) * P.w.nx + i;
  if (DIM == 3) return curv3d_at<T>(p, Y, Z, (T)P.dx, (T)P.dy, (T)P.dz);
  if (ST == DCL_CURV_LSS) return curv2d_lss_at<T>(p, Y, (T)P.dx, (T)P.dy);
  return curv2d_at<T>(p, Y, (T)P.dx, (T)P.dy);
}

template <typename T, int DIM, int ST>
__global__ __launch_bounds__(256) void curv_dense(const T* __restrict__ phi, T* __restrict__ kappa,
                                                  CurvP P) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= P.w.nx || j >= P.w.ny) return;
  kappa[(k * P.w.ny + j) * P.w.nx + i] = curv_node<T, DIM, ST>(phi, P, i, j, k);
}

template <typename T, int DIM, int ST, bool COMPACT>
__global__ __launch_bounds__(256) void curv_band(const T* __restrict__ phi, T* __restrict__ out,
                                                 CurvP P, const dcl_icoor* __restrict__ band,
                                                 int64_t n) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  const dcl_icoor q = band[b];
  const int i = q.x - P.w.px, j = q.y - P.w.py, k = DIM == 3 ? q.z - P.w.pz : 0;
  const bool inside = i >= 0 && i < P.w.nx && j >= 0 && j < P.w.ny && k >= 0 && k < P.w.nz;
  const T v = inside ? curv_node<T, DIM, ST>(phi, P, i, j, k) : T(0);
  if (COMPACT) out[b] = v;
  else if (inside) out[(k * P.w.ny + j) * P.w.nx + i] = v;
}

template <typename T, int DIM, int ST>
static int launch_curv(const void* phi, void* kappa, const CurvP& P, const dcl_icoor* band,
                       int64_t n, void* bv, hipStream_t s) {
  const T* p = static_cast<const T*>(phi);
  if (!band) {
    dim3 grid((P.w.nx + 63) / 64, (P.w.ny + 3) / 4, DIM == 3 ? P.w.nz : 1);
    hipLaunchKernelGGL((curv_dense<T, DIM, ST>), grid, dim3(64, 4), 0, s, p,
                       static_cast<T*>(kappa), P);
    DCL_LAUNCHED("curv_dense");
    return 0;
  }
  if (n == 0) return 0;
  if (bv)
    hipLaunchKernelGGL((curv_band<T, DIM, ST, true>), dim3(grid1d(n, 256)), dim3(256), 0, s, p,
                       static_cast<T*>(bv), P, band, n);
  else
    hipLaunchKernelGGL((curv_band<T, DIM, ST, false>), dim3(grid1d(n, 256)), dim3(256), 0, s, p,
                       static_cast<T*>(kappa), P, band, n);
  DCL_LAUNCHED("curv_band");
  return 0;
}

template <typename T>
static int dispatch_curv(int dim, int st, const void* phi, void* kappa, const CurvP& P,
                         const dcl_icoor* band, int64_t n, void* bv, hipStream_t s) {
  if (dim == 3) return launch_curv<T, 3, DCL_CURV_STD>(phi, kappa, P, band, n, bv, s);
  if (st == DCL_CURV_LSS) return launch_curv<T, 2, DCL_CURV_LSS>(phi, kappa, P, band, n, bv, s);
  return launch_curv<T, 2, DCL_CURV_STD>(phi, kappa, P, band, n, bv, s);
}

int curv_launch(const void* phi, void* kappa, const dcl_window* win, const dcl_icoor* band,
                int64_t n, void* bv, int precision, int stencil, hipStream_t s) {
  CurvP P;
  P.w = kwin(*win);
  P.dx = win->d.x;
  P.dy = win->d.y;
  P.dz = win->d.z;
  const int dim = win->n.z < 2 ? 2 : 3;
  if (precision == DCL_F64) return dispatch_curv<double>(dim, stencil, phi, kappa, P, band, n, bv, s);
  return dispatch_curv<float>(dim, stencil, phi, kappa, P, band, n, bv, s);
}

int check_curv_args(const void* phi, void* kappa, const dcl_window* win, const dcl_icoor* band,
                    int64_t n, void* bv, int precision, int stencil) {
  DCL_TRY(check_precision(precision));
  if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
  DCL_TRY(check_window(win, "phi", 1));
  if (stencil != DCL_CURV_STD && stencil != DCL_CURV_LSS)
    return set_error(DCL_ERR_ARG_WRONG, "stencil %d is neither DCL_CURV_STD nor DCL_CURV_LSS", stencil);
  if (stencil == DCL_CURV_LSS && win->n.z >= 2)
    return set_error(DCL_ERR_SUP, "the least-squares curvature stencil is 2D only");
  if (band && n < 0) return set_error(DCL_ERR_ARG_SIZ, "band_len < 0");
  if (!band || !bv) {
    if (!kappa) return set_error(DCL_ERR_ARG_NULL, "kappa is NULL");
    if (kappa == phi) return set_error(DCL_ERR_ARG_WRONG, "kappa and phi must be distinct buffers");
  }
  return 0;
}

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_curvature(const void* phi, void* kappa, const dcl_window* win,
                             const dcl_icoor* band, int64_t band_len, void* band_values,
                             int precision, int stencil, void* stream) {
  DCL_TRY(check_curv_args(phi, kappa, win, band, band_len, band_values, precision, stencil));
  return curv_launch(phi, kappa, win, band, band_len, band_values, precision, stencil,
                     as_stream(stream));
}
