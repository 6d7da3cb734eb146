// dcl_curv.hpp -- pointwise curvature stencils shared by the dense, band,
// fused and batched kernels.
#pragma once

#include "dcl_common.hpp"

namespace dcl {

// reciprocal spacings of the 3D stencil: 1/(2d), 1/d^2, 1/(4 d d')
template <typename T>
struct CurvR {
  T x2, y2, z2, xx, yy, zz, xy, xz, yz;
};

template <typename T>
__host__ __device__ inline CurvR<T> curv_recips(double dx, double dy, double dz) {
  CurvR<T> r;
  r.x2 = (T)(1.0 / (2.0 * dx));
  r.y2 = (T)(1.0 / (2.0 * dy));
  r.z2 = (T)(1.0 / (2.0 * dz));
  r.xx = (T)(1.0 / (dx * dx));
  r.yy = (T)(1.0 / (dy * dy));
  r.zz = (T)(1.0 / (dz * dz));
  r.xy = (T)(1.0 / (4. * dx * dy));
  r.xz = (T)(1.0 / (4. * dx * dz));
  r.yz = (T)(1.0 / (4. * dy * dz));
  return r;
}

// GridFunction3D_Curv (Common/Grid3D.c:83-108) from three planes k-1, k, k+1
// (pointers at node (i, j) of each plane, row stride Y): the 7-point
// derivatives, the 12 edge points of the cross terms, NaN -> 0.  Divisions by
// the spacings are multiplications by their reciprocals.
// c / sqrt(g^3), g = |grad phi|^2, as c y^3 with y = 1/sqrt(g): in FP32 the
// hardware reciprocal square root (a few ulp); in FP64 its estimate refined
// by two Newton steps y <- y + y (1 - g y^2) / 2 (within an ulp or two of the
// reference's quotient and square root, a quarter of their instructions).
// At g = 0 the result is NaN (0 * inf), which the callers map to 0 as the
// reference maps its 0 / 0; c is 0 there, every term carrying a gradient
// component.
template <typename T>
__device__ __forceinline__ T grad_pow3_div(T c, T g) {
  if constexpr (sizeof(T) == 4) {
    const T q = __builtin_amdgcn_rsqf(g);
    return c * (q * q * q);
  } else {
    double y = __builtin_amdgcn_rsq(g);
    double e = fma(-(g * y), y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-(g * y), y, 1.0);
    y = fma(0.5 * y, e, y);
    return c * (y * y * y);
  }
}

template <typename T>
__device__ __forceinline__ T curv3d_planes(const T* pm, const T* p, const T* pp, int Y, const CurvR<T>& r) {
  const T c0 = p[0];
  const T px = (p[1] - p[-1]) * r.x2;
  const T py = (p[Y] - p[-Y]) * r.y2;
  const T pz = (pp[0] - pm[0]) * r.z2;
  const T px2 = px * px, py2 = py * py, pz2 = pz * pz;
  const T ppp = px2 + py2 + pz2;
  const T pxx = (p[-1] - T(2.) * c0 + p[1]) * r.xx;
  const T pyy = (p[-Y] - T(2.) * c0 + p[Y]) * r.yy;
  const T pzz = (pm[0] - T(2.) * c0 + pp[0]) * r.zz;
  const T pxy = (p[-Y - 1] + p[Y + 1] - p[Y - 1] - p[-Y + 1]) * r.xy;
  const T pxz = (pm[-1] + pp[1] - pp[-1] - pm[1]) * r.xz;
  const T pyz = (pm[-Y] + pp[Y] - pp[-Y] - pm[Y]) * r.yz;
  T c = px2 * (pyy + pzz) + py2 * (pxx + pzz) + pz2 * (pxx + pyy);
  c -= T(2.) * (px * py * pxy + px * pz * pxz + py * pz * pyz);
  c = grad_pow3_div<T>(c, ppp);
  if (c != c) c = T(0);
  return c;
}

// the same at a node of a window with plane stride Z
template <typename T>
__device__ __forceinline__ T curv3d_at(const T* __restrict__ p, int Y, int Z, const CurvR<T>& r) {
  return curv3d_planes<T>(p - Z, p, p + Z, Y, r);
}

template <typename T>
__device__ __forceinline__ T curv2d_at(const T* __restrict__ p, int Y, T dx, T dy) {
  const T px = (p[1] - p[-1]) / (T(2.) * dx);
  const T py = (p[Y] - p[-Y]) / (T(2.) * dy);
  const T px2 = px * px, py2 = py * py;
  const T pxx = (p[-1] - T(2.) * p[0] + p[1]) / (dx * dx);
  const T pyy = (p[-Y] - T(2.) * p[0] + p[Y]) / (dy * dy);
  const T pxy = (p[-Y - 1] + p[Y + 1] - p[Y - 1] - p[-Y + 1]) / (T(4.) * dx * dy);
  T k = grad_pow3_div<T>(pxx * py2 - T(2.) * px * py * pxy + pyy * px2, px2 + py2);
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

}  // namespace dcl
