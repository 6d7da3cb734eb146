// dcl_common.hpp -- shared device helpers of the gfx950 level-set kernels.
//
// Coordinates (band nodes, departure points, particles) are global grid-index
// doubles; values (phi, velocity, weights, sums) use the storage type T.
// In-window offsets are 32-bit: the host rejects windows whose element count
// (times dof) reaches 2^31.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dcell_lsm.h"

namespace dcl {

// window geometry as the kernels see it
struct KWin {
  int px, py, pz;
  int nx, ny, nz;
};

// velocity as the kernels see it: component c of node (x, y, z) (local) is
// element ((z ny + y) nx + x) dof + udof + c cs; cs = 1 for components
// stored per node, the window's node count for component planes (dof 1)
struct KVel {
  KWin w;
  int dof, udof, cs;
};

__host__ inline KWin kwin(const dcl_window& w) {
  KWin k;
  k.px = w.p.x; k.py = w.p.y; k.pz = w.p.z;
  k.nx = w.n.x; k.ny = w.n.y; k.nz = w.n.z < 1 ? 1 : w.n.z;
  return k;
}

__host__ inline KVel kvel(const dcl_velocity& v) {
  KVel k;
  k.w = kwin(v.win);
  k.dof = v.dof;
  k.udof = v.udof;
  k.cs = v.dof == 1 ? k.w.nx * k.w.ny * k.w.nz : 1;
  return k;
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) {
  return fmin(fmax(v, lo), hi);  // NaN -> lo: a bad coordinate never leaves the window
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// linear split of a global coordinate: local base index in [0, n-2], fraction in [0, 1]
__device__ __forceinline__ int lin_split(double s, int p, int n, double& f) {
  s = clampd(s, (double)p, (double)(p + n - 1));
  double fl = floor(s);
  int b = (int)fl;
  if (b > p + n - 2) b = p + n - 2;
  f = s - (double)b;
  return b - p;
}

// cubic split: local base index in [1, n-3] (stencil base-1 .. base+2)
__device__ __forceinline__ int cub_split(double s, int p, int n, double& f) {
  s = clampd(s, (double)(p + 1), (double)(p + n - 2));
  double fl = floor(s);
  int b = (int)fl;
  if (b > p + n - 3) b = p + n - 3;
  f = s - (double)b;
  return b - p;
}

// Catmull-Rom (Keys a = -1/2) weights of the stencil offsets -1, 0, 1, 2.
// The weights and the stencil sums below are fixed sequences of IEEE
// operations (contraction off, fused multiply-adds written out), so every
// kernel that inlines them -- LDS-staged or gathering from memory -- gives
// the same bits.
template <typename T>
__device__ __forceinline__ void cr_weights(T t, T w[4]) {
#pragma clang fp contract(off)
  const T t2 = t * t, t3 = t2 * t;
  w[0] = T(0.5) * (fma(T(2), t2, -t3) - t);
  w[1] = T(0.5) * (fma(T(3), t3, -(T(5) * t2)) + T(2));
  w[2] = T(0.5) * (fma(T(4), t2, t) - T(3) * t3);
  w[3] = T(0.5) * (t3 - t2);
}

// Catmull-Rom tricubic sum over 16 rows of 4 values; row(j, k, r) fills r
// with the row at stencil offsets (y, z) = (j, k).  Every product-sum is
// contracted to an fma.
template <typename T, typename Row>
__device__ __forceinline__ T cubic_sum(const T wx[4], const T wy[4], const T wz[4], Row row) {
  T sum = T(0);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    T sy = T(0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      T r[4];
      row(j, k, r);
      T sx = T(0);
      sx = fma(wx[0], r[0], sx);
      sx = fma(wx[1], r[1], sx);
      sx = fma(wx[2], r[2], sx);
      sx = fma(wx[3], r[3], sx);
      sy = fma(wy[j], sx, sy);
    }
    sum = fma(wz[k], sy, sum);
  }
  return sum;
}

template <typename T>
__device__ __forceinline__ void cr_dweights(T t, T dw[4]) {
  const T t2 = t * t;
  dw[0] = T(0.5) * (T(4) * t - T(3) * t2 - T(1));
  dw[1] = T(0.5) * (T(9) * t2 - T(10) * t);
  dw[2] = T(0.5) * (T(8) * t + T(1) - T(9) * t2);
  dw[3] = T(0.5) * (T(3) * t2 - T(2) * t);
}

// The velocity arithmetic below is written out operation by operation
// (contraction off, fma where written), so the kernels that read the faces
// from memory, from LDS or through buffer loads give the same bits.
template <typename T>
__device__ __forceinline__ T vmid(T a, T b) {  // mean of two faces
#pragma clang fp contract(off)
  return fma(T(0.5), a, T(0.5) * b);
}

template <typename T>
__device__ __forceinline__ T vlerp(T a, T b, T t) {  // a + t (b - a)
#pragma clang fp contract(off)
  return fma(t, b - a, a);
}

// trilinear over the corners a[z][y][x] with fractions (x1, y1, z1)
template <typename T>
__device__ __forceinline__ T vtri(T a000, T a100, T a010, T a110, T a001, T a101, T a011, T a111, T x1, T y1,
                                  T z1) {
  const T r00 = vlerp(a000, a100, x1), r10 = vlerp(a010, a110, x1);
  const T r01 = vlerp(a001, a101, x1), r11 = vlerp(a011, a111, x1);
  return vlerp(vlerp(r00, r10, y1), vlerp(r01, r11, y1), z1);
}

// Velocity at an integer node (the first stage of every backtrace).  With
// the fractions 1/2 along the staggered axis and 0 elsewhere, the eight terms
// of InterpolateVelocity3D (Common/InterpolateVelocity.c:35-67) reduce
// exactly to the mean of the two faces around the node; the node is clamped
// into [vp, vp+vn-2] first.
template <typename T, int DIM>
__device__ __forceinline__ void node_vel(const T* __restrict__ vel, const KVel& v, int X, int Y,
                                         int Z, T V[DIM]) {
  const int x = clampi(X, v.w.px, v.w.px + v.w.nx - 2) - v.w.px;
  const int y = clampi(Y, v.w.py, v.w.py + v.w.ny - 2) - v.w.py;
  const int sx = v.dof, sy = v.w.nx * v.dof;
  if (DIM == 3) {
    const int z = clampi(Z, v.w.pz, v.w.pz + v.w.nz - 2) - v.w.pz;
    const int sz = v.w.ny * sy;
    const T* b = vel + (z * v.w.ny + y) * sy + x * sx + v.udof;
    V[0] = vmid(b[0], b[sx]);
    V[1] = vmid(b[v.cs], b[sy + v.cs]);
    V[2] = vmid(b[2 * v.cs], b[sz + 2 * v.cs]);
  } else {
    const T* b = vel + y * sy + x * sx + v.udof;
    V[0] = vmid(b[0], b[sx]);
    V[1] = vmid(b[v.cs], b[sy + v.cs]);
  }
}

// MAC split of a point: along each axis the split at X (components not
// staggered along it) and at X + 1/2 (the one that is), in double, after
// clamping X into [vp, vp+vn-2]; local base indices and T fractions
template <typename T, int DIM>
__device__ __forceinline__ void mac_split(const KVel& v, const double Xin[DIM], int b0[3], int b1[3], T f0[3],
                                          T f1[3]) {
  const int lo[3] = {v.w.px, v.w.py, v.w.pz};
  const int nn[3] = {v.w.nx, v.w.ny, v.w.nz};
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const double X = clampd(Xin[a], (double)lo[a], (double)(lo[a] + nn[a] - 2));
    const double fl0 = floor(X), fl1 = floor(X + 0.5);
    b0[a] = (int)fl0 - lo[a];
    b1[a] = (int)fl1 - lo[a];
    f0[a] = (T)(X - fl0);
    f1[a] = (T)(X - fl1 + 0.5);
  }
}

// MAC-staggered trilinear velocity at an arbitrary point: InterpolateVelocity3D
// / 2D (Common/InterpolateVelocity.c:7-67).  The fraction split is done in
// double (mac_split); along each axis only two splits exist, so they are
// shared between the components.  Weights and the trilinear sum are in T.
template <typename T, int DIM>
__device__ __forceinline__ void mac_vel(const T* __restrict__ vel, const KVel& v,
                                        const double Xin[DIM], T V[DIM]) {
  int b0[3], b1[3];
  T f0[3], f1[3];
  mac_split<T, DIM>(v, Xin, b0, b1, f0, f1);
  const int sx = v.dof, sy = v.w.nx * v.dof, sz = v.w.ny * sy;
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    const int ix = c == 0 ? b1[0] : b0[0], iy = c == 1 ? b1[1] : b0[1];
    const T x1 = c == 0 ? f1[0] : f0[0], y1 = c == 1 ? f1[1] : f0[1];
    if (DIM == 3) {
      const int iz = c == 2 ? b1[2] : b0[2];
      const T z1 = c == 2 ? f1[2] : f0[2];
      const T* b = vel + iz * sz + iy * sy + ix * sx + v.udof + c * v.cs;
      V[c] = vtri(b[0], b[sx], b[sy], b[sy + sx], b[sz], b[sz + sx], b[sz + sy], b[sz + sy + sx], x1, y1, z1);
    } else {
      const T* b = vel + iy * sy + ix * sx + v.udof + c * v.cs;
      V[c] = vlerp(vlerp(b[0], b[sx], x1), vlerp(b[sy], b[sy + sx], x1), y1);
    }
  }
}

template <typename T>
__device__ __forceinline__ T linear_gather3(const T* b, int Y, int Z, T sx, T sy, T sz, T ox, T oy, T oz) {
  T sum = T(0);
  sum += b[0] * ox * oy * oz;
  sum += b[Z] * ox * oy * sz;
  sum += b[Y] * ox * sy * oz;
  sum += b[Y + Z] * ox * sy * sz;
  sum += b[1] * sx * oy * oz;
  sum += b[1 + Z] * sx * oy * sz;
  sum += b[1 + Y] * sx * sy * oz;
  sum += b[1 + Y + Z] * sx * sy * sz;
  return sum;
}

template <typename T>
__device__ __forceinline__ T trilinear(const T* __restrict__ phi, const KWin& w, const double S[3]) {
  double fx, fy, fz;
  const int xs = lin_split(S[0], w.px, w.nx, fx);
  const int ys = lin_split(S[1], w.py, w.ny, fy);
  const int zs = lin_split(S[2], w.pz, w.nz, fz);
  const int Y = w.nx, Z = w.nx * w.ny;
  return linear_gather3<T>(phi + zs * Z + ys * Y + xs, Y, Z, (T)fx, (T)fy, (T)fz, (T)(1.0 - fx),
                           (T)(1.0 - fy), (T)(1.0 - fz));
}

// Bilinear2D (Common/Grid.c:190-212): the (2x-1) form, x outer, y inner, /4
template <typename T>
__device__ __forceinline__ T bilinear_ref2d(const T* __restrict__ phi, const KWin& w, const double S[2]) {
  double fx, fy;
  const int xi = lin_split(S[0], w.px, w.nx, fx);
  const int yi = lin_split(S[1], w.py, w.ny, fy);
  const T xf = (T)(2. * fx - 1.), yf = (T)(2. * fy - 1.);
  const T xb0 = T(1) - xf, xb1 = T(1) + xf, yb0 = T(1) - yf, yb1 = T(1) + yf;
  const T* b = phi + yi * w.nx + xi;
  T sum = T(0);
  sum += b[0] * xb0 * yb0;
  sum += b[w.nx] * xb0 * yb1;
  sum += b[1] * xb1 * yb0;
  sum += b[1 + w.nx] * xb1 * yb1;
  return sum / T(4);
}

// GridInterpolate3D / 2D (Common/Grid.c:353-411), order k, j, i
template <typename T, int DIM>
__device__ __forceinline__ T grid_interp(const T* __restrict__ phi, const KWin& w, const double X[DIM]) {
  double fx, fy, fz = 0;
  const int xs = lin_split(X[0], w.px, w.nx, fx);
  const int ys = lin_split(X[1], w.py, w.ny, fy);
  const T wx1 = (T)fx, wx0 = (T)(1.0 - fx), wy1 = (T)fy, wy0 = (T)(1.0 - fy);
  const int Y = w.nx;
  if (DIM == 3) {
    const int zs = lin_split(X[2], w.pz, w.nz, fz);
    const T wz1 = (T)fz, wz0 = (T)(1.0 - fz);
    const int Z = w.nx * w.ny;
    const T* b = phi + zs * Z + ys * Y + xs;
    T v = T(0);
    v += wx0 * wy0 * wz0 * b[0];
    v += wx1 * wy0 * wz0 * b[1];
    v += wx0 * wy1 * wz0 * b[Y];
    v += wx1 * wy1 * wz0 * b[Y + 1];
    v += wx0 * wy0 * wz1 * b[Z];
    v += wx1 * wy0 * wz1 * b[Z + 1];
    v += wx0 * wy1 * wz1 * b[Z + Y];
    v += wx1 * wy1 * wz1 * b[Z + Y + 1];
    return v;
  } else {
    const T* b = phi + ys * Y + xs;
    T v = T(0);
    v += wx0 * wy0 * b[0];
    v += wx1 * wy0 * b[1];
    v += wx0 * wy1 * b[Y];
    v += wx1 * wy1 * b[Y + 1];
    return v;
  }
}

// Catmull-Rom tricubic sum over a 4x4x4 stencil whose (0,0,0) corner is at b
template <typename T>
__device__ __forceinline__ T cubic_gather3(const T* b, int Y, int Z, const T wx[4], const T wy[4],
                                           const T wz[4]) {
  return cubic_sum<T>(wx, wy, wz, [&](int j, int k, T r[4]) {
    const T* q = b + k * Z + j * Y;
    r[0] = q[0];
    r[1] = q[1];
    r[2] = q[2];
    r[3] = q[3];
  });
}

// Catmull-Rom tricubic gather over the 4x4x4 stencil around S
template <typename T>
__device__ __forceinline__ T tricubic(const T* __restrict__ phi, const KWin& w, const double S[3]) {
  double tx, ty, tz;
  const int bx = cub_split(S[0], w.px, w.nx, tx);
  const int by = cub_split(S[1], w.py, w.ny, ty);
  const int bz = cub_split(S[2], w.pz, w.nz, tz);
  T wx[4], wy[4], wz[4];
  cr_weights<T>((T)tx, wx);
  cr_weights<T>((T)ty, wy);
  cr_weights<T>((T)tz, wz);
  const int Y = w.nx, Z = w.nx * w.ny;
  return cubic_gather3<T>(phi + (bz - 1) * Z + (by - 1) * Y + (bx - 1), Y, Z, wx, wy, wz);
}

template <typename T>
__device__ __forceinline__ T bicubic(const T* __restrict__ phi, const KWin& w, const double S[2]) {
  double tx, ty;
  const int bx = cub_split(S[0], w.px, w.nx, tx);
  const int by = cub_split(S[1], w.py, w.ny, ty);
  T wx[4], wy[4];
  cr_weights<T>((T)tx, wx);
  cr_weights<T>((T)ty, wy);
  const int Y = w.nx;
  const T* b = phi + (by - 1) * Y + (bx - 1);
  T sum = T(0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const T* q = b + j * Y;
    T sx = T(0);
    sx = fma(wx[0], q[0], sx);
    sx = fma(wx[1], q[1], sx);
    sx = fma(wx[2], q[2], sx);
    sx = fma(wx[3], q[3], sx);
    sum = fma(wy[j], sx, sum);
  }
  return sum;
}

// Ordered-integer keys: max/min on the key is max/min on the float value, so
// float extrema can use the integer atomics (deterministic, order-free).
__device__ __forceinline__ unsigned long long okey(double v) {
  unsigned long long b = __double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double okey_val(unsigned long long k) {
  unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double(b);
}
__device__ __forceinline__ unsigned int okey(float v) {
  unsigned int b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float okey_val(unsigned int k) {
  unsigned int b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(b);
}

template <typename T> struct KeyOf;
template <> struct KeyOf<double> { typedef unsigned long long type; };
template <> struct KeyOf<float> { typedef unsigned int type; };

}  // namespace dcl
