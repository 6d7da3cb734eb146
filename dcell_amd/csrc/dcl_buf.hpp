// dcl_buf.hpp -- the 3D SL node update with buffer loads.
//
// Every gather of the update (the node velocity, the RK2 midpoint's 3 x 8
// MAC corners, the 4 x 4 x 4 Catmull-Rom stencil or the 2 x 2 x 2 trilinear
// one) is a set of loads at one per-lane base offset plus offsets that are
// the same for every lane (corner and row strides).  Buffer loads take the
// per-lane part as a 32-bit VGPR offset and the uniform part as an SGPR
// offset, so no per-load 64-bit address arithmetic is spent, and each
// stencil row of 4 consecutive values is one 16-byte load (two in FP64).
// Buffer descriptors also bound every access to the window's bytes.
#pragma once

#include "dcl_sl.hpp"

namespace dcl {

typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// raw (untyped) buffer of `bytes` bytes starting at p
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// The loaded dwords are reinterpreted as a whole vector.  (hipcc 7.2 loads
// only the first dword when __builtin_bit_cast is applied to a single element
// of the returned vector, so that form is not used.)
template <typename T>
struct BL;

template <>
struct BL<float> {
  static __device__ __forceinline__ float one(rsrc_t r, int v, int s) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
  }
  static __device__ __forceinline__ void two(rsrc_t r, int v, int s, float o[2]) {
    const f32x2 x = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
    o[0] = x.x;
    o[1] = x.y;
  }
  static __device__ __forceinline__ void four(rsrc_t r, int v, int s, float o[4]) {
    const f32x4 x = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
    o[0] = x.x;
    o[1] = x.y;
    o[2] = x.z;
    o[3] = x.w;
  }
};

template <>
struct BL<double> {
  static __device__ __forceinline__ double one(rsrc_t r, int v, int s) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
  }
  static __device__ __forceinline__ void two(rsrc_t r, int v, int s, double o[2]) {
    const f64x2 x = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
    o[0] = x.x;
    o[1] = x.y;
  }
  static __device__ __forceinline__ void four(rsrc_t r, int v, int s, double o[4]) {
    const f64x2 x = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
    const f64x2 y = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, v + 16, s, 0));
    o[0] = x.x;
    o[1] = x.y;
    o[2] = y.x;
    o[3] = y.y;
  }
};

// node velocity (see node_vel): the two faces around the node per component.
// With 3 components per node, (u, v, w) of the node and u of its +x
// neighbour are one 16-byte load in FP32, two in FP64 (the texture path's
// cost is per load instruction, not per byte); with component planes, u of
// the node and of its +x neighbour are one load.
template <typename T>
__device__ __forceinline__ void node_vel_buf(rsrc_t rv, const KVel& v, int X, int Y, int Z, T V[3]) {
  const int x = clampi(X, v.w.px, v.w.px + v.w.nx - 2) - v.w.px;
  const int y = clampi(Y, v.w.py, v.w.py + v.w.ny - 2) - v.w.py;
  const int z = clampi(Z, v.w.pz, v.w.pz + v.w.nz - 2) - v.w.pz;
  const int sx = v.dof, sy = v.w.nx * v.dof, sz = v.w.ny * sy;
  const int e = (int)sizeof(T);
  const int vo = (((z * v.w.ny + y) * v.w.nx + x) * v.dof + v.udof) * e;
  if (v.dof == 3) {
    T q[4];
    BL<T>::four(rv, vo, 0, q);
    V[0] = vmid(q[0], q[3]);
    V[1] = vmid(q[1], BL<T>::one(rv, vo + e, sy * e));
    V[2] = vmid(q[2], BL<T>::one(rv, vo + 2 * e, sz * e));
    return;
  }
  const int ce = v.cs * e;  // bytes between components
  if (v.dof == 1) {
    T p[2];
    BL<T>::two(rv, vo, 0, p);
    V[0] = vmid(p[0], p[1]);
  } else {
    V[0] = vmid(BL<T>::one(rv, vo, 0), BL<T>::one(rv, vo, sx * e));
  }
  V[1] = vmid(BL<T>::one(rv, vo + ce, 0), BL<T>::one(rv, vo + ce, sy * e));
  V[2] = vmid(BL<T>::one(rv, vo + 2 * ce, 0), BL<T>::one(rv, vo + 2 * ce, sz * e));
}

// MAC trilinear velocity of the split (b0, b1, f0, f1) (see mac_vel).  With 3
// FP32 components per node, each x pair of faces of one component is the
// first and last value of one 16-byte load; with component planes it is two
// neighbouring values, one load (16 bytes in FP64).
template <typename T>
__device__ __forceinline__ void mac_vel_buf_split(rsrc_t rv, const KVel& v, const int b0[3], const int b1[3],
                                                  const T f0[3], const T f1[3], T V[3]) {
  const int sx = v.dof, sy = v.w.nx * v.dof, sz = v.w.ny * sy;
  const int e = (int)sizeof(T);
  const bool quad = sizeof(T) == 4 && v.dof == 3;
  const bool planes = v.dof == 1;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int ix = c == 0 ? b1[0] : b0[0], iy = c == 1 ? b1[1] : b0[1], iz = c == 2 ? b1[2] : b0[2];
    const T x1 = c == 0 ? f1[0] : f0[0], y1 = c == 1 ? f1[1] : f0[1], z1 = c == 2 ? f1[2] : f0[2];
    const int vo = (iz * sz + iy * sy + ix * sx + v.udof + c * v.cs) * e;
    T a000, a100, a010, a110, a001, a101, a011, a111;
    if (planes) {
      T p[2];
      BL<T>::two(rv, vo, 0, p);
      a000 = p[0], a100 = p[1];
      BL<T>::two(rv, vo, sy * e, p);
      a010 = p[0], a110 = p[1];
      BL<T>::two(rv, vo, sz * e, p);
      a001 = p[0], a101 = p[1];
      BL<T>::two(rv, vo, (sz + sy) * e, p);
      a011 = p[0], a111 = p[1];
    } else if (quad) {
      T q[4];
      BL<T>::four(rv, vo, 0, q);
      a000 = q[0], a100 = q[3];
      BL<T>::four(rv, vo, sy * e, q);
      a010 = q[0], a110 = q[3];
      BL<T>::four(rv, vo, sz * e, q);
      a001 = q[0], a101 = q[3];
      BL<T>::four(rv, vo, (sz + sy) * e, q);
      a011 = q[0], a111 = q[3];
    } else {
      a000 = BL<T>::one(rv, vo, 0), a100 = BL<T>::one(rv, vo, sx * e);
      a010 = BL<T>::one(rv, vo, sy * e), a110 = BL<T>::one(rv, vo, (sy + sx) * e);
      a001 = BL<T>::one(rv, vo, sz * e), a101 = BL<T>::one(rv, vo, (sz + sx) * e);
      a011 = BL<T>::one(rv, vo, (sz + sy) * e), a111 = BL<T>::one(rv, vo, (sz + sy + sx) * e);
    }
    V[c] = vtri(a000, a100, a010, a110, a001, a101, a011, a111, x1, y1, z1);
  }
}

template <typename T>
__device__ __forceinline__ void mac_vel_buf(rsrc_t rv, const KVel& v, const double Xin[3], T V[3]) {
  int b0[3], b1[3];
  T f0[3], f1[3];
  mac_split<T, 3>(v, Xin, b0, b1, f0, f1);
  mac_vel_buf_split<T>(rv, v, b0, b1, f0, f1, V);
}

// node velocity (see node_vel) with component planes for a thread that
// walks +z through the nodes of one (X, Y): the w face above the node at one
// plane is the face below it at the next.  carry (uniform): wc holds this
// node's lower w face, loaded as the upper one at the previous plane; wc is
// set to this node's upper face.
template <typename T>
__device__ __forceinline__ void node_vel_planes_z(rsrc_t rv, const KVel& v, int X, int Y, int Z, bool carry, T& wc,
                                                  T V[3]) {
  const int x = clampi(X, v.w.px, v.w.px + v.w.nx - 2) - v.w.px;
  const int y = clampi(Y, v.w.py, v.w.py + v.w.ny - 2) - v.w.py;
  const int z = clampi(Z, v.w.pz, v.w.pz + v.w.nz - 2) - v.w.pz;
  const int e = (int)sizeof(T);
  const int sy = v.w.nx, sz = v.w.ny * sy;
  const int vo = ((z * v.w.ny + y) * v.w.nx + x) * e;
  const int ce = v.cs * e;
  T p[2];
  BL<T>::two(rv, vo, 0, p);
  V[0] = vmid(p[0], p[1]);
  V[1] = vmid(BL<T>::one(rv, vo + ce, 0), BL<T>::one(rv, vo + ce, sy * e));
  T lo;
  if (carry)
    lo = wc;
  else
    lo = BL<T>::one(rv, vo + 2 * ce, 0);
  wc = BL<T>::one(rv, vo + 2 * ce, sz * e);
  V[2] = vmid(lo, wc);
}

// node_vel_planes_z split for a thread that loads the next plane's node
// faces while it works the current one: the faces of node (X, Y, Z) other
// than its lower w face (the u pair, the v pair, the upper w face), and the
// lower w face alone
template <typename T>
struct NodeFaces {
  T u0, u1, v0, v1, wh;
};
template <typename T>
__device__ __forceinline__ int node_face_offset(const KVel& v, int X, int Y, int Z) {
  const int x = clampi(X, v.w.px, v.w.px + v.w.nx - 2) - v.w.px;
  const int y = clampi(Y, v.w.py, v.w.py + v.w.ny - 2) - v.w.py;
  const int z = clampi(Z, v.w.pz, v.w.pz + v.w.nz - 2) - v.w.pz;
  return ((z * v.w.ny + y) * v.w.nx + x) * (int)sizeof(T);
}
template <typename T>
__device__ __forceinline__ NodeFaces<T> node_faces_planes(rsrc_t rv, const KVel& v, int X, int Y, int Z) {
  const int e = (int)sizeof(T);
  const int vo = node_face_offset<T>(v, X, Y, Z), ce = v.cs * e;
  NodeFaces<T> f;
  T p[2];
  BL<T>::two(rv, vo, 0, p);
  f.u0 = p[0];
  f.u1 = p[1];
  f.v0 = BL<T>::one(rv, vo + ce, 0);
  f.v1 = BL<T>::one(rv, vo + ce, v.w.nx * e);
  f.wh = BL<T>::one(rv, vo + 2 * ce, v.w.nx * v.w.ny * e);
  return f;
}
template <typename T>
__device__ __forceinline__ T node_wlow_planes(rsrc_t rv, const KVel& v, int X, int Y, int Z) {
  return BL<T>::one(rv, node_face_offset<T>(v, X, Y, Z) + 2 * v.cs * (int)sizeof(T), 0);
}

// departure point of node (X, Y, Z) (see sl_departure) from its node
// velocity V, the RK2 midpoint velocity read with buffer loads
template <typename T, int BT>
__device__ __forceinline__ void sl_dep3_from(rsrc_t rv, const SLP& P, int X, int Y, int Z, const T V[3],
                                             double S[3]) {
  const double Xd[3] = {(double)X, (double)Y, (double)Z};
  if (BT == DCL_RK2) {
    double Xm[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) Xm[a] = Xd[a] - (double)V[a] * P.ch[a];
    T V2[3];
    mac_vel_buf<T>(rv, P.v, Xm, V2);
#pragma unroll
    for (int a = 0; a < 3; ++a) S[a] = Xd[a] - (double)V2[a] * P.c[a];
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) S[a] = Xd[a] - (double)V[a] * P.c[a];
  }
}

// departure point of node (X, Y, Z) (see sl_departure) with buffer loads
template <typename T, int BT>
__device__ __forceinline__ void sl_dep3_buf(rsrc_t rv, const SLP& P, int X, int Y, int Z, double S[3]) {
  T V[3];
  node_vel_buf<T>(rv, P.v, X, Y, Z, V);
  sl_dep3_from<T, BT>(rv, P, X, Y, Z, V, S);
}

// Bilinear3D (Common/Grid3D.c:17-28) sum in the order i, j, k over the
// [y][z] rows of (x, x+1)
template <typename T>
__device__ __forceinline__ T linear_sum(const T a00[2], const T a01[2], const T a10[2], const T a11[2], T sx,
                                        T sy, T sz, T ox, T oy, T oz) {
  T sum = T(0);
  sum += a00[0] * ox * oy * oz;
  sum += a01[0] * ox * oy * sz;
  sum += a10[0] * ox * sy * oz;
  sum += a11[0] * ox * sy * sz;
  sum += a00[1] * sx * oy * oz;
  sum += a01[1] * sx * oy * sz;
  sum += a10[1] * sx * sy * oz;
  sum += a11[1] * sx * sy * sz;
  return sum;
}

// gather of phi0 at the departure point S with buffer loads
template <typename T, int INTERP>
__device__ __forceinline__ T gather3_buf(rsrc_t rp, const KWin& w, const double S[3]) {
  const int Yr = w.nx, Zr = w.nx * w.ny;
  const int e = (int)sizeof(T);
  double fx, fy, fz;
  if (INTERP == DCL_CUBIC) {
    const int bx = cub_split(S[0], w.px, w.nx, fx);
    const int by = cub_split(S[1], w.py, w.ny, fy);
    const int bz = cub_split(S[2], w.pz, w.nz, fz);
    T wx[4], wy[4], wz[4];
    cr_weights<T>((T)fx, wx);
    cr_weights<T>((T)fy, wy);
    cr_weights<T>((T)fz, wz);
    const int vo = ((bz - 1) * Zr + (by - 1) * Yr + (bx - 1)) * e;
    return cubic_sum<T>(wx, wy, wz, [&](int j, int k, T r[4]) { BL<T>::four(rp, vo, (k * Zr + j * Yr) * e, r); });
  } else {
    const int bx = lin_split(S[0], w.px, w.nx, fx);
    const int by = lin_split(S[1], w.py, w.ny, fy);
    const int bz = lin_split(S[2], w.pz, w.nz, fz);
    const int vo = (bz * Zr + by * Yr + bx) * e;
    T a00[2], a01[2], a10[2], a11[2];
    BL<T>::two(rp, vo, 0, a00);
    BL<T>::two(rp, vo, Zr * e, a01);
    BL<T>::two(rp, vo, Yr * e, a10);
    BL<T>::two(rp, vo, (Yr + Zr) * e, a11);
    return linear_sum<T>(a00, a01, a10, a11, (T)fx, (T)fy, (T)fz, (T)(1.0 - fx), (T)(1.0 - fy), (T)(1.0 - fz));
  }
}

// the full 3D node update (sl_point<T, 3, INTERP, BT>) with buffer loads
template <typename T, int INTERP, int BT>
__device__ __forceinline__ T sl_point3_buf(rsrc_t rp, rsrc_t rv, const SLP& P, int X, int Y, int Z) {
  double S[3];
  sl_dep3_buf<T, BT>(rv, P, X, Y, Z, S);
  return gather3_buf<T, INTERP>(rp, P.w, S);
}

// windows the buffer-load kernels take: their per-lane byte offsets are
// signed 32-bit values, so every buffer stays below 2 GiB (larger windows
// take the pointer-based kernels)
__host__ inline bool buf_ok(long long phi_bytes, long long vel_bytes) {
  return phi_bytes < 0x80000000ll && vel_bytes < 0x80000000ll;
}

}  // namespace dcl
