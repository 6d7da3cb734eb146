// dcl_ring.hpp -- dense 3D SL with phi0 staged in LDS.
//
// A block owns a TX x TY column of nodes (TX TY = 256 threads, one node
// each) and walks a chunk of ZC planes in z.
// phi0 over the column plus a halo of H = 3 nodes per side is held in LDS as
// a ring of 2H + 2 = 8 planes: at plane k the ring holds planes k-H .. k+H,
// and the plane k+H+1 is fetched into registers while plane k is computed,
// then stored into the slot of plane k-H-1, so one barrier per plane
// suffices.  Each staged value is loaded once per block (the column's halo
// makes it (TX+6)(TY+6) / (TX TY) values per node, 1.9 for 16 x 16 and 2.7
// for 64 x 4, plus 2H planes per chunk),
// and the 64 (tricubic) or 8 (trilinear) stencil reads of a node come from
// LDS.  H = 3 holds every stencil of a departure less than 2 cells from its
// node along each axis (stencil base-1 .. base+2, base = floor(S)); a node
// whose stencil leaves the ring (a faster flow, or the window's clamps)
// gathers from global memory instead.  Velocity is read with buffer loads.
#pragma once

#include "dcl_buf.hpp"

namespace dcl {

constexpr int RG_H = 3, RG_R = 8;  // halo, ring slots (2 RG_H + 2)

// geometry of a TX x TY column: the box of BX x BY values per plane sits in
// LDS with row pitch RP and plane (slot) pitch PP, by default packed
template <int TX_, int TY_, int RP_ = TX_ + 2 * RG_H, int PP_ = RP_ * (TY_ + 2 * RG_H), int VRP_ = TX_ + 3,
          int VPP_ = VRP_ * (TY_ + 3)>
struct RingGeom {
  static constexpr int TX = TX_, TY = TY_;
  static constexpr int BX = TX + 2 * RG_H, BY = TY + 2 * RG_H;
  static constexpr int PLANE = BX * BY;  // staged values per plane
  static constexpr int RP = RP_, PP = PP_;
  static constexpr int VRP = VRP_, VPP = VPP_;  // the velocity ring's pitches
  static_assert(RP >= BX && PP >= RP * BY, "ring pitches too small for the box");
  static constexpr int NT = TX * TY;
  static constexpr int PF = (PLANE + NT - 1) / NT;  // staged values per thread and plane
};

// The geometry the kernels use for a TX x TY column of T: the FP32 32 x 8
// column pads the phi0 ring's plane pitch to 544 values and the velocity
// ring's to 392 (from 532 and 385).  Where the departures vary across a wave
// (velocity fields with several periods per cell, as in batches of small
// cells) the lanes of one LDS read sit in different planes and rows, and the
// unpadded pitches put many of them on the same banks; padded, the kernel is
// 4-10% faster on such fields and as fast on smooth ones, and four blocks of
// 40.9 KB still fit a CU's 160 KB.
template <typename T, int TX, int TY>
struct RingFor {
  typedef RingGeom<TX, TY> type;
};
template <>
struct RingFor<float, 32, 8> {
  typedef RingGeom<32, 8, 38, 544, 35, 392> type;
};

// byte offsets within a plane of the box values this thread stages (x, y
// clamped to the window), -1 past the box, and their LDS positions within a
// slot; the same for every plane
template <typename T, typename G>
__device__ __forceinline__ void ring_offsets(const KWin& w, int x0, int y0, int tid, int xo[G::PF],
                                             int so[G::PF]) {
#pragma unroll
  for (int q = 0; q < G::PF; ++q) {
    const int e = tid + q * G::NT;
    const int bx = e % G::BX, by = e / G::BX;
    const int gx = clampi(x0 - RG_H + bx, 0, w.nx - 1), gy = clampi(y0 - RG_H + by, 0, w.ny - 1);
    xo[q] = e < G::PLANE ? (gy * w.nx + gx) * (int)sizeof(T) : -1;
    so[q] = G::RP == G::BX ? e : by * G::RP + bx;
  }
}

// the box's values of plane z into registers; zo = z nx ny sizeof(T) (uniform)
template <typename T, typename G>
__device__ __forceinline__ void ring_fetch(rsrc_t rp, const int xo[G::PF], int zo, T pf[G::PF]) {
#pragma unroll
  for (int q = 0; q < G::PF; ++q)
    if (xo[q] >= 0) pf[q] = BL<T>::one(rp, xo[q], zo);
}

template <typename T, typename G>
__device__ __forceinline__ void ring_store(T* ring, int slot, int tid, const int so[G::PF], const T pf[G::PF]) {
#pragma unroll
  for (int q = 0; q < G::PF; ++q) {
    const int e = tid + q * G::NT;
    if (e < G::PLANE) ring[slot * G::PP + so[q]] = pf[q];
  }
}

// gather of phi0 at S for the node in plane k of the column at (x0, y0); the
// ring slot of plane z is (z - z0 + H) mod 8
template <typename T, int INTERP, typename G>
__device__ __forceinline__ T gather3_ring(rsrc_t rp, const T* ring, const KWin& w, const double S[3], int x0,
                                          int y0, int z0, int k) {
  double fx, fy, fz;
  if (INTERP == DCL_CUBIC) {
    const int bx = cub_split(S[0], w.px, w.nx, fx);
    const int by = cub_split(S[1], w.py, w.ny, fy);
    const int bz = cub_split(S[2], w.pz, w.nz, fz);
    T wx[4], wy[4], wz[4];
    cr_weights<T>((T)fx, wx);
    cr_weights<T>((T)fy, wy);
    cr_weights<T>((T)fz, wz);
    const int lx = bx - 1 - (x0 - RG_H), ly = by - 1 - (y0 - RG_H), lz = bz - 1 - (k - RG_H);
    if ((unsigned)lx <= (unsigned)(G::BX - 4) && (unsigned)ly <= (unsigned)(G::BY - 4) &&
        (unsigned)lz <= (unsigned)(2 * RG_H - 3)) {
      const T* b = ring + ly * G::RP + lx;
      const int zs = bz - 1 - z0 + RG_H;
      return cubic_sum<T>(wx, wy, wz, [&](int j, int kk, T r[4]) {
        const T* q = b + ((zs + kk) & (RG_R - 1)) * G::PP + j * G::RP;
        r[0] = q[0];
        r[1] = q[1];
        r[2] = q[2];
        r[3] = q[3];
      });
    }
    const int Yr = w.nx, Zr = w.nx * w.ny, e = (int)sizeof(T);
    const int vo = ((bz - 1) * Zr + (by - 1) * Yr + (bx - 1)) * e;
    return cubic_sum<T>(wx, wy, wz, [&](int j, int kk, T r[4]) { BL<T>::four(rp, vo, (kk * Zr + j * Yr) * e, r); });
  } else {
    const int bx = lin_split(S[0], w.px, w.nx, fx);
    const int by = lin_split(S[1], w.py, w.ny, fy);
    const int bz = lin_split(S[2], w.pz, w.nz, fz);
    const T sx = (T)fx, sy = (T)fy, sz = (T)fz;
    const T ox = (T)(1.0 - fx), oy = (T)(1.0 - fy), oz = (T)(1.0 - fz);
    const int lx = bx - (x0 - RG_H), ly = by - (y0 - RG_H), lz = bz - (k - RG_H);
    T a00[2], a01[2], a10[2], a11[2];
    if ((unsigned)lx <= (unsigned)(G::BX - 2) && (unsigned)ly <= (unsigned)(G::BY - 2) &&
        (unsigned)lz <= (unsigned)(2 * RG_H - 1)) {
      const int zs = bz - z0 + RG_H;
      const T* p0 = ring + (zs & (RG_R - 1)) * G::PP + ly * G::RP + lx;
      const T* p1 = ring + ((zs + 1) & (RG_R - 1)) * G::PP + ly * G::RP + lx;
      a00[0] = p0[0];
      a00[1] = p0[1];
      a10[0] = p0[G::RP];
      a10[1] = p0[G::RP + 1];
      a01[0] = p1[0];
      a01[1] = p1[1];
      a11[0] = p1[G::RP];
      a11[1] = p1[G::RP + 1];
    } else {
      const int Yr = w.nx, Zr = w.nx * w.ny, e = (int)sizeof(T);
      const int vo = (bz * Zr + by * Yr + bx) * e;
      BL<T>::two(rp, vo, 0, a00);
      BL<T>::two(rp, vo, Zr * e, a01);
      BL<T>::two(rp, vo, Yr * e, a10);
      BL<T>::two(rp, vo, (Yr + Zr) * e, a11);
    }
    return linear_sum<T>(a00, a01, a10, a11, sx, sy, sz, ox, oy, oz);
  }
}

// Velocity ring (optional): with the velocity window equal to phi's and 3
// components per node, the faces over the column plus 1 node below and 2
// above in x and y are staged too, one component after the other (so the x
// pair of faces of a MAC corner is two neighbouring words), for the planes
// k-1 .. k+2 that the node velocity and the RK2 midpoint (less than one cell
// away) read, plus the plane being prefetched: VR_R = 5 slots, plane z in
// slot (z - z0 + 1) mod 5.  A lane whose midpoint corners leave the box reads
// them from memory.
constexpr int VR_R = 5;

// the velocity box: VX x VY nodes per plane, each component's values in LDS
// with row pitch RP and plane pitch PP
template <int TX_, int TY_, int RP_ = TX_ + 3, int PP_ = RP_ * (TY_ + 3)>
struct VelGeom {
  static constexpr int VX = TX_ + 3, VY = TY_ + 3;
  static constexpr int PLANE = VX * VY;
  static constexpr int RP = RP_, PP = PP_;
  static_assert(RP >= VX && PP >= RP * VY, "velocity ring pitches too small for the box");
  static constexpr int NT = TX_ * TY_;
  static constexpr int PF = (PLANE + NT - 1) / NT;
};

template <typename T, typename V>
__device__ __forceinline__ void vring_offsets(const KWin& w, int x0, int y0, int tid, int xo[V::PF],
                                              int so[V::PF]) {
#pragma unroll
  for (int q = 0; q < V::PF; ++q) {
    const int e = tid + q * V::NT;
    const int bx = e % V::VX, by = e / V::VX;
    const int gx = clampi(x0 - 1 + bx, 0, w.nx - 1), gy = clampi(y0 - 1 + by, 0, w.ny - 1);
    xo[q] = e < V::PLANE ? (gy * w.nx + gx) * 3 * (int)sizeof(T) : -1;
    so[q] = V::RP == V::VX ? e : by * V::RP + bx;
  }
}

// the faces (u, v, w) of the box nodes of plane z; zo = 3 z nx ny sizeof(T)
template <typename T, typename V>
__device__ __forceinline__ void vring_fetch(rsrc_t rv, const int xo[V::PF], int zo, T pf[V::PF][3]) {
#pragma unroll
  for (int q = 0; q < V::PF; ++q) {
    if (xo[q] >= 0) {
      if (sizeof(T) == 4) {
        typedef float f32x3 __attribute__((ext_vector_type(3)));
        const f32x3 x = __builtin_bit_cast(f32x3, __builtin_amdgcn_raw_buffer_load_b96(rv, xo[q], zo, 0));
        pf[q][0] = (T)x.x;
        pf[q][1] = (T)x.y;
        pf[q][2] = (T)x.z;
      } else {
        T a[2];
        BL<T>::two(rv, xo[q], zo, a);
        pf[q][0] = a[0];
        pf[q][1] = a[1];
        pf[q][2] = BL<T>::one(rv, xo[q] + 2 * (int)sizeof(T), zo);
      }
    }
  }
}

template <typename T, typename V>
__device__ __forceinline__ void vring_store(T* vr, int slot, int tid, const int so[V::PF], const T pf[V::PF][3]) {
#pragma unroll
  for (int q = 0; q < V::PF; ++q) {
    const int e = tid + q * V::NT;
    if (e < V::PLANE) {
#pragma unroll
      for (int c = 0; c < 3; ++c) vr[(slot * 3 + c) * V::PP + so[q]] = pf[q][c];
    }
  }
}

// the departure point of interior node (i, j, k) (local) with the velocity
// read from the ring; sl[m] is the slot of plane k-1+m
template <typename T, int BT, typename V>
__device__ __forceinline__ void sl_dep3_ring(rsrc_t rv, const T* vr, const int sl[4], const SLP& P, int i, int j,
                                             int k, int x0, int y0, double S[3]) {
  // node velocity: faces at (i, i+1), (j, j+1), (k, k+1); interior nodes need no clamp
  const int bx = i - (x0 - 1), by = j - (y0 - 1);
  const T* p0 = vr + sl[1] * 3 * V::PP + by * V::RP + bx;
  const T* p1 = vr + sl[2] * 3 * V::PP + by * V::RP + bx;
  T Vn[3];
  Vn[0] = vmid(p0[0], p0[1]);
  Vn[1] = vmid(p0[V::PP], p0[V::PP + V::RP]);
  Vn[2] = vmid(p0[2 * V::PP], p1[2 * V::PP]);
  const int X = P.w.px + i, Y = P.w.py + j, Z = P.w.pz + k;
  const double Xd[3] = {(double)X, (double)Y, (double)Z};
  if (BT != DCL_RK2) {
#pragma unroll
    for (int a = 0; a < 3; ++a) S[a] = Xd[a] - (double)Vn[a] * P.c[a];
    return;
  }
  double Xm[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) Xm[a] = Xd[a] - (double)Vn[a] * P.ch[a];
  int b0[3], b1[3];
  T f0[3], f1[3];
  mac_split<T, 3>(P.v, Xm, b0, b1, f0, f1);
  T V2[3];
  const int lo[3] = {x0 - 1, y0 - 1, k - 1};
  const int hi[3] = {x0 - 1 + V::VX - 1, y0 - 1 + V::VY - 1, k + 2};
  bool in = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) in = in && b0[a] >= lo[a] && b1[a] + 1 <= hi[a];
  if (in) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int ix = c == 0 ? b1[0] : b0[0], iy = c == 1 ? b1[1] : b0[1], iz = c == 2 ? b1[2] : b0[2];
      const T x1 = c == 0 ? f1[0] : f0[0], y1 = c == 1 ? f1[1] : f0[1], z1 = c == 2 ? f1[2] : f0[2];
      const int rel = iz - (k - 1);  // 0 .. 2
      const int sa = rel == 0 ? sl[0] : rel == 1 ? sl[1] : sl[2];
      const int sb = rel == 0 ? sl[1] : rel == 1 ? sl[2] : sl[3];
      const int o = (iy - lo[1]) * V::RP + (ix - lo[0]) + c * V::PP;
      const T* qa = vr + sa * 3 * V::PP + o;
      const T* qb = vr + sb * 3 * V::PP + o;
      V2[c] = vtri(qa[0], qa[1], qa[V::RP], qa[V::RP + 1], qb[0], qb[1], qb[V::RP], qb[V::RP + 1], x1, y1, z1);
    }
  } else {
    mac_vel_buf_split<T>(rv, P.v, b0, b1, f0, f1, V2);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) S[a] = Xd[a] - (double)V2[a] * P.c[a];
}

// a phi value of the column's output: FP64 nontemporally (the 256^3 step
// with the curvature after it measured 1% faster, the FP32 kernels as much
// slower)
template <typename T>
__device__ __forceinline__ void ring_put(T* p, T v) {
  if constexpr (sizeof(T) == 8)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// the SL step of one column: nodes (x0 .. x0+TX-1, y0 .. y0+TY-1, z0 ..
// z0+zc-1) of the window, clipped to it; nodes outside [2, n-3] copy phi0.
// Every thread of the block calls it (it holds barriers) and it ends on a
// barrier, so a block may run it for one column after another with the same
// ring (RG_R * G::PP values).
template <typename T, int INTERP, int BT, typename G, bool VL>
__device__ __forceinline__ void ring_column(const T* __restrict__ phi0, T* __restrict__ phi, rsrc_t rp,
                                            rsrc_t rv, const SLP& P, int x0, int y0, int z0, int zc, T* ring,
                                            T* vr) {
  typedef VelGeom<G::TX, G::TY, G::VRP, G::VPP> V;
  const int nx = P.w.nx, ny = P.w.ny, nz = P.w.nz;
  const int z1 = min(nz, z0 + zc);
  const int tid = threadIdx.x + G::TX * threadIdx.y;
  // bytes per plane of phi0 (a third of velocity's); products are taken
  // unsigned, the buffer offsets being 32-bit (windows below 2 GiB)
  const unsigned pz = (unsigned)(nx * ny) * (unsigned)sizeof(T);
  int xo[G::PF], so[G::PF], vxo[V::PF], vso[V::PF];
  ring_offsets<T, G>(P.w, x0, y0, tid, xo, so);
  if (VL) vring_offsets<T, V>(P.w, x0, y0, tid, vxo, vso);
  for (int z = max(z0 - RG_H, 0); z <= min(z0 + RG_H, nz - 1); ++z) {
    T pf[G::PF];
    ring_fetch<T, G>(rp, xo, (int)((unsigned)z * pz), pf);
    ring_store<T, G>(ring, (z - z0 + RG_H) & (RG_R - 1), tid, so, pf);
  }
  if (VL) {
    for (int z = max(z0 - 1, 0); z <= min(z0 + 2, nz - 1); ++z) {
      T vf[V::PF][3];
      vring_fetch<T, V>(rv, vxo, (int)(3u * (unsigned)z * pz), vf);
      vring_store<T, V>(vr, (z - z0 + 1) % VR_R, tid, vso, vf);
    }
  }
  __syncthreads();
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  const bool in_w = i < nx && j < ny;
  const bool in_xy = i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3;
  // component planes: the node's upper w face carries to the next plane
  // (node_vel_planes_z) when the previous plane was computed here and the
  // velocity window's clamp did not hold z there.  In FP64 the next plane's
  // node faces are loaded while this plane is worked (nf), so a plane waits
  // on one round trip to memory (the RK2 midpoint's corners) instead of
  // two: 0.327 -> 0.320 ms at 256^3 (in FP32 it was slower)
  const bool planes = P.v.dof == 1;
  T wc = T(0);
  NodeFaces<T> nf = {T(0), T(0), T(0), T(0), T(0)};
  for (int k = z0; k < z1; ++k) {
    const int zn = k + RG_H + 1;
    const bool pre = zn < nz && k + 1 < z1;
    T pf[G::PF];
    if (pre) ring_fetch<T, G>(rp, xo, (int)((unsigned)zn * pz), pf);
    const bool vpre = VL && k + 3 < nz && k + 1 < z1;
    T vf[V::PF][3];
    if (vpre) vring_fetch<T, V>(rv, vxo, (int)(3u * (unsigned)(k + 3) * pz), vf);
    if (in_w) {
      const int id = (k * ny + j) * nx + i;
      if (!(in_xy && k >= 2 && k <= nz - 3)) {
        ring_put<T>(phi + id, phi0[id]);
      } else {
        double S[3];
        if (VL) {
          const int b = (k - z0) % VR_R;
          const int sl[4] = {b, (b + 1) % VR_R, (b + 2) % VR_R, (b + 3) % VR_R};
          sl_dep3_ring<T, BT, V>(rv, vr, sl, P, i, j, k, x0, y0, S);
        } else if (planes) {
          const int vz0 = P.v.w.pz, vz1 = P.v.w.pz + P.v.w.nz - 2;
          const bool pf = k - 1 >= max(z0, 2);  // the previous plane was computed here
          const bool carry = pf && clampi(P.w.pz + k, vz0, vz1) == clampi(P.w.pz + k - 1, vz0, vz1) + 1;
          const int X = P.w.px + i, Y = P.w.py + j, Z = P.w.pz + k;
          T V[3];
          if constexpr (sizeof(T) == 8) {
            NodeFaces<T> f;
            if (pf)
              f = nf;
            else
              f = node_faces_planes<T>(rv, P.v, X, Y, Z);
            T lo;
            if (carry)
              lo = wc;
            else
              lo = node_wlow_planes<T>(rv, P.v, X, Y, Z);
            V[0] = vmid(f.u0, f.u1);
            V[1] = vmid(f.v0, f.v1);
            V[2] = vmid(lo, f.wh);
            wc = f.wh;
          } else {
            node_vel_planes_z<T>(rv, P.v, X, Y, Z, carry, wc, V);
          }
          sl_dep3_from<T, BT>(rv, P, X, Y, Z, V, S);
          if constexpr (sizeof(T) == 8) {
            if (k + 1 < z1 && k + 1 <= nz - 3) nf = node_faces_planes<T>(rv, P.v, X, Y, Z + 1);
          }
        } else {
          sl_dep3_buf<T, BT>(rv, P, P.w.px + i, P.w.py + j, P.w.pz + k, S);
        }
        ring_put<T>(phi + id, gather3_ring<T, INTERP, G>(rp, ring, P.w, S, x0, y0, z0, k));
      }
    }
    if (pre) ring_store<T, G>(ring, (zn - z0 + RG_H) & (RG_R - 1), tid, so, pf);
    if (vpre) vring_store<T, V>(vr, (k + 3 - z0 + 1) % VR_R, tid, vso, vf);
    __syncthreads();
  }
}

// Waves per SIMD the ring kernels are compiled for: the FP32 velocity-ring
// kernel's LDS (40 KB per 256-thread block) admits 4 blocks per CU, and
// asking for 4 waves per SIMD brings its registers from 136 to 104 (no
// spills), 7% faster at 384^3; the others keep the compiler's choice.
template <typename T, bool VL>
constexpr int ring_waves() {
  return sizeof(T) == 4 && VL ? 4 : 1;
}

// grid: tiles_x * tiles_y * nzc blocks of TX x TY threads (XCD-remapped so
// each XCD walks neighbouring columns), each walking zch planes.  VL:
// velocity staged too (velocity window = phi window, 3 components, udof 0).
template <typename T, int INTERP, int BT, int TX, int TY, bool VL>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(ring_waves<T, VL>(), 8))) void sl_ring3d(
    const T* __restrict__ phi0, T* __restrict__ phi, const T* __restrict__ vel, SLP P, unsigned tiles_x,
    unsigned tiles_y, unsigned nzc, int zch, unsigned pbytes, unsigned vbytes) {
  typedef typename RingFor<T, TX, TY>::type G;
  typedef VelGeom<TX, TY, G::VRP, G::VPP> V;
  __shared__ T ring[RG_R * G::PP];
  __shared__ T vr[VL ? VR_R * 3 * V::PP : 1];
  const unsigned nb = tiles_x * tiles_y * nzc;
  unsigned lb = blockIdx.x;
  {
    const unsigned per = nb / 8u;
    if (per != 0 && lb < per * 8u) lb = (lb % 8u) * per + lb / 8u;
  }
  const unsigned tx = lb % tiles_x, ty = (lb / tiles_x) % tiles_y, zc = lb / (tiles_x * tiles_y);
  ring_column<T, INTERP, BT, G, VL>(phi0, phi, make_rsrc(phi0, pbytes), make_rsrc(vel, vbytes), P, (int)tx * TX,
                                    (int)ty * TY, (int)zc * zch, zch, ring, vr);
}

}  // namespace dcl
