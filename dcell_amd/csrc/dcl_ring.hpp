// dcl_ring.hpp -- dense 3D SL with phi0 staged in LDS.
//
// A block owns a 16 x 16 column of nodes and walks a chunk of ZC planes in z.
// phi0 over the column plus a halo of H = 3 nodes per side is held in LDS as
// a ring of 2H + 2 = 8 planes: at plane k the ring holds planes k-H .. k+H,
// and the plane k+H+1 is fetched into registers while plane k is computed,
// then stored into the slot of plane k-H-1, so one barrier per plane
// suffices.  Each staged value is loaded once per block (the column's halo
// makes it (16+6)^2 / 16^2 = 1.9 values per node, plus 2H planes per chunk),
// and the 64 (tricubic) or 8 (trilinear) stencil reads of a node come from
// LDS.  H = 3 holds every stencil of a departure less than 2 cells from its
// node along each axis (stencil base-1 .. base+2, base = floor(S)); a node
// whose stencil leaves the ring (a faster flow, or the window's clamps)
// gathers from global memory instead.  Velocity is read with buffer loads.
#pragma once

#include "dcl_buf.hpp"

namespace dcl {

constexpr int RG_TX = 16, RG_TY = 16, RG_H = 3, RG_R = 8;  // RG_R = 2 RG_H + 2 ring slots
constexpr int RG_BX = RG_TX + 2 * RG_H, RG_BY = RG_TY + 2 * RG_H;
constexpr int RG_PLANE = RG_BX * RG_BY;
constexpr int RG_NT = RG_TX * RG_TY;
constexpr int RG_PF = (RG_PLANE + RG_NT - 1) / RG_NT;  // staged values per thread and plane

// the box's values of plane z (x, y clamped to the window) into registers
template <typename T>
__device__ __forceinline__ void ring_fetch(rsrc_t rp, const KWin& w, int x0, int y0, int z, int tid,
                                           T pf[RG_PF]) {
#pragma unroll
  for (int q = 0; q < RG_PF; ++q) {
    const int e = tid + q * RG_NT;
    if (e < RG_PLANE) {
      const int bx = e % RG_BX, by = e / RG_BX;
      const int gx = clampi(x0 - RG_H + bx, 0, w.nx - 1), gy = clampi(y0 - RG_H + by, 0, w.ny - 1);
      pf[q] = BL<T>::one(rp, ((z * w.ny + gy) * w.nx + gx) * (int)sizeof(T), 0);
    }
  }
}

template <typename T>
__device__ __forceinline__ void ring_store(T* ring, int slot, int tid, const T pf[RG_PF]) {
#pragma unroll
  for (int q = 0; q < RG_PF; ++q) {
    const int e = tid + q * RG_NT;
    if (e < RG_PLANE) ring[slot * RG_PLANE + e] = pf[q];
  }
}

// gather of phi0 at S for the node in plane k of the column at (x0, y0); the
// ring slot of plane z is (z - z0 + H) mod 8
template <typename T, int INTERP>
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
    if ((unsigned)lx <= (unsigned)(RG_BX - 4) && (unsigned)ly <= (unsigned)(RG_BY - 4) &&
        (unsigned)lz <= (unsigned)(2 * RG_H - 3)) {
      const T* b = ring + ly * RG_BX + lx;
      const int zs = bz - 1 - z0 + RG_H;
      return cubic_sum<T>(wx, wy, wz, [&](int j, int kk, T r[4]) {
        const T* q = b + ((zs + kk) & (RG_R - 1)) * RG_PLANE + j * RG_BX;
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
    if ((unsigned)lx <= (unsigned)(RG_BX - 2) && (unsigned)ly <= (unsigned)(RG_BY - 2) &&
        (unsigned)lz <= (unsigned)(2 * RG_H - 1)) {
      const int zs = bz - z0 + RG_H;
      const T* p0 = ring + (zs & (RG_R - 1)) * RG_PLANE + ly * RG_BX + lx;
      const T* p1 = ring + ((zs + 1) & (RG_R - 1)) * RG_PLANE + ly * RG_BX + lx;
      a00[0] = p0[0];
      a00[1] = p0[1];
      a10[0] = p0[RG_BX];
      a10[1] = p0[RG_BX + 1];
      a01[0] = p1[0];
      a01[1] = p1[1];
      a11[0] = p1[RG_BX];
      a11[1] = p1[RG_BX + 1];
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

// the SL step of one column: nodes (x0 .. x0+15, y0 .. y0+15, z0 .. z0+ZC-1)
// of the window, clipped to it; nodes outside [2, n-3] copy phi0.  Every
// thread of the block calls it (it holds barriers) and it ends on a barrier,
// so a block may run it for one column after another with the same ring.
template <typename T, int INTERP, int BT, int ZC>
__device__ __forceinline__ void ring_column(const T* __restrict__ phi0, T* __restrict__ phi, rsrc_t rp,
                                            rsrc_t rv, const SLP& P, int x0, int y0, int z0, T* ring) {
  const int nx = P.w.nx, ny = P.w.ny, nz = P.w.nz;
  const int z1 = min(nz, z0 + ZC);
  const int tid = threadIdx.x + RG_TX * threadIdx.y;
  for (int z = max(z0 - RG_H, 0); z <= min(z0 + RG_H, nz - 1); ++z) {
    T pf[RG_PF];
    ring_fetch<T>(rp, P.w, x0, y0, z, tid, pf);
    ring_store<T>(ring, (z - z0 + RG_H) & (RG_R - 1), tid, pf);
  }
  __syncthreads();
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  const bool in_w = i < nx && j < ny;
  const bool in_xy = i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3;
  for (int k = z0; k < z1; ++k) {
    const int zn = k + RG_H + 1;
    const bool pre = zn < nz && k + 1 < z1;
    T pf[RG_PF];
    if (pre) ring_fetch<T>(rp, P.w, x0, y0, zn, tid, pf);
    if (in_w) {
      const int id = (k * ny + j) * nx + i;
      if (!(in_xy && k >= 2 && k <= nz - 3)) {
        phi[id] = phi0[id];
      } else {
        double S[3];
        sl_dep3_buf<T, BT>(rv, P, P.w.px + i, P.w.py + j, P.w.pz + k, S);
        phi[id] = gather3_ring<T, INTERP>(rp, ring, P.w, S, x0, y0, z0, k);
      }
    }
    if (pre) ring_store<T>(ring, (zn - z0 + RG_H) & (RG_R - 1), tid, pf);
    __syncthreads();
  }
}

// grid: tiles_x * tiles_y * nzc blocks of RG_TX x RG_TY threads (XCD-remapped
// so each XCD walks neighbouring columns)
template <typename T, int INTERP, int BT, int ZC, int WPE>
__global__ __launch_bounds__(RG_NT) __attribute__((amdgpu_waves_per_eu(WPE))) void sl_ring3d(const T* __restrict__ phi0, T* __restrict__ phi,
                                                   const T* __restrict__ vel, SLP P, unsigned tiles_x,
                                                   unsigned tiles_y, unsigned nzc, unsigned pbytes,
                                                   unsigned vbytes) {
  __shared__ T ring[RG_R * RG_PLANE];
  const unsigned nb = tiles_x * tiles_y * nzc;
  unsigned lb = blockIdx.x;
  {
    const unsigned per = nb / 8u;
    if (per != 0 && lb < per * 8u) lb = (lb % 8u) * per + lb / 8u;
  }
  const unsigned tx = lb % tiles_x, ty = (lb / tiles_x) % tiles_y, zc = lb / (tiles_x * tiles_y);
  ring_column<T, INTERP, BT, ZC>(phi0, phi, make_rsrc(phi0, pbytes), make_rsrc(vel, vbytes), P,
                                 (int)tx * RG_TX, (int)ty * RG_TY, (int)zc * ZC, ring);
}

}  // namespace dcl
