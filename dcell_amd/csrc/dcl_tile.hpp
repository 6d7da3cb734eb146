// dcl_tile.hpp -- SL and curvature for one tile of a 3D window, used by the
// persistent batched kernels.
//
// SL: the tile is TX x TY x TZ nodes, worked by TX x TY threads that each
// walk TZ nodes in z with the buffer-load node update (dcl_buf.hpp); nodes
// outside [2, n-3] copy phi0.  (Staging phi0 in LDS was measured slower: the
// gathers hit L1/L2 well and the update is bound by its instruction count.)
//
// Curvature: the tile plus a one-node halo of the new phi is copied into LDS
// and the 19-point stencil reads come from there.
#pragma once

#include "dcl_buf.hpp"
#include "dcl_curv.hpp"

namespace dcl {

// one SL tile with origin (x0, y0, z0) in local window coordinates: each of
// the TX x TY threads walks TZ nodes in z, gathering with buffer loads
template <typename T, int INTERP, int BT, int TX, int TY, int TZ>
__device__ __forceinline__ void sl_tile(rsrc_t rp, rsrc_t rv, const T* __restrict__ phi0, T* __restrict__ phi,
                                        const SLP& P, int x0, int y0, int z0) {
  const int nx = P.w.nx, ny = P.w.ny, nz = P.w.nz;
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const bool in_xy = i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3;
  for (int kz = 0; kz < TZ; ++kz) {
    const int k = z0 + kz;
    if (k >= nz) break;
    const int id = (k * ny + j) * nx + i;
    if (!(in_xy && k >= 2 && k <= nz - 3)) phi[id] = phi0[id];
    else phi[id] = sl_point3_buf<T, INTERP, BT>(rp, rv, P, P.w.px + i, P.w.py + j, P.w.pz + k);
  }
}

template <typename T, int TX, int TY, int TZ>
struct CurvBox {
  static constexpr int BX = TX + 2, BY = TY + 2, BZ = TZ + 2;
  static constexpr int SIZE = BX * BY * BZ;
};

// one curvature tile: kappa at the tile's nodes, 0 where the stencil leaves the window
template <typename T, int TX, int TY, int TZ>
__device__ __forceinline__ void curv_tile(const T* __restrict__ phi, T* __restrict__ kappa, const KWin& w,
                                          const CurvR<T>& r, int x0, int y0, int z0, T* box) {
  typedef CurvBox<T, TX, TY, TZ> C;
  const int nx = w.nx, ny = w.ny, nz = w.nz;
  const int Y = nx, Z = nx * ny;
  const int tid = threadIdx.x + TX * threadIdx.y;
  for (int e = tid; e < C::SIZE; e += TX * TY) {
    const int bx = e % C::BX, q = e / C::BX;
    const int by = q % C::BY, bz = q / C::BY;
    const int gx = clampi(x0 - 1 + bx, 0, nx - 1), gy = clampi(y0 - 1 + by, 0, ny - 1);
    const int gz = clampi(z0 - 1 + bz, 0, nz - 1);
    box[e] = phi[gz * Z + gy * Y + gx];
  }
  __syncthreads();
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const bool in_xy = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
  for (int kz = 0; kz < TZ; ++kz) {
    const int k = z0 + kz;
    if (k >= nz) break;
    T c = T(0);
    if (in_xy && k >= 1 && k <= nz - 2)
      c = curv3d_at<T>(box + ((kz + 1) * C::BY + threadIdx.y + 1) * C::BX + threadIdx.x + 1, C::BX,
                       C::BX * C::BY, r);
    kappa[k * Z + j * Y + i] = c;
  }
}

// hardware block id -> tile id such that each XCD (ids are dealt round-robin
// over the 8 XCDs) walks a contiguous run of tiles
__device__ __forceinline__ unsigned xcd_tile(unsigned hw, unsigned nblocks) {
  const unsigned per = nblocks / 8u;
  if (per == 0 || hw >= per * 8u) return hw;
  return (hw % 8u) * per + hw / 8u;
}

}  // namespace dcl
