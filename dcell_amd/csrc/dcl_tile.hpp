// dcl_tile.hpp -- LDS-tiled SL and curvature for one tile of a 3D window,
// shared by the dense kernels and the persistent batched kernels.
//
// SL: the tile is TX x TY x TZ nodes, worked by TX x TY threads that each
// walk TZ nodes in z.  phi0 over the tile plus a halo of H nodes per side is
// copied into LDS with coalesced loads (H = 4 for the Catmull-Rom stencil,
// departures up to 2 cells; H = 3 for trilinear), then every stencil read of
// a node comes from LDS; a departure whose stencil leaves the staged box
// falls back to global loads, so any velocity is handled.  Velocity is read
// from global memory.  Nodes outside [2, n-3] copy phi0.
//
// Curvature: the tile plus a one-node halo of the new phi is copied into LDS
// and the 19-point stencil reads come from there.
#pragma once

#include "dcl_curv.hpp"
#include "dcl_sl.hpp"

namespace dcl {

template <int INTERP>
struct Halo {
  static constexpr int H = INTERP == DCL_CUBIC ? 4 : 3;
};

template <typename T, int INTERP, int TX, int TY, int TZ>
struct SLBox {
  static constexpr int H = Halo<INTERP>::H;
  static constexpr int BX = TX + 2 * H, BY = TY + 2 * H, BZ = TZ + 2 * H;
  static constexpr int SIZE = BX * BY * BZ;
};

// one SL tile with origin (x0, y0, z0) in local window coordinates; all
// threads of the block call it (it synchronizes once after staging)
template <typename T, int INTERP, int BT, int TX, int TY, int TZ>
__device__ __forceinline__ void sl_tile(const T* __restrict__ phi0, T* __restrict__ phi,
                                        const T* __restrict__ vel, const SLP& P, int x0, int y0, int z0,
                                        T* box) {
  typedef SLBox<T, INTERP, TX, TY, TZ> C;
  const int nx = P.w.nx, ny = P.w.ny, nz = P.w.nz;
  const int ox = x0 - C::H, oy = y0 - C::H, oz = z0 - C::H;
  const int Y = nx, Z = nx * ny;
  const int tid = threadIdx.x + TX * threadIdx.y;
  for (int e = tid; e < C::SIZE; e += TX * TY) {
    const int bx = e % C::BX, r = e / C::BX;
    const int by = r % C::BY, bz = r / C::BY;
    const int gx = clampi(ox + bx, 0, nx - 1), gy = clampi(oy + by, 0, ny - 1), gz = clampi(oz + bz, 0, nz - 1);
    box[e] = phi0[gz * Z + gy * Y + gx];
  }
  __syncthreads();
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const bool in_xy = i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3;
  for (int kz = 0; kz < TZ; ++kz) {
    const int k = z0 + kz;
    if (k >= nz) break;
    const int id = k * Z + j * Y + i;
    if (!(in_xy && k >= 2 && k <= nz - 3)) {
      phi[id] = box[((k - oz) * C::BY + (j - oy)) * C::BX + (i - ox)];
      continue;
    }
    double S[3];
    sl_departure<T, 3, BT>(vel, P, P.w.px + i, P.w.py + j, P.w.pz + k, S);
    T v;
    if (INTERP == DCL_CUBIC) {
      double fx, fy, fz;
      const int bx = cub_split(S[0], P.w.px, nx, fx) - 1;
      const int by = cub_split(S[1], P.w.py, ny, fy) - 1;
      const int bz = cub_split(S[2], P.w.pz, nz, fz) - 1;
      T wx[4], wy[4], wz[4];
      cr_weights<T>((T)fx, wx);
      cr_weights<T>((T)fy, wy);
      cr_weights<T>((T)fz, wz);
      const int lx = bx - ox, ly = by - oy, lz = bz - oz;
      if (lx >= 0 && lx + 3 < C::BX && ly >= 0 && ly + 3 < C::BY && lz >= 0 && lz + 3 < C::BZ)
        v = cubic_gather3<T>(box + (lz * C::BY + ly) * C::BX + lx, C::BX, C::BX * C::BY, wx, wy, wz);
      else
        v = cubic_gather3<T>(phi0 + bz * Z + by * Y + bx, Y, Z, wx, wy, wz);
    } else {
      double fx, fy, fz;
      const int bx = lin_split(S[0], P.w.px, nx, fx);
      const int by = lin_split(S[1], P.w.py, ny, fy);
      const int bz = lin_split(S[2], P.w.pz, nz, fz);
      const T sx = (T)fx, sy = (T)fy, sz = (T)fz;
      const T qx = (T)(1.0 - fx), qy = (T)(1.0 - fy), qz = (T)(1.0 - fz);
      const int lx = bx - ox, ly = by - oy, lz = bz - oz;
      if (lx >= 0 && lx + 1 < C::BX && ly >= 0 && ly + 1 < C::BY && lz >= 0 && lz + 1 < C::BZ)
        v = linear_gather3<T>(box + (lz * C::BY + ly) * C::BX + lx, C::BX, C::BX * C::BY, sx, sy, sz, qx,
                              qy, qz);
      else
        v = linear_gather3<T>(phi0 + bz * Z + by * Y + bx, Y, Z, sx, sy, sz, qx, qy, qz);
    }
    phi[id] = v;
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
