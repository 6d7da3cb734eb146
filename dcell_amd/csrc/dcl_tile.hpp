// dcl_tile.hpp -- SL and curvature for one tile of a 3D window, used by the
// persistent batched kernels.
//
// SL: the tile is TX x TY x TZ nodes, worked by TX x TY threads that each
// walk TZ nodes in z with the buffer-load node update (dcl_buf.hpp); nodes
// outside [2, n-3] copy phi0.  (Staging phi0 in LDS was measured slower: the
// gathers hit L1/L2 well and the update is bound by its instruction count.)
//
// Curvature: a rolling four-plane LDS ring of the new phi with a one-node
// halo and two planes of prefetch; the 19-point stencil reads come from
// there (curv_column2, also the dense 3D curvature kernel).
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

// Curvature of one TX x (TY RY) x TZ column worked by TX x TY threads with
// RY node rows each (rows ty, ty + TY, ...): planes k-1, k, k+1 of phi with a
// one-node halo sit in a four-slot LDS ring (pl, 4 PF TX TY values, PF =
// ceil((TX+2)(TY RY+2) / (TX TY))) while planes k+2 and k+3 are in flight
// into registers, two planes of prefetch per block.  Lanes past a plane's
// values load a clamped in-window value into the slot's padding, so the
// loads and LDS stores of the steady state are unconditional and the wait
// for plane k+2 leaves plane k+3's loads in flight.  kappa is 0 where the
// stencil leaves the window; the arithmetic is curv3d_planes (dcl_curv.hpp).
// Every thread of the block calls it, and it ends on a barrier.
template <typename T, int TX, int TY, int RY, int TZ>
__device__ __forceinline__ void curv_column2(rsrc_t rp, T* __restrict__ kappa, const KWin& w, const CurvR<T>& r,
                                             int x0, int y0, int z0, T* pl) {
  constexpr int PX = TX + 2, PY = TY * RY + 2, PLANE = PX * PY, NT = TX * TY, PF = (PLANE + NT - 1) / NT;
  constexpr int SLOT = PF * NT;
  const int nx = w.nx, ny = w.ny, nz = w.nz;
  const int tid = threadIdx.x + TX * threadIdx.y;
  const unsigned pz = (unsigned)(nx * ny) * (unsigned)sizeof(T);
  int xo[PF];
#pragma unroll
  for (int q = 0; q < PF; ++q) {
    const int e = min(tid + q * NT, PLANE - 1);
    const int bx = e % PX, by = e / PX;
    const int gx = clampi(x0 - 1 + bx, 0, nx - 1), gy = clampi(y0 - 1 + by, 0, ny - 1);
    xo[q] = (gy * nx + gx) * (int)sizeof(T);
  }
  auto fetch = [&](int k, T (&v)[PF]) {
    const int zo = (int)((unsigned)clampi(k, 0, nz - 1) * pz);
#pragma unroll
    for (int q = 0; q < PF; ++q) v[q] = BL<T>::one(rp, xo[q], zo);
  };
  auto store = [&](int slot, const T (&v)[PF]) {
#pragma unroll
    for (int q = 0; q < PF; ++q) pl[slot * SLOT + tid + q * NT] = v[q];
  };
  const int i = x0 + threadIdx.x;
  const bool in_x = i >= 1 && i <= nx - 2;
  auto compute = [&](int kz) {
    const int k = z0 + kz;
    const T* s0 = pl + (kz & 3) * SLOT;
    const T* s1 = pl + ((kz + 1) & 3) * SLOT;
    const T* s2 = pl + ((kz + 2) & 3) * SLOT;
#pragma unroll
    for (int ry = 0; ry < RY; ++ry) {
      const int j = y0 + threadIdx.y + ry * TY;
      if (i < nx && j < ny) {
        const int lo = (threadIdx.y + ry * TY + 1) * PX + threadIdx.x + 1;
        T c = T(0);
        if (in_x && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2)
          c = curv3d_planes<T>(s0 + lo, s1 + lo, s2 + lo, PX, r);
        // FP64 kappa is stored nontemporally: measured right after a kernel
        // that wrote the window (the SL step), 0.069 against 0.072 ms at
        // 256^3; in FP32 plain stores are as fast or faster
        T* dst = kappa + ((size_t)k * ny + j) * nx + i;
        if (sizeof(T) == 8)
          __builtin_nontemporal_store(c, dst);
        else
          *dst = c;
      }
    }
  };
  T va[PF], vb[PF], vc[PF];
  fetch(z0 - 1, va);
  fetch(z0, vb);
  fetch(z0 + 1, vc);
  store(0, va);
  store(1, vb);
  store(2, vc);
  fetch(z0 + 2, va);  // plane k+2 of iteration 0
  __syncthreads();
  const int kmax = min(TZ, nz - z0);
  int kz = 0;
  // steady state: both iterations of a pair fetch plane k+3 and store plane k+2
  for (; kz + 3 < kmax; kz += 2) {
    fetch(z0 + kz + 3, vb);
    compute(kz);
    store((kz + 3) & 3, va);
    __syncthreads();
    fetch(z0 + kz + 4, va);
    compute(kz + 1);
    store((kz + 4) & 3, vb);
    __syncthreads();
  }
  // the last one to three planes (va holds plane z0 + kz + 2)
  for (; kz < kmax; ++kz) {
    if (kz + 2 < kmax) fetch(z0 + kz + 3, vb);
    compute(kz);
    if (kz + 1 < kmax) store((kz + 3) & 3, va);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PF; ++q) va[q] = vb[q];
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
