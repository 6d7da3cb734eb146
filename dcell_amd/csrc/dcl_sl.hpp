// dcl_sl.hpp -- the per-node semi-Lagrangian update shared by the dense,
// band and batched kernels: MAC velocity at the node, Euler or RK2 departure
// point in double, trilinear / Bilinear2D or Catmull-Rom gather of phi0.
#pragma once

#include "dcl_common.hpp"

namespace dcl {

struct SLP {
  KWin w;
  KVel v;
  double c[3];   // dt / dh per axis
  double ch[3];  // dt / (2 dh) per axis
};

// departure point of node (X, Y, Z): S = X - V(X) dt/dh (Euler) or
// S = X - V(Xm) dt/dh with Xm = X - V(X) dt/(2 dh) (RK2)
template <typename T, int DIM, int BT>
__device__ __forceinline__ void sl_departure(const T* __restrict__ vel, const SLP& P, int X, int Y, int Z,
                                             double S[3]) {
  T V[DIM];
  node_vel<T, DIM>(vel, P.v, X, Y, Z, V);
  const double Xd[3] = {(double)X, (double)Y, (double)Z};
  S[2] = Xd[2];
  if (BT == DCL_RK2) {
    double Xm[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < DIM; ++a) Xm[a] = Xd[a] - (double)V[a] * P.ch[a];
    T V2[DIM];
    mac_vel<T, DIM>(vel, P.v, Xm, V2);
#pragma unroll
    for (int a = 0; a < DIM; ++a) S[a] = Xd[a] - (double)V2[a] * P.c[a];
  } else {
#pragma unroll
    for (int a = 0; a < DIM; ++a) S[a] = Xd[a] - (double)V[a] * P.c[a];
  }
}

template <typename T, int DIM, int INTERP, int BT>
__device__ __forceinline__ T sl_point(const T* __restrict__ phi0, const T* __restrict__ vel,
                                      const SLP& P, int X, int Y, int Z) {
  double S[3];
  sl_departure<T, DIM, BT>(vel, P, X, Y, Z, S);
  if (INTERP == DCL_CUBIC) return DIM == 3 ? tricubic<T>(phi0, P.w, S) : bicubic<T>(phi0, P.w, S);
  return DIM == 3 ? trilinear<T>(phi0, P.w, S) : bilinear_ref2d<T>(phi0, P.w, S);
}

}  // namespace dcl
