// sl_tiled.hip -- dense 3D SL step with phi0 staged in LDS (dcl_tile.hpp):
// one block per 32 x 8 x TZ tile (TZ = 4 in FP64, 8 in FP32, sized so the
// staged box fits 64 KB of LDS), blocks numbered so each XCD walks a
// contiguous run of tiles.
#include "dcl_host.hpp"
#include "dcl_tile.hpp"

namespace dcl {

template <typename T>
struct DenseTile {
  static constexpr int TX = 32, TY = 8, TZ = sizeof(T) == 8 ? 4 : 8;
};

template <typename T, int INTERP, int BT>
__global__ __launch_bounds__(256) void sl_tiled3d(const T* __restrict__ phi0, T* __restrict__ phi,
                                                  const T* __restrict__ vel, SLP P, int ntx, int nty,
                                                  int ntz) {
  typedef DenseTile<T> D;
  __shared__ T box[SLBox<T, INTERP, D::TX, D::TY, D::TZ>::SIZE];
  const unsigned lb = xcd_tile(blockIdx.x, (unsigned)(ntx * nty * ntz));
  const int tx = lb % ntx, ty = (lb / ntx) % nty, tz = lb / (ntx * nty);
  sl_tile<T, INTERP, BT, D::TX, D::TY, D::TZ>(phi0, phi, vel, P, tx * D::TX, ty * D::TY, tz * D::TZ, box);
}

template <typename T, int INTERP, int BT>
static int launch_tiled(const void* phi0, void* phi, const void* vel, const SLP& P, hipStream_t s) {
  typedef DenseTile<T> D;
  const int ntx = (P.w.nx + D::TX - 1) / D::TX, nty = (P.w.ny + D::TY - 1) / D::TY;
  const int ntz = (P.w.nz + D::TZ - 1) / D::TZ;
  hipLaunchKernelGGL((sl_tiled3d<T, INTERP, BT>), dim3(ntx * nty * ntz), dim3(D::TX, D::TY), 0, s,
                     static_cast<const T*>(phi0), static_cast<T*>(phi), static_cast<const T*>(vel), P,
                     ntx, nty, ntz);
  DCL_LAUNCHED("sl_tiled3d");
  return 0;
}

template <typename T>
static int dispatch_tiled(int interp, int bt, const void* phi0, void* phi, const void* vel, const SLP& P,
                          hipStream_t s) {
  if (interp == DCL_CUBIC && bt == DCL_RK2) return launch_tiled<T, DCL_CUBIC, DCL_RK2>(phi0, phi, vel, P, s);
  if (interp == DCL_CUBIC) return launch_tiled<T, DCL_CUBIC, DCL_EULER>(phi0, phi, vel, P, s);
  if (bt == DCL_RK2) return launch_tiled<T, DCL_LINEAR, DCL_RK2>(phi0, phi, vel, P, s);
  return launch_tiled<T, DCL_LINEAR, DCL_EULER>(phi0, phi, vel, P, s);
}

int sl_tiled_launch(const void* phi0, void* phi, const void* vel, const SLP& P, int precision, int interp,
                    int bt, hipStream_t s) {
  if (precision == DCL_F64) return dispatch_tiled<double>(interp, bt, phi0, phi, vel, P, s);
  return dispatch_tiled<float>(interp, bt, phi0, phi, vel, P, s);
}

}  // namespace dcl
