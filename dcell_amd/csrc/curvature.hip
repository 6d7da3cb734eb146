// curvature.hip -- finite-difference curvature of phi on gfx950.
//
//   3D: GridFunction3D_Curv, 19 points (Common/Grid3D.c:83-108)
//   2D: GridFunction2D_Curv, 9 points (Common/Grid.c:231-254)
//   2D least squares: GridFunction2D_CurvLSS, 81 points
//       (LevelSetMethod/Curvature.c:3-186)
// kappa = 0 where the stencil leaves the window (and, for the first two, where
// the formula gives NaN, as in the reference).
#include <stdlib.h>
#include <string.h>

#include "dcl_host.hpp"
#include "dcl_tile.hpp"

namespace dcl {

struct CurvP {
  KWin w;
  double dx, dy, dz;
};

// dense 3D windows use the 2.5D LDS kernel; DCL_DENSE_KERNEL=gather selects
// the plain per-node kernel that windows of 2 GiB and more take (read at
// every call)
static bool use_tiled_curv() {
  const char* e = getenv("DCL_DENSE_KERNEL");
  return !(e && strcmp(e, "gather") == 0);
}

template <typename T, int DIM, int ST>
__device__ __forceinline__ T curv_node(const T* __restrict__ phi, const CurvP& P, int i, int j, int k) {
  const int m = ST == DCL_CURV_LSS ? 4 : 1;
  if (i < m || i > P.w.nx - 1 - m || j < m || j > P.w.ny - 1 - m) return T(0);
  if (DIM == 3 && (k < 1 || k > P.w.nz - 2)) return T(0);
  const int Y = P.w.nx, Z = P.w.nx * P.w.ny;
  const T* p = phi + (k * P.w.ny + j) * P.w.nx + i;
  if (DIM == 3) return curv3d_at<T>(p, Y, Z, curv_recips<T>(P.dx, P.dy, P.dz));
  if (ST == DCL_CURV_LSS) return curv2d_lss_at<T>(p, Y, (T)P.dx, (T)P.dy);
  return curv2d_at<T>(p, Y, (T)P.dx, (T)P.dy);
}

template <typename T, int DIM, int ST>
__global__ __launch_bounds__(256) void curv_dense(const T* __restrict__ phi, T* __restrict__ kappa,
                                                  CurvP P) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= P.w.nx || j >= P.w.ny) return;
  kappa[(k * P.w.ny + j) * P.w.nx + i] = curv_node<T, DIM, ST>(phi, P, i, j, k);
}

// Dense 3D curvature, 2.5D tiling: a block of 64 x 4 threads owns a
// 64 x 8 x 16 column of nodes (two rows per thread) and walks it with
// curv_column2 (a four-plane LDS ring of phi with a one-node halo, two planes
// of prefetch in flight, one barrier per plane).
constexpr int CTX = 64, CTY = 4, CRY = 2, CTZ = 16;
constexpr int CPF = ((CTX + 2) * (CTY * CRY + 2) + CTX * CTY - 1) / (CTX * CTY);

template <typename T>
__global__ __launch_bounds__(256) void curv_tiled3d(const T* __restrict__ phi, T* __restrict__ kappa,
                                                    CurvP P, int ntx, int nty, int ntz, unsigned pbytes) {
  __shared__ T pl[4 * CPF * CTX * CTY];
  const unsigned nb = (unsigned)(ntx * nty * ntz);
  const unsigned per = nb / 8u;
  const unsigned lb = (per == 0 || blockIdx.x >= per * 8u) ? blockIdx.x : (blockIdx.x % 8u) * per + blockIdx.x / 8u;
  const int tx = lb % ntx, ty = (lb / ntx) % nty, tz = lb / (ntx * nty);
  curv_column2<T, CTX, CTY, CRY, CTZ>(make_rsrc(phi, pbytes), kappa, P.w, curv_recips<T>(P.dx, P.dy, P.dz),
                                      tx * CTX, ty * CTY * CRY, tz * CTZ, pl);
}

template <typename T, int DIM, int ST, bool COMPACT>
__global__ __launch_bounds__(256) void curv_band(const T* __restrict__ phi, T* __restrict__ out,
                                                 CurvP P, const dcl_icoor* __restrict__ band,
                                                 int64_t n) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  const dcl_icoor q = band[b];
  const int i = q.x - P.w.px, j = q.y - P.w.py, k = DIM == 3 ? q.z - P.w.pz : 0;
  const bool inside = i >= 0 && i < P.w.nx && j >= 0 && j < P.w.ny && k >= 0 && k < P.w.nz;
  const T v = inside ? curv_node<T, DIM, ST>(phi, P, i, j, k) : T(0);
  if (COMPACT) out[b] = v;
  else if (inside) out[(k * P.w.ny + j) * P.w.nx + i] = v;
}

template <typename T, int DIM, int ST>
static int launch_curv(const void* phi, void* kappa, const CurvP& P, const dcl_icoor* band,
                       int64_t n, void* bv, hipStream_t s) {
  const T* p = static_cast<const T*>(phi);
  const long long pbytes = (long long)P.w.nx * P.w.ny * P.w.nz * (long long)sizeof(T);
  if (!band && DIM == 3 && use_tiled_curv() && buf_ok(pbytes, 0)) {
    const int ntx = (P.w.nx + CTX - 1) / CTX, nty = (P.w.ny + CTY * CRY - 1) / (CTY * CRY),
              ntz = (P.w.nz + CTZ - 1) / CTZ;
    hipLaunchKernelGGL((curv_tiled3d<T>), dim3(ntx * nty * ntz), dim3(CTX, CTY), 0, s, p,
                       static_cast<T*>(kappa), P, ntx, nty, ntz, (unsigned)pbytes);
    DCL_LAUNCHED("curv_tiled3d");
    return 0;
  }
  if (!band) {
    dim3 grid((P.w.nx + 63) / 64, (P.w.ny + 3) / 4, DIM == 3 ? P.w.nz : 1);
    hipLaunchKernelGGL((curv_dense<T, DIM, ST>), grid, dim3(64, 4), 0, s, p,
                       static_cast<T*>(kappa), P);
    DCL_LAUNCHED("curv_dense");
    return 0;
  }
  if (n == 0) return 0;
  if (bv)
    hipLaunchKernelGGL((curv_band<T, DIM, ST, true>), dim3(grid1d(n, 256)), dim3(256), 0, s, p,
                       static_cast<T*>(bv), P, band, n);
  else
    hipLaunchKernelGGL((curv_band<T, DIM, ST, false>), dim3(grid1d(n, 256)), dim3(256), 0, s, p,
                       static_cast<T*>(kappa), P, band, n);
  DCL_LAUNCHED("curv_band");
  return 0;
}

template <typename T>
static int dispatch_curv(int dim, int st, const void* phi, void* kappa, const CurvP& P,
                         const dcl_icoor* band, int64_t n, void* bv, hipStream_t s) {
  if (dim == 3) return launch_curv<T, 3, DCL_CURV_STD>(phi, kappa, P, band, n, bv, s);
  if (st == DCL_CURV_LSS) return launch_curv<T, 2, DCL_CURV_LSS>(phi, kappa, P, band, n, bv, s);
  return launch_curv<T, 2, DCL_CURV_STD>(phi, kappa, P, band, n, bv, s);
}

int curv_launch(const void* phi, void* kappa, const dcl_window* win, const dcl_icoor* band,
                int64_t n, void* bv, int precision, int stencil, hipStream_t s) {
  CurvP P;
  P.w = kwin(*win);
  P.dx = win->d.x;
  P.dy = win->d.y;
  P.dz = win->d.z;
  const int dim = win->n.z < 2 ? 2 : 3;
  if (precision == DCL_F64) return dispatch_curv<double>(dim, stencil, phi, kappa, P, band, n, bv, s);
  return dispatch_curv<float>(dim, stencil, phi, kappa, P, band, n, bv, s);
}

int check_curv_args(const void* phi, void* kappa, const dcl_window* win, const dcl_icoor* band,
                    int64_t n, void* bv, int precision, int stencil) {
  DCL_TRY(check_precision(precision));
  if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
  DCL_TRY(check_window(win, "phi", 1));
  if (stencil != DCL_CURV_STD && stencil != DCL_CURV_LSS)
    return set_error(DCL_ERR_ARG_WRONG, "stencil %d is neither DCL_CURV_STD nor DCL_CURV_LSS", stencil);
  if (stencil == DCL_CURV_LSS && win->n.z >= 2)
    return set_error(DCL_ERR_SUP, "the least-squares curvature stencil is 2D only");
  if (band && n < 0) return set_error(DCL_ERR_ARG_SIZ, "band_len < 0");
  if (!band || !bv) {
    if (!kappa) return set_error(DCL_ERR_ARG_NULL, "kappa is NULL");
    if (kappa == phi) return set_error(DCL_ERR_ARG_WRONG, "kappa and phi must be distinct buffers");
  }
  return 0;
}

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_curvature(const void* phi, void* kappa, const dcl_window* win,
                             const dcl_icoor* band, int64_t band_len, void* band_values,
                             int precision, int stencil, void* stream) {
  DCL_TRY(check_curv_args(phi, kappa, win, band, band_len, band_values, precision, stencil));
  return curv_launch(phi, kappa, win, band, band_len, band_values, precision, stencil,
                     as_stream(stream));
}
