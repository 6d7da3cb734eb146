// reduce.hip -- reductions and stream compaction on gfx950.
//
//   dcl_max_velocity      LevelSet_MaxVelocity (LevelSetMethod/Advect.c:70-111)
//   dcl_band_compact      |phi| < width in raster order: per-block wavefront
//                         ballots counted, offsets scanned, then each lane's
//                         slot from mbcnt over its wave's ballot
//   dcl_interface_metrics delta-weighted area and curvature with
//                         LevelSetDiracDelta3D / 2D (NormalDirection.c:36-57)
//   dcl_error_norms       max / L2 differences
// Sums are reduced in a fixed order (block partials, then one block), so the
// results do not depend on scheduling.
#include <mutex>

#include "dcl_host.hpp"

namespace dcl {

constexpr int RB = 256;      // threads per reduction block
constexpr int RGRID = 2048;  // blocks of the grid-stride reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  return v;
}

// reduce NV values per thread over a 256-thread block; kinds: 0 sum, 1 max
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const int (&kind)[NV], double* out) {
  __shared__ double sh[NV][RB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    v[k] = kind[k] ? wave_max(v[k]) : wave_sum(v[k]);
    if (lane == 0) sh[k][wv] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double a = sh[k][0];
      for (int i = 1; i < RB / 64; ++i) a = kind[k] ? (sh[k][i] > a ? sh[k][i] : a) : a + sh[k][i];
      out[k] = a;
    }
  }
}

// one block folds nb partial rows of NV values in a fixed order
template <int NV>
__global__ __launch_bounds__(RB) void fold_partials_k(const double* __restrict__ part, int nb,
                                                      double* __restrict__ out, int k0, int k1,
                                                      int k2, int k3) {
  const int kinds[4] = {k0, k1, k2, k3};
  double v[NV];
  int kind[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    kind[k] = kinds[k];
    v[k] = 0.0;
  }
  for (int i = threadIdx.x; i < nb; i += RB)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double x = part[i * NV + k];
      v[k] = kind[k] ? (x > v[k] ? x : v[k]) : v[k] + x;
    }
  block_reduce<NV>(v, kind, out);
}

// ------------------------------------------------------------ max |V|
template <typename T, int DIM>
__global__ __launch_bounds__(RB) void maxvel_k(const T* __restrict__ vel, KVel v,
                                               const char* __restrict__ nodes, int64_t n,
                                               int64_t stride, double* __restrict__ part) {
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    const int* q = reinterpret_cast<const int*>(nodes + i * stride);
    T V[DIM];
    node_vel<T, DIM>(vel, v, q[0], q[1], DIM == 3 ? q[2] : v.w.pz, V);
    T s = V[0] * V[0] + V[1] * V[1];
    if (DIM == 3) s += V[2] * V[2];
    const double mag = (double)sqrt(s);
    m = mag > m ? mag : m;  // the reference's "mag > maxVel ? mag : maxVel"
  }
  double vals[1] = {m};
  const int kind[1] = {1};
  block_reduce<1>(vals, kind, part + blockIdx.x);
}

// ------------------------------------------------------------ band compaction
template <typename T>
__global__ __launch_bounds__(RB) void band_count_k(const T* __restrict__ phi, int64_t n, T width,
                                                   unsigned* __restrict__ counts) {
  __shared__ unsigned sh[RB / 64];
  const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
  const bool pred = i < n && fabs(phi[i]) < width;
  const unsigned long long bal = __ballot(pred);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = (unsigned)__popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

template <typename T>
__global__ __launch_bounds__(RB) void band_write_k(const T* __restrict__ phi, KWin w, int64_t n, T width,
                                                   const unsigned* __restrict__ offsets,
                                                   const unsigned long long* __restrict__ tile_sums,
                                                   dcl_icoor* __restrict__ out, int64_t cap) {
  __shared__ unsigned sh[RB / 64];
  const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
  const bool pred = i < n && fabs(phi[i]) < width;
  const unsigned long long bal = __ballot(pred);
  const unsigned lane_pre = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[wv] = (unsigned)__popcll(bal);
  __syncthreads();
  unsigned wave_pre = 0;
  for (int k = 0; k < wv; ++k) wave_pre += sh[k];
  if (!pred) return;
  const int64_t pos = (int64_t)(offsets[blockIdx.x] + tile_sums[blockIdx.x / SCAN_TILE]) + wave_pre + lane_pre;
  if (pos >= cap) return;
  const int64_t plane = (int64_t)w.nx * w.ny;
  const int z = (int)(i / plane), r = (int)(i - (int64_t)z * plane);
  const int y = r / w.nx, x = r - y * w.nx;
  dcl_icoor q;
  q.x = w.px + x;
  q.y = w.py + y;
  q.z = w.pz + z;
  out[pos] = q;
}

// ------------------------------------------------------------ interface metrics
template <typename T, int DIM>
__global__ __launch_bounds__(RB) void metrics_k(const T* __restrict__ phi, const T* __restrict__ kappa,
                                                KWin w, double dx, double dy, double dz, double eps,
                                                double* __restrict__ part) {
  const int ix = w.nx - 2, iy = w.ny - 2, iz = DIM == 3 ? w.nz - 2 : 1;
  const int64_t n = (int64_t)ix * iy * iz;
  const int Y = w.nx, Z = w.nx * w.ny;
  const double dV = DIM == 3 ? dx * dy * dz : dx * dy;
  const double pi = 3.14159265358979323846;
  double area = 0, kint = 0, cnt = 0;
  for (int64_t t = (int64_t)blockIdx.x * RB + threadIdx.x; t < n; t += (int64_t)gridDim.x * RB) {
    const int x = (int)(t % ix) + 1;
    const int64_t r = t / ix;
    const int y = (int)(r % iy) + 1;
    const int z = DIM == 3 ? (int)(r / iy) + 1 : 0;
    const int id = z * Z + y * Y + x;
    const double ph = (double)phi[id];
    if (ph < -eps || ph > eps) continue;
    const double px = ((double)phi[id + 1] - (double)phi[id - 1]) / (2.0 * dx);
    const double py = ((double)phi[id + Y] - (double)phi[id - Y]) / (2.0 * dy);
    double g2 = px * px + py * py;
    if (DIM == 3) {
      const double pz = ((double)phi[id + Z] - (double)phi[id - Z]) / (2.0 * dz);
      g2 += pz * pz;
    }
    const double delta = sqrt(g2) * (1 + cos(pi * ph / eps)) / (2 * eps);
    area += delta * dV;
    if (kappa) kint += delta * (double)kappa[id] * dV;
    cnt += 1;
  }
  double v[3] = {area, kint, cnt};
  const int kind[3] = {0, 0, 0};
  block_reduce<3>(v, kind, part + 3 * blockIdx.x);
}

// ------------------------------------------------------------ error norms
template <typename TA, typename TB>
__global__ __launch_bounds__(RB) void norms_k(const TA* __restrict__ a, const TB* __restrict__ b,
                                              int64_t n, double* __restrict__ part) {
  double mx = 0, ss = 0, mb = 0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    const double bb = (double)b[i];
    const double d = fabs((double)a[i] - bb);
    mx = d > mx ? d : mx;
    ss += d * d;
    const double ab = fabs(bb);
    mb = ab > mb ? ab : mb;
  }
  double v[3] = {mx, ss, mb};
  const int kind[3] = {1, 0, 1};
  block_reduce<3>(v, kind, part + 3 * blockIdx.x);
}

static int fetch(double* host, const double* dev, int count, hipStream_t s) {
  DCL_HIP(hipMemcpyAsync(host, dev, sizeof(double) * count, hipMemcpyDeviceToHost, s));
  DCL_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_max_velocity(const dcl_velocity* vel, const void* nodes, int64_t count,
                                int64_t stride_bytes, int dim, int precision, double* max_out,
                                void* stream) {
  DCL_TRY(check_precision(precision));
  if (!max_out) return set_error(DCL_ERR_ARG_NULL, "max_out is NULL");
  if (dim != 2 && dim != 3) return set_error(DCL_ERR_ARG_WRONG, "dim %d is neither 2 nor 3", dim);
  DCL_TRY(check_velocity(vel, dim));
  if (count < 0) return set_error(DCL_ERR_ARG_SIZ, "count < 0");
  if (count > 0 && !nodes) return set_error(DCL_ERR_ARG_NULL, "nodes is NULL");
  if (stride_bytes < 12 || stride_bytes % 4)
    return set_error(DCL_ERR_ARG_WRONG, "stride %lld must be a multiple of 4 and >= 12",
                     (long long)stride_bytes);
  *max_out = 0.0;
  if (count == 0) return 0;
  const KVel v = kvel(*vel);
  const int nb = (int)(grid1d(count, RB) < (unsigned)RGRID ? grid1d(count, RB) : RGRID);
  void* sc = nullptr;
  DCL_TRY(scratch(sizeof(double) * (nb + 1), &sc));
  double* part = static_cast<double*>(sc);
  hipStream_t s = as_stream(stream);
  const char* nd = static_cast<const char*>(nodes);
#define DCL_MV(T, D)                                                                        \
  hipLaunchKernelGGL((maxvel_k<T, D>), dim3(nb), dim3(RB), 0, s, static_cast<const T*>(vel->data), \
                     v, nd, count, stride_bytes, part)
  if (precision == DCL_F64) {
    if (dim == 3) DCL_MV(double, 3); else DCL_MV(double, 2);
  } else {
    if (dim == 3) DCL_MV(float, 3); else DCL_MV(float, 2);
  }
#undef DCL_MV
  DCL_LAUNCHED("maxvel");
  hipLaunchKernelGGL((fold_partials_k<1>), dim3(1), dim3(RB), 0, s, part, nb, part + nb, 1, 0, 0, 0);
  DCL_LAUNCHED("fold_partials");
  return fetch(max_out, part + nb, 1, s);
}

// scratch of a band compaction: total | tile sums | block counts
static size_t band_scratch_bytes(int64_t n) {
  const int64_t nb = grid1d(n, RB);
  return 16 + sizeof(unsigned long long) * scan_tiles(nb) + sizeof(unsigned) * nb;
}

// The allocation-free band compaction's scratch: one array per device, held
// until the process ends, sized for windows of up to BAND_STATIC_MAX_NODES
// nodes.  Host code uses the device address that hipGetSymbolAddress returns
// for it on each device, never the array's name.
constexpr int64_t BAND_STATIC_MAX_NODES = (int64_t)1 << 28;
constexpr size_t BAND_STATIC_BYTES = (4u << 20) + (64u << 10);
static __device__ __align__(256) unsigned char band_static_scratch[BAND_STATIC_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_static_mu;
static void* g_static_ptr[MAX_DEVICES];

static int band_static_scratch_ptr(void** ptr) {
  int dev = 0;
  DCL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return set_error(DCL_ERR_SUP, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void* p = nullptr;
    DCL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(band_static_scratch)));
    g_static_ptr[dev] = p;
  }
  *ptr = g_static_ptr[dev];
  return 0;
}

template <typename T>
static int band_compact_t(const void* phi, const dcl_window* win, double width, dcl_icoor* out,
                          int64_t cap, int64_t* count_out, void* sc, hipStream_t s) {
  const int64_t n = win_nodes(*win);
  const int64_t nb = grid1d(n, RB);
  const int64_t ntiles = scan_tiles(nb);
  unsigned long long* total = static_cast<unsigned long long*>(sc);
  unsigned long long* sums = total + 2;
  unsigned* counts = reinterpret_cast<unsigned*>(sums + ntiles);
  const T* p = static_cast<const T*>(phi);
  hipLaunchKernelGGL((band_count_k<T>), dim3(nb), dim3(RB), 0, s, p, n, (T)width, counts);
  DCL_LAUNCHED("band_count");
  const ScanArrays sa = {{counts, nullptr}, {sums, nullptr}, {total, nullptr}};
  DCL_TRY(scan_counts(sa, 1, nb, s));
  if (cap > 0) {
    hipLaunchKernelGGL((band_write_k<T>), dim3(nb), dim3(RB), 0, s, p, kwin(*win), n, (T)width, counts, sums,
                       out, cap);
    DCL_LAUNCHED("band_write");
  }
  unsigned long long tot = 0;
  DCL_HIP(hipMemcpyAsync(&tot, total, sizeof(tot), hipMemcpyDeviceToHost, s));
  DCL_HIP(hipStreamSynchronize(s));
  *count_out = (int64_t)tot;
  return 0;
}

static int check_band_compact(const void* phi, const dcl_window* win, int precision, const dcl_icoor* band_out,
                              int64_t capacity, const int64_t* count_out) {
  DCL_TRY(check_precision(precision));
  if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
  if (!count_out) return set_error(DCL_ERR_ARG_NULL, "count_out is NULL");
  if (capacity < 0) return set_error(DCL_ERR_ARG_SIZ, "capacity < 0");
  if (capacity > 0 && !band_out) return set_error(DCL_ERR_ARG_NULL, "band_out is NULL");
  return check_window(win, "phi", 1);
}

extern "C" int dcl_band_compact(const void* phi, const dcl_window* win, double width, int precision,
                                dcl_icoor* band_out, int64_t capacity, int64_t* count_out,
                                void* stream) {
  DCL_TRY(check_band_compact(phi, win, precision, band_out, capacity, count_out));
  hipStream_t s = as_stream(stream);
  void* sc = nullptr;
  DCL_TRY(scratch(band_scratch_bytes(win_nodes(*win)), &sc));
  if (precision == DCL_F64) return band_compact_t<double>(phi, win, width, band_out, capacity, count_out, sc, s);
  return band_compact_t<float>(phi, win, width, band_out, capacity, count_out, sc, s);
}

extern "C" int dcl_band_compact_static(const void* phi, const dcl_window* win, double width, int precision,
                                       dcl_icoor* band_out, int64_t capacity, int64_t* count_out,
                                       void* stream) {
  DCL_TRY(check_band_compact(phi, win, precision, band_out, capacity, count_out));
  const int64_t n = win_nodes(*win);
  if (n > BAND_STATIC_MAX_NODES || band_scratch_bytes(n) > BAND_STATIC_BYTES)
    return set_error(DCL_ERR_ARG_SIZ,
                     "window of %lld nodes: the allocation-free band compaction takes windows of up to "
                     "%lld nodes (its static scratch holds %zu bytes)",
                     (long long)n, (long long)BAND_STATIC_MAX_NODES, BAND_STATIC_BYTES);
  void* sc = nullptr;
  DCL_TRY(band_static_scratch_ptr(&sc));
  hipStream_t s = as_stream(stream);
  if (precision == DCL_F64) return band_compact_t<double>(phi, win, width, band_out, capacity, count_out, sc, s);
  return band_compact_t<float>(phi, win, width, band_out, capacity, count_out, sc, s);
}

extern "C" int dcl_interface_metrics(const void* phi, const void* kappa, const dcl_window* win,
                                     double eps, int precision, double out[4], void* stream) {
  DCL_TRY(check_precision(precision));
  if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
  if (!out) return set_error(DCL_ERR_ARG_NULL, "out is NULL");
  DCL_TRY(check_window(win, "phi", 3));
  if (!(eps > 0)) return set_error(DCL_ERR_ARG_WRONG, "eps must be positive");
  const int dim = win->n.z < 2 ? 2 : 3;
  const int nb = RGRID;
  void* sc = nullptr;
  DCL_TRY(scratch(sizeof(double) * 3 * (nb + 1), &sc));
  double* part = static_cast<double*>(sc);
  hipStream_t s = as_stream(stream);
  const KWin w = kwin(*win);
#define DCL_MT(T, D)                                                                             \
  hipLaunchKernelGGL((metrics_k<T, D>), dim3(nb), dim3(RB), 0, s, static_cast<const T*>(phi),       \
                     static_cast<const T*>(kappa), w, win->d.x, win->d.y, win->d.z, eps, part)
  if (precision == DCL_F64) {
    if (dim == 3) DCL_MT(double, 3); else DCL_MT(double, 2);
  } else {
    if (dim == 3) DCL_MT(float, 3); else DCL_MT(float, 2);
  }
#undef DCL_MT
  DCL_LAUNCHED("interface_metrics");
  hipLaunchKernelGGL((fold_partials_k<3>), dim3(1), dim3(RB), 0, s, part, nb, part + 3 * nb, 0, 0, 0, 0);
  DCL_LAUNCHED("fold_partials");
  double r[3];
  DCL_TRY(fetch(r, part + 3 * nb, 3, s));
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[0] > 0 ? r[1] / r[0] : 0.0;
  out[3] = r[2];
  return 0;
}

extern "C" int dcl_error_norms(const void* a, int prec_a, const void* b, int prec_b, int64_t n,
                               double out[4], void* stream) {
  DCL_TRY(check_precision(prec_a));
  DCL_TRY(check_precision(prec_b));
  if (!out) return set_error(DCL_ERR_ARG_NULL, "out is NULL");
  if (n < 0) return set_error(DCL_ERR_ARG_SIZ, "n < 0");
  if (n > 0 && (!a || !b)) return set_error(DCL_ERR_ARG_NULL, "a or b is NULL");
  out[0] = out[1] = out[2] = 0;
  out[3] = (double)n;
  if (n == 0) return 0;
  const int nb = (int)(grid1d(n, RB) < (unsigned)RGRID ? grid1d(n, RB) : RGRID);
  void* sc = nullptr;
  DCL_TRY(scratch(sizeof(double) * 3 * (nb + 1), &sc));
  double* part = static_cast<double*>(sc);
  hipStream_t s = as_stream(stream);
#define DCL_NM(TA, TB)                                                                         \
  hipLaunchKernelGGL((norms_k<TA, TB>), dim3(nb), dim3(RB), 0, s, static_cast<const TA*>(a),       \
                     static_cast<const TB*>(b), n, part)
  if (prec_a == DCL_F64 && prec_b == DCL_F64) DCL_NM(double, double);
  else if (prec_a == DCL_F64) DCL_NM(double, float);
  else if (prec_b == DCL_F64) DCL_NM(float, double);
  else DCL_NM(float, float);
#undef DCL_NM
  DCL_LAUNCHED("error_norms");
  hipLaunchKernelGGL((fold_partials_k<3>), dim3(1), dim3(RB), 0, s, part, nb, part + 3 * nb, 1, 0, 1, 0);
  DCL_LAUNCHED("fold_partials");
  return fetch(out, part + 3 * nb, 3, s);
}
