// pls.hip -- particle level set on gfx950.
//
//   dcl_pls_advect       ParticleLS_AdvectParticles3D / 2D (ParticleLS3D.c:6-36,
//                        ParticleLS2D.c:6-36): RK2 midpoint, one thread per particle
//   dcl_pls_correct      ParticleLS_ErrorCorrection3D / 2D (ParticleLS3D.c:40-116,
//                        ParticleLS2D.c:40-115)
//   dcl_pls_adjust_radii ParticleLS_AdjustRadii (ParticleLS.c:158-185)
//
// The correction keeps phi+ and phi- as ordered-integer keys of the particle
// extrema only (the reference copies phi into two whole grids first,
// ParticleLS3D.c:60-61); max(phi, P+) is the reference's phi+ because max is
// order-free.  Escaped particles scatter with integer atomic max/min and
// append the touched node to an event list; the merge then visits only
// touched nodes (dense) or the band, and puts the keys back to their neutral
// values, so no full-grid pass is needed per call.
#include <new>

#include "dcl_buf.hpp"
#include "dcl_host.hpp"

namespace dcl {

template <typename T>
__device__ __forceinline__ T psign(T a) { return a >= T(0) ? (a == T(0) ? T(0) : T(1)) : T(-1); }

// |X - a|^2 summed left to right without FMA contraction, so the FP64 corner
// values round exactly as the reference's expression (ParticleLS3D.c:85-87)
__device__ __forceinline__ double dist2(double dx, double dy, double dz, bool three) {
#pragma clang fp contract(off)
  double r = dx * dx + dy * dy;
  if (three) r = r + dz * dz;
  return r;
}

// --------------------------------------------------------------- advect
// BUF (3D, velocity window below 2 GiB): the MAC interpolations read with
// buffer loads (32-bit offsets; with 3 FP32 components an x pair of faces is
// one 16-byte load), the same arithmetic as the pointer path
template <typename T, int DIM, bool BUF>
__global__ __launch_bounds__(256) void pls_advect_k(T* __restrict__ P, int64_t n,
                                                    const T* __restrict__ vel, KVel v,
                                                    double dhx, double dhy, double dhz, double dt,
                                                    unsigned vbytes) {
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= n) return;
  T* p = P + 4 * m;
  const double dh[3] = {dhx, dhy, dhz};
  double X[3] = {(double)p[0], (double)p[1], (double)p[2]};
  T V1[3], V2[3];
  const rsrc_t rv = make_rsrc(vel, vbytes);
  if (BUF && DIM == 3)
    mac_vel_buf<T>(rv, v, X, V1);
  else
    mac_vel<T, DIM>(vel, v, X, V1);
  double Xm[3] = {X[0], X[1], X[2]};
#pragma unroll
  for (int a = 0; a < DIM; ++a) Xm[a] = X[a] + dt / 2 * (double)V1[a] / dh[a];
  if (BUF && DIM == 3)
    mac_vel_buf<T>(rv, v, Xm, V2);
  else
    mac_vel<T, DIM>(vel, v, Xm, V2);
#pragma unroll
  for (int a = 0; a < DIM; ++a) p[a] = (T)(X[a] + dt * (double)V2[a] / dh[a]);
}

// ------------------------------------------------------------- correct
struct PlsK {
  KWin w;
  int flags;
  unsigned epoch;
  long long max_events;
};

template <typename T>
__global__ void pls_reset_counter(unsigned long long* counter) { *counter = 0; }

// escape test + corner scatter, one thread per particle
template <typename T, int DIM>
__global__ __launch_bounds__(256) void pls_escape_k(const T* __restrict__ P, int64_t n,
                                                    const T* __restrict__ phi,
                                                    typename KeyOf<T>::type* __restrict__ kpos,
                                                    typename KeyOf<T>::type* __restrict__ kneg,
                                                    int* __restrict__ events,
                                                    unsigned long long* __restrict__ counter, PlsK K) {
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= n) return;
  const T* p = P + 4 * m;
  const double X[3] = {(double)p[0], (double)p[1], (double)p[2]};
  const T r = p[3];
  const T dist = grid_interp<T, DIM>(phi, K.w, X);
  const T sign = psign(r);
  if (!(sign * dist < T(0) && fabs(dist) > fabs(r))) return;
  const int Ax = (int)floor(X[0]);
  const int Ay = (int)floor(X[1]);
  const int Az = DIM == 3 ? (int)floor((K.flags & DCL_PLS_REF_AZ) ? X[1] : X[2]) : K.w.pz;
#pragma unroll
  for (int k = 0; k < (DIM == 3 ? 2 : 1); ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int ax = Ax + i, ay = Ay + j, az = Az + k;
        const int lx = ax - K.w.px, ly = ay - K.w.py, lz = DIM == 3 ? az - K.w.pz : 0;
        if (lx < 0 || lx >= K.w.nx || ly < 0 || ly >= K.w.ny || lz < 0 || lz >= K.w.nz) continue;
        const double r2 = dist2(X[0] - ax, X[1] - ay, X[2] - az, DIM == 3);
        const T phi_p = sign * (sign * r - (T)sqrt(r2));
        const int id = (lz * K.w.ny + ly) * K.w.nx + lx;
        if (sign > T(0)) atomicMax(&kpos[id], okey(phi_p));
        else atomicMin(&kneg[id], okey(phi_p));
        const unsigned long long e = atomicAdd(counter, 1ull);
        if ((long long)e < K.max_events) events[e] = id;
      }
}

template <typename T>
__device__ __forceinline__ T merged(T ph, typename KeyOf<T>::type kp, typename KeyOf<T>::type kn) {
  const T pp = okey_val(kp), pn = okey_val(kn);
  const T phi_pos = pp < ph ? ph : pp;   // PetscMax(phi_p, phi_pos)
  const T phi_neg = pn < ph ? pn : ph;   // PetscMin(phi_p, phi_neg)
  return fabs(phi_pos) <= fabs(phi_neg) ? phi_pos : phi_neg;
}

// dense merge: the first event of a node (per call epoch) owns it
template <typename T>
__global__ __launch_bounds__(256) void pls_merge_events_k(T* __restrict__ phi,
                                                          typename KeyOf<T>::type* __restrict__ kpos,
                                                          typename KeyOf<T>::type* __restrict__ kneg,
                                                          unsigned* __restrict__ claim,
                                                          const int* __restrict__ events,
                                                          const unsigned long long* __restrict__ counter,
                                                          PlsK K) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long ne = (long long)*counter;
  if (e >= ne || e >= K.max_events) return;
  const int id = events[e];
  if (atomicExch(&claim[id], K.epoch) == K.epoch) return;
  typedef typename KeyOf<T>::type Key;
  const Key kp = kpos[id], kn = kneg[id];
  phi[id] = merged<T>(phi[id], kp, kn);
  kpos[id] = okey(-(T)INFINITY);
  kneg[id] = okey((T)INFINITY);
}

// band merge (the reference's loop over ls->band, ParticleLS3D.c:104-113)
template <typename T, bool COMPACT>
__global__ __launch_bounds__(256) void pls_merge_band_k(T* phi, T* out,
                                                        const typename KeyOf<T>::type* __restrict__ kpos,
                                                        const typename KeyOf<T>::type* __restrict__ kneg,
                                                        const dcl_icoor* __restrict__ band, int64_t n,
                                                        PlsK K, int dim) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  const dcl_icoor q = band[b];
  const int lx = q.x - K.w.px, ly = q.y - K.w.py, lz = dim == 3 ? q.z - K.w.pz : 0;
  const bool inside = lx >= 0 && lx < K.w.nx && ly >= 0 && ly < K.w.ny && lz >= 0 && lz < K.w.nz;
  if (!inside) {
    if (COMPACT) out[b] = T(0);
    return;
  }
  const int id = (lz * K.w.ny + ly) * K.w.nx + lx;
  const T v = merged<T>(phi[id], kpos[id], kneg[id]);
  if (COMPACT) out[b] = v;
  else phi[id] = v;
}

// put the keys of touched nodes back to neutral (band mode); idempotent writes
template <typename T>
__global__ __launch_bounds__(256) void pls_reset_events_k(typename KeyOf<T>::type* __restrict__ kpos,
                                                          typename KeyOf<T>::type* __restrict__ kneg,
                                                          const int* __restrict__ events,
                                                          const unsigned long long* __restrict__ counter,
                                                          PlsK K) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long ne = (long long)*counter;
  if (e >= ne || e >= K.max_events) return;
  const int id = events[e];
  kpos[id] = okey(-(T)INFINITY);
  kneg[id] = okey((T)INFINITY);
}

template <typename T>
__global__ __launch_bounds__(256) void pls_fill_keys_k(typename KeyOf<T>::type* kpos,
                                                       typename KeyOf<T>::type* kneg,
                                                       unsigned* claim, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  kpos[i] = okey(-(T)INFINITY);
  kneg[i] = okey((T)INFINITY);
  claim[i] = 0u;
}

// --------------------------------------------------------------- radii
template <typename T, int DIM>
__global__ __launch_bounds__(256) void pls_radii_k(T* __restrict__ P, int64_t n,
                                                   const T* __restrict__ phi, KWin w, T rmin, T rmax) {
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= n) return;
  T* p = P + 4 * m;
  const double X[3] = {(double)p[0], (double)p[1], (double)p[2]};
  const T dist = grid_interp<T, DIM>(phi, w, X);
  const T sign = psign(p[3]);
  const T sphi = sign * dist;
  if (rmin < sphi && sphi < rmax) p[3] = sign * sphi;
  else if (sphi > rmax) p[3] = sign * rmax;
  else if (sphi < rmin) p[3] = sign * rmin;
}

}  // namespace dcl

using namespace dcl;

struct dcl_pls_ws_s {
  int precision;
  int64_t max_nodes, max_events;
  void* kpos;
  void* kneg;
  unsigned* claim;
  int* events;
  unsigned long long* counter;
  unsigned epoch;
};

static int check_particles(const void* particles, int64_t count, int dim, int precision) {
  DCL_TRY(check_precision(precision));
  if (count < 0) return set_error(DCL_ERR_ARG_SIZ, "particle count %lld < 0", (long long)count);
  if (count > 0 && !particles) return set_error(DCL_ERR_ARG_NULL, "particles is NULL");
  if (dim != 2 && dim != 3) return set_error(DCL_ERR_ARG_WRONG, "dim %d is neither 2 nor 3", dim);
  return 0;
}

extern "C" int dcl_pls_advect(void* particles, int64_t count, const dcl_velocity* vel, dcl_coor dh,
                              double dt, int dim, int precision, void* stream) {
  DCL_TRY(check_particles(particles, count, dim, precision));
  DCL_TRY(check_velocity(vel, dim));
  if (!(dh.x > 0 && dh.y > 0 && (dim == 2 || dh.z > 0)))
    return set_error(DCL_ERR_ARG_WRONG, "dh must be positive");
  if (count == 0) return 0;
  const KVel v = kvel(*vel);
  hipStream_t s = as_stream(stream);
  const unsigned g = grid1d(count, 256);
  const long long vb = win_nodes(vel->win) * vel->dof * (long long)(precision == DCL_F64 ? 8 : 4);
  const bool buf = dim == 3 && buf_ok(0, vb);
#define DCL_ADV(T, D)                                                                                     \
  if (buf)                                                                                              \
    hipLaunchKernelGGL((pls_advect_k<T, D, true>), dim3(g), dim3(256), 0, s, static_cast<T*>(particles),  \
                       count, static_cast<const T*>(vel->data), v, dh.x, dh.y, dh.z, dt, (unsigned)vb);   \
  else                                                                                                  \
    hipLaunchKernelGGL((pls_advect_k<T, D, false>), dim3(g), dim3(256), 0, s, static_cast<T*>(particles), \
                       count, static_cast<const T*>(vel->data), v, dh.x, dh.y, dh.z, dt, 0u)
  if (precision == DCL_F64) {
    if (dim == 3) DCL_ADV(double, 3); else DCL_ADV(double, 2);
  } else {
    if (dim == 3) DCL_ADV(float, 3); else DCL_ADV(float, 2);
  }
#undef DCL_ADV
  DCL_LAUNCHED("pls_advect");
  return 0;
}

extern "C" int dcl_pls_ws_create(int64_t max_nodes, int64_t max_events, int precision, void* stream,
                                 dcl_pls_ws* out) {
  DCL_TRY(check_precision(precision));
  if (!out) return set_error(DCL_ERR_ARG_NULL, "ws is NULL");
  *out = nullptr;
  if (max_nodes <= 0 || max_nodes >= ((int64_t)1 << 31))
    return set_error(DCL_ERR_ARG_SIZ, "max_nodes %lld out of range", (long long)max_nodes);
  if (max_events <= 0 || max_events >= ((int64_t)1 << 40))
    return set_error(DCL_ERR_ARG_SIZ, "max_events %lld out of range", (long long)max_events);
  dcl_pls_ws ws = new (std::nothrow) dcl_pls_ws_s();
  if (!ws) return set_error(DCL_ERR_MEM, "host allocation failed");
  ws->precision = precision;
  ws->max_nodes = max_nodes;
  ws->max_events = max_events;
  ws->epoch = 0;
  const size_t kb = precision == DCL_F64 ? 8 : 4;
  int rc = 0;
  void* p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  const size_t bytes[5] = {kb * max_nodes, kb * max_nodes, 4 * (size_t)max_nodes, 4 * (size_t)max_events, 8};
  for (int i = 0; i < 5 && !rc; ++i) rc = dcl_malloc(&p[i], bytes[i]);
  if (rc) {
    for (int i = 0; i < 5; ++i) dcl_free(p[i]);
    delete ws;
    return rc;
  }
  ws->kpos = p[0];
  ws->kneg = p[1];
  ws->claim = static_cast<unsigned*>(p[2]);
  ws->events = static_cast<int*>(p[3]);
  ws->counter = static_cast<unsigned long long*>(p[4]);
  hipStream_t s = as_stream(stream);
  const unsigned g = grid1d(max_nodes, 256);
  if (precision == DCL_F64)
    hipLaunchKernelGGL((pls_fill_keys_k<double>), dim3(g), dim3(256), 0, s,
                       static_cast<unsigned long long*>(ws->kpos),
                       static_cast<unsigned long long*>(ws->kneg), ws->claim, max_nodes);
  else
    hipLaunchKernelGGL((pls_fill_keys_k<float>), dim3(g), dim3(256), 0, s,
                       static_cast<unsigned*>(ws->kpos), static_cast<unsigned*>(ws->kneg), ws->claim,
                       max_nodes);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    for (int i = 0; i < 5; ++i) dcl_free(p[i]);
    delete ws;
    return hip_fail(e, "pls workspace init", __FILE__, __LINE__);
  }
  *out = ws;
  return 0;
}

extern "C" int dcl_pls_ws_destroy(dcl_pls_ws ws) {
  if (!ws) return 0;
  DCL_HIP(hipDeviceSynchronize());
  dcl_free(ws->kpos);
  dcl_free(ws->kneg);
  dcl_free(ws->claim);
  dcl_free(ws->events);
  dcl_free(ws->counter);
  delete ws;
  return 0;
}

template <typename T>
static int pls_correct_t(const void* particles, int64_t count, void* phi, const dcl_window* win,
                         const dcl_icoor* band, int64_t band_len, void* band_values, int flags,
                         int dim, dcl_pls_ws ws, hipStream_t s) {
  typedef typename KeyOf<T>::type Key;
  PlsK K;
  K.w = kwin(*win);
  K.flags = flags;
  K.epoch = ++ws->epoch == 0 ? ++ws->epoch : ws->epoch;
  K.max_events = ws->max_events;
  Key* kp = static_cast<Key*>(ws->kpos);
  Key* kn = static_cast<Key*>(ws->kneg);
  T* ph = static_cast<T*>(phi);
  hipLaunchKernelGGL((pls_reset_counter<T>), dim3(1), dim3(1), 0, s, ws->counter);
  DCL_LAUNCHED("pls_reset_counter");
  if (count > 0) {
    const unsigned g = grid1d(count, 256);
    if (dim == 3)
      hipLaunchKernelGGL((pls_escape_k<T, 3>), dim3(g), dim3(256), 0, s, static_cast<const T*>(particles),
                         count, ph, kp, kn, ws->events, ws->counter, K);
    else
      hipLaunchKernelGGL((pls_escape_k<T, 2>), dim3(g), dim3(256), 0, s, static_cast<const T*>(particles),
                         count, ph, kp, kn, ws->events, ws->counter, K);
    DCL_LAUNCHED("pls_escape");
  }
  const int64_t corners = dim == 3 ? 8 : 4;
  const unsigned ge = grid1d(count * corners > 0 ? count * corners : 1, 256);
  if (!band) {
    hipLaunchKernelGGL((pls_merge_events_k<T>), dim3(ge), dim3(256), 0, s, ph, kp, kn, ws->claim,
                       ws->events, ws->counter, K);
    DCL_LAUNCHED("pls_merge_events");
    return 0;
  }
  if (band_len > 0) {
    const unsigned gb = grid1d(band_len, 256);
    if (band_values)
      hipLaunchKernelGGL((pls_merge_band_k<T, true>), dim3(gb), dim3(256), 0, s, ph,
                         static_cast<T*>(band_values), kp, kn, band, band_len, K, dim);
    else
      hipLaunchKernelGGL((pls_merge_band_k<T, false>), dim3(gb), dim3(256), 0, s, ph, ph, kp, kn, band,
                         band_len, K, dim);
    DCL_LAUNCHED("pls_merge_band");
  }
  hipLaunchKernelGGL((pls_reset_events_k<T>), dim3(ge), dim3(256), 0, s, kp, kn, ws->events,
                     ws->counter, K);
  DCL_LAUNCHED("pls_reset_events");
  return 0;
}

extern "C" int dcl_pls_correct(const void* particles, int64_t count, void* phi, const dcl_window* win,
                               const dcl_icoor* band, int64_t band_len, void* band_values, int flags,
                               int dim, int precision, dcl_pls_ws ws, void* stream) {
  DCL_TRY(check_particles(particles, count, dim, precision));
  if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
  if (!ws) return set_error(DCL_ERR_ARG_NULL, "workspace is NULL");
  DCL_TRY(check_window(win, "phi", 2));
  if ((win->n.z < 2 ? 2 : 3) != dim) return set_error(DCL_ERR_ARG_WRONG, "window is not %dD", dim);
  if (ws->precision != precision)
    return set_error(DCL_ERR_ARG_WRONG, "workspace precision differs from the call's");
  if (win_nodes(*win) > ws->max_nodes)
    return set_error(DCL_ERR_ARG_SIZ, "window has %lld nodes, workspace holds %lld",
                     (long long)win_nodes(*win), (long long)ws->max_nodes);
  const int64_t corners = dim == 3 ? 8 : 4;
  if (count > ws->max_events / corners)
    return set_error(DCL_ERR_ARG_SIZ, "%lld particles may touch %lld corners; workspace holds %lld events",
                     (long long)count, (long long)(count * corners), (long long)ws->max_events);
  if (band && band_len < 0) return set_error(DCL_ERR_ARG_SIZ, "band_len < 0");
  hipStream_t s = as_stream(stream);
  if (precision == DCL_F64)
    return pls_correct_t<double>(particles, count, phi, win, band, band_len, band_values, flags, dim, ws, s);
  return pls_correct_t<float>(particles, count, phi, win, band, band_len, band_values, flags, dim, ws, s);
}

extern "C" int dcl_pls_adjust_radii(void* particles, int64_t count, const void* phi,
                                    const dcl_window* win, double rmin, double rmax, int dim,
                                    int precision, void* stream) {
  DCL_TRY(check_particles(particles, count, dim, precision));
  if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
  DCL_TRY(check_window(win, "phi", 2));
  if ((win->n.z < 2 ? 2 : 3) != dim) return set_error(DCL_ERR_ARG_WRONG, "window is not %dD", dim);
  if (count == 0) return 0;
  const KWin w = kwin(*win);
  hipStream_t s = as_stream(stream);
  const unsigned g = grid1d(count, 256);
#define DCL_RAD(T, D)                                                                           \
  hipLaunchKernelGGL((pls_radii_k<T, D>), dim3(g), dim3(256), 0, s, static_cast<T*>(particles), \
                     count, static_cast<const T*>(phi), w, (T)rmin, (T)rmax)
  if (precision == DCL_F64) {
    if (dim == 3) DCL_RAD(double, 3); else DCL_RAD(double, 2);
  } else {
    if (dim == 3) DCL_RAD(float, 3); else DCL_RAD(float, 2);
  }
#undef DCL_RAD
  DCL_LAUNCHED("pls_adjust_radii");
  return 0;
}
