// adjoint.hip -- reverse-mode derivative of the dense 3D SL step on gfx950.
//
// Forward (sl_advect.hip, dense): phi[X] = I(phi0; S(X)), S = X - c V(Xm),
// Xm = X - (c/2) V(X) for RK2 (S = X - c V(X) for Euler), c = dt/dh, with
// V the MAC-staggered trilinear velocity and I the trilinear or Catmull-Rom
// gather.  Given lam_next = dL/dphi, one thread per node recomputes its
// forward state and scatters
//   dL/dphi0 += g w_ijk                               (transposed gather)
//   dL/dvel  += MAC weights x (-c g dI/dS)            (at Xm, or at X for Euler)
//             + node weights x (-(c/2) (dV/dXm)^T ...) (RK2 first stage)
// where a clamped coordinate contributes no derivative.  The transposed
// gather of one block lands in a small box of phi0 around the block's nodes,
// so it is accumulated in LDS first and flushed to HBM with one atomic per
// touched word; stencil entries that fall outside the block's box go to HBM
// directly.  x-neighbour lanes whose stencils overlap sum their shared words
// with DPP row shifts before the LDS atomic, and likewise the shared corners
// of their MAC velocity scatter before the HBM atomic.  Outer two layers:
// phi = phi0 there, so dL/dphi0 += lam_next.
#include <algorithm>
#include <type_traits>

#include "dcl_buf.hpp"
#include "dcl_host.hpp"

namespace dcl {

struct AdjP {
  KWin w;
  KVel v;
  double c[3];
  double ch[3];
};

template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

// MAC velocity at an arbitrary point with its spatial derivative and the
// data needed to scatter into the velocity: per component c the local base
// offset of its 2x2x2 corner cube and the three fractions.
template <typename T>
struct MacState {
  T V[3];
  T dV[3][3];   // dV[c][axis]
  int base[3];  // element offset of corner (0,0,0) of component c (includes c + udof)
  T f[3][3];    // fractions [c][axis]
};

template <typename T>
__device__ __forceinline__ void mac_state(const T* __restrict__ vel, const KVel& v, const double Xin[3],
                                          MacState<T>& m, bool clamped[3]) {
  double X[3];
  const int lo[3] = {v.w.px, v.w.py, v.w.pz};
  const int nn[3] = {v.w.nx, v.w.ny, v.w.nz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double hi = (double)(lo[a] + nn[a] - 2);
    X[a] = clampd(Xin[a], (double)lo[a], hi);
    clamped[a] = !(Xin[a] >= (double)lo[a] && Xin[a] <= hi);
  }
  const int sx = v.dof, sy = v.w.nx * v.dof, sz = v.w.ny * sy;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int bi[3];
    T fr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double sh = a == c ? 0.5 : 0.0;
      const double fl = floor(X[a] + sh);
      bi[a] = (int)fl - lo[a];
      fr[a] = (T)(X[a] - fl + sh);
      m.f[c][a] = fr[a];
    }
    const int base = bi[2] * sz + bi[1] * sy + bi[0] * sx + v.udof + c;
    m.base[c] = base;
    const T* b = vel + base;
    const T f000 = b[0], f100 = b[sx], f010 = b[sy], f110 = b[sy + sx];
    const T f001 = b[sz], f101 = b[sz + sx], f011 = b[sz + sy], f111 = b[sz + sy + sx];
    const T x1 = fr[0], x0 = T(1) - fr[0], y1 = fr[1], y0 = T(1) - fr[1], z1 = fr[2], z0 = T(1) - fr[2];
    // values along x for the four (y, z) rows
    const T r00 = x0 * f000 + x1 * f100, r10 = x0 * f010 + x1 * f110;
    const T r01 = x0 * f001 + x1 * f101, r11 = x0 * f011 + x1 * f111;
    m.V[c] = z0 * (y0 * r00 + y1 * r10) + z1 * (y0 * r01 + y1 * r11);
    const T d00 = f100 - f000, d10 = f110 - f010, d01 = f101 - f001, d11 = f111 - f011;
    m.dV[c][0] = clamped[0] ? T(0) : z0 * (y0 * d00 + y1 * d10) + z1 * (y0 * d01 + y1 * d11);
    m.dV[c][1] = clamped[1] ? T(0) : z0 * (r10 - r00) + z1 * (r11 - r01);
    m.dV[c][2] = clamped[2] ? T(0) : (y0 * r01 + y1 * r11) - (y0 * r00 + y1 * r10);
  }
}

// DPP row shifts (VALU, within rows of 16 lanes): the value of the lane d
// below / above in the same row, 0 where that lane is outside the row.
// Every lane of the wave executes them.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_t(float v) {
  return __int_as_float(dpp_i32<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_t(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i32<CTRL>((int)(unsigned)(unsigned long long)b);
  const int hi = dpp_i32<CTRL>((int)(unsigned)((unsigned long long)b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ int dpp_t(int v) {
  return dpp_i32<CTRL>(v);
}
// row_shr:d (0x110 + d): lane l reads lane l - d; row_shl:d (0x100 + d): l + d
template <typename V>
__device__ __forceinline__ V lane_from_lower(V v, int d) {
  return d == 1 ? dpp_t<0x111>(v) : d == 2 ? dpp_t<0x112>(v) : dpp_t<0x113>(v);
}
template <typename V>
__device__ __forceinline__ V lane_from_upper(V v, int d) {
  return d == 1 ? dpp_t<0x101>(v) : d == 2 ? dpp_t<0x102>(v) : dpp_t<0x103>(v);
}

// mac_scatter with x-neighbour lanes combined: where the lane d = 1 below in
// the 16-lane row has its corner cube of component c one x cell lower, its
// x1 corners are this lane's x0 corners, which this lane adds; that lane
// skips them.  Every lane of the wave calls it (act false: nothing to add).
template <typename T>
__device__ __forceinline__ void mac_scatter_lanes(T* __restrict__ gvel, const KVel& v, const MacState<T>& m,
                                                  const T H[3], bool act) {
  const int sx = v.dof, sy = v.w.nx * v.dof, sz = v.w.ny * sy;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const bool on = act && H[c] != T(0);
    const int key = on ? m.base[c] + 1 : 0;
    const int kl = lane_from_lower(key, 1), ku = lane_from_upper(key, 1);
    const bool lo = on && kl != 0 && kl + sx == key, up = on && ku != 0 && key + sx == ku;
    T a[4] = {T(0), T(0), T(0), T(0)}, b[4] = {T(0), T(0), T(0), T(0)};  // x0 / x1 corners: y0z0 y1z0 y0z1 y1z1
    if (on) {
      const T x1 = m.f[c][0], x0 = T(1) - x1, y1 = m.f[c][1], y0 = T(1) - y1, z1 = m.f[c][2], z0 = T(1) - z1;
      const T h = H[c];
      a[0] = h * x0 * y0 * z0;
      b[0] = h * x1 * y0 * z0;
      a[1] = h * x0 * y1 * z0;
      b[1] = h * x1 * y1 * z0;
      a[2] = h * x0 * y0 * z1;
      b[2] = h * x1 * y0 * z1;
      a[3] = h * x0 * y1 * z1;
      b[3] = h * x1 * y1 * z1;
    }
    T bl[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bl[q] = lane_from_lower(b[q], 1);
    if (on) {
      T* p = gvel + m.base[c];
      const int off[4] = {0, sy, sz, sz + sy};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        atomic_add(p + off[q], lo ? a[q] + bl[q] : a[q]);
        if (!up) atomic_add(p + off[q] + sx, b[q]);
      }
    }
  }
}

// A wave's lanes are a 16 x 4 plane of nodes (its 16-lane DPP rows are x
// rows, lanes ABX apart are y neighbours); adj_finish_k's blocks stack ABZ
// such planes.
constexpr int ABX = 16, ABY = 4, ABZ = 4;
static_assert(ABX * ABY == 64, "a wave is one z plane of the block: lanes ABX apart are y neighbours");

// One wave per block: its 16 x 4 lanes own a 16 x 4 column of nodes and walk
// AWZ planes of it, scattering lam into a wave-private LDS box (the column
// plus the reach WR) with plain read-add-write, lanes that share a stencil
// corner found with a tag table and given LDS atomics, stencils beyond the
// box added into lam with HBM atomics (as adj_step_k below); the velocity
// gradient goes to HBM with atomics, x- and y-neighbour lanes combining the
// MAC corners and node faces they share.
constexpr int AWZ = 4, WR = 3;  // node planes per wave, box reach
constexpr int WBX = ABX + 2 * WR, WBY = ABY + 2 * WR, WBZ = AWZ + 2 * WR, WBOX = WBX * WBY * WBZ;
constexpr int NTAG = 1024;

template <typename T, int INTERP, int BT>
__global__ __launch_bounds__(64) void sl_adjoint_k(const T* __restrict__ phi0, const T* __restrict__ vel,
                                                   const T* __restrict__ lam_next, T* __restrict__ lam,
                                                   T* __restrict__ gvel, AdjP P) {
  __shared__ T box_s[WBOX];
  __shared__ int tag_s[NTAG];
  typedef __attribute__((address_space(3))) T lds_T;
  typedef __attribute__((address_space(3))) int lds_int;
  volatile lds_T* box = (volatile lds_T*)box_s;
  volatile lds_int* tag = (volatile lds_int*)tag_s;
  const int lane = threadIdx.x;
  for (int t = lane; t < WBOX; t += 64) box[t] = T(0);

  const int ox = blockIdx.x * ABX, oy = blockIdx.y * ABY, oz = blockIdx.z * AWZ;
  const int i = ox + (lane & (ABX - 1)), j = oy + lane / ABX;
  const int nx = P.w.nx, ny = P.w.ny, nz = P.w.nz;
  const int Y = nx, Z = nx * ny;
  const int NP = INTERP == DCL_CUBIC ? 4 : 2;
  const int OFF = INTERP == DCL_CUBIC ? -1 : 0;
#pragma unroll 1
  for (int kz = 0; kz < AWZ; ++kz) {
    const int k = oz + kz;
    const bool valid = i < nx && j < ny && k < nz;
    const bool inner = valid && i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3 && k >= 2 && k <= nz - 3;
    if (valid && !inner) {
      const int id = k * Z + j * Y + i;
      const T g = lam_next[id];
      if (g != T(0)) atomic_add(&lam[id], g);
    }
    // The stencil loop runs with every lane of the wave (lanes without work
    // carry g = 0 and no box key), so that x-neighbour lanes can combine
    // their contributions to the same LDS word before writing it.
    T g = T(0);
    if (inner) g = lam_next[k * Z + j * Y + i];
    const bool act = inner && g != T(0);
    double Xm[3] = {0.0, 0.0, 0.0};
    int b[3] = {0, 0, 0};
    bool cs[3] = {false, false, false};
    T w[3][4], dw[3][4];
    int bx0 = 0, by0 = 0, bz0 = 0;
    bool in_box = false;
    if (act) {
      const int X = P.w.px + i, Yg = P.w.py + j, Zg = P.w.pz + k;
      const double Xd[3] = {(double)X, (double)Yg, (double)Zg};
      T V1[3];
      node_vel<T, 3>(vel, P.v, X, Yg, Zg, V1);
      // the midpoint value now (the forward's arithmetic); its derivative
      // state only after the stencil loop, to keep it out of the loop's
      // registers
      double S[3];
      if (BT == DCL_RK2) {
#pragma unroll
        for (int a = 0; a < 3; ++a) Xm[a] = Xd[a] - (double)V1[a] * P.ch[a];
        T V2[3];
        mac_vel<T, 3>(vel, P.v, Xm, V2);
#pragma unroll
        for (int a = 0; a < 3; ++a) S[a] = Xd[a] - (double)V2[a] * P.c[a];
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) S[a] = Xd[a] - (double)V1[a] * P.c[a];
      }
      // split + weights
      const int pp[3] = {P.w.px, P.w.py, P.w.pz}, nn[3] = {nx, ny, nz};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double f;
        if (INTERP == DCL_CUBIC) {
          const double lo = pp[a] + 1, hi = pp[a] + nn[a] - 2;
          cs[a] = !(S[a] >= lo && S[a] <= hi);
          b[a] = cub_split(S[a], pp[a], nn[a], f);
          cr_weights<T>((T)f, w[a]);
          cr_dweights<T>((T)f, dw[a]);
        } else {
          const double lo = pp[a], hi = pp[a] + nn[a] - 1;
          cs[a] = !(S[a] >= lo && S[a] <= hi);
          b[a] = lin_split(S[a], pp[a], nn[a], f);
          w[a][0] = (T)(1.0 - f);
          w[a][1] = (T)f;
          dw[a][0] = T(-1);
          dw[a][1] = T(1);
        }
      }
      bx0 = b[0] + OFF - (ox - WR);
      by0 = b[1] + OFF - (oy - WR);
      bz0 = b[2] + OFF - (oz - WR);
      in_box = bx0 >= 0 && bx0 + NP <= WBX && by0 >= 0 && by0 + NP <= WBY && bz0 >= 0 && bz0 + NP <= WBZ;
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) w[a][q] = dw[a][q] = T(0);
    }
    // the box word of the lane's stencil corner, plus one (0: no box word).
    // Lanes d = 1 .. NP-1 apart in a 16-lane row whose keys differ by d
    // write the same word in every stencil row, the lower lane with its
    // entry d and the upper lane with its entry 0: the lower one leaves it
    // to the upper one, which adds both.
    const int key = in_box ? (bz0 * WBY + by0) * WBX + bx0 + 1 : 0;
    const int slot = key & (NTAG - 1);
    if (in_box) tag[slot] = lane;
    const bool lost = in_box && tag[slot] != lane;
    if (lost) tag[slot] = -1;
    const bool shared = in_box && tag[slot] == -1;
    const bool plain = in_box && !shared;
    bool lo_m[4] = {false, false, false, false}, up_m[4] = {false, false, false, false};
#pragma unroll
    for (int d = 1; d < NP; ++d) {
      const int kl = lane_from_lower(key, d), ku = lane_from_upper(key, d);
      lo_m[d] = key != 0 && kl != 0 && kl + d == key;
      up_m[d] = key != 0 && ku != 0 && key + d == ku;
    }
    // transposed gather into the box (or HBM outside it) + gradient gather
    T G[3] = {T(0), T(0), T(0)};
    const T* src = phi0 + (b[2] + OFF) * Z + (b[1] + OFF) * Y + (b[0] + OFF);
    // one stencil plane at a time: bounds the loads in flight (registers)
#pragma unroll 1
    for (int kk = 0; kk < NP; ++kk)
#pragma unroll
      for (int jj = 0; jj < NP; ++jj) {
        const T wyz = w[1][jj] * w[2][kk];
        T add[4];
#pragma unroll
        for (int ii = 0; ii < NP; ++ii) add[ii] = g * w[0][ii] * wyz;
        T word0 = add[0];
#pragma unroll
        for (int d = 1; d < NP; ++d) {
          const T o = lane_from_lower(add[d], d);
          if (lo_m[d]) word0 += o;
        }
        if (act) {
          const T* r = src + kk * Z + jj * Y;
          T rx = T(0), rd = T(0);
#pragma unroll
          for (int ii = 0; ii < NP; ++ii) {
            const T f = r[ii];
            rx += w[0][ii] * f;
            rd += dw[0][ii] * f;
          }
          G[0] += rd * wyz;
          G[1] += rx * dw[1][jj] * w[2][kk];
          G[2] += rx * w[1][jj] * dw[2][kk];
        }
        const int wo = ((bz0 + kk) * WBY + (by0 + jj)) * WBX + bx0;
#pragma unroll
        for (int ii = 0; ii < NP; ++ii) {
          const T v = ii == 0 ? word0 : add[ii];
          const bool mine = ii == 0 || !up_m[ii];
          if (plain && mine) box[wo + ii] = box[wo + ii] + v;
          if (shared && mine) atomicAdd(&box_s[wo + ii], v);
          if (act && !in_box)
            atomic_add(&lam[(b[2] + OFF + kk) * Z + (b[1] + OFF + jj) * Y + (b[0] + OFF + ii)], add[ii]);
        }
      }
    T H[3] = {T(0), T(0), T(0)}, H1[3] = {T(0), T(0), T(0)};
    MacState<T> m2;
    if (act) {
#pragma unroll
      for (int a = 0; a < 3; ++a) H[a] = cs[a] ? T(0) : (T)(-P.c[a]) * g * G[a];
      if (BT == DCL_RK2) {
        bool cm[3] = {false, false, false};
        mac_state<T>(vel, P.v, Xm, m2, cm);
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
          const T K = H[0] * m2.dV[0][bb] + H[1] * m2.dV[1][bb] + H[2] * m2.dV[2][bb];
          H1[bb] = (T)(-P.ch[bb]) * K;
        }
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) H1[a] = H[a];
      }
    }
    if (BT == DCL_RK2) mac_scatter_lanes<T>(gvel, P.v, m2, H, act);
    // node velocity: V_a = (f(X) + f(X + e_a)) / 2.  The x face at X + e_x
    // is the next x lane's face at its node, and the y face at X + e_y that
    // of the lane ABX above (the node at y + 1 in the same wave): that lane
    // adds both.
    const int sx = P.v.dof, sy = P.w.nx * P.v.dof, sz = P.w.ny * sy;
    const int gofs = (k * P.w.ny + j) * sy + i * sx + P.v.udof;
    const T h0 = T(0.5) * H1[0], h1 = T(0.5) * H1[1], h2 = T(0.5) * H1[2];
    const int nkey = act ? gofs + 1 : 0;
    const int nkl = lane_from_lower(nkey, 1), nku = lane_from_upper(nkey, 1);
    const bool nlo = act && nkl != 0 && nkl + sx == nkey, nup = act && nku != 0 && nkey + sx == nku;
    const T h0l = lane_from_lower(h0, 1);
    const int ykl = __shfl_up(nkey, ABX), yku = __shfl_down(nkey, ABX);
    const bool ylo = act && lane >= ABX && ykl != 0 && ykl + sy == nkey;
    const bool yup = act && lane < 64 - ABX && yku != 0 && nkey + sy == yku;
    const T h1l = __shfl_up(h1, ABX);
    if (act) {
      T* gb = gvel + gofs;
      atomic_add(gb, nlo ? h0 + h0l : h0);
      if (!nup) atomic_add(gb + sx, h0);
      atomic_add(gb + 1, ylo ? h1 + h1l : h1);
      if (!yup) atomic_add(gb + sy + 1, h1);
      atomic_add(gb + 2, h2);
      atomic_add(gb + sz + 2, h2);
    }
  }
  // flush the box: one HBM atomic per nonzero word inside the window
  for (int t = lane; t < WBOX; t += 64) {
    const T v = box[t];
    if (v == T(0)) continue;
    const int lx = t % WBX, ly = (t / WBX) % WBY, lz = t / (WBX * WBY);
    const int gx = ox - WR + lx, gy = oy - WR + ly, gz = oz - WR + lz;
    if (gx < 0 || gx >= nx || gy < 0 || gy >= ny || gz < 0 || gz >= nz) continue;
    atomic_add(&lam[gz * Z + gy * Y + gx], v);
  }
}

// ---------------------------------------------------------------------------
// Steady velocity (one velocity for every step, the differentiation form):
// the departure geometry of a node is the same at every step, and the
// velocity gradient of a step is linear in the node's sensitivity
// H = -c g dI/dS.  So
//   prepare (once):  each interior node's stencil base, fractions and
//                    clamped axes -> a record (no phi);
//   step (per step): the transposed gather of lam_next into lam from the
//                    record, and hsum[node] += H (the node's own entries, no
//                    atomics, no velocity reads);
//   finish (once):   dL/dvel += the MAC / node-velocity scatter of hsum, the
//                    same scatter as sl_adjoint_k's per step.
// The record holds {bx, by, bz, flags} (int4, flags: bit a = axis a clamped,
// bit 3 = interior node) and the three fractions, from which the weights are
// formed exactly as the forward forms them: in double (a double4), except
// for the FP32 tricubic, whose weights start from the fractions rounded to
// float (a float4: the record is sized for the double4 and this uses half
// of its fraction area).
constexpr size_t REC_BYTES = sizeof(int4) + sizeof(double4);
template <typename T, int INTERP>
using Frac4 = typename std::conditional<sizeof(T) == 4 && INTERP == DCL_CUBIC, float4, double4>::type;

template <typename T, int INTERP, int BT>
__global__ __launch_bounds__(256) void adj_prepare_k(const T* __restrict__ vel, AdjP P, int4* __restrict__ rb,
                                                     void* __restrict__ rfv) {
  // the fraction record's type depends on INTERP; it is passed untyped so the
  // kernel's mangled name holds no expression over DCL_CUBIC (an enumerator
  // of an unnamed enum, whose internal name the host and device compilations
  // number differently)
  Frac4<T, INTERP>* __restrict__ rf = static_cast<Frac4<T, INTERP>*>(rfv);
  const int64_t n = (int64_t)P.w.nx * P.w.ny * P.w.nz;
  for (int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x; id < n; id += (int64_t)gridDim.x * 256) {
    const int i = (int)(id % P.w.nx), j = (int)((id / P.w.nx) % P.w.ny), k = (int)(id / ((int64_t)P.w.nx * P.w.ny));
    const bool inner = i >= 2 && i <= P.w.nx - 3 && j >= 2 && j <= P.w.ny - 3 && k >= 2 && k <= P.w.nz - 3;
    int4 q = make_int4(0, 0, 0, 0);
    Frac4<T, INTERP> f4 = {0, 0, 0, 0};
    if (inner) {
      const int X = P.w.px + i, Yg = P.w.py + j, Zg = P.w.pz + k;
      const double Xd[3] = {(double)X, (double)Yg, (double)Zg};
      T V1[3];
      node_vel<T, 3>(vel, P.v, X, Yg, Zg, V1);
      double S[3];
      if (BT == DCL_RK2) {
        double Xm[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) Xm[a] = Xd[a] - (double)V1[a] * P.ch[a];
        T V2[3];
        mac_vel<T, 3>(vel, P.v, Xm, V2);
#pragma unroll
        for (int a = 0; a < 3; ++a) S[a] = Xd[a] - (double)V2[a] * P.c[a];
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) S[a] = Xd[a] - (double)V1[a] * P.c[a];
      }
      const int pp[3] = {P.w.px, P.w.py, P.w.pz}, nn[3] = {P.w.nx, P.w.ny, P.w.nz};
      int b[3], flags = 8;
      double fr[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double f;
        if (INTERP == DCL_CUBIC) {
          const double lo = pp[a] + 1, hi = pp[a] + nn[a] - 2;
          if (!(S[a] >= lo && S[a] <= hi)) flags |= 1 << a;
          b[a] = cub_split(S[a], pp[a], nn[a], f);
        } else {
          const double lo = pp[a], hi = pp[a] + nn[a] - 1;
          if (!(S[a] >= lo && S[a] <= hi)) flags |= 1 << a;
          b[a] = lin_split(S[a], pp[a], nn[a], f);
        }
        fr[a] = f;
      }
      q = make_int4(b[0], b[1], b[2], flags);
      f4.x = fr[0];
      f4.y = fr[1];
      f4.z = fr[2];
    }
    rb[id] = q;
    rf[id] = f4;
  }
}

// One adjoint step from the record.  A block is one wave: its 16 x 4 lanes
// own a 16 x 4 column of nodes and walk AWZ planes of it, scattering into an
// LDS box of lam (the column plus the reach WR) that no other wave touches.
// So the scatter is a plain LDS read-add-write per lane: a float atomic in
// LDS costs about 3 cycles per active lane (a 64-lane ds_add_f32 takes ~190
// cycles of a CU's LDS, a ds_write 5), and that was this step's cost.  Two
// lanes of one instruction may write the same word only when their stencil
// corners coincide (departures that converge); such lanes are found per
// plane with a tag table and take LDS atomics instead.  x-neighbour lanes
// whose corners are 1..3 words apart combine their shared words with DPP row
// shifts before writing (as in sl_adjoint_k).  Stencils that leave the box
// (departures of 2 cells or more) add into lam with HBM atomics.  The box and
// tag accesses are volatile: the lanes read words other lanes wrote, which
// the compiler must not forward from a lane's own earlier store.
template <typename T, int INTERP>
__global__ __launch_bounds__(64) void adj_step_k(const T* __restrict__ phi0, const int4* __restrict__ rb,
                                                 const void* __restrict__ rfv,
                                                 const T* __restrict__ lam_next, T* __restrict__ lam,
                                                 T* __restrict__ hsum, AdjP P) {
  constexpr int NP = INTERP == DCL_CUBIC ? 4 : 2;
  constexpr int OFF = INTERP == DCL_CUBIC ? -1 : 0;
  const Frac4<T, INTERP>* __restrict__ rf = static_cast<const Frac4<T, INTERP>*>(rfv);
  __shared__ T box_s[WBOX];
  __shared__ int tag_s[NTAG];
  typedef __attribute__((address_space(3))) T lds_T;
  typedef __attribute__((address_space(3))) int lds_int;
  volatile lds_T* box = (volatile lds_T*)box_s;
  volatile lds_int* tag = (volatile lds_int*)tag_s;
  const int lane = threadIdx.x;
  for (int t = lane; t < WBOX; t += 64) box[t] = T(0);

  const int ox = blockIdx.x * ABX, oy = blockIdx.y * ABY, oz = blockIdx.z * AWZ;
  const int i = ox + (lane & (ABX - 1)), j = oy + lane / ABX;
  const int nx = P.w.nx, ny = P.w.ny, nz = P.w.nz;
  const int Y = nx, Z = nx * ny;
  const rsrc_t rp = make_rsrc(phi0, (unsigned)((int64_t)nx * ny * nz * (int64_t)sizeof(T)));
  // a node's record, lam_next and sensitivity sum so far, loaded together,
  // the next plane's while this plane's stencils are worked
  auto inner_at = [&](int k) {
    return i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3 && k >= 2 && k <= nz - 3;
  };
  T ng = T(0), nh[3] = {T(0), T(0), T(0)};
  int4 nq = make_int4(0, 0, 0, 0);
  Frac4<T, INTERP> nf = {0, 0, 0, 0};
  auto load_node = [&](int k) {
    ng = T(0);
    nh[0] = nh[1] = nh[2] = T(0);
    nq = make_int4(0, 0, 0, 0);
    nf = Frac4<T, INTERP>{0, 0, 0, 0};
    if (inner_at(k)) {
      const int id = k * Z + j * Y + i;
      ng = lam_next[id];
      nq = rb[id];
      nf = rf[id];
      const T* hs = hsum + 3 * (int64_t)id;
#pragma unroll
      for (int a = 0; a < 3; ++a) nh[a] = hs[a];
    }
  };
  load_node(oz);
#pragma unroll 1
  for (int kz = 0; kz < AWZ; ++kz) {
    const int k = oz + kz;
    const bool valid = i < nx && j < ny && k < nz;
    const bool inner = valid && inner_at(k);
    const int id = k * Z + j * Y + i;
    if (valid && !inner) {
      const T gb = lam_next[id];
      if (gb != T(0)) atomic_add(&lam[id], gb);
    }
    const T g = ng, h0[3] = {nh[0], nh[1], nh[2]};
    const int4 q = nq;
    const Frac4<T, INTERP> f4 = nf;
    if (kz + 1 < AWZ) load_node(k + 1);
    T* hs = hsum + 3 * (int64_t)id;
    const bool act = inner && g != T(0);
    const int b[3] = {q.x, q.y, q.z};
    T w[3][4], dw[3][4];
    const double fr[3] = {(double)f4.x, (double)f4.y, (double)f4.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (act) {
        if (INTERP == DCL_CUBIC) {
          cr_weights<T>((T)fr[a], w[a]);
          cr_dweights<T>((T)fr[a], dw[a]);
        } else {
          w[a][0] = (T)(1.0 - fr[a]);
          w[a][1] = (T)fr[a];
          dw[a][0] = T(-1);
          dw[a][1] = T(1);
        }
      } else {
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) w[a][qq] = dw[a][qq] = T(0);
      }
    }
    const int bx0 = b[0] + OFF - (ox - WR), by0 = b[1] + OFF - (oy - WR), bz0 = b[2] + OFF - (oz - WR);
    const bool in_box =
        act && bx0 >= 0 && bx0 + NP <= WBX && by0 >= 0 && by0 + NP <= WBY && bz0 >= 0 && bz0 + NP <= WBZ;
    const int key = in_box ? (bz0 * WBY + by0) * WBX + bx0 + 1 : 0;
    // lanes sharing a stencil corner: each tags its slot, lanes that lose
    // mark the slot, and every lane whose slot ends marked takes atomics
    const int slot = key & (NTAG - 1);
    if (in_box) tag[slot] = lane;
    const bool lost = in_box && tag[slot] != lane;
    if (lost) tag[slot] = -1;
    const bool shared = in_box && tag[slot] == -1;
    const bool plain = in_box && !shared;
    bool lo_m[4] = {false, false, false, false}, up_m[4] = {false, false, false, false};
#pragma unroll
    for (int d = 1; d < NP; ++d) {
      const int kl = lane_from_lower(key, d), ku = lane_from_upper(key, d);
      lo_m[d] = key != 0 && kl != 0 && kl + d == key;
      up_m[d] = key != 0 && ku != 0 && key + d == ku;
    }
    T G[3] = {T(0), T(0), T(0)};
    // the stencil rows with buffer loads (a row of 4 values is one 16-byte
    // load in FP32, two in FP64; inactive lanes read the window's first row),
    // all stencil planes in flight at once in FP32, one at a time in FP64
    const int vo = act ? ((b[2] + OFF) * Z + (b[1] + OFF) * Y + (b[0] + OFF)) * (int)sizeof(T) : 0;
    constexpr int KP = sizeof(T) == 4 ? NP : 1;
#pragma unroll 1
    for (int k0 = 0; k0 < NP; k0 += KP) {
      T rows[KP][4][4];
#pragma unroll
      for (int dk = 0; dk < KP; ++dk)
#pragma unroll
        for (int jj = 0; jj < NP; ++jj) {
          const int so = ((k0 + dk) * Z + jj * Y) * (int)sizeof(T);
          if (NP == 4) BL<T>::four(rp, vo, so, rows[dk][jj]);
          else BL<T>::two(rp, vo, so, rows[dk][jj]);
        }
#pragma unroll
      for (int dk = 0; dk < KP; ++dk)
#pragma unroll
        for (int jj = 0; jj < NP; ++jj) {
          const int kk = k0 + dk;
          const T wyz = w[1][jj] * w[2][kk];
          T add[4];
#pragma unroll
          for (int ii = 0; ii < NP; ++ii) add[ii] = g * w[0][ii] * wyz;
          T word0 = add[0];
#pragma unroll
          for (int d = 1; d < NP; ++d) {
            const T o = lane_from_lower(add[d], d);
            if (lo_m[d]) word0 += o;
          }
          if (act) {
            T rx = T(0), rd = T(0);
#pragma unroll
            for (int ii = 0; ii < NP; ++ii) {
              const T f = rows[dk][jj][ii];
              rx += w[0][ii] * f;
              rd += dw[0][ii] * f;
            }
            G[0] += rd * wyz;
            G[1] += rx * dw[1][jj] * w[2][kk];
            G[2] += rx * w[1][jj] * dw[2][kk];
          }
          const int wo = ((bz0 + kk) * WBY + (by0 + jj)) * WBX + bx0;
#pragma unroll
          for (int ii = 0; ii < NP; ++ii) {
            const T v = ii == 0 ? word0 : add[ii];
            const bool mine = ii == 0 || !up_m[ii];
            if (plain && mine) box[wo + ii] = box[wo + ii] + v;
            if (shared && mine) atomicAdd(&box_s[wo + ii], v);
            if (act && !in_box)
              atomic_add(&lam[(b[2] + OFF + kk) * Z + (b[1] + OFF + jj) * Y + (b[0] + OFF + ii)], add[ii]);
          }
        }
    }
    if (act) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (!(q.w & (1 << a))) hs[a] = h0[a] + (T)(-P.c[a]) * g * G[a];
    }
  }
  for (int t = lane; t < WBOX; t += 64) {
    const T v = box[t];
    if (v == T(0)) continue;
    const int lx = t % WBX, ly = (t / WBX) % WBY, lz = t / (WBX * WBY);
    const int gx = ox - WR + lx, gy = oy - WR + ly, gz = oz - WR + lz;
    if (gx < 0 || gx >= nx || gy < 0 || gy >= ny || gz < 0 || gz >= nz) continue;
    atomic_add(&lam[gz * Z + gy * Y + gx], v);
  }
}

// dL/dvel += the velocity scatter of the accumulated sensitivities H = hsum
// (the tail of sl_adjoint_k: the MAC corners at the midpoint and, through the
// midpoint's dependence on the node velocity, the node's faces), 16 x 4 x 4
// blocks for the same lane combining.
template <typename T, int BT>
__global__ __launch_bounds__(256) void adj_finish_k(const T* __restrict__ vel, const T* __restrict__ hsum,
                                                    T* __restrict__ gvel, AdjP P) {
  const int tid = threadIdx.x + ABX * (threadIdx.y + ABY * threadIdx.z);
  const int i = blockIdx.x * ABX + threadIdx.x, j = blockIdx.y * ABY + threadIdx.y, k = blockIdx.z * ABZ + threadIdx.z;
  const int nx = P.w.nx, ny = P.w.ny, nz = P.w.nz;
  const bool inner = i < nx && j < ny && k < nz && i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3 && k >= 2 &&
                     k <= nz - 3;
  T H[3] = {T(0), T(0), T(0)};
  if (inner) {
    const T* hs = hsum + 3 * (int64_t)((k * ny + j) * nx + i);
    H[0] = hs[0];
    H[1] = hs[1];
    H[2] = hs[2];
  }
  const bool act = inner && (H[0] != T(0) || H[1] != T(0) || H[2] != T(0));
  T H1[3] = {T(0), T(0), T(0)};
  MacState<T> m2;
  if (act) {
    if (BT == DCL_RK2) {
      const int X = P.w.px + i, Yg = P.w.py + j, Zg = P.w.pz + k;
      const double Xd[3] = {(double)X, (double)Yg, (double)Zg};
      T V1[3];
      node_vel<T, 3>(vel, P.v, X, Yg, Zg, V1);
      double Xm[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) Xm[a] = Xd[a] - (double)V1[a] * P.ch[a];
      bool cm[3] = {false, false, false};
      mac_state<T>(vel, P.v, Xm, m2, cm);
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const T K = H[0] * m2.dV[0][bb] + H[1] * m2.dV[1][bb] + H[2] * m2.dV[2][bb];
        H1[bb] = (T)(-P.ch[bb]) * K;
      }
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) H1[a] = H[a];
    }
  }
  if (BT == DCL_RK2) mac_scatter_lanes<T>(gvel, P.v, m2, H, act);
  const int sx = P.v.dof, sy = P.w.nx * P.v.dof, sz = P.w.ny * sy;
  const int gofs = (k * P.w.ny + j) * sy + i * sx + P.v.udof;
  const T h0 = T(0.5) * H1[0], h1 = T(0.5) * H1[1], h2 = T(0.5) * H1[2];
  const int nkey = act ? gofs + 1 : 0;
  const int nkl = lane_from_lower(nkey, 1), nku = lane_from_upper(nkey, 1);
  const bool nlo = act && nkl != 0 && nkl + sx == nkey, nup = act && nku != 0 && nkey + sx == nku;
  const T h0l = lane_from_lower(h0, 1);
  const int lane = tid & 63;
  const int ykl = __shfl_up(nkey, ABX), yku = __shfl_down(nkey, ABX);
  const bool ylo = act && lane >= ABX && ykl != 0 && ykl + sy == nkey;
  const bool yup = act && lane < 64 - ABX && yku != 0 && nkey + sy == yku;
  const T h1l = __shfl_up(h1, ABX);
  if (act) {
    T* gb = gvel + gofs;
    atomic_add(gb, nlo ? h0 + h0l : h0);
    if (!nup) atomic_add(gb + sx, h0);
    atomic_add(gb + 1, ylo ? h1 + h1l : h1);
    if (!yup) atomic_add(gb + sy + 1, h1);
    atomic_add(gb + 2, h2);
    atomic_add(gb + sz + 2, h2);
  }
}

template <typename T, int INTERP, int BT>
static int launch_adj(const void* phi0, const void* vel, const void* lam_next, void* lam, void* gvel,
                      const AdjP& P, hipStream_t s) {
  dim3 grid((P.w.nx + ABX - 1) / ABX, (P.w.ny + ABY - 1) / ABY, (P.w.nz + AWZ - 1) / AWZ);
  hipLaunchKernelGGL((sl_adjoint_k<T, INTERP, BT>), grid, dim3(64), 0, s,
                     static_cast<const T*>(phi0), static_cast<const T*>(vel),
                     static_cast<const T*>(lam_next), static_cast<T*>(lam), static_cast<T*>(gvel), P);
  DCL_LAUNCHED("sl_adjoint");
  return 0;
}

template <typename T>
static int dispatch_adj(int interp, int bt, const void* phi0, const void* vel, const void* ln, void* lam,
                        void* gv, const AdjP& P, hipStream_t s) {
  if (interp == DCL_CUBIC && bt == DCL_RK2) return launch_adj<T, DCL_CUBIC, DCL_RK2>(phi0, vel, ln, lam, gv, P, s);
  if (interp == DCL_CUBIC) return launch_adj<T, DCL_CUBIC, DCL_EULER>(phi0, vel, ln, lam, gv, P, s);
  if (bt == DCL_RK2) return launch_adj<T, DCL_LINEAR, DCL_RK2>(phi0, vel, ln, lam, gv, P, s);
  return launch_adj<T, DCL_LINEAR, DCL_EULER>(phi0, vel, ln, lam, gv, P, s);
}

}  // namespace dcl

using namespace dcl;

static int adj_params(const dcl_window* win, const dcl_velocity* vel, double dt, AdjP& P) {
  DCL_TRY(check_window(win, "phi", 5));
  if (win->n.z < 2) return set_error(DCL_ERR_SUP, "the adjoint is 3D only");
  DCL_TRY(check_velocity(vel, 3));
  const dcl_window& vw = vel->win;
  if (vw.p.x != win->p.x || vw.p.y != win->p.y || vw.p.z != win->p.z || vw.n.x != win->n.x ||
      vw.n.y != win->n.y || vw.n.z != win->n.z)
    return set_error(DCL_ERR_SUP, "the adjoint needs the velocity window to equal phi's");
  P.w = kwin(*win);
  P.v = kvel(*vel);
  const double d[3] = {win->d.x, win->d.y, win->d.z};
  for (int a = 0; a < 3; ++a) {
    P.c[a] = dt / d[a];
    P.ch[a] = dt / (2.0 * d[a]);
  }
  return 0;
}

static int check_scheme(int interp, int backtrace) {
  if (interp != DCL_LINEAR && interp != DCL_CUBIC) return set_error(DCL_ERR_ARG_WRONG, "bad interp");
  if (backtrace != DCL_EULER && backtrace != DCL_RK2) return set_error(DCL_ERR_ARG_WRONG, "bad backtrace");
  return 0;
}

extern "C" int dcl_sl_adjoint_record_bytes(const dcl_window* win, int64_t* bytes) {
  if (!bytes) return set_error(DCL_ERR_ARG_NULL, "bytes is NULL");
  DCL_TRY(check_window(win, "phi", 5));
  *bytes = win_nodes(*win) * (int64_t)REC_BYTES;
  return 0;
}

template <typename T>
static int adj_prepare_t(int interp, int bt, const void* vel, const AdjP& P, void* record, hipStream_t s) {
  const int64_t n = (int64_t)P.w.nx * P.w.ny * P.w.nz;
  int4* rb = static_cast<int4*>(record);
  const unsigned nb = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  const T* v = static_cast<const T*>(vel);
#define DCL_PREP(I, B)                                                                  \
  hipLaunchKernelGGL((adj_prepare_k<T, I, B>), dim3(nb), dim3(256), 0, s, v, P, rb, static_cast<void*>(rb + n))
  if (interp == DCL_CUBIC && bt == DCL_RK2) DCL_PREP(DCL_CUBIC, DCL_RK2);
  else if (interp == DCL_CUBIC) DCL_PREP(DCL_CUBIC, DCL_EULER);
  else if (bt == DCL_RK2) DCL_PREP(DCL_LINEAR, DCL_RK2);
  else DCL_PREP(DCL_LINEAR, DCL_EULER);
#undef DCL_PREP
  DCL_LAUNCHED("adj_prepare");
  return 0;
}

extern "C" int dcl_sl_adjoint_prepare(const dcl_window* win, const dcl_velocity* vel, double dt, int precision,
                                      int interp, int backtrace, void* record, void* stream) {
  DCL_TRY(check_precision(precision));
  DCL_TRY(check_scheme(interp, backtrace));
  if (!record) return set_error(DCL_ERR_ARG_NULL, "record is NULL");
  AdjP P;
  DCL_TRY(adj_params(win, vel, dt, P));
  hipStream_t s = as_stream(stream);
  if (precision == DCL_F64) return adj_prepare_t<double>(interp, backtrace, vel->data, P, record, s);
  return adj_prepare_t<float>(interp, backtrace, vel->data, P, record, s);
}

template <typename T>
static int adj_step_t(int interp, const void* phi0, const void* record, const void* lam_next, void* lam,
                      void* hsum, const AdjP& P, hipStream_t s) {
  const int64_t n = (int64_t)P.w.nx * P.w.ny * P.w.nz;
  const int4* rb = static_cast<const int4*>(record);
  dim3 grid((P.w.nx + ABX - 1) / ABX, (P.w.ny + ABY - 1) / ABY, (P.w.nz + AWZ - 1) / AWZ);
  const T* p0 = static_cast<const T*>(phi0);
  const T* ln = static_cast<const T*>(lam_next);
  if (interp == DCL_CUBIC)
    hipLaunchKernelGGL((adj_step_k<T, DCL_CUBIC>), grid, dim3(64), 0, s, p0, rb,
                       static_cast<const void*>(rb + n), ln, static_cast<T*>(lam), static_cast<T*>(hsum), P);
  else
    hipLaunchKernelGGL((adj_step_k<T, DCL_LINEAR>), grid, dim3(64), 0, s, p0, rb,
                       static_cast<const void*>(rb + n), ln, static_cast<T*>(lam), static_cast<T*>(hsum), P);
  DCL_LAUNCHED("adj_step");
  return 0;
}

extern "C" int dcl_sl_adjoint_step(const void* phi0, const dcl_window* win, const void* record,
                                   const void* lam_next, void* lam, void* hsum, double dt, int precision,
                                   int interp, void* stream) {
  DCL_TRY(check_precision(precision));
  if (!phi0 || !record || !lam_next || !lam || !hsum) return set_error(DCL_ERR_ARG_NULL, "a buffer is NULL");
  if (interp != DCL_LINEAR && interp != DCL_CUBIC) return set_error(DCL_ERR_ARG_WRONG, "bad interp");
  DCL_TRY(check_window(win, "phi", 5));
  if (win->n.z < 2) return set_error(DCL_ERR_SUP, "the adjoint is 3D only");
  // phi0's stencils are read with buffer loads at signed 32-bit byte offsets
  if (win_nodes(*win) * (precision == DCL_F64 ? 8 : 4) >= ((int64_t)1 << 31))
    return set_error(DCL_ERR_ARG_SIZ, "dcl_sl_adjoint_step takes phi windows below 2 GiB");
  if (lam == lam_next || lam == phi0 || hsum == lam || hsum == lam_next)
    return set_error(DCL_ERR_ARG_WRONG, "lam and hsum must not alias inputs or each other");
  AdjP P;
  P.w = kwin(*win);
  const double d[3] = {win->d.x, win->d.y, win->d.z};
  for (int a = 0; a < 3; ++a) {
    P.c[a] = dt / d[a];
    P.ch[a] = dt / (2.0 * d[a]);
  }
  hipStream_t s = as_stream(stream);
  if (precision == DCL_F64) return adj_step_t<double>(interp, phi0, record, lam_next, lam, hsum, P, s);
  return adj_step_t<float>(interp, phi0, record, lam_next, lam, hsum, P, s);
}

extern "C" int dcl_sl_adjoint_finish(const dcl_window* win, const dcl_velocity* vel, const void* hsum, void* gvel,
                                     double dt, int precision, int backtrace, void* stream) {
  DCL_TRY(check_precision(precision));
  if (backtrace != DCL_EULER && backtrace != DCL_RK2) return set_error(DCL_ERR_ARG_WRONG, "bad backtrace");
  if (!hsum || !gvel) return set_error(DCL_ERR_ARG_NULL, "a buffer is NULL");
  AdjP P;
  DCL_TRY(adj_params(win, vel, dt, P));
  if (gvel == vel->data || gvel == hsum) return set_error(DCL_ERR_ARG_WRONG, "gvel must not alias inputs");
  dim3 grid((P.w.nx + ABX - 1) / ABX, (P.w.ny + ABY - 1) / ABY, (P.w.nz + ABZ - 1) / ABZ);
  hipStream_t s = as_stream(stream);
#define DCL_FIN(T, B)                                                                                      \
  hipLaunchKernelGGL((adj_finish_k<T, B>), grid, dim3(ABX, ABY, ABZ), 0, s, static_cast<const T*>(vel->data), \
                     static_cast<const T*>(hsum), static_cast<T*>(gvel), P)
  if (precision == DCL_F64) {
    if (backtrace == DCL_RK2) DCL_FIN(double, DCL_RK2); else DCL_FIN(double, DCL_EULER);
  } else {
    if (backtrace == DCL_RK2) DCL_FIN(float, DCL_RK2); else DCL_FIN(float, DCL_EULER);
  }
#undef DCL_FIN
  DCL_LAUNCHED("adj_finish");
  return 0;
}

extern "C" int dcl_sl_adjoint(const void* phi0, const dcl_window* win, const dcl_velocity* vel,
                              const void* lam_next, void* lam, void* gvel, double dt, int precision,
                              int interp, int backtrace, void* stream) {
  DCL_TRY(check_precision(precision));
  if (!phi0 || !lam_next || !lam || !gvel) return set_error(DCL_ERR_ARG_NULL, "a buffer is NULL");
  DCL_TRY(check_window(win, "phi", 5));
  if (win->n.z < 2) return set_error(DCL_ERR_SUP, "the adjoint is 3D only");
  DCL_TRY(check_velocity(vel, 3));
  const dcl_window& vw = vel->win;
  if (vw.p.x != win->p.x || vw.p.y != win->p.y || vw.p.z != win->p.z || vw.n.x != win->n.x ||
      vw.n.y != win->n.y || vw.n.z != win->n.z)
    return set_error(DCL_ERR_SUP, "the adjoint needs the velocity window to equal phi's");
  if (interp != DCL_LINEAR && interp != DCL_CUBIC) return set_error(DCL_ERR_ARG_WRONG, "bad interp");
  if (backtrace != DCL_EULER && backtrace != DCL_RK2) return set_error(DCL_ERR_ARG_WRONG, "bad backtrace");
  if (lam == lam_next || lam == phi0) return set_error(DCL_ERR_ARG_WRONG, "lam must not alias inputs");
  AdjP P;
  P.w = kwin(*win);
  P.v = kvel(*vel);
  const double d[3] = {win->d.x, win->d.y, win->d.z};
  for (int a = 0; a < 3; ++a) {
    P.c[a] = dt / d[a];
    P.ch[a] = dt / (2.0 * d[a]);
  }
  hipStream_t s = as_stream(stream);
  if (precision == DCL_F64)
    return dispatch_adj<double>(interp, backtrace, phi0, vel->data, lam_next, lam, gvel, P, s);
  return dispatch_adj<float>(interp, backtrace, phi0, vel->data, lam_next, lam, gvel, P, s);
}
