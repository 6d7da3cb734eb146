// reinit_sweep.hip -- the parallel reinitialisation on gfx950
// (dcl_reinitialize_parallel): LevelSetReinitialize
// (LevelSetMethod/FastMarching.c:10-146) computed by sweeps over the whole
// window instead of one accepted node at a time.
//
// The boundary values are the fast marching's (the distance to the
// orthogonal projection at every projection record) and so is the update:
// AccumulateCoef (FastMarching2D.c:33-94) over the axes, then the larger
// root.  What the fast marching decides by its order -- which neighbours
// are already accepted when a node is pushed, and which push's value the
// node keeps -- a sweep decides from the neighbours' values: each push is
// replayed in the order of the pushing neighbours' values (the boundary
// nodes' pushes first), with the neighbours accepted by then, and the node
// keeps the least value pushed before its own value comes up.  Every sweep
// recomputes every node from the previous sweep's values (two buffers), so
// the result does not depend on scheduling; sweeps end when no node moves by
// more than band_width * 1e-9.  Nodes whose order is decided by their last
// bits can trade values in a cycle: once the count of moving nodes has
// fallen and then stops falling, a node's value may only go down, which ends
// every cycle.
//
// Work skips what cannot change: the window is cut into tiles of 64 x 2 x 2
// nodes (2D: 64 x 4), and a tile is swept only if it or a tile next to it
// changed in the previous sweep (the update reads two nodes along each axis,
// which never reach past the next tile).  A skipped tile's output buffer
// already holds its values: they did not change in the previous sweep.
#include <math.h>
#include <string.h>

#include "dcl_host.hpp"
#include "dcl_reinit.hpp"

namespace dcl {

constexpr int RS_B = 256;
constexpr double RS_TOL = 1e-9;  // settled: no node moves by more than band_width * RS_TOL

template <int DIM>
struct RsTile {
  static constexpr int X = 64, Y = DIM == 3 ? 2 : 4, Z = DIM == 3 ? 2 : 1;
};

struct RsCell {
  double* buf[2];       // buf[0]: the caller's phi, buf[1]: scratch
  unsigned char* mask;  // 0: a - side node, 1: a + side node, 2: a boundary node
  KWin w;
  int d3;
  double bw, inf, tol;
  int tx, ty, tz;  // tiles per axis
  long long tile0;  // the level set's first tile in the batch
  const dcl_irregular_node* nodes;
  int64_t count;
};

struct RsState {
  int k;        // sweeps done
  int settle;   // values may only go down
  int falling;  // the count of moving nodes has fallen
  int done;     // 1 settled, 2 NaN update, 3 out of sweeps
  long long last;          // the previous sweep's count, -1 before the first
  unsigned long long cnt;  // nodes moved by more than tol in the current sweep
  int nan;
  int pad;
};

enum { RS_ACTIVE = 0, RS_SETTLED = 1, RS_NAN = 2, RS_SWEEPS = 3 };

// FMM_InitializeEikonal (FastMarching.c:46-58): phi to +-PHI_INF; the mask
// keeps each node's side
__global__ __launch_bounds__(RS_B) void rs_reset_k(double* __restrict__ phi, unsigned char* __restrict__ mask,
                                                   int64_t n, double inf) {
  for (int64_t i = (int64_t)blockIdx.x * RS_B + threadIdx.x; i < n; i += (int64_t)gridDim.x * RS_B) {
    const bool pos = phi[i] > 0;
    mask[i] = pos ? 1 : 0;
    phi[i] = pos ? inf : -inf;
  }
}

// FMM_InitializeEikonal (FastMarching.c:60-74): the boundary values
__global__ __launch_bounds__(RS_B) void rs_bc_k(RsCell c) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * RS_B + threadIdx.x;
  if (i >= c.count) return;
  const dcl_irregular_node n = c.nodes[i];
  if (n.axis != -1 || !ri_inside(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z)) return;
  const double r = c.d3 ? sqrt(n.op.x * n.op.x + n.op.y * n.op.y + n.op.z * n.op.z)
                         : sqrt(n.op.x * n.op.x + n.op.y * n.op.y);
  const int64_t j = ri_index(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z);
  c.buf[0][j] = n.sign_center * r;
  c.mask[j] = 2;
}

// a neighbour as the side `sign` sees it: its value q = sign * phi, the time
// it is accepted (-1: before the side's pops, its value for a node of this
// side, INFINITY: never) and whether it pushes its neighbours in this side
struct RsNb {
  double q, t;
  bool push;
};

__device__ __forceinline__ RsNb rs_view(const RsCell& c, const double* __restrict__ in, int sign, int x, int y,
                                        int z) {
  RsNb r;
  r.q = c.inf;
  r.t = INFINITY;
  r.push = false;
  if (!ri_inside(c.w, c.d3, x, y, z)) return r;
  const int64_t i = ri_index(c.w, c.d3, x, y, z);
  const double q = sign * in[i];
  const int m = c.mask[i];
  r.q = q;
  if (m == 2) {
    // a boundary node pushes in this side when sign * phi >= 0
    if (0 <= q && q < c.inf) {
      r.t = -1;
      r.push = true;
    }
  } else if (m == (sign > 0 ? 1 : 0)) {
    if (q < c.inf) {
      r.t = q;
      r.push = true;
    }
  } else if (sign > 0 && 0 <= q && q < c.inf) {
    r.t = -1;  // a - side node: final before the + side starts
  }
  return r;
}

// the node's value from its neighbours' (NaN for a NaN update)
template <int DIM>
__device__ __forceinline__ double rs_node(const RsCell& c, const double* __restrict__ in, int sign, int X, int Y,
                                          int Z) {
#pragma clang fp contract(off)
  RsNb nb[DIM][4];  // per axis: -2, -1, +1, +2
#pragma unroll
  for (int a = 0; a < DIM; ++a)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int s = k < 2 ? k - 2 : k - 1;
      nb[a][k] = rs_view(c, in, sign, X + (a == 0 ? s : 0), Y + (a == 1 ? s : 0), Z + (a == 2 ? s : 0));
    }
  double cand = INFINITY, last = -INFINITY;
  // at most 2 * DIM distinct push times, then one pass that finds none
  for (int guard = 0; guard <= 2 * DIM; ++guard) {
    double T = INFINITY;
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int k = 1; k <= 2; ++k)
        if (nb[a][k].push && nb[a][k].t > last && nb[a][k].t < T) T = nb[a][k].t;
    if (T == INFINITY || !(T < cand)) break;
    double A = 0, B = 0, C = -1, loc[7];
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      loc[0] = loc[6] = loc[3] = c.inf;
      loc[1] = nb[a][0].t <= T ? nb[a][0].q : c.inf;
      loc[2] = nb[a][1].t <= T ? nb[a][1].q : c.inf;
      loc[4] = nb[a][2].t <= T ? nb[a][2].q : c.inf;
      loc[5] = nb[a][3].t <= T ? nb[a][3].q : c.inf;
      ri_accumulate(c.inf, loc, A, B, C);
    }
    const double disc = sqrt(B * B - 4 * A * C);
    double d1 = (-B + disc) / (2 * A);
    if (d1 != d1) d1 = -B / (2 * A);
    if (d1 != d1) return NAN;
    if (d1 < cand) cand = d1;
    last = T;
  }
  return cand <= c.bw ? sign * cand : sign * c.inf;
}

// one sweep of every level set still sweeping: one block per tile
template <int DIM>
__global__ __launch_bounds__(RS_B) void rs_sweep_k(const RsCell* __restrict__ cells, int ncells,
                                                   RsState* __restrict__ st, unsigned char* __restrict__ chg0,
                                                   unsigned char* __restrict__ chg1) {
#pragma clang fp contract(off)
  using TL = RsTile<DIM>;
  const long long g = blockIdx.x;
  int lo = 0, hi = ncells - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cells[mid].tile0 <= g) lo = mid;
    else hi = mid - 1;
  }
  const RsCell c = cells[lo];
  RsState* s = st + lo;
  if (s->done) return;
  const int k = s->k, settle = s->settle;
  const unsigned char* prev = (k & 1) ? chg1 : chg0;  // flags of sweep k
  unsigned char* cur = (k & 1) ? chg0 : chg1;         // flags of sweep k + 1
  const int tid = threadIdx.x;
  const long long t = g - c.tile0;
  const int ti = (int)(t % c.tx), tj = (int)((t / c.tx) % c.ty), tk = (int)(t / ((long long)c.tx * c.ty));
  bool act = k == 0;
  if (!act && tid < (DIM == 3 ? 27 : 9)) {
    const int a = ti + tid % 3 - 1, b = tj + (tid / 3) % 3 - 1, e = DIM == 3 ? tk + tid / 9 - 1 : 0;
    if (a >= 0 && a < c.tx && b >= 0 && b < c.ty && e >= 0 && e < c.tz)
      act = prev[c.tile0 + ((long long)e * c.ty + b) * c.tx + a] != 0;
  }
  if (!__syncthreads_or(act)) {
    if (tid == 0) cur[g] = 0;
    return;
  }
  const double* __restrict__ in = (k & 1) ? c.buf[1] : c.buf[0];
  double* __restrict__ out = (k & 1) ? c.buf[0] : c.buf[1];
  const int lx = ti * TL::X + (tid & 63);
  const int ly = tj * TL::Y + (tid >> 6) % TL::Y;
  const int lz = DIM == 3 ? tk * TL::Z + (tid >> 6) / TL::Y : 0;
  bool moved = false, flip = false, nan = false;
  if (lx < c.w.nx && ly < c.w.ny && lz < c.w.nz) {
    const int64_t i = ((int64_t)lz * c.w.ny + ly) * c.w.nx + lx;
    const double old = in[i];
    double v = old;
    const int m = c.mask[i];
    if (m != 2) {
      const int sg = m ? 1 : -1;
      const double u = rs_node<DIM>(c, in, sg, c.w.px + lx, c.w.py + ly, c.w.pz + lz);
      nan = u != u;
      if (!settle || sg * u < sg * v) v = u;
    }
    moved = fabs(v - old) > c.tol;
    flip = __double_as_longlong(v) != __double_as_longlong(old);
    out[i] = v;
  }
  const int nm = __syncthreads_count(moved);
  const int nf = __syncthreads_or(flip);
  const int nn = __syncthreads_or(nan);
  if (tid == 0) {
    cur[g] = nf ? 1 : 0;
    if (nm) atomicAdd(&s->cnt, (unsigned long long)nm);
    if (nn) atomicOr(&s->nan, 1);
  }
}

// after a sweep: each level set's count decides whether it has settled and
// whether its values may only go down from the next sweep on
__global__ __launch_bounds__(RS_B) void rs_state_k(RsState* __restrict__ st, int ncells, int max_sweeps,
                                                   int* __restrict__ active) {
  const int i = blockIdx.x * RS_B + threadIdx.x;
  if (i >= ncells) return;
  RsState s = st[i];
  if (s.done) return;
  const unsigned long long c = s.cnt;
  s.cnt = 0;
  s.k++;
  if (s.nan) {
    s.done = RS_NAN;
  } else if (c == 0) {
    s.done = RS_SETTLED;
  } else {
    if (s.last >= 0 && (long long)c < s.last) s.falling = 1;
    else if (s.last >= 0 && s.falling) s.settle = 1;
    s.last = (long long)c;
    if (s.k >= max_sweeps) s.done = RS_SWEEPS;
  }
  st[i] = s;
  if (!s.done) atomicAdd(active, 1);
}

// the band: boundary nodes in record order, then each side in window order
__global__ __launch_bounds__(RS_B) void rs_bc_count_k(RsCell c, unsigned* __restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * RS_B + threadIdx.x;
  bool p = false;
  if (i < c.count) {
    const dcl_irregular_node& n = c.nodes[i];
    p = n.axis == -1 && ri_inside(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z);
  }
  const int cnt = __syncthreads_count(p);
  if (threadIdx.x == 0) counts[blockIdx.x] = (unsigned)cnt;
}

// the exclusive rank of a true predicate within the block (RS_B threads)
__device__ __forceinline__ unsigned rs_rank(bool p, unsigned* sh) {
  const unsigned long long bal = __ballot(p);
  const unsigned lane_pre =
      __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[wv] = (unsigned)__popcll(bal);
  __syncthreads();
  unsigned pre = 0;
  for (int k = 0; k < wv; ++k) pre += sh[k];
  __syncthreads();
  return pre + lane_pre;
}

__global__ __launch_bounds__(RS_B) void rs_bc_write_k(RsCell c, const unsigned* __restrict__ offsets,
                                                      const unsigned long long* __restrict__ tile_sums,
                                                      dcl_icoor* __restrict__ band, int64_t cap) {
  __shared__ unsigned sh[RS_B / 64];
  const int64_t i = (int64_t)blockIdx.x * RS_B + threadIdx.x;
  bool p = false;
  dcl_icoor q = {0, 0, 0};
  if (i < c.count) {
    const dcl_irregular_node& n = c.nodes[i];
    p = n.axis == -1 && ri_inside(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z);
    q = n.pos;
  }
  const unsigned r = rs_rank(p, sh);
  if (!p) return;
  const int64_t pos = (int64_t)(offsets[blockIdx.x] + tile_sums[blockIdx.x / SCAN_TILE]) + r;
  if (pos < cap) band[pos] = q;
}

__global__ __launch_bounds__(RS_B) void rs_side_count_k(const double* __restrict__ phi,
                                                        const unsigned char* __restrict__ mask, int64_t n, double inf,
                                                        unsigned* __restrict__ c0, unsigned* __restrict__ c1) {
  const int64_t i = (int64_t)blockIdx.x * RS_B + threadIdx.x;
  const bool in = i < n && fabs(phi[i]) < inf;
  const int m = i < n ? mask[i] : 2;
  const int n0 = __syncthreads_count(in && m == 0);
  const int n1 = __syncthreads_count(in && m == 1);
  if (threadIdx.x == 0) {
    c0[blockIdx.x] = (unsigned)n0;
    c1[blockIdx.x] = (unsigned)n1;
  }
}

__global__ __launch_bounds__(RS_B) void rs_side_write_k(const double* __restrict__ phi,
                                                        const unsigned char* __restrict__ mask, KWin w, int d3,
                                                        int64_t n, double inf, ScanArrays sa,
                                                        const unsigned long long* __restrict__ nbc,
                                                        dcl_icoor* __restrict__ band, int64_t cap) {
  __shared__ unsigned sh[RS_B / 64];
  const int64_t i = (int64_t)blockIdx.x * RS_B + threadIdx.x;
  const bool in = i < n && fabs(phi[i]) < inf;
  const int m = i < n ? mask[i] : 2;
  const unsigned r0 = rs_rank(in && m == 0, sh);
  const unsigned r1 = rs_rank(in && m == 1, sh);
  if (!in || m == 2) return;
  const int64_t base = (int64_t)*nbc + (m ? (int64_t)*sa.total[0] : 0);
  const unsigned* cnt = m ? sa.counts[1] : sa.counts[0];
  const unsigned long long* sums = m ? sa.sums[1] : sa.sums[0];
  const int64_t pos = base + (int64_t)(cnt[blockIdx.x] + sums[blockIdx.x / SCAN_TILE]) + (m ? r1 : r0);
  if (pos >= cap) return;
  const int64_t plane = (int64_t)w.nx * w.ny;
  const int z = d3 ? (int)(i / plane) : 0, rr = (int)(i - (int64_t)z * plane);
  const int y = rr / w.nx, x = rr - y * w.nx;
  dcl_icoor q;
  q.x = w.px + x;
  q.y = w.py + y;
  q.z = w.pz + z;
  band[pos] = q;
}

static size_t rs_al(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_reinitialize_parallel(dcl_reinit_cell* cells, int ncells, int max_sweeps, int* sweeps,
                                         void* stream) {
  if (ncells < 0) return set_error(DCL_ERR_ARG_SIZ, "ncells < 0");
  if (ncells > 0 && !cells) return set_error(DCL_ERR_ARG_NULL, "cells is NULL");
  if (max_sweeps < 0) return set_error(DCL_ERR_ARG_SIZ, "max_sweeps < 0");
  if (ncells == 0) return 0;
  if (ncells > (1 << 20)) return set_error(DCL_ERR_ARG_SIZ, "at most 2^20 level sets per call");
  if (max_sweeps == 0) max_sweeps = 256;
  hipStream_t s = as_stream(stream);
  int d3 = -1;
  long long tiles = 0;
  int64_t max_nodes = 0, max_count = 0;
  size_t cell_bytes = 0;
  for (int i = 0; i < ncells; ++i) {
    const dcl_reinit_cell& u = cells[i];
    if (!u.phi) return set_error(DCL_ERR_ARG_NULL, "level set %d: phi is NULL", i);
    DCL_TRY(check_window(&u.win, "phi", 1));
    if (u.count < 0 || (u.count > 0 && !u.nodes))
      return set_error(DCL_ERR_ARG_NULL, "level set %d: bad irregular-node list", i);
    if (u.count > INT32_MAX) return set_error(DCL_ERR_ARG_SIZ, "level set %d: more than 2^31 - 1 records", i);
    if (u.band_capacity < 0 || (u.band_capacity > 0 && !u.band))
      return set_error(DCL_ERR_ARG_NULL, "level set %d: bad band output", i);
    if (!(u.band_width > 0) || !(u.phi_inf > u.band_width))
      return set_error(DCL_ERR_ARG_WRONG, "level set %d: need 0 < band_width < phi_inf", i);
    const int c3 = u.win.n.z >= 2;
    if (d3 >= 0 && c3 != d3) return set_error(DCL_ERR_ARG_WRONG, "a batch mixes 2D and 3D level sets");
    d3 = c3;
    const int64_t n = win_nodes(u.win);
    const long long t = d3 ? (long long)((u.win.n.x + 63) / 64) * ((u.win.n.y + 1) / 2) * ((u.win.n.z + 1) / 2)
                           : (long long)((u.win.n.x + 63) / 64) * ((u.win.n.y + 3) / 4);
    tiles += t;
    max_nodes = n > max_nodes ? n : max_nodes;
    max_count = u.count > max_count ? u.count : max_count;
    cell_bytes += rs_al(sizeof(double) * (size_t)n) + rs_al((size_t)n);
  }
  if (tiles > INT32_MAX) return set_error(DCL_ERR_ARG_SIZ, "more than 2^31 - 1 tiles in a batch");
  // scan space for the compactions, reused level set by level set
  const int64_t nbn = (max_nodes + RS_B - 1) / RS_B, nbr = (max_count + RS_B - 1) / RS_B + 1;
  const size_t scan_bytes = 2 * rs_al(sizeof(unsigned) * (size_t)nbn) +
                            2 * rs_al(sizeof(unsigned long long) * (size_t)scan_tiles(nbn)) +
                            rs_al(sizeof(unsigned) * (size_t)nbr) +
                            rs_al(sizeof(unsigned long long) * (size_t)scan_tiles(nbr));
  const size_t bytes = rs_al(sizeof(RsCell) * ncells) + rs_al(sizeof(RsState) * ncells) +
                       rs_al(sizeof(int) * (size_t)(max_sweeps + 1)) + 2 * rs_al((size_t)tiles) +
                       rs_al(sizeof(unsigned long long) * 3 * (size_t)ncells) + scan_bytes + cell_bytes;
  void* sc = nullptr;
  DCL_TRY(scratch(bytes, &sc));
  char* base = static_cast<char*>(sc);
  auto take = [&](size_t b) {
    char* p = base;
    base += rs_al(b);
    return p;
  };
  RsCell* dcells = reinterpret_cast<RsCell*>(take(sizeof(RsCell) * ncells));
  RsState* dstate = reinterpret_cast<RsState*>(take(sizeof(RsState) * ncells));
  int* active = reinterpret_cast<int*>(take(sizeof(int) * (size_t)(max_sweeps + 1)));
  unsigned char* chg0 = reinterpret_cast<unsigned char*>(take((size_t)tiles));
  unsigned char* chg1 = reinterpret_cast<unsigned char*>(take((size_t)tiles));
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 3 * ncells));
  unsigned* sc_c0 = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * (size_t)nbn));
  unsigned* sc_c1 = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * (size_t)nbn));
  unsigned long long* sc_s0 = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * scan_tiles(nbn)));
  unsigned long long* sc_s1 = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * scan_tiles(nbn)));
  unsigned* sc_cr = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * (size_t)nbr));
  unsigned long long* sc_sr = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * scan_tiles(nbr)));

  RsCell* host = static_cast<RsCell*>(malloc(sizeof(RsCell) * ncells));
  RsState* hstate = static_cast<RsState*>(malloc(sizeof(RsState) * ncells));
  unsigned long long* htot = static_cast<unsigned long long*>(malloc(sizeof(unsigned long long) * 3 * ncells));
  if (!host || !hstate || !htot) {
    free(host);
    free(hstate);
    free(htot);
    return set_error(DCL_ERR_MEM, "host memory");
  }
  long long t0 = 0;
  for (int i = 0; i < ncells; ++i) {
    const dcl_reinit_cell& u = cells[i];
    RsCell& c = host[i];
    const int64_t n = win_nodes(u.win);
    c.buf[0] = static_cast<double*>(u.phi);
    c.buf[1] = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n));
    c.mask = reinterpret_cast<unsigned char*>(take((size_t)n));
    c.w = kwin(u.win);
    c.d3 = d3;
    c.bw = u.band_width;
    c.inf = u.phi_inf;
    c.tol = u.band_width * RS_TOL;
    c.tx = (u.win.n.x + 63) / 64;
    c.ty = d3 ? (u.win.n.y + 1) / 2 : (u.win.n.y + 3) / 4;
    c.tz = d3 ? (u.win.n.z + 1) / 2 : 1;
    c.tile0 = t0;
    t0 += (long long)c.tx * c.ty * c.tz;
    c.nodes = u.nodes;
    c.count = u.count;
    RsState& q = hstate[i];
    memset(&q, 0, sizeof(q));
    q.last = -1;
  }
  int rc = 0;
  {
    hipError_t e = hipMemcpyAsync(dcells, host, sizeof(RsCell) * ncells, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dstate, hstate, sizeof(RsState) * ncells, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(active, 0, sizeof(int) * (size_t)(max_sweeps + 1), s);
    if (e != hipSuccess) rc = hip_fail(e, "reinit upload", __FILE__, __LINE__);
  }
  for (int i = 0; i < ncells && !rc; ++i) {
    const RsCell& c = host[i];
    const int64_t n = win_nodes(cells[i].win);
    const unsigned rb = grid1d(n, RS_B) < 4096u ? grid1d(n, RS_B) : 4096u;
    hipLaunchKernelGGL(rs_reset_k, dim3(rb), dim3(RS_B), 0, s, c.buf[0], c.mask, n, c.inf);
    if (c.count > 0) hipLaunchKernelGGL(rs_bc_k, dim3(grid1d(c.count, RS_B)), dim3(RS_B), 0, s, c);
    if (hipGetLastError() != hipSuccess) rc = set_error(DCL_ERR_LIB, "reinit reset / boundary launch failed");
  }
  // the sweeps; the count of level sets still sweeping is read every few
  // sweeps (a level set that has settled skips the sweeps after)
  for (int j = 0; j < max_sweeps && !rc; ++j) {
    if (d3)
      hipLaunchKernelGGL(rs_sweep_k<3>, dim3((unsigned)tiles), dim3(RS_B), 0, s, dcells, ncells, dstate, chg0, chg1);
    else
      hipLaunchKernelGGL(rs_sweep_k<2>, dim3((unsigned)tiles), dim3(RS_B), 0, s, dcells, ncells, dstate, chg0, chg1);
    hipLaunchKernelGGL(rs_state_k, dim3(grid1d(ncells, RS_B)), dim3(RS_B), 0, s, dstate, ncells, max_sweeps,
                       active + j);
    if (hipGetLastError() != hipSuccess) {
      rc = set_error(DCL_ERR_LIB, "reinit sweep launch failed");
      break;
    }
    if ((j & 3) == 3 || j == max_sweeps - 1) {
      int a = 0;
      hipError_t e = hipMemcpyAsync(&a, active + j, sizeof(int), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) rc = hip_fail(e, "reinit sweeps", __FILE__, __LINE__);
      if (a == 0) break;
    }
  }
  if (!rc) {
    hipError_t e = hipMemcpyAsync(hstate, dstate, sizeof(RsState) * ncells, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "reinit sweeps", __FILE__, __LINE__);
  }
  for (int i = 0; i < ncells && !rc; ++i) {
    const RsState& q = hstate[i];
    if (sweeps) sweeps[i] = q.k;
    if (q.done == RS_NAN) rc = set_error(DCL_ERR_LIB, "level set %d: NaN in a reinitialisation update", i);
    else if (q.done != RS_SETTLED)
      rc = set_error(DCL_ERR_LIB, "level set %d: reinitialisation did not settle in %d sweeps", i, max_sweeps);
  }
  // the result is in the buffer the last sweep wrote; then the band
  for (int i = 0; i < ncells && !rc; ++i) {
    const RsCell& c = host[i];
    dcl_reinit_cell& u = cells[i];
    const int64_t n = win_nodes(u.win);
    if (hstate[i].k & 1) {
      hipError_t e = hipMemcpyAsync(c.buf[0], c.buf[1], sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) {
        rc = hip_fail(e, "reinit copy", __FILE__, __LINE__);
        break;
      }
    }
    unsigned long long* tot = totals + 3 * i;
    const int64_t br = (c.count + RS_B - 1) / RS_B;
    if (br > 0) {
      hipLaunchKernelGGL(rs_bc_count_k, dim3((unsigned)br), dim3(RS_B), 0, s, c, sc_cr);
      const ScanArrays sr = {{sc_cr, nullptr}, {sc_sr, nullptr}, {tot, nullptr}};
      rc = scan_counts(sr, 1, br, s);
      if (rc) break;
      if (u.band_capacity > 0)
        hipLaunchKernelGGL(rs_bc_write_k, dim3((unsigned)br), dim3(RS_B), 0, s, c, sc_cr, sc_sr, u.band,
                           u.band_capacity);
    } else {
      // no records: a one-block count of nothing gives the zero total
      hipLaunchKernelGGL(rs_bc_count_k, dim3(1), dim3(RS_B), 0, s, c, sc_cr);
      const ScanArrays sr = {{sc_cr, nullptr}, {sc_sr, nullptr}, {tot, nullptr}};
      rc = scan_counts(sr, 1, 1, s);
      if (rc) break;
    }
    const int64_t bn = (n + RS_B - 1) / RS_B;
    hipLaunchKernelGGL(rs_side_count_k, dim3((unsigned)bn), dim3(RS_B), 0, s, c.buf[0], c.mask, n, c.inf, sc_c0,
                       sc_c1);
    const ScanArrays sn = {{sc_c0, sc_c1}, {sc_s0, sc_s1}, {tot + 1, tot + 2}};
    rc = scan_counts(sn, 2, bn, s);
    if (rc) break;
    if (u.band_capacity > 0)
      hipLaunchKernelGGL(rs_side_write_k, dim3((unsigned)bn), dim3(RS_B), 0, s, c.buf[0], c.mask, c.w, c.d3, n,
                         c.inf, sn, tot, u.band, u.band_capacity);
    if (hipGetLastError() != hipSuccess) rc = set_error(DCL_ERR_LIB, "reinit band launch failed");
  }
  if (!rc) {
    hipError_t e = hipMemcpyAsync(htot, totals, sizeof(unsigned long long) * 3 * ncells, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "reinit band", __FILE__, __LINE__);
  }
  for (int i = 0; i < ncells && !rc; ++i) cells[i].band_count = (int64_t)(htot[3 * i] + htot[3 * i + 1] + htot[3 * i + 2]);
  if (rc) (void)hipStreamSynchronize(s);
  free(host);
  free(hstate);
  free(htot);
  return rc;
}
