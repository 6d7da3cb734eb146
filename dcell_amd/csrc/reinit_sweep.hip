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
// Work skips what cannot change.  A node whose stencil (the node and two
// nodes each way along each axis) held no change in the previous sweep would
// come out as it is, so a change marks the nodes whose stencil holds it, and
// each sweep recomputes only the marked nodes; the output buffer already
// holds the others' values (both buffers start equal, and a node not
// recomputed did not change in the sweep before).  The boundary values are
// the changes before the first sweep.  A sweep is three launches: one lists
// the marked nodes (each chunk of 1024 nodes carries a flag set with its
// marks, so a block reads its chunks' flags at once and scans only the
// marked chunks, four marks per thread, clearing them), one updates the
// listed nodes, one thread each, so that whole waves run the chain of
// updates, and one wave per level set sums the moved nodes and decides.
// Lists are gathered per block in LDS and published with one atomic each:
// an atomic per wave on one counter would queue.
//
// A batch of level sets runs in one launch per step, over all nodes (in
// raster order) and all records, each block finding its level set by a
// binary search.
#include <math.h>
#include <string.h>

#include "dcl_buf.hpp"
#include "dcl_host.hpp"
#include "dcl_reinit.hpp"

namespace dcl {

constexpr int RS_B = 256;
constexpr double RS_TOL = 1e-9;  // settled: no node moves by more than band_width * RS_TOL
constexpr int RS_OUT = 3;        // the mask of a stencil node outside the window
constexpr int RS_CH = 1024;      // nodes per chunk of the mark scan (four per thread)
constexpr int RS_STAGE = 2048;   // list entries a block gathers before publishing them
constexpr int RS_NEED_GRID = 2048;  // blocks of the mark scan
constexpr int RS_GRID = 4096;       // blocks of the update, each taking listed nodes in turn
constexpr int RS_GROUP = 4;         // sweeps queued between looks at the count of level sets still sweeping

struct RsCell {
  double* buf[2];        // buf[0]: the caller's phi, buf[1]: scratch
  unsigned char* mask;   // 0: a - side node, 1: a + side node, 2: a boundary node
  unsigned char* nd[2];  // marks: sweep k + 1 recomputes the nodes marked in nd[k & 1]
  unsigned char* cd[2];  // per chunk of RS_CH nodes: it holds a mark in nd[k & 1]
  KWin w;
  int d3;
  int small;  // below 2^28 nodes: buffer loads
  double bw, inf, tol;
  long long qb0, qb1;  // the level set's raster blocks (RS_B nodes each) in the batch
  long long cb0;       // its first chunk (RS_CH nodes) in the batch
  long long rb0, rb1;  // its record blocks (RS_B records each) in the batch
  const dcl_irregular_node* nodes;
  int64_t count;
  dcl_icoor* band;
  int64_t band_cap;
};

struct RsState {
  int k;        // sweeps done
  int settle;   // values may only go down
  int falling;  // the count of moving nodes has fallen
  int done;     // 1 settled, 2 NaN update, 3 out of sweeps
  long long last;  // the previous sweep's count, -1 before the first
  int nan;
  int pad;
};

enum { RS_ACTIVE = 0, RS_SETTLED = 1, RS_NAN = 2, RS_SWEEPS = 3 };
enum { RS_KEY_RASTER = 0, RS_KEY_RECORD = 1, RS_KEY_CHUNK = 2 };

// the level set of batch block g: the last one whose first block is <= g
template <int KEY>
__device__ __forceinline__ int rs_find(const RsCell* __restrict__ cells, int ncells, long long g) {
  int lo = 0, hi = ncells - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const long long f = KEY == RS_KEY_RASTER ? cells[mid].qb0 : KEY == RS_KEY_RECORD ? cells[mid].rb0 : cells[mid].cb0;
    if (f <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int64_t rs_nodes(const RsCell& c) { return (int64_t)c.w.nx * c.w.ny * c.w.nz; }

// a node changed: the nodes whose stencil holds it (itself and two nodes
// each way along each axis) are recomputed in the next sweep
template <int DIM>
__device__ __forceinline__ void rs_mark(const RsCell& c, unsigned char* nd, unsigned char* cd, int64_t i, int lx,
                                        int ly, int lz) {
  const int64_t sy = c.w.nx, sz = (int64_t)c.w.nx * c.w.ny;
  nd[i] = 1;
  cd[i / RS_CH] = 1;
#pragma unroll
  for (int o = -2; o <= 2; ++o) {
    if (o == 0) continue;
    if (lx + o >= 0 && lx + o < c.w.nx) {
      nd[i + o] = 1;
      cd[(i + o) / RS_CH] = 1;
    }
    if (ly + o >= 0 && ly + o < c.w.ny) {
      nd[i + o * sy] = 1;
      cd[(i + o * sy) / RS_CH] = 1;
    }
    if (DIM == 3 && lz + o >= 0 && lz + o < c.w.nz) {
      nd[i + o * sz] = 1;
      cd[(i + o * sz) / RS_CH] = 1;
    }
  }
}

// FMM_InitializeEikonal (FastMarching.c:46-58): phi to +-PHI_INF in both
// buffers; the mask keeps each node's side
__global__ __launch_bounds__(RS_B) void rs_reset_k(const RsCell* __restrict__ cells, int ncells, long long nq) {
  const long long g = blockIdx.x;
  if (g >= nq) return;
  const RsCell& c = cells[rs_find<RS_KEY_RASTER>(cells, ncells, g)];
  const int64_t i = (g - c.qb0) * RS_B + threadIdx.x;
  if (i >= rs_nodes(c)) return;
  const bool pos = c.buf[0][i] > 0;
  const double v = pos ? c.inf : -c.inf;
  c.mask[i] = pos ? 1 : 0;
  c.buf[0][i] = v;
  c.buf[1][i] = v;
  c.nd[0][i] = 0;
  c.nd[1][i] = 0;
  if (i % RS_CH == 0) {
    c.cd[0][i / RS_CH] = 0;
    c.cd[1][i / RS_CH] = 0;
  }
}

// FMM_InitializeEikonal (FastMarching.c:60-74): the boundary values, which
// mark the first sweep's nodes
template <int DIM>
__global__ __launch_bounds__(RS_B) void rs_bc_k(const RsCell* __restrict__ cells, int ncells, long long nr) {
#pragma clang fp contract(off)
  const long long g = blockIdx.x;
  if (g >= nr) return;
  const RsCell& c = cells[rs_find<RS_KEY_RECORD>(cells, ncells, g)];
  const int64_t i = (g - c.rb0) * RS_B + threadIdx.x;
  if (i >= c.count) return;
  const dcl_irregular_node n = c.nodes[i];
  if (n.axis != -1 || !ri_inside(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z)) return;
  const double r = c.d3 ? sqrt(n.op.x * n.op.x + n.op.y * n.op.y + n.op.z * n.op.z)
                         : sqrt(n.op.x * n.op.x + n.op.y * n.op.y);
  const int64_t j = ri_index(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z);
  const double v = n.sign_center * r;
  c.buf[0][j] = v;
  c.buf[1][j] = v;
  c.mask[j] = 2;
  rs_mark<DIM>(c, c.nd[0], c.cd[0], j, n.pos.x - c.w.px, n.pos.y - c.w.py, c.d3 ? n.pos.z - c.w.pz : 0);
}

// a byte or a double of a level set's node array (buffer loads below 2^28
// nodes)
__device__ __forceinline__ unsigned rs_byte(const RsCell& c, const unsigned char* p, int64_t i) {
  if (c.small) return __builtin_amdgcn_raw_buffer_load_b8(make_rsrc(p, (unsigned)rs_nodes(c)), (int)i, 0, 0);
  return p[i];
}

__device__ __forceinline__ double rs_f64(const RsCell& c, const double* p, int64_t i) {
  if (c.small) return BL<double>::one(make_rsrc(p, (unsigned)(8 * rs_nodes(c))), (int)(8 * i), 0);
  return p[i];
}

// a neighbour as the side `sign` sees it: its value q = sign * phi, the time
// it is accepted (-1: before the side's pops, its value for a node of this
// side, INFINITY: never) and whether it pushes its neighbours in this side
struct RsNb {
  double q, t;
  bool push;
};

__device__ __forceinline__ RsNb rs_view(double inf, int sign, double v, int m) {
  RsNb r;
  r.q = inf;
  r.t = INFINITY;
  r.push = false;
  if (m == RS_OUT) return r;
  const double q = sign * v;
  r.q = q;
  if (m == 2) {
    // a boundary node pushes in this side when sign * phi >= 0
    if (0 <= q && q < inf) {
      r.t = -1;
      r.push = true;
    }
  } else if (m == (sign > 0 ? 1 : 0)) {
    if (q < inf) {
      r.t = q;
      r.push = true;
    }
  } else if (sign > 0 && 0 <= q && q < inf) {
    r.t = -1;  // a - side node: final before the + side starts
  }
  return r;
}

// the acceptance-time code of a neighbour as the side `sign` sees it (see
// rs_view): its time, or RS_T_STATIC for a node accepted from the start that
// does not push in this side
constexpr double RS_T_STATIC = -2.0;

__device__ __forceinline__ double rs_tcode(double inf, int sign, double v, int m) {
  const RsNb n = rs_view(inf, sign, v, m);
  return n.push ? n.t : (n.t < INFINITY ? RS_T_STATIC : INFINITY);
}

// the node's value from its stencil: nv the values and nt the
// acceptance-time codes (for the node's side) of the nodes two each way
// along each axis (-2, -1, +1, +2); NaN for a NaN update.  An axis's
// AccumulateCoef terms depend only on which of its four nodes are accepted,
// so they are kept from one push to the next and recomputed when that
// changes; the sums over the axes are redone in axis order each push (no
// partial sum is ever -0, so adding a kept term gives the same bits as
// adding it fresh).
template <int DIM>
__device__ __forceinline__ double rs_node(double inf, double bw, int sign, const double (&nv)[DIM][4],
                                          const double (&nt)[DIM][4]) {
#pragma clang fp contract(off)
  double cand = INFINITY, last = -INFINITY;
  double ta[DIM], tb[DIM], tc[DIM];
  unsigned sig[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    ta[a] = tb[a] = tc[a] = 0.0;
    sig[a] = ~0u;
  }
  // at most 2 * DIM distinct push times, then one pass that finds none
#pragma unroll 1
  for (int guard = 0; guard <= 2 * DIM; ++guard) {
    double T = INFINITY;
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int k = 1; k <= 2; ++k) {
        const double t = nt[a][k];
        if (t != RS_T_STATIC && t > last && t < T) T = t;
      }
    if (T == INFINITY || !(T < cand)) break;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      unsigned sg = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) sg |= nt[a][k] <= T ? 1u << k : 0u;
      if (sg != sig[a]) {
        sig[a] = sg;
        double loc[7];
        loc[0] = loc[6] = loc[3] = inf;
#pragma unroll
        for (int k = 0; k < 4; ++k) loc[k < 2 ? k + 1 : k + 2] = (sg >> k) & 1u ? sign * nv[a][k] : inf;
        double x = 0, y = 0, z = 0;
        ri_accumulate(inf, loc, x, y, z);
        ta[a] = x;
        tb[a] = y;
        tc[a] = z;
      }
    }
    double A = 0, B = 0, C = -1;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      A += ta[a];
      B += tb[a];
      C += tc[a];
    }
    const double disc = sqrt(B * B - 4 * A * C);
    double d1 = (-B + disc) / (2 * A);
    if (d1 != d1) d1 = -B / (2 * A);
    if (d1 != d1) return NAN;
    if (d1 < cand) cand = d1;
    last = T;
  }
  return cand <= bw ? sign * cand : sign * inf;
}

// the block-wide exclusive prefix of v (RS_B threads) and the total
__device__ __forceinline__ int rs_scan(int v, unsigned* wsum, int* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(x, o, 64);
    if (lane >= o) x += t;
  }
  if (lane == 63) wsum[wv] = (unsigned)x;
  __syncthreads();
  int pre = 0;
  for (int w = 0; w < wv; ++w) pre += (int)wsum[w];
  *total = (int)(wsum[0] + wsum[1] + wsum[2] + wsum[3]);
  __syncthreads();
  return pre + x - v;
}

// the nodes a sweep recomputes: the marked ones, read and cleared four at a
// time (the marks of the sweep after next are written into these bytes)
__global__ __launch_bounds__(RS_B) void rs_need_k(const RsCell* __restrict__ cells, int ncells,
                                                  const RsState* __restrict__ st, long long nch,
                                                  int2* __restrict__ nlist, int* __restrict__ ncount) {
  __shared__ int2 stage[RS_STAGE];
  __shared__ unsigned wsum[RS_B / 64];
  __shared__ int at;
  __shared__ int live[RS_B];  // this round's chunks holding marks, as batch chunk indices
  const int tid = threadIdx.x;
  int held = 0;  // entries in stage (uniform)
  // the block's chunks are blockIdx.x + q * gridDim.x; their flags are read
  // RS_B at a time, one per thread, and the marked ones visited in turn
  const long long per = (nch - blockIdx.x + gridDim.x - 1) / gridDim.x;
  for (long long q0 = 0; q0 < per; q0 += RS_B) {
    const long long g = blockIdx.x + (q0 + tid) * gridDim.x;
    bool has = false;
    if (q0 + tid < per) {
      const int ci = rs_find<RS_KEY_CHUNK>(cells, ncells, g);
      const RsCell& c = cells[ci];
      if (!st[ci].done) {
        unsigned char* cdk = (st[ci].k & 1) ? c.cd[1] : c.cd[0];
        has = cdk[g - c.cb0] != 0;
        if (has) cdk[g - c.cb0] = 0;
      }
    }
    int nlive;
    const int slot = rs_scan(has ? 1 : 0, wsum, &nlive);
    if (has) live[slot] = (int)g;
    __syncthreads();
    for (int e = 0; e < nlive; ++e) {
      if (held + RS_CH > RS_STAGE) {
        if (tid == 0) at = atomicAdd(ncount, held);
        __syncthreads();
        for (int q = tid; q < held; q += RS_B) nlist[at + q] = stage[q];
        __syncthreads();
        held = 0;
      }
      const long long gc = live[e];
      const int ci = rs_find<RS_KEY_CHUNK>(cells, ncells, gc);
      const RsCell& c = cells[ci];
      unsigned char* nd = (st[ci].k & 1) ? c.nd[1] : c.nd[0];
      const int64_t n = rs_nodes(c);
      const int64_t i0 = (gc - c.cb0) * RS_CH + 4 * tid;
      unsigned w = 0;
      if (i0 < n) {
        // the mark arrays are padded to 256 bytes, so a whole word is in them
        if (c.small)
          w = __builtin_amdgcn_raw_buffer_load_b32(make_rsrc(nd, (unsigned)((n + 3) & ~3ll)), (int)i0, 0, 0);
        else
          w = *reinterpret_cast<const unsigned*>(nd + i0);
        if (i0 + 4 > n) w &= n - i0 == 1 ? 0xffu : n - i0 == 2 ? 0xffffu : 0xffffffu;
        if (w) *reinterpret_cast<unsigned*>(nd + i0) = 0u;
      }
      const int cnt = ((w & 0xffu) != 0) + ((w & 0xff00u) != 0) + ((w & 0xff0000u) != 0) + ((w >> 24) != 0);
      int tot;
      int pos = held + rs_scan(cnt, wsum, &tot);
#pragma unroll
      for (int b2 = 0; b2 < 4; ++b2)
        if ((w >> (8 * b2)) & 0xffu) stage[pos++] = make_int2(ci, (int)(i0 + b2));
      held += tot;
      __syncthreads();
    }
  }
  if (held > 0) {
    if (tid == 0) at = atomicAdd(ncount, held);
    __syncthreads();
    for (int q = tid; q < held; q += RS_B) nlist[at + q] = stage[q];
  }
}

// the listed nodes' new values, one thread each.  The per-sweep counts of
// moving nodes are added into nparts partial sums per level set, each on
// its own 128-byte line (part p of level set i at p * pstride + i).
template <int DIM>
__global__ __launch_bounds__(RS_B) void rs_update_k(const RsCell* __restrict__ cells, RsState* __restrict__ st,
                                                    const int2* __restrict__ nlist, const int* __restrict__ ncount,
                                                    unsigned long long* __restrict__ parts, int nparts, int pstride) {
#pragma clang fp contract(off)
  const int n = *ncount;
  for (int e0 = blockIdx.x * RS_B; e0 < n; e0 += gridDim.x * RS_B) {
    const int e = e0 + threadIdx.x;
    bool moved = false;
    int ci = -1;
    if (e < n) {
      const int2 w = nlist[e];
      ci = w.x;
      const RsCell& c = cells[ci];
      const int k = st[ci].k, settle = st[ci].settle;
      const double* __restrict__ in = (k & 1) ? c.buf[1] : c.buf[0];
      double* __restrict__ out = (k & 1) ? c.buf[0] : c.buf[1];
      unsigned char* __restrict__ ncur = (k & 1) ? c.nd[0] : c.nd[1];
      unsigned char* __restrict__ ccur = (k & 1) ? c.cd[0] : c.cd[1];
      const int64_t i = w.y;
      const int64_t sy = c.w.nx, sz = (int64_t)c.w.nx * c.w.ny;
      const int lz = DIM == 3 ? (int)(i / sz) : 0;
      const int r2 = (int)(i - lz * sz);
      const int ly = r2 / c.w.nx, lx = r2 - ly * c.w.nx;
      const int m = rs_byte(c, c.mask, i);
      if (m != 2) {
        const double old = rs_f64(c, in, i);
        const int sg = m ? 1 : -1;
        double nv[DIM][4], nt[DIM][4];
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          const int p = a == 0 ? lx : a == 1 ? ly : lz;
          const int lim = a == 0 ? c.w.nx : a == 1 ? c.w.ny : c.w.nz;
          const int64_t st_ = a == 0 ? 1 : a == 1 ? sy : sz;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int o = q < 2 ? q - 2 : q - 1;
            double v = 0.0;
            int mm = RS_OUT;
            if (p + o >= 0 && p + o < lim) {
              v = rs_f64(c, in, i + o * st_);
              mm = rs_byte(c, c.mask, i + o * st_);
            }
            nv[a][q] = v;
            nt[a][q] = rs_tcode(c.inf, sg, v, mm);
          }
        }
        const double u = rs_node<DIM>(c.inf, c.bw, sg, nv, nt);
        if (u != u) atomicOr(&st[ci].nan, 1);
        double v = old;
        if (!settle || sg * u < sg * v) v = u;
        moved = fabs(v - old) > c.tol;
        out[i] = v;
        if (__double_as_longlong(v) != __double_as_longlong(old)) rs_mark<DIM>(c, ncur, ccur, i, lx, ly, lz);
      }
    }
    // one atomic per wave where the wave's nodes are of one level set
    const int c0 = __builtin_amdgcn_readfirstlane(ci);
    const bool uni = __ballot(ci != c0 && ci >= 0) == 0ull;
    const unsigned long long mb = __ballot(moved);
    if (uni) {
      if (mb && (threadIdx.x & 63) == 0)
        atomicAdd(&parts[(long long)((e0 / RS_B + (threadIdx.x >> 6)) % nparts) * pstride + c0],
                  (unsigned long long)__popcll(mb));
    } else if (moved) {
      atomicAdd(&parts[(long long)((e0 / RS_B) % nparts) * pstride + ci], 1ull);
    }
  }
}

// after a sweep, one wave per level set: the count of moving nodes decides
// whether it has settled and whether its values may only go down from the
// next sweep on
__global__ __launch_bounds__(RS_B) void rs_state_k(RsState* __restrict__ st, int ncells, int max_sweeps,
                                                   int* __restrict__ active, unsigned long long* __restrict__ parts,
                                                   int nparts, int pstride) {
  const int i = blockIdx.x * (RS_B / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= ncells) return;
  if (st[i].done) return;
  unsigned long long c = 0;
  if (lane < nparts) {
    c = parts[(long long)lane * pstride + i];
    parts[(long long)lane * pstride + i] = 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if (lane != 0) return;
  RsState s = st[i];
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

__device__ __forceinline__ long long rs_prefix(const unsigned* counts, const unsigned long long* sums, long long b) {
  return (long long)counts[b] + (long long)sums[b / SCAN_TILE];
}

// rs_final_k and rs_side_write_k run one raster block (RS_B nodes) per wave,
// nodes e * 64 + lane for e = 0 .. RS_B / 64 - 1 (raster order: e, then
// lane); rs_final_k issues its loads before its stores and keeps each row's
// ballot of each side, from which rs_side_write_k ranks the band nodes.
constexpr int RS_WPB = RS_B / 64;  // raster blocks per workgroup
constexpr int RS_ROWS = RS_B / 64; // 64-node rows per raster block

// the result into the caller's buffer where the last sweep wrote the other
// one, then the band counts: each side per raster block (the block past the
// last level set counts nothing, so the scans' last entry is the total), and
// each row's ballot of each side for rs_side_write_k
__global__ __launch_bounds__(RS_B) void rs_final_k(const RsCell* __restrict__ cells, int ncells, long long nq,
                                                   long long nq1, const RsState* __restrict__ st,
                                                   unsigned* __restrict__ c0, unsigned* __restrict__ c1,
                                                   unsigned long long* __restrict__ sides) {
  const long long g = (long long)blockIdx.x * RS_WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  unsigned n0 = 0, n1 = 0;
  if (g < nq) {  // uniform over the wave
    const int ci = rs_find<RS_KEY_RASTER>(cells, ncells, g);
    const RsCell& c = cells[ci];
    const int64_t n = rs_nodes(c);
    const bool odd = st[ci].k & 1;
    const int64_t i0 = (g - c.qb0) * RS_B + lane;
    const double* src = c.buf[odd ? 1 : 0];
    // every row's loads first: the copy's stores could alias later rows'
    // loads and would otherwise serialise them
    double v[RS_ROWS];
    int m[RS_ROWS];
#pragma unroll
    for (int e = 0; e < RS_ROWS; ++e) {
      const int64_t i = i0 + e * 64;
      v[e] = i < n ? src[i] : 0.0;
      m[e] = i < n ? (int)c.mask[i] : 2;
    }
    unsigned long long mine = 0;  // lane 2e / 2e + 1: row e's - / + side ballot
#pragma unroll
    for (int e = 0; e < RS_ROWS; ++e) {
      const int64_t i = i0 + e * 64;
      if (odd && i < n) c.buf[0][i] = v[e];
      const bool fin = fabs(v[e]) < c.inf;
      const unsigned long long b0 = __ballot(m[e] == 0 && fin), b1 = __ballot(m[e] == 1 && fin);
      n0 += (unsigned)__popcll(b0);
      n1 += (unsigned)__popcll(b1);
      if (lane == 2 * e) mine = b0;
      if (lane == 2 * e + 1) mine = b1;
    }
    if (lane < 2 * RS_ROWS) sides[g * (2 * RS_ROWS) + lane] = mine;
  }
  if (lane == 0 && g < nq1) {
    c0[g] = n0;
    c1[g] = n1;
  }
}

__device__ __forceinline__ bool rs_is_bc(const RsCell& c, int64_t i, dcl_icoor* pos) {
  if (i >= c.count) return false;
  const dcl_irregular_node& n = c.nodes[i];
  *pos = n.pos;
  return n.axis == -1 && ri_inside(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z);
}

// boundary records per record block
__global__ __launch_bounds__(RS_B) void rs_bc_count_k(const RsCell* __restrict__ cells, int ncells, long long nr,
                                                      unsigned* __restrict__ counts) {
  const long long g = blockIdx.x;
  bool p = false;
  if (g < nr) {
    const RsCell& c = cells[rs_find<RS_KEY_RECORD>(cells, ncells, g)];
    dcl_icoor q;
    p = rs_is_bc(c, (g - c.rb0) * RS_B + threadIdx.x, &q);
  }
  const int cnt = __syncthreads_count(p);
  if (threadIdx.x == 0) counts[g] = (unsigned)cnt;
}

// the band of each level set: [boundary nodes in record order | - side |
// + side], each side in raster order
__global__ __launch_bounds__(RS_B) void rs_bc_write_k(const RsCell* __restrict__ cells, int ncells, long long nr,
                                                      const unsigned* __restrict__ rc,
                                                      const unsigned long long* __restrict__ rs) {
  __shared__ unsigned sh[RS_B / 64];
  const long long g = blockIdx.x;
  if (g >= nr) return;
  const RsCell& c = cells[rs_find<RS_KEY_RECORD>(cells, ncells, g)];
  dcl_icoor q = {0, 0, 0};
  const bool p = rs_is_bc(c, (g - c.rb0) * RS_B + threadIdx.x, &q);
  const unsigned r = rs_rank(p, sh);
  if (!p) return;
  const int64_t pos = rs_prefix(rc, rs, g) - rs_prefix(rc, rs, c.rb0) + r;
  if (pos < c.band_cap) c.band[pos] = q;
}

// the band nodes of raster block g, ranked by rs_final_k's ballots
__device__ __forceinline__ void rs_side_block(const RsCell* __restrict__ cells, int ncells, long long g, int lane,
                                              const unsigned* __restrict__ rc,
                                              const unsigned long long* __restrict__ rs,
                                              const unsigned* __restrict__ c0,
                                              const unsigned long long* __restrict__ s0,
                                              const unsigned* __restrict__ c1,
                                              const unsigned long long* __restrict__ s1,
                                              const unsigned long long* __restrict__ sides) {
  const RsCell& c = cells[rs_find<RS_KEY_RASTER>(cells, ncells, g)];
  const long long p0 = rs_prefix(c0, s0, g), p1 = rs_prefix(c1, s1, g);
  const long long nbc = rs_prefix(rc, rs, c.rb1) - rs_prefix(rc, rs, c.rb0);
  const long long n0 = rs_prefix(c0, s0, c.qb1) - rs_prefix(c0, s0, c.qb0);
  const long long base0 = nbc + p0 - rs_prefix(c0, s0, c.qb0);
  const long long base1 = nbc + n0 + p1 - rs_prefix(c1, s1, c.qb0);
  const int64_t plane = (int64_t)c.w.nx * c.w.ny;
  const int64_t i0 = (g - c.qb0) * RS_B + lane;
  const unsigned long long lt = (1ull << lane) - 1ull;  // the lanes below this one
  const unsigned long long* sb = sides + g * (2 * RS_ROWS);
  long long r0 = 0, r1 = 0;  // band nodes of each side in the rows before
#pragma unroll
  for (int e = 0; e < RS_ROWS; ++e) {
    const unsigned long long b0 = sb[2 * e], b1 = sb[2 * e + 1];
    const bool k0 = (b0 >> lane) & 1ull, k1 = (b1 >> lane) & 1ull;
    if (k0 || k1) {
      const int64_t pos = k0 ? base0 + r0 + __popcll(b0 & lt) : base1 + r1 + __popcll(b1 & lt);
      if (pos < c.band_cap) {
        const int64_t i = i0 + e * 64;
        const int z = c.d3 ? (int)(i / plane) : 0, rr = (int)(i - (int64_t)z * plane);
        const int y = rr / c.w.nx, x = rr - y * c.w.nx;
        dcl_icoor q;
        q.x = c.w.px + x;
        q.y = c.w.py + y;
        q.z = c.w.pz + z;
        c.band[pos] = q;
      }
    }
    r0 += __popcll(b0);
    r1 += __popcll(b1);
  }
}

// the two sides of the band, from the counts' scans and rs_final_k's
// ballots: one raster block per wave; a block without band nodes reads its
// two counts and stops
__global__ __launch_bounds__(RS_B) void rs_side_write_k(const RsCell* __restrict__ cells, int ncells, long long nq,
                                                        const unsigned* __restrict__ rc,
                                                        const unsigned long long* __restrict__ rs,
                                                        const unsigned* __restrict__ c0,
                                                        const unsigned long long* __restrict__ s0,
                                                        const unsigned* __restrict__ c1,
                                                        const unsigned long long* __restrict__ s1,
                                                        const unsigned long long* __restrict__ sides) {
  const long long g = (long long)blockIdx.x * RS_WPB + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= nq) return;  // uniform over the wave; no block barrier below
  if (rs_prefix(c0, s0, g + 1) == rs_prefix(c0, s0, g) && rs_prefix(c1, s1, g + 1) == rs_prefix(c1, s1, g)) return;
  rs_side_block(cells, ncells, g, lane, rc, rs, c0, s0, c1, s1, sides);
}

// per level set: boundary nodes, - side, + side
__global__ __launch_bounds__(RS_B) void rs_totals_k(const RsCell* __restrict__ cells, int ncells,
                                                    const unsigned* __restrict__ rc,
                                                    const unsigned long long* __restrict__ rs,
                                                    const unsigned* __restrict__ c0,
                                                    const unsigned long long* __restrict__ s0,
                                                    const unsigned* __restrict__ c1,
                                                    const unsigned long long* __restrict__ s1,
                                                    unsigned long long* __restrict__ totals) {
  const int i = blockIdx.x * RS_B + threadIdx.x;
  if (i >= ncells) return;
  const RsCell& c = cells[i];
  totals[3 * i] = (unsigned long long)(rs_prefix(rc, rs, c.rb1) - rs_prefix(rc, rs, c.rb0));
  totals[3 * i + 1] = (unsigned long long)(rs_prefix(c0, s0, c.qb1) - rs_prefix(c0, s0, c.qb0));
  totals[3 * i + 2] = (unsigned long long)(rs_prefix(c1, s1, c.qb1) - rs_prefix(c1, s1, c.qb0));
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
  long long nodes = 0, nq = 0, nr = 0, nch = 0;
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
    nodes += n;
    nq += (n + RS_B - 1) / RS_B;
    nch += (n + RS_CH - 1) / RS_CH;
    nr += (u.count + RS_B - 1) / RS_B;
    cell_bytes += rs_al(sizeof(double) * (size_t)n) + 3 * rs_al((size_t)n) +
                  2 * rs_al((size_t)((n + RS_CH - 1) / RS_CH));
  }
  if (nodes >= INT32_MAX || nr >= INT32_MAX)
    return set_error(DCL_ERR_ARG_SIZ, "a batch of more than 2^31 - 1 nodes or record blocks");
  // one more raster and record block, counting nothing: the scans' totals
  const long long nq1 = nq + 1, nr1 = nr + 1;
  const int nparts = ncells >= 4096 ? 1 : (4096 / ncells < 64 ? 4096 / ncells : 64);
  const int pstride = ncells < 16 ? 16 : ncells;
  const size_t bytes = rs_al(sizeof(RsCell) * ncells) + rs_al(sizeof(RsState) * ncells) +
                       rs_al(sizeof(unsigned long long) * (size_t)nparts * pstride) +
                       2 * rs_al(sizeof(int) * (size_t)(max_sweeps + 1)) + rs_al(sizeof(int2) * (size_t)nodes) +
                       rs_al(sizeof(unsigned long long) * 3 * (size_t)ncells) +
                       2 * rs_al(sizeof(unsigned) * (size_t)nq1) +
                       2 * rs_al(sizeof(unsigned long long) * (size_t)scan_tiles(nq1)) +
                       rs_al(sizeof(unsigned) * (size_t)nr1) +
                       rs_al(sizeof(unsigned long long) * (size_t)scan_tiles(nr1)) +
                       rs_al(sizeof(unsigned long long) * 3) +
                       rs_al(sizeof(unsigned long long) * 2 * RS_ROWS * (size_t)nq) + cell_bytes;
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
  unsigned long long* parts =
      reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * (size_t)nparts * pstride));
  int* active = reinterpret_cast<int*>(take(sizeof(int) * (size_t)(max_sweeps + 1)));
  int* ncount = reinterpret_cast<int*>(take(sizeof(int) * (size_t)(max_sweeps + 1)));
  int2* nlist = reinterpret_cast<int2*>(take(sizeof(int2) * (size_t)nodes));
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 3 * ncells));
  unsigned* c0 = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * (size_t)nq1));
  unsigned* c1 = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * (size_t)nq1));
  unsigned long long* s0 = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * scan_tiles(nq1)));
  unsigned long long* s1 = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * scan_tiles(nq1)));
  unsigned* rc = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * (size_t)nr1));
  unsigned long long* rsum = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * scan_tiles(nr1)));
  unsigned long long* grand = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 3));
  unsigned long long* sides =
      reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 2 * RS_ROWS * (size_t)nq));

  RsCell* host = static_cast<RsCell*>(malloc(sizeof(RsCell) * ncells));
  RsState* hstate = static_cast<RsState*>(malloc(sizeof(RsState) * ncells));
  unsigned long long* htot = static_cast<unsigned long long*>(malloc(sizeof(unsigned long long) * 3 * ncells));
  if (!host || !hstate || !htot) {
    free(host);
    free(hstate);
    free(htot);
    return set_error(DCL_ERR_MEM, "host memory");
  }
  long long q0 = 0, r0 = 0, h0 = 0;
  for (int i = 0; i < ncells; ++i) {
    const dcl_reinit_cell& u = cells[i];
    RsCell& c = host[i];
    const int64_t n = win_nodes(u.win);
    c.buf[0] = static_cast<double*>(u.phi);
    c.buf[1] = reinterpret_cast<double*>(take(sizeof(double) * (size_t)n));
    c.mask = reinterpret_cast<unsigned char*>(take((size_t)n));
    c.nd[0] = reinterpret_cast<unsigned char*>(take((size_t)n));
    c.nd[1] = reinterpret_cast<unsigned char*>(take((size_t)n));
    c.cd[0] = reinterpret_cast<unsigned char*>(take((size_t)((n + RS_CH - 1) / RS_CH)));
    c.cd[1] = reinterpret_cast<unsigned char*>(take((size_t)((n + RS_CH - 1) / RS_CH)));
    c.w = kwin(u.win);
    c.d3 = d3;
    c.small = n < (1 << 28);
    c.bw = u.band_width;
    c.inf = u.phi_inf;
    c.tol = u.band_width * RS_TOL;
    c.qb0 = q0;
    q0 += (n + RS_B - 1) / RS_B;
    c.qb1 = q0;
    c.cb0 = h0;
    h0 += (n + RS_CH - 1) / RS_CH;
    c.rb0 = r0;
    r0 += (u.count + RS_B - 1) / RS_B;
    c.rb1 = r0;
    c.nodes = u.nodes;
    c.count = u.count;
    c.band = u.band;
    c.band_cap = u.band_capacity;
    RsState& q = hstate[i];
    memset(&q, 0, sizeof(q));
    q.last = -1;
  }
  int rc_ = 0;
  {
    hipError_t e = hipMemcpyAsync(dcells, host, sizeof(RsCell) * ncells, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dstate, hstate, sizeof(RsState) * ncells, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(active, 0, sizeof(int) * (size_t)(max_sweeps + 1), s);
    if (e == hipSuccess) e = hipMemsetAsync(ncount, 0, sizeof(int) * (size_t)(max_sweeps + 1), s);
    if (e == hipSuccess) e = hipMemsetAsync(parts, 0, sizeof(unsigned long long) * (size_t)nparts * pstride, s);
    if (e != hipSuccess) rc_ = hip_fail(e, "reinit upload", __FILE__, __LINE__);
  }
  if (!rc_) {
    hipLaunchKernelGGL(rs_reset_k, dim3((unsigned)nq), dim3(RS_B), 0, s, dcells, ncells, nq);
    if (nr > 0) {
      if (d3) hipLaunchKernelGGL(rs_bc_k<3>, dim3((unsigned)nr), dim3(RS_B), 0, s, dcells, ncells, nr);
      else hipLaunchKernelGGL(rs_bc_k<2>, dim3((unsigned)nr), dim3(RS_B), 0, s, dcells, ncells, nr);
    }
    if (hipGetLastError() != hipSuccess) rc_ = set_error(DCL_ERR_LIB, "reinit reset / boundary launch failed");
  }
  // the sweeps, in groups of RS_GROUP: once a group is queued, the count of
  // level sets still sweeping after the group before it is awaited (through
  // pinned memory), so the GPU always holds the next group; a level set
  // that has settled skips the sweeps queued after it
  static thread_local int* pin = nullptr;
  if (!pin && !rc_) {
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&pin), 2 * sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) {
      pin = nullptr;
      rc_ = hip_fail(e, "reinit pinned count", __FILE__, __LINE__);
    }
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (int q = 0; q < 2 && !rc_; ++q) {
    hipError_t e = hipEventCreateWithFlags(&ev[q], hipEventDisableTiming);
    if (e != hipSuccess) rc_ = hip_fail(e, "reinit events", __FILE__, __LINE__);
  }
  const unsigned ng = (unsigned)(nch < RS_NEED_GRID ? nch : RS_NEED_GRID);
  bool settled = false;
  for (int j0 = 0, grp = 0; j0 < max_sweeps && !rc_ && !settled; j0 += RS_GROUP, ++grp) {
    const int j1 = j0 + RS_GROUP < max_sweeps ? j0 + RS_GROUP : max_sweeps;
    for (int j = j0; j < j1; ++j) {
      hipLaunchKernelGGL(rs_need_k, dim3(ng), dim3(RS_B), 0, s, dcells, ncells, dstate, nch, nlist, ncount + j);
      if (d3)
        hipLaunchKernelGGL(rs_update_k<3>, dim3(RS_GRID), dim3(RS_B), 0, s, dcells, dstate, nlist, ncount + j, parts,
                           nparts, pstride);
      else
        hipLaunchKernelGGL(rs_update_k<2>, dim3(RS_GRID), dim3(RS_B), 0, s, dcells, dstate, nlist, ncount + j, parts,
                           nparts, pstride);
      hipLaunchKernelGGL(rs_state_k, dim3(grid1d(ncells, RS_B / 64)), dim3(RS_B), 0, s, dstate, ncells, max_sweeps,
                         active + j, parts, nparts, pstride);
    }
    if (hipGetLastError() != hipSuccess) {
      rc_ = set_error(DCL_ERR_LIB, "reinit sweep launch failed");
      break;
    }
    hipError_t e = hipMemcpyAsync(pin + (grp & 1), active + (j1 - 1), sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(ev[grp & 1], s);
    if (e == hipSuccess && grp > 0) {
      e = hipEventSynchronize(ev[(grp - 1) & 1]);
      if (e == hipSuccess && pin[(grp - 1) & 1] == 0) settled = true;
    }
    if (e != hipSuccess) rc_ = hip_fail(e, "reinit sweeps", __FILE__, __LINE__);
  }
  for (int q = 0; q < 2; ++q)
    if (ev[q]) (void)hipEventDestroy(ev[q]);
  if (!rc_) {
    hipError_t e = hipMemcpyAsync(hstate, dstate, sizeof(RsState) * ncells, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc_ = hip_fail(e, "reinit sweeps", __FILE__, __LINE__);
  }
  for (int i = 0; i < ncells && !rc_; ++i) {
    const RsState& q = hstate[i];
    if (sweeps) sweeps[i] = q.k;
    if (q.done == RS_NAN) rc_ = set_error(DCL_ERR_LIB, "level set %d: NaN in a reinitialisation update", i);
    else if (q.done != RS_SETTLED)
      rc_ = set_error(DCL_ERR_LIB, "level set %d: reinitialisation did not settle in %d sweeps", i, max_sweeps);
  }
  // the result, then the band
  if (!rc_) {
    hipLaunchKernelGGL(rs_final_k, dim3((unsigned)((nq1 + RS_WPB - 1) / RS_WPB)), dim3(RS_B), 0, s, dcells,
                       ncells, nq, nq1, dstate, c0, c1, sides);
    hipLaunchKernelGGL(rs_bc_count_k, dim3((unsigned)nr1), dim3(RS_B), 0, s, dcells, ncells, nr, rc);
    if (hipGetLastError() != hipSuccess) rc_ = set_error(DCL_ERR_LIB, "reinit band launch failed");
  }
  if (!rc_) {
    const ScanArrays sn = {{c0, c1}, {s0, s1}, {grand, grand + 1}};
    rc_ = scan_counts(sn, 2, nq1, s);
  }
  if (!rc_) {
    const ScanArrays sr = {{rc, nullptr}, {rsum, nullptr}, {grand + 2, nullptr}};
    rc_ = scan_counts(sr, 1, nr1, s);
  }
  if (!rc_) {
    if (nr > 0) hipLaunchKernelGGL(rs_bc_write_k, dim3((unsigned)nr), dim3(RS_B), 0, s, dcells, ncells, nr, rc, rsum);
    hipLaunchKernelGGL(rs_side_write_k, dim3((unsigned)((nq + RS_WPB - 1) / RS_WPB)), dim3(RS_B), 0, s, dcells,
                       ncells, nq, rc, rsum, c0, s0, c1, s1, sides);
    hipLaunchKernelGGL(rs_totals_k, dim3(grid1d(ncells, RS_B)), dim3(RS_B), 0, s, dcells, ncells, rc, rsum, c0, s0,
                       c1, s1, totals);
    if (hipGetLastError() != hipSuccess) rc_ = set_error(DCL_ERR_LIB, "reinit band launch failed");
  }
  if (!rc_) {
    hipError_t e = hipMemcpyAsync(htot, totals, sizeof(unsigned long long) * 3 * ncells, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc_ = hip_fail(e, "reinit band", __FILE__, __LINE__);
  }
  for (int i = 0; i < ncells && !rc_; ++i)
    cells[i].band_count = (int64_t)(htot[3 * i] + htot[3 * i + 1] + htot[3 * i + 2]);
  if (rc_) (void)hipStreamSynchronize(s);
  free(host);
  free(hstate);
  free(htot);
  return rc_;
}
