// reinit.hip -- reinitialisation on gfx950: LevelSetReinitialize
// (LevelSetMethod/FastMarching.c:10-146) with the second-order neighbour
// updates of FastMarching3D.c:14-101 / FastMarching2D.c:33-185 and the
// binary min-heap of Common/Heap.c:59-139.
//
// The fast marching method accepts nodes one at a time in heap order, and
// which neighbours count as known when a node is updated (first or second
// order, which side) depends on that order, equal keys included.  The GPU
// keeps that order exactly: one wave runs one sign pass of one level set,
// lane 0 owns the heap (the reference's sift-up and sift-down with the same
// comparisons, so equal keys pop in the same order) and the band list, and
// lanes 0..5 (0..3 in 2D) compute the updates of the accepted node's
// neighbours side by side.  The parallelism is across passes and level sets:
// the - pass writes only nodes whose phi is -PHI_INF and reads only
// non-negative values of -phi (the + pass the mirror image), so the two
// passes of a level set run as two concurrent waves with the same result as
// the reference's sequential order, and a batch of level sets (DCell's
// cells) runs as 2 x ncells waves.
//
// Steps: reset phi to +-PHI_INF (reinit_reset_k), write the boundary values
// at the projection records (reinit_bc_k), then the passes (reinit_fmm_k);
// the - pass also lists the boundary nodes in record order.  The band is
// assembled as [boundary nodes | - pass | + pass] by copies.
//
// Departures from the reference, where it reads out of bounds or leaves a
// value unset: a neighbour outside the window is not pushed and a record
// outside it is skipped (the reference reads them unchecked; LevelSetResize
// keeps the band away from the edge), a pushed 2D node's z is the window's
// (the reference leaves it unset), and a NaN update or a full heap stops the
// pass with an error instead of exiting.
//
// The heap's first 4096 slots (12 levels) sit in LDS, where nearly every pop
// and push spends most of its steps; deeper levels are in memory.
#include <math.h>

#include "dcl_host.hpp"
#include "dcl_reinit.hpp"

namespace dcl {

// a heap entry: the tentative value and the node's window-local
// coordinates packed 21 bits each (windows are below 2^21 per axis)
struct RiEntry {
  double v;
  unsigned long long key;
};

__device__ __forceinline__ unsigned long long ri_pack(int lx, int ly, int lz) {
  return (unsigned long long)lx | ((unsigned long long)ly << 21) | ((unsigned long long)lz << 42);
}

// the heap's first RI_LDS slots (12 levels) live in LDS, the rest in memory
constexpr int RI_LDS = 4096;

struct RiHeap {
  RiEntry* lds;
  RiEntry* mem;
};

__device__ __forceinline__ RiEntry ri_get(const RiHeap& h, int64_t i) { return i < RI_LDS ? h.lds[i] : h.mem[i]; }
__device__ __forceinline__ void ri_set(const RiHeap& h, int64_t i, const RiEntry& e) {
  if (i < RI_LDS)
    h.lds[i] = e;
  else
    h.mem[i] = e;
}

struct RiCell {
  double* phi;
  KWin w;
  int d3;
  const dcl_irregular_node* nodes;
  int64_t count;
  double bw, inf;
  RiEntry* heap[2];
  int64_t heap_cap;
  dcl_icoor* list[3];  // boundary nodes, - pass, + pass
  int64_t list_cap;
  long long* stat;  // boundary count, - pass count, + pass count, error
};

enum { RI_OK = 0, RI_NAN = 1, RI_HEAP = 2, RI_STEPS = 3 };

__global__ __launch_bounds__(256) void reinit_reset_k(double* __restrict__ phi, int64_t n, double inf) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double v = phi[i];
    phi[i] = v > 0 ? inf : -inf;
  }
}

// FMM_InitializeEikonal (FastMarching.c:60-74): phi at each projection record
// is the signed distance to its orthogonal projection
__global__ __launch_bounds__(256) void reinit_bc_k(RiCell c) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= c.count) return;
  const dcl_irregular_node n = c.nodes[i];
  if (n.axis != -1) return;
  if (!ri_inside(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z)) return;
  const double r = c.d3 ? sqrt(n.op.x * n.op.x + n.op.y * n.op.y + n.op.z * n.op.z)
                         : sqrt(n.op.x * n.op.x + n.op.y * n.op.y);
  c.phi[ri_index(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z)] = n.sign_center * r;
}

// phi as this pass sees it.  Lane 0 of the wave writes the accepted nodes
// and the other lanes read them back after a workgroup-scope fence (one
// wave, one CU: its L1 holds the writes).  The other pass may write nodes of
// the opposite sign at any time; whether a read sees the old -+PHI_INF or
// the new value, it is of the opposite sign, which this pass treats the same
// either way, so no wider coherence is needed.
__device__ __forceinline__ double ri_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// HeapInsert (Common/Heap.c:59-91), lane 0 only; the item is stored once at
// its final place (the reference moves it level by level; same final tree)
__device__ __forceinline__ void ri_heap_insert(const RiHeap& H, int64_t& len, const RiEntry& item) {
  int64_t c = len;
  len = c + 1;
  while (c != 1) {
    const int64_t p = c >> 1;
    const RiEntry parent = ri_get(H, p);
    if (parent.v - item.v < 0) break;
    ri_set(H, c, parent);
    c = p;
  }
  ri_set(H, c, item);
}

// HeapPop (Common/Heap.c:94-139), lane 0 only
__device__ __forceinline__ RiEntry ri_heap_pop(const RiHeap& H, int64_t& hlen) {
  const RiEntry top = ri_get(H, 1);
  int64_t len = hlen;
  if (len == 2) {
    hlen = 1;
    return top;
  }
  len--;
  const RiEntry bubble = ri_get(H, len);
  hlen = len;
  int64_t p = 1, l = 2;
  while (l < len) {
    const int64_t r = l + 1;
    int64_t s = l;
    RiEntry child = ri_get(H, l);
    if (r < len) {
      const RiEntry right = ri_get(H, r);
      if (right.v - child.v < 0) {
        s = r;
        child = right;
      }
    }
    if (bubble.v - child.v <= 0) break;
    ri_set(H, p, child);
    p = s;
    l = 2 * p;
  }
  ri_set(H, p, bubble);
  return top;
}

// STAR (Common/Common.h:49-51)
__device__ __forceinline__ int ri_star(int n, int a) { return (n >> 1) == a ? ((n & 1) ? -1 : 1) : 0; }

// FMM_PushNeighbors3D / 2D for node (X, Y, Z): every lane of the wave calls
// it; lanes below NN compute one neighbour each, lane 0 inserts them in
// STAR order.  Returns an error code (uniform).
template <int NN>
__device__ __forceinline__ int ri_push(const RiCell& c, int sign, int X, int Y, int Z, const RiHeap& H,
                                      int64_t& hlen) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x;
  const int i = X + ri_star(lane, 0), j = Y + ri_star(lane, 1), k = c.d3 ? Z + ri_star(lane, 2) : Z;
  bool valid = false, nan = false;
  double d1 = 0.0;
  if (lane < NN && ri_inside(c.w, c.d3, i, j, k)) {
    if (!(sign * ri_load(c.phi + ri_index(c.w, c.d3, i, j, k)) < c.bw)) {
      valid = true;
      double a = 0, b = 0, cc = -1, loc[7];
#pragma unroll
      for (int axis = 0; axis < (NN == 6 ? 3 : 2); ++axis) {
#pragma unroll
        for (int m = 0; m < 7; ++m) {
          const int s = m - 3;
          const int ii = i + (axis == 0 ? s : 0), jj = j + (axis == 1 ? s : 0), kk = k + (axis == 2 ? s : 0);
          loc[m] = ri_inside(c.w, c.d3, ii, jj, kk) ? sign * ri_load(c.phi + ri_index(c.w, c.d3, ii, jj, kk)) : c.inf;
        }
        ri_accumulate(c.inf, loc, a, b, cc);
      }
      const double disc = sqrt(b * b - 4 * a * cc);
      d1 = (-b + disc) / (2 * a);
      if (d1 != d1) d1 = -b / (2 * a);
      nan = d1 != d1;
    }
  }
  if (__ballot(nan)) return RI_NAN;
  const uint64_t vm = __ballot(valid);
#pragma unroll
  for (int n = 0; n < NN; ++n) {
    if (!((vm >> n) & 1ull)) continue;
    RiEntry e;
    e.v = __shfl(d1, n);
    e.key = ri_pack(__shfl(i, n) - c.w.px, __shfl(j, n) - c.w.py, c.d3 ? __shfl(k, n) - c.w.pz : 0);
    if (hlen >= c.heap_cap) return RI_HEAP;
    if (lane == 0) ri_heap_insert(H, hlen, e);
    else hlen++;
  }
  return RI_OK;
}

// one wave per (level set, sign): FMM_InitializeEikonal's pushes
// (FastMarching.c:76-93) and FMM_SolveEikonal (:97-146)
template <int NN>
__global__ __launch_bounds__(64) void reinit_fmm_k(const RiCell* __restrict__ cells) {
#pragma clang fp contract(off)
  const RiCell c = cells[blockIdx.x >> 1];
  const int pass = blockIdx.x & 1;
  const int sign = pass ? 1 : -1;
  const int lane = threadIdx.x;
  __shared__ RiEntry lds_heap[RI_LDS];
  const RiHeap H = {lds_heap, pass ? c.heap[1] : c.heap[0]};
  dcl_icoor* accepted = pass ? c.list[2] : c.list[1];
  int64_t hlen = 1, nbc = 0, nacc = 0;
  int err = RI_OK;
  // the boundary nodes, in record order
  for (int64_t base = 0; base < c.count && err == RI_OK; base += 64) {
    const int64_t r = base + lane;
    bool bc = false;
    int x = 0, y = 0, z = 0;
    double v = -1.0;
    if (r < c.count) {
      const dcl_irregular_node n = c.nodes[r];
      bc = n.axis == -1 && ri_inside(c.w, c.d3, n.pos.x, n.pos.y, n.pos.z);
      x = n.pos.x, y = n.pos.y, z = n.pos.z;
      if (bc) v = sign * ri_load(c.phi + ri_index(c.w, c.d3, x, y, z));
    }
    const uint64_t bm = __ballot(bc);
    if (pass == 0 && bc) {
      const int64_t at = nbc + __popcll(bm & ((1ull << lane) - 1ull));
      if (at < c.list_cap) {
        dcl_icoor q;
        q.x = x, q.y = y, q.z = z;
        c.list[0][at] = q;
      }
    }
    nbc += __popcll(bm);
    uint64_t pm = __ballot(bc && v >= 0.);
    while (pm && err == RI_OK) {
      const int l = __builtin_ctzll(pm);
      pm &= pm - 1;
      err = ri_push<NN>(c, sign, __shfl(x, l), __shfl(y, l), __shfl(z, l), H, hlen);
    }
  }
  // FMM_SolveEikonal; every pop either ends the pass or is followed by at
  // most NN pushes, and a node is accepted again only at a value below its
  // last, so a bounded number of pops always suffices: the cap only keeps a
  // corrupted input from running without end
  const int64_t max_pops = 64 * ((int64_t)c.w.nx * c.w.ny * (c.d3 ? c.w.nz : 1) + c.count) + (1 << 20);
  int64_t pops = 0;
  while (hlen > 1 && err == RI_OK) {
    if (++pops > max_pops) {
      err = RI_STEPS;
      break;
    }
    RiEntry top;
    if (lane == 0) top = ri_heap_pop(H, hlen);
    else hlen--;
    const double v = __shfl(top.v, 0);
    const unsigned long long key = __shfl(top.key, 0);
    const int x = c.w.px + (int)(key & 0x1fffffull), y = c.w.py + (int)((key >> 21) & 0x1fffffull);
    const int z = c.d3 ? c.w.pz + (int)(key >> 42) : c.w.pz;  // 2D: the window's z (see orc_reinitialize)
    if (v > c.bw) break;
    double* q = c.phi + ri_index(c.w, c.d3, x, y, z);
    double cur = 0.0;
    if (lane == 0) cur = ri_load(q);
    cur = __shfl(cur, 0);
    if (v >= sign * cur) continue;
    if (lane == 0) {
      if (nacc < c.list_cap) {
        dcl_icoor p;
        p.x = x, p.y = y, p.z = z;
        accepted[nacc] = p;
      }
      __hip_atomic_store(q, sign * v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    ++nacc;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    err = ri_push<NN>(c, sign, x, y, z, H, hlen);
  }
  if (lane == 0) {
    if (pass == 0) c.stat[0] = nbc;
    c.stat[1 + pass] = nacc;
    if (err != RI_OK) c.stat[3] = err;
  }
}

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_reinitialize(dcl_reinit_cell* cells, int ncells, int64_t heap_capacity, void* stream) {
  if (ncells < 0) return set_error(DCL_ERR_ARG_SIZ, "ncells < 0");
  if (ncells > 0 && !cells) return set_error(DCL_ERR_ARG_NULL, "cells is NULL");
  if (heap_capacity < 0) return set_error(DCL_ERR_ARG_SIZ, "heap_capacity < 0");
  if (ncells == 0) return 0;
  if (ncells > (1 << 20)) return set_error(DCL_ERR_ARG_SIZ, "at most 2^20 level sets per call");
  hipStream_t s = as_stream(stream);
  // scratch: cell descriptors | stats | per cell: 2 heaps, 3 lists
  size_t bytes = ((sizeof(RiCell) * ncells + 255) / 256) * 256 + ((sizeof(long long) * 4 * ncells + 255) / 256) * 256;
  for (int i = 0; i < ncells; ++i) {
    dcl_reinit_cell& u = cells[i];
    if (!u.phi) return set_error(DCL_ERR_ARG_NULL, "level set %d: phi is NULL", i);
    DCL_TRY(check_window(&u.win, "phi", 1));
    if (u.count < 0 || (u.count > 0 && !u.nodes))
      return set_error(DCL_ERR_ARG_NULL, "level set %d: bad irregular-node list", i);
    if (u.count > INT32_MAX) return set_error(DCL_ERR_ARG_SIZ, "level set %d: more than 2^31 - 1 records", i);
    if (u.band_capacity < 0 || (u.band_capacity > 0 && !u.band))
      return set_error(DCL_ERR_ARG_NULL, "level set %d: bad band output", i);
    if (!(u.band_width > 0) || !(u.phi_inf > u.band_width))
      return set_error(DCL_ERR_ARG_WRONG, "level set %d: need 0 < band_width < phi_inf", i);
    const int64_t hc = heap_capacity ? heap_capacity : 2 * u.count + 4096;
    bytes += 2 * (((sizeof(RiEntry) * (size_t)(hc + 1)) + 255) / 256) * 256;
    bytes += 3 * (((sizeof(dcl_icoor) * (size_t)(u.band_capacity > 0 ? u.band_capacity : 1)) + 255) / 256) * 256;
  }
  void* sc = nullptr;
  DCL_TRY(scratch(bytes, &sc));
  char* base = static_cast<char*>(sc);
  RiCell* dcells = reinterpret_cast<RiCell*>(base);
  base += ((sizeof(RiCell) * ncells + 255) / 256) * 256;
  long long* stats = reinterpret_cast<long long*>(base);
  base += ((sizeof(long long) * 4 * ncells + 255) / 256) * 256;
  RiCell* host = static_cast<RiCell*>(malloc(sizeof(RiCell) * ncells));
  long long* hstat = static_cast<long long*>(malloc(sizeof(long long) * 4 * ncells));
  if (!host || !hstat) {
    free(host);
    free(hstat);
    return set_error(DCL_ERR_MEM, "host memory");
  }
  int d3 = -1, mixed = 0;
  for (int i = 0; i < ncells; ++i) {
    const dcl_reinit_cell& u = cells[i];
    RiCell& c = host[i];
    c.phi = static_cast<double*>(u.phi);
    c.w = kwin(u.win);
    c.d3 = u.win.n.z >= 2;
    if (d3 >= 0 && d3 != c.d3) mixed = 1;
    d3 = c.d3;
    c.nodes = u.nodes;
    c.count = u.count;
    c.bw = u.band_width;
    c.inf = u.phi_inf;
    c.heap_cap = heap_capacity ? heap_capacity : 2 * u.count + 4096;
    for (int p = 0; p < 2; ++p) {
      c.heap[p] = reinterpret_cast<RiEntry*>(base);
      base += ((sizeof(RiEntry) * (size_t)(c.heap_cap + 1) + 255) / 256) * 256;
    }
    c.list_cap = u.band_capacity;
    for (int p = 0; p < 3; ++p) {
      c.list[p] = reinterpret_cast<dcl_icoor*>(base);
      base += ((sizeof(dcl_icoor) * (size_t)(u.band_capacity > 0 ? u.band_capacity : 1) + 255) / 256) * 256;
    }
    c.stat = stats + 4 * i;
  }
  int rc = 0;
  if (mixed) rc = set_error(DCL_ERR_ARG_WRONG, "a batch mixes 2D and 3D level sets");
  if (!rc) {
    hipError_t e = hipMemcpyAsync(dcells, host, sizeof(RiCell) * ncells, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(stats, 0, sizeof(long long) * 4 * ncells, s);
    if (e != hipSuccess) rc = hip_fail(e, "reinit upload", __FILE__, __LINE__);
  }
  for (int i = 0; i < ncells && !rc; ++i) {
    const RiCell& c = host[i];
    const int64_t n = win_nodes(cells[i].win);
    const unsigned rb = grid1d(n, 256) < 4096u ? grid1d(n, 256) : 4096u;
    hipLaunchKernelGGL(reinit_reset_k, dim3(rb), dim3(256), 0, s, c.phi, n, c.inf);
    if (c.count > 0) hipLaunchKernelGGL(reinit_bc_k, dim3(grid1d(c.count, 256)), dim3(256), 0, s, c);
    if (hipGetLastError() != hipSuccess) rc = set_error(DCL_ERR_LIB, "reinit reset / boundary launch failed");
  }
  if (!rc) {
    if (d3)
      hipLaunchKernelGGL(reinit_fmm_k<6>, dim3(2 * ncells), dim3(64), 0, s, dcells);
    else
      hipLaunchKernelGGL(reinit_fmm_k<4>, dim3(2 * ncells), dim3(64), 0, s, dcells);
    if (hipGetLastError() != hipSuccess) rc = set_error(DCL_ERR_LIB, "reinit fast marching launch failed");
  }
  if (!rc) {
    hipError_t e = hipMemcpyAsync(hstat, stats, sizeof(long long) * 4 * ncells, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "reinit", __FILE__, __LINE__);
  }
  // the band: [boundary nodes | - pass | + pass], clipped to the capacity
  for (int i = 0; i < ncells && !rc; ++i) {
    const long long* st = hstat + 4 * i;
    if (st[3] == RI_NAN) {
      rc = set_error(DCL_ERR_LIB, "level set %d: NaN in a fast marching update", i);
      break;
    }
    if (st[3] == RI_STEPS) {
      rc = set_error(DCL_ERR_LIB, "level set %d: fast marching did not finish", i);
      break;
    }
    if (st[3] == RI_HEAP) {
      rc = set_error(DCL_ERR_ARG_SIZ, "level set %d: fast marching heap full (capacity %lld)", i,
                     (long long)host[i].heap_cap);
      break;
    }
    dcl_reinit_cell& u = cells[i];
    int64_t at = 0;
    for (int p = 0; p < 3 && !rc; ++p) {
      const int64_t len = st[p];
      const int64_t take = at >= u.band_capacity ? 0 : (len < u.band_capacity - at ? len : u.band_capacity - at);
      if (take > 0) {
        hipError_t e = hipMemcpyAsync(u.band + at, host[i].list[p], sizeof(dcl_icoor) * (size_t)take,
                                      hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) rc = hip_fail(e, "reinit band copy", __FILE__, __LINE__);
      }
      at += len;
    }
    u.band_count = at;
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "reinit", __FILE__, __LINE__);
  }
  free(host);
  free(hstat);
  return rc;
}
