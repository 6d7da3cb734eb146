// scan.hip -- exclusive scan of per-block counts for the compaction kernels
// (the band list, the irregular-node list), in two levels over tiles of
// SCAN_TILE counts: scan_tiles_k scans every tile in place (one block per
// tile, one count per thread) and writes the tile's sum; scan_sums_k scans the
// tile sums in place (one block, a running carry over chunks of SCAN_TILE) and
// writes the grand total.  The exclusive prefix of count i is then
// counts[i] + sums[i / SCAN_TILE].  Each launch scans up to two arrays of the
// same length (blockIdx.y / blockIdx.x picks the array).
#include "dcl_host.hpp"

namespace dcl {

// exclusive scan of v over a block of SCAN_TILE threads; *tot gets the sum.
// sh: 32 entries of LDS, free again when the function returns.
__device__ static unsigned long long scan_block_excl(unsigned long long v, unsigned long long* sh,
                                                     unsigned long long* tot) {
  constexpr int WAVES = SCAN_TILE / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long t = __shfl_up(x, o, 64);
    if (lane >= o) x += t;
  }
  if (lane == 63) sh[wv] = x;
  __syncthreads();
  if (wv == 0) {
    unsigned long long y = lane < WAVES ? sh[lane] : 0ull;
#pragma unroll
    for (int o = 1; o < WAVES; o <<= 1) {
      const unsigned long long t = __shfl_up(y, o, 64);
      if (lane >= o) y += t;
    }
    if (lane < WAVES) sh[WAVES + lane] = y;  // inclusive prefix of the wave sums
  }
  __syncthreads();
  const unsigned long long wave_excl = wv ? sh[WAVES + wv - 1] : 0ull;
  *tot = sh[2 * WAVES - 1];
  __syncthreads();
  return wave_excl + x - v;
}

__global__ __launch_bounds__(SCAN_TILE) void scan_tiles_k(ScanArrays a, int64_t n) {
  __shared__ unsigned long long sh[2 * (SCAN_TILE / 64)];
  unsigned* c = a.counts[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  const unsigned v = i < n ? c[i] : 0u;
  unsigned long long tot;
  const unsigned long long e = scan_block_excl(v, sh, &tot);
  if (i < n) c[i] = (unsigned)e;  // a tile's prefix stays below SCAN_TILE * 2^16
  if (threadIdx.x == 0) a.sums[blockIdx.y][blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_TILE) void scan_sums_k(ScanArrays a, int64_t ntiles) {
  __shared__ unsigned long long sh[2 * (SCAN_TILE / 64)];
  unsigned long long* s = a.sums[blockIdx.x];
  unsigned long long carry = 0;
  for (int64_t base = 0; base < ntiles; base += SCAN_TILE) {
    const int64_t i = base + threadIdx.x;
    const unsigned long long v = i < ntiles ? s[i] : 0ull;
    unsigned long long tot;
    const unsigned long long e = scan_block_excl(v, sh, &tot);
    if (i < ntiles) s[i] = carry + e;
    carry += tot;
  }
  if (threadIdx.x == 0) *a.total[blockIdx.x] = carry;
}

int scan_counts(const ScanArrays& a, int narrays, int64_t n, hipStream_t s) {
  if (narrays < 1 || narrays > 2) return set_error(DCL_ERR_ARG_SIZ, "scan of %d arrays", narrays);
  const int64_t ntiles = scan_tiles(n);
  hipLaunchKernelGGL(scan_tiles_k, dim3((unsigned)ntiles, (unsigned)narrays), dim3(SCAN_TILE), 0, s, a, n);
  DCL_LAUNCHED("scan_tiles");
  hipLaunchKernelGGL(scan_sums_k, dim3((unsigned)narrays), dim3(SCAN_TILE), 0, s, a, ntiles);
  DCL_LAUNCHED("scan_sums");
  return 0;
}

}  // namespace dcl
