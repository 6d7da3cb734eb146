// formats.hip -- DCell's output formats packed on gfx950, so a device-resident
// grid or node list leaves the GPU as the file's bytes in one copy.
//
//   dcl_real64_pack        the PETSc binary vector GridWrite writes
//                          (Common/Serialize.c:23-55: VecView to a binary
//                          viewer): a big-endian int32 class id (1211214)
//                          and int32 count, then the values as big-endian
//                          doubles -- what Visualization/DCellVTK.py:24-32
//                          reads back (8-byte header, '>f8').
//   dcl_irregular_node_rows the rows of LevelSetWriteIrregularNodeList_2D /
//                          _3D (LevelSetMethod/IrregularNodes.c:145-219),
//                          big-endian doubles: 3D, 23 per record, every
//                          record; 2D, 11 per record, the crossing records
//                          only (axis != -1).
//
// Both are byte rearrangements, bound by HBM: one thread per value (per row
// field), 8-byte loads and stores; the rows' records and surface quantities
// pass through LDS so that their loads stay contiguous.  The 2D rows skip
// records, so their positions come from a scan of per-block counts.
#include "dcl_host.hpp"

namespace dcl {

constexpr int FB = 256;
constexpr int RB = 64;  // node-list records per block of the row kernels (one wave)
constexpr unsigned REAL64_CLASSID = 1211214u;  // PETSc's VEC_FILE_CLASSID

__device__ __forceinline__ unsigned long long be64(double v) {
  return __builtin_bswap64((unsigned long long)__double_as_longlong(v));
}

template <typename T>
__global__ __launch_bounds__(FB) void real64_pack_k(const T* __restrict__ src, int64_t count,
                                                    unsigned long long* __restrict__ dst) {
  // dst[0] is the header: class id and count, each a big-endian int32
  const int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x;
  if (i == 0) {
    const unsigned long long h = (unsigned long long)__builtin_bswap32(REAL64_CLASSID) |
                                 ((unsigned long long)__builtin_bswap32((unsigned)count) << 32);
    dst[0] = h;
  }
  if (i < count) dst[1 + i] = be64((double)src[i]);
}

// where row field f of a record comes from (IrregularNodes.c:145-172 /
// 178-219): the record's 8-byte word w (code w), its int32 word w (code 16 +
// w: pos), the surface's 8-byte word w (code 32 + w), or the IIM field w the
// row carries that this library does not compute (code 64 + w; 3D: f1, f1_n,
// f1_t, f1_nn, f1_tt, f1_nt, numNei; 2D: f1, fa1, f1_n, f1_nn), zero when
// extra is NULL.  Record words: pos at int32 0-2, X at doubles 5-7; surface
// words: n 0-2, s 3-5, r 6-8, k 9.
template <int DIM>
__device__ __forceinline__ int row_source(int f) {
  if (DIM == 3) {
    if (f < 3) return 5 + f;                    // X
    if (f < 6) return 16 + (f - 3);             // pos
    if (f < 9) return 32 + (f - 6);             // n
    if (f == 9) return 32 + 9;                  // k
    if (f < 16) return 64 + (f - 10);           // f1 .. f1_nt
    if (f < 22) return 32 + 3 + (f - 16);       // s, r
    return 64 + 6;                              // numNei
  }
  if (f < 2) return 5 + f;                      // X
  if (f < 4) return 16 + (f - 2);               // pos
  if (f < 6) return 32 + (f - 4);               // n
  if (f < 8) return 64 + (f - 6);               // f1, fa1
  if (f == 8) return 32 + 9;                    // k
  return 64 + 2 + (f - 9);                      // f1_n, f1_nn
}

constexpr int REC_W = sizeof(dcl_irregular_node) / 8;  // 11
constexpr int SURF_W = sizeof(dcl_surface) / 8;        // 13
static_assert(sizeof(dcl_irregular_node) == 88 && sizeof(dcl_surface) == 104, "record layouts");

// 2D: crossing records per block of RB records (one wave)
__global__ __launch_bounds__(RB) void rows_count_k(const dcl_irregular_node* __restrict__ nodes, int64_t count,
                                                   unsigned* __restrict__ counts) {
  const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
  const unsigned long long bal = __ballot(r < count && nodes[r].axis != -1);
  if (threadIdx.x == 0) counts[blockIdx.x] = (unsigned)__popcll(bal);
}

// one block of FB threads per RB records: the block's records and surface
// quantities are staged in LDS with contiguous 8-byte loads, then each kept
// record's row is written at its position, one thread per field over the
// block's rows (the stores contiguous)
template <int DIM>
__global__ __launch_bounds__(FB) void rows_write_k(const dcl_irregular_node* __restrict__ nodes,
                                                   const dcl_surface* __restrict__ surf,
                                                   const double* __restrict__ extra, int64_t count,
                                                   const unsigned* __restrict__ offsets,
                                                   const unsigned long long* __restrict__ tile_sums,
                                                   unsigned long long* __restrict__ dst, int64_t cap) {
  constexpr int NF = DIM == 3 ? 23 : 11;
  constexpr int NX = DIM == 3 ? 7 : 4;
  __shared__ double lrec[RB * REC_W];
  __shared__ double lsurf[RB * SURF_W];
  __shared__ int slot[RB];  // the block's kept records, in order
  __shared__ int kept_s;
  const int64_t r0 = (int64_t)blockIdx.x * RB;
  const int nrec = (int)(count - r0 < RB ? count - r0 : RB);
  const double* grec = reinterpret_cast<const double*>(nodes + r0);
  const double* gsurf = reinterpret_cast<const double*>(surf + r0);
  for (int i = threadIdx.x; i < nrec * REC_W; i += FB) lrec[i] = grec[i];
  for (int i = threadIdx.x; i < nrec * SURF_W; i += FB) lsurf[i] = gsurf[i];
  __syncthreads();
  if (threadIdx.x < RB) {  // wave 0: the kept records in order
    const bool keep =
        (int)threadIdx.x < nrec &&
        (DIM == 3 || reinterpret_cast<const dcl_irregular_node*>(lrec)[threadIdx.x].axis != -1);
    const unsigned long long bal = __ballot(keep);
    const unsigned pre =
        __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
    if (keep) slot[pre] = threadIdx.x;
    if (threadIdx.x == 0) kept_s = __popcll(bal);
  }
  __syncthreads();
  const int kept = kept_s;
  const int64_t base = DIM == 3 ? r0 : (int64_t)(offsets[blockIdx.x] + tile_sums[blockIdx.x / SCAN_TILE]);
  for (int q = threadIdx.x; q < kept * NF; q += FB) {
    const int k = q / NF, f = q - k * NF;
    const int64_t row = base + k;
    if (row >= cap) continue;
    const int j = slot[k];
    const int src = row_source<DIM>(f);
    double v;
    if (src < 16)
      v = lrec[j * REC_W + src];
    else if (src < 32)
      v = (double)reinterpret_cast<const int*>(lrec + j * REC_W)[src - 16];
    else if (src < 64)
      v = lsurf[j * SURF_W + (src - 32)];
    else
      v = extra ? extra[(r0 + j) * NX + (src - 64)] : 0.0;
    dst[row * NF + f] = be64(v);
  }
}

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_real64_pack(const void* src, int precision, int64_t count, void* dst, void* stream) {
  DCL_TRY(check_precision(precision));
  if (count < 0 || count > INT32_MAX) return set_error(DCL_ERR_ARG_SIZ, "count must be in [0, 2^31 - 1]");
  if (!dst || (count > 0 && !src)) return set_error(DCL_ERR_ARG_NULL, "src or dst is NULL");
  hipStream_t s = as_stream(stream);
  const unsigned nb = count > 0 ? grid1d(count, FB) : 1u;
  if (precision == DCL_F64)
    hipLaunchKernelGGL(real64_pack_k<double>, dim3(nb), dim3(FB), 0, s, static_cast<const double*>(src), count,
                       static_cast<unsigned long long*>(dst));
  else
    hipLaunchKernelGGL(real64_pack_k<float>, dim3(nb), dim3(FB), 0, s, static_cast<const float*>(src), count,
                       static_cast<unsigned long long*>(dst));
  DCL_LAUNCHED("real64_pack");
  return 0;
}

extern "C" int dcl_irregular_node_rows(const dcl_irregular_node* nodes, const dcl_surface* surf,
                                       const double* extra, int64_t count, int dim, void* dst, int64_t capacity,
                                       int64_t* rows_out, void* stream) {
  if (dim != 2 && dim != 3) return set_error(DCL_ERR_ARG_WRONG, "dim must be 2 or 3");
  if (count < 0 || count > INT32_MAX) return set_error(DCL_ERR_ARG_SIZ, "count must be in [0, 2^31 - 1]");
  if (capacity < 0) return set_error(DCL_ERR_ARG_SIZ, "capacity < 0");
  if (!rows_out) return set_error(DCL_ERR_ARG_NULL, "rows_out is NULL");
  if (count > 0 && (!nodes || !surf)) return set_error(DCL_ERR_ARG_NULL, "nodes or surf is NULL");
  if (capacity > 0 && !dst) return set_error(DCL_ERR_ARG_NULL, "dst is NULL");
  hipStream_t s = as_stream(stream);
  if (count == 0) {
    *rows_out = 0;
    return 0;
  }
  const int64_t nb = (count + RB - 1) / RB;
  unsigned long long* out = static_cast<unsigned long long*>(dst);
  if (dim == 3) {
    if (capacity > 0)
      hipLaunchKernelGGL(rows_write_k<3>, dim3((unsigned)nb), dim3(FB), 0, s, nodes, surf, extra, count, nullptr,
                         nullptr, out, capacity);
    DCL_LAUNCHED("irregular_rows");
    DCL_HIP(hipStreamSynchronize(s));
    *rows_out = count;
    return 0;
  }
  const int64_t nt = scan_tiles(nb);
  void* sc = nullptr;
  DCL_TRY(scratch(16 + sizeof(unsigned long long) * (size_t)nt + sizeof(unsigned) * (size_t)nb, &sc));
  unsigned long long* total = static_cast<unsigned long long*>(sc);
  unsigned long long* sums = total + 2;
  unsigned* counts = reinterpret_cast<unsigned*>(sums + nt);
  hipLaunchKernelGGL(rows_count_k, dim3((unsigned)nb), dim3(RB), 0, s, nodes, count, counts);
  DCL_LAUNCHED("irregular_rows_count");
  const ScanArrays sa = {{counts, nullptr}, {sums, nullptr}, {total, nullptr}};
  DCL_TRY(scan_counts(sa, 1, nb, s));
  if (capacity > 0)
    hipLaunchKernelGGL(rows_write_k<2>, dim3((unsigned)nb), dim3(FB), 0, s, nodes, surf, extra, count, counts, sums,
                       out, capacity);
  DCL_LAUNCHED("irregular_rows");
  unsigned long long tot = 0;
  DCL_HIP(hipMemcpyAsync(&tot, total, sizeof(tot), hipMemcpyDeviceToHost, s));
  DCL_HIP(hipStreamSynchronize(s));
  *rows_out = (int64_t)tot;
  return 0;
}
