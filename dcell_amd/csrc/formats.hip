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
// field), 8-byte loads and stores.  The 2D rows skip records, so their
// positions come from a scan of per-block counts.
#include "dcl_host.hpp"

namespace dcl {

constexpr int FB = 256;
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

// the row of record r (IrregularNodes.c:145-172 / 178-219); extra: the IIM
// fields the row carries that this library does not compute (3D: f1, f1_n,
// f1_t, f1_nn, f1_tt, f1_nt, numNei; 2D: f1, fa1, f1_n, f1_nn), zero when NULL
template <int DIM>
__device__ __forceinline__ double row_field(const dcl_irregular_node& n, const dcl_surface& s,
                                            const double* __restrict__ extra, int64_t r, int f) {
  constexpr int NX = DIM == 3 ? 7 : 4;
  auto ex = [&](int k) { return extra ? extra[r * NX + k] : 0.0; };
  if (DIM == 3) {
    switch (f) {
      case 0: return n.X.x;
      case 1: return n.X.y;
      case 2: return n.X.z;
      case 3: return n.pos.x;
      case 4: return n.pos.y;
      case 5: return n.pos.z;
      case 6: return s.n.x;
      case 7: return s.n.y;
      case 8: return s.n.z;
      case 9: return s.k;
      case 10: return ex(0);  // f1
      case 11: return ex(1);  // f1_n
      case 12: return ex(2);  // f1_t
      case 13: return ex(3);  // f1_nn
      case 14: return ex(4);  // f1_tt
      case 15: return ex(5);  // f1_nt
      case 16: return s.s.x;
      case 17: return s.s.y;
      case 18: return s.s.z;
      case 19: return s.r.x;
      case 20: return s.r.y;
      case 21: return s.r.z;
      default: return ex(6);  // numNei
    }
  }
  switch (f) {
    case 0: return n.X.x;
    case 1: return n.X.y;
    case 2: return n.pos.x;
    case 3: return n.pos.y;
    case 4: return s.n.x;
    case 5: return s.n.y;
    case 6: return ex(0);  // f1
    case 7: return ex(1);  // fa1
    case 8: return s.k;
    case 9: return ex(2);  // f1_n
    default: return ex(3);  // f1_nn
  }
}

// 2D: crossing records per block of FB records
__global__ __launch_bounds__(FB) void rows_count_k(const dcl_irregular_node* __restrict__ nodes, int64_t count,
                                                   unsigned* __restrict__ counts) {
  const int64_t r = (int64_t)blockIdx.x * FB + threadIdx.x;
  const int c = __syncthreads_count(r < count && nodes[r].axis != -1);
  if (threadIdx.x == 0) counts[blockIdx.x] = (unsigned)c;
}

// one block per FB records: each kept record's row at its position, one
// thread per field over the block's rows
template <int DIM>
__global__ __launch_bounds__(FB) void rows_write_k(const dcl_irregular_node* __restrict__ nodes,
                                                   const dcl_surface* __restrict__ surf,
                                                   const double* __restrict__ extra, int64_t count,
                                                   const unsigned* __restrict__ offsets,
                                                   const unsigned long long* __restrict__ tile_sums,
                                                   unsigned long long* __restrict__ dst, int64_t cap) {
  constexpr int NF = DIM == 3 ? 23 : 11;
  __shared__ int slot[FB];  // the block's kept records, in order
  __shared__ unsigned wsum[FB / 64];
  const int64_t r0 = (int64_t)blockIdx.x * FB;
  const int64_t r = r0 + threadIdx.x;
  const bool keep = r < count && (DIM == 3 || nodes[r].axis != -1);
  const unsigned long long bal = __ballot(keep);
  const unsigned lane_pre =
      __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = (unsigned)__popcll(bal);
  __syncthreads();
  unsigned pre = 0;
  for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) pre += wsum[w];
  const int kept = (int)(wsum[0] + wsum[1] + wsum[2] + wsum[3]);
  if (keep) slot[pre + lane_pre] = threadIdx.x;
  __syncthreads();
  const int64_t base = DIM == 3 ? r0 : (int64_t)(offsets[blockIdx.x] + tile_sums[blockIdx.x / SCAN_TILE]);
  for (int q = threadIdx.x; q < kept * NF; q += FB) {
    const int k = q / NF, f = q - k * NF;
    const int64_t row = base + k;
    if (row >= cap) continue;
    const int64_t rr = r0 + slot[k];
    dst[row * NF + f] = be64(row_field<DIM>(nodes[rr], surf[rr], extra, rr, f));
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
  const int64_t nb = (count + FB - 1) / FB;
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
  hipLaunchKernelGGL(rows_count_k, dim3((unsigned)nb), dim3(FB), 0, s, nodes, count, counts);
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
