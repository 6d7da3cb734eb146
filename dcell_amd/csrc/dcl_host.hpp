// dcl_host.hpp -- host-side helpers shared by the C-ABI translation units:
// the per-thread error message, HIP error mapping and argument checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "dcell_lsm.h"
#include "dcl_common.hpp"

namespace dcl {

int set_error(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// window element count (times dof) must fit the kernels' 32-bit offsets
inline int64_t win_nodes(const dcl_window& w) {
  const int64_t nz = w.n.z < 2 ? 1 : w.n.z;
  return (int64_t)w.n.x * w.n.y * nz;
}

int check_window(const dcl_window* w, const char* name, int min_n, int dof = 1);
// planes: the entry point also takes component planes (dof 1)
int check_velocity(const dcl_velocity* v, int dim, bool planes = false);
int check_precision(int precision);

inline unsigned grid1d(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// compute units of the current device (queried once per device; 256 if the
// query fails)
int device_cus();

// Device scratch owned by the calling host thread on the current device
// (grown on demand, kept until the process ends).  Only the synchronous
// entry points (reductions with host results) use it, and they synchronize
// their stream before returning, so the buffer is free again on return.
int scratch(size_t bytes, void** ptr);

// Exclusive scan of per-block counts (scan.hip), in two levels over tiles of
// SCAN_TILE counts: afterwards the exclusive prefix of count i is
// counts[i] + sums[i / SCAN_TILE] and *total is the sum.  A tile's prefix must
// stay below 2^32 (counts below 2^22).  Asynchronous on stream s; scans
// narrays (1 or 2) arrays of the same length n at once.
constexpr int SCAN_TILE = 1024;
struct ScanArrays {
  unsigned* counts[2];
  unsigned long long* sums[2];   // scan_tiles(n) entries each
  unsigned long long* total[2];  // one entry each
};
inline int64_t scan_tiles(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }
int scan_counts(const ScanArrays& a, int narrays, int64_t n, hipStream_t s);

}  // namespace dcl

#define DCL_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e__ = (call);                                                   \
    if (e__ != hipSuccess) return dcl::hip_fail(e__, #call, __FILE__, __LINE__); \
  } while (0)

#define DCL_LAUNCHED(name)                                                     \
  do {                                                                         \
    hipError_t e__ = hipGetLastError();                                        \
    if (e__ != hipSuccess) return dcl::hip_fail(e__, name, __FILE__, __LINE__); \
  } while (0)

#define DCL_TRY(call)          \
  do {                         \
    int rc__ = (call);         \
    if (rc__) return rc__;     \
  } while (0)
