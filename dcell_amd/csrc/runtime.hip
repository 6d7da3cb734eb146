// runtime.hip -- error reporting, argument checks and the memory helpers of
// the C ABI (dcell_lsm.h).
#include <atomic>
#include <stdio.h>
#include <string.h>

#include "dcl_host.hpp"

namespace dcl {

static thread_local char g_msg[512] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  return set_error(DCL_ERR_LIB, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

int check_window(const dcl_window* w, const char* name, int min_n, int dof) {
  if (!w) return set_error(DCL_ERR_ARG_NULL, "%s window is NULL", name);
  const bool two = w->n.z < 2;
  if (w->n.x < min_n || w->n.y < min_n || (!two && w->n.z < min_n))
    return set_error(DCL_ERR_ARG_SIZ, "%s window %dx%dx%d: every axis needs at least %d nodes",
                     name, w->n.x, w->n.y, w->n.z, min_n);
  if (win_nodes(*w) * dof >= (int64_t)1 << 31)
    return set_error(DCL_ERR_ARG_SIZ, "%s window has %lld elements; the limit is 2^31 - 1", name,
                     (long long)(win_nodes(*w) * dof));
  if (!(w->d.x > 0) || !(w->d.y > 0) || (!two && !(w->d.z > 0)))
    return set_error(DCL_ERR_ARG_WRONG, "%s window spacing must be positive", name);
  return 0;
}

int check_velocity(const dcl_velocity* v, int dim, bool planes) {
  if (!v) return set_error(DCL_ERR_ARG_NULL, "velocity is NULL");
  if (!v->data) return set_error(DCL_ERR_ARG_NULL, "velocity data is NULL");
  if ((v->win.n.z < 2 ? 2 : 3) != dim)
    return set_error(DCL_ERR_ARG_WRONG, "velocity window is %dD, phi is %dD",
                     v->win.n.z < 2 ? 2 : 3, dim);
  if (v->dof == 1) {
    if (!planes) return set_error(DCL_ERR_SUP, "this entry point takes velocity components stored per node");
    if (v->udof != 0) return set_error(DCL_ERR_ARG_WRONG, "component planes (dof 1) take udof 0, not %d", v->udof);
    return check_window(&v->win, "velocity", 2, dim);
  }
  if (v->udof < 0 || v->udof + dim > v->dof)
    return set_error(DCL_ERR_ARG_WRONG, "velocity dof %d cannot hold %d components from udof %d",
                     v->dof, dim, v->udof);
  return check_window(&v->win, "velocity", 2, v->dof);
}

int check_precision(int precision) {
  if (precision != DCL_F64 && precision != DCL_F32)
    return set_error(DCL_ERR_ARG_WRONG, "precision %d is neither DCL_F64 nor DCL_F32", precision);
  return 0;
}

int device_cus() {
  static std::atomic<int> cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = cached[dev].load(std::memory_order_relaxed);
  if (c > 0) return c;
  if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
  cached[dev].store(c, std::memory_order_relaxed);
  return c;
}

struct ScratchSlot {
  void* ptr = nullptr;
  size_t bytes = 0;
};

int scratch(size_t bytes, void** ptr) {
  static thread_local ScratchSlot slots[16];
  int dev = 0;
  DCL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 16) return set_error(DCL_ERR_SUP, "device %d out of range", dev);
  ScratchSlot& s = slots[dev];
  if (s.bytes < bytes) {
    if (s.ptr) {
      DCL_HIP(hipDeviceSynchronize());
      DCL_HIP(hipFree(s.ptr));
      s.ptr = nullptr;
      s.bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    hipError_t e = hipMalloc(&s.ptr, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return set_error(DCL_ERR_MEM, "scratch hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    s.bytes = want;
  }
  *ptr = s.ptr;
  return 0;
}

// Streaming device copy: 16-byte elements, each thread keeps SC_U loads in
// flight (non-temporal loads and stores: the data is touched once), grid
// stride over the buffer.
constexpr int SC_U = 4, SC_BLOCKS = 16384;
typedef unsigned int sc_u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void stream_copy_k(const sc_u32x4* __restrict__ src, sc_u32x4* __restrict__ dst,
                                                     int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + (SC_U - 1) * stride < n; i += SC_U * stride) {
    sc_u32x4 v[SC_U];
#pragma unroll
    for (int u = 0; u < SC_U; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < SC_U; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace dcl

using namespace dcl;

extern "C" {

int dcl_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
  if (!bytes) return 0;
  if (!dst || !src) return set_error(DCL_ERR_ARG_NULL, "dst or src is NULL");
  if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
    return set_error(DCL_ERR_ARG_WRONG, "dcl_stream_copy needs 16-byte aligned buffers and sizes");
  const int64_t n = (int64_t)(bytes / 16);
  const int64_t need = (n + 255) / 256;
  const unsigned blocks = (unsigned)(need < SC_BLOCKS ? need : SC_BLOCKS);
  hipLaunchKernelGGL(stream_copy_k, dim3(blocks), dim3(256), 0, as_stream(stream),
                     static_cast<const sc_u32x4*>(src), static_cast<sc_u32x4*>(dst), n);
  DCL_LAUNCHED("stream_copy");
  return 0;
}


int dcl_abi_version(void) { return DCL_ABI_VERSION; }

const char* dcl_last_error(void) { return g_msg; }

int dcl_device_count(int* count) {
  if (!count) return set_error(DCL_ERR_ARG_NULL, "count is NULL");
  DCL_HIP(hipGetDeviceCount(count));
  return 0;
}

int dcl_malloc(void** ptr, size_t bytes) {
  if (!ptr) return set_error(DCL_ERR_ARG_NULL, "ptr is NULL");
  *ptr = nullptr;
  if (bytes == 0) return 0;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();
    return set_error(DCL_ERR_MEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  return 0;
}

int dcl_free(void* ptr) {
  if (ptr) DCL_HIP(hipFree(ptr));
  return 0;
}

int dcl_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
  if (!bytes) return 0;
  DCL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
  return 0;
}

int dcl_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
  if (!bytes) return 0;
  DCL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
  DCL_HIP(hipStreamSynchronize(as_stream(stream)));
  return 0;
}

int dcl_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  if (!bytes) return 0;
  DCL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return 0;
}

int dcl_host_malloc(void** ptr, size_t bytes) {
  if (!ptr) return set_error(DCL_ERR_ARG_NULL, "ptr is NULL");
  *ptr = nullptr;
  if (bytes == 0) return 0;
  hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();
    return set_error(DCL_ERR_MEM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  return 0;
}

int dcl_host_free(void* ptr) {
  if (ptr) DCL_HIP(hipHostFree(ptr));
  return 0;
}

int dcl_stream_synchronize(void* stream) {
  DCL_HIP(hipStreamSynchronize(as_stream(stream)));
  return 0;
}

}  // extern "C"
