// velocity.hip -- the level set's velocity window from a device-resident
// fluid field: LevelSetGetVelocity (LevelSetMethod/Advect.c:145-192).  The
// fluid stores [p, u, v, w] ([p, u, v] in 2D) per node over its whole domain;
// the level set's velocity grid holds (u, v, w) over the level set's window.
// The reference fetches the window's box, clipped to the domain, with a
// Global Arrays get; here one thread per node of that box copies the
// velocity components (U_FACE..W_FACE).  An HBM-bound strided copy: each
// 3D node reads one 32-byte (16-byte in FP32) fluid record, of which the
// velocity is 24 (12) bytes, and writes 24 (12) bytes, per node
// (u, v, w) or, for dcl_get_velocity_planes, one plane per component (the
// layout the FP64 SL kernels gather fastest from).
#include "dcl_host.hpp"

namespace dcl {

constexpr int GVB = 256;

struct GvBox {
  int x0, y0, z0;  // first global node of the box
  int bx, by, bz;  // box extent
  int fx, fy;      // fluid domain extent in x, y
  KWin w;          // the velocity window
};

template <typename T, int DIM, bool PL>
__global__ __launch_bounds__(GVB) void get_velocity_k(const T* __restrict__ fluid, GvBox g, T* __restrict__ vel) {
  const int64_t t = (int64_t)blockIdx.x * GVB + threadIdx.x;
  const int64_t plane = (int64_t)g.bx * g.by;
  if (t >= plane * g.bz) return;
  const int z = (int)(t / plane), r = (int)(t - z * plane);
  const int y = r / g.bx, x = r - y * g.bx;
  const int X = g.x0 + x, Y = g.y0 + y, Z = g.z0 + z;
  const size_t src = (((size_t)Z * g.fy + Y) * g.fx + X) * (DIM + 1);
  const size_t node = ((size_t)(Z - g.w.pz) * g.w.ny + (Y - g.w.py)) * g.w.nx + (X - g.w.px);
  // component c of the node at dst + c * cs
  const size_t dst = PL ? node : node * DIM;
  const size_t cs = PL ? (size_t)g.w.nx * g.w.ny * g.w.nz : 1;
  if (DIM == 3 && sizeof(T) == 8) {
    // a node's record is 32 bytes, 32-byte aligned: two 16-byte loads
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 a = *reinterpret_cast<const d2*>(fluid + src);
    const d2 b = *reinterpret_cast<const d2*>(fluid + src + 2);
    vel[dst] = (T)a.y;
    vel[dst + cs] = (T)b.x;
    vel[dst + 2 * cs] = (T)b.y;
  } else if (DIM == 3) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 a = *reinterpret_cast<const f4*>(fluid + src);
    vel[dst] = (T)a.y;
    vel[dst + cs] = (T)a.z;
    vel[dst + 2 * cs] = (T)a.w;
  } else {
    vel[dst] = fluid[src + 1];
    vel[dst + cs] = fluid[src + 2];
  }
}

template <typename T, bool PL>
static int get_velocity_t(const void* fluid, const dcl_icoor& fn, const dcl_window* win, void* vel, hipStream_t s) {
  const bool d3 = fn.z >= 2;
  const int nz = d3 ? win->n.z : 1;
  GvBox g;
  // the window's box clipped to the domain (Advect.c:160-186)
  g.x0 = win->p.x < 0 ? 0 : win->p.x;
  g.y0 = win->p.y < 0 ? 0 : win->p.y;
  g.z0 = d3 ? (win->p.z < 0 ? 0 : win->p.z) : 0;
  const int x1 = win->p.x + win->n.x < fn.x ? win->p.x + win->n.x : fn.x;
  const int y1 = win->p.y + win->n.y < fn.y ? win->p.y + win->n.y : fn.y;
  const int z1 = d3 ? (win->p.z + nz < fn.z ? win->p.z + nz : fn.z) : 1;
  g.bx = x1 - g.x0;
  g.by = y1 - g.y0;
  g.bz = z1 - g.z0;
  if (g.bx <= 0 || g.by <= 0 || g.bz <= 0) return 0;  // no overlap: nothing to copy
  g.fx = fn.x;
  g.fy = fn.y;
  g.w = kwin(*win);
  if (!d3) {
    g.w.pz = 0;
    g.w.nz = 1;
  }
  const int64_t n = (int64_t)g.bx * g.by * g.bz;
  if (d3)
    hipLaunchKernelGGL((get_velocity_k<T, 3, PL>), dim3(grid1d(n, GVB)), dim3(GVB), 0, s, static_cast<const T*>(fluid),
                       g, static_cast<T*>(vel));
  else
    hipLaunchKernelGGL((get_velocity_k<T, 2, PL>), dim3(grid1d(n, GVB)), dim3(GVB), 0, s, static_cast<const T*>(fluid),
                       g, static_cast<T*>(vel));
  DCL_LAUNCHED("get_velocity");
  return 0;
}

}  // namespace dcl

using namespace dcl;

static int get_velocity(const void* fluid, const dcl_icoor* fluid_n, const dcl_window* win, int precision, void* vel,
                        void* stream, bool planes) {
  DCL_TRY(check_precision(precision));
  if (!fluid || !vel) return set_error(DCL_ERR_ARG_NULL, "fluid or vel is NULL");
  if (!fluid_n) return set_error(DCL_ERR_ARG_NULL, "fluid_n is NULL");
  const bool d3 = fluid_n->z >= 2;
  if (fluid_n->x < 1 || fluid_n->y < 1)
    return set_error(DCL_ERR_ARG_SIZ, "fluid domain %dx%dx%d is empty", fluid_n->x, fluid_n->y, fluid_n->z);
  DCL_TRY(check_window(win, "velocity", 1, d3 ? 3 : 2));
  if (d3 != (win->n.z >= 2))
    return set_error(DCL_ERR_ARG_WRONG, "fluid domain is %dD, velocity window %dD", d3 ? 3 : 2,
                     win->n.z >= 2 ? 3 : 2);
  const int64_t fnodes = (int64_t)fluid_n->x * fluid_n->y * (d3 ? fluid_n->z : 1);
  if (fnodes * (d3 ? 4 : 3) >= (int64_t)1 << 40)
    return set_error(DCL_ERR_ARG_SIZ, "fluid domain too large");
  if (d3 && (reinterpret_cast<uintptr_t>(fluid) % (precision == DCL_F64 ? 32 : 16)) != 0)
    return set_error(DCL_ERR_ARG_WRONG, "fluid must be aligned to one node record (%d bytes)",
                     precision == DCL_F64 ? 32 : 16);
  hipStream_t s = as_stream(stream);
  if (precision == DCL_F64)
    return planes ? get_velocity_t<double, true>(fluid, *fluid_n, win, vel, s)
                  : get_velocity_t<double, false>(fluid, *fluid_n, win, vel, s);
  return planes ? get_velocity_t<float, true>(fluid, *fluid_n, win, vel, s)
                : get_velocity_t<float, false>(fluid, *fluid_n, win, vel, s);
}

extern "C" int dcl_get_velocity(const void* fluid, const dcl_icoor* fluid_n, const dcl_window* win, int precision,
                                void* vel, void* stream) {
  return get_velocity(fluid, fluid_n, win, precision, vel, stream, false);
}

extern "C" int dcl_get_velocity_planes(const void* fluid, const dcl_icoor* fluid_n, const dcl_window* win,
                                       int precision, void* vel, void* stream) {
  return get_velocity(fluid, fluid_n, win, precision, vel, stream, true);
}
