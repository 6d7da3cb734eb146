// irregular.hip -- the irregular-node list on gfx950:
// LevelSetUpdateIrregularNodeList (LevelSetMethod/IrregularNodes.c:8-140,
// LevelSetMethod/IrregularNodes3D.c:6-117) with the orthogonal projections of
// LevelSetMethod/OrthogonalProjection3D.c:6-208 and
// LevelSetMethod/OrthogonalProjection.c:6-184.
//
// Per band node: a projection record when phi changes sign towards one of the
// 6 (4) axis neighbours -- Newton on the tricubic (bicubic) polynomial of each
// of the 8 (4) cells around the node, the nearest root in its cell, with the
// quadratic / linear fallbacks -- then one record per sign change towards -x,
// -y, -z and per sign change of the face averages.  The records are written in
// the reference's append order: a count pass (records and projections per
// block, one thread per band node), an exclusive scan of the block counts, a
// write pass that places each node's records at its block's offset plus its
// prefix in the block (staged in LDS and copied out coalesced) and lists the
// projections as tasks, then a projection pass in which a group of 8 (4) lanes
// takes a task, one lane per cell.
//
// The tricubic (bicubic) polynomial of TricubicInverse.h /
// TricubicInterpolation.h is evaluated in its separable Catmull-Rom form (the
// same polynomial).  All arithmetic is double, one IEEE operation at a time
// (contraction off) in a fixed order, so the Newton iterations take the same
// path on every run and every device.  Stencil reads clamp the index into
// the window; the reference requires band nodes away from its edge
// (IrregularNodes.c:5) and reads unchecked.
#include <float.h>
#include <stddef.h>

#include "dcl_host.hpp"

namespace dcl {

struct IrrP {
  KWin w;
};

// STAR (Common/Common.h:49-51): +x, -x, +y, -y, +z, -z; component c of entry m
__host__ __device__ constexpr int irr_star(int m, int c) { return (m >> 1) == c ? ((m & 1) ? -1 : 1) : 0; }
// STAGGERED_GRID (Common/Common.h:42-47): none, -x, -y, -z
__host__ __device__ constexpr int irr_stag(int f, int c) { return f == c + 1 ? -1 : 0; }

template <typename T>
__device__ __forceinline__ double irr_at(const T* __restrict__ phi, const KWin& w, int x, int y, int z) {
  const int lx = clampi(x - w.px, 0, w.nx - 1), ly = clampi(y - w.py, 0, w.ny - 1);
  const int lz = w.nz < 2 ? 0 : clampi(z - w.pz, 0, w.nz - 1);
  return (double)phi[((size_t)lz * w.ny + ly) * w.nx + lx];
}

__device__ __forceinline__ void irr_crw(double t, double w[4], double d[4]) {
#pragma clang fp contract(off)
  const double t2 = t * t, t3 = t2 * t;
  w[0] = 0.5 * ((2.0 * t2 - t3) - t);
  w[1] = 0.5 * ((3.0 * t3 - 5.0 * t2) + 2.0);
  w[2] = 0.5 * ((4.0 * t2 + t) - 3.0 * t3);
  w[3] = 0.5 * (t3 - t2);
  d[0] = 0.5 * ((4.0 * t - 3.0 * t2) - 1.0);
  d[1] = 0.5 * (9.0 * t2 - 10.0 * t);
  d[2] = 0.5 * ((8.0 * t + 1.0) - 9.0 * t2);
  d[3] = 0.5 * (3.0 * t2 - 2.0 * t);
}

// The node's neighbourhood is staged in LDS: nb[(c * 5 + b) * 5 + a] holds
// phi at node offset (a - 2, b - 2, c - 2) (nb[b * 5 + a] in 2D), read with
// clamped indices.  Cell (i, j, k) of the node (i, j, k in {0, 1}) has its
// 4x4x4 stencil at offsets (i + a, j + b, k + c), a, b, c in 0..3.

// p, px, py, pz at (x, y, z) of the tricubic of cell (i, j, k)
__device__ __forceinline__ void irr_tricubic(const double* nb, int i, int j, int k, double x, double y, double z,
                                             double e[4]) {
#pragma clang fp contract(off)
  double wx[4], dx[4], wy[4], dy[4], wz[4], dz[4];
  irr_crw(x, wx, dx);
  irr_crw(y, wy, dy);
  irr_crw(z, wz, dz);
  double p = 0.0, px = 0.0, py = 0.0, pz = 0.0;
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {
    double sp = 0.0, spx = 0.0, spy = 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const double* row = nb + ((k + c) * 5 + j + b) * 5 + i;
      double rx = 0.0, rdx = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double f = row[a];
        rx = rx + wx[a] * f;
        rdx = rdx + dx[a] * f;
      }
      sp = sp + wy[b] * rx;
      spx = spx + wy[b] * rdx;
      spy = spy + dy[b] * rx;
    }
    p = p + wz[c] * sp;
    px = px + wz[c] * spx;
    py = py + wz[c] * spy;
    pz = pz + dz[c] * sp;
  }
  e[0] = p;
  e[1] = px;
  e[2] = py;
  e[3] = pz;
}

// p, px, py at (x, y) of the bicubic of cell (i, j)
__device__ __forceinline__ void irr_bicubic(const double* nb, int i, int j, double x, double y, double e[3]) {
#pragma clang fp contract(off)
  double wx[4], dx[4], wy[4], dy[4];
  irr_crw(x, wx, dx);
  irr_crw(y, wy, dy);
  double p = 0.0, px = 0.0, py = 0.0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const double* row = nb + (j + b) * 5 + i;
    double rx = 0.0, rdx = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const double f = row[a];
      rx = rx + wx[a] * f;
      rdx = rdx + dx[a] * f;
    }
    p = p + wy[b] * rx;
    px = px + wy[b] * rdx;
    py = py + dy[b] * rx;
  }
  e[0] = p;
  e[1] = px;
  e[2] = py;
}

// OrthogonalProjection3D_Linear (OrthogonalProjection3D.c:186-208)
__device__ void irr_proj3_linear(const double s[27], double op[3]) {
#pragma clang fp contract(off)
  double mindist = 2.0;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const double nb = s[((1 + irr_star(m, 2)) * 3 + 1 + irr_star(m, 1)) * 3 + 1 + irr_star(m, 0)];
    if (s[13] * nb <= 0.0) {
      const double dist = s[13] / (s[13] - nb);
      if (dist < mindist) {
        mindist = dist;
        op[0] = dist * irr_star(m, 0);
        op[1] = dist * irr_star(m, 1);
        op[2] = dist * irr_star(m, 2);
      }
    }
  }
}

// OrthogonalProjection3D_Quadratic (OrthogonalProjection3D.c:130-182); s[(k*3+j)*3+i]
__device__ void irr_proj3_quadratic(const double s[27], double op[3]) {
#pragma clang fp contract(off)
#define S3(k, j, i) s[((k) * 3 + (j)) * 3 + (i)]
  const double p = S3(1, 1, 1);
  const double px = (S3(1, 1, 2) - S3(1, 1, 0)) / 2.0;
  const double py = (S3(1, 2, 1) - S3(1, 0, 1)) / 2.0;
  const double pz = (S3(2, 1, 1) - S3(0, 1, 1)) / 2.0;
  const double px2 = px * px, py2 = py * py, pz2 = pz * pz;
  const double pxx = (S3(1, 1, 2) - 2.0 * p) + S3(1, 1, 0);
  const double pyy = (S3(1, 2, 1) - 2.0 * p) + S3(1, 0, 1);
  const double pzz = (S3(2, 1, 1) - 2.0 * p) + S3(0, 1, 1);
  const double pxy = (((S3(1, 0, 0) + S3(1, 2, 2)) - S3(1, 2, 0)) - S3(1, 0, 2)) / 4.0;
  const double pxz = (((S3(0, 1, 0) + S3(2, 1, 2)) - S3(2, 1, 0)) - S3(0, 1, 2)) / 4.0;
  const double pyz = (((S3(0, 0, 1) + S3(2, 2, 1)) - S3(2, 0, 1)) - S3(0, 2, 1)) / 4.0;
#undef S3
  const double pn = sqrt((px2 + py2) + pz2);
  const double c = (((((px2 * pxx + py2 * pyy) + pz2 * pzz) + 2.0 * px * py * pxy) + 2.0 * px * pz * pxz) +
                    2.0 * py * pz * pyz) /
                   ((px2 + py2) + pz2);
  double a = 0.0;
  for (int m = 0; m < 5; ++m) {
    const double f = (p + pn * a) + 0.5 * a * a * c;
    const double df = pn + a * c;
    a = a - f / df;
  }
  if (fabs(a) < 1.0) {
    op[0] = a * px / pn;
    op[1] = a * py / pn;
    op[2] = a * pz / pn;
  } else {
    irr_proj3_linear(s, op);
  }
}

// One cell of OrthogonalProjection3D (OrthogonalProjection3D.c:6-126):
// Newton on the tricubic of cell q (q = (k * 2 + j) * 2 + i, the reference's
// loop order) from the node's corner of the cell.  Returns the distance to
// the root (DBL_MAX when Newton does not converge within MAXITER or the root
// leaves the cell) and its offset from the node in op.
__device__ double irr_cell3(const double* nb, int q, double op[3]) {
#pragma clang fp contract(off)
  const double TOL = 1e-3, TOL2 = TOL * TOL;
  const int MAXITER = 20;
  const int k = q >> 2, j = (q >> 1) & 1, i = q & 1;
  const double x0x = 1 - i, x0y = 1 - j, x0z = 1 - k;
  double x = x0x, y = x0y, z = x0z;
  int m;
#pragma unroll 1
  for (m = 0; m < MAXITER; ++m) {
    double e[4];
    irr_tricubic(nb, i, j, k, x, y, z, e);
    const double p = e[0], px = e[1], py = e[2], pz = e[3];
    const double pp = (px * px + py * py) + pz * pz;
    const double d1x = -p * px / pp, d1y = -p * py / pp, d1z = -p * pz / pp;
    const double xp = ((x0x - x) * px + (x0y - y) * py) + (x0z - z) * pz;
    const double d2x = (x0x - x) - xp / pp * px;
    const double d2y = (x0y - y) - xp / pp * py;
    const double d2z = (x0z - z) - xp / pp * pz;
    x += d1x + 0.5 * d2x;
    y += d1y + 0.5 * d2y;
    z += d1z + 0.5 * d2z;
    const double norm = ((((d1x * d1x + d1y * d1y) + d1z * d1z) + d2x * d2x) + d2y * d2y) + d2z * d2z;
    if (norm < TOL2) break;
  }
  op[0] = x + i - 1;
  op[1] = y + j - 1;
  op[2] = z + k - 1;
  if (m == MAXITER) return DBL_MAX;
  if (x < -TOL || 1 + TOL < x || y < -TOL || 1 + TOL < y || z < -TOL || 1 + TOL < z) return DBL_MAX;
  return sqrt(((x - x0x) * (x - x0x) + (y - x0y) * (y - x0y)) + (z - x0z) * (z - x0z));
}

// OrthogonalProjection2D_Linear (OrthogonalProjection.c:174-193); s[j*3+i]
__device__ void irr_proj2_linear(const double s[9], double op[3], double* mindist) {
#pragma clang fp contract(off)
  *mindist = 2.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const double nb = s[(1 + irr_star(m, 1)) * 3 + 1 + irr_star(m, 0)];
    if (s[4] * nb <= 0.0) {
      const double dist = s[4] / (s[4] - nb);
      if (dist < *mindist) {
        *mindist = dist;
        op[0] = dist * irr_star(m, 0);
        op[1] = dist * irr_star(m, 1);
      }
    }
  }
}

// One cell of OrthogonalProjection2D_Bicubic (OrthogonalProjection.c:27-128),
// q = j * 2 + i; as irr_cell3.
__device__ double irr_cell2(const double* nb, int q, double op[3]) {
#pragma clang fp contract(off)
  const double TOL = 1e-3, TOL2 = TOL * TOL;
  const int MAXITER = 20;
  const int j = q >> 1, i = q & 1;
  const double x0x = 1 - i, x0y = 1 - j;
  double x = x0x, y = x0y;
  int m;
#pragma unroll 1
  for (m = 0; m < MAXITER; ++m) {
    double e[3];
    irr_bicubic(nb, i, j, x, y, e);
    const double p = e[0], px = e[1], py = e[2];
    const double pp = px * px + py * py;
    const double d1x = -p * px / pp, d1y = -p * py / pp;
    const double xp = (x0x - x) * px + (x0y - y) * py;
    const double d2x = (x0x - x) - xp / pp * px;
    const double d2y = (x0y - y) - xp / pp * py;
    x += d1x + 0.5 * d2x;
    y += d1y + 0.5 * d2y;
    const double norm = ((d1x * d1x + d1y * d1y) + d2x * d2x) + d2y * d2y;
    if (norm < TOL2) break;
  }
  op[0] = x + i - 1;
  op[1] = y + j - 1;
  op[2] = 0.0;
  if (m == MAXITER) return DBL_MAX;
  if (x < -TOL || 1 + TOL < x || y < -TOL || 1 + TOL < y) return DBL_MAX;
  return sqrt((x - x0x) * (x - x0x) + (y - x0y) * (y - x0y));
}

__device__ __forceinline__ void irr_init(dcl_irregular_node& r, const dcl_icoor& q) {
  r.pos = q;
  r.axis = 0;
  r.shift = 0;
  r.sign_center = 0;
  r.sign_face = 0;
  r.pad = 0;
  r.d = 0.0;
  r.X.x = r.X.y = r.X.z = 0.0;
  r.op.x = r.op.y = r.op.z = 0.0;
}

// The 3x3x3 (3x3) neighbourhood of band node q, s[(k * 3 + j) * 3 + i] =
// phi at offset (i - 1, j - 1, k - 1); clamped reads only near the edges.
template <typename T, int DIM>
__device__ __forceinline__ void irr_load(const T* __restrict__ phi, const KWin& w, const dcl_icoor q, double* s) {
  const int lx = q.x - w.px, ly = q.y - w.py, lz = DIM == 3 ? q.z - w.pz : 0;
  const bool inner = lx >= 1 && lx <= w.nx - 2 && ly >= 1 && ly <= w.ny - 2 &&
                     (DIM == 2 || (lz >= 1 && lz <= w.nz - 2));
  if (inner) {
    const T* c = phi + ((size_t)lz * w.ny + ly) * w.nx + lx;
    const int64_t sy = w.nx, sz = (int64_t)w.nx * w.ny;
#pragma unroll
    for (int K = 0; K < (DIM == 3 ? 3 : 1); ++K)
#pragma unroll
      for (int J = 0; J < 3; ++J)
#pragma unroll
        for (int I = 0; I < 3; ++I)
          s[(K * 3 + J) * 3 + I] = (double)c[(DIM == 3 ? (K - 1) * sz : 0) + (J - 1) * sy + (I - 1)];
  } else {
#pragma unroll
    for (int K = 0; K < (DIM == 3 ? 3 : 1); ++K)
#pragma unroll
      for (int J = 0; J < 3; ++J)
#pragma unroll
        for (int I = 0; I < 3; ++I)
          s[(K * 3 + J) * 3 + I] = irr_at(phi, w, q.x + I - 1, q.y + J - 1, DIM == 3 ? q.z + K - 1 : w.pz);
  }
}

#define IRR_S3(s, dx, dy, dz) (s)[((1 + (dz)) * 3 + 1 + (dy)) * 3 + 1 + (dx)]
#define IRR_S2(s, dx, dy) (s)[(1 + (dy)) * 3 + 1 + (dx)]

// a projection record: phi changes sign towards a STAR (axis) neighbour
// (IrregularNodes3D.c, IrregularNodes.c)
template <int DIM>
__device__ __forceinline__ bool irr_has_proj(const double* s) {
  if (DIM == 3) {
    const double c = IRR_S3(s, 0, 0, 0);
    return c * IRR_S3(s, 1, 0, 0) <= 0.0 || c * IRR_S3(s, -1, 0, 0) <= 0.0 || c * IRR_S3(s, 0, 1, 0) <= 0.0 ||
           c * IRR_S3(s, 0, -1, 0) <= 0.0 || c * IRR_S3(s, 0, 0, 1) <= 0.0 || c * IRR_S3(s, 0, 0, -1) <= 0.0;
  }
  const double c = IRR_S2(s, 0, 0);
  return c * IRR_S2(s, 1, 0) <= 0.0 || c * IRR_S2(s, -1, 0) <= 0.0 || c * IRR_S2(s, 0, -1) <= 0.0 ||
         c * IRR_S2(s, 0, 1) <= 0.0;
}

// The records of band node q from its neighbourhood s (irr_load).  EMIT
// false: only count them.  Records go to out[base + 0 ..] while below cap;
// the projection record (always the node's first) without X and op, which
// irr_project_k adds.  Returns the number of records.
template <bool EMIT>
__device__ __forceinline__ int irr_node3(const double* s, const dcl_icoor q, dcl_irregular_node* out, int64_t base,
                                         int64_t cap) {
#pragma clang fp contract(off)
  const double c = IRR_S3(s, 0, 0, 0);
  int n = 0;
  if (irr_has_proj<3>(s)) {
    if (EMIT && base < cap) {
      dcl_irregular_node r;
      irr_init(r, q);
      r.axis = -1;
      r.shift = -1;
      r.sign_center = c > 0.0 ? 1 : -1;
      out[base] = r;
    }
    ++n;
  }
  dcl_irregular_node r;
#pragma unroll
  for (int f = 1; f <= 3; ++f) {
    const double nb = IRR_S3(s, irr_stag(f, 0), irr_stag(f, 1), irr_stag(f, 2));
    if (c * nb <= 0.0) {
      if (EMIT && base + n < cap) {
        irr_init(r, q);
        r.d = c / (c - nb);
        r.sign_center = c > 0.0 ? 1 : -1;
        r.sign_face = r.d < 0.5 ? -r.sign_center : r.sign_center;
        r.shift = 0;
        r.axis = f - 1;
        r.X.x = q.x + r.d * irr_stag(f, 0);
        r.X.y = q.y + r.d * irr_stag(f, 1);
        r.X.z = q.z + r.d * irr_stag(f, 2);
        out[base + n] = r;
      }
      ++n;
    }
  }
#pragma unroll
  for (int face = 1; face <= 3; ++face) {
    const int gx = irr_stag(face, 0), gy = irr_stag(face, 1), gz = irr_stag(face, 2);
    const double hi = (IRR_S3(s, gx, gy, gz) + c) / 2.0;
#pragma unroll
    for (int ax = 1; ax <= 3; ++ax) {
      if (ax == face) continue;
      const int lx = irr_stag(ax, 0), ly = irr_stag(ax, 1), lz = irr_stag(ax, 2);
      const double low = (IRR_S3(s, gx + lx, gy + ly, gz + lz) + IRR_S3(s, lx, ly, lz)) / 2.0;
      if (hi * low <= 0.0) {
        if (EMIT && base + n < cap) {
          irr_init(r, q);
          r.d = hi / (hi - low);
          if (r.d != r.d) r.d = 0.0;
          r.sign_center = hi > 0.0 ? 1 : -1;
          r.sign_face = r.d < 0.5 ? -r.sign_center : r.sign_center;
          r.shift = face;
          r.axis = ax - 1;
          r.X.x = (q.x + gx / 2.0) + r.d * lx;
          r.X.y = (q.y + gy / 2.0) + r.d * ly;
          r.X.z = (q.z + gz / 2.0) + r.d * lz;
          out[base + n] = r;
        }
        ++n;
      }
    }
  }
  return n;
}

template <bool EMIT>
__device__ __forceinline__ int irr_node2(const double* s, const dcl_icoor q, dcl_irregular_node* out, int64_t base,
                                         int64_t cap) {
#pragma clang fp contract(off)
  const double c = IRR_S2(s, 0, 0);
  int n = 0;
  if (irr_has_proj<2>(s)) {
    if (EMIT && base < cap) {
      dcl_irregular_node r;
      irr_init(r, q);
      r.axis = -1;
      r.shift = -1;
      r.sign_center = c > 0.0 ? 1 : -1;
      out[base] = r;
    }
    ++n;
  }
  dcl_irregular_node r;
#pragma unroll
  for (int f = 1; f <= 2; ++f) {
    const double nb = IRR_S2(s, irr_stag(f, 0), irr_stag(f, 1));
    if (c * nb <= 0.0) {
      if (EMIT && base + n < cap) {
        irr_init(r, q);
        r.d = c / (c - nb);
        r.sign_center = c > 0.0 ? 1 : -1;
        r.sign_face = r.d < 0.5 ? -r.sign_center : r.sign_center;
        r.shift = 0;
        r.axis = f - 1;
        r.X.x = q.x + r.d * irr_stag(f, 0);
        r.X.y = q.y + r.d * irr_stag(f, 1);
        out[base + n] = r;
      }
      ++n;
    }
  }
#pragma unroll
  for (int face = 1; face <= 2; ++face) {  // U face, then V face (IrregularNodes.c:107-137)
    const double hi = face == 1 ? (IRR_S2(s, -1, 0) + c) / 2.0 : (IRR_S2(s, 0, -1) + c) / 2.0;
    const double low = face == 1 ? (IRR_S2(s, -1, -1) + IRR_S2(s, 0, -1)) / 2.0
                                 : (IRR_S2(s, -1, -1) + IRR_S2(s, -1, 0)) / 2.0;
    if (hi * low <= 0.0) {
      if (EMIT && base + n < cap) {
        irr_init(r, q);
        r.d = hi / (hi - low);
        r.sign_center = hi > 0.0 ? 1 : -1;
        r.sign_face = r.d < 0.5 ? -r.sign_center : r.sign_center;
        r.shift = face;
        r.axis = face == 1 ? 1 : 0;
        if (face == 1) {
          r.X.x = q.x - 1 / 2.0;
          r.X.y = (q.y + 0 / 2.0) - r.d;
        } else {
          r.X.x = (q.x + 0 / 2.0) - r.d;
          r.X.y = q.y - 1 / 2.0;
        }
        out[base + n] = r;
      }
      ++n;
    }
  }
  return n;
}

constexpr int IRB = 256;
static_assert(sizeof(dcl_irregular_node) == 88 && offsetof(dcl_irregular_node, X) == 40 &&
                  offsetof(dcl_irregular_node, op) == 64,
              "record layout: X and op are the last six doubles");
constexpr int IRR_STAGE = 256;  // records of one block staged in LDS (22.5 KB)

template <typename T, int DIM>
__global__ __launch_bounds__(IRB) void irr_count_k(const T* __restrict__ phi, KWin w,
                                                   const dcl_icoor* __restrict__ band, int64_t nb,
                                                   unsigned* __restrict__ counts, unsigned* __restrict__ pcounts) {
  __shared__ unsigned sh[IRB / 64];
  const int64_t b = (int64_t)blockIdx.x * IRB + threadIdx.x;
  unsigned v = 0;  // records in the low 16 bits, projections above
  if (b < nb) {
    double s[DIM == 3 ? 27 : 9];
    irr_load<T, DIM>(phi, w, band[b], s);
    const int n = DIM == 3 ? irr_node3<false>(s, band[b], nullptr, 0, 0) : irr_node2<false>(s, band[b], nullptr, 0, 0);
    v = (unsigned)n | ((unsigned)irr_has_proj<DIM>(s) << 16);
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = sh[0] + sh[1] + sh[2] + sh[3];
    counts[blockIdx.x] = t & 0xffffu;
    pcounts[blockIdx.x] = t >> 16;
  }
}

// a projection to compute: the node and the slot of its record
struct IrrTask {
  long long slot;
  int x, y, z, pad;
};

// Every record; the projections without X and op, which go to the task list
// in band order (each task carries its record's slot).
template <typename T, int DIM>
__global__ __launch_bounds__(IRB) void irr_write_k(const T* __restrict__ phi, KWin w,
                                                   const dcl_icoor* __restrict__ band, int64_t nb,
                                                   const unsigned* __restrict__ offsets,
                                                   const unsigned* __restrict__ poffsets,
                                                   const unsigned long long* __restrict__ sums,
                                                   const unsigned long long* __restrict__ psums,
                                                   dcl_irregular_node* __restrict__ out, int64_t cap,
                                                   IrrTask* __restrict__ tasks) {
  __shared__ unsigned sh[IRB];
  __shared__ dcl_irregular_node stage[IRR_STAGE];
  const int64_t b = (int64_t)blockIdx.x * IRB + threadIdx.x;
  dcl_icoor q = {0, 0, 0};
  double s[DIM == 3 ? 27 : 9];
  int n = 0;
  bool proj = false;
  if (b < nb) {
    q = band[b];
    irr_load<T, DIM>(phi, w, q, s);
    n = DIM == 3 ? irr_node3<false>(s, q, nullptr, 0, 0) : irr_node2<false>(s, q, nullptr, 0, 0);
    proj = irr_has_proj<DIM>(s);
  }
  // exclusive prefix of the records (low 16 bits) and the projections (high
  // 16 bits) within the block (Hillis-Steele in LDS)
  const unsigned mine = (unsigned)n | ((unsigned)proj << 16);
  sh[threadIdx.x] = mine;
  __syncthreads();
  for (int o = 1; o < IRB; o <<= 1) {
    const unsigned add = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  const int64_t boff = (int64_t)(offsets[blockIdx.x] + sums[blockIdx.x / SCAN_TILE]);
  const unsigned btotal = sh[IRB - 1] & 0xffffu;
  const unsigned excl = sh[threadIdx.x] - mine;
  const int64_t local = excl & 0xffffu;
  const int64_t base = boff + local;
  if (btotal <= IRR_STAGE) {
    // the block's records are contiguous: build them in LDS, then copy them
    // out with coalesced 8-byte stores
    const int64_t bcap = cap - boff;
    if (b < nb && n > 0 && local < bcap) {
      if (DIM == 3)
        irr_node3<true>(s, q, stage, local, bcap);
      else
        irr_node2<true>(s, q, stage, local, bcap);
    }
    __syncthreads();
    const int64_t nrec = bcap < (int64_t)btotal ? (bcap > 0 ? bcap : 0) : (int64_t)btotal;
    constexpr int WORDS = sizeof(dcl_irregular_node) / 8;
    const unsigned long long* src = reinterpret_cast<const unsigned long long*>(stage);
    unsigned long long* dst = reinterpret_cast<unsigned long long*>(out + boff);
    for (int64_t i = threadIdx.x; i < nrec * WORDS; i += IRB) dst[i] = src[i];
  } else if (b < nb && n > 0 && base < cap) {
    if (DIM == 3)
      irr_node3<true>(s, q, out, base, cap);
    else
      irr_node2<true>(s, q, out, base, cap);
  }
  // the projection task of the node, at the block's offset plus its prefix
  // (also when its record is beyond cap: irr_project_k skips those)
  if (proj) {
    IrrTask t;
    t.slot = base;
    t.x = q.x;
    t.y = q.y;
    t.z = q.z;
    t.pad = 0;
    tasks[poffsets[blockIdx.x] + psums[blockIdx.x / SCAN_TILE] + (excl >> 16)] = t;
  }
}

// LDS writes of one lane visible to the other lanes of its wave
__device__ __forceinline__ void irr_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The projection records of the task list.  Each task is handled by a group
// of CELLS lanes, one per cell around the node, over the node's
// neighbourhood staged in LDS (nbv); a wave takes GROUPS consecutive tasks,
// the grid strides over the list.  The nearest converged root wins, ties to
// the lower cell -- the reference's sequential loop with a strict
// `dist < mindist` -- and lane 0 of the group falls back to the quadratic /
// linear projection as the reference does; the group writes X and op.
template <typename T, int DIM>
__global__ __launch_bounds__(IRB) void irr_project_k(const T* __restrict__ phi, KWin w,
                                                     const IrrTask* __restrict__ tasks,
                                                     const unsigned long long* __restrict__ ntask,
                                                     dcl_irregular_node* __restrict__ out, int64_t cap) {
#pragma clang fp contract(off)
  constexpr int CELLS = DIM == 3 ? 8 : 4, GROUPS = 64 / CELLS, NBS = DIM == 3 ? 128 : 32;
  constexpr int NBV = DIM == 3 ? 125 : 25;
  __shared__ double cache[IRB / 64 * GROUPS * NBS];
  const int lane = threadIdx.x & 63, grp = lane / CELLS, sub = lane % CELLS;
  double* nbv = cache + ((threadIdx.x >> 6) * GROUPS + grp) * NBS;
  const unsigned total = (unsigned)*ntask;  // at most one task per band node, fewer than 2^31
  const unsigned stride = gridDim.x * (IRB / 64) * GROUPS;
  for (unsigned first = (blockIdx.x * (IRB / 64) + (threadIdx.x >> 6)) * GROUPS; first < total; first += stride) {
    const unsigned id = first + grp;
    IrrTask t = {0, 0, 0, 0, 0};
    if (id < total) t = tasks[id];
    const bool act = id < total && t.slot < cap;
    if (act) {
      const int lx = t.x - w.px, ly = t.y - w.py, lz = DIM == 3 ? t.z - w.pz : 0;
      const bool inner = lx >= 2 && lx <= w.nx - 3 && ly >= 2 && ly <= w.ny - 3 &&
                         (DIM == 2 || (lz >= 2 && lz <= w.nz - 3));
      if (inner) {  // the same values without the clamps
        const T* c0 = phi + ((size_t)lz * w.ny + ly) * w.nx + lx;
        const int64_t sy = w.nx, sz = (int64_t)w.nx * w.ny;
        for (int e = sub; e < NBV; e += CELLS) {
          const int a = e % 5, bb = (e / 5) % 5, c = e / 25;
          nbv[e] = (double)c0[(DIM == 3 ? (c - 2) * sz : 0) + (bb - 2) * sy + (a - 2)];
        }
      } else {
        for (int e = sub; e < NBV; e += CELLS) {
          const int a = e % 5, bb = (e / 5) % 5, c = e / 25;
          nbv[e] = irr_at(phi, w, t.x - 2 + a, t.y - 2 + bb, DIM == 3 ? t.z - 2 + c : w.pz);
        }
      }
    }
    irr_wave_sync();
    double op[3] = {0.0, 0.0, 0.0};
    double dist = DBL_MAX;
    if (act) dist = DIM == 3 ? irr_cell3(nbv, sub, op) : irr_cell2(nbv, sub, op);
    if (!(dist < DBL_MAX)) dist = DBL_MAX;  // not converged, outside the cell, or NaN
    int who = sub;
#pragma unroll
    for (int o = 1; o < CELLS; o <<= 1) {
      const double od = __shfl_xor(dist, o);
      const int ow = __shfl_xor(who, o);
      const double o0 = __shfl_xor(op[0], o), o1 = __shfl_xor(op[1], o), o2 = __shfl_xor(op[2], o);
      if (od < dist || (od == dist && ow < who)) {
        dist = od;
        who = ow;
        op[0] = o0;
        op[1] = o1;
        op[2] = o2;
      }
    }
    if (act && sub == 0) {
      const bool none = dist == DBL_MAX;
      if (none) op[0] = op[1] = op[2] = 0.0;
      if (DIM == 3) {
        if (none) {
          double s[27];
#pragma unroll
          for (int K = 0; K < 3; ++K)
#pragma unroll
            for (int J = 0; J < 3; ++J)
#pragma unroll
              for (int I = 0; I < 3; ++I) s[(K * 3 + J) * 3 + I] = nbv[((K + 1) * 5 + J + 1) * 5 + I + 1];
          irr_proj3_quadratic(s, op);  // OrthogonalProjection3D.c:105-107
        }
      } else {
        double s[9];
#pragma unroll
        for (int J = 0; J < 3; ++J)
#pragma unroll
          for (int I = 0; I < 3; ++I) s[J * 3 + I] = nbv[(J + 1) * 5 + I + 1];
        double opl[3] = {0.0, 0.0, 0.0}, distl;
        irr_proj2_linear(s, opl, &distl);
        if (!(dist < distl)) {  // OrthogonalProjection.c:16-20
          op[0] = opl[0];
          op[1] = opl[1];
          op[2] = opl[2];
        }
      }
    }
    // lane 0's op to its group; X = pos + op and op, one double per lane
    const int lead = lane - sub;
    op[0] = __shfl(op[0], lead);
    op[1] = __shfl(op[1], lead);
    op[2] = __shfl(op[2], lead);
    if (act) {
      double* xo = &out[t.slot].X.x;  // X.x, X.y, X.z, op.x, op.y, op.z
#pragma unroll
      for (int k = sub; k < 6; k += CELLS) {
        double v;
        if (k == 0)
          v = t.x + op[0];
        else if (k == 1)
          v = t.y + op[1];
        else if (k == 2)
          v = DIM == 3 ? t.z + op[2] : 0.0;
        else
          v = op[k - 3];
        xo[k] = v;
      }
    }
    irr_wave_sync();
  }
}

template <typename T, int DIM>
static int irr_launch(const T* p, const KWin& w, const dcl_icoor* band, int64_t nb, const ScanArrays& sa,
                      IrrTask* tasks, dcl_irregular_node* out, int64_t cap, hipStream_t s) {
  const int64_t nblk = (nb + IRB - 1) / IRB;
  hipLaunchKernelGGL((irr_count_k<T, DIM>), dim3(nblk), dim3(IRB), 0, s, p, w, band, nb, sa.counts[0],
                     sa.counts[1]);
  DCL_LAUNCHED("irr_count");
  DCL_TRY(scan_counts(sa, 2, nblk, s));  // records, projections
  if (cap <= 0) return 0;
  hipLaunchKernelGGL((irr_write_k<T, DIM>), dim3(nblk), dim3(IRB), 0, s, p, w, band, nb, sa.counts[0],
                     sa.counts[1], sa.sums[0], sa.sums[1], out, cap, tasks);
  DCL_LAUNCHED("irr_write");
  // at most one task per band node; at most 1024 blocks stride over the list
  constexpr int GROUPS = DIM == 3 ? 8 : 16;
  const int64_t want = (nb + (IRB / 64) * GROUPS - 1) / ((IRB / 64) * GROUPS);
  const int64_t pblk = want < 1024 ? want : 1024;
  hipLaunchKernelGGL((irr_project_k<T, DIM>), dim3(pblk), dim3(IRB), 0, s, p, w, tasks, sa.total[1], out, cap);
  DCL_LAUNCHED("irr_project");
  return 0;
}

template <typename T>
static int irregular_t(const void* phi, const dcl_window* win, const dcl_icoor* band, int64_t nb,
                       dcl_irregular_node* out, int64_t cap, int64_t* count_out, hipStream_t s) {
  const int64_t nblk = (nb + IRB - 1) / IRB;
  const int64_t ntiles = scan_tiles(nblk);
  // scratch: 2 totals | 2 x tile sums | 2 x block counts (16-byte padded) | tasks (one per band node at most)
  const size_t counts_bytes = ((sizeof(unsigned) * nblk + 15) / 16) * 16;
  const size_t sums_off = 16, counts_off = sums_off + 2 * sizeof(unsigned long long) * ntiles;
  const size_t tasks_off = ((counts_off + 2 * counts_bytes + 15) / 16) * 16;
  void* sc = nullptr;
  DCL_TRY(scratch(tasks_off + sizeof(IrrTask) * (size_t)nb, &sc));
  char* base = static_cast<char*>(sc);
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(base);
  unsigned long long* sums = reinterpret_cast<unsigned long long*>(base + sums_off);
  unsigned* counts = reinterpret_cast<unsigned*>(base + counts_off);
  const ScanArrays sa = {{counts, reinterpret_cast<unsigned*>(base + counts_off + counts_bytes)},
                         {sums, sums + ntiles},
                         {totals, totals + 1}};
  IrrTask* tasks = reinterpret_cast<IrrTask*>(base + tasks_off);
  const T* p = static_cast<const T*>(phi);
  const KWin w = kwin(*win);
  const int rc = win->n.z >= 2 ? irr_launch<T, 3>(p, w, band, nb, sa, tasks, out, cap, s)
                               : irr_launch<T, 2>(p, w, band, nb, sa, tasks, out, cap, s);
  if (rc) return rc;
  unsigned long long tot = 0;
  DCL_HIP(hipMemcpyAsync(&tot, totals, sizeof(tot), hipMemcpyDeviceToHost, s));
  DCL_HIP(hipStreamSynchronize(s));
  *count_out = (int64_t)tot;
  return 0;
}

}  // namespace dcl

using namespace dcl;

extern "C" int dcl_irregular_nodes(const void* phi, const dcl_window* win, const dcl_icoor* band, int64_t band_len,
                                   int precision, dcl_irregular_node* out, int64_t capacity, int64_t* count_out,
                                   void* stream) {
  DCL_TRY(check_precision(precision));
  if (!phi) return set_error(DCL_ERR_ARG_NULL, "phi is NULL");
  if (!count_out) return set_error(DCL_ERR_ARG_NULL, "count_out is NULL");
  if (band_len < 0) return set_error(DCL_ERR_ARG_SIZ, "band_len < 0");
  if (band_len > 0 && !band) return set_error(DCL_ERR_ARG_NULL, "band is NULL");
  if (capacity < 0) return set_error(DCL_ERR_ARG_SIZ, "capacity < 0");
  if (capacity > 0 && !out) return set_error(DCL_ERR_ARG_NULL, "out is NULL");
  DCL_TRY(check_window(win, "phi", 1));
  *count_out = 0;
  if (band_len == 0) return 0;
  if (band_len > INT32_MAX) return set_error(DCL_ERR_ARG_SIZ, "band too long (at most 2^31 - 1 nodes)");
  hipStream_t s = as_stream(stream);
  if (precision == DCL_F64) return irregular_t<double>(phi, win, band, band_len, out, capacity, count_out, s);
  return irregular_t<float>(phi, win, band, band_len, out, capacity, count_out, s);
}
