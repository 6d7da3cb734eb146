/*
 * surface_oracle.c -- CPU FP64 restatement of the surface quantities at the
 * irregular nodes, IIMUpdateSurfaceQuantities_3D / _2D
 * (IIM/SurfaceQuantities.c:31-118): the unit normal from Bilinear3D
 * (Common/Grid3D.c:4-31) / Bilinear2D (Common/Grid.c:190-212) of the
 * central-difference functionals (Common/Grid3D.c:38-108, Common/Grid.c:
 * 219-254), the tangents of LocalCoor3DTangential (IIM/LocalCoor.c:38-65),
 * the principal curvatures from the Hessian rotated by IIMGlobalToLocal_1st /
 * _2nd (IIM/Rotations.c:45-63), and the curvature functional.  The user's
 * interfacial force callback (iim->F) is not part of it.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Nothing in
 * dcell_amd/ links or calls this file.
 *
 * Every expression keeps the reference's operand order, one IEEE operation at
 * a time (built with -ffp-contract=off), as the GPU kernel
 * (dcell_amd/csrc/surface.hip) does, so the two agree bit for bit.  Reads
 * clamp the index into the window; the reference reads unchecked.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "lsm_oracle.h"

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

static inline double at3(const double *phi, const orc_win *w, int x, int y, int z) {
  const int lx = clampi(x - w->px, 0, w->nx - 1), ly = clampi(y - w->py, 0, w->ny - 1);
  const int lz = clampi(z - w->pz, 0, w->nz - 1);
  return phi[((int64_t)lz * w->ny + ly) * w->nx + lx];
}

static inline double at2(const double *phi, const orc_win *w, int x, int y) {
  const int lx = clampi(x - w->px, 0, w->nx - 1), ly = clampi(y - w->py, 0, w->ny - 1);
  return phi[(int64_t)ly * w->nx + lx];
}

enum { GF_DX, GF_DY, GF_DZ, GF_DXX, GF_DYY, GF_DZZ, GF_DXY, GF_DXZ, GF_DYZ, GF_CURV };

/* the GridFunction3D_* functionals (Common/Grid3D.c:38-108) at node (i, j, k) */
static double gf3(int kind, const double *phi, const orc_win *w, int i, int j, int k) {
#define P(a, b, c) at3(phi, w, (a), (b), (c))
  const double dx = w->dx, dy = w->dy, dz = w->dz;
  switch (kind) {
    case GF_DX: return (P(i + 1, j, k) - P(i - 1, j, k)) / (2.0 * dx);
    case GF_DY: return (P(i, j + 1, k) - P(i, j - 1, k)) / (2.0 * dy);
    case GF_DZ: return (P(i, j, k + 1) - P(i, j, k - 1)) / (2.0 * dz);
    case GF_DXX: return (P(i - 1, j, k) - 2. * P(i, j, k) + P(i + 1, j, k)) / (dx * dx);
    case GF_DYY: return (P(i, j - 1, k) - 2. * P(i, j, k) + P(i, j + 1, k)) / (dy * dy);
    case GF_DZZ: return (P(i, j, k - 1) - 2. * P(i, j, k) + P(i, j, k + 1)) / (dz * dz);
    case GF_DXY:
      return (P(i - 1, j - 1, k) + P(i + 1, j + 1, k) - P(i - 1, j + 1, k) - P(i + 1, j - 1, k)) / (4. * dx * dy);
    case GF_DXZ:
      return (P(i - 1, j, k - 1) + P(i + 1, j, k + 1) - P(i - 1, j, k + 1) - P(i + 1, j, k - 1)) / (4. * dx * dz);
    case GF_DYZ:
      return (P(i, j - 1, k - 1) + P(i, j + 1, k + 1) - P(i, j - 1, k + 1) - P(i, j + 1, k - 1)) / (4. * dy * dz);
    default: {
      const double px = (P(i + 1, j, k) - P(i - 1, j, k)) / (2.0 * dx);
      const double py = (P(i, j + 1, k) - P(i, j - 1, k)) / (2.0 * dy);
      const double pz = (P(i, j, k + 1) - P(i, j, k - 1)) / (2.0 * dz);
      const double px2 = px * px, py2 = py * py, pz2 = pz * pz;
      const double ppp = px2 + py2 + pz2;
      const double pxx = (P(i - 1, j, k) - 2. * P(i, j, k) + P(i + 1, j, k)) / (dx * dx);
      const double pyy = (P(i, j - 1, k) - 2. * P(i, j, k) + P(i, j + 1, k)) / (dy * dy);
      const double pzz = (P(i, j, k - 1) - 2. * P(i, j, k) + P(i, j, k + 1)) / (dz * dz);
      const double pxy =
          (P(i - 1, j - 1, k) + P(i + 1, j + 1, k) - P(i - 1, j + 1, k) - P(i + 1, j - 1, k)) / (4. * dx * dy);
      const double pxz =
          (P(i - 1, j, k - 1) + P(i + 1, j, k + 1) - P(i - 1, j, k + 1) - P(i + 1, j, k - 1)) / (4. * dx * dz);
      const double pyz =
          (P(i, j - 1, k - 1) + P(i, j + 1, k + 1) - P(i, j - 1, k + 1) - P(i, j + 1, k - 1)) / (4. * dy * dz);
      double c = px2 * (pyy + pzz) + py2 * (pxx + pzz) + pz2 * (pxx + pyy);
      c -= 2. * (px * py * pxy + px * pz * pxz + py * pz * pyz);
      c /= sqrt(ppp * ppp * ppp);
      if (c != c) c = 0;
      return c;
    }
  }
#undef P
}

/* Bilinear3D (Common/Grid3D.c:4-31): loop order i, j, k */
static double bilinear3_gf(int kind, const double *phi, const orc_win *w, const double p[3]) {
  const int xs = (int)floor(p[0]), ys = (int)floor(p[1]), zs = (int)floor(p[2]);
  const double sx = p[0] - xs, sy = p[1] - ys, sz = p[2] - zs;
  double sum = 0.0;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        const double xb = i * sx + (1 - i) * (1 - sx);
        const double yb = j * sy + (1 - j) * (1 - sy);
        const double zb = k * sz + (1 - k) * (1 - sz);
        sum += gf3(kind, phi, w, xs + i, ys + j, zs + k) * xb * yb * zb;
      }
  return sum;
}

/* the GridFunction2D_* functionals (Common/Grid.c:219-254) at node (i, j) */
static double gf2(int kind, const double *phi, const orc_win *w, int i, int j) {
#define P(a, b) at2(phi, w, (a), (b))
  const double dx = w->dx, dy = w->dy;
  switch (kind) {
    case GF_DX: return (P(i + 1, j) - P(i - 1, j)) / (2. * dx);
    case GF_DY: return (P(i, j + 1) - P(i, j - 1)) / (2. * dy);
    default: {
      const double px = (P(i + 1, j) - P(i - 1, j)) / (2. * dx);
      const double py = (P(i, j + 1) - P(i, j - 1)) / (2. * dy);
      const double px2 = px * px, py2 = py * py;
      const double pxx = (P(i - 1, j) - 2. * P(i, j) + P(i + 1, j)) / (dx * dx);
      const double pyy = (P(i, j - 1) - 2. * P(i, j) + P(i, j + 1)) / (dy * dy);
      const double pxy = (P(i - 1, j - 1) + P(i + 1, j + 1) - P(i - 1, j + 1) - P(i + 1, j - 1)) / (4. * dx * dy);
      double k = (pxx * py2 - 2. * px * py * pxy + pyy * px2) / sqrt((px2 + py2) * (px2 + py2) * (px2 + py2));
      if (k != k) k = 0;
      return k;
    }
  }
#undef P
}

/* Bilinear2D (Common/Grid.c:190-212): the (2x - 1) form, loop order x, y */
static double bilinear2_gf(int kind, const double *phi, const orc_win *w, double x, double y) {
  const int xi = (int)floor(x), yi = (int)floor(y);
  const double xf = (2. * (x - xi) - 1.), yf = (2. * (y - yi) - 1.);
  double sum = 0.;
  for (int k = 0; k < 2; ++k)
    for (int l = 0; l < 2; ++l) {
      const double xb = 1. + (2. * k - 1.) * xf;
      const double yb = 1. + (2. * l - 1.) * yf;
      sum += gf2(kind, phi, w, xi + k, yi + l) * xb * yb;
    }
  sum /= 4.;
  return sum;
}

/* LocalCoor3DTangential (IIM/LocalCoor.c:38-65) */
static void tangential(const double n[3], double s[3], double r[3]) {
  double h;
  if (n[0] * n[0] + n[1] * n[1] >= n[0] * n[0] + n[2] * n[2]) {
    h = sqrt(n[0] * n[0] + n[1] * n[1]);
    s[0] = n[1] / h;
    s[1] = -n[0] / h;
    s[2] = 0;
    r[0] = n[0] * n[2];
    r[1] = n[1] * n[2];
    r[2] = -n[0] * n[0] - n[1] * n[1];
  } else {
    h = sqrt(n[0] * n[0] + n[2] * n[2]);
    s[0] = n[2] / h;
    s[1] = 0;
    s[2] = -n[0] / h;
    r[0] = -n[0] * n[1];
    r[1] = n[0] * n[0] + n[2] * n[2];
    r[2] = -n[1] * n[2];
  }
  h = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  r[0] /= h;
  r[1] /= h;
  r[2] /= h;
}

static void surface3(const double *phi, const orc_win *w, const double X[3], orc_surface *o) {
  double n[3], s[3], r[3];
  n[0] = bilinear3_gf(GF_DX, phi, w, X);
  n[1] = bilinear3_gf(GF_DY, phi, w, X);
  n[2] = bilinear3_gf(GF_DZ, phi, w, X);
  const double h = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= h;
  n[1] /= h;
  n[2] /= h;
  tangential(n, s, r);
  /* IIMGlobalToLocal_1st of the gradient: e = |grad phi| along n */
  const double jx = n[0] * h, jy = n[1] * h, jz = n[2] * h;
  const double je = jx * n[0] + jy * n[1] + jz * n[2];
  const double xx = bilinear3_gf(GF_DXX, phi, w, X);
  const double xy = bilinear3_gf(GF_DXY, phi, w, X);
  const double xz = bilinear3_gf(GF_DXZ, phi, w, X);
  const double yy = bilinear3_gf(GF_DYY, phi, w, X);
  const double yz = bilinear3_gf(GF_DYZ, phi, w, X);
  const double zz = bilinear3_gf(GF_DZZ, phi, w, X);
  /* IIMGlobalToLocal_2nd: the nn, nt and tt entries */
  const double nn = s[0] * (xx * s[0] + xy * s[1] + xz * s[2]) + s[1] * (xy * s[0] + yy * s[1] + yz * s[2]) +
                    s[2] * (xz * s[0] + yz * s[1] + zz * s[2]);
  const double nt = r[0] * (xx * s[0] + xy * s[1] + xz * s[2]) + r[1] * (xy * s[0] + yy * s[1] + yz * s[2]) +
                    r[2] * (xz * s[0] + yz * s[1] + zz * s[2]);
  const double tt = r[0] * (xx * r[0] + xy * r[1] + xz * r[2]) + r[1] * (xy * r[0] + yy * r[1] + yz * r[2]) +
                    r[2] * (xz * r[0] + yz * r[1] + zz * r[2]);
  memcpy(o->n, n, sizeof n);
  memcpy(o->s, s, sizeof s);
  memcpy(o->r, r, sizeof r);
  o->k_nn = -nn / je;
  o->k_tt = -tt / je;
  o->k_nt = -nt / je;
  o->k = bilinear3_gf(GF_CURV, phi, w, X);
}

static void surface2(const double *phi, const orc_win *w, const double X[3], orc_surface *o) {
  memset(o, 0, sizeof *o);
  o->k = bilinear2_gf(GF_CURV, phi, w, X[0], X[1]);
  o->k_nn = o->k;
  o->k_tt = 0;
  o->k_nt = 0;
  double nx = bilinear2_gf(GF_DX, phi, w, X[0], X[1]);
  double ny = bilinear2_gf(GF_DY, phi, w, X[0], X[1]);
  const double h = sqrt(nx * nx + ny * ny);
  nx = nx / h;
  ny = ny / h;
  o->n[0] = nx;
  o->n[1] = ny;
  o->s[0] = ny;
  o->s[1] = -nx;
}

void orc_surface_quantities(const double *phi, const orc_win *w, const orc_irregular_node *nodes, int64_t count,
                            orc_surface *out) {
  const int d3 = w->nz >= 2;
  for (int64_t i = 0; i < count; ++i) {
    if (d3)
      surface3(phi, w, nodes[i].X, &out[i]);
    else
      surface2(phi, w, nodes[i].X, &out[i]);
  }
}
