/*
 * irregular_oracle.c -- CPU FP64 restatement of the irregular-node list,
 * LevelSetUpdateIrregularNodeList (LevelSetMethod/IrregularNodes.c:8-140,
 * LevelSetMethod/IrregularNodes3D.c:6-117) with the orthogonal projections of
 * LevelSetMethod/OrthogonalProjection3D.c:6-208 and
 * LevelSetMethod/OrthogonalProjection.c:6-184.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Nothing in
 * dcell_amd/ links or calls this file.
 *
 * The tricubic (bicubic) polynomial of TricubicInverse.h /
 * TricubicInterpolation.h is evaluated in its separable Catmull-Rom form (the
 * two are the same polynomial; tests/test_oracle.py checks this evaluator
 * against the reference's compiled fragments).  Every expression is written
 * as one IEEE operation at a time (built with -ffp-contract=off) in the same
 * order as the GPU kernel (dcell_amd/csrc/irregular.hip), so the Newton
 * iterations of both take the same path and the records agree bit for bit.
 *
 * Stencil reads clamp the index into the window; the reference reads the
 * 5x5x5 (5x5) neighbourhood of a band node unchecked and requires band nodes
 * away from the window's edge (IrregularNodes.c:5).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "lsm_oracle.h"

#define MAXR DBL_MAX /* PETSC_MAX_REAL */

/* STAR (Common/Common.h:49-51) and STAGGERED_GRID (Common/Common.h:42-47) */
static const int STAR[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
static const int STAG[4][3] = {{0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* phi at global node (x, y, z), the index clamped into the window */
static inline double at(const double *phi, const orc_win *w, int x, int y, int z) {
  const int lx = clampi(x - w->px, 0, w->nx - 1), ly = clampi(y - w->py, 0, w->ny - 1);
  const int lz = w->nz < 2 ? 0 : clampi(z - w->pz, 0, w->nz - 1);
  return phi[((int64_t)lz * w->ny + ly) * (int64_t)w->nx + lx];
}

/* Catmull-Rom weights of the offsets -1..2 and their derivatives at t */
static void crw(double t, double w[4], double d[4]) {
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

/* p, px, py, pz at (x, y, z) of the tricubic on the 4x4x4 stencil st[k][j][i] */
void orc_irr_tricubic(const double *st, double x, double y, double z, double out[4]) {
  double wx[4], dx[4], wy[4], dy[4], wz[4], dz[4];
  crw(x, wx, dx);
  crw(y, wy, dy);
  crw(z, wz, dz);
  double p = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  for (int k = 0; k < 4; ++k) {
    double sp = 0.0, spx = 0.0, spy = 0.0;
    for (int j = 0; j < 4; ++j) {
      double rx = 0.0, rdx = 0.0;
      for (int i = 0; i < 4; ++i) {
        const double f = st[(k * 4 + j) * 4 + i];
        rx = rx + wx[i] * f;
        rdx = rdx + dx[i] * f;
      }
      sp = sp + wy[j] * rx;
      spx = spx + wy[j] * rdx;
      spy = spy + dy[j] * rx;
    }
    p = p + wz[k] * sp;
    px = px + wz[k] * spx;
    py = py + wz[k] * spy;
    pz = pz + dz[k] * sp;
  }
  out[0] = p;
  out[1] = px;
  out[2] = py;
  out[3] = pz;
}

/* p, px, py at (x, y) of the bicubic on the 4x4 stencil st[j][i] */
void orc_irr_bicubic(const double *st, double x, double y, double out[3]) {
  double wx[4], dx[4], wy[4], dy[4];
  crw(x, wx, dx);
  crw(y, wy, dy);
  double p = 0.0, px = 0.0, py = 0.0;
  for (int j = 0; j < 4; ++j) {
    double rx = 0.0, rdx = 0.0;
    for (int i = 0; i < 4; ++i) {
      const double f = st[j * 4 + i];
      rx = rx + wx[i] * f;
      rdx = rdx + dx[i] * f;
    }
    p = p + wy[j] * rx;
    px = px + wy[j] * rdx;
    py = py + dy[j] * rx;
  }
  out[0] = p;
  out[1] = px;
  out[2] = py;
}

/* OrthogonalProjection3D_Linear (OrthogonalProjection3D.c:186-208) */
static void proj3_linear(const double s[3][3][3], double op[3]) {
  double mindist = 2.0;
  for (int m = 0; m < 6; ++m) {
    const double nb = s[1 + STAR[m][2]][1 + STAR[m][1]][1 + STAR[m][0]];
    if (s[1][1][1] * nb <= 0.0) {
      const double dist = s[1][1][1] / (s[1][1][1] - nb);
      if (dist < mindist) {
        mindist = dist;
        op[0] = dist * STAR[m][0];
        op[1] = dist * STAR[m][1];
        op[2] = dist * STAR[m][2];
      }
    }
  }
}

/* OrthogonalProjection3D_Quadratic (OrthogonalProjection3D.c:130-182) */
static void proj3_quadratic(const double s[3][3][3], double op[3]) {
  const double p = s[1][1][1];
  const double px = (s[1][1][2] - s[1][1][0]) / 2.0;
  const double py = (s[1][2][1] - s[1][0][1]) / 2.0;
  const double pz = (s[2][1][1] - s[0][1][1]) / 2.0;
  const double px2 = px * px, py2 = py * py, pz2 = pz * pz;
  const double pxx = (s[1][1][2] - 2.0 * p) + s[1][1][0];
  const double pyy = (s[1][2][1] - 2.0 * p) + s[1][0][1];
  const double pzz = (s[2][1][1] - 2.0 * p) + s[0][1][1];
  const double pxy = (((s[1][0][0] + s[1][2][2]) - s[1][2][0]) - s[1][0][2]) / 4.0;
  const double pxz = (((s[0][1][0] + s[2][1][2]) - s[2][1][0]) - s[0][1][2]) / 4.0;
  const double pyz = (((s[0][0][1] + s[2][2][1]) - s[2][0][1]) - s[0][2][1]) / 4.0;
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
    proj3_linear(s, op);
  }
}

/* OrthogonalProjection3D (OrthogonalProjection3D.c:6-126): Newton on the
 * tricubic of each of the 8 cells around the node, nearest converged root */
static void proj3(const double *phi, const orc_win *w, int X, int Y, int Z, const double s[3][3][3],
                  double op[3]) {
  const double TOL = 1e-3, TOL2 = TOL * TOL;
  const int MAXITER = 20;
  double mindist = MAXR;
  for (int k = 0; k <= 1; ++k)
    for (int j = 0; j <= 1; ++j)
      for (int i = 0; i <= 1; ++i) {
        /* the 4x4x4 stencil of the cell with local origin at node offset (i-1, j-1, k-1) */
        double st[64];
        for (int c = 0; c < 4; ++c)
          for (int b = 0; b < 4; ++b)
            for (int a = 0; a < 4; ++a)
              st[(c * 4 + b) * 4 + a] = at(phi, w, X - 2 + i + a, Y - 2 + j + b, Z - 2 + k + c);
        const double x0x = 1 - i, x0y = 1 - j, x0z = 1 - k;
        double x = x0x, y = x0y, z = x0z;
        int m;
        for (m = 0; m < MAXITER; ++m) {
          double e[4];
          orc_irr_tricubic(st, x, y, z, e);
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
        double dist;
        if (m == MAXITER) {
          dist = MAXR;
        } else if (x < -TOL || 1 + TOL < x || y < -TOL || 1 + TOL < y || z < -TOL || 1 + TOL < z) {
          dist = MAXR;
        } else {
          dist = sqrt(((x - x0x) * (x - x0x) + (y - x0y) * (y - x0y)) + (z - x0z) * (z - x0z));
        }
        if (dist < mindist) {
          mindist = dist;
          op[0] = x + i - 1;
          op[1] = y + j - 1;
          op[2] = z + k - 1;
        }
      }
  if (fabs(mindist - MAXR) < 1) proj3_quadratic(s, op);
}

/* OrthogonalProjection2D_Linear (OrthogonalProjection.c:174-193) */
static void proj2_linear(const double s[3][3], double op[3], double *mindist) {
  *mindist = 2.0;
  for (int m = 0; m < 4; ++m) {
    const double nb = s[1 + STAR[m][1]][1 + STAR[m][0]];
    if (s[1][1] * nb <= 0.0) {
      const double dist = s[1][1] / (s[1][1] - nb);
      if (dist < *mindist) {
        *mindist = dist;
        op[0] = dist * STAR[m][0];
        op[1] = dist * STAR[m][1];
      }
    }
  }
}

/* OrthogonalProjection2D (OrthogonalProjection.c:6-22) with _Bicubic (:27-128) */
static void proj2(const double *phi, const orc_win *w, int X, int Y, const double s[3][3], double op[3]) {
  const double TOL = 1e-3, TOL2 = TOL * TOL;
  const int MAXITER = 20;
  double mindist = MAXR, opb[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j <= 1; ++j)
    for (int i = 0; i <= 1; ++i) {
      double st[16];
      for (int b = 0; b < 4; ++b)
        for (int a = 0; a < 4; ++a) st[b * 4 + a] = at(phi, w, X - 2 + i + a, Y - 2 + j + b, w->pz);
      const double x0x = 1 - i, x0y = 1 - j;
      double x = x0x, y = x0y;
      int m;
      for (m = 0; m < MAXITER; ++m) {
        double e[3];
        orc_irr_bicubic(st, x, y, e);
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
      double dist;
      if (m == MAXITER) {
        dist = MAXR;
      } else if (x < -TOL || 1 + TOL < x || y < -TOL || 1 + TOL < y) {
        dist = MAXR;
      } else {
        dist = sqrt((x - x0x) * (x - x0x) + (y - x0y) * (y - x0y));
      }
      if (dist < mindist) {
        mindist = dist;
        opb[0] = x + i - 1;
        opb[1] = y + j - 1;
        opb[2] = 0.0;
      }
    }
  double opl[3] = {0.0, 0.0, 0.0}, distl;
  proj2_linear(s, opl, &distl);
  if (mindist < distl) {
    op[0] = opb[0], op[1] = opb[1], op[2] = opb[2];
  } else {
    op[0] = opl[0], op[1] = opl[1], op[2] = opl[2];
  }
}

static void put(orc_irregular_node *out, int64_t cap, int64_t *cnt, const orc_irregular_node *r) {
  if (*cnt < cap) out[*cnt] = *r;
  ++*cnt;
}

static void node_init(orc_irregular_node *r, const orc_icoor *q) {
  memset(r, 0, sizeof(*r));
  r->pos = *q;
}

/* LevelSetUpdateIrregularNodeList_3D (IrregularNodes3D.c:6-117) */
static void nodes3(const double *phi, const orc_win *w, const orc_icoor *q, orc_irregular_node *out, int64_t cap,
                   int64_t *cnt) {
  double s[3][3][3];
  for (int K = -1; K <= 1; ++K)
    for (int J = -1; J <= 1; ++J)
      for (int I = -1; I <= 1; ++I) s[K + 1][J + 1][I + 1] = at(phi, w, q->x + I, q->y + J, q->z + K);
  const double c = s[1][1][1];
  orc_irregular_node r;
  for (int m = 0; m < 6; ++m) {  /* projection record for the FMM boundary condition */
    if (c * s[1 + STAR[m][2]][1 + STAR[m][1]][1 + STAR[m][0]] <= 0.0) {
      node_init(&r, q);
      proj3(phi, w, q->x, q->y, q->z, s, r.op);
      r.axis = -1;
      r.shift = -1;
      r.sign_center = c > 0.0 ? 1 : -1;
      r.X[0] = q->x + r.op[0];
      r.X[1] = q->y + r.op[1];
      r.X[2] = q->z + r.op[2];
      put(out, cap, cnt, &r);
      break;
    }
  }
  for (int f = 1; f <= 3; ++f) {  /* cell-centred crossings towards -x, -y, -z */
    const double nb = s[1 + STAG[f][2]][1 + STAG[f][1]][1 + STAG[f][0]];
    if (c * nb <= 0.0) {
      node_init(&r, q);
      r.d = c / (c - nb);
      r.sign_center = c > 0.0 ? 1 : -1;
      r.sign_face = r.d < 0.5 ? -r.sign_center : r.sign_center;
      r.shift = 0;
      r.axis = f - 1;
      r.X[0] = q->x + r.d * STAG[f][0];
      r.X[1] = q->y + r.d * STAG[f][1];
      r.X[2] = q->z + r.d * STAG[f][2];
      put(out, cap, cnt, &r);
    }
  }
  for (int face = 1; face <= 3; ++face) {  /* crossings on the faces, along the two other axes */
    const int *g = STAG[face];
    const double hi = (s[1 + g[2]][1 + g[1]][1 + g[0]] + c) / 2.0;
    for (int ax = 1; ax <= 3; ++ax) {
      if (ax == face) continue;
      const int *lo = STAG[ax];
      const double low = (s[1 + g[2] + lo[2]][1 + g[1] + lo[1]][1 + g[0] + lo[0]] + s[1 + lo[2]][1 + lo[1]][1 + lo[0]]) / 2.0;
      if (hi * low <= 0.0) {
        node_init(&r, q);
        r.d = hi / (hi - low);
        if (r.d != r.d) r.d = 0.0;
        r.sign_center = hi > 0.0 ? 1 : -1;
        r.sign_face = r.d < 0.5 ? -r.sign_center : r.sign_center;
        r.shift = face;
        r.axis = ax - 1;
        r.X[0] = (q->x + g[0] / 2.0) + r.d * lo[0];
        r.X[1] = (q->y + g[1] / 2.0) + r.d * lo[1];
        r.X[2] = (q->z + g[2] / 2.0) + r.d * lo[2];
        put(out, cap, cnt, &r);
      }
    }
  }
}

/* LevelSetUpdateIrregularNodeList_2D (IrregularNodes.c:30-140) */
static void nodes2(const double *phi, const orc_win *w, const orc_icoor *q, orc_irregular_node *out, int64_t cap,
                   int64_t *cnt) {
  static const int nei[4][2] = {{1, 0}, {-1, 0}, {0, -1}, {0, 1}};
  double s[3][3];
  for (int J = -1; J <= 1; ++J)
    for (int I = -1; I <= 1; ++I) s[J + 1][I + 1] = at(phi, w, q->x + I, q->y + J, w->pz);
  const double c = s[1][1];
  orc_irregular_node r;
  for (int m = 0; m < 4; ++m) {
    if (c * s[1 + nei[m][1]][1 + nei[m][0]] <= 0.0) {
      node_init(&r, q);
      proj2(phi, w, q->x, q->y, s, r.op);
      r.axis = -1;
      r.shift = -1;
      r.sign_center = c > 0.0 ? 1 : -1;
      r.X[0] = q->x + r.op[0];
      r.X[1] = q->y + r.op[1];
      put(out, cap, cnt, &r);
      break;
    }
  }
  for (int f = 1; f <= 2; ++f) {
    const double nb = s[1 + STAG[f][1]][1 + STAG[f][0]];
    if (c * nb <= 0.0) {
      node_init(&r, q);
      r.d = c / (c - nb);
      r.sign_center = c > 0.0 ? 1 : -1;
      r.sign_face = r.d < 0.5 ? -r.sign_center : r.sign_center;
      r.shift = 0;
      r.axis = f - 1;
      r.X[0] = q->x + r.d * STAG[f][0];
      r.X[1] = q->y + r.d * STAG[f][1];
      put(out, cap, cnt, &r);
    }
  }
  /* U face (IrregularNodes.c:107-121), then V face (:123-137) */
  for (int face = 1; face <= 2; ++face) {
    const double hi = face == 1 ? (s[1][0] + c) / 2.0 : (s[0][1] + c) / 2.0;
    const double low = face == 1 ? (s[0][0] + s[0][1]) / 2.0 : (s[0][0] + s[1][0]) / 2.0;
    if (hi * low <= 0.0) {
      node_init(&r, q);
      r.d = hi / (hi - low);
      r.sign_center = hi > 0.0 ? 1 : -1;
      r.sign_face = r.d < 0.5 ? -r.sign_center : r.sign_center;
      r.shift = face;
      r.axis = face == 1 ? 1 : 0;
      if (face == 1) {
        r.X[0] = q->x + STAG[1][0] / 2.0;
        r.X[1] = (q->y + STAG[1][1] / 2.0) - r.d;
      } else {
        r.X[0] = (q->x + STAG[2][0] / 2.0) - r.d;
        r.X[1] = q->y + STAG[2][1] / 2.0;
      }
      put(out, cap, cnt, &r);
    }
  }
}

int64_t orc_irregular_nodes(const double *phi, const orc_win *w, const orc_icoor *band, int64_t band_len,
                            orc_irregular_node *out, int64_t capacity) {
  int64_t cnt = 0;
  for (int64_t b = 0; b < band_len; ++b) {
    if (w->nz < 2)
      nodes2(phi, w, &band[b], out, capacity, &cnt);
    else
      nodes3(phi, w, &band[b], out, capacity, &cnt);
  }
  return cnt;
}
