/*
 * ref_tricubic.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Compiles the reference's own tricubic fragments where they lie
 * (LevelSetMethod/TricubicInverse.h: the 64 coefficients a[] from a 5x5x5
 * stencil; LevelSetMethod/TricubicInterpolation.h: p, px, py, pz of that
 * polynomial) inside a function that provides the locals they expect, exactly
 * as OrthogonalProjection3D.c:22-48 does (loop indices i, j, k in {0, 1}; local
 * coordinates x, y, z in [0, 1] measured from node (1+i, 1+j, 1+k)).  Neither
 * fragment includes anything, so no PETSc or GA header is involved.
 *
 * The include path points at the reference tree (oracle/Makefile); the output
 * goes to oracle/_ref/, which is not committed.
 */
#include <string.h>

/* st: 4x4x4 stencil [k][j][i]; x, y, z: local coordinates; out = {p, px, py, pz} */
void ref_tricubic_eval(const double *st, double x, double y, double z, double out[4]) {
  double phi[5][5][5];
  double a[64];
  int i = 0, j = 0, k = 0;
  memset(phi, 0, sizeof(phi));
  for (int kk = 0; kk < 4; ++kk)
    for (int jj = 0; jj < 4; ++jj)
      for (int ii = 0; ii < 4; ++ii) phi[kk][jj][ii] = st[(kk * 4 + jj) * 4 + ii];
#include "TricubicInverse.h"
  double x2 = x * x, x3 = x2 * x;
  double y2 = y * y, y3 = y2 * y;
  double z2 = z * z, z3 = z2 * z;
  double p, px, py, pz;
#include "TricubicInterpolation.h"
  out[0] = p;
  out[1] = px;
  out[2] = py;
  out[3] = pz;
  (void)i; (void)j; (void)k;
}
