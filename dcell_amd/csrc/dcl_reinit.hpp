// dcl_reinit.hpp -- device helpers shared by the two reinitialisations
// (reinit.hip: the fast marching in the reference's order; reinit_sweep.hip:
// the parallel sweeps): window indexing and AccumulateCoef.
#pragma once

#include "dcl_common.hpp"

namespace dcl {

__device__ __forceinline__ int64_t ri_index(const KWin& w, int d3, int x, int y, int z) {
  return ((int64_t)(d3 ? z - w.pz : 0) * w.ny + (y - w.py)) * w.nx + (x - w.px);
}

__device__ __forceinline__ bool ri_inside(const KWin& w, int d3, int x, int y, int z) {
  return x >= w.px && x < w.px + w.nx && y >= w.py && y < w.py + w.ny && (!d3 || (z >= w.pz && z < w.pz + w.nz));
}

// AccumulateCoef (FastMarching2D.c:33-94): one axis of max(D+, D-, 0)^2
__device__ __forceinline__ void ri_accumulate(double bw, const double* phi7, double& a, double& b, double& c) {
#pragma clang fp contract(off)
  const double* p = &phi7[3];
  const double D2p = (3 * p[0] - 4 * p[1] + p[2]) / 2.;
  const double D2m = (3 * p[0] - 4 * p[-1] + p[-2]) / 2.;
  const double D1p = p[0] - p[1];
  const double D1m = p[0] - p[-1];
  double Dm, am, bm, cm, Dp, ap, bp, cp, t;
  if (0 <= p[1] && p[1] < bw && 0 <= p[2] && p[2] < bw && D2p > 0) {
    t = (4 * p[1] - p[2]) / 3;
    ap = 9. / 4.;
    bp = -9. / 2. * t;
    cp = 9. / 4. * t * t;
    Dp = D2p;
  } else if (0 <= p[1] && p[1] < bw && D1p > 0) {
    ap = 1.;
    bp = -2. * p[1];
    cp = p[1] * p[1];
    Dp = D1p;
  } else {
    ap = 0;
    bp = 0;
    cp = 0;
    Dp = 0;
  }
  if (0 <= p[-1] && p[-1] < bw && 0 <= p[-2] && p[-2] < bw && D2m > 0) {
    t = (4 * p[-1] - p[-2]) / 3;
    am = 9. / 4.;
    bm = -9. / 2. * t;
    cm = 9. / 4. * t * t;
    Dm = D2m;
  } else if (0 <= p[-1] && p[-1] < bw && D1m > 0) {
    am = 1;
    bm = -2. * p[-1];
    cm = p[-1] * p[-1];
    Dm = D1m;
  } else {
    am = 0;
    bm = 0;
    cm = 0;
    Dm = 0;
  }
  if (Dm > Dp) {
    a += am;
    b += bm;
    c += cm;
  } else {
    a += ap;
    b += bp;
    c += cp;
  }
}

}  // namespace dcl
