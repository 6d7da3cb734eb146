/*
 * advect_sl.c -- drop-in for the level-set advection entry points of
 * LevelSetMethod/AdvectSL.c, AdvectRK2.c and Advect.c:33-111.  Same names,
 * signatures, call order and counters; the band loops run as gfx950 kernels
 * through the C ABI (dcell_lsm.h).  phi0, the velocity grid and the band are
 * uploaded per call (the host may have changed them: reinitialisation and the
 * irregular-node update run on the host between steps), the band values come
 * back compactly and are written into phi at the band nodes.
 */
#include <string.h>

#include "LevelSetMethod.h"
#include "LSM_private.h"
#include "dcell_gpu.h"
#include "dropin.h"

static int g_interp = DCL_LINEAR;
static int g_backtrace = DCL_EULER;

PetscErrorCode LevelSetAdvectSLSetScheme(int interp, int backtrace) {
  if ((interp != DCL_LINEAR && interp != DCL_CUBIC) || (backtrace != DCL_EULER && backtrace != DCL_RK2))
    return DCL_ERR_ARG_WRONG;
  g_interp = interp;
  g_backtrace = backtrace;
  return 0;
}

/* write compact band values into g at the band nodes that lie in its window,
 * split over threads (a node listed twice gets the same value from both of
 * its entries, so the order of the writes does not matter) */
static void scatter_band(Grid g, const iCoor *band, int len, const PetscReal *vals) {
  const int nz = g->is2D ? 1 : g->n.z;
  const int nt = dl_threads();
#pragma omp parallel for num_threads(nt) if (nt > 1 && len > 65536) schedule(static)
  for (int b = 0; b < len; ++b) {
    const int x = band[b].x - g->p.x, y = band[b].y - g->p.y, z = g->is2D ? 0 : band[b].z - g->p.z;
    if (x < 0 || x >= g->n.x || y < 0 || y >= g->n.y || z < 0 || z >= nz) continue;
    g->v1[((long)z * g->n.y + y) * g->n.x + x] = vals[b];
  }
}

/* The band loop of LevelSetAdvectSL_2D / _3D (AdvectSL.c:23-89).  Only the
 * part of phi0 and of the velocity grid that the band's departures can reach
 * goes to the GPU: the band's bounding box grown by a margin that covers
 * dt max|v| / dh (max over the uploaded velocity box, which every
 * interpolated velocity is a convex combination of) plus the widest stencil,
 * so every stencil the kernels read lies inside the boxes and the results are
 * those of the whole windows (where a box meets a window edge, it ends there
 * as the window does). */
static PetscErrorCode advect_band(LevelSet ls, Grid velgrid, PetscReal dt, const char *where) {
  const int len = ArrayLength(ls->band);
  const iCoor *band = (const iCoor *)ArrayGetData(ls->band);
  void *dphi0, *dvel, *dband, *dout, *hout;
  if (len == 0) return 0;
  const int dim = ls->phi0->is2D ? 2 : 3;
  dl_box bb, pb;
  dl_bbox(band, len, sizeof(iCoor), dim, &bb);
  dcl_window vw, pw;
  int M = 0;
  DL_CHECK(where, dl_velocity_box(velgrid, &bb, dt, ls->phi0->d, dim, &dvel, &vw, &M));
  if (!dl_box_around(ls->phi0, &bb, M, 0, &pb)) return dl_fail(where, DCL_ERR_ARG_SIZ);
  DL_CHECK(where, dl_upload_box(DL_PHI0, ls->phi0, &pb, &dphi0, &pw, NULL, 0));
  DL_CHECK(where, dl_slot(DL_BAND, sizeof(iCoor) * (size_t)len, &dband));
  DL_CHECK(where, dl_htod(dband, band, sizeof(iCoor) * (size_t)len, NULL));
  DL_CHECK(where, dl_slot(DL_OUT, sizeof(PetscReal) * (size_t)len, &dout));
  dcl_velocity V;
  V.data = dvel;
  V.win = vw;
  V.dof = velgrid->dof;
  V.udof = 0;
  DL_CHECK(where, dcl_sl_advect(dphi0, NULL, &pw, &V, (const dcl_icoor *)dband, len, dout, dt, DCL_F64,
                                g_interp, g_backtrace, NULL));
  DL_CHECK(where, dl_host(sizeof(PetscReal) * (size_t)len, &hout));
  DL_CHECK(where, dl_dtoh(hout, dout, sizeof(PetscReal) * (size_t)len, NULL));
  scatter_band(ls->phi, band, len, (const PetscReal *)hout);
  return 0;
}

PetscErrorCode LevelSetAdvectSL(LevelSet ls, Grid velgrid, PetscReal dt) {
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = PetscLogEventBegin(EVENT_LevelSetAdvectSL, 0, 0, 0, 0); CHKERRQ(ierr);
  dl_scope_begin();
  if (ls->phi->is2D) {
    ierr = LevelSetAdvectSL_2D(ls, velgrid, dt);
  } else {
    ierr = LevelSetAdvectSL_3D(ls, velgrid, dt);
  }
  dl_scope_end();
  CHKERRQ(ierr);
  ierr = PetscLogEventEnd(EVENT_LevelSetAdvectSL, 0, 0, 0, 0); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}

PetscErrorCode LevelSetAdvectSL_2D(LevelSet ls, Grid velgrid, PetscReal dt) {
  return advect_band(ls, velgrid, dt, "LevelSetAdvectSL_2D");
}

PetscErrorCode LevelSetAdvectSL_3D(LevelSet ls, Grid velgrid, PetscReal dt) {
  return advect_band(ls, velgrid, dt, "LevelSetAdvectSL_3D");
}

/* AdvectRK2.c:5-14: phi0 <- phi, SL over dt (the caller passes dt/2) */
PetscErrorCode LevelSetAdvectSLRK2HalfStep(LevelSet ls, Grid velgrid, PetscReal dt) {
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = GridCopy(ls->phi, ls->phi0); CHKERRQ(ierr);
  ierr = LevelSetAdvectSL(ls, velgrid, dt); CHKERRQ(ierr);
  ierr = LevelSetUpdateIrregularNodeList(ls); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}

/* AdvectRK2.c:18-26: SL from the same phi0 with the mid-step velocity */
PetscErrorCode LevelSetAdvectSLRK2FullStep(LevelSet ls, Grid velgrid, PetscReal dt) {
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = LevelSetAdvectSL(ls, velgrid, dt); CHKERRQ(ierr);
  ierr = LevelSetUpdateIrregularNodeList(ls); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}

/* Advect.c:33-50; the velocity grid is uploaded once for the SL step and
 * the CFL count */
static PetscErrorCode advect_and_reinit(LevelSet ls, Grid velgrid, PetscReal dt) {
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = GridCopy(ls->phi, ls->phi0); CHKERRQ(ierr);
  ierr = LevelSetAdvectSL(ls, velgrid, dt); CHKERRQ(ierr);
  ierr = LevelSetCFLIncrement(ls, velgrid, dt); CHKERRQ(ierr);
  ierr = LevelSetUpdateIrregularNodeList(ls); CHKERRQ(ierr);
  if (ls->AdvectCount > ls->AdvectThres || ls->CFLcount > ls->CFLthres) {
    ls->CFLcount = 0;
    ls->AdvectCount = 0;
    ierr = LevelSetReinitialize(ls); CHKERRQ(ierr);
  }
  PetscFunctionReturn(0);
}

PetscErrorCode LevelSetAdvectAndReinit(LevelSet ls, Grid velgrid, PetscReal dt) {
  dl_scope_begin();
  PetscErrorCode ierr = advect_and_reinit(ls, velgrid, dt);
  dl_scope_end();
  return ierr;
}

/* Advect.c:54-66 */
PetscErrorCode LevelSetCFLIncrement(LevelSet ls, Grid velgrid, PetscReal dt) {
  Coor dh = ls->phi->d;
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = LevelSet_MaxVelocity(ls, velgrid); CHKERRQ(ierr);
  ls->CFLcount += dt * ls->maxVel / dh.x;
  ls->AdvectCount++;
  PetscFunctionReturn(0);
}

/* Advect.c:70-111: max |V| over the irregular nodes, on the GPU; the
 * velocity grid over the nodes' box (their node velocities read the faces at
 * the node and one above) */
PetscErrorCode LevelSet_MaxVelocity(LevelSet ls, Grid velgrid) {
  const int len = ArrayLength(ls->irregularNodes);
  const IrregularNode *nodes = (const IrregularNode *)ArrayGetData(ls->irregularNodes);
  void *dvel, *dnodes, *hpos;
  double maxVel = 0;
  ls->maxVel = 0;
  if (len == 0) return 0;
  const int dim = ls->phi->is2D ? 2 : 3;
  DL_CHECK("LevelSet_MaxVelocity", dl_host(sizeof(iCoor) * (size_t)len, &hpos));
  iCoor *pos = (iCoor *)hpos;
  for (int i = 0; i < len; ++i) pos[i] = nodes[i].pos;
  DL_CHECK("LevelSet_MaxVelocity", dl_slot(DL_NODES, sizeof(iCoor) * (size_t)len, &dnodes));
  DL_CHECK("LevelSet_MaxVelocity", dl_htod(dnodes, pos, sizeof(iCoor) * (size_t)len, NULL));
  dl_box bb;
  dl_bbox(pos, len, sizeof(iCoor), dim, &bb);
  dcl_window vw;
  int M = 0;
  DL_CHECK("LevelSet_MaxVelocity", dl_velocity_box(velgrid, &bb, 0.0, ls->phi->d, dim, &dvel, &vw, &M));
  dcl_velocity V;
  V.data = dvel;
  V.win = vw;
  V.dof = velgrid->dof;
  V.udof = 0;
  DL_CHECK("LevelSet_MaxVelocity", dcl_max_velocity(&V, dnodes, len, sizeof(iCoor), dim, DCL_F64, &maxVel, NULL));
  ls->maxVel = maxVel;
  return 0;
}

PetscErrorCode LevelSetCurvature(LevelSet ls, Grid kappa) {
  Grid phi = ls->phi;
  void *dphi, *dk;
  if (kappa->dof != 1 || kappa->n.x != phi->n.x || kappa->n.y != phi->n.y ||
      (!phi->is2D && kappa->n.z != phi->n.z) || kappa->is2D != phi->is2D)
    return dl_fail("LevelSetCurvature", DCL_ERR_ARG_SIZ);
  DL_CHECK("LevelSetCurvature", dl_upload_grid(DL_PHI, phi, &dphi));
  DL_CHECK("LevelSetCurvature", dl_slot(DL_KAPPA, sizeof(PetscReal) * (size_t)phi->SIZE, &dk));
  const dcl_window w = dl_window(phi);
  DL_CHECK("LevelSetCurvature", dcl_curvature(dphi, dk, &w, NULL, 0, NULL, DCL_F64, DCL_CURV_STD, NULL));
  DL_CHECK("LevelSetCurvature", dl_dtoh(kappa->v1, dk, sizeof(PetscReal) * (size_t)phi->SIZE, NULL));
  return 0;
}

PetscErrorCode LevelSetGPUFinalize(void) { return dl_release(); }
