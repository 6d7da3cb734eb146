/*
 * particle_ls.c -- drop-in for the particle-level-set entry points of
 * LevelSetMethod/ParticleLS3D.c, ParticleLS2D.c and ParticleLS.c:158-185,
 * 299-388.  Same names, signatures and step order; particle advection, the
 * escaped-particle correction and the radius update run as gfx950 kernels
 * through the C ABI (dcell_lsm.h).  struct _Particle {Coor X; PetscReal
 * radius} is dcl_particle, so the particle array is uploaded as it is.
 */
#include <math.h>

#include "LevelSetMethod.h"
#include "LSM_private.h"
#include "dcell_gpu.h"
#include "dropin.h"

static int g_ref_az = 1;
static dcl_pls_ws g_ws;
static long g_ws_nodes, g_ws_events;

PetscErrorCode LevelSetPLSSetReferenceCornerZ(PetscBool reference) {
  g_ref_az = reference == PETSC_TRUE;
  return 0;
}

static int ensure_ws(long nodes, long events) {
  if (g_ws && nodes <= g_ws_nodes && events <= g_ws_events) return 0;
  if (g_ws) {
    int rc = dcl_pls_ws_destroy(g_ws);
    g_ws = NULL;
    if (rc) return rc;
  }
  long n = nodes > g_ws_nodes ? nodes : g_ws_nodes;
  long e = events > g_ws_events ? events : g_ws_events;
  if (e < 1024) e = 1024;
  int rc = dcl_pls_ws_create(n, e, DCL_F64, NULL, &g_ws);
  if (rc) return rc;
  g_ws_nodes = n;
  g_ws_events = e;
  return 0;
}

static PetscErrorCode upload_particles(ParticleLS pls, void **dev, int *len) {
  *len = ArrayLength(pls->particles);
  const size_t bytes = sizeof(struct _Particle) * (size_t)*len;
  int rc = dl_slot(DL_PART, bytes ? bytes : 1, dev);
  if (rc) return rc;
  return dl_htod(*dev, ArrayGetData(pls->particles), bytes, NULL);
}

/* the box of the nodes around the particles' positions */
static void particle_bbox(ParticleLS pls, int len, int dim, dl_box *b) {
  const struct _Particle *p = (const struct _Particle *)ArrayGetData(pls->particles);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < len; ++i) {
    const double x[3] = {p[i].X.x, p[i].X.y, p[i].X.z};
    for (int a = 0; a < dim; ++a) {
      if (x[a] < lo[a]) lo[a] = x[a];
      if (x[a] > hi[a]) hi[a] = x[a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    if (a >= dim || !(lo[a] <= hi[a])) {
      b->lo[a] = b->hi[a] = 0;
      continue;
    }
    /* clamped into int range; the box is clipped to the window afterwards */
    const double l = floor(lo[a]), h = floor(hi[a]) + 1;
    b->lo[a] = l < -1e9 ? -1000000000 : (int)l;
    b->hi[a] = h > 1e9 ? 1000000000 : (int)h;
  }
}

static PetscErrorCode advect_particles(ParticleLS pls, Coor dh, Grid velgrid, PetscReal dt, int dim,
                                       const char *where) {
  void *dpart, *dvel;
  int len;
  PetscErrorCode ierr;
  ierr = PetscLogEventBegin(EVENT_ParticleLS_AdvectParticles, 0, 0, 0, 0); CHKERRQ(ierr);
  DL_CHECK(where, upload_particles(pls, &dpart, &len));
  if (len > 0) {
    /* the velocity over the particles' box grown by their reach in one step
     * (each RK2 stage moves a particle by at most dt max|v| / dh) */
    dcl_window vw;
    if (dim == 2 && !velgrid->is2D) {
      /* the 2D particle step over a 3D velocity grid (the RK2 PLS steps, as
       * in the reference) reads its first plane: the whole grid goes up */
      dl_velocity_invalidate();
      DL_CHECK(where, dl_upload_grid(DL_VEL, velgrid, &dvel));
      vw = dl_window(velgrid);
    } else {
      dl_box bb;
      particle_bbox(pls, len, dim, &bb);
      int M = 0;
      DL_CHECK(where, dl_velocity_box(velgrid, &bb, dt, dh, dim, &dvel, &vw, &M));
    }
    dcl_velocity V;
    V.data = dvel;
    V.win = vw;
    V.dof = velgrid->dof;
    V.udof = 0;
    dcl_coor h = {dh.x, dh.y, dh.z};
    DL_CHECK(where, dcl_pls_advect(dpart, len, &V, h, dt, dim, DCL_F64, NULL));
    DL_CHECK(where, dl_dtoh(ArrayGetData(pls->particles), dpart,
                                    sizeof(struct _Particle) * (size_t)len, NULL));
  }
  ierr = PetscLogEventEnd(EVENT_ParticleLS_AdvectParticles, 0, 0, 0, 0); CHKERRQ(ierr);
  return 0;
}

/* ParticleLS3D.c:6-36 */
PetscErrorCode ParticleLS_AdvectParticles3D(ParticleLS pls, Coor dh, Grid velgrid, PetscReal dt) {
  return advect_particles(pls, dh, velgrid, dt, 3, "ParticleLS_AdvectParticles3D");
}

/* ParticleLS2D.c:6-36 */
PetscErrorCode ParticleLS_AdvectParticles2D(ParticleLS pls, Coor dh, Grid velgrid, PetscReal dt) {
  return advect_particles(pls, dh, velgrid, dt, 2, "ParticleLS_AdvectParticles2D");
}

static PetscErrorCode correct(ParticleLS pls, LevelSet ls, int dim, const char *where) {
  void *dpart, *dphi, *dband, *dout, *hout;
  int len;
  const int blen = ArrayLength(ls->band);
  const iCoor *band = (const iCoor *)ArrayGetData(ls->band);
  PetscErrorCode ierr;
  ierr = PetscLogEventBegin(EVENT_ParticleLS_ErrorCorrection, 0, 0, 0, 0); CHKERRQ(ierr);
  DL_CHECK(where, upload_particles(pls, &dpart, &len));
  if (blen > 0) {
    /* phi over the box of the band and of the particles' corner cubes (with
     * the reference's A.z = floor(X.y), the z range of the corners is the y
     * range of the particles), grown by one node */
    const int flags = (dim == 3 && g_ref_az) ? DCL_PLS_REF_AZ : 0;
    dl_box bb, pbb, box;
    dl_bbox(band, blen, sizeof(iCoor), dim, &bb);
    if (len > 0) {
      particle_bbox(pls, len, dim, &pbb);
      if (flags) {
        if (pbb.lo[1] < pbb.lo[2]) pbb.lo[2] = pbb.lo[1];
        if (pbb.hi[1] > pbb.hi[2]) pbb.hi[2] = pbb.hi[1];
      }
      for (int a = 0; a < 3; ++a) {
        if (pbb.lo[a] < bb.lo[a]) bb.lo[a] = pbb.lo[a];
        if (pbb.hi[a] > bb.hi[a]) bb.hi[a] = pbb.hi[a];
      }
    }
    if (!dl_box_around(ls->phi, &bb, 1, 0, &box)) return dl_fail(where, DCL_ERR_ARG_SIZ);
    dcl_window w;
    DL_CHECK(where, dl_upload_box(DL_PHI, ls->phi, &box, &dphi, &w, NULL, 0));
    DL_CHECK(where, ensure_ws((long)w.n.x * w.n.y * w.n.z, 8L * (len > 0 ? len : 1)));
    DL_CHECK(where, dl_slot(DL_BAND, sizeof(iCoor) * (size_t)blen, &dband));
    DL_CHECK(where, dl_htod(dband, band, sizeof(iCoor) * (size_t)blen, NULL));
    DL_CHECK(where, dl_slot(DL_OUT, sizeof(PetscReal) * (size_t)blen, &dout));
    DL_CHECK(where, dcl_pls_correct(dpart, len, dphi, &w, (const dcl_icoor *)dband, blen, dout, flags, dim,
                                    DCL_F64, g_ws, NULL));
    DL_CHECK(where, dl_host(sizeof(PetscReal) * (size_t)blen, &hout));
    DL_CHECK(where, dl_dtoh(hout, dout, sizeof(PetscReal) * (size_t)blen, NULL));
    Grid g = ls->phi;
    const int nz = g->is2D ? 1 : g->n.z;
    const PetscReal *v = (const PetscReal *)hout;
    for (int b = 0; b < blen; ++b) {
      const int x = band[b].x - g->p.x, y = band[b].y - g->p.y, z = g->is2D ? 0 : band[b].z - g->p.z;
      if (x < 0 || x >= g->n.x || y < 0 || y >= g->n.y || z < 0 || z >= nz) continue;
      g->v1[((long)z * g->n.y + y) * g->n.x + x] = v[b];
    }
  }
  ierr = PetscLogEventEnd(EVENT_ParticleLS_ErrorCorrection, 0, 0, 0, 0); CHKERRQ(ierr);
  return 0;
}

/* ParticleLS3D.c:40-116 */
PetscErrorCode ParticleLS_ErrorCorrection3D(ParticleLS pls, LevelSet ls) {
  return correct(pls, ls, 3, "ParticleLS_ErrorCorrection3D");
}

/* ParticleLS2D.c:40-115 */
PetscErrorCode ParticleLS_ErrorCorrection2D(ParticleLS pls, LevelSet ls) {
  return correct(pls, ls, 2, "ParticleLS_ErrorCorrection2D");
}

/* ParticleLS.c:158-185 */
PetscErrorCode ParticleLS_AdjustRadii(ParticleLS pls, LevelSet ls) {
  void *dpart, *dphi;
  int len;
  PetscErrorCode ierr;
  ierr = PetscLogEventBegin(EVENT_ParticleLS_AdjustRadii, 0, 0, 0, 0); CHKERRQ(ierr);
  DL_CHECK("ParticleLS_AdjustRadii", upload_particles(pls, &dpart, &len));
  if (len > 0) {
    /* phi over the particles' box grown by one node (trilinear phi at each) */
    dl_box bb, box;
    particle_bbox(pls, len, ls->phi->is2D ? 2 : 3, &bb);
    if (!dl_box_around(ls->phi, &bb, 1, 0, &box)) return dl_fail("ParticleLS_AdjustRadii", DCL_ERR_ARG_SIZ);
    dcl_window w;
    DL_CHECK("ParticleLS_AdjustRadii", dl_upload_box(DL_PHI, ls->phi, &box, &dphi, &w, NULL, 0));
    DL_CHECK("ParticleLS_AdjustRadii", dcl_pls_adjust_radii(dpart, len, dphi, &w, pls->R_MIN, pls->R_MAX,
                                                            ls->phi->is2D ? 2 : 3, DCL_F64, NULL));
    DL_CHECK("ParticleLS_AdjustRadii", dl_dtoh(ArrayGetData(pls->particles), dpart,
                                                       sizeof(struct _Particle) * (size_t)len, NULL));
  }
  ierr = PetscLogEventEnd(EVENT_ParticleLS_AdjustRadii, 0, 0, 0, 0); CHKERRQ(ierr);
  return 0;
}

/* ParticleLS.c:299-329; the velocity grid is uploaded once for the
 * particles, the SL step and the CFL count */
static PetscErrorCode advect_pls(LevelSet ls, Grid velgrid, PetscReal dt) {
  ParticleLS pls = ls->pls;
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = pls->AdvectParticles(pls, ls->phi->d, velgrid, dt); CHKERRQ(ierr);
  ierr = GridCopy(ls->phi, ls->phi0); CHKERRQ(ierr);
  ierr = LevelSetAdvectSL(ls, velgrid, dt); CHKERRQ(ierr);
  ierr = pls->ErrorCorrection(pls, ls); CHKERRQ(ierr);
  ierr = LevelSetUpdateIrregularNodeList(ls); CHKERRQ(ierr);
  ierr = LevelSetCFLIncrement(ls, velgrid, dt); CHKERRQ(ierr);
  if (ls->AdvectCount >= ls->AdvectThres || ls->CFLcount >= ls->CFLthres) {
    ierr = LevelSetReinitialize(ls); CHKERRQ(ierr);
    ierr = ParticleLS_ReseedParticles(pls, ls); CHKERRQ(ierr);
    ls->CFLcount = 0;
    ls->AdvectCount = 0;
  }
  ierr = ParticleLS_AdjustRadii(pls, ls); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}

PetscErrorCode LevelSetAdvectPLS(LevelSet ls, Grid velgrid, PetscReal dt) {
  dl_scope_begin();
  PetscErrorCode ierr = advect_pls(ls, velgrid, dt);
  dl_scope_end();
  return ierr;
}

/* ParticleLS.c:333-350; the particle step is ParticleLS_AdvectParticles2D,
 * as in the reference */
static PetscErrorCode pls_rk2_half(LevelSet ls, Grid velgrid, PetscReal dt) {
  ParticleLS pls = ls->pls;
  PetscErrorCode ierr;
  PetscFunctionBegin;
  if (pls->particles0 == NULL) {
    ierr = ArrayDuplicate(pls->particles, &pls->particles0); CHKERRQ(ierr);
  }
  ierr = ArrayCopy(pls->particles, pls->particles0); CHKERRQ(ierr);
  ierr = ParticleLS_AdvectParticles2D(pls, ls->phi->d, velgrid, dt); CHKERRQ(ierr);
  ierr = GridCopy(ls->phi, ls->phi0); CHKERRQ(ierr);
  ierr = LevelSetAdvectSL(ls, velgrid, dt); CHKERRQ(ierr);
  ierr = pls->ErrorCorrection(pls, ls); CHKERRQ(ierr);
  ierr = LevelSetUpdateIrregularNodeList(ls); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}

PetscErrorCode LevelSetAdvectPLSRK2HalfStep(LevelSet ls, Grid velgrid, PetscReal dt) {
  dl_scope_begin();
  PetscErrorCode ierr = pls_rk2_half(ls, velgrid, dt);
  dl_scope_end();
  return ierr;
}

/* ParticleLS.c:353-375 */
static PetscErrorCode pls_rk2_full(LevelSet ls, Grid velgrid, PetscReal dt) {
  ParticleLS pls = ls->pls;
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = ArrayCopy(pls->particles0, pls->particles); CHKERRQ(ierr);
  ierr = ParticleLS_AdvectParticles2D(pls, ls->phi->d, velgrid, dt); CHKERRQ(ierr);
  ierr = LevelSetAdvectSL(ls, velgrid, dt); CHKERRQ(ierr);
  ierr = pls->ErrorCorrection(pls, ls); CHKERRQ(ierr);
  ierr = LevelSetUpdateIrregularNodeList(ls); CHKERRQ(ierr);
  ierr = LevelSetCFLIncrement(ls, velgrid, dt); CHKERRQ(ierr);
  if (ls->AdvectCount >= ls->AdvectThres || ls->CFLcount >= ls->CFLthres) {
    ierr = LevelSetReinitialize(ls); CHKERRQ(ierr);
    ierr = pls->ErrorCorrection(pls, ls); CHKERRQ(ierr);
    ierr = ParticleLS_ReseedParticles(pls, ls); CHKERRQ(ierr);
    ls->CFLcount = 0;
    ls->AdvectCount = 0;
  }
  ierr = ParticleLS_AdjustRadii(pls, ls); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}

PetscErrorCode LevelSetAdvectPLSRK2FullStep(LevelSet ls, Grid velgrid, PetscReal dt) {
  dl_scope_begin();
  PetscErrorCode ierr = pls_rk2_full(ls, velgrid, dt);
  dl_scope_end();
  return ierr;
}
