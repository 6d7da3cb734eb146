/*
 * reinit.c -- LevelSetReinitialize (LevelSetMethod/FastMarching.c:10-41) over
 * the C ABI: LevelSetResize first, as the reference does, then phi and the
 * irregular nodes go to the GPU, dcl_reinitialize_parallel recomputes phi in
 * parallel sweeps that replay the fast marching's pushes in value order (the
 * same band, phi equal to the fast marching's at most nodes and within its
 * discretisation error elsewhere, the band's two sides in window order), and
 * phi (the whole window: +-PHI_INF outside the new band) and the new band come
 * back into the level set.  With DCELL_GPU_REINIT=exact in the environment
 * dcl_reinitialize runs the second-order fast marching itself in the
 * reference's heap order instead (bit for bit the reference, band order
 * included): a sequential chain of dependent steps, slower on the GPU than
 * the reference's on one host core, kept as the parity check.
 */
#include <stdlib.h>
#include <string.h>

#include "LevelSetMethod.h"
#include "LSM_private.h"
#include "dropin.h"

PetscErrorCode LevelSetReinitialize(LevelSet ls) {
  static const char *where = "LevelSetReinitialize";
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = PetscLogEventBegin(EVENT_LevelSetReinitialize, 0, 0, 0, 0); CHKERRQ(ierr);
  ierr = LevelSetResize(ls); CHKERRQ(ierr);
  void *dphi, *drec, *dband, *hrec;
  DL_CHECK(where, dl_upload_grid(DL_PHI, ls->phi, &dphi));
  const int len = ArrayLength(ls->irregularNodes);
  const IrregularNode *nodes = (const IrregularNode *)ArrayGetData(ls->irregularNodes);
  const size_t rbytes = sizeof(dcl_irregular_node) * (size_t)(len > 0 ? len : 1);
  DL_CHECK(where, dl_host(rbytes, &hrec));
  dcl_irregular_node *r = (dcl_irregular_node *)hrec;
  memset(r, 0, rbytes);
  for (int i = 0; i < len; ++i) { /* the fields the boundary condition reads */
    r[i].pos.x = nodes[i].pos.x;
    r[i].pos.y = nodes[i].pos.y;
    r[i].pos.z = nodes[i].pos.z;
    r[i].axis = nodes[i].axis;
    r[i].sign_center = nodes[i].signCenter;
    r[i].op.x = nodes[i].op.x;
    r[i].op.y = nodes[i].op.y;
    r[i].op.z = nodes[i].op.z;
  }
  DL_CHECK(where, dl_slot(DL_IRREG, rbytes, &drec));
  DL_CHECK(where, dl_htod(drec, r, rbytes, NULL));
  const long nodes_w = dl_nodes(ls->phi);
  DL_CHECK(where, dl_slot(DL_BAND, sizeof(iCoor) * (size_t)nodes_w, &dband));
  dcl_reinit_cell cell;
  memset(&cell, 0, sizeof cell);
  cell.phi = dphi;
  cell.win = dl_window(ls->phi);
  cell.nodes = (const dcl_irregular_node *)drec;
  cell.count = len;
  cell.band_width = ls->bandWidth;
  cell.phi_inf = ls->PHI_INF;
  cell.band = (dcl_icoor *)dband;
  cell.band_capacity = nodes_w;
  const char *mode = getenv("DCELL_GPU_REINIT");
  if (mode && strcmp(mode, "exact") == 0) {
    DL_CHECK(where, dcl_reinitialize(&cell, 1, 0, NULL));
  } else {
    DL_CHECK(where, dcl_reinitialize_parallel(&cell, 1, 0, NULL, NULL));
  }
  DL_CHECK(where, dl_dtoh(ls->phi->v1, dphi, sizeof(double) * (size_t)nodes_w, NULL));
  ierr = ArraySetSize(ls->band, (int)cell.band_count); CHKERRQ(ierr);
  if (cell.band_count > 0) {
    DL_CHECK(where, dl_dtoh(ArrayGetData(ls->band), dband, sizeof(iCoor) * (size_t)cell.band_count, NULL));
  }
  ierr = PetscLogEventEnd(EVENT_LevelSetReinitialize, 0, 0, 0, 0); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}
