/*
 * surface.c -- IIMUpdateSurfaceQuantities_2D / _3D (IIM/SurfaceQuantities.c:
 * 31-118) over the C ABI: the X of every irregular node goes to the GPU,
 * dcl_surface_quantities computes the unit normal, the tangents, k, k_nn,
 * k_tt and k_nt, the drop-in writes them into the nodes and then calls the
 * interfacial force iim->F on each node in list order, on the host, as the
 * reference does inside its loop (the force functions read only their own
 * node).
 */
#include <math.h>
#include <string.h>

#include "ImmersedInterfaceMethod.h"
#include "LSM_private.h"
#include "dropin.h"

static PetscErrorCode surface_quantities(IIM iim, LevelSet ls, const char *where) {
  const int len = ArrayLength(ls->irregularNodes);
  if (len == 0) return 0;
  IrregularNode *nodes = (IrregularNode *)ArrayGetData(ls->irregularNodes);
  void *dphi, *drec, *dout, *hrec;
  /* phi over the box of the records' interface points grown by 4 nodes (each
   * reads the 4x4x4 neighbourhood of its cell) */
  const int dim = ls->phi->is2D ? 2 : 3;
  dl_box bb, pb;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < len; ++i) {
    const double x[3] = {nodes[i].X.x, nodes[i].X.y, nodes[i].X.z};
    for (int a = 0; a < dim; ++a) {
      if (x[a] < lo[a]) lo[a] = x[a];
      if (x[a] > hi[a]) hi[a] = x[a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    const int ok = a < dim && lo[a] <= hi[a] && lo[a] > -1e9 && hi[a] < 1e9;
    bb.lo[a] = ok ? (int)floor(lo[a]) : -1000000000;
    bb.hi[a] = ok ? (int)floor(hi[a]) + 1 : 1000000000;
  }
  if (!dl_box_around(ls->phi, &bb, 4, 0, &pb)) return dl_fail(where, DCL_ERR_ARG_SIZ);
  dcl_window w;
  DL_CHECK(where, dl_upload_box(DL_PHI, ls->phi, &pb, &dphi, &w, NULL, 0));
  /* the records' X, packed as the C ABI's irregular-node records */
  const size_t rbytes = sizeof(dcl_irregular_node) * (size_t)len;
  const size_t sbytes = sizeof(dcl_surface) * (size_t)len;
  DL_CHECK(where, dl_host(rbytes > sbytes ? rbytes : sbytes, &hrec));
  dcl_irregular_node *r = (dcl_irregular_node *)hrec;
  memset(r, 0, rbytes);
  for (int i = 0; i < len; ++i) {
    r[i].X.x = nodes[i].X.x;
    r[i].X.y = nodes[i].X.y;
    r[i].X.z = nodes[i].X.z;
  }
  DL_CHECK(where, dl_slot(DL_IRREG, rbytes, &drec));
  DL_CHECK(where, dl_htod(drec, r, rbytes, NULL));
  DL_CHECK(where, dl_slot(DL_SURF, sbytes, &dout));
  DL_CHECK(where, dcl_surface_quantities(dphi, &w, (const dcl_irregular_node *)drec, len, DCL_F64,
                                         (dcl_surface *)dout, NULL));
  DL_CHECK(where, dl_dtoh(hrec, dout, sbytes, NULL));
  const dcl_surface *q = (const dcl_surface *)hrec;
  const int is2D = ls->phi->is2D;
  for (int i = 0; i < len; ++i) {
    IrregularNode *n = &nodes[i];
    n->k = q[i].k;
    n->k_nn = q[i].k_nn;
    n->k_tt = q[i].k_tt;
    n->k_nt = q[i].k_nt;
    n->nx = q[i].n.x;
    n->ny = q[i].n.y;
    n->sx = q[i].s.x;
    n->sy = q[i].s.y;
    if (!is2D) { /* the 2D reference leaves nz, sz and r as they are */
      n->nz = q[i].n.z;
      n->sz = q[i].s.z;
      n->rx = q[i].r.x;
      n->ry = q[i].r.y;
      n->rz = q[i].r.z;
    }
    if (iim->F) iim->F(n, iim->context);
  }
  return 0;
}

PetscErrorCode IIMUpdateSurfaceQuantities_2D(IIM iim, LevelSet ls) {
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = PetscLogEventBegin(EVENT_IIMUpdateSurfaceQuantities, 0, 0, 0, 0); CHKERRQ(ierr);
  iim->dh = ls->phi->d; /* SurfaceQuantities.c:44 */
  ierr = surface_quantities(iim, ls, "IIMUpdateSurfaceQuantities_2D"); CHKERRQ(ierr);
  ierr = PetscLogEventEnd(EVENT_IIMUpdateSurfaceQuantities, 0, 0, 0, 0); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}

PetscErrorCode IIMUpdateSurfaceQuantities_3D(IIM iim, LevelSet ls) {
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = PetscLogEventBegin(EVENT_IIMUpdateSurfaceQuantities, 0, 0, 0, 0); CHKERRQ(ierr);
  ierr = surface_quantities(iim, ls, "IIMUpdateSurfaceQuantities_3D"); CHKERRQ(ierr);
  ierr = PetscLogEventEnd(EVENT_IIMUpdateSurfaceQuantities, 0, 0, 0, 0); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}
