/*
 * irregular.c -- LevelSetUpdateIrregularNodeList (LevelSetMethod/
 * IrregularNodes.c:8-28) and its 2D / 3D halves (IrregularNodes.c:30-140,
 * IrregularNodes3D.c:6-117) over the C ABI: the records of every band node
 * are computed on the GPU (dcl_irregular_nodes, which takes the dimension
 * from the window) in the reference's append order and appended to
 * ls->irregularNodes with the fields the reference sets (pos, axis, shift,
 * signCenter, signFace, d, X, op); the other fields are zero, as after the
 * reference's ArrayZero (IrregularNodes.c:17).
 */
#include <string.h>

#include "LevelSetMethod.h"
#include "LSM_private.h"
#include "dropin.h"

/* appends the records of ls->band to ls->irregularNodes */
static PetscErrorCode append_records(LevelSet ls, const char *where) {
  PetscErrorCode ierr;
  const int len = ArrayLength(ls->band);
  if (len == 0) return 0;
  void *dphi, *dband, *dout, *hrec;
  /* phi over the band's box grown by 4 nodes: a record reads the node's
   * 3x3x3 neighbourhood and its projection the 5x5x5 one, so the box holds
   * every value read, and a node is as far from the box's edge as from the
   * window's wherever the box ends inside the window */
  dl_box bb, pb;
  dl_bbox(ArrayGetData(ls->band), len, sizeof(iCoor), ls->phi->is2D ? 2 : 3, &bb);
  if (!dl_box_around(ls->phi, &bb, 4, 0, &pb)) return dl_fail(where, DCL_ERR_ARG_SIZ);
  dcl_window w;
  DL_CHECK(where, dl_upload_box(DL_PHI, ls->phi, &pb, &dphi, &w, NULL, 0));
  DL_CHECK(where, dl_slot(DL_BAND, sizeof(iCoor) * (size_t)len, &dband));
  DL_CHECK(where, dl_htod(dband, ArrayGetData(ls->band), sizeof(iCoor) * (size_t)len, NULL));
  /* at most 1 projection + 3 (2) crossings + 6 (2) face crossings per node */
  const int64_t cap = 10 * (int64_t)len;
  DL_CHECK(where, dl_slot(DL_IRREG, sizeof(dcl_irregular_node) * (size_t)cap, &dout));
  int64_t count = 0;
  DL_CHECK(where, dcl_irregular_nodes(dphi, &w, (const dcl_icoor *)dband, len, DCL_F64,
                                      (dcl_irregular_node *)dout, cap, &count, NULL));
  if (count > cap) return dl_fail(where, DCL_ERR_ARG_SIZ);
  DL_CHECK(where, dl_host(sizeof(dcl_irregular_node) * (size_t)count, &hrec));
  DL_CHECK(where, dl_dtoh(hrec, dout, sizeof(dcl_irregular_node) * (size_t)count, NULL));
  const int base = ArrayLength(ls->irregularNodes);
  ierr = ArraySetSize(ls->irregularNodes, base + (int)count); CHKERRQ(ierr);
  IrregularNode *nodes = (IrregularNode *)ArrayGetData(ls->irregularNodes) + base;
  memset(nodes, 0, sizeof(IrregularNode) * (size_t)count);
  const dcl_irregular_node *r = (const dcl_irregular_node *)hrec;
  for (int64_t i = 0; i < count; ++i) {
    IrregularNode *n = &nodes[i];
    n->pos.x = r[i].pos.x;
    n->pos.y = r[i].pos.y;
    n->pos.z = r[i].pos.z;
    n->axis = r[i].axis;
    n->shift = (VelFace)r[i].shift;
    n->signCenter = r[i].sign_center;
    n->signFace = r[i].sign_face;
    n->d = r[i].d;
    n->X.x = r[i].X.x;
    n->X.y = r[i].X.y;
    n->X.z = r[i].X.z;
    n->op.x = r[i].op.x;
    n->op.y = r[i].op.y;
    n->op.z = r[i].op.z;
  }
  return 0;
}

PetscErrorCode LevelSetUpdateIrregularNodeList_2D(LevelSet ls) {
  return append_records(ls, "LevelSetUpdateIrregularNodeList_2D");
}

PetscErrorCode LevelSetUpdateIrregularNodeList_3D(LevelSet ls) {
  return append_records(ls, "LevelSetUpdateIrregularNodeList_3D");
}

PetscErrorCode LevelSetUpdateIrregularNodeList(LevelSet ls) {
  PetscErrorCode ierr;
  PetscFunctionBegin;
  ierr = PetscLogEventBegin(EVENT_LevelSetUpdateIrregularNodeList, 0, 0, 0, 0); CHKERRQ(ierr);
  ierr = ArraySetSize(ls->irregularNodes, 0); CHKERRQ(ierr);
  ierr = ArrayZero(ls->irregularNodes); CHKERRQ(ierr);
  if (ls->phi->is2D) {
    ierr = LevelSetUpdateIrregularNodeList_2D(ls); CHKERRQ(ierr);
  } else {
    ierr = LevelSetUpdateIrregularNodeList_3D(ls); CHKERRQ(ierr);
  }
  ierr = PetscLogEventEnd(EVENT_LevelSetUpdateIrregularNodeList, 0, 0, 0, 0); CHKERRQ(ierr);
  PetscFunctionReturn(0);
}
