/*
 * LevelSetMethod.h -- DCell's public level-set interface
 * (LevelSetMethod/LevelSetMethod.h:10-94): the LevelSet and IrregularNode
 * structs, whose fields callers read directly, and the prototypes.  Kept
 * identical so code written against the reference compiles unchanged.
 *
 * libdcell_levelset implements the hot-path entry points on the GPU:
 * LevelSetAdvectSL, LevelSetAdvectSLRK2HalfStep/FullStep,
 * LevelSetAdvectAndReinit, LevelSetAdvectPLS, LevelSetAdvectPLSRK2HalfStep/
 * FullStep (and, in LSM_private.h, the particle advection, correction and
 * radius functions, LevelSetCFLIncrement and LevelSet_MaxVelocity).  Added
 * GPU entry points are in dcell_gpu.h.  The remaining prototypes (FMM
 * reinitialisation, irregular-node lists, initialisations, the implicit step)
 * belong to the rest of DCell.
 */
#ifndef LEVELSETMETHOD_H_
#define LEVELSETMETHOD_H_

#include "Grid.h"
#include "Common.h"

typedef struct _ParticleLS *ParticleLS;
typedef struct _LevelSet *LevelSet;

struct _LevelSet {
  Grid phi;              /* the level set */
  Grid phi0;             /* phi at the start of the step */
  LevelSet psi;          /* implicit step */
  Grid tmp;              /* buffer while resizing */
  Array band;            /* narrow band coordinates [x, y, z] (iCoor) */
  Array irregularNodes;
  PetscReal bandWidth;
  PetscReal PHI_INF;
  PetscReal CFLthres;
  PetscReal CFLcount;
  PetscInt AdvectThres;
  PetscInt AdvectCount;
  PetscReal maxVel;
  ParticleLS pls;
  PetscErrorCode (*Advect)(LevelSet ls, Grid velgrid, PetscReal dt);
};

typedef struct _IrregularNode IrregularNode;
struct _IrregularNode {
  iCoor pos;
  int axis;
  VelFace shift;
  int signCenter;
  int signFace;
  int GAP;
  PetscReal nx, ny, nz;
  PetscReal sx, sy, sz;
  PetscReal rx, ry, rz;
  PetscReal d;
  PetscReal k;
  PetscReal k_nn, k_tt, k_nt;
  PetscReal F1, f1, f1_n, f1_t, f1_nn, f1_tt, f1_nt;
  PetscReal F2, f2, f2_n, f2_t, f2_nn, f2_tt, f2_nt;
  PetscReal F3, f3, f3_n, f3_t, f3_nn, f3_tt, f3_nt;
  PetscReal ftx, fty, ftz;
  PetscReal fa1, fa2;
  Coor X;
  Coor op;
  int numNei;
  PetscReal uj;
};

/* Level Set */
PetscErrorCode LevelSetCreate(Coor dh, iCoor pos, iCoor size, LevelSet *levelset);
PetscErrorCode LevelSetDestroy(LevelSet ls);
PetscErrorCode LevelSetSetBandWidth(LevelSet ls, PetscReal bandwidth);
PetscErrorCode LevelSetDuplicate(LevelSet ls, LevelSet *copy);
PetscErrorCode LevelSetResize(LevelSet ls);
PetscErrorCode LevelSetUpdateIrregularNodeList(LevelSet ls);
PetscErrorCode LevelSetWriteIrregularNodeList(LevelSet ls, int idx);
PetscErrorCode LevelSetNormalDirection(LevelSet ls, Coor X, Coor *n);
PetscReal LevelSetDiracDelta2D(PetscReal **phi, const Coor dh, const Coor X);
PetscReal LevelSetDiracDelta3D(PetscReal ***phi, const Coor dh, const Coor X);

PetscErrorCode LevelSetAdvect(LevelSet ls, int ga, PetscReal dt);
PetscErrorCode LevelSetAdvectSL(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetAdvectAndReinit(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetAdvectPLS(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetAdvectSLRK2HalfStep(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetAdvectSLRK2FullStep(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetAdvectImplicit(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetAdvectImplicitInit(LevelSet ls, PetscInt *n);
PetscErrorCode LevelSetAdvectImplicitRHS(LevelSet ls, int ga, PetscReal dt, PetscReal *g);
PetscErrorCode LevelSetAdvectImplicitUpdate(LevelSet ls, PetscReal lambda, PetscReal *dpsi);
PetscErrorCode LevelSetAdvectImplicitReinit(LevelSet ls, PetscReal dt);

PetscErrorCode LevelSetInitializeParticles(LevelSet ls);
PetscErrorCode ParticleLSDestroy(ParticleLS pls);
PetscErrorCode ParticleLSWriteParticles(ParticleLS pls, int t);
PetscErrorCode LevelSetAdvectPLSRK2HalfStep(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetAdvectPLSRK2FullStep(LevelSet ls, Grid velgrid, PetscReal dt);

PetscErrorCode LevelSetReinitialize(LevelSet ls);

/* Level Set Initializations */
PetscErrorCode LevelSetInitializeToCircle(Coor dh, Coor center, PetscReal radius, LevelSet *lset);
PetscErrorCode LevelSetInitializeToSphere(Coor dh, Coor center, PetscReal radius, LevelSet *lset);
PetscErrorCode LevelSetInitializeToStar2D(Coor dh, Coor center, PetscReal radius, PetscReal amp,
                                          PetscReal numPetals, LevelSet *lset);
PetscErrorCode LevelSetInitializeToStar3D(Coor dh, Coor center, PetscReal radius, PetscReal amp,
                                          PetscReal numPetals, LevelSet *lset);
PetscErrorCode LevelSetInitializeFromImage(LevelSet ls);

#endif /* LEVELSETMETHOD_H_ */
