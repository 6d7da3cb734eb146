/*
 * dcell_gpu.h -- level-set entry points added by the MI355X build, at the
 * level of DCell's LevelSet API (LevelSetMethod.h stays unchanged).  None has
 * a counterpart in the reference; each is implemented in libdcell_levelset
 * over the C ABI of dcell_lsm.h.
 */
#ifndef DCELL_GPU_H_
#define DCELL_GPU_H_

#include "LevelSetMethod.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Interpolant and backtrace of LevelSetAdvectSL for this process:
 * interp DCL_LINEAR (the reference's Bilinear3D/2D, the default) or
 * DCL_CUBIC (Catmull-Rom); backtrace DCL_EULER (the reference's, the
 * default) or DCL_RK2.  Values as in dcell_lsm.h. */
PetscErrorCode LevelSetAdvectSLSetScheme(int interp, int backtrace);

/* Corner base of the escaped-particle correction: PETSC_TRUE reproduces the
 * reference's A.z = floor(X.y) (ParticleLS3D.c:76, the default, for parity);
 * PETSC_FALSE uses floor(X.z). */
PetscErrorCode LevelSetPLSSetReferenceCornerZ(PetscBool reference);

/* Curvature of ls->phi at every node of its window into kappa (a dof-1 grid
 * of the same window): GridFunction3D_Curv / GridFunction2D_Curv evaluated
 * on the GPU for the whole window at once.  Nodes where the stencil leaves
 * the window get 0. */
PetscErrorCode LevelSetCurvature(LevelSet ls, Grid kappa);

/* Frees the device buffers the drop-in keeps between calls. */
PetscErrorCode LevelSetGPUFinalize(void);

/* Bytes the drop-in has copied host -> device and device -> host since the
 * library was loaded (running totals; either pointer may be NULL). */
PetscErrorCode LevelSetGPUTransferStats(unsigned long long *htod_bytes, unsigned long long *dtoh_bytes);

#ifdef __cplusplus
}
#endif
#endif /* DCELL_GPU_H_ */
