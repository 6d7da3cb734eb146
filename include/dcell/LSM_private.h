/*
 * LSM_private.h -- DCell's private level-set declarations
 * (LevelSetMethod/LSM_private.h:4-77): the particle-level-set types, whose
 * layout callers rely on (particles are read back as struct _Particle), the
 * log events and the private function prototypes.  The log events are
 * declared extern here and defined once by the container library.
 */
#ifndef LSM_PRIVATE_H_
#define LSM_PRIVATE_H_

extern PetscLogEvent EVENT_LevelSetUpdateIrregularNodeList;
extern PetscLogEvent EVENT_LevelSetReinitialize;
extern PetscLogEvent EVENT_LevelSetWriteIrregularNodeList;
extern PetscLogEvent EVENT_LevelSetGetVelocity;
extern PetscLogEvent EVENT_LevelSetAdvectSL;
extern PetscLogEvent EVENT_ParticleLS_AdvectParticles;
extern PetscLogEvent EVENT_ParticleLS_ErrorCorrection;
extern PetscLogEvent EVENT_ParticleLS_AdjustRadii;
extern PetscLogEvent EVENT_ParticleLS_ReseedParticles;

PetscErrorCode LevelSetRegisterEvents(void);
PetscErrorCode LevelSetUpdateIrregularNodeList_2D(LevelSet ls);
PetscErrorCode LevelSetUpdateIrregularNodeList_3D(LevelSet ls);
PetscErrorCode LevelSetGetVelocity(LevelSet ls, int ga, Grid velgrid);
PetscErrorCode LevelSetGatherVelocity(LevelSet ls, int ga, Grid velgrid);
PetscErrorCode LevelSetAdvectSL_2D(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetAdvectSL_3D(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSetCFLIncrement(LevelSet ls, Grid velgrid, PetscReal dt);
PetscErrorCode LevelSet_MaxVelocity(LevelSet ls, Grid velgrid);

typedef struct _Particle {
  Coor X;
  PetscReal radius;
} *Particle;

struct _ParticleLS {
  /* parameters */
  PetscReal R_MIN;   /* minimum particle radius */
  PetscReal R_MAX;   /* maximum particle radius */
  PetscReal D_INIT;  /* initial placement within +/- D_INIT of the zero level set */
  PetscInt S_INIT;   /* initial seeding density */
  int L_MAX;         /* max iterations when attracting to phi_goal */
  PetscReal G_TOL;   /* tolerance of phi == phi_goal */
  PetscReal E_DIST;  /* escaped-particle distance when deleted */

  /* actions */
  PetscErrorCode (*AdvectParticles)(ParticleLS pls, Coor dh, Grid velgrid, PetscReal dt);
  PetscErrorCode (*ErrorCorrection)(ParticleLS pls, LevelSet ls);

  /* objects */
  PetscRandom rnd;
  Array particles;
  Array particles0;  /* particle positions at the start of an RK2 step */
};

PetscErrorCode ParticleLS_AdjustRadii(ParticleLS pls, LevelSet ls);
PetscErrorCode ParticleLS_AdvectParticles2D(ParticleLS pls, Coor dh, Grid velgrid, PetscReal dt);
PetscErrorCode ParticleLS_AdvectParticles3D(ParticleLS pls, Coor dh, Grid velgrid, PetscReal dt);
PetscErrorCode ParticleLS_ErrorCorrection2D(ParticleLS pls, LevelSet ls);
PetscErrorCode ParticleLS_ErrorCorrection3D(ParticleLS pls, LevelSet ls);
PetscErrorCode ParticleLS_ReseedParticles(ParticleLS pls, LevelSet ls);

#endif /* LSM_PRIVATE_H_ */
