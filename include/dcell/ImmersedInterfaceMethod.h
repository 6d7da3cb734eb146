/*
 * ImmersedInterfaceMethod.h -- the part of DCell's IIM interface
 * (IIM/ImmersedInterfaceMethod.h) that the drop-in implements: the IIM
 * object's layout (IIM/ImmersedInterfaceMethod.h:20-34, kept field for field
 * because callers and the rest of DCell read it directly) and the surface
 * quantities at the irregular nodes (IIM/SurfaceQuantities.c:31-118).
 */
#ifndef DCELL_IIM_H
#define DCELL_IIM_H

#include "LevelSetMethod.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _LocalCoor *LocalCoor;
typedef struct _LeastSq *LeastSq;
typedef struct _SpatialIndex *SpatialIndex;

/* per-node interfacial force, called after the node's surface quantities */
typedef void (*InterfacialForce)(IrregularNode *n, void *context);

typedef struct _IIM {
  LocalCoor lc;
  LeastSq lsq;
  PetscInt Np;
  PetscReal eps;      /* radius around a node considered a neighbour */
  PetscReal mu;       /* fluid viscosity */
  InterfacialForce F; /* normal and tangential force functions */
  void *context;      /* passed to F */
  Array idx, coor, val;
  Array debug;
  Coor dh;            /* grid widths */
  SpatialIndex sidx;
} *IIM;

extern PetscLogEvent EVENT_IIMUpdateSurfaceQuantities;

/* IIM/SurfaceQuantities.c:31-62 and :64-118: normal, tangents and
 * curvatures of every irregular node of ls on the GPU, then iim->F per
 * node on the host, in list order */
PetscErrorCode IIMUpdateSurfaceQuantities_2D(IIM iim, LevelSet ls);
PetscErrorCode IIMUpdateSurfaceQuantities_3D(IIM iim, LevelSet ls);

#ifdef __cplusplus
}
#endif
#endif
