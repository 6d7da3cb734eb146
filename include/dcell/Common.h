/*
 * Common.h -- the parts of DCell's Common/Common.h interface the level-set
 * update uses: coordinate tuples, the staggered-grid shift table, and the
 * growable Array.  Types, field order and function signatures are those of
 * the reference (Common/Common.h:18-51, :75-95) so that code written against
 * the reference compiles unchanged; without PETSc the scalar types come from
 * petsc_compat.h.
 */
#ifndef COMMON_H_
#define COMMON_H_

#ifdef DCELL_WITH_PETSC
#include "petsc.h"
#else
#include "petsc_compat.h"
#endif

#define UNUSED(x) (void)(x)

/* arbitrary tuple of at most three reals; in 2D z is ignored */
typedef struct {
  PetscReal x, y, z;
} Coor;

typedef struct {
  PetscInt x, y, z;
} iCoor;

/* half-cell shift of the u, v, w faces of the MAC grid */
static const PetscReal Tensor1[][3] = {
  {0.5, 0, 0},
  {0, 0.5, 0},
  {0, 0, 0.5}
};

typedef enum { CELL_CENTER, U_FACE, V_FACE, W_FACE } VelFace;
static const iCoor STAGGERED_GRID[4] = {
  {0, 0, 0},
  {-1, 0, 0},
  {0, -1, 0},
  {0, 0, -1}
};

PetscErrorCode InterpolateVelocity2D(const int udof, PetscReal ***field, const Coor X, Coor *vel);
PetscErrorCode InterpolateVelocity3D(const int udof, PetscReal ****field, const Coor X, Coor *vel);

/* growable array (Common/Array.c) */
typedef struct _Array *Array;
PetscErrorCode ArrayCreate(const char name[], int elemSize, Array *array);
PetscErrorCode ArrayDestroy(Array a);
PetscErrorCode ArraySetName(Array a, const char name[]);
PetscErrorCode ArraySetSize(Array a, int size);
PetscErrorCode ArraySetMaxSize(Array a, int maxsize);
PetscErrorCode ArrayAppend(Array a, void *elem);
PetscErrorCode ArrayAppendPtr(Array a, void *elem);
PetscErrorCode ArrayGet(Array a, int i, void *elem);
PetscErrorCode ArrayGetP(Array a, int i, void *elem);
PetscErrorCode ArrayWrite(Array a, int t);
PetscErrorCode ArrayZero(Array a);
PetscErrorCode ArrayDelete1(Array a, int idx);
PetscErrorCode ArrayCopy(Array src, Array copy);
PetscErrorCode ArrayDuplicate(Array a, Array *newarray);
PetscErrorCode ArrayGetCoor(Array a, iCoor pos, void *elem);
PetscErrorCode ArrayGetCoorP(Array a, iCoor pos, void *elem);
PetscErrorCode ArraySetCoor(Array a, iCoor shift, iCoor size);
int ArrayLength(Array a);
void *ArrayGetData(Array a);

#endif /* COMMON_H_ */
