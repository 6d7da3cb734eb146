/*
 * compat.c -- PETSc-free containers behind DCell's level-set interface, so
 * the drop-in can be built and tested without the rest of DCell: Grid
 * (Common/Grid.c:8-148, 415-423), Array (Common/Array.c), LevelSetCreate /
 * Destroy (LevelSetMethod/LevelSet.c:6-79), the log events, and weak no-op
 * stand-ins for the steps outside the level-set update (irregular-node
 * rebuild, FMM reinitialisation, particle reseeding).  A DCell build links
 * the reference's own files instead and does not use this library.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>

#include "LevelSetMethod.h"
#include "LSM_private.h"

PetscLogEvent EVENT_LevelSetUpdateIrregularNodeList;
PetscLogEvent EVENT_IIMUpdateSurfaceQuantities;
PetscLogEvent EVENT_LevelSetReinitialize;
PetscLogEvent EVENT_LevelSetWriteIrregularNodeList;
PetscLogEvent EVENT_LevelSetGetVelocity;
PetscLogEvent EVENT_LevelSetAdvectSL;
PetscLogEvent EVENT_ParticleLS_AdvectParticles;
PetscLogEvent EVENT_ParticleLS_ErrorCorrection;
PetscLogEvent EVENT_ParticleLS_AdjustRadii;
PetscLogEvent EVENT_ParticleLS_ReseedParticles;

PetscErrorCode LevelSetRegisterEvents(void) { return 0; }

/* ------------------------------------------------------------------ Grid */
/* pointer tables indexed with global coordinates, as VecGetArray2d/3d/4d
 * build them (Grid_MakeGrid, Common/Grid.c:69-93) */
static void grid_free_tables(Grid g) {
  if (!g->grid) return;
  const int nz = g->is2D ? 1 : g->n.z;
  if (g->is2D) {
    if (g->dof == 1) free((PetscReal **)g->grid + g->p.y);
    else {
      PetscReal ***t = (PetscReal ***)g->grid;
      free(t[g->p.y] + g->p.x);
      free(t + g->p.y);
    }
  } else {
    PetscReal ***t = (PetscReal ***)g->grid;
    if (g->dof == 1) {
      free(t[g->p.z] + g->p.y);
      free(t + g->p.z);
    } else {
      PetscReal ****t4 = (PetscReal ****)g->grid;
      free(t4[g->p.z][g->p.y] + g->p.x);
      free(t4[g->p.z] + g->p.y);
      free(t4 + g->p.z);
    }
  }
  (void)nz;
  g->grid = NULL;
}

static PetscErrorCode grid_make_tables(Grid g) {
  const int nx = g->n.x, ny = g->n.y, nz = g->is2D ? 1 : g->n.z, dof = g->dof;
  if (g->is2D && dof == 1) {
    PetscReal **rows = malloc(sizeof(PetscReal *) * ny);
    if (!rows) return PETSC_ERR_MEM;
    for (int j = 0; j < ny; ++j) rows[j] = g->v1 + (long)j * nx - g->p.x;
    g->grid = (PetscReal *)(rows - g->p.y);
  } else if (g->is2D) {
    PetscReal ***rows = malloc(sizeof(PetscReal **) * ny);
    PetscReal **cells = malloc(sizeof(PetscReal *) * (size_t)ny * nx);
    if (!rows || !cells) { free(rows); free(cells); return PETSC_ERR_MEM; }
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) cells[(long)j * nx + i] = g->v1 + ((long)j * nx + i) * dof;
      rows[j] = cells + (long)j * nx - g->p.x;
    }
    g->grid = (PetscReal *)(rows - g->p.y);
  } else if (dof == 1) {
    PetscReal ***planes = malloc(sizeof(PetscReal **) * nz);
    PetscReal **rows = malloc(sizeof(PetscReal *) * (size_t)nz * ny);
    if (!planes || !rows) { free(planes); free(rows); return PETSC_ERR_MEM; }
    for (int k = 0; k < nz; ++k) {
      for (int j = 0; j < ny; ++j) rows[(long)k * ny + j] = g->v1 + ((long)k * ny + j) * nx - g->p.x;
      planes[k] = rows + (long)k * ny - g->p.y;
    }
    g->grid = (PetscReal *)(planes - g->p.z);
  } else {
    PetscReal ****planes = malloc(sizeof(PetscReal ***) * nz);
    PetscReal ***rows = malloc(sizeof(PetscReal **) * (size_t)nz * ny);
    PetscReal **cells = malloc(sizeof(PetscReal *) * (size_t)nz * ny * nx);
    if (!planes || !rows || !cells) { free(planes); free(rows); free(cells); return PETSC_ERR_MEM; }
    for (int k = 0; k < nz; ++k) {
      for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i)
          cells[((long)k * ny + j) * nx + i] = g->v1 + (((long)k * ny + j) * nx + i) * dof;
        rows[(long)k * ny + j] = cells + ((long)k * ny + j) * nx - g->p.x;
      }
      planes[k] = rows + (long)k * ny - g->p.y;
    }
    g->grid = (PetscReal *)(planes - g->p.z);
  }
  return 0;
}

PetscErrorCode GridCreate(Coor dh, iCoor pos, iCoor size, int dof, Grid *grid) {
  Grid g = calloc(1, sizeof(struct _Grid));
  if (!g) return PETSC_ERR_MEM;
  g->d = dh;
  g->n = size;
  g->p = pos;
  g->dof = dof;
  strcpy(g->name, "grid");
  g->is2D = size.z < 2 ? PETSC_TRUE : PETSC_FALSE;
  g->SIZE = g->is2D ? size.x * size.y * dof : size.x * size.y * size.z * dof;
  g->MAXSIZE = g->SIZE;
  g->v1 = calloc(g->SIZE > 0 ? g->SIZE : 1, sizeof(PetscReal));
  if (!g->v1) { free(g); return PETSC_ERR_MEM; }
  PetscErrorCode ierr = grid_make_tables(g);
  if (ierr) { free(g->v1); free(g); return ierr; }
  g->Interpolate = NULL;
  *grid = g;
  return 0;
}

PetscErrorCode GridDestroy(Grid g) {
  if (!g) return 0;
  grid_free_tables(g);
  free(g->v1);
  free(g);
  return 0;
}

PetscErrorCode GridDuplicate(Grid g, Grid *newg) { return GridCreate(g->d, g->p, g->n, g->dof, newg); }

PetscErrorCode GridSetDx(Grid g, Coor d) {
  g->d = d;
  return 0;
}

PetscErrorCode GridSetName(Grid g, const char *name) {
  strncpy(g->name, name, PETSC_MAX_PATH_LEN - 1);
  g->name[PETSC_MAX_PATH_LEN - 1] = 0;
  return 0;
}

PetscErrorCode GridGet(Grid g, void *grid) {
  *((void **)grid) = g->grid;
  return 0;
}

PetscErrorCode GridGetBounds(Grid g, iCoor *p, iCoor *q) {
  *p = g->p;
  q->x = g->p.x + g->n.x;
  q->y = g->p.y + g->n.y;
  q->z = g->p.z + g->n.z;
  return 0;
}

/* Common/Grid.c:118-152: new window, storage grown if needed, zeroed */
PetscErrorCode GridResize(Grid g, iCoor pos, iCoor size) {
  grid_free_tables(g);
  if (g->is2D) {
    pos.z = 0;
    size.z = 0;
  }
  g->p = pos;
  g->n = size;
  g->SIZE = g->is2D ? size.x * size.y * g->dof : size.x * size.y * size.z * g->dof;
  if (g->SIZE > g->MAXSIZE) {
    free(g->v1);
    g->v1 = malloc(sizeof(PetscReal) * g->SIZE);
    if (!g->v1) return PETSC_ERR_MEM;
    g->MAXSIZE = g->SIZE;
  }
  memset(g->v1, 0, sizeof(PetscReal) * g->MAXSIZE);
  return grid_make_tables(g);
}

/* Common/Grid.c:415-423 */
PetscErrorCode GridCopy(Grid g, Grid copy) {
  PetscErrorCode ierr = GridResize(copy, g->p, g->n);
  CHKERRQ(ierr);
  memcpy(copy->v1, g->v1, sizeof(PetscReal) * g->SIZE);
  return 0;
}

/* ----------------------------------------------------------------- Array */
struct _Array {
  MPI_Comm comm;
  void *dataArray;
  int len;
  int ELEMSIZE;
  int MAXSIZE;
  PetscReal scale;
  char name[64];
  iCoor p, q;
  iCoor size;
};

PetscErrorCode ArrayCreate(const char name[], int elemSize, Array *array) {
  const int initSize = 1024;
  Array a = calloc(1, sizeof(struct _Array));
  if (!a) return PETSC_ERR_MEM;
  a->dataArray = calloc(initSize, elemSize);
  if (!a->dataArray) { free(a); return PETSC_ERR_MEM; }
  a->ELEMSIZE = elemSize;
  a->MAXSIZE = initSize;
  a->len = 0;
  a->scale = 1.1;
  strncpy(a->name, name, sizeof(a->name) - 1);
  *array = a;
  return 0;
}

PetscErrorCode ArrayDestroy(Array a) {
  if (!a) return 0;
  free(a->dataArray);
  free(a);
  return 0;
}

PetscErrorCode ArraySetName(Array a, const char name[]) {
  strncpy(a->name, name, sizeof(a->name) - 1);
  return 0;
}

PetscErrorCode ArraySetSize(Array a, int size) {
  if (size < 0) return PETSC_ERR_ARG_OUTOFRANGE;
  a->len = size;
  if (a->MAXSIZE < size) {
    const int newsize = (int)(a->scale * size) > size ? (int)(a->scale * size) : size;
    void *tmp = calloc(newsize, a->ELEMSIZE);
    if (!tmp) return PETSC_ERR_MEM;
    memcpy(tmp, a->dataArray, (size_t)a->ELEMSIZE * a->MAXSIZE);
    free(a->dataArray);
    a->dataArray = tmp;
    a->MAXSIZE = newsize;
  }
  return 0;
}

PetscErrorCode ArraySetMaxSize(Array a, int maxsize) {
  const int len = a->len;
  PetscErrorCode ierr = ArraySetSize(a, maxsize);
  CHKERRQ(ierr);
  return ArraySetSize(a, len);
}

PetscErrorCode ArrayGet(Array a, int i, void *elem) {
  if (i < 0 || i >= a->len) return PETSC_ERR_ARG_OUTOFRANGE;
  *((void **)elem) = (char *)a->dataArray + (size_t)a->ELEMSIZE * i;
  return 0;
}

PetscErrorCode ArrayAppend(Array a, void *elem) {
  PetscErrorCode ierr = ArraySetSize(a, a->len + 1);
  CHKERRQ(ierr);
  return ArrayGet(a, a->len - 1, elem);
}

PetscErrorCode ArrayZero(Array a) {
  memset(a->dataArray, 0, (size_t)a->ELEMSIZE * a->MAXSIZE);
  return 0;
}

PetscErrorCode ArrayDelete1(Array a, int idx) {
  if (idx < 0 || idx >= a->len) return PETSC_ERR_ARG_OUTOFRANGE;
  memcpy((char *)a->dataArray + (size_t)a->ELEMSIZE * idx,
         (char *)a->dataArray + (size_t)a->ELEMSIZE * (a->len - 1), a->ELEMSIZE);
  a->len--;
  return 0;
}

PetscErrorCode ArrayCopy(Array src, Array copy) {
  PetscErrorCode ierr = ArraySetSize(copy, src->len);
  CHKERRQ(ierr);
  memcpy(copy->dataArray, src->dataArray, (size_t)src->len * src->ELEMSIZE);
  return 0;
}

PetscErrorCode ArrayDuplicate(Array a, Array *newarray) {
  PetscErrorCode ierr = ArrayCreate(a->name, a->ELEMSIZE, newarray);
  CHKERRQ(ierr);
  return ArraySetSize(*newarray, a->MAXSIZE);
}

int ArrayLength(Array a) { return a->len; }

void *ArrayGetData(Array a) { return a->dataArray; }

/* -------------------------------------------------------------- LevelSet */
PetscErrorCode LevelSetSetBandWidth(LevelSet ls, PetscReal bandwidth) {
  ls->bandWidth = bandwidth;
  ls->PHI_INF = ls->bandWidth + 1.5;
  return 0;
}

/* LevelSet.c:6-37 (PETSc options not read) */
PetscErrorCode LevelSetCreate(Coor dh, iCoor pos, iCoor size, LevelSet *levelset) {
  PetscErrorCode ierr;
  LevelSet ls = calloc(1, sizeof(struct _LevelSet));
  if (!ls) return PETSC_ERR_MEM;
  ierr = GridCreate(dh, pos, size, 1, &ls->phi); CHKERRQ(ierr);
  ierr = GridSetName(ls->phi, "phi"); CHKERRQ(ierr);
  ierr = GridCreate(dh, pos, size, 1, &ls->phi0); CHKERRQ(ierr);
  ierr = GridSetName(ls->phi0, "phi0"); CHKERRQ(ierr);
  ierr = GridCreate(dh, pos, size, 1, &ls->tmp); CHKERRQ(ierr);
  ierr = GridSetName(ls->tmp, "tmp"); CHKERRQ(ierr);
  ls->CFLthres = 3;
  ls->CFLcount = 0;
  ls->AdvectCount = 0;
  ls->AdvectThres = 1000;
  ierr = LevelSetSetBandWidth(ls, 6); CHKERRQ(ierr);
  ierr = ArrayCreate("irregularNodes", sizeof(IrregularNode), &ls->irregularNodes); CHKERRQ(ierr);
  ierr = ArrayCreate("band", sizeof(iCoor), &ls->band); CHKERRQ(ierr);
  /* the default step (LevelSet.c:33) lives in the drop-in library, which is
   * loaded after this one: look it up when a level set is created */
  ls->Advect = (PetscErrorCode(*)(LevelSet, Grid, PetscReal))dlsym(RTLD_DEFAULT, "LevelSetAdvectAndReinit");
  *levelset = ls;
  return 0;
}

PetscErrorCode ParticleLSDestroy(ParticleLS pls) {
  if (!pls) return 0;
  ArrayDestroy(pls->particles);
  if (pls->particles0) ArrayDestroy(pls->particles0);
  free(pls);
  return 0;
}

PetscErrorCode LevelSetDestroy(LevelSet ls) {
  if (!ls) return 0;
  if (ls->pls) ParticleLSDestroy(ls->pls);
  GridDestroy(ls->phi);
  GridDestroy(ls->phi0);
  GridDestroy(ls->tmp);
  ArrayDestroy(ls->irregularNodes);
  ArrayDestroy(ls->band);
  free(ls);
  return 0;
}

/* Steps outside the level-set update: no-ops here, the rest of DCell
 * provides them (LevelSetMethod/Resize.c:5-120, ParticleLS.c:202-295).  The
 * irregular-node list and the reinitialisation are the drop-in's
 * (libdcell_levelset.so); standalone, the window is taken as large enough
 * (DCell's LevelSetResize grows it when the band nears its edge). */
__attribute__((weak)) PetscErrorCode LevelSetResize(LevelSet ls) {
  (void)ls;
  return 0;
}

__attribute__((weak)) PetscErrorCode ParticleLS_ReseedParticles(ParticleLS pls, LevelSet ls) {
  (void)pls;
  (void)ls;
  return 0;
}
