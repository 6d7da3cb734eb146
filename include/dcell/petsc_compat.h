/*
 * petsc_compat.h -- the PETSc types and macros DCell's level-set interface
 * uses, for builds without PETSc.  Only what the drop-in (libdcell_levelset)
 * and its standalone containers (libdcell_compat) need: scalar types with
 * PETSc's widths (PetscReal = double, PetscInt = 32-bit int), opaque handles,
 * the error-return macros and the small math macros.  In a DCell build that
 * links PETSc, <petsc.h> provides all of these and this file is not used.
 */
#ifndef DCELL_PETSC_COMPAT_H
#define DCELL_PETSC_COMPAT_H

#include <math.h>
#include <stddef.h>
#include <stdint.h>

typedef double PetscReal;
typedef double PetscScalar;
typedef int PetscInt;
typedef int PetscErrorCode;
typedef int PetscLogEvent;
typedef enum { PETSC_FALSE, PETSC_TRUE } PetscBool;
typedef void *MPI_Comm;
typedef struct _p_Vec *Vec;
typedef struct _p_Mat *Mat;
typedef struct _p_DM *DM;
typedef struct _p_PetscViewer *PetscViewer;
typedef struct _p_PetscRandom *PetscRandom;

#define PETSC_MAX_PATH_LEN 4096
#define PETSC_PI 3.14159265358979323846
#define PETSC_MAX_REAL 1.7976931348623157e+308

#define PETSC_ERR_MEM 55
#define PETSC_ERR_SUP 56
#define PETSC_ERR_ARG_OUTOFRANGE 63

#define PetscFunctionBegin do { } while (0)
#define PetscFunctionReturn(rc) return (rc)
#define CHKERRQ(ierr) do { if (ierr) return (ierr); } while (0)
#define SETERRQ(comm, code, msg) return (code)

#define PetscLogEventBegin(e, a, b, c, d) 0
#define PetscLogEventEnd(e, a, b, c, d) 0
#define PetscInfo(obj, msg) 0

#define PetscAbs(a) (((a) >= 0) ? (a) : -(a))
#define PetscSign(a) (((a) >= 0) ? ((a) == 0 ? 0 : 1) : -1)
#define PetscMax(a, b) (((a) < (b)) ? (b) : (a))
#define PetscMin(a, b) (((a) < (b)) ? (a) : (b))
#define PetscSqr(a) ((a) * (a))
#define PetscSqrtScalar(a) sqrt(a)

#endif
