/*
 * dropin.h -- internals shared by the C drop-in translation units: device
 * buffers kept between calls, host staging, window conversion and error
 * reporting.  The drop-in is called from one thread per process, like the
 * reference (which keeps its velocity grid and correction grids in statics,
 * LevelSetMethod/Advect.c:12, ParticleLS3D.c:51); inside a call, box
 * uploads pack rows on a small OpenMP team.
 */
#ifndef DCELL_DROPIN_H
#define DCELL_DROPIN_H

#include <stddef.h>

#include "Grid.h"
#include "dcell_lsm.h"

enum {
  DL_PHI0,
  DL_PHI,
  DL_VEL,
  DL_BAND,
  DL_OUT,
  DL_PART,
  DL_NODES,
  DL_KAPPA,
  DL_IRREG,
  DL_SURF,
  DL_NSLOTS
};

/* device buffer of at least `bytes` for a slot (grown, never shrunk) */
int dl_slot(int slot, size_t bytes, void **ptr);
/* host scratch buffer of at least `bytes` (pageable: a copy from it is done
 * with the buffer when dl_htod returns) */
int dl_host(size_t bytes, void **ptr);
/* dcl_memcpy_htod / dcl_memcpy_dtoh with the bytes counted for
 * LevelSetGPUTransferStats */
int dl_htod(void *dst, const void *src, size_t bytes, void *stream);
int dl_dtoh(void *dst, const void *src, size_t bytes, void *stream);
/* host threads that pack box uploads: DCELL_GPU_HOST_THREADS, default 8
 * (at most 64), read at the first upload */
int dl_threads(void);
/* release every buffer */
int dl_release(void);

dcl_window dl_window(Grid g);
/* upload the whole window of g into a slot */
int dl_upload_grid(int slot, Grid g, void **dev);
/* nodes of a window */
long dl_nodes(Grid g);

/* An index box of global node coordinates, bounds inclusive. */
typedef struct {
  int lo[3], hi[3];
} dl_box;

/* the bounding box of n coordinate triples (int32, `stride` bytes apart) */
void dl_bbox(const void *coords, long n, size_t stride, int dim, dl_box *b);
/* b grown by m nodes below and m + extra_hi above along the grid's axes,
 * clipped to g's window; returns 0 if the result is empty */
int dl_box_around(Grid g, const dl_box *b, int m, int extra_hi, dl_box *out);
/* Upload the box of g (every dof of its nodes) into a slot, packed as a
 * window of its own (its p = box lo), through page-locked staging with the
 * copy of each packed chunk overlapping the packing of the next (the rows of
 * a chunk are packed by dl_threads() threads).  *w gets the box's window
 * (spacing of g).  With maxabs, the largest |value| of each of the first
 * ncomp components over the box (NaN if any value is NaN).  Returns when the
 * copies are enqueued on the default stream. */
int dl_upload_box(int slot, Grid g, const dl_box *b, void **dev, dcl_window *w, double *maxabs, int ncomp);

/* A scope around a public entry point: within the outermost one, the
 * velocity grid that dl_velocity_box uploads is reused by later calls (the
 * reference does not change the velocity grid inside one of its steps). */
void dl_scope_begin(void);
void dl_scope_end(void);
/* The velocity grid over a box around the nodes of b that every departure
 * of a step of length dt from those nodes stays inside (the box grows until
 * its margin covers dt max|v| / dh plus the widest stencil), uploaded once
 * per scope; *M gets the margin used.  dh: the level set's spacing.  The
 * first box tried uses the margin the previous call on the same grid needed,
 * so a steady velocity costs one upload per call. */
int dl_velocity_box(Grid velgrid, const dl_box *b, double dt, Coor dh, int dim, void **dev, dcl_window *w,
                    int *M);
/* forget the velocity upload (the slot is about to hold something else) */
void dl_velocity_invalidate(void);

/* report a failed C-ABI call on stderr (the reference's callers drop the
 * return code of ls->Advect, LevelSetMethod/Advect.c:26) and pass it on */
int dl_fail(const char *where, int rc);

#define DL_CHECK(where, call)           \
  do {                                  \
    int rc__ = (call);                  \
    if (rc__) return dl_fail(where, rc__); \
  } while (0)

#endif
