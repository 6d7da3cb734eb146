/*
 * dropin.h -- internals shared by the C drop-in translation units: device
 * buffers kept between calls, host staging, window conversion and error
 * reporting.  The drop-in is single-threaded per process, like the
 * reference (which keeps its velocity grid and correction grids in statics,
 * LevelSetMethod/Advect.c:12, ParticleLS3D.c:51).
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
/* host staging buffer of at least `bytes` */
int dl_host(size_t bytes, void **ptr);
/* release every buffer */
int dl_release(void);

dcl_window dl_window(Grid g);
/* upload the whole window of g into a slot */
int dl_upload_grid(int slot, Grid g, void **dev);
/* nodes of a window */
long dl_nodes(Grid g);

/* report a failed C-ABI call on stderr (the reference's callers drop the
 * return code of ls->Advect, LevelSetMethod/Advect.c:26) and pass it on */
int dl_fail(const char *where, int rc);

#define DL_CHECK(where, call)           \
  do {                                  \
    int rc__ = (call);                  \
    if (rc__) return dl_fail(where, rc__); \
  } while (0)

#endif
