/*
 * dropin_device.c -- device buffers and host staging of the C drop-in.
 */
#include <stdio.h>
#include <stdlib.h>

#include "dropin.h"

static void *g_dev[DL_NSLOTS];
static size_t g_dev_bytes[DL_NSLOTS];
static void *g_host;
static size_t g_host_bytes;

int dl_slot(int slot, size_t bytes, void **ptr) {
  if (slot < 0 || slot >= DL_NSLOTS) return DCL_ERR_ARG_WRONG;
  if (g_dev_bytes[slot] < bytes) {
    int rc;
    if (g_dev[slot]) {
      rc = dcl_stream_synchronize(NULL);
      if (rc) return rc;
      rc = dcl_free(g_dev[slot]);
      if (rc) return rc;
      g_dev[slot] = NULL;
      g_dev_bytes[slot] = 0;
    }
    size_t want = bytes + bytes / 8 + 256;
    rc = dcl_malloc(&g_dev[slot], want);
    if (rc) return rc;
    g_dev_bytes[slot] = want;
  }
  *ptr = g_dev[slot];
  return 0;
}

int dl_host(size_t bytes, void **ptr) {
  if (g_host_bytes < bytes) {
    free(g_host);
    size_t want = bytes + bytes / 8 + 256;
    g_host = malloc(want);
    if (!g_host) {
      g_host_bytes = 0;
      return DCL_ERR_MEM;
    }
    g_host_bytes = want;
  }
  *ptr = g_host;
  return 0;
}

int dl_release(void) {
  int rc = dcl_stream_synchronize(NULL);
  for (int i = 0; i < DL_NSLOTS; ++i) {
    if (g_dev[i]) dcl_free(g_dev[i]);
    g_dev[i] = NULL;
    g_dev_bytes[i] = 0;
  }
  free(g_host);
  g_host = NULL;
  g_host_bytes = 0;
  return rc;
}

dcl_window dl_window(Grid g) {
  dcl_window w;
  w.p.x = g->p.x;
  w.p.y = g->p.y;
  w.p.z = g->p.z;
  w.n.x = g->n.x;
  w.n.y = g->n.y;
  w.n.z = g->is2D ? 1 : g->n.z;
  w.d.x = g->d.x;
  w.d.y = g->d.y;
  w.d.z = g->d.z;
  return w;
}

long dl_nodes(Grid g) { return (long)g->n.x * g->n.y * (g->is2D ? 1 : g->n.z); }

int dl_upload_grid(int slot, Grid g, void **dev) {
  const size_t bytes = sizeof(PetscReal) * (size_t)g->SIZE;
  int rc = dl_slot(slot, bytes, dev);
  if (rc) return rc;
  return dcl_memcpy_htod(*dev, g->v1, bytes, NULL);
}

int dl_fail(const char *where, int rc) {
  fprintf(stderr, "[dcell-gpu] %s failed (%d): %s\n", where, rc, dcl_last_error());
  return rc;
}
