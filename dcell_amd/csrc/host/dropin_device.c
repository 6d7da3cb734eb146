/*
 * dropin_device.c -- device buffers and host staging of the C drop-in.
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dropin.h"
#include "dcell_gpu.h"

static unsigned long long g_htod_bytes, g_dtoh_bytes;

int dl_htod(void *dst, const void *src, size_t bytes, void *stream) {
  g_htod_bytes += bytes;
  return dcl_memcpy_htod(dst, src, bytes, stream);
}

int dl_dtoh(void *dst, const void *src, size_t bytes, void *stream) {
  g_dtoh_bytes += bytes;
  return dcl_memcpy_dtoh(dst, src, bytes, stream);
}

PetscErrorCode LevelSetGPUTransferStats(unsigned long long *htod_bytes, unsigned long long *dtoh_bytes) {
  if (htod_bytes) *htod_bytes = g_htod_bytes;
  if (dtoh_bytes) *dtoh_bytes = g_dtoh_bytes;
  return 0;
}

int dl_threads(void) {
  static int n;
  if (!n) {
    const char *e = getenv("DCELL_GPU_HOST_THREADS");
    int v = e ? atoi(e) : 8;
    n = v < 1 ? 1 : v > 64 ? 64 : v;
  }
  return n;
}

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

/* page-locked staging ring of the box uploads: each packed chunk takes the
 * next DL_CHUNK bytes and its copy overlaps the packing of the next chunk;
 * when the ring wraps, the default stream is synchronised first, so no
 * chunk is overwritten while its copy may still run */
static char *g_pin;
static size_t g_pin_off;
enum { DL_CHUNK = 4 << 20, DL_RING = 16 * DL_CHUNK };

static int dl_stage(size_t bytes, char **ptr) {
  if (!g_pin) {
    int rc = dcl_host_malloc((void **)&g_pin, DL_RING);
    if (rc) return rc;
    g_pin_off = 0;
  }
  if (g_pin_off + bytes > DL_RING) {
    int rc = dcl_stream_synchronize(NULL);
    if (rc) return rc;
    g_pin_off = 0;
  }
  *ptr = g_pin + g_pin_off;
  g_pin_off += bytes;
  return 0;
}

static struct {
  int depth;
  Grid grid;
  const void *data;
  dl_box box;
  int valid;
  void *dev;
  dcl_window w;
  double maxabs[3];
  Grid last_grid; /* the grid of the last successful dl_velocity_box */
  int last_m;     /* and the margin it needed */
} g_vel;

int dl_release(void) {
  int rc = dcl_stream_synchronize(NULL);
  if (g_pin) dcl_host_free(g_pin);
  g_pin = NULL;
  g_pin_off = 0;
  g_vel.valid = 0;
  g_vel.last_grid = NULL;
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
  return dl_htod(*dev, g->v1, bytes, NULL);
}

int dl_fail(const char *where, int rc) {
  fprintf(stderr, "[dcell-gpu] %s failed (%d): %s\n", where, rc, dcl_last_error());
  return rc;
}

void dl_bbox(const void *coords, long n, size_t stride, int dim, dl_box *b) {
  int lo[3] = {2147483647, 2147483647, 2147483647}, hi[3] = {-2147483647 - 1, -2147483647 - 1, -2147483647 - 1};
  const char *p = (const char *)coords;
  const int nt = dl_threads();
#pragma omp parallel for num_threads(nt) if (nt > 1 && n > 65536) schedule(static) \
    reduction(min : lo[:3]) reduction(max : hi[:3])
  for (long i = 0; i < n; ++i) {
    const int *q = (const int *)(p + i * stride);
    for (int a = 0; a < dim; ++a) {
      if (q[a] < lo[a]) lo[a] = q[a];
      if (q[a] > hi[a]) hi[a] = q[a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    b->lo[a] = lo[a];
    b->hi[a] = hi[a];
  }
  if (dim == 2) b->lo[2] = b->hi[2] = 0;
}

int dl_box_around(Grid g, const dl_box *b, int m, int extra_hi, dl_box *out) {
  const int p[3] = {g->p.x, g->p.y, g->p.z};
  const int n[3] = {g->n.x, g->n.y, g->is2D ? 1 : g->n.z};
  for (int a = 0; a < 3; ++a) {
    if (a == 2 && g->is2D) {
      out->lo[2] = out->hi[2] = p[2];
      continue;
    }
    long lo = (long)b->lo[a] - m, hi = (long)b->hi[a] + m + extra_hi;
    if (lo < p[a]) lo = p[a];
    if (hi > (long)p[a] + n[a] - 1) hi = (long)p[a] + n[a] - 1;
    if (lo > hi) return 0;
    out->lo[a] = (int)lo;
    out->hi[a] = (int)hi;
  }
  return 1;
}

int dl_upload_box(int slot, Grid g, const dl_box *b, void **dev, dcl_window *w, double *maxabs, int ncomp) {
  const int dof = g->dof;
  const long bx = b->hi[0] - b->lo[0] + 1, by = b->hi[1] - b->lo[1] + 1, bz = b->hi[2] - b->lo[2] + 1;
  const size_t row = sizeof(PetscReal) * (size_t)(bx * dof);
  if (row > DL_CHUNK) return DCL_ERR_ARG_SIZ;
  const size_t total = row * (size_t)(by * bz);
  int rc = dl_slot(slot, total, dev);
  if (rc) return rc;
  double mx[4] = {0, 0, 0, 0};
  int nan = 0;
  if (ncomp > 4) ncomp = 4;
  if (!maxabs) ncomp = 0;
  const long gx = g->n.x, gy = g->n.y;
  const long rows_per_chunk = DL_CHUNK / (long)row;
  const long nrows = by * bz;
  const int nt = dl_threads();
  size_t sent = 0;
  for (long r0 = 0; r0 < nrows; r0 += rows_per_chunk) {
    const long r1 = r0 + rows_per_chunk < nrows ? r0 + rows_per_chunk : nrows;
    char *st;
    rc = dl_stage(row * (size_t)(r1 - r0), &st);
    if (rc) return rc;
#pragma omp parallel num_threads(nt) if (nt > 1 && r1 - r0 > 1)
    {
      double lm[4] = {0, 0, 0, 0};
      int ln = 0;
#pragma omp for schedule(static)
      for (long r = r0; r < r1; ++r) {
        const long y = b->lo[1] + r % by, z = b->lo[2] + r / by;
        const long lz = g->is2D ? 0 : z - g->p.z;
        const PetscReal *src = g->v1 + ((lz * gy + (y - g->p.y)) * gx + (b->lo[0] - g->p.x)) * dof;
        memcpy(st + (r - r0) * row, src, row);
        for (int c = 0; c < ncomp; ++c) {
          double m = lm[c];
          for (long i = 0; i < bx; ++i) {
            const double v = fabs(src[i * dof + c]);
            ln |= v != v;
            m = v > m ? v : m;
          }
          lm[c] = m;
        }
      }
#pragma omp critical(dl_maxabs)
      {
        for (int c = 0; c < ncomp; ++c)
          if (lm[c] > mx[c]) mx[c] = lm[c];
        nan |= ln;
      }
    }
    rc = dl_htod((char *)*dev + sent, st, row * (size_t)(r1 - r0), NULL);
    if (rc) return rc;
    sent += row * (size_t)(r1 - r0);
  }
  for (int c = 0; c < ncomp; ++c) maxabs[c] = nan ? NAN : mx[c];
  w->p.x = b->lo[0];
  w->p.y = b->lo[1];
  w->p.z = b->lo[2];
  w->n.x = (int)bx;
  w->n.y = (int)by;
  w->n.z = g->is2D ? 1 : (int)bz;
  w->d.x = g->d.x;
  w->d.y = g->d.y;
  w->d.z = g->d.z;
  return 0;
}

void dl_scope_begin(void) {
  if (g_vel.depth++ == 0) g_vel.valid = 0;
}

void dl_scope_end(void) {
  if (--g_vel.depth <= 0) {
    g_vel.depth = 0;
    g_vel.valid = 0;
  }
}

void dl_velocity_invalidate(void) { g_vel.valid = 0; }

static int box_inside(const dl_box *a, const dl_box *b) {
  for (int k = 0; k < 3; ++k)
    if (a->lo[k] < b->lo[k] || a->hi[k] > b->hi[k]) return 0;
  return 1;
}

int dl_velocity_box(Grid velgrid, const dl_box *b, double dt, Coor dh, int dim, void **dev, dcl_window *w,
                    int *M) {
  const double d[3] = {dh.x, dh.y, dh.z};
  int m = 4;
  if (g_vel.last_grid == velgrid && g_vel.last_m > m) m = g_vel.last_m;
  for (int iter = 0; iter < 16; ++iter) {
    dl_box vb;
    if (!dl_box_around(velgrid, b, m, 1, &vb)) return DCL_ERR_ARG_SIZ;
    if (!(g_vel.valid && g_vel.grid == velgrid && g_vel.data == velgrid->v1 && box_inside(&vb, &g_vel.box))) {
      double mx[3] = {0, 0, 0};
      int rc = dl_upload_box(DL_VEL, velgrid, &vb, &g_vel.dev, &g_vel.w, mx, dim);
      if (rc) return rc;
      g_vel.valid = g_vel.depth > 0;
      g_vel.grid = velgrid;
      g_vel.data = velgrid->v1;
      g_vel.box = vb;
      memcpy(g_vel.maxabs, mx, sizeof mx);
    }
    /* the uploaded box holds b grown by m; the margin must hold every
     * departure: |dt V / dh| <= D along each axis (V interpolated from values
     * in the box), plus the widest stencil (cubic base - 1 .. base + 2) and
     * the MAC corner above */
    double D = 0;
    for (int a = 0; a < dim; ++a) {
      const double x = dt * g_vel.maxabs[a] / d[a];
      if (!(x <= 1e6)) return DCL_ERR_ARG_WRONG; /* NaN or absurd velocities */
      if (x > D) D = x;
    }
    const int need = (int)ceil(D) + 3;
    if (need <= m) {
      *dev = g_vel.dev;
      *w = g_vel.w;
      *M = m;
      g_vel.last_grid = velgrid;
      g_vel.last_m = need;
      return 0;
    }
    m = need;
    g_vel.valid = 0;
  }
  return DCL_ERR_ARG_SIZ;
}
