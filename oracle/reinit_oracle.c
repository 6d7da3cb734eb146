/*
 * reinit_oracle.c -- CPU FP64 restatement of the reinitialisation,
 * LevelSetReinitialize (LevelSetMethod/FastMarching.c:10-41) with
 * FMM_InitializeEikonal (:43-95), FMM_SolveEikonal (:97-146), the
 * second-order neighbour updates FMM_PushNeighbors3D
 * (LevelSetMethod/FastMarching3D.c:14-101) / FMM_PushNeighbors2D
 * (FastMarching2D.c:98-185) with AccumulateCoef (FastMarching2D.c:33-94),
 * and the binary min-heap of Common/Heap.c:59-139.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Nothing in
 * dcell_amd/ links or calls this file.
 *
 * Every step keeps the reference's order: the heap's sift-up and sift-down
 * with its comparisons (so equal keys pop in the same order), the neighbour
 * order of STAR, the axis order x, y, z, and each expression's operand order
 * (built with -ffp-contract=off).  Departures from the reference, all where
 * it reads out of bounds or leaves a value unset: a neighbour outside the
 * window is not pushed and a record outside it is skipped (the reference
 * reads them unchecked; LevelSetResize keeps the band away from the edge,
 * LevelSetMethod/Resize.c:13-14), a pushed 2D node's z is the window's, and
 * a NaN update is reported (return -1) instead of exiting the process.
 *
 * The sign passes: the - pass only writes nodes whose phi is -PHI_INF and
 * only reads non-negative values of -phi, and the + pass the mirror image,
 * so each pass is unaffected by the other; the band is [boundary nodes in
 * list order | - pass in acceptance order | + pass in acceptance order].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lsm_oracle.h"

typedef struct {
  double v;
  int32_t x, y, z;
} hnode;

typedef struct {
  hnode *t; /* 1-based */
  int64_t len, cap;
} heap;

/* HeapInsert (Common/Heap.c:59-91) */
static int heap_insert(heap *h, hnode item) {
  if (h->len + 1 > h->cap) {
    int64_t cap = h->cap ? 2 * h->cap : 1024;
    hnode *t = realloc(h->t, sizeof(hnode) * (size_t)cap);
    if (!t) return -1;
    h->t = t;
    h->cap = cap;
  }
  int64_t c = h->len;
  h->len = c + 1;
  h->t[c] = item;
  while (c != 1) {
    const int64_t p = c / 2;
    const hnode parent = h->t[p];
    if (parent.v - item.v < 0) break;
    h->t[c] = parent;
    h->t[p] = item;
    c = p;
  }
  return 0;
}

/* HeapPop (Common/Heap.c:94-139); h->len counts the unused slot 0 */
static hnode heap_pop(heap *h) {
  const hnode top = h->t[1];
  int64_t len = h->len;
  if (len == 2) {
    h->len = 1;
    return top;
  }
  int64_t p = 1;
  len--;
  const hnode bubble = h->t[p] = h->t[len];
  h->len = len;
  int64_t l = 2 * p;
  while (l < len) {
    const int64_t r = 2 * p + 1;
    int64_t s = l;
    if (r < len && h->t[r].v - h->t[l].v < 0) s = r;
    if (h->t[p].v - h->t[s].v <= 0) break;
    h->t[p] = h->t[s];
    h->t[s] = bubble;
    p = s;
    l = 2 * p;
  }
  return top;
}

/* AccumulateCoef (FastMarching2D.c:33-94): max(D+, D-, 0)^2 of one axis
 * from the 7 values p[-3..3] around the node */
static void accumulate(double bw, const double *phi7, double *a, double *b, double *c) {
  const double *p = &phi7[3];
  const double D2p = (3 * p[0] - 4 * p[1] + p[2]) / 2.;
  const double D2m = (3 * p[0] - 4 * p[-1] + p[-2]) / 2.;
  const double D1p = p[0] - p[1];
  const double D1m = p[0] - p[-1];
  double Dm, am, bm, cm, Dp, ap, bp, cp, t;
  if (0 <= p[1] && p[1] < bw && 0 <= p[2] && p[2] < bw && D2p > 0) {
    t = (4 * p[1] - p[2]) / 3;
    ap = 9. / 4.;
    bp = -9. / 2. * t;
    cp = 9. / 4. * t * t;
    Dp = D2p;
  } else if (0 <= p[1] && p[1] < bw && D1p > 0) {
    ap = 1.;
    bp = -2. * p[1];
    cp = p[1] * p[1];
    Dp = D1p;
  } else {
    ap = 0;
    bp = 0;
    cp = 0;
    Dp = 0;
  }
  if (0 <= p[-1] && p[-1] < bw && 0 <= p[-2] && p[-2] < bw && D2m > 0) {
    t = (4 * p[-1] - p[-2]) / 3;
    am = 9. / 4.;
    bm = -9. / 2. * t;
    cm = 9. / 4. * t * t;
    Dm = D2m;
  } else if (0 <= p[-1] && p[-1] < bw && D1m > 0) {
    am = 1;
    bm = -2. * p[-1];
    cm = p[-1] * p[-1];
    Dm = D1m;
  } else {
    am = 0;
    bm = 0;
    cm = 0;
    Dm = 0;
  }
  if (Dm > Dp) {
    *a += am;
    *b += bm;
    *c += cm;
  } else {
    *a += ap;
    *b += bp;
    *c += cp;
  }
}

typedef struct {
  double *phi;
  const orc_win *w;
  int d3;
  double band_width, phi_inf;
} fmm;

static inline int inside(const fmm *f, int x, int y, int z) {
  const orc_win *w = f->w;
  return x >= w->px && x < w->px + w->nx && y >= w->py && y < w->py + w->ny &&
         (!f->d3 || (z >= w->pz && z < w->pz + w->nz));
}

static inline double *at(const fmm *f, int x, int y, int z) {
  const orc_win *w = f->w;
  const int64_t lz = f->d3 ? z - w->pz : 0;
  return &f->phi[(lz * w->ny + (y - w->py)) * w->nx + (x - w->px)];
}

/* STAR (Common/Common.h:49-51) */
static const int STAR[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

/* FMM_PushNeighbors3D / 2D: one update per neighbour on the wavefront */
static int push_neighbors(const fmm *f, heap *h, int sign, int X, int Y, int Z) {
  const int nn = f->d3 ? 6 : 4;
  for (int n = 0; n < nn; ++n) {
    /* in 2D the reference leaves a pushed node's pos.z unset
     * (FastMarching2D.c:177-180); here it is the window's z */
    const int i = X + STAR[n][0], j = Y + STAR[n][1], k = f->d3 ? Z + STAR[n][2] : f->w->pz;
    if (!inside(f, i, j, k)) continue;
    if (sign * *at(f, i, j, k) < f->band_width) continue;
    double a = 0, b = 0, c = -1, loc[7];
    for (int axis = 0; axis < (f->d3 ? 3 : 2); ++axis) {
      for (int m = 0; m < 7; ++m) {
        const int s = m - 7 / 2;
        const int ii = i + (axis == 0 ? s : 0), jj = j + (axis == 1 ? s : 0), kk = k + (axis == 2 ? s : 0);
        loc[m] = inside(f, ii, jj, kk) ? sign * *at(f, ii, jj, kk) : f->phi_inf;
      }
      accumulate(f->phi_inf, loc, &a, &b, &c);
    }
    const double disc = sqrt(b * b - 4 * a * c);
    double d1 = (-b + disc) / (2 * a);
    if (d1 != d1) d1 = -b / (2 * a);
    if (d1 != d1) return -1;
    const hnode item = {d1, i, j, k};
    if (heap_insert(h, item)) return -2;
  }
  return 0;
}

int64_t orc_reinitialize(double *phi, const orc_win *w, const orc_irregular_node *nodes, int64_t count,
                         double band_width, double phi_inf, orc_icoor *band, int64_t band_cap) {
  fmm f = {phi, w, w->nz >= 2, band_width, phi_inf};
  const int64_t N = (int64_t)w->nx * w->ny * (f.d3 ? w->nz : 1);
  for (int64_t i = 0; i < N; ++i) phi[i] = phi[i] > 0 ? phi_inf : -phi_inf;
  int64_t nb = 0;
#define APPEND(X, Y, Z)                   \
  do {                                    \
    if (nb < band_cap) {                  \
      band[nb].x = (X);                   \
      band[nb].y = (Y);                   \
      band[nb].z = (Z);                   \
    }                                     \
    ++nb;                                 \
  } while (0)
  /* the boundary condition: the distance to the orthogonal projection */
  for (int64_t i = 0; i < count; ++i) {
    const orc_irregular_node *n = &nodes[i];
    if (n->axis != -1 || !inside(&f, n->pos.x, n->pos.y, n->pos.z)) continue;
    APPEND(n->pos.x, n->pos.y, n->pos.z);
    const double r = f.d3 ? sqrt(n->op[0] * n->op[0] + n->op[1] * n->op[1] + n->op[2] * n->op[2])
                          : sqrt(n->op[0] * n->op[0] + n->op[1] * n->op[1]);
    *at(&f, n->pos.x, n->pos.y, n->pos.z) = n->sign_center * r;
  }
  heap h = {NULL, 1, 0};
  int rc = 0;
  for (int sign = -1; sign <= 1 && rc == 0; sign += 2) {
    for (int64_t i = 0; i < count && rc == 0; ++i) {
      const orc_irregular_node *n = &nodes[i];
      if (n->axis != -1 || !inside(&f, n->pos.x, n->pos.y, n->pos.z)) continue;
      if (sign * *at(&f, n->pos.x, n->pos.y, n->pos.z) >= 0.)
        rc = push_neighbors(&f, &h, sign, n->pos.x, n->pos.y, n->pos.z);
    }
    /* FMM_SolveEikonal */
    while (rc == 0 && h.len > 1) {
      const hnode node = heap_pop(&h);
      if (node.v > band_width) break;
      double *q = at(&f, node.x, node.y, node.z);
      if (node.v >= sign * *q) continue;
      APPEND(node.x, node.y, node.z);
      *q = sign * node.v;
      rc = push_neighbors(&f, &h, sign, node.x, node.y, node.z);
    }
    h.len = 1; /* HeapClear */
  }
#undef APPEND
  free(h.t);
  return rc ? rc : nb;
}

/* ------------------------------------------------------------------------
 * The parallel reinitialisation (dcl_reinitialize_parallel), restated: the
 * checker of the GPU sweeps, itself checked against orc_reinitialize above.
 *
 * Each node of a sign pass gets, from its neighbours' current values, the
 * value the fast marching above would give it if its neighbours were
 * accepted in the order of their values: the pushes it receives are replayed
 * in that order (the boundary nodes' pushes first, then one push per
 * neighbour value), each with the neighbours accepted by then, and the node
 * keeps the least update made before its own value comes up.  Sweeps repeat
 * this at every node from the previous sweep's values (Jacobi) until a sweep
 * changes nothing.  Where the fast marching accepts nodes in value order
 * (no ties decided by the heap's layout, no update below an accepted
 * neighbour) the result is its result bit for bit.
 *
 * mask: 0 a - pass node, 1 a + pass node, 2 a boundary node.
 * ------------------------------------------------------------------------ */

/* test hook, so the tests can show their pins catch a plausible mistake:
 * 1 counts every usable neighbour as accepted before the node whatever its
 * value; 2 replays pushes past the node's own value; 3 keeps the last push
 * instead of the least; 0 in every other use */
int orc_sweep_plant = 0;

/* one neighbour as pass `sign` sees it: its value q (sign * phi) and the
 * time it is accepted (-1: before the pass's pops; its value for a node of
 * this pass; INFINITY: never) */
typedef struct {
  double q, t;
  int pushes; /* accepted in this pass, so it pushes its neighbours */
} sw_nb;

static sw_nb sw_view(const fmm *f, const double *in, const uint8_t *mask, int sign, int x, int y, int z) {
  sw_nb r = {f->phi_inf, INFINITY, 0};
  if (!inside(f, x, y, z)) return r;
  const int64_t i = at(f, x, y, z) - f->phi;
  const double q = sign * in[i];
  const int m = mask[i];
  r.q = q;
  if (m == 2) {
    /* a boundary node: pushes in this pass when sign * phi >= 0 */
    if (0 <= q && q < f->phi_inf) {
      r.t = -1;
      r.pushes = 1;
    }
  } else if (m == (sign > 0)) {
    if (q < f->phi_inf) {
      r.t = q;
      r.pushes = 1;
    }
  } else if (sign > 0 && 0 <= q && q < f->phi_inf) {
    r.t = -1; /* a - pass node, final when the + pass starts */
  }
  if (orc_sweep_plant == 1 && r.t < INFINITY) r.t = -1;
  return r;
}

/* returns the node's new value, or NAN for a NaN update */
#define SW_TOL 1e-9

static double sw_node(const fmm *f, const double *in, const uint8_t *mask, int sign, int X, int Y, int Z) {
  const int na = f->d3 ? 3 : 2;
  sw_nb nb[3][4]; /* per axis: -2, -1, +1, +2 */
  for (int a = 0; a < na; ++a)
    for (int k = 0; k < 4; ++k) {
      const int s = k < 2 ? k - 2 : k - 1;
      nb[a][k] = sw_view(f, in, mask, sign, X + (a == 0 ? s : 0), Y + (a == 1 ? s : 0), Z + (a == 2 ? s : 0));
    }
  double cand = INFINITY, last = -INFINITY;
  for (;;) {
    /* the next push: the least acceptance time above `last` among the
     * neighbours that push */
    double T = INFINITY;
    for (int a = 0; a < na; ++a)
      for (int k = 1; k <= 2; ++k)
        if (nb[a][k].pushes && nb[a][k].t > last && nb[a][k].t < T) T = nb[a][k].t;
    if (T == INFINITY || (!(T < cand) && orc_sweep_plant != 2)) break;
    double A = 0, B = 0, Cc = -1, loc[7];
    for (int a = 0; a < na; ++a) {
      loc[0] = loc[6] = loc[3] = f->phi_inf;
      loc[1] = nb[a][0].t <= T ? nb[a][0].q : f->phi_inf;
      loc[2] = nb[a][1].t <= T ? nb[a][1].q : f->phi_inf;
      loc[4] = nb[a][2].t <= T ? nb[a][2].q : f->phi_inf;
      loc[5] = nb[a][3].t <= T ? nb[a][3].q : f->phi_inf;
      accumulate(f->phi_inf, loc, &A, &B, &Cc);
    }
    const double disc = sqrt(B * B - 4 * A * Cc);
    double d1 = (-B + disc) / (2 * A);
    if (d1 != d1) d1 = -B / (2 * A);
    if (d1 != d1) return NAN;
    if (d1 < cand || orc_sweep_plant == 3) cand = d1;
    last = T;
  }
  return cand <= f->band_width ? sign * cand : sign * f->phi_inf;
}

int64_t orc_reinitialize_sweep(double *phi, const orc_win *w, const orc_irregular_node *nodes, int64_t count,
                               double band_width, double phi_inf, orc_icoor *band, int64_t band_cap,
                               int max_sweeps, int *sweeps) {
  fmm f = {phi, w, w->nz >= 2, band_width, phi_inf};
  const int64_t N = (int64_t)w->nx * w->ny * (f.d3 ? w->nz : 1);
  uint8_t *mask = malloc((size_t)N);
  double *other = malloc(sizeof(double) * (size_t)N);
  if (!mask || !other) {
    free(mask);
    free(other);
    return -2;
  }
  for (int64_t i = 0; i < N; ++i) {
    mask[i] = phi[i] > 0;
    phi[i] = phi[i] > 0 ? phi_inf : -phi_inf;
  }
  int64_t nb = 0;
  for (int64_t i = 0; i < count; ++i) {
    const orc_irregular_node *n = &nodes[i];
    if (n->axis != -1 || !inside(&f, n->pos.x, n->pos.y, n->pos.z)) continue;
    if (nb < band_cap) {
      band[nb].x = n->pos.x;
      band[nb].y = n->pos.y;
      band[nb].z = n->pos.z;
    }
    ++nb;
    const double r = f.d3 ? sqrt(n->op[0] * n->op[0] + n->op[1] * n->op[1] + n->op[2] * n->op[2])
                          : sqrt(n->op[0] * n->op[0] + n->op[1] * n->op[1]);
    double *q = at(&f, n->pos.x, n->pos.y, n->pos.z);
    *q = n->sign_center * r;
    mask[q - phi] = 2;
  }
  /* settled: no node moves by more than tol in a sweep (nodes whose order
   * is decided by their last bits can trade rounding errors forever) */
  const double tol = band_width * SW_TOL;
  double *in = phi, *out = other;
  int it = 0, rc = 0, settle = 0, falling = 0;
  int64_t changed = 1, last = -1;
  while (changed && rc == 0) {
    if (it >= max_sweeps) {
      rc = -3;
      break;
    }
    changed = 0;
#pragma omp parallel for schedule(static) reduction(+ : changed) reduction(| : rc)
    for (int64_t i = 0; i < N; ++i) {
      double v = in[i];
      if (mask[i] != 2) {
        const int X = w->px + (int)(i % w->nx), Y = w->py + (int)((i / w->nx) % w->ny);
        const int Z = f.d3 ? w->pz + (int)(i / ((int64_t)w->nx * w->ny)) : w->pz;
        fmm g = f;
        g.phi = in;
        const int sg = mask[i] ? 1 : -1;
        const double u = sw_node(&g, in, mask, sg, X, Y, Z);
        if (u != u) rc |= 1;
        if (!settle || sg * u < sg * v) v = u;
      }
      changed += fabs(v - in[i]) > tol;
      out[i] = v;
    }
    if (rc) rc = -1;
    /* a sweep that moves no fewer nodes than the one before, once the count
     * has started to fall, leaves only nodes trading places in a cycle: from
     * then on a node's value only goes down, which ends every cycle */
    if (last >= 0 && changed < last) falling = 1;
    else if (last >= 0 && falling) settle = 1;
    last = changed;
    double *t = in;
    in = out;
    out = t;
    ++it;
  }
  if (in != phi) memcpy(phi, in, sizeof(double) * (size_t)N);
  *sweeps = it;
  if (rc == 0) {
    for (int p = 0; p < 2; ++p)
      for (int64_t i = 0; i < N; ++i)
        if (mask[i] == p && fabs(phi[i]) < phi_inf) {
          if (nb < band_cap) {
            band[nb].x = w->px + (int)(i % w->nx);
            band[nb].y = w->py + (int)((i / w->nx) % w->ny);
            band[nb].z = f.d3 ? w->pz + (int)(i / ((int64_t)w->nx * w->ny)) : w->pz;
          }
          ++nb;
        }
  }
  free(mask);
  free(other);
  return rc ? rc : nb;
}
