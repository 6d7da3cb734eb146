/*
 * lsm_oracle.h -- CPU FP64 restatement of DCell's level-set update hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in dcell_amd/ links or calls it.
 *
 * Conventions (shared with the GPU path, see the boundary policy below):
 *   - A window is the reference's Grid geometry: global offset p, size n,
 *     spacing d (Common/Grid.h:11-13).  Scalars are row-major, x fastest:
 *     v[((z-p.z)*n.y + (y-p.y))*n.x + (x-p.x)]; vector grids append *dof + c
 *     (Common/Grid.c:69-93).  n.z < 2 means 2D (Common/Grid.c:21), then z = p.z.
 *   - Positions (band points, departure points, particles) are in global
 *     grid-index units (LevelSetMethod/AdvectSL.c:74-82).
 *
 * Boundary policy (the reference has none; it reads outside the window):
 *   - velocity sample points are clamped into [vp, vp+vn-2] per axis;
 *   - linear phi samples: coordinate clamped into [p, p+n-1], base index
 *     floor() limited to p+n-2;
 *   - cubic phi samples: coordinate clamped into [p+1, p+n-2], base index
 *     limited to p+n-3;
 *   - dense SL updates nodes in [p+2, p+n-3] per axis and carries the outer
 *     two layers over from phi0;
 *   - curvature is evaluated where its stencil fits in the window, 0 elsewhere;
 *   - PLS corner writes outside the window are skipped.
 * For inputs the reference handles without reading out of bounds every clamp
 * is a no-op, so the restatement equals the reference there.
 */
#ifndef DCELL_LSM_ORACLE_H
#define DCELL_LSM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int32_t x, y, z; } orc_icoor;
typedef struct {
  int32_t px, py, pz;
  int32_t nx, ny, nz;
  double dx, dy, dz;
} orc_win;

enum { ORC_LINEAR = 0, ORC_CUBIC = 1 };
enum { ORC_EULER = 0, ORC_RK2 = 1 };
enum { ORC_CURV_STD = 0, ORC_CURV_LSS = 1 };
enum { ORC_PLS_REF_AZ = 1 };  /* reproduce A.z = floor(X.y) (ParticleLS3D.c:76) */

void orc_set_threads(int n);
int orc_get_threads(void);

/* pointwise interpolants */
void orc_interp_vel3d(const double *vel, const orc_win *vw, int vdof, int udof,
                      const double X[3], double V[3]);
void orc_interp_vel2d(const double *vel, const orc_win *vw, int vdof, int udof,
                      const double X[2], double V[2]);
double orc_bilinear3d(const double *phi, const orc_win *w, const double S[3]);
double orc_bilinear2d(const double *phi, const orc_win *w, const double S[2]);
double orc_gridinterp3d(const double *phi, const orc_win *w, const double X[3]);
double orc_gridinterp2d(const double *phi, const orc_win *w, const double X[2]);
void orc_cr_weights(double t, double w[4]);
void orc_cr_dweights(double t, double dw[4]);
double orc_tricubic(const double *phi, const orc_win *w, const double S[3]);
double orc_bicubic(const double *phi, const orc_win *w, const double S[2]);
/* Catmull-Rom value + gradient on a bare 4x4x4 stencil [k][j][i], local t in [0,1] */
void orc_cr_stencil3(const double *st64, double x, double y, double z, double out[4]);

/* SL step.  band == NULL: dense over the window.  band_out != NULL (band mode):
 * write the band values compactly instead of into phi. */
int orc_sl_advect(const double *phi0, double *phi, const double *vel,
                  const orc_icoor *band, int64_t band_len, double *band_out,
                  const orc_win *w, const orc_win *vw, int vdof, int udof,
                  double dt, int interp, int backtrace);

/* curvature */
double orc_curv3d_at(const double *phi, const orc_win *w, int x, int y, int z);
double orc_curv2d_at(const double *phi, const orc_win *w, int x, int y);
double orc_curv2d_lss_at(const double *phi, const orc_win *w, int x, int y);
int orc_curvature(const double *phi, double *kappa, const orc_icoor *band,
                  int64_t band_len, double *band_out, const orc_win *w, int stencil);

/* particle level set; particles are {x, y, z, radius} (LSM_private.h:45-48) */
int orc_pls_advect(double *particles, int64_t count, const double *vel,
                   const orc_win *vw, int vdof, int udof, const double dh[3],
                   double dt, int dim);
int orc_pls_correct(const double *particles, int64_t count, double *phi,
                    const orc_icoor *band, int64_t band_len, const orc_win *w,
                    int flags, int dim);
int orc_pls_adjust_radii(double *particles, int64_t count, const double *phi,
                         const orc_win *w, double rmin, double rmax, int dim);

/* reductions */
double orc_max_velocity(const double *vel, const orc_win *vw, int vdof, int udof,
                        const orc_icoor *nodes, int64_t count, int64_t stride_bytes,
                        int dim);
int64_t orc_band_compact(const double *phi, const orc_win *w, double width,
                         orc_icoor *out, int64_t capacity);
void orc_interface_metrics(const double *phi, const double *kappa, const orc_win *w,
                           double eps, double out[4]);

/* adjoint of the dense cubic/linear SL step with respect to phi0 and vel.
 * lam_next = dL/dphi (after the step); lam += dL/dphi0; gvel += dL/dvel. */
int orc_sl_adjoint(const double *phi0, const double *vel, const double *lam_next,
                   double *lam, double *gvel, const orc_win *w, const orc_win *vw,
                   int vdof, int udof, double dt, int interp, int backtrace);

/* error norms of a against b: {max|a-b|, sum (a-b)^2, max|b|, n} over mask (NULL = all) */
void orc_error_norms(const double *a, const double *b, const uint8_t *mask,
                     int64_t n, double out[4]);

/* Irregular-node list (LevelSetMethod/IrregularNodes.c:8-140,
 * IrregularNodes3D.c:6-117); irregular_oracle.c.  One record per node the
 * reference appends, in its order; the fields the reference sets, others 0.
 * Returns the number of records (records past capacity are not written). */
typedef struct {
  orc_icoor pos;
  int32_t axis;        /* 0..2, -1 for the orthogonal projection record */
  int32_t shift;       /* -1 projection, 0 cell centre, 1..3 U / V / W face */
  int32_t sign_center;
  int32_t sign_face;
  int32_t pad;
  double d;
  double X[3];
  double op[3];
} orc_irregular_node;

int64_t orc_irregular_nodes(const double *phi, const orc_win *w, const orc_icoor *band, int64_t band_len,
                            orc_irregular_node *out, int64_t capacity);
/* the evaluators of the projections' Newton iterations: p and gradient */
void orc_irr_tricubic(const double *st64, double x, double y, double z, double out[4]);
void orc_irr_bicubic(const double *st16, double x, double y, double out[3]);

/* Surface quantities at irregular nodes (IIM/SurfaceQuantities.c:31-118);
 * surface_oracle.c.  At each record's X: unit normal n, tangents s, r (2D:
 * s = (ny, -nx), r = 0), curvature k, principal curvatures k_nn, k_tt, k_nt
 * (2D: k_nn = k, k_tt = k_nt = 0). */
typedef struct {
  double n[3], s[3], r[3];
  double k, k_nn, k_tt, k_nt;
} orc_surface;

void orc_surface_quantities(const double *phi, const orc_win *w, const orc_irregular_node *nodes, int64_t count,
                            orc_surface *out);

/* Reinitialisation (LevelSetMethod/FastMarching.c:10-146); reinit_oracle.c.
 * phi (in/out): reset to +-phi_inf, then the boundary values at the
 * projection records and the second-order fast marching out to band_width
 * on each side.  band: the new band in the reference's order.  Returns its
 * length (entries past band_cap are not written), or -1 on a NaN update. */
int64_t orc_reinitialize(double *phi, const orc_win *w, const orc_irregular_node *nodes, int64_t count,
                         double band_width, double phi_inf, orc_icoor *band, int64_t band_cap);

/* The parallel reinitialisation's sweeps (reinit_oracle.c): same inputs;
 * the band is [boundary nodes in record order | - pass nodes | + pass nodes],
 * each pass in window order.  *sweeps: the sweeps run, the last one changing
 * nothing.  Returns the band length, -1 on a NaN update, -3 when max_sweeps
 * sweeps did not settle. */
int64_t orc_reinitialize_sweep(double *phi, const orc_win *w, const orc_irregular_node *nodes, int64_t count,
                               double band_width, double phi_inf, orc_icoor *band, int64_t band_cap,
                               int max_sweeps, int *sweeps);

#ifdef __cplusplus
}
#endif
#endif
