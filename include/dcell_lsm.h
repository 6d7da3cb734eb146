/*
 * dcell_lsm.h -- C ABI of the MI355X (gfx950) level-set update of DCell.
 *
 * Plain C: raw device pointers, sizes and a stream handle (a hipStream_t
 * passed as void*, NULL = the default stream).  Every entry point returns 0 on
 * success or one of the DCL_ERR_* codes below, which take PETSc's numeric
 * values so a DCell caller can CHKERRQ() them; dcl_last_error() gives the
 * message of the last failure on the calling thread.  Entry points only
 * enqueue work on the stream unless they document a host result.
 *
 * Layout (the reference's own, Common/Grid.c:69-93): a window is a dense box
 * of nodes with global offset p and size n; scalars are row-major, x fastest,
 * v[((z-p.z)*n.y + (y-p.y))*n.x + (x-p.x)]; vector grids are array-of-structs
 * [z][y][x][dof].  n.z < 2 means 2D (Common/Grid.c:21).  Positions (band
 * nodes, departure points, particles) are global grid-index coordinates
 * (LevelSetMethod/AdvectSL.c:74-82).  Velocity is MAC-staggered: component c
 * of node i sits half a cell below i along axis c (Common/Common.h:26-30,
 * Common/InterpolateVelocity.c:45-50).
 *
 * Precision: DCL_F64 stores and computes in double.  DCL_F32 stores phi,
 * velocity and curvature in float and computes weights and sums in float;
 * departure coordinates and their index/fraction split stay in double.
 *
 * Boundary policy (the reference reads outside the window where a departure
 * point leaves it; these clamps are no-ops wherever it does not):
 *   velocity sample points are clamped into [vp, vp+vn-2] per axis;
 *   linear phi samples into [p, p+n-1] with base index <= p+n-2;
 *   cubic phi samples into [p+1, p+n-2] with base index <= p+n-3;
 *   dense SL updates nodes in [p+2, p+n-3] and carries the two outer layers
 *   over from phi0; curvature is 0 where its stencil leaves the window;
 *   particle-correction corner writes outside the window are dropped.
 *
 * Calls that share device buffers must be ordered by the caller (same
 * stream).  The library keeps no global state besides the per-thread error
 * message and the objects created by the *_create functions.
 */
#ifndef DCELL_LSM_H
#define DCELL_LSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DCL_ABI_VERSION 1

/* return codes (PETSc's values for the same conditions) */
#define DCL_SUCCESS 0
#define DCL_ERR_MEM 55       /* device or host allocation failed */
#define DCL_ERR_SUP 56       /* unsupported combination of options */
#define DCL_ERR_ARG_SIZ 60   /* a size or count is out of range */
#define DCL_ERR_ARG_WRONG 62 /* an option value is invalid */
#define DCL_ERR_LIB 76       /* the HIP runtime reported an error */
#define DCL_ERR_ARG_NULL 85  /* a required pointer is NULL */

/* iCoor (Common/Common.h:32-34) */
typedef struct { int32_t x, y, z; } dcl_icoor;
/* Coor (Common/Common.h:28-30) */
typedef struct { double x, y, z; } dcl_coor;
/* the geometry fields of struct _Grid (Common/Grid.h:11-13) */
typedef struct {
  dcl_icoor p; /* global index of the first node */
  dcl_icoor n; /* nodes per axis; n.z < 2 is a 2D window */
  dcl_coor d;  /* spacing in units of length */
} dcl_window;

/* struct _Particle (LevelSetMethod/LSM_private.h:45-48): position in global
 * index units and a signed radius in grid units; 32 bytes */
typedef struct { double x, y, z, radius; } dcl_particle;
/* the DCL_F32 form of a particle; 16 bytes */
typedef struct { float x, y, z, radius; } dcl_particle_f32;

/* a velocity grid: the reference's velgrid (LevelSetMethod/Advect.c:12-22).
 * dof 1 stores the components in planes instead, [c][z][y][x] (u, v, w, each
 * plane the window's node count), the layout of a solver that keeps its MAC
 * components apart; dcl_sl_advect and dcl_levelset_step take it (an x pair
 * of one component's faces is then one load in the gathers), the other entry
 * points return DCL_ERR_SUP for it. */
typedef struct {
  const void *data; /* device, [z][y][x][dof] (or [c][z][y][x] for dof 1) of the precision in use */
  dcl_window win;
  int32_t dof;      /* components stored per node: 2 (2D), 3 (3D), 4 fluid [p,u,v,w]; 1: planes */
  int32_t udof;     /* index of u within a node: 0 for a velgrid, 1 for the fluid (U_FACE) */
} dcl_velocity;

enum { DCL_F64 = 0, DCL_F32 = 1 };
enum { DCL_LINEAR = 0, DCL_CUBIC = 1 };
enum { DCL_EULER = 0, DCL_RK2 = 1 };
enum { DCL_CURV_STD = 0, DCL_CURV_LSS = 1 };
/* reproduce the reference's corner base A.z = floor(X.y) (ParticleLS3D.c:76) */
enum { DCL_PLS_REF_AZ = 1 };

/* ---------------------------------------------------------------- runtime */
int dcl_abi_version(void);
const char *dcl_last_error(void);
int dcl_device_count(int *count);
int dcl_malloc(void **ptr, size_t bytes);
int dcl_free(void *ptr);
int dcl_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int dcl_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int dcl_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int dcl_stream_synchronize(void *stream);
/* device-to-device copy by a streaming kernel (16-byte non-temporal loads
 * and stores, several in flight per thread); 16-byte aligned buffers and
 * size.  The benchmark's measured HBM ceiling.  Asynchronous. */
int dcl_stream_copy(void *dst, const void *src, size_t bytes, void *stream);
/* page-locked host memory (staging of asynchronous copies) */
int dcl_host_malloc(void **ptr, size_t bytes);
int dcl_host_free(void *ptr);

/* ------------------------------------------------------------- SL advect */
/* One semi-Lagrangian step phi0 -> phi.
 * Replaces the loop of LevelSetAdvectSL_3D / _2D (LevelSetMethod/AdvectSL.c:23-89),
 * reached from LevelSetAdvectSL (AdvectSL.c:6-19).
 *   band == NULL: dense; every node of the window is written (interior nodes
 *     advected, the two outer layers copied from phi0).
 *   band != NULL: the band_len nodes of band (global coordinates, the
 *     reference's ls->band, LevelSetMethod.h:15) are advected; their values go
 *     to phi at those nodes, or, when band_values != NULL, to band_values[b]
 *     (phi is then not touched and may be NULL).
 *   interp DCL_LINEAR is the reference's Bilinear3D / Bilinear2D gather
 *     (Common/Grid3D.c:4-31, Common/Grid.c:190-212); DCL_CUBIC is the
 *     Catmull-Rom tricubic / bicubic of LevelSetMethod/TricubicInverse.h +
 *     TricubicInterpolation.h (OrthogonalProjection.c:48-79 in 2D).
 *   backtrace DCL_EULER is the reference's S = X - V dt/dh (AdvectSL.c:80-82);
 *     DCL_RK2 is the particle midpoint rule (ParticleLS3D.c:25-32) run
 *     backwards: Xm = X - dt/2 V(X)/dh, S = X - dt V(Xm)/dh.
 * vel->win may differ from *win (sample points are clamped into vel->win). */
int dcl_sl_advect(const void *phi0, void *phi, const dcl_window *win,
                  const dcl_velocity *vel, const dcl_icoor *band, int64_t band_len,
                  void *band_values, double dt, int precision, int interp,
                  int backtrace, void *stream);

/* ------------------------------------------------------------ curvature */
/* Curvature of phi at every node (band == NULL) or at band nodes.
 * DCL_CURV_STD: GridFunction3D_Curv, 19 points (Common/Grid3D.c:83-108) in 3D;
 *   GridFunction2D_Curv, 9 points (Common/Grid.c:231-254) in 2D; NaN -> 0.
 * DCL_CURV_LSS (2D only): GridFunction2D_CurvLSS, the 9x9 least-squares
 *   stencil of LevelSetMethod/Curvature.c:3-186, whose pxx and pyy are half the
 *   second derivatives, as in the reference; no NaN guard.
 * kappa is in 1/length: phi's units cancel, dh is physical. */
int dcl_curvature(const void *phi, void *kappa, const dcl_window *win,
                  const dcl_icoor *band, int64_t band_len, void *band_values,
                  int precision, int stencil, void *stream);

/* One dense level-set update: SL step phi0 -> phi (as dcl_sl_advect, dense)
 * and the curvature of the new phi into kappa (as dcl_curvature,
 * DCL_CURV_STD), fused where the shapes allow.  This is the unit measured as
 * one grid-point update.  No counterpart in the reference, which samples
 * curvature pointwise at interface nodes (IIM/SurfaceQuantities.c:111). */
int dcl_levelset_step(const void *phi0, void *phi, void *kappa, const dcl_window *win,
                      const dcl_velocity *vel, double dt, int precision, int interp,
                      int backtrace, void *stream);

/* ------------------------------------------------------ particle level set */
/* RK2 midpoint particle advection, ParticleLS_AdvectParticles3D / 2D
 * (LevelSetMethod/ParticleLS3D.c:6-36, ParticleLS2D.c:6-36).  particles is a
 * device array of dcl_particle (DCL_F64) or dcl_particle_f32 (DCL_F32); dh is
 * ls->phi->d; dim is 2 or 3. */
int dcl_pls_advect(void *particles, int64_t count, const dcl_velocity *vel,
                   dcl_coor dh, double dt, int dim, int precision, void *stream);

/* Workspace of the escaped-particle correction: the phi+ / phi- extremum
 * grids (kept at their neutral value between calls) and the list of touched
 * nodes.  max_nodes = window nodes; max_events bounds 2^dim * escaped
 * particles per call (the call fails with DCL_ERR_ARG_SIZ beyond it, leaving
 * phi untouched). */
typedef struct dcl_pls_ws_s *dcl_pls_ws;
int dcl_pls_ws_create(int64_t max_nodes, int64_t max_events, int precision, void *stream,
                      dcl_pls_ws *ws);
int dcl_pls_ws_destroy(dcl_pls_ws ws);

/* Escaped-particle correction, ParticleLS_ErrorCorrection3D / 2D
 * (LevelSetMethod/ParticleLS3D.c:40-116, ParticleLS2D.c:40-115): particles on
 * the wrong side of phi by more than their radius raise phi+ (or lower phi-)
 * at the corners of their cell to s(s r - |X - a|); then phi = phi+ where
 * |phi+| <= |phi-| else phi-, over the band (band != NULL) or over every node
 * (band == NULL).  With band_values != NULL the merged band values are written
 * there and phi is left as it was.  flags: DCL_PLS_REF_AZ. */
int dcl_pls_correct(const void *particles, int64_t count, void *phi,
                    const dcl_window *win, const dcl_icoor *band, int64_t band_len,
                    void *band_values, int flags, int dim, int precision,
                    dcl_pls_ws ws, void *stream);

/* ParticleLS_AdjustRadii (LevelSetMethod/ParticleLS.c:158-185):
 * r = s * clamp(s phi(X), rmin, rmax) with the reference's strict bounds. */
int dcl_pls_adjust_radii(void *particles, int64_t count, const void *phi,
                         const dcl_window *win, double rmin, double rmax, int dim,
                         int precision, void *stream);

/* ------------------------------------------------------------ reductions */
/* max |V| over nodes, LevelSet_MaxVelocity (LevelSetMethod/Advect.c:70-111).
 * nodes: device array read with a byte stride (12 for dcl_icoor; the size of
 * IrregularNode to pass ls->irregularNodes data as it is).  Synchronous: the
 * result lands in *max_out (host). */
int dcl_max_velocity(const dcl_velocity *vel, const void *nodes, int64_t count,
                     int64_t stride_bytes, int dim, int precision, double *max_out,
                     void *stream);

/* Nodes with |phi| < width in raster order (z, y, x), written to band_out
 * (device, capacity entries).  Synchronous: *count_out (host) is the total
 * number found, which may exceed capacity (then only capacity are written).
 * Replaces the band list the reference rebuilds in the FMM
 * (LevelSetMethod/FastMarching.c:33,126-129) for GPU-resident use. */
int dcl_band_compact(const void *phi, const dcl_window *win, double width,
                     int precision, dcl_icoor *band_out, int64_t capacity,
                     int64_t *count_out, void *stream);

/* dcl_band_compact without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and synchronisation as dcl_band_compact; its
 * per-block counts and scan sums live in a static device array of
 * 4 MiB + 64 KiB per GPU (held until the process ends), so windows of up to
 * 2^28 nodes are accepted and larger ones fail with DCL_ERR_ARG_SIZ.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * host threads, other streams) race and must be serialised by the caller. */
int dcl_band_compact_static(const void *phi, const dcl_window *win, double width,
                            int precision, dcl_icoor *band_out, int64_t capacity,
                            int64_t *count_out, void *stream);

/* One record of the irregular-node list: the fields that
 * LevelSetUpdateIrregularNodeList sets (LevelSetMethod/IrregularNodes3D.c:
 * 53-113, IrregularNodes.c:69-137); the reference's IrregularNode
 * (LevelSetMethod.h:30-54) carries them among fields filled later.  shift -1
 * and axis -1: the orthogonal projection of the node (op, X = pos + op);
 * shift 0: a crossing towards -x/-y/-z (axis 0..2) at fraction d; shift 1..3:
 * a crossing of the U/V/W face averages along axis.  88 bytes. */
typedef struct {
  dcl_icoor pos;
  int32_t axis, shift, sign_center, sign_face, pad;
  double d;
  dcl_coor X;  /* global index units */
  dcl_coor op;
} dcl_irregular_node;

/* LevelSetUpdateIrregularNodeList (LevelSetMethod/IrregularNodes.c:8-28): the
 * records of every band node in the reference's append order, with the
 * orthogonal projections of OrthogonalProjection3D.c:6-208 /
 * OrthogonalProjection.c:6-184 (Newton on the tricubic / bicubic of the
 * cells around the node, quadratic and linear fallbacks) computed in double.
 * phi: device window (DCL_F64 or DCL_F32, read as double); band: device
 * nodes; out: device, capacity records.  Synchronous: *count_out (host) is
 * the total, which may exceed capacity (then only capacity are written).
 * Stencil reads clamp into the window.  band_len < 2^31. */
int dcl_irregular_nodes(const void *phi, const dcl_window *win, const dcl_icoor *band,
                        int64_t band_len, int precision, dcl_irregular_node *out,
                        int64_t capacity, int64_t *count_out, void *stream);

/* One level set of a reinitialisation batch (see dcl_reinitialize). */
typedef struct {
  void *phi;                       /* device FP64 window, rewritten in place */
  dcl_window win;
  const dcl_irregular_node *nodes; /* device: phi's irregular-node records */
  int64_t count;
  double band_width, phi_inf;      /* grid units; the reference's are 6 and 7.5 */
  dcl_icoor *band;                 /* device: the new band, band_capacity entries */
  int64_t band_capacity;
  int64_t band_count;              /* out: the new band's length (may exceed capacity) */
} dcl_reinit_cell;

/* LevelSetReinitialize (LevelSetMethod/FastMarching.c:10-146) of ncells
 * level sets: phi reset to +-phi_inf, the signed distance to the orthogonal
 * projection at every projection record (axis -1) as the boundary values,
 * then the second-order fast marching (FastMarching3D.c:14-101,
 * FastMarching2D.c:33-185) out to band_width on each side, with the
 * reference's heap order (Common/Heap.c:59-139), so phi and the band come
 * out as the reference's: band = [projection nodes in record order | - side
 * in acceptance order | + side in acceptance order].  Neighbours outside the
 * window are not pushed (the reference reads them unchecked).
 * heap_capacity: heap entries per side and level set (0: 2 * count + 4096;
 * a full heap is an error).  All level sets of a call are 2D or all 3D.
 * Synchronous: band_count is set on return.  Each side of each level set is
 * one wave, so a batch runs its level sets side by side. */
int dcl_reinitialize(dcl_reinit_cell *cells, int ncells, int64_t heap_capacity, void *stream);

/* The same reinitialisation computed in parallel sweeps, for level sets too
 * large for the fast marching's one-node-at-a-time order (it replaces
 * LevelSetReinitialize, LevelSetMethod/FastMarching.c:10-146, where a
 * result within the fast marching's own discretisation error will do).
 * Same boundary values and the same second-order update
 * (FastMarching2D.c:33-94); each sweep gives every node the value the fast
 * marching would give it if its neighbours were accepted in the order of
 * their values (its pushes replayed in that order), from the previous
 * sweep's values, until no node moves by more than band_width * 1e-9.
 * phi comes out bit for bit as dcl_reinitialize's at most nodes (98% of the
 * band of a 256^3 sphere: wherever the fast marching accepts in value order
 * and no tie is decided by its heap's layout); elsewhere the two differ by a
 * fraction of either's distance error (the same band set but at band_width).  band = [projection nodes in record order |
 * - side | + side], each side in window order (z, y, x).  max_sweeps: the
 * sweeps allowed per level set (0: 256; not settling is an error); sweeps
 * (host, ncells entries, may be NULL) gets the sweeps each level set took.
 * Synchronous; band_count is set on return.  Uses device scratch of about
 * 19.3 bytes per window node (a second buffer, masks, marks, the node
 * list, the band's row ballots), kept by the calling thread for later
 * calls. */
int dcl_reinitialize_parallel(dcl_reinit_cell *cells, int ncells, int max_sweeps, int *sweeps,
                              void *stream);

/* LevelSetGetVelocity (LevelSetMethod/Advect.c:145-192) from a fluid field
 * resident on the device.  fluid: the fluid's [z][y][x][4] array (p, u, v,
 * w) over its domain of fluid_n nodes, node (0,0,0) at global (0,0,0), 3D
 * records aligned to their size (32 bytes in DCL_F64, 16 in DCL_F32); in 2D
 * (fluid_n->z < 2) [y][x][3] (p, u, v).  Copies the velocity components
 * (U_FACE..W_FACE) at the nodes of win inside the domain into vel, the
 * [z][y][x][3] ([y][x][2]) array over win; nodes of win outside the domain
 * are not written (the reference's get is clipped to the domain,
 * Advect.c:160-186).  Asynchronous on stream. */
int dcl_get_velocity(const void *fluid, const dcl_icoor *fluid_n, const dcl_window *win, int precision,
                     void *vel, void *stream);

/* dcl_get_velocity writing vel as component planes, the dof-1 layout of
 * dcl_velocity ([c][z][y][x], 2D [c][y][x], each plane over win), which
 * dcl_sl_advect and dcl_levelset_step take; same copy, same clipping.  The
 * reference's velocity grid is per node (dof 3); this layout is the FP64 SL
 * kernels' fastest (one load per pair of x-neighbouring values). */
int dcl_get_velocity_planes(const void *fluid, const dcl_icoor *fluid_n, const dcl_window *win, int precision,
                            void *vel, void *stream);

/* The surface quantities of one irregular-node record: the fields that
 * IIMUpdateSurfaceQuantities_3D / _2D (IIM/SurfaceQuantities.c:31-118) set
 * in the reference's IrregularNode (LevelSetMethod.h:30-54): unit normal n
 * (nx, ny, nz), tangents s and r (LocalCoor3DTangential,
 * IIM/LocalCoor.c:38-65), curvature k and the principal curvatures k_nn,
 * k_tt, k_nt.  In 2D: s = (ny, -nx), r = 0, k_nn = k, k_tt = k_nt = 0.
 * 104 bytes. */
typedef struct {
  dcl_coor n, s, r;
  double k, k_nn, k_tt, k_nt;
} dcl_surface;

/* IIMUpdateSurfaceQuantities_3D / _2D (IIM/SurfaceQuantities.c:31-118)
 * without the user's force callback (iim->F): at the X of each of the count
 * records in nodes (device, as written by dcl_irregular_nodes), the
 * Bilinear3D / Bilinear2D (Common/Grid3D.c:4-31, Common/Grid.c:190-212)
 * interpolations of the central-difference functionals
 * (Common/Grid3D.c:38-108, Common/Grid.c:219-254), in double.  phi: device
 * window (DCL_F64, or DCL_F32 read as double); out: device, count records.
 * Asynchronous on stream.  Reads clamp into the window.  count < 2^31. */
int dcl_surface_quantities(const void *phi, const dcl_window *win, const dcl_irregular_node *nodes,
                           int64_t count, int precision, dcl_surface *out, void *stream);

/* ---------------------------------------------------------------- formats */
/* The bytes of the .Real64 file GridWrite writes for a grid's values
 * (Common/Serialize.c:23-55, VecView to a PETSc binary viewer): a big-endian
 * int32 class id 1211214 and int32 count, then count big-endian doubles --
 * the layout Visualization/DCellVTK.py:24-32 reads.  src: device, count
 * values (DCL_F64, or DCL_F32 widened to double); dst: device, 8 + 8 * count
 * bytes, 8-byte aligned.  Asynchronous on stream.  count < 2^31. */
int dcl_real64_pack(const void *src, int precision, int64_t count, void *dst, void *stream);

/* The rows LevelSetWriteIrregularNodeList writes for a node list
 * (LevelSetMethod/IrregularNodes.c:145-219), big-endian doubles.  3D: one
 * row of 23 per record, every record: X (3), pos (3), n (3), k, f1, f1_n,
 * f1_t, f1_nn, f1_tt, f1_nt, s (3), r (3), numNei.  2D: one row of 11 per
 * crossing record (axis != -1): X.x, X.y, pos.x, pos.y, n.x, n.y, f1, fa1, k,
 * f1_n, f1_nn.  nodes and surf (device, count each) as dcl_irregular_nodes
 * and dcl_surface_quantities write them; extra (device, may be NULL for
 * zeros): the IIM fields of each record that this library does not compute,
 * 7 per record in 3D (f1, f1_n, f1_t, f1_nn, f1_tt, f1_nt, numNei), 4 in 2D
 * (f1, fa1, f1_n, f1_nn).  dst: device, capacity rows.  Synchronous:
 * *rows_out (host) is the number of rows, which may exceed capacity (then
 * only capacity are written).  count < 2^31. */
int dcl_irregular_node_rows(const dcl_irregular_node *nodes, const dcl_surface *surf, const double *extra,
                            int64_t count, int dim, void *dst, int64_t capacity, int64_t *rows_out,
                            void *stream);

/* Interface measure with LevelSetDiracDelta3D / 2D
 * (LevelSetMethod/NormalDirection.c:36-57) at the nodes [p+1, p+n-2]:
 * out[0] = sum delta dV (area; length in 2D), out[1] = sum delta kappa dV
 * (kappa may be NULL), out[2] = out[1]/out[0], out[3] = nodes with
 * |phi| <= eps.  Synchronous, out is host memory. */
int dcl_interface_metrics(const void *phi, const void *kappa, const dcl_window *win,
                          double eps, int precision, double out[4], void *stream);

/* max|a-b|, sum (a-b)^2, max|b|, n over n elements, a and b of their own
 * precisions.  Synchronous, out is host memory. */
int dcl_error_norms(const void *a, int prec_a, const void *b, int prec_b, int64_t n,
                    double out[4], void *stream);

/* ---------------------------------------------------------------- batches */
/* Many independent level sets (the reference's DCells, which each rank
 * advects one after another, DWorld/DCellsArray.c:57-60) updated together:
 * the cells' windows are packed back to back in one phi buffer, one kappa
 * buffer and one velocity buffer (dim components per node, as a velgrid),
 * cut into tiles, and one step is two persistent launches over the packed
 * tile list (SL then curvature, as dcl_levelset_step per cell), each XCD
 * working through a chunk of the list of about equal cost, so cells of
 * uneven sizes load the whole chip.  A batch holds 3D cells or 2D cells, not
 * both.  No counterpart in the reference.  The velocity of each cell shares
 * the cell's window. */
typedef struct dcl_batch_s *dcl_batch;

/* cell-size classes of a 3D batch (largest axis <= 48, <= 80, larger) and
 * the ring z-chunks the autotuner tries for each (16, 32, 64, 128) */
#define DCL_BATCH_CLASSES 3
#define DCL_BATCH_ZCHUNKS 4
/* tiles per claim the autotuner tries: 0 (static stride), 1, 4, 16 */
#define DCL_BATCH_CLAIMS 4

int dcl_batch_create(int32_t ncells, const dcl_window *wins, int precision, int interp,
                     int backtrace, dcl_batch *batch);
int dcl_batch_destroy(dcl_batch batch);
/* element offsets of every cell in the packed phi/kappa buffer and in the
 * packed velocity buffer; ncells + 1 entries each (the last is the total) */
int dcl_batch_layout(dcl_batch batch, int64_t *phi_offsets, int64_t *vel_offsets);
/* the step size of every cell (host array of ncells); required before stepping */
int dcl_batch_set_dt(dcl_batch batch, const double *dt);
/* configuration of the launches (3D): shape 0..5 = curvature and SL tile
 * shape (tile[] of dcl_batch_get_config), shape 6..11 = the curvature tiles
 * of shape - 6 with the SL step over 32 x 8 columns of the cell class's
 * z-chunk staged in an LDS ring; occupancy 0..2 -> 2, 4, 8 resident blocks
 * per CU.  2D batches use 64 x 4 node tiles whatever the shape. */
int dcl_batch_set_config(dcl_batch batch, int shape, int occupancy);
int dcl_batch_get_config(dcl_batch batch, int *shape, int *occupancy, int32_t tile[3],
                         int *blocks);
/* Work distribution of both launches: the tile list is cut into one chunk
 * of about equal cost per XCD; with tiles_per_claim 0 (the default) the
 * blocks of an XCD take its chunk's tiles in a fixed stride, with k > 0 they
 * claim k tiles at a time from a per-chunk atomic counter and then take
 * tiles from the other chunks */
int dcl_batch_set_claim(dcl_batch batch, int tiles_per_claim);
int dcl_batch_get_claim(dcl_batch batch, int *tiles_per_claim);
/* ring z-chunk of one size class (default 32) */
int dcl_batch_set_class_zchunk(dcl_batch batch, int cls, int zchunk);
/* per size class: the largest cell axis it holds (-1: unbounded), its cell
 * count and its ring z-chunk; any pointer may be NULL */
int dcl_batch_get_classes(dcl_batch batch, int32_t max_edge[DCL_BATCH_CLASSES],
                          int32_t cells[DCL_BATCH_CLASSES], int32_t zchunk[DCL_BATCH_CLASSES]);
/* one step phi0 -> phi, kappa(phi) for every cell */
int dcl_batch_step(dcl_batch batch, const void *phi0, void *phi, void *kappa, const void *vel,
                   void *stream);
/* Band mode, the band loop of LevelSetAdvectSL_2D / _3D
 * (LevelSetMethod/AdvectSL.c:23-89) for every cell at once: band is a device
 * array of count entries {x, y, z, cell} (int32; global node coordinates of
 * that cell's window, z ignored in 2D).  phi gets the SL value and then kappa
 * the curvature of phi at every listed node; other nodes are left as they
 * are.  No allocation. */
int dcl_batch_band_step(dcl_batch batch, const void *phi0, void *phi, void *kappa,
                        const void *vel, const int32_t *band, int64_t count, void *stream);
/* Times every configuration (reps steps each, after one warm-up step) on
 * these buffers with the static stride and keeps the fastest, then, when that
 * uses ring columns, each size class's z-chunk in turn, then the tiles per
 * claim.  ms_table (optional, 36 entries, shape-major, -1 where not timed),
 * zc_table (optional, DCL_BATCH_CLASSES x DCL_BATCH_ZCHUNKS, -1 where not
 * timed) and claim_table (optional, DCL_BATCH_CLAIMS entries) receive the
 * per-step times.  Synchronous. */
int dcl_batch_autotune(dcl_batch batch, const void *phi0, void *phi, void *kappa,
                       const void *vel, int reps, double *ms_table, double *zc_table,
                       double *claim_table, double *best_ms, void *stream);
/* One step with these buffers recorded as an instantiated hipGraph (two
 * kernel nodes); launch it with dcl_batch_graph_launch. */
int dcl_batch_graph_create(dcl_batch batch, const void *phi0, void *phi, void *kappa,
                           const void *vel, void **graph_exec);
int dcl_batch_graph_launch(void *graph_exec, void *stream);
int dcl_batch_graph_destroy(void *graph_exec);

/* --------------------------------------------------------------- adjoint */
/* Reverse-mode derivative of one dense 3D SL step (dcl_sl_advect, dense):
 * given lam_next = dL/dphi, accumulate lam += dL/dphi0 and gvel += dL/dvel
 * (gvel has the layout of vel->data).  Transposed gather with atomics in LDS
 * and then HBM.  No counterpart in the reference.  Requires vel->win == *win. */
int dcl_sl_adjoint(const void *phi0, const dcl_window *win, const dcl_velocity *vel,
                   const void *lam_next, void *lam, void *gvel, double dt,
                   int precision, int interp, int backtrace, void *stream);

/* The adjoint of N SL steps with one velocity (a steady velocity, as in the
 * differentiation form), split so that the velocity work is done once:
 *   dcl_sl_adjoint_prepare: every interior node's departure stencil (base,
 *     fractions, clamped axes) into `record` (a device buffer of
 *     dcl_sl_adjoint_record_bytes bytes; 48 per node);
 *   dcl_sl_adjoint_step, once per step in reverse order: lam += dL/dphi0 of
 *     the step from lam_next = dL/dphi, and hsum (3 per node, device, start
 *     at zero) += the step's departure sensitivities -c lam_next dI/dS;
 *   dcl_sl_adjoint_finish: gvel += dL/dvel of all the steps, formed from hsum
 *     with the velocity scatter of dcl_sl_adjoint (MAC corners at the RK2
 *     midpoint and the node's faces).
 * The sum of dcl_sl_adjoint over the steps equals these to rounding.  No
 * counterpart in the reference; the velocity window must equal phi's, and
 * dcl_sl_adjoint_step takes phi windows below 2 GiB (DCL_ERR_ARG_SIZ). */
int dcl_sl_adjoint_record_bytes(const dcl_window *win, int64_t *bytes);
int dcl_sl_adjoint_prepare(const dcl_window *win, const dcl_velocity *vel, double dt, int precision,
                           int interp, int backtrace, void *record, void *stream);
int dcl_sl_adjoint_step(const void *phi0, const dcl_window *win, const void *record,
                        const void *lam_next, void *lam, void *hsum, double dt, int precision,
                        int interp, void *stream);
int dcl_sl_adjoint_finish(const dcl_window *win, const dcl_velocity *vel, const void *hsum,
                          void *gvel, double dt, int precision, int backtrace, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* DCELL_LSM_H */
