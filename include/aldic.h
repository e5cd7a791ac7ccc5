/*
 * aldic.h -- C ABI of the MI355X (gfx950) implementation of ALDIC's local
 * matching path.  Plain pointers and sizes only; every entry point returns
 * ALDIC_OK or an error code, with a message from aldic_last_error().
 *
 * Data layout is the reference's MATLAB layout: images are M x N
 * column-major double arrays whose first dimension is image x (the
 * reference loads images transposed, gui/loadStereoImages.m:79); node
 * coordinates are 1-based; U is interleaved [u1 v1 u2 v2 ...];
 * F is [F11 F21 F12 F22 ...] per node (LocalICGNQuadtree.m:159-163).
 *
 * Entry points with host pointers copy in and out (the MATLAB/MEX-facing
 * drop-in).  Entry points with the _dev suffix take device pointers that
 * are already resident in HBM and run asynchronously on the context's
 * stream.
 *
 * Per-subset failures are not errors: they are reported through convIt
 * exactly as the reference reports them (-1 where MATLAB would throw,
 * 101 not converged, 102 >= 40% masked, 103 empty subset).
 *
 * Image, plan and network handles may be destroyed before or after the
 * context that made them; every other call needs a live context.
 */
#ifndef ALDIC_H
#define ALDIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALDIC_OK 0
#define ALDIC_EINVAL 1  /* bad argument */
#define ALDIC_EDEVICE 2 /* HIP runtime error */
#define ALDIC_ENOMEM 3  /* device allocation failed */

typedef struct aldic_ctx aldic_ctx;
typedef struct aldic_images aldic_images;
typedef struct aldic_plan aldic_plan;

/* Message of the last error raised on this thread. */
const char* aldic_last_error(void);

/* A context owns a HIP stream (default priority) and reusable device
 * buffers for the host-pointer entry points.  Contexts are independent;
 * one context must not be used from two threads at once.  Replaces the
 * MEX build/load step (func/ensureMexCompiled.m:1-88). */
int aldic_ctx_create(int device, aldic_ctx** out);
int aldic_ctx_destroy(aldic_ctx* ctx);
/* hipStream_t of the context, for callers that order their own work. */
void* aldic_ctx_stream(aldic_ctx* ctx);
int aldic_ctx_synchronize(aldic_ctx* ctx);

/* Per-stage kernel timing with HIP events on the context stream, the
 * counterpart of the reference's per-stage tic/toc (LocalICGNQuadtree.m:
 * 69/107, Subpb1Quadtree.m:73/110, funIntegerSearch.m:130/182).  When
 * enabled, every launch of a stage is bracketed by two events;
 * aldic_ctx_stage_times synchronises the stream and returns, per stage,
 * the summed milliseconds and the number of launches (arrays of
 * ALDIC_NSTAGES entries, either may be NULL), optionally resetting them. */
#define ALDIC_STAGE_SEARCH 0
#define ALDIC_STAGE_GRID2NODES 1
#define ALDIC_STAGE_ICGN_C0 2      /* +0..3: fast kernel per size class, */
                                   /* +4: generic (> 65 px), +5: handed-on subsets */
#define ALDIC_STAGE_SUBPB1_C0 8    /* same layout for Subpb1 */
#define ALDIC_STAGE_CNN 14         /* learned starting point: patches + network */
#define ALDIC_STAGE_FALLBACK 15    /* gate, node search and IC-GN rerun */
#define ALDIC_STAGE_SUBPB2 16      /* global FE step: assembly and solve */
#define ALDIC_NSTAGES 17
/* Ordering between contexts (one stream each) on one device, for callers
 * that pipeline independent frames over two contexts: aldic_ctx_record
 * marks the point reached by the work submitted so far to ctx (mark
 * 0 .. ALDIC_NMARKS-1, re-recording replaces it); aldic_ctx_wait makes
 * work submitted later to ctx wait for the last recording of `mark` on
 * other.  No counterpart in the reference. */
#define ALDIC_NMARKS 8
int aldic_ctx_record(aldic_ctx* ctx, int32_t mark);
int aldic_ctx_wait(aldic_ctx* ctx, aldic_ctx* other, int32_t mark);

/* HIP graph of a context's work (the stereo/temporal local step repeated
 * per frame): everything submitted to ctx between capture_begin and
 * capture_end is recorded, not run; aldic_graph_launch replays it on a
 * context's stream with the arguments as captured (device-resident counts
 * are read when the graph runs).  Stage timing must be off while
 * capturing.  Run the sequence once before capturing it: the context's
 * staging buffers are pinned from capture_begin until every graph captured
 * on it is destroyed, and a call that would have to grow one meanwhile
 * fails with ALDIC_EINVAL.  Waits on marks recorded outside the capture
 * cannot be captured; order graphs across contexts with record/wait
 * between launches.  No counterpart in the reference. */
typedef struct aldic_graph aldic_graph;
int aldic_ctx_capture_begin(aldic_ctx* ctx);
int aldic_ctx_capture_end(aldic_ctx* ctx, aldic_graph** out);
int aldic_graph_launch(aldic_graph* graph, aldic_ctx* ctx);
int aldic_graph_destroy(aldic_graph* graph);

/* The DICpara fields the local matching path reads, as a plain struct
 * (setDICparaDefaults.m:18-56; the drivers' tolerances StereoMatch_STAQ.m:71,
 * TemporalMatch_quadtree_ST1.m:190; the ALM constants ibid. :266-298, :333).
 * A search distance holds n = 0 (empty, []), 1, 2 or 4 values.  winsize and
 * winstepsize 0 mean unset ([]).  beta 0 means "derive": 1e-3 *
 * winstepsize^2 * mu (TemporalMatch_quadtree_ST1.m:298).  The GUI, plotting,
 * strain and smoothing fields are not on this path. */
typedef struct {
  int32_t winsize, winstepsize, winsizeMin;
  int64_t gridx[2], gridy[2];
  int32_t StereoSearchDistance[4], nStereoSearchDistance;
  int32_t TemporalSearchDistance[4], nTemporalSearchDistance;
  int32_t NewFFTSearchDistance[4], nNewFFTSearchDistance;
  int32_t fixSearchDistanceOrNot;
  int32_t InitFFTSearchMethod; /* 1: whole-field search, the one built */
  int32_t DICIncOrNot, UseGlobal, GaussPtOrder, ClusterNo;
  double tol_stereo, tol_temporal;
  double mu, beta, alpha, tol2;
} aldic_params;
/* setDICparaDefaults.m's values (winsize, winstepsize unset). */
int aldic_params_default(aldic_params* p);
/* validateDICpara.m's rules on these fields; ALDIC_EINVAL with the
 * reference's message when one fails. */
int aldic_params_validate(const aldic_params* p);
/* The search region [L R U D] of IntegerSearchQuadtree.m:52-90 and the
 * repmat of funIntegerSearch.m:12 for ImgSeqNum (-1 stereo, >= 2 temporal);
 * ALDIC_EINVAL where the reference would prompt with input(). */
int aldic_params_search_region(const aldic_params* p, int32_t ImgSeqNum, int64_t region[4]);

/* How the integer searches on this context (every aldic_integer_search*
 * and aldic_search_* entry) find the integer ZNCC peak; results follow
 * funIntegerSearch.m:154-169 either way.  ALDIC_SEARCH_DIRECT: direct sums
 * per lag.  ALDIC_SEARCH_FFT: the numerators of each 64 x 64 block of lags
 * as one 128 x 128 FFT cross-correlation in LDS (MATLAB's normxcorr2 also
 * works in the frequency domain when that is faster), for template sides up
 * to 65 px (larger ones stay direct); the window sums and the 3 x 3
 * sub-pixel fit are direct.  ALDIC_SEARCH_AUTO (default): FFT blocks for a
 * whole grid of at least 1024 points with lag windows of at least 32 x 32,
 * direct sums otherwise (an FFT block holds most of a CU's LDS; listed and
 * small searches that run beside other kernels keep the direct tiles, which
 * share CUs).  The FP32
 * numerators differ from the direct sums by rounding, so a near-tie for the
 * peak can resolve differently. */
#define ALDIC_SEARCH_DIRECT 0
#define ALDIC_SEARCH_FFT 1
#define ALDIC_SEARCH_AUTO 2
int aldic_ctx_set_search_method(aldic_ctx* ctx, int32_t method);

int aldic_ctx_set_timing(aldic_ctx* ctx, int32_t enable);
int aldic_ctx_stage_times(aldic_ctx* ctx, double* ms, int32_t* launches, int32_t reset);

/* Device memory for the _dev entry points (plumbing with no counterpart in
 * the reference, whose MATLAB workers own every array).  Copies are
 * synchronous on the context stream. */
int aldic_dev_malloc(aldic_ctx* ctx, int64_t bytes, void** out);
int aldic_dev_free(aldic_ctx* ctx, void* p);
int aldic_memcpy_htod(aldic_ctx* ctx, void* dst, const void* src, int64_t bytes);
int aldic_memcpy_dtoh(aldic_ctx* ctx, void* dst, const void* src, int64_t bytes);

/* Z = ba_interp2_spline(F, X, Y): F is M x N x P, X (dim-2 coordinate) and
 * Y (dim-1 coordinate) hold ND points, Z is ND x P.  Same class for all
 * arrays.  Replaces mexFunction / interpolate<REAL>,
 * ba_interp2_spline.cpp:359-435. */
int aldic_interp_spline_f64(aldic_ctx* ctx, const double* F, int64_t M, int64_t N, int64_t P,
                            const double* X, const double* Y, int64_t ND, double* Z);
int aldic_interp_spline_f32(aldic_ctx* ctx, const float* F, int64_t M, int64_t N, int64_t P,
                            const float* X, const float* Y, int64_t ND, float* Z);

/* Df = funImgGradient(ImgRef, ImgDef, mask): DfDx, DfDy are (M-6) x (N-6)
 * (Df.DfCropWidth = 3).  mask may be NULL (all ones).  Replaces
 * func/funImgGradient.m:1-87 (finite-difference branch). */
int aldic_img_gradient(aldic_ctx* ctx, const double* ImgRef, const double* mask, int64_t M,
                       int64_t N, double* DfDx, double* DfDy);

/* Upload the normalised reference and deformed images, the reference mask
 * (Df.ImgRefMask, NULL = all ones) and Df.DfDx/Df.DfDy ((M-6) x (N-6), NULL
 * = computed on the device by the funImgGradient rule) as resident FP32
 * planes.  This is the state that LocalICGNQuadtree, Subpb1Quadtree and
 * IntegerSearchQuadtree receive as (Df, ImgRef, ImgDef). */
int aldic_images_create(aldic_ctx* ctx, const double* ImgRef, const double* ImgDef,
                        const double* mask, const double* DfDx, const double* DfDy, int64_t M,
                        int64_t N, aldic_images** out);
/* Same, from FP32 planes already on the device (copied into the handle):
 * fm = ImgRef .* mask, g = ImgDef, mask (NULL = ones), fx/fy NULL = computed
 * from fm on the device. */
int aldic_images_create_dev(aldic_ctx* ctx, const float* fm, const float* g, const float* mask,
                            const float* fx, const float* fy, int64_t M, int64_t N,
                            aldic_images** out);
/* Images from raw 8- or 16-bit pixels as loadStereoImages decodes them
 * (gui/loadStereoImages.m:79, double(raw)': a row-major H x W pixel buffer
 * is the reference's M = W by N = H column-major image): each is
 * normalised on the device by funNormalizeImg (func/funNormalizeImg.m:
 * 22-40: (Img - mean) / std over the ROI gridx x gridy, clamped to the
 * image; mean and std from exact integer sums) in FP64 and then prepared as
 * aldic_images_create with Df computed on the device.  Uploads 1 or 2 bytes
 * per pixel instead of 8.  stats (may be NULL) receives mean and std of
 * ImgRef then ImgDef. */
int aldic_images_create_raw(aldic_ctx* ctx, const void* ImgRef, const void* ImgDef, int32_t bits,
                            const double* mask, int64_t M, int64_t N, const int64_t gridx[2],
                            const int64_t gridy[2], aldic_images** out, double* stats);
/* funNormalizeImg of one raw device image into an FP64 device image (M x N,
 * same memory order); mean_std (host, may be NULL) receives mean and std. */
int aldic_normalize_raw_dev(aldic_ctx* ctx, const void* d_raw, int32_t bits, int64_t M, int64_t N,
                            const int64_t gridx[2], const int64_t gridy[2], double* d_out,
                            double* mean_std);
/* Device pointers of the resident FP32 planes (M x N each): fm = ImgRef
 * .* mask, g = ImgDef, mask, and Df.DfDx / Df.DfDy on the full grid (zero
 * outside rows/columns 4..M-3, 4..N-3).  Any may be NULL.  The planes
 * belong to the handle. */
int aldic_images_planes(const aldic_images* im, const float** fm, const float** g,
                        const float** mask, const float** fx, const float** fy);

/* Interpolant of ImgDef used by every solver on these images.  No
 * counterpart on the reference's path: its MEX evaluates a local 4 x 4
 * kernel on raw pixels (ba_interp2_spline.cpp:310-325), which is
 * ALDIC_INTERP_LOCAL, the default and the only mode with reference parity.
 * ALDIC_INTERP_BSPLINE is the interpolating cubic B-spline (the exact spline
 * whose coefficients the reference's kernel approximates cell by cell, cf.
 * func/BicubicBsplineBuildFullTable.m:1-24): its coefficient plane is made
 * here by a separable recursive prefilter with mirror boundaries, and the
 * solvers evaluate the spline on it.  It removes the local kernel's
 * position bias (up to 0.0069 px on a ramp) and changes results by about
 * that much; the reference's "g == 0" mask tests then see the spline's
 * values.  Calling it again (e.g. after writing g through
 * aldic_images_planes) remakes the coefficients; returns when they are
 * complete. */
#define ALDIC_INTERP_LOCAL 0
#define ALDIC_INTERP_BSPLINE 1
int aldic_images_set_interp(aldic_images* im, int32_t mode);
/* The prefilter alone: interpolating cubic B-spline coefficients of an
 * M x N FP32 device plane (column-major, mirror boundaries) into d_out,
 * stream-ordered on ctx.  d_in and d_out must not overlap.  No counterpart
 * in the reference. */
int aldic_bspline_prefilter_dev(aldic_ctx* ctx, const float* d_in, float* d_out, int64_t M,
                                int64_t N);
int aldic_images_destroy(aldic_images* im);

/* Per-node part of LocalICGNQuadtree (func_quadtree/LocalICGNQuadtree.m:
 * 67-163, one funICGNQuadtree per node, func_quadtree/funICGNQuadtree.m:
 * 1-454), first-order, 'GaussNewton'.  coords is n x 2 column-major
 * (coordinatesFEM), U0 is 2n; winsize holds one value (winsize_per_node=0,
 * DICpara.winsize) or one per node.  Outputs: U (2n), F (4n), H21 (n x 21
 * column-major, HtempPar; may be NULL), convIt (n, ConvItPerEle), markHole
 * (n; may be NULL), zncc (n, ZNCC of the last evaluated warp, no
 * counterpart in the reference; may be NULL).  The bad-point statistics and
 * the NaN fill of LocalICGNQuadtree.m:167-325 are the caller's (host). */
int aldic_local_icgn(aldic_ctx* ctx, const aldic_images* im, const double* coords,
                     const double* U0, int64_t n, const int32_t* winsize,
                     int32_t winsize_per_node, double tol, double* U, double* F, double* H21,
                     int32_t* convIt, uint8_t* markHole, double* zncc);

/* Per-node part of Subpb1Quadtree (func_quadtree/Subpb1Quadtree.m:71-165,
 * funICGN_Subpb1quadtree.m:1-476).  winsize_xy is n x 2 column-major
 * (DICpara.winsize_List).  udual (4n) is read by the reference but does
 * not enter the update.  Outputs U (2n) and convIt (n; 0 where MATLAB
 * would throw, as the reference leaves it). */
int aldic_subpb1(aldic_ctx* ctx, const aldic_images* im, const double* coords,
                 const double* UOld, const double* FOld, const double* udual,
                 const double* vdual, int64_t n, const int32_t* winsize_xy, double mu,
                 double tol, double* U, int32_t* convIt);

/* A plan fixes the node set of repeated device-resident solves: it sorts
 * nodes into size classes (one wave per subset up to 33 px, four waves up
 * to 65 px, a workgroup beyond) and owns the scratch for subsets handed
 * from the fast to the generic kernel.  winsize is a host array (one value
 * or n values), or n x 2 (winsize_List) when rect is nonzero. */
int aldic_plan_create(aldic_ctx* ctx, const int32_t* winsize, int64_t n, int32_t winsize_per_node,
                      int32_t rect, aldic_plan** out);
int aldic_plan_destroy(aldic_plan* plan);

/* Device-resident LocalICGN: d_coords n x 2 (double), d_U0 2n (float);
 * outputs float/int device arrays as aldic_local_icgn (d_F, d_H21, d_hole,
 * d_zncc may be NULL).  Asynchronous on the context stream.  d_winsize
 * must hold the winsize values the plan was built from (the plan's size
 * classes are fixed on the host); a node whose device value is larger than
 * its class holds is solved by the workgroup kernel instead, never with a
 * clipped window. */
int aldic_local_icgn_dev(aldic_ctx* ctx, const aldic_images* im, const aldic_plan* plan,
                         const double* d_coords, const float* d_U0, const int32_t* d_winsize,
                         double tol, float* d_U, float* d_F, float* d_H21, int32_t* d_convIt,
                         uint8_t* d_hole, float* d_zncc);

/* Autotuner of the fast IC-GN kernel: per size class, times each launch
 * shape (waves per subset x pixels per lane; ALDIC_NVARIANTS of them) over
 * reps launches on the given nodes and stores the fastest in the plan (used
 * by both solvers).  chosen (4 entries, may be NULL) receives the picks.
 * The timed launches are real solves: outputs are overwritten. */
#define ALDIC_NVARIANTS 3
int aldic_local_icgn_autotune_dev(aldic_ctx* ctx, const aldic_images* im, aldic_plan* plan,
                                  const double* d_coords, const float* d_U0,
                                  const int32_t* d_winsize, double tol, float* d_U, float* d_F,
                                  int32_t* d_convIt, float* d_zncc, int32_t reps, int32_t* chosen);
/* Fix the launch shape of one size class (solver 0: IC-GN, 1: Subpb1). */
int aldic_plan_set_variant(aldic_plan* plan, int32_t solver, int32_t cls, int32_t variant);
/* Stopping rule of the plan's IC-GN solves (the _dev entry points; Subpb1
 * keeps the reference's).  rule 0 (default): the reference's, ||b_k|| /
 * ||b_1|| < tol or ||b_k|| < tol, tested before the update, which is then
 * not applied (funICGNQuadtree.m:376-389).  rule 1: the step-size rule for
 * throughput runs, no counterpart in the reference: the update is applied
 * and the subset is converged when it moved the subset by less than tol_dp
 * px, sqrt(du^2 + dv^2 + (du/dx^2 + dv/dx^2) hx^2 + (du/dy^2 + dv/dy^2)
 * hy^2) < tol_dp with (hx, hy) the subset's half widths; the status is the
 * number of evaluations, 101 if the cap comes first.  rule 2: the
 * a-posteriori form of rule 1, no counterpart in the reference either:
 * with m_k that measure of the k-th update, converged when m_k < tol_dp or
 * when the iteration contracts (m_k <= m_{k-1} / 2) and the distance to its
 * fixed point predicted from the observed ratio q = m_k / m_{k-1},
 * m_k q / (1 - q), is below tol_dp (m_k^2 < tol_dp (m_{k-1} - m_k)): it
 * saves the evaluation rule 1 spends confirming a step it could predict.
 * The ZNCC output is that of the last evaluated warp, before the last
 * update. */
int aldic_plan_set_stopping(aldic_plan* plan, int32_t rule, double tol_dp);
/* on = 1: the plan's solves (aldic_local_icgn*_dev, aldic_subpb1_dev) run
 * the four fast size classes as one persistent launch fed by a size-sorted
 * work queue instead of one launch per class on two streams.  Its launch
 * shapes are fixed (four waves per 65 and per 33 px subset, one wave per 17
 * and per 11 px subset: variants 0, 0, 2, 0 of classes 0..3; the plan's
 * autotuned ones are not used); with the same shapes the per-class launches
 * give bit-identical results.  No counterpart in the reference (its
 * parfor). */
int aldic_plan_set_persistent(aldic_plan* plan, int32_t on);

/* Same as aldic_local_icgn_dev with a starting guess of the gradients:
 * P0 = [F0 U0] instead of [0 0 0 0 U0] (funICGNQuadtree.m:41).  d_F0 is 4n
 * floats [u_x v_x u_y v_y] per node, or NULL for zeros.  max_iter caps the
 * loop at k <= max_iter (0: the reference's 100, funICGNQuadtree.m:135);
 * a subset stopped by the cap reports 101.  The learned starting point
 * feeds this; it has no counterpart in the reference. */
int aldic_local_icgn_f0_dev(aldic_ctx* ctx, const aldic_images* im, const aldic_plan* plan,
                            const double* d_coords, const float* d_U0, const float* d_F0,
                            const int32_t* d_winsize, double tol, int32_t max_iter, float* d_U,
                            float* d_F, float* d_H21, int32_t* d_convIt, uint8_t* d_hole,
                            float* d_zncc);

/* Device-resident Subpb1. */
int aldic_subpb1_dev(aldic_ctx* ctx, const aldic_images* im, const aldic_plan* plan,
                     const double* d_coords, const float* d_UOld, const float* d_FOld,
                     const float* d_vdual, const int32_t* d_winsize_xy, double mu, double tol,
                     float* d_U, int32_t* d_convIt);

/* funIntegerSearchWholeField (func/funIntegerSearch.m:100-203) as called by
 * IntegerSearchQuadtree (func_quadtree/IntegerSearchQuadtree.m:93) with
 * InitFFTSearchMethod = 1.  region = [L R U D] after the repmat of
 * funIntegerSearch.m:12.  Grid outputs are nY x nX column-major; pass NULL
 * outputs to query nX, nY.  ipeak_row/col are the 1-based integer peak in
 * the cropped lag window (0 where normxcorr2 would throw, u = v = NaN).
 * qfactors (may be NULL) receives cc.qfactors, nX*nY x 2 column-major
 * (PCE, PPE of each point's full normxcorr2 map, compute_qFactor :211-264,
 * normalised over the grid :184-187), rows in the reference's point order
 * (x index fastest). */
int aldic_integer_search(aldic_ctx* ctx, const double* f, const double* g, int64_t M, int64_t N,
                         const int64_t gridx[2], const int64_t gridy[2], int32_t winsize,
                         int32_t winstepsize, const int64_t region[4], double* x, double* y,
                         double* u, double* v, double* ccmax, int32_t* ipeak_row,
                         int32_t* ipeak_col, double* qfactors, int64_t* nX, int64_t* nY);

/* Initial guess of arbitrary nodes from the search grid (as the reference
 * carries U0 from the search grid onto quadtree nodes,
 * func/StereoMatch_STAQ.m:20 and :59): bilinear interpolation of the nY x nX
 * grids d_u, d_v (node centres x0 + ci*step, y0 + cj*step), constant
 * beyond the outer nodes.  d_coords n x 2 (double), d_U0 2n (float). */
int aldic_grid_to_nodes_dev(aldic_ctx* ctx, const double* d_u, const double* d_v, int64_t nX,
                            int64_t nY, double x0, double y0, double step,
                            const double* d_coords, int64_t n, float* d_U0);
/* A warp field held on a regular grid (point t = ci + cj*nX at (x0 +
 * ci*step, y0 + cj*step), FP32 device arrays in the solvers' layout: d_Ug
 * 2 and d_Fg 4 values per point) -> its bilinear interpolant at every node
 * (d_U0 2n, d_F0 4n, may be NULL), constant beyond the outer points.  The
 * learned start uses it to run the network on a grid coarser than the
 * nodes.  No counterpart in the reference. */
int aldic_field_to_nodes_dev(aldic_ctx* ctx, const float* d_Ug, const float* d_Fg, int64_t nX,
                             int64_t nY, double x0, double y0, double step,
                             const double* d_coords, int64_t n, float* d_U0, float* d_F0);

/* Device-resident integer search on FP32 planes d_f, d_g (M x N); outputs
 * are device arrays of nX*nY entries (nY x nX column-major).  The per-point
 * ZNCC lag maps live in a context-owned scratch buffer. */
int aldic_integer_search_dev(aldic_ctx* ctx, const float* d_f, const float* d_g, int64_t M,
                             int64_t N, const int64_t gridx[2], const int64_t gridy[2],
                             int32_t winsize, int32_t winstepsize, const int64_t region[4],
                             double* d_x, double* d_y, double* d_u, double* d_v,
                             double* d_ccmax, int32_t* d_ipeak_row, int32_t* d_ipeak_col);

/* aldic_integer_search_dev without any allocation: the per-point scratch
 * (clamps, mean-removed template, per-tile bests) lives in a file-scope
 * device array of the library, ALDIC_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends.  A grid whose scratch would exceed it is
 * ALDIC_EINVAL (aldic_last_error gives the size needed).  Outputs as
 * aldic_integer_search_dev.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (two contexts, two threads) race.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (work under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
#define ALDIC_NOALLOC_SCRATCH_BYTES (32 << 20)
int aldic_integer_search_noalloc_dev(aldic_ctx* ctx, const float* d_f, const float* d_g,
                                     int64_t M, int64_t N, const int64_t gridx[2],
                                     const int64_t gridy[2], int32_t winsize,
                                     int32_t winstepsize, const int64_t region[4], double* d_x,
                                     double* d_y, double* d_u, double* d_v, double* d_ccmax,
                                     int32_t* d_ipeak_row, int32_t* d_ipeak_col);

/* ---------------------------------------------------------------------------
 * Learned starting point (no counterpart in the reference: it stands beside
 * IntegerSearchQuadtree as the source of IC-GN's P0) and its fallback.
 * ------------------------------------------------------------------------- */
typedef struct aldic_cnn aldic_cnn;

/* A network with room for batches of up to max_batch (multiple of 16)
 * 48 x 48 patch pairs.  Parameters start at zero: load them with
 * aldic_cnn_set_params. */
int aldic_cnn_create(aldic_ctx* ctx, int32_t max_batch, aldic_cnn** out);
int aldic_cnn_destroy(aldic_cnn* net);
/* FP32 parameter vector: per layer W [cout][Kp] (K padded to 16) then bias.
 * The layout is fixed by the architecture: net may be NULL. */
int aldic_cnn_num_params(const aldic_cnn* net, int64_t* n);
/* info = {w_off, b_off, cout, K, Kp, cin, stride, input side, is_conv};
 * net may be NULL. */
int aldic_cnn_layout(const aldic_cnn* net, int32_t layer, int64_t info[9]);
/* Output units: outputs are [u_x v_x u_y v_y]/sf and [du dv]/su. */
int aldic_cnn_scales(double* sf, double* su);
int aldic_cnn_set_params(aldic_cnn* net, const float* host, int64_t n);  /* resets Adam */
int aldic_cnn_get_params(aldic_cnn* net, float* host, int64_t n);
int aldic_cnn_get_grads(aldic_cnn* net, float* host, int64_t n);
int aldic_cnn_set_loss_weights(aldic_cnn* net, const float w[6]);
/* d_x: B x 48 x 48 x 2 bf16 (NHWC); d_out: B x 6 FP32 (may be NULL). */
int aldic_cnn_forward_dev(aldic_cnn* net, const void* d_x, int64_t B, float* d_out);
/* One step of mean-squared error on normalised labels (B x 6 FP32):
 * forward, backward into the gradient buffer and, when update != 0, Adam
 * with rate lr.  d_loss (device float, may be NULL) receives the loss. */
int aldic_cnn_train_step_dev(aldic_cnn* net, const void* d_x, const float* d_labels, int64_t B,
                             float lr, int32_t update, float* d_loss);
/* Patch pairs of the nodes: reference at the rounded node, deformed at the
 * node + round(d_prior) (2n floats, NULL = 0), each normalised. */
int aldic_cnn_patches_dev(aldic_ctx* ctx, const aldic_images* im, const double* d_coords,
                          int64_t n, const float* d_prior, void* d_out);
/* Patches + network + output scaling: d_U0 = round(prior) + du, d_F0 = F. */
int aldic_cnn_predict_dev(aldic_cnn* net, const aldic_images* im, const double* d_coords,
                          int64_t n, const float* d_prior, float* d_U0, float* d_F0);
/* Training images: seeded speckle mosaic of tile x tile tiles, each tile of
 * the deformed image under its own affine displacement about its centre
 * node (d_warps: 6 floats [u_x v_x u_y v_y u v] per tile, tiles along dim 1
 * first; NULL = reference image). */
int aldic_speckle_mosaic_dev(aldic_ctx* ctx, int64_t M, int64_t N, int32_t tile,
                             const float* d_warps, uint32_t seed, uint32_t noise_seed,
                             double density, double radius, double scale, double noise,
                             double top, float* d_img);
/* Gate: nodes with convIt outside [1, max_iter] or ZNCC below min_zncc
 * (d_zncc may be NULL) are listed in d_list; the count goes to d_count. */
int aldic_gate_dev(aldic_ctx* ctx, int64_t n, const int32_t* d_convIt, const float* d_zncc,
                   int32_t max_iter, double min_zncc, int32_t* d_list, int32_t* d_count);
/* Integer search (funIntegerSearch.m:139-203 per node, the whole image as
 * ROI) of the template (winsize+1)^2 centred on each listed node; writes
 * U0 (and F0 = 0, d_F0 may be NULL) of those nodes.  At most max_count
 * entries; the count is read on the device. */
int aldic_search_nodes_dev(aldic_ctx* ctx, const aldic_images* im, const double* d_coords,
                           int64_t n, const int32_t* d_list, const int32_t* d_count,
                           int64_t max_count, int32_t winsize, const int64_t region[4],
                           float* d_U0, float* d_F0);
/* The search-init starting point of the listed nodes only: the points of
 * the funIntegerSearch grid (gridx, gridy, winsize, winstepsize, region as
 * aldic_integer_search_dev) that aldic_grid_to_nodes_dev would read for
 * them are searched, and U0 (F0 = 0, d_F0 may be NULL) of the listed
 * nodes is interpolated from them -- the same U0 the whole-grid path gives
 * those nodes.  d_f, d_g are the FP32 planes the search runs on. */
int aldic_search_grid_nodes_dev(aldic_ctx* ctx, const float* d_f, const float* d_g, int64_t M,
                                int64_t N, const int64_t gridx[2], const int64_t gridy[2],
                                int32_t winsize, int32_t winstepsize, const int64_t region[4],
                                const double* d_coords, int64_t n, const int32_t* d_list,
                                const int32_t* d_count, float* d_U0, float* d_F0);
/* IC-GN (funICGNQuadtree.m) of the listed nodes only, from [F0 U0]: list
 * entries are distinct node indices of the plan; other nodes' outputs are
 * left as they are. */
int aldic_local_icgn_list_dev(aldic_ctx* ctx, const aldic_images* im, const aldic_plan* plan,
                              const double* d_coords, const float* d_U0, const float* d_F0,
                              const int32_t* d_winsize, const int32_t* d_list,
                              const int32_t* d_count, double tol, float* d_U, float* d_F,
                              float* d_H21, int32_t* d_convIt, uint8_t* d_hole, float* d_zncc);

/* ---- the global step of the ALM loop ---- */

/* The FE Q4 mesh of Subpb2Quadtree (DICmesh): coords n x 2 (coordinatesFEM,
 * column-major), elements ne x 8 (elementsFEM, column-major, 1-based; 0 in
 * columns 5..8 = no hanging node on edge 2-3, 3-4, 4-1, 1-2), dirichlet nd
 * 1-based nodes (DICmesh.dirichlet), neumann nn x 4 (node1, node2, n_x, n_y;
 * DICmesh.neumann, NULL when empty), GaussPtOrder 2..5.  The sparsity of the
 * global matrix and the order of its assembly sums are fixed here. */
typedef struct aldic_fem aldic_fem;
int aldic_fem_create(aldic_ctx* ctx, const double* coords, int64_t n, const int32_t* elements,
                     int64_t ne, const int32_t* dirichlet, int64_t nd, const double* neumann,
                     int64_t nn, int32_t gauss_order, aldic_fem** out);
int aldic_fem_destroy(aldic_fem* fem);
/* Nonzeros of the scalar matrix, free nodes, Dirichlet nodes. */
int aldic_fem_info(const aldic_fem* fem, int64_t* nnz, int64_t* nfree, int64_t* ndir);
/* Uhat = Subpb2Quadtree(DICmesh, GaussPtOrder, beta, mu, U, F, udual, vdual,
 * alpha, winstepsize, 0) (func_quadtree/Subpb2Quadtree.m:1-397): U (2n), F
 * (4n) and Uhat (2n) in FP64.  udual and vdual are not arguments: the
 * reference multiplies both by zero (:42-44).  Nodes in no element (and
 * the first node of unique(elementsFEM), the reference's :314-318) get 0,
 * Dirichlet nodes get U, the rest the solution of the free system, here by
 * Jacobi-preconditioned conjugate gradients to a relative residual of 1e-13
 * per component (iterations, may be NULL, receives the count; max_iter <= 0
 * picks 2n).  One call at a time per mesh. */
int aldic_subpb2(aldic_ctx* ctx, aldic_fem* fem, double beta, double mu, double alpha,
                 double winstepsize, const double* U, const double* F, double* Uhat,
                 int32_t* iterations);
int aldic_subpb2_dev(aldic_ctx* ctx, aldic_fem* fem, double beta, double mu, double alpha,
                     double winstepsize, const double* d_U, const double* d_F, double* d_Uhat,
                     int32_t max_iter, int32_t* iterations);
/* The same solve with no host synchronisation, so that it can be recorded
 * in a graph with the local step around it (the ALM iteration of
 * TemporalMatch_quadtree_ST1.m:337-380): all ceil(max_iter / 16) chunks of
 * 16 CG iterations the limit allows are launched (1 <= max_iter <= 65536;
 * the chunks after convergence return at once), and the iteration count is
 * written to the device int d_iterations (may be NULL) when the solve
 * ends.  No counterpart in the reference (its solve is a host `\`). */
int aldic_subpb2_dev_async(aldic_ctx* ctx, aldic_fem* fem, double beta, double mu, double alpha,
                           double winstepsize, const double* d_U, const double* d_F,
                           double* d_Uhat, int32_t max_iter, int32_t* d_iterations);

/* One iteration of the augmented-Lagrangian loop of
 * TemporalMatch_quadtree_ST1.m:337-387 on the device, with no host
 * synchronisation (the iterations of a frame can be recorded in one
 * graph):
 *   USubpb1 = Subpb1(UOld = USubpb2, FOld = FSubpb2 = d_F1, vdual) on the
 *     plan's nodes (rect plan, winsize_List in d_winsize_xy), then the bad
 *     points of Subpb1Quadtree.m:167-182 (ConvIt < 0 or > 99 or > max(mean
 *     + 0.25 std, 10)) -- where the reference fills them by natural-
 *     neighbour interpolation (host code, aldic_nan_fill), this step gives
 *     a bad or thrown node the global estimate USubpb2 it started from;
 *   USubpb2 = Subpb2(USubpb1, d_F1) on the mesh (Jacobi-PCG, at most
 *     cg_max_iter iterations), hole-edge nodes (d_hole nonzero, may be
 *     NULL) keep USubpb1 (:359);
 *   when check_update: UpdateY = ||USubpb2_prev - USubpb2|| / sqrt(2n) and
 *     UpdateY2 likewise for USubpb1 into d_update[0..1] (:363-366);
 *   vdual = USubpb2 - USubpb1 (:379); when check_update and either norm is
 *     below tol2 the loop is done (:381-384) and later iterations on the
 *     same d_state do nothing.
 * In/out (FP64 device arrays): d_U1 (USubpb1, 2n: on input the previous
 * USubpb1, e.g. LocalICGN's), d_U2 (USubpb2, 2n), d_vdual (2n).  d_state:
 * int32[2] = {done, bad Subpb1 nodes of the last executed iteration};
 * first != 0 clears done and sets d_update to {1e4, NaN}.  The plan and
 * the mesh must hold the same nodes.  No counterpart in the reference as
 * a single call; it replaces the body of its ALM while loop. */
int aldic_alm_iteration_dev(aldic_ctx* ctx, const aldic_images* im, const aldic_plan* plan,
                            aldic_fem* fem, const double* d_coords, const int32_t* d_winsize_xy,
                            double mu, double beta, double alpha, double winstepsize, double tol,
                            double tol2, int32_t check_update, int32_t first, const double* d_F1,
                            const uint8_t* d_hole, double* d_U1, double* d_U2, double* d_vdual,
                            double* d_update, int32_t* d_state, int32_t cg_max_iter);

/* ---- host post-processing of the drivers (no device work) ---- */

/* scatteredInterpolant(xs, ys, vals, 'natural', 'nearest') evaluated at
 * (xq, yq): Sibson natural-neighbour interpolation inside the convex hull of
 * the sites, linear along its edges, the nearest site's value outside (the
 * first site in input order on a tie).  vals is n x nv column-major (nv
 * value columns over one set of sites, e.g. u and v), out is nq x nv.
 * Duplicate sites are merged with their values averaged.  Fewer than three
 * distinct sites, or collinear sites, is ALDIC_EINVAL (the reference's
 * interpolant construction fails there).  Replaces the scatteredInterpolant
 * calls of LocalICGNQuadtree.m:235-241, :302-305 and
 * Subpb1Quadtree.m:221-227, :239-242. */
int aldic_natural_interp(const double* xs, const double* ys, const double* vals, int64_t n,
                         int32_t nv, const double* xq, const double* yq, int64_t nq, double* out);
/* The NaN fill that follows the bad-point rules of the drivers, in place on
 * U (2n, interleaved; NaN marks a bad node).  Pass 1, when a node is NaN
 * and mask (M x N, column-major, nonzero = inside) is given: per
 * 8-connected mask region, the NaN nodes whose rounded position lies in it
 * are interpolated from the region's valid nodes when there are at least
 * three (a region whose sites are degenerate is skipped, the reference's
 * catch).  Pass 2, when a node is still NaN or (F given, 4n) some F(4k+1)
 * is NaN: U at every node becomes the interpolant of all valid nodes.
 * LocalICGNQuadtree.m:200-247 + :290-325 (F given, mask = Df.ImgRefMask);
 * Subpb1Quadtree.m:189-254 (F = NULL, mask = DICpara.ImgRefMask).
 * regions_filled (may be NULL) receives the number of regions pass 1
 * interpolated in. */
int aldic_nan_fill(const double* coords, int64_t n, double* U, const double* F,
                   const uint8_t* mask, int64_t M, int64_t N, int32_t* regions_filled);

#ifdef __cplusplus
}
#endif
#endif
