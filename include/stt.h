/*
 * stt.h — C-ABI of the MI355X space-time tomography hot path.
 *
 * The reference program (gmzang/SpaceTimeTomography) exposes this path only as
 * in-process C++ classes (ST_Tomography, TypicalRayProducer, RayCaster,
 * VolumeData, warping::warp).  Each entry point below replaces one of those
 * members or free functions; the citation next to it names the reference code
 * whose behaviour it reproduces.  Plain pointers and sizes only: no torch or
 * HIP types cross this boundary.
 *
 * Conventions
 *   - Every function returns an int status (STT_OK on success).  On failure
 *     stt_last_error() returns a message for the calling thread.  No C++
 *     exception crosses the ABI.
 *   - Volumes are planar fp32, x fastest, then y, then z (CImg layout,
 *     CImg.h:11159-11161).  A flow field is three such planes (u, v, w) in
 *     voxel units.  Projections are W x H x P fp32, pixel (i, j) of view p at
 *     i + j*W + p*W*H (the frame's slice of the projection stack,
 *     ST_Tomography.cpp:355).
 *   - Host pointers are copied in or out; "_dev" variants take device pointers
 *     that the library borrows and never frees.
 *   - One context per thread.  All device work of a context is issued on the
 *     context's own HIP streams (its stream, plus seven frame-group streams
 *     that the batched iteration forks onto and joins before it returns);
 *     functions that return data synchronise them.
 */
#ifndef STT_H
#define STT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STT_OK            0
#define STT_ERR_INVALID   1  /* null pointer, out-of-range index, size mismatch */
#define STT_ERR_HIP       2  /* HIP runtime error */
#define STT_ERR_NOMEM     3  /* device or host allocation failed */
#define STT_ERR_STATE     4  /* call out of order (e.g. no views or projections set) */
#define STT_ERR_NODEVICE  5  /* no HIP device visible */

/* Backprojector (ST_Tomography.cpp:221-226, main.cpp:43 "--bp"). */
#define STT_BP_VOXEL      0  /* voxel-driven, bicubic detector sampling (ST_Tomography.cpp:687-731) */
#define STT_BP_RAY        1  /* ray-driven scatter + normalise (ST_Tomography.cpp:733-765) */

/* Spatial prior (main.cpp:42 "--prior").  The reference only implements ATV
 * (per-component clamp, ST_Tomography.cpp:523-525).  STV is this library's
 * isotropic spatio-temporal TV (4-component gradient incl. the forward
 * temporal difference, projected onto the lambda L2 ball). */
#define STT_PRIOR_ATV     0
#define STT_PRIOR_TV      1  /* isotropic spatial TV: 3-component L2-ball projection */
#define STT_PRIOR_STV     2

/* Interpolation of the semi-Lagrangian warp (warping.h:49-88). */
#define STT_INTERP_CUBIC_CLAMPED 0  /* CImg Catmull-Rom clamped to the 8-corner min/max (reference) */
#define STT_INTERP_TRILINEAR     1  /* trilinear, same trace; differentiable path */

/* Scan and volume geometry: the geometric part of the ST_Tomography
 * constructor parameter set (ST_Tomography.h:27-29, main.cpp:26-37). */
typedef struct stt_geometry {
    int    vol[3];      /* volume size in voxels (main.cpp -f, "volxyz"), each 4..16384 */
    float  spacing[3];  /* voxel spacing in mm (main.cpp -c; VolumeData::m_Spacings) */
    int    det[2];      /* detector W, H in pixels (main.cpp -u, "imgwh") */
    double ds;          /* detector pitch in mm (main.cpp -d) */
    double sid;         /* source to iso-centre distance in mm (main.cpp -g) */
    double sdd;         /* source to detector distance in mm (main.cpp -j) */
    double oxyz[3];     /* volume-centre offset in mm (main.cpp --oxyz; ST_Tomography.cpp:244-246) */
} stt_geometry;

/* Solver parameters: the numerical part of the constructor parameter set.
 * stt_params_reference() fills the values the reference actually runs with
 * (its hard-coded globals, ST_Tomography.cpp:93-114, 219). */
typedef struct stt_params {
    double sigma;        /* primal-dual sigma (CP_sigma, ST_Tomography.cpp:102) */
    double tau;          /* primal-dual tau (CP_tau, :100); SART factor = sqrt(tau) (:110-111) */
    double lambda_tv;    /* TV weight (TV_w, main.cpp -l) */
    double k_tp;         /* temporal-prior weight (TP_w, main.cpp -k) */
    double e_bc;         /* brightness-constancy weight (BC_w, main.cpp -e) */
    double huber;        /* huber epsilon (main.cpp --huber; m_huberfactor :219) */
    double alpha;        /* SART step (main.cpp -a) */
    float  g_max;        /* upper clamp; <= 0 means max(frame-0 projections)/4 (:356-357) */
    int    prior;        /* STT_PRIOR_* */
    int    bp;           /* STT_BP_* */
    int    interp;       /* STT_INTERP_* for the PD warps */
    int    roi_cut[8];   /* region of interest cut[] (:114); all zero = whole interior */
    int    roi_jtop;     /* the constant 7 in imROI's j test (:324); 0 = none */
    int    fp64_reductions; /* 1: residual norms and G_Max reduction in FP64 */
} stt_params;

/* Per-kernel timing of the most recent stt_iterate (HIP events on the
 * context stream).  Times in milliseconds, summed over launches. */
typedef struct stt_timing {
    double total_ms;
    double fp_ms;        int fp_launches;
    double bp_ms;        int bp_launches;
    double dual_ms;      int dual_launches;
    double primal_ms;    int primal_launches;
    double aux_ms;       int aux_launches;   /* layout copies feeding the projector */
    /* Preparation before the iterations: re-tracing the flows written since
     * the last PD step (trace_launches of them; once per flow change, not per
     * iteration), plus any rebuild of the launch tables. */
    double trace_ms;     int trace_launches;
} stt_timing;

typedef struct stt_ctx stt_ctx;

/* ---- library / errors --------------------------------------------------- */
const char* stt_last_error(void);
int  stt_version(void);
/* Fills p with the reference's effective values (sigma=0.1125, tau=1,
 * huber=0.1, alpha=0.1, lambda=0.1, k=e=0 from main.cpp:67-95 and
 * ST_Tomography.cpp:99-114, rose ROI cut[]).  Pure host function. */
void stt_params_reference(stt_params* p);
/* Same but with the whole-interior ROI (cut = 0, jtop = 0). */
void stt_params_default(stt_params* p);

/* ---- context (ST_Tomography ctor + SetData, ST_Tomography.cpp:200-361) -- */
/* nframes volumes plus one padding copy of the last frame (VolumeData.cpp:126);
 * views_per_frame[f] views for frame f (uneven counts allowed). */
int  stt_ctx_create(const stt_geometry* g, const stt_params* p, int nframes,
                    const int* views_per_frame, stt_ctx** out);
int  stt_ctx_destroy(stt_ctx* ctx);
int  stt_ctx_set_params(stt_ctx* ctx, const stt_params* p);
int  stt_ctx_get_gmax(stt_ctx* ctx, float* g_max);

/* View tables.  Angles in degrees about +y (InitViews, ST_Tomography.cpp:887-934):
 * each view is Quaternion(0,1,0, theta*3.1415926/180). */
int  stt_set_views(stt_ctx* ctx, int frame, const float* angles_deg, int n);
/* Explicit rotations, x,y,z,w per view (pQuaternion[frame][view]). */
int  stt_set_view_quaternions(stt_ctx* ctx, int frame, const float* q_xyzw, int n);

/* Explicit camera per view, as TypicalRayProducer::SetViewMatrix /
 * SetProjectionMatrix take it (TypicalRayProducer.h:22-23): column-major 4x4
 * modelview (View = T * S * R, Scene.cpp:181-207) and projection (frustum,
 * Scene.cpp:69-138) matrices, n of each; the per-view record then follows
 * TypicalRayProducer::_calculateMatrix (TypicalRayProducer.cpp:142-208). */
int  stt_set_view_matrices(stt_ctx* ctx, int frame, const float* view16, const float* proj16, int n);
/* The Scene's modelview and projection matrices for a rotation (x,y,z,w) in
 * this geometry (camera of ST_Tomography.cpp:229-246).  Host only. */
int  stt_scene_matrices(const stt_geometry* g, const float* q_xyzw, float* view16, float* proj16);

/* Data (VolumeData::m_Volumelist / m_Flowlist, Projslist). frame in [0, nframes];
 * frame == nframes is the padding copy. */
int  stt_set_volume(stt_ctx* ctx, int frame, const float* vol);
int  stt_get_volume(stt_ctx* ctx, int frame, float* vol);
/* pair in [0, nframes-1].  The context keeps the end points of the PD warps'
 * traces of each flow (24 bytes per voxel and frame, allocated on the first
 * PD step; without that memory the kernels trace in place).  Every write of a
 * flow -- stt_set_flow, stt_set_buffer(STT_BUF_FLOW), stt_optical_flow -- and
 * every change of p.interp makes the next PD step re-trace it. */
int  stt_set_flow(stt_ctx* ctx, int pair, const float* uvw);
int  stt_set_projections(stt_ctx* ctx, int frame, const float* b); /* W*H*views[frame] */
/* Device-pointer variants (borrowed, copied device-to-device, never freed):
 * the same data entry points for a caller whose buffers already live on the
 * device (a torch tensor's data_ptr(), a buffer of its own).  Same sizes,
 * ranges and errors as the host forms; stt_set_projections_dev derives
 * G_Max from frame 0 as stt_set_projections does. */
int  stt_set_volume_dev(stt_ctx* ctx, int frame, const float* vol_dev);
int  stt_get_volume_dev(stt_ctx* ctx, int frame, float* vol_dev);
int  stt_set_flow_dev(stt_ctx* ctx, int pair, const float* uvw_dev);
int  stt_set_projections_dev(stt_ctx* ctx, int frame, const float* b_dev);

/* Dual / auxiliary state readout (X_bar, X_prev, Y_tp, Y_bc: one volume each;
 * Y_tv: four planes, the three spatial components then STV's temporal
 * component, which stays zero for the other priors; FLOW: the three planes
 * of flow pair `frame`, 0 <= frame < nframes). */
#define STT_BUF_XBAR  0
#define STT_BUF_XPREV 1
#define STT_BUF_YTV   2
#define STT_BUF_YTP   3
#define STT_BUF_YBC   4
#define STT_BUF_FLOW  5
int  stt_get_buffer(stt_ctx* ctx, int which, int frame, float* out);
int  stt_set_buffer(stt_ctx* ctx, int which, int frame, const float* in);

/* ---- ImageFormationModel operators -------------------------------------- */
/* Forward projection of view `view` of frame `frame` from the current volume
 * (GenerateImage_GetCorrections -> GenerateImage_MT -> ProduceRay ->
 * CastingRay, TypicalRayProducer.core.h:32-69,123-263, RayCaster.core.h:42-187).
 * Ax: line integrals; raylen: tfar-tnear for hit rays, left unchanged for
 * missed rays (the reference keeps the previous value); weight: sample
 * counts.  Any output pointer may be NULL.  raylen is read and written. */
int  stt_forward_project(stt_ctx* ctx, int frame, int view, float* Ax,
                         float* raylen, float* weight);
/* Voxel-driven backprojection of a correction image (ST_Tomography.cpp:687-731):
 * x += alpha*bicubic(corr), clamp [0, G_Max]. */
int  stt_backproject_voxel(stt_ctx* ctx, int frame, int view, const float* corr,
                           double alpha);
/* Ray-driven backprojection (ST_Tomography.cpp:733-765 with the scatter of
 * SingleScalarVolume.core.h:128-143): x += contrib/wsum, clamp (0/0 -> 0).
 * The sums are exact 64-bit fixed-point accumulations of the reference's
 * fp32 addends (identical on every run), divided in FP64.  contrib/wsum may
 * be NULL; when given they receive the sums rounded to fp32.  In a sweep or
 * iteration the ray lengths are reset after every ray-BP view, as the
 * reference's TypicalRayProducer::Initial() does. */
int  stt_backproject_ray(stt_ctx* ctx, int frame, int view, const float* corr,
                         float* contrib, float* wsum);
/* Matched transpose of stt_forward_project: out += A^T y (the exact adjoint of
 * the sampled line integral including the sd factor).  VJP of the projector
 * with respect to the volume. */
int  stt_fp_vjp(stt_ctx* ctx, int frame, int view, const float* y, float* out);
/* Device-pointer forms of the three operators above (GenerateImage_MT and the
 * voxel BP of ST_Tomography.cpp:687-731, the projector's transpose):
 * the caller's device buffers are read and written in place of host arrays,
 * with the same semantics (raylen read and written; out accumulated). */
int  stt_forward_project_dev(stt_ctx* ctx, int frame, int view, float* Ax_dev,
                             float* raylen_dev, float* weight_dev);
int  stt_backproject_voxel_dev(stt_ctx* ctx, int frame, int view,
                               const float* corr_dev, double alpha);
int  stt_fp_vjp_dev(stt_ctx* ctx, int frame, int view, const float* y_dev,
                    float* out_dev);

/* One SART sweep of frame `frame` (ST_Tomography::SART, :616-808):
 * views in order[0..n) (n <= views[frame]; NULL = identity order), then
 * x *= imROI. */
int  stt_sart_sweep(stt_ctx* ctx, int frame, const int* order, int n);

/* ---- ST_Tomo proximal steps --------------------------------------------- */
/* Semi-Lagrangian warp warping::warp(dim, aabb, vel, src, dt) (warping.h:116-145)
 * on the context's volume grid; host pointers; aabb = [0, dims]. */
int  stt_warp(stt_ctx* ctx, const float* uvw, const float* src, float dt,
              float* out, int interp);
/* The same warp on the context's own buffers: out (host, V floats) =
 * warp(flow[pair], volume[pair + 1], dt).  Render dumps this after each flow
 * solve, I2 warped back by its flow with dt = -1 (ST_Tomography.cpp:462);
 * here the flow and the frame stay on the device.  pair in [0, nframes-2]. */
int  stt_warp_frame(stt_ctx* ctx, int pair, float dt, int interp, float* out);
/* VJPs of the trilinear warp (stt_warp with STT_INTERP_TRILINEAR): given
 * g = dL/d(out), accumulate dL/d(src) into gsrc and dL/d(uvw) into guvw
 * (host arrays, either may be NULL).  The clamps of the sampling grid to
 * [1, n-2] have zero derivative; the first velocity sample is the node value
 * of u at clamp(i, 1, n-2). */
int  stt_warp_vjp(stt_ctx* ctx, const float* uvw, const float* src, float dt,
                  const float* g, float* gsrc, float* guvw);
/* Device-pointer forms of stt_warp (warping.h:116-145) and stt_warp_vjp. */
int  stt_warp_dev(stt_ctx* ctx, const float* uvw_dev, const float* src_dev,
                  float dt, float* out_dev, int interp);
int  stt_warp_vjp_dev(stt_ctx* ctx, const float* uvw_dev, const float* src_dev,
                      float dt, const float* g_dev, float* gsrc_dev,
                      float* guvw_dev);

/* Initialise X_bar <- x for all frames and zero the duals
 * (ST_Tomography.cpp:494-500 first pass, SetData :339-350). */
int  stt_pd_init(stt_ctx* ctx);
/* Dual step of one frame (ST_Tomography.cpp:502-534). */
int  stt_pd_dual(stt_ctx* ctx, int frame);
/* Primal step of one frame (ST_Tomography.cpp:536-578). */
int  stt_pd_primal(stt_ctx* ctx, int frame);
/* X_bar = 2x - X_prev on the ROI (ST_Tomography.cpp:586-604). */
int  stt_pd_extrapolate(stt_ctx* ctx, int frame);
/* Memory knob with no counterpart in the reference (which re-traces in every
 * warp, warping.h:100-114): keep = 1 (the default) stores the PD traces' end
 * points, 24 bytes per voxel and frame, allocated on the next PD step;
 * keep = 0 frees them and the PD kernels trace in place (the same float end
 * points, so bit-identical results, at the cost of the velocity sampling in
 * every PD step).  At 512^3 x 16 frames the stored traces take 51.5 GB. */
int  stt_set_trace_storage(stt_ctx* ctx, int keep);

/* One or more ST iterations (one "No_PO" pass over all frames,
 * ST_Tomography.cpp:477-608): per frame dual, primal, SART sweep over all of
 * the frame's views in view-table order, extrapolation.
 *   batched = 0: the reference's per-frame order, one frame per launch.
 *   batched = 1: all duals, then all primals, then every frame's view k in one
 *                launch (k = 0..max views); identical results.
 *   use_graph = 1: capture one iteration in a hipGraph and replay it.
 * timing may be NULL; when given, kernels are bracketed with HIP events. */
int  stt_iterate(stt_ctx* ctx, int n_iters, int batched, int use_graph,
                 stt_timing* timing);
/* Groups of contexts.  Frames of different grids cannot share a context
 * (the PD steps couple neighbouring frames of one grid), so independent
 * solves of mixed volume sizes -- 128^3 and 192^3 frames with uneven view
 * counts -- take one context per size.  A group runs their ST iterations as
 * one unit: each context's batched iteration on its own streams, forked from
 * and joined into the first context's stream, and with use_graph = 1 all of
 * it captured in one hipGraph launched once per iteration.  Each context's
 * results are those of stt_iterate(ctx, n, 1, ...).  No counterpart in the
 * reference, which solves one grid.  The contexts must share a device and
 * outlive the group; one group per thread.  total_ms (may be NULL) receives
 * the time of the n_iters iterations (HIP events). */
typedef struct stt_group stt_group;
int  stt_group_create(stt_ctx* const* ctxs, int n, stt_group** out);
int  stt_group_destroy(stt_group* grp);
int  stt_group_iterate(stt_group* grp, int n_iters, int use_graph, double* total_ms);
/* Projector shape of this context (other contexts keep theirs; the context's
 * captured graph is rebuilt): 0 = slab projector (16x16 rays per block,
 * volume boxes staged in LDS; default); -32 or -8 = slab projector with 32x8
 * or 8x32 rays per block; 64, 32, 16 or 8 = level-synchronous projector with
 * that many rays per detector row of a 64-lane wave (these keep an x<->z
 * transposed copy of every frame volume while selected).  Sample counts and
 * ray lengths do not depend on the shape. */
#define STT_PROJECTOR_SHAPES 7
int  stt_set_projector_tile(stt_ctx* ctx, int rays_per_row);
/* Times every projector shape on the launches a SART sweep issues (for the
 * first view indices k: the layout copy and the projector over every frame's
 * view k, corrections and ray lengths into scratch buffers, so no solver
 * state changes) and keeps the fastest; *rays_per_row (may be NULL) receives
 * the choice and ms_per_shape[STT_PROJECTOR_SHAPES] (may be NULL) the time
 * per view index of shapes 0/-32/-8/64/32/16/8. */
int  stt_autotune(stt_ctx* ctx, int* rays_per_row, double* ms_per_shape);
/* Staged box ("brick") of the slab projector shapes, in KiB of LDS per
 * block: 16, 20, 24 (default) or 28.  A larger box lets a slab span more
 * planes (fewer staging rounds per ray) and leaves fewer blocks per CU.  The
 * box only decides where a tap is read from (LDS or global memory), never
 * its value, so projections are bit-identical for every size.  No
 * counterpart in the reference (its projector reads the volume directly);
 * the context's captured graph is rebuilt. */
#define STT_PROJECTOR_BOXES 4
int  stt_set_projector_box(stt_ctx* ctx, int kib);
/* Times the box sizes 16/20/24/28 KiB with the context's current slab shape
 * on the same launches as stt_autotune and keeps the fastest; *kib (may be
 * NULL) receives it and ms_per_box[STT_PROJECTOR_BOXES] (may be NULL) the
 * time per view index of each size.  STT_ERR_INVALID while a
 * level-synchronous shape is selected. */
int  stt_autotune_box(stt_ctx* ctx, int* kib, double* ms_per_box);

/* Measurement helper: the projector launches of one iteration (every
 * frame's views, layout copies included, SART mode) timed with HIP events,
 * averaged over reps.  graph_form = 0: the batched eager form, one launch per
 * view index over all frames; 1: the launch shape stt_iterate's hipGraph
 * uses (single-frame launches on the frame-group streams), captured and
 * replayed.  Corrections and ray lengths go to scratch buffers: no solver
 * state changes. */
int  stt_time_projector(stt_ctx* ctx, int graph_form, int reps, double* ms_per_iteration);

/* Blocks until the context stream is idle. */
int  stt_synchronize(stt_ctx* ctx);

/* FP64 reductions on the device, fixed summation order.
 * r2 = sum over frames f and views k of ||b_fk - A_fk x_f||^2 (one forward
 * projection per view into scratch; no solver state changes). */
int  stt_residual_norm(stt_ctx* ctx, double* r2);
/* Same *r2 as stt_residual_norm, bit for bit, without allocating: no
 * hipMalloc, hipFree or other allocating call.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid it).  Its
 * scratch -- one work item, one view's line integrals and ray lengths, the
 * FP64 partials and running sums -- is a file-scope device array of about
 * 8 MiB that each device keeps until the process ends, so detectors are
 * limited to 2^20 pixels (STT_ERR_INVALID beyond).  The context's device
 * must be current (STT_ERR_STATE otherwise).  Every call on a device shares
 * that device's array: calls that overlap on one device race (contexts on
 * different streams included), so serialise them. */
int  stt_residual_norm_noalloc(stt_ctx* ctx, double* r2);
/* Power iteration (iters steps from a fixed pseudo-random start) for the
 * norm of the PD coupling operator K over the context's F frames:
 * K v = (forward differences of each frame along x, y, z on the dual's
 * region -- the grid interior inside the ROI cut --, v[f+1] - v[f] over the
 * whole volume with v[F] = 0).  *L = sqrt(lambda_max(K^T K)) estimate. */
int  stt_power_iteration(stt_ctx* ctx, int iters, double* L);

/* ---- acquisition view tables (host only, no device needed) -------------- */
/* The reference's rounds of views (InitViews, ST_Tomography.cpp:887-934 with
 * RotateVolume :865-884): round j starts at start_deg[j] and each further
 * view rotates the previous one by 360/views_per_round degrees about +y,
 * composed in fp32 as the reference does.  Writes nrounds*views_per_round
 * quaternions (x,y,z,w), ready for stt_set_view_quaternions. */
int  stt_view_rounds(const float* start_deg, int nrounds, int views_per_round,
                     float* q_xyzw);
/* Scanning angles of Low_discrepency_squence(Np, Ns, p, Nz)
 * (low discrepancy sequence/Low_discrepency_squence.m:15-44): van der
 * Corput radical inverse in base p over ns circle segments, plus a 0-degree
 * projection before every nz-th entry.  Writes np angles in degrees. */
int  stt_lds_sequence(int np, int ns, int p, int nz, float* out_deg);

/* ---- optical flow between two frames (optical_flow.h) --------------------- */
/* Parameters of L1Huber_OF_MultiScale3D as ST_Tomography::Render calls it
 * (ST_Tomography.cpp:441-456); stt_of_params_reference() fills those values. */
typedef struct stt_of_params {
    float tol;      /* residual threshold of a warp's iterations (precision) */
    float lambda;   /* TV weight (smoothness) */
    float huber;    /* Huber epsilon (norm 4) */
    float eta;      /* pyramid scale per level */
    float sigma;    /* blur of both images before the pyramid */
    int max_iter;   /* iteration cap per warp */
    int norm;       /* 4 = Huber-smoothed TV, otherwise TV */
    int warps;      /* warps per level */
    int levels;     /* pyramid levels */
    int cut[6];     /* ROI margins at full resolution (x0 x1 y0 y1 z0 z1) */
} stt_of_params;
void stt_of_params_reference(stt_of_params* p);

/* One pyramid level of L1Huber_OF (optical_flow.h:83-623) on host buffers:
 * I1, I2 of dims[0]*dims[1]*dims[2] voxels (x fastest); V = 3 planes
 * (u, v, w) and Y = 10 dual planes (y11 y12 y13 y21 y22 y23 y31 y32 y33 y4),
 * both updated on the ROI (margins roi[6]) and kept elsewhere; ext = the
 * world extent of the grid (aabb [0, ext]; ext = dims at full resolution).
 * iters[w] and residual[w] (may be NULL) receive each warp's iteration count
 * and last convergence residual. */
int  stt_of_level(const int dims[3], const float ext[3], const float* I1,
                  const float* I2, float* V, float* Y, const stt_of_params* p,
                  const int roi[6], int* iters, double* residual);

/* L1Huber_OF_MultiScale3D (optical_flow.h:630-734) on host buffers: blur
 * both images, pyramid of p->levels levels (sizes int(dims * eta^k)),
 * coarse to fine, V upsampled between levels; Y starts at zero as Render
 * passes it.  V (3 planes, full resolution, voxel units) in/out.  iters
 * (levels * warps, coarsest level first; may be NULL) receives the
 * iteration counts.  Replaces the flow sub-optimisation call of Render
 * (ST_Tomography.cpp:452-455). */
int  stt_of_multiscale(const int dims[3], const float* I1, const float* I2,
                       float* V, const stt_of_params* p, int* iters);
/* The same on the context's device-resident frames: I1 = volume `pair`,
 * I2 = volume `pair + 1`, V = flow `pair` (updated in place). */
int  stt_optical_flow(stt_ctx* ctx, int pair, const stt_of_params* p,
                      int* iters);
/* CImg get_blur(sigma) (CImg.h:33282: order-0 Deriche along x, y, z, Neumann
 * boundary) and get_resize(new_dims, spectrum, 5) (CImg.h:27958: moving
 * average on shrinking axes, cubic on growing ones), on host buffers of
 * `spectrum` planes. */
int  stt_blur(const int dims[3], int spectrum, const float* img, float sigma,
              float* out);
int  stt_resize(const int dims[3], int spectrum, const float* img,
                const int new_dims[3], float* out);

/* ---- host-side synthetic inputs (no device needed) ----------------------- */
/* Additive ellipsoid phantom on an nx*ny*nz grid (seeded). */
int  stt_synth_phantom(int nx, int ny, int nz, int n_ellipsoids, uint64_t seed,
                       float* out);
/* The same phantom pulled back through the smooth deformation D(x,t) of
 * frame t (sum of 4 sinusoidal modes, max |D| <= amp voxels). */
int  stt_synth_phantom_frame(int nx, int ny, int nz, int n_ellipsoids,
                             uint64_t seed, uint64_t deform_seed, float amp,
                             int t, float* out);
/* Flow of frame pair (t, t+1): u = D(x, t+1) - D(x, t), three planes. */
int  stt_synth_flow(int nx, int ny, int nz, uint64_t deform_seed, float amp,
                    int t, float* uvw);
/* View tables (degrees): kind 0 = even over 360, 1 = golden angle
 * ((offset+k)*180*(3-sqrt 5)) mod 360, 2 = van der Corput base 2 scaled to 360. */
int  stt_synth_angles(int kind, int n, int offset, float* out);

/* ---- on-disk formats (host only, no device needed) ----------------------- */
/* The reference reads and writes its data through CImg (built with libtiff):
 * the projection stack and initial volumes with CImg(filename) -> load_tiff
 * (ST_Tomography.cpp:355, VolumeData.cpp:119-122), reconstructions with
 * save_tiff (:805), warped frames with save_tiff and flows with save_analyze
 * (:464-468).  These functions replace those calls.  Images are returned in
 * CImg's layout: img[x + W*(y + H*(z + D*c))], converted to fp32. */
typedef struct stt_image_info {
    int   width, height, depth, spectrum;  /* CImg dims x, y, z (TIFF pages), c */
    int   bits;           /* bits per sample on disk (first page) */
    int   sample_format;  /* 1 unsigned integer, 2 signed integer, 3 IEEE float */
    float voxel[3];       /* voxel size: Analyze pixdim[1..3]; TIFF 1/resolution or 1 */
} stt_image_info;
/* Dimensions of the image CImg::load(path) would return (CImg.h:46206-46225):
 * ".tif"/".tiff" -> TIFF, ".hdr"/".nii" (and ".img") -> Analyze 7.5/NIfTI-1,
 * extension matched case-insensitively. */
int  stt_image_query(const char* path, stt_image_info* info);
/* Reads the image into out (n floats, n >= W*H*D*C).
 * TIFF (CImg::load_tiff, CImg.h:47445-47765): every page becomes one z
 * slice; samples convert to fp32 by CImg's type rule (8 bit: uint8 if
 * SampleFormat = 1 else int8; 16 bit: uint16 / int16; 32 bit: uint32 /
 * int32 / float; 64-bit float is accepted as well).  Strips or tiles,
 * chunky or planar, either byte order, classic TIFF or BigTIFF; compression
 * none, PackBits or LZW, predictor none or horizontal.  Pages of different
 * sizes, palette/YCbCr/CMYK photometrics and bilevel images are rejected.
 * Analyze (CImg::_load_analyze, CImg.h:47829-47955): datatypes 2, 4, 8, 16,
 * 64 times the header's scale factor (0 read as 1), byte order from the
 * header size; ".hdr" pairs with ".img", ".nii" reads at vox_offset. */
int  stt_load_image(const char* path, float* out, size_t n);
/* CImg::save_tiff of a float image (CImg.h:51554-51644): one page per z
 * slice, c samples per pixel (chunky), 32-bit IEEE float, uncompressed,
 * rows per strip = max(8192 / (W*c*4), 1) (libtiff's default strip size),
 * tags Orientation 1, FillOrder 1, Photometric 1 (2 for c = 3 or 4),
 * Software "CImg"; BigTIFF when the data reach 2 GiB. */
int  stt_save_tiff(const char* path, const float* img, int w, int h, int d, int c);
/* CImg::save_analyze of a float image (CImg.h:51715-51790): 348-byte header
 * + ".img" data for ".hdr"/".img"/no extension (both names derived as CImg
 * derives them), one 352-byte-header file for ".nii"; datatype 16, pixdim
 * voxel[0..2] (NULL -> 1, 1, 1).  Dimensions must fit a short. */
int  stt_save_analyze(const char* path, const float* img, int w, int h, int d, int c,
                      const float* voxel);
/* The start angles of ST_Tomography's constructor (ST_Tomography.cpp:258-267):
 * whitespace-separated decimal numbers read frame-major into
 * deg[f*nrounds + r].  A missing file, a malformed or non-finite number or
 * fewer than nframes*nrounds values is an error; values after those are
 * ignored. */
int  stt_read_start_angles(const char* path, int nframes, int nrounds, float* deg);
/* Block sizes of CImg::get_split(axis, nparts) with nparts > 0
 * (CImg.h:31391-31420), used to cut the projection stack into frames
 * (ST_Tomography.cpp:355): sizes[0..nparts-1], summing to size. */
int  stt_split_sizes(int size, int nparts, int* sizes);

#ifdef __cplusplus
}
#endif
#endif /* STT_H */
