/*
 * stt_oracle.h — CPU oracle of the space-time tomography hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference program's
 * algorithm (gmzang/SpaceTimeTomography, C++/OpenMP) on the CPU, in fp32 with
 * the reference's expression order and in fp64.  It is the checker of the
 * HIP library and the timed CPU baseline; the shipped path never links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   geometry, ProduceRay, CastingRay, trilinear, ray-BP scatter, matrix
 *   inverse, quaternion, CImg cubic  -> oracle/_ref (the reference's own
 *   sources compiled here) and the committed fixtures in tests/golden/;
 *   voxel BP bicubic, SART correction, PD steps, extrapolation, ROI, warp
 *   trace -> closed forms, invariants and brute force (tests).
 */
#ifndef STT_ORACLE_H
#define STT_ORACLE_H
#include "../include/stt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-view matrices in fp32 exactly as the reference builds them:
 * out[0:16] P2M, [16:32] M2V, [32:48] MV, [48:64] V2M, [64:80] MV*V2M,
 * [80:84] EyePos, [84:88] M2V*EyePos, [88] sample distance. */
int o_view_matrices(const stt_geometry* g, const float* q_xyzw, float* out);

/* Quaternion(0,1,0, theta*3.1415926/180) (ST_Tomography.cpp:899-911). */
void o_quat_y(float theta_deg, float* q_xyzw);

/* Forward projection of one view (GenerateImage_MT).  raylen is in/out
 * (missed rays keep their value).  Any output may be NULL. */
int o_fp_f32(const stt_geometry* g, const float* q, const float* vol,
             float* Ax, float* raylen, float* weight);
int o_fp_f64(const stt_geometry* g, const float* q, const double* vol,
             double* Ax, double* raylen, double* weight);
/* Ray-driven scatter of corr (BP mode of GenerateImage_MT), sequential so the
 * summation order is fixed: contrib += w*corr[pix], wsum += w. */
int o_bp_ray_scatter_f32(const stt_geometry* g, const float* q, const float* corr,
                         float* contrib, float* wsum);
int o_bp_ray_scatter_f64(const stt_geometry* g, const float* q, const double* corr,
                         double* contrib, double* wsum);
/* Ray-BP normalisation x += contrib/wsum, clamp [0,gmax] (ST_Tomography.cpp:746-762). */
int o_bp_ray_apply_f32(int n, const float* contrib, const float* wsum, float gmax,
                       float* vol);
/* Voxel-driven BP (ST_Tomography.cpp:687-731). */
int o_bp_voxel_f32(const stt_geometry* g, const float* q, const float* corr,
                   double alpha, float gmax, float* vol);
int o_bp_voxel_f64(const stt_geometry* g, const float* q, const double* corr,
                   double alpha, double gmax, double* vol);
/* SART correction of one view (ST_Tomography.cpp:667-676, pY == 0). */
int o_correction_f32(int n, float factor, const float* b, const float* Ax,
                     const float* raylen, float* corr);

/* CImg _cubic_atXYZ (CImg.h:13429-13520) on a w*h*d fp32 image. */
float o_cubic_atxyz_f32(const float* img, int w, int h, int d, float fx, float fy, float fz);
/* warping::warp (warping.h:116-145). dims = {nx, ny, nz}; aabb = [0, dims]. */
int o_warp_f32(const int* dims, const float* uvw, const float* src, float dt,
               float* out, int interp, int nthreads);
int o_warp_f64(const int* dims, const double* uvw, const double* src, double dt,
               double* out, int interp, int nthreads);

/* imROI of SetData (ST_Tomography.cpp:303-329) as 0/1 floats. */
int o_roi_mask(const int* dims, const int* cut, int jtop, float* out);

/* ---- full solver state ---------------------------------------------------- */
typedef struct o_ctx o_ctx;
/* f64 = 1: every array and all arithmetic in double (the FP64 oracle). */
o_ctx* o_ctx_create(const stt_geometry* g, const stt_params* p, int nframes,
                    const int* views_per_frame, int f64);
void   o_ctx_destroy(o_ctx* c);
int    o_ctx_set_view_quaternions(o_ctx* c, int frame, const float* q, int n);
int    o_ctx_set_volume(o_ctx* c, int frame, const float* vol);
int    o_ctx_get_volume(o_ctx* c, int frame, float* vol);
int    o_ctx_get_volume_f64(o_ctx* c, int frame, double* vol);
int    o_ctx_set_flow(o_ctx* c, int pair, const float* uvw);
int    o_ctx_set_projections(o_ctx* c, int frame, const float* b);
/* which: STT_BUF_* */
int    o_ctx_get_buffer(o_ctx* c, int which, int frame, float* out);
int    o_ctx_set_buffer(o_ctx* c, int which, int frame, const float* in);
float  o_ctx_gmax(o_ctx* c);
/* The shared detector raylen buffer (one TypicalRayProducer for all frames). */
int    o_ctx_get_raylen(o_ctx* c, float* out);
int    o_ctx_sart(o_ctx* c, int frame, const int* order, int n);
int    o_ctx_pd_init(o_ctx* c);
int    o_ctx_pd_dual(o_ctx* c, int frame);
int    o_ctx_pd_primal(o_ctx* c, int frame);
int    o_ctx_pd_extrapolate(o_ctx* c, int frame);
/* n ST iterations in the reference's sequential frame order.
 * faithful = 1 reproduces the reference's cost structure: two full volume
 * copies per view (SingleScalarVolume.core.h:41-44,218) and a 4-thread warp
 * (warping.h:134). */
int    o_ctx_iterate(o_ctx* c, int n, int faithful);
/* Per-frame stale-raylen buffers instead of one shared buffer (the batched
 * layout); identical results whenever b is 0 on rays that miss the volume. */
void   o_ctx_set_per_frame_raylen(o_ctx* c, int on);
/* CPU-baseline sample: wall time of PD dual+primal of one frame and of a
 * SART sweep over its first nviews views (steady_clock seconds). */
int    o_ctx_time_sample(o_ctx* c, int frame, int nviews, int faithful,
                         double* t_pd, double* t_views);

#ifdef __cplusplus
}
#endif
#endif
