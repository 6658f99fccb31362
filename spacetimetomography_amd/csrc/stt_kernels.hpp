// stt_kernels.hpp — host-side launchers of the HIP kernels in stt_kernels.hip.
#pragma once
#include <hip/hip_runtime_api.h>

#include "stt_device.hpp"

namespace stt {

enum { FP_MODE_FP = 0, FP_MODE_SART = 1, FP_MODE_SCATTER = 2, FP_MODE_RAYBP = 3 };

// Projector over nitems work items (one per frame), on stream s.
//   FP_MODE_FP      : ax, raylen (hit rays), weight
//   FP_MODE_SART    : corr = (F b - F Ax) / (F len + 1), raylen updated
//   FP_MODE_SCATTER : contrib += factor*corr*w (fp32 atomics; the projector's transpose)
//   FP_MODE_RAYBP   : acc_c += corr*w, acc_w += w in exact 64-bit fixed point
//                     (ray BP; launch_raybp_prep first sets the scale)
// Projector shape: g.fp_tile, per context (0 = slab projector, default; 64,
// 32, 16 or 8 = level-synchronous projector with that many rays per wave row).
hipError_t launch_fp(int mode, const WorkItem* items, int nitems, const VolGeom& g, float factor, hipStream_t s);
// Refreshes the transposed volume copy of every item whose view samples it
// (level-synchronous projector only; a no-op for the slab projector).
hipError_t launch_transpose_xz(const WorkItem* items, int nitems, const VolGeom& g, hipStream_t s);
hipError_t launch_bp_voxel(const WorkItem* items, int nitems, const VolGeom& g, const PdConst& pc, hipStream_t s);
// Ray BP, per view: scale of the fixed-point sums and the reset of the ray
// lengths (after the SART pass), then scatter (FP_MODE_RAYBP), then apply.
hipError_t launch_raybp_prep(const WorkItem* items, int nitems, const VolGeom& g, hipStream_t s);
hipError_t launch_bp_ray_apply(const WorkItem* items, int nitems, const VolGeom& g, const PdConst& pc, hipStream_t s);
// End points of the PD traces of one flow (step +1 into tb, -1 into tf).
hipError_t launch_trace(const float* uvw, float* tb, float* tf, const VolGeom& g, int interp, hipStream_t s);
hipError_t launch_pd_dual(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc, hipStream_t s);
hipError_t launch_pd_primal(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc, hipStream_t s);
hipError_t launch_extrapolate(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc,
                              hipStream_t s);
hipError_t launch_warp(const float* uvw, const float* src, float* out, const VolGeom& g, float dt, int cubic,
                       hipStream_t s);
// Adjoint of the trilinear warp: gsrc += g dout/dsrc, guvw += g dout/duvw (atomics; either may be null).
hipError_t launch_warp_vjp(const float* uvw, const float* src, const float* gout, float* gsrc, float* guvw,
                           const VolGeom& g, float dt, hipStream_t s);
// FP64 reductions: kReductionBlocks partials per call, then one ordered final sum.
constexpr int kReductionBlocks = 256;
// partial[t * stride], t < kReductionBlocks
hipError_t launch_residual_partial(const float* b, const float* ax, int n, double* partial, int stride,
                                   hipStream_t s);
hipError_t launch_dot_partial(const float* a, const float* c, size_t n, double* partial, hipStream_t s);
hipError_t launch_sum_partials(const double* partial, int n, double* out, hipStream_t s);
hipError_t launch_scale_by_inv_sqrt(float* x, size_t n, const double* nrm2, hipStream_t s);

// Allocation-free residual norm (stt_residual_norm_noalloc).  A file-scope
// device array of kNoAllocScratchBytes holds, per call: one work item, the
// current view's line integrals and ray lengths (detectors up to
// kNoAllocMaxDet pixels), its kReductionBlocks FP64 partials, the blocks'
// running sums and the result.  noalloc_scratch() returns the array's
// address on the current device (hipError_t as int).
constexpr size_t kNoAllocMaxDet = size_t(1) << 20;
constexpr size_t kNoAllocItemOff = 0;                                    // one WorkItem (256 B reserved)
constexpr size_t kNoAllocPartOff = 256;                                  // kReductionBlocks doubles
constexpr size_t kNoAllocAccOff = kNoAllocPartOff + 8 * kReductionBlocks;  // kReductionBlocks doubles
constexpr size_t kNoAllocOutOff = kNoAllocAccOff + 8 * kReductionBlocks;   // 1 double
constexpr size_t kNoAllocAxOff = 8192;                                   // kNoAllocMaxDet floats
constexpr size_t kNoAllocRlOff = kNoAllocAxOff + 4 * kNoAllocMaxDet;      // kNoAllocMaxDet floats
constexpr size_t kNoAllocScratchBytes = kNoAllocRlOff + 4 * kNoAllocMaxDet;
constexpr int kNoAllocMaxDevices = 64;
static_assert(sizeof(WorkItem) <= kNoAllocPartOff, "work item slot too small");
static_assert(kNoAllocOutOff + 8 <= kNoAllocAxOff, "scratch layout overlaps");
int noalloc_scratch(char** ptr);
hipError_t launch_item_make(WorkItem* dst, const ViewRec* view, float* vol, float* volT, const float* b, float* ax,
                            float* raylen, hipStream_t s);
hipError_t launch_zero_f64(double* p, int n, hipStream_t s);
hipError_t launch_add_partials(double* acc, const double* part, int n, hipStream_t s);
hipError_t launch_fill_hash(float* x, size_t n, unsigned seed, hipStream_t s);
// Power-iteration operator K (masked spatial forward differences + time difference) and K^T.
hipError_t launch_pi_apply_k(const float* v, float* y, int F, const VolGeom& g, const PdConst& pc, hipStream_t s);
hipError_t launch_pi_apply_kt(const float* y, float* w, int F, const VolGeom& g, hipStream_t s);

// Optical flow, one pyramid level: aabb-aware clamped-cubic warp of I2 by V,
// gradient + data term + step sizes, primal (mode 0 plain, 1 keep K^T y,
// 2 residual) and dual (res: residual) steps, ROI write-back.  Residual
// kernels write one FP64 partial per block (of_partials(g) blocks).
int of_partials(const VolGeom& g);
// The primal covers the ROI's y and z range only: its number of partials.
int of_primal_partials(const VolGeom& g, const int* roi);
hipError_t launch_of_warp(const float* v, const float* img, float* out, const VolGeom& g, const float* ext,
                          hipStream_t s);
hipError_t launch_of_prep(const float* W, const float* I1, const float* v, float* G, const VolGeom& g,
                          const int* roi, hipStream_t s);
hipError_t launch_of_primal(int mode, const float* vin, float* vout, const float* Y, float* kty, const float* G,
                            double* partial, const VolGeom& g, const int* roi, hipStream_t s);
// first: 0 normal, 1 first iteration of a level, 2 first iteration of a later warp
hipError_t launch_of_dual(bool res, int first, const float* v, const float* vold, float* Y, const float* k4old,
                          const float* G, double* partial, const VolGeom& g, float sigma1, float hf, float lambda,
                          int huber, hipStream_t s);
hipError_t launch_of_k4(const float* v, const float* G, float* k4, const VolGeom& g, hipStream_t s);
hipError_t launch_of_roi_copy(float* dst, const float* src, int planes, const VolGeom& g, const int* roi,
                              hipStream_t s);

// The flow pyramid's image operators: order-0 Deriche along one axis (one
// thread per line) and one axis of CImg's interpolation-5 resize (tables from
// the host: moving-average segments when shrinking, cubic taps when growing).
struct DericheCoef {
    float a0, a1, a2, a3, b1, b2, coefp, coefn;
};
hipError_t launch_deriche(const float* in, float* out, const int* dims, int axis, const DericheCoef& c, hipStream_t s);
hipError_t launch_resize_axis(const float* in, float* out, const int* dims, int axis, int ns, const int* seg0,
                              const int* seg_s, const int* seg_d, const int* sidx, const double* frac,
                              hipStream_t s);

}  // namespace stt
