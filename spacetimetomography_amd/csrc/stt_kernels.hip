// stt_kernels.hip — CDNA4 (gfx950) kernels of the space-time tomography path.
//
// Kernels (one thread = one ray or one voxel, 64-wide wavefronts):
//   k_fp<MODE>     cone-beam ray caster, one ray per lane.  MODE_FP writes the
//                  line integral, ray length and sample count; MODE_SART fuses
//                  the SART correction into the epilogue; MODE_SCATTER replays
//                  the same samples and scatters trilinear weights (ray-driven
//                  backprojector and the matched adjoint of the projector).
//   k_bp_voxel     voxel-driven backprojector, bicubic detector sampling,
//                  atomic-free; on a sweep's last view it also applies the ROI
//                  mask and the primal-dual extrapolation.
//   k_bp_ray_apply ray-BP normalisation x += contrib/wsum.
//   k_pd_dual      dual step with the backward warp fused per voxel.
//   k_pd_primal    primal step with the forward warp fused per voxel.
//   k_warp         stand-alone semi-Lagrangian warp.
// A launch covers several frames at once (grid.y / grid.z carry the frame),
// which is the batched form; the per-frame form launches one frame.
#include <hip/hip_runtime.h>

#include <mutex>

#include "stt_device.hpp"
#include "stt_kernels.hpp"

namespace stt {

// Reference min/max macros are ternaries (ST_Tomography.cpp:167-173); NaN
// behaviour differs from fminf/fmaxf, so the ray set-up keeps the ternaries.
__device__ __forceinline__ float tmin(float a, float b) { return (a < b) ? a : b; }
__device__ __forceinline__ float tmax(float a, float b) { return (a > b) ? a : b; }

struct Ray {
    float o[3], d[3];
    float tnear, tfar;
};

// ProduceRay (TypicalRayProducer.core.h:32-69).  No FMA contraction: the ray
// parameters decide the sample set, so they follow the reference's fp32
// expression order exactly.
__device__ __forceinline__ void produce_ray(const ViewRec& v, const VolGeom& g, float px, float py, Ray& r) {
#pragma clang fp contract(off)
    float m[4];
#pragma unroll
    for (int row = 0; row < 4; ++row) m[row] = (v.P2M[row] * px + v.P2M[4 + row] * py) + v.P2M[12 + row];
    const float winv = 1.0f / m[3];
    float dir[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dir[c] = m[c] * winv - v.eye[c];
    const float linv = 1.0f / sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.o[c] = v.o[c];
        r.d[c] = v.m2v[c] * (dir[c] * linv);
    }
    const float hi[3] = {g.bx, g.by, g.bz};
    float tn = -INFINITY, tf = INFINITY;
    float tmn[3], tmx[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float t1 = (0.0f - r.o[c]) / r.d[c];
        const float t2 = (hi[c] - r.o[c]) / r.d[c];
        tmn[c] = tmin(t1, t2);
        tmx[c] = tmax(t1, t2);
    }
    tn = tmax(tmax(tmn[0], tmn[1]), tmn[2]);
    r.tnear = tmax(tn, 0.0f);
    tf = tmin(tmin(tmx[0], tmx[1]), tmx[2]);
    r.tfar = tf;
}

// Global-address-space views of pointers read from work tables, so loads and
// atomics compile to global_* instead of flat_* instructions.
typedef __attribute__((address_space(1))) float gfloat;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f32x4 gf32x4;
typedef __attribute__((address_space(1))) unsigned long long gu64;
__device__ __forceinline__ gfloat* gptr(float* p) { return (gfloat*)p; }
__device__ __forceinline__ const gfloat* gptr(const float* p) { return (const gfloat*)p; }

// Raw buffer over one padded volume: 32-bit byte offsets, the +1 row and
// +1 slice taps through SGPR offsets, hardware range check on num_records.
struct VolView {
    __amdgpu_buffer_rsrc_t rsrc;
    unsigned nx4, sxy4, sxynx4;
    float bx, by, bz;
};

// Buffer resource word 3 of the gfx9 family (32-bit data format, raw access).
constexpr int kBufWord3 = 0x00020000;

__device__ __forceinline__ VolView vol_view(const float* vol, const VolGeom& g) {
    VolView v;
    v.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)vol, (short)0, (int)g.vol_bytes, kBufWord3);
    v.nx4 = 4u * (unsigned)g.nx;
    v.sxy4 = 4u * (unsigned)(g.nx * g.ny);
    v.sxynx4 = v.sxy4 + v.nx4;
    v.bx = g.bx;
    v.by = g.by;
    v.bz = g.bz;
    return v;
}

// T_VolumeAccess::GetScalarValue trilinear (SingleScalarVolume.core.h:87-146):
// clamp to [0, dims-1], taps at floor and floor+1.  The volume allocation
// carries a zero tail of nx*ny + nx + 1 floats, so the +1 taps need no clamp:
// where the reference clamps a tap, its weight is exactly 0.  Split into the
// tap fetch and the interpolation so the caster can keep the next sample's
// loads in flight while it interpolates the current one.
struct Taps {
    float2 r00, r10, r01, r11;  // (x, x+1) pairs of rows (y, z), (y+1, z), (y, z+1), (y+1, z+1)
    float fx, fy, fz;
};

__device__ __forceinline__ void fetch_taps(const VolView& v, float px, float py, float pz, Taps& T) {
    px = __builtin_amdgcn_fmed3f(px, 0.0f, v.bx);
    py = __builtin_amdgcn_fmed3f(py, 0.0f, v.by);
    pz = __builtin_amdgcn_fmed3f(pz, 0.0f, v.bz);
    const unsigned ix = (unsigned)px, iy = (unsigned)py, iz = (unsigned)pz;
    T.fx = __builtin_amdgcn_fractf(px);
    T.fy = __builtin_amdgcn_fractf(py);
    T.fz = __builtin_amdgcn_fractf(pz);
    const unsigned off = (ix << 2) + __umul24(iy, v.nx4) + __umul24(iz, v.sxy4);
    T.r00 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, off, 0, 0));
    T.r10 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, off, v.nx4, 0));
    T.r01 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, off, v.sxy4, 0));
    T.r11 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, off, v.sxynx4, 0));
}

__device__ __forceinline__ float lerp_taps(const Taps& T) {
    const float a = fmaf(T.fx, T.r00.y - T.r00.x, T.r00.x);
    const float b = fmaf(T.fx, T.r10.y - T.r10.x, T.r10.x);
    const float c = fmaf(T.fx, T.r01.y - T.r01.x, T.r01.x);
    const float d = fmaf(T.fx, T.r11.y - T.r11.x, T.r11.x);
    const float ab = fmaf(T.fy, b - a, a);
    const float cd = fmaf(T.fy, d - c, c);
    return fmaf(T.fz, cd - ab, ab);
}

__device__ __forceinline__ float sample_tri(const VolView& v, float px, float py, float pz) {
    Taps T;
    fetch_taps(v, px, py, pz, T);
    return lerp_taps(T);
}

// Raw buffer over the transposed copy (index z + nz*(y + ny*x), zero tail of
// nz*ny + nz + 1 floats): the contiguous tap pair runs along z.
__device__ __forceinline__ VolView vol_view_t(const float* volT, const VolGeom& g) {
    VolView v;
    v.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)volT, (short)0, (int)g.volT_bytes, kBufWord3);
    v.nx4 = 4u * (unsigned)g.nz;
    v.sxy4 = 4u * (unsigned)(g.nz * g.ny);
    v.sxynx4 = v.sxy4 + v.nx4;
    v.bx = g.bx;
    v.by = g.by;
    v.bz = g.bz;
    return v;
}

// Same trilinear sample from the transposed copy: interpolation runs along
// z, then y, then x (fractions placed so lerp_taps applies unchanged).
__device__ __forceinline__ void fetch_taps_t(const VolView& v, float px, float py, float pz, Taps& T) {
    px = __builtin_amdgcn_fmed3f(px, 0.0f, v.bx);
    py = __builtin_amdgcn_fmed3f(py, 0.0f, v.by);
    pz = __builtin_amdgcn_fmed3f(pz, 0.0f, v.bz);
    const unsigned ix = (unsigned)px, iy = (unsigned)py, iz = (unsigned)pz;
    T.fx = __builtin_amdgcn_fractf(pz);
    T.fy = __builtin_amdgcn_fractf(py);
    T.fz = __builtin_amdgcn_fractf(px);
    const unsigned off = (iz << 2) + __umul24(iy, v.nx4) + __umul24(ix, v.sxy4);
    T.r00 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, off, 0, 0));
    T.r10 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, off, v.nx4, 0));
    T.r01 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, off, v.sxy4, 0));
    T.r11 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, off, v.sxynx4, 0));
}

// Trilinear scatter of one sample (SingleScalarVolume.core.h:87-143): clamp
// to [0, dims-1], the reference's 8 fp32 weights w[n], then per corner the
// contribution w[n]*corr and the weight w[n] (UpdateContributionSum, :73-84).
// Accumulators carry the volume's zero tail, so the +1 corners need no clamp
// (their weights are 0 wherever the reference clamps them).
struct TriCorners {
    unsigned base;
    unsigned off[8];
    float w[8];
};

__device__ __forceinline__ void tri_corners(const VolGeom& g, float px, float py, float pz, TriCorners& c) {
    px = __builtin_amdgcn_fmed3f(px, 0.0f, g.bx);
    py = __builtin_amdgcn_fmed3f(py, 0.0f, g.by);
    pz = __builtin_amdgcn_fmed3f(pz, 0.0f, g.bz);
    const unsigned ix = (unsigned)px, iy = (unsigned)py, iz = (unsigned)pz;
    const float t0 = __builtin_amdgcn_fractf(px), t1 = __builtin_amdgcn_fractf(py), t2 = __builtin_amdgcn_fractf(pz);
    const float m0 = 1.0f - t0, m1 = 1.0f - t1, m2 = 1.0f - t2;
    c.w[0] = m0 * m1 * m2;
    c.w[1] = t0 * m1 * m2;
    c.w[2] = m0 * t1 * m2;
    c.w[3] = t0 * t1 * m2;
    c.w[4] = m0 * m1 * t2;
    c.w[5] = t0 * m1 * t2;
    c.w[6] = m0 * t1 * t2;
    c.w[7] = t0 * t1 * t2;
    const unsigned nx = (unsigned)g.nx, sxy = (unsigned)(g.nx * g.ny);
    c.base = ix + __umul24(iy, nx) + __umul24(iz, sxy);
    c.off[0] = 0u;
    c.off[1] = 1u;
    c.off[2] = nx;
    c.off[3] = nx + 1u;
    c.off[4] = sxy;
    c.off[5] = sxy + 1u;
    c.off[6] = sxy + nx;
    c.off[7] = sxy + nx + 1u;
}

// fp32 scatter with hardware atomics: the projector's transpose (A^T y).
struct SinkF32 {
    gfloat* acc;
    float val;
    __device__ __forceinline__ void operator()(const VolGeom& g, float px, float py, float pz) const {
        TriCorners c;
        tri_corners(g, px, py, pz, c);
        gfloat* a = acc + c.base;
#pragma unroll
        for (int n = 0; n < 8; ++n) __builtin_amdgcn_global_atomic_fadd_f32(a + c.off[n], val * c.w[n]);
    }
};

// a * 2^s rounded to the nearest integer; exact wherever a's 24-bit mantissa
// fits above 2^-s (|a| * 2^s < 2^62 by the choice of s)
__device__ __forceinline__ unsigned long long fixed_of(float a, double scale) {
    return (unsigned long long)__double2ll_rn((double)a * scale);
}

// Ray-BP scatter into exact 64-bit fixed-point sums.  Integer addition is
// associative, so the sums -- and the volume update -- do not depend on the
// order in which the lanes' atomics land: the result is the same on every
// run.  Each addend is the reference's fp32 value (w[n]*corr, w[n]).
struct SinkFixed {
    unsigned long long* acc_c;
    unsigned long long* acc_w;
    float val;
    double cscale, wscale;
    __device__ __forceinline__ void operator()(const VolGeom& g, float px, float py, float pz) const {
        TriCorners c;
        tri_corners(g, px, py, pz, c);
        gu64* ac = (gu64*)acc_c + c.base;
        gu64* aw = (gu64*)acc_w + c.base;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            __hip_atomic_fetch_add(ac + c.off[n], fixed_of(c.w[n] * val, cscale), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(aw + c.off[n], fixed_of(c.w[n], wscale), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
};

// CastingRay (RayCaster.core.h:42-187) with one block = the whole volume
// (RayCasterPrepare :199-201).  Samples: the entry point once more than the
// grid (:60-61), then t_k = tnear + k*sd below the next block boundary; the only
// interior boundaries are the planes at voxel coordinate 0.5 crossed in the
// negative direction, where the grid restarts at ceil((t_b - tnear)/sd).
struct SinkNone {
    __device__ __forceinline__ void operator()(const VolGeom&, float, float, float) const {}
};

template <bool SCATTER, class Sink>
__device__ __forceinline__ float cast_ray(const VolView& vol, const Sink& sink, const VolGeom& g, const Ray& r,
                                          int& nsamp) {
    const float sd = g.sd;
    const int dimv[3] = {g.nx, g.ny, g.nz};
    float pos1[3];
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int c = 0; c < 3; ++c) pos1[c] = r.o[c] + r.d[c] * r.tnear;
    }
    float step[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) step[c] = r.d[c] * sd;
    float density = 0.0f;
    int n = 1;
    if (SCATTER) sink(g, pos1[0], pos1[1], pos1[2]);
    else density = sample_tri(vol, pos1[0], pos1[1], pos1[2]) * sd;
    float tMax[3], tDelta[3];
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float dc = r.d[c];
            if (dc == 0.0f) {
                tMax[c] = r.tfar;
                tDelta[c] = 0.0f;
            } else if (dc < 0.0f) {
                const int blk = ((int)ceilf(pos1[c] - 0.5f) - 1) / dimv[c];
                tMax[c] = ((float)(blk * dimv[c]) + 0.5f - r.o[c]) / dc;
                tDelta[c] = (float)(-dimv[c]) / dc;
            } else {
                const int blk = ((int)floorf(pos1[c] - 0.5f)) / dimv[c];
                tMax[c] = ((float)((blk + 1) * dimv[c]) + 0.5f - r.o[c]) / dc;
                tDelta[c] = (float)dimv[c] / dc;
            }
        }
    }
    float t1 = r.tnear;
    for (int seg = 0; seg < 8 && t1 < r.tfar; ++seg) {
        float tb = tmin(tmin(tMax[0], tMax[1]), tMax[2]);
        if (tb > r.tfar) tb = r.tfar;
        float t, px, py, pz;
        {
#pragma clang fp contract(off)
            t = ceilf((t1 - r.tnear) / sd) * sd + r.tnear;
            px = r.o[0] + r.d[0] * t;
            py = r.o[1] + r.d[1] * t;
            pz = r.o[2] + r.d[2] * t;
        }
        if (SCATTER) {
            while (t < tb && n < g.max_samples) {
                sink(g, px, py, pz);
                ++n;
                t += sd;
                px += step[0];
                py += step[1];
                pz += step[2];
            }
        } else if (t < tb) {
            // one sample ahead: the taps of sample k+1 are in flight while
            // sample k is interpolated
            Taps cur;
            fetch_taps(vol, px, py, pz, cur);
            while (n < g.max_samples) {
                ++n;
                t += sd;
                px += step[0];
                py += step[1];
                pz += step[2];
                const bool more = t < tb;
                Taps nxt;
                if (more) fetch_taps(vol, px, py, pz, nxt);
                density = fmaf(lerp_taps(cur), sd, density);
                if (!more) break;
                cur = nxt;
            }
        }
        if (tMax[0] < tMax[1]) {
            if (tMax[0] < tMax[2]) tMax[0] += tDelta[0];
            else tMax[2] += tDelta[2];
        } else {
            if (tMax[1] < tMax[2]) tMax[1] += tDelta[1];
            else tMax[2] += tDelta[2];
        }
        t1 = tb;
    }
    nsamp = n;
    return density;
}

// Detector tile of TX x TY rays per block and the work item (frame) it
// belongs to; the grid is (tiles, items).  Blocks are dealt to the 8 XCDs
// round-robin in dispatch order (x fastest); the remapping gives each XCD a
// contiguous eighth of the (item, tile) range, so an XCD's L2 serves whole
// frames of a multi-frame launch (and a contiguous band of detector rows,
// one slab of the volume, of a single-frame launch).  8 frames of 256^3 in
// one launch.
template <int TX, int TY>
__device__ __forceinline__ bool tile_pixel(const VolGeom& g, int& i, int& j, int& item) {
    const int tiles_x = (g.W + TX - 1) / TX;
    const int tiles_y = (g.H + TY - 1) / TY;
    const int nb = tiles_x * tiles_y;
    const int n = nb * (int)gridDim.y;
    int b = (int)blockIdx.x + nb * (int)blockIdx.y;
    if ((n & 7) == 0) b = (b & 7) * (n >> 3) + (b >> 3);
    item = b / nb;
    b -= item * nb;
    const int tx = b % tiles_x, ty = b / tiles_x;
    i = tx * TX + (threadIdx.x % TX);
    j = ty * TY + (threadIdx.x / TX);
    return i < g.W && j < g.H;
}

template <int MODE, int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_fp(const WorkItem* __restrict__ items, VolGeom g, float factor) {
    int i, j;
    int item;
    if (!tile_pixel<TX, TY>(g, i, j, item)) return;
    const WorkItem& it = items[item];
    const ViewRec& v = *it.view;
    const int k = i + j * g.W;
    Ray r;
    produce_ray(v, g, (float)i, (float)j, r);
    const bool hit = r.tnear < r.tfar;
    if (MODE == FP_MODE_SCATTER) {
        // the projector's transpose: factor*y[k] scattered with the weights
        if (hit) {
            int ns;
            const SinkF32 sink{gptr(it.contrib), gptr(it.corr)[k] * factor};
            cast_ray<true>(VolView{}, sink, g, r, ns);
        }
        return;
    }
    if (MODE == FP_MODE_RAYBP) {
        // Ray BP: the BP-mode pass of GenerateImage_MT over the same samples
        // (ST_Tomography.cpp:733-739).  Its ray lengths equal the SART pass's
        // and are cleared right after (Initial(), :763), so none is written.
        if (hit) {
            int ns;
            const SinkFixed sink{it.acc_c, it.acc_w, gptr(it.corr)[k], ldexp(1.0, *it.acc_shift),
                                 ldexp(1.0, g.acc_wshift)};
            cast_ray<true>(VolView{}, sink, g, r, ns);
        }
        return;
    }
    float value = 0.0f;
    int ns = 0;
    if (hit) value = cast_ray<false>(vol_view(it.vol, g), SinkNone{}, g, r, ns);
    if (MODE == FP_MODE_FP) {
        if (hit && it.raylen) gptr(it.raylen)[k] = r.tfar - r.tnear;
        if (it.ax) gptr(it.ax)[k] = value;
        if (it.weight) gptr(it.weight)[k] = (float)ns;
    } else {
        // SART correction (ST_Tomography.cpp:667-676); pY is zero when read.
#pragma clang fp contract(off)
        float len;
        gfloat* rl = gptr(it.raylen);
        if (hit) {
            len = r.tfar - r.tnear;
            rl[k] = len;
        } else {
            len = rl[k];  // stale value of the last view whose ray k hit
        }
        const float num = factor * gptr(it.b)[k] - factor * value;
        const float den = factor * len + 1.0f;
        gptr(it.corr)[k] = num / den;
    }
}

// imROI (ST_Tomography.cpp:303-329).
__device__ __forceinline__ bool roi_in(const VolGeom& g, const PdConst& pc, int i, int j, int k) {
    const int* c = pc.cut;
    if (((i < c[6]) && (k < c[7])) || ((i < c[6]) && (k > (g.nz - c[7])))) return false;
    if (((i > (g.nx - c[6])) && (k < c[7])) || ((i > (g.nx - c[6])) && (k > (g.nz - c[7])))) return false;
    if (((k < c[6]) && (i < c[7])) || ((k < c[6]) && (i > (g.nx - c[7])))) return false;
    if (((k > (g.nz - c[6])) && (i < c[7])) || ((k > (g.nz - c[6])) && (i > (g.nx - c[7])))) return false;
    if (i < c[0] || i > g.nx - c[1] || j < c[2] || j > g.ny - pc.jtop || k < c[4] || k > g.nz - c[5]) return false;
    return true;
}
// ROI test of the primal-dual loops (ST_Tomography.cpp:514,565,593).
__device__ __forceinline__ bool cut_out(const VolGeom& g, const PdConst& pc, int i, int j, int k) {
    const int* c = pc.cut;
    return i < c[0] || i > g.nx - c[1] || j < c[2] || j > g.ny - c[3] || k < c[4] || k > g.nz - c[5];
}

// Sweep epilogue: x *= imROI (ST_Tomography.cpp:793-804), then the
// extrapolation X_bar = 2x - X_prev on the ROI (:586-604).
__device__ __forceinline__ float finalize_voxel(const WorkItem& it, const VolGeom& g, const PdConst& pc, int x, int y,
                                                int z, size_t q, float xv) {
    xv *= roi_in(g, pc, x, y, z) ? 1.0f : 0.0f;
    if (it.xbar && !cut_out(g, pc, x, y, z)) gptr(it.xbar)[q] = 2.0f * xv - gptr(it.xprev)[q];
    return xv;
}

// Sample generator of CastingRay (RayCaster.core.h:42-187): the entry sample,
// then for each block segment the grid t = ceil((t1 - tnear)/sd)*sd + tnear,
// t += sd while t < t_b, the DDA advancing one tMax per finished segment.
// Same samples, same order as cast_ray; only the interleaving across the
// lanes of a wave changes.  The first segment starts at t = tnear, on the
// entry point itself, so when it is not empty the entry sample and that first
// sample coincide and are taken once with weight 2 (RayCaster.core.h:60-61,
// 137-145); a ray whose first segment is empty (a block boundary before the
// entry, the pre-entry case) takes the entry sample on its own.
struct RayGen {
    float o[3], d[3], step[3];
    float tnear, tfar;
    float tMax[3], tDelta[3];
    float t1, tb, t;
    float pend[3];  // position of the pending (next) sample
    float w;        // weight (sample multiplicity) of the pending sample
    int segs;
};

// Opens the next non-empty block segment after t1; false when the ray ends.
__device__ __forceinline__ bool raygen_open(RayGen& G, float sd) {
#pragma nounroll
    for (int guard = 0; guard < 8; ++guard) {
        if (!(G.t1 < G.tfar) || G.segs >= 8) return false;
        float tb = tmin(tmin(G.tMax[0], G.tMax[1]), G.tMax[2]);
        if (tb > G.tfar) tb = G.tfar;
        G.tb = tb;
        {
#pragma clang fp contract(off)
            G.t = ceilf((G.t1 - G.tnear) / sd) * sd + G.tnear;
            G.pend[0] = G.o[0] + G.d[0] * G.t;
            G.pend[1] = G.o[1] + G.d[1] * G.t;
            G.pend[2] = G.o[2] + G.d[2] * G.t;
        }
        ++G.segs;
        if (G.t < G.tb) return true;
        // empty segment: advance the DDA and try the next one
        if (G.tMax[0] < G.tMax[1]) {
            if (G.tMax[0] < G.tMax[2]) G.tMax[0] += G.tDelta[0];
            else G.tMax[2] += G.tDelta[2];
        } else {
            if (G.tMax[1] < G.tMax[2]) G.tMax[1] += G.tDelta[1];
            else G.tMax[2] += G.tDelta[2];
        }
        G.t1 = G.tb;
    }
    return false;
}

// Current segment exhausted: advance the DDA and open the next one.
__device__ __forceinline__ bool raygen_next_segment(RayGen& G, float sd) {
    if (G.tMax[0] < G.tMax[1]) {
        if (G.tMax[0] < G.tMax[2]) G.tMax[0] += G.tDelta[0];
        else G.tMax[2] += G.tDelta[2];
    } else {
        if (G.tMax[1] < G.tMax[2]) G.tMax[1] += G.tDelta[1];
        else G.tMax[2] += G.tDelta[2];
    }
    G.t1 = G.tb;
    return raygen_open(G, sd);
}

// Sets up the ray.  `entry_alone` is set when the entry sample has to be taken
// on its own before the generator's pending sample; the return value says
// whether the generator holds a pending sample.  A ray that misses the volume
// takes no sample at all.
__device__ __forceinline__ bool raygen_init(RayGen& G, const Ray& r, const VolGeom& g, bool hit, bool& entry_alone) {
    const int dimv[3] = {g.nx, g.ny, g.nz};
    float pos1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        G.o[c] = r.o[c];
        G.d[c] = r.d[c];
        G.step[c] = r.d[c] * g.sd;
    }
    G.tnear = r.tnear;
    G.tfar = r.tfar;
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int c = 0; c < 3; ++c) pos1[c] = r.o[c] + r.d[c] * r.tnear;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float dc = r.d[c];
            if (dc == 0.0f) {
                G.tMax[c] = r.tfar;
                G.tDelta[c] = 0.0f;
            } else if (dc < 0.0f) {
                const int blk = ((int)ceilf(pos1[c] - 0.5f) - 1) / dimv[c];
                G.tMax[c] = ((float)(blk * dimv[c]) + 0.5f - r.o[c]) / dc;
                G.tDelta[c] = (float)(-dimv[c]) / dc;
            } else {
                const int blk = ((int)floorf(pos1[c] - 0.5f)) / dimv[c];
                G.tMax[c] = ((float)((blk + 1) * dimv[c]) + 0.5f - r.o[c]) / dc;
                G.tDelta[c] = (float)dimv[c] / dc;
            }
        }
    }
    G.t1 = r.tnear;
    G.segs = 0;
    G.tb = G.t = 0.0f;
    G.w = 1.0f;
    entry_alone = false;
    if (!hit) return false;
    const bool open = raygen_open(G, g.sd);
    // The first segment opens at t = ceil(0)*sd + tnear = tnear, the entry
    // point; only when it is non-empty do the two samples coincide.
    if (open && G.segs == 1) {
        G.w = 2.0f;
        return true;
    }
    entry_alone = true;
    return open;
}

__device__ __forceinline__ float wave_min_f32(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// Level-synchronous projector.  A wave holds 64 rays of a detector tile and
// advances a depth level along the view's dominant axis D by the smallest
// per-sample step of its rays; a lane takes its next sample once the level
// has reached it.  Rays of one tile are nearly parallel, so nearly every lane
// samples on every step while the wave's taps stay within one or two planes,
// close together across the transverse axis that the memory layout keeps
// contiguous (x-fastest for z-dominant views, the transposed copy for
// x-dominant views).  Samples, their order and each ray's sum are those of
// CastingRay.
template <int D, bool DOWN, bool TRANSPOSED>
__device__ __forceinline__ float march_ray(const VolView& vv, const VolGeom& g, const Ray& r, bool hit, int& nsamp) {
    RayGen G;
    bool entry_alone;
    bool have = raygen_init(G, r, g, hit, entry_alone);
    float density = 0.0f;
    int n = 0;
    if (entry_alone) {  // pre-entry case: the entry sample first, on its own
        float e0, e1, e2;
        {
#pragma clang fp contract(off)
            e0 = r.o[0] + r.d[0] * r.tnear;
            e1 = r.o[1] + r.d[1] * r.tnear;
            e2 = r.o[2] + r.d[2] * r.tnear;
        }
        Taps T;
        if (TRANSPOSED) fetch_taps_t(vv, e0, e1, e2, T);
        else fetch_taps(vv, e0, e1, e2, T);
        density = lerp_taps(T) * g.sd;
        n = 1;
    }
    const float s = wave_min_f32(have ? fabsf(G.step[D]) : INFINITY);
    float lvl = DOWN ? wave_max_f32(have ? G.pend[D] : -INFINITY) : wave_min_f32(have ? G.pend[D] : INFINITY);
    const float dl = DOWN ? -s : s;
    const int limit = 4 * (g.nx + g.ny + g.nz) + 2 * g.max_samples + 16;
    for (int it = 0; it < limit && __any(have); ++it) {
        if (have && (DOWN ? G.pend[D] >= lvl : G.pend[D] <= lvl)) {
            Taps T;
            if (TRANSPOSED) fetch_taps_t(vv, G.pend[0], G.pend[1], G.pend[2], T);
            else fetch_taps(vv, G.pend[0], G.pend[1], G.pend[2], T);
            const float w = G.w;
            n += (int)w;
            G.w = 1.0f;
            G.t += g.sd;
            G.pend[0] += G.step[0];
            G.pend[1] += G.step[1];
            G.pend[2] += G.step[2];
            if (!(G.t < G.tb)) have = raygen_next_segment(G, g.sd);
            have = have && n < g.max_samples;
            density = fmaf(lerp_taps(T), w * g.sd, density);
        }
        lvl += dl;
    }
    nsamp = n;
    return density;
}

template <int MODE, int TX>
__global__ __launch_bounds__(256) void k_fp_level(const WorkItem* __restrict__ items, VolGeom g, float factor) {
    constexpr int TY = 256 / TX;
    int i, j;
    int item;
    const bool valid = tile_pixel<TX, TY>(g, i, j, item);
    const WorkItem& it = items[item];
    const ViewRec& v = *it.view;
    const int march = v.march;
    const bool down = (march & 4) != 0;
    Ray r;
    produce_ray(v, g, (float)(valid ? i : 0), (float)(valid ? j : 0), r);
    const bool hit = valid && r.tnear < r.tfar;
    int ns = 0;
    float value;
    const int D = march & 3;
    if (MODE == FP_MODE_SART && (march & 8)) {
        const VolView vt = vol_view_t(it.volT, g);
        value = down ? march_ray<0, true, true>(vt, g, r, hit, ns) : march_ray<0, false, true>(vt, g, r, hit, ns);
    } else {
        const VolView vv = vol_view(it.vol, g);
        if (D == 0) value = down ? march_ray<0, true, false>(vv, g, r, hit, ns) : march_ray<0, false, false>(vv, g, r, hit, ns);
        else if (D == 1) value = down ? march_ray<1, true, false>(vv, g, r, hit, ns) : march_ray<1, false, false>(vv, g, r, hit, ns);
        else value = down ? march_ray<2, true, false>(vv, g, r, hit, ns) : march_ray<2, false, false>(vv, g, r, hit, ns);
    }
    if (!valid) return;
    const int k = i + j * g.W;
    if (MODE == FP_MODE_FP) {
        if (hit && it.raylen) gptr(it.raylen)[k] = r.tfar - r.tnear;
        if (it.ax) gptr(it.ax)[k] = hit ? value : 0.0f;
        if (it.weight) gptr(it.weight)[k] = (float)ns;
    } else {
        // SART correction (ST_Tomography.cpp:667-676); pY is zero when read.
#pragma clang fp contract(off)
        float len;
        gfloat* rl = gptr(it.raylen);
        if (hit) {
            len = r.tfar - r.tnear;
            rl[k] = len;
        } else {
            len = rl[k];  // stale value of the last view whose ray k hit
        }
        const float num = factor * gptr(it.b)[k] - factor * (hit ? value : 0.0f);
        const float den = factor * len + 1.0f;
        gptr(it.corr)[k] = num / den;
    }
}

// ---------------------------------------------------------------------------
// Slab projector.  A 256-thread block owns a 16 x 16 detector tile (wave w =
// tile rows 4w..4w+3) and marches it through the volume in slabs of up to
// kSlabK planes along the view's dominant axis D.  All rays of a view leave
// the same source point, so at a plane c the tile's rays lie between the
// extreme slopes dA/dD, dB/dD of its rays on the two transverse axes; the
// block stages the box of voxels those positions can touch (with one voxel of
// margin and the +1 taps) from the x-fastest volume into LDS with row loads,
// then every lane takes the samples of its ray whose base plane lies in the
// slab and reads the 8 taps from LDS.  The sample sequence is CastingRay's
// (the same generator as march_ray), only the tap source changes: a sample
// whose taps are not inside the staged box (behind the slab -- the pre-entry
// samples of grazing rays --, off the box by rounding, or a slab too wide for
// LDS) reads them from global memory.  No layout copy of the volume is needed
// for any view direction.
constexpr int kSlabCap = 6144;  // default floats of staged box per block (24 KB); the box size is a template
                                // argument (16, 20, 24 or 28 KB: stt_set_projector_box / stt_autotune_box)
constexpr int kSlabK = 23;      // most sample planes per slab (box depth <= 24)
constexpr int kSlabKMin = 8;    // fewest sample planes per slab (before the LDS cap)
constexpr float kSlabShear = 0.7f;  // slab depth: the box may widen by this fraction of its near-plane width
                                    // (0.35, 0.5, 0.9, 1.2: the iteration 1.5-6% slower)
constexpr int kStagePasses = 4; // staging covers x-y windows up to 4*256 elements
constexpr bool kSlabUncheckedPath = true;  // sampling without the per-sample box test (see slab_ray)

struct SlabBox {
    int lo[3], n[3];   // box origin and extent per volume axis (x, y, z)
    int cbeg, cend;    // sample base planes of the slab: [cbeg, cend) (up) / (cend, cbeg] (down)
    int px, pxy;       // LDS row and plane pitch (odd, so lanes on one x spread over the banks)
    int base;          // LDS index of (lo), subtracted from tap indices
    bool staged;
    bool inner;        // every sample of the slab lies inside [0, dims-1]^3: no clamp needed
};

// A slab box passes from the wave that computes it to the whole block
// through an LDS slot of kBoxInts packed ints, read back as wave-uniform
// (scalar) values: 16-bit lo[0..2] and n[0..2] (every volume dimension is
// at most 16384, so a staged box's values fit), then cbeg as a signed
// 24-bit value (|cbeg| <= nx + ny + nz + kSlabK), the depth K = |cend - cbeg|
// (at most kSlabK) and the two flags.  The row and plane pitches and the
// base index follow from them on the scalar unit.  Boxes that are not staged
// use only cbeg and cend.
constexpr int kBoxInts = 4;
__device__ __forceinline__ void box_put(int* slot, const SlabBox& b) {
    if ((threadIdx.x & 63) == 0) {
        const int K = abs(b.cend - b.cbeg);
        const unsigned w[kBoxInts] = {
            ((unsigned)b.lo[0] & 0xFFFFu) | ((unsigned)b.lo[1] << 16),
            ((unsigned)b.lo[2] & 0xFFFFu) | ((unsigned)b.n[2] << 16),
            ((unsigned)b.n[0] & 0xFFFFu) | ((unsigned)b.n[1] << 16),
            ((unsigned)b.cbeg & 0xFFFFFFu) | ((unsigned)(K & 0x1F) << 24) | ((unsigned)b.staged << 29) |
                ((unsigned)b.inner << 30)};
#pragma unroll
        for (int k = 0; k < kBoxInts; ++k) slot[k] = (int)w[k];
    }
}
template <bool DOWN>
__device__ __forceinline__ SlabBox box_get(const int* slot) {
    const unsigned v0 = (unsigned)__builtin_amdgcn_readfirstlane(slot[0]);
    const unsigned v1 = (unsigned)__builtin_amdgcn_readfirstlane(slot[1]);
    const unsigned v2 = (unsigned)__builtin_amdgcn_readfirstlane(slot[2]);
    const unsigned v3 = (unsigned)__builtin_amdgcn_readfirstlane(slot[3]);
    SlabBox b;
    b.lo[0] = (int)(v0 & 0xFFFFu);
    b.lo[1] = (int)(v0 >> 16);
    b.lo[2] = (int)(v1 & 0xFFFFu);
    b.n[0] = (int)(v2 & 0xFFFFu);
    b.n[1] = (int)(v2 >> 16);
    b.n[2] = (int)(v1 >> 16);
    b.cbeg = (int)(v3 << 8) >> 8;
    const int K = (int)((v3 >> 24) & 0x1Fu);
    b.cend = DOWN ? b.cbeg - K : b.cbeg + K;
    b.staged = ((v3 >> 29) & 1u) != 0;
    b.inner = ((v3 >> 30) & 1u) != 0;
    b.px = b.n[0] | 1;
    b.pxy = (b.px * b.n[1]) | 1;
    b.base = b.lo[0] + b.px * b.lo[1] + b.pxy * b.lo[2];
    return b;
}

// Block-wide min/max of NV values (wave shuffles, then one LDS round).
template <int NW, int NV>
__device__ __forceinline__ void block_minmax(float v[NV], const bool is_max[NV], float* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        v[k] = is_max[k] ? wave_max_f32(v[k]) : wave_min_f32(v[k]);
        if (lane == 0) red[wv * NV + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        float a = red[k];
#pragma unroll
        for (int w = 1; w < NW; ++w) a = is_max[k] ? fmaxf(a, red[w * NV + k]) : fminf(a, red[w * NV + k]);
        v[k] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, a)));
    }
}

// Box of the slab whose sample base planes start at c0 (uniform arithmetic);
// CAP floats of LDS hold it.
// PADDED: the projector's x-fastest LDS layout (odd row and plane pitches);
// false: a dense layout with the slab axis slowest (the ray-BP scatter box),
// whose footprint is exactly n[A] * n[B] * n[D].
template <int D, bool DOWN, int CAP, bool PADDED = true>
__device__ __forceinline__ SlabBox slab_box(const VolGeom& g, int c0, const float o[3], float sAmin, float sAmax,
                                            float sBmin, float sBmax, float dfirst) {
    constexpr int A = D == 0 ? 1 : 0;
    constexpr int B = D == 2 ? 1 : 2;
    const int dims[3] = {g.nx, g.ny, g.nz};
    SlabBox bx;
    // positions of the slab's samples along D: [c0, c0 + K) going up,
    // [c0 - K + 1, c0 + 1) going down, transverse extremes at the two ends.
    // Samples before a ray's entry point lie outside the volume along D (the
    // clamp maps them to the first plane); dfirst, the block's extreme first
    // sample along D, extends the near end over them.
    const float e0 = (DOWN ? fmaxf((float)(c0 + 1), dfirst) : fminf((float)c0, dfirst)) - o[D];
    // Depth: an oblique bundle crosses an axis-aligned box diagonally, so the
    // box's transverse extent grows by the largest |slope| per plane; the slab
    // ends where that growth reaches kSlabShear of the near-plane width.
    int K;
    {
        const float wA = fabsf(e0) * (sAmax - sAmin) + 2.0f, gA = fmaxf(fabsf(sAmin), fabsf(sAmax));
        const float wB = fabsf(e0) * (sBmax - sBmin) + 2.0f, gB = fmaxf(fabsf(sBmin), fabsf(sBmax));
        const float kf = kSlabShear * fminf(wA * __builtin_amdgcn_rcpf(gA), wB * __builtin_amdgcn_rcpf(gB));
        K = (int)fminf(fmaxf(kf, (float)kSlabKMin), (float)kSlabK);
    }
    const float e1 = (float)(DOWN ? c0 - K + 1 : c0 + K) - o[D];
    const float a0 = fminf(fminf(e0 * sAmin, e0 * sAmax), fminf(e1 * sAmin, e1 * sAmax)) + o[A];
    const float a1 = fmaxf(fmaxf(e0 * sAmin, e0 * sAmax), fmaxf(e1 * sAmin, e1 * sAmax)) + o[A];
    const float b0 = fminf(fminf(e0 * sBmin, e0 * sBmax), fminf(e1 * sBmin, e1 * sBmax)) + o[B];
    const float b1 = fmaxf(fmaxf(e0 * sBmin, e0 * sBmax), fmaxf(e1 * sBmin, e1 * sBmax)) + o[B];
    // Cells of the samples' clamped positions, plus the +1 taps.  A sample
    // lies within box_eps of its exact line, and the sampler clamps it to
    // [0, dims-1] before taking its cell; clamping is monotone, so the cells
    // of the slab's samples lie between the cells of the clamped, widened
    // extremes.  (Clipping only the integer bounds instead let a block whose
    // samples all lie beyond a face -- pre-entry samples of grazing rays,
    // clamped onto it -- get lo = dims above their cell dims-1, or a box one
    // cell wide at the low face missing the +1 tap.)
    const float eps = g.box_eps;
    const float bA = D == 0 ? g.by : g.bx, bB = D == 2 ? g.by : g.bz;
    const int loA = (int)floorf(fminf(fmaxf(a0 - eps, 0.0f), bA));
    const int loB = (int)floorf(fminf(fmaxf(b0 - eps, 0.0f), bB));
    const int hiA = (int)floorf(fminf(fmaxf(a1 + eps, 0.0f), bA)) + 1;
    const int hiB = (int)floorf(fminf(fmaxf(b1 + eps, 0.0f), bB)) + 1;
    const int nA = hiA - loA + 1;
    const int nB = hiB - loB + 1;
    if (nA > 0 && nB > 0) {
        // thinner slab when the box would not fit; its true box only shrinks,
        // so the extents above stay a valid (conservative) bound
        const int per_plane = PADDED ? (nA | 1) * nB : nA * nB;
        if (per_plane * (K + 1) > CAP) {
            K = (int)((float)CAP * __builtin_amdgcn_rcpf((float)per_plane)) - 1;
            while (K > 0 && per_plane * (K + 1) > CAP) --K;
            K = max(K, 1);
        }
    }
    // box along D: base planes of the slab and their +1 taps
    int dlo, dhi;
    if (DOWN) {
        dlo = max(c0 - K + 1, 0);
        dhi = min(c0 + 1, dims[D]);
        bx.cbeg = c0;
        bx.cend = c0 - K;
    } else {
        dlo = max(c0, 0);
        dhi = min(c0 + K, dims[D]);
        bx.cbeg = c0;
        bx.cend = c0 + K;
    }
    bx.lo[D] = dlo;
    bx.n[D] = dhi - dlo + 1;
    bx.lo[A] = loA;
    bx.n[A] = nA;
    bx.lo[B] = loB;
    bx.n[B] = nB;
    bx.px = bx.n[0] | 1;
    bx.pxy = (bx.px * bx.n[1]) | 1;
    if (PADDED)
        bx.staged = nA > 0 && nB > 0 && bx.n[D] > 0 && bx.pxy * bx.n[2] <= CAP &&
                    bx.px * bx.n[1] <= 256 * kStagePasses;  // the smallest block stages 256 per pass
    else
        bx.staged = nA > 0 && nB > 0 && bx.n[D] > 0 && nA <= CAP && nB <= CAP &&
                    nA * nB * bx.n[D] <= CAP;  // each factor bounded first: no int overflow
    bx.base = bx.lo[0] + bx.px * bx.lo[1] + bx.pxy * bx.lo[2];
    // Inner slab: the sampler's clamp to [0, dims-1] is the identity on all
    // of its samples.  Transversally the samples lie within eps of the
    // extremes a0..a1, b0..b1.  Along D a sample belongs to the slab when its
    // clamped coordinate c lies in [c0, c0+K) (up) or [c0-K+1, c0+1) (down);
    // with that range inside (0, bD) the clamp cannot have moved it (a
    // position below 0 clamps to 0, one above bD to bD, both outside the
    // range), so c is the position itself.
    const float bD = D == 0 ? g.bx : (D == 1 ? g.by : g.bz);
    bx.inner = bx.staged && a0 - eps >= 0.0f && a1 + eps <= bA && b0 - eps >= 0.0f && b1 + eps <= bB &&
               (DOWN ? (c0 - K >= 0 && (float)(c0 + 1) <= bD) : (c0 >= 1 && (float)(c0 + K) <= bD));
    return bx;
}

// Cooperative load of the box into LDS with gfx950's direct-to-LDS buffer
// loads (buffer_load_dword ... lds): each lane's dword lands at the wave's
// LDS base (M0) + 4 * lane, with no register round trip and no LDS write
// instruction.  The box's x-y window is linearised with the LDS row pitch px
// (element e = col + px * row, col < px), so the 64 lanes of a wave holding
// window elements e0 .. e0+63 fill 64 consecutive LDS dwords of a plane.
// Each thread owns fixed window elements (tid + NT*m) and issues the loads
// of all the box's planes back to back (plane steps move only the
// wave-uniform offsets); the caller waits for them and synchronises before
// sampling.  The padding column (col = n[0] when n[0] is even) loads the next
// voxel along x, and indices up to dims (the +1 taps of the last voxel,
// weight 0) read the volume's zero tail or the next row: every address stays
// inside the allocation, which the buffer range check also enforces.
template <int NT>
__device__ __forceinline__ void slab_stage(const VolView& vv, const VolGeom& g, const SlabBox& bx, float* box) {
    const int tid = threadIdx.x;
    const int wave0 = __builtin_amdgcn_readfirstlane(tid & ~63);
    const int px = bx.px, nbz = bx.n[2];
    const int npl = px * bx.n[1];
    // e / px for e < 2^16 (px == 1: the rows are taken directly)
    const unsigned magic = px > 1 ? (unsigned)(0xFFFFFFFFull / (unsigned)px) + 1u : 0u;
    const unsigned nx = (unsigned)g.nx, nxy = (unsigned)(g.nx * g.ny);
    const unsigned zbase = 4u * nxy * (unsigned)bx.lo[2];
#pragma unroll
    for (int m = 0; m < kStagePasses * 256 / NT; ++m) {
        if (NT * m >= npl) break;  // uniform
        const int e = tid + NT * m;
        const unsigned row = px > 1 ? __umulhi((unsigned)e, magic) : (unsigned)e;
        const unsigned col = (unsigned)e - row * (unsigned)px;
        const unsigned voff = 4u * ((unsigned)bx.lo[0] + col + nx * ((unsigned)bx.lo[1] + row));
        if (e < npl) {
            typedef __attribute__((address_space(3))) void lds_void;
            lds_void* dst = (lds_void*)(box + wave0 + NT * m);
#pragma unroll 4
            for (int z = 0; z < nbz; ++z)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(vv.rsrc, (lds_void*)((float __attribute__((address_space(3)))*)dst + bx.pxy * z),
                                                         4, voff, zbase + 4u * nxy * (unsigned)z, 0, 0);
        }
    }
}

// Waits for the direct-to-LDS loads of slab_stage (vmcnt(0); expcnt and
// lgkmcnt left at their maxima).
__device__ __forceinline__ void slab_stage_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Samples of one slab from its staged box, without a per-sample box test
// (every sample of the wave's rays lies in the box).  A lane leaves the loop
// at its first sample past the slab, or at max_samples, or when its ray ends
// (nlim = -1).  Without COUNT, an ended ray's lane gets an unreachable plane
// limit instead (the loop still ends: every sample advances t towards a
// finite segment end).  CLAMP = false on inner slabs, where the clamp to
// [0, dims-1] is the identity (SlabBox::inner).
template <int D, bool DOWN, bool COUNT, bool CLAMP>
__device__ __forceinline__ void slab_walk(RayGen& G, const VolGeom& g, const SlabBox& bx, const float* box, float lim,
                                          float& nlim, float& llim, float& nf, float& density) {
    constexpr float kEnded = DOWN ? INFINITY : -INFINITY;
    const char* bb = (const char*)(box - bx.base);
    const unsigned px4 = 4u * (unsigned)bx.px, pxy4 = 4u * (unsigned)bx.pxy;
    const float sdw = g.sd;
#pragma unroll 1
    for (;;) {
        const float cx = CLAMP ? __builtin_amdgcn_fmed3f(G.pend[0], 0.0f, g.bx) : G.pend[0];
        const float cy = CLAMP ? __builtin_amdgcn_fmed3f(G.pend[1], 0.0f, g.by) : G.pend[1];
        const float cz = CLAMP ? __builtin_amdgcn_fmed3f(G.pend[2], 0.0f, g.bz) : G.pend[2];
        const float cc[3] = {cx, cy, cz};
        if (COUNT) {
            if (!((DOWN ? cc[D] >= lim : cc[D] < lim) && nf < nlim)) break;
        } else if (!(DOWN ? cc[D] >= llim : cc[D] < llim)) {
            break;
        }
        const unsigned off =
            ((unsigned)(int)cx << 2) + __umul24((unsigned)(int)cy, px4) + __umul24((unsigned)(int)cz, pxy4);
        const float* b0 = (const float*)(bb + off);
        Taps T;
        T.fx = __builtin_amdgcn_fractf(cx);
        T.fy = __builtin_amdgcn_fractf(cy);
        T.fz = __builtin_amdgcn_fractf(cz);
        // The fractions are taken here, before the position advances: left
        // to the compiler they sink below the update and each iteration ends
        // with three register copies of the position (an empty asm statement,
        // no instruction).
        asm volatile("" : "+v"(T.fx), "+v"(T.fy), "+v"(T.fz));
        T.r00 = make_float2(b0[0], b0[1]);
        T.r10 = make_float2(b0[bx.px], b0[bx.px + 1]);
        T.r01 = make_float2(b0[bx.pxy], b0[bx.pxy + 1]);
        T.r11 = make_float2(b0[bx.pxy + bx.px], b0[bx.pxy + bx.px + 1]);
        if (COUNT) nf += 1.0f;
        G.t += sdw;
        G.pend[0] += G.step[0];
        G.pend[1] += G.step[1];
        G.pend[2] += G.step[2];
        if (__builtin_expect(!(G.t < G.tb), 0)) {
            if (!raygen_next_segment(G, sdw)) {
                nlim = -1.0f;
                llim = kEnded;
            }
        }
        density = fmaf(lerp_taps(T), sdw, density);
    }
}

// COUNT: the sample count is an output (FP mode); the SART pass needs only
// the line integral, so its unchecked loop keeps no count.
template <int D, bool DOWN, int NT, bool COUNT, int CAP>
__device__ __forceinline__ float slab_ray(const WorkItem& it, const VolGeom& g, const Ray& r, bool hit, int& nsamp,
                                          float* box, float* red, int* slots) {
    constexpr int A = D == 0 ? 1 : 0;
    constexpr int B = D == 2 ? 1 : 2;
    const VolView vv = vol_view(it.vol, g);
    const float bnd[3] = {g.bx, g.by, g.bz};
    RayGen G;
    bool entry_alone;
    bool have = raygen_init(G, r, g, hit, entry_alone);
    float density = 0.0f;
    int n = 0;
    if (entry_alone) {  // pre-entry case: the entry sample first, on its own
        float e0, e1, e2;
        {
#pragma clang fp contract(off)
            e0 = r.o[0] + r.d[0] * r.tnear;
            e1 = r.o[1] + r.d[1] * r.tnear;
            e2 = r.o[2] + r.d[2] * r.tnear;
        }
        Taps T;
        fetch_taps(vv, e0, e1, e2, T);
        density = lerp_taps(T) * g.sd;
        n = 1;
    } else if (have) {
        // The entry sample coincides with the pending one and counts twice
        // (weight 2); its first copy is taken here, so every sample of the
        // marching loops has weight 1.
        Taps T;
        fetch_taps(vv, G.pend[0], G.pend[1], G.pend[2], T);
        density = lerp_taps(T) * g.sd;
        n = 1;
        G.w = 1.0f;
    }
    // Slopes of the rays that march with the slabs, and the first base plane.
    const bool fwd = DOWN ? r.d[D] < 0.0f : r.d[D] > 0.0f;
    float q[6];
    const bool qmax[6] = {false, true, false, true, DOWN, DOWN};
    {
        const float sa = r.d[A] / r.d[D], sb = r.d[B] / r.d[D];
        const bool use = have && fwd;
        q[0] = use ? sa : INFINITY;
        q[1] = use ? sa : -INFINITY;
        q[2] = use ? sb : INFINITY;
        q[3] = use ? sb : -INFINITY;
        const float pD = floorf(__builtin_amdgcn_fmed3f(G.pend[D], 0.0f, bnd[D]));
        q[4] = have ? pD : (DOWN ? -INFINITY : INFINITY);
        q[5] = use ? G.pend[D] : (DOWN ? -INFINITY : INFINITY);  // unclamped
    }
    block_minmax<NT / 64, 6>(q, qmax, red);
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, r.o[c])));
    const bool slopes_ok = q[0] <= q[1] && q[2] <= q[3];
    const bool any_ray = DOWN ? q[4] > -INFINITY : q[4] < INFINITY;  // some lane of the block samples
    int c0 = (int)fminf(fmaxf(q[4], -1.0f), (float)g.nx + g.ny + g.nz);
    const int dimD = D == 0 ? g.nx : (D == 1 ? g.ny : g.nz);
    const float sdw = g.sd;
    // Every sample of a ray that marches with the slabs lies inside the box
    // of its slab.  The box spans the block's slope range from the common eye
    // over the slab's extent along D (the first slab's extended back to the
    // samples before the entry points), widened by box_eps: a sample deviates
    // from its exact line by the rounding of at most max_samples additions,
    // which box_eps bounds, so its cell never leaves the box.  The bound
    // holds while |pend| <= 2*max(dims): a ray's positions lie between its
    // first sample and its exit point, so the first sample is checked.  Waves
    // whose sampling lanes all march forward within that bound take the loop
    // without a per-sample box test.
    const float cmax = 2.0f * (fmaxf(g.bx, fmaxf(g.by, g.bz)) + 1.0f);
    const bool near = fmaxf(fabsf(G.pend[0]), fmaxf(fabsf(G.pend[1]), fabsf(G.pend[2]))) <= cmax;
    const bool wave_in_box = kSlabUncheckedPath && __builtin_amdgcn_ballot_w64(have && !(fwd && near)) == 0;
    const float maxn = (float)g.max_samples;
    float nf = (float)n;
    // The boxes are uniform arithmetic on the block's slopes, which gfx950
    // can only do on the vector ALUs: one wave computes each box and hands it
    // to the others through an LDS slot.  The first slab's box (which alone
    // reaches back to samples before an entry point, dfirst) is computed
    // here; the box of slab s+1 depends only on where slab s ends, so wave
    // (s mod NT/64) computes it while slab s is sampled, and the barrier that
    // ends slab s publishes it.
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (any_ray) {
        if (wv == 0) {
            SlabBox b0 = slab_box<D, DOWN, CAP>(g, c0, o, q[0], q[1], q[2], q[3], q[5]);
            if (!slopes_ok) b0.staged = false;
            box_put(slots, b0);
        }
        __syncthreads();
    }
#pragma unroll 1
    for (int guard = 0; any_ray && guard <= dimD + 1; ++guard) {
        const SlabBox bx = box_get<DOWN>(slots + kBoxInts * (guard & 1));
        if (bx.staged) slab_stage<NT>(vv, g, bx, box);
        if (wv == (guard & (NT / 64 - 1))) {
            SlabBox nb = slab_box<D, DOWN, CAP>(g, bx.cend, o, q[0], q[1], q[2], q[3], DOWN ? -INFINITY : INFINITY);
            if (!slopes_ok) nb.staged = false;
            box_put(slots + kBoxInts * ((guard + 1) & 1), nb);
        }
        slab_stage_wait();
        __syncthreads();
        if (bx.staged && wave_in_box) {
            // Samples of the slab are those whose clamped base plane p = (int)cD
            // satisfies p < cend (up) or p > cend (down); cD >= 0, so the test
            // runs on the float.
            const float lim = DOWN ? (float)(bx.cend + 1) : (float)bx.cend;
            constexpr float kEnded = DOWN ? INFINITY : -INFINITY;
            float nlim = have ? maxn : -1.0f;
            float llim = have ? lim : kEnded;
            if (bx.inner) slab_walk<D, DOWN, COUNT, false>(G, g, bx, box, lim, nlim, llim, nf, density);
            else slab_walk<D, DOWN, COUNT, true>(G, g, bx, box, lim, nlim, llim, nf, density);
            have = COUNT ? nf < nlim : llim != kEnded;
            c0 = bx.cend;
            if (!__syncthreads_or(have)) break;
            continue;
        }
        n = (int)nf;
        // One wave-uniform loop per slab: each iteration every lane whose
        // pending sample lies in the slab (or behind it) takes it.
#pragma unroll 1
        for (int it = 0; it <= g.max_samples; ++it) {
            const float cx = __builtin_amdgcn_fmed3f(G.pend[0], 0.0f, g.bx);
            const float cy = __builtin_amdgcn_fmed3f(G.pend[1], 0.0f, g.by);
            const float cz = __builtin_amdgcn_fmed3f(G.pend[2], 0.0f, g.bz);
            const int ix = (int)cx, iy = (int)cy, iz = (int)cz;
            const int ic[3] = {ix, iy, iz};
            const int p = ic[D];
            const bool take = have && (DOWN ? p > bx.cend : p < bx.cend);
            if (__builtin_amdgcn_ballot_w64(take) == 0) break;
            if (take) {
                Taps T;
                T.fx = __builtin_amdgcn_fractf(cx);
                T.fy = __builtin_amdgcn_fractf(cy);
                T.fz = __builtin_amdgcn_fractf(cz);
                const bool in_slab = DOWN ? p <= bx.cbeg : p >= bx.cbeg;
                const bool in_box = bx.staged && in_slab && (unsigned)(ic[A] - bx.lo[A]) < (unsigned)(bx.n[A] - 1) &&
                                    (unsigned)(ic[B] - bx.lo[B]) < (unsigned)(bx.n[B] - 1);
                if (__builtin_expect(in_box, 1)) {
                    const int li = ix + __mul24(iy, bx.px) + __mul24(iz, bx.pxy) - bx.base;
                    const float* b0 = box + li;
                    T.r00 = make_float2(b0[0], b0[1]);
                    T.r10 = make_float2(b0[bx.px], b0[bx.px + 1]);
                    T.r01 = make_float2(b0[bx.pxy], b0[bx.pxy + 1]);
                    T.r11 = make_float2(b0[bx.pxy + bx.px], b0[bx.pxy + bx.px + 1]);
                } else {
                    const unsigned off =
                        ((unsigned)ix << 2) + __umul24((unsigned)iy, vv.nx4) + __umul24((unsigned)iz, vv.sxy4);
                    T.r00 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(vv.rsrc, off, 0, 0));
                    T.r10 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(vv.rsrc, off, vv.nx4, 0));
                    T.r01 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(vv.rsrc, off, vv.sxy4, 0));
                    T.r11 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(vv.rsrc, off, vv.sxynx4, 0));
                }
                const float w = G.w;
                n += (int)w;
                G.w = 1.0f;
                G.t += sdw;
                G.pend[0] += G.step[0];
                G.pend[1] += G.step[1];
                G.pend[2] += G.step[2];
                if (__builtin_expect(!(G.t < G.tb), 0)) have = raygen_next_segment(G, sdw);
                have = have && n < g.max_samples;
                density = fmaf(lerp_taps(T), w * sdw, density);
            }
        }
        nf = (float)n;
        c0 = bx.cend;
        if (!__syncthreads_or(have)) break;
    }
    nsamp = (int)nf;
    return density;
}

template <int MODE, int TW, int TH, int CAP>
__global__ __launch_bounds__(TW* TH) void k_fp_slab(const WorkItem* __restrict__ items, VolGeom g, float factor) {
    constexpr int NT = TW * TH;
    __shared__ float box[CAP];
    __shared__ float red[6 * NT / 64];
    __shared__ int slots[2 * kBoxInts];
    int i, j;
    int item;
    const bool valid = tile_pixel<TW, TH>(g, i, j, item);
    const WorkItem& it = items[item];
    const ViewRec& v = *it.view;
    const int march = v.march & 7;
    Ray r;
    produce_ray(v, g, (float)(valid ? i : 0), (float)(valid ? j : 0), r);
    const bool hit = valid && r.tnear < r.tfar;
    int ns = 0;
    float value;
    switch (march) {
        case 0: value = slab_ray<0, false, NT, MODE == FP_MODE_FP, CAP>(it, g, r, hit, ns, box, red, slots); break;
        case 4: value = slab_ray<0, true, NT, MODE == FP_MODE_FP, CAP>(it, g, r, hit, ns, box, red, slots); break;
        case 1: value = slab_ray<1, false, NT, MODE == FP_MODE_FP, CAP>(it, g, r, hit, ns, box, red, slots); break;
        case 5: value = slab_ray<1, true, NT, MODE == FP_MODE_FP, CAP>(it, g, r, hit, ns, box, red, slots); break;
        case 6: value = slab_ray<2, true, NT, MODE == FP_MODE_FP, CAP>(it, g, r, hit, ns, box, red, slots); break;
        default: value = slab_ray<2, false, NT, MODE == FP_MODE_FP, CAP>(it, g, r, hit, ns, box, red, slots); break;
    }
    if (!valid) return;
    const int k = i + j * g.W;
    if (MODE == FP_MODE_FP) {
        if (hit && it.raylen) gptr(it.raylen)[k] = r.tfar - r.tnear;
        if (it.ax) gptr(it.ax)[k] = hit ? value : 0.0f;
        if (it.weight) gptr(it.weight)[k] = (float)ns;
    } else {
        // SART correction (ST_Tomography.cpp:667-676); pY is zero when read.
#pragma clang fp contract(off)
        float len;
        gfloat* rl = gptr(it.raylen);
        if (hit) {
            len = r.tfar - r.tnear;
            rl[k] = len;
        } else {
            len = rl[k];  // stale value of the last view whose ray k hit
        }
        const float num = factor * gptr(it.b)[k] - factor * (hit ? value : 0.0f);
        const float den = factor * len + 1.0f;
        gptr(it.corr)[k] = num / den;
    }
}

// ---------------------------------------------------------------------------
// Ray-driven BP scatter through the slab boxes (the BP-mode pass of
// GenerateImage_MT, ST_Tomography.cpp:733-739; per sample and tap
// SingleScalarVolume.core.h:73-84,128-143: contribution w*corr and weight w).
// The addends are SinkFixed's exact 64-bit fixed-point values.  A block
// marches its 16x16 rays through the same slabs as the slab projector and
// adds each slab's addends into an LDS copy of the slab box with 64-bit LDS
// atomics; after the slab every non-zero box voxel is added to the global
// sums with one atomic per sum.  Integer addition is associative, so the
// sums equal those of one global atomic per sample and tap (k_fp<RAYBP>) bit
// for bit, with about a fifth of the global atomics.  Samples outside the
// box (waves that do not march with the slabs, boxes too large for LDS) go
// straight to the global sums.
constexpr int kRayBpCap = 2048;  // voxels of an LDS slab box: two 64-bit sums each (32 KB)

// SingleScalarVolume's eight trilinear weights of a clamped position, the
// products in tri_corners' order.
__device__ __forceinline__ void tri_weights(float cx, float cy, float cz, float w[8]) {
    const float t0 = __builtin_amdgcn_fractf(cx), t1 = __builtin_amdgcn_fractf(cy), t2 = __builtin_amdgcn_fractf(cz);
    const float m0 = 1.0f - t0, m1 = 1.0f - t1, m2 = 1.0f - t2;
    w[0] = m0 * m1 * m2;
    w[1] = t0 * m1 * m2;
    w[2] = m0 * t1 * m2;
    w[3] = t0 * t1 * m2;
    w[4] = m0 * m1 * t2;
    w[5] = t0 * m1 * t2;
    w[6] = m0 * t1 * t2;
    w[7] = t0 * t1 * t2;
}

// One sample's eight taps into the LDS box at index li (cell (x,y,z) - lo);
// sx, sy, sz: the box's strides along the volume axes.
__device__ __forceinline__ void lds_scatter(unsigned long long* lc, unsigned long long* lw, int li, int sx, int sy,
                                            int sz, const float w[8], float val, double cscale, double wscale) {
    const int off[8] = {0, sx, sy, sy + sx, sz, sz + sx, sz + sy, sz + sy + sx};
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        __hip_atomic_fetch_add(lc + li + off[n], fixed_of(w[n] * val, cscale), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(lw + li + off[n], fixed_of(w[n], wscale), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

template <int D, bool DOWN, int NT, int CAP>
__device__ __forceinline__ void slab_scatter(const VolGeom& g, const Ray& r, bool hit, const SinkFixed& sink,
                                             unsigned long long* lc, unsigned long long* lw, float* red, int* slots) {
    constexpr int A = D == 0 ? 1 : 0;
    constexpr int B = D == 2 ? 1 : 2;
    const float bnd[3] = {g.bx, g.by, g.bz};
    RayGen G;
    bool entry_alone;
    bool have = raygen_init(G, r, g, hit, entry_alone);
    int n = 0;
    if (entry_alone) {  // pre-entry case: the entry sample first, on its own
        float e0, e1, e2;
        {
#pragma clang fp contract(off)
            e0 = r.o[0] + r.d[0] * r.tnear;
            e1 = r.o[1] + r.d[1] * r.tnear;
            e2 = r.o[2] + r.d[2] * r.tnear;
        }
        sink(g, e0, e1, e2);
        n = 1;
    } else if (have) {
        // the entry sample coincides with the pending one and is scattered
        // twice (CastingRay's entry sample, then the grid's first); the first
        // copy here, so every sample of the loops below has weight 1
        sink(g, G.pend[0], G.pend[1], G.pend[2]);
        n = 1;
        G.w = 1.0f;
    }
    // slopes of the rays that march with the slabs and the first base plane
    // (as slab_ray)
    const bool fwd = DOWN ? r.d[D] < 0.0f : r.d[D] > 0.0f;
    float q[6];
    const bool qmax[6] = {false, true, false, true, DOWN, DOWN};
    {
        const float sa = r.d[A] / r.d[D], sb = r.d[B] / r.d[D];
        const bool use = have && fwd;
        q[0] = use ? sa : INFINITY;
        q[1] = use ? sa : -INFINITY;
        q[2] = use ? sb : INFINITY;
        q[3] = use ? sb : -INFINITY;
        const float pD = floorf(__builtin_amdgcn_fmed3f(G.pend[D], 0.0f, bnd[D]));
        q[4] = have ? pD : (DOWN ? -INFINITY : INFINITY);
        q[5] = use ? G.pend[D] : (DOWN ? -INFINITY : INFINITY);
    }
    block_minmax<NT / 64, 6>(q, qmax, red);
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, r.o[c])));
    const bool slopes_ok = q[0] <= q[1] && q[2] <= q[3];
    const bool any_ray = DOWN ? q[4] > -INFINITY : q[4] < INFINITY;
    int c0 = (int)fminf(fmaxf(q[4], -1.0f), (float)g.nx + g.ny + g.nz);
    const int dimD = D == 0 ? g.nx : (D == 1 ? g.ny : g.nz);
    const float sdw = g.sd;
    const float cmax = 2.0f * (fmaxf(g.bx, fmaxf(g.by, g.bz)) + 1.0f);
    const bool near = fmaxf(fabsf(G.pend[0]), fmaxf(fabsf(G.pend[1]), fabsf(G.pend[2]))) <= cmax;
    const bool wave_in_box = kSlabUncheckedPath && __builtin_amdgcn_ballot_w64(have && !(fwd && near)) == 0;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const unsigned nx = (unsigned)g.nx, nxy = (unsigned)(g.nx * g.ny);
    if (any_ray) {
        if (wv == 0) {
            SlabBox b0 = slab_box<D, DOWN, CAP, false>(g, c0, o, q[0], q[1], q[2], q[3], q[5]);
            if (!slopes_ok) b0.staged = false;
            box_put(slots, b0);
        }
        __syncthreads();
    }
#pragma unroll 1
    for (int guard = 0; any_ray && guard <= dimD + 1; ++guard) {
        const SlabBox bx = box_get<DOWN>(slots + kBoxInts * (guard & 1));
        if (wv == (guard & (NT / 64 - 1))) {
            SlabBox nb = slab_box<D, DOWN, CAP, false>(g, bx.cend, o, q[0], q[1], q[2], q[3], DOWN ? -INFINITY : INFINITY);
            if (!slopes_ok) nb.staged = false;
            box_put(slots + kBoxInts * ((guard + 1) & 1), nb);
        }
        __syncthreads();  // the box is published; the previous slab's flush has cleared the LDS sums
        // dense box layout, the slab axis D slowest: index of cell c is
        // ((c[D] - lo[D]) * n[B] + c[B] - lo[B]) * n[A] + c[A] - lo[A]
        int st[3];
        st[A] = 1;
        st[B] = bx.n[A];
        st[D] = bx.n[A] * bx.n[B];
        const int sx = st[0], sy = st[1], sz = st[2];
        const int base = bx.lo[0] * sx + bx.lo[1] * sy + bx.lo[2] * sz;
        if (bx.staged && wave_in_box) {
            // every sample of the slab lies in the box (as slab_walk)
            const float lim = DOWN ? (float)(bx.cend + 1) : (float)bx.cend;
            constexpr float kEnded = DOWN ? INFINITY : -INFINITY;
            float llim = have ? lim : kEnded;
#pragma unroll 1
            for (;;) {
                const float cx = __builtin_amdgcn_fmed3f(G.pend[0], 0.0f, g.bx);
                const float cy = __builtin_amdgcn_fmed3f(G.pend[1], 0.0f, g.by);
                const float cz = __builtin_amdgcn_fmed3f(G.pend[2], 0.0f, g.bz);
                const float cc[3] = {cx, cy, cz};
                if (!(DOWN ? cc[D] >= llim : cc[D] < llim)) break;
                const int li = __mul24((int)cx, sx) + __mul24((int)cy, sy) + __mul24((int)cz, sz) - base;
                float w[8];
                tri_weights(cx, cy, cz, w);
                lds_scatter(lc, lw, li, sx, sy, sz, w, sink.val, sink.cscale, sink.wscale);
                G.t += sdw;
                G.pend[0] += G.step[0];
                G.pend[1] += G.step[1];
                G.pend[2] += G.step[2];
                if (__builtin_expect(!(G.t < G.tb), 0)) {
                    if (!raygen_next_segment(G, sdw)) llim = kEnded;
                }
            }
            have = llim != kEnded;
        } else {
            // one wave-uniform loop per slab: each iteration every lane whose
            // pending sample lies in the slab (or behind it) takes it (as
            // slab_ray's checked loop)
#pragma unroll 1
            for (int itn = 0; itn <= g.max_samples; ++itn) {
                const float cx = __builtin_amdgcn_fmed3f(G.pend[0], 0.0f, g.bx);
                const float cy = __builtin_amdgcn_fmed3f(G.pend[1], 0.0f, g.by);
                const float cz = __builtin_amdgcn_fmed3f(G.pend[2], 0.0f, g.bz);
                const int ix = (int)cx, iy = (int)cy, iz = (int)cz;
                const int ic[3] = {ix, iy, iz};
                const int p = ic[D];
                const bool take = have && (DOWN ? p > bx.cend : p < bx.cend);
                if (__builtin_amdgcn_ballot_w64(take) == 0) break;
                if (take) {
                    const bool in_slab = DOWN ? p <= bx.cbeg : p >= bx.cbeg;
                    const bool in_box = bx.staged && in_slab && (unsigned)(ic[A] - bx.lo[A]) < (unsigned)(bx.n[A] - 1) &&
                                        (unsigned)(ic[B] - bx.lo[B]) < (unsigned)(bx.n[B] - 1);
                    if (in_box) {
                        float w[8];
                        tri_weights(cx, cy, cz, w);
                        lds_scatter(lc, lw, __mul24(ix, sx) + __mul24(iy, sy) + __mul24(iz, sz) - base, sx, sy, sz, w,
                                    sink.val, sink.cscale, sink.wscale);
                    } else {
                        sink(g, G.pend[0], G.pend[1], G.pend[2]);
                    }
                    ++n;
                    G.w = 1.0f;
                    G.t += sdw;
                    G.pend[0] += G.step[0];
                    G.pend[1] += G.step[1];
                    G.pend[2] += G.step[2];
                    if (__builtin_expect(!(G.t < G.tb), 0)) have = raygen_next_segment(G, sdw);
                    have = have && n < g.max_samples;
                }
            }
        }
        const bool more = __syncthreads_or(have);  // every LDS addend of the slab is in
        if (bx.staged) {
            // flush: rows (B, D) of the box to the waves, A to the lanes; each
            // non-zero voxel once into the global sums, then cleared (taps past
            // the volume's last cell carry weight 0 and stay zero)
            const int nA = bx.n[A], nB = bx.n[B], rows = nB * bx.n[D];
            const unsigned gst[3] = {1u, nx, nxy};  // global strides per volume axis
            gu64* ac = (gu64*)sink.acc_c;
            gu64* aw = (gu64*)sink.acc_w;
#pragma unroll 1
            for (int row = wv; row < rows; row += NT / 64) {
                const int jb = row % nB, jd = row / nB;  // wave-uniform
                const int e0 = row * nA;
                const unsigned q0 = gst[A] * (unsigned)bx.lo[A] + gst[B] * (unsigned)(bx.lo[B] + jb) +
                                    gst[D] * (unsigned)(bx.lo[D] + jd);
                for (int ia = lane; ia < nA; ia += 64) {
                    const unsigned long long cv = lc[e0 + ia], wvv = lw[e0 + ia];
                    if ((cv | wvv) != 0ull) {
                        const unsigned q = q0 + gst[A] * (unsigned)ia;
                        __hip_atomic_fetch_add(ac + q, cv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_fetch_add(aw + q, wvv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        lc[e0 + ia] = 0ull;
                        lw[e0 + ia] = 0ull;
                    }
                }
            }
        }
        c0 = bx.cend;
        if (!more) break;
    }
}

template <int TW, int TH>
__global__ __launch_bounds__(TW* TH) void k_raybp_slab(const WorkItem* __restrict__ items, VolGeom g) {
    constexpr int NT = TW * TH;
    __shared__ unsigned long long acc[2 * kRayBpCap];
    __shared__ float red[6 * NT / 64];
    __shared__ int slots[2 * kBoxInts];
    for (int e = threadIdx.x; e < 2 * kRayBpCap; e += NT) acc[e] = 0ull;  // published by the first barrier below
    int i, j;
    int item;
    const bool valid = tile_pixel<TW, TH>(g, i, j, item);
    const WorkItem& it = items[item];
    const ViewRec& v = *it.view;
    Ray r;
    produce_ray(v, g, (float)(valid ? i : 0), (float)(valid ? j : 0), r);
    const bool hit = valid && r.tnear < r.tfar;
    const SinkFixed sink{it.acc_c, it.acc_w, valid ? gptr(it.corr)[i + j * g.W] : 0.0f, ldexp(1.0, *it.acc_shift),
                         ldexp(1.0, g.acc_wshift)};
    unsigned long long* lc = acc;
    unsigned long long* lw = acc + kRayBpCap;
    switch (v.march & 7) {
        case 0: slab_scatter<0, false, NT, kRayBpCap>(g, r, hit, sink, lc, lw, red, slots); break;
        case 4: slab_scatter<0, true, NT, kRayBpCap>(g, r, hit, sink, lc, lw, red, slots); break;
        case 1: slab_scatter<1, false, NT, kRayBpCap>(g, r, hit, sink, lc, lw, red, slots); break;
        case 5: slab_scatter<1, true, NT, kRayBpCap>(g, r, hit, sink, lc, lw, red, slots); break;
        case 6: slab_scatter<2, true, NT, kRayBpCap>(g, r, hit, sink, lc, lw, red, slots); break;
        default: slab_scatter<2, false, NT, kRayBpCap>(g, r, hit, sink, lc, lw, red, slots); break;
    }
}

// x <-> z transpose of the volumes of the x-dominant views in a work table
// (index x + nx*(y + ny*z) -> z + nz*(y + ny*x)) through a 32x33 LDS tile.
__global__ __launch_bounds__(256) void k_transpose_xz(const WorkItem* __restrict__ items, VolGeom g) {
    const WorkItem& it = items[blockIdx.z / g.ny];
    if (!(it.view->march & 8)) return;  // uniform per block
    const int y = blockIdx.z % g.ny;
    __shared__ float tile[32][33];
    const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const gfloat* in = gptr((const float*)it.vol);
    gfloat* out = gptr(it.volT);
    const size_t sxy = (size_t)g.nx * g.ny;
#pragma unroll
    for (int jj = ty; jj < 32; jj += 8) {
        const int x = x0 + tx, z = z0 + jj;
        if (x < g.nx && z < g.nz) tile[jj][tx] = in[(size_t)x + (size_t)g.nx * y + sxy * z];
    }
    __syncthreads();
    const size_t szy = (size_t)g.nz * g.ny;
#pragma unroll
    for (int jj = ty; jj < 32; jj += 8) {
        const int x = x0 + jj, z = z0 + tx;
        if (x < g.nx && z < g.nz) out[(size_t)z + (size_t)g.nz * y + szy * x] = tile[tx][jj];
    }
}

// Catmull-Rom weights of cubicInterpolate (ST_Tomography.cpp:183-186),
// identical polynomial to CImg's cubic (CImg.h:13429-13520).
__device__ __forceinline__ void cr_weights(float t, float w[4]) {
#pragma clang fp contract(off)
    const float t2 = t * t, t3 = t2 * t;
    w[0] = 0.5f * (-t + 2.0f * t2 - t3);
    w[1] = 1.0f + 0.5f * (-5.0f * t2 + 3.0f * t3);
    w[2] = 0.5f * (t + 4.0f * t2 - 3.0f * t3);
    w[3] = 0.5f * (-t2 + t3);
}

// Detector position of a voxel centre with the reference's arithmetic
// (ST_Tomography.cpp:698-703): float matrix-vector product without
// contraction, IEEE division, m_halfFrustumWidth/m_ds in double.
__device__ __noinline__ float2 exact_detector_position(const ViewRec& v, float fx, float fy, float fz) {
#pragma clang fp contract(off)
    const float* M = v.mvv2m;
    const float e0 = ((M[0] * fx + M[1] * fy) + M[2] * fz) + M[3];
    const float e1 = ((M[4] * fx + M[5] * fy) + M[6] * fz) + M[7];
    const float e2 = ((M[8] * fx + M[9] * fy) + M[10] * fz) + M[11];
    const float tx = (e0 / e2) * v.sdd;
    const float ty = (e1 / e2) * v.sdd;
    const float ux = (float)(v.halfW - (double)tx);
    const float uy = (float)(v.halfH - (double)ty);
    return make_float2((float)((double)ux / v.ds), (float)((double)uy / v.ds));
}

// Voxel-driven backprojection (ST_Tomography.cpp:687-731): project the voxel
// centre, bicubic (Catmull-Rom) sample of the correction image, x += alpha*v,
// clamp [0, G_Max].  Atomic-free, one thread per voxel column segment.
//
// A 256-thread block owns a 32 x 8 x 16 voxel tile (a wave's two half-waves
// take rows y0 + 2w and y0 + 2w + 1, lane = x within a half-wave, each
// thread 16 consecutive z).  The tile is a box, so the detector positions of
// its voxels lie inside the bounding box of its eight corners' projections;
// the block stages that box of the correction image in LDS with row loads and
// reads every voxel's 4x4 window from there: 16 taps per voxel from LDS
// instead of four unaligned 16-byte global rows.  Stacking the tile along z
// keeps the box small: it grows with the tile's x and y extent, not with
// its 16 z steps.
//
// Blocks whose box lies at least two pixels inside the [2, W-2) x [2, H-2)
// window and fits in LDS take a fast path: every voxel is updated and its
// window is inside the box (one pixel of margin absorbs the rounding of the
// corner projections).  Every other block takes the general path: per-voxel
// skip test with the reference's own arithmetic near the window edges, and a
// window read from global memory when it is not inside the staged box.
constexpr int kBpTileX = 32;  // 64 x 4 staged 22% fewer blocks at 512^3 / 1024^2 (the rest read global memory);
constexpr int kBpTileY = 8;   // 16 x 16 was slower at both sizes
constexpr int kBpVoxZ = 16;  // voxels per thread along z (8 and 32 measured slower)
constexpr int kBpLdsFloats = 4096;  // 16 KB staged detector footprint per block (22 KB: the iteration 3% slower, as
                                    // fewer blocks share a CU with the other streams' projector blocks; 12 KB: more fall back)
constexpr int kBpMaxCols = 192;     // three 64-lane column chunks
constexpr int kBpMaxRows = 64;      // 16 rows per wave

// Catmull-Rom weights (Horner form of cr_weights' polynomials).
__device__ __forceinline__ void cr_weights_h(float t, float w[4]) {
    const float t2 = t * t;
    w[0] = t * fmaf(t, fmaf(-0.5f, t, 1.0f), -0.5f);
    w[1] = fmaf(t2, fmaf(1.5f, t, -2.5f), 1.0f);
    w[2] = t * fmaf(t, fmaf(-1.5f, t, 2.0f), 0.5f);
    w[3] = t2 * fmaf(0.5f, t, -0.5f);
}

__device__ __forceinline__ float bp_bicubic(float dx, float dy, const float4 rows[4]) {
    float wy[4], wx[4];
    cr_weights_h(dy, wy);
    cr_weights_h(dx, wx);
    float h[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) h[b] = fmaf(wx[3], rows[b].w, fmaf(wx[2], rows[b].z, fmaf(wx[1], rows[b].y, wx[0] * rows[b].x)));
    return fmaf(wy[3], h[3], fmaf(wy[2], h[2], fmaf(wy[1], h[1], wy[0] * h[0])));
}

__global__ __launch_bounds__(256) void k_bp_voxel(const WorkItem* __restrict__ items, VolGeom g, PdConst pc) {
    __shared__ float win[kBpLdsFloats];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nzt = (g.nz + kBpVoxZ - 1) / kBpVoxZ;
    const int x0 = blockIdx.x * kBpTileX, y0 = blockIdx.y * kBpTileY;
    const int z0 = (blockIdx.z % nzt) * kBpVoxZ;
    const WorkItem& it = items[blockIdx.z / nzt];
    const ViewRec& v = *it.view;
    const float* A = v.bp;
    const int xe = min(x0 + kBpTileX, g.nx) - 1, ye = min(y0 + kBpTileY, g.ny) - 1, ze = min(z0 + kBpVoxZ, g.nz) - 1;

    // Footprint: lanes 0..7 project one tile corner each.
    int lx, ly, hx, hy;
    {
        const float cx = (float)((lane & 1) ? xe : x0);
        const float cy = (float)((lane & 2) ? ye : y0);
        const float cz = (float)((lane & 4) ? ze : z0);
        const float ex = fmaf(A[0], cx, fmaf(A[1], cy, fmaf(A[2], cz, A[3])));
        const float ey = fmaf(A[4], cx, fmaf(A[5], cy, fmaf(A[6], cz, A[7])));
        const float ez = fmaf(A[8], cx, fmaf(A[9], cy, fmaf(A[10], cz, A[11])));
        const float r = __builtin_amdgcn_rcpf(ez);
        const float pxf = fminf(fmaxf(fmaf(-ex, r, v.bp[12]), -8.0f), (float)g.W + 8.0f);
        const float pyf = fminf(fmaxf(fmaf(-ey, r, v.bp[13]), -8.0f), (float)g.H + 8.0f);
        lx = (int)floorf(pxf);
        ly = (int)floorf(pyf);
        hx = lx;
        hy = ly;
#pragma unroll
        for (int m = 1; m <= 4; m <<= 1) {
            lx = min(lx, __shfl_xor(lx, m, 64));
            ly = min(ly, __shfl_xor(ly, m, 64));
            hx = max(hx, __shfl_xor(hx, m, 64));
            hy = max(hy, __shfl_xor(hy, m, 64));
        }
        lx = __builtin_amdgcn_readfirstlane(lx);
        ly = __builtin_amdgcn_readfirstlane(ly);
        hx = __builtin_amdgcn_readfirstlane(hx);
        hy = __builtin_amdgcn_readfirstlane(hy);
    }
    // A voxel's window spans px-1..px+2; one more pixel of margin each side.
    const int bx0 = max(lx - 2, 1), by0 = max(ly - 2, 1);
    const int bx1 = min(hx + 3, g.W - 1), by1 = min(hy + 3, g.H - 1);
    const int cols = bx1 - bx0 + 1, rows = by1 - by0 + 1;
    const bool staged = cols > 0 && rows > 0 && cols <= kBpMaxCols && rows <= kBpMaxRows && cols * rows <= kBpLdsFloats;
    const bool full_tile = x0 + kBpTileX <= g.nx && y0 + kBpTileY <= g.ny && z0 + kBpVoxZ <= g.nz;
    const bool interior = staged && full_tile && lx >= 4 && ly >= 4 && hx <= g.W - 5 && hy <= g.H - 5;

    // This thread's voxels are loaded before the footprint, so the two
    // latencies overlap.
    const int x = x0 + (lane & 31), y = y0 + 2 * wv + (lane >> 5);
    const int nz_here = (x < g.nx && y < g.ny) ? min(kBpVoxZ, g.nz - z0) : 0;
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc((void*)it.vol, (short)0, (int)g.vol_bytes, kBufWord3);
    const int vo = 4 * (x + g.nx * y);
    const int plane4 = 4 * g.nx * g.ny;
    float xs[kBpVoxZ];
#pragma unroll
    for (int i = 0; i < kBpVoxZ; ++i)
        xs[i] = i < nz_here ? __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rv, vo, (z0 + i) * plane4, 0)) : 0.0f;

    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc((void*)it.corr, (short)0, (int)g.det_bytes, kBufWord3);
    if (staged) {
        // Rows wv, wv+4, ... of the box; up to 16 rows x 3 column chunks per
        // wave, all loads in flight before the first LDS write.  Lanes past
        // the box's last column re-read that column (every address stays
        // inside the image) and store nothing.
        int co[3];
#pragma unroll
        for (int h = 0; h < 3; ++h) co[h] = 4 * min(lane + 64 * h, cols - 1);
        float vals[16][3];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int rr = wv + 4 * k;
            if (rr < rows) {
                const int so = 4 * ((by0 + rr) * g.W + bx0);
#pragma unroll
                for (int h = 0; h < 3; ++h)
                    vals[k][h] = (64 * h < cols) ? __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rc, co[h], so, 0)) : 0.0f;
            }
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int rr = wv + 4 * k;
            if (rr < rows) {
#pragma unroll
                for (int h = 0; h < 3; ++h)
                    if (lane + 64 * h < cols) win[rr * cols + lane + 64 * h] = vals[k][h];
            }
        }
    }
    __syncthreads();
    if (nz_here == 0) return;

    const float fx = (float)x, fy = (float)y;
    // voxel i's rows are e + i * A[column z]: one rounding each, no drift
    // along the column
    const float ex = fmaf(A[0], fx, fmaf(A[1], fy, fmaf(A[2], (float)z0, A[3])));
    const float ey = fmaf(A[4], fx, fmaf(A[5], fy, fmaf(A[6], (float)z0, A[7])));
    const float ez = fmaf(A[8], fx, fmaf(A[9], fy, fmaf(A[10], (float)z0, A[11])));
    bool any = false;
    if (interior) {
        const int base = -(by0 + 1) * cols - (bx0 + 1);
#pragma unroll
        for (int i = 0; i < kBpVoxZ; ++i) {
            const float r = __builtin_amdgcn_rcpf(fmaf(A[10], (float)i, ez));
            const float pxf = fmaf(-fmaf(A[2], (float)i, ex), r, v.bp[12]);
            const float pyf = fmaf(-fmaf(A[6], (float)i, ey), r, v.bp[13]);
            const int px = (int)pxf, py = (int)pyf;
            const float* w0 = win + (__mul24(py, cols) + px + base);
            float4 rw[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) rw[b] = make_float4(w0[b * cols], w0[b * cols + 1], w0[b * cols + 2], w0[b * cols + 3]);
            const float gv = bp_bicubic(__builtin_amdgcn_fractf(pxf), __builtin_amdgcn_fractf(pyf), rw);
            xs[i] = __builtin_amdgcn_fmed3f(fmaf(pc.alpha, gv, xs[i]), 0.0f, pc.gmax);
        }
        any = true;
    } else {
        const float hxw = (float)(g.W - 2), hyw = (float)(g.H - 2);
#pragma unroll
        for (int i = 0; i < kBpVoxZ; ++i) {
            if (i < nz_here) {
                const float fz = (float)(z0 + i);
                const float r = __builtin_amdgcn_rcpf(fmaf(A[10], (float)i, ez));
                float pxf = fmaf(-fmaf(A[2], (float)i, ex), r, v.bp[12]);
                float pyf = fmaf(-fmaf(A[6], (float)i, ey), r, v.bp[13]);
                // Within 0.01 px of the window's edges the update/skip decision
                // is taken with the reference's own arithmetic (rare).
                if (fabsf(pxf - 2.0f) < 0.01f || fabsf(pxf - hxw) < 0.01f || fabsf(pyf - 2.0f) < 0.01f ||
                    fabsf(pyf - hyw) < 0.01f) {
                    const float2 e = exact_detector_position(v, fx, fy, fz);
                    pxf = e.x;
                    pyf = e.y;
                }
                const int px = (int)pxf, py = (int)pyf;
                if (!(px < 2 || px >= g.W - 2 || py < 2 || py >= g.H - 2)) {
                    any = true;
                    const int wx0 = px - 1 - bx0, wy0 = py - 1 - by0;
                    float4 rw[4];
                    if (staged && wx0 >= 0 && wx0 + 3 < cols && wy0 >= 0 && wy0 + 3 < rows) {
                        const float* w0 = win + wy0 * cols + wx0;
#pragma unroll
                        for (int b = 0; b < 4; ++b)
                            rw[b] = make_float4(w0[b * cols], w0[b * cols + 1], w0[b * cols + 2], w0[b * cols + 3]);
                    } else {
                        const int off = 4 * ((py - 1) * g.W + (px - 1));
#pragma unroll
                        for (int b = 0; b < 4; ++b)
                            rw[b] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rc, off, b * 4 * g.W, 0));
                    }
                    const float gv = bp_bicubic(pxf - (float)px, pyf - (float)py, rw);
                    xs[i] = __builtin_amdgcn_fmed3f(fmaf(pc.alpha, gv, xs[i]), 0.0f, pc.gmax);
                }
            }
        }
    }
    if (it.last) {
#pragma unroll
        for (int i = 0; i < kBpVoxZ; ++i)
            if (i < nz_here) {
                const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * (z0 + i));
                xs[i] = finalize_voxel(it, g, pc, x, y, z0 + i, q, xs[i]);
            }
    } else if (!any) {
        return;
    }
#pragma unroll
    for (int i = 0; i < kBpVoxZ; ++i)
        if (i < nz_here) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, xs[i]), rv, vo, (z0 + i) * plane4, 0);
}

// Ray-BP set-up of one view, one block per work item, after the SART pass
// has written the correction: the scale of the fixed-point contribution sums
// and the reference's Initial() of the ray lengths.
//   Scale: every addend w*corr has |w| <= 1 and one ray adds at most
//   max_samples of them to a voxel, so a view's sums stay below
//   D * max_samples * max|corr|.  acc_shift = 62 - e with that bound < 2^e
//   keeps every sum below 2^62 in the int64 accumulators (the weights use
//   the same bound with max|corr| = 1: VolGeom::acc_wshift).
//   Ray lengths: TypicalRayProducer::Initial() (TypicalRayProducer.cpp:78-96,
//   called at ST_Tomography.cpp:763) zeroes them after every ray-BP view, so
//   the next view's missed rays read 0 rather than a stale length.
__global__ __launch_bounds__(1024) void k_raybp_prep(const WorkItem* __restrict__ items, VolGeom g) {
    __shared__ float red[16];
    const WorkItem& it = items[blockIdx.x];
    const int D = g.W * g.H;
    const gfloat* corr = gptr(it.corr);
    gfloat* rl = it.raylen ? gptr(it.raylen) : nullptr;
    float m = 0.0f;
    for (int k = threadIdx.x; k < D; k += 1024) {
        m = fmaxf(m, fabsf(corr[k]));
        if (rl) rl[k] = 0.0f;
    }
    m = wave_max_f32(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = red[0];
        for (int w = 1; w < 16; ++w) a = fmaxf(a, red[w]);
        int e = 0;
        const double bound = (double)a * (double)D * (double)g.max_samples;
        int shift = 0;
        if (bound > 0.0 && bound < 1e300) {
            (void)frexp(bound, &e);  // bound < 2^e
            shift = 62 - e;
        } else if (bound == 0.0) {
            shift = 0;  // every addend is 0
        }
        *it.acc_shift = shift;
    }
}

// Ray-BP normalisation (ST_Tomography.cpp:746-762): x += contrib/wsum in
// double (no alpha) with the macro clamp, so a voxel no ray reached (0/0 ->
// NaN) becomes 0; then the accumulators are cleared (Initial(), :763).
__global__ __launch_bounds__(256) void k_bp_ray_apply(const WorkItem* __restrict__ items, VolGeom g, PdConst pc) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z % g.nz;
    const WorkItem& it = items[blockIdx.z / g.nz];
    if (x >= g.nx || y >= g.ny) return;
    const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    gu64* ac = (gu64*)it.acc_c;
    gu64* aw = (gu64*)it.acc_w;
    gfloat* vol = gptr(it.vol);
    const double contrib = ldexp((double)(long long)ac[q], -*it.acc_shift);
    const double wsum = ldexp((double)(long long)aw[q], -g.acc_wshift);
    const float m = (float)(contrib / wsum);
    float xv = vol[q] + m;
    xv = tmax(tmin(pc.gmax, xv), 0.0f);  // ternaries: NaN from 0/0 becomes 0 as in the reference
    ac[q] = 0ull;
    aw[q] = 0ull;
    if (it.last) xv = finalize_voxel(it, g, pc, x, y, z, q, xv);
    vol[q] = xv;
}

// ---------------------------------------------------------------------------
// Warp (warping.h:10-145) on the grid aabb = [0, dims].

// Grid point of a warp sample: world_to_cell_grid for aabb = [0, dims]
// (warping.h:31-36), clamped to [1, dim-2] (warping.h:60-62).  Holds the
// base cell, Catmull-Rom weights per axis (CImg _cubic_atXYZ neighbours
// p = x-1, c = x, n = x+1, a = min(x+2, dim-1); where CImg takes n = x because
// the fraction is 0 the weight of n is 0, so x+1 is read instead) and the byte
// offsets of the 16 rows of the 4x4x4 neighbourhood.
struct SamplePt {
    float fx, fy, fz;
    float wx[4], wy[4], wz[4];
    unsigned xoff;      // byte offset of column x-1 (cubic) or x (trilinear)
    unsigned yoff[4];   // byte offsets of rows y-1, y, y+1, min(y+2, ny-1)
    unsigned zoff[4];   // byte offsets of planes z-1 .. min(z+2, nz-1)
    bool xedge;         // x+2 == nx: the fourth column is clamped to nx-1
};

// world_to_cell_grid (warping.h:31-36) for aabb = [0, e]: the ST path has
// e = dims, the flow pyramid keeps the finest level's extent on every level.
template <bool CUBIC>
__device__ __forceinline__ void sample_point_e(const VolGeom& g, const float e[3], float px, float py, float pz,
                                               SamplePt& s) {
    const float d[3] = {(float)g.nx, (float)g.ny, (float)g.nz};
    const float p[3] = {px, py, pz};
    float gc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) gc[c] = __builtin_amdgcn_fmed3f((d[c] * p[c]) / e[c] - 0.5f, 1.0f, d[c] - 2.0f);
    const int ix = (int)gc[0], iy = (int)gc[1], iz = (int)gc[2];
    s.fx = gc[0] - (float)ix;
    s.fy = gc[1] - (float)iy;
    s.fz = gc[2] - (float)iz;
    const unsigned ny4 = 4u * (unsigned)g.nx, nz4 = 4u * (unsigned)(g.nx * g.ny);
    if (CUBIC) {
        cr_weights(s.fx, s.wx);
        cr_weights(s.fy, s.wy);
        cr_weights(s.fz, s.wz);
        s.xoff = 4u * (unsigned)(ix - 1);
        s.xedge = ix + 2 >= g.nx;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            s.yoff[k] = __umul24((unsigned)min(iy - 1 + k, g.ny - 1), ny4);
            s.zoff[k] = __umul24((unsigned)min(iz - 1 + k, g.nz - 1), nz4);
        }
    } else {
        s.xoff = 4u * (unsigned)ix;
        s.xedge = false;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            s.yoff[k] = __umul24((unsigned)(iy + k), ny4);
            s.zoff[k] = __umul24((unsigned)(iz + k), nz4);
        }
    }
}

template <bool CUBIC>
__device__ __forceinline__ void sample_point(const VolGeom& g, float px, float py, float pz, SamplePt& s) {
    const float d[3] = {(float)g.nx, (float)g.ny, (float)g.nz};
    sample_point_e<CUBIC>(g, d, px, py, pz, s);
}

// sample_image (warping.h:49-88) on channel plane `plane_off` (bytes) of an
// image buffer: clamped Catmull-Rom bounded by the min/max of the 8 corners
// of the cell (order x, y, z as CImg), or trilinear at the same point.
template <bool CUBIC>
__device__ __forceinline__ float sample_img(__amdgpu_buffer_rsrc_t rs, unsigned plane_off, const SamplePt& s) {
    if (!CUBIC) {
        float2 r[2][2];
#pragma unroll
        for (int kz = 0; kz < 2; ++kz)
#pragma unroll
            for (int ky = 0; ky < 2; ++ky)
                r[kz][ky] = __builtin_bit_cast(
                    float2, __builtin_amdgcn_raw_buffer_load_b64(rs, s.xoff + s.yoff[ky] + s.zoff[kz], plane_off, 0));
        // explicit fmaf: the same rounding wherever this is inlined (the
        // stored trace end points and the in-place traces must agree)
        const float gx = 1.0f - s.fx, gy = 1.0f - s.fy;
        const float c00 = fmaf(r[0][0].y, s.fx, r[0][0].x * gx), c10 = fmaf(r[0][1].y, s.fx, r[0][1].x * gx);
        const float c01 = fmaf(r[1][0].y, s.fx, r[1][0].x * gx), c11 = fmaf(r[1][1].y, s.fx, r[1][1].x * gx);
        const float e0 = fmaf(c10, s.fy, c00 * gy), e1 = fmaf(c11, s.fy, c01 * gy);
        return fmaf(e1, s.fz, e0 * (1.0f - s.fz));
    }
    float mn = INFINITY, mx = -INFINITY;
    float acc_z = 0.0f;
#pragma unroll
    for (int kz = 0; kz < 4; ++kz) {
        float acc_y = 0.0f;
#pragma unroll
        for (int ky = 0; ky < 4; ++ky) {
            float4 row = __builtin_bit_cast(
                float4, __builtin_amdgcn_raw_buffer_load_b128(rs, s.xoff + s.yoff[ky] + s.zoff[kz], plane_off, 0));
            if (s.xedge) row.w = row.z;
            if ((ky == 1 || ky == 2) && (kz == 1 || kz == 2)) {
                mn = fminf(mn, fminf(row.y, row.z));
                mx = fmaxf(mx, fmaxf(row.y, row.z));
            }
            const float rv = fmaf(s.wx[3], row.w, fmaf(s.wx[2], row.z, fmaf(s.wx[1], row.y, s.wx[0] * row.x)));
            acc_y = fmaf(s.wy[ky], rv, acc_y);
        }
        acc_z = fmaf(s.wz[kz], acc_y, acc_z);
    }
    return fmaxf(mn, fminf(mx, acc_z));
}

// trace (warping.h:100-114) from the cell centre of (i,j,k) with trace step
// tdt (= -dt of warp): the end point q at which warp samples the image.  The
// first velocity sample sits exactly on grid node clamp(i,1,n-2), where the
// clamped cubic returns the node value, so it is read directly.  The three
// velocity planes share one buffer.
template <bool CUBIC>
__device__ __forceinline__ void trace_voxel(const float* uvw, const VolGeom& g, int i, int j, int k, float tdt,
                                            float q[3]) {
    const unsigned V = (unsigned)(g.nx * g.ny * g.nz);
    const __amdgpu_buffer_rsrc_t rs_vel =
        __builtin_amdgcn_make_buffer_rsrc((void*)uvw, (short)0, (int)(12u * V), kBufWord3);
    const int ci = min(max(i, 1), g.nx - 2), cj = min(max(j, 1), g.ny - 2), ck = min(max(k, 1), g.nz - 2);
    const size_t nq = (size_t)ci + (size_t)g.nx * ((size_t)cj + (size_t)g.ny * ck);
    const gfloat* u = gptr(uvw);
    const float p[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
    const float v1[3] = {u[nq], u[V + nq], u[2 * (size_t)V + nq]};
    float p2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) p2[c] = p[c] + tdt * v1[c];
    SamplePt s;
    sample_point<CUBIC>(g, p2[0], p2[1], p2[2], s);
    float v2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v2[c] = sample_img<CUBIC>(rs_vel, 4u * V * (unsigned)c, s);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float p0 = p2[c] - tdt * v2[c];
        q[c] = p2[c] + 0.5f * (p[c] - p0);
    }
}

// The image sample of warp at a traced end point q.
template <bool CUBIC>
__device__ __forceinline__ float sample_at(const float* img, const VolGeom& g, const float q[3]) {
    const unsigned V = (unsigned)(g.nx * g.ny * g.nz);
    const __amdgpu_buffer_rsrc_t rs_img = __builtin_amdgcn_make_buffer_rsrc((void*)img, (short)0, (int)(4u * V), kBufWord3);
    SamplePt s;
    sample_point<CUBIC>(g, q[0], q[1], q[2], s);
    return sample_img<CUBIC>(rs_img, 0u, s);
}

// warp of img by uvw at voxel (i,j,k): trace, then sample.
template <bool CUBIC>
__device__ __forceinline__ float warp_voxel(const float* img, const float* uvw, const VolGeom& g, int i, int j, int k,
                                            float tdt) {
    float q[3];
    trace_voxel<CUBIC>(uvw, g, i, j, k, tdt, q);
    return sample_at<CUBIC>(img, g, q);
}

// The PD steps' warp of img at voxel (i,j,k): from the stored end points of
// the trace when the context keeps them (three planes of V floats), else
// traced here.  Both give the same float end point, so the same sample.  The
// form is a template argument (pd_warp_kind), so each PD kernel instance
// carries one sampling path and the registers of that path only.
template <int WK>
__device__ __forceinline__ float pd_warp(const float* img, const float* uvw, const float* tr, const VolGeom& g, int i,
                                         int j, int k, size_t qi, float tdt) {
    constexpr bool kCubic = (WK & 1) == 0;
    if (WK < 2) {
        const size_t V = (size_t)g.nx * g.ny * g.nz;
        const gfloat* t = gptr(tr);
        const float q[3] = {t[qi], t[V + qi], t[2 * V + qi]};
        return sample_at<kCubic>(img, g, q);
    }
    return warp_voxel<kCubic>(img, uvw, g, i, j, k, tdt);
}

// 0: stored end points, clamped cubic; 1: stored, trilinear; 2: traced here,
// clamped cubic; 3: traced here, trilinear.
inline int pd_warp_kind(const PdConst& pc) { return (pc.stored_traces ? 0 : 2) + (pc.interp == 0 ? 0 : 1); }

// End points of the two PD traces of flow u_f at every voxel: step +1 (the
// dual's backward warp, warp(u_f, ., dt = -1)) into tb, step -1 (the primal's
// forward warp, dt = +1) into tf; three planes each.
template <bool CUBIC>
__global__ __launch_bounds__(256) void k_trace(const float* __restrict__ uvw, VolGeom g, float* __restrict__ tb,
                                              float* __restrict__ tf) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const size_t V = (size_t)g.nx * g.ny * g.nz;
    const size_t qi = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    float q[3];
    trace_voxel<CUBIC>(uvw, g, x, y, z, 1.0f, q);
    gfloat* b = gptr(tb);
    b[qi] = q[0];
    b[V + qi] = q[1];
    b[2 * V + qi] = q[2];
    trace_voxel<CUBIC>(uvw, g, x, y, z, -1.0f, q);
    gfloat* f = gptr(tf);
    f[qi] = q[0];
    f[V + qi] = q[1];
    f[2 * V + qi] = q[2];
}

__global__ __launch_bounds__(256) void k_warp(const float* __restrict__ uvw, const float* __restrict__ src,
                                              float* __restrict__ out, VolGeom g, float tdt, int cubic) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    out[q] = cubic ? warp_voxel<true>(src, uvw, g, x, y, z, tdt) : warp_voxel<false>(src, uvw, g, x, y, z, tdt);
}

// Dual step of frame f (ST_Tomography.cpp:502-534); grid.z = z + nz*(f - f0).
template <int WK>
__global__ __launch_bounds__(256) void k_pd_dual(const FrameBufs* __restrict__ fb, int f0, VolGeom g, PdConst pc) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z % g.nz;
    const int f = f0 + blockIdx.z / g.nz;
    if (x >= g.nx || y >= g.ny) return;
    const size_t V = (size_t)g.nx * g.ny * g.nz;
    const size_t sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t q = (size_t)x + sy * y + sz * z;
    const FrameBufs& B = fb[f];
    const FrameBufs& B1 = fb[f + 1];
    const gfloat* xb0 = gptr(B.xbar);
    const gfloat* xb1 = gptr(B1.xbar);
    gfloat* Ytv = gptr(B.ytv);
    gfloat* Ytp = gptr(B.ytp);
    gfloat* Ybc = gptr(B.ybc);
    // m_backward_warp = warp(u_f, X_bar[f+1], dt = -1): trace step +1 (:503)
    const float bw = pd_warp<WK>(B1.xbar, B.flow, B.trace_bw, g, x, y, z, q, 1.0f);
    const float xq = xb0[q];
    float ytp = Ytp[q] + pc.sigma * (xb1[q] - xq);  // (:504)
    float ybc = Ybc[q] + pc.sigma * (bw - xq);      // (:505)
    if (x >= 1 && x < g.nx - 1 && y >= 1 && y < g.ny - 1 && z >= 1 && z < g.nz - 1 && !cut_out(g, pc, x, y, z)) {
        const float h = pc.huber_factor;
        float y0 = Ytv[q] + pc.sigma * (xb0[q + 1] - xq) * h;
        float y1 = Ytv[V + q] + pc.sigma * (xb0[q + sy] - xq) * h;
        float y2 = Ytv[2 * V + q] + pc.sigma * (xb0[q + sz] - xq) * h;
        if (pc.prior == 0) {  // ATV per-component clamp (:523-525)
            y0 = tmax(-pc.lambda, tmin(pc.lambda, y0));
            y1 = tmax(-pc.lambda, tmin(pc.lambda, y1));
            y2 = tmax(-pc.lambda, tmin(pc.lambda, y2));
        } else if (pc.prior == 1) {  // isotropic TV: projection onto the lambda L2 ball
            const float nrm = sqrtf(y0 * y0 + y1 * y1 + y2 * y2);
            const float sc = nrm > pc.lambda ? pc.lambda / nrm : 1.0f;
            y0 *= sc; y1 *= sc; y2 *= sc;
        } else {  // STV: 4-vector with the forward temporal difference, lambda L2 ball
            float y3 = Ytv[3 * V + q] + pc.sigma * (xb1[q] - xq) * h;
            const float nrm = sqrtf(y0 * y0 + y1 * y1 + y2 * y2 + y3 * y3);
            const float sc = nrm > pc.lambda ? pc.lambda / nrm : 1.0f;
            y0 *= sc; y1 *= sc; y2 *= sc; y3 *= sc;
            Ytv[3 * V + q] = y3;
        }
        Ytv[q] = y0;
        Ytv[V + q] = y1;
        Ytv[2 * V + q] = y2;
        ytp = ytp * pc.ktp / (pc.ktp + 2.0f * pc.sigma);  // (:528)
        ybc = tmax(-pc.ebc, tmin(pc.ebc, ybc));           // (:530)
    }
    Ytp[q] = ytp;
    Ybc[q] = ybc;
}

// Primal step of frame f (ST_Tomography.cpp:536-578).
template <int WK>
__global__ __launch_bounds__(256) void k_pd_primal(const FrameBufs* __restrict__ fb, int f0, VolGeom g, PdConst pc) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z % g.nz;
    const int f = f0 + blockIdx.z / g.nz;
    if (x >= g.nx || y >= g.ny) return;
    const size_t V = (size_t)g.nx * g.ny * g.nz;
    const size_t sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t q = (size_t)x + sy * y + sz * z;
    const FrameBufs& B = fb[f];
    gfloat* X = gptr(B.x);
    const gfloat* Ytp = gptr(B.ytp);
    const gfloat* Ybc = gptr(B.ybc);
    float xv = X[q];
    gptr(B.xprev)[q] = xv;  // (:536)
    if (f == 0) {           // (:540-545)
        xv = xv - pc.tau * (-Ytp[q]);
        xv = xv - pc.tau * (-Ybc[q]);
    } else {                // (:546-555): forward warp of Y_bc[f-1] with u_{f-1}, trace step -1
        const FrameBufs& Bm = fb[f - 1];
        const float fw = pd_warp<WK>(Bm.ybc, Bm.flow, Bm.trace_fw, g, x, y, z, q, -1.0f);
        xv = xv - pc.tau * (gptr(Bm.ytp)[q] - Ytp[q]);
        xv = xv - pc.tau * (fw - Ybc[q]);
    }
    if (x >= 1 && x < g.nx - 1 && y >= 1 && y < g.ny - 1 && z >= 1 && z < g.nz - 1 && !cut_out(g, pc, x, y, z)) {
        const gfloat* t = gptr(B.ytv);
        const float y1 = -(t[q] - t[q - 1] + t[V + q] - t[V + q - sy] + t[2 * V + q] - t[2 * V + q - sz]);  // (:570-571)
        xv = xv - pc.tau * y1;
        if (pc.prior == 2) {  // STV: adjoint of the forward temporal difference, y_t[f-1] - y_t[f]
            const float ytm = f > 0 ? gptr(fb[f - 1].ytv)[3 * V + q] : 0.0f;
            xv = xv - pc.tau * (ytm - t[3 * V + q]);
        }
        xv = tmax(tmin(pc.gmax, xv), 0.0f);  // (:575)
    }
    X[q] = xv;
}

// Extrapolation X_bar = 2x - X_prev on the ROI (ST_Tomography.cpp:586-604),
// the stand-alone form of the sweep epilogue.
__global__ __launch_bounds__(256) void k_extrapolate(const FrameBufs* __restrict__ fb, int f0, VolGeom g, PdConst pc) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z % g.nz;
    const int f = f0 + blockIdx.z / g.nz;
    if (x >= g.nx || y >= g.ny || cut_out(g, pc, x, y, z)) return;
    const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    const FrameBufs& B = fb[f];
    gptr(B.xbar)[q] = 2.0f * gptr(B.x)[q] - gptr(B.xprev)[q];
}

// ---------------------------------------------------------------------------
// FP64 reductions.  Fixed partition (grid-stride over kReductionBlocks
// blocks, one partial per block, one final block), so every run sums in the
// same order.

__device__ __forceinline__ double block_sum_f64(double v, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    return t;  // valid in thread 0
}

// partial[blockIdx.x * stride] = sum over k of (b[k] - ax[k])^2 (ax = 0 for a
// missed ray).  With one view per slot and stride = number of views, block
// t's partials of all views are contiguous, so k_sum_partials adds each
// block's partials over the views in view order.
__global__ __launch_bounds__(256) void k_residual_partial(const float* __restrict__ b, const float* __restrict__ ax,
                                                          int n, double* __restrict__ partial, int stride) {
    __shared__ double red[4];
    double acc = 0.0;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        const double r = (double)b[k] - (double)ax[k];
        acc = fma(r, r, acc);
    }
    const double t = block_sum_f64(acc, red);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * stride] = t;
}

// partial[blockIdx.x] = sum of a[k]*c[k] (a == c: squared norm).
__global__ __launch_bounds__(256) void k_dot_partial(const float* __restrict__ a, const float* __restrict__ c, size_t n,
                                                     double* __restrict__ partial) {
    __shared__ double red[4];
    double acc = 0.0;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
        acc = fma((double)a[k], (double)c[k], acc);
    const double t = block_sum_f64(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// out[0] = sum of partial[0..n) in index order (one block).
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
    __shared__ double red[4];
    // each thread sums a contiguous run, so the order is fixed
    const int per = (n + 255) / 256;
    double acc = 0.0;
    for (int k = threadIdx.x * per; k < min(n, (int)(threadIdx.x + 1) * per); ++k) acc += partial[k];
    const double t = block_sum_f64(acc, red);
    if (threadIdx.x == 0) out[0] = t;
}

// ---- allocation-free residual norm -----------------------------------------
// File-scope scratch of stt_residual_norm_noalloc, one copy per device.  In
// host code the name `scratch` is only a host-side stand-in: its device
// address comes from hipGetSymbolAddress (noalloc_scratch), and only that
// address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// One work item of the projector, assembled on the device from the
// context's device pointers (no host-to-device copy of a host temporary).
__global__ void k_item_make(WorkItem* __restrict__ dst, const ViewRec* view, float* vol, float* volT,
                            const float* b, float* ax, float* raylen) {
    WorkItem it{};
    it.view = view;
    it.vol = vol;
    it.volT = volT;
    it.b = b;
    it.ax = ax;
    it.raylen = raylen;
    *dst = it;
}

__global__ __launch_bounds__(256) void k_zero_f64(double* __restrict__ p, int n) {
    for (int k = threadIdx.x; k < n; k += 256) p[k] = 0.0;
}

// acc[t] += part[t]: block t's running sum over the views, added in view
// order exactly as k_sum_partials adds a block-major partial array.
__global__ __launch_bounds__(256) void k_add_partials(double* __restrict__ acc, const double* __restrict__ part,
                                                      int n) {
    for (int k = threadIdx.x; k < n; k += 256) acc[k] += part[k];
}

// x *= s (s = 1/||x|| read from device memory as a double).
__global__ __launch_bounds__(256) void k_scale_by_inv_sqrt(float* __restrict__ x, size_t n, const double* __restrict__ nrm2) {
    const double v = nrm2[0];
    const float s = v > 0.0 ? (float)(1.0 / sqrt(v)) : 0.0f;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) x[k] *= s;
}

// Deterministic start vector of the power iteration: x[k] in [-1, 1) from a
// 32-bit integer hash of k.
__global__ __launch_bounds__(256) void k_fill_hash(float* __restrict__ x, size_t n, unsigned seed) {
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        unsigned h = (unsigned)k * 0x9E3779B1u ^ seed;
        h ^= h >> 16;
        h *= 0x85EBCA6Bu;
        h ^= h >> 13;
        h *= 0xC2B2AE35u;
        h ^= h >> 16;
        x[k] = (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
    }
}

// PD coupling operator of the power iteration: K v = (masked forward
// differences along x, y, z on the dual's region (interior of the grid
// inside the ROI cut, ST_Tomography.cpp:509-516), forward time difference
// v[f+1] - v[f] over the whole volume with v[F] = 0).  v: F volumes, y: F
// blocks of 4 planes.  grid.z = z + nz*f.
__global__ __launch_bounds__(256) void k_pi_apply_k(const float* __restrict__ v, float* __restrict__ y, int F, VolGeom g,
                                                    PdConst pc) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int yy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z % g.nz;
    const int f = blockIdx.z / g.nz;
    if (x >= g.nx || yy >= g.ny) return;
    const size_t V = (size_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t q = (size_t)x + sy * yy + sz * z;
    const float* vf = v + V * f;
    float* yf = y + 4 * V * f;
    const float c = vf[q];
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    if (x >= 1 && x < g.nx - 1 && yy >= 1 && yy < g.ny - 1 && z >= 1 && z < g.nz - 1 && !cut_out(g, pc, x, yy, z)) {
        gx = vf[q + 1] - c;
        gy = vf[q + sy] - c;
        gz = vf[q + sz] - c;
    }
    yf[q] = gx;
    yf[V + q] = gy;
    yf[2 * V + q] = gz;
    yf[3 * V + q] = (f + 1 < F ? v[V * (f + 1) + q] : 0.0f) - c;
}

// w = K^T y: minus the backward divergence of the spatial components plus
// the adjoint of the time difference, y_t[f-1] - y_t[f] (no y_t[-1]).
__global__ __launch_bounds__(256) void k_pi_apply_kt(const float* __restrict__ y, float* __restrict__ w, int F, VolGeom g) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int yy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z % g.nz;
    const int f = blockIdx.z / g.nz;
    if (x >= g.nx || yy >= g.ny) return;
    const size_t V = (size_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t q = (size_t)x + sy * yy + sz * z;
    const float* yf = y + 4 * V * f;
    float r = -(yf[q] + yf[V + q] + yf[2 * V + q]);
    if (x > 0) r += yf[q - 1];
    if (yy > 0) r += yf[V + q - sy];
    if (z > 0) r += yf[2 * V + q - sz];
    r += (f > 0 ? y[4 * V * (f - 1) + 3 * V + q] : 0.0f) - yf[3 * V + q];
    w[V * f + q] = r;
}

// Adjoint of the trilinear warp (warping.h:100-145 with trilinear sampling):
// for each output voxel, scatter g * d(out)/d(src) into gsrc and
// g * d(out)/d(uvw) into guvw.  The trace is p2 = p + tdt*v1 (v1 = node value
// of u at clamp(i,1,n-2)), v2 = S_u(p2), q = p2 + (p - (p2 - tdt*v2))/2,
// out = S_src(q); the grid clamps to [1, n-2] have zero derivative outside.
struct TriPt {
    int ix, iy, iz;
    float fx, fy, fz;
    float dx, dy, dz;  // 1 inside the clamp range, 0 where the coordinate is clamped
};

__device__ __forceinline__ TriPt tri_point(const VolGeom& g, const float p[3]) {
    const float d[3] = {(float)g.nx, (float)g.ny, (float)g.nz};
    float gc[3], in[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float raw = (d[c] * p[c]) / d[c] - 0.5f;
        gc[c] = __builtin_amdgcn_fmed3f(raw, 1.0f, d[c] - 2.0f);
        in[c] = (raw > 1.0f && raw < d[c] - 2.0f) ? 1.0f : 0.0f;
    }
    TriPt t;
    t.ix = (int)gc[0];
    t.iy = (int)gc[1];
    t.iz = (int)gc[2];
    t.fx = gc[0] - (float)t.ix;
    t.fy = gc[1] - (float)t.iy;
    t.fz = gc[2] - (float)t.iz;
    t.dx = in[0];
    t.dy = in[1];
    t.dz = in[2];
    return t;
}

// Trilinear value and gradient (w.r.t. the sample position) of one plane.
__device__ __forceinline__ float tri_eval(const gfloat* img, const VolGeom& g, const TriPt& t, float grad[3]) {
    const size_t sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t b = (size_t)t.ix + sy * t.iy + sz * t.iz;
    const float c000 = img[b], c100 = img[b + 1], c010 = img[b + sy], c110 = img[b + sy + 1];
    const float c001 = img[b + sz], c101 = img[b + sz + 1], c011 = img[b + sz + sy], c111 = img[b + sz + sy + 1];
    const float gx = 1.0f - t.fx, gy = 1.0f - t.fy, gz = 1.0f - t.fz;
    const float c00 = c000 * gx + c100 * t.fx, c10 = c010 * gx + c110 * t.fx;
    const float c01 = c001 * gx + c101 * t.fx, c11 = c011 * gx + c111 * t.fx;
    const float e0 = c00 * gy + c10 * t.fy, e1 = c01 * gy + c11 * t.fy;
    grad[0] = t.dx * (gy * gz * (c100 - c000) + t.fy * gz * (c110 - c010) + gy * t.fz * (c101 - c001) +
                      t.fy * t.fz * (c111 - c011));
    grad[1] = t.dy * (gz * (c10 - c00) + t.fz * (c11 - c01));
    grad[2] = t.dz * (e1 - e0);
    return e0 * gz + e1 * t.fz;
}

__device__ __forceinline__ void tri_scatter(gfloat* img, const VolGeom& g, const TriPt& t, float val) {
    const size_t sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t b = (size_t)t.ix + sy * t.iy + sz * t.iz;
    const float gx = 1.0f - t.fx, gy = 1.0f - t.fy, gz = 1.0f - t.fz;
    __builtin_amdgcn_global_atomic_fadd_f32(img + b, val * gx * gy * gz);
    __builtin_amdgcn_global_atomic_fadd_f32(img + b + 1, val * t.fx * gy * gz);
    __builtin_amdgcn_global_atomic_fadd_f32(img + b + sy, val * gx * t.fy * gz);
    __builtin_amdgcn_global_atomic_fadd_f32(img + b + sy + 1, val * t.fx * t.fy * gz);
    __builtin_amdgcn_global_atomic_fadd_f32(img + b + sz, val * gx * gy * t.fz);
    __builtin_amdgcn_global_atomic_fadd_f32(img + b + sz + 1, val * t.fx * gy * t.fz);
    __builtin_amdgcn_global_atomic_fadd_f32(img + b + sz + sy, val * gx * t.fy * t.fz);
    __builtin_amdgcn_global_atomic_fadd_f32(img + b + sz + sy + 1, val * t.fx * t.fy * t.fz);
}

__global__ __launch_bounds__(256) void k_warp_vjp(const float* __restrict__ uvw, const float* __restrict__ src,
                                                  const float* __restrict__ gout, float* __restrict__ gsrc,
                                                  float* __restrict__ guvw, VolGeom g, float tdt) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const size_t V = (size_t)g.nx * g.ny * g.nz;
    const size_t o = (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k);
    const float go = gptr(gout)[o];
    if (go == 0.0f) return;
    const gfloat* u = gptr(uvw);
    const int ci = min(max(i, 1), g.nx - 2), cj = min(max(j, 1), g.ny - 2), ck = min(max(k, 1), g.nz - 2);
    const size_t nq = (size_t)ci + (size_t)g.nx * ((size_t)cj + (size_t)g.ny * ck);
    const float p[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
    const float v1[3] = {u[nq], u[V + nq], u[2 * V + nq]};
    float p2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) p2[c] = p[c] + tdt * v1[c];
    const TriPt t2 = tri_point(g, p2);
    float v2[3], J[3][3];  // J[c][d] = d v2_c / d p2_d
#pragma unroll
    for (int c = 0; c < 3; ++c) v2[c] = tri_eval(u + V * c, g, t2, J[c]);
    float q[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float p0 = p2[c] - tdt * v2[c];
        q[c] = p2[c] + 0.5f * (p[c] - p0);
    }
    const TriPt tq = tri_point(g, q);
    float a[3];
    (void)tri_eval(gptr(src), g, tq, a);
    if (gsrc) tri_scatter(gptr(gsrc), g, tq, go);
    if (!guvw) return;
    // dL/dq = go * grad S_src(q); q = (p + p2)/2 + tdt*v2/2
    float dq[3], dv2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dq[c] = go * a[c];
        dv2[c] = 0.5f * tdt * dq[c];
    }
    gfloat* gu = gptr(guvw);
#pragma unroll
    for (int c = 0; c < 3; ++c) tri_scatter(gu + V * c, g, t2, dv2[c]);
    // dL/dp2 = dq/2 + J^T dv2; dL/dv1 = tdt * dL/dp2 (node value at clamp(i))
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float dp2 = 0.5f * dq[d] + J[0][d] * dv2[0] + J[1][d] * dv2[1] + J[2][d] * dv2[2];
        __builtin_amdgcn_global_atomic_fadd_f32(gu + V * d + nq, tdt * dp2);
    }
}

// ---------------------------------------------------------------------------
// Launchers.

// ---------------------------------------------------------------------------
// 3-D L1/Huber-TV optical flow, one pyramid level (optical_flow.h:83-623).
// Planes of n voxels: V (3: u v w), Y (10: y11 y12 y13 y21 y22 y23 y31 y32
// y33 y4), Kx (10: the previous iteration's K V), Kty (3, kept only where a
// residual needs it), G (4: ux uy uz ut), T (4: tau1 tau2 tau3 sigma2).  fp32
// in the reference's expression order; residual sums in FP64, fixed order.

// Workgroups go to the eight XCDs in turn by linear id, and each XCD has its
// own L2.  The flow PD kernels read +-1 neighbours across y rows and z
// slices, so their tiles are remapped: XCD k takes one contiguous run of a
// tile order and finds most neighbour rows in its own L2.
//   ZCHUNK = false: x fastest, then y, then z (the primal, whose grid covers
//     the ROI box);
//   ZCHUNK = true: chunks of kZc slices, z fastest inside a chunk, then x,
//     then y (the dual: one 256^2 slice of its 30 planes is about twice an
//     XCD's L2, so slice-by-slice order loses the z neighbours before the
//     next slice reads them).
// Both steps are bijections for any grid size (runs of G/8 or G/8 + 1
// tiles; the last chunk may be shorter).  `lin` is the plain linear index
// x + gx (y + gy z) of the tile, which orders the per-block partial sums.
// Measured at 256^3 on MI355X: dual 435 -> 414 us (fetched bytes 1.19 ->
// 1.04 of its algorithmic reads), primal 229 -> 218 us; the warp and the
// gradient kernels ran slower remapped and keep the plain order.
struct OfTile {
    unsigned x, y, z, lin;
};
template <bool ZCHUNK>
__device__ __forceinline__ OfTile of_xcd_tile() {
    constexpr unsigned kXcds = 8, kZc = 8;
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z, G = gx * gy * gz;
    const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned xcd = L % kXcds, q = G / kXcds, r = G % kXcds;
    const unsigned T = xcd * q + min(xcd, r) + L / kXcds;
    if (!ZCHUNK) return OfTile{T % gx, (T / gx) % gy, T / (gx * gy), T};
    const unsigned zc = T / (gx * gy * kZc), h = min(kZc, gz - kZc * zc);
    const unsigned rem = T - zc * gx * gy * kZc, t2 = rem / h;
    const unsigned x = t2 % gx, y = t2 / gx, z = kZc * zc + rem % h;
    return OfTile{x, y, z, x + gx * (y + gy * z)};
}

// warp(dim, aabb, V, I2, -dt) with dt = 1 and aabb = [0, ext] (:245): trace
// from the cell centre in world units, clamped-cubic samples.
__global__ __launch_bounds__(256) void k_of_warp(const float* __restrict__ v, const float* __restrict__ img,
                                                 float* __restrict__ out, VolGeom g, float ex, float ey, float ez) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const unsigned V = (unsigned)(g.nx * g.ny * g.nz);
    const __amdgpu_buffer_rsrc_t rs_vel = __builtin_amdgcn_make_buffer_rsrc((void*)v, (short)0, (int)(12u * V), kBufWord3);
    const __amdgpu_buffer_rsrc_t rs_img = __builtin_amdgcn_make_buffer_rsrc((void*)img, (short)0, (int)(4u * V), kBufWord3);
    const float e[3] = {ex, ey, ez};
    const int ijk[3] = {x, y, z};
    const int dd[3] = {g.nx, g.ny, g.nz};
    float p[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)  // cell_grid_to_world (warping.h:24-29), evaluated in double
        p[c] = (float)(0.0 + ((double)(float)ijk[c] + 0.5) * (double)(e[c] - 0.0f) / (double)(float)dd[c]);
    SamplePt s;
    sample_point_e<true>(g, e, p[0], p[1], p[2], s);
    float p2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) p2[c] = p[c] + 1.0f * sample_img<true>(rs_vel, 4u * V * (unsigned)c, s);
    sample_point_e<true>(g, e, p2[0], p2[1], p2[2], s);
    float q[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float p0 = p2[c] - 1.0f * sample_img<true>(rs_vel, 4u * V * (unsigned)c, s);
        q[c] = p2[c] + 0.5f * (p[c] - p0);
    }
    sample_point_e<true>(g, e, q[0], q[1], q[2], s);
    out[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] = sample_img<true>(rs_img, 0u, s);
}

struct OfRoi {
    int x0, x1, y0, y1, z0, z1;  // margins: the ROI is [x0, nx - x1) x [y0, ny - y1) x [z0, nz - z1)
};
__device__ __forceinline__ bool of_in_roi(const VolGeom& g, const OfRoi& r, int x, int y, int z) {
    return x >= r.x0 && x < g.nx - r.x1 && y >= r.y0 && y < g.ny - r.y1 && z >= r.z0 && z < g.nz - r.z1;
}

// Gradient of the warped image (CImg central differences, clamped
// neighbours, CImg.h:34570-34580) and the linearised data term on the ROI
// (:254-268); step sizes on every voxel (:273-287).
__global__ __launch_bounds__(256) void k_of_prep(const float* __restrict__ W, const float* __restrict__ I1,
                                                 const float* __restrict__ v, float* __restrict__ G, VolGeom g,
                                                 OfRoi roi) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const size_t n = (size_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t q = (size_t)x + sy * y + sz * z;
    if (of_in_roi(g, roi, x, y, z)) {
        const size_t xp = x + 1 < g.nx ? q + 1 : q, xm = x > 0 ? q - 1 : q;
        const size_t yp = y + 1 < g.ny ? q + sy : q, ym = y > 0 ? q - sy : q;
        const size_t zp = z + 1 < g.nz ? q + sz : q, zm = z > 0 ? q - sz : q;
        const float ux = (W[xp] - W[xm]) / 2.0f;
        const float uy = (W[yp] - W[ym]) / 2.0f;
        const float uz = (W[zp] - W[zm]) / 2.0f;
        G[q] = ux;
        G[n + q] = uy;
        G[2 * n + q] = uz;
        G[3 * n + q] = W[q] - I1[q] - ux * v[q] - uy * v[n + q] - uz * v[2 * n + q];
    }
}

// The per-voxel steps of :273-287 are functions of the gradient planes, which
// the primal and the dual read anyway; they are recomputed there with the
// reference's expressions instead of being stored (stepsize = 1).  Outside the
// ROI the gradient keeps its zero fill, so tau = 1/4 and sigma2 = 1/0 = inf,
// as the stored planes held.
__device__ __forceinline__ float of_tau(float u) { return 1.0f / (4.0f / 1.0f + fabsf(u)); }
__device__ __forceinline__ float of_sigma2(float ux, float uy, float uz) {
    return 1.0f / (fabsf(ux) + fabsf(uy) + fabsf(uz));
}

// tools.cpp differences (forward: 0 on the last index; backward: a at the
// first index, -a[n-2] at the last).
// Both loads are unconditional (the neighbour index is clamped to the voxel
// itself at the border) and the border cases are selects on the values, so a
// wave issues every load of a stencil before it waits on any of them.
__device__ __forceinline__ float of_dm(const float* a, size_t q, size_t st, int i, int n) {
    const float c = a[q];
    const float p = a[i > 0 ? q - st : q];
    const float d = c - p;
    return (i > 0 && i < n - 1) ? d : (i == 0 ? c : -p);
}
__device__ __forceinline__ float of_dp(const float* a, size_t q, size_t st, int i, int n) {
    const float c = a[q];
    const float nb = a[i < n - 1 ? q + st : q];
    const float d = nb - c;
    return i < n - 1 ? d : 0.0f;
}

__device__ __forceinline__ void of_block_partial(double acc, double* partial, unsigned tile) {
    __shared__ double red[4];
    const double t = block_sum_f64(acc, red);
    if (threadIdx.x == 0) partial[tile] = t;
}

// Primal step on the ROI (:300-340).  MODE 1 keeps K^T y for the next
// iteration's residual; MODE 2 adds this iteration's primal residual.
// The flow alternates between two buffers: iteration k reads vin (the
// previous iterate) and writes vout on the ROI; outside the ROI both buffers
// hold the level's input flow, which no step changes.
template <int MODE>
__global__ __launch_bounds__(256) void k_of_primal(const float* __restrict__ vin, float* __restrict__ vout,
                                                   const float* __restrict__ Y, float* __restrict__ kty,
                                                   const float* __restrict__ G, double* __restrict__ partial,
                                                   VolGeom g, OfRoi roi) {
#pragma clang fp contract(off)
    // the grid covers the ROI's box in y and z; x keeps whole 64-voxel rows
    // so that a wave's loads stay aligned to cache lines
    const OfTile tl = of_xcd_tile<false>();
    const int x = tl.x * 64 + (threadIdx.x & 63);
    const int y = roi.y0 + tl.y * 4 + (threadIdx.x >> 6);
    const int z = roi.z0 + tl.z;
    double acc = 0.0;
    if (x < g.nx && y < g.ny && of_in_roi(g, roi, x, y, z)) {
        const size_t n = (size_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (size_t)g.nx * g.ny;
        const size_t q = (size_t)x + sy * y + sz * z;
        const float y4 = Y[9 * n + q];
        float r = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float* yc = Y + (size_t)(3 * c) * n;
            const float gc = G[c * n + q];
            const float kt = -1.0f * of_dm(yc, q, 1, x, g.nx) - 1.0f * of_dm(yc + n, q, sy, y, g.ny) -
                             1.0f * of_dm(yc + 2 * n, q, sz, z, g.nz) + gc * y4;
            const float tc = of_tau(gc);
            const float vo = vin[c * n + q];
            const float vn = vo - tc * kt;
            vout[c * n + q] = vn;
            if (MODE == 1) kty[c * n + q] = kt;
            if (MODE == 2) r = r + fabsf((vo - vn) / tc - kty[c * n + q] + kt);
        }
        acc = (double)r;
    }
    if (MODE == 2) of_block_partial(acc, partial, tl.lin);
}

// Dual step on every voxel (:343-480); RES adds the dual residual.
// The reference keeps Kx = K v of the previous iteration in ten planes
// (:379-422) and extrapolates with 2 K v_new - Kx_old.  K is linear, so the
// old values are recomputed here from the previous iterate vold with the same
// fp32 operations that produced them, which gives the same floats while
// moving 3 planes instead of 20.  Two iterations differ:
//   FIRST = 1  first iteration of a level: Kx_old = 0 (:229-242);
//   FIRST = 2  first iteration of a later warp: the nine gradient terms are
//              K vold as usual, but Kx4_old was formed with the previous
//              warp's image gradient, kept in k4old (k_of_k4).
template <bool RES, int FIRST>
__global__ __launch_bounds__(256) void k_of_dual(const float* __restrict__ v, const float* __restrict__ vold,
                                                 float* __restrict__ Y, const float* __restrict__ k4old,
                                                 const float* __restrict__ G, double* __restrict__ partial, VolGeom g,
                                                 float sigma1, float hf, float lambda, int huber) {
#pragma clang fp contract(off)
    const OfTile tl = of_xcd_tile<true>();
    const int x = tl.x * 64 + (threadIdx.x & 63);
    const int y = tl.y * 4 + (threadIdx.x >> 6);
    const int z = tl.z;
    double acc = 0.0;
    if (x < g.nx && y < g.ny) {
        const size_t n = (size_t)g.nx * g.ny * g.nz, sy = g.nx, sz = (size_t)g.nx * g.ny;
        const size_t q = (size_t)x + sy * y + sz * z;
        // every load before the first store, so a wave has them all in flight
        float kn[3][3], ko[3][3], yo[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float* vc = v + (size_t)c * n;
            const float* oc = vold + (size_t)c * n;
            kn[c][0] = 1.0f * of_dp(vc, q, 1, x, g.nx);
            kn[c][1] = 1.0f * of_dp(vc, q, sy, y, g.ny);
            kn[c][2] = 1.0f * of_dp(vc, q, sz, z, g.nz);
#pragma unroll
            for (int e = 0; e < 3; ++e) ko[c][e] = 0.0f;
            if (FIRST != 1) {
                ko[c][0] = 1.0f * of_dp(oc, q, 1, x, g.nx);
                ko[c][1] = 1.0f * of_dp(oc, q, sy, y, g.ny);
                ko[c][2] = 1.0f * of_dp(oc, q, sz, z, g.nz);
            }
#pragma unroll
            for (int e = 0; e < 3; ++e) yo[c][e] = Y[(size_t)(3 * c + e) * n + q];
        }
        const float g0 = G[q], g1 = G[n + q], g2 = G[2 * n + q], g3 = G[3 * n + q];
        const float v0 = v[q], v1 = v[n + q], v2 = v[2 * n + q];
        const float k4 = g0 * v0 + g1 * v1 + g2 * v2;
        const float k4o = FIRST == 1 ? 0.0f
                          : FIRST == 2 ? k4old[q]
                                       : g0 * vold[q] + g1 * vold[n + q] + g2 * vold[2 * n + q];
        const float s2 = of_sigma2(g0, g1, g2);
        const float y4o = Y[9 * n + q];
        float r = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float t[3];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                t[e] = yo[c][e] + sigma1 * (kn[c][e] + kn[c][e] - ko[c][e]);
                if (huber) t[e] = t[e] * hf;
            }
            const float div = fmaxf(1.0f, sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) / lambda);
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const float yn = t[e] / div;
                Y[(size_t)(3 * c + e) * n + q] = yn;
                if (RES) r = r + fabsf((yo[c][e] - yn) / sigma1 - ko[c][e] + kn[c][e]);
            }
        }
        // std::max(-1, std::min(1, a)): min(1, a) keeps 1 unless a < 1, so NaN gives 1
        const float a = y4o + s2 * (k4 + k4 - k4o + g3);
        const float mn = a < 1.0f ? a : 1.0f;
        const float y4n = -1.0f < mn ? mn : -1.0f;
        Y[9 * n + q] = y4n;
        if (RES) r = r + fabsf((y4o - y4n) / s2 - k4o + k4);
        acc = (double)r;
    }
    if (RES) of_block_partial(acc, partial, tl.lin);
}

// Kx4 of the last iterate with the current image gradient, before the next
// warp replaces the gradient (the value the reference's Kx4 plane carries
// into that warp's first iteration).
__global__ __launch_bounds__(256) void k_of_k4(const float* __restrict__ v, const float* __restrict__ G,
                                               float* __restrict__ k4, VolGeom g) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    k4[q] = G[q] * v[q] + G[n + q] * v[n + q] + G[2 * n + q] * v[2 * n + q];
}

// dst = src on the ROI, for `planes` planes (the write-back of :503-531).
__global__ __launch_bounds__(256) void k_of_roi_copy(float* __restrict__ dst, const float* __restrict__ src, int planes,
                                                     VolGeom g, OfRoi roi) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || !of_in_roi(g, roi, x, y, z)) return;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    for (int m = 0; m < planes; ++m) dst[(size_t)m * n + q] = src[(size_t)m * n + q];
}

// ---- the flow pyramid's image operators (CImg 1.8.0) ----------------------
// One line per thread of the order-0 Deriche filter with Neumann boundary
// (_cimg_deriche_apply, CImg.h:32917-32989): forward pass into out, backward
// pass adds its response.  Lines run along `axis` of a d0 x d1 x d2 volume.
__global__ __launch_bounds__(256) void k_deriche(const float* __restrict__ in, float* __restrict__ out, int d0, int d1,
                                                 int d2, int axis, DericheCoef c) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int dims[3] = {d0, d1, d2};
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    if (t >= dims[o1] * dims[o2]) return;
    const int u = t % dims[o1], v = t / dims[o1];
    const size_t st[3] = {1, (size_t)d0, (size_t)d0 * d1};
    const size_t base = (size_t)u * st[o1] + (size_t)v * st[o2], off = st[axis];
    const int N = dims[axis];
    float xp = in[base], yb = c.coefp * xp, yp = yb;
    for (int m = 0; m < N; ++m) {
        const float xc = in[base + m * off];
        const float yc = c.a0 * xc + c.a1 * xp - c.b1 * yp - c.b2 * yb;
        out[base + m * off] = yc;
        xp = xc;
        yb = yp;
        yp = yc;
    }
    float xn = in[base + (size_t)(N - 1) * off], xa = xn, yn = c.coefn * xn, ya = yn;
    for (int n = N - 1; n >= 0; --n) {
        const float xc = in[base + n * off];
        const float yc = c.a2 * xn + c.a3 * xa - c.b1 * yn - c.b2 * ya;
        xa = xn;
        xn = xc;
        ya = yn;
        yn = yc;
        out[base + n * off] = out[base + n * off] + yc;
    }
}

// One axis of get_resize(..., 5): shrinking axes take the moving average
// (CImg.h:28135-28150) whose output t sums in[seg_s[k]] * seg_d[k] over
// k in [seg0[t], seg0[t+1]) and divides by the old length; growing axes take
// the cubic of CImg.h:28441-28475 at source index sidx[t] + frac[t], in double.
__global__ __launch_bounds__(256) void k_resize_axis(const float* __restrict__ in, float* __restrict__ out, int d0,
                                                     int d1, int d2, int axis, int ns, const int* __restrict__ seg0,
                                                     const int* __restrict__ seg_s, const int* __restrict__ seg_d,
                                                     const int* __restrict__ sidx, const double* __restrict__ frac) {
#pragma clang fp contract(off)
    const int dims[3] = {d0, d1, d2};
    int nd[3] = {d0, d1, d2};
    nd[axis] = ns;
    const size_t total = (size_t)nd[0] * nd[1] * nd[2];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int c[3];
    c[0] = (int)(i % nd[0]);
    c[1] = (int)((i / nd[0]) % nd[1]);
    c[2] = (int)(i / ((size_t)nd[0] * nd[1]));
    const int t = c[axis];
    const size_t st[3] = {1, (size_t)d0, (size_t)d0 * d1};
    c[axis] = 0;
    const size_t base = (size_t)c[0] * st[0] + (size_t)c[1] * st[1] + (size_t)c[2] * st[2];
    const size_t off = st[axis];
    const int w = dims[axis];
    if (w > ns) {
        float acc = 0.0f;
        for (int k = seg0[t]; k < seg0[t + 1]; ++k) acc += in[base + (size_t)seg_s[k] * off] * (float)(unsigned)seg_d[k];
        out[i] = acc / (float)(unsigned)w;
        return;
    }
    const long si = sidx[t];
    const double tt = frac[t], val1 = in[base + si * off];
    const double val0 = si > 0 ? (double)in[base + (si - 1) * off] : val1;
    const double val2 = si <= w - 2 ? (double)in[base + (si + 1) * off] : val1;
    const double val3 = si < w - 2 ? (double)in[base + (si + 2) * off] : val2;
    const double val = val1 + (double)0.5f * (tt * (-val0 + val2) + tt * tt * (2 * val0 - 5 * val1 + 4 * val2 - val3) +
                                              tt * tt * tt * (-val0 + 3 * val1 - 3 * val2 + val3));
    out[i] = (float)val;
}

static inline dim3 vol_grid(const VolGeom& g, int nitems) {
    return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz * nitems);
}

// Projector shape, VolGeom::fp_tile of the launching context (set by
// stt_set_projector_tile() / the tuner): 0 = slab projector (16x16 rays per
// block, LDS-staged volume boxes), -32 / -8 = slab projector with 32x8 / 8x32
// rays per block; 64, 32, 16, 8 = level-synchronous
// projector with that many rays per wave row (64x4, 32x8, 16x16 or 8x32 rays
// per 256-thread block, transposed copy for x-dominant views in the SART
// sweep).

template <int MODE, int CAP>
static void launch_fp_slab_cap(int tw, dim3 grid, const WorkItem* items, const VolGeom& g, float factor, hipStream_t s) {
    switch (tw) {
        case 32: hipLaunchKernelGGL((k_fp_slab<MODE, 32, 8, CAP>), grid, dim3(256), 0, s, items, g, factor); break;
        case 8: hipLaunchKernelGGL((k_fp_slab<MODE, 8, 32, CAP>), grid, dim3(256), 0, s, items, g, factor); break;
        default: hipLaunchKernelGGL((k_fp_slab<MODE, 16, 16, CAP>), grid, dim3(256), 0, s, items, g, factor); break;
    }
}

template <int MODE>
static hipError_t launch_fp_slab(const WorkItem* items, int nitems, const VolGeom& g, float factor, hipStream_t s) {
    // fp_tile 0: 16 x 16 rays per block; -32: 32 x 8; -8: 8 x 32.  fp_box_kib:
    // the staged box (16, 20, 24 or 28 KiB of LDS per block)
    const int tw = g.fp_tile == 0 ? 16 : -g.fp_tile, th = 256 / tw;
    const dim3 grid(((g.W + tw - 1) / tw) * ((g.H + th - 1) / th), nitems);
    switch (g.fp_box_kib) {
        case 16: launch_fp_slab_cap<MODE, 4096>(tw, grid, items, g, factor, s); break;
        case 20: launch_fp_slab_cap<MODE, 5120>(tw, grid, items, g, factor, s); break;
        case 28: launch_fp_slab_cap<MODE, 7168>(tw, grid, items, g, factor, s); break;
        default: launch_fp_slab_cap<MODE, kSlabCap>(tw, grid, items, g, factor, s); break;
    }
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_fp_level(const WorkItem* items, int nitems, const VolGeom& g, float factor, hipStream_t s) {
    const int tx = g.fp_tile, ty = 256 / tx;
    const dim3 grid(((g.W + tx - 1) / tx) * ((g.H + ty - 1) / ty), nitems);
    switch (tx) {
        case 64: hipLaunchKernelGGL((k_fp_level<MODE, 64>), grid, dim3(256), 0, s, items, g, factor); break;
        case 16: hipLaunchKernelGGL((k_fp_level<MODE, 16>), grid, dim3(256), 0, s, items, g, factor); break;
        case 8: hipLaunchKernelGGL((k_fp_level<MODE, 8>), grid, dim3(256), 0, s, items, g, factor); break;
        default: hipLaunchKernelGGL((k_fp_level<MODE, 32>), grid, dim3(256), 0, s, items, g, factor); break;
    }
    return hipGetLastError();
}

hipError_t launch_fp(int mode, const WorkItem* items, int nitems, const VolGeom& g, float factor, hipStream_t s) {
    constexpr int TX = 16, TY = 16;
    const int nb = ((g.W + TX - 1) / TX) * ((g.H + TY - 1) / TY);
    dim3 grid(nb, nitems);
    switch (mode) {
        case FP_MODE_FP:
            return g.fp_tile <= 0 ? launch_fp_slab<FP_MODE_FP>(items, nitems, g, factor, s)
                                  : launch_fp_level<FP_MODE_FP>(items, nitems, g, factor, s);
        case FP_MODE_SART:
            return g.fp_tile <= 0 ? launch_fp_slab<FP_MODE_SART>(items, nitems, g, factor, s)
                                  : launch_fp_level<FP_MODE_SART>(items, nitems, g, factor, s);
        case FP_MODE_SCATTER: hipLaunchKernelGGL((k_fp<FP_MODE_SCATTER, TX, TY>), grid, dim3(TX * TY), 0, s, items, g, factor); break;
        case FP_MODE_RAYBP:
            // slab shapes: LDS-accumulated slab scatter; level shapes: one
            // global atomic per sample and tap (same sums bit for bit)
            if (g.fp_tile <= 0) hipLaunchKernelGGL((k_raybp_slab<TX, TY>), grid, dim3(TX * TY), 0, s, items, g);
            else hipLaunchKernelGGL((k_fp<FP_MODE_RAYBP, TX, TY>), grid, dim3(TX * TY), 0, s, items, g, factor);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_transpose_xz(const WorkItem* items, int nitems, const VolGeom& g, hipStream_t s) {
    if (g.fp_tile <= 0) return hipSuccess;  // slab projectors read the x-fastest volume
    const dim3 grid((g.nx + 31) / 32, (g.nz + 31) / 32, g.ny * nitems);
    hipLaunchKernelGGL(k_transpose_xz, grid, dim3(256), 0, s, items, g);
    return hipGetLastError();
}

hipError_t launch_bp_voxel(const WorkItem* items, int nitems, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    const dim3 grid((g.nx + kBpTileX - 1) / kBpTileX, (g.ny + kBpTileY - 1) / kBpTileY,
                    ((g.nz + kBpVoxZ - 1) / kBpVoxZ) * nitems);
    hipLaunchKernelGGL(k_bp_voxel, grid, dim3(256), 0, s, items, g, pc);
    return hipGetLastError();
}

hipError_t launch_raybp_prep(const WorkItem* items, int nitems, const VolGeom& g, hipStream_t s) {
    hipLaunchKernelGGL(k_raybp_prep, dim3(nitems), dim3(1024), 0, s, items, g);
    return hipGetLastError();
}

hipError_t launch_bp_ray_apply(const WorkItem* items, int nitems, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    hipLaunchKernelGGL(k_bp_ray_apply, vol_grid(g, nitems), dim3(256), 0, s, items, g, pc);
    return hipGetLastError();
}

hipError_t launch_pd_dual(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    const dim3 grid = vol_grid(g, nf);
    switch (pd_warp_kind(pc)) {
        case 0: hipLaunchKernelGGL(k_pd_dual<0>, grid, dim3(256), 0, s, fb, f0, g, pc); break;
        case 1: hipLaunchKernelGGL(k_pd_dual<1>, grid, dim3(256), 0, s, fb, f0, g, pc); break;
        case 2: hipLaunchKernelGGL(k_pd_dual<2>, grid, dim3(256), 0, s, fb, f0, g, pc); break;
        default: hipLaunchKernelGGL(k_pd_dual<3>, grid, dim3(256), 0, s, fb, f0, g, pc); break;
    }
    return hipGetLastError();
}

hipError_t launch_pd_primal(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    const dim3 grid = vol_grid(g, nf);
    switch (pd_warp_kind(pc)) {
        case 0: hipLaunchKernelGGL(k_pd_primal<0>, grid, dim3(256), 0, s, fb, f0, g, pc); break;
        case 1: hipLaunchKernelGGL(k_pd_primal<1>, grid, dim3(256), 0, s, fb, f0, g, pc); break;
        case 2: hipLaunchKernelGGL(k_pd_primal<2>, grid, dim3(256), 0, s, fb, f0, g, pc); break;
        default: hipLaunchKernelGGL(k_pd_primal<3>, grid, dim3(256), 0, s, fb, f0, g, pc); break;
    }
    return hipGetLastError();
}

hipError_t launch_trace(const float* uvw, float* tb, float* tf, const VolGeom& g, int interp, hipStream_t s) {
    if (interp == 0)
        hipLaunchKernelGGL(k_trace<true>, vol_grid(g, 1), dim3(256), 0, s, uvw, g, tb, tf);
    else
        hipLaunchKernelGGL(k_trace<false>, vol_grid(g, 1), dim3(256), 0, s, uvw, g, tb, tf);
    return hipGetLastError();
}

hipError_t launch_extrapolate(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_extrapolate, vol_grid(g, nf), dim3(256), 0, s, fb, f0, g, pc);
    return hipGetLastError();
}

hipError_t launch_warp_vjp(const float* uvw, const float* src, const float* gout, float* gsrc, float* guvw,
                           const VolGeom& g, float dt, hipStream_t s) {
    hipLaunchKernelGGL(k_warp_vjp, vol_grid(g, 1), dim3(256), 0, s, uvw, src, gout, gsrc, guvw, g, -dt);
    return hipGetLastError();
}

hipError_t launch_residual_partial(const float* b, const float* ax, int n, double* partial, int stride,
                                   hipStream_t s) {
    hipLaunchKernelGGL(k_residual_partial, dim3(kReductionBlocks), dim3(256), 0, s, b, ax, n, partial, stride);
    return hipGetLastError();
}

hipError_t launch_dot_partial(const float* a, const float* c, size_t n, double* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_dot_partial, dim3(kReductionBlocks), dim3(256), 0, s, a, c, n, partial);
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double* partial, int n, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, n, out);
    return hipGetLastError();
}

// Device address of `scratch` on the current device: looked up once per
// device (hipGetSymbolAddress, with that device current) and kept.
int noalloc_scratch(char** ptr) {
    static std::mutex mu;
    static char* addr[kNoAllocMaxDevices] = {};
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        addr[dev] = (char*)p;
    }
    *ptr = addr[dev];
    return (int)hipSuccess;
}

hipError_t launch_item_make(WorkItem* dst, const ViewRec* view, float* vol, float* volT, const float* b, float* ax,
                            float* raylen, hipStream_t s) {
    hipLaunchKernelGGL(k_item_make, dim3(1), dim3(1), 0, s, dst, view, vol, volT, b, ax, raylen);
    return hipGetLastError();
}

hipError_t launch_zero_f64(double* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_add_partials(double* acc, const double* part, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_add_partials, dim3(1), dim3(256), 0, s, acc, part, n);
    return hipGetLastError();
}

hipError_t launch_scale_by_inv_sqrt(float* x, size_t n, const double* nrm2, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_by_inv_sqrt, dim3(1024), dim3(256), 0, s, x, n, nrm2);
    return hipGetLastError();
}

hipError_t launch_fill_hash(float* x, size_t n, unsigned seed, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_hash, dim3(1024), dim3(256), 0, s, x, n, seed);
    return hipGetLastError();
}

hipError_t launch_pi_apply_k(const float* v, float* y, int F, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    hipLaunchKernelGGL(k_pi_apply_k, vol_grid(g, F), dim3(256), 0, s, v, y, F, g, pc);
    return hipGetLastError();
}

hipError_t launch_pi_apply_kt(const float* y, float* w, int F, const VolGeom& g, hipStream_t s) {
    hipLaunchKernelGGL(k_pi_apply_kt, vol_grid(g, F), dim3(256), 0, s, y, w, F, g);
    return hipGetLastError();
}

hipError_t launch_warp(const float* uvw, const float* src, float* out, const VolGeom& g, float dt, int cubic,
                       hipStream_t s) {
    // warp(..., dt) traces with -dt (warping.h:126)
    hipLaunchKernelGGL(k_warp, vol_grid(g, 1), dim3(256), 0, s, uvw, src, out, g, -dt, cubic);
    return hipGetLastError();
}


// ---- optical flow launchers ------------------------------------------------

int of_partials(const VolGeom& g) { return ((g.nx + 63) / 64) * ((g.ny + 3) / 4) * g.nz; }

// The primal's grid: 64 x 4 blocks over every x and the ROI's y and z range.
static dim3 of_primal_grid(const VolGeom& g, const int* roi) {
    const int ex = g.nx - roi[0] - roi[1], ey = g.ny - roi[2] - roi[3], ez = g.nz - roi[4] - roi[5];
    if (ex <= 0 || ey <= 0 || ez <= 0) return dim3(0, 0, 0);
    return dim3((g.nx + 63) / 64, (ey + 3) / 4, ez);
}

int of_primal_partials(const VolGeom& g, const int* roi) {
    const dim3 d = of_primal_grid(g, roi);
    return (int)(d.x * d.y * d.z);
}

static inline OfRoi of_roi(const int* r) { return OfRoi{r[0], r[1], r[2], r[3], r[4], r[5]}; }

hipError_t launch_of_warp(const float* v, const float* img, float* out, const VolGeom& g, const float* ext,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_of_warp, vol_grid(g, 1), dim3(256), 0, s, v, img, out, g, ext[0], ext[1], ext[2]);
    return hipGetLastError();
}

hipError_t launch_of_prep(const float* W, const float* I1, const float* v, float* G, const VolGeom& g,
                          const int* roi, hipStream_t s) {
    hipLaunchKernelGGL(k_of_prep, vol_grid(g, 1), dim3(256), 0, s, W, I1, v, G, g, of_roi(roi));
    return hipGetLastError();
}

hipError_t launch_of_primal(int mode, const float* vin, float* vout, const float* Y, float* kty, const float* G,
                            double* partial, const VolGeom& g, const int* roi, hipStream_t s) {
    const dim3 grid = of_primal_grid(g, roi);
    if (grid.x == 0) return hipSuccess;  // empty ROI: no primal voxel, no partials
    const OfRoi r = of_roi(roi);
    switch (mode) {
        case 1: hipLaunchKernelGGL(k_of_primal<1>, grid, dim3(256), 0, s, vin, vout, Y, kty, G, partial, g, r); break;
        case 2: hipLaunchKernelGGL(k_of_primal<2>, grid, dim3(256), 0, s, vin, vout, Y, kty, G, partial, g, r); break;
        default: hipLaunchKernelGGL(k_of_primal<0>, grid, dim3(256), 0, s, vin, vout, Y, kty, G, partial, g, r); break;
    }
    return hipGetLastError();
}

hipError_t launch_of_dual(bool res, int first, const float* v, const float* vold, float* Y, const float* k4old,
                          const float* G, double* partial, const VolGeom& g, float sigma1, float hf, float lambda,
                          int huber, hipStream_t s) {
    const dim3 grid = vol_grid(g, 1);
#define STT_OF_DUAL(R, F)                                                                                          \
    hipLaunchKernelGGL((k_of_dual<R, F>), grid, dim3(256), 0, s, v, vold, Y, k4old, G, partial, g, sigma1, hf, \
                       lambda, huber)
    if (first == 1) {
        if (res) STT_OF_DUAL(true, 1);
        else STT_OF_DUAL(false, 1);
    } else if (first == 2) {
        if (res) STT_OF_DUAL(true, 2);
        else STT_OF_DUAL(false, 2);
    } else {
        if (res) STT_OF_DUAL(true, 0);
        else STT_OF_DUAL(false, 0);
    }
#undef STT_OF_DUAL
    return hipGetLastError();
}

hipError_t launch_of_k4(const float* v, const float* G, float* k4, const VolGeom& g, hipStream_t s) {
    hipLaunchKernelGGL(k_of_k4, vol_grid(g, 1), dim3(256), 0, s, v, G, k4, g);
    return hipGetLastError();
}

hipError_t launch_of_roi_copy(float* dst, const float* src, int planes, const VolGeom& g, const int* roi,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_of_roi_copy, vol_grid(g, 1), dim3(256), 0, s, dst, src, planes, g, of_roi(roi));
    return hipGetLastError();
}


hipError_t launch_deriche(const float* in, float* out, const int* dims, int axis, const DericheCoef& c, hipStream_t s) {
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    const int lines = dims[o1] * dims[o2];
    hipLaunchKernelGGL(k_deriche, dim3((lines + 255) / 256), dim3(256), 0, s, in, out, dims[0], dims[1], dims[2], axis, c);
    return hipGetLastError();
}

hipError_t launch_resize_axis(const float* in, float* out, const int* dims, int axis, int ns, const int* seg0,
                              const int* seg_s, const int* seg_d, const int* sidx, const double* frac,
                              hipStream_t s) {
    int nd[3] = {dims[0], dims[1], dims[2]};
    nd[axis] = ns;
    const size_t total = (size_t)nd[0] * nd[1] * nd[2];
    hipLaunchKernelGGL(k_resize_axis, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, dims[0],
                       dims[1], dims[2], axis, ns, seg0, seg_s, seg_d, sidx, frac);
    return hipGetLastError();
}

}  // namespace stt
