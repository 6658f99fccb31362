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

// T_VolumeAccess::GetScalarValue trilinear (SingleScalarVolume.core.h:87-146).
// The volume allocation carries a zero tail of nx*ny + nx + 1 floats, so the
// +1 taps need no clamp: where the reference clamps a tap its weight is 0.
__device__ __forceinline__ float sample_tri(const float* __restrict__ vol, const VolGeom& g, float px, float py,
                                            float pz) {
    px = fminf(fmaxf(px, 0.0f), g.bx);
    py = fminf(fmaxf(py, 0.0f), g.by);
    pz = fminf(fmaxf(pz, 0.0f), g.bz);
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    const float fx = px - (float)ix, fy = py - (float)iy, fz = pz - (float)iz;
    const int sxy = g.nx * g.ny;
    const float* p = vol + (ix + g.nx * iy + sxy * iz);
    const float v000 = p[0], v100 = p[1];
    const float v010 = p[g.nx], v110 = p[g.nx + 1];
    const float v001 = p[sxy], v101 = p[sxy + 1];
    const float v011 = p[sxy + g.nx], v111 = p[sxy + g.nx + 1];
    const float gx = 1.0f - fx;
    const float a = v000 * gx + v100 * fx;
    const float b = v010 * gx + v110 * fx;
    const float c = v001 * gx + v101 * fx;
    const float d = v011 * gx + v111 * fx;
    const float gy = 1.0f - fy;
    const float ab = a * gy + b * fy;
    const float cd = c * gy + d * fy;
    return ab * (1.0f - fz) + cd * fz;
}

// Scatter of the 8 trilinear weights (SingleScalarVolume.core.h:99-105, 128-143).
__device__ __forceinline__ void scatter_tri(float* __restrict__ contrib, float* __restrict__ wsum, const VolGeom& g,
                                            float px, float py, float pz, float val) {
    px = fminf(fmaxf(px, 0.0f), g.bx);
    py = fminf(fmaxf(py, 0.0f), g.by);
    pz = fminf(fmaxf(pz, 0.0f), g.bz);
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    const float t0 = px - (float)ix, t1 = py - (float)iy, t2 = pz - (float)iz;
    const float m0 = 1.0f - t0, m1 = 1.0f - t1, m2 = 1.0f - t2;
    const float w[8] = {m0 * m1 * m2, t0 * m1 * m2, m0 * t1 * m2, t0 * t1 * m2,
                        m0 * m1 * t2, t0 * m1 * t2, m0 * t1 * t2, t0 * t1 * t2};
    const int sxy = g.nx * g.ny;
    const int base = ix + g.nx * iy + sxy * iz;
    const int off[8] = {0, 1, g.nx, g.nx + 1, sxy, sxy + 1, sxy + g.nx, sxy + g.nx + 1};
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        atomicAdd(contrib + base + off[n], val * w[n]);
        if (wsum) atomicAdd(wsum + base + off[n], w[n]);
    }
}

// CastingRay (RayCaster.core.h:42-187) with one block = the whole volume
// (RayCasterPrepare :199-201).  Samples: the entry point once more than the
// grid (:60-61), then t_k = tnear + k*sd below the next block boundary; the only
// interior boundaries are the planes at voxel coordinate 0.5 crossed in the
// negative direction, where the grid restarts at ceil((t_b - tnear)/sd).
template <bool SCATTER>
__device__ __forceinline__ float cast_ray(const float* __restrict__ vol, float* __restrict__ contrib,
                                          float* __restrict__ wsum, float sval, const VolGeom& g, const Ray& r,
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
    if (SCATTER) scatter_tri(contrib, wsum, g, pos1[0], pos1[1], pos1[2], sval);
    else density = sample_tri(vol, g, pos1[0], pos1[1], pos1[2]) * sd;
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
        while (t < tb && n < g.max_samples) {
            if (SCATTER) scatter_tri(contrib, wsum, g, px, py, pz, sval);
            else density = fmaf(sample_tri(vol, g, px, py, pz), sd, density);
            ++n;
            t += sd;
            px += step[0];
            py += step[1];
            pz += step[2];
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

// Detector tile of TX x TY rays per 256-thread block.  Blocks are dealt to the
// 8 XCDs round-robin; remapping makes each XCD own a contiguous band of
// detector rows, so its L2 sees one slab of the volume.
template <int TX, int TY>
__device__ __forceinline__ bool tile_pixel(const VolGeom& g, int& i, int& j) {
    const int tiles_x = (g.W + TX - 1) / TX;
    const int tiles_y = (g.H + TY - 1) / TY;
    const int nb = tiles_x * tiles_y;
    int b = blockIdx.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    const int tx = b % tiles_x, ty = b / tiles_x;
    i = tx * TX + (threadIdx.x % TX);
    j = ty * TY + (threadIdx.x / TX);
    return i < g.W && j < g.H;
}

template <int MODE, int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_fp(const WorkItem* __restrict__ items, VolGeom g, float factor) {
    int i, j;
    if (!tile_pixel<TX, TY>(g, i, j)) return;
    const WorkItem& it = items[blockIdx.y];
    const ViewRec& v = *it.view;
    const int k = i + j * g.W;
    Ray r;
    produce_ray(v, g, (float)i, (float)j, r);
    const bool hit = r.tnear < r.tfar;
    if (MODE == FP_MODE_SCATTER) {
        // Ray BP / adjoint: the BP-mode pass of GenerateImage_MT.
        if (hit) {
            int ns;
            const float c = it.corr[k] * factor;
            cast_ray<true>(nullptr, it.contrib, it.wsum, c, g, r, ns);
            if (it.raylen) it.raylen[k] = r.tfar - r.tnear;
        }
        return;
    }
    float value = 0.0f;
    int ns = 0;
    if (hit) value = cast_ray<false>(it.vol, nullptr, nullptr, 0.0f, g, r, ns);
    if (MODE == FP_MODE_FP) {
        if (hit && it.raylen) it.raylen[k] = r.tfar - r.tnear;
        if (it.ax) it.ax[k] = value;
        if (it.weight) it.weight[k] = (float)ns;
    } else {
        // SART correction (ST_Tomography.cpp:667-676); pY is zero when read.
#pragma clang fp contract(off)
        float len;
        if (hit) {
            len = r.tfar - r.tnear;
            it.raylen[k] = len;
        } else {
            len = it.raylen[k];  // stale value of the last view whose ray k hit
        }
        const float num = factor * it.b[k] - factor * value;
        const float den = factor * len + 1.0f;
        it.corr[k] = num / den;
    }
}

// Catmull-Rom of cubicInterpolate (ST_Tomography.cpp:183-186).
__device__ __forceinline__ float cubic_cr(float p0, float p1, float p2, float p3, float x) {
    return p1 + 0.5f * x * (p2 - p0 + x * (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3 + x * (3.0f * (p1 - p2) + p3 - p0)));
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
    if (it.xbar && !cut_out(g, pc, x, y, z)) it.xbar[q] = 2.0f * xv - it.xprev[q];
    return xv;
}

// Voxel-driven backprojection (ST_Tomography.cpp:687-731).  64x4 voxels of one
// z-slice per block; grid.z = z + nz*item.
__global__ __launch_bounds__(256) void k_bp_voxel(const WorkItem* __restrict__ items, VolGeom g, PdConst pc) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z % g.nz;
    const WorkItem& it = items[blockIdx.z / g.nz];
    if (x >= g.nx || y >= g.ny) return;
    const ViewRec& v = *it.view;
    const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    const float fx = (float)x, fy = (float)y, fz = (float)z;
    const float* M = v.MVV2M;
    const float e0 = M[0] * fx + M[4] * fy + M[8] * fz + M[12];
    const float e1 = M[1] * fx + M[5] * fy + M[9] * fz + M[13];
    const float e2 = M[2] * fx + M[6] * fy + M[10] * fz + M[14];
    float pxf = v.halfW - (e0 / e2) * v.sdd;
    float pyf = v.halfH - (e1 / e2) * v.sdd;
    pxf = pxf / v.ds;
    pyf = pyf / v.ds;
    const int px = (int)pxf, py = (int)pyf;
    const bool inside = !(px < 2 || px >= g.W - 2 || py < 2 || py >= g.H - 2);
    if (!inside && !it.last) return;
    float xv = it.vol[q];
    if (inside) {
        const float dx = pxf - (float)px, dy = pyf - (float)py;
        const float* C = it.corr + (size_t)(py - 1) * g.W + (px - 1);
        float col[4][4];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int a = 0; a < 4; ++a) col[a][b] = C[b * g.W + a];
        float arr[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) arr[a] = cubic_cr(col[a][0], col[a][1], col[a][2], col[a][3], dy);
        const float gv = cubic_cr(arr[0], arr[1], arr[2], arr[3], dx);
        xv = xv + pc.alpha * gv;
        xv = tmax(tmin(pc.gmax, xv), 0.0f);
    }
    if (it.last) xv = finalize_voxel(it, g, pc, x, y, z, q, xv);
    it.vol[q] = xv;
}

// Ray-BP normalisation (ST_Tomography.cpp:746-762): x += contrib/wsum with the
// macro clamp (0/0 -> NaN -> 0); clears the accumulators (Initial(), :763).
__global__ __launch_bounds__(256) void k_bp_ray_apply(const WorkItem* __restrict__ items, VolGeom g, PdConst pc) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z % g.nz;
    const WorkItem& it = items[blockIdx.z / g.nz];
    if (x >= g.nx || y >= g.ny) return;
    const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    const float m = (float)((double)it.contrib[q] / (double)it.wsum[q]);
    float xv = it.vol[q] + m;
    xv = tmax(tmin(pc.gmax, xv), 0.0f);
    it.contrib[q] = 0.0f;
    it.wsum[q] = 0.0f;
    if (it.last) xv = finalize_voxel(it, g, pc, x, y, z, q, xv);
    it.vol[q] = xv;
}

// ---------------------------------------------------------------------------
// Warp (warping.h:10-145) on the grid aabb = [0, dims].

// CImg _cubic_atXYZ weights for one axis (CImg.h:13429-13520) in weight form.
__device__ __forceinline__ void cr_weights(float t, float w[4]) {
    const float t2 = t * t, t3 = t2 * t;
    w[0] = 0.5f * (-t + 2.0f * t2 - t3);
    w[1] = 1.0f + 0.5f * (-5.0f * t2 + 3.0f * t3);
    w[2] = 0.5f * (t + 4.0f * t2 - 3.0f * t3);
    w[3] = 0.5f * (-t2 + t3);
}

struct SamplePt {
    int ix, iy, iz;         // base cell of the clamped grid point
    int xs[4], ys[4], zs[4];
    float wx[4], wy[4], wz[4];
    float fx, fy, fz;
};

// world_to_cell_grid + clamp to [1, dim-2] (warping.h:31-36, 54-63).
__device__ __forceinline__ void sample_point(const VolGeom& g, float px, float py, float pz, bool cubic, SamplePt& s) {
    const float dims[3] = {(float)g.nx, (float)g.ny, (float)g.nz};
    const float p[3] = {px, py, pz};
    float gc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = (dims[c] * p[c]) / dims[c] - 0.5f;
        v = fmaxf(1.0f, fminf(dims[c] - 2.0f, v));
        gc[c] = v;
    }
    s.ix = (int)gc[0]; s.iy = (int)gc[1]; s.iz = (int)gc[2];
    s.fx = gc[0] - (float)s.ix; s.fy = gc[1] - (float)s.iy; s.fz = gc[2] - (float)s.iz;
    if (cubic) {
        const int n[3] = {g.nx, g.ny, g.nz};
        const int base[3] = {s.ix, s.iy, s.iz};
        const float f[3] = {s.fx, s.fy, s.fz};
        int* idx[3] = {s.xs, s.ys, s.zs};
        float* w[3] = {s.wx, s.wy, s.wz};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int b = base[c];
            idx[c][0] = b - 1 < 0 ? 0 : b - 1;
            idx[c][1] = b;
            idx[c][2] = f[c] > 0.0f ? b + 1 : b;
            idx[c][3] = b + 2 >= n[c] ? n[c] - 1 : b + 2;
            cr_weights(f[c], w[c]);
        }
    }
}

// sample_image (warping.h:49-88): clamped Catmull-Rom, bounded by the min/max
// of the 8 corners of the cell; or trilinear at the same point.
__device__ __forceinline__ float sample_img(const float* __restrict__ img, const VolGeom& g, const SamplePt& s,
                                            bool cubic) {
    const int sy = g.nx, sz = g.nx * g.ny;
    const float* c0 = img + (s.ix + sy * s.iy + sz * s.iz);
    const float a000 = c0[0], a100 = c0[1], a010 = c0[sy], a110 = c0[sy + 1];
    const float a001 = c0[sz], a101 = c0[sz + 1], a011 = c0[sz + sy], a111 = c0[sz + sy + 1];
    if (!cubic) {
        const float gx = 1.0f - s.fx, gy = 1.0f - s.fy;
        const float c00 = a000 * gx + a100 * s.fx, c10 = a010 * gx + a110 * s.fx;
        const float c01 = a001 * gx + a101 * s.fx, c11 = a011 * gx + a111 * s.fx;
        const float e0 = c00 * gy + c10 * s.fy, e1 = c01 * gy + c11 * s.fy;
        return e0 * (1.0f - s.fz) + e1 * s.fz;
    }
    const float mn = fminf(fminf(fminf(a000, a100), fminf(a010, a110)), fminf(fminf(a001, a101), fminf(a011, a111)));
    const float mx = fmaxf(fmaxf(fmaxf(a000, a100), fmaxf(a010, a110)), fmaxf(fmaxf(a001, a101), fmaxf(a011, a111)));
    float acc_z = 0.0f;
#pragma unroll
    for (int kz = 0; kz < 4; ++kz) {
        float acc_y = 0.0f;
#pragma unroll
        for (int ky = 0; ky < 4; ++ky) {
            const float* row = img + (size_t)sy * s.ys[ky] + (size_t)sz * s.zs[kz];
            const float r = s.wx[0] * row[s.xs[0]] + s.wx[1] * row[s.xs[1]] + s.wx[2] * row[s.xs[2]] +
                            s.wx[3] * row[s.xs[3]];
            acc_y = fmaf(s.wy[ky], r, acc_y);
        }
        acc_z = fmaf(s.wz[kz], acc_y, acc_z);
    }
    return fmaxf(mn, fminf(mx, acc_z));
}

// trace (warping.h:100-114) from the cell centre of (i,j,k) with trace step
// tdt (= -dt of warp), then the image sample.  The first velocity sample sits
// exactly on grid node clamp(i,1,n-2), where the clamped cubic returns the node
// value, so it is read directly.
__device__ __forceinline__ float warp_voxel(const float* __restrict__ img, const float* __restrict__ uvw,
                                            const VolGeom& g, int i, int j, int k, float tdt, bool cubic) {
    const size_t V = (size_t)g.nx * g.ny * g.nz;
    const int ci = min(max(i, 1), g.nx - 2), cj = min(max(j, 1), g.ny - 2), ck = min(max(k, 1), g.nz - 2);
    const size_t nq = (size_t)ci + (size_t)g.nx * ((size_t)cj + (size_t)g.ny * ck);
    const float p[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
    float v1[3] = {uvw[nq], uvw[V + nq], uvw[2 * V + nq]};
    float p2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) p2[c] = p[c] + tdt * v1[c];
    SamplePt s;
    sample_point(g, p2[0], p2[1], p2[2], cubic, s);
    float v2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v2[c] = sample_img(uvw + c * V, g, s, cubic);
    float q[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float p0 = p2[c] - tdt * v2[c];
        q[c] = p2[c] + 0.5f * (p[c] - p0);
    }
    sample_point(g, q[0], q[1], q[2], cubic, s);
    return sample_img(img, g, s, cubic);
}

__global__ __launch_bounds__(256) void k_warp(const float* __restrict__ uvw, const float* __restrict__ src,
                                              float* __restrict__ out, VolGeom g, float tdt, int cubic) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny) return;
    const size_t q = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    out[q] = warp_voxel(src, uvw, g, x, y, z, tdt, cubic != 0);
}

// Dual step of frame f (ST_Tomography.cpp:502-534); grid.z = z + nz*(f - f0).
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
    const float* xb0 = B.xbar;
    const float* xb1 = B1.xbar;
    // m_backward_warp = warp(u_f, X_bar[f+1], dt = -1): trace step +1 (:503)
    const float bw = warp_voxel(xb1, B.flow, g, x, y, z, 1.0f, pc.interp == 0);
    const float xq = xb0[q];
    float ytp = B.ytp[q] + pc.sigma * (xb1[q] - xq);  // (:504)
    float ybc = B.ybc[q] + pc.sigma * (bw - xq);      // (:505)
    if (x >= 1 && x < g.nx - 1 && y >= 1 && y < g.ny - 1 && z >= 1 && z < g.nz - 1 && !cut_out(g, pc, x, y, z)) {
        const float h = pc.huber_factor;
        float y0 = B.ytv[q] + pc.sigma * (xb0[q + 1] - xq) * h;
        float y1 = B.ytv[V + q] + pc.sigma * (xb0[q + sy] - xq) * h;
        float y2 = B.ytv[2 * V + q] + pc.sigma * (xb0[q + sz] - xq) * h;
        if (pc.prior == 0) {  // ATV per-component clamp (:523-525)
            y0 = tmax(-pc.lambda, tmin(pc.lambda, y0));
            y1 = tmax(-pc.lambda, tmin(pc.lambda, y1));
            y2 = tmax(-pc.lambda, tmin(pc.lambda, y2));
        } else {  // isotropic TV: projection onto the lambda L2 ball
            const float nrm = sqrtf(y0 * y0 + y1 * y1 + y2 * y2);
            const float sc = nrm > pc.lambda ? pc.lambda / nrm : 1.0f;
            y0 *= sc; y1 *= sc; y2 *= sc;
        }
        B.ytv[q] = y0;
        B.ytv[V + q] = y1;
        B.ytv[2 * V + q] = y2;
        ytp = ytp * pc.ktp / (pc.ktp + 2.0f * pc.sigma);  // (:528)
        ybc = tmax(-pc.ebc, tmin(pc.ebc, ybc));           // (:530)
    }
    B.ytp[q] = ytp;
    B.ybc[q] = ybc;
}

// Primal step of frame f (ST_Tomography.cpp:536-578).
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
    float xv = B.x[q];
    B.xprev[q] = xv;  // (:536)
    if (f == 0) {     // (:540-545)
        xv = xv - pc.tau * (-B.ytp[q]);
        xv = xv - pc.tau * (-B.ybc[q]);
    } else {          // (:546-555): forward warp of Y_bc[f-1] with u_{f-1}, trace step -1
        const FrameBufs& Bm = fb[f - 1];
        const float fw = warp_voxel(Bm.ybc, Bm.flow, g, x, y, z, -1.0f, pc.interp == 0);
        xv = xv - pc.tau * (Bm.ytp[q] - B.ytp[q]);
        xv = xv - pc.tau * (fw - B.ybc[q]);
    }
    if (x >= 1 && x < g.nx - 1 && y >= 1 && y < g.ny - 1 && z >= 1 && z < g.nz - 1 && !cut_out(g, pc, x, y, z)) {
        const float* t = B.ytv;
        const float y1 = -(t[q] - t[q - 1] + t[V + q] - t[V + q - sy] + t[2 * V + q] - t[2 * V + q - sz]);  // (:570-571)
        xv = xv - pc.tau * y1;
        xv = tmax(tmin(pc.gmax, xv), 0.0f);  // (:575)
    }
    B.x[q] = xv;
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
    B.xbar[q] = 2.0f * B.x[q] - B.xprev[q];
}

// ---------------------------------------------------------------------------
// Launchers.

static inline dim3 vol_grid(const VolGeom& g, int nitems) {
    return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz * nitems);
}

hipError_t launch_fp(int mode, const WorkItem* items, int nitems, const VolGeom& g, float factor, hipStream_t s) {
    constexpr int TX = 16, TY = 16;
    const int nb = ((g.W + TX - 1) / TX) * ((g.H + TY - 1) / TY);
    dim3 grid(nb, nitems);
    switch (mode) {
        case FP_MODE_FP: hipLaunchKernelGGL((k_fp<FP_MODE_FP, TX, TY>), grid, dim3(TX * TY), 0, s, items, g, factor); break;
        case FP_MODE_SART: hipLaunchKernelGGL((k_fp<FP_MODE_SART, TX, TY>), grid, dim3(TX * TY), 0, s, items, g, factor); break;
        case FP_MODE_SCATTER: hipLaunchKernelGGL((k_fp<FP_MODE_SCATTER, TX, TY>), grid, dim3(TX * TY), 0, s, items, g, factor); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bp_voxel(const WorkItem* items, int nitems, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    hipLaunchKernelGGL(k_bp_voxel, vol_grid(g, nitems), dim3(256), 0, s, items, g, pc);
    return hipGetLastError();
}

hipError_t launch_bp_ray_apply(const WorkItem* items, int nitems, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    hipLaunchKernelGGL(k_bp_ray_apply, vol_grid(g, nitems), dim3(256), 0, s, items, g, pc);
    return hipGetLastError();
}

hipError_t launch_pd_dual(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    hipLaunchKernelGGL(k_pd_dual, vol_grid(g, nf), dim3(256), 0, s, fb, f0, g, pc);
    return hipGetLastError();
}

hipError_t launch_pd_primal(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc, hipStream_t s) {
    hipLaunchKernelGGL(k_pd_primal, vol_grid(g, nf), dim3(256), 0, s, fb, f0, g, pc);
    return hipGetLastError();
}

hipError_t launch_extrapolate(const FrameBufs* fb, int f0, int nf, const VolGeom& g, const PdConst& pc,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_extrapolate, vol_grid(g, nf), dim3(256), 0, s, fb, f0, g, pc);
    return hipGetLastError();
}

hipError_t launch_warp(const float* uvw, const float* src, float* out, const VolGeom& g, float dt, int cubic,
                       hipStream_t s) {
    // warp(..., dt) traces with -dt (warping.h:126)
    hipLaunchKernelGGL(k_warp, vol_grid(g, 1), dim3(256), 0, s, uvw, src, out, g, -dt, cubic);
    return hipGetLastError();
}

}  // namespace stt
