// ref_driver.cpp — calls the reference's own projector code, compiled here.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
// the reference sources where they lie (include paths point into the
// read-only reference tree; nothing is copied).  The translation unit
// #includes the reference's static "kernels" the way ST_Tomography.cpp does
// (ST_Tomography.cpp:839-841) and exposes them through extern "C" so the
// tests can pin the CPU oracle against the reference's own arithmetic:
//   Matrix4x4 / Quaternion / Vector4        (Math/*.cpp, compiled as-is)
//   ProduceRay                              (TypicalRayProducer.core.h:32-69)
//   CastingRay                              (RayCaster.core.h:42-187)
//   T_VolumeAccess trilinear + BP scatter   (SingleScalarVolume.core.h:41-146)
//   CImg<float>::_cubic_atXYZ               (CImg.h:13429)
//   dxp3..dzm3 finite differences           (tools.cpp:136-200, compiled as-is)
//   CImg get_gradient / get_blur / get_resize (CImg.h:34514, 33282, 27958),
//                                           the optical flow's image operators
// What the reference keeps in classes that need OpenGL headers absent from
// this image (Scene.cpp, TypicalRayProducer.cpp) is restated here as glue:
// the per-view matrix chain of Scene::_calculateMatrix (Scene.cpp:181-207)
// and TypicalRayProducer::_calculateMatrix (TypicalRayProducer.cpp:142-208),
// evaluated with the reference's own Matrix4x4/Quaternion operators, and the
// pixel loop of GenerateImage_MT (TypicalRayProducer.core.h:168-224).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "VolumeData.h"   // CImg container (reference)
#include "Quaternion.h"
#include "RayCaster.h"

// The including translation unit of the reference provides function-like
// min/max before the .core.h files; same semantics here.
#ifndef max
#define max(a, b) (((a) > (b)) ? (a) : (b))
#endif
#ifndef min
#define min(a, b) (((a) < (b)) ? (a) : (b))
#endif

#include "SingleScalarVolume.core.h"
#include "RayCaster.core.h"
#include "TypicalRayProducer.core.h"

#include "tools.h"
#include "../include/stt.h"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

struct RefView {
    Matrix4x4 P2M, M2V, MV, V2M;
    Vector4 eye;
};

RefView ref_view(const stt_geometry* g, const float* q) {
    // Scene members as set by the ST_Tomography ctor and _calculateCamera
    // (ST_Tomography.cpp:229-246, Scene.cpp:77,131,133).
    float sid = (float)g->sid, sdd = (float)g->sdd;
    double halfW = (g->det[0] - 1) * g->ds * 0.5f;
    double halfH = (g->det[1] - 1) * g->ds * 0.5f;
    float tr[3];
    tr[0] = (float)(-g->oxyz[0]);
    tr[2] = (float)(+g->oxyz[1]);
    tr[1] = (float)(-g->oxyz[2]);
    float ztb = -1 * sid;
    float nearp = sdd, farp = 10000.0;
    Quaternion rot;
    rot.x = q[0]; rot.y = q[1]; rot.z = q[2]; rot.w = q[3];
    float scale = 1.0f;
    Matrix4x4 view = Matrix4x4::TranslateMatrix(tr[0], tr[1], ztb + tr[2]) *
                     Matrix4x4::ScaleMatrix(scale, scale, scale) * rot.ToMatrix();
    Matrix4x4 proj;
    float* pm = proj.ele;
    pm[0] = nearp / halfW;
    pm[1] = pm[2] = pm[3] = 0.0f;
    pm[4] = 0.0f;
    pm[5] = nearp / halfH;
    pm[6] = pm[7] = pm[8] = pm[9] = 0.0f;
    pm[10] = -(farp + nearp) / (farp - nearp);
    pm[11] = -1.0f;
    pm[12] = pm[13] = 0.0f;
    pm[14] = -2.0f * farp * nearp / (farp - nearp);
    pm[15] = 0.0f;
    float volsize[3];
    for (int c = 0; c < 3; ++c) volsize[c] = (g->vol[c] - 1) * g->spacing[c];
    RefView r;
    r.MV = view * Matrix4x4::TranslateMatrix(-volsize[0] * 0.5f, -volsize[1] * 0.5f, -volsize[2] * 0.5f);
    Matrix4x4 m2v;
    for (int i = 0; i < 16; ++i) m2v.ele[i] = 0.0f;
    m2v.ele[0] = 1.0f / g->spacing[0];
    m2v.ele[5] = 1.0f / g->spacing[1];
    m2v.ele[10] = 1.0f / g->spacing[2];
    m2v.ele[15] = 1.0f;
    r.M2V = m2v;
    r.V2M = m2v;
    r.V2M.Inverse();
    Matrix4x4 v2p;
    for (int i = 0; i < 16; ++i) v2p.ele[i] = 0.0f;
    v2p.ele[0] = 0.5f * (g->det[0] - 1);
    v2p.ele[5] = 0.5f * (g->det[1] - 1);
    v2p.ele[10] = 0.5f;
    v2p.ele[12] = v2p.ele[0];
    v2p.ele[13] = v2p.ele[5];
    v2p.ele[14] = v2p.ele[10];
    v2p.ele[15] = 1.0;
    Matrix4x4 m2p = v2p * proj * r.MV;
    r.P2M = m2p;
    r.P2M.Inverse();
    r.eye[0] = r.P2M[8] / r.P2M[11];
    r.eye[1] = r.P2M[9] / r.P2M[11];
    r.eye[2] = r.P2M[10] / r.P2M[11];
    r.eye[3] = 1.0f;
    return r;
}

// Loads the reference's file-scope statics for one view, as
// GenerateImage_GetCorrections, SingleScalarVolumePrepare and RayCasterPrepare do.
void ref_prepare(const stt_geometry* g, const RefView& r, const float* vol, bool bp_mode, float* contrib,
                 float* wsum, float* corr, RayCaster& rc) {
    P2MMatrix = r.P2M;
    M2VMatrix = r.M2V;
    for (int c = 0; c < 3; ++c) Volume_Dims[c] = g->vol[c];
    CroppingBounds[0] = CroppingBounds[2] = CroppingBounds[4] = 0.0f;
    CroppingBounds[1] = (float)(Volume_Dims[0] - 1);
    CroppingBounds[3] = (float)(Volume_Dims[1] - 1);
    CroppingBounds[5] = (float)(Volume_Dims[2] - 1);
    EyePos = r.eye;
    rc.SetSampleDistance(std::sqrt(g->spacing[0] * g->spacing[0] + g->spacing[1] * g->spacing[1] +
                                   g->spacing[2] * g->spacing[2]));
    int dims[3] = {g->vol[0], g->vol[1], g->vol[2]};
    RayCasterPrepare(&rc, dims);
    if (va) delete va;
    CImg<float> img;
    if (vol) img.assign(vol, dims[0], dims[1], dims[2], 1);
    else img.assign(dims[0], dims[1], dims[2], 1, 0.0f);
    float sp[3] = {g->spacing[0], g->spacing[1], g->spacing[2]};
    va = new T_VolumeAccess(img, dims, sp, bp_mode, contrib, wsum, corr);
}

}  // namespace

REF_API int r_view_matrices(const stt_geometry* g, const float* q, float* out) {
    RefView r = ref_view(g, q);
    std::memcpy(out, r.P2M.ele, 64);
    std::memcpy(out + 16, r.M2V.ele, 64);
    std::memcpy(out + 32, r.MV.ele, 64);
    std::memcpy(out + 48, r.V2M.ele, 64);
    Matrix4x4 mvv2m = r.MV * r.V2M;
    std::memcpy(out + 64, mvv2m.ele, 64);
    std::memcpy(out + 80, r.eye.ele, 16);
    Vector4 o = r.M2V * r.eye;
    std::memcpy(out + 84, o.ele, 16);
    out[88] = std::sqrt(g->spacing[0] * g->spacing[0] + g->spacing[1] * g->spacing[1] + g->spacing[2] * g->spacing[2]);
    return 0;
}

// Quaternion(ax,ay,az,angle)*identity as InitViews builds it (ST_Tomography.cpp:899-911).
REF_API void r_quat_y(float theta_deg, float* q) {
    double angle = theta_deg * 3.1415926 / 180;
    Vector4 v;
    v[0] = 0; v[1] = 1; v[2] = v[3] = 0.0f;
    double length = v.Length();
    v *= 1.0f / length;
    Quaternion old;
    Quaternion r = Quaternion(v[0], v[1], v[2], length * angle) * old;
    q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
}

// Rotation of RotateVolume (ST_Tomography.cpp:865-884): q(deg) * q_old.
REF_API void r_quat_rotate_y(const float* q_old, float deg, float* q) {
    double angle = deg * 3.1415926 / 180;
    Quaternion old;
    old.x = q_old[0]; old.y = q_old[1]; old.z = q_old[2]; old.w = q_old[3];
    Vector4 v;
    v[0] = 0; v[1] = 1; v[2] = v[3] = 0.0f;
    double length = v.Length();
    v *= 1.0f / length;
    Quaternion r = Quaternion(v[0], v[1], v[2], length * angle) * old;
    q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
}

REF_API float r_matrix_inverse(const float* in, float* out) {
    Matrix4x4 m;
    std::memcpy(m.ele, in, 64);
    float det = m.Inverse();
    std::memcpy(out, m.ele, 64);
    return det;
}

// One forward projection: pixel loop of GenerateImage_MT (TypicalRayProducer.core.h:168-224)
// around the reference's ProduceRay/CastingRay/T_VolumeAccess.
REF_API int r_fp(const stt_geometry* g, const float* q, const float* vol, float* Ax, float* raylen, float* weight,
                 float* tnear_out, float* tfar_out) {
    RefView r = ref_view(g, q);
    RayCaster rc;
    ref_prepare(g, r, vol, false, nullptr, nullptr, nullptr, rc);
    const int W = g->det[0], H = g->det[1];
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            int k = i + j * W;
            Vector4 voxel_zero, voxel_unit, CameraDirection;
            float tnear, tfar;
            ProduceRay(i, j, voxel_zero, voxel_unit, CameraDirection, tnear, tfar);
            double value = 0.0f;
            float SumWeight_Ray = 0.0f;
            if (tnear_out) tnear_out[k] = tnear;
            if (tfar_out) tfar_out[k] = tfar;
            if (tnear < tfar) {
                if (raylen) raylen[k] = tfar - tnear;
                value = CastingRay(voxel_zero, voxel_unit, CameraDirection, tnear, tfar, SumWeight_Ray, k);
            }
            if (weight) weight[k] = (float)SumWeight_Ray;
            if (Ax) Ax[k] = value;
        }
    delete va;
    va = 0;
    return 0;
}

// Ray-driven BP scatter: BP mode of the same loop (SingleScalarVolume.core.h:128-143),
// run sequentially (the reference's OpenMP loop races on the accumulators).
REF_API int r_bp_ray_scatter(const stt_geometry* g, const float* q, const float* corr, float* contrib, float* wsum) {
    RefView r = ref_view(g, q);
    RayCaster rc;
    const int W = g->det[0], H = g->det[1];
    std::vector<float> c(corr, corr + (size_t)W * H);
    ref_prepare(g, r, nullptr, true, contrib, wsum, c.data(), rc);
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            int k = i + j * W;
            Vector4 voxel_zero, voxel_unit, CameraDirection;
            float tnear, tfar;
            ProduceRay(i, j, voxel_zero, voxel_unit, CameraDirection, tnear, tfar);
            float SumWeight_Ray = 0.0f;
            if (tnear < tfar) CastingRay(voxel_zero, voxel_unit, CameraDirection, tnear, tfar, SumWeight_Ray, k);
        }
    delete va;
    va = 0;
    return 0;
}

REF_API float r_cubic_atxyz(const float* img, int w, int h, int d, float fx, float fy, float fz) {
    CImg<float> im(img, w, h, d, 1, true);
    return im._cubic_atXYZ(fx, fy, fz, 0);
}

// tools.cpp finite differences over a whole volume: 0/1/2 = dxp3/dyp3/dzp3,
// 3/4/5 = dxm3/dym3/dzm3.
REF_API int r_of_diff(const int* dims, const float* a, int which, float* out) {
    int d[3] = {dims[0], dims[1], dims[2]};
    float* data = const_cast<float*>(a);
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                float v;
                switch (which) {
                    case 0: v = dxp3(data, d, i, j, k); break;
                    case 1: v = dyp3(data, d, i, j, k); break;
                    case 2: v = dzp3(data, d, i, j, k); break;
                    case 3: v = dxm3(data, d, i, j, k); break;
                    case 4: v = dym3(data, d, i, j, k); break;
                    default: v = dzm3(data, d, i, j, k); break;
                }
                out[index3DtoLinear(d, i, j, k)] = v;
            }
    return 0;
}

REF_API int r_gradient_central(const int* dims, const float* img, float* gx, float* gy, float* gz) {
    CImg<float> im(img, dims[0], dims[1], dims[2], 1, true);
    CImgList<float> g = im.get_gradient("xyz", 0);
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    std::memcpy(gx, g[0]._data, n * sizeof(float));
    std::memcpy(gy, g[1]._data, n * sizeof(float));
    std::memcpy(gz, g[2]._data, n * sizeof(float));
    return 0;
}

REF_API int r_blur(const int* dims, int spectrum, const float* img, float sigma, float* out) {
    CImg<float> im(img, dims[0], dims[1], dims[2], spectrum, true);
    CImg<float> b = im.get_blur(sigma);
    std::memcpy(out, b._data, b.size() * sizeof(float));
    return 0;
}

REF_API int r_resize(const int* dims, int spectrum, const float* img, int nx, int ny, int nz, int interp, float* out) {
    CImg<float> im(img, dims[0], dims[1], dims[2], spectrum, true);
    CImg<float> r = im.get_resize(nx, ny, nz, spectrum, interp);
    std::memcpy(out, r._data, r.size() * sizeof(float));
    return 0;
}

// ---- on-disk formats through CImg itself (CImg.h save_analyze / _load_analyze,
// get_split).  TIFF needs libtiff headers, absent here, so it is not built.

template <typename T>
static int save_analyze_as(const char* path, const float* img, int w, int h, int d, int c, const float* voxel) {
    CImg<T> im(w, h, d, c);
    for (size_t i = 0; i < im.size(); ++i) im._data[i] = (T)img[i];
    im.save_analyze(path, voxel);
    return 0;
}

// datatype: 2 uchar, 4 short, 8 int, 16 float, 64 double (CImg picks it from T)
REF_API int r_save_analyze(const char* path, const float* img, int w, int h, int d, int c, const float* voxel,
                           int datatype) {
    try {
        switch (datatype) {
            case 2: return save_analyze_as<unsigned char>(path, img, w, h, d, c, voxel);
            case 4: return save_analyze_as<short>(path, img, w, h, d, c, voxel);
            case 8: return save_analyze_as<int>(path, img, w, h, d, c, voxel);
            case 16: return save_analyze_as<float>(path, img, w, h, d, c, voxel);
            case 64: return save_analyze_as<double>(path, img, w, h, d, c, voxel);
            default: return 1;
        }
    } catch (...) {
        return 2;
    }
}

// Loads with CImg<float>::load_analyze; dims[4] = w, h, d, c; voxel[3].
// out may be NULL to query the dims only.
REF_API int r_load_analyze(const char* path, int* dims, float* voxel, float* out, size_t n) {
    try {
        CImg<float> im;
        float vs[3] = {0, 0, 0};
        im.load_analyze(path, vs);
        dims[0] = im.width();
        dims[1] = im.height();
        dims[2] = im.depth();
        dims[3] = im.spectrum();
        if (voxel) std::memcpy(voxel, vs, sizeof vs);
        if (out) {
            if (n < im.size()) return 3;
            std::memcpy(out, im._data, im.size() * sizeof(float));
        }
        return 0;
    } catch (...) {
        return 2;
    }
}

// Depths of CImg::get_split('z', nb) over an image of depth `size`.
REF_API int r_split_sizes(int size, int nb, int* sizes) {
    try {
        CImg<float> im(1, 1, size, 1, 0.f);
        CImgList<float> parts = im.get_split('z', nb);
        for (unsigned int k = 0; k < parts.size(); ++k) sizes[k] = parts[k].depth();
        return (int)parts.size();
    } catch (...) {
        return -1;
    }
}
