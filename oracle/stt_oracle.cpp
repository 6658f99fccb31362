// stt_oracle.cpp — CPU oracle of the space-time tomography hot path.
//
// TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline); never part of
// the shipped path.  Every function restates the reference file:line cited
// next to it.  Citations are relative to the reference repository root
// (SpaceTimeTomography/...).  Built with -ffp-contract=off so fp32 results
// follow the reference's plain IEEE expression order.
//
// R = float reproduces the reference's mixed precision (double literals in
// cubicInterpolate, double m_halfFrustumWidth/m_ds/m_alpha); R = double is
// the FP64 oracle (all arithmetic in double).
//
// Pins of the optical-flow solver: its warp, gradient, blur and resize by
// closed forms and the reference's own code compiled in oracle/_ref; its fixed
// point by a duality-gap certificate; its stopping residual
// err = (d^2 + p^2) / nPx (optical_flow.h:359-362, 465-486, 500-502) by an
// FP64 recomputation of the adaptive-PDHG primal-dual residual from three
// consecutive iterates (tests/test_of_oracle.py).
#include "stt_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <functional>
#include <type_traits>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Reference min/max macros (ST_Tomography.cpp:167-173, RayCaster.core.h:22-28):
// ternaries, not fmin/fmax, so NaN handling matches.
template <class A, class B> inline typename std::common_type<A, B>::type MAXM(A a, B b) { return (a > b) ? a : b; }
template <class A, class B> inline typename std::common_type<A, B>::type MINM(A a, B b) { return (a < b) ? a : b; }

template <class R> struct M4 { R e[16]; };

// Matrix4x4::IdentityMatrix (Matrix4x4.h IdentityMatrix)
template <class R> M4<R> ident() {
    M4<R> m;
    for (int i = 0; i < 16; ++i) m.e[i] = R(0);
    m.e[0] = m.e[5] = m.e[10] = m.e[15] = R(1);
    return m;
}

// operator*(Matrix4x4, Matrix4x4), column-major (Math/Matrix4x4.h:66-91)
template <class R> M4<R> mul(const M4<R>& A, const M4<R>& B) {
    M4<R> r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row)
            r.e[4 * c + row] = A.e[row] * B.e[4 * c + 0] + A.e[4 + row] * B.e[4 * c + 1] +
                               A.e[8 + row] * B.e[4 * c + 2] + A.e[12 + row] * B.e[4 * c + 3];
    return r;
}

// operator*(Matrix4x4, Vector4) (Math/Matrix4x4.h:294-304)
template <class R> void mulv(const M4<R>& M, const R v[4], R out[4]) {
    R r[4];
    for (int row = 0; row < 4; ++row)
        r[row] = M.e[row] * v[0] + M.e[4 + row] * v[1] + M.e[8 + row] * v[2] + M.e[12 + row] * v[3];
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}

// Det3x3 macro and Matrix4x4::Inverse, cofactor expansion (Math/Matrix4x4.cpp:79-135)
template <class R> inline R det3(R a, R b, R c, R d, R e, R f, R g, R h, R i) {
    return a * e * i + b * f * g + c * d * h - c * e * g - a * f * h - b * d * i;
}
template <class R> R inverse(M4<R>& m) {
    R* el = m.e;
    R a1 = el[0], b1 = el[1], c1 = el[2], d1 = el[3];
    R a2 = el[4], b2 = el[5], c2 = el[6], d2 = el[7];
    R a3 = el[8], b3 = el[9], c3 = el[10], d3 = el[11];
    R a4 = el[12], b4 = el[13], c4 = el[14], d4 = el[15];
    R dt1 = det3(b2, b3, b4, c2, c3, c4, d2, d3, d4);
    R dt2 = det3(a2, a3, a4, c2, c3, c4, d2, d3, d4);
    R dt3 = det3(a2, a3, a4, b2, b3, b4, d2, d3, d4);
    R dt4 = det3(a2, a3, a4, b2, b3, b4, c2, c3, c4);
    R det = a1 * dt1 - b1 * dt2 + c1 * dt3 - d1 * dt4;
    if (det == R(0)) return R(0);
    R inv = R(1) / det;
    el[0] = dt1 * inv;
    el[4] = -dt2 * inv;
    el[8] = dt3 * inv;
    el[12] = -dt4 * inv;
    el[1] = -det3(b1, b3, b4, c1, c3, c4, d1, d3, d4) * inv;
    el[5] = det3(a1, a3, a4, c1, c3, c4, d1, d3, d4) * inv;
    el[9] = -det3(a1, a3, a4, b1, b3, b4, d1, d3, d4) * inv;
    el[13] = det3(a1, a3, a4, b1, b3, b4, c1, c3, c4) * inv;
    el[2] = det3(b1, b2, b4, c1, c2, c4, d1, d2, d4) * inv;
    el[6] = -det3(a1, a2, a4, c1, c2, c4, d1, d2, d4) * inv;
    el[10] = det3(a1, a2, a4, b1, b2, b4, d1, d2, d4) * inv;
    el[14] = -det3(a1, a2, a4, b1, b2, b4, c1, c2, c4) * inv;
    el[3] = -det3(b1, b2, b3, c1, c2, c3, d1, d2, d3) * inv;
    el[7] = det3(a1, a2, a3, c1, c2, c3, d1, d2, d3) * inv;
    el[11] = -det3(a1, a2, a3, b1, b2, b3, d1, d2, d3) * inv;
    el[15] = det3(a1, a2, a3, b1, b2, b3, c1, c2, c3) * inv;
    return det;
}

// Matrix4x4::TranslateMatrix (Math/Matrix4x4.cpp:137-142)
template <class R> M4<R> translate(R tx, R ty, R tz) {
    M4<R> m = ident<R>();
    m.e[12] = tx; m.e[13] = ty; m.e[14] = tz;
    return m;
}

// Quaternion::ToMatrix (Math/Quaternion.h:46-70); q = {x, y, z, w}
template <class R> M4<R> quat_to_matrix(const R q[4]) {
    const R x = q[0], y = q[1], z = q[2], w = q[3];
    M4<R> mat;
    R* m = mat.e;
    m[0] = R(1) - R(2) * (y * y + z * z);
    m[1] = R(2) * (x * y + z * w);
    m[2] = R(2) * (z * x - y * w);
    m[3] = R(0);
    m[4] = R(2) * (x * y - z * w);
    m[5] = R(1) - R(2) * (z * z + x * x);
    m[6] = R(2) * (y * z + x * w);
    m[7] = R(0);
    m[8] = R(2) * (z * x + y * w);
    m[9] = R(2) * (y * z - x * w);
    m[10] = R(1) - R(2) * (x * x + y * y);
    m[11] = R(0);
    m[12] = R(0); m[13] = R(0); m[14] = R(0); m[15] = R(1);
    return mat;
}

// Per-view state assembled as in Scene::_calculateCamera/_calculateMatrix
// (Scene.cpp:69-138,181-207), TypicalRayProducer::_calculateMatrix
// (TypicalRayProducer.cpp:142-208) and GenerateImage_GetCorrections
// (TypicalRayProducer.core.h:230-263).
template <class R> struct View {
    M4<R> P2M, M2V, MV, V2M, MVV2M;
    R eye[4];
    R o[4];
    double halfW, halfH, ds;  // double members of Scene (Scene.h:114-116)
    R sdd;                    // float m_sdd (Scene.h:112)
    R sd;                     // sample distance ||spacing|| (ST_Tomography.cpp:300)
    R bounds[6];              // CroppingBounds [0, dims-1] (TypicalRayProducer.core.h:241-244)
    int dims[3];
    int W, H;
};

template <class R> View<R> make_view(const stt_geometry& g, const float qf[4]) {
    View<R> v;
    const R q[4] = {R(qf[0]), R(qf[1]), R(qf[2]), R(qf[3])};
    // ST_Tomography ctor (ST_Tomography.cpp:229-246): halves in double, translate from oxyz.
    v.halfW = (g.det[0] - 1) * g.ds * 0.5f;
    v.halfH = (g.det[1] - 1) * g.ds * 0.5f;
    v.ds = g.ds;
    // float Scene::m_sid / m_sdd (Scene.h:111-112); the FP64 oracle keeps full precision.
    const R sid = sizeof(R) == 4 ? R(float(g.sid)) : R(g.sid);
    const R sdd = sizeof(R) == 4 ? R(float(g.sdd)) : R(g.sdd);
    v.sdd = sdd;
    R tr[3];
    tr[0] = R(-g.oxyz[0]);
    tr[2] = R(+g.oxyz[1]);
    tr[1] = R(-g.oxyz[2]);
    if (sizeof(R) == 4) { tr[0] = R(float(tr[0])); tr[1] = R(float(tr[1])); tr[2] = R(float(tr[2])); }
    // _calculateCamera (Scene.cpp:77,131,133)
    const R ztb = R(-1) * sid;
    const R nearp = sdd;
    const R farp = R(10000.0);
    // _calculateMatrix (Scene.cpp:187-206): View = T * S(1) * R(q)
    M4<R> S = ident<R>();
    M4<R> view = mul(mul(translate<R>(tr[0], tr[1], ztb + tr[2]), S), quat_to_matrix<R>(q));
    M4<R> proj;
    R* pm = proj.e;
    for (int i = 0; i < 16; ++i) pm[i] = R(0);
    pm[0] = R(nearp / v.halfW);
    pm[5] = R(nearp / v.halfH);
    pm[10] = -(farp + nearp) / (farp - nearp);
    pm[11] = R(-1);
    pm[14] = R(-2) * farp * nearp / (farp - nearp);
    pm[15] = R(0);
    // TypicalRayProducer::SetVolSize (TypicalRayProducer.cpp:116-118) and _calculateMatrix (:142-208)
    R volsize[3];
    for (int c = 0; c < 3; ++c) volsize[c] = R(g.vol[c] - 1) * R(g.spacing[c]);
    v.MV = mul(view, translate<R>(-volsize[0] * R(0.5), -volsize[1] * R(0.5), -volsize[2] * R(0.5)));
    M4<R> m2v;
    for (int i = 0; i < 16; ++i) m2v.e[i] = R(0);
    m2v.e[0] = R(1) / R(g.spacing[0]);
    m2v.e[5] = R(1) / R(g.spacing[1]);
    m2v.e[10] = R(1) / R(g.spacing[2]);
    m2v.e[15] = R(1);
    v.M2V = m2v;
    v.V2M = m2v;
    inverse(v.V2M);
    M4<R> v2p;
    for (int i = 0; i < 16; ++i) v2p.e[i] = R(0);
    v2p.e[0] = R(0.5) * R(g.det[0] - 1);
    v2p.e[5] = R(0.5) * R(g.det[1] - 1);
    v2p.e[10] = R(0.5);
    v2p.e[12] = v2p.e[0];
    v2p.e[13] = v2p.e[5];
    v2p.e[14] = v2p.e[10];
    v2p.e[15] = R(1);
    M4<R> m2p = mul(mul(v2p, proj), v.MV);
    v.P2M = m2p;
    inverse(v.P2M);
    // EyePos (TypicalRayProducer.core.h:255-258)
    v.eye[0] = v.P2M.e[8] / v.P2M.e[11];
    v.eye[1] = v.P2M.e[9] / v.P2M.e[11];
    v.eye[2] = v.P2M.e[10] / v.P2M.e[11];
    v.eye[3] = R(1);
    mulv(v.M2V, v.eye, v.o);  // voxel_zero = M2VMatrix*EyePos (TypicalRayProducer.core.h:43)
    v.MVV2M = mul(v.MV, v.V2M);  // m_model_view*m_VolumeToModelMatrix (ST_Tomography.cpp:699)
    // sample distance sqrt(s.s) in float (ST_Tomography.cpp:300)
    if (sizeof(R) == 4)
        v.sd = R(std::sqrt(float(g.spacing[0] * g.spacing[0] + g.spacing[1] * g.spacing[1] + g.spacing[2] * g.spacing[2])));
    else
        v.sd = std::sqrt(R(g.spacing[0]) * g.spacing[0] + R(g.spacing[1]) * g.spacing[1] + R(g.spacing[2]) * g.spacing[2]);
    for (int c = 0; c < 3; ++c) {
        v.dims[c] = g.vol[c];
        v.bounds[2 * c] = R(0);
        v.bounds[2 * c + 1] = R(g.vol[c] - 1);
    }
    v.W = g.det[0];
    v.H = g.det[1];
    return v;
}

// ProduceRay (TypicalRayProducer.core.h:32-69)
template <class R> inline void produce_ray(const View<R>& v, R pixX, R pixY, R o[4], R d[4], R& tnear, R& tfar) {
    R pv[4] = {pixX, pixY, R(0), R(1)};
    R mp[4];
    mulv(v.P2M, pv, mp);
    R winv = R(1) / mp[3];
    for (int i = 0; i < 4; ++i) mp[i] *= winv;
    R dir[4];
    for (int i = 0; i < 4; ++i) dir[i] = mp[i] - v.eye[i];
    R linv = R(1) / std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);  // Vector4::Normalize (Vector4.h:125-131)
    dir[0] *= linv; dir[1] *= linv; dir[2] *= linv;
    for (int i = 0; i < 4; ++i) o[i] = v.o[i];
    mulv(v.M2V, dir, d);
    R t1, t2;
    t1 = (v.bounds[0] - o[0]) / d[0];
    t2 = (v.bounds[1] - o[0]) / d[0];
    R tminX = MINM(t1, t2), tmaxX = MAXM(t1, t2);
    t1 = (v.bounds[2] - o[1]) / d[1];
    t2 = (v.bounds[3] - o[1]) / d[1];
    R tminY = MINM(t1, t2), tmaxY = MAXM(t1, t2);
    t1 = (v.bounds[4] - o[2]) / d[2];
    t2 = (v.bounds[5] - o[2]) / d[2];
    R tminZ = MINM(t1, t2), tmaxZ = MAXM(t1, t2);
    tnear = MAXM(MAXM(tminX, tminY), tminZ);
    tnear = MAXM(tnear, R(0));
    tfar = MINM(MINM(tmaxX, tmaxY), tmaxZ);
}

// T_VolumeAccess::GetScalarValue(Vector4 ...) trilinear (SingleScalarVolume.core.h:60-71, 87-146).
// scatter != nullptr selects BP mode: the 8 weights go to contrib/wsum, returns 0.
template <class R> struct Scatter {
    R* contrib;
    R* wsum;
    R corr;
};
template <class R> inline R trilinear_sample(const View<R>& v, const R* vol, const R p[4], R& sumw, Scatter<R>* sc) {
    const int* D = v.dims;
    R pos[3];
    for (int c = 0; c < 3; ++c) pos[c] = MAXM(MINM(p[c], R(D[c] - 1)), R(0));
    int ip[3] = {int(pos[0]), int(pos[1]), int(pos[2])};
    auto at = [&](int x, int y, int z) -> R {
        x = MAXM(MINM(x, D[0] - 1), 0);
        y = MAXM(MINM(y, D[1] - 1), 0);
        z = MAXM(MINM(z, D[2] - 1), 0);
        return vol[(size_t)x + (size_t)D[0] * ((size_t)y + (size_t)D[1] * z)];
    };
    R t0 = pos[0] - R(ip[0]), t1 = pos[1] - R(ip[1]), t2 = pos[2] - R(ip[2]);
    R mt0 = R(1) - t0, mt1 = R(1) - t1, mt2 = R(1) - t2;
    if (sc) {
        const R w[8] = {mt0 * mt1 * mt2, t0 * mt1 * mt2, mt0 * t1 * mt2, t0 * t1 * mt2,
                        mt0 * mt1 * t2,  t0 * mt1 * t2,  mt0 * t1 * t2,  t0 * t1 * t2};
        const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        for (int n = 0; n < 8; ++n) {  // UpdateContributionSum (SingleScalarVolume.core.h:73-84)
            int x = MAXM(MINM(ip[0] + off[n][0], D[0] - 1), 0);
            int y = MAXM(MINM(ip[1] + off[n][1], D[1] - 1), 0);
            int z = MAXM(MINM(ip[2] + off[n][2], D[2] - 1), 0);
            size_t k = (size_t)x + (size_t)D[0] * ((size_t)y + (size_t)D[1] * z);
            sc->contrib[k] += sc->corr * w[n];
            sc->wsum[k] += w[n];
        }
        return R(0);
    }
    sumw += R(1);
    R f2 = pos[0] - R(ip[0]);
    R f1 = R(1) - f2;
    R a = at(ip[0], ip[1], ip[2]) * f1 + at(ip[0] + 1, ip[1], ip[2]) * f2;
    R b = at(ip[0], ip[1] + 1, ip[2]) * f1 + at(ip[0] + 1, ip[1] + 1, ip[2]) * f2;
    R c = at(ip[0], ip[1], ip[2] + 1) * f1 + at(ip[0] + 1, ip[1], ip[2] + 1) * f2;
    R d = at(ip[0], ip[1] + 1, ip[2] + 1) * f1 + at(ip[0] + 1, ip[1] + 1, ip[2] + 1) * f2;
    f2 = pos[1] - R(ip[1]);
    f1 = R(1) - f2;
    a = a * f1 + b * f2;
    c = c * f1 + d * f2;
    f2 = pos[2] - R(ip[2]);
    f1 = R(1) - f2;
    return a * f1 + c * f2;
}

// CastingRay (RayCaster.core.h:42-187) with bsize = dims (RayCasterPrepare :191-207).
template <class R> R casting_ray(const View<R>& v, const R* vol, const R o[4], const R d[4], R tnear, R tfar,
                                 R& sumw, Scatter<R>* sc) {
    const R sd = v.sd;
    const int* bs = v.dims;
    R density = R(0);
    R t1 = tnear;
    R pos1[4];
    for (int i = 0; i < 4; ++i) pos1[i] = o[i] + d[i] * t1;
    R step[4];
    for (int i = 0; i < 4; ++i) step[i] = d[i] * sd;
    density += trilinear_sample(v, vol, pos1, sumw, sc) * sd;
    R tMax[3], tDelta[3] = {R(0), R(0), R(0)};
    for (int c = 0; c < 3; ++c) {
        if (d[c] == R(0)) {
            tMax[c] = tfar;
        } else if (d[c] < R(0)) {
            int blk = (int(std::ceil(pos1[c] - R(0.5))) - 1) / bs[c];
            tMax[c] = (R(blk * bs[c]) + R(0.5) - o[c]) / d[c];
            tDelta[c] = R(-bs[c]) / d[c];
        } else {
            int blk = int(std::floor(pos1[c] - R(0.5))) / bs[c];
            tMax[c] = (R((blk + 1) * bs[c]) + R(0.5) - o[c]) / d[c];
            tDelta[c] = R(bs[c]) / d[c];
        }
    }
    while (t1 < tfar) {
        R tb = MINM(MINM(tMax[0], tMax[1]), tMax[2]);
        if (tb > tfar) tb = tfar;
        R t = std::ceil((t1 - tnear) / sd) * sd + tnear;
        R pos[4];
        for (int i = 0; i < 4; ++i) pos[i] = o[i] + d[i] * t;
        while (t < tb) {
            density += trilinear_sample(v, vol, pos, sumw, sc) * sd;
            t += sd;
            for (int i = 0; i < 4; ++i) pos[i] += step[i];
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
    return density;
}

// GenerateImage_MT (TypicalRayProducer.core.h:123-225); pixel centres are integers (:195).
template <class R> void fp_view(const View<R>& v, const R* vol, R* Ax, R* raylen, R* weight, const R* corr,
                                R* contrib, R* wsum) {
    const int W = v.W, H = v.H;
    if (corr) {  // BP mode: sequential over pixels so the summation order is fixed (the reference races here)
        for (int j = 0; j < H; ++j)
            for (int i = 0; i < W; ++i) {
                const int k = i + j * W;
                R o[4], d[4], tn, tf;
                produce_ray(v, R(i), R(j), o, d, tn, tf);
                R sw = R(0);
                if (tn < tf) {
                    if (raylen) raylen[k] = tf - tn;
                    Scatter<R> sc{contrib, wsum, corr[k]};
                    casting_ray(v, vol, o, d, tn, tf, sw, &sc);
                }
                if (weight) weight[k] = R(0);
                if (Ax) Ax[k] = R(0);
            }
        return;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < H; ++j) {
        for (int i = 0; i < W; ++i) {
            const int k = i + j * W;
            R o[4], d[4], tn, tf;
            produce_ray(v, R(i), R(j), o, d, tn, tf);
            R value = R(0);
            R sw = R(0);
            if (tn < tf) {
                if (raylen) raylen[k] = tf - tn;
                value = casting_ray<R>(v, vol, o, d, tn, tf, sw, nullptr);
            }
            if (weight) weight[k] = sw;
            if (Ax) Ax[k] = value;
        }
    }
}

// cubicInterpolate / bicubicInterpolate (ST_Tomography.cpp:183-197).  For fp32 the
// reference evaluates in double (double literals) and rounds to float.
inline float cubic_ref(const float p[4], float x) {
    return p[1] + 0.5 * x * (p[2] - p[0] + x * (2.0 * p[0] - 5.0 * p[1] + 4.0 * p[2] - p[3] + x * (3.0 * (p[1] - p[2]) + p[3] - p[0])));
}
inline double cubic_ref(const double p[4], double x) {
    return p[1] + 0.5 * x * (p[2] - p[0] + x * (2.0 * p[0] - 5.0 * p[1] + 4.0 * p[2] - p[3] + x * (3.0 * (p[1] - p[2]) + p[3] - p[0])));
}
template <class R> inline R bicubic_ref(R p[4][4], R x, R y) {
    R arr[4];
    arr[0] = cubic_ref(p[0], y);
    arr[1] = cubic_ref(p[1], y);
    arr[2] = cubic_ref(p[2], y);
    arr[3] = cubic_ref(p[3], y);
    return cubic_ref(arr, x);
}

// Voxel-driven backprojection (ST_Tomography.cpp:687-731).
template <class R> void bp_voxel(const View<R>& v, const R* corr, double alpha, R gmax, R* vol) {
    const int* D = v.dims;
    const int W = v.W, H = v.H;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < D[2]; ++z)
        for (int y = 0; y < D[1]; ++y)
            for (int x = 0; x < D[0]; ++x) {
                const size_t pos = (size_t)x + (size_t)D[0] * ((size_t)y + (size_t)D[1] * z);
                R vec[4] = {R(x), R(y), R(z), R(1)};
                R e[4];
                mulv(v.MVV2M, vec, e);
                R px_f = R(v.halfW - (e[0] / e[2]) * v.sdd);
                R py_f = R(v.halfH - (e[1] / e[2]) * v.sdd);
                px_f = R(px_f / v.ds);
                py_f = R(py_f / v.ds);
                int px = int(px_f), py = int(py_f);
                R dx = px_f - R(px), dy = py_f - R(py);
                if (px < 2 || px >= W - 2 || py < 2 || py >= H - 2) continue;
                const R* pC = corr + (size_t)py * W + px;
                const int iy = W;
                R pimg[4][4] = {{pC[-1 - iy], pC[-1], pC[-1 + iy], pC[-1 + 2 * iy]},
                                {pC[-iy], pC[0], pC[iy], pC[2 * iy]},
                                {pC[1 - iy], pC[1], pC[1 + iy], pC[1 + 2 * iy]},
                                {pC[2 - iy], pC[2], pC[2 + iy], pC[2 + 2 * iy]}};
                R gv = bicubic_ref(pimg, dx, dy);
                R xv = R(vol[pos] + alpha * gv);
                vol[pos] = MAXM(MINM(gmax, xv), R(0));
            }
}

// CImg<float>::_cubic_atXYZ (CImg.h:13429-13520): Neumann clamp, Catmull-Rom, order x, y, z.
template <class R> R cubic_atxyz(const R* img, int w, int h, int dd, R fx, R fy, R fz) {
    auto cut = [](R v, int lo, int hi) -> R { return v < R(lo) ? R(lo) : (v > R(hi) ? R(hi) : v); };
    const R nfx = cut(fx, 0, w - 1), nfy = cut(fy, 0, h - 1), nfz = cut(fz, 0, dd - 1);
    const int x = int(nfx), y = int(nfy), z = int(nfz);
    const R dx = nfx - R(x), dy = nfy - R(y), dz = nfz - R(z);
    const int px = x - 1 < 0 ? 0 : x - 1, nx = dx > 0 ? x + 1 : x, ax = x + 2 >= w ? w - 1 : x + 2;
    const int py = y - 1 < 0 ? 0 : y - 1, ny = dy > 0 ? y + 1 : y, ay = y + 2 >= h ? h - 1 : y + 2;
    const int pz = z - 1 < 0 ? 0 : z - 1, nz = dz > 0 ? z + 1 : z, az = z + 2 >= dd ? dd - 1 : z + 2;
    auto I = [&](int xi, int yi, int zi) -> R { return img[(size_t)xi + (size_t)w * ((size_t)yi + (size_t)h * zi)]; };
    auto cub = [](R t, R p, R c, R n, R a) -> R {
        return c + R(0.5) * (t * (-p + n) + t * t * (R(2) * p - R(5) * c + R(4) * n - a) + t * t * t * (-p + R(3) * c - R(3) * n + a));
    };
    const int xs[4] = {px, x, nx, ax}, ys[4] = {py, y, ny, ay}, zs[4] = {pz, z, nz, az};
    R plane[4];
    for (int kz = 0; kz < 4; ++kz) {
        R row[4];
        for (int ky = 0; ky < 4; ++ky)
            row[ky] = cub(dx, I(xs[0], ys[ky], zs[kz]), I(xs[1], ys[ky], zs[kz]), I(xs[2], ys[ky], zs[kz]), I(xs[3], ys[ky], zs[kz]));
        plane[kz] = cub(dy, row[0], row[1], row[2], row[3]);
    }
    return cub(dz, plane[0], plane[1], plane[2], plane[3]);
}

// warping::world_to_cell_grid for aabb = [0, ext] (warping.h:31-36): float ops, then -0.5 in double.
// The ST path uses ext = dims; the flow pyramid keeps the finest level's extent.
template <class R> inline R to_cell(int dim, R world, R ext = R(-1)) {
    R lo = R(0), hi = ext < R(0) ? R(dim) : ext;
    if (sizeof(R) == 4) return R(double(R(dim) * (world - lo) / (hi - lo)) - 0.5);
    return R(dim) * (world - lo) / (hi - lo) - R(0.5);
}

// warping::sample_image (warping.h:49-88): clamp to [1, dim-2], 8-corner min/max, clamped cubic.
// interp = STT_INTERP_TRILINEAR replaces the cubic by trilinear at the same point.
template <class R> R sample_image(const int* dims, const R* img, const R p[3], int interp, const R* ext = nullptr) {
    R g[3];
    for (int c = 0; c < 3; ++c) {
        g[c] = to_cell<R>(dims[c], p[c], ext ? ext[c] : R(-1));
        g[c] = std::max(R(1), std::min(R(dims[c] - 2), g[c]));
    }
    const int ix = int(g[0]), iy = int(g[1]), iz = int(g[2]);
    const int nx = dims[0], ny = dims[1];
    auto I = [&](int x, int y, int z) -> R { return img[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)]; };
    if (interp == STT_INTERP_TRILINEAR) {
        R fx = g[0] - R(ix), fy = g[1] - R(iy), fz = g[2] - R(iz);
        R c00 = I(ix, iy, iz) * (R(1) - fx) + I(ix + 1, iy, iz) * fx;
        R c10 = I(ix, iy + 1, iz) * (R(1) - fx) + I(ix + 1, iy + 1, iz) * fx;
        R c01 = I(ix, iy, iz + 1) * (R(1) - fx) + I(ix + 1, iy, iz + 1) * fx;
        R c11 = I(ix, iy + 1, iz + 1) * (R(1) - fx) + I(ix + 1, iy + 1, iz + 1) * fx;
        R c0 = c00 * (R(1) - fy) + c10 * fy;
        R c1 = c01 * (R(1) - fy) + c11 * fy;
        return c0 * (R(1) - fz) + c1 * fz;
    }
    R mn = I(ix, iy, iz), mx = mn;
    const int co[7][3] = {{1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int n = 0; n < 7; ++n) {
        R val = I(ix + co[n][0], iy + co[n][1], iz + co[n][2]);
        mn = std::min(mn, val);
        mx = std::max(mx, val);
    }
    R cv = cubic_atxyz<R>(img, dims[0], dims[1], dims[2], g[0], g[1], g[2]);
    return std::max(mn, std::min(mx, cv));
}

// warping::trace (warping.h:100-114) and warp (:116-145); velocity planes u, v, w.
template <class R>
void warp(const int* dims, const R* uvw, const R* src, R dt, R* out, int interp, int nthreads, const R* ext = nullptr) {
    const size_t V = (size_t)dims[0] * dims[1] * dims[2];
    const R* vel[3] = {uvw, uvw + V, uvw + 2 * V};
    auto job = [&](int z) {
        for (int y = 0; y < dims[1]; ++y)
            for (int x = 0; x < dims[0]; ++x) {
                // cell_grid_to_world (warping.h:24-29): (i + 0.5) * (hi - lo) / dim in double
                R p[3];
                const int ijk[3] = {x, y, z};
                for (int c = 0; c < 3; ++c) {
                    const R e = ext ? ext[c] : R(dims[c]);
                    p[c] = R(0.0 + (double(R(ijk[c])) + 0.5) * double(e - R(0)) / double(R(dims[c])));
                }
                const R tdt = -dt;
                R v1[3], p2[3], v2[3], p0[3], q[3];
                for (int c = 0; c < 3; ++c) v1[c] = sample_image<R>(dims, vel[c], p, interp, ext);
                for (int c = 0; c < 3; ++c) p2[c] = p[c] + tdt * v1[c];
                for (int c = 0; c < 3; ++c) v2[c] = sample_image<R>(dims, vel[c], p2, interp, ext);
                for (int c = 0; c < 3; ++c) p0[c] = p2[c] - tdt * v2[c];
                for (int c = 0; c < 3; ++c) q[c] = p2[c] + R(0.5) * (p[c] - p0[c]);
                out[(size_t)x + (size_t)dims[0] * ((size_t)y + (size_t)dims[1] * z)] = sample_image<R>(dims, src, q, interp, ext);
            }
    };
    if (nthreads > 0) {  // the reference's interleaved z-slices on std::threads (warping.h:122-142)
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t)
            th.emplace_back([&, t] {
                for (int z = t; z < dims[2]; z += nthreads) job(z);
            });
        for (auto& t : th) t.join();
    } else {
#pragma omp parallel for schedule(static)
        for (int z = 0; z < dims[2]; ++z) job(z);
    }
}

// imROI (ST_Tomography.cpp:303-329)
inline bool roi_in(const int* D, const int* cut, int jtop, int i, int j, int k) {
    if (((i < cut[6]) && (k < cut[7])) || ((i < cut[6]) && (k > (D[2] - cut[7])))) return false;
    if (((i > (D[0] - cut[6])) && (k < cut[7])) || ((i > (D[0] - cut[6])) && (k > (D[2] - cut[7])))) return false;
    if (((k < cut[6]) && (i < cut[7])) || ((k < cut[6]) && (i > (D[0] - cut[7])))) return false;
    if (((k > (D[2] - cut[6])) && (i < cut[7])) || ((k > (D[2] - cut[6])) && (i > (D[0] - cut[7])))) return false;
    if (i < cut[0] || i > D[0] - cut[1] || j < cut[2] || j > D[1] - jtop || k < cut[4] || k > D[2] - cut[5]) return false;
    return true;
}
// ROI test of the PD loops (ST_Tomography.cpp:514, 565, 593)
inline bool cut_out(const int* D, const int* cut, int i, int j, int k) {
    return i < cut[0] || i > D[0] - cut[1] || j < cut[2] || j > D[1] - cut[3] || k < cut[4] || k > D[2] - cut[5];
}

}  // namespace

// ---------------------------------------------------------------------------
// Full solver state (ST_Tomography members and globals, ST_Tomography.cpp:44-166, 339-350)
struct o_ctx_base {
    virtual ~o_ctx_base() {}
};

template <class R> struct OState : o_ctx_base {
    stt_geometry g;
    stt_params p;
    int F;
    int D[3];
    size_t V;
    int W, H;
    std::vector<int> P;
    std::vector<std::vector<float>> quats;  // per frame, 4*P
    std::vector<std::vector<View<R>>> views;
    std::vector<std::vector<R>> x, xbar, xprev, ytv, ytp, ybc, flow, b;
    std::vector<R> raylen;                    // shared detector buffer of the single TypicalRayProducer
    std::vector<std::vector<R>> raylen_frame; // per-frame variant
    bool per_frame_raylen = false;
    R gmax = R(10);
    bool gmax_set = false;
    bool pd_ready = false;  // X_bar initialised from x (first PD pass)
    // effective parameters (float globals in the reference)
    R sigma, tau, lam, ktp, ebc, hub, factor;
    double alpha;

    void derive_params() {
        sigma = R(p.sigma);
        tau = R(p.tau);
        lam = R(p.lambda_tv);
        ktp = R(p.k_tp);
        ebc = R(p.e_bc);
        hub = R(p.huber);
        alpha = p.alpha;
        // Factor = sqrtf(2*SART_lambda), SART_lambda = CP_tau/2 (ST_Tomography.cpp:110-111)
        if (sizeof(R) == 4) factor = R(std::sqrt(float(2) * (float(tau) / float(2))));
        else factor = std::sqrt(R(2) * (tau / R(2)));
        if (p.g_max > 0) { gmax = R(p.g_max); gmax_set = true; }
    }
    R huberfactor() const {  // m_huberfactor (ST_Tomography.cpp:219)
        if (sizeof(R) == 4) return R(1.0f / (1.0f + double(float(sigma)) * p.huber / p.lambda_tv));
        return R(1) / (R(1) + sigma * hub / lam);
    }
};

struct o_ctx {
    int f64;
    o_ctx_base* st;
};

template <class R> static void sart_frame(OState<R>& s, int f, const int* order, int n, bool faithful) {
    const size_t Dn = (size_t)s.W * s.H;
    std::vector<R> Ax(Dn), corr(Dn), weight(Dn);
    std::vector<R>& rl = s.per_frame_raylen ? s.raylen_frame[f] : s.raylen;
    std::vector<R> contrib, wsum, copy1, copy2;
    if (s.p.bp == STT_BP_RAY) { contrib.assign(s.V, R(0)); wsum.assign(s.V, R(0)); }
    R* vol = s.x[f].data();
    for (int i = 0; i < n; ++i) {
        const int view = order ? order[i] : i;
        const View<R>& v = s.views[f][view];
        if (faithful) {  // T_VolumeAccess(CImg data) by value + idata = data (SingleScalarVolume.core.h:41-44,218)
            copy1.assign(vol, vol + s.V);
            copy2 = copy1;
            fp_view<R>(v, copy2.data(), Ax.data(), rl.data(), weight.data(), nullptr, nullptr, nullptr);
        } else {
            fp_view<R>(v, vol, Ax.data(), rl.data(), weight.data(), nullptr, nullptr, nullptr);
        }
        // correction (ST_Tomography.cpp:667-676); pY is identically zero when read (:618, :678-682)
        const R* bv = s.b[f].data() + (size_t)view * Dn;
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)Dn; ++k) {
            R num = s.factor * bv[k] - s.factor * Ax[k] - R(0);
            R den = R(s.factor * rl[k] + 1.0);
            corr[k] = num / den;
        }
        if (s.p.bp == STT_BP_VOXEL) {
            bp_voxel<R>(v, corr.data(), s.alpha, s.gmax, vol);
        } else {
            // BP-mode cast (ST_Tomography.cpp:735-739), normalise (:746-762), Initial() (:763)
            fp_view<R>(v, vol, nullptr, rl.data(), nullptr, corr.data(), contrib.data(), wsum.data());
#pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)s.V; ++k) {
                R m = R((1.0 * contrib[k]) / wsum[k]);
                R xv = vol[k] + m;
                vol[k] = MAXM(MINM(s.gmax, xv), R(0));
                contrib[k] = R(0);
                wsum[k] = R(0);
            }
            std::fill(rl.begin(), rl.end(), R(0));
        }
    }
    // x *= imROI (ST_Tomography.cpp:793-804)
#pragma omp parallel for schedule(static)
    for (int z = 0; z < s.D[2]; ++z)
        for (int y = 0; y < s.D[1]; ++y)
            for (int xx = 0; xx < s.D[0]; ++xx) {
                size_t k = (size_t)xx + (size_t)s.D[0] * ((size_t)y + (size_t)s.D[1] * z);
                vol[k] *= roi_in(s.D, s.p.roi_cut, s.p.roi_jtop, xx, y, z) ? R(1) : R(0);
            }
}

template <class R> static void pd_dual(OState<R>& s, int f, int warp_threads) {
    const size_t V = s.V;
    const int* D = s.D;
    std::vector<R> bw(V);
    // m_backward_warp = warp(dims, aabb, flow[f], X_bar[f+1], -dt) (ST_Tomography.cpp:503)
    warp<R>(D, s.flow[f].data(), s.xbar[f + 1].data(), R(-1), bw.data(), s.p.interp, warp_threads);
    R* ytp = s.ytp[f].data();
    R* ybc = s.ybc[f].data();
    const R* xb0 = s.xbar[f].data();
    const R* xb1 = s.xbar[f + 1].data();
    // Y_tp += sigma*(X_bar[f+1] - X_bar[f]); Y_bc += sigma*(bw - X_bar[f]) (:504-505)
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)V; ++k) {
        ytp[k] += s.sigma * (xb1[k] - xb0[k]);
        ybc[k] += s.sigma * (bw[k] - xb0[k]);
    }
    const R h = s.huberfactor();
    R* ytv = s.ytv[f].data();
    const size_t sy = D[0], sz = (size_t)D[0] * D[1];
#pragma omp parallel for schedule(static)
    for (int k = 1; k < D[2] - 1; ++k)
        for (int j = 1; j < D[1] - 1; ++j)
            for (int i = 1; i < D[0] - 1; ++i) {
                if (cut_out(D, s.p.roi_cut, i, j, k)) continue;
                size_t q = (size_t)i + sy * j + sz * k;
                R g0 = s.sigma * (xb0[q + 1] - xb0[q]) * h;
                R g1 = s.sigma * (xb0[q + sy] - xb0[q]) * h;
                R g2 = s.sigma * (xb0[q + sz] - xb0[q]) * h;
                R y0 = ytv[q] + g0, y1 = ytv[q + V] + g1, y2 = ytv[q + 2 * V] + g2;
                if (s.p.prior == STT_PRIOR_ATV) {  // per-component clamp (:523-525)
                    y0 = MAXM(-s.lam, MINM(s.lam, y0));
                    y1 = MAXM(-s.lam, MINM(s.lam, y1));
                    y2 = MAXM(-s.lam, MINM(s.lam, y2));
                } else if (s.p.prior == STT_PRIOR_TV) {  // isotropic TV: projection onto the lambda L2 ball
                    R nrm = std::sqrt(y0 * y0 + y1 * y1 + y2 * y2);
                    R sc = nrm > s.lam ? s.lam / nrm : R(1);
                    y0 *= sc; y1 *= sc; y2 *= sc;
                } else {  // STV: 4-vector with the forward temporal difference, lambda L2 ball
                    R y3 = ytv[q + 3 * V] + s.sigma * (xb1[q] - xb0[q]) * h;
                    R nrm = std::sqrt(y0 * y0 + y1 * y1 + y2 * y2 + y3 * y3);
                    R sc = nrm > s.lam ? s.lam / nrm : R(1);
                    y0 *= sc; y1 *= sc; y2 *= sc; y3 *= sc;
                    ytv[q + 3 * V] = y3;
                }
                ytv[q] = y0; ytv[q + V] = y1; ytv[q + 2 * V] = y2;
                ytp[q] = ytp[q] * s.ktp / (s.ktp + R(2) * s.sigma);   // (:528)
                ybc[q] = MAXM(-s.ebc, MINM(s.ebc, ybc[q]));          // (:530)
            }
}

template <class R> static void pd_primal(OState<R>& s, int f, int warp_threads) {
    const size_t V = s.V;
    const int* D = s.D;
    R* x = s.x[f].data();
    std::copy(x, x + V, s.xprev[f].begin());  // X_prev = x (:536)
    const R* ytp = s.ytp[f].data();
    const R* ybc = s.ybc[f].data();
    if (f == 0) {  // (:540-545)
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)V; ++k) {
            x[k] -= s.tau * (-ytp[k]);
            x[k] -= s.tau * (-ybc[k]);
        }
    } else {  // (:546-555)
        std::vector<R> fw(V);
        warp<R>(D, s.flow[f - 1].data(), s.ybc[f - 1].data(), R(+1), fw.data(), s.p.interp, warp_threads);
        const R* ytpm = s.ytp[f - 1].data();
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)V; ++k) {
            x[k] -= s.tau * (ytpm[k] - ytp[k]);
            x[k] -= s.tau * (fw[k] - ybc[k]);
        }
    }
    const R* ytv = s.ytv[f].data();
    // STV's temporal dual component of this frame and of the previous one
    // (adjoint of the forward difference x[f+1] - x[f]: y[f-1] - y[f], with
    // no y[-1])
    const R* yt = ytv + 3 * V;
    const R* ytm = f > 0 ? s.ytv[f - 1].data() + 3 * V : nullptr;
    const bool stv = s.p.prior == STT_PRIOR_STV;
    const size_t sy = D[0], sz = (size_t)D[0] * D[1];
#pragma omp parallel for schedule(static)
    for (int k = 1; k < D[2] - 1; ++k)
        for (int j = 1; j < D[1] - 1; ++j)
            for (int i = 1; i < D[0] - 1; ++i) {
                if (cut_out(D, s.p.roi_cut, i, j, k)) continue;
                size_t q = (size_t)i + sy * j + sz * k;
                // (:570-575)
                R y1 = -(ytv[q] - ytv[q - 1] + ytv[V + q] - ytv[V + q - sy] + ytv[2 * V + q] - ytv[2 * V + q - sz]);
                x[q] -= (s.tau * y1);
                if (stv) x[q] -= s.tau * ((ytm ? ytm[q] : R(0)) - yt[q]);
                x[q] = MAXM(MINM(s.gmax, x[q]), R(0));
            }
}

template <class R> static void pd_extrapolate(OState<R>& s, int f) {
    const int* D = s.D;
    R* xb = s.xbar[f].data();
    const R* x = s.x[f].data();
    const R* xp = s.xprev[f].data();
#pragma omp parallel for schedule(static)
    for (int z = 0; z < D[2]; ++z)
        for (int y = 0; y < D[1]; ++y)
            for (int xx = 0; xx < D[0]; ++xx) {
                if (cut_out(D, s.p.roi_cut, xx, y, z)) continue;
                size_t q = (size_t)xx + (size_t)D[0] * ((size_t)y + (size_t)D[1] * z);
                xb[q] = R(2) * x[q] - xp[q];  // (:598)
            }
}

template <class R> static void pd_init_state(OState<R>& s) {
    // first pass of the PD loop: X_bar <- x for every frame (ST_Tomography.cpp:494-500);
    // duals and X_prev start at zero (SetData, :339-350)
    for (int f = 0; f <= s.F; ++f) {
        s.xbar[f] = s.x[f];
        std::fill(s.ytv[f].begin(), s.ytv[f].end(), 0);
        std::fill(s.ytp[f].begin(), s.ytp[f].end(), 0);
        std::fill(s.ybc[f].begin(), s.ybc[f].end(), 0);
        std::fill(s.xprev[f].begin(), s.xprev[f].end(), 0);
    }
    s.pd_ready = true;
}

template <class R> static o_ctx_base* make_state(const stt_geometry* g, const stt_params* p, int F, const int* P) {
    auto* s = new OState<R>();
    s->g = *g;
    s->p = *p;
    s->F = F;
    for (int c = 0; c < 3; ++c) s->D[c] = g->vol[c];
    s->V = (size_t)g->vol[0] * g->vol[1] * g->vol[2];
    s->W = g->det[0];
    s->H = g->det[1];
    s->P.assign(P, P + F);
    s->quats.resize(F);
    s->views.resize(F);
    s->x.assign(F + 1, std::vector<R>(s->V, R(0)));
    s->xbar = s->x;
    s->xprev = s->x;
    s->ytp = s->x;
    s->ybc = s->x;
    s->ytv.assign(F + 1, std::vector<R>(4 * s->V, R(0)));  // 3 spatial planes + STV's temporal plane
    s->flow.assign(F, std::vector<R>(3 * s->V, R(0)));
    s->b.resize(F);
    for (int f = 0; f < F; ++f) s->b[f].assign((size_t)s->W * s->H * P[f], R(0));
    s->raylen.assign((size_t)s->W * s->H, R(0));
    s->raylen_frame.assign(F, std::vector<R>((size_t)s->W * s->H, R(0)));
    s->derive_params();
    return s;
}

// ---------------------------------------------------------------------------
extern "C" {

int o_view_matrices(const stt_geometry* g, const float* q, float* out) {
    View<float> v = make_view<float>(*g, q);
    std::memcpy(out, v.P2M.e, 64);
    std::memcpy(out + 16, v.M2V.e, 64);
    std::memcpy(out + 32, v.MV.e, 64);
    std::memcpy(out + 48, v.V2M.e, 64);
    std::memcpy(out + 64, v.MVV2M.e, 64);
    std::memcpy(out + 80, v.eye, 16);
    std::memcpy(out + 84, v.o, 16);
    out[88] = v.sd;
    return 0;
}

// Quaternion(ax,ay,az,angle) (Math/Quaternion.cpp:7-35) for axis +y with the
// angle of InitViews (ST_Tomography.cpp:899-911): theta*3.1415926/180 in double -> float.
void o_quat_y(float theta_deg, float* q) {
    float angle = float(double(theta_deg) * 3.1415926 / 180);
    float ax = 0.f, ay = 1.f, az = 0.f;
    float length = std::sqrt(ax * ax + ay * ay + az * az);
    float li = 1.0f / length;
    float rad = angle / 2;
    float w = std::cos(rad), sc = std::sin(rad);
    // * m_OldRotation (identity): Quaternion product (Math/Quaternion.h:76-86)
    float x = ax * li * sc, y = ay * li * sc, z = az * li * sc;
    q[3] = w * 1.f - x * 0.f - y * 0.f - z * 0.f;
    q[0] = w * 0.f + x * 1.f + y * 0.f - z * 0.f;
    q[1] = w * 0.f + y * 1.f + z * 0.f - x * 0.f;
    q[2] = w * 0.f + z * 1.f + x * 0.f - y * 0.f;
}

int o_fp_f32(const stt_geometry* g, const float* q, const float* vol, float* Ax, float* raylen, float* weight) {
    View<float> v = make_view<float>(*g, q);
    fp_view<float>(v, vol, Ax, raylen, weight, nullptr, nullptr, nullptr);
    return 0;
}
int o_fp_f64(const stt_geometry* g, const float* q, const double* vol, double* Ax, double* raylen, double* weight) {
    View<double> v = make_view<double>(*g, q);
    fp_view<double>(v, vol, Ax, raylen, weight, nullptr, nullptr, nullptr);
    return 0;
}
int o_bp_ray_scatter_f32(const stt_geometry* g, const float* q, const float* corr, float* contrib, float* wsum) {
    View<float> v = make_view<float>(*g, q);
    fp_view<float>(v, contrib /*unused in BP mode*/, nullptr, nullptr, nullptr, corr, contrib, wsum);
    return 0;
}
int o_bp_ray_scatter_f64(const stt_geometry* g, const float* q, const double* corr, double* contrib, double* wsum) {
    View<double> v = make_view<double>(*g, q);
    fp_view<double>(v, contrib, nullptr, nullptr, nullptr, corr, contrib, wsum);
    return 0;
}
int o_bp_ray_apply_f32(int n, const float* contrib, const float* wsum, float gmax, float* vol) {
    for (int k = 0; k < n; ++k) {
        float m = float((1.0 * contrib[k]) / wsum[k]);
        float xv = vol[k] + m;
        vol[k] = MAXM(MINM(gmax, xv), 0.0f);
    }
    return 0;
}
int o_bp_voxel_f32(const stt_geometry* g, const float* q, const float* corr, double alpha, float gmax, float* vol) {
    View<float> v = make_view<float>(*g, q);
    bp_voxel<float>(v, corr, alpha, gmax, vol);
    return 0;
}
int o_bp_voxel_f64(const stt_geometry* g, const float* q, const double* corr, double alpha, double gmax, double* vol) {
    View<double> v = make_view<double>(*g, q);
    bp_voxel<double>(v, corr, alpha, gmax, vol);
    return 0;
}
int o_correction_f32(int n, float factor, const float* b, const float* Ax, const float* raylen, float* corr) {
    for (int k = 0; k < n; ++k) {
        float num = factor * b[k] - factor * Ax[k] - 0.0f;
        float den = float(factor * raylen[k] + 1.0);
        corr[k] = num / den;
    }
    return 0;
}
float o_cubic_atxyz_f32(const float* img, int w, int h, int d, float fx, float fy, float fz) {
    return cubic_atxyz<float>(img, w, h, d, fx, fy, fz);
}
int o_warp_ext_f32(const int* dims, const float* ext, const float* uvw, const float* src, float dt, float* out) {
    warp<float>(dims, uvw, src, dt, out, STT_INTERP_CUBIC_CLAMPED, 0, ext);
    return 0;
}

// ---------------------------------------------------------------------------
// 3-D L1/Huber-TV optical flow, one pyramid level (optical_flow.h:83-623),
// fp32 in the reference's expression order.  Two choices of this oracle:
// the convergence sums are accumulated in double in a fixed order (the
// reference adds them in float across OpenMP threads, one of them without a
// reduction clause, :325-333), and the gradient planes outside the ROI are
// zero (the reference never writes them; fresh allocations read zero).

static inline size_t of_ix(const int* d, int i, int j, int k) { return (size_t)i + (size_t)d[0] * ((size_t)j + (size_t)d[1] * k); }
// tools.cpp:136-147: forward differences, 0 on the last index
static inline float of_dxp(const float* a, const int* d, int i, int j, int k) {
    return i < d[0] - 1 ? a[of_ix(d, i + 1, j, k)] - a[of_ix(d, i, j, k)] : 0.0f;
}
static inline float of_dyp(const float* a, const int* d, int i, int j, int k) {
    return j < d[1] - 1 ? a[of_ix(d, i, j + 1, k)] - a[of_ix(d, i, j, k)] : 0.0f;
}
static inline float of_dzp(const float* a, const int* d, int i, int j, int k) {
    return k < d[2] - 1 ? a[of_ix(d, i, j, k + 1)] - a[of_ix(d, i, j, k)] : 0.0f;
}
// tools.cpp:153-200: backward differences; a at the first index, -a[n-2] at the last
static inline float of_dxm(const float* a, const int* d, int i, int j, int k) {
    if (i > 0 && i < d[0] - 1) return a[of_ix(d, i, j, k)] - a[of_ix(d, i - 1, j, k)];
    if (i == 0) return a[of_ix(d, i, j, k)];
    return -a[of_ix(d, i - 1, j, k)];
}
static inline float of_dym(const float* a, const int* d, int i, int j, int k) {
    if (j > 0 && j < d[1] - 1) return a[of_ix(d, i, j, k)] - a[of_ix(d, i, j - 1, k)];
    if (j == 0) return a[of_ix(d, i, j, k)];
    return -a[of_ix(d, i, j - 1, k)];
}
static inline float of_dzm(const float* a, const int* d, int i, int j, int k) {
    if (k > 0 && k < d[2] - 1) return a[of_ix(d, i, j, k)] - a[of_ix(d, i, j, k - 1)];
    if (k == 0) return a[of_ix(d, i, j, k)];
    return -a[of_ix(d, i, j, k - 1)];
}
static inline float of_min(float a, float b) { return b < a ? b : a; }  // std::min
static inline float of_max(float a, float b) { return a < b ? b : a; }  // std::max

int o_of_diff(const int* d, const float* a, int which, float* out) {
    const int n = d[0] * d[1] * d[2];
    (void)n;
#pragma omp parallel for
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                float v;
                switch (which) {
                    case 0: v = of_dxp(a, d, i, j, k); break;
                    case 1: v = of_dyp(a, d, i, j, k); break;
                    case 2: v = of_dzp(a, d, i, j, k); break;
                    case 3: v = of_dxm(a, d, i, j, k); break;
                    case 4: v = of_dym(a, d, i, j, k); break;
                    default: v = of_dzm(a, d, i, j, k); break;
                }
                out[of_ix(d, i, j, k)] = v;
            }
    return 0;
}

// CImg get_gradient("xyz", 0) (CImg.h:34570-34580): (I[+1] - I[-1]) / 2 with
// neighbours clamped at the borders (cimg_for3x3x3).
int o_gradient_central(const int* d, const float* I, float* gx, float* gy, float* gz) {
#pragma omp parallel for
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const int ip = std::min(i + 1, d[0] - 1), im = std::max(i - 1, 0);
                const int jp = std::min(j + 1, d[1] - 1), jm = std::max(j - 1, 0);
                const int kp = std::min(k + 1, d[2] - 1), km = std::max(k - 1, 0);
                const size_t q = of_ix(d, i, j, k);
                gx[q] = (I[of_ix(d, ip, j, k)] - I[of_ix(d, im, j, k)]) / 2;
                gy[q] = (I[of_ix(d, i, jp, k)] - I[of_ix(d, i, jm, k)]) / 2;
                gz[q] = (I[of_ix(d, i, j, kp)] - I[of_ix(d, i, j, km)]) / 2;
            }
    return 0;
}

// L1Huber_OF (optical_flow.h:83-623).  V: 3 planes (u, v, w), updated in the
// ROI; Y: 10 planes (y11 y12 y13 y21 y22 y23 y31 y32 y33 y4), written back in
// the ROI (:503-531).  roi = {x0, x1, y0, y1, z0, z1} margins (the global ROI[]).
// iters_out[w] = iterations of warp w; err_out[w] = its last residual.
int o_of_level(const int* d, const float* ext, const float* I1, const float* I2, float* V, float* Y, float tol,
               float lambda, int max_iter, int norm, int warps, float huber, const int* roi, int* iters_out,
               double* err_out) {
    const size_t n = (size_t)d[0] * d[1] * d[2];
    std::vector<float> v(3 * n), y(10 * n), Kx(10 * n, 0.0f), Kty(3 * n, 0.0f), KtyOld(3 * n, 0.0f), vOld(3 * n, 0.0f);
    std::vector<float> ux(n, 0.0f), uy(n, 0.0f), uz(n, 0.0f), ut(n, 0.0f), tau(3 * n), sig2(n), W(n), g0(n), g1(n), g2(n);
    std::copy(V, V + 3 * n, v.begin());
    std::copy(Y, Y + 10 * n, y.begin());
    const float stepsize = 1.0f;
    const float sigma1 = of_min(stepsize / 3.0f, of_min(stepsize / 3.0f, stepsize / 3.0f));  // :170
    const float hf = 1.0f / (1.0f + sigma1 * huber / lambda);                                 // :183
    auto in_roi = [&](int i, int j, int k) {
        return i >= roi[0] && i < d[0] - roi[1] && j >= roi[2] && j < d[1] - roi[3] && k >= roi[4] && k < d[2] - roi[5];
    };
    for (int w = 0; w < warps; ++w) {
        // _warpI2 = warp(dim, aabb, V, image2, -dt) and its central gradient (:245-247)
        warp<float>(d, v.data(), I2, -1.0f, W.data(), STT_INTERP_CUBIC_CLAMPED, 0, ext);
        o_gradient_central(d, W.data(), g0.data(), g1.data(), g2.data());
#pragma omp parallel for
        for (int k = 0; k < d[2]; ++k)
            for (int j = 0; j < d[1]; ++j)
                for (int i = 0; i < d[0]; ++i) {
                    if (!in_roi(i, j, k)) continue;
                    const size_t q = of_ix(d, i, j, k);
                    ux[q] = g0[q];
                    uy[q] = g1[q];
                    uz[q] = g2[q];
                    ut[q] = W[q] - I1[q] - ux[q] * v[q] - uy[q] * v[n + q] - uz[q] * v[2 * n + q];  // :264
                }
        for (size_t q = 0; q < n; ++q) {  // :273-287
            tau[q] = 4.0f / stepsize + std::fabs(ux[q]);
            tau[n + q] = 4.0f / stepsize + std::fabs(uy[q]);
            tau[2 * n + q] = 4.0f / stepsize + std::fabs(uz[q]);
            sig2[q] = std::fabs(ux[q]) + std::fabs(uy[q]) + std::fabs(uz[q]);
            tau[q] = 1.0f / tau[q];
            tau[n + q] = 1.0f / tau[n + q];
            tau[2 * n + q] = 1.0f / tau[2 * n + q];
            sig2[q] = 1.0f / sig2[q];
        }
        int it = 0;
        double err = 1.0;
        while (err > tol && it <= max_iter) {
            ++it;
            const bool res = it % 50 == 0;
            double pacc = 0.0, dacc = 0.0;
            // primal: Kty = -div y + grad(I) y4, v -= tau Kty, on the ROI (:300-340)
#pragma omp parallel for reduction(+ : pacc)
            for (int k = 0; k < d[2]; ++k)
                for (int j = 0; j < d[1]; ++j)
                    for (int i = 0; i < d[0]; ++i) {
                        if (!in_roi(i, j, k)) continue;
                        const size_t q = of_ix(d, i, j, k);
                        const float u3[3] = {ux[q], uy[q], uz[q]};
                        float t3[3];
                        for (int c = 0; c < 3; ++c) {
                            if (res) {
                                vOld[c * n + q] = v[c * n + q];
                                KtyOld[c * n + q] = Kty[c * n + q];
                            }
                            const float* yc = y.data() + 3 * c * n;
                            const float kt = -stepsize * of_dxm(yc, d, i, j, k) - stepsize * of_dym(yc + n, d, i, j, k) -
                                             stepsize * of_dzm(yc + 2 * n, d, i, j, k) + u3[c] * y[9 * n + q];
                            Kty[c * n + q] = kt;
                            v[c * n + q] -= tau[c * n + q] * kt;
                            if (res)
                                t3[c] = std::fabs((vOld[c * n + q] - v[c * n + q]) / tau[c * n + q] - KtyOld[c * n + q] + kt);
                        }
                        if (res) pacc += (double)(t3[0] + t3[1] + t3[2]);
                    }
            // dual on every voxel (:343-480)
#pragma omp parallel for reduction(+ : dacc)
            for (int k = 0; k < d[2]; ++k)
                for (int j = 0; j < d[1]; ++j)
                    for (int i = 0; i < d[0]; ++i) {
                        const size_t q = of_ix(d, i, j, k);
                        float yo[10], ko[10], kn[10];
                        for (int m = 0; m < 10; ++m) {
                            yo[m] = y[m * n + q];
                            ko[m] = Kx[m * n + q];
                        }
                        for (int c = 0; c < 3; ++c) {
                            const float* vc = v.data() + c * n;
                            kn[3 * c] = stepsize * of_dxp(vc, d, i, j, k);
                            kn[3 * c + 1] = stepsize * of_dyp(vc, d, i, j, k);
                            kn[3 * c + 2] = stepsize * of_dzp(vc, d, i, j, k);
                        }
                        kn[9] = ux[q] * v[q] + uy[q] * v[n + q] + uz[q] * v[2 * n + q];
                        for (int c = 0; c < 3; ++c) {
                            float t[3];
                            for (int e = 0; e < 3; ++e) {
                                const int m = 3 * c + e;
                                t[e] = yo[m] + sigma1 * (kn[m] + kn[m] - ko[m]);
                                if (norm == 4) t[e] = t[e] * hf;
                            }
                            const float div = of_max(1.0f, std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) / lambda);
                            for (int e = 0; e < 3; ++e) y[(3 * c + e) * n + q] = t[e] / div;
                        }
                        y[9 * n + q] = of_max(-1.0f, of_min(1.0f, yo[9] + sig2[q] * (kn[9] + kn[9] - ko[9] + ut[q])));
                        for (int m = 0; m < 10; ++m) Kx[m * n + q] = kn[m];
                        if (res) {
                            float acc = 0.0f;
                            for (int m = 0; m < 9; ++m) acc += std::fabs((yo[m] - y[m * n + q]) / sigma1 - ko[m] + kn[m]);
                            acc += std::fabs((yo[9] - y[9 * n + q]) / sig2[q] - ko[9] + kn[9]);
                            dacc += (double)acc;
                        }
                    }
            if (res) err = (dacc * dacc + pacc * pacc) / (double)n;
        }
        iters_out[w] = it;
        err_out[w] = err;
    }
    // write back on the ROI (:503-531)
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                if (!in_roi(i, j, k)) continue;
                const size_t q = of_ix(d, i, j, k);
                for (int m = 0; m < 10; ++m) Y[m * n + q] = y[m * n + q];
                for (int c = 0; c < 3; ++c) V[c * n + q] = v[c * n + q];
            }
    return 0;
}

// ---- the flow pyramid's image operators (CImg 1.8.0) ----------------------
// Deriche coefficients of deriche(sigma, 0, axis) (CImg.h:32917-32968).
struct DericheCoef {
    float a0, a1, a2, a3, b1, b2, coefp, coefn;
};
static DericheCoef deriche_coef(float sigma) {
    const float nsigma = sigma >= 0 ? sigma : -sigma;  // (percent form not used)
    const float nnsigma = nsigma < 0.1f ? 0.1f : nsigma, alpha = 1.695f / nnsigma, ema = (float)std::exp(-alpha),
                ema2 = (float)std::exp(-2 * alpha), b1 = -2 * ema, b2 = ema2;
    const float k = (1 - ema) * (1 - ema) / (1 + 2 * alpha * ema - ema2);
    DericheCoef c;
    c.a0 = k;
    c.a1 = k * (alpha - 1) * ema;
    c.a2 = k * (alpha + 1) * ema;
    c.a3 = -k * ema2;
    c.b1 = b1;
    c.b2 = b2;
    c.coefp = (c.a0 + c.a1) / (1 + b1 + b2);
    c.coefn = (c.a2 + c.a3) / (1 + b1 + b2);
    return c;
}
// One line of _cimg_deriche_apply (Neumann boundary), in place.
static void deriche_line(float* x, long N, long off, const DericheCoef& c, std::vector<float>& Y) {
    Y.resize(N);
    float* px = x;
    float yb = 0, yp = 0, xp = *px;
    yb = yp = c.coefp * xp;
    for (long m = 0; m < N; ++m) {
        const float xc = *px;
        px += off;
        const float yc = Y[m] = c.a0 * xc + c.a1 * xp - c.b1 * yp - c.b2 * yb;
        xp = xc;
        yb = yp;
        yp = yc;
    }
    float xn = *(px - off), xa = xn, yn = c.coefn * xn, ya = yn;
    for (long n = N - 1; n >= 0; --n) {
        const float xc = *(px -= off);
        const float yc = c.a2 * xn + c.a3 * xa - c.b1 * yn - c.b2 * ya;
        xa = xn;
        xn = xc;
        ya = yn;
        yn = yc;
        *px = Y[n] + yc;
    }
}
// get_blur(sigma) (CImg.h:33249-33284): order-0 Deriche along x, y, z.
int o_blur(const int* d, int spectrum, const float* img, float sigma, float* out) {
    const size_t n = (size_t)d[0] * d[1] * d[2];
    std::copy(img, img + n * spectrum, out);
    if (sigma < 0.1f) return 0;
    const DericheCoef c = deriche_coef(sigma);
    for (int ch = 0; ch < spectrum; ++ch) {
        float* im = out + ch * n;
        std::vector<float> Y;
        if (d[0] > 1)
            for (int z = 0; z < d[2]; ++z)
                for (int y = 0; y < d[1]; ++y) deriche_line(im + of_ix(d, 0, y, z), d[0], 1, c, Y);
        if (d[1] > 1)
            for (int z = 0; z < d[2]; ++z)
                for (int x = 0; x < d[0]; ++x) deriche_line(im + of_ix(d, x, 0, z), d[1], d[0], c, Y);
        if (d[2] > 1)
            for (int y = 0; y < d[1]; ++y)
                for (int x = 0; x < d[0]; ++x) deriche_line(im + of_ix(d, x, y, 0), d[2], (long)d[0] * d[1], c, Y);
    }
    return 0;
}

// One axis of get_resize(..., interpolation 5), boundary 0 (CImg.h:28432-28520):
// moving average (type 2, CImg.h:28135-28150) when shrinking, cubic when growing.
static void resize_axis(const std::vector<float>& src, const int* d, int axis, int ns, std::vector<float>& dst,
                        int* dn) {
    for (int a = 0; a < 3; ++a) dn[a] = d[a];
    dn[axis] = ns;
    const int w = d[axis];
    const size_t st_s = axis == 0 ? 1 : (axis == 1 ? (size_t)d[0] : (size_t)d[0] * d[1]);
    const size_t st_d = axis == 0 ? 1 : (axis == 1 ? (size_t)dn[0] : (size_t)dn[0] * dn[1]);
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;  // the two other axes
    dst.assign((size_t)dn[0] * dn[1] * dn[2], 0.0f);
    auto base_s = [&](int u, int v) {
        int c[3] = {0, 0, 0};
        c[o1] = u;
        c[o2] = v;
        return of_ix(d, c[0], c[1], c[2]);
    };
    auto base_d = [&](int u, int v) {
        int c[3] = {0, 0, 0};
        c[o1] = u;
        c[o2] = v;
        return of_ix(dn, c[0], c[1], c[2]);
    };
    if (w > ns) {  // moving average
        for (int v = 0; v < d[o2]; ++v)
            for (int u = 0; u < d[o1]; ++u) {
                const float* ps = src.data() + base_s(u, v);
                float* pd = dst.data() + base_d(u, v);
                unsigned a = (unsigned)w * ns, b = w, c = ns, s = 0, t = 0;
                while (a) {
                    const unsigned dd = std::min(b, c);
                    a -= dd;
                    b -= dd;
                    c -= dd;
                    pd[t * st_d] += ps[s * st_s] * dd;
                    if (!b) {
                        pd[t * st_d] /= w;
                        ++t;
                        b = w;
                    }
                    if (!c) {
                        ++s;
                        c = ns;
                    }
                }
            }
        return;
    }
    // cubic: source index and fraction per output from CImg's double recurrence
    const double f = ns > 1 ? (w - 1.0) / (ns - 1) : 0;
    std::vector<unsigned> sidx(ns);
    std::vector<double> frac(ns);
    double curr = 0;
    for (int x = 0; x < ns; ++x) {
        sidx[x] = (unsigned)curr;
        frac[x] = curr - (unsigned)curr;
        curr = std::min(w - 1.0, curr + f);
    }
    for (int v = 0; v < d[o2]; ++v)
        for (int u = 0; u < d[o1]; ++u) {
            const float* ps = src.data() + base_s(u, v);
            float* pd = dst.data() + base_d(u, v);
            for (int x = 0; x < ns; ++x) {
                const long si = sidx[x];
                const double t = frac[x], val1 = ps[si * st_s], val0 = si > 0 ? (double)ps[(si - 1) * st_s] : val1,
                             val2 = si <= w - 2 ? (double)ps[(si + 1) * st_s] : val1,
                             val3 = si < w - 2 ? (double)ps[(si + 2) * st_s] : val2,
                             val = val1 + 0.5f * (t * (-val0 + val2) + t * t * (2 * val0 - 5 * val1 + 4 * val2 - val3) +
                                                  t * t * t * (-val0 + 3 * val1 - 3 * val2 + val3));
                pd[x * st_d] = (float)val;
            }
        }
}
int o_resize(const int* d, int spectrum, const float* img, const int* nd, float* out) {
    const size_t n = (size_t)d[0] * d[1] * d[2], nn = (size_t)nd[0] * nd[1] * nd[2];
    for (int ch = 0; ch < spectrum; ++ch) {
        std::vector<float> cur(img + ch * n, img + (ch + 1) * n), nxt;
        int cd[3] = {d[0], d[1], d[2]}, ndim[3];
        for (int a = 0; a < 3; ++a) {
            if (nd[a] == cd[a]) continue;
            resize_axis(cur, cd, a, nd[a], nxt, ndim);
            cur.swap(nxt);
            for (int k = 0; k < 3; ++k) cd[k] = ndim[k];
        }
        std::copy(cur.begin(), cur.end(), out + ch * nn);
    }
    return 0;
}

// L1Huber_OF_MultiScale3D (optical_flow.h:630-734) with Y1..Y4 = 0 as Render
// passes them (ST_Tomography.cpp:452-455).  V: 3 planes in/out at full size.
// level_dims (levels x 3) and iters (levels x warps) may be null.
int o_of_multiscale(const int* D, const float* I1, const float* I2, float* V, float tol, float lambda, int max_iter,
                    int norm, int warps, float eta, float sigma, int levels, float huber, const int* cut,
                    int* level_dims, int* iters) {
    const size_t N = (size_t)D[0] * D[1] * D[2];
    std::vector<float> b1(N), b2(N);
    o_blur(D, 1, I1, sigma, b1.data());
    o_blur(D, 1, I2, sigma, b2.data());
    std::vector<std::array<int, 3>> dims(levels);
    std::vector<std::vector<float>> im1(levels), im2(levels), vs(levels);
    float t_eta = 1.0f;
    for (int s = 0; s < levels; ++s) {  // finest first (:660-676)
        std::array<int, 3> nd;
        for (int a = 0; a < 3; ++a) nd[a] = (int)((float)D[a] * t_eta);
        dims[s] = nd;
        const size_t m = (size_t)nd[0] * nd[1] * nd[2];
        im1[s].resize(m);
        im2[s].resize(m);
        vs[s].resize(3 * m);
        o_resize(D, 1, b1.data(), nd.data(), im1[s].data());
        o_resize(D, 1, b2.data(), nd.data(), im2[s].data());
        o_resize(D, 3, V, nd.data(), vs[s].data());
        t_eta = t_eta * eta;
    }
    const float ext[3] = {(float)D[0], (float)D[1], (float)D[2]};
    std::vector<float> Y;  // Y1..Y4 = 0, resized zeros stay zero
    for (int k = 0; k < levels; ++k) {  // coarsest first (:690-729)
        const int s = levels - 1 - k;
        const int* nd = dims[s].data();
        const size_t m = (size_t)nd[0] * nd[1] * nd[2];
        if (k == 0) Y.assign(10 * m, 0.0f);
        int roi[6];
        for (int a = 0; a < 6; ++a) roi[a] = (int)((float)cut[a] * ((float)nd[a / 2] / (float)D[a / 2]));  // :700-705
        std::vector<double> errs(warps);
        o_of_level(nd, ext, im1[s].data(), im2[s].data(), vs[s].data(), Y.data(), tol, lambda, max_iter, norm, warps,
                   huber, roi, iters ? iters + k * warps : std::vector<int>(warps).data(), errs.data());
        if (level_dims)
            for (int a = 0; a < 3; ++a) level_dims[3 * k + a] = nd[a];
        if (k + 1 < levels) {  // :712-722
            const int* nn = dims[s - 1].data();
            const size_t mn = (size_t)nn[0] * nn[1] * nn[2];
            std::vector<float> Yn(10 * mn);
            o_resize(nd, 3, vs[s].data(), nn, vs[s - 1].data());
            o_resize(nd, 10, Y.data(), nn, Yn.data());
            Y.swap(Yn);
        }
    }
    std::copy(vs[0].begin(), vs[0].end(), V);
    return 0;
}

int o_warp_f32(const int* dims, const float* uvw, const float* src, float dt, float* out, int interp, int nthreads) {
    warp<float>(dims, uvw, src, dt, out, interp, nthreads);
    return 0;
}
int o_warp_f64(const int* dims, const double* uvw, const double* src, double dt, double* out, int interp, int nthreads) {
    warp<double>(dims, uvw, src, dt, out, interp, nthreads);
    return 0;
}
int o_roi_mask(const int* dims, const int* cut, int jtop, float* out) {
    for (int z = 0; z < dims[2]; ++z)
        for (int y = 0; y < dims[1]; ++y)
            for (int x = 0; x < dims[0]; ++x)
                out[(size_t)x + (size_t)dims[0] * ((size_t)y + (size_t)dims[1] * z)] = roi_in(dims, cut, jtop, x, y, z) ? 1.f : 0.f;
    return 0;
}

o_ctx* o_ctx_create(const stt_geometry* g, const stt_params* p, int nframes, const int* P, int f64) {
    if (!g || !p || nframes < 1 || !P) return nullptr;
    o_ctx* c = new o_ctx;
    c->f64 = f64;
    c->st = f64 ? make_state<double>(g, p, nframes, P) : make_state<float>(g, p, nframes, P);
    return c;
}
void o_ctx_destroy(o_ctx* c) {
    if (!c) return;
    delete c->st;
    delete c;
}

#define DISPATCH(c, expr)                                          \
    do {                                                           \
        if ((c)->f64) { auto& s = *static_cast<OState<double>*>((c)->st); expr; } \
        else { auto& s = *static_cast<OState<float>*>((c)->st); expr; }           \
    } while (0)

int o_ctx_set_view_quaternions(o_ctx* c, int f, const float* q, int n) {
    int rc = 0;
    DISPATCH(c, {
        if (f < 0 || f >= s.F || n != s.P[f]) { rc = 1; break; }
        s.quats[f].assign(q, q + 4 * n);
        s.views[f].clear();
        using RT = std::remove_reference_t<decltype(s.gmax)>;
        for (int i = 0; i < n; ++i) s.views[f].push_back(make_view<RT>(s.g, q + 4 * i));
    });
    return rc;
}
int o_ctx_set_volume(o_ctx* c, int f, const float* vol) {
    int rc = 0;
    DISPATCH(c, {
        if (f < 0 || f > s.F) { rc = 1; break; }
        for (size_t k = 0; k < s.V; ++k) s.x[f][k] = vol[k];
    });
    return rc;
}
int o_ctx_get_volume(o_ctx* c, int f, float* vol) {
    int rc = 0;
    DISPATCH(c, {
        if (f < 0 || f > s.F) { rc = 1; break; }
        for (size_t k = 0; k < s.V; ++k) vol[k] = float(s.x[f][k]);
    });
    return rc;
}
int o_ctx_get_volume_f64(o_ctx* c, int f, double* vol) {
    int rc = 0;
    DISPATCH(c, {
        if (f < 0 || f > s.F) { rc = 1; break; }
        for (size_t k = 0; k < s.V; ++k) vol[k] = double(s.x[f][k]);
    });
    return rc;
}
int o_ctx_set_flow(o_ctx* c, int pair, const float* uvw) {
    int rc = 0;
    DISPATCH(c, {
        if (pair < 0 || pair >= s.F) { rc = 1; break; }
        for (size_t k = 0; k < 3 * s.V; ++k) s.flow[pair][k] = uvw[k];
    });
    return rc;
}
int o_ctx_set_projections(o_ctx* c, int f, const float* b) {
    int rc = 0;
    DISPATCH(c, {
        if (f < 0 || f >= s.F) { rc = 1; break; }
        for (size_t k = 0; k < s.b[f].size(); ++k) s.b[f][k] = b[k];
        if (f == 0 && !s.gmax_set) {  // G_Max = max(frame-0 projections)/4 (ST_Tomography.cpp:356-357)
            float mx = -3.4e38f;
            for (size_t k = 0; k < s.b[0].size(); ++k) mx = std::max(mx, float(b[k]));
            s.gmax = decltype(s.gmax)(mx / 4.0f);
        }
    });
    return rc;
}
int o_ctx_get_buffer(o_ctx* c, int which, int f, float* out) {
    int rc = 0;
    DISPATCH(c, {
        if (f < 0 || f > s.F) { rc = 1; break; }
        const auto* src = &s.xbar[f];
        switch (which) {
            case STT_BUF_XBAR: src = &s.xbar[f]; break;
            case STT_BUF_XPREV: src = &s.xprev[f]; break;
            case STT_BUF_YTV: src = &s.ytv[f]; break;
            case STT_BUF_YTP: src = &s.ytp[f]; break;
            case STT_BUF_YBC: src = &s.ybc[f]; break;
            default: rc = 1;
        }
        if (rc) break;
        for (size_t k = 0; k < src->size(); ++k) out[k] = float((*src)[k]);
    });
    return rc;
}
int o_ctx_set_buffer(o_ctx* c, int which, int f, const float* in) {
    int rc = 0;
    DISPATCH(c, {
        if (f < 0 || f > s.F) { rc = 1; break; }
        auto* dst = &s.xbar[f];
        switch (which) {
            case STT_BUF_XBAR: dst = &s.xbar[f]; break;
            case STT_BUF_XPREV: dst = &s.xprev[f]; break;
            case STT_BUF_YTV: dst = &s.ytv[f]; break;
            case STT_BUF_YTP: dst = &s.ytp[f]; break;
            case STT_BUF_YBC: dst = &s.ybc[f]; break;
            default: rc = 1;
        }
        if (rc) break;
        for (size_t k = 0; k < dst->size(); ++k) (*dst)[k] = in[k];
    });
    return rc;
}
float o_ctx_gmax(o_ctx* c) {
    float r = 0;
    DISPATCH(c, { r = float(s.gmax); });
    return r;
}
int o_ctx_get_raylen(o_ctx* c, float* out) {
    DISPATCH(c, {
        for (size_t k = 0; k < s.raylen.size(); ++k) out[k] = float(s.raylen[k]);
    });
    return 0;
}
void o_ctx_set_per_frame_raylen(o_ctx* c, int on) {
    DISPATCH(c, { s.per_frame_raylen = on != 0; });
}
int o_ctx_sart(o_ctx* c, int f, const int* order, int n) {
    int rc = 0;
    DISPATCH(c, {
        if (f < 0 || f >= s.F || n > s.P[f] || (int)s.views[f].size() != s.P[f]) { rc = 1; break; }
        sart_frame(s, f, order, n, false);
    });
    return rc;
}
int o_ctx_pd_init(o_ctx* c) {
    DISPATCH(c, { pd_init_state(s); });
    return 0;
}
int o_ctx_pd_dual(o_ctx* c, int f) {
    DISPATCH(c, { pd_dual(s, f, 0); });
    return 0;
}
int o_ctx_pd_primal(o_ctx* c, int f) {
    DISPATCH(c, { pd_primal(s, f, 0); });
    return 0;
}
int o_ctx_pd_extrapolate(o_ctx* c, int f) {
    DISPATCH(c, { pd_extrapolate(s, f); });
    return 0;
}
// One pass of `for (No_PO ...)` body (ST_Tomography.cpp:492-607), n times.
int o_ctx_iterate(o_ctx* c, int n, int faithful) {
    int rc = 0;
    DISPATCH(c, {
        for (int f = 0; f < s.F; ++f)
            if ((int)s.views[f].size() != s.P[f]) rc = 1;
        if (rc) break;
        const int wt = faithful ? 4 : 0;
        if (!s.pd_ready) pd_init_state(s);
        for (int it = 0; it < n; ++it)
            for (int f = 0; f < s.F; ++f) {
                pd_dual(s, f, wt);
                pd_primal(s, f, wt);
                sart_frame(s, f, nullptr, s.P[f], faithful != 0);
                pd_extrapolate(s, f);
            }
    });
    return rc;
}

// Timed sample of one frame's share of an ST iteration for the CPU baseline:
// PD dual + primal of `frame` (two warps when frame > 0), then a SART sweep
// over its first nviews views.  Seconds via steady_clock.
int o_ctx_time_sample(o_ctx* c, int frame, int nviews, int faithful, double* t_pd, double* t_views) {
    int rc = 0;
    DISPATCH(c, {
        if (frame < 0 || frame >= s.F || nviews < 1 || nviews > s.P[frame] ||
            (int)s.views[frame].size() != s.P[frame]) { rc = 1; break; }
        const int wt = faithful ? 4 : 0;
        auto t0 = std::chrono::steady_clock::now();
        pd_dual(s, frame, wt);
        pd_primal(s, frame, wt);
        auto t1 = std::chrono::steady_clock::now();
        sart_frame(s, frame, nullptr, nviews, faithful != 0);
        auto t2 = std::chrono::steady_clock::now();
        *t_pd = std::chrono::duration<double>(t1 - t0).count();
        *t_views = std::chrono::duration<double>(t2 - t1).count();
    });
    return rc;
}

}  // extern "C"
