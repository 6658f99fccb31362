// stt_geometry.cpp — per-view cone-beam geometry on the host.
//
// Builds the 256-byte ViewRec the kernels read.  The reference assembles the
// same quantities with its Matrix4x4 / Quaternion classes every view
// (Scene::_calculateMatrix, Scene.cpp:181-207; TypicalRayProducer::
// _calculateMatrix, TypicalRayProducer.cpp:142-208; EyePos,
// TypicalRayProducer.core.h:255-258).  The ray parameters decide which
// samples a ray takes, so this file reproduces the reference's fp32
// arithmetic operation for operation (float, no FMA contraction: compiled with
// -ffp-contract=off) and is checked bit-for-bit against the reference's own
// matrix code in the tests.
#include <cmath>
#include <cstring>

#include "stt_device.hpp"
#include "stt_host.hpp"

namespace stt {
namespace {

// 4x4 column-major fp32 matrix; element (row, col) at e[4*col + row].
struct Mat4 {
    float e[16];
    static Mat4 zero() {
        Mat4 m;
        std::memset(m.e, 0, sizeof(m.e));
        return m;
    }
    static Mat4 identity() {
        Mat4 m = zero();
        m.e[0] = m.e[5] = m.e[10] = m.e[15] = 1.0f;
        return m;
    }
    static Mat4 translation(float tx, float ty, float tz) {
        Mat4 m = identity();
        m.e[12] = tx;
        m.e[13] = ty;
        m.e[14] = tz;
        return m;
    }
    float at(int r, int c) const { return e[4 * c + r]; }
};

// Product with the reference's summation order a0*b0 + a1*b1 + a2*b2 + a3*b3
// (Math/Matrix4x4.h:66-91).
Mat4 operator*(const Mat4& A, const Mat4& B) {
    Mat4 R;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            R.e[4 * c + r] = A.at(r, 0) * B.at(0, c) + A.at(r, 1) * B.at(1, c) + A.at(r, 2) * B.at(2, c) +
                             A.at(r, 3) * B.at(3, c);
    return R;
}

void apply(const Mat4& M, const float v[4], float out[4]) {
    float t[4];
    for (int r = 0; r < 4; ++r) t[r] = M.at(r, 0) * v[0] + M.at(r, 1) * v[1] + M.at(r, 2) * v[2] + M.at(r, 3) * v[3];
    for (int r = 0; r < 4; ++r) out[r] = t[r];
}

// 3x3 minor a*e*i + b*f*g + c*d*h - c*e*g - a*f*h - b*d*i, evaluated left to
// right as the reference's Det3x3 macro (Math/Matrix4x4.cpp:79).
inline float minor3(float a, float b, float c, float d, float e, float f, float g, float h, float i) {
    return a * e * i + b * f * g + c * d * h - c * e * g - a * f * h - b * d * i;
}

// Adjugate / determinant inverse in fp32 (Math/Matrix4x4.cpp:80-135); a
// singular matrix is returned unchanged, as there.
Mat4 inverted(const Mat4& m) {
    const float* s = m.e;
    // columns of the stored matrix as the reference names them (a = col 0 ...)
    const float a1 = s[0], b1 = s[1], c1 = s[2], d1 = s[3];
    const float a2 = s[4], b2 = s[5], c2 = s[6], d2 = s[7];
    const float a3 = s[8], b3 = s[9], c3 = s[10], d3 = s[11];
    const float a4 = s[12], b4 = s[13], c4 = s[14], d4 = s[15];
    const float m1 = minor3(b2, b3, b4, c2, c3, c4, d2, d3, d4);
    const float m2 = minor3(a2, a3, a4, c2, c3, c4, d2, d3, d4);
    const float m3 = minor3(a2, a3, a4, b2, b3, b4, d2, d3, d4);
    const float m4 = minor3(a2, a3, a4, b2, b3, b4, c2, c3, c4);
    const float det = a1 * m1 - b1 * m2 + c1 * m3 - d1 * m4;
    if (det == 0.0f) return m;
    const float k = 1.0f / det;
    Mat4 r;
    float* o = r.e;
    o[0] = m1 * k;
    o[4] = -m2 * k;
    o[8] = m3 * k;
    o[12] = -m4 * k;
    o[1] = -minor3(b1, b3, b4, c1, c3, c4, d1, d3, d4) * k;
    o[5] = minor3(a1, a3, a4, c1, c3, c4, d1, d3, d4) * k;
    o[9] = -minor3(a1, a3, a4, b1, b3, b4, d1, d3, d4) * k;
    o[13] = minor3(a1, a3, a4, b1, b3, b4, c1, c3, c4) * k;
    o[2] = minor3(b1, b2, b4, c1, c2, c4, d1, d2, d4) * k;
    o[6] = -minor3(a1, a2, a4, c1, c2, c4, d1, d2, d4) * k;
    o[10] = minor3(a1, a2, a4, b1, b2, b4, d1, d2, d4) * k;
    o[14] = -minor3(a1, a2, a4, b1, b2, b4, c1, c2, c4) * k;
    o[3] = -minor3(b1, b2, b3, c1, c2, c3, d1, d2, d3) * k;
    o[7] = minor3(a1, a2, a3, c1, c2, c3, d1, d2, d3) * k;
    o[11] = -minor3(a1, a2, a3, b1, b2, b3, d1, d2, d3) * k;
    o[15] = minor3(a1, a2, a3, b1, b2, b3, c1, c2, c3) * k;
    return r;
}

// Rotation matrix of a unit quaternion (x, y, z, w) (Math/Quaternion.h:46-70).
Mat4 rotation(const float q[4]) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    Mat4 m = Mat4::identity();
    m.e[0] = 1.0f - 2.0f * (y * y + z * z);
    m.e[1] = 2.0f * (x * y + z * w);
    m.e[2] = 2.0f * (z * x - y * w);
    m.e[4] = 2.0f * (x * y - z * w);
    m.e[5] = 1.0f - 2.0f * (z * z + x * x);
    m.e[6] = 2.0f * (y * z + x * w);
    m.e[8] = 2.0f * (z * x + y * w);
    m.e[9] = 2.0f * (y * z - x * w);
    m.e[10] = 1.0f - 2.0f * (x * x + y * y);
    return m;
}

}  // namespace

void quat_about_y(float theta_deg, float q[4]) {
    // Quaternion(0,1,0, angle) * identity with angle = theta*3.1415926/180
    // taken to float (ST_Tomography.cpp:899-911, Math/Quaternion.cpp:7-35,
    // Quaternion.h:76-86).  The axis is already unit length.
    const float angle = (float)((double)theta_deg * 3.1415926 / 180);
    const float half = angle / 2;
    const float w = std::cos(half), s = std::sin(half);
    const float ax[3] = {0.0f, 1.0f * s, 0.0f};
    // product with the identity quaternion, term by term
    q[3] = w * 1.0f - ax[0] * 0.0f - ax[1] * 0.0f - ax[2] * 0.0f;
    q[0] = w * 0.0f + ax[0] * 1.0f + ax[1] * 0.0f - ax[2] * 0.0f;
    q[1] = w * 0.0f + ax[1] * 1.0f + ax[2] * 0.0f - ax[0] * 0.0f;
    q[2] = w * 0.0f + ax[2] * 1.0f + ax[0] * 0.0f - ax[1] * 0.0f;
}

void quat_rotate_y(const float q_old[4], double deg, float q[4]) {
    // RotateVolume (ST_Tomography.cpp:865-884): Quaternion(0,1,0, angle) *
    // m_OldRotation with angle = deg*3.1415926/180 taken to float, and the
    // Hamilton product in the order of Math/Quaternion.h:76-86.
    const float angle = (float)(deg * 3.1415926 / 180);
    const float half = angle / 2;
    const float pw = std::cos(half), ps = std::sin(half);
    const float px = 0.0f * ps, py = 1.0f * ps, pz = 0.0f * ps;
    const float qx = q_old[0], qy = q_old[1], qz = q_old[2], qw = q_old[3];
    q[3] = pw * qw - px * qx - py * qy - pz * qz;
    q[0] = pw * qx + px * qw + py * qz - pz * qy;
    q[1] = pw * qy + py * qw + pz * qx - px * qz;
    q[2] = pw * qz + pz * qw + px * qy - py * qx;
}

void scene_matrices(const stt_geometry& g, const float q[4], float view_out[16], float proj_out[16]) {
    // Camera of the ST_Tomography ctor and Scene::_calculateCamera /
    // _calculateMatrix (ST_Tomography.cpp:229-246; Scene.cpp:77,131,133,181-207).
    const double halfW = (g.det[0] - 1) * g.ds * 0.5f;
    const double halfH = (g.det[1] - 1) * g.ds * 0.5f;
    const float sid = (float)g.sid, sdd = (float)g.sdd;
    const float shift_x = (float)(-g.oxyz[0]);
    const float shift_y = (float)(-g.oxyz[2]);
    const float shift_z = (float)(+g.oxyz[1]);
    const float z_base = -1 * sid;
    const float n = sdd, f = 10000.0f;

    const Mat4 view = Mat4::translation(shift_x, shift_y, z_base + shift_z) * Mat4::identity() * rotation(q);
    Mat4 proj = Mat4::zero();
    proj.e[0] = (float)(n / halfW);
    proj.e[5] = (float)(n / halfH);
    proj.e[10] = -(f + n) / (f - n);
    proj.e[11] = -1.0f;
    proj.e[14] = -2.0f * f * n / (f - n);
    std::memcpy(view_out, view.e, sizeof(view.e));
    std::memcpy(proj_out, proj.e, sizeof(proj.e));
}

ViewRec build_view_from(const stt_geometry& g, const float view16[16], const float proj16[16]) {
    // TypicalRayProducer::_calculateMatrix (TypicalRayProducer.cpp:142-208)
    // from the modelview and projection matrices the solver sets
    // (SetViewMatrix / SetProjectionMatrix).
    const double halfW = (g.det[0] - 1) * g.ds * 0.5f;
    const double halfH = (g.det[1] - 1) * g.ds * 0.5f;
    const float sdd = (float)g.sdd;
    Mat4 view, proj;
    std::memcpy(view.e, view16, sizeof(view.e));
    std::memcpy(proj.e, proj16, sizeof(proj.e));

    float half_extent[3];
    for (int c = 0; c < 3; ++c) half_extent[c] = -((g.vol[c] - 1) * g.spacing[c]) * 0.5f;
    const Mat4 modelview = view * Mat4::translation(half_extent[0], half_extent[1], half_extent[2]);
    Mat4 m2v = Mat4::zero();
    for (int c = 0; c < 3; ++c) m2v.e[5 * c] = 1.0f / g.spacing[c];
    m2v.e[15] = 1.0f;
    const Mat4 v2m = inverted(m2v);
    Mat4 v2p = Mat4::zero();
    v2p.e[0] = 0.5f * (g.det[0] - 1);
    v2p.e[5] = 0.5f * (g.det[1] - 1);
    v2p.e[10] = 0.5f;
    v2p.e[12] = v2p.e[0];
    v2p.e[13] = v2p.e[5];
    v2p.e[14] = v2p.e[10];
    v2p.e[15] = 1.0f;
    const Mat4 p2m = inverted(v2p * proj * modelview);

    ViewRec rec;
    std::memset(&rec, 0, sizeof(rec));
    std::memcpy(rec.P2M, p2m.e, sizeof(rec.P2M));
    const Mat4 mvv2m = modelview * v2m;
    // voxel-BP projection rows folded with the detector scale (double, then fp32)
    const double k = (double)sdd / g.ds;
    for (int c = 0; c < 4; ++c) {
        rec.bp[c] = (float)(k * mvv2m.at(0, c));
        rec.bp[4 + c] = (float)(k * mvv2m.at(1, c));
        rec.bp[8 + c] = mvv2m.at(2, c);
    }
    rec.bp[12] = (float)(halfW / g.ds);
    rec.bp[13] = (float)(halfH / g.ds);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) rec.mvv2m[4 * r + c] = mvv2m.at(r, c);
    rec.eye[0] = p2m.e[8] / p2m.e[11];
    rec.eye[1] = p2m.e[9] / p2m.e[11];
    rec.eye[2] = p2m.e[10] / p2m.e[11];
    rec.eye[3] = 1.0f;
    apply(m2v, rec.eye, rec.o);
    // dominant axis of the ray from the source to the volume centre (voxel units)
    int dom = 0;
    float dmax = -1.0f, dsgn = 1.0f;
    for (int c = 0; c < 3; ++c) {
        const float dc = 0.5f * (g.vol[c] - 1) - rec.o[c];
        if (std::fabs(dc) > dmax) {
            dmax = std::fabs(dc);
            dom = c;
            dsgn = dc < 0 ? -1.0f : 1.0f;
        }
    }
    rec.march = dom | (dsgn < 0 ? 4 : 0) | (dom == 0 ? 8 : 0);
    for (int c = 0; c < 3; ++c) rec.m2v[c] = m2v.e[5 * c];
    rec.m2v[3] = 1.0f;
    rec.halfW = halfW;
    rec.halfH = halfH;
    rec.sdd = sdd;
    rec.ds = g.ds;
    return rec;
}

ViewRec build_view(const stt_geometry& g, const float q[4]) {
    float view[16], proj[16];
    scene_matrices(g, q, view, proj);
    return build_view_from(g, view, proj);
}

float sample_distance(const stt_geometry& g) {
    // sqrt(s.s) in float (ST_Tomography.cpp:300)
    return std::sqrt(g.spacing[0] * g.spacing[0] + g.spacing[1] * g.spacing[1] + g.spacing[2] * g.spacing[2]);
}

}  // namespace stt
