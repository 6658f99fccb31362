// stt_synth.cpp — seeded synthetic inputs (host only, no device needed).
//
// The reference reads its inputs from files (TIFF volumes, a projection
// stack, lds.txt start angles; VolumeData.cpp:103-129, ST_Tomography.cpp:
// 258-276,355).  The benchmark and tests use these generators instead:
//   * an additive ellipsoid phantom (random centres, semi-axes, rotations and
//     densities from one seed);
//   * a smooth deformation D(x,t) = sum of 4 sinusoidal modes, |D| <= amp,
//     with frame t the phantom pulled back through it and the flow of the
//     frame pair (t, t+1) u = D(x,t+1) - D(x,t) in voxel units (warp's
//     backward sampling X[t](x) ~ X[t+1](x + u), ST_Tomography.cpp:503);
//   * view tables: even, golden-angle, and van der Corput base 2 (the
//     acquisition's low-discrepancy sequence, Low_discrepency_squence.m:15-44,
//     with Ns = 1 and no reference-angle inserts).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/stt.h"

namespace {

struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

struct Ellipsoid {
    double c[3];
    double inv_a[3];
    double R[9];  // rows: body axes in volume coordinates
    float rho;
};

std::vector<Ellipsoid> make_ellipsoids(int nx, int ny, int nz, int n, uint64_t seed) {
    SplitMix rng(seed);
    const double N[3] = {(double)nx, (double)ny, (double)nz};
    const double Nm = (N[0] + N[1] + N[2]) / 3.0;
    std::vector<Ellipsoid> es(n);
    for (auto& e : es) {
        for (int c = 0; c < 3; ++c) e.c[c] = 0.5 * (N[c] - 1) + rng.uniform(-0.3, 0.3) * N[c];
        for (int c = 0; c < 3; ++c) e.inv_a[c] = 1.0 / (rng.uniform(0.05, 0.3) * Nm);
        // uniform random unit quaternion (Shoemake)
        const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
        const double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
        const double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
        double* R = e.R;
        R[0] = 1 - 2 * (qy * qy + qz * qz); R[1] = 2 * (qx * qy - qz * qw);     R[2] = 2 * (qx * qz + qy * qw);
        R[3] = 2 * (qx * qy + qz * qw);     R[4] = 1 - 2 * (qx * qx + qz * qz); R[5] = 2 * (qy * qz - qx * qw);
        R[6] = 2 * (qx * qz - qy * qw);     R[7] = 2 * (qy * qz + qx * qw);     R[8] = 1 - 2 * (qx * qx + qy * qy);
        e.rho = (float)rng.uniform(0.1, 1.0);
    }
    return es;
}

inline float phantom_at(const std::vector<Ellipsoid>& es, double x, double y, double z) {
    float v = 0.0f;
    for (const auto& e : es) {
        const double d[3] = {x - e.c[0], y - e.c[1], z - e.c[2]};
        double s = 0;
        for (int r = 0; r < 3; ++r) {
            const double b = (e.R[3 * r] * d[0] + e.R[3 * r + 1] * d[1] + e.R[3 * r + 2] * d[2]) * e.inv_a[r];
            s += b * b;
        }
        if (s <= 1.0) v += e.rho;
    }
    return v;
}

struct Deformation {
    double A[4][3];
    double k[4][3];
    double phi[4], omega[4];
    double N[3];
    void eval(double x, double y, double z, double t, double out[3]) const {
        out[0] = out[1] = out[2] = 0.0;
        for (int m = 0; m < 4; ++m) {
            const double arg = 2 * M_PI * (k[m][0] * x / N[0] + k[m][1] * y / N[1] + k[m][2] * z / N[2]) + phi[m] +
                               omega[m] * t;
            const double s = std::sin(arg);
            for (int c = 0; c < 3; ++c) out[c] += A[m][c] * s;
        }
    }
};

Deformation make_deformation(int nx, int ny, int nz, uint64_t seed, float amp) {
    SplitMix rng(seed);
    Deformation D;
    D.N[0] = nx; D.N[1] = ny; D.N[2] = nz;
    double total = 0;
    for (int m = 0; m < 4; ++m) {
        double v[3], n2 = 0;
        for (int c = 0; c < 3; ++c) { v[c] = rng.uniform(-1, 1); n2 += v[c] * v[c]; }
        const double mag = rng.uniform(0.5, 1.0) / std::sqrt(n2 + 1e-12);
        for (int c = 0; c < 3; ++c) D.A[m][c] = v[c] * mag;
        total += std::sqrt(D.A[m][0] * D.A[m][0] + D.A[m][1] * D.A[m][1] + D.A[m][2] * D.A[m][2]);
        for (int c = 0; c < 3; ++c) D.k[m][c] = (double)(int)(rng.uniform() * 3.0);  // 0, 1 or 2 periods
        if (D.k[m][0] == 0 && D.k[m][1] == 0 && D.k[m][2] == 0) D.k[m][m % 3] = 1;
        D.phi[m] = rng.uniform(0, 2 * M_PI);
        D.omega[m] = rng.uniform(0.2, 0.6);
    }
    // |D| <= sum_m |A_m| = amp
    const double s = total > 0 ? amp / total : 0.0;
    for (int m = 0; m < 4; ++m)
        for (int c = 0; c < 3; ++c) D.A[m][c] *= s;
    return D;
}

bool bad_dims(int nx, int ny, int nz) { return nx < 1 || ny < 1 || nz < 1; }

}  // namespace

extern "C" {

int stt_synth_phantom(int nx, int ny, int nz, int n_ellipsoids, uint64_t seed, float* out) {
    if (bad_dims(nx, ny, nz) || !out || n_ellipsoids < 0) return STT_ERR_INVALID;
    const auto es = make_ellipsoids(nx, ny, nz, n_ellipsoids, seed);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                out[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)] = phantom_at(es, x, y, z);
    return STT_OK;
}

int stt_synth_phantom_frame(int nx, int ny, int nz, int n_ellipsoids, uint64_t seed, uint64_t deform_seed, float amp,
                            int t, float* out) {
    if (bad_dims(nx, ny, nz) || !out || n_ellipsoids < 0) return STT_ERR_INVALID;
    const auto es = make_ellipsoids(nx, ny, nz, n_ellipsoids, seed);
    const Deformation D = make_deformation(nx, ny, nz, deform_seed, amp);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                double d[3];
                D.eval(x, y, z, t, d);
                out[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)] = phantom_at(es, x - d[0], y - d[1], z - d[2]);
            }
    return STT_OK;
}

int stt_synth_flow(int nx, int ny, int nz, uint64_t deform_seed, float amp, int t, float* uvw) {
    if (bad_dims(nx, ny, nz) || !uvw) return STT_ERR_INVALID;
    const Deformation D = make_deformation(nx, ny, nz, deform_seed, amp);
    const size_t V = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                double a[3], b[3];
                D.eval(x, y, z, t, a);
                D.eval(x, y, z, t + 1, b);
                const size_t q = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                for (int c = 0; c < 3; ++c) uvw[c * V + q] = (float)(b[c] - a[c]);
            }
    return STT_OK;
}

int stt_synth_angles(int kind, int n, int offset, float* out) {
    if (n < 0 || !out || offset < 0) return STT_ERR_INVALID;
    for (int k = 0; k < n; ++k) {
        double a;
        const long long i = (long long)offset + k;
        switch (kind) {
            case 0: a = 360.0 * k / n; break;                                           // even over 360
            case 1: a = std::fmod(i * 180.0 * (3.0 - std::sqrt(5.0)), 360.0); break;    // golden angle
            case 2: {                                                                   // van der Corput base 2
                double r = 0, q = 0.5;
                for (long long b = i; b != 0; b /= 2, q /= 2) r += (b % 2) * q;
                a = 360.0 * r;
                break;
            }
            default: return STT_ERR_INVALID;
        }
        out[k] = (float)a;
    }
    return STT_OK;
}

}  // extern "C"
