// slab_box_model.cpp -- host model of the slab projector's box containment.
//
// Test infrastructure.  Restates, with the same fp32 operations, what
// k_fp_slab does on the way to a sample (spacetimetomography_amd/csrc/
// stt_kernels.hip): produce_ray, the RayGen sample generator, the block-wide
// slope and first-plane reductions of a 16x16 ray tile, slab_box and the
// slab-by-slab walk of every ray.  For each wave that the unsampled-box path
// would take (all its sampling lanes march forward, near the volume), every
// sample's clamped cell and its +1 taps must lie inside the staged box; the
// model counts the samples that do not.  The view records come from the
// library's own build_view (stt_geometry.cpp, compiled in).
//
// bound = 0: the transverse box bounds as first written (floor of the line
// extremes widened by box_eps, clipped to [0, dims] at one end only);
// bound = 1: the extremes clamped to [0, dims-1] before the floor, the
// range the clamped sample positions actually take.
//
// slab_model_ex also models the ray-BP scatter's box (k_raybp_slab): a
// capacity in voxels and the dense layout with the slab axis slowest.
//
// Inner slabs (SlabBox::inner, walked without the clamp to [0, dims-1]): for
// every sample of such a slab the clamped and unclamped positions must be
// equal, and the unclamped slab-membership test must take the same samples
// and stop at the same one; the model counts the samples where either fails.
#include <cmath>
#include <cstring>
#include <vector>

#include "../include/stt.h"
#include "../spacetimetomography_amd/csrc/stt_device.hpp"
#include "../spacetimetomography_amd/csrc/stt_host.hpp"

namespace {

using stt::ViewRec;

inline float tmin(float a, float b) { return (a < b) ? a : b; }
inline float tmax(float a, float b) { return (a > b) ? a : b; }
inline float med3(float v, float lo, float hi) { return std::fmin(std::fmax(v, lo), hi); }

struct Geo {
    int dims[3];
    int W, H;
    float sd, bnd[3];
    int max_samples;
    float box_eps;
};

struct Ray {
    float o[3], d[3], tnear, tfar;
};

void produce_ray(const ViewRec& v, const Geo& g, float px, float py, Ray& r) {
    float m[4];
    for (int row = 0; row < 4; ++row) m[row] = (v.P2M[row] * px + v.P2M[4 + row] * py) + v.P2M[12 + row];
    const float winv = 1.0f / m[3];
    float dir[3];
    for (int c = 0; c < 3; ++c) dir[c] = m[c] * winv - v.eye[c];
    const float linv = 1.0f / std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (int c = 0; c < 3; ++c) {
        r.o[c] = v.o[c];
        r.d[c] = v.m2v[c] * (dir[c] * linv);
    }
    float tmn[3], tmx[3];
    for (int c = 0; c < 3; ++c) {
        const float t1 = (0.0f - r.o[c]) / r.d[c];
        const float t2 = (g.bnd[c] - r.o[c]) / r.d[c];
        tmn[c] = tmin(t1, t2);
        tmx[c] = tmax(t1, t2);
    }
    const float tn = tmax(tmax(tmn[0], tmn[1]), tmn[2]);
    r.tnear = tmax(tn, 0.0f);
    r.tfar = tmin(tmin(tmx[0], tmx[1]), tmx[2]);
}

struct Gen {
    float o[3], d[3], step[3], tnear, tfar, tMax[3], tDelta[3], t1, tb, t, pend[3], w;
    int segs;
};

void dda_advance(Gen& G) {
    if (G.tMax[0] < G.tMax[1]) {
        if (G.tMax[0] < G.tMax[2]) G.tMax[0] += G.tDelta[0];
        else G.tMax[2] += G.tDelta[2];
    } else {
        if (G.tMax[1] < G.tMax[2]) G.tMax[1] += G.tDelta[1];
        else G.tMax[2] += G.tDelta[2];
    }
}

bool gen_open(Gen& G, float sd) {
    for (int guard = 0; guard < 8; ++guard) {
        if (!(G.t1 < G.tfar) || G.segs >= 8) return false;
        float tb = tmin(tmin(G.tMax[0], G.tMax[1]), G.tMax[2]);
        if (tb > G.tfar) tb = G.tfar;
        G.tb = tb;
        G.t = std::ceil((G.t1 - G.tnear) / sd) * sd + G.tnear;
        for (int c = 0; c < 3; ++c) G.pend[c] = G.o[c] + G.d[c] * G.t;
        ++G.segs;
        if (G.t < G.tb) return true;
        dda_advance(G);
        G.t1 = G.tb;
    }
    return false;
}

bool gen_next(Gen& G, float sd) {
    dda_advance(G);
    G.t1 = G.tb;
    return gen_open(G, sd);
}

bool gen_init(Gen& G, const Ray& r, const Geo& g, bool hit, bool& entry_alone) {
    for (int c = 0; c < 3; ++c) {
        G.o[c] = r.o[c];
        G.d[c] = r.d[c];
        G.step[c] = r.d[c] * g.sd;
    }
    G.tnear = r.tnear;
    G.tfar = r.tfar;
    float pos1[3];
    for (int c = 0; c < 3; ++c) pos1[c] = r.o[c] + r.d[c] * r.tnear;
    for (int c = 0; c < 3; ++c) {
        const float dc = r.d[c];
        if (dc == 0.0f) {
            G.tMax[c] = r.tfar;
            G.tDelta[c] = 0.0f;
        } else if (dc < 0.0f) {
            const int blk = ((int)std::ceil(pos1[c] - 0.5f) - 1) / g.dims[c];
            G.tMax[c] = ((float)(blk * g.dims[c]) + 0.5f - r.o[c]) / dc;
            G.tDelta[c] = (float)(-g.dims[c]) / dc;
        } else {
            const int blk = ((int)std::floor(pos1[c] - 0.5f)) / g.dims[c];
            G.tMax[c] = ((float)((blk + 1) * g.dims[c]) + 0.5f - r.o[c]) / dc;
            G.tDelta[c] = (float)g.dims[c] / dc;
        }
    }
    G.t1 = r.tnear;
    G.segs = 0;
    G.tb = G.t = 0.0f;
    G.w = 1.0f;
    entry_alone = false;
    if (!hit) return false;
    const bool open = gen_open(G, g.sd);
    if (open && G.segs == 1) {
        G.w = 2.0f;
        return true;
    }
    entry_alone = true;
    return open;
}

struct Box {
    int lo[3], n[3], cbeg, cend;
    bool staged, inner;
};

constexpr int kSlabK = 23, kSlabKMin = 8, kStagePasses = 4;
constexpr float kSlabShear = 0.7f;

Box slab_box(const Geo& g, int D, bool down, int c0, const float o[3], float sAmin, float sAmax, float sBmin,
             float sBmax, float dfirst, int bound, int cap, bool dense) {
    const int A = D == 0 ? 1 : 0, B = D == 2 ? 1 : 2;
    Box bx;
    const float e0 = (down ? std::fmax((float)(c0 + 1), dfirst) : std::fmin((float)c0, dfirst)) - o[D];
    int K;
    {
        const float wA = std::fabs(e0) * (sAmax - sAmin) + 2.0f, gA = std::fmax(std::fabs(sAmin), std::fabs(sAmax));
        const float wB = std::fabs(e0) * (sBmax - sBmin) + 2.0f, gB = std::fmax(std::fabs(sBmin), std::fabs(sBmax));
        const float kf = kSlabShear * std::fmin(wA * (1.0f / gA), wB * (1.0f / gB));
        K = (int)std::fmin(std::fmax(kf, (float)kSlabKMin), (float)kSlabK);
    }
    const float e1 = (float)(down ? c0 - K + 1 : c0 + K) - o[D];
    const float a0 = std::fmin(std::fmin(e0 * sAmin, e0 * sAmax), std::fmin(e1 * sAmin, e1 * sAmax)) + o[A];
    const float a1 = std::fmax(std::fmax(e0 * sAmin, e0 * sAmax), std::fmax(e1 * sAmin, e1 * sAmax)) + o[A];
    const float b0 = std::fmin(std::fmin(e0 * sBmin, e0 * sBmax), std::fmin(e1 * sBmin, e1 * sBmax)) + o[B];
    const float b1 = std::fmax(std::fmax(e0 * sBmin, e0 * sBmax), std::fmax(e1 * sBmin, e1 * sBmax)) + o[B];
    const float lim = 1.0e6f, eps = g.box_eps;
    int loA, loB, hiA, hiB;
    if (bound == 0) {
        loA = std::max((int)std::floor(std::fmin(std::fmax(a0 - eps, -lim), lim)), 0);
        loB = std::max((int)std::floor(std::fmin(std::fmax(b0 - eps, -lim), lim)), 0);
        hiA = std::min((int)std::floor(std::fmin(std::fmax(a1 + eps, -lim), lim)) + 1, g.dims[A]);
        hiB = std::min((int)std::floor(std::fmin(std::fmax(b1 + eps, -lim), lim)) + 1, g.dims[B]);
    } else {
        loA = (int)std::floor(med3(a0 - eps, 0.0f, g.bnd[A]));
        loB = (int)std::floor(med3(b0 - eps, 0.0f, g.bnd[B]));
        hiA = (int)std::floor(med3(a1 + eps, 0.0f, g.bnd[A])) + 1;
        hiB = (int)std::floor(med3(b1 + eps, 0.0f, g.bnd[B])) + 1;
    }
    const int nA = hiA - loA + 1, nB = hiB - loB + 1;
    if (nA > 0 && nB > 0) {
        const int per_plane = dense ? nA * nB : (nA | 1) * nB;
        if (per_plane * (K + 1) > cap) {
            K = (int)((float)cap * (1.0f / (float)per_plane)) - 1;
            while (K > 0 && per_plane * (K + 1) > cap) --K;
            K = std::max(K, 1);
        }
    }
    int dlo, dhi;
    if (down) {
        dlo = std::max(c0 - K + 1, 0);
        dhi = std::min(c0 + 1, g.dims[D]);
        bx.cbeg = c0;
        bx.cend = c0 - K;
    } else {
        dlo = std::max(c0, 0);
        dhi = std::min(c0 + K, g.dims[D]);
        bx.cbeg = c0;
        bx.cend = c0 + K;
    }
    bx.lo[D] = dlo;
    bx.n[D] = dhi - dlo + 1;
    bx.lo[A] = loA;
    bx.n[A] = nA;
    bx.lo[B] = loB;
    bx.n[B] = nB;
    const int px = bx.n[0] | 1, pxy = (px * bx.n[1]) | 1;
    if (dense)  // the ray-BP scatter box: slab axis slowest, no padding
        bx.staged = nA > 0 && nB > 0 && bx.n[D] > 0 && nA <= cap && nB <= cap && nA * nB * bx.n[D] <= cap;
    else
        bx.staged = nA > 0 && nB > 0 && bx.n[D] > 0 && pxy * bx.n[2] <= cap && bx.n[0] * bx.n[1] <= 256 * kStagePasses;
    const float bD = g.bnd[D];
    bx.inner = bx.staged && a0 - eps >= 0.0f && a1 + eps <= g.bnd[A] && b0 - eps >= 0.0f && b1 + eps <= g.bnd[B] &&
               (down ? (c0 - K >= 0 && (float)(c0 + 1) <= bD) : (c0 >= 1 && (float)(c0 + K) <= bD));
    return bx;
}

struct Lane {
    Gen G;
    bool have, fwd, near;
    int n;
};

}  // namespace

// tw: rays per tile row (16, 32 or 8; 256 / tw rows), the k_fp_slab tile.
// cap: floats (projector) or voxels (ray-BP scatter) of the LDS box;
// dense: the scatter's layout (slab axis slowest, footprint n[A] n[B] n[D])
// instead of the projector's padded x-fastest one.
extern "C" int slab_model_ex(const stt_geometry* geom, const float* angles, int nviews, int bound, int tw, int cap,
                             int dense, long long* out) {
    if (tw != 16 && tw != 32 && tw != 8) return 1;
    const int th = 256 / tw;
    Geo g;
    for (int c = 0; c < 3; ++c) g.dims[c] = geom->vol[c];
    g.W = geom->det[0];
    g.H = geom->det[1];
    g.sd = stt::sample_distance(*geom);
    for (int c = 0; c < 3; ++c) g.bnd[c] = (float)(g.dims[c] - 1);
    g.max_samples = 2 * (g.dims[0] + g.dims[1] + g.dims[2]) + 16;
    {
        const float cmax = 2.0f * (float)std::max(g.dims[0], std::max(g.dims[1], g.dims[2]));
        const float ulp = std::nextafter(cmax, INFINITY) - cmax;
        g.box_eps = 1.0f / 32.0f + 0.5f * ulp * (float)g.max_samples;
    }
    long long checked = 0, bad = 0, waves = 0, samples = 0, inner_samples = 0, inner_bad = 0;
    for (int vi = 0; vi < nviews; ++vi) {
        float q[4];
        stt::quat_about_y(angles[vi], q);
        const ViewRec v = stt::build_view(*geom, q);
        const int march = v.march & 7;
        const int D = march & 3;
        const bool down = (march & 4) != 0;
        const int A = D == 0 ? 1 : 0, B = D == 2 ? 1 : 2;
        const int tiles_x = (g.W + tw - 1) / tw, tiles_y = (g.H + th - 1) / th;
        for (int tile = 0; tile < tiles_x * tiles_y; ++tile) {
            const int tx = tile % tiles_x, ty = tile / tiles_x;
            std::vector<Lane> L(256);
            float q6[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, down ? -INFINITY : INFINITY,
                           down ? -INFINITY : INFINITY};
            float o[3] = {v.o[0], v.o[1], v.o[2]};
            for (int t = 0; t < 256; ++t) {
                const int i = tx * tw + t % tw, j = ty * th + t / tw;
                const bool valid = i < g.W && j < g.H;
                Ray r;
                produce_ray(v, g, (float)(valid ? i : 0), (float)(valid ? j : 0), r);
                const bool hit = valid && r.tnear < r.tfar;
                Lane& ln = L[t];
                bool entry_alone;
                ln.have = gen_init(ln.G, r, g, hit, entry_alone);
                ln.n = (entry_alone || ln.have) ? 1 : 0;
                if (!entry_alone && ln.have) ln.G.w = 1.0f;
                ln.fwd = down ? r.d[D] < 0.0f : r.d[D] > 0.0f;
                const bool use = ln.have && ln.fwd;
                const float sa = r.d[A] / r.d[D], sb = r.d[B] / r.d[D];
                if (use) {
                    q6[0] = std::fmin(q6[0], sa);
                    q6[1] = std::fmax(q6[1], sa);
                    q6[2] = std::fmin(q6[2], sb);
                    q6[3] = std::fmax(q6[3], sb);
                    q6[5] = down ? std::fmax(q6[5], ln.G.pend[D]) : std::fmin(q6[5], ln.G.pend[D]);
                }
                if (ln.have) {
                    const float pD = std::floor(med3(ln.G.pend[D], 0.0f, g.bnd[D]));
                    q6[4] = down ? std::fmax(q6[4], pD) : std::fmin(q6[4], pD);
                }
                const float cmax = 2.0f * (std::fmax(g.bnd[0], std::fmax(g.bnd[1], g.bnd[2])) + 1.0f);
                ln.near = std::fmax(std::fabs(ln.G.pend[0]), std::fmax(std::fabs(ln.G.pend[1]),
                                                                       std::fabs(ln.G.pend[2]))) <= cmax;
            }
            bool wave_in_box[4];
            for (int w = 0; w < 4; ++w) {
                wave_in_box[w] = true;
                for (int t = 64 * w; t < 64 * w + 64; ++t)
                    if (L[t].have && !(L[t].fwd && L[t].near)) wave_in_box[w] = false;
            }
            const bool slopes_ok = q6[0] <= q6[1] && q6[2] <= q6[3];
            const bool any_ray = down ? q6[4] > -INFINITY : q6[4] < INFINITY;
            int c0 = (int)std::fmin(std::fmax(q6[4], -1.0f), (float)g.dims[0] + g.dims[1] + g.dims[2]);
            const int dimD = g.dims[D];
            for (int guard = 0; any_ray && guard <= dimD + 1; ++guard) {
                const float dfirst = guard == 0 ? q6[5] : (down ? -INFINITY : INFINITY);
                Box bx = slab_box(g, D, down, c0, o, q6[0], q6[1], q6[2], q6[3], dfirst, bound, cap, dense != 0);
                if (!slopes_ok) bx.staged = false;
                bool any_left = false;
                for (int t = 0; t < 256; ++t) {
                    Lane& ln = L[t];
                    const bool unchecked = bx.staged && wave_in_box[t / 64];
                    if (unchecked && (t % 64) == 0) ++waves;
                    for (int it = 0; ln.have && it <= g.max_samples; ++it) {
                        float cc[3];
                        for (int c = 0; c < 3; ++c) cc[c] = med3(ln.G.pend[c], 0.0f, g.bnd[c]);
                        const int p = (int)cc[D];
                        const bool member = down ? p > bx.cend : p < bx.cend;
                        if (unchecked && bx.inner) {
                            // inner slab: the unclamped walk takes the same samples and stops at the same one
                            const float lim = down ? (float)(bx.cend + 1) : (float)bx.cend;
                            const bool member_raw = down ? ln.G.pend[D] >= lim : ln.G.pend[D] < lim;
                            if (member_raw != member) ++inner_bad;
                            if (member) {
                                ++inner_samples;
                                for (int c = 0; c < 3; ++c)
                                    if (cc[c] != ln.G.pend[c]) {
                                        ++inner_bad;
                                        break;
                                    }
                            }
                        }
                        if (!member) break;
                        if (unchecked) {
                            ++checked;
                            for (int c = 0; c < 3; ++c) {
                                const int rel = (int)cc[c] - bx.lo[c];
                                if (rel < 0 || rel > bx.n[c] - 2) {
                                    ++bad;
                                    break;
                                }
                            }
                        }
                        ++samples;
                        ln.n += 1;
                        ln.G.t += g.sd;
                        for (int c = 0; c < 3; ++c) ln.G.pend[c] += ln.G.step[c];
                        if (!(ln.G.t < ln.G.tb)) ln.have = gen_next(ln.G, g.sd);
                        ln.have = ln.have && ln.n < g.max_samples;
                    }
                    any_left = any_left || ln.have;
                }
                c0 = bx.cend;
                if (!any_left) break;
            }
        }
    }
    out[0] = checked;
    out[1] = bad;
    out[2] = waves;
    out[3] = samples;
    out[4] = inner_samples;
    out[5] = inner_bad;
    return 0;
}

// The projector's box: 6144 floats, padded layout.
extern "C" int slab_model(const stt_geometry* geom, const float* angles, int nviews, int bound, int tw,
                          long long* out) {
    return slab_model_ex(geom, angles, nviews, bound, tw, 6144, 0, out);
}
