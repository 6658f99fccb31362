// stt_capi.cpp — the C-ABI of include/stt.h on one HIP device.
//
// A context owns every array of the reference's solver state in HBM
// (volumes x[0..F], X_bar, X_prev, duals Y_tv/Y_tp/Y_bc, flows, projection
// stacks, detector buffers) plus a device table of per-view geometry records
// and pre-built work tables for the two execution forms:
//   per-frame : the reference's order (ST_Tomography.cpp:492-607), one frame
//               per launch, one shared stale ray-length buffer (the single
//               TypicalRayProducer of the reference);
//   batched   : all duals, all primals, then view k of every frame in one
//               launch; per-frame ray-length buffers.
// All work goes to the context's stream; an iteration can be captured in a
// hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <functional>
#include <vector>

#include "../../include/stt.h"
#include "stt_device.hpp"
#include "stt_host.hpp"
#include "stt_kernels.hpp"

using namespace stt;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace

int stt::set_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail(e_ == hipErrorOutOfMemory ? STT_ERR_NOMEM : STT_ERR_HIP,                        \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                             \
    } while (0)

// zero tail behind every volume so the projector's +1 trilinear taps need no clamp
size_t tail_pad(const int* d) {
    size_t p = (size_t)d[0] * d[1] + d[0] + 2;
    return (p + 63) / 64 * 64;
}

// Scale of the ray-BP weight sums: a view adds weights w <= 1, at most
// max_samples per ray to one voxel, so a sum stays below D * max_samples <
// 2^e; at 2^(62 - e) every int64 weight sum stays below 2^62.
int raybp_weight_shift(const VolGeom& g) {
    int e = 0;
    (void)std::frexp((double)g.W * (double)g.H * (double)g.max_samples, &e);
    return 62 - e;
}

// Owners released on every return path of the entry points that allocate.
struct DevMem {
    void* p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};
// Stream-ordered scratch: freed with hipFreeAsync on the same stream on
// every return path, after the work queued before the return.
struct StreamMem {
    void* p = nullptr;
    hipStream_t s = nullptr;
    hipError_t alloc(size_t bytes, hipStream_t stream) {
        s = stream;
        return hipMallocAsync(&p, bytes, stream);
    }
    ~StreamMem() {
        if (p) (void)hipFreeAsync(p, s);
    }
};
struct OwnedStream {
    hipStream_t s = nullptr;
    ~OwnedStream() {
        if (s) (void)hipStreamDestroy(s);
    }
};
struct OwnedEvent {
    hipEvent_t e = nullptr;
    ~OwnedEvent() {
        if (e) (void)hipEventDestroy(e);
    }
};
struct OwnedGraphExec {
    hipGraphExec_t e = nullptr;
    ~OwnedGraphExec() {
        if (e) (void)hipGraphExecDestroy(e);
    }
};

// ---- optical flow, one level on device buffers ----------------------------
// V (3 planes) is updated in place on the ROI; Y (10 planes) receives the
// dual variables on the ROI (optical_flow.h:503-531).
int of_level_device(const VolGeom& g, const float ext[3], const float* I1, const float* I2, float* V, float* Y,
                    const stt_of_params& p, const int roi[6], hipStream_t s, int* iters, double* residual) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const int nb = of_partials(g);
    const int nbp = of_primal_partials(g, roi);  // <= nb
    DevMem mem;
    // W | G(4) | V'(3) | K4(1) | Kty(3) | y(10) floats, then 2 nb + 2 doubles
    const size_t nf = 22 * n;
    HIPCHK(hipMalloc(&mem.p, nf * sizeof(float) + (2 * (size_t)nb + 2) * sizeof(double)));
    float* W = (float*)mem.p;
    float* G = W + n;
    float* V2 = G + 4 * n;
    float* K4 = V2 + 3 * n;
    float* Kty = K4 + n;
    float* y = Kty + 3 * n;
    double* part = (double*)(W + nf);
    double* sums = part + 2 * (size_t)nb;
    HIPCHK(hipMemsetAsync(G, 0, 4 * n * sizeof(float), s));
    HIPCHK(hipMemsetAsync(Kty, 0, 3 * n * sizeof(float), s));
    HIPCHK(hipMemcpyAsync(y, Y, 10 * n * sizeof(float), hipMemcpyDeviceToDevice, s));
    // iteration k reads buf[(k-1) & 1] and writes buf[k & 1]; both start as the input flow
    HIPCHK(hipMemcpyAsync(V2, V, 3 * n * sizeof(float), hipMemcpyDeviceToDevice, s));
    float* const buf[2] = {V, V2};
    const float third = 1.0f / 3.0f;
    const float sigma1 = std::min(third, std::min(third, third));  // :170
    const float hf = 1.0f / (1.0f + sigma1 * p.huber / p.lambda);  // :183
    const int huber = p.norm == 4;
    // first: 1 for a level's first iteration, 2 for a later warp's first iteration
    auto iteration = [&](int it, int first) -> hipError_t {
        const int mode = it % 50 == 49 ? 1 : (it % 50 == 0 ? 2 : 0);
        const float* vin = buf[(it - 1) & 1];
        float* vout = buf[it & 1];
        hipError_t e = launch_of_primal(mode, vin, vout, y, Kty, G, part, g, roi, s);
        if (e == hipSuccess)
            e = launch_of_dual(mode == 2, first, vout, vin, y, K4, G, part + nb, g, sigma1, hf, p.lambda, huber, s);
        if (e == hipSuccess && mode == 2) e = launch_sum_partials(part, nbp, sums, s);
        if (e == hipSuccess && mode == 2) e = launch_sum_partials(part + nb, nb, sums + 1, s);
        return e;
    };
    // 50 iterations (the residual's period) as one graph, one per kind of
    // first iteration; 50 is even, so every chunk starts and ends on buf[0]
    OwnedGraphExec chunk[3];
    auto chunk_for = [&](int first) -> hipError_t {
        if (chunk[first].e) return hipSuccess;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        for (int i = 1; i <= 50 && e == hipSuccess; ++i) e = iteration(i, i == 1 ? first : 0);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = ce;
        if (e == hipSuccess) e = hipGraphInstantiate(&chunk[first].e, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        return e;
    };
    double host[2];
    int cur = 0;  // buffer holding the current flow
    for (int w = 0; w < p.warps; ++w) {
        if (cur) {  // the last warp stopped after an odd iteration count
            HIPCHK(hipMemcpyAsync(V, V2, 3 * n * sizeof(float), hipMemcpyDeviceToDevice, s));
            cur = 0;
        }
        const int first = w == 0 ? 1 : 2;
        if (w > 0) HIPCHK(launch_of_k4(V, G, K4, g, s));
        HIPCHK(launch_of_warp(V, I2, W, g, ext, s));
        HIPCHK(launch_of_prep(W, I1, V, G, g, roi, s));
        int it = 0;
        double err = 1.0;
        while (err > p.tol && it <= p.max_iter) {  // :290
            if (it % 50 == 0 && it + 50 <= p.max_iter + 1) {
                const int f = it == 0 ? first : 0;
                HIPCHK(chunk_for(f));
                HIPCHK(hipGraphLaunch(chunk[f].e, s));
                it += 50;
            } else {
                ++it;
                HIPCHK(iteration(it, it == 1 ? first : 0));
            }
            if (it % 50 == 0) {
                HIPCHK(hipMemcpyAsync(host, sums, sizeof(host), hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                err = (host[1] * host[1] + host[0] * host[0]) / (double)n;  // :485
            }
        }
        cur = it & 1;
        if (iters) iters[w] = it;
        if (residual) residual[w] = err;
    }
    if (cur) HIPCHK(hipMemcpyAsync(V, V2, 3 * n * sizeof(float), hipMemcpyDeviceToDevice, s));
    HIPCHK(launch_of_roi_copy(Y, y, 10, g, roi, s));
    HIPCHK(hipStreamSynchronize(s));
    return STT_OK;
}

// ---- the flow pyramid's image operators (CImg 1.8.0) on device buffers ----
// Order-0 Deriche coefficients (CImg.h:32946-32968), fp32 with std::exp as the
// reference computes them on the host.
DericheCoef deriche_coef(float sigma) {
    const float nnsigma = sigma < 0.1f ? 0.1f : sigma, alpha = 1.695f / nnsigma, ema = (float)std::exp(-alpha),
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

// get_blur(sigma): Deriche along x, y, z (CImg.h:33249-33284); tmp holds n floats.
int blur_device(const float* in, float* out, const int* d, int spectrum, float sigma, float* tmp, hipStream_t s) {
    const size_t n = (size_t)d[0] * d[1] * d[2];
    for (int ch = 0; ch < spectrum; ++ch) {
        const float* src = in + ch * n;
        float* dst = out + ch * n;
        if (sigma < 0.1f) {
            HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, s));
            continue;
        }
        const DericheCoef c = deriche_coef(sigma);
        HIPCHK(launch_deriche(src, dst, d, 0, c, s));
        HIPCHK(launch_deriche(dst, tmp, d, 1, c, s));
        HIPCHK(launch_deriche(tmp, dst, d, 2, c, s));
    }
    return STT_OK;
}

// Tables of one axis of get_resize(..., 5) from w to ns samples: the moving
// average's (source, weight) runs per output (CImg.h:28139-28150) or the
// cubic's source index and fraction (CImg's double recurrence, :28448-28454).
struct ResizeTables {
    std::vector<int> seg0, seg_s, seg_d, sidx;
    std::vector<double> frac;
};
ResizeTables resize_tables(int w, int ns) {
    ResizeTables t;
    if (w > ns) {
        t.seg0.assign(ns + 1, 0);
        unsigned a = (unsigned)w * ns, b = w, c = ns, src = 0, out = 0;
        while (a) {
            const unsigned dd = std::min(b, c);
            a -= dd;
            b -= dd;
            c -= dd;
            t.seg_s.push_back((int)src);
            t.seg_d.push_back((int)dd);
            if (!b) {
                ++out;
                t.seg0[out] = (int)t.seg_s.size();
                b = w;
            }
            if (!c) {
                ++src;
                c = ns;
            }
        }
    } else {
        const double f = ns > 1 ? (w - 1.0) / (ns - 1) : 0;
        double curr = 0;
        for (int x = 0; x < ns; ++x) {
            t.sidx.push_back((int)(unsigned)curr);
            t.frac.push_back(curr - (unsigned)curr);
            curr = std::min(w - 1.0, curr + f);
        }
    }
    return t;
}

// get_resize(nd, spectrum, 5) of a d-sized image with `spectrum` planes;
// axes x, y, z in turn; scratch holds 2 * max(n(d), n(nd)) floats.
int resize_device(const float* in, const int* d, int spectrum, const int* nd, float* out, float* scratch,
                  hipStream_t s) {
    const size_t n = (size_t)d[0] * d[1] * d[2], nn = (size_t)nd[0] * nd[1] * nd[2];
    const size_t half = std::max(n, nn);
    ResizeTables tab[3];
    size_t ints = 0, dbls = 0;
    for (int a = 0; a < 3; ++a) {
        if (nd[a] == d[a]) continue;
        tab[a] = resize_tables(d[a], nd[a]);
        ints += tab[a].seg0.size() + tab[a].seg_s.size() + tab[a].seg_d.size() + tab[a].sidx.size();
        dbls += tab[a].frac.size();
    }
    DevMem tm;
    if (ints + dbls) HIPCHK(hipMalloc(&tm.p, dbls * sizeof(double) + ints * sizeof(int) + 64));
    double* fbase = (double*)tm.p;
    int* ibase = (int*)(fbase + dbls);
    const int* tp[3][4] = {};
    const double* fp[3] = {};
    {
        double* fcur = fbase;
        int* icur = ibase;
        for (int a = 0; a < 3; ++a) {
            if (nd[a] == d[a]) continue;
            const std::vector<int>* v[4] = {&tab[a].seg0, &tab[a].seg_s, &tab[a].seg_d, &tab[a].sidx};
            for (int k = 0; k < 4; ++k) {
                tp[a][k] = icur;
                if (!v[k]->empty()) HIPCHK(hipMemcpy(icur, v[k]->data(), v[k]->size() * sizeof(int), hipMemcpyHostToDevice));
                icur += v[k]->size();
            }
            fp[a] = fcur;
            if (!tab[a].frac.empty())
                HIPCHK(hipMemcpy(fcur, tab[a].frac.data(), tab[a].frac.size() * sizeof(double), hipMemcpyHostToDevice));
            fcur += tab[a].frac.size();
        }
    }
    int last = -1;
    for (int a = 0; a < 3; ++a)
        if (nd[a] != d[a]) last = a;
    for (int ch = 0; ch < spectrum; ++ch) {
        const float* cur = in + ch * n;
        int cd[3] = {d[0], d[1], d[2]};
        int ping = 0;
        if (last < 0) {
            HIPCHK(hipMemcpyAsync(out + ch * nn, cur, n * sizeof(float), hipMemcpyDeviceToDevice, s));
            continue;
        }
        for (int a = 0; a < 3; ++a) {
            if (nd[a] == cd[a]) continue;
            float* dst = a == last ? out + ch * nn : scratch + ping * half;
            HIPCHK(launch_resize_axis(cur, dst, cd, a, nd[a], tp[a][0], tp[a][1], tp[a][2], tp[a][3], fp[a], s));
            cd[a] = nd[a];
            cur = dst;
            ping ^= 1;
        }
    }
    HIPCHK(hipStreamSynchronize(s));  // the tables are freed on return
    return STT_OK;
}

// L1Huber_OF_MultiScale3D (optical_flow.h:630-734) on device buffers, with
// Y1..Y4 = 0 as Render passes them (ST_Tomography.cpp:452-455).  V (3
// planes, full size) in/out.  iters: levels x warps, coarsest level first.
int of_multiscale_device(const int* D, const float* I1, const float* I2, float* V, const stt_of_params& p,
                         hipStream_t s, int* iters) {
    const size_t N = (size_t)D[0] * D[1] * D[2];
    const int L = p.levels;
    std::vector<std::array<int, 3>> dims(L);
    std::vector<size_t> sz(L), off(L);
    float t_eta = 1.0f;
    size_t total = 0;
    for (int k = 0; k < L; ++k) {  // finest first (:660-676)
        for (int a = 0; a < 3; ++a) dims[k][a] = (int)((float)D[a] * t_eta);
        for (int a = 0; a < 3; ++a)
            if (dims[k][a] < 4) return fail(STT_ERR_INVALID, "a pyramid level is smaller than 4 voxels");
        sz[k] = (size_t)dims[k][0] * dims[k][1] * dims[k][2];
        off[k] = total;
        total += 5 * sz[k];  // im1, im2, V (3)
        t_eta = t_eta * p.eta;
    }
    DevMem mem;
    // levels | blurred I1, I2 | Y (10, finest size) | Y' (10) | scratch (2 N, >= every resize)
    const size_t nf = total + 2 * N + 20 * N + 2 * N;
    HIPCHK(hipMalloc(&mem.p, nf * sizeof(float)));
    float* lv = (float*)mem.p;
    float* b1 = lv + total;
    float* b2 = b1 + N;
    float* Y = b2 + N;
    float* Yn = Y + 10 * N;
    float* scratch = Yn + 10 * N;
    int rc = blur_device(I1, b1, D, 1, p.sigma, scratch, s);
    if (!rc) rc = blur_device(I2, b2, D, 1, p.sigma, scratch, s);
    for (int k = 0; k < L && !rc; ++k) {
        float* im1 = lv + off[k];
        rc = resize_device(b1, D, 1, dims[k].data(), im1, scratch, s);
        if (!rc) rc = resize_device(b2, D, 1, dims[k].data(), im1 + sz[k], scratch, s);
        if (!rc) rc = resize_device(V, D, 3, dims[k].data(), im1 + 2 * sz[k], scratch, s);
    }
    if (rc) return rc;
    const float ext[3] = {(float)D[0], (float)D[1], (float)D[2]};
    for (int j = 0; j < L; ++j) {  // coarsest first (:690-729)
        const int k = L - 1 - j;
        const int* nd = dims[k].data();
        const VolGeom g = [&] {
            VolGeom v{};
            v.nx = nd[0];
            v.ny = nd[1];
            v.nz = nd[2];
            return v;
        }();
        if (j == 0) HIPCHK(hipMemsetAsync(Y, 0, 10 * sz[k] * sizeof(float), s));
        int roi[6];
        for (int a = 0; a < 6; ++a) roi[a] = (int)((float)p.cut[a] * ((float)nd[a / 2] / (float)D[a / 2]));  // :700-705
        float* im1 = lv + off[k];
        rc = of_level_device(g, ext, im1, im1 + sz[k], im1 + 2 * sz[k], Y, p, roi, s,
                             iters ? iters + j * p.warps : nullptr, nullptr);
        if (rc) return rc;
        if (k > 0) {  // :712-722
            float* vnext = lv + off[k - 1] + 2 * sz[k - 1];
            rc = resize_device(im1 + 2 * sz[k], nd, 3, dims[k - 1].data(), vnext, scratch, s);
            if (!rc) rc = resize_device(Y, nd, 10, dims[k - 1].data(), Yn, scratch, s);
            if (rc) return rc;
            std::swap(Y, Yn);
        }
    }
    HIPCHK(hipMemcpyAsync(V, lv + off[0] + 2 * sz[0], 3 * N * sizeof(float), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return STT_OK;
}

}  // namespace

struct stt_ctx {
    stt_geometry g{};
    stt_params p{};
    int F = 0;
    std::vector<int> P;
    std::vector<int> view_off;  // prefix sums of P
    int total_views = 0;
    int maxP = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    VolGeom vg{};
    PdConst pc{};
    size_t V = 0, Vpad = 0, VpadT = 0, D = 0;
    float gmax = 10.0f;
    bool gmax_from_data = true;
    bool pd_ready = false;

    // device buffers
    std::vector<float*> x, xT, xbar, xprev, ytv, ytp, ybc, flow, b, raylen_f, corr_f;
    // ray BP: per-frame exact 64-bit fixed-point sums and their scale slots
    std::vector<unsigned long long*> acc_c_f, acc_w_f;
    int* acc_shift_dev = nullptr;
    float* raylen_shared = nullptr;
    float* zero_flow = nullptr;
    // End points of the PD warps' traces of each flow u_f (warping.h:100-114):
    // trace_bw[f] with step +1 for the dual's backward warp of X_bar[f+1],
    // trace_fw[f] with step -1 for frame f+1's forward warp of Y_bc[f].  They
    // depend on the flow and the interpolation only, which stay fixed over the
    // PD passes of an outer iteration, so the velocity sampling runs once per
    // flow instead of in every pass.  flow_gen[f] counts the writes of u_f;
    // empty vectors (allocation refused) mean the kernels trace themselves.
    std::vector<float*> trace_bw, trace_fw;
    std::vector<uint64_t> flow_gen, trace_gen;
    int trace_interp = -1;
    bool trace_tried = false;
    bool keep_traces = true;  // stt_set_trace_storage
    ViewRec* views_dev = nullptr;
    std::vector<ViewRec> views_host;
    std::vector<char> views_set;
    FrameBufs* fb_dev = nullptr;
    // work tables: plain (frame-major, one item per view) and batched (per step k)
    WorkItem* items_plain = nullptr;
    WorkItem* items_batch = nullptr;
    std::vector<int> batch_off, batch_cnt;
    // the batched sweep split into frame groups, one per stream: group h
    // holds frames [split_f[h], split_f[h+1]), items in split_items, per k
    static constexpr int kGroups = 8;  // 4 and 2: the iteration 5% and 7% slower at 8 frames
    WorkItem* split_items = nullptr;
    int ngroups = 1;
    int split_f[kGroups + 1] = {};
    std::vector<int> split_off[kGroups], split_cnt[kGroups];
    int split_maxP[kGroups] = {};
    hipStream_t streams[kGroups] = {};  // streams[0] is the context stream
    hipEvent_t ev_fork = nullptr, ev_join[kGroups] = {};
    WorkItem* items_scratch = nullptr;  // ad-hoc single launches
    int scratch_cap = 0;
    bool tables_dirty = true;
    // graph
    hipGraphExec_t graph_exec = nullptr;
    int graph_batched = -1;
    uint64_t graph_gen = 0;  // bumped whenever captured launches would go stale (groups compare it)
    std::vector<void*> allocs;

    ~stt_ctx() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (void* a : allocs) (void)hipFree(a);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        for (int h = 1; h < kGroups; ++h) {
            if (ev_join[h]) (void)hipEventDestroy(ev_join[h]);
            if (streams[h]) (void)hipStreamDestroy(streams[h]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    int alloc(float** ptr, size_t nfloats) {
        void* d = nullptr;
        HIPCHK(hipMalloc(&d, nfloats * sizeof(float)));
        allocs.push_back(d);
        HIPCHK(hipMemsetAsync(d, 0, nfloats * sizeof(float), stream));
        *ptr = (float*)d;
        return STT_OK;
    }
    template <class T>
    int alloc_t(T** ptr, size_t n) {
        void* d = nullptr;
        HIPCHK(hipMalloc(&d, n * sizeof(T)));
        allocs.push_back(d);
        HIPCHK(hipMemsetAsync(d, 0, n * sizeof(T), stream));
        *ptr = (T*)d;
        return STT_OK;
    }

    void derive_consts() {
        pc.sigma = (float)p.sigma;
        pc.tau = (float)p.tau;
        pc.lambda = (float)p.lambda_tv;
        pc.ktp = (float)p.k_tp;
        pc.ebc = (float)p.e_bc;
        // m_huberfactor = 1/(1 + sigma*huber/lambda) (ST_Tomography.cpp:219)
        pc.huber_factor = (float)(1.0f / (1.0f + (double)pc.sigma * p.huber / p.lambda_tv));
        // Factor = sqrtf(2*SART_lambda), SART_lambda = tau/2 (:110-111)
        pc.factor = std::sqrt(2.0f * (pc.tau / 2.0f));
        pc.alpha = (float)p.alpha;
        for (int i = 0; i < 8; ++i) pc.cut[i] = p.roi_cut[i];
        pc.jtop = p.roi_jtop;
        pc.prior = p.prior;
        pc.interp = p.interp;
        pc.nframes = F;
        if (p.g_max > 0) {
            gmax = p.g_max;
            gmax_from_data = false;
        } else {
            gmax_from_data = true;
        }
        pc.gmax = gmax;
    }

    WorkItem make_item(int f, int view, bool shared_raylen) const {
        WorkItem it{};
        it.view = views_dev + view_off[f] + view;
        it.vol = x[f];
        it.volT = xT[f];
        it.b = b[f] + (size_t)view * D;
        it.raylen = shared_raylen ? raylen_shared : raylen_f[f];
        it.corr = corr_f[f];
        it.contrib = nullptr;
        it.acc_c = acc_c_f.empty() ? nullptr : acc_c_f[f];
        it.acc_w = acc_w_f.empty() ? nullptr : acc_w_f[f];
        it.acc_shift = acc_shift_dev ? acc_shift_dev + f : nullptr;
        it.xprev = xprev[f];
        it.xbar = xbar[f];
        it.frame = f;
        it.last = (view == P[f] - 1) ? 1 : 0;
        return it;
    }

    int build_tables() {
        if (!tables_dirty) return STT_OK;
        std::vector<WorkItem> plain;
        for (int f = 0; f < F; ++f)
            for (int k = 0; k < P[f]; ++k) plain.push_back(make_item(f, k, true));
        HIPCHK(hipMemcpyAsync(items_plain, plain.data(), plain.size() * sizeof(WorkItem), hipMemcpyHostToDevice, stream));
        std::vector<WorkItem> batch;
        batch_off.assign(maxP, 0);
        batch_cnt.assign(maxP, 0);
        for (int k = 0; k < maxP; ++k) {
            batch_off[k] = (int)batch.size();
            for (int f = 0; f < F; ++f)
                if (k < P[f]) batch.push_back(make_item(f, k, false));
            batch_cnt[k] = (int)batch.size() - batch_off[k];
        }
        HIPCHK(hipMemcpyAsync(items_batch, batch.data(), batch.size() * sizeof(WorkItem), hipMemcpyHostToDevice, stream));
        // frame groups with about equal view counts (contiguous frames)
        ngroups = std::min(kGroups, F);
        split_f[0] = 0;
        for (int h = 1, f = 0, acc = 0; h < ngroups; ++h) {
            while (f < F - (ngroups - h) && (long)acc * ngroups < (long)h * total_views) acc += P[f++];
            split_f[h] = std::max(f, split_f[h - 1] + 1);
            f = split_f[h];
        }
        split_f[ngroups] = F;
        std::vector<WorkItem> split;
        for (int h = 0; h < ngroups; ++h) {
            split_maxP[h] = 0;
            for (int f = split_f[h]; f < split_f[h + 1]; ++f) split_maxP[h] = std::max(split_maxP[h], P[f]);
            split_off[h].assign(split_maxP[h], 0);
            split_cnt[h].assign(split_maxP[h], 0);
            for (int k = 0; k < split_maxP[h]; ++k) {
                split_off[h][k] = (int)split.size();
                for (int f = split_f[h]; f < split_f[h + 1]; ++f)
                    if (k < P[f]) split.push_back(make_item(f, k, false));
                split_cnt[h][k] = (int)split.size() - split_off[h][k];
            }
        }
        HIPCHK(hipMemcpyAsync(split_items, split.data(), split.size() * sizeof(WorkItem), hipMemcpyHostToDevice, stream));
        std::vector<FrameBufs> fbs(F + 1);
        for (int f = 0; f <= F; ++f) {
            fbs[f].x = x[f];
            fbs[f].xbar = xbar[f];
            fbs[f].xprev = xprev[f];
            fbs[f].ytv = ytv[f];
            fbs[f].ytp = ytp[f];
            fbs[f].ybc = ybc[f];
            fbs[f].flow = f < F ? flow[f] : zero_flow;
            fbs[f].trace_bw = f < (int)trace_bw.size() ? trace_bw[f] : nullptr;
            fbs[f].trace_fw = f < (int)trace_fw.size() ? trace_fw[f] : nullptr;
        }
        pc.stored_traces = (int)trace_bw.size() == F && (int)trace_fw.size() == F;
        HIPCHK(hipMemcpyAsync(fb_dev, fbs.data(), fbs.size() * sizeof(FrameBufs), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        tables_dirty = false;
        return STT_OK;
    }

    int ready_check() const {
        for (int f = 0; f < F; ++f)
            if (!views_set[f]) return fail(STT_ERR_STATE, "views of frame " + std::to_string(f) + " not set");
        return STT_OK;
    }

    // Ad-hoc launch table (synchronous upload; callers synchronise anyway).
    int upload_scratch(const std::vector<WorkItem>& items) {
        if ((int)items.size() > scratch_cap) return fail(STT_ERR_INVALID, "too many work items");
        HIPCHK(hipMemcpyAsync(items_scratch, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));  // the host vector dies with the caller's expression
        return STT_OK;
    }

    int launch_bp_for(const WorkItem* items_dev, int n, bool ray_bp) {
        if (ray_bp) {
            HIPCHK(launch_raybp_prep(items_dev, n, vg, stream));
            HIPCHK(launch_fp(FP_MODE_RAYBP, items_dev, n, vg, 1.0f, stream));
            HIPCHK(launch_bp_ray_apply(items_dev, n, vg, pc, stream));
        } else {
            HIPCHK(launch_bp_voxel(items_dev, n, vg, pc, stream));
        }
        return STT_OK;
    }

    struct Timer {
        stt_timing* t;
        hipStream_t s;
        std::vector<hipEvent_t> ev;
        std::vector<int> kind;
        explicit Timer(stt_timing* tt, hipStream_t ss) : t(tt), s(ss) {}
        ~Timer() {
            for (auto e : ev) (void)hipEventDestroy(e);
        }
        void mark(int k) {
            if (!t) return;
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            (void)hipEventRecord(e, s);
            ev.push_back(e);
            kind.push_back(k);
        }
        // kind of interval i = kind recorded at its end mark
        void collect() {
            if (!t || ev.size() < 2) return;
            (void)hipEventSynchronize(ev.back());
            for (size_t i = 1; i < ev.size(); ++i) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, ev[i - 1], ev[i]);
                switch (kind[i]) {
                    case 1: t->fp_ms += ms; t->fp_launches++; break;
                    case 2: t->bp_ms += ms; t->bp_launches++; break;
                    case 3: t->dual_ms += ms; t->dual_launches++; break;
                    case 4: t->primal_ms += ms; t->primal_launches++; break;
                    case 5: t->aux_ms += ms; t->aux_launches++; break;
                    default: break;
                }
            }
        }
    };

    // One ST iteration (one No_PO pass over all frames).
    int issue_iteration(bool batched, Timer* tm) {
        const bool ray_bp = p.bp == STT_BP_RAY;
        auto mark = [&](int k) {
            if (tm) tm->mark(k);
        };
        if (batched && !tm && ngroups >= 2) {
            // The frames' sweeps are independent: the frame groups run on
            // their own streams (default priority), so one group's
            // backprojection overlaps another group's projector.  The duals
            // and primals of all frames run first, as one launch each (a
            // group's primal run beside the other groups' sweeps made the
            // iteration 0.7% slower, and no faster once the primal read
            // stored trace end points; per-group duals and primals, each
            // group's first primal waiting for the previous group's last
            // dual, were 0.5% slower).  Joined before returning.
            HIPCHK(launch_pd_dual(fb_dev, 0, F, vg, pc, stream));
            HIPCHK(launch_pd_primal(fb_dev, 0, F, vg, pc, stream));
            HIPCHK(hipEventRecord(ev_fork, stream));
            for (int h = 1; h < ngroups; ++h) HIPCHK(hipStreamWaitEvent(streams[h], ev_fork, 0));
            for (int h = 0; h < ngroups; ++h) {
                hipStream_t sh = streams[h];
                for (int k = 0; k < split_maxP[h]; ++k) {
                    const WorkItem* it = split_items + split_off[h][k];
                    const int n = split_cnt[h][k];
                    HIPCHK(launch_transpose_xz(it, n, vg, sh));
                    HIPCHK(launch_fp(FP_MODE_SART, it, n, vg, pc.factor, sh));
                    if (ray_bp) {
                        HIPCHK(launch_raybp_prep(it, n, vg, sh));
                        HIPCHK(launch_fp(FP_MODE_RAYBP, it, n, vg, 1.0f, sh));
                        HIPCHK(launch_bp_ray_apply(it, n, vg, pc, sh));
                    } else {
                        HIPCHK(launch_bp_voxel(it, n, vg, pc, sh));
                    }
                }
            }
            for (int h = 1; h < ngroups; ++h) {
                HIPCHK(hipEventRecord(ev_join[h], streams[h]));
                HIPCHK(hipStreamWaitEvent(stream, ev_join[h], 0));
            }
        } else if (batched) {
            HIPCHK(launch_pd_dual(fb_dev, 0, F, vg, pc, stream));
            mark(3);
            HIPCHK(launch_pd_primal(fb_dev, 0, F, vg, pc, stream));
            mark(4);
            for (int k = 0; k < maxP; ++k) {
                const WorkItem* it = items_batch + batch_off[k];
                HIPCHK(launch_transpose_xz(it, batch_cnt[k], vg, stream));
                mark(5);
                HIPCHK(launch_fp(FP_MODE_SART, it, batch_cnt[k], vg, pc.factor, stream));
                mark(1);
                int rc = launch_bp_for(it, batch_cnt[k], ray_bp);
                if (rc) return rc;
                mark(2);
            }
        } else {
            int off = 0;
            for (int f = 0; f < F; ++f) {
                HIPCHK(launch_pd_dual(fb_dev, f, 1, vg, pc, stream));
                mark(3);
                HIPCHK(launch_pd_primal(fb_dev, f, 1, vg, pc, stream));
                mark(4);
                for (int k = 0; k < P[f]; ++k) {
                    const WorkItem* it = items_plain + off + k;
                    HIPCHK(launch_transpose_xz(it, 1, vg, stream));
                    mark(5);
                    HIPCHK(launch_fp(FP_MODE_SART, it, 1, vg, pc.factor, stream));
                    mark(1);
                    int rc = launch_bp_for(it, 1, ray_bp);
                    if (rc) return rc;
                    mark(2);
                }
                off += P[f];
            }
        }
        return STT_OK;
    }

    int pd_init() {
        for (int f = 0; f <= F; ++f) {
            HIPCHK(hipMemcpyAsync(xbar[f], x[f], V * sizeof(float), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemsetAsync(ytv[f], 0, 4 * V * sizeof(float), stream));
            HIPCHK(hipMemsetAsync(ytp[f], 0, V * sizeof(float), stream));
            HIPCHK(hipMemsetAsync(ybc[f], 0, V * sizeof(float), stream));
            HIPCHK(hipMemsetAsync(xprev[f], 0, V * sizeof(float), stream));
        }
        pd_ready = true;
        return STT_OK;
    }

    void flow_changed(int f) { ++flow_gen[f]; }

    // Brings the stored trace end points up to date with the flows and the
    // interpolation (stream-ordered before the PD steps that read them).  The
    // first call allocates them, 24 bytes per voxel and frame; if the device
    // refuses that memory, the PD kernels keep tracing in place.
    int last_retraced = 0;  // flows re-traced by the latest refresh_traces
    int refresh_traces() {
        last_retraced = 0;
        if (!trace_tried && keep_traces) {
            trace_tried = true;
            std::vector<void*> got;
            bool ok = true;
            for (int f = 0; f < 2 * F && ok; ++f) {
                void* a = nullptr;
                ok = hipMalloc(&a, 3 * V * sizeof(float)) == hipSuccess;
                if (ok) got.push_back(a);
            }
            if (ok) {
                for (int f = 0; f < F; ++f) {
                    trace_bw.push_back((float*)got[2 * f]);
                    trace_fw.push_back((float*)got[2 * f + 1]);
                }
                allocs.insert(allocs.end(), got.begin(), got.end());
                tables_dirty = true;
                invalidate_graph();  // the PD launches switch to the stored-trace kernels
            } else {
                for (void* a : got) (void)hipFree(a);  // none of it is kept
                (void)hipGetLastError();              // the refused allocation is not an error of the next call
            }
        }
        if (!trace_bw.empty()) {
            for (int f = 0; f < F; ++f) {
                if (trace_gen[f] == flow_gen[f] && trace_interp == pc.interp) continue;
                HIPCHK(launch_trace(flow[f], trace_bw[f], trace_fw[f], vg, pc.interp, stream));
                trace_gen[f] = flow_gen[f];
                ++last_retraced;
            }
            trace_interp = pc.interp;
        }
        return build_tables();
    }

    // Takes an allocation off the list freed at destruction (the caller frees it).
    void forget_alloc(const void* a) {
        const auto at = std::find(allocs.begin(), allocs.end(), a);
        if (at != allocs.end()) allocs.erase(at);
    }

    // The x <-> z transposed volumes feed only the level-synchronous
    // projector's x-dominant views (k_transpose_xz, then k_fp_level reading
    // volT), so they exist only while a level shape is selected: (F+1) * V
    // floats, 9.1 GB at 512^3 x 16 frames, that the slab projector never reads.
    int ensure_transposed() {
        if (xT[0]) return STT_OK;
        for (int f = 0; f <= F; ++f) {
            const int rc = alloc(&xT[f], VpadT);
            if (rc) {
                (void)drop_transposed();
                return rc;
            }
        }
        tables_dirty = true;
        invalidate_graph();
        return STT_OK;
    }
    int drop_transposed() {
        bool any = false;
        for (float* a : xT) any = any || a;
        if (!any) return STT_OK;
        HIPCHK(hipStreamSynchronize(stream));
        hipError_t first = hipSuccess;  // every copy is released and forgotten even if one free fails
        for (float*& a : xT) {
            if (!a) continue;
            forget_alloc(a);
            const hipError_t e = hipFree(a);
            if (first == hipSuccess) first = e;
            a = nullptr;
        }
        tables_dirty = true;
        invalidate_graph();
        HIPCHK(first);
        return STT_OK;
    }

    // Release the stored trace end points (stt_set_trace_storage(ctx, 0)).
    int drop_traces() {
        if (trace_bw.empty()) return STT_OK;
        HIPCHK(hipStreamSynchronize(stream));
        hipError_t first = hipSuccess;  // every buffer is released and forgotten even if one free fails
        for (int f = 0; f < F; ++f)
            for (float* a : {trace_bw[f], trace_fw[f]}) {
                forget_alloc(a);
                const hipError_t e = hipFree(a);
                if (first == hipSuccess) first = e;
            }
        trace_bw.clear();
        trace_fw.clear();
        trace_gen.assign(F, 0);
        tables_dirty = true;
        invalidate_graph();
        HIPCHK(first);
        return STT_OK;
    }

    void invalidate_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
        graph_batched = -1;
        ++graph_gen;
    }

    // Everything stt_iterate checks and prepares before it issues work.
    int prepare_iteration() {
        int rc = ready_check();
        if (rc) return rc;
        rc = refresh_traces();
        if (rc) return rc;
        if (!pd_ready && (rc = pd_init())) return rc;
        return STT_OK;
    }
};

// Several contexts iterated as one unit (stt_group_*).
struct stt_group {
    std::vector<stt_ctx*> ctxs;
    std::vector<uint64_t> gens;  // the contexts' graph_gen when the graph was captured
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t exec = nullptr;

    ~stt_group() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (fork) (void)hipEventDestroy(fork);
        for (hipEvent_t e : join)
            if (e) (void)hipEventDestroy(e);
    }
};

#define CTX_CHECK(ctx) \
    if (!(ctx)) return fail(STT_ERR_INVALID, "null context")

namespace {

int valid_geometry(const stt_geometry* g) {
    if (!g) return fail(STT_ERR_INVALID, "null geometry");
    for (int c = 0; c < 3; ++c) {
        if (g->vol[c] < 4) return fail(STT_ERR_INVALID, "volume dims must be >= 4");
        if (g->vol[c] > 16384) return fail(STT_ERR_INVALID, "volume dims must be <= 16384");
        if (!(g->spacing[c] > 0.0f)) return fail(STT_ERR_INVALID, "voxel spacing must be > 0");
    }
    if (g->det[0] < 5 || g->det[1] < 5) return fail(STT_ERR_INVALID, "detector must be at least 5x5");
    if (!(g->ds > 0) || !(g->sdd > 0) || !(g->sid > 0)) return fail(STT_ERR_INVALID, "ds, sid, sdd must be > 0");
    const double v = (double)g->vol[0] * g->vol[1] * g->vol[2];
    if ((v + tail_pad(g->vol)) * 4.0 >= 4294967295.0) return fail(STT_ERR_INVALID, "a volume must stay below 4 GiB");
    // a flow field's three planes are addressed through one 32-bit buffer range
    if (v * 12.0 >= 4294967295.0) return fail(STT_ERR_INVALID, "a 3-plane flow field must stay below 4 GiB");
    if ((double)g->det[0] * g->det[1] * 4.0 >= 4294967295.0) return fail(STT_ERR_INVALID, "detector too large");
    if ((double)g->vol[0] * g->vol[1] * 4.0 >= 16777216.0 || (double)g->vol[2] * g->vol[1] * 4.0 >= 16777216.0)
        return fail(STT_ERR_INVALID, "a volume slice must stay below 4M voxels (24-bit offsets)");
    return STT_OK;
}

}  // namespace

extern "C" {

const char* stt_last_error(void) { return g_last_error.c_str(); }

int stt_version(void) { return 100; }

void stt_params_default(stt_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->sigma = 0.9 / 8.0;  // CP_sigma = .9/(L*CP_tau), L = 8 (ST_Tomography.cpp:99-102)
    p->tau = 1.0;          // CP_tau (:100)
    p->lambda_tv = 0.1;    // main.cpp:70
    p->k_tp = 0.0;         // main.cpp:71
    p->e_bc = 0.0;         // main.cpp:72
    p->huber = 0.1;        // main.cpp:73
    p->alpha = 0.1;        // main.cpp:76
    p->g_max = 0.0f;
    p->prior = STT_PRIOR_ATV;
    p->bp = STT_BP_VOXEL;
    p->interp = STT_INTERP_CUBIC_CLAMPED;
    p->roi_jtop = 0;
    p->fp64_reductions = 1;
}

void stt_params_reference(stt_params* p) {
    stt_params_default(p);
    if (!p) return;
    const int cut[8] = {8, 8, 8, 45, 8, 8, 30, 20};  // ST_Tomography.cpp:114
    for (int i = 0; i < 8; ++i) p->roi_cut[i] = cut[i];
    p->roi_jtop = 7;  // ST_Tomography.cpp:324
}

int stt_ctx_create(const stt_geometry* g, const stt_params* p, int nframes, const int* views_per_frame,
                   stt_ctx** out) {
    if (!out) return fail(STT_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int rc = valid_geometry(g);
    if (rc) return rc;
    if (!p) return fail(STT_ERR_INVALID, "null params");
    if (nframes < 1 || !views_per_frame) return fail(STT_ERR_INVALID, "nframes must be >= 1 with a view count per frame");
    for (int f = 0; f < nframes; ++f)
        if (views_per_frame[f] < 1) return fail(STT_ERR_INVALID, "every frame needs at least one view");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(STT_ERR_NODEVICE, "no HIP device visible");
    auto* c = new stt_ctx();
    c->g = *g;
    c->p = *p;
    c->F = nframes;
    c->P.assign(views_per_frame, views_per_frame + nframes);
    c->view_off.assign(nframes + 1, 0);
    for (int f = 0; f < nframes; ++f) c->view_off[f + 1] = c->view_off[f] + c->P[f];
    c->total_views = c->view_off[nframes];
    c->maxP = *std::max_element(c->P.begin(), c->P.end());
    (void)hipGetDevice(&c->device);
    auto bail = [&](int code) {
        delete c;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess)
        return bail(fail(STT_ERR_HIP, "stream or event creation failed"));
    c->streams[0] = c->stream;
    for (int h = 1; h < stt_ctx::kGroups; ++h)
        if (hipStreamCreateWithFlags(&c->streams[h], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_join[h], hipEventDisableTiming) != hipSuccess)
            return bail(fail(STT_ERR_HIP, "stream or event creation failed"));
    c->V = (size_t)g->vol[0] * g->vol[1] * g->vol[2];
    c->Vpad = c->V + tail_pad(g->vol);
    const int volT[3] = {g->vol[2], g->vol[1], g->vol[0]};
    c->VpadT = c->V + tail_pad(volT);
    c->D = (size_t)g->det[0] * g->det[1];
    c->vg.nx = g->vol[0];
    c->vg.ny = g->vol[1];
    c->vg.nz = g->vol[2];
    c->vg.W = g->det[0];
    c->vg.H = g->det[1];
    c->vg.sd = sample_distance(*g);
    c->vg.bx = (float)(g->vol[0] - 1);
    c->vg.by = (float)(g->vol[1] - 1);
    c->vg.bz = (float)(g->vol[2] - 1);
    c->vg.max_samples = 2 * (g->vol[0] + g->vol[1] + g->vol[2]) + 16;
    {
        // accumulated rounding of pend += step over max_samples steps, with
        // |pend| <= 2*max(dims) (half an ulp per addition), plus 1/32 cell
        const float cmax = 2.0f * (float)std::max(g->vol[0], std::max(g->vol[1], g->vol[2]));
        const float ulp = std::nextafter(cmax, INFINITY) - cmax;
        c->vg.box_eps = 1.0f / 32.0f + 0.5f * ulp * (float)c->vg.max_samples;
    }
    c->vg.acc_wshift = raybp_weight_shift(c->vg);
    c->vg.fp_box_kib = 24;
    c->vg.vol_bytes = (unsigned)(c->Vpad * sizeof(float));
    c->vg.volT_bytes = (unsigned)(c->VpadT * sizeof(float));
    c->vg.det_bytes = (unsigned)((size_t)g->det[0] * g->det[1] * sizeof(float));
    c->derive_consts();
    const int F = nframes;
    c->x.resize(F + 1);
    c->xT.assign(F + 1, nullptr);  // allocated while a level projector shape is selected
    c->xbar.resize(F + 1);
    c->xprev.resize(F + 1);
    c->ytv.resize(F + 1);
    c->ytp.resize(F + 1);
    c->ybc.resize(F + 1);
    c->flow.resize(F);
    c->b.resize(F);
    c->raylen_f.resize(F);
    c->corr_f.resize(F);
    for (int f = 0; f <= F; ++f) {
        if ((rc = c->alloc(&c->x[f], c->Vpad))) return bail(rc);
        if ((rc = c->alloc(&c->xbar[f], c->V))) return bail(rc);
        if ((rc = c->alloc(&c->xprev[f], c->V))) return bail(rc);
        if ((rc = c->alloc(&c->ytv[f], 4 * c->V))) return bail(rc);  // 3 spatial + temporal (STV) planes
        if ((rc = c->alloc(&c->ytp[f], c->V))) return bail(rc);
        if ((rc = c->alloc(&c->ybc[f], c->V))) return bail(rc);
    }
    for (int f = 0; f < F; ++f) {
        if ((rc = c->alloc(&c->flow[f], 3 * c->V))) return bail(rc);
        if ((rc = c->alloc(&c->b[f], c->D * c->P[f]))) return bail(rc);
        if ((rc = c->alloc(&c->raylen_f[f], c->D))) return bail(rc);
        if ((rc = c->alloc(&c->corr_f[f], c->D))) return bail(rc);
    }
    if (p->bp == STT_BP_RAY) {
        c->acc_c_f.resize(F);
        c->acc_w_f.resize(F);
        for (int f = 0; f < F; ++f) {
            if ((rc = c->alloc_t(&c->acc_c_f[f], c->Vpad))) return bail(rc);
            if ((rc = c->alloc_t(&c->acc_w_f[f], c->Vpad))) return bail(rc);
        }
        if ((rc = c->alloc_t(&c->acc_shift_dev, (size_t)F))) return bail(rc);
    }
    if ((rc = c->alloc(&c->raylen_shared, c->D))) return bail(rc);
    if ((rc = c->alloc(&c->zero_flow, 3 * c->V))) return bail(rc);
    if ((rc = c->alloc_t(&c->views_dev, (size_t)c->total_views))) return bail(rc);
    if ((rc = c->alloc_t(&c->fb_dev, (size_t)F + 1))) return bail(rc);
    if ((rc = c->alloc_t(&c->items_plain, (size_t)c->total_views))) return bail(rc);
    if ((rc = c->alloc_t(&c->items_batch, (size_t)c->total_views))) return bail(rc);
    if ((rc = c->alloc_t(&c->split_items, (size_t)c->total_views))) return bail(rc);
    c->scratch_cap = std::max(std::max(64, c->maxP + F), 4 * F);  // 4*F: the tuner's batched launches
    if ((rc = c->alloc_t(&c->items_scratch, (size_t)c->scratch_cap))) return bail(rc);
    c->views_host.resize(c->total_views);
    c->views_set.assign(F, 0);
    c->flow_gen.assign(F, 1);
    c->trace_gen.assign(F, 0);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(STT_ERR_HIP, "stream sync failed"));
    *out = c;
    return STT_OK;
}

int stt_ctx_destroy(stt_ctx* ctx) {
    CTX_CHECK(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return STT_OK;
}

int stt_ctx_set_params(stt_ctx* ctx, const stt_params* p) {
    CTX_CHECK(ctx);
    if (!p) return fail(STT_ERR_INVALID, "null params");
    if (p->bp != ctx->p.bp) return fail(STT_ERR_INVALID, "the backprojector is fixed at context creation");
    const float old_gmax = ctx->gmax;
    const bool from_data = ctx->gmax_from_data;
    ctx->p = *p;
    ctx->derive_consts();
    if (p->g_max <= 0 && from_data) {
        ctx->gmax = old_gmax;  // keep the data-derived value
        ctx->pc.gmax = old_gmax;
    }
    ctx->invalidate_graph();
    return STT_OK;
}

int stt_ctx_get_gmax(stt_ctx* ctx, float* g_max) {
    CTX_CHECK(ctx);
    if (!g_max) return fail(STT_ERR_INVALID, "null output");
    *g_max = ctx->gmax;
    return STT_OK;
}

int stt_set_view_quaternions(stt_ctx* ctx, int frame, const float* q, int n) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    if (!q || n != ctx->P[frame]) return fail(STT_ERR_INVALID, "view count must equal views_per_frame[frame]");
    ViewRec* dst = ctx->views_host.data() + ctx->view_off[frame];
    for (int i = 0; i < n; ++i) dst[i] = build_view(ctx->g, q + 4 * i);
    HIPCHK(hipMemcpyAsync(ctx->views_dev + ctx->view_off[frame], dst, n * sizeof(ViewRec), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->views_set[frame] = 1;
    ctx->tables_dirty = true;
    ctx->invalidate_graph();
    return STT_OK;
}

int stt_set_view_matrices(stt_ctx* ctx, int frame, const float* view16, const float* proj16, int n) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    if (!view16 || !proj16 || n != ctx->P[frame])
        return fail(STT_ERR_INVALID, "need n = views_per_frame[frame] view and projection matrices");
    ViewRec* dst = ctx->views_host.data() + ctx->view_off[frame];
    for (int i = 0; i < n; ++i) dst[i] = build_view_from(ctx->g, view16 + 16 * i, proj16 + 16 * i);
    HIPCHK(hipMemcpyAsync(ctx->views_dev + ctx->view_off[frame], dst, n * sizeof(ViewRec), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->views_set[frame] = 1;
    ctx->tables_dirty = true;
    ctx->invalidate_graph();
    return STT_OK;
}

int stt_scene_matrices(const stt_geometry* g, const float* q_xyzw, float* view16, float* proj16) {
    int rc = valid_geometry(g);
    if (rc) return rc;
    if (!q_xyzw || !view16 || !proj16) return fail(STT_ERR_INVALID, "null argument");
    scene_matrices(*g, q_xyzw, view16, proj16);
    return STT_OK;
}

int stt_set_views(stt_ctx* ctx, int frame, const float* angles_deg, int n) {
    CTX_CHECK(ctx);
    if (!angles_deg) return fail(STT_ERR_INVALID, "null angles");
    if (n < 0) return fail(STT_ERR_INVALID, "negative view count");
    std::vector<float> q(4 * (size_t)n);
    for (int i = 0; i < n; ++i) quat_about_y(angles_deg[i], &q[4 * i]);
    return stt_set_view_quaternions(ctx, frame, q.data(), n);
}

static int buffer_of(stt_ctx* ctx, int which, int frame, float** ptr, size_t* n) {
    if (frame < 0 || frame > ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    switch (which) {
        case STT_BUF_XBAR: *ptr = ctx->xbar[frame]; *n = ctx->V; break;
        case STT_BUF_XPREV: *ptr = ctx->xprev[frame]; *n = ctx->V; break;
        case STT_BUF_YTV: *ptr = ctx->ytv[frame]; *n = 4 * ctx->V; break;
        case STT_BUF_YTP: *ptr = ctx->ytp[frame]; *n = ctx->V; break;
        case STT_BUF_YBC: *ptr = ctx->ybc[frame]; *n = ctx->V; break;
        case STT_BUF_FLOW:
            if (frame >= ctx->F) return fail(STT_ERR_INVALID, "flow pair out of range");
            *ptr = ctx->flow[frame];
            *n = 3 * ctx->V;
            break;
        default: return fail(STT_ERR_INVALID, "unknown buffer id");
    }
    return STT_OK;
}

int stt_get_buffer(stt_ctx* ctx, int which, int frame, float* out) {
    CTX_CHECK(ctx);
    if (!out) return fail(STT_ERR_INVALID, "null output");
    float* src;
    size_t n;
    int rc = buffer_of(ctx, which, frame, &src, &n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, src, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_set_buffer(stt_ctx* ctx, int which, int frame, const float* in) {
    CTX_CHECK(ctx);
    if (!in) return fail(STT_ERR_INVALID, "null input");
    float* dst;
    size_t n;
    int rc = buffer_of(ctx, which, frame, &dst, &n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(dst, in, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (which == STT_BUF_FLOW) ctx->flow_changed(frame);
    return STT_OK;
}

int stt_set_volume(stt_ctx* ctx, int frame, const float* vol) {
    CTX_CHECK(ctx);
    if (!vol) return fail(STT_ERR_INVALID, "null volume");
    if (frame < 0 || frame > ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    HIPCHK(hipMemcpyAsync(ctx->x[frame], vol, ctx->V * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_get_volume(stt_ctx* ctx, int frame, float* vol) {
    CTX_CHECK(ctx);
    if (!vol) return fail(STT_ERR_INVALID, "null volume");
    if (frame < 0 || frame > ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    HIPCHK(hipMemcpyAsync(vol, ctx->x[frame], ctx->V * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_set_volume_dev(stt_ctx* ctx, int frame, const float* vol_dev) {
    CTX_CHECK(ctx);
    if (!vol_dev) return fail(STT_ERR_INVALID, "null volume");
    if (frame < 0 || frame > ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    HIPCHK(hipMemcpyAsync(ctx->x[frame], vol_dev, ctx->V * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_get_volume_dev(stt_ctx* ctx, int frame, float* vol_dev) {
    CTX_CHECK(ctx);
    if (!vol_dev) return fail(STT_ERR_INVALID, "null volume");
    if (frame < 0 || frame > ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    HIPCHK(hipMemcpyAsync(vol_dev, ctx->x[frame], ctx->V * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

namespace {

// The buffer entry points below serve both the host-pointer form and the
// "_dev" form (device pointers, borrowed): `in` and `out` are the copy kinds
// of the caller's inputs and outputs.
int set_flow_impl(stt_ctx* ctx, int pair, const float* uvw, hipMemcpyKind in) {
    CTX_CHECK(ctx);
    if (!uvw) return fail(STT_ERR_INVALID, "null flow");
    if (pair < 0 || pair >= ctx->F) return fail(STT_ERR_INVALID, "flow pair out of range");
    HIPCHK(hipMemcpyAsync(ctx->flow[pair], uvw, 3 * ctx->V * sizeof(float), in, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->flow_changed(pair);
    return STT_OK;
}

int set_projections_impl(stt_ctx* ctx, int frame, const float* b, hipMemcpyKind in) {
    CTX_CHECK(ctx);
    if (!b) return fail(STT_ERR_INVALID, "null projections");
    if (frame < 0 || frame >= ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    const size_t n = ctx->D * ctx->P[frame];
    HIPCHK(hipMemcpyAsync(ctx->b[frame], b, n * sizeof(float), in, ctx->stream));
    if (frame == 0 && ctx->gmax_from_data) {
        // G_Max = max(frame-0 projections)/4 (ST_Tomography.cpp:356-357)
        std::vector<float> staged;
        const float* hb = b;
        if (in != hipMemcpyHostToDevice) {  // device input: its values, read back once
            staged.resize(n);
            HIPCHK(hipMemcpyAsync(staged.data(), ctx->b[frame], n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            hb = staged.data();
        }
        float mx = -3.402823466e38f;
        for (size_t k = 0; k < n; ++k) mx = std::max(mx, hb[k]);
        ctx->gmax = mx / 4.0f;
        ctx->pc.gmax = ctx->gmax;
        ctx->invalidate_graph();
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int forward_project_impl(stt_ctx* ctx, int frame, int view, float* Ax, float* raylen, float* weight, hipMemcpyKind in,
                         hipMemcpyKind out) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F || view < 0 || view >= ctx->P[frame])
        return fail(STT_ERR_INVALID, "frame/view out of range");
    if (!ctx->views_set[frame]) return fail(STT_ERR_STATE, "views of this frame not set");
    const size_t D = ctx->D;
    float *dax = nullptr, *drl = nullptr, *dw = nullptr;
    StreamMem dax_mem;
    HIPCHK(dax_mem.alloc(3 * D * sizeof(float), ctx->stream));
    dax = (float*)dax_mem.p;
    drl = dax + D;
    dw = drl + D;
    if (raylen) HIPCHK(hipMemcpyAsync(drl, raylen, D * sizeof(float), in, ctx->stream));
    else HIPCHK(hipMemsetAsync(drl, 0, D * sizeof(float), ctx->stream));
    WorkItem it = ctx->make_item(frame, view, true);
    it.ax = dax;
    it.raylen = drl;
    it.weight = dw;
    int rc = ctx->upload_scratch({it});
    if (rc) return rc;
    HIPCHK(launch_fp(FP_MODE_FP, ctx->items_scratch, 1, ctx->vg, 1.0f, ctx->stream));
    if (Ax) HIPCHK(hipMemcpyAsync(Ax, dax, D * sizeof(float), out, ctx->stream));
    if (raylen) HIPCHK(hipMemcpyAsync(raylen, drl, D * sizeof(float), out, ctx->stream));
    if (weight) HIPCHK(hipMemcpyAsync(weight, dw, D * sizeof(float), out, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int backproject_voxel_impl(stt_ctx* ctx, int frame, int view, const float* corr, double alpha, hipMemcpyKind in) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F || view < 0 || view >= ctx->P[frame])
        return fail(STT_ERR_INVALID, "frame/view out of range");
    if (!corr) return fail(STT_ERR_INVALID, "null correction");
    if (!ctx->views_set[frame]) return fail(STT_ERR_STATE, "views of this frame not set");
    HIPCHK(hipMemcpyAsync(ctx->corr_f[frame], corr, ctx->D * sizeof(float), in, ctx->stream));
    WorkItem it = ctx->make_item(frame, view, true);
    it.last = 0;
    int rc = ctx->upload_scratch({it});
    if (rc) return rc;
    PdConst pc = ctx->pc;
    pc.alpha = (float)alpha;
    HIPCHK(launch_bp_voxel(ctx->items_scratch, 1, ctx->vg, pc, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

}  // namespace

int stt_set_flow(stt_ctx* ctx, int pair, const float* uvw) {
    return set_flow_impl(ctx, pair, uvw, hipMemcpyHostToDevice);
}
int stt_set_flow_dev(stt_ctx* ctx, int pair, const float* uvw_dev) {
    return set_flow_impl(ctx, pair, uvw_dev, hipMemcpyDeviceToDevice);
}
int stt_set_projections(stt_ctx* ctx, int frame, const float* b) {
    return set_projections_impl(ctx, frame, b, hipMemcpyHostToDevice);
}
int stt_set_projections_dev(stt_ctx* ctx, int frame, const float* b_dev) {
    return set_projections_impl(ctx, frame, b_dev, hipMemcpyDeviceToDevice);
}
int stt_forward_project(stt_ctx* ctx, int frame, int view, float* Ax, float* raylen, float* weight) {
    return forward_project_impl(ctx, frame, view, Ax, raylen, weight, hipMemcpyHostToDevice, hipMemcpyDeviceToHost);
}
int stt_forward_project_dev(stt_ctx* ctx, int frame, int view, float* Ax_dev, float* raylen_dev, float* weight_dev) {
    return forward_project_impl(ctx, frame, view, Ax_dev, raylen_dev, weight_dev, hipMemcpyDeviceToDevice,
                                hipMemcpyDeviceToDevice);
}
int stt_backproject_voxel(stt_ctx* ctx, int frame, int view, const float* corr, double alpha) {
    return backproject_voxel_impl(ctx, frame, view, corr, alpha, hipMemcpyHostToDevice);
}
int stt_backproject_voxel_dev(stt_ctx* ctx, int frame, int view, const float* corr_dev, double alpha) {
    return backproject_voxel_impl(ctx, frame, view, corr_dev, alpha, hipMemcpyDeviceToDevice);
}

int stt_backproject_ray(stt_ctx* ctx, int frame, int view, const float* corr, float* contrib, float* wsum) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F || view < 0 || view >= ctx->P[frame])
        return fail(STT_ERR_INVALID, "frame/view out of range");
    if (!corr) return fail(STT_ERR_INVALID, "null correction");
    if (!ctx->views_set[frame]) return fail(STT_ERR_STATE, "views of this frame not set");
    const size_t Vp = ctx->Vpad, V = ctx->V;
    DevMem acc;
    HIPCHK(hipMalloc(&acc.p, 2 * Vp * sizeof(unsigned long long) + 256));
    auto* ac = (unsigned long long*)acc.p;
    auto* aw = ac + Vp;
    int* shift = (int*)(aw + Vp);
    HIPCHK(hipMemsetAsync(acc.p, 0, 2 * Vp * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->corr_f[frame], corr, ctx->D * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    WorkItem it = ctx->make_item(frame, view, true);
    it.acc_c = ac;
    it.acc_w = aw;
    it.acc_shift = shift;
    it.raylen = nullptr;  // the caller's correction stands alone: no ray lengths to reset
    it.last = 0;
    int rc = ctx->upload_scratch({it});
    if (rc) return rc;
    HIPCHK(launch_raybp_prep(ctx->items_scratch, 1, ctx->vg, ctx->stream));
    HIPCHK(launch_fp(FP_MODE_RAYBP, ctx->items_scratch, 1, ctx->vg, 1.0f, ctx->stream));
    std::vector<long long> hc, hw;
    int hs = 0;
    if (contrib || wsum) {
        hc.resize(V);
        hw.resize(V);
        HIPCHK(hipMemcpyAsync(hc.data(), ac, V * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(hw.data(), aw, V * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(&hs, shift, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(launch_bp_ray_apply(ctx->items_scratch, 1, ctx->vg, ctx->pc, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; contrib && k < V; ++k) contrib[k] = (float)std::ldexp((double)hc[k], -hs);
    for (size_t k = 0; wsum && k < V; ++k) wsum[k] = (float)std::ldexp((double)hw[k], -ctx->vg.acc_wshift);
    return STT_OK;
}

namespace {

int fp_vjp_impl(stt_ctx* ctx, int frame, int view, const float* y, float* out, hipMemcpyKind in, hipMemcpyKind back) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F || view < 0 || view >= ctx->P[frame])
        return fail(STT_ERR_INVALID, "frame/view out of range");
    if (!y || !out) return fail(STT_ERR_INVALID, "null argument");
    if (!ctx->views_set[frame]) return fail(STT_ERR_STATE, "views of this frame not set");
    float* acc = nullptr;
    StreamMem acc_mem;
    HIPCHK(acc_mem.alloc((ctx->Vpad + ctx->D) * sizeof(float), ctx->stream));
    acc = (float*)acc_mem.p;
    HIPCHK(hipMemsetAsync(acc, 0, ctx->Vpad * sizeof(float), ctx->stream));
    HIPCHK(hipMemcpyAsync(acc, out, ctx->V * sizeof(float), in, ctx->stream));
    float* dy = acc + ctx->Vpad;
    HIPCHK(hipMemcpyAsync(dy, y, ctx->D * sizeof(float), in, ctx->stream));
    WorkItem it = ctx->make_item(frame, view, true);
    it.corr = dy;
    it.contrib = acc;
    it.raylen = nullptr;
    int rc = ctx->upload_scratch({it});
    if (rc) return rc;
    // A x = sd * sum_s trilinear(x, p_s), so A^T y scatters y*sd with the trilinear weights.
    HIPCHK(launch_fp(FP_MODE_SCATTER, ctx->items_scratch, 1, ctx->vg, ctx->vg.sd, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, acc, ctx->V * sizeof(float), back, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

}  // namespace

int stt_fp_vjp(stt_ctx* ctx, int frame, int view, const float* y, float* out) {
    return fp_vjp_impl(ctx, frame, view, y, out, hipMemcpyHostToDevice, hipMemcpyDeviceToHost);
}
int stt_fp_vjp_dev(stt_ctx* ctx, int frame, int view, const float* y_dev, float* out_dev) {
    return fp_vjp_impl(ctx, frame, view, y_dev, out_dev, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice);
}

int stt_sart_sweep(stt_ctx* ctx, int frame, const int* order, int n) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    if (n < 1 || n > ctx->P[frame]) return fail(STT_ERR_INVALID, "view count out of range");
    if (!ctx->views_set[frame]) return fail(STT_ERR_STATE, "views of this frame not set");
    std::vector<WorkItem> items;
    for (int i = 0; i < n; ++i) {
        const int v = order ? order[i] : i;
        if (v < 0 || v >= ctx->P[frame]) return fail(STT_ERR_INVALID, "view index out of range in order");
        WorkItem it = ctx->make_item(frame, v, true);
        it.last = (i == n - 1) ? 1 : 0;
        it.xbar = nullptr;  // ROI mask only; the extrapolation belongs to the PD loop
        items.push_back(it);
    }
    const bool ray_bp = ctx->p.bp == STT_BP_RAY;
    for (int i = 0; i < n; ++i) {
        int rc = ctx->upload_scratch({items[i]});
        if (rc) return rc;
        HIPCHK(launch_transpose_xz(ctx->items_scratch, 1, ctx->vg, ctx->stream));
        HIPCHK(launch_fp(FP_MODE_SART, ctx->items_scratch, 1, ctx->vg, ctx->pc.factor, ctx->stream));
        rc = ctx->launch_bp_for(ctx->items_scratch, 1, ray_bp);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(ctx->stream));  // the scratch slot is reused by the next view
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

namespace {

int warp_impl(stt_ctx* ctx, const float* uvw, const float* src, float dt, float* out, int interp, hipMemcpyKind in,
              hipMemcpyKind back) {
    CTX_CHECK(ctx);
    if (!uvw || !src || !out) return fail(STT_ERR_INVALID, "null argument");
    if (interp != STT_INTERP_CUBIC_CLAMPED && interp != STT_INTERP_TRILINEAR)
        return fail(STT_ERR_INVALID, "unknown interpolation");
    const size_t V = ctx->V;
    float* buf = nullptr;
    StreamMem buf_mem;
    HIPCHK(buf_mem.alloc(5 * V * sizeof(float), ctx->stream));
    buf = (float*)buf_mem.p;
    HIPCHK(hipMemcpyAsync(buf, uvw, 3 * V * sizeof(float), in, ctx->stream));
    HIPCHK(hipMemcpyAsync(buf + 3 * V, src, V * sizeof(float), in, ctx->stream));
    HIPCHK(launch_warp(buf, buf + 3 * V, buf + 4 * V, ctx->vg, dt, interp == STT_INTERP_CUBIC_CLAMPED,
                       ctx->stream));
    HIPCHK(hipMemcpyAsync(out, buf + 4 * V, V * sizeof(float), back, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

}  // namespace

int stt_warp(stt_ctx* ctx, const float* uvw, const float* src, float dt, float* out, int interp) {
    return warp_impl(ctx, uvw, src, dt, out, interp, hipMemcpyHostToDevice, hipMemcpyDeviceToHost);
}
int stt_warp_dev(stt_ctx* ctx, const float* uvw_dev, const float* src_dev, float dt, float* out_dev, int interp) {
    return warp_impl(ctx, uvw_dev, src_dev, dt, out_dev, interp, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice);
}

int stt_warp_frame(stt_ctx* ctx, int pair, float dt, int interp, float* out) {
    CTX_CHECK(ctx);
    if (!out) return fail(STT_ERR_INVALID, "null output");
    if (pair < 0 || pair + 1 >= ctx->F) return fail(STT_ERR_INVALID, "frame pair out of range");
    if (interp != STT_INTERP_CUBIC_CLAMPED && interp != STT_INTERP_TRILINEAR)
        return fail(STT_ERR_INVALID, "unknown interpolation");
    const size_t V = ctx->V;
    float* buf = nullptr;
    StreamMem buf_mem;
    HIPCHK(buf_mem.alloc(V * sizeof(float), ctx->stream));
    buf = (float*)buf_mem.p;
    HIPCHK(launch_warp(ctx->flow[pair], ctx->x[pair + 1], buf, ctx->vg, dt, interp == STT_INTERP_CUBIC_CLAMPED,
                       ctx->stream));
    HIPCHK(hipMemcpyAsync(out, buf, V * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

namespace {

int warp_vjp_impl(stt_ctx* ctx, const float* uvw, const float* src, float dt, const float* g, float* gsrc, float* guvw,
                  hipMemcpyKind in, hipMemcpyKind back) {
    CTX_CHECK(ctx);
    if (!uvw || !src || !g) return fail(STT_ERR_INVALID, "null argument");
    if (!gsrc && !guvw) return STT_OK;
    const size_t V = ctx->V;
    float* buf = nullptr;
    StreamMem buf_mem;
    HIPCHK(buf_mem.alloc(9 * V * sizeof(float), ctx->stream));
    buf = (float*)buf_mem.p;
    float *d_uvw = buf, *d_src = buf + 3 * V, *d_g = buf + 4 * V, *d_gsrc = buf + 5 * V, *d_guvw = buf + 6 * V;
    HIPCHK(hipMemcpyAsync(d_uvw, uvw, 3 * V * sizeof(float), in, ctx->stream));
    HIPCHK(hipMemcpyAsync(d_src, src, V * sizeof(float), in, ctx->stream));
    HIPCHK(hipMemcpyAsync(d_g, g, V * sizeof(float), in, ctx->stream));
    if (gsrc) HIPCHK(hipMemcpyAsync(d_gsrc, gsrc, V * sizeof(float), in, ctx->stream));
    if (guvw) HIPCHK(hipMemcpyAsync(d_guvw, guvw, 3 * V * sizeof(float), in, ctx->stream));
    HIPCHK(launch_warp_vjp(d_uvw, d_src, d_g, gsrc ? d_gsrc : nullptr, guvw ? d_guvw : nullptr, ctx->vg, dt,
                           ctx->stream));
    if (gsrc) HIPCHK(hipMemcpyAsync(gsrc, d_gsrc, V * sizeof(float), back, ctx->stream));
    if (guvw) HIPCHK(hipMemcpyAsync(guvw, d_guvw, 3 * V * sizeof(float), back, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

}  // namespace

int stt_warp_vjp(stt_ctx* ctx, const float* uvw, const float* src, float dt, const float* g, float* gsrc,
                 float* guvw) {
    return warp_vjp_impl(ctx, uvw, src, dt, g, gsrc, guvw, hipMemcpyHostToDevice, hipMemcpyDeviceToHost);
}
int stt_warp_vjp_dev(stt_ctx* ctx, const float* uvw_dev, const float* src_dev, float dt, const float* g_dev,
                     float* gsrc_dev, float* guvw_dev) {
    return warp_vjp_impl(ctx, uvw_dev, src_dev, dt, g_dev, gsrc_dev, guvw_dev, hipMemcpyDeviceToDevice,
                         hipMemcpyDeviceToDevice);
}

int stt_pd_init(stt_ctx* ctx) {
    CTX_CHECK(ctx);
    int rc = ctx->pd_init();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_set_trace_storage(stt_ctx* ctx, int keep) {
    CTX_CHECK(ctx);
    if (keep != 0 && keep != 1) return fail(STT_ERR_INVALID, "keep must be 0 or 1");
    ctx->keep_traces = keep == 1;
    if (keep) {
        ctx->trace_tried = !ctx->trace_bw.empty();  // a refused allocation is tried again
        return STT_OK;
    }
    return ctx->drop_traces();
}

int stt_pd_dual(stt_ctx* ctx, int frame) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    int rc = ctx->refresh_traces();
    if (rc) return rc;
    HIPCHK(launch_pd_dual(ctx->fb_dev, frame, 1, ctx->vg, ctx->pc, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_pd_primal(stt_ctx* ctx, int frame) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    int rc = ctx->refresh_traces();
    if (rc) return rc;
    HIPCHK(launch_pd_primal(ctx->fb_dev, frame, 1, ctx->vg, ctx->pc, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_pd_extrapolate(stt_ctx* ctx, int frame) {
    CTX_CHECK(ctx);
    if (frame < 0 || frame >= ctx->F) return fail(STT_ERR_INVALID, "frame out of range");
    int rc = ctx->build_tables();
    if (rc) return rc;
    HIPCHK(launch_extrapolate(ctx->fb_dev, frame, 1, ctx->vg, ctx->pc, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_iterate(stt_ctx* ctx, int n_iters, int batched, int use_graph, stt_timing* timing) {
    CTX_CHECK(ctx);
    if (n_iters < 0) return fail(STT_ERR_INVALID, "negative iteration count");
    if (timing) std::memset(timing, 0, sizeof(*timing));
    OwnedEvent p0, p1, e0, e1;
    if (timing) {
        HIPCHK(hipEventCreate(&p0.e));
        HIPCHK(hipEventCreate(&p1.e));
        HIPCHK(hipEventCreate(&e0.e));
        HIPCHK(hipEventCreate(&e1.e));
        HIPCHK(hipEventRecord(p0.e, ctx->stream));
    }
    int rc = ctx->prepare_iteration();
    if (rc) return rc;
    if (timing) {
        HIPCHK(hipEventRecord(p1.e, ctx->stream));
        HIPCHK(hipEventRecord(e0.e, ctx->stream));
    }
    if (use_graph) {
        if (!ctx->graph_exec || ctx->graph_batched != (batched ? 1 : 0)) {
            ctx->invalidate_graph();
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            rc = ctx->issue_iteration(batched != 0, nullptr);
            hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ce != hipSuccess) return fail(STT_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return fail(STT_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
            ctx->graph_batched = batched ? 1 : 0;
        }
        for (int it = 0; it < n_iters; ++it) HIPCHK(hipGraphLaunch(ctx->graph_exec, ctx->stream));
    } else {
        for (int it = 0; it < n_iters; ++it) {
            stt_ctx::Timer tm(timing, ctx->stream);
            tm.mark(0);
            rc = ctx->issue_iteration(batched != 0, timing ? &tm : nullptr);
            if (rc) return rc;
            tm.collect();
        }
    }
    if (timing) {
        HIPCHK(hipEventRecord(e1.e, ctx->stream));
        HIPCHK(hipEventSynchronize(e1.e));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0.e, e1.e));
        timing->total_ms = ms;
        HIPCHK(hipEventElapsedTime(&ms, p0.e, p1.e));
        timing->trace_ms = ms;
        timing->trace_launches = ctx->last_retraced;
    }
    return STT_OK;
}

int stt_time_projector(stt_ctx* ctx, int graph_form, int reps, double* ms_per_iteration) {
    CTX_CHECK(ctx);
    if (!ms_per_iteration || reps < 1) return fail(STT_ERR_INVALID, "need reps >= 1 and an output");
    int rc = ctx->ready_check();
    if (rc) return rc;
    if ((rc = ctx->build_tables())) return rc;
    const size_t D = ctx->D;
    // every frame's views in the form's launch tables, with the corrections
    // and ray lengths redirected to scratch (no solver state changes)
    std::vector<WorkItem> items;
    std::vector<std::array<int, 3>> launches;  // (group, first item, count)
    if (graph_form) {
        for (int h = 0; h < ctx->ngroups; ++h)
            for (int k = 0; k < ctx->split_maxP[h]; ++k) {
                launches.push_back({h, (int)items.size(), ctx->split_cnt[h][k]});
                for (int f = ctx->split_f[h]; f < ctx->split_f[h + 1]; ++f)
                    if (k < ctx->P[f]) items.push_back(ctx->make_item(f, k, false));
            }
    } else {
        for (int k = 0; k < ctx->maxP; ++k) {
            launches.push_back({0, (int)items.size(), ctx->batch_cnt[k]});
            for (int f = 0; f < ctx->F; ++f)
                if (k < ctx->P[f]) items.push_back(ctx->make_item(f, k, false));
        }
    }
    StreamMem buf_mem, tab_mem;
    HIPCHK(buf_mem.alloc(2 * items.size() * D * sizeof(float), ctx->stream));
    HIPCHK(tab_mem.alloc(items.size() * sizeof(WorkItem), ctx->stream));
    auto* buf = (float*)buf_mem.p;
    auto* tab = (WorkItem*)tab_mem.p;
    for (size_t n = 0; n < items.size(); ++n) {
        items[n].corr = buf + 2 * n * D;
        items[n].raylen = buf + (2 * n + 1) * D;
        items[n].last = 0;
    }
    HIPCHK(hipMemsetAsync(buf, 0, 2 * items.size() * D * sizeof(float), ctx->stream));
    HIPCHK(hipMemcpyAsync(tab, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    auto issue = [&](bool multi) -> hipError_t {
        hipError_t e = hipSuccess;
        if (multi) {
            e = hipEventRecord(ctx->ev_fork, ctx->stream);
            for (int h = 1; h < ctx->ngroups && e == hipSuccess; ++h) e = hipStreamWaitEvent(ctx->streams[h], ctx->ev_fork, 0);
        }
        for (const auto& l : launches) {
            hipStream_t s = multi ? ctx->streams[l[0]] : ctx->stream;
            if (e == hipSuccess) e = launch_transpose_xz(tab + l[1], l[2], ctx->vg, s);
            if (e == hipSuccess) e = launch_fp(FP_MODE_SART, tab + l[1], l[2], ctx->vg, ctx->pc.factor, s);
        }
        if (multi)
            for (int h = 1; h < ctx->ngroups && e == hipSuccess; ++h) {
                e = hipEventRecord(ctx->ev_join[h], ctx->streams[h]);
                if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->ev_join[h], 0);
            }
        return e;
    };
    OwnedGraphExec exec;
    if (graph_form) {
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t ie = issue(ctx->ngroups >= 2);
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        if (ie != hipSuccess || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(STT_ERR_HIP, std::string("projector graph capture: ") +
                                         hipGetErrorString(ie != hipSuccess ? ie : ce));
        }
        const hipError_t ge = hipGraphInstantiate(&exec.e, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(ge);
        HIPCHK(hipGraphLaunch(exec.e, ctx->stream));  // warm-up
    } else {
        HIPCHK(issue(false));  // warm-up
    }
    OwnedEvent t0, t1;
    HIPCHK(hipEventCreate(&t0.e));
    HIPCHK(hipEventCreate(&t1.e));
    HIPCHK(hipEventRecord(t0.e, ctx->stream));
    for (int r = 0; r < reps; ++r) {
        if (graph_form) HIPCHK(hipGraphLaunch(exec.e, ctx->stream));
        else HIPCHK(issue(false));
    }
    HIPCHK(hipEventRecord(t1.e, ctx->stream));
    HIPCHK(hipEventSynchronize(t1.e));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, t0.e, t1.e));
    *ms_per_iteration = ms / reps;
    return STT_OK;
}

int stt_group_create(stt_ctx* const* ctxs, int n, stt_group** out) {
    if (!out) return fail(STT_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if (!ctxs || n < 1) return fail(STT_ERR_INVALID, "a group needs at least one context");
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return fail(STT_ERR_INVALID, "null context in group");
        if (ctxs[i]->device != ctxs[0]->device) return fail(STT_ERR_INVALID, "group contexts must share one device");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return fail(STT_ERR_INVALID, "a context appears twice in the group");
    }
    auto* g = new stt_group();
    g->ctxs.assign(ctxs, ctxs + n);
    g->join.assign(n, nullptr);
    hipError_t e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
    for (int i = 1; i < n && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        delete g;
        return fail(STT_ERR_HIP, std::string("group events: ") + hipGetErrorString(e));
    }
    *out = g;
    return STT_OK;
}

int stt_group_destroy(stt_group* grp) {
    delete grp;
    return STT_OK;
}

namespace {

// Every context's batched iteration on its own stream, forked from and
// joined into the first context's stream.
int issue_group(stt_group* grp) {
    stt_ctx* c0 = grp->ctxs[0];
    const int n = (int)grp->ctxs.size();
    if (n > 1) {
        HIPCHK(hipEventRecord(grp->fork, c0->stream));
        for (int i = 1; i < n; ++i) HIPCHK(hipStreamWaitEvent(grp->ctxs[i]->stream, grp->fork, 0));
    }
    for (int i = 0; i < n; ++i) {
        const int rc = grp->ctxs[i]->issue_iteration(true, nullptr);
        if (rc) return rc;
    }
    for (int i = 1; i < n; ++i) {
        HIPCHK(hipEventRecord(grp->join[i], grp->ctxs[i]->stream));
        HIPCHK(hipStreamWaitEvent(c0->stream, grp->join[i], 0));
    }
    return STT_OK;
}

}  // namespace

int stt_group_iterate(stt_group* grp, int n_iters, int use_graph, double* total_ms) {
    if (!grp) return fail(STT_ERR_INVALID, "null group");
    if (n_iters < 0) return fail(STT_ERR_INVALID, "negative iteration count");
    for (stt_ctx* c : grp->ctxs) {
        const int rc = c->prepare_iteration();
        if (rc) return rc;
    }
    stt_ctx* c0 = grp->ctxs[0];
    // the contexts' own set-up (table uploads, PD initialisation) is on their
    // streams; the first context's stream waits for it before the group starts
    for (size_t i = 1; i < grp->ctxs.size(); ++i) {
        HIPCHK(hipEventRecord(grp->join[i], grp->ctxs[i]->stream));
        HIPCHK(hipStreamWaitEvent(c0->stream, grp->join[i], 0));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (total_ms) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, c0->stream));
    }
    if (use_graph) {
        bool stale = grp->exec == nullptr;
        for (size_t i = 0; i < grp->ctxs.size() && !stale; ++i) stale = grp->gens[i] != grp->ctxs[i]->graph_gen;
        if (stale) {
            if (grp->exec) (void)hipGraphExecDestroy(grp->exec);
            grp->exec = nullptr;
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
            const int rc = issue_group(grp);
            const hipError_t ce = hipStreamEndCapture(c0->stream, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ce != hipSuccess) return fail(STT_ERR_HIP, std::string("group graph capture: ") + hipGetErrorString(ce));
            const hipError_t ie = hipGraphInstantiate(&grp->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return fail(STT_ERR_HIP, std::string("group graph instantiate: ") + hipGetErrorString(ie));
            grp->gens.clear();
            for (stt_ctx* c : grp->ctxs) grp->gens.push_back(c->graph_gen);
        }
        for (int it = 0; it < n_iters; ++it) HIPCHK(hipGraphLaunch(grp->exec, c0->stream));
    } else {
        for (int it = 0; it < n_iters; ++it) {
            const int rc = issue_group(grp);
            if (rc) return rc;
        }
    }
    if (total_ms) {
        HIPCHK(hipEventRecord(e1, c0->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *total_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    // later calls on any member context see the group's work finished
    for (size_t i = 1; i < grp->ctxs.size(); ++i) {
        HIPCHK(hipEventRecord(grp->fork, c0->stream));
        HIPCHK(hipStreamWaitEvent(grp->ctxs[i]->stream, grp->fork, 0));
    }
    return STT_OK;
}

int stt_set_projector_tile(stt_ctx* ctx, int rays_per_row) {
    CTX_CHECK(ctx);
    if (rays_per_row != 0 && rays_per_row != -32 && rays_per_row != -8 && rays_per_row != 64 && rays_per_row != 32 &&
        rays_per_row != 16 && rays_per_row != 8)
        return fail(STT_ERR_INVALID, "projector shape must be 0, -32 or -8 (slab), 64, 32, 16 or 8 (level-synchronous)");
    const int rc = rays_per_row > 0 ? ctx->ensure_transposed() : ctx->drop_transposed();
    if (rc) return rc;
    ctx->vg.fp_tile = rays_per_row;
    ctx->invalidate_graph();
    return STT_OK;
}

namespace {

// Times the projector launches a SART sweep issues (for the first view
// indices k: the layout copy and the SART projector over every frame's view
// k, batched form) under each of n configurations that set(i) applies to the
// context's VolGeom, and returns the fastest index.  Corrections and ray
// lengths go to scratch buffers, so no solver state changes (the layout
// copy is rebuilt before every use).  ms[i] (may be NULL) receives the time
// per view index.
int time_projector_configs(stt_ctx* ctx, int n, const std::function<void(int)>& set, double* ms, int* best_index) {
    const size_t D = ctx->D;
    const int nk = std::min(ctx->maxP, 4);
    std::vector<WorkItem> items;
    std::vector<int> off, cnt;
    for (int k = 0; k < nk; ++k) {
        off.push_back((int)items.size());
        for (int f = 0; f < ctx->F; ++f)
            if (k < ctx->P[f]) items.push_back(ctx->make_item(f, k, false));
        cnt.push_back((int)items.size() - off.back());
    }
    StreamMem buf_mem;
    HIPCHK(buf_mem.alloc(2 * items.size() * D * sizeof(float), ctx->stream));
    float* buf = (float*)buf_mem.p;
    for (size_t i = 0; i < items.size(); ++i) {
        items[i].corr = buf + 2 * i * D;
        items[i].raylen = buf + (2 * i + 1) * D;
    }
    HIPCHK(hipMemsetAsync(buf, 0, 2 * items.size() * D * sizeof(float), ctx->stream));
    int rc = ctx->upload_scratch(items);
    if (rc) return rc;
    OwnedEvent e0, e1;
    HIPCHK(hipEventCreate(&e0.e));
    HIPCHK(hipEventCreate(&e1.e));
    auto sweep = [&]() -> hipError_t {
        for (int k = 0; k < nk; ++k) {
            const WorkItem* it = ctx->items_scratch + off[k];
            hipError_t e = launch_transpose_xz(it, cnt[k], ctx->vg, ctx->stream);
            if (e == hipSuccess) e = launch_fp(FP_MODE_SART, it, cnt[k], ctx->vg, ctx->pc.factor, ctx->stream);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    double best_ms = 1e30;
    *best_index = 0;
    for (int i = 0; i < n; ++i) {
        set(i);
        HIPCHK(sweep());  // warm-up
        HIPCHK(hipEventRecord(e0.e, ctx->stream));
        for (int rep = 0; rep < 2; ++rep) HIPCHK(sweep());
        HIPCHK(hipEventRecord(e1.e, ctx->stream));
        HIPCHK(hipEventSynchronize(e1.e));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0.e, e1.e));
        const double per = t / (2.0 * nk);
        if (ms) ms[i] = per;
        if (per < best_ms) {
            best_ms = per;
            *best_index = i;
        }
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

const int kProjectorBoxKib[STT_PROJECTOR_BOXES] = {16, 20, 24, 28};

}  // namespace

int stt_autotune(stt_ctx* ctx, int* rays_per_row, double* ms_per_shape) {
    CTX_CHECK(ctx);
    int rc = ctx->ready_check();
    if (rc) return rc;
    const int shapes[STT_PROJECTOR_SHAPES] = {0, -32, -8, 64, 32, 16, 8};
    if ((rc = ctx->ensure_transposed())) return rc;  // the level shapes under test need the copies
    const int old = ctx->vg.fp_tile;
    int bi = 0;
    rc = time_projector_configs(ctx, STT_PROJECTOR_SHAPES, [&](int i) { ctx->vg.fp_tile = shapes[i]; }, ms_per_shape,
                                &bi);
    if (rc) {
        ctx->vg.fp_tile = old;
        return rc;
    }
    const int best = shapes[bi];
    ctx->vg.fp_tile = best;
    if (best != old) ctx->invalidate_graph();
    if (best <= 0 && (rc = ctx->drop_transposed())) return rc;
    if (rays_per_row) *rays_per_row = best;
    return STT_OK;
}

int stt_set_projector_box(stt_ctx* ctx, int kib) {
    CTX_CHECK(ctx);
    if (kib != 16 && kib != 20 && kib != 24 && kib != 28)
        return fail(STT_ERR_INVALID, "projector box must be 16, 20, 24 or 28 KiB");
    if (ctx->vg.fp_box_kib != kib) {
        ctx->vg.fp_box_kib = kib;
        ctx->invalidate_graph();
    }
    return STT_OK;
}

int stt_autotune_box(stt_ctx* ctx, int* kib, double* ms_per_box) {
    CTX_CHECK(ctx);
    int rc = ctx->ready_check();
    if (rc) return rc;
    if (ctx->vg.fp_tile > 0) return fail(STT_ERR_INVALID, "box sizes apply to the slab projector shapes (0, -32, -8)");
    const int old = ctx->vg.fp_box_kib;
    int bi = 0;
    rc = time_projector_configs(ctx, STT_PROJECTOR_BOXES, [&](int i) { ctx->vg.fp_box_kib = kProjectorBoxKib[i]; },
                                ms_per_box, &bi);
    if (rc) {
        ctx->vg.fp_box_kib = old;
        return rc;
    }
    ctx->vg.fp_box_kib = kProjectorBoxKib[bi];
    if (ctx->vg.fp_box_kib != old) ctx->invalidate_graph();
    if (kib) *kib = ctx->vg.fp_box_kib;
    return STT_OK;
}

int stt_view_rounds(const float* start_deg, int nrounds, int views_per_round, float* q_xyzw) {
    if (!start_deg || !q_xyzw || nrounds < 1 || views_per_round < 1)
        return fail(STT_ERR_INVALID, "need start angles, rounds >= 1, views per round >= 1");
    // InitViews (ST_Tomography.cpp:887-934): each round starts from the
    // identity rotated by its start angle; every further view of the round
    // rotates the previous one by 360/views_per_round degrees (RotateVolume,
    // degreeEachProj = 360.0/projsEachRound, :140-141, :865-884).
    const double step = 360.0 / views_per_round;
    for (int j = 0; j < nrounds; ++j) {
        float* q = q_xyzw + 4 * (size_t)j * views_per_round;
        quat_about_y(start_deg[j], q);
        for (int k = 1; k < views_per_round; ++k) quat_rotate_y(q + 4 * (k - 1), step, q + 4 * k);
    }
    return STT_OK;
}

int stt_lds_sequence(int np, int ns, int p, int nz, float* out_deg) {
    if (!out_deg || np < 0 || ns < 1 || p < 2 || nz < 1) return fail(STT_ERR_INVALID, "need np >= 0, ns >= 1, p >= 2, nz >= 1");
    // Low_discrepency_squence.m:15-44: the radical inverse of i = 0, 1, ...
    // in base p scaled to one of ns circle segments, replicated over the ns
    // segments, with an extra 0-degree projection before every nz-th entry.
    int len = 0;
    for (long long i = 0; len < np; ++i) {
        if (len > 0 && len % nz == 0) {
            out_deg[len++] = 0.0f;
            if (len >= np) break;
        }
        double r = 0.0, q = 1.0 / p;
        for (long long b = i; b != 0; b /= p) {
            r += (double)(b % p) * q;
            q /= p;
        }
        r = 360.0 * r / ns;
        for (int k = 0; len < np && k < ns; ++k) out_deg[len++] = (float)(r + k * 360.0 / ns);
    }
    return STT_OK;
}

int stt_synchronize(stt_ctx* ctx) {
    CTX_CHECK(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_residual_norm(stt_ctx* ctx, double* r2) {
    CTX_CHECK(ctx);
    if (!r2) return fail(STT_ERR_INVALID, "null argument");
    int rc = ctx->ready_check();
    if (rc) return rc;
    const size_t D = ctx->D;
    const int F = ctx->F;
    // one projector launch per view index k over every frame's view k; the
    // line integrals and ray lengths go to scratch (no solver state changes)
    std::vector<WorkItem> items;
    std::vector<int> off, cnt;
    for (int k = 0; k < ctx->maxP; ++k) {
        off.push_back((int)items.size());
        for (int f = 0; f < F; ++f)
            if (k < ctx->P[f]) items.push_back(ctx->make_item(f, k, false));
        cnt.push_back((int)items.size() - off.back());
    }
    const int nitems = (int)items.size();
    float* ax = nullptr;
    double* partial = nullptr;
    WorkItem* dev_items = nullptr;
    StreamMem ax_mem;
    HIPCHK(ax_mem.alloc(2 * (size_t)F * D * sizeof(float), ctx->stream));
    ax = (float*)ax_mem.p;
    StreamMem partial_mem;
    HIPCHK(partial_mem.alloc(((size_t)nitems * kReductionBlocks + 1) * sizeof(double), ctx->stream));
    partial = (double*)partial_mem.p;
    StreamMem dev_items_mem;
    HIPCHK(dev_items_mem.alloc((size_t)nitems * sizeof(WorkItem), ctx->stream));
    dev_items = (WorkItem*)dev_items_mem.p;
    for (int k = 0; k < ctx->maxP; ++k)
        for (int n = 0; n < cnt[k]; ++n) {
            WorkItem& it = items[off[k] + n];
            it.ax = ax + (size_t)n * D;
            it.raylen = ax + ((size_t)F + n) * D;
            it.weight = nullptr;
        }
    HIPCHK(hipMemcpyAsync(dev_items, items.data(), (size_t)nitems * sizeof(WorkItem), hipMemcpyHostToDevice, ctx->stream));
    for (int k = 0; k < ctx->maxP; ++k) {
        HIPCHK(launch_fp(FP_MODE_FP, dev_items + off[k], cnt[k], ctx->vg, 1.0f, ctx->stream));
        for (int n = 0; n < cnt[k]; ++n) {
            const WorkItem& it = items[off[k] + n];
            // block-major: block t of every view at partial[t * nitems + view]
            HIPCHK(launch_residual_partial(it.b, it.ax, (int)D, partial + (off[k] + n), nitems, ctx->stream));
        }
    }
    double* out = partial + (size_t)nitems * kReductionBlocks;
    HIPCHK(launch_sum_partials(partial, nitems * kReductionBlocks, out, ctx->stream));
    HIPCHK(hipMemcpyAsync(r2, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_residual_norm_noalloc(stt_ctx* ctx, double* r2) {
    CTX_CHECK(ctx);
    if (!r2) return fail(STT_ERR_INVALID, "null argument");
    int rc = ctx->ready_check();
    if (rc) return rc;
    if (ctx->D > kNoAllocMaxDet)
        return fail(STT_ERR_INVALID, "stt_residual_norm_noalloc: detector of " + std::to_string(ctx->D) +
                                         " pixels exceeds its scratch limit of " + std::to_string(kNoAllocMaxDet));
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != ctx->device)
        return fail(STT_ERR_STATE, "stt_residual_norm_noalloc: make the context's device " +
                                       std::to_string(ctx->device) + " current");
    char* base = nullptr;
    HIPCHK((hipError_t)noalloc_scratch(&base));
    auto* item = (WorkItem*)(base + kNoAllocItemOff);
    auto* part = (double*)(base + kNoAllocPartOff);
    auto* acc = (double*)(base + kNoAllocAccOff);
    auto* out = (double*)(base + kNoAllocOutOff);
    auto* ax = (float*)(base + kNoAllocAxOff);
    auto* rl = (float*)(base + kNoAllocRlOff);
    const size_t D = ctx->D;
    // The views in stt_residual_norm's order (view index k, then frame), one
    // at a time: projection into scratch, its partials, then each block's
    // running sum -- the additions k_sum_partials makes over the block-major
    // partials of all views, in the same order, so *r2 is the same double.
    HIPCHK(launch_zero_f64(acc, kReductionBlocks, ctx->stream));
    for (int k = 0; k < ctx->maxP; ++k)
        for (int f = 0; f < ctx->F; ++f) {
            if (k >= ctx->P[f]) continue;
            const float* b = ctx->b[f] + (size_t)k * D;
            HIPCHK(launch_item_make(item, ctx->views_dev + ctx->view_off[f] + k, ctx->x[f], ctx->xT[f], b, ax, rl,
                                    ctx->stream));
            HIPCHK(launch_fp(FP_MODE_FP, item, 1, ctx->vg, 1.0f, ctx->stream));
            HIPCHK(launch_residual_partial(b, ax, (int)D, part, 1, ctx->stream));
            HIPCHK(launch_add_partials(acc, part, kReductionBlocks, ctx->stream));
        }
    HIPCHK(launch_sum_partials(acc, kReductionBlocks, out, ctx->stream));
    HIPCHK(hipMemcpyAsync(r2, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return STT_OK;
}

int stt_power_iteration(stt_ctx* ctx, int iters, double* L) {
    CTX_CHECK(ctx);
    if (!L || iters < 1) return fail(STT_ERR_INVALID, "need iters >= 1 and an output");
    const int F = ctx->F;
    const size_t V = ctx->V, n = (size_t)F * V;
    float *v = nullptr, *y = nullptr;
    double* red = nullptr;
    StreamMem v_mem;
    HIPCHK(v_mem.alloc(2 * n * sizeof(float), ctx->stream));
    v = (float*)v_mem.p;
    StreamMem y_mem;
    HIPCHK(y_mem.alloc(4 * n * sizeof(float), ctx->stream));
    y = (float*)y_mem.p;
    StreamMem red_mem;
    HIPCHK(red_mem.alloc((kReductionBlocks + 1) * sizeof(double), ctx->stream));
    red = (double*)red_mem.p;
    float* w = v + n;
    double* nrm2 = red + kReductionBlocks;
    auto norm2 = [&](const float* a) -> hipError_t {
        hipError_t e = launch_dot_partial(a, a, n, red, ctx->stream);
        if (e == hipSuccess) e = launch_sum_partials(red, kReductionBlocks, nrm2, ctx->stream);
        return e;
    };
    HIPCHK(launch_fill_hash(v, n, 0x5EEDu, ctx->stream));
    HIPCHK(norm2(v));
    HIPCHK(launch_scale_by_inv_sqrt(v, n, nrm2, ctx->stream));
    for (int it = 0; it < iters; ++it) {
        HIPCHK(launch_pi_apply_k(v, y, F, ctx->vg, ctx->pc, ctx->stream));
        HIPCHK(launch_pi_apply_kt(y, w, F, ctx->vg, ctx->stream));
        HIPCHK(norm2(w));  // ||K^T K v||^2 with ||v|| = 1
        HIPCHK(launch_scale_by_inv_sqrt(w, n, nrm2, ctx->stream));
        std::swap(v, w);
    }
    double lam2 = 0.0;
    HIPCHK(hipMemcpyAsync(&lam2, nrm2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *L = std::sqrt(std::sqrt(lam2));  // lambda_max(K^T K) = ||K^T K v||; ||K|| = sqrt of it
    return STT_OK;
}

void stt_of_params_reference(stt_of_params* p) {
    if (!p) return;
    // ST_Tomography.cpp:443-456
    p->tol = 0.1f;
    p->lambda = 1.2f;
    p->huber = 1.0f;
    p->eta = 0.65f;
    p->sigma = 0.5f;
    p->max_iter = 3000;
    p->norm = 4;
    p->warps = 3;
    p->levels = 3;
    const int cut[6] = {8, 8, 8, 45, 8, 8};  // the global cut[] Render passes (ST_Tomography.cpp:114, :455)
    for (int k = 0; k < 6; ++k) p->cut[k] = cut[k];
}

namespace {
int check_dims(const int* d, int lo) {
    for (int c = 0; c < 3; ++c)
        if (d[c] < lo || d[c] > 4096) return fail(STT_ERR_INVALID, "each dimension must be in [" + std::to_string(lo) + ", 4096]");
    if ((size_t)d[0] * d[1] * d[2] * 16 * sizeof(float) >= (size_t)1 << 32)
        return fail(STT_ERR_INVALID, "a volume must stay below 2^26 voxels");
    return STT_OK;
}
int check_device() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(STT_ERR_NODEVICE, "no HIP device");
    return STT_OK;
}
int check_of_params(const stt_of_params* p) {
    if (!p) return fail(STT_ERR_INVALID, "null flow parameters");
    if (p->warps < 0 || p->max_iter < 0 || !(p->lambda > 0.0f) || p->levels < 1 || !(p->eta > 0.0f && p->eta <= 1.0f))
        return fail(STT_ERR_INVALID, "bad flow parameters");
    for (int k = 0; k < 6; ++k)
        if (p->cut[k] < 0) return fail(STT_ERR_INVALID, "ROI margins must be >= 0");
    return STT_OK;
}
}  // namespace

int stt_blur(const int dims[3], int spectrum, const float* img, float sigma, float* out) {
    if (!dims || !img || !out || spectrum < 1) return fail(STT_ERR_INVALID, "bad argument");
    int rc = check_dims(dims, 1);
    if (!rc) rc = check_device();
    if (rc) return rc;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    OwnedStream st;
    HIPCHK(hipStreamCreate(&st.s));
    DevMem buf;
    HIPCHK(hipMalloc(&buf.p, (2 * (size_t)spectrum + 1) * n * sizeof(float)));
    float* din = (float*)buf.p;
    float* dout = din + spectrum * n;
    HIPCHK(hipMemcpyAsync(din, img, spectrum * n * sizeof(float), hipMemcpyHostToDevice, st.s));
    rc = blur_device(din, dout, dims, spectrum, sigma, dout + spectrum * n, st.s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, dout, spectrum * n * sizeof(float), hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipStreamSynchronize(st.s));
    return STT_OK;
}

int stt_resize(const int dims[3], int spectrum, const float* img, const int new_dims[3], float* out) {
    if (!dims || !new_dims || !img || !out || spectrum < 1) return fail(STT_ERR_INVALID, "bad argument");
    int rc = check_dims(dims, 2);
    if (!rc) rc = check_dims(new_dims, 2);
    if (!rc) rc = check_device();
    if (rc) return rc;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2], nn = (size_t)new_dims[0] * new_dims[1] * new_dims[2];
    OwnedStream st;
    HIPCHK(hipStreamCreate(&st.s));
    DevMem buf;
    HIPCHK(hipMalloc(&buf.p, (spectrum * (n + nn) + 2 * std::max(n, nn)) * sizeof(float)));
    float* din = (float*)buf.p;
    float* dout = din + spectrum * n;
    HIPCHK(hipMemcpyAsync(din, img, spectrum * n * sizeof(float), hipMemcpyHostToDevice, st.s));
    rc = resize_device(din, dims, spectrum, new_dims, dout, dout + spectrum * nn, st.s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, dout, spectrum * nn * sizeof(float), hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipStreamSynchronize(st.s));
    return STT_OK;
}

int stt_of_multiscale(const int dims[3], const float* I1, const float* I2, float* V, const stt_of_params* p,
                      int* iters) {
    if (!dims || !I1 || !I2 || !V) return fail(STT_ERR_INVALID, "null argument");
    int rc = check_dims(dims, 4);
    if (!rc) rc = check_of_params(p);
    if (!rc) rc = check_device();
    if (rc) return rc;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    OwnedStream st;
    HIPCHK(hipStreamCreate(&st.s));
    DevMem buf;
    HIPCHK(hipMalloc(&buf.p, 5 * n * sizeof(float)));
    float* d1 = (float*)buf.p;
    float* d2 = d1 + n;
    float* dv = d2 + n;
    HIPCHK(hipMemcpyAsync(d1, I1, n * sizeof(float), hipMemcpyHostToDevice, st.s));
    HIPCHK(hipMemcpyAsync(d2, I2, n * sizeof(float), hipMemcpyHostToDevice, st.s));
    HIPCHK(hipMemcpyAsync(dv, V, 3 * n * sizeof(float), hipMemcpyHostToDevice, st.s));
    rc = of_multiscale_device(dims, d1, d2, dv, *p, st.s, iters);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(V, dv, 3 * n * sizeof(float), hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipStreamSynchronize(st.s));
    return STT_OK;
}

int stt_optical_flow(stt_ctx* ctx, int pair, const stt_of_params* p, int* iters) {
    CTX_CHECK(ctx);
    if (pair < 0 || pair + 1 >= ctx->F) return fail(STT_ERR_INVALID, "frame pair out of range");
    int rc = check_of_params(p);
    if (rc) return rc;
    const int dims[3] = {ctx->g.vol[0], ctx->g.vol[1], ctx->g.vol[2]};
    rc = check_dims(dims, 4);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->flow_changed(pair);  // also when the solve fails part-way: the buffer may hold a partial flow
    return of_multiscale_device(dims, ctx->x[pair], ctx->x[pair + 1], ctx->flow[pair], *p, ctx->stream, iters);
}

int stt_of_level(const int dims[3], const float ext[3], const float* I1, const float* I2, float* V, float* Y,
                 const stt_of_params* p, const int roi[6], int* iters, double* residual) {
    if (!dims || !ext || !I1 || !I2 || !V || !Y || !p || !roi) return fail(STT_ERR_INVALID, "null argument");
    for (int c = 0; c < 3; ++c) {
        if (dims[c] < 4 || dims[c] > 4096) return fail(STT_ERR_INVALID, "each dimension must be in [4, 4096]");
        if (!(ext[c] > 0.0f)) return fail(STT_ERR_INVALID, "extent must be positive");
    }
    for (int k = 0; k < 6; ++k)
        if (roi[k] < 0) return fail(STT_ERR_INVALID, "ROI margins must be >= 0");
    if (p->warps < 0 || p->max_iter < 0 || !(p->lambda > 0.0f)) return fail(STT_ERR_INVALID, "bad flow parameters");
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    if (n * 16 * sizeof(float) >= (size_t)1 << 32) return fail(STT_ERR_INVALID, "a level must stay below 2^26 voxels");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(STT_ERR_NODEVICE, "no HIP device");
    VolGeom g{};
    g.nx = dims[0];
    g.ny = dims[1];
    g.nz = dims[2];
    OwnedStream st;
    HIPCHK(hipStreamCreate(&st.s));
    DevMem buf;
    HIPCHK(hipMalloc(&buf.p, 15 * n * sizeof(float)));
    float* dI1 = (float*)buf.p;
    float* dI2 = dI1 + n;
    float* dV = dI2 + n;
    float* dY = dV + 3 * n;
    HIPCHK(hipMemcpyAsync(dI1, I1, n * sizeof(float), hipMemcpyHostToDevice, st.s));
    HIPCHK(hipMemcpyAsync(dI2, I2, n * sizeof(float), hipMemcpyHostToDevice, st.s));
    HIPCHK(hipMemcpyAsync(dV, V, 3 * n * sizeof(float), hipMemcpyHostToDevice, st.s));
    HIPCHK(hipMemcpyAsync(dY, Y, 10 * n * sizeof(float), hipMemcpyHostToDevice, st.s));
    const int rc = of_level_device(g, ext, dI1, dI2, dV, dY, *p, roi, st.s, iters, residual);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(V, dV, 3 * n * sizeof(float), hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipMemcpyAsync(Y, dY, 10 * n * sizeof(float), hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipStreamSynchronize(st.s));
    return STT_OK;
}

}  // extern "C"
