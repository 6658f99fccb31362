// stt_device.hpp — records shared by the host library and the HIP kernels.
//
// Everything the kernels need about one projection sits in a 256-byte
// ViewRec built once per view on the host; a launch covers several frames
// (the batched form) through a table of WorkItems, one per frame taking part.
#pragma once
#include <cstdint>

namespace stt {

// One projection's cone geometry, in the precision the reference uses.
// Matrices are column-major (Matrix4x4::ele order).
struct alignas(256) ViewRec {
    float P2M[16];     // pixel -> model (TypicalRayProducer.cpp:204-205)
    // Voxel BP fast path: rows 0 and 1 of MV*V2M (ST_Tomography.cpp:699) scaled
    // by sdd/ds, row 2 unscaled, then halfW/ds and halfH/ds, so a voxel's
    // detector position is px = bp[12] - (row0.v)/(row2.v) (and py alike).
    float bp[16];
    float eye[4];      // EyePos (TypicalRayProducer.core.h:255-258)
    float o[4];        // voxel_zero = M2V * EyePos
    float m2v[4];      // diagonal of M2V (1/spacing)
    // Voxel BP exact path (detector-edge decisions): rows 0..2 of MV*V2M as the
    // reference multiplies them, float m_sdd, double m_halfFrustumWidth/Height
    // and m_ds (Scene.h:111-116).
    float mvv2m[12];
    float sdd;
    // Plane marching of the projector: bits 0-1 dominant axis D of the view's
    // central ray (0 x, 1 y, 2 z), bit 2 set when that ray runs towards -D,
    // bit 3 set when the view samples the transposed (z-fastest) volume copy.
    int march;
    double halfW, halfH, ds;
};
static_assert(sizeof(ViewRec) == 256, "ViewRec must stay one 256-byte record");

// Static per-context constants passed by value to every kernel.
struct VolGeom {
    int nx, ny, nz;      // volume dims
    int W, H;            // detector
    float sd;            // sample distance ||spacing|| (ST_Tomography.cpp:300)
    float bx, by, bz;    // CroppingBounds upper faces dims-1 (TypicalRayProducer.core.h:241-244)
    int max_samples;     // hard cap of the sample loop (never reached by a valid ray)
    float box_eps;       // bound (cells) on how far a sample drifts off its exact line: max_samples
                         // rounded additions at coordinates of magnitude <= 2*max(dims), plus slack
    int acc_wshift;      // ray-BP weight sums are int64 multiples of 2^-acc_wshift (see k_raybp_prep)
    int fp_tile;         // projector shape of this context (host-side launch choice): 0 = slab projector
                         // (16x16 rays per block), -32 / -8 = slab projector with 32x8 / 8x32 rays per block,
                         // 64/32/16/8 = level-synchronous projector with that many rays per wave row
    unsigned vol_bytes;  // bytes of one padded volume allocation (buffer range)
    unsigned volT_bytes; // bytes of one padded transposed (z-fastest) volume
    unsigned det_bytes;  // bytes of one detector image
    int fp_box_kib;      // slab projector's staged box per block: 16, 20, 24 (default) or 28 KiB of LDS
};

// Solver constants (reference globals ST_Tomography.cpp:88-111, 219).
struct PdConst {
    float sigma, tau, lambda, ktp, ebc, huber_factor, gmax, factor;
    float alpha;
    int cut[8];
    int jtop;
    int prior;
    int interp;
    int nframes;
    int stored_traces;  // 1: FrameBufs::trace_bw/trace_fw hold every frame's PD trace end points
};

// One frame's share of a projector / backprojector launch.
struct WorkItem {
    const ViewRec* view;
    float* vol;          // x[f]
    float* volT;         // transposed copy of x[f] (z fastest) for x-dominant views
    const float* b;      // measured projection of this view (W*H)
    float* raylen;       // stale ray-length buffer (read/write)
    float* corr;         // correction image / ray-BP input (W*H)
    float* ax;           // forward projection output (stand-alone FP)
    float* weight;       // sample-count output (stand-alone FP)
    float* contrib;      // fp32 scatter target of the projector's transpose (padded V)
    unsigned long long* acc_c;  // ray-BP sums in 64-bit fixed point (two's complement, padded V):
    unsigned long long* acc_w;  //   contributions w*corr at 2^-acc_shift[0], weights at 2^-acc_wshift
    int* acc_shift;             //   this view's contribution scale, written by k_raybp_prep
    float* xprev;        // X_prev[f] (finalize step)
    float* xbar;         // X_bar[f] (finalize step)
    int frame;
    int last;            // 1: last view of the sweep -> x *= imROI and extrapolate
};

// Per-frame buffer table of the primal-dual kernels (F+1 entries).
struct FrameBufs {
    float* x;
    float* xbar;
    float* xprev;
    float* ytv;   // 3 planes
    float* ytp;
    float* ybc;
    float* flow;      // u_f, 3 planes (zero for the last frame)
    float* trace_bw;  // end points of u_f traced by +1 (dual's warp), 3 planes; null: trace in the kernel
    float* trace_fw;  // end points of u_f traced by -1 (next frame's primal warp), 3 planes; or null
};

}  // namespace stt
