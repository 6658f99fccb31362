"""The benchmark configurations as parity cases on the MI355X.

- **Configuration 1.** Plain single-frame SART+TV, with the oracle run in
  fp32 and fp64:
  - 64^3, 10-ellipsoid phantom (seed 1);
  - 64 even views over 360 degrees, 128x128 detector;
  - alpha 0.3, lambda 0.01, 20 iterations, identity view order.
- **Configuration 2.** Space-time STV at its stated size, 256^3 x 8 frames
  with 45 golden-angle views on a 512x512 detector: the plain per-frame
  form against the batched and the batched hipGraph forms (the oracle
  cannot run this size in seconds).
- **Configuration 3.** The differentiable case at its real size:
  - 128^3 x 10 frames, 32 views/frame on a 256x256 detector;
  - projector-adjoint and warp-adjoint dot products;
  - a seeded smooth flow with |u| <= 3;
  - the flow VJP against FP64 central differences at 128^3 (local
    directions: see the test).
- **Configuration 4.** Uneven batched solves: 32 frames with seeded view
  counts U{8..64} (seed 7):
  - one batched hipGraph ST iteration against the oracle, with the volume
    reduced to 40^3 (384x384 detector kept);
  - the stated mix, 19 frames of 128^3 and 13 of 192^3 (seed 8): one group
    iteration against each context alone, and two frames against the oracle.

  Mixed 128^3/192^3 volumes run as one context per size in one stt_group
  (one hipGraph per group iteration). A context holds one grid, because the
  PD coupling warps frame f+1 onto frame f.
- **Configuration 5.** 512^3 x 16 frames, 1024x1024, 60 van der Corput views,
  k = 0.1, e = 0.05, FP64 reductions, at its stated size: slab projector
  against the oracle on two views, batched eager against batched hipGraph
  bit for bit, bounds; and its solver at 96^3 x 4 frames against the FP64
  oracle.
"""
import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from helpers import rel_l2, smooth_flow

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["compat", "spec-shuffled"])
def test_config1_plain_sart_tv_20_iterations(mode):
    """Configuration 1: 64^3, 64 even views about +y, 128x128, SART + TV, 20
    iterations against the fp32 and FP64 oracles.  "compat": the reference's
    effective solver constants (sigma 0.1125, tau 1) and the identity view
    order; "spec-shuffled": sigma = tau = 0.2 and the views swept in a
    seeded shuffle (seed 3)."""
    print(f"[config1] {mode}: 64^3, 64 even views, 128x128, alpha 0.3, lambda 0.01, 20 iterations", flush=True)
    dims, det = (64, 64, 64), (128, 128)
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 600.0, 1000.0)
    p = stt.params_default()
    # k = e = 0 (no temporal terms) and huber 0.1 (main.cpp:73)
    sigma, tau = (0.1125, 1.0) if mode == "compat" else (0.2, 0.2)
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = sigma, tau, 0.01, 0.0, 0.0, 0.1, 0.3
    P = 64
    ang = list(stt.synth_angles(0, P))
    if mode != "compat":  # the sweep follows the view table: a shuffled table is a shuffled sweep
        ang = [ang[k] for k in np.random.default_rng(3).permutation(P)]
    truth = stt.synth_phantom(dims, 10, seed=1)
    ctx = stt.Context(g, p, [P])
    ctx.set_views(0, ang)
    ctx.set_volume(0, truth)
    b = np.stack([ctx.forward_project(0, k)[0] for k in range(P)])
    zero = np.zeros_like(truth)
    ctx.set_volume(0, zero)
    ctx.set_volume(1, zero)
    ctx.set_projections(0, b)
    ocs = []
    for f64 in (False, True):
        oc = O.OracleCtx(g, p, [P], f64mode=f64)
        oc.set_views(0, ang)
        oc.set_volume(0, zero)
        oc.set_volume(1, zero)
        oc.set_projections(0, b.reshape(-1))
        ocs.append(oc)
    ctx.pd_init()
    ctx.iterate(20, batched=False)
    got = ctx.get_volume(0)
    for oc in ocs:
        oc.pd_init()
        oc.iterate(20)
    ref32, ref64 = ocs[0].get_volume(0), ocs[1].get_volume_f64(0)
    e32, e64 = rel_l2(got, ref32), rel_l2(got, ref64)
    print(f"  relL2 vs fp32 oracle {e32:.2e}, vs fp64 oracle {e64:.2e}, fp32 oracle vs fp64 {rel_l2(ref32, ref64):.2e}, "
          f"reconstruction error vs phantom {rel_l2(got, truth):.3f}", flush=True)
    assert e32 <= 1e-4, e32
    assert rel_l2(got, truth) < rel_l2(zero, truth)  # it reconstructs
    ctx.close()


def test_config2_plain_vs_batched_at_full_size():
    """One ST iteration at 256^3 x 8 frames (45 golden-angle views per frame,
    512x512 detector, STV, sigma = tau = 0.2, voxel BP, clamped-cubic warps
    of a smooth flow of amplitude 4 voxels) in the plain per-frame form, the
    batched form and the batched hipGraph form.  What holds at any size:
    batched and hipGraph are the same launches, so bit-identical; plain
    differs only in which frame's stale ray length a missed ray reuses; every
    voxel stays in [0, G_Max]."""
    print("[config2] 256^3 x 8 frames, 45 views, 512x512: plain vs batched vs hipGraph, one iteration", flush=True)
    dims, det, F, P = (256, 256, 256), (512, 512), 8, 45
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 600.0, 1000.0)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01, 0.3
    p.prior = stt.STT_PRIOR_STV
    p.bp = stt.STT_BP_VOXEL
    p.interp = stt.STT_INTERP_CUBIC_CLAMPED
    truth = [stt.synth_phantom_frame(dims, f, n_ellipsoids=10, seed=1, deform_seed=2, amp=4.0) for f in range(F)]
    flows = [stt.synth_flow(dims, f, deform_seed=2, amp=4.0) for f in range(F - 1)]
    angles = [stt.synth_angles(1, P, offset=P * f) for f in range(F)]
    x0 = [(0.5 * t).astype(np.float32) for t in truth]
    ctx = stt.Context(g, p, [P] * F)
    b = []
    for f in range(F):
        ctx.set_views(f, angles[f])
        ctx.set_volume(f, truth[f])
        b.append(np.concatenate([ctx.forward_project(f, k)[0] for k in range(P)]))
    ctx.close()
    out, gmax = {}, None
    for form, batched, graph in (("plain", False, False), ("batched", True, False), ("graph", True, True)):
        ctx = stt.Context(g, p, [P] * F)
        for f in range(F):
            ctx.set_views(f, angles[f])
            ctx.set_projections(f, b[f])
            ctx.set_volume(f, x0[f])
        ctx.set_volume(F, x0[F - 1])
        for f in range(F - 1):
            ctx.set_flow(f, flows[f])
        ctx.iterate(1, batched=batched, use_graph=graph)
        out[form] = [ctx.get_volume(f) for f in range(F)]
        gmax = ctx.gmax()
        ctx.close()
        print(f"[config2] {form} done", flush=True)
    for f in range(F):
        assert np.array_equal(out["graph"][f], out["batched"][f]), f
        assert rel_l2(out["batched"][f], out["plain"][f]) <= 1e-6, (f, rel_l2(out["batched"][f], out["plain"][f]))
        v = out["graph"][f]
        assert np.isfinite(v).all() and v.min() >= 0.0 and v.max() <= gmax
        assert not np.array_equal(v, x0[f])  # the iteration moved the volume


def test_config3_adjoint_dot_products_at_128():
    print("[config3] 128^3 x 10 frames, 32 views, 256x256: FP adjoint and warp adjoint dot tests", flush=True)
    dims, det = (128, 128, 128), (256, 256)
    V = int(np.prod(dims))
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 600.0, 1000.0)
    F, P = 10, 32
    ctx = stt.Context(g, stt.params_default(), [P] * F)
    rng_x, rng_y = np.random.default_rng(5), np.random.default_rng(6)
    worst = 0.0
    for f in (0, 9):  # first and last frame of the stack, 4 views each
        ctx.set_views(f, list(stt.synth_angles(1, P, offset=P * f)))
        x = rng_x.standard_normal(V).astype(np.float32)
        ctx.set_volume(f, x)
        for k in range(0, P, 8):
            ax = ctx.forward_project(f, k)[0]
            y = rng_y.standard_normal(det[0] * det[1]).astype(np.float32)
            aty = ctx.fp_vjp(f, k, y)
            lhs = float(np.dot(ax.astype(np.float64), y))
            rhs = float(np.dot(x.astype(np.float64), aty))
            rel = abs(lhs - rhs) / (np.linalg.norm(ax) * np.linalg.norm(y))
            worst = max(worst, rel)
    print(f"  projector: worst |<Ax,y> - <x,A^T y>| / (|Ax||y|) = {worst:.2e}", flush=True)
    assert worst <= 1e-6
    uvw = smooth_flow(dims, 3.0, seed=4)
    src = rng_x.standard_normal(V).astype(np.float32)
    gout = rng_y.standard_normal(V).astype(np.float32)
    for dt in (-1.0, 1.0):
        wsrc = ctx.warp(uvw, src, dt, stt.STT_INTERP_TRILINEAR)
        gsrc, _ = ctx.warp_vjp(uvw, src, dt, gout)
        lhs = float(np.dot(wsrc.astype(np.float64), gout))
        rhs = float(np.dot(src.astype(np.float64), gsrc))
        rel = abs(lhs - rhs) / (np.linalg.norm(wsrc) * np.linalg.norm(gout))
        print(f"  warp dt={dt}: |<Wx,g> - <x,W^T g>| / (|Wx||g|) = {rel:.2e}", flush=True)
        assert rel <= 1e-6
    ctx.close()


def test_config4_uneven_batched_iteration():
    rng = np.random.default_rng(7)
    F = 32
    P = [int(v) for v in rng.integers(8, 65, F)]
    dims, det = (40, 40, 40), (384, 384)
    print(f"[config4] F={F} views U{{8..64}} (total {sum(P)}), {dims}, {det}, batched + hipGraph", flush=True)
    g = stt.Geometry.make(dims, det, (1, 1, 1), 0.25, 600.0, 1000.0)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01, 0.3
    ctx = stt.Context(g, p, P)
    oc = O.OracleCtx(g, p, P)
    zero = np.zeros(int(np.prod(dims)), np.float32)
    for f in range(F):
        ang = list(stt.synth_angles(2, P[f], offset=sum(P[:f])))
        truth = stt.synth_phantom_frame(dims, f, n_ellipsoids=6, seed=8, deform_seed=2, amp=1.5)
        ctx.set_views(f, ang)
        oc.set_views(f, ang)
        ctx.set_volume(f, truth)
        b = np.stack([ctx.forward_project(f, k)[0] for k in range(P[f])])
        ctx.set_projections(f, b)
        oc.set_projections(f, b.reshape(-1))
        if f < F - 1:
            u = stt.synth_flow(dims, f, deform_seed=2, amp=1.5)
            ctx.set_flow(f, u)
            oc.set_flow(f, u)
    for f in range(F + 1):
        ctx.set_volume(f, zero)
        oc.set_volume(f, zero)
    ctx.pd_init()
    oc.pd_init()
    ctx.autotune()
    ctx.iterate(1, batched=True, use_graph=True)
    oc.per_frame_raylen(True)
    oc.iterate(1)
    worst = max(rel_l2(ctx.get_volume(f), oc.get_volume(f)) for f in range(F))
    print(f"  worst frame relL2 vs oracle {worst:.2e}", flush=True)
    assert worst <= 1e-4
    ctx.close()


def _mixed_context(dims, P, seed, det=(96, 80)):
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 600.0, 1000.0)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.01, 0.3
    p.prior = stt.STT_PRIOR_ATV  # SART(1 sweep) + PD ATV, k = e = 0: independent solves
    ctx = stt.Context(g, p, P)
    zero = np.zeros(int(np.prod(dims)), np.float32)
    for f in range(len(P)):
        ctx.set_views(f, list(stt.synth_angles(2, P[f], offset=7 * f)))
        ctx.set_volume(f, stt.synth_phantom_frame(dims, f, n_ellipsoids=5, seed=seed, deform_seed=2, amp=1.0))
        ctx.set_projections(f, np.stack([ctx.forward_project(f, k)[0] for k in range(P[f])]))
    for f in range(len(P) + 1):
        ctx.set_volume(f, zero)
    return ctx


@pytest.mark.parametrize("use_graph", [True, False])
def test_config4_mixed_sizes_in_one_group(use_graph):
    # frames of two grids (scaled-down 128^3 / 192^3 mix) with uneven views,
    # one context per grid, iterated as one group
    shapes = [((24, 24, 24), [5, 9, 7]), ((36, 36, 36), [8, 6])]
    print(f"[config4-mixed] {shapes} use_graph={use_graph}", flush=True)
    grouped = [_mixed_context(d, P, seed=3 + i) for i, (d, P) in enumerate(shapes)]
    alone = [_mixed_context(d, P, seed=3 + i) for i, (d, P) in enumerate(shapes)]
    with stt.Group(grouped) as grp:
        ms = grp.iterate(2, use_graph=use_graph)
        ms2 = grp.iterate(1, use_graph=use_graph)  # a replay of the same graph
    for ctx in alone:
        ctx.iterate(3, batched=True, use_graph=use_graph)
    print(f"  group: 2 iterations {ms:.2f} ms, then 1 in {ms2:.2f} ms", flush=True)
    for (d, P), a, b in zip(shapes, grouped, alone):
        for f in range(len(P)):
            assert np.array_equal(a.get_volume(f), b.get_volume(f)), (d, f)
        assert not np.array_equal(a.get_volume(0), np.zeros(int(np.prod(d)), np.float32))
    for c in grouped + alone:
        c.close()


def _loss64(dims, uvw, src, dt, gout):
    out = O.warp(dims, uvw.astype(np.float64), src, dt, interp=stt.STT_INTERP_TRILINEAR, f64mode=True)
    return float(np.dot(out, gout.astype(np.float64)))


@pytest.mark.parametrize("dt", [-1.0, 1.0])
def test_config3_flow_vjp_finite_differences_at_128(dt):
    """The warp's flow VJP at configuration 3's size (128^3, seeded smooth
    flow |u| <= 3, seed 4) against FP64 central differences of the oracle
    warp.  Trilinear sampling has kinks at cell faces; a perturbation of the
    whole field moves millions of samples, some across a face (and the GPU's
    fp32 trace puts about 1e-5 of them on the other side of a face than the
    FP64 trace), which shifts a whole-field difference quotient by about 1e-3
    at this size whatever the step (measured: 2e-7 at 24^3, 1e-3 at 48^3 to
    128^3, also between steps 1e-6 and 1e-4 in FP64 alone).  So the checks
    are local: directional derivatives along 8 smooth bumps of the flow
    (h = 1e-6, a few thousand samples each) and 64 single coordinates
    (h = 1e-3), at most two of each may land on a kink."""
    print(f"[config3-fd] 128^3 dt={dt}: flow VJP vs FP64 central differences", flush=True)
    dims = (128, 128, 128)
    V = int(np.prod(dims))
    g = stt.Geometry.make(dims, (16, 16), (1, 1, 1), 1.0, 600.0, 1000.0)
    ctx = stt.Context(g, stt.params_default(), [1])
    rng = np.random.default_rng(40)
    uvw = (smooth_flow(dims, 3.0, seed=4) * np.float32(0.9731)
           + np.float32(1e-3) * rng.standard_normal(3 * V).astype(np.float32))
    src = smooth_flow(dims, 1.0, seed=41)[:V] + rng.random(V, dtype=np.float32)
    gout = rng.standard_normal(V).astype(np.float32)
    _, guvw = ctx.warp_vjp(uvw, src, dt, gout)
    zz, yy, xx = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in dims[::-1]), indexing="ij")
    hd = 1e-6
    berrs = []
    for _ in range(8):
        c = rng.uniform(8, 120, 3)
        bump = np.exp(-((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) / (2 * 2.0 ** 2)).reshape(-1)
        bump[bump < 1e-6] = 0.0
        w = rng.standard_normal(3)
        delta = np.concatenate([w[0] * bump, w[1] * bump, w[2] * bump])
        fd = (_loss64(dims, uvw + hd * delta, src, dt, gout) - _loss64(dims, uvw - hd * delta, src, dt, gout)) / (2 * hd)
        an = float(np.dot(guvw.astype(np.float64), delta))
        berrs.append(abs(fd - an) / max(abs(fd), 1e-3))
    berrs = np.array(berrs)
    print(f"  8 bumps: rel errors {np.array2string(berrs, precision=2)}", flush=True)
    assert np.count_nonzero(berrs > 1e-4) <= 2, berrs
    h = 1e-3
    coords = rng.choice(3 * V, 64, replace=False)
    errs = []
    for c in coords:
        u_p = uvw.astype(np.float64)
        u_m = u_p.copy()
        u_p[c] += h
        u_m[c] -= h
        fdc = (_loss64(dims, u_p, src, dt, gout) - _loss64(dims, u_m, src, dt, gout)) / (2 * h)
        errs.append(abs(fdc - guvw[c]) / max(1.0, abs(fdc)))
    errs = np.array(errs)
    print(f"  64 coordinates: median rel {np.median(errs):.2e}, max {errs.max():.2e}", flush=True)
    assert np.count_nonzero(errs > 1e-3) <= 2, errs
    ctx.close()


def _config4_contexts():
    """Configuration 4 as stated: 32 frames, view counts U{8..64} (seed 7),
    sizes drawn from {128^3, 192^3} (seed 8), 384x384 detector, van der
    Corput views, SART (one sweep) + PD ATV with k = e = 0 (independent
    solves); one context per size.  Returns the views per size and, per
    size, the inputs to build identical contexts."""
    F = 32
    P = [int(v) for v in np.random.default_rng(7).integers(8, 65, F)]
    sizes = [int(v) for v in np.random.default_rng(8).choice([128, 192], F)]
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.0, 0.0, 0.01, 0.3
    p.prior = stt.STT_PRIOR_ATV
    specs = []
    for n in sorted(set(sizes)):
        idx = [f for f in range(F) if sizes[f] == n]
        dims = (n, n, n)
        g = stt.Geometry.make(dims, (384, 384), (1, 1, 1), 1.0, 600.0, 1000.0)
        Pn = [P[f] for f in idx]
        ang = [list(stt.synth_angles(2, P[f], offset=sum(P[:f]))) for f in idx]
        ctx = stt.Context(g, p, Pn)
        b = []
        for j, f in enumerate(idx):
            ctx.set_views(j, ang[j])
            ctx.set_volume(j, stt.synth_phantom_frame(dims, f, n_ellipsoids=10, seed=1, deform_seed=2, amp=3.0))
            b.append(np.stack([ctx.forward_project(j, k)[0] for k in range(P[f])]))
        ctx.close()
        specs.append(dict(dims=dims, g=g, P=Pn, ang=ang, b=b))
    return p, specs


def _config4_build(p, s):
    ctx = stt.Context(s["g"], p, s["P"])
    zero = np.zeros(int(np.prod(s["dims"])), np.float32)
    for j in range(len(s["P"])):
        ctx.set_views(j, s["ang"][j])
        ctx.set_projections(j, s["b"][j])
    for j in range(len(s["P"]) + 1):
        ctx.set_volume(j, zero)
    return ctx


def test_config4_real_mix_group_and_oracle():
    """The 128^3/192^3 mix at its stated size: one group iteration (one
    hipGraph for both contexts) equals each context iterated alone, bit for
    bit; two frames of the 128^3 context equal the oracle's solve of those
    frames (the frames are independent solves: k = e = 0)."""
    p, specs = _config4_contexts()
    print("[config4-real] " + ", ".join(f"{len(s['P'])} frames of {s['dims'][0]}^3 ({sum(s['P'])} views)"
                                        for s in specs) + ", 384x384", flush=True)
    grouped = [_config4_build(p, s) for s in specs]
    alone = [_config4_build(p, s) for s in specs]
    with stt.Group(grouped) as grp:
        ms = grp.iterate(1, use_graph=True)
    print(f"  group iteration {ms:.1f} ms", flush=True)
    for c in alone:
        c.iterate(1, batched=True, use_graph=True)
    for s, a, b in zip(specs, grouped, alone):
        for j in range(len(s["P"])):
            assert np.array_equal(a.get_volume(j), b.get_volume(j)), (s["dims"], j)
    s = specs[0]
    assert s["dims"] == (128, 128, 128)
    oc = O.OracleCtx(s["g"], p, s["P"][:2])
    zero = np.zeros(int(np.prod(s["dims"])), np.float32)
    for j in range(2):
        oc.set_views(j, s["ang"][j])
        oc.set_projections(j, s["b"][j].reshape(-1))
    for j in range(3):
        oc.set_volume(j, zero)
    oc.per_frame_raylen(True)
    oc.iterate(1)
    for j in range(2):
        err = rel_l2(grouped[0].get_volume(j), oc.get_volume(j))
        print(f"  128^3 frame {j} ({s['P'][j]} views): relL2 vs oracle {err:.2e}", flush=True)
        assert err <= 1e-4, (j, err)
    for c in grouped + alone:
        c.close()


def _config5_params():
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01, 0.3
    p.prior = stt.STT_PRIOR_STV
    p.bp = stt.STT_BP_VOXEL
    p.interp = stt.STT_INTERP_CUBIC_CLAMPED
    p.fp64_reductions = 1
    return p


def test_config5_full_scale_batched_equals_graph():
    """Configuration 5 at its stated size: 512^3 fp32 x 16 frames, a 3x512^3
    flow per frame pair, 1024x1024 detector, 60 van der Corput views per
    frame, k = 0.1, e = 0.05, huber 0.01, STV, FP64 reductions (about 150 GB
    of HBM).  The slab projector's line integrals of two views equal the
    oracle's (sample counts bit for bit); one iteration in the batched eager
    form and one in the batched hipGraph form, from the same state, give
    bit-identical volumes, every voxel in [0, G_Max]."""
    dims, det, F, P = (512, 512, 512), (1024, 1024), 16, 60
    print(f"[config5] {dims} x {F} frames, {det}, {P} van der Corput views/frame", flush=True)
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 600.0, 1000.0)
    ctx = stt.Context(g, _config5_params(), [P] * F)
    V = int(np.prod(dims))
    for f in range(F):
        ang = list(stt.synth_angles(2, P, offset=P * f))
        truth = stt.synth_phantom_frame(dims, f, n_ellipsoids=10, seed=1, deform_seed=2, amp=4.0)
        ctx.set_views(f, ang)
        ctx.set_volume(f, truth)
        b = np.empty((P, det[0] * det[1]), np.float32)
        for k in range(P):
            if f == 0 and k in (0, 31):
                b[k], _, w = ctx.forward_project(f, k)
                oax, _, ow = O.fp(g, O.quat_y(ang[k]), truth)
                assert np.array_equal(w, ow), k
                err = rel_l2(b[k], oax)
                print(f"  frame 0 view {k} ({ang[k]:.2f} deg): slab vs oracle relL2 {err:.2e}", flush=True)
                assert err <= 1e-5, (k, err)
            else:
                b[k] = ctx.forward_project(f, k)[0]
        ctx.set_projections(f, b)
        if f < F - 1:
            ctx.set_flow(f, stt.synth_flow(dims, f, deform_seed=2, amp=4.0))
        print(f"  frame {f} set", flush=True)
    zero = np.zeros(V, np.float32)
    eager = []
    for graph in (False, True):
        for f in range(F + 1):
            ctx.set_volume(f, zero)
        ctx.pd_init()
        ctx.iterate(1, batched=True, use_graph=graph)
        gmax = ctx.gmax()
        for f in range(F):
            v = ctx.get_volume(f)
            assert np.isfinite(v).all() and v.min() >= 0.0 and v.max() <= gmax, f
            assert v.max() > 0.0, f
            if graph:
                assert np.array_equal(v, eager[f]), f
            else:
                eager.append(v)
        print(f"  {'hipGraph' if graph else 'eager'} iteration done", flush=True)
    ctx.close()


def test_config5_parameters_vs_fp64_oracle():
    """Configuration 5's solver (STV, k = 0.1, e = 0.05, huber 0.01, van der
    Corput views, clamped-cubic warps) at the largest size the FP64 oracle
    finishes in seconds -- 96^3 x 4 frames, 60 views/frame, 192x192
    detector -- one batched hipGraph iteration: the fp32 GPU result against
    the FP64 oracle (and the fp32 oracle against the FP64 one)."""
    dims, det, F, P = (96, 96, 96), (192, 192), 4, 60
    print(f"[config5-fp64] {dims} x {F} frames, {det}, {P} views: GPU vs FP64 oracle", flush=True)
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 600.0, 1000.0)
    p = _config5_params()
    ctx = stt.Context(g, p, [P] * F)
    ocs = [O.OracleCtx(g, p, [P] * F, f64mode=m) for m in (False, True)]
    for f in range(F):
        ang = list(stt.synth_angles(2, P, offset=P * f))
        truth = stt.synth_phantom_frame(dims, f, n_ellipsoids=10, seed=1, deform_seed=2, amp=2.0)
        ctx.set_views(f, ang)
        ctx.set_volume(f, truth)
        b = np.stack([ctx.forward_project(f, k)[0] for k in range(P)])
        ctx.set_projections(f, b)
        for oc in ocs:
            oc.set_views(f, ang)
            oc.set_projections(f, b.reshape(-1))
        if f < F - 1:
            u = stt.synth_flow(dims, f, deform_seed=2, amp=2.0)
            ctx.set_flow(f, u)
            for oc in ocs:
                oc.set_flow(f, u)
    zero = np.zeros(int(np.prod(dims)), np.float32)
    for f in range(F + 1):
        ctx.set_volume(f, zero)
        for oc in ocs:
            oc.set_volume(f, zero)
    ctx.iterate(1, batched=True, use_graph=True)
    for oc in ocs:
        oc.per_frame_raylen(True)
        oc.iterate(1)
    worst64 = worst32 = 0.0
    for f in range(F):
        got = ctx.get_volume(f)
        r64 = ocs[1].get_volume_f64(f)
        e64, e32 = rel_l2(got, r64), rel_l2(got, ocs[0].get_volume(f))
        print(f"  frame {f}: relL2 vs FP64 oracle {e64:.2e}, vs fp32 oracle {e32:.2e}, "
              f"fp32 oracle vs FP64 {rel_l2(ocs[0].get_volume(f), r64):.2e}", flush=True)
        worst64, worst32 = max(worst64, e64), max(worst32, e32)
    assert worst32 <= 1e-4 and worst64 <= 1e-4, (worst32, worst64)
    ctx.close()
