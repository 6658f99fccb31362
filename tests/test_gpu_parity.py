"""Parity of the HIP path with the CPU oracle (run on an MI355X).

Every test prints the case it runs before touching the device.  Tolerances:
integer-valued outputs (sample counts) and the ray lengths are bit-exact;
fp32 line integrals within relative L2 1e-5; volumes after one or more
sweeps / iterations within relative L2 1e-4 (FMA contraction and float-vs-
double cubic evaluation are the only sources of difference).
"""
import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from helpers import rel_l2, smooth_flow

pytestmark = pytest.mark.gpu

FP_CASES = [
    # vol, det, spacing, ds, sid, sdd, oxyz, angle
    ((24, 24, 24), (40, 36), (1.0, 1.0, 1.0), 1.0, 60.0, 100.0, (0.0, 0.0, 0.0), 0.0),
    ((24, 20, 28), (48, 40), (1.0, 1.0, 1.0), 0.8, 55.0, 110.0, (1.5, -2.0, 0.5), 37.0),
    ((20, 26, 22), (36, 44), (0.7, 1.3, 0.9), 0.6, 50.0, 95.0, (0.0, 1.0, -1.0), 215.0),
    ((32, 32, 32), (30, 30), (1.0, 1.0, 1.0), 1.0, 70.0, 120.0, (0.0, 0.0, 0.0), 90.0),   # partial misses
    ((17, 19, 23), (41, 29), (1.1, 0.9, 1.0), 1.2, 45.0, 90.0, (-1.0, 0.5, 2.0), 300.0),
]


def _ctx(vol, det, spacing=(1, 1, 1), ds=1.0, sid=60.0, sdd=100.0, oxyz=(0, 0, 0), P=(1,), params=None):
    g = stt.Geometry.make(vol, det, spacing, ds, sid, sdd, oxyz)
    p = params if params is not None else stt.params_default()
    return g, stt.Context(g, p, list(P))


PROJECTOR_SHAPES = [0, -32, -8, 64, 32, 16, 8]  # slab projector (three tiles), level-synchronous tiles


@pytest.mark.parametrize("case", range(len(FP_CASES)))
def test_fp_matches_oracle(case):
    vol, det, sp, ds, sid, sdd, oxyz, ang = FP_CASES[case]
    print(f"[fp] case={case} vol={vol} det={det} angle={ang}", flush=True)
    rng = np.random.default_rng(100 + case)
    g, ctx = _ctx(vol, det, sp, ds, sid, sdd, oxyz, P=(1,))
    x = rng.random(int(np.prod(vol)), dtype=np.float32)
    ctx.set_volume(0, x)
    ctx.set_views(0, [ang])
    rl0 = rng.random(det[0] * det[1], dtype=np.float32)  # stale values must survive on missed rays
    q = O.quat_y(ang)
    oax, orl, ow = O.fp(g, q, x, raylen=rl0)
    ax64, _, _ = O.fp(g, q, x.astype(np.float64), f64mode=True)
    try:
        for shape in PROJECTOR_SHAPES:
            print(f"  projector shape {shape}", flush=True)
            ctx.set_projector_tile(shape)
            ax, rl, w = ctx.forward_project(0, 0, raylen=rl0)
            assert np.array_equal(w, ow), f"shape {shape}: sample counts differ at {np.count_nonzero(w != ow)} rays"
            assert np.array_equal(rl, orl), f"shape {shape}: ray lengths differ"
            assert rel_l2(ax, oax) <= 1e-5, (shape, rel_l2(ax, oax))
            print(f"  relL2 vs fp32 oracle {rel_l2(ax, oax):.2e}, vs fp64 oracle {rel_l2(ax, ax64):.2e}")
    finally:
        ctx.set_projector_tile(0)
        ctx.close()


def test_fp_shapes_agree_at_full_size():
    """256^3 phantom, 512x512 detector (the iteration benchmark's geometry):
    the slab projector reads its volume from LDS boxes, the level projector
    straight from HBM.  Views around the 45-degree switch of the dominant axis
    included.  Sample counts must be identical and values agree to fp32
    rounding; two views are also checked against the oracle."""
    print("[fp-full] 256^3, 512x512, slab (16x16, 32x8, 8x32) vs level-16 vs oracle", flush=True)
    dims, det = (256, 256, 256), (512, 512)
    angles = [0.0, 17.3, 44.99, 45.0, 45.01, 90.0, 134.9, 180.0, 226.1, 271.7, 315.0, 359.9]
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 600.0, 1000.0)
    ctx = stt.Context(g, stt.params_default(), [len(angles)])
    x = stt.synth_phantom(dims, 10, seed=3)
    ctx.set_volume(0, x)
    ctx.set_views(0, angles)
    try:
        for v, a in enumerate(angles):
            ctx.set_projector_tile(16)
            ax1, rl1, w1 = ctx.forward_project(0, v)
            for slab in (-32, -8, 0):
                ctx.set_projector_tile(slab)
                ax0, rl0, w0 = ctx.forward_project(0, v)
                assert np.array_equal(w0, w1), (a, slab, np.count_nonzero(w0 != w1))
                assert np.array_equal(rl0, rl1), (a, slab)
                e = rel_l2(ax0, ax1)
                assert e <= 1e-6, (a, slab, e)
            if v in (1, 4):
                oax, _, ow = O.fp(g, O.quat_y(a), x)
                assert np.array_equal(w0, ow), a
                assert rel_l2(ax0, oax) <= 1e-5, (a, rel_l2(ax0, oax))
                print(f"  angle {a}: slab vs oracle {rel_l2(ax0, oax):.2e}", flush=True)
            print(f"  angle {a}: slab vs level {e:.2e}", flush=True)
    finally:
        ctx.set_projector_tile(0)
        ctx.close()


def test_slab_equals_level_on_fine_detector():
    """Regression: the slab projector against the level projector at 40^3 on
    a 384x384 detector of 0.25 mm pixels, where grazing rays near the
    volume's edge once sampled outside the staged box.  Sample counts and ray
    lengths bit-exact, values within 1e-6 of the level projector's."""
    print("[slab-vs-level] 40^3, 384x384 at 0.25 mm, 12 views", flush=True)
    dims, det = (40, 40, 40), (384, 384)
    g = stt.Geometry.make(dims, det, (1, 1, 1), 0.25, 600.0, 1000.0)
    ctx = stt.Context(g, stt.params_default(), [12])
    ang = list(stt.synth_angles(2, 12, offset=0))
    ctx.set_views(0, ang)
    ctx.set_volume(0, stt.synth_phantom_frame(dims, 0, n_ellipsoids=6, seed=8, deform_seed=2, amp=1.5))
    try:
        for k in range(12):
            ctx.set_projector_tile(16)
            a1, r1, w1 = ctx.forward_project(0, k)
            for slab in (0, -32, -8):
                ctx.set_projector_tile(slab)
                a0, r0, w0 = ctx.forward_project(0, k)
                assert np.array_equal(w0, w1) and np.array_equal(r0, r1), (k, slab)
                err = float(np.max(np.abs(a0 - a1)) / max(float(np.max(np.abs(a1))), 1e-30))
                assert err <= 1e-6, (k, slab, ang[k], err)
    finally:
        ctx.set_projector_tile(0)
        ctx.close()


def test_projector_box_sizes_bit_identical():
    """The slab projector's staged box size (16/20/24/28 KiB) only decides
    where a tap is read from, LDS or global memory: for every slab tile the
    forward projections (values, sample counts, ray lengths) and a SART
    sweep from the same start are bit-identical across sizes.  The box
    autotuner keeps one of the sizes and refuses a level-synchronous shape."""
    dims, det = (96, 96, 96), (192, 192)
    print(f"[fp-box] {dims}, {det}, boxes {stt.PROJECTOR_BOXES}", flush=True)
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 300.0, 500.0)
    angles = [0.0, 30.0, 44.99, 45.0, 91.0, 200.0]
    p = stt.params_default()
    p.alpha = 0.3
    ctx = stt.Context(g, p, [len(angles)])
    ctx.set_views(0, angles)
    x0 = stt.synth_phantom_frame(dims, 0, n_ellipsoids=6, seed=5, deform_seed=2, amp=2.0)
    ctx.set_volume(0, x0)
    b = np.concatenate([ctx.forward_project(0, k)[0] for k in range(len(angles))])
    ctx.set_projections(0, b * 1.05)
    try:
        for tile in (0, -32, -8):
            ctx.set_projector_tile(tile)
            ref = None
            for kib in stt.PROJECTOR_BOXES:
                ctx.set_projector_box(kib)
                ctx.set_volume(0, x0)
                out = [ctx.forward_project(0, k) for k in range(len(angles))]
                ctx.sart_sweep(0)
                swept = ctx.get_volume(0)
                if ref is None:
                    ref = (out, swept)
                    continue
                for k in range(len(angles)):
                    for a, r in zip(out[k], ref[0][k]):
                        assert np.array_equal(a, r), (tile, kib, angles[k])
                assert np.array_equal(swept, ref[1]), (tile, kib)
            print(f"  tile {tile}: projections and sweep identical across boxes", flush=True)
        ctx.set_projector_tile(0)
        kib, ms = ctx.autotune_box()
        print(f"  autotune_box -> {kib} KiB, ms per view index {[round(v, 4) for v in ms]}", flush=True)
        assert kib in stt.PROJECTOR_BOXES and all(v > 0 for v in ms)
        ctx.set_projector_tile(16)
        with pytest.raises(stt.SttError):
            ctx.autotune_box()
        with pytest.raises(stt.SttError):
            ctx.set_projector_box(32)
    finally:
        ctx.set_projector_tile(0)
        ctx.set_projector_box(24)
        ctx.close()


def test_fp_constant_volume_closed_form():
    """Constant volume c: every hit ray integrates to c * sd * (number of samples)."""
    print("[fp-const] 32^3, 48x48", flush=True)
    g, ctx = _ctx((32, 32, 32), (48, 48), P=(3,))
    c = 0.75
    ctx.set_volume(0, np.full(32 ** 3, c, np.float32))
    ctx.set_views(0, [0.0, 45.0, 123.4])
    sd = np.float32(np.sqrt(3.0))
    for v in range(3):
        ax, rl, w = ctx.forward_project(0, v)
        hit = w > 0
        np.testing.assert_allclose(ax[hit], c * sd * w[hit], rtol=2e-6)
        assert np.all(ax[~hit] == 0)
    ctx.close()


def test_bp_voxel_matches_oracle():
    print("[bp-voxel] 28x24x30 det 48x44", flush=True)
    rng = np.random.default_rng(7)
    vol, det = (28, 24, 30), (48, 44)
    g, ctx = _ctx(vol, det, ds=0.9, oxyz=(0.5, -1.0, 0.0), P=(2,))
    x = rng.random(int(np.prod(vol)), dtype=np.float32)
    ctx.set_views(0, [12.0, 200.0])
    ctx.set_volume(0, x)
    gmax = ctx.gmax()
    for v, ang in enumerate([12.0, 200.0]):
        corr = rng.standard_normal(det[0] * det[1]).astype(np.float32)
        ctx.backproject_voxel(0, v, corr, 0.3)
        x = O.bp_voxel(g, O.quat_y(ang), corr, 0.3, gmax, x)
    got = ctx.get_volume(0)
    assert rel_l2(got, x) <= 1e-5, rel_l2(got, x)
    ctx.close()


def _quat_axis(axis, deg):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    h = np.radians(deg) / 2.0
    return np.array([*(a * np.sin(h)), np.cos(h)], np.float32)


# (vol, det, sid, sdd): magnification about 1.5-2 (most tiles take the
# interior path), about 3 (wider staged windows, partly past the detector's
# edges), about 4-5 (windows too large for LDS: the general path's global
# window reads)
BP_TILE_CASES = [
    ((64, 64, 64), (136, 136), 300.0, 500.0),
    ((72, 48, 80), (220, 150), 180.0, 500.0),
    ((64, 64, 64), (330, 330), 150.0, 500.0),
]


@pytest.mark.parametrize("case", range(len(BP_TILE_CASES)))
def test_bp_voxel_tiles_match_oracle(case):
    """The voxel backprojector against the oracle on volumes with interior
    tiles, at three magnifications, for views about the y axis and views
    about two oblique axes (quaternions), one view at a time."""
    vol, det, sid, sdd = BP_TILE_CASES[case]
    print(f"[bp-tiles] case={case} vol={vol} det={det} sid={sid} sdd={sdd}", flush=True)
    rng = np.random.default_rng(70 + case)
    qs = [O.quat_y(17.0), O.quat_y(135.0), O.quat_y(250.0),
          _quat_axis((0.3, 1.0, 0.2), 40.0), _quat_axis((1.0, 0.4, -0.5), 200.0)]
    g, ctx = _ctx(vol, det, sid=sid, sdd=sdd, P=(len(qs),))
    x0 = (0.5 * rng.random(int(np.prod(vol)))).astype(np.float32)
    ctx.set_view_quaternions(0, np.stack(qs))
    gmax = ctx.gmax()
    for v, q in enumerate(qs):
        # each view from the same start: the error of one backprojection
        ctx.set_volume(0, x0)
        corr = rng.standard_normal(det[0] * det[1]).astype(np.float32)
        ctx.backproject_voxel(0, v, corr, 0.3)
        got = ctx.get_volume(0)
        ref = O.bp_voxel(g, q, corr, 0.3, gmax, x0)
        ref64 = O.bp_voxel_f64(g, q, corr, 0.3, gmax, x0)
        err, err64, o64 = rel_l2(got, ref), rel_l2(got, ref64), rel_l2(ref, ref64)
        print(f"  view {v}: rel L2 {err:.2e} vs fp32 oracle, {err64:.2e} vs FP64 oracle "
              f"(fp32 oracle vs FP64 {o64:.2e}); voxels changed {np.count_nonzero(got != x0)}",
              flush=True)
        # White-noise corrections at magnification up to 5 make a voxel's
        # value sensitive to the fp32 rounding of its detector position: the
        # bar is the fp32 oracle's own distance from FP64 (at most 1.5x it),
        # and 2e-5 from the fp32 oracle.
        assert err <= 2e-5, (v, err)
        assert err64 <= 1.5 * o64 + 1e-6, (v, err64, o64)
    ctx.close()


def test_bp_ray_matches_oracle():
    print("[bp-ray] 20^3 det 32x32", flush=True)
    rng = np.random.default_rng(8)
    vol, det = (20, 20, 20), (32, 32)
    p = stt.params_default()
    p.bp = stt.STT_BP_RAY
    g, ctx = _ctx(vol, det, sid=50.0, sdd=90.0, P=(1,), params=p)
    ctx.set_views(0, [33.0])
    corr = rng.standard_normal(det[0] * det[1]).astype(np.float32)
    c, w = ctx.backproject_ray(0, 0, corr)
    oc, ow = O.bp_ray_scatter(g, O.quat_y(33.0), corr)
    assert rel_l2(w, ow) <= 1e-6
    assert rel_l2(c, oc) <= 1e-5
    ctx.close()


def test_fp_vjp_dot_product():
    """<A x, y> == <x, A^T y> for the matched ray-driven transpose: relative
    to ||A x|| ||y|| for signed inputs (the dot product itself cancels), and
    relative to the products themselves for non-negative inputs."""
    print("[fp-vjp] 24^3 det 40x40", flush=True)
    rng = np.random.default_rng(9)
    vol, det = (24, 24, 24), (40, 40)
    g, ctx = _ctx(vol, det, P=(2,))
    ctx.set_views(0, [10.0, 250.0])
    for signed in (True, False):
        if signed:
            x = rng.standard_normal(int(np.prod(vol))).astype(np.float32)
            y = rng.standard_normal(det[0] * det[1]).astype(np.float32)
        else:
            x = rng.random(int(np.prod(vol)), dtype=np.float32)
            y = rng.random(det[0] * det[1], dtype=np.float32)
        ctx.set_volume(0, x)
        for v in range(2):
            ax, _, _ = ctx.forward_project(0, v)
            aty = ctx.fp_vjp(0, v, y)
            lhs = float(np.dot(ax.astype(np.float64), y.astype(np.float64)))
            rhs = float(np.dot(x.astype(np.float64), aty.astype(np.float64)))
            scale = (np.linalg.norm(ax.astype(np.float64)) * np.linalg.norm(y.astype(np.float64))
                     if signed else max(abs(lhs), abs(rhs)))
            print(f"  signed={signed} view={v} <Ax,y>={lhs:.9g} <x,A^Ty>={rhs:.9g} rel={abs(lhs - rhs) / scale:.2e}")
            assert abs(lhs - rhs) <= 1e-6 * scale, (lhs, rhs, scale)
    ctx.close()


@pytest.mark.parametrize("interp", [stt.STT_INTERP_CUBIC_CLAMPED, stt.STT_INTERP_TRILINEAR])
@pytest.mark.parametrize("dt", [-1.0, 1.0])
def test_warp_matches_oracle(interp, dt):
    print(f"[warp] interp={interp} dt={dt}", flush=True)
    dims = (24, 20, 28)
    g, ctx = _ctx(dims, (16, 16), P=(1,))
    rng = np.random.default_rng(11)
    src = rng.random(int(np.prod(dims)), dtype=np.float32)
    uvw = smooth_flow(dims, 3.0, seed=12)
    out = ctx.warp(uvw, src, dt, interp)
    ref = O.warp(dims, uvw, src, dt, interp)
    assert rel_l2(out, ref) <= 1e-5, rel_l2(out, ref)
    ctx.close()


@pytest.mark.parametrize("interp", [stt.STT_INTERP_CUBIC_CLAMPED, stt.STT_INTERP_TRILINEAR])
def test_warp_frame_equals_host_warp(interp):
    """stt_warp_frame on the context's flow and frame equals stt_warp on
    the same arrays passed from the host, bit for bit; bad pairs are refused."""
    print(f"[warp-frame] interp={interp} 26x22x18 F=3", flush=True)
    dims = (26, 22, 18)
    g = stt.Geometry.make(dims, (32, 32), (1, 1, 1), 1.0, 60.0, 100.0)
    ctx = stt.Context(g, stt.params_default(), [2, 2, 2])
    for f in range(3):
        ctx.set_volume(f, stt.synth_phantom_frame(dims, f, n_ellipsoids=5, seed=4, deform_seed=5, amp=2.0))
    for f in range(2):
        ctx.set_flow(f, stt.synth_flow(dims, f, deform_seed=5, amp=2.0))
    for pair in range(2):
        got = ctx.warp_frame(pair, -1.0, interp)
        ref = ctx.warp(ctx.get_buffer(stt.STT_BUF_FLOW, pair), ctx.get_volume(pair + 1), -1.0, interp)
        assert np.array_equal(got, ref), pair
    for bad in (-1, 2):
        with pytest.raises(stt.SttError):
            ctx.warp_frame(bad)
    ctx.close()


def test_warp_zero_flow_identity():
    print("[warp-zero]", flush=True)
    dims = (16, 16, 16)
    g, ctx = _ctx(dims, (16, 16), P=(1,))
    src = np.random.default_rng(13).random(16 ** 3, dtype=np.float32)
    out = ctx.warp(np.zeros(3 * 16 ** 3, np.float32), src, -1.0).reshape(16, 16, 16)
    s = src.reshape(16, 16, 16)
    idx = np.clip(np.arange(16), 1, 14)
    assert np.array_equal(out, s[np.ix_(idx, idx, idx)])
    ctx.close()


def _setup_pair(vol, det, P, angles_fn, params, F=1, seed=21, flows=True, x0=None):
    g = stt.Geometry.make(vol, det, (1, 1, 1), 1.0, 60.0, 100.0)
    ctx = stt.Context(g, params, list(P))
    oc = O.OracleCtx(g, params, list(P))
    rng = np.random.default_rng(seed)
    V = int(np.prod(vol))
    for f in range(F):
        ang = angles_fn(f)
        ctx.set_views(f, ang)
        oc.set_views(f, ang)
        truth = stt.synth_phantom_frame(vol, f, n_ellipsoids=6, seed=1, deform_seed=2, amp=2.0)
        b = np.concatenate([O.fp(g, O.quat_y(a), truth)[0] for a in ang])
        ctx.set_projections(f, b)
        oc.set_projections(f, b)
    for f in range(F + 1):
        x = (x0 if x0 is not None else rng.random(V, dtype=np.float32) * 0.2)
        ctx.set_volume(f, x)
        oc.set_volume(f, x)
    if flows:
        for f in range(F - 1):
            u = stt.synth_flow(vol, f, deform_seed=2, amp=2.0)
            ctx.set_flow(f, u)
            oc.set_flow(f, u)
    return g, ctx, oc


@pytest.mark.parametrize("shape", [0, 16])
def test_sart_sweep_matches_oracle(shape):
    print(f"[sart] 32^3 det 48x48 16 views projector shape {shape}", flush=True)
    p = stt.params_default()
    p.alpha = 0.3
    ang = list(np.linspace(0, 360, 16, endpoint=False))
    g, ctx, oc = _setup_pair((32, 32, 32), (48, 48), (16,), lambda f: ang, p)
    assert abs(ctx.gmax() - oc.gmax()) <= 1e-6 * abs(oc.gmax())
    order = np.random.default_rng(3).permutation(16).astype(np.int32)
    try:
        ctx.set_projector_tile(shape)
        ctx.sart_sweep(0, order)
    finally:
        ctx.set_projector_tile(0)
    oc.sart(0, order)
    got, ref = ctx.get_volume(0), oc.get_volume(0)
    assert rel_l2(got, ref) <= 1e-4, rel_l2(got, ref)
    ctx.close()


@pytest.mark.parametrize("interp,prior", [
    (stt.STT_INTERP_CUBIC_CLAMPED, stt.STT_PRIOR_ATV),
    (stt.STT_INTERP_TRILINEAR, stt.STT_PRIOR_ATV),
    (stt.STT_INTERP_CUBIC_CLAMPED, stt.STT_PRIOR_TV),
    (stt.STT_INTERP_CUBIC_CLAMPED, stt.STT_PRIOR_STV),
])
def test_pd_steps_match_oracle(interp, prior):
    print(f"[pd-steps] interp={interp} prior={prior}", flush=True)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01
    p.interp = interp
    p.prior = prior
    p.roi_cut[:] = [2, 3, 1, 4, 2, 2, 5, 4]
    p.roi_jtop = 3
    vol = (24, 22, 26)
    g, ctx, oc = _setup_pair(vol, (32, 32), (4, 4, 4), lambda f: [0, 90, 180, 270], p, F=3)
    ctx.pd_init()
    oc.pd_init()
    for f in range(3):
        ctx.pd_dual(f)
        oc.pd_dual(f)
    for f in range(3):
        ctx.pd_primal(f)
        oc.pd_primal(f)
        ctx.pd_extrapolate(f)
        oc.pd_extrapolate(f)
    for f in range(3):
        for which in (stt.STT_BUF_YTV, stt.STT_BUF_YTP, stt.STT_BUF_YBC, stt.STT_BUF_XBAR, stt.STT_BUF_XPREV):
            a, b = ctx.get_buffer(which, f), oc.get_buffer(which, f)
            assert rel_l2(a, b) <= 1e-5, (f, which, rel_l2(a, b))
        assert rel_l2(ctx.get_volume(f), oc.get_volume(f)) <= 1e-5
    ctx.close()


@pytest.mark.parametrize("mode,prior", [
    ("plain", stt.STT_PRIOR_ATV), ("batched", stt.STT_PRIOR_ATV), ("graph", stt.STT_PRIOR_ATV),
    ("graph", stt.STT_PRIOR_STV),
])
def test_st_iteration_matches_oracle(mode, prior):
    print(f"[st-iter] mode={mode} prior={prior} F=3 28^3 det 40x40", flush=True)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01, 0.3
    p.prior = prior
    P = (6, 5, 7)
    ga = lambda f: list(stt.synth_angles(1, P[f], offset=sum(P[:f])))
    g, ctx, oc = _setup_pair((28, 28, 28), (40, 40), P, ga, p, F=3)
    batched = mode != "plain"
    ctx.iterate(2, batched=batched, use_graph=(mode == "graph"))
    if batched:
        oc.per_frame_raylen(True)
    oc.iterate(2)
    for f in range(3):
        got, ref = ctx.get_volume(f), oc.get_volume(f)
        assert rel_l2(got, ref) <= 1e-4, (f, rel_l2(got, ref))
        xb, oxb = ctx.get_buffer(stt.STT_BUF_XBAR, f), oc.get_buffer(stt.STT_BUF_XBAR, f)
        assert rel_l2(xb, oxb) <= 1e-4
    ctx.close()


def test_flow_change_between_iterations_matches_oracle():
    """The PD warps' stored trace end points follow each flow write: after new
    flows (one through stt_set_flow, one through stt_set_buffer) the next
    iteration equals the oracle's on the new flows."""
    print("[trace-refresh] F=3 26^3 det 36x36, flows replaced after one iteration", flush=True)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01, 0.3
    P = (5, 6, 4)
    vol = (26, 26, 26)
    ga = lambda f: list(stt.synth_angles(1, P[f], offset=sum(P[:f])))
    g, ctx, oc = _setup_pair(vol, (36, 36), P, ga, p, F=3)
    oc.per_frame_raylen(True)
    ctx.iterate(1, batched=True, use_graph=True)
    oc.iterate(1)
    u0 = stt.synth_flow(vol, 4, deform_seed=9, amp=1.5)
    u1 = stt.synth_flow(vol, 5, deform_seed=9, amp=2.5)
    ctx.set_flow(0, u0)
    oc.set_flow(0, u0)
    ctx.set_buffer(stt.STT_BUF_FLOW, 1, u1)
    oc.set_flow(1, u1)
    ctx.iterate(1, batched=True, use_graph=True)
    oc.iterate(1)
    for f in range(3):
        got, ref = ctx.get_volume(f), oc.get_volume(f)
        assert rel_l2(got, ref) <= 1e-4, (f, rel_l2(got, ref))
        assert rel_l2(ctx.get_buffer(stt.STT_BUF_YBC, f), oc.get_buffer(stt.STT_BUF_YBC, f)) <= 1e-4
    ctx.close()


def test_interp_change_recomputes_traces():
    """Switching the warp interpolation with stt_ctx_set_params re-traces the
    flows: the PD steps then equal the oracle's run with the new interpolation."""
    print("[trace-interp] F=3 24x22x26, cubic traces, then trilinear steps", flush=True)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01
    p.interp = stt.STT_INTERP_CUBIC_CLAMPED
    vol = (24, 22, 26)
    g, ctx, _ = _setup_pair(vol, (32, 32), (4, 4, 4), lambda f: [0, 90, 180, 270], p, F=3)
    ctx.iterate(0)  # PD initialisation and the cubic traces, no step
    p2 = stt.params_default()
    p2.sigma, p2.tau, p2.lambda_tv, p2.k_tp, p2.e_bc, p2.huber = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01
    p2.interp = stt.STT_INTERP_TRILINEAR
    ctx.set_params(p2)
    g2, ctx2, oc = _setup_pair(vol, (32, 32), (4, 4, 4), lambda f: [0, 90, 180, 270], p2, F=3)
    ctx2.close()
    oc.pd_init()
    for f in range(3):
        ctx.pd_dual(f)
        oc.pd_dual(f)
    for f in range(3):
        ctx.pd_primal(f)
        oc.pd_primal(f)
    for f in range(3):
        for which in (stt.STT_BUF_YBC, stt.STT_BUF_YTV):
            a, b = ctx.get_buffer(which, f), oc.get_buffer(which, f)
            assert rel_l2(a, b) <= 1e-5, (f, which, rel_l2(a, b))
        assert rel_l2(ctx.get_volume(f), oc.get_volume(f)) <= 1e-5
    ctx.close()


@pytest.mark.parametrize("interp", [0, 1])
def test_trace_storage_off_is_bit_identical(interp):
    """The PD kernels that trace in place (stt_set_trace_storage(ctx, 0)) give
    the stored-trace kernels' results bit for bit, in every form of the
    iteration, and storage switched off and on again mid-run changes nothing."""
    print(f"[trace-storage] F=3 24x22x26, interp {interp}, STV", flush=True)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01, 0.3
    p.interp, p.prior = interp, stt.STT_PRIOR_STV
    vol, P = (24, 22, 26), (4, 4, 4)
    ctxs = []
    for _ in range(2):
        ctxs.append(_setup_pair(vol, (32, 32), P, lambda f: [0, 90, 180, 270], p, F=3)[1])
    kept, traced = ctxs
    traced.set_trace_storage(False)
    for batched, graph in ((False, False), (True, False), (True, True)):
        kept.iterate(1, batched=batched, use_graph=graph)
        traced.iterate(1, batched=batched, use_graph=graph)
    traced.set_trace_storage(True)
    kept.set_trace_storage(False)
    for c in ctxs:
        c.iterate(1, batched=True, use_graph=True)
    for f in range(3):
        assert np.array_equal(kept.get_volume(f), traced.get_volume(f)), f
        for which in (stt.STT_BUF_YBC, stt.STT_BUF_YTV, stt.STT_BUF_XBAR):
            assert np.array_equal(kept.get_buffer(which, f), traced.get_buffer(which, f)), (f, which)
    for c in ctxs:
        c.close()


def test_batched_equals_plain():
    print("[batched-vs-plain] F=4", flush=True)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01, 0.3
    P = (5, 8, 3, 6)
    vol, det = (24, 24, 24), (36, 36)
    ga = lambda f: list(stt.synth_angles(1, P[f], offset=sum(P[:f])))
    outs = []
    for batched in (False, True):
        g, ctx, oc = _setup_pair(vol, det, P, ga, p, F=4, seed=5)
        ctx.iterate(3, batched=batched, use_graph=batched)
        outs.append([ctx.get_volume(f) for f in range(4)])
        ctx.close()
    for f in range(4):
        assert rel_l2(outs[1][f], outs[0][f]) <= 1e-6


def test_fp_and_ray_bp_vs_reference_fixture():
    """The HIP projector against vectors written by the reference's own code."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))
    for i in range(int(gold["ncases"])):
        geo = gold[f"c{i}_geo"]
        vol = tuple(int(v) for v in geo[0:3])
        det = (int(geo[3]), int(geo[4]))
        print(f"[fixture] case {i} vol={vol} det={det}", flush=True)
        p = stt.params_default()
        p.bp = stt.STT_BP_RAY
        g = stt.Geometry.make(vol, det, tuple(gold[f"c{i}_spacing"]), geo[5], geo[6], geo[7], tuple(geo[8:11]))
        ctx = stt.Context(g, p, [1])
        ctx.set_view_quaternions(0, gold[f"c{i}_q"])
        ctx.set_volume(0, gold[f"c{i}_x"])
        ax, rl, w = ctx.forward_project(0, 0, raylen=gold[f"c{i}_raylen_in"])
        assert np.array_equal(w, gold[f"c{i}_weight"]), f"case {i}: sample counts differ"
        assert np.array_equal(rl, gold[f"c{i}_raylen"]), f"case {i}: ray lengths differ"
        assert rel_l2(ax, gold[f"c{i}_ax"]) <= 1e-5
        c, s = ctx.backproject_ray(0, 0, gold[f"c{i}_corr"])
        assert rel_l2(s, gold[f"c{i}_wsum"]) <= 1e-6
        assert rel_l2(c, gold[f"c{i}_contrib"]) <= 1e-5
        ctx.close()
