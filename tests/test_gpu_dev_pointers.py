"""The device-pointer ("_dev") entry points of the C-ABI against their host
forms: each one, given device buffers holding the same values, leaves the
same context state and returns the same outputs (same kernels, only the copy
kinds differ).  Bit for bit where the kernel's additions have a fixed order;
the projector transpose and the warp VJP scatter with float atomics, whose
order varies from launch to launch, so those two are held to fp32 rounding
(max difference <= 1e-5 of the largest magnitude).  The device buffers come
from hipMalloc of the HIP runtime libstt.so links, as a caller's own buffers
would."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from helpers import smooth_flow

pytestmark = pytest.mark.gpu

_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_H2D, _D2H = 1, 2


class DevBuf:
    """A float32 device buffer of n elements (hipMalloc / hipFree)."""

    def __init__(self, n, init=None):
        self.n = int(n)
        self.ptr = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.ptr), 4 * self.n) == 0
        if init is not None:
            self.upload(init)

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32).reshape(-1)
        assert a.size == self.n
        assert _hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), 4 * self.n, _H2D) == 0

    def download(self):
        out = np.empty(self.n, np.float32)
        assert _hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, 4 * self.n, _D2H) == 0
        return out

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


VOL, DET, P = (24, 20, 28), (40, 36), (3, 2)


def same_to_fp32_rounding(got, ref):
    """Equal up to reordered fp32 atomic additions."""
    return float(np.abs(got.astype(np.float64) - ref).max()) <= 1e-5 * float(np.abs(ref).max())


@pytest.fixture()
def case():
    g = stt.Geometry.make(VOL, DET, (1, 1, 1), 1.0, 90.0, 150.0)
    ctx = stt.Context(g, stt.params_default(), list(P))
    rng = np.random.default_rng(11)
    off = 0
    bs = []
    for f in range(len(P)):
        ang = list(stt.synth_angles(1, P[f], offset=off))
        off += P[f]
        ctx.set_views(f, ang)
        truth = stt.synth_phantom_frame(VOL, f, n_ellipsoids=5, seed=3, deform_seed=2, amp=2.0)
        b = np.concatenate([O.fp(g, O.quat_y(a), truth)[0] for a in ang])
        bs.append((b + 0.05 * rng.standard_normal(b.size)).astype(np.float32))
        ctx.set_projections(f, bs[-1])
        ctx.set_volume(f, (rng.random(ctx.V, dtype=np.float32) * 0.3).astype(np.float32))
    bufs = []

    def dev(n, init=None):
        bufs.append(DevBuf(n, init))
        return bufs[-1]

    yield ctx, rng, bs, dev
    for b in bufs:
        b.free()
    ctx.close()


def test_volume_and_flow_buffers(case):
    ctx, rng, _, dev = case
    print("[dev] set/get volume, set flow", flush=True)
    x = (rng.random(ctx.V, dtype=np.float32) * 2 - 0.5).astype(np.float32)
    d = dev(ctx.V, x)
    ctx.set_volume_dev(1, d.ptr)
    assert np.array_equal(ctx.get_volume(1), x)
    out = dev(ctx.V, np.zeros(ctx.V, np.float32))
    ctx.get_volume_dev(1, out.ptr)
    assert np.array_equal(out.download(), x)
    uvw = smooth_flow(VOL, 1.5, 4)
    ctx.set_flow(0, uvw)
    host_form = ctx.get_buffer(stt.STT_BUF_FLOW, 0)
    ctx.set_flow(0, np.zeros_like(uvw))
    ctx.set_flow_dev(0, dev(3 * ctx.V, uvw).ptr)
    assert np.array_equal(ctx.get_buffer(stt.STT_BUF_FLOW, 0), host_form)
    assert np.array_equal(host_form, uvw)
    # the flow changed flag makes stt_warp_frame use the new flow in both forms
    ctx.set_volume(1, x)
    w_dev = ctx.warp_frame(0)
    ctx.set_flow(0, uvw)
    assert np.array_equal(ctx.warp_frame(0), w_dev)


def test_projections_set_g_max_and_the_residual(case):
    ctx, _, bs, dev = case
    print("[dev] set projections", flush=True)
    g_host, r_host = ctx.gmax(), ctx.residual_norm()
    scaled = (bs[0] * 1.5).astype(np.float32)
    ctx.set_projections_dev(0, dev(scaled.size, scaled).ptr)
    assert ctx.gmax() == np.float32(scaled.max() / np.float32(4.0))
    assert ctx.residual_norm() != r_host
    ctx.set_projections_dev(0, dev(bs[0].size, bs[0]).ptr)
    assert ctx.gmax() == g_host and ctx.residual_norm() == r_host
    ctx.set_projections_dev(1, dev(bs[1].size, bs[1]).ptr)
    assert ctx.gmax() == g_host and ctx.residual_norm() == r_host


def test_projector_backprojector_and_vjp(case):
    ctx, rng, _, dev = case
    D = ctx.D
    for frame, view in ((0, 2), (1, 0)):
        print(f"[dev] projector frame {frame} view {view}", flush=True)
        rl0 = (rng.random(D, dtype=np.float32) * 3).astype(np.float32)
        ax, rl, w = ctx.forward_project(frame, view, rl0)
        dax, drl, dw = dev(D, np.zeros(D, np.float32)), dev(D, rl0), dev(D, np.zeros(D, np.float32))
        ctx.forward_project_dev(frame, view, dax.ptr, drl.ptr, dw.ptr)
        assert np.array_equal(dax.download(), ax)
        assert np.array_equal(drl.download(), rl)
        assert np.array_equal(dw.download(), w)
        assert not np.array_equal(rl, rl0)  # hit rays overwrite the given lengths (missed rays keep them)
        only_ax = dev(D, np.zeros(D, np.float32))
        ctx.forward_project_dev(frame, view, only_ax.ptr, None, None)
        assert np.array_equal(only_ax.download(), ax)

        y = rng.standard_normal(D).astype(np.float32)
        acc0 = rng.standard_normal(ctx.V).astype(np.float32)
        host_vjp = ctx.fp_vjp(frame, view, y, acc0)
        out = dev(ctx.V, acc0)
        ctx.fp_vjp_dev(frame, view, dev(D, y).ptr, out.ptr)
        assert same_to_fp32_rounding(out.download(), host_vjp)

        x0 = ctx.get_volume(frame)
        corr = (rng.standard_normal(D) * 0.1).astype(np.float32)
        ctx.backproject_voxel(frame, view, corr, 0.7)
        host_bp = ctx.get_volume(frame)
        assert not np.array_equal(host_bp, x0)
        ctx.set_volume(frame, x0)
        ctx.backproject_voxel_dev(frame, view, dev(D, corr).ptr, 0.7)
        assert np.array_equal(ctx.get_volume(frame), host_bp)


@pytest.mark.parametrize("interp", [stt.STT_INTERP_CUBIC_CLAMPED, stt.STT_INTERP_TRILINEAR])
def test_warp_and_warp_vjp(case, interp):
    ctx, rng, _, dev = case
    V = ctx.V
    print(f"[dev] warp interp {interp}", flush=True)
    uvw = smooth_flow(VOL, 2.0, 7)
    src = stt.synth_phantom(VOL, 6, seed=5)
    host = ctx.warp(uvw, src, -1.0, interp)
    out = dev(V, np.zeros(V, np.float32))
    ctx.warp_dev(dev(3 * V, uvw).ptr, dev(V, src).ptr, -1.0, out.ptr, interp)
    assert np.array_equal(out.download(), host)

    g = rng.standard_normal(V).astype(np.float32)
    gs0 = rng.standard_normal(V).astype(np.float32)
    gu0 = rng.standard_normal(3 * V).astype(np.float32)
    hs, hu = ctx.warp_vjp(uvw, src, 0.5, g, gs0.copy(), gu0.copy())
    ds, du = dev(V, gs0), dev(3 * V, gu0)
    ctx.warp_vjp_dev(dev(3 * V, uvw).ptr, dev(V, src).ptr, 0.5, dev(V, g).ptr, ds.ptr, du.ptr)
    assert same_to_fp32_rounding(ds.download(), hs) and same_to_fp32_rounding(du.download(), hu)
    only_src = dev(V, gs0)
    ctx.warp_vjp_dev(dev(3 * V, uvw).ptr, dev(V, src).ptr, 0.5, dev(V, g).ptr, only_src.ptr, None)
    assert same_to_fp32_rounding(only_src.download(), hs)


def test_null_device_arguments_are_rejected(case):
    ctx, _, _, _ = case
    print("[dev] null arguments", flush=True)
    for call in (lambda: ctx.set_flow_dev(0, None), lambda: ctx.set_projections_dev(0, None),
                 lambda: ctx.backproject_voxel_dev(0, 0, None, 1.0), lambda: ctx.fp_vjp_dev(0, 0, None, None),
                 lambda: ctx.warp_dev(None, None, 1.0, None)):
        with pytest.raises(stt.SttError) as e:
            call()
        assert e.value.code == stt.STT_ERR_INVALID
