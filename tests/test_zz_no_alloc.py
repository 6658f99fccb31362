"""The allocation-free residual norm (stt_residual_norm_noalloc) against
stt_residual_norm, through the C-ABI: two calls in a row, the second on a
smaller problem, each equal to the allocating entry point's FP64 result bit
for bit (same additions in the same order); a detector beyond the scratch
limit is refused with an error."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from spacetimetomography_amd._lib import lib

pytestmark = pytest.mark.gpu

def _ctx(vol, det, P, seed):
    g = stt.Geometry.make(vol, det, (1, 1, 1), 1.0, 90.0, 150.0)
    ctx = stt.Context(g, stt.params_default(), list(P))
    rng = np.random.default_rng(seed)
    off = 0
    for f in range(len(P)):
        ang = list(stt.synth_angles(1, P[f], offset=off))
        off += P[f]
        ctx.set_views(f, ang)
        truth = stt.synth_phantom_frame(vol, f, n_ellipsoids=5, seed=seed, deform_seed=2, amp=2.0)
        b = np.concatenate([O.fp(g, O.quat_y(a), truth)[0] for a in ang])
        ctx.set_projections(f, (b + 0.05 * rng.standard_normal(b.size)).astype(np.float32))
        ctx.set_volume(f, (rng.random(int(np.prod(vol)), dtype=np.float32) * 0.3).astype(np.float32))
    return ctx


def _noalloc(ctx):
    r2 = C.c_double(-1.0)
    rc = lib.stt_residual_norm_noalloc(ctx.h, C.byref(r2))
    return rc, r2.value


def test_noalloc_residual_norm_equals_allocating_entry_point():
    for vol, det, P, seed in (((48, 48, 48), (96, 80), (5, 4, 6), 3), ((24, 20, 28), (40, 36), (3, 2), 4)):
        print(f"[noalloc] {vol} det {det} P={P}", flush=True)
        ctx = _ctx(vol, det, P, seed)
        try:
            ref = ctx.residual_norm()
            rc, got = _noalloc(ctx)
            assert rc == stt.STT_OK, lib.stt_last_error()
            print(f"  r2 = {got!r} (allocating entry point {ref!r})", flush=True)
            assert ref > 0 and got == ref
            rc, again = _noalloc(ctx)  # the scratch keeps the last call's contents: cleared per call
            assert rc == stt.STT_OK and again == ref
        finally:
            ctx.close()


def test_noalloc_refuses_detector_beyond_scratch():
    print("[noalloc-limit] 8^3, 1100x1000 detector", flush=True)
    ctx = _ctx((8, 8, 8), (1100, 1000), (1,), 5)
    try:
        rc, _ = _noalloc(ctx)
        assert rc == stt.STT_ERR_INVALID
        assert b"scratch limit" in lib.stt_last_error()
    finally:
        ctx.close()
