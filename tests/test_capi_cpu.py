"""The C-ABI library: loads, exports every declared entry point, and its
host-only functions and argument checks behave (no device needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import spacetimetomography_amd as stt
from conftest import HAVE_GPU, ROOT


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(stt_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(stt.lib, n)]
    assert not missing, missing


def test_params_reference_values():
    p = stt.params_reference()
    assert p.sigma == pytest.approx(0.1125)      # CP_sigma = .9/(8*1) (ST_Tomography.cpp:99-102)
    assert p.tau == 1.0 and p.alpha == pytest.approx(0.1) and p.huber == pytest.approx(0.1)
    assert list(p.roi_cut) == [8, 8, 8, 45, 8, 8, 30, 20] and p.roi_jtop == 7
    d = stt.params_default()
    assert list(d.roi_cut) == [0] * 8 and d.roi_jtop == 0


def test_invalid_arguments_are_rejected_before_device_use():
    out = C.c_void_p()
    p = stt.params_default()
    P = (C.c_int * 1)(4)
    g = stt.Geometry.make((2, 32, 32), (64, 64))  # dims < 4
    assert stt.lib.stt_ctx_create(C.byref(g), C.byref(p), 1, P, C.byref(out)) == stt.STT_ERR_INVALID
    assert b"dims" in stt.lib.stt_last_error()
    g = stt.Geometry.make((16385, 8, 8), (64, 64))  # the slab projector packs box extents in 16 bits
    assert stt.lib.stt_ctx_create(C.byref(g), C.byref(p), 1, P, C.byref(out)) == stt.STT_ERR_INVALID
    assert b"16384" in stt.lib.stt_last_error()
    g = stt.Geometry.make((32, 32, 32), (64, 64), spacing=(1, 0, 1))
    assert stt.lib.stt_ctx_create(C.byref(g), C.byref(p), 1, P, C.byref(out)) == stt.STT_ERR_INVALID
    g = stt.Geometry.make((32, 32, 32), (64, 64))
    assert stt.lib.stt_ctx_create(None, C.byref(p), 1, P, C.byref(out)) == stt.STT_ERR_INVALID
    assert stt.lib.stt_ctx_create(C.byref(g), C.byref(p), 0, P, C.byref(out)) == stt.STT_ERR_INVALID
    Z = (C.c_int * 1)(0)
    assert stt.lib.stt_ctx_create(C.byref(g), C.byref(p), 1, Z, C.byref(out)) == stt.STT_ERR_INVALID
    assert stt.lib.stt_ctx_destroy(None) == stt.STT_ERR_INVALID
    assert stt.lib.stt_iterate(None, 1, 1, 0, None) == stt.STT_ERR_INVALID
    grp = C.c_void_p()
    assert stt.lib.stt_group_create(None, 1, C.byref(grp)) == stt.STT_ERR_INVALID
    none2 = (C.c_void_p * 2)(None, None)
    assert stt.lib.stt_group_create(none2, 2, C.byref(grp)) == stt.STT_ERR_INVALID
    assert b"null context" in stt.lib.stt_last_error()
    assert stt.lib.stt_group_create(none2, 0, C.byref(grp)) == stt.STT_ERR_INVALID
    assert stt.lib.stt_group_iterate(None, 1, 1, None) == stt.STT_ERR_INVALID
    assert stt.lib.stt_group_destroy(None) == stt.STT_OK
    assert stt.lib.stt_warp_frame(None, 0, -1.0, 0, None) == stt.STT_ERR_INVALID
    # the device-pointer forms check the context before any copy
    assert stt.lib.stt_set_volume_dev(None, 0, None) == stt.STT_ERR_INVALID
    assert stt.lib.stt_get_volume_dev(None, 0, None) == stt.STT_ERR_INVALID
    assert stt.lib.stt_set_flow_dev(None, 0, None) == stt.STT_ERR_INVALID
    assert stt.lib.stt_set_projections_dev(None, 0, None) == stt.STT_ERR_INVALID
    assert stt.lib.stt_forward_project_dev(None, 0, 0, None, None, None) == stt.STT_ERR_INVALID
    assert stt.lib.stt_backproject_voxel_dev(None, 0, 0, None, 1.0) == stt.STT_ERR_INVALID
    assert stt.lib.stt_fp_vjp_dev(None, 0, 0, None, None) == stt.STT_ERR_INVALID
    assert stt.lib.stt_warp_dev(None, None, None, 1.0, None, 0) == stt.STT_ERR_INVALID
    assert stt.lib.stt_warp_vjp_dev(None, None, None, 1.0, None, None, None) == stt.STT_ERR_INVALID
    assert b"null context" in stt.lib.stt_last_error()


@pytest.mark.skipif(HAVE_GPU, reason="checks the no-device path")
def test_no_device_is_reported():
    out = C.c_void_p()
    g = stt.Geometry.make((32, 32, 32), (64, 64))
    p = stt.params_default()
    P = (C.c_int * 1)(4)
    assert stt.lib.stt_ctx_create(C.byref(g), C.byref(p), 1, P, C.byref(out)) == stt.STT_ERR_NODEVICE


def test_synth_phantom_is_seeded_and_bounded():
    a = stt.synth_phantom((24, 20, 16), 10, seed=1)
    b = stt.synth_phantom((24, 20, 16), 10, seed=1)
    c = stt.synth_phantom((24, 20, 16), 10, seed=2)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() <= 10.0 and a.max() > 0


def test_synth_deformation_amplitude_and_frame_zero():
    dims = (20, 18, 16)
    for t in range(3):
        u = stt.synth_flow(dims, t, deform_seed=2, amp=4.0).reshape(3, -1)
        # |D(t+1) - D(t)| <= 2*amp
        assert np.sqrt((u.astype(np.float64) ** 2).sum(0)).max() <= 8.0 + 1e-5
    f0 = stt.synth_phantom_frame(dims, 0, 6, seed=1, deform_seed=2, amp=0.0)
    assert np.array_equal(f0, stt.synth_phantom(dims, 6, seed=1))


def test_synth_angles():
    even = stt.synth_angles(0, 8)
    np.testing.assert_allclose(even, np.arange(8) * 45.0)
    gold = stt.synth_angles(1, 4, offset=2)
    np.testing.assert_allclose(gold, np.mod((2 + np.arange(4)) * 180 * (3 - np.sqrt(5)), 360), rtol=1e-6)
    vdc = stt.synth_angles(2, 8)
    np.testing.assert_allclose(vdc, [0, 180, 90, 270, 45, 225, 135, 315])
