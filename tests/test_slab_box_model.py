"""Host model of the slab projector's staged-box containment
(tests/slab_box_model.cpp): the sample walk and slab boxes of k_fp_slab with
the same fp32 operations, geometry from the library's build_view.  Every
sample that the unchecked sampling path takes must have its clamped cell
and +1 taps inside the staged box.

The bound the kernel uses (1: line extremes clamped to [0, dims-1] before
the floor) leaves no sample outside at the geometry where the first bound
(0) failed on the device -- 40^3 on a 384^2 detector of 0.25 mm pixels,
where blocks of grazing rays had all their samples in a slab beyond a face --
nor at the other geometries below; the first bound's misses there are
reproduced, so the model sees the failure it guards against.  Slabs marked
inner (walked without the clamp) must take exactly the clamped walk's
samples at unclamped positions."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import spacetimetomography_amd as stt

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("slabmodel") / "slab_model.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", so,
                    os.path.join(HERE, "slab_box_model.cpp"),
                    os.path.join(ROOT, "spacetimetomography_amd", "csrc", "stt_geometry.cpp")], check=True)
    lib = C.CDLL(so)
    lib.slab_model_ex.restype = C.c_int
    lib.slab_model_ex.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_longlong)]

    def run(vol, det, spacing, ds, sid, sdd, angles, bound, oxyz=(0, 0, 0), tw=16, cap=6144, dense=False):
        g = stt.Geometry.make(vol, det, spacing, ds, sid, sdd, oxyz)
        a = np.ascontiguousarray(angles, np.float32)
        out = (C.c_longlong * 6)()
        assert lib.slab_model_ex(C.byref(g), a.ctypes.data_as(C.POINTER(C.c_float)), len(a), bound, tw, cap,
                                 int(dense), out) == 0
        return dict(checked=out[0], outside=out[1], waves=out[2], samples=out[3], inner=out[4], inner_bad=out[5])

    return run


@pytest.mark.parametrize("tw", [16, 32, 8])
def test_fine_detector_grazing_rays(model, tw):
    ang = list(stt.synth_angles(2, 12, offset=0))
    old = model((40, 40, 40), (384, 384), (1, 1, 1), 0.25, 600.0, 1000.0, ang, 0, tw=tw)
    new = model((40, 40, 40), (384, 384), (1, 1, 1), 0.25, 600.0, 1000.0, ang, 1, tw=tw)
    print(old, new)
    assert old["outside"] > 0           # the failure seen on the device
    assert new["outside"] == 0
    assert new["checked"] > 0.99 * new["samples"]  # nearly every sample takes the unchecked path
    assert new["samples"] == old["samples"]
    assert new["inner_bad"] == 0


@pytest.mark.parametrize("tw", [16, 32, 8])
@pytest.mark.parametrize("case", [
    ((64, 48, 56), (200, 160), (1.0, 1.0, 1.0), 0.5, 120.0, 200.0, (0.0, 0.0, 0.0)),
    ((40, 36, 44), (128, 112), (0.9, 1.2, 1.0), 0.6, 80.0, 140.0, (2.0, -1.5, 1.0)),
    ((32, 32, 32), (96, 96), (1.0, 1.0, 1.0), 0.2, 60.0, 100.0, (0.0, 3.0, 0.0)),
])
def test_other_geometries(model, case, tw):
    vol, det, sp, ds, sid, sdd, oxyz = case
    ang = list(stt.synth_angles(1, 10, offset=3)) + [0.0, 45.0, 90.0, 135.0, 180.0, 270.0]
    new = model(vol, det, sp, ds, sid, sdd, ang, 1, oxyz, tw=tw)
    print(new)
    assert new["outside"] == 0 and new["checked"] > 0 and new["inner_bad"] == 0 and new["inner"] > 0


@pytest.mark.parametrize("tw", [16, 32, 8])
def test_iteration_geometry_inner_slabs(model, tw):
    """256^3 on the 512x512 detector at 600/1000 mm (the iteration
    benchmark's geometry), views on both sides of the dominant-axis switch:
    no sample outside its box, inner slabs exact, and most samples taken on
    inner slabs (the clamp-free walk is the common case)."""
    new = model((256, 256, 256), (512, 512), (1, 1, 1), 1.0, 600.0, 1000.0, [17.3, 44.99, 45.01, 226.1], 1, tw=tw)
    print(new, "inner fraction", new["inner"] / new["samples"])
    assert new["outside"] == 0 and new["inner_bad"] == 0
    assert new["inner"] > 0.5 * new["samples"]


def test_ray_bp_scatter_boxes_stage_every_sample(model):
    """The ray-BP scatter's 2048-voxel LDS box in its dense layout (slab axis
    slowest, footprint n[A] n[B] n[D]) fits every box the slab depth cap
    accepts: on the iteration benchmark's geometry (256^3, 512^2 at
    600/1000 mm) and 12 golden-angle views, every sample of every wave goes
    through a staged box, with no tap outside it.  The projector's padded
    x-fastest layout at the same capacity leaves a fifth of the samples in
    refused boxes (the x-dominant views, whose fastest LDS axis is the slab
    depth), which is what the dense layout removed."""
    ang = list(stt.synth_angles(1, 12, offset=0))
    geo = ((256, 256, 256), (512, 512), (1, 1, 1), 1.0, 600.0, 1000.0)
    dense = model(*geo, ang, 1, cap=2048, dense=True)
    padded = model(*geo, ang, 1, cap=2048, dense=False)
    print(dense, padded)
    assert dense["outside"] == 0 and padded["outside"] == 0
    assert dense["samples"] == padded["samples"]
    assert dense["checked"] == dense["samples"]
    assert padded["checked"] < 0.9 * padded["samples"]
