"""The CPU oracle against the reference's own code.

tests/golden/ref_vectors.npz holds outputs of the reference's ProduceRay /
CastingRay / T_VolumeAccess / Matrix4x4 / Quaternion / CImg code compiled in
this container (tests/golden/make_golden.py).  The oracle must reproduce them
bit for bit.  When oracle/_ref is built (reference tree present), further
random cases are compared live.
"""
import ctypes as C
import os

import numpy as np
import pytest

import _oracle as O
from spacetimetomography_amd._lib import Geometry

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))
NC = int(GOLD["ncases"])


def geometry_of(i):
    geo = GOLD[f"c{i}_geo"]
    vol = tuple(int(v) for v in geo[0:3])
    det = (int(geo[3]), int(geo[4]))
    return Geometry.make(vol, det, tuple(GOLD[f"c{i}_spacing"]), geo[5], geo[6], geo[7], tuple(geo[8:11])), geo[11]


@pytest.mark.parametrize("i", range(NC))
def test_view_matrices_bit_exact(i):
    g, ang = geometry_of(i)
    q = O.quat_y(ang)
    assert np.array_equal(q, GOLD[f"c{i}_q"])
    assert np.array_equal(O.view_matrices(g, q), GOLD[f"c{i}_mats"])


@pytest.mark.parametrize("i", range(NC))
def test_forward_projection_bit_exact(i):
    g, _ = geometry_of(i)
    ax, rl, w = O.fp(g, GOLD[f"c{i}_q"], GOLD[f"c{i}_x"], raylen=GOLD[f"c{i}_raylen_in"])
    assert np.array_equal(w, GOLD[f"c{i}_weight"])
    assert np.array_equal(rl, GOLD[f"c{i}_raylen"])
    assert np.array_equal(ax, GOLD[f"c{i}_ax"])


def test_fixture_covers_pre_entry_samples_and_misses():
    extra = missed = 0
    for i in range(NC):
        g, _ = geometry_of(i)
        sd = np.float32(np.sqrt(np.sum(GOLD[f"c{i}_spacing"].astype(np.float32) ** 2)))
        tn, tf, w = GOLD[f"c{i}_tnear"], GOLD[f"c{i}_tfar"], GOLD[f"c{i}_weight"]
        hit = tn < tf
        extra += int(np.sum((w > 1 + np.ceil((tf - tn) / sd)) & hit))
        missed += int(np.sum(~hit))
    assert extra > 100 and missed > 100


@pytest.mark.parametrize("i", range(NC))
def test_ray_bp_scatter_bit_exact(i):
    g, _ = geometry_of(i)
    c, w = O.bp_ray_scatter(g, GOLD[f"c{i}_q"], GOLD[f"c{i}_corr"])
    assert np.array_equal(w, GOLD[f"c{i}_wsum"])
    assert np.array_equal(c, GOLD[f"c{i}_contrib"])


def test_cubic_atxyz_bit_exact():
    img = GOLD["cubic_img"]
    got = np.array([O.cubic_atxyz(img, (7, 9, 5), *map(float, p)) for p in GOLD["cubic_pts"]], np.float32)
    assert np.array_equal(got, GOLD["cubic_vals"])


def test_quaternions_bit_exact():
    got = np.stack([O.quat_y(a) for a in GOLD["quat_angles"]])
    assert np.array_equal(got, GOLD["quats"])


@pytest.mark.skipif(not O.HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("seed", range(8))
def test_live_random_geometry_vs_reference_build(seed):
    rng = np.random.default_rng(1000 + seed)
    vol = tuple(int(v) for v in rng.integers(6, 18, 3))
    det = (int(rng.integers(12, 40)), int(rng.integers(12, 40)))
    sp = tuple(float(v) for v in rng.uniform(0.6, 1.4, 3))
    sid = float(rng.uniform(15, 60))
    g = Geometry.make(vol, det, sp, float(rng.uniform(0.4, 1.4)), sid, sid * float(rng.uniform(1.3, 2.2)),
                      tuple(float(v) for v in rng.uniform(-3, 3, 3)))
    ang = float(rng.uniform(0, 360))
    q = O.quat_y(ang)
    qr = np.zeros(4, np.float32)
    O.ref.r_quat_y(ang, O.f32(qr))
    assert np.array_equal(q, qr)
    m = np.zeros(89, np.float32)
    O.ref.r_view_matrices(C.byref(g), O.f32(q), O.f32(m))
    assert np.array_equal(O.view_matrices(g, q), m)
    x = rng.random(int(np.prod(vol)), dtype=np.float32)
    D = det[0] * det[1]
    ax, rl, w = (np.zeros(D, np.float32) for _ in range(3))
    O.ref.r_fp(C.byref(g), O.f32(q), O.f32(x), O.f32(ax), O.f32(rl), O.f32(w), None, None)
    oax, orl, ow = O.fp(g, q, x)
    assert np.array_equal(ow, w) and np.array_equal(orl, rl) and np.array_equal(oax, ax)
    corr = rng.standard_normal(D).astype(np.float32)
    c, s = (np.zeros(x.size, np.float32) for _ in range(2))
    O.ref.r_bp_ray_scatter(C.byref(g), O.f32(q), O.f32(corr), O.f32(c), O.f32(s))
    oc, os_ = O.bp_ray_scatter(g, q, corr)
    assert np.array_equal(oc, c) and np.array_equal(os_, s)
