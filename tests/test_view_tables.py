"""Acquisition view tables: the reference's rounds of views and its
low-discrepancy angle sequence (host functions, no device)."""
import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt


def _gold():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))


def test_view_rounds_match_reference_composition_fixture():
    """10 views from 17 degrees, composed in fp32 as InitViews/RotateVolume
    compose them (fixture written by the reference's own Quaternion code)."""
    g = _gold()
    q = stt.view_rounds([float(g["round_start"])], 10)
    assert np.array_equal(q, g["round_quats"])


@pytest.mark.skipif(not O.HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
def test_view_rounds_match_reference_build():
    rng = np.random.default_rng(21)
    starts = rng.uniform(0, 360, 5).astype(np.float32)
    got = stt.view_rounds(starts, 10)
    for j, s in enumerate(starts):
        q = np.zeros(4, np.float32)
        O.ref.r_quat_y(float(s), O.f32(q))
        assert np.array_equal(got[10 * j], q)
        for k in range(1, 10):
            qn = np.zeros(4, np.float32)
            O.ref.r_quat_rotate_y(O.f32(np.ascontiguousarray(q)), 36.0, O.f32(qn))
            assert np.array_equal(got[10 * j + k], qn), (j, k)
            q = qn


def test_view_rounds_rotate_by_36_degrees():
    """Each view is the previous one turned by 36 degrees about +y (up to fp32
    drift of the composition), so the rotation angle advances by 36."""
    q = stt.view_rounds([0.0, 12.5], 10).astype(np.float64)
    ang = np.degrees(2 * np.arctan2(q[:, 1], q[:, 3])) % 360
    expect = np.concatenate([(36.0 * np.arange(10)) % 360, (12.5 + 36.0 * np.arange(10)) % 360])
    d = (ang - expect + 180) % 360 - 180
    assert np.abs(d).max() < 1e-3
    assert np.allclose(q[:, 0], 0) and np.allclose(q[:, 2], 0)


def test_lds_sequence_van_der_corput_closed_forms():
    # base 2: radical inverses 0, 1/2, 1/4, 3/4, 1/8, 5/8, 3/8, 7/8
    assert np.allclose(stt.lds_sequence(8, 1, 2), [0, 180, 90, 270, 45, 225, 135, 315])
    # base 3: 0, 1/3, 2/3, 1/9, 4/9, 7/9, 2/9, 5/9, 8/9
    assert np.allclose(stt.lds_sequence(9, 1, 3), np.array([0, 3, 6, 1, 4, 7, 2, 5, 8]) * 40.0)
    # two segments: each radical inverse scaled to 180 degrees, then +180
    assert np.allclose(stt.lds_sequence(6, 2, 2), [0, 180, 90, 270, 45, 225])
    # a 0-degree reference projection before every 3rd entry (at lengths 3, 6, ...)
    assert np.allclose(stt.lds_sequence(8, 1, 2, 3), [0, 180, 90, 0, 270, 45, 0, 225])
    # the reference's test script: 100 projections, Nz = 50 -> one insert at 50
    s = stt.lds_sequence(100, 1, 2, 50)
    assert s[50] == 0 and np.allclose(s[:50], stt.lds_sequence(50, 1, 2))


def test_synth_angles_vdc_is_the_lds_sequence():
    assert np.array_equal(stt.synth_angles(2, 60), stt.lds_sequence(60, 1, 2))
