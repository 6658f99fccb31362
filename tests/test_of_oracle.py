"""Optical flow oracle (one pyramid level) against the reference's own
finite differences and CImg gradient (compiled into oracle/_ref), the
discrete gradient/divergence adjoint pair, and fixed points of the solver."""
import numpy as np
import pytest

import _oracle as O

needs_ref = pytest.mark.skipif(not O.HAVE_REF, reason="oracle/_ref not built (reference tree absent)")


def _vol(dims, seed):
    return np.random.default_rng(seed).standard_normal(int(np.prod(dims))).astype(np.float32)


@needs_ref
@pytest.mark.parametrize("which", range(6))
def test_differences_match_reference_tools(which):
    dims = (7, 6, 5)
    a = _vol(dims, which)
    ref = np.zeros_like(a)
    O.ref.r_of_diff(O.i32(np.asarray(dims, np.int32)), O.f32(a), which, O.f32(ref))
    assert np.array_equal(O.of_diff(dims, a, which), ref)


@pytest.mark.parametrize("axis", range(3))
def test_forward_backward_differences_are_adjoint(axis):
    """<dp a, b> = -<a, dm b>: tools.cpp's backward difference (first index
    b, last index -b[n-2]) is minus the adjoint of its forward difference."""
    dims = (9, 7, 6)
    a, b = _vol(dims, 10 + axis), _vol(dims, 20 + axis)
    lhs = float(np.dot(O.of_diff(dims, a, axis).astype(np.float64), b))
    rhs = -float(np.dot(a.astype(np.float64), O.of_diff(dims, b, 3 + axis)))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1.0), (lhs, rhs)


def test_central_gradient_closed_form():
    dims = (8, 7, 6)
    img = _vol(dims, 3)
    gx, gy, gz = O.gradient_central(dims, img)
    a = img.reshape(dims[::-1])  # (z, y, x)
    for g, ax in ((gx, 2), (gy, 1), (gz, 0)):
        pad = np.pad(a, [(1, 1) if k == ax else (0, 0) for k in range(3)], mode="edge")
        sl_p = [slice(None)] * 3
        sl_m = [slice(None)] * 3
        sl_p[ax] = slice(2, None)
        sl_m[ax] = slice(0, -2)
        exp = ((pad[tuple(sl_p)] - pad[tuple(sl_m)]) / np.float32(2)).astype(np.float32)
        assert np.array_equal(g.reshape(dims[::-1]), exp)


@needs_ref
def test_central_gradient_matches_cimg():
    dims = (8, 7, 6)
    img = _vol(dims, 4)
    ref = [np.zeros_like(img) for _ in range(3)]
    O.ref.r_gradient_central(O.i32(np.asarray(dims, np.int32)), O.f32(img), O.f32(ref[0]), O.f32(ref[1]), O.f32(ref[2]))
    for g, r in zip(O.gradient_central(dims, img), ref):
        assert np.array_equal(g, r)


def test_warp_ext_equals_warp_when_extent_is_dims():
    dims = (12, 10, 9)
    rng = np.random.default_rng(5)
    uvw = (rng.standard_normal(3 * int(np.prod(dims))) * 0.7).astype(np.float32)
    src = _vol(dims, 6)
    a = O.warp_ext(dims, np.asarray(dims, np.float32), uvw, src, -1.0)
    b = O.warp(dims, uvw, src, -1.0)
    assert np.array_equal(a, b)


def _smooth_pair(dims, shift):
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) + 0.5 for n in dims[::-1]], indexing="ij")

    def f(xx, yy, zz):
        return (np.sin(0.45 * xx) * np.cos(0.37 * yy) + 0.6 * np.sin(0.31 * zz + 0.2 * xx)).astype(np.float32)

    return f(x, y, z).ravel(), f(x - shift[0], y - shift[1], z - shift[2]).ravel()


def test_of_level_zero_flow_is_a_fixed_point_for_identical_images():
    dims = (14, 12, 10)
    I1, _ = _smooth_pair(dims, (0, 0, 0))
    n = int(np.prod(dims))
    V, Y, its, errs = O.of_level(dims, np.asarray(dims, np.float32), I1, I1, np.zeros(3 * n, np.float32),
                                 np.zeros(10 * n, np.float32), max_iter=120, warps=2)
    assert np.array_equal(V, np.zeros(3 * n, np.float32))
    assert np.all(np.abs(Y[: 9 * n]) == 0)


def test_of_level_recovers_a_translation():
    """I2(x) = I1(x - s): the warp samples I2(x + v), so the flow is v = s."""
    dims = (20, 18, 16)
    s = (0.6, -0.4, 0.3)
    I1, I2 = _smooth_pair(dims, s)
    n = int(np.prod(dims))
    V, _, its, _ = O.of_level(dims, np.asarray(dims, np.float32), I1, I2, np.zeros(3 * n, np.float32),
                              np.zeros(10 * n, np.float32), tol=1e-6, lam=0.05, max_iter=400, warps=3)
    inner = np.zeros(dims[::-1], bool)
    inner[3:-3, 3:-3, 3:-3] = True
    for c in range(3):
        m = float(V[c * n:(c + 1) * n].reshape(dims[::-1])[inner].mean())
        assert abs(m - s[c]) < 0.12, (c, m, s[c])


@needs_ref
@pytest.mark.parametrize("spectrum", [1, 3])
def test_blur_matches_cimg(spectrum):
    dims = (13, 9, 11)
    img = _vol(dims, 30)[:int(np.prod(dims))]
    img = np.tile(img, spectrum) * np.repeat(np.arange(1, spectrum + 1, dtype=np.float32), int(np.prod(dims)))
    ref = np.zeros_like(img)
    O.ref.r_blur(O.i32(np.asarray(dims, np.int32)), spectrum, O.f32(img), 0.5, O.f32(ref))
    assert np.array_equal(O.blur(dims, img, 0.5, spectrum), ref)


@needs_ref
@pytest.mark.parametrize("new", [(8, 6, 7), (20, 15, 18), (13, 5, 16), (166 // 7, 9, 11)])
def test_resize_matches_cimg(new):
    """get_resize(..., interpolation 5): moving average on shrinking axes, cubic on growing ones."""
    dims = (13, 9, 11)
    img = _vol(dims, 31)
    for spectrum in (1, 3):
        im = np.tile(img, spectrum)
        ref = np.zeros(int(np.prod(new)) * spectrum, np.float32)
        O.ref.r_resize(O.i32(np.asarray(dims, np.int32)), spectrum, O.f32(im), new[0], new[1], new[2], 5, O.f32(ref))
        assert np.array_equal(O.resize(dims, im, new, spectrum), ref), (new, spectrum)


def test_resize_moving_average_closed_form():
    """Halving an axis averages pairs; resizing a constant keeps it."""
    dims = (8, 4, 6)
    img = _vol(dims, 32)
    out = O.resize(dims, img, (4, 4, 6)).reshape(6, 4, 4)
    a = img.reshape(6, 4, 8)
    np.testing.assert_allclose(out, (a[:, :, 0::2] + a[:, :, 1::2]) / 2, rtol=1e-6)
    const = np.full(int(np.prod(dims)), 2.5, np.float32)
    for new in ((5, 3, 4), (11, 7, 9)):
        np.testing.assert_allclose(O.resize(dims, const, new), 2.5, rtol=1e-6)


def test_of_multiscale_level_sizes_and_translation():
    """Pyramid sizes int(dims * eta^k), coarsest first; the flow stays in
    full-resolution voxel units on every level (the warp keeps the finest
    extent), so a translation of the whole image is recovered as is."""
    dims = (24, 22, 20)
    s = (0.8, -0.6, 0.5)
    I1, I2 = _smooth_pair(dims, s)
    n = int(np.prod(dims))
    V, ld, its = O.of_multiscale(dims, I1, I2, np.zeros(3 * n, np.float32), lam=0.05, tol=1e-6, max_iter=150,
                                 warps=2, cut=(1, 1, 1, 1, 1, 1))
    assert ld.tolist() == [[10, 9, 8], [15, 14, 13], [24, 22, 20]]
    inner = np.zeros(dims[::-1], bool)
    inner[4:-4, 4:-4, 4:-4] = True
    for c in range(3):
        m = float(V[c * n:(c + 1) * n].reshape(dims[::-1])[inner].mean())
        assert abs(m - s[c]) < 0.15, (c, m, s[c])
