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


def _fwd(a, dims, ax):
    """tools.cpp forward difference along x/y/z (0 at the last index)."""
    a = a.reshape(dims[::-1])
    out = np.zeros_like(a)
    axis = 2 - ax
    lo, hi = [slice(None)] * 3, [slice(None)] * 3
    lo[axis], hi[axis] = slice(0, -1), slice(1, None)
    out[tuple(lo)] = a[tuple(hi)] - a[tuple(lo)]
    return out.ravel()


def _bwd_interior(a, dims, ax):
    """a[i] - a[i-1] (the backward difference away from the first index)."""
    a = a.reshape(dims[::-1])
    out = np.zeros_like(a)
    axis = 2 - ax
    lo, hi = [slice(None)] * 3, [slice(None)] * 3
    lo[axis], hi[axis] = slice(0, -1), slice(1, None)
    out[tuple(hi)] = a[tuple(hi)] - a[tuple(lo)]
    return out.ravel()


@pytest.mark.parametrize("norm", [4, 1])
def test_of_level_fixed_point_has_a_zero_duality_gap(norm):
    """The solver loop (optical_flow.h:300-480) is a primal-dual iteration for
        min_v  sum_c lam*H_eps(|D v_c|) + sum_ROI |grad(I2w) . v + ut|,
    the Huber-TV / L1 energy its prox steps define: the TV duals are scaled
    by 1/(1 + sigma1*eps/lam) and projected on the lam-ball per flow
    component (:441-455; eps = 0 without Huber, norm != 4), the data dual is
    clamped to [-1, 1] with the linearisation ut (:264, :458).  At its fixed
    point every Fenchel-Young gap vanishes,
        lam*H_eps(|z_c|) + eps/(2 lam)|y_c|^2 - y_c . z_c = 0,  z_c = D v_c,
        |z4 + ut| - y4 (z4 + ut) = 0,                         z4 = grad(I2w) . v,
    and the primal step stops moving (K^T y = 0 on the ROI).  The energy and
    the gaps are evaluated here from the returned v and y, independently of
    the oracle's update code: a wrong Huber factor, a lost data term or a
    flipped divergence leaves a gap of order one, and so does a perturbed v
    (both checked).  One warp, so the linearisation is that of the input
    flow (zero)."""
    dims = (14, 12, 10)
    n = int(np.prod(dims))
    roi = (2, 2, 2, 2, 2, 2)
    lam, eps = 0.05, 0.1
    I1, I2 = _smooth_pair(dims, (0.5, -0.3, 0.2))
    V0 = np.zeros(3 * n, np.float32)

    def solve(huber):
        V, Y, _, _ = O.of_level(dims, np.asarray(dims, np.float32), I1, I2, V0, np.zeros(10 * n, np.float32), tol=0.0,
                                lam=lam, max_iter=1000, norm=norm, warps=1, huber=huber, roi=roi)
        return V.astype(np.float64).reshape(3, n), Y.astype(np.float64).reshape(10, n)

    W = O.warp_ext(dims, np.asarray(dims, np.float32), V0, I2, -1.0)
    box = np.zeros(dims[::-1], bool)
    box[roi[4]:dims[2] - roi[5], roi[2]:dims[1] - roi[3], roi[0]:dims[0] - roi[1]] = True
    inner = np.zeros(dims[::-1], bool)
    inner[roi[4] + 1:dims[2] - roi[5] - 1, roi[2] + 1:dims[1] - roi[3] - 1, roi[0] + 1:dims[0] - roi[1] - 1] = True
    box, inner = box.ravel(), inner.ravel()
    u = np.stack([g.astype(np.float64) * box for g in O.gradient_central(dims, W)])
    ut = (W.astype(np.float64) - I1) * box  # the input flow is zero
    e_judge = eps if norm == 4 else 0.0

    def certificate(v, y):
        gap = 0.0
        energy = 0.0
        for c in range(3):
            z = np.stack([_fwd(v[c], dims, ax) for ax in range(3)])
            yc = y[3 * c:3 * c + 3]
            s = np.sqrt((z * z).sum(0))
            h = np.where(s <= e_judge, s * s / (2 * e_judge) if e_judge else s, s - e_judge / 2)
            energy += lam * h[box].sum()
            fy = lam * h + (e_judge / (2 * lam)) * (yc * yc).sum(0) - (yc * z).sum(0)
            gap += fy[box].sum()
        r = (u * v).sum(0) + ut
        energy += np.abs(r[box]).sum()
        gap += (np.abs(r) - y[9] * r)[box].sum()
        kty = np.stack([-(_bwd_interior(y[3 * c], dims, 0) + _bwd_interior(y[3 * c + 1], dims, 1) +
                          _bwd_interior(y[3 * c + 2], dims, 2)) + u[c] * y[9] for c in range(3)])
        return gap / energy, float(np.abs(kty[:, inner]).max())

    v, y = solve(eps)
    assert np.abs(y[9][box]).max() <= 1.0 + 1e-6
    for c in range(3):  # the TV duals lie in the lam-ball
        assert np.sqrt((y[3 * c:3 * c + 3][:, box] ** 2).sum(0)).max() <= lam * (1 + 1e-5)
    rel_gap, feas = certificate(v, y)
    print(f"norm {norm}: relative duality gap {rel_gap:.2e}, |K^T y| {feas:.2e}")
    assert abs(rel_gap) <= 1e-5 and feas <= 1e-5
    # sensitivity: a perturbed flow, and (with Huber) the fixed point of another epsilon
    rng = np.random.default_rng(0)
    assert certificate(v + 0.02 * rng.standard_normal(v.shape) * box, y)[0] > 1e-2
    if norm == 4:
        assert certificate(*solve(0.5))[0] > 1e-2


@needs_ref
@pytest.mark.parametrize("norm,lam", [(4, 0.05), (1, 0.2)])
def test_of_level_stopping_residual_is_the_primal_dual_residual(norm, lam):
    """The solver's convergence test (optical_flow.h:359-362, 465-486,
    500-502) is the primal-dual residual of adaptive PDHG (Goldstein, Li,
    Yuan, Esser, Baraniuk, "Adaptive primal-dual splitting methods for
    statistical learning and image processing", 2015) with the loop's
    diagonal steps: for the iterates v^k, y^k after iteration k, at every
    50th iteration
        P = (v^{k-1} - v^k) / tau - K^T (y^{k-2} - y^{k-1})   (the primal step of
            iteration k reads y^{k-1}; on the ROI),
        D = (y^{k-1} - y^k) / sigma - K (v^{k-1} - v^k)   (every voxel; sigma1 on
            the nine TV rows, sigma2 on the data row),
        err = (|D|_1^2 + |P|_1^2) / n.
    Recomputed here in FP64 from the oracle's iterates after 48, 49 and 50
    iterations, with K = (forward differences of each flow component, grad(I2w)
    . v) assembled from the differences pinned against the reference's
    tools.cpp and CImg's central gradient; tau = 1/(4 + |grad_c|), sigma1 =
    1/3, sigma2 = 1/sum|grad| (optical_flow.h:169, 285-291).  Readings that
    differ in one term -- the data row dropped from D, sigma2 on the TV rows,
    P without its K^T term, squares instead of absolute values -- move err by
    far more than the tolerance, which the test checks too."""
    dims = (10, 9, 8)
    n = int(np.prod(dims))
    ext = np.asarray(dims, np.float32)
    I1, I2 = _smooth_pair(dims, (0.5, -0.3, 0.2))
    V0 = np.zeros(3 * n, np.float32)
    Y0 = np.zeros(10 * n, np.float32)
    roi = (0,) * 6  # the write-back covers every voxel, so the returned iterates are the loop's

    def run(max_iter):
        V, Y, its, errs = O.of_level(dims, ext, I1, I2, V0, Y0, tol=0.0, lam=lam, max_iter=max_iter, norm=norm,
                                     warps=1, huber=0.1, roi=roi)
        assert its[0] == max_iter + 1  # while (err > tol && iterations <= maxIterations) ++iterations
        return V.astype(np.float64).reshape(3, n), Y.astype(np.float64).reshape(10, n), float(errs[0])

    _, y48, _ = run(47)
    v49, y49, _ = run(48)
    v50, y50, err = run(49)
    W = O.warp_ext(dims, ext, V0, I2, -1.0)
    u = np.stack([g.astype(np.float64) for g in O.gradient_central(dims, W)])
    tau = 1.0 / (4.0 + np.abs(u))
    # sigma2 is infinite where the warped image is flat (the volume's corners,
    # whose clamped warp samples equal their neighbours'): the data row's
    # (y4 change) / sigma2 is 0 there, as in the loop
    with np.errstate(divide="ignore"):
        sigma1, sigma2 = 1.0 / 3.0, 1.0 / np.abs(u).sum(0)

    def diff(a, which):
        return O.of_diff(dims, np.ascontiguousarray(a, np.float32), which).astype(np.float64)

    def K(v):  # 10 rows: D_x, D_y, D_z of each component, then grad(I2w) . v
        rows = [diff(v[c], ax) for c in range(3) for ax in range(3)]
        return np.stack(rows + [(u * v).sum(0)])

    def KT(y):  # -div of each component's TV dual + grad_c * y4
        return np.stack([-(diff(y[3 * c], 3) + diff(y[3 * c + 1], 4) + diff(y[3 * c + 2], 5)) + u[c] * y[9]
                         for c in range(3)])

    P = (v49 - v50) / tau - (KT(y48) - KT(y49))
    Dtv = (y49[:9] - y50[:9]) / sigma1 - (K(v49)[:9] - K(v50)[:9])
    Ddat = (y49[9] - y50[9]) / sigma2 - (K(v49)[9] - K(v50)[9])

    def residual(p, d):
        return (d * d + p * p) / n

    p1, d1 = np.abs(P).sum(), np.abs(Dtv).sum() + np.abs(Ddat).sum()
    ref = residual(p1, d1)
    rel = abs(err - ref) / ref
    print(f"norm {norm}: oracle err {err:.8e}, recomputed {ref:.8e}, relative difference {rel:.1e} "
          f"(|P|_1 {p1:.4e}, |D|_1 {d1:.4e})")
    tol = 1e-5  # fp32 iterates and per-voxel terms against their FP64 recomputation
    assert rel <= tol, (err, ref)
    variants = {
        "data row dropped": residual(p1, np.abs(Dtv).sum()),
        "sigma2 on the TV rows": residual(
            p1, np.abs((y49[:9] - y50[:9]) / sigma2 - (K(v49)[:9] - K(v50)[:9])).sum() + np.abs(Ddat).sum()),
        "P without K^T": residual(np.abs((v49 - v50) / tau).sum(), d1),
        "squares": residual((P * P).sum(), (Dtv * Dtv).sum() + (Ddat * Ddat).sum()),
    }
    for name, val in variants.items():
        assert abs(val - ref) > 100 * tol * ref, (name, val, ref)
