"""Closed-form and invariant checks of the oracle parts that no reference run
pins (the reference's voxel backprojector, SART correction, primal-dual steps,
warp trace and ROI live in files that need headers absent from this image).

Each check is chosen so that a plausible slip fails it: a transposed axis, a
sign or an off-by-one in the detector mapping, a dropped term in the dual or
primal update, a wrong tap in the interpolation.
"""
import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from helpers import rel_l2, smooth_flow


def camera_coords(g, theta_deg, pts_vox):
    """Pinhole camera of the reference, written from the geometry: the volume
    rotates by theta about +y around its centre, then sits at camera offset
    (-ox, -oz, -sid + oy); the source is the camera origin looking down -z."""
    sp = np.array(list(g.spacing), np.float64)
    dims = np.array(list(g.vol), np.float64)
    m = (pts_vox - (dims - 1) / 2) * sp
    th = np.deg2rad(theta_deg)
    c, s = np.cos(th), np.sin(th)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    t = np.array([-g.oxyz[0], -g.oxyz[2], -g.sid + g.oxyz[1]])
    return m @ R.T + t


def detector_coords(g, e):
    halfW = (g.det[0] - 1) * g.ds / 2
    halfH = (g.det[1] - 1) * g.ds / 2
    px = (halfW - e[:, 0] / e[:, 2] * g.sdd) / g.ds
    py = (halfH - e[:, 1] / e[:, 2] * g.sdd) / g.ds
    return px, py


@pytest.mark.parametrize("kind", ["linear", "quadratic"])
def test_voxel_bp_reproduces_polynomial_images(kind):
    """Catmull-Rom reproduces polynomials up to degree 2, so backprojecting
    c(i, j) adds alpha * c(px, py) at each voxel's projected position."""
    g = stt.Geometry.make((14, 12, 16), (40, 36), (0.9, 1.1, 1.0), 0.8, 40.0, 75.0, (0.6, -0.4, 1.1))
    theta = 63.0
    W, H = g.det[0], g.det[1]
    jj, ii = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    if kind == "linear":
        img = 0.3 + 0.02 * ii - 0.015 * jj
        f = lambda x, y: 0.3 + 0.02 * x - 0.015 * y
    else:
        img = 0.1 + 1e-3 * (ii - 17.0) ** 2 + 5e-4 * (jj - 12.0) * (ii - 3.0)
        f = lambda x, y: 0.1 + 1e-3 * (x - 17.0) ** 2 + 5e-4 * (y - 12.0) * (x - 3.0)
    corr = img.astype(np.float32).reshape(-1)
    x0 = np.full(int(np.prod(list(g.vol))), 5.0, np.float32)
    out = O.bp_voxel(g, O.quat_y(theta), corr, 0.25, 1e9, x0)
    z, y, x = np.meshgrid(*(np.arange(n) for n in (g.vol[2], g.vol[1], g.vol[0])), indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64)
    px, py = detector_coords(g, camera_coords(g, theta, pts))
    inside = (px >= 2) & (px < W - 2) & (py >= 2) & (py < H - 2)
    away = inside & (px > 2.01) & (px < W - 2.01) & (py > 2.01) & (py < H - 2.01)
    assert away.sum() > 0.3 * away.size
    np.testing.assert_allclose(out[away], 5.0 + 0.25 * f(px[away], py[away]), rtol=0, atol=2e-5)
    outside = ~((px >= 1.99) & (px < W - 1.99) & (py >= 1.99) & (py < H - 1.99))
    assert np.all(out[outside] == 5.0)


def test_forward_projection_of_gaussian_matches_line_integral():
    """FP of a Gaussian blob against its analytic line integral
    sqrt(2 pi) s exp(-d^2 / 2 s^2) (d = distance from the blob centre to the ray)."""
    g = stt.Geometry.make((48, 48, 48), (64, 64), (1.0, 1.0, 1.0), 0.9, 80.0, 140.0, (1.5, -1.0, 2.0))
    theta, s = 221.0, 4.0
    dims = np.array(list(g.vol))
    z, y, x = np.meshgrid(*(np.arange(n) for n in dims[::-1]), indexing="ij")
    r2 = ((x - 23.5) ** 2 + (y - 23.5) ** 2 + (z - 23.5) ** 2).astype(np.float64)
    vol = np.exp(-r2 / (2 * s * s)).astype(np.float32).reshape(-1)
    ax, rl, w = O.fp(g, O.quat_y(theta), vol)
    jj, ii = np.meshgrid(np.arange(g.det[1]), np.arange(g.det[0]), indexing="ij")
    halfW, halfH = (g.det[0] - 1) * g.ds / 2, (g.det[1] - 1) * g.ds / 2
    u = np.stack([ii.ravel() * g.ds - halfW, jj.ravel() * g.ds - halfH, np.full(ii.size, -g.sdd)], 1)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    c = camera_coords(g, theta, np.array([[23.5, 23.5, 23.5]]))[0]
    d2 = np.sum((c - (u @ c)[:, None] * u) ** 2, axis=1)
    exact = np.sqrt(2 * np.pi) * s * np.exp(-d2 / (2 * s * s))
    core = exact > 0.3 * exact.max()
    assert core.sum() > 50
    assert rel_l2(ax[core], exact[core]) < 1.5e-2
    # the peak sits on the projected blob centre (orientation / offset of the detector)
    pk = np.argmax(ax)
    pc = detector_coords(g, c[None, :])
    assert abs(pk % g.det[0] - pc[0][0]) <= 1.0 and abs(pk // g.det[0] - pc[1][0]) <= 1.0


def test_forward_projection_constant_volume_counts_samples():
    g = stt.Geometry.make((20, 16, 18), (36, 30), (1.0, 1.0, 1.0), 0.7, 45.0, 80.0)
    c = 0.625
    ax, rl, w = O.fp(g, O.quat_y(17.0), np.full(20 * 16 * 18, c, np.float32))
    sd = np.float32(np.sqrt(3.0))
    hit = w > 0
    np.testing.assert_allclose(ax[hit], c * sd * w[hit], rtol=2e-6)
    # entry point counted twice: at least 1 + ceil(len / sd) samples
    assert np.all(w[hit] >= 1 + np.ceil(rl[hit] / sd))


def test_correction_formula():
    rng = np.random.default_rng(1)
    b, ax, rl = (rng.random(500, dtype=np.float32) for _ in range(3))
    for factor in (1.0, np.sqrt(0.2)):
        out = np.zeros(500, np.float32)
        O.orc.o_correction_f32(500, factor, O.f32(b), O.f32(ax), O.f32(rl), O.f32(out))
        b64, a64, r64 = (v.astype(np.float64) for v in (b, ax, rl))
        np.testing.assert_allclose(out, (factor * b64 - factor * a64) / (factor * r64 + 1), rtol=1e-6, atol=1e-7)


def test_warp_zero_flow_is_identity_with_edge_replication():
    dims = (10, 12, 9)
    src = np.random.default_rng(2).random(int(np.prod(dims)), dtype=np.float32)
    out = O.warp(dims, np.zeros(3 * src.size, np.float32), src, -1.0).reshape(dims[::-1])
    s = src.reshape(dims[::-1])
    iz, iy, ix = (np.clip(np.arange(n), 1, n - 2) for n in dims[::-1])
    assert np.array_equal(out, s[np.ix_(iz, iy, ix)])


@pytest.mark.parametrize("interp", [0, 1])
def test_warp_constant_integer_shift(interp):
    dims = (16, 14, 12)
    V = int(np.prod(dims))
    src = np.random.default_rng(3).random(V, dtype=np.float32)
    uvw = np.zeros(3 * V, np.float32)
    uvw[:V] = 2.0     # u
    uvw[2 * V:] = -1.0  # w
    out = O.warp(dims, uvw, src, -1.0, interp).reshape(dims[::-1])  # backward warp samples at x + u
    s = src.reshape(dims[::-1])
    # interior voxels whose displaced point stays inside [1, n-2]
    assert np.array_equal(out[2:10, 1:13, 1:12], s[1:9, 1:13, 3:14])
    out_f = O.warp(dims, uvw, src, +1.0, interp).reshape(dims[::-1])  # forward warp samples at x - u
    assert np.array_equal(out_f[1:10, 1:13, 3:14], s[2:11, 1:13, 1:12])


def test_warp_reproduces_quadratic_under_fractional_shift():
    """Clamped Catmull-Rom is exact for quadratics (the 8-corner clamp is
    inactive for a monotone field) and trilinear for linear fields."""
    dims = (20, 18, 16)
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in dims[::-1]), indexing="ij")
    V = x.size
    uvw = np.zeros(3 * V, np.float32)
    uvw[:V] = 0.375
    uvw[V:2 * V] = -0.25
    lin = (0.5 + 0.1 * x + 0.05 * y - 0.02 * z).astype(np.float32).reshape(-1)
    quad = (0.2 + 0.01 * x * x + 0.02 * y).astype(np.float32).reshape(-1)
    for interp, img, fn in ((1, lin, lambda a, b, c: 0.5 + 0.1 * a + 0.05 * b - 0.02 * c),
                            (0, quad, lambda a, b, c: 0.2 + 0.01 * a * a + 0.02 * b)):
        out = O.warp(dims, uvw, img, -1.0, interp).reshape(dims[::-1])
        sl = (slice(2, -2), slice(2, -2), slice(2, -2))
        expect = fn(x + 0.375, y - 0.25, z)
        np.testing.assert_allclose(out[sl], expect[sl], atol=2e-5)


def _affine_warp_case(dims, seed):
    """An affine velocity v(p) = A p + b (world = cell index + 0.5, voxel
    units) and a linear image I(g) = c.g + d on the cell grid.  Trilinear and
    clamped Catmull-Rom both reproduce linear data exactly and the 8-corner
    clamp never binds on it, so warping.h:100-114's error-compensated trace
    has a closed form: with tau = -dt,
        p2 = p + tau v(p),  p0 = p2 - tau v(p2) = p - tau^2 A v(p),
        x* = p2 + (p - p0)/2 = p + tau v(p) + tau^2 A v(p) / 2."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-0.06, 0.06, (3, 3))
    b = rng.uniform(-1.5, 1.5, 3)
    c = rng.uniform(-0.5, 0.5, 3)
    d = 0.7
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in dims[::-1]), indexing="ij")
    P = np.stack([x, y, z], -1) + 0.5                      # cell centres, world coordinates
    v = P @ A.T + b
    uvw = np.concatenate([v[..., k].reshape(-1) for k in range(3)]).astype(np.float32)
    img = ((np.stack([x, y, z], -1) @ c) + d).astype(np.float32).reshape(-1)
    return A, b, c, d, P, uvw, img


@pytest.mark.parametrize("interp", [0, 1])
@pytest.mark.parametrize("dt", [-1.0, 1.0])
def test_warp_trace_affine_flow_closed_form(interp, dt):
    """A non-uniform (affine) flow makes v(p2) differ from v(p), so the
    compensation term (p - p0)/2 is active.  Dropping it, flipping its sign
    or sampling v(p) twice moves x* by tau^2 |A v|/2 or more, about 0.05
    voxels here, against an fp32 tolerance of 2e-5."""
    dims = (22, 20, 18)
    A, b, c, d, P, uvw, img = _affine_warp_case(dims, seed=11)
    tau = -dt
    vp = P @ A.T + b
    xstar = P + tau * vp + 0.5 * tau * tau * (vp @ A.T)
    expect = (xstar - 0.5) @ c + d                          # I at grid coordinate x* - 0.5
    p2 = P + tau * vp
    # only voxels whose trace stays where no coordinate clamp ([1, dim-2]) binds
    lim = np.array(dims, np.float64) - 2
    ok = np.ones(P.shape[:3], bool)
    for q in (P, p2, xstar):
        g = q - 0.5
        ok &= np.all((g >= 1.05) & (g <= lim - 0.05), -1)
    assert ok.sum() > 1500
    # the three plausible slips all move the result well beyond tolerance
    wrong = {"dropped": P + tau * vp,
             "flipped": P + tau * vp - 0.5 * tau * tau * (vp @ A.T)}
    for name, q in wrong.items():
        assert np.abs(((q - 0.5) @ c + d) - expect)[ok].max() > 50 * 2e-5, name
    for f64mode in (False, True):
        out = O.warp(dims, uvw, img, dt, interp, f64mode).reshape(dims[::-1])
        err = np.abs(out - expect)[ok].max()
        assert err <= (2e-5 if not f64mode else 2e-6), (f64mode, err)


def test_bp_ray_apply_closed_form():
    """x += contrib / wsum, then the reference's macro clamp max(min(G, x), 0)
    (ST_Tomography.cpp:746-762, :167-173).  A voxel no ray reached has
    contrib = wsum = 0: 0/0 is NaN, min(G, NaN) keeps NaN and max(NaN, 0)
    gives 0, so it is zeroed (not left alone)."""
    rng = np.random.default_rng(5)
    n, G = 4000, 2.5
    x = rng.uniform(0, G, n).astype(np.float32)
    w = rng.uniform(0.05, 4.0, n).astype(np.float32)
    c = (w * rng.normal(0, 1.0, n)).astype(np.float32)
    w[::7] = 0.0
    c[::7] = 0.0                       # unreached voxels
    w[3::11] = 0.0
    c[3::11] = np.where(np.arange(n)[3::11] % 2 == 0, 1.0, -1.0)  # +-x/0: +-inf
    out = x.copy()
    assert O.orc.o_bp_ray_apply_f32(n, O.f32(c), O.f32(w), G, O.f32(out)) == 0
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = c.astype(np.float64) / w
    expect = np.clip(x + ratio, 0.0, G)
    expect[np.isnan(ratio)] = 0.0
    unreached = (w == 0) & (c == 0)
    assert unreached.sum() > 400 and np.all(out[unreached] == 0.0)
    assert np.all(out[(w == 0) & (c > 0)] == np.float32(G))
    assert np.all(out[(w == 0) & (c < 0)] == 0.0)
    np.testing.assert_allclose(out, expect, rtol=1e-6, atol=1e-6)
    # no alpha: a unit ratio moves an interior voxel by exactly one
    one = np.array([1.25], np.float32)
    assert O.orc.o_bp_ray_apply_f32(1, O.f32(np.array([0.5], np.float32)), O.f32(np.array([0.5], np.float32)),
                                    G, O.f32(one)) == 0 and one[0] == np.float32(2.25)


def _pd_ctx(dims, params, F=2):
    g = stt.Geometry.make(dims, (16, 16))
    oc = O.OracleCtx(g, params, [1] * F)
    for f in range(F):
        oc.set_views(f, [0.0])
    return oc


def test_dual_step_closed_form():
    """Without active clamps the dual step adds sigma*h*grad+(X_bar) on the
    interior ROI and sigma*(X_bar[f+1] - X_bar[f]) / sigma*(warp - X_bar[f])
    everywhere, then scales Y_tp by k/(k + 2 sigma) on the interior ROI."""
    dims = (9, 8, 10)
    V = int(np.prod(dims))
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber = 0.3, 0.5, 1e6, 0.7, 1e6, 2.0
    p.roi_cut[:] = [2, 0, 0, 0, 0, 3, 0, 0]
    oc = _pd_ctx(dims, p)
    rng = np.random.default_rng(4)
    x0, x1 = rng.random(V, dtype=np.float32), rng.random(V, dtype=np.float32)
    oc.set_volume(0, x0)
    oc.set_volume(1, x1)
    oc.pd_init()
    oc.pd_dual(0)
    X0 = x0.reshape(dims[::-1]).astype(np.float64)
    X1 = x1.reshape(dims[::-1]).astype(np.float64)
    sig, h = 0.3, 1.0 / (1.0 + 0.3 * 2.0 / 1e6)
    nz, ny, nx = dims[::-1]
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    inner = (i >= 1) & (i < nx - 1) & (j >= 1) & (j < ny - 1) & (k >= 1) & (k < nz - 1)
    inner &= ~((i < 2) | (i > nx - 0) | (k > nz - 3))
    gx = np.zeros_like(X0); gy = np.zeros_like(X0); gz = np.zeros_like(X0)
    gx[:, :, :-1] = X0[:, :, 1:] - X0[:, :, :-1]
    gy[:, :-1, :] = X0[:, 1:, :] - X0[:, :-1, :]
    gz[:-1, :, :] = X0[1:, :, :] - X0[:-1, :, :]
    ytv = oc.get_buffer(stt.STT_BUF_YTV, 0).reshape(4, nz, ny, nx)
    assert np.all(ytv[3] == 0)  # STV's temporal plane is unused by ATV
    for c, gg in enumerate((gx, gy, gz)):
        np.testing.assert_allclose(ytv[c], np.where(inner, sig * h * gg, 0.0), atol=1e-6)
    ytp = oc.get_buffer(stt.STT_BUF_YTP, 0).reshape(nz, ny, nx)
    base = sig * (X1 - X0)
    np.testing.assert_allclose(ytp, np.where(inner, base * 0.7 / (0.7 + 0.6), base), atol=1e-6)
    # zero flow: the backward warp is X_bar[1] with edge replication
    s = X1[np.ix_(*(np.clip(np.arange(n), 1, n - 2) for n in (nz, ny, nx)))]
    ybc = oc.get_buffer(stt.STT_BUF_YBC, 0).reshape(nz, ny, nx)
    np.testing.assert_allclose(ybc, sig * (s - X0), atol=1e-6)


def test_dual_prox_clamps():
    dims = (8, 8, 8)
    V = 512
    p = stt.params_default()
    p.sigma, p.lambda_tv, p.k_tp, p.e_bc, p.huber = 5.0, 0.05, 0.0, 0.02, 0.0
    oc = _pd_ctx(dims, p)
    rng = np.random.default_rng(5)
    oc.set_volume(0, rng.random(V, dtype=np.float32))
    oc.set_volume(1, rng.random(V, dtype=np.float32) * 3)
    oc.pd_init()
    oc.pd_dual(0)
    inner = np.zeros((8, 8, 8), bool)
    inner[1:7, 1:7, 1:7] = True
    ytv = oc.get_buffer(stt.STT_BUF_YTV, 0).reshape(4, 8, 8, 8)[:3]
    assert np.abs(ytv).max() <= 0.05 + 1e-7 and np.isclose(np.abs(ytv[:, inner]).max(), 0.05)
    ytp = oc.get_buffer(stt.STT_BUF_YTP, 0).reshape(8, 8, 8)
    assert np.all(ytp[inner] == 0)          # k = 0 removes the temporal dual
    ybc = oc.get_buffer(stt.STT_BUF_YBC, 0).reshape(8, 8, 8)
    assert np.abs(ybc[inner]).max() <= 0.02 + 1e-7
    assert np.abs(ybc[~inner]).max() > 0.02  # no prox outside the interior


def test_primal_tv_term_is_minus_adjoint_of_forward_difference():
    """<grad+ u, Y> = -<u, div- Y> for u zero on the boundary and Y zero off
    the interior: the primal step's TV term is the exact adjoint of the dual's
    forward difference."""
    dims = (9, 10, 8)
    nz, ny, nx = dims[::-1]
    V = nx * ny * nz
    p = stt.params_default()
    p.tau, p.sigma, p.k_tp, p.e_bc, p.g_max = 0.5, 0.0, 0.0, 0.0, 1e6
    oc = _pd_ctx(dims, p, F=1)
    rng = np.random.default_rng(6)
    Y = rng.standard_normal((3, nz, ny, nx)).astype(np.float32)
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    Y[:, ~inner] = 0
    x = np.full(V, 50.0, np.float32)
    oc.set_volume(0, x)
    oc.pd_init()
    oc.set_buffer(stt.STT_BUF_YTV, 0, np.concatenate([Y.reshape(-1), np.zeros(V, np.float32)]))
    oc.pd_primal(0)  # sigma = 0 duals: only the TV term moves x
    xn = oc.get_volume(0).reshape(nz, ny, nx).astype(np.float64)
    divY = np.where(inner, (xn - 50.0) / 0.5, 0.0)  # x -= tau * (-div Y)
    u = rng.standard_normal((nz, ny, nx))
    u[~inner] = 0
    g = [np.zeros_like(u) for _ in range(3)]
    g[0][:, :, :-1] = u[:, :, 1:] - u[:, :, :-1]
    g[1][:, :-1, :] = u[:, 1:, :] - u[:, :-1, :]
    g[2][:-1, :, :] = u[1:, :, :] - u[:-1, :, :]
    lhs = sum(float(np.sum(g[c][inner] * Y[c][inner])) for c in range(3))
    rhs = -float(np.sum(u * divY))
    assert abs(lhs - rhs) <= 1e-4 * max(1.0, abs(lhs))


def _grad4(X0, X1):
    """Forward differences of frame X0 along x, y, z and to the next frame."""
    g = [np.zeros_like(X0) for _ in range(4)]
    g[0][:, :, :-1] = X0[:, :, 1:] - X0[:, :, :-1]
    g[1][:, :-1, :] = X0[:, 1:, :] - X0[:, :-1, :]
    g[2][:-1, :, :] = X0[1:, :, :] - X0[:-1, :, :]
    g[3] = X1 - X0
    return g


def test_stv_dual_projects_spacetime_gradient_on_l2_ball():
    """From zero duals, STV's dual is sigma*h*(grad_xyz, d_t) of X_bar scaled
    onto the lambda L2 ball of the 4-vector; ATV leaves the temporal plane at 0."""
    dims = (9, 8, 10)
    nz, ny, nx = dims[::-1]
    V = nx * ny * nz
    p = stt.params_default()
    p.prior = stt.STT_PRIOR_STV
    p.sigma, p.lambda_tv, p.k_tp, p.e_bc, p.huber = 0.8, 0.3, 0.0, 1e6, 0.5
    oc = _pd_ctx(dims, p)
    rng = np.random.default_rng(14)
    x0, x1 = rng.random(V, dtype=np.float32), rng.random(V, dtype=np.float32)
    oc.set_volume(0, x0)
    oc.set_volume(1, x1)
    oc.pd_init()
    oc.pd_dual(0)
    X0 = x0.reshape(nz, ny, nx).astype(np.float64)
    X1 = x1.reshape(nz, ny, nx).astype(np.float64)
    h = 1.0 / (1.0 + 0.8 * 0.5 / 0.3)
    g = np.stack([0.8 * h * c for c in _grad4(X0, X1)])
    nrm = np.sqrt((g ** 2).sum(axis=0))
    expect = g * np.minimum(1.0, 0.3 / np.maximum(nrm, 1e-30))
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    ytv = oc.get_buffer(stt.STT_BUF_YTV, 0).reshape(4, nz, ny, nx)
    np.testing.assert_allclose(ytv, np.where(inner, expect, 0.0), atol=2e-6)
    # both regimes occur: some 4-vectors inside the ball, some projected onto it
    assert (nrm[inner] < 0.3).any() and (nrm[inner] > 0.3).any()
    assert np.abs(np.sqrt((ytv[:, inner] ** 2).sum(axis=0)) - np.minimum(nrm[inner], 0.3)).max() < 1e-6


def test_stv_reduces_to_isotropic_tv_on_a_static_sequence():
    dims = (8, 9, 7)
    nz, ny, nx = dims[::-1]
    V = nx * ny * nz
    rng = np.random.default_rng(15)
    x0 = rng.random(V, dtype=np.float32)
    out = {}
    for prior in (stt.STT_PRIOR_TV, stt.STT_PRIOR_STV):
        p = stt.params_default()
        p.prior = prior
        p.sigma, p.lambda_tv, p.k_tp, p.e_bc, p.huber = 2.0, 0.1, 0.0, 1e6, 0.0
        oc = _pd_ctx(dims, p)
        oc.set_volume(0, x0)
        oc.set_volume(1, x0)
        oc.pd_init()
        oc.pd_dual(0)
        out[prior] = oc.get_buffer(stt.STT_BUF_YTV, 0).reshape(4, -1)
    np.testing.assert_array_equal(out[stt.STT_PRIOR_STV], out[stt.STT_PRIOR_TV])
    assert np.abs(out[stt.STT_PRIOR_TV][:3]).max() > 0.05


def test_stv_primal_temporal_term_is_adjoint_of_forward_time_difference():
    """<D_t u, y> = <u, D_t^T y> with (D_t u)[f] = u[f+1] - u[f] (the padding
    frame is constant), where D_t^T y is read off the primal step of every
    frame: x[f] -= tau * (y[f-1] - y[f])."""
    dims = (7, 8, 6)
    nz, ny, nx = dims[::-1]
    V = nx * ny * nz
    F = 3
    p = stt.params_default()
    p.prior = stt.STT_PRIOR_STV
    p.tau, p.sigma, p.k_tp, p.e_bc, p.g_max = 0.25, 0.0, 0.0, 0.0, 1e6
    oc = _pd_ctx(dims, p, F=F)
    rng = np.random.default_rng(16)
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    for f in range(F + 1):
        oc.set_volume(f, np.full(V, 40.0, np.float32))
    oc.pd_init()
    Yt = rng.standard_normal((F, nz, ny, nx)).astype(np.float32)
    Yt[:, ~inner] = 0
    for f in range(F):
        oc.set_buffer(stt.STT_BUF_YTV, f, np.concatenate([np.zeros(3 * V, np.float32), Yt[f].reshape(-1)]))
    adj = []
    for f in range(F):
        oc.pd_primal(f)
        adj.append((40.0 - oc.get_volume(f).reshape(nz, ny, nx).astype(np.float64)) / 0.25)
    u = rng.standard_normal((F, nz, ny, nx))
    u[:, ~inner] = 0
    Dtu = [(u[f + 1] if f + 1 < F else 0.0) - u[f] for f in range(F)]
    lhs = sum(float(np.sum(Dtu[f] * Yt[f])) for f in range(F))
    rhs = sum(float(np.sum(u[f] * adj[f])) for f in range(F))
    assert abs(lhs - rhs) <= 1e-4 * max(1.0, abs(lhs))
    # and the term is zero off the interior
    assert all(np.abs(a[~inner]).max() == 0 for a in adj)


def test_roi_mask_rose_cut():
    dims = (64, 64, 64)
    m = O.roi_mask(dims, [8, 8, 8, 45, 8, 8, 30, 20], 7).reshape(64, 64, 64)  # [z, y, x]
    assert m[32, 32, 32] == 1
    assert m[32, 32, 7] == 0 and m[32, 32, 8] == 1 and m[32, 32, 56] == 1 and m[32, 32, 57] == 0   # i in [8, 56]
    assert m[32, 7, 32] == 0 and m[32, 57, 32] == 1 and m[32, 58, 32] == 0                        # j in [8, 57]
    assert m[10, 32, 29] == 0 and m[10, 32, 30] == 1                                              # x-z chamfer
    assert m[32, 32, 32] == 1 and m[29, 32, 10] == 0 and m[30, 32, 10] == 1
    full = O.roi_mask(dims, [0] * 8, 0)
    assert np.all(full == 1)


def test_sart_sweep_moves_towards_data():
    """One SART sweep lowers the data residual of a consistent problem."""
    g = stt.Geometry.make((20, 20, 20), (32, 32), (1, 1, 1), 1.0, 60.0, 100.0)
    p = stt.params_default()
    p.alpha, p.g_max = 0.3, 100.0
    P = 12
    ang = list(np.linspace(0, 360, P, endpoint=False))
    oc = O.OracleCtx(g, p, [P])
    oc.set_views(0, ang)
    truth = stt.synth_phantom((20, 20, 20), 5, seed=3)
    b = np.concatenate([O.fp(g, O.quat_y(a), truth)[0] for a in ang])
    oc.set_projections(0, b)
    oc.set_volume(0, np.zeros(8000, np.float32))

    def resid(x):
        return float(np.linalg.norm(np.concatenate([O.fp(g, O.quat_y(a), x)[0] for a in ang]) - b))

    r0 = resid(np.zeros(8000, np.float32))
    oc.sart(0)
    r1 = resid(oc.get_volume(0))
    oc.sart(0)
    r2 = resid(oc.get_volume(0))
    assert r1 < 0.6 * r0 and r2 < r1
