"""Warp adjoint, FP64 residual norm and power iteration on the MI355X.

The warp VJP has no reference counterpart.  It is checked as follows:
  - volume adjoint: a dot-product test against the GPU forward warp;
  - flow adjoint: central differences of the FP64 oracle warp, along a
    smooth direction (h = 1e-6) and on 64 random coordinates (h = 1e-3).
The residual norm is compared with the FP64 oracle projector's residual.
The power iteration is compared with the largest singular value of the
explicit coupling matrix K (numpy SVD) on tiny grids.
"""
import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from helpers import smooth_flow

pytestmark = pytest.mark.gpu


def _ctx(dims, F=1, P=None, params=None, det=(16, 16)):
    g = stt.Geometry.make(dims, det)
    p = params if params is not None else stt.params_default()
    return g, stt.Context(g, p, list(P) if P is not None else [1] * F)


@pytest.mark.parametrize("dt", [-1.0, 1.0])
def test_warp_vjp_volume_dot_product(dt):
    print(f"[warp-vjp-dot] dt={dt} 24x20x28", flush=True)
    dims = (24, 20, 28)
    V = int(np.prod(dims))
    g, ctx = _ctx(dims)
    rng = np.random.default_rng(31)
    uvw = smooth_flow(dims, 3.0, seed=32)
    src = rng.standard_normal(V).astype(np.float32)
    gout = rng.standard_normal(V).astype(np.float32)
    wsrc = ctx.warp(uvw, src, dt, stt.STT_INTERP_TRILINEAR)
    gsrc, _ = ctx.warp_vjp(uvw, src, dt, gout)
    lhs = float(np.dot(wsrc.astype(np.float64), gout))
    rhs = float(np.dot(src.astype(np.float64), gsrc))
    scale = float(np.linalg.norm(wsrc) * np.linalg.norm(gout))
    assert abs(lhs - rhs) <= 1e-5 * scale, (lhs, rhs, scale)
    ctx.close()


def _loss64(dims, uvw, src, dt, gout):
    out = O.warp(dims, uvw.astype(np.float64), src, dt, interp=stt.STT_INTERP_TRILINEAR, f64mode=True)
    return float(np.dot(out, gout.astype(np.float64)))


@pytest.mark.parametrize("dt", [-1.0, 1.0])
def test_warp_vjp_flow_fp64_finite_differences(dt):
    print(f"[warp-vjp-fd] dt={dt} 20x18x22", flush=True)
    dims = (20, 18, 22)
    V = int(np.prod(dims))
    g, ctx = _ctx(dims)
    rng = np.random.default_rng(33)
    # no flow value is an integer, so no trace point sits exactly on a grid
    # node where the trilinear interpolant has a kink
    uvw = (smooth_flow(dims, 3.0, seed=34) * np.float32(0.9731)
           + np.float32(1e-3) * rng.standard_normal(3 * V).astype(np.float32))
    src = smooth_flow(dims, 1.0, seed=35)[:V] + rng.random(V, dtype=np.float32)  # smooth + texture
    gout = rng.standard_normal(V).astype(np.float32)
    _, guvw = ctx.warp_vjp(uvw, src, dt, gout)
    # directional derivative along a smooth direction; a step of 1e-6 (FP64)
    # keeps the samples off the trilinear kinks that a whole-field step crosses
    delta = smooth_flow(dims, 1.0, seed=36).astype(np.float64)
    hd = 1e-6
    fd = (_loss64(dims, uvw + hd * delta, src, dt, gout) - _loss64(dims, uvw - hd * delta, src, dt, gout)) / (2 * hd)
    an = float(np.dot(guvw.astype(np.float64), delta))
    assert abs(fd - an) <= 1e-4 * abs(fd), (fd, an)
    # 64 random coordinates at h = 1e-3
    h = 1e-3
    coords = rng.choice(3 * V, 64, replace=False)
    errs = []
    for c in coords:
        e = np.zeros(3 * V)
        e[c] = 1.0
        fdc = (_loss64(dims, uvw + h * e, src, dt, gout) - _loss64(dims, uvw - h * e, src, dt, gout)) / (2 * h)
        errs.append(abs(fdc - guvw[c]) / max(1.0, abs(fdc)))
    errs = np.array(errs)
    # a coordinate whose perturbation moves a sample across a cell face sees
    # the trilinear kink; allow at most two such coordinates
    assert np.count_nonzero(errs > 1e-3) <= 2, errs
    ctx.close()


def test_residual_norm_matches_fp64_oracle():
    print("[residual-norm] 24^3 F=2 P=(4,3) det 32x32", flush=True)
    dims, det = (24, 24, 24), (32, 32)
    P = (4, 3)
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 60.0, 100.0)
    ctx = stt.Context(g, stt.params_default(), list(P))
    angles = [list(stt.synth_angles(1, P[f], offset=sum(P[:f]))) for f in range(2)]
    r2_ref = 0.0
    for f in range(2):
        truth = stt.synth_phantom_frame(dims, f, n_ellipsoids=6, seed=1, deform_seed=2, amp=1.5)
        ctx.set_views(f, angles[f])
        ctx.set_volume(f, truth)
        b = np.stack([ctx.forward_project(f, k)[0] for k in range(P[f])])
        ctx.set_projections(f, b)
        x = (0.5 * truth).astype(np.float32)
        ctx.set_volume(f, x)
        for k, a in enumerate(angles[f]):
            ax = O.fp(g, O.quat_y(a), x, f64mode=True)[0]
            r2_ref += float(np.sum((b[k].astype(np.float64) - ax) ** 2))
    r2 = ctx.residual_norm()
    assert r2_ref > 0
    assert abs(r2 - r2_ref) <= 1e-5 * r2_ref, (r2, r2_ref)
    ctx.close()


def _coupling_matrix(dims, F, cut):
    nx, ny, nz = dims
    V = nx * ny * nz

    def idx(f, x, y, z):
        return f * V + x + nx * (y + ny * z)

    rows = []
    for f in range(F):
        for c in range(4):
            for z in range(nz):
                for y in range(ny):
                    for x in range(nx):
                        r = np.zeros(F * V)
                        q = idx(f, x, y, z)
                        if c < 3:
                            inner = 1 <= x < nx - 1 and 1 <= y < ny - 1 and 1 <= z < nz - 1
                            cutout = (x < cut[0] or x > nx - cut[1] or y < cut[2] or y > ny - cut[3]
                                      or z < cut[4] or z > nz - cut[5])
                            if inner and not cutout:
                                nb = [(x + 1, y, z), (x, y + 1, z), (x, y, z + 1)][c]
                                r[idx(f, *nb)] += 1.0
                                r[q] -= 1.0
                        else:
                            if f + 1 < F:
                                r[idx(f + 1, x, y, z)] += 1.0
                            r[q] -= 1.0
                        rows.append(r)
    return np.array(rows)


@pytest.mark.parametrize("dims,F,cut", [((5, 6, 4), 2, [0] * 8), ((7, 6, 5), 3, [2, 1, 0, 2, 0, 0, 0, 0])])
def test_power_iteration_matches_svd(dims, F, cut):
    print(f"[power-iteration] dims={dims} F={F} cut={cut[:6]}", flush=True)
    p = stt.params_default()
    p.roi_cut[:] = cut
    g, ctx = _ctx(dims, F=F, params=p)
    L = ctx.power_iteration(3000)
    K = _coupling_matrix(dims, F, cut)
    smax = float(np.linalg.norm(K, 2))
    assert L <= smax * (1 + 1e-5), (L, smax)
    assert abs(L - smax) <= 1e-3 * smax, (L, smax)
    ctx.close()
