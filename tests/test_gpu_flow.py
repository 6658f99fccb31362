"""Optical flow (one pyramid level) on the MI355X against the oracle.

The oracle restates optical_flow.h:83-623 in fp32 with the reference's
expression order; GPU and oracle sum the convergence residual in FP64 in
different fixed orders, so their iteration counts may differ only where a
residual lands on the threshold."""
import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _pair(dims, shift, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) + 0.5 for n in dims[::-1]], indexing="ij")
    k = rng.uniform(0.2, 0.5, (3, 3))

    def f(xx, yy, zz):
        return (np.sin(k[0, 0] * xx + k[0, 1] * yy) * np.cos(k[1, 1] * yy) + 0.6 * np.sin(k[2, 2] * zz + k[2, 0] * xx)
                + 0.3 * np.cos(k[1, 2] * zz)).astype(np.float32)

    return f(x, y, z).ravel(), f(x - shift[0], y - shift[1], z - shift[2]).ravel()


@pytest.mark.parametrize("case", [
    # dims, ext, roi, max_iter, warps
    ((20, 18, 16), None, (1,) * 6, 120, 2),
    ((17, 22, 13), (26.0, 33.0, 20.0), (2, 1, 1, 3, 1, 2), 60, 2),
    ((24, 20, 18), None, (1,) * 6, 21, 1),
    # 31 iterations per warp: odd counts and warp starts outside the 50-iteration graphs
    ((14, 12, 10), None, (1, 2, 1, 1, 2, 1), 30, 3),
])
def test_of_level_matches_oracle(case):
    dims, ext, roi, max_iter, warps = case
    ext = np.asarray(dims if ext is None else ext, np.float32)
    print(f"[of-level] dims={dims} ext={ext.tolist()} roi={roi} max_iter={max_iter} warps={warps}", flush=True)
    n = int(np.prod(dims))
    I1, I2 = _pair(dims, (0.7, -0.5, 0.4), seed=sum(dims))
    rng = np.random.default_rng(3)
    V0 = (0.2 * rng.standard_normal(3 * n)).astype(np.float32)
    Y0 = (0.05 * rng.standard_normal(10 * n)).astype(np.float32)
    p = stt.of_params_reference()
    p.tol, p.lambda_, p.max_iter, p.warps = 1e-9, 0.3, max_iter, warps
    V, Y, its, res = stt.of_level(dims, ext, I1, I2, V0, Y0, p, roi)
    oV, oY, oits, ores = O.of_level(dims, ext, I1, I2, V0, Y0, tol=p.tol, lam=p.lambda_, max_iter=max_iter,
                                    warps=warps, roi=roi)
    print(f"  iterations gpu {its.tolist()} oracle {oits.tolist()}; residual gpu {res.tolist()} oracle {ores.tolist()}")
    print(f"  relL2 V {rel_l2(V, oV):.2e}  Y {rel_l2(Y, oY):.2e}", flush=True)
    # fp32 volumes within 1e-4; the residual sums absolute values of small
    # differences, which amplifies rounding (the shared cubic sampler uses FMA)
    assert np.array_equal(its, oits)
    assert rel_l2(V, oV) <= 1e-4
    assert rel_l2(Y, oY) <= 1e-4
    np.testing.assert_allclose(res, ores, rtol=1e-4)


def test_of_level_stops_at_the_residual_threshold():
    dims = (16, 16, 16)
    n = int(np.prod(dims))
    I1, I2 = _pair(dims, (0.5, 0.2, -0.3), seed=9)
    p = stt.of_params_reference()
    p.max_iter, p.warps, p.tol = 3000, 1, 1e-3
    V, Y, its, res = stt.of_level(dims, dims, I1, I2, np.zeros(3 * n, np.float32), np.zeros(10 * n, np.float32), p)
    print(f"[of-stop] iterations {its.tolist()} residual {res.tolist()}", flush=True)
    assert its[0] % 50 == 0 and its[0] <= 3001 and res[0] <= 1e-3


@pytest.mark.parametrize("spectrum", [1, 3])
def test_blur_matches_oracle(spectrum):
    dims = (21, 17, 19)
    n = int(np.prod(dims))
    img = np.random.default_rng(40 + spectrum).standard_normal(n * spectrum).astype(np.float32)
    print(f"[blur] dims={dims} spectrum={spectrum}", flush=True)
    assert np.array_equal(stt.blur(dims, img, 0.5, spectrum), O.blur(dims, img, 0.5, spectrum))


@pytest.mark.parametrize("new", [(13, 11, 12), (34, 27, 30), (21, 9, 25), (166, 17, 19)])
def test_resize_matches_oracle(new):
    dims = (21, 17, 19)
    n = int(np.prod(dims))
    img = np.random.default_rng(50).standard_normal(3 * n).astype(np.float32)
    print(f"[resize] {dims} -> {new}", flush=True)
    for spectrum in (1, 3):
        a = stt.resize(dims, img[:n * spectrum], new, spectrum)
        b = O.resize(dims, img[:n * spectrum], new, spectrum)
        assert np.array_equal(a, b), (new, spectrum, float(np.abs(a - b).max()))


def _ms_params(max_iter, warps, cut):
    p = stt.of_params_reference()
    p.max_iter, p.warps, p.tol, p.lambda_ = max_iter, warps, 1e-9, 0.3
    p.cut[:] = cut
    return p


def test_of_multiscale_matches_oracle():
    dims = (26, 22, 20)
    n = int(np.prod(dims))
    I1, I2 = _pair(dims, (0.9, -0.6, 0.5), seed=77)
    V0 = np.zeros(3 * n, np.float32)
    cut = (2, 2, 2, 5, 2, 2)
    p = _ms_params(60, 2, cut)
    print(f"[of-multiscale] dims={dims} levels={p.levels} eta={p.eta} max_iter=60 warps=2 cut={cut}", flush=True)
    V, its = stt.of_multiscale(dims, I1, I2, V0, p)
    oV, ld, oits = O.of_multiscale(dims, I1, I2, V0, tol=p.tol, lam=p.lambda_, max_iter=60, warps=2, cut=cut)
    print(f"  levels {ld.tolist()} iterations gpu {its.tolist()} oracle {oits.tolist()} relL2 V {rel_l2(V, oV):.2e}",
          flush=True)
    assert np.array_equal(its, oits)
    assert rel_l2(V, oV) <= 1e-4


def test_context_optical_flow_equals_host_form():
    dims = (24, 20, 18)
    n = int(np.prod(dims))
    I1, I2 = _pair(dims, (0.5, 0.3, -0.4), seed=78)
    g = stt.Geometry.make(dims, (16, 16))
    ctx = stt.Context(g, stt.params_default(), [1, 1, 1])
    ctx.set_volume(1, I1)
    ctx.set_volume(2, I2)
    V0 = (0.1 * np.random.default_rng(79).standard_normal(3 * n)).astype(np.float32)
    ctx.set_flow(1, V0)
    p = _ms_params(40, 1, (1,) * 6)
    print(f"[ctx-flow] dims={dims} pair=1", flush=True)
    its = ctx.optical_flow(1, p)
    got = ctx.get_buffer(stt.STT_BUF_FLOW, 1)
    ref, rits = stt.of_multiscale(dims, I1, I2, V0, p)
    ctx.close()
    assert np.array_equal(its, rits)
    assert np.array_equal(got, ref)
