"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, when built,
of the reference-compiled checker (oracle/_ref/libstt_ref.so).

Test infrastructure only: the oracle is the checker, never the thing tested.
"""
import ctypes as C
import os

import numpy as np

from spacetimetomography_amd._lib import Geometry, Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libstt_ref.so")

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_P = C.c_void_p
_G = C.POINTER(Geometry)

orc = C.CDLL(ORACLE_PATH)
_sigs = {
    "o_view_matrices": (C.c_int, [_G, _F, _F]),
    "o_quat_y": (None, [C.c_float, _F]),
    "o_fp_f32": (C.c_int, [_G, _F, _F, _F, _F, _F]),
    "o_fp_f64": (C.c_int, [_G, _F, _D, _D, _D, _D]),
    "o_bp_ray_scatter_f32": (C.c_int, [_G, _F, _F, _F, _F]),
    "o_bp_ray_scatter_f64": (C.c_int, [_G, _F, _D, _D, _D]),
    "o_bp_ray_apply_f32": (C.c_int, [C.c_int, _F, _F, C.c_float, _F]),
    "o_bp_voxel_f32": (C.c_int, [_G, _F, _F, C.c_double, C.c_float, _F]),
    "o_bp_voxel_f64": (C.c_int, [_G, _F, _D, C.c_double, C.c_double, _D]),
    "o_correction_f32": (C.c_int, [C.c_int, C.c_float, _F, _F, _F, _F]),
    "o_cubic_atxyz_f32": (C.c_float, [_F, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]),
    "o_warp_f32": (C.c_int, [_I, _F, _F, C.c_float, _F, C.c_int, C.c_int]),
    "o_warp_f64": (C.c_int, [_I, _D, _D, C.c_double, _D, C.c_int, C.c_int]),
    "o_roi_mask": (C.c_int, [_I, _I, C.c_int, _F]),
    "o_ctx_create": (_P, [_G, C.POINTER(Params), C.c_int, _I, C.c_int]),
    "o_ctx_destroy": (None, [_P]),
    "o_ctx_set_view_quaternions": (C.c_int, [_P, C.c_int, _F, C.c_int]),
    "o_ctx_set_volume": (C.c_int, [_P, C.c_int, _F]),
    "o_ctx_get_volume": (C.c_int, [_P, C.c_int, _F]),
    "o_ctx_get_volume_f64": (C.c_int, [_P, C.c_int, _D]),
    "o_ctx_set_flow": (C.c_int, [_P, C.c_int, _F]),
    "o_ctx_set_projections": (C.c_int, [_P, C.c_int, _F]),
    "o_ctx_get_buffer": (C.c_int, [_P, C.c_int, C.c_int, _F]),
    "o_ctx_set_buffer": (C.c_int, [_P, C.c_int, C.c_int, _F]),
    "o_ctx_gmax": (C.c_float, [_P]),
    "o_ctx_time_sample": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _D, _D]),
    "o_ctx_get_raylen": (C.c_int, [_P, _F]),
    "o_ctx_sart": (C.c_int, [_P, C.c_int, _I, C.c_int]),
    "o_ctx_pd_init": (C.c_int, [_P]),
    "o_ctx_pd_dual": (C.c_int, [_P, C.c_int]),
    "o_ctx_pd_primal": (C.c_int, [_P, C.c_int]),
    "o_ctx_pd_extrapolate": (C.c_int, [_P, C.c_int]),
    "o_ctx_iterate": (C.c_int, [_P, C.c_int, C.c_int]),
    "o_ctx_set_per_frame_raylen": (None, [_P, C.c_int]),
    "o_warp_ext_f32": (C.c_int, [_I, _F, _F, _F, C.c_float, _F]),
    "o_of_diff": (C.c_int, [_I, _F, C.c_int, _F]),
    "o_gradient_central": (C.c_int, [_I, _F, _F, _F, _F]),
    "o_of_level": (C.c_int, [_I, _F, _F, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_float,
                             _I, _I, _D]),
    "o_blur": (C.c_int, [_I, C.c_int, _F, C.c_float, _F]),
    "o_resize": (C.c_int, [_I, C.c_int, _F, _I, _F]),
    "o_of_multiscale": (C.c_int, [_I, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_float,
                                  C.c_float, C.c_int, C.c_float, _I, _I, _I]),
}
for _n, (_r, _a) in _sigs.items():
    _f = getattr(orc, _n)
    _f.restype = _r
    _f.argtypes = _a

# The reference-compiled checker is loaded on first use (O.ref), so test
# processes that never call it -- every GPU test -- do not map it.
HAVE_REF = os.path.exists(REF_PATH)
_REF_SIGS = {
    "r_view_matrices": (C.c_int, [_G, _F, _F]),
    "r_quat_y": (None, [C.c_float, _F]),
    "r_quat_rotate_y": (None, [_F, C.c_float, _F]),
    "r_matrix_inverse": (C.c_float, [_F, _F]),
    "r_fp": (C.c_int, [_G, _F, _F, _F, _F, _F, _F, _F]),
    "r_bp_ray_scatter": (C.c_int, [_G, _F, _F, _F, _F]),
    "r_cubic_atxyz": (C.c_float, [_F, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]),
    "r_of_diff": (C.c_int, [_I, _F, C.c_int, _F]),
    "r_gradient_central": (C.c_int, [_I, _F, _F, _F, _F]),
    "r_blur": (C.c_int, [_I, C.c_int, _F, C.c_float, _F]),
    "r_resize": (C.c_int, [_I, C.c_int, _F, C.c_int, C.c_int, C.c_int, C.c_int, _F]),
    "r_save_analyze": (C.c_int, [C.c_char_p, _F, C.c_int, C.c_int, C.c_int, C.c_int, _F, C.c_int]),
    "r_load_analyze": (C.c_int, [C.c_char_p, _I, _F, _F, C.c_size_t]),
    "r_split_sizes": (C.c_int, [C.c_int, C.c_int, _I]),
}
_ref_lib = None


def __getattr__(name):
    global _ref_lib
    if name != "ref":
        raise AttributeError(name)
    if _ref_lib is None and HAVE_REF:
        lib = C.CDLL(REF_PATH)
        for n, (r, a) in _REF_SIGS.items():
            f = getattr(lib, n)
            f.restype = r
            f.argtypes = a
        _ref_lib = lib
    return _ref_lib


def f32(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_F)


def f64(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def i32(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)


def quat_y(theta):
    q = np.zeros(4, np.float32)
    orc.o_quat_y(float(theta), f32(q))
    return q


def quats_y(angles):
    return np.stack([quat_y(a) for a in angles]).astype(np.float32)


def view_matrices(g, q):
    out = np.zeros(89, np.float32)
    orc.o_view_matrices(C.byref(g), f32(np.ascontiguousarray(q, np.float32)), f32(out))
    return out


def fp(g, q, vol, raylen=None, f64mode=False):
    D = g.det[0] * g.det[1]
    q = np.ascontiguousarray(q, np.float32)
    if f64mode:
        v = np.ascontiguousarray(vol, np.float64)
        ax = np.zeros(D, np.float64); rl = np.zeros(D, np.float64) if raylen is None else raylen.astype(np.float64)
        w = np.zeros(D, np.float64)
        orc.o_fp_f64(C.byref(g), f32(q), f64(v), f64(ax), f64(rl), f64(w))
        return ax, rl, w
    v = np.ascontiguousarray(vol, np.float32)
    ax = np.zeros(D, np.float32)
    rl = np.zeros(D, np.float32) if raylen is None else raylen.astype(np.float32).copy()
    w = np.zeros(D, np.float32)
    orc.o_fp_f32(C.byref(g), f32(q), f32(v), f32(ax), f32(rl), f32(w))
    return ax, rl, w


def bp_ray_scatter(g, q, corr):
    V = g.vol[0] * g.vol[1] * g.vol[2]
    c = np.zeros(V, np.float32); w = np.zeros(V, np.float32)
    orc.o_bp_ray_scatter_f32(C.byref(g), f32(np.ascontiguousarray(q, np.float32)),
                             f32(np.ascontiguousarray(corr, np.float32)), f32(c), f32(w))
    return c, w


def bp_voxel(g, q, corr, alpha, gmax, vol):
    v = np.ascontiguousarray(vol, np.float32).copy()
    orc.o_bp_voxel_f32(C.byref(g), f32(np.ascontiguousarray(q, np.float32)),
                       f32(np.ascontiguousarray(corr, np.float32)), alpha, gmax, f32(v))
    return v


def bp_voxel_f64(g, q, corr, alpha, gmax, vol):
    v = np.ascontiguousarray(vol, np.float64).copy()
    orc.o_bp_voxel_f64(C.byref(g), f32(np.ascontiguousarray(q, np.float32)),
                       f64(np.ascontiguousarray(corr, np.float64)), alpha, gmax, f64(v))
    return v


def warp(dims, uvw, src, dt, interp=0, f64mode=False):
    d = np.array(dims, np.int32)
    if f64mode:
        out = np.zeros(int(np.prod(dims)), np.float64)
        orc.o_warp_f64(i32(d), f64(np.ascontiguousarray(uvw, np.float64)), f64(np.ascontiguousarray(src, np.float64)),
                       dt, f64(out), interp, 0)
        return out
    out = np.zeros(int(np.prod(dims)), np.float32)
    orc.o_warp_f32(i32(d), f32(np.ascontiguousarray(uvw, np.float32)), f32(np.ascontiguousarray(src, np.float32)),
                   dt, f32(out), interp, 0)
    return out


def cubic_atxyz(img, dims, fx, fy, fz):
    return orc.o_cubic_atxyz_f32(f32(np.ascontiguousarray(img, np.float32)), dims[0], dims[1], dims[2], fx, fy, fz)


def roi_mask(dims, cut, jtop):
    out = np.zeros(int(np.prod(dims)), np.float32)
    orc.o_roi_mask(i32(np.array(dims, np.int32)), i32(np.array(cut, np.int32)), jtop, f32(out))
    return out


class OracleCtx:
    def __init__(self, g, p, P, f64mode=False):
        self.g, self.p = g, p
        self.P = np.ascontiguousarray(P, np.int32)
        self.F = len(self.P)
        self.V = g.vol[0] * g.vol[1] * g.vol[2]
        self.h = orc.o_ctx_create(C.byref(g), C.byref(p), self.F, i32(self.P), int(f64mode))
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            orc.o_ctx_destroy(self.h)
            self.h = None

    def set_views(self, f, angles):
        q = np.ascontiguousarray(quats_y(angles).reshape(-1))
        assert orc.o_ctx_set_view_quaternions(self.h, f, f32(q), len(angles)) == 0

    def set_volume(self, f, v):
        assert orc.o_ctx_set_volume(self.h, f, f32(np.ascontiguousarray(v, np.float32))) == 0

    def get_volume(self, f):
        out = np.zeros(self.V, np.float32)
        assert orc.o_ctx_get_volume(self.h, f, f32(out)) == 0
        return out

    def get_volume_f64(self, f):
        out = np.zeros(self.V, np.float64)
        assert orc.o_ctx_get_volume_f64(self.h, f, f64(out)) == 0
        return out

    def set_flow(self, pair, uvw):
        assert orc.o_ctx_set_flow(self.h, pair, f32(np.ascontiguousarray(uvw, np.float32))) == 0

    def set_projections(self, f, b):
        assert orc.o_ctx_set_projections(self.h, f, f32(np.ascontiguousarray(b, np.float32))) == 0

    def get_buffer(self, which, f):
        n = 4 * self.V if which == 2 else self.V
        out = np.zeros(n, np.float32)
        assert orc.o_ctx_get_buffer(self.h, which, f, f32(out)) == 0
        return out

    def set_buffer(self, which, f, a):
        assert orc.o_ctx_set_buffer(self.h, which, f, f32(np.ascontiguousarray(a, np.float32))) == 0

    def gmax(self):
        return orc.o_ctx_gmax(self.h)

    def sart(self, f, order=None):
        n = int(self.P[f]) if order is None else len(order)
        o = None if order is None else i32(np.ascontiguousarray(order, np.int32))
        assert orc.o_ctx_sart(self.h, f, o, n) == 0

    def pd_init(self):
        orc.o_ctx_pd_init(self.h)

    def pd_dual(self, f):
        orc.o_ctx_pd_dual(self.h, f)

    def pd_primal(self, f):
        orc.o_ctx_pd_primal(self.h, f)

    def pd_extrapolate(self, f):
        orc.o_ctx_pd_extrapolate(self.h, f)

    def iterate(self, n=1, faithful=False):
        assert orc.o_ctx_iterate(self.h, n, int(faithful)) == 0

    def per_frame_raylen(self, on=True):
        orc.o_ctx_set_per_frame_raylen(self.h, int(on))


def warp_ext(dims, ext, uvw, src, dt):
    """warping::warp with aabb = [0, ext] (the flow pyramid's coarse levels)."""
    out = np.zeros(int(np.prod(dims)), np.float32)
    orc.o_warp_ext_f32(i32(np.asarray(dims, np.int32)), f32(np.asarray(ext, np.float32)), f32(uvw), f32(src), float(dt),
                       f32(out))
    return out


def of_diff(dims, a, which):
    """tools.cpp differences: 0/1/2 = forward x/y/z, 3/4/5 = backward x/y/z."""
    out = np.zeros(int(np.prod(dims)), np.float32)
    orc.o_of_diff(i32(np.asarray(dims, np.int32)), f32(a), int(which), f32(out))
    return out


def gradient_central(dims, img):
    g = [np.zeros(int(np.prod(dims)), np.float32) for _ in range(3)]
    orc.o_gradient_central(i32(np.asarray(dims, np.int32)), f32(img), f32(g[0]), f32(g[1]), f32(g[2]))
    return g


def of_level(dims, ext, I1, I2, V, Y, tol=0.1, lam=1.2, max_iter=3000, norm=4, warps=3, huber=1.0, roi=(1,) * 6):
    """L1Huber_OF on one level; returns (V, Y, iterations per warp, residual per warp)."""
    V = np.ascontiguousarray(V, np.float32).copy()
    Y = np.ascontiguousarray(Y, np.float32).copy()
    its = np.zeros(warps, np.int32)
    errs = np.zeros(warps, np.float64)
    orc.o_of_level(i32(np.asarray(dims, np.int32)), f32(np.asarray(ext, np.float32)), f32(I1), f32(I2), f32(V), f32(Y),
                   float(tol), float(lam), int(max_iter), int(norm), int(warps), float(huber),
                   i32(np.asarray(roi, np.int32)), i32(its), f64(errs))
    return V, Y, its, errs


def blur(dims, img, sigma, spectrum=1):
    out = np.zeros(int(np.prod(dims)) * spectrum, np.float32)
    orc.o_blur(i32(np.asarray(dims, np.int32)), spectrum, f32(img), float(sigma), f32(out))
    return out


def resize(dims, img, new_dims, spectrum=1):
    out = np.zeros(int(np.prod(new_dims)) * spectrum, np.float32)
    orc.o_resize(i32(np.asarray(dims, np.int32)), spectrum, f32(img), i32(np.asarray(new_dims, np.int32)), f32(out))
    return out


def of_multiscale(dims, I1, I2, V, tol=0.1, lam=1.2, max_iter=3000, norm=4, warps=3, eta=0.65, sigma=0.5, levels=3,
                  huber=1.0, cut=(8, 8, 8, 45, 8, 8)):
    """L1Huber_OF_MultiScale3D; returns (V, level dims, iterations per level and warp)."""
    V = np.ascontiguousarray(V, np.float32).copy()
    ld = np.zeros(3 * levels, np.int32)
    its = np.zeros(levels * warps, np.int32)
    orc.o_of_multiscale(i32(np.asarray(dims, np.int32)), f32(I1), f32(I2), f32(V), float(tol), float(lam),
                        int(max_iter), int(norm), int(warps), float(eta), float(sigma), int(levels), float(huber),
                        i32(np.asarray(cut, np.int32)), i32(ld), i32(its))
    return V, ld.reshape(levels, 3), its.reshape(levels, warps)
