"""ctypes binding of the C-ABI declared in include/stt.h.

The shared library lives in-tree at spacetimetomography_amd/lib/libstt.so and
is built by spacetimetomography_amd/build.py.  There is no fallback: if the
library is missing, importing this module raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libstt.so")

STT_OK = 0
STT_ERR_INVALID = 1
STT_ERR_HIP = 2
STT_ERR_NOMEM = 3
STT_ERR_STATE = 4
STT_ERR_NODEVICE = 5

STT_BP_VOXEL, STT_BP_RAY = 0, 1
STT_PRIOR_ATV, STT_PRIOR_TV, STT_PRIOR_STV = 0, 1, 2
STT_INTERP_CUBIC_CLAMPED, STT_INTERP_TRILINEAR = 0, 1
STT_BUF_XBAR, STT_BUF_XPREV, STT_BUF_YTV, STT_BUF_YTP, STT_BUF_YBC, STT_BUF_FLOW = range(6)
PROJECTOR_SHAPES = (0, -32, -8, 64, 32, 16, 8)  # stt_autotune order (STT_PROJECTOR_SHAPES)
PROJECTOR_SHAPES_N = len(PROJECTOR_SHAPES)
PROJECTOR_BOXES = (16, 20, 24, 28)  # stt_autotune_box order (STT_PROJECTOR_BOXES), KiB


class Geometry(C.Structure):
    """stt_geometry: vol, spacing, det, ds, sid, sdd, oxyz (main.cpp flag set)."""
    _fields_ = [("vol", C.c_int * 3), ("spacing", C.c_float * 3), ("det", C.c_int * 2),
                ("ds", C.c_double), ("sid", C.c_double), ("sdd", C.c_double),
                ("oxyz", C.c_double * 3)]

    @classmethod
    def make(cls, vol, det, spacing=(1.0, 1.0, 1.0), ds=1.0, sid=600.0, sdd=1000.0,
             oxyz=(0.0, 0.0, 0.0)):
        g = cls()
        g.vol[:] = [int(v) for v in vol]
        g.spacing[:] = [float(s) for s in spacing]
        g.det[:] = [int(d) for d in det]
        g.ds, g.sid, g.sdd = float(ds), float(sid), float(sdd)
        g.oxyz[:] = [float(o) for o in oxyz]
        return g


class Params(C.Structure):
    """stt_params (the numerical part of the ST_Tomography constructor)."""
    _fields_ = [("sigma", C.c_double), ("tau", C.c_double), ("lambda_tv", C.c_double),
                ("k_tp", C.c_double), ("e_bc", C.c_double), ("huber", C.c_double),
                ("alpha", C.c_double), ("g_max", C.c_float), ("prior", C.c_int),
                ("bp", C.c_int), ("interp", C.c_int), ("roi_cut", C.c_int * 8),
                ("roi_jtop", C.c_int), ("fp64_reductions", C.c_int)]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double),
                ("fp_ms", C.c_double), ("fp_launches", C.c_int),
                ("bp_ms", C.c_double), ("bp_launches", C.c_int),
                ("dual_ms", C.c_double), ("dual_launches", C.c_int),
                ("primal_ms", C.c_double), ("primal_launches", C.c_int),
                ("aux_ms", C.c_double), ("aux_launches", C.c_int),
                ("trace_ms", C.c_double), ("trace_launches", C.c_int)]


class ImageInfo(C.Structure):
    """stt_image_info: CImg dims (x, y, z, c), on-disk sample layout, voxel size."""
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("depth", C.c_int), ("spectrum", C.c_int),
                ("bits", C.c_int), ("sample_format", C.c_int), ("voxel", C.c_float * 3)]


class OfParams(C.Structure):
    """stt_of_params (L1Huber_OF_MultiScale3D's arguments)."""
    _fields_ = [("tol", C.c_float), ("lambda_", C.c_float), ("huber", C.c_float), ("eta", C.c_float),
                ("sigma", C.c_float), ("max_iter", C.c_int), ("norm", C.c_int), ("warps", C.c_int),
                ("levels", C.c_int), ("cut", C.c_int * 6)]


if not os.path.exists(LIB_PATH):
    raise ImportError(f"libstt.so not built at {LIB_PATH}; run spacetimetomography_amd/build.py")

lib = C.CDLL(LIB_PATH)

_P = C.c_void_p
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)

_SIGS = {
    "stt_last_error": (C.c_char_p, []),
    "stt_version": (C.c_int, []),
    "stt_params_reference": (None, [C.POINTER(Params)]),
    "stt_params_default": (None, [C.POINTER(Params)]),
    "stt_ctx_create": (C.c_int, [C.POINTER(Geometry), C.POINTER(Params), C.c_int, _I, C.POINTER(_P)]),
    "stt_ctx_destroy": (C.c_int, [_P]),
    "stt_ctx_set_params": (C.c_int, [_P, C.POINTER(Params)]),
    "stt_ctx_get_gmax": (C.c_int, [_P, _F]),
    "stt_set_views": (C.c_int, [_P, C.c_int, _F, C.c_int]),
    "stt_set_view_quaternions": (C.c_int, [_P, C.c_int, _F, C.c_int]),
    "stt_set_view_matrices": (C.c_int, [_P, C.c_int, _F, _F, C.c_int]),
    "stt_scene_matrices": (C.c_int, [C.POINTER(Geometry), _F, _F, _F]),
    "stt_set_volume": (C.c_int, [_P, C.c_int, _F]),
    "stt_get_volume": (C.c_int, [_P, C.c_int, _F]),
    "stt_set_flow": (C.c_int, [_P, C.c_int, _F]),
    "stt_set_projections": (C.c_int, [_P, C.c_int, _F]),
    "stt_set_volume_dev": (C.c_int, [_P, C.c_int, _P]),
    "stt_get_volume_dev": (C.c_int, [_P, C.c_int, _P]),
    "stt_set_flow_dev": (C.c_int, [_P, C.c_int, _P]),
    "stt_set_projections_dev": (C.c_int, [_P, C.c_int, _P]),
    "stt_forward_project_dev": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P]),
    "stt_backproject_voxel_dev": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_double]),
    "stt_fp_vjp_dev": (C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    "stt_warp_dev": (C.c_int, [_P, _P, _P, C.c_float, _P, C.c_int]),
    "stt_warp_vjp_dev": (C.c_int, [_P, _P, _P, C.c_float, _P, _P, _P]),
    "stt_get_buffer": (C.c_int, [_P, C.c_int, C.c_int, _F]),
    "stt_set_buffer": (C.c_int, [_P, C.c_int, C.c_int, _F]),
    "stt_forward_project": (C.c_int, [_P, C.c_int, C.c_int, _F, _F, _F]),
    "stt_backproject_voxel": (C.c_int, [_P, C.c_int, C.c_int, _F, C.c_double]),
    "stt_backproject_ray": (C.c_int, [_P, C.c_int, C.c_int, _F, _F, _F]),
    "stt_fp_vjp": (C.c_int, [_P, C.c_int, C.c_int, _F, _F]),
    "stt_sart_sweep": (C.c_int, [_P, C.c_int, _I, C.c_int]),
    "stt_warp": (C.c_int, [_P, _F, _F, C.c_float, _F, C.c_int]),
    "stt_warp_frame": (C.c_int, [_P, C.c_int, C.c_float, C.c_int, _F]),
    "stt_warp_vjp": (C.c_int, [_P, _F, _F, C.c_float, _F, _F, _F]),
    "stt_pd_init": (C.c_int, [_P]),
    "stt_pd_dual": (C.c_int, [_P, C.c_int]),
    "stt_pd_primal": (C.c_int, [_P, C.c_int]),
    "stt_pd_extrapolate": (C.c_int, [_P, C.c_int]),
    "stt_set_trace_storage": (C.c_int, [_P, C.c_int]),
    "stt_iterate": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(Timing)]),
    "stt_synchronize": (C.c_int, [_P]),
    "stt_time_projector": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "stt_set_projector_tile": (C.c_int, [_P, C.c_int]),
    "stt_autotune": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "stt_set_projector_box": (C.c_int, [_P, C.c_int]),
    "stt_autotune_box": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "stt_residual_norm": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "stt_residual_norm_noalloc": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "stt_power_iteration": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    "stt_synth_phantom": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, _F]),
    "stt_synth_phantom_frame": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                          C.c_float, C.c_int, _F]),
    "stt_synth_flow": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_float, C.c_int, _F]),
    "stt_synth_angles": (C.c_int, [C.c_int, C.c_int, C.c_int, _F]),
    "stt_view_rounds": (C.c_int, [_F, C.c_int, C.c_int, _F]),
    "stt_lds_sequence": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _F]),
    "stt_of_params_reference": (None, [C.POINTER(OfParams)]),
    "stt_of_level": (C.c_int, [_I, _F, _F, _F, _F, _F, C.POINTER(OfParams), _I, _I, C.POINTER(C.c_double)]),
    "stt_of_multiscale": (C.c_int, [_I, _F, _F, _F, C.POINTER(OfParams), _I]),
    "stt_optical_flow": (C.c_int, [_P, C.c_int, C.POINTER(OfParams), _I]),
    "stt_blur": (C.c_int, [_I, C.c_int, _F, C.c_float, _F]),
    "stt_resize": (C.c_int, [_I, C.c_int, _F, _I, _F]),
    "stt_group_create": (C.c_int, [C.POINTER(_P), C.c_int, C.POINTER(_P)]),
    "stt_group_destroy": (C.c_int, [_P]),
    "stt_group_iterate": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "stt_image_query": (C.c_int, [C.c_char_p, C.POINTER(ImageInfo)]),
    "stt_load_image": (C.c_int, [C.c_char_p, _F, C.c_size_t]),
    "stt_save_tiff": (C.c_int, [C.c_char_p, _F, C.c_int, C.c_int, C.c_int, C.c_int]),
    "stt_save_analyze": (C.c_int, [C.c_char_p, _F, C.c_int, C.c_int, C.c_int, C.c_int, _F]),
    "stt_read_start_angles": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _F]),
    "stt_split_sizes": (C.c_int, [C.c_int, C.c_int, _I]),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


class SttError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"stt error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != STT_OK:
        raise SttError(rc, lib.stt_last_error().decode(errors="replace"))


def dptr(x):
    """A device address for the "_dev" entry points (None passes NULL): an
    int or a c_void_p holding memory allocated through the HIP runtime that
    libstt.so links (ROCm's libamdhip64.so.7).  A runtime bundled elsewhere in
    the process (PyTorch wheels carry their own) has its own address space
    bookkeeping, so its pointers do not qualify.  The buffer is borrowed; work
    that writes it must have finished, because the library uses its own
    stream."""
    if x is None:
        return None
    return x if isinstance(x, C.c_void_p) else C.c_void_p(int(x))


def fptr(a):
    """float* of a C-contiguous float32 numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "need a contiguous float32 array"
    return a.ctypes.data_as(_F)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)


def params_default():
    p = Params()
    lib.stt_params_default(C.byref(p))
    return p


def params_reference():
    p = Params()
    lib.stt_params_reference(C.byref(p))
    return p


def synth_phantom(dims, n_ellipsoids=10, seed=1):
    nx, ny, nz = dims
    out = np.empty(nx * ny * nz, np.float32)
    check(lib.stt_synth_phantom(nx, ny, nz, n_ellipsoids, seed, fptr(out)))
    return out


def synth_phantom_frame(dims, t, n_ellipsoids=10, seed=1, deform_seed=2, amp=4.0):
    nx, ny, nz = dims
    out = np.empty(nx * ny * nz, np.float32)
    check(lib.stt_synth_phantom_frame(nx, ny, nz, n_ellipsoids, seed, deform_seed, amp, t, fptr(out)))
    return out


def synth_flow(dims, t, deform_seed=2, amp=4.0):
    nx, ny, nz = dims
    out = np.empty(3 * nx * ny * nz, np.float32)
    check(lib.stt_synth_flow(nx, ny, nz, deform_seed, amp, t, fptr(out)))
    return out


def synth_angles(kind, n, offset=0):
    out = np.empty(n, np.float32)
    check(lib.stt_synth_angles(kind, n, offset, fptr(out)))
    return out


def view_rounds(start_deg, views_per_round=10):
    """Quaternions (n, 4) of the reference's rounds of views: each round from
    its start angle, then views_per_round-1 rotations by 360/views_per_round."""
    st = np.ascontiguousarray(start_deg, np.float32).reshape(-1)
    out = np.zeros((len(st) * views_per_round, 4), np.float32)
    check(lib.stt_view_rounds(fptr(st), len(st), views_per_round, fptr(out)))
    return out


def lds_sequence(np_, ns=1, p=2, nz=10**9):
    """Low_discrepency_squence(Np, Ns, p, Nz) angles in degrees."""
    out = np.zeros(np_, np.float32)
    check(lib.stt_lds_sequence(np_, ns, p, nz, fptr(out)))
    return out


def of_params_reference():
    p = OfParams()
    lib.stt_of_params_reference(C.byref(p))
    return p


def of_level(dims, ext, I1, I2, V, Y, params, roi=(1,) * 6):
    """One level of L1Huber_OF on the device; returns (V, Y, iterations, residuals) per warp."""
    dims = np.asarray(dims, np.int32)
    ext = np.asarray(ext, np.float32)
    roi = np.asarray(roi, np.int32)
    n = int(np.prod(dims))
    I1 = np.ascontiguousarray(I1, np.float32)
    I2 = np.ascontiguousarray(I2, np.float32)
    V = np.ascontiguousarray(V, np.float32).copy()
    Y = np.ascontiguousarray(Y, np.float32).copy()
    if I1.size != n or I2.size != n or V.size != 3 * n or Y.size != 10 * n:
        raise ValueError("I1, I2 need n voxels, V 3n, Y 10n")
    its = np.zeros(max(params.warps, 1), np.int32)
    res = np.zeros(max(params.warps, 1), np.float64)
    check(lib.stt_of_level(dims.ctypes.data_as(_I), fptr(ext), fptr(I1), fptr(I2), fptr(V), fptr(Y), C.byref(params),
                           roi.ctypes.data_as(_I), its.ctypes.data_as(_I),
                           res.ctypes.data_as(C.POINTER(C.c_double))))
    return V, Y, its[:params.warps], res[:params.warps]


def of_multiscale(dims, I1, I2, V, params):
    """L1Huber_OF_MultiScale3D on the device; returns (V, iterations[levels, warps], coarsest level first)."""
    dims = np.asarray(dims, np.int32)
    n = int(np.prod(dims))
    I1 = np.ascontiguousarray(I1, np.float32)
    I2 = np.ascontiguousarray(I2, np.float32)
    V = np.ascontiguousarray(V, np.float32).copy()
    if I1.size != n or I2.size != n or V.size != 3 * n:
        raise ValueError("I1, I2 need n voxels, V 3n")
    its = np.zeros(max(params.levels * params.warps, 1), np.int32)
    check(lib.stt_of_multiscale(dims.ctypes.data_as(_I), fptr(I1), fptr(I2), fptr(V), C.byref(params),
                                its.ctypes.data_as(_I)))
    return V, its[:params.levels * params.warps].reshape(params.levels, params.warps)


def blur(dims, img, sigma, spectrum=1):
    dims = np.asarray(dims, np.int32)
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros_like(img)
    check(lib.stt_blur(dims.ctypes.data_as(_I), int(spectrum), fptr(img), float(sigma), fptr(out)))
    return out


def resize(dims, img, new_dims, spectrum=1):
    dims = np.asarray(dims, np.int32)
    nd = np.asarray(new_dims, np.int32)
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros(int(np.prod(nd)) * spectrum, np.float32)
    check(lib.stt_resize(dims.ctypes.data_as(_I), int(spectrum), fptr(img), nd.ctypes.data_as(_I), fptr(out)))
    return out


def _path(path):
    return os.fsencode(path)


def image_query(path):
    """Dimensions and sample layout of the image CImg::load(path) would return."""
    info = ImageInfo()
    check(lib.stt_image_query(_path(path), C.byref(info)))
    return info


def load_image(path):
    """CImg::load of a .tif/.tiff/.hdr/.img/.nii file as float32 (c, z, y, x)."""
    info = image_query(path)
    shape = (info.spectrum, info.depth, info.height, info.width)
    out = np.empty(shape, np.float32)
    check(lib.stt_load_image(_path(path), fptr(out), out.size))
    return out


def _cimg_dims(img):
    """(w, h, d, c) of an array shaped (x), (y, x), (z, y, x) or (c, z, y, x)."""
    img = np.ascontiguousarray(img, np.float32)
    shape = (1,) * (4 - img.ndim) + img.shape
    c, d, h, w = shape
    return img, w, h, d, c


def save_tiff(path, img):
    """CImg::save_tiff of a float image: one page per z slice."""
    img, w, h, d, c = _cimg_dims(img)
    check(lib.stt_save_tiff(_path(path), fptr(img), w, h, d, c))


def save_analyze(path, img, voxel=None):
    """CImg::save_analyze of a float image (.hdr/.img pair or single .nii)."""
    img, w, h, d, c = _cimg_dims(img)
    vox = None if voxel is None else np.asarray(voxel, np.float32)
    check(lib.stt_save_analyze(_path(path), fptr(img), w, h, d, c, fptr(vox)))


def read_start_angles(path, nframes, nrounds):
    """lds.txt start angles, frame-major: array (nframes, nrounds) in degrees."""
    out = np.empty((nframes, nrounds), np.float32)
    check(lib.stt_read_start_angles(_path(path), int(nframes), int(nrounds), fptr(out)))
    return out


def split_sizes(size, nparts):
    """Block sizes of CImg::get_split(axis, nparts) over an axis of `size`."""
    out = np.empty(int(nparts), np.int32)
    check(lib.stt_split_sizes(int(size), int(nparts), iptr(out)))
    return out


class Context:
    """Thin owner of an stt_ctx (one HIP stream, all solver state in HBM)."""

    def __init__(self, geometry, params, views_per_frame):
        self.g = geometry
        self.p = params
        self.P = np.ascontiguousarray(views_per_frame, dtype=np.int32)
        self.F = len(self.P)
        self.V = int(np.prod(list(geometry.vol)))
        self.D = int(geometry.det[0] * geometry.det[1])
        h = _P()
        check(lib.stt_ctx_create(C.byref(geometry), C.byref(params), self.F, iptr(self.P), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.stt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_views(self, frame, angles_deg):
        a = np.ascontiguousarray(angles_deg, np.float32)
        check(lib.stt_set_views(self.h, frame, fptr(a), len(a)))

    def set_view_quaternions(self, frame, q):
        q = np.ascontiguousarray(q, np.float32).reshape(-1)
        check(lib.stt_set_view_quaternions(self.h, frame, fptr(q), len(q) // 4))

    def set_volume(self, frame, vol):
        check(lib.stt_set_volume(self.h, frame, fptr(np.ascontiguousarray(vol, np.float32).reshape(-1))))

    def get_volume(self, frame):
        out = np.empty(self.V, np.float32)
        check(lib.stt_get_volume(self.h, frame, fptr(out)))
        return out

    def set_flow(self, pair, uvw):
        check(lib.stt_set_flow(self.h, pair, fptr(np.ascontiguousarray(uvw, np.float32).reshape(-1))))

    def set_projections(self, frame, b):
        check(lib.stt_set_projections(self.h, frame, fptr(np.ascontiguousarray(b, np.float32).reshape(-1))))

    # Device-pointer forms (stt_*_dev): arguments are device addresses (dptr).
    def set_volume_dev(self, frame, vol_dev):
        check(lib.stt_set_volume_dev(self.h, frame, dptr(vol_dev)))

    def get_volume_dev(self, frame, out_dev):
        check(lib.stt_get_volume_dev(self.h, frame, dptr(out_dev)))

    def set_flow_dev(self, pair, uvw_dev):
        check(lib.stt_set_flow_dev(self.h, pair, dptr(uvw_dev)))

    def set_projections_dev(self, frame, b_dev):
        check(lib.stt_set_projections_dev(self.h, frame, dptr(b_dev)))

    def forward_project_dev(self, frame, view, ax_dev=None, raylen_dev=None, weight_dev=None):
        check(lib.stt_forward_project_dev(self.h, frame, view, dptr(ax_dev), dptr(raylen_dev), dptr(weight_dev)))

    def backproject_voxel_dev(self, frame, view, corr_dev, alpha):
        check(lib.stt_backproject_voxel_dev(self.h, frame, view, dptr(corr_dev), alpha))

    def fp_vjp_dev(self, frame, view, y_dev, out_dev):
        check(lib.stt_fp_vjp_dev(self.h, frame, view, dptr(y_dev), dptr(out_dev)))

    def warp_dev(self, uvw_dev, src_dev, dt, out_dev, interp=STT_INTERP_CUBIC_CLAMPED):
        check(lib.stt_warp_dev(self.h, dptr(uvw_dev), dptr(src_dev), dt, dptr(out_dev), interp))

    def warp_vjp_dev(self, uvw_dev, src_dev, dt, g_dev, gsrc_dev=None, guvw_dev=None):
        check(lib.stt_warp_vjp_dev(self.h, dptr(uvw_dev), dptr(src_dev), dt, dptr(g_dev), dptr(gsrc_dev),
                                   dptr(guvw_dev)))

    def get_buffer(self, which, frame):
        n = 4 * self.V if which == STT_BUF_YTV else (3 * self.V if which == STT_BUF_FLOW else self.V)
        out = np.empty(n, np.float32)
        check(lib.stt_get_buffer(self.h, which, frame, fptr(out)))
        return out

    def set_buffer(self, which, frame, data):
        n = 4 * self.V if which == STT_BUF_YTV else (3 * self.V if which == STT_BUF_FLOW else self.V)
        a = np.ascontiguousarray(data, np.float32).reshape(-1)
        if a.size != n:
            raise ValueError(f"buffer {which} holds {n} floats, got {a.size}")
        check(lib.stt_set_buffer(self.h, which, frame, fptr(a)))

    def set_params(self, params):
        """stt_ctx_set_params: new PD/SART/warp parameters (the backprojector stays)."""
        check(lib.stt_ctx_set_params(self.h, C.byref(params)))

    def gmax(self):
        g = C.c_float()
        check(lib.stt_ctx_get_gmax(self.h, C.byref(g)))
        return g.value

    def forward_project(self, frame, view, raylen=None):
        ax = np.empty(self.D, np.float32)
        rl = np.zeros(self.D, np.float32) if raylen is None else np.ascontiguousarray(raylen, np.float32).copy()
        w = np.empty(self.D, np.float32)
        check(lib.stt_forward_project(self.h, frame, view, fptr(ax), fptr(rl), fptr(w)))
        return ax, rl, w

    def backproject_voxel(self, frame, view, corr, alpha):
        check(lib.stt_backproject_voxel(self.h, frame, view, fptr(np.ascontiguousarray(corr, np.float32)), alpha))

    def backproject_ray(self, frame, view, corr):
        c = np.zeros(self.V, np.float32)
        w = np.zeros(self.V, np.float32)
        check(lib.stt_backproject_ray(self.h, frame, view, fptr(np.ascontiguousarray(corr, np.float32)),
                                      fptr(c), fptr(w)))
        return c, w

    def fp_vjp(self, frame, view, y, out=None):
        o = np.zeros(self.V, np.float32) if out is None else np.ascontiguousarray(out, np.float32).copy()
        check(lib.stt_fp_vjp(self.h, frame, view, fptr(np.ascontiguousarray(y, np.float32)), fptr(o)))
        return o

    def sart_sweep(self, frame, order=None):
        n = int(self.P[frame]) if order is None else len(order)
        o = None if order is None else np.ascontiguousarray(order, np.int32)
        check(lib.stt_sart_sweep(self.h, frame, iptr(o), n))

    def warp(self, uvw, src, dt, interp=STT_INTERP_CUBIC_CLAMPED):
        out = np.empty(self.V, np.float32)
        check(lib.stt_warp(self.h, fptr(np.ascontiguousarray(uvw, np.float32)),
                           fptr(np.ascontiguousarray(src, np.float32)), dt, fptr(out), interp))
        return out

    def warp_frame(self, pair, dt=-1.0, interp=STT_INTERP_CUBIC_CLAMPED):
        """stt_warp_frame: volume pair+1 warped by flow pair, on the device buffers."""
        out = np.empty(self.V, np.float32)
        check(lib.stt_warp_frame(self.h, pair, dt, interp, fptr(out)))
        return out

    def optical_flow(self, pair, params):
        """L1Huber_OF_MultiScale3D between volumes pair and pair + 1 into flow pair (in place)."""
        its = np.zeros(max(params.levels * params.warps, 1), np.int32)
        check(lib.stt_optical_flow(self.h, pair, C.byref(params), its.ctypes.data_as(_I)))
        return its[:params.levels * params.warps].reshape(params.levels, params.warps)

    def warp_vjp(self, uvw, src, dt, g, gsrc=None, guvw=None):
        """Adjoint of the trilinear warp: returns (gsrc, guvw) with the VJP
        added to the given accumulators (zeros when not given)."""
        gsrc = np.zeros(self.V, np.float32) if gsrc is None else np.ascontiguousarray(gsrc, np.float32)
        guvw = np.zeros(3 * self.V, np.float32) if guvw is None else np.ascontiguousarray(guvw, np.float32)
        check(lib.stt_warp_vjp(self.h, fptr(np.ascontiguousarray(uvw, np.float32)),
                               fptr(np.ascontiguousarray(src, np.float32)), dt,
                               fptr(np.ascontiguousarray(g, np.float32)), fptr(gsrc), fptr(guvw)))
        return gsrc, guvw

    def residual_norm(self):
        r2 = C.c_double()
        check(lib.stt_residual_norm(self.h, C.byref(r2)))
        return r2.value

    def power_iteration(self, iters):
        L = C.c_double()
        check(lib.stt_power_iteration(self.h, iters, C.byref(L)))
        return L.value

    def pd_init(self):
        check(lib.stt_pd_init(self.h))

    def pd_dual(self, frame):
        check(lib.stt_pd_dual(self.h, frame))

    def pd_primal(self, frame):
        check(lib.stt_pd_primal(self.h, frame))

    def pd_extrapolate(self, frame):
        check(lib.stt_pd_extrapolate(self.h, frame))

    def set_trace_storage(self, keep):
        """Keep (True) or free (False) the stored PD trace end points."""
        check(lib.stt_set_trace_storage(self.h, int(bool(keep))))

    def iterate(self, n=1, batched=True, use_graph=False, timing=False):
        t = Timing() if timing else None
        check(lib.stt_iterate(self.h, n, int(batched), int(use_graph), C.byref(t) if t is not None else None))
        return t

    def synchronize(self):
        check(lib.stt_synchronize(self.h))

    def time_projector(self, graph_form, reps=3):
        """ms of one iteration's projector launches (stt_time_projector)."""
        ms = C.c_double()
        check(lib.stt_time_projector(self.h, int(graph_form), int(reps), C.byref(ms)))
        return ms.value

    def set_projector_tile(self, rays_per_row):
        check(lib.stt_set_projector_tile(self.h, rays_per_row))

    def autotune(self):
        best = C.c_int()
        ms = (C.c_double * PROJECTOR_SHAPES_N)()
        check(lib.stt_autotune(self.h, C.byref(best), ms))
        return best.value, list(ms)

    def set_projector_box(self, kib):
        check(lib.stt_set_projector_box(self.h, kib))

    def autotune_box(self):
        best = C.c_int()
        ms = (C.c_double * len(PROJECTOR_BOXES))()
        check(lib.stt_autotune_box(self.h, C.byref(best), ms))
        return best.value, list(ms)


class Group:
    """stt_group: several contexts (e.g. one per volume size) iterated as one
    unit, one hipGraph for all of them with use_graph."""

    def __init__(self, contexts):
        self.contexts = list(contexts)  # keeps the contexts alive
        arr = (_P * len(self.contexts))(*[c.h for c in self.contexts])
        h = _P()
        check(lib.stt_group_create(arr, len(self.contexts), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.stt_group_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def iterate(self, n=1, use_graph=True):
        """n ST iterations of every context; returns their time in ms."""
        ms = C.c_double()
        check(lib.stt_group_iterate(self.h, int(n), int(use_graph), C.byref(ms)))
        return ms.value
