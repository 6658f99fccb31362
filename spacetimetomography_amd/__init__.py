"""MI355X space-time tomography hot path (HIP/gfx950) behind a C-ABI.

The native library is spacetimetomography_amd/lib/libstt.so; this package
only binds it (ctypes).  See include/stt.h for the entry points.
"""
from ._lib import (  # noqa: F401
    Context, Group, Geometry, Params, Timing, SttError, check, lib, LIB_PATH,
    params_default, params_reference, synth_phantom, synth_phantom_frame, synth_flow, synth_angles,
    view_rounds, lds_sequence, OfParams, of_params_reference, of_level, of_multiscale, blur, resize,
    ImageInfo, image_query, load_image, save_tiff, save_analyze, read_start_angles, split_sizes,
    STT_OK, STT_ERR_INVALID, STT_ERR_HIP, STT_ERR_NOMEM, STT_ERR_STATE, STT_ERR_NODEVICE,
    STT_BP_VOXEL, STT_BP_RAY, STT_PRIOR_ATV, STT_PRIOR_TV, STT_PRIOR_STV,
    STT_INTERP_CUBIC_CLAMPED, STT_INTERP_TRILINEAR,
    STT_BUF_XBAR, STT_BUF_XPREV, STT_BUF_YTV, STT_BUF_YTP, STT_BUF_YBC, STT_BUF_FLOW,
    PROJECTOR_SHAPES, PROJECTOR_BOXES,
)
