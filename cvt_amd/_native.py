"""ctypes binding of ``libcvt_amd_pm.so`` (the C ABI of include/pm_stereo.h).

The library is built in-tree by ``cvt_amd/build.py``.  There is no fallback:
a missing library raises, and a host without a HIP device gets PM_ERR_HIP
from ``pm_create``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(HERE, "libcvt_amd_pm.so")

PM_OK, PM_ERR_ARG, PM_ERR_STATE, PM_ERR_HIP, PM_ERR_NOMEM, PM_ERR_IO = range(6)
PM_PP_NONE, PM_PP_LRCHECK, PM_PP_LRCHECK_FILL = range(3)
PM_SMOOTH_NONE, PM_SMOOTH_HUBER, PM_SMOOTH_DEPTHMAP = range(3)
PM_COST_CENSUS, PM_COST_COLOR_GRAD = range(2)
PM_SCHEDULE_JACOBI, PM_SCHEDULE_RED_BLACK = range(2)

_p = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_u64 = ctypes.c_uint64


class PmParams(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_int),
                ("disparity_max", ctypes.c_float),
                ("color_weight", ctypes.c_float),
                ("cost_kind", ctypes.c_int),
                ("cost_weight", ctypes.c_float),
                ("schedule", ctypes.c_int),
                ("smoothing", ctypes.c_int),
                ("theta_min", ctypes.c_float),
                ("schedule_iterations", ctypes.c_int),
                ("device", ctypes.c_int),
                ("stream", ctypes.c_void_p)]


class PmBatchPair(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("channels", ctypes.c_int), ("stride", ctypes.c_int),
                ("disparity_max", ctypes.c_float),
                ("d_left", ctypes.c_void_p), ("d_right", ctypes.c_void_p),
                ("d_gt", ctypes.c_void_p), ("gt_stride", ctypes.c_int)]


class PmBatchTuning(ctypes.Structure):
    _fields_ = [("tile_rows", ctypes.c_int), ("workgroups", ctypes.c_int),
                ("ms", ctypes.c_float)]


# name -> (restype, argtypes); every symbol the headers under include/
# declare
SIGNATURES = {
    "pm_params_default": (None, [ctypes.POINTER(PmParams)]),
    "pm_create": (_i, [ctypes.POINTER(PmParams), ctypes.POINTER(_p)]),
    "pm_destroy": (None, [_p]),
    "pm_last_error": (ctypes.c_char_p, [_p]),
    "pm_set_images": (_i, [_p, _p, _p, _i, _i, _i, _i]),
    "pm_set_images_device": (_i, [_p, _p, _p, _i, _i, _i, _i]),
    "pm_setup": (_i, [_p]),
    "pm_setup_prior": (_i, [_p, _p, _p]),
    "pm_iterate": (_i, [_p, _i]),
    "pm_iteration": (_i, [_p]),
    "pm_get_planes": (_i, [_p, _i, _p]),
    "pm_set_planes": (_i, [_p, _i, _p, _i]),
    "pm_cost": (_i, [_p, _i, _p, _p]),
    "pm_disparity": (_i, [_p, _i, _f, _f, _p, _p]),
    "pm_disparity_device": (_i, [_p, _i, _f, _f, _p]),
    "pm_normals": (_i, [_p, _i, _p]),
    "pm_sync": (_i, [_p]),
    "pm_get_texels": (_i, [_p, _i, _p]),
    "pm_draws": (_i, [_p, _u64, _i, _p]),
    "pm_klt": (_i, [_p, _i, _p, _i, _p]),
    "pm_iterate_counted": (_i, [_p, _i, _p, _p]),
    "pm_copy_probe": (_i, [_p, _p, _p, ctypes.c_size_t]),
    "pm_math_selftest": (_i, [_p, _p, _i, _p]),
    "pm_get_smooth": (_i, [_p, _i, _p]),
    "pm_get_tensor": (_i, [_p, _i, _p]),
    "pm_rof": (_i, [_p, _i, _p, _p, _i, _p]),
    "pm_rof_device": (_i, [_p, _i, _p, _p, _i, _p]),
    # include/pm_net.h
    "pm_net_create": (_i, [_i, _p, _u64, ctypes.POINTER(_p)]),
    "pm_net_destroy": (None, [_p]),
    "pm_net_last_error": (ctypes.c_char_p, [_p]),
    "pm_net_num_params": (ctypes.c_size_t, [_p]),
    "pm_net_get_params": (_i, [_p, _p, ctypes.c_size_t]),
    "pm_net_set_params": (_i, [_p, _p, ctypes.c_size_t]),
    "pm_net_get_grads": (_i, [_p, _p, ctypes.c_size_t]),
    "pm_net_forward": (_i, [_p, _p, _i, _i, _i, _p]),
    "pm_net_train_step": (_i, [_p, _p, _p, _i, _i, _i, _f, _f, _f,
                               ctypes.POINTER(_f)]),
    "pm_setup_proposal": (_i, [_p, _p]),
    "pm_setup_proposal_views": (_i, [_p, _p, _i]),
    # include/pm_batch.h
    "pm_batch_create": (_i, [ctypes.POINTER(PmParams), _p, _i,
                             ctypes.POINTER(_p)]),
    "pm_batch_destroy": (None, [_p]),
    "pm_batch_last_error": (ctypes.c_char_p, [_p]),
    "pm_batch_size": (_i, [_p]),
    "pm_batch_set_pair": (_i, [_p, _i, _p]),
    "pm_batch_run": (_i, [_p, _i, _i, _f, _f, _f, _f]),
    "pm_batch_get_disparity": (_i, [_p, _i, _p]),
    "pm_batch_get_planes": (_i, [_p, _i, _i, _p]),
    "pm_batch_get_error_counts": (_i, [_p, _p]),
    "pm_batch_disparity_ptr": (_i, [_p, _i, ctypes.POINTER(_p)]),
    "pm_batch_autotune": (_i, [_p, _i, _p, _i, ctypes.POINTER(_i)]),
    "pm_batch_get_tuning": (_i, [_p, _p]),
    "pm_batch_set_tuning": (_i, [_p, _p]),
    "pm_batch_set_proposal": (_i, [_p, _p]),
    "pm_batch_set_proposal_views": (_i, [_p, _p, _i]),
    "pm_batch_sync": (_i, [_p]),
    # include/pm_io.h
    "pm_io_last_error": (ctypes.c_char_p, []),
    "pm_pfm_info": (_i, [ctypes.c_char_p, _p, _p, _p]),
    "pm_pfm_read": (_i, [ctypes.c_char_p, _p, _i, _i, _i, _i]),
    "pm_pfm_write": (_i, [ctypes.c_char_p, _p, _i, _i, _i, _i]),
    "pm_eval_error": (_i, [_p, _i, _p, _i, _i, _i, _f, _f, _p, _p, _p]),
    "pm_eval_error_device": (_i, [_p, _i, _p, _i, _i, _i, _f, _f, _p, _p]),
    "pm_error_histogram": (_i, [_p, _i, _p, _i, _i, _i, _i, _p, _i, _p]),
    "pm_default_histogram": (_i, [_i, _p, _p]),
    "pm_disparity_planes_noalloc": (_i, [_p, _p, _i, _i, _i, _f, _f, _p, _i,
                                         _p]),
}

_lib = None


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(SO_PATH):
            raise ImportError(
                "cvt_amd: native library %s is missing; build it with "
                "`python cvt_amd/build.py` (or __graft_entry__.build())"
                % SO_PATH)
        L = ctypes.CDLL(SO_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
