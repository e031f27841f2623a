"""cvt_amd -- MI355X-native PatchMatch stereo engine.

Drop-in for the setup/iterate path of cvt::PMHuberStereo: hand-written HIP
kernels for gfx950 behind the C ABI of include/pm_stereo.h, a C++ class
(include/cvt_amd/PMHuberStereo.h) and this Python interface; PFM disparity
maps and cvt's disparity error metrics in ``pfm`` and ``disparity``.
"""
from . import disparity, net, pfm  # noqa: F401
from .stereo import (DISPARITY_PP_LRCHECK, DISPARITY_PP_LRCHECK_FILL,  # noqa
                     DISPARITY_PP_NONE, PM_COST_CENSUS, PM_COST_COLOR_GRAD,
                     PM_SCHEDULE_JACOBI, PM_SCHEDULE_RED_BLACK,
                     PM_SMOOTH_DEPTHMAP, PM_SMOOTH_HUBER, PM_SMOOTH_NONE,
                     Context, PMError, PMHuberStereo)

__all__ = ["Context", "PMHuberStereo", "PMError", "DISPARITY_PP_NONE",
           "DISPARITY_PP_LRCHECK", "DISPARITY_PP_LRCHECK_FILL",
           "PM_SMOOTH_NONE", "PM_SMOOTH_HUBER", "PM_SMOOTH_DEPTHMAP",
           "PM_COST_CENSUS", "PM_COST_COLOR_GRAD", "PM_SCHEDULE_JACOBI",
           "PM_SCHEDULE_RED_BLACK"]
