"""cvt::Disparity's error metrics (cvt/vision/Disparity.{h,cpp}) over the
engine library's C ABI (include/pm_io.h)."""
import ctypes

import numpy as np

from . import _native


class IOError_(RuntimeError):
    """A failed pm_io call (status code and message)."""

    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def _check(rc):
    if rc != _native.PM_OK:
        raise IOError_(rc, _native.lib().pm_io_last_error().decode())


def _gray(a, name):
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim != 2:
        raise ValueError("%s must be a 2-D GRAY_FLOAT map" % name)
    return a


def evaluate_error(disparity, gt, abs_thres=1.0, rel_thres=0.0):
    """Disparity::evaluateError (Disparity.cpp:201-233): (error, valid, all)
    over the pixels with gt > 0."""
    d, g = _gray(disparity, "disparity"), _gray(gt, "gt")
    if d.shape != g.shape:
        raise ValueError("disparity and gt differ in size")
    H, W = g.shape
    e, v, a = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    _check(_native.lib().pm_eval_error(
        d.ctypes.data, W, g.ctypes.data, W, W, H, float(abs_thres),
        float(rel_thres), ctypes.byref(e), ctypes.byref(v), ctypes.byref(a)))
    return e.value, v.value, a.value


def evaluate_error_device(d_disparity, d_gt, W, H, d_counts, abs_thres=1.0,
                          rel_thres=0.0, stride_d=None, stride_gt=None,
                          stream=None):
    """Disparity::evaluateError on the device (pm_eval_error_device):
    device pointers of W x H float maps; writes (error, valid, all) as
    three uint64 at d_counts in stream order."""
    _check(_native.lib().pm_eval_error_device(
        ctypes.c_void_p(d_disparity), stride_d or W, ctypes.c_void_p(d_gt),
        stride_gt or W, W, H, float(abs_thres), float(rel_thres),
        ctypes.c_void_p(d_counts), stream))


def evaluate_error_rate(disparity, gt, abs_thres=1.0, rel_thres=0.0):
    """Disparity::evaluateError's (error, density) form (Disparity.h:49-56):
    error / all and valid / all."""
    e, v, a = evaluate_error(disparity, gt, abs_thres, rel_thres)
    a = np.float32(a)
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float32(e) / a), float(np.float32(v) / a)


def bad_pixel_rate(disparity, gt, abs_thres=1.0):
    """bad-N px: the error rate at abs_thres with no relative threshold;
    LR-rejected pixels (disparity 0) count as errors."""
    return evaluate_error_rate(disparity, gt, abs_thres, 0.0)[0]


def default_histogram(relative=False):
    """Disparity::defaultHistogramAbsolute / defaultHistogramRelative bin
    edges (Disparity.cpp:320-355)."""
    n = ctypes.c_int()
    L = _native.lib()
    _check(L.pm_default_histogram(int(bool(relative)), None, ctypes.byref(n)))
    bins = np.zeros(n.value, np.float32)
    _check(L.pm_default_histogram(int(bool(relative)), bins.ctypes.data,
                                  ctypes.byref(n)))
    return bins


def histogram_error(gt, test, bins=None, relative=False):
    """Disparity::histogramErrorAbsolute / histogramErrorRelative
    (Disparity.cpp:236-318): fraction of the pixels (both maps finite and
    > 0) whose error lies in (bins[i-1], bins[i]]."""
    g, t = _gray(gt, "gt"), _gray(test, "test")
    if g.shape != t.shape:
        raise ValueError("gt and test differ in size")
    bins = default_histogram(relative) if bins is None else \
        np.ascontiguousarray(bins, np.float32)
    out = np.zeros(bins.size, np.float32)
    H, W = g.shape
    _check(_native.lib().pm_error_histogram(
        g.ctypes.data, W, t.ctypes.data, W, W, H, int(bool(relative)),
        bins.ctypes.data, bins.size, out.ctypes.data))
    return bins, out
