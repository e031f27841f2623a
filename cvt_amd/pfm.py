"""PFM disparity-map files (the format of cvt's PFM loader and saver
plugins, plugins/iloader/PFM/PFMLoader.cpp, plugins/isaver/PFM/
PFMSaver.cpp) over the engine library's C ABI (include/pm_io.h)."""
import ctypes
import os

import numpy as np

from . import _native
from .disparity import _check


def _path(p):
    return os.fsencode(p)


def load_pfm(path):
    """(H, W) float32 for a "Pf" file, (H, W, 4) RGBA for "PF"; top row
    first."""
    L = _native.lib()
    W, H, C = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(L.pm_pfm_info(_path(path), ctypes.byref(W), ctypes.byref(H),
                         ctypes.byref(C)))
    shape = (H.value, W.value) if C.value == 1 else (H.value, W.value, 4)
    out = np.empty(shape, np.float32)
    _check(L.pm_pfm_read(_path(path), out.ctypes.data, W.value, H.value,
                         C.value, W.value * C.value))
    return out


def save_pfm(path, img):
    """Write a (H, W) map as "Pf" or a (H, W, 4) image as "PF" (alpha
    dropped)."""
    img = np.ascontiguousarray(img, np.float32)
    if img.ndim == 2:
        H, W = img.shape
        C = 1
    elif img.ndim == 3 and img.shape[2] == 4:
        H, W, C = img.shape
    else:
        raise ValueError("expected (H, W) or (H, W, 4) float32")
    _check(_native.lib().pm_pfm_write(_path(path), img.ctypes.data, W, H, C,
                                      W * C))
