"""Python interface of batched PatchMatch (include/pm_batch.h).

``Batch`` runs plain PatchMatch on many stereo pairs of uneven sizes at
once: every kernel of the sequence is one launch over all pairs, the
cost-evaluating kernels walk the pairs' tile blocks in a persistent grid,
and the whole run is one HIP graph launch.  Each pair's result is bit for
bit what a single-pair ``Context`` with the same parameters computes.

Pairs are given as dicts with the device pointers of their images (torch
tensors' ``data_ptr()`` work): ``width``, ``height``, ``channels`` (1 or 4),
``stride`` (floats per row, default width x channels), ``disparity_max``,
``left``, ``right`` and optionally ``gt`` / ``gt_stride`` (ground-truth
disparity for the on-device error counts).
"""
import ctypes

import numpy as np

from . import _native
from ._native import PM_COST_CENSUS, PM_PP_LRCHECK, PmBatchPair, PmBatchTuning
from .stereo import PMError


def _pair(q):
    W, H, C = int(q["width"]), int(q["height"]), int(q.get("channels", 4))
    return PmBatchPair(width=W, height=H, channels=C,
                       stride=int(q.get("stride", W * C)),
                       disparity_max=float(q["disparity_max"]),
                       d_left=int(q["left"]), d_right=int(q["right"]),
                       d_gt=int(q["gt"]) if q.get("gt") else None,
                       gt_stride=int(q.get("gt_stride", W)))


class Batch:
    def __init__(self, pairs, radius=5, color_weight=12.0,
                 cost_kind=PM_COST_CENSUS, cost_weight=16.0, device=0,
                 stream=None):
        self._lib = _native.lib()
        p = _native.PmParams()
        self._lib.pm_params_default(ctypes.byref(p))
        p.radius = int(radius)
        p.color_weight = float(color_weight)
        p.cost_kind = int(cost_kind)
        p.cost_weight = float(cost_weight)
        p.device = int(device)
        p.stream = stream
        arr = (PmBatchPair * len(pairs))(*[_pair(q) for q in pairs])
        h = ctypes.c_void_p()
        rc = self._lib.pm_batch_create(ctypes.byref(p), arr, len(pairs),
                                       ctypes.byref(h))
        self._h = h
        if rc != _native.PM_OK:
            msg = (self._lib.pm_batch_last_error(h).decode() if h
                   else "pm_batch_create")
            if h:
                self._lib.pm_batch_destroy(h)
            self._h = None
            raise PMError(rc, msg)
        self.shapes = [(int(q["height"]), int(q["width"])) for q in pairs]

    def close(self):
        if self._h:
            self._lib.pm_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.shapes)

    def _check(self, rc):
        if rc != _native.PM_OK:
            raise PMError(rc, self._lib.pm_batch_last_error(self._h).decode())

    def set_pair(self, i, pair):
        self._check(self._lib.pm_batch_set_pair(self._h, int(i),
                                                ctypes.byref(_pair(pair))))

    def run(self, iterations, pp=PM_PP_LRCHECK, maxdispdiff=1.0, maxdeg=5.0,
            err_abs=-1.0, err_rel=0.0):
        """Stream-ordered: setup, iterations, post-processing (and error
        counts with err_abs >= 0) of every pair, one graph launch."""
        self._check(self._lib.pm_batch_run(self._h, int(iterations), int(pp),
                                           float(maxdispdiff), float(maxdeg),
                                           float(err_abs), float(err_rel)))

    def disparity(self, i):
        H, W = self.shapes[i]
        out = np.empty((H, W), np.float32)
        self._check(self._lib.pm_batch_get_disparity(
            self._h, int(i), out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def disparity_ptr(self, i):
        p = ctypes.c_void_p()
        self._check(self._lib.pm_batch_disparity_ptr(self._h, int(i),
                                                     ctypes.byref(p)))
        return p.value

    def planes(self, i, view):
        H, W = self.shapes[i]
        out = np.empty((H, W, 4), np.float32)
        self._check(self._lib.pm_batch_get_planes(
            self._h, int(i), int(view), out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def error_counts(self):
        """(n, 3) uint64: error, valid, all per pair (Disparity::
        evaluateError), from the last run with err_abs >= 0."""
        out = np.zeros((len(self), 3), np.uint64)
        self._check(self._lib.pm_batch_get_error_counts(
            self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def autotune(self, reps=2):
        """Times every candidate launch shape and keeps the fastest;
        returns the candidates as dicts (tile_rows, workgroups, ms)."""
        table = (PmBatchTuning * 32)()
        n = ctypes.c_int()
        self._check(self._lib.pm_batch_autotune(self._h, int(reps), table, 32,
                                                ctypes.byref(n)))
        return [{"tile_rows": t.tile_rows, "workgroups": t.workgroups,
                 "ms": t.ms} for t in table[:n.value]]

    @property
    def tuning(self):
        t = PmBatchTuning()
        self._check(self._lib.pm_batch_get_tuning(self._h, ctypes.byref(t)))
        return {"tile_rows": t.tile_rows, "workgroups": t.workgroups,
                "ms": t.ms}

    def set_tuning(self, tile_rows, workgroups=0):
        t = PmBatchTuning(tile_rows=int(tile_rows),
                          workgroups=int(workgroups), ms=0.0)
        self._check(self._lib.pm_batch_set_tuning(self._h, ctypes.byref(t)))

    def set_proposal(self, net, views=2):
        """Seed every pair from a ProposalNet (cvt_amd.net) in each run's
        setup (views=1: the left views only); None restores the random
        init.  The net must outlive its use by the batch."""
        self._net = net
        self._check(self._lib.pm_batch_set_proposal_views(
            self._h, net.handle if net is not None else None, int(views)))

    def sync(self):
        self._check(self._lib.pm_batch_sync(self._h))
