"""Python host interface of the MI355X PatchMatch stereo engine.

``PMHuberStereo`` mirrors cvt's class (cvt/vision/PMHuberStereo.h:42-110):
``disparityMap(left, right, patchsize, depthmax, iterations, pp,
normalmap, initdisparity, initnormalmap)`` with the reference's defaults,
plus the public setup / iterate / planes interface.  ``Context`` is the
thin object over one C-ABI context (include/pm_stereo.h); every call goes
to the HIP engine, there is no CPU path.

Images are numpy float32 arrays, ``(H, W)`` for GRAY or ``(H, W, 4)`` for
RGBA; plane fields are ``(H, W, 4)`` = (a, b, c, cost).
"""
import ctypes

import numpy as np

from . import _native
from ._native import (PM_COST_CENSUS, PM_COST_COLOR_GRAD,  # noqa
                      PM_PP_LRCHECK, PM_PP_LRCHECK_FILL, PM_PP_NONE,
                      PM_SCHEDULE_JACOBI, PM_SCHEDULE_RED_BLACK,
                      PM_SMOOTH_DEPTHMAP, PM_SMOOTH_HUBER, PM_SMOOTH_NONE,
                      PmParams)

DISPARITY_PP_NONE = PM_PP_NONE
DISPARITY_PP_LRCHECK = PM_PP_LRCHECK
DISPARITY_PP_LRCHECK_FILL = PM_PP_LRCHECK_FILL


class PMError(RuntimeError):
    """A failed engine call: ``code`` is the pm_status, the message the
    engine's pm_last_error text."""

    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _image(img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    if img.ndim == 2:
        C = 1
    elif img.ndim == 3 and img.shape[2] == 4:
        C = 4
    else:
        raise ValueError("image must be (H, W) GRAY or (H, W, 4) RGBA, got %s"
                         % (img.shape,))
    H, W = img.shape[:2]
    return img, W, H, C


class Context:
    """One engine context (pm_create ... pm_destroy)."""

    def __init__(self, radius=5, disparity_max=64.0, color_weight=12.0,
                 smoothing=PM_SMOOTH_NONE, theta_min=1e-4,
                 schedule_iterations=0, device=0, stream=None,
                 cost_kind=PM_COST_CENSUS, cost_weight=16.0,
                 schedule=PM_SCHEDULE_JACOBI):
        self._lib = _native.lib()
        p = PmParams()
        self._lib.pm_params_default(ctypes.byref(p))
        p.radius = int(radius)
        p.disparity_max = float(disparity_max)
        p.color_weight = float(color_weight)
        p.cost_kind = int(cost_kind)
        p.cost_weight = float(cost_weight)
        p.schedule = int(schedule)
        p.smoothing = int(smoothing)
        p.theta_min = float(theta_min)
        p.schedule_iterations = int(schedule_iterations)
        p.device = int(device)
        p.stream = stream
        h = ctypes.c_void_p()
        rc = self._lib.pm_create(ctypes.byref(p), ctypes.byref(h))
        self._h = h
        if rc != _native.PM_OK:
            msg = self._lib.pm_last_error(h).decode() if h else "pm_create"
            if h:
                self._lib.pm_destroy(h)
            self._h = None
            raise PMError(rc, msg)
        self.radius = p.radius
        self.disparity_max = p.disparity_max
        self.shape = None
        self.channels = None

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if self._h:
            self._lib.pm_destroy(self._h)
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

    def _check(self, rc):
        if rc != _native.PM_OK:
            raise PMError(rc, self._lib.pm_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    # -- inputs ------------------------------------------------------------
    def set_images(self, left, right):
        left, W, H, C = _image(left)
        right, W2, H2, C2 = _image(right)
        if (W, H, C) != (W2, H2, C2):
            raise ValueError("left/right images differ in size or format")
        self._check(self._lib.pm_set_images(self._h, _ptr(left), _ptr(right),
                                            W, H, C, W * C))
        self.shape, self.channels = (H, W), C

    def set_images_device(self, d_left, d_right, W, H, channels, stride=None):
        """Device pointers (ints) of two W x H images in HBM."""
        stride = W * channels if stride is None else stride
        self._check(self._lib.pm_set_images_device(
            self._h, ctypes.c_void_p(d_left), ctypes.c_void_p(d_right), W, H,
            channels, stride))
        self.shape, self.channels = (H, W), channels

    def setup(self):
        self._check(self._lib.pm_setup(self._h))

    def setup_prior(self, disparity, normals=None):
        H, W = self.shape
        d = np.ascontiguousarray(disparity, np.float32)
        assert d.shape == (H, W)
        n = None
        if normals is not None:
            n = np.ascontiguousarray(normals, np.float32)
            assert n.shape == (H, W, 4)
        self._check(self._lib.pm_setup_prior(self._h, _ptr(d),
                                             _ptr(n) if n is not None
                                             else None))

    def setup_proposal(self, net, views=2):
        """Setup seeded by a plane-proposal network (cvt_amd.net.ProposalNet)
        with a per-pixel fallback to the random init (pm_setup_proposal);
        views=1 seeds the left view only, the right view taking the random
        init (pm_setup_proposal_views)."""
        self._check(self._lib.pm_setup_proposal_views(self._h, net.handle,
                                                      int(views)))

    # -- iterate -------------------------------------------------------------
    def iterate(self, n=1):
        self._check(self._lib.pm_iterate(self._h, int(n)))

    @property
    def iteration(self):
        return self._lib.pm_iteration(self._h)

    def planes(self, view=0):
        H, W = self.shape
        out = np.empty((H, W, 4), np.float32)
        self._check(self._lib.pm_get_planes(self._h, int(view), _ptr(out)))
        return out

    def set_planes(self, view, planes, iteration):
        H, W = self.shape
        p = np.ascontiguousarray(planes, np.float32)
        assert p.shape == (H, W, 4)
        self._check(self._lib.pm_set_planes(self._h, int(view), _ptr(p),
                                            int(iteration)))

    def cost(self, view, planes):
        H, W = self.shape
        p = np.ascontiguousarray(planes, np.float32)
        assert p.shape == (H, W, 4)
        out = np.empty((H, W), np.float32)
        self._check(self._lib.pm_cost(self._h, int(view), _ptr(p), _ptr(out)))
        return out

    # -- outputs -------------------------------------------------------------
    def disparity(self, pp=PM_PP_LRCHECK, maxdispdiff=1.0, maxdeg=5.0,
                  checked=False):
        H, W = self.shape
        d = np.empty((H, W), np.float32)
        chk = np.empty((H, W, 4), np.float32) if checked else None
        self._check(self._lib.pm_disparity(self._h, int(pp),
                                           float(maxdispdiff), float(maxdeg),
                                           _ptr(d),
                                           _ptr(chk) if checked else None))
        return (d, chk) if checked else d

    def disparity_device(self, d_out, pp=PM_PP_LRCHECK, maxdispdiff=1.0,
                         maxdeg=5.0):
        self._check(self._lib.pm_disparity_device(
            self._h, int(pp), float(maxdispdiff), float(maxdeg),
            ctypes.c_void_p(d_out)))

    def normals(self, pp=PM_PP_LRCHECK):
        H, W = self.shape
        out = np.empty((H, W, 4), np.float32)
        self._check(self._lib.pm_normals(self._h, int(pp), _ptr(out)))
        return out

    def sync(self):
        self._check(self._lib.pm_sync(self._h))

    # -- diagnostics ---------------------------------------------------------
    def texels(self, view=0):
        H, W = self.shape
        out = np.empty((H, W, 8), np.float32)
        self._check(self._lib.pm_get_texels(self._h, int(view), _ptr(out)))
        return out

    def draws(self, k, n):
        out = np.empty(int(n), np.float32)
        self._check(self._lib.pm_draws(self._h, int(k), int(n), _ptr(out)))
        return out

    def klt(self, view, planes, passes=1):
        H, W = self.shape
        p = np.ascontiguousarray(planes, np.float32)
        out = np.empty((H, W, 4), np.float32)
        self._check(self._lib.pm_klt(self._h, int(view), _ptr(p),
                                     int(passes), _ptr(out)))
        return out

    def iterate_counted(self, n=1):
        """iterate(n) that returns (cost_taps, klt_taps) evaluated."""
        a = ctypes.c_uint64()
        b = ctypes.c_uint64()
        self._check(self._lib.pm_iterate_counted(self._h, int(n),
                                                 ctypes.byref(a),
                                                 ctypes.byref(b)))
        return a.value, b.value

    def copy_probe(self, d_src, d_dst, nbytes):
        self._check(self._lib.pm_copy_probe(self._h, ctypes.c_void_p(d_src),
                                            ctypes.c_void_p(d_dst),
                                            int(nbytes)))

    def math_selftest(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty((x.size, 5), np.float32)
        self._check(self._lib.pm_math_selftest(self._h, _ptr(x), x.size,
                                               _ptr(out)))
        return out

    # -- PM-Huber coupling -------------------------------------------------
    def smooth(self, view=0):
        """The smoothed (normal, disparity / disparity_max) field a view's
        accept rule pulls towards (PM_SMOOTH_HUBER)."""
        H, W = self.shape
        out = np.empty((H, W, 4), np.float32)
        self._check(self._lib.pm_get_smooth(self._h, int(view), _ptr(out)))
        return out

    def tensor(self, view=0):
        """A view's diffusion tensor (PM_SMOOTH_HUBER)."""
        H, W = self.shape
        out = np.empty((H, W, 4), np.float32)
        self._check(self._lib.pm_get_tensor(self._h, int(view), _ptr(out)))
        return out

    def rof(self, view, f, lam, iters=40):
        """`iters` diffusion-tensor ROF steps on f with a view's tensor."""
        H, W = self.shape
        f = np.ascontiguousarray(f, np.float32)
        lam = np.ascontiguousarray(lam, np.float32)
        assert f.shape == (H, W, 4) and lam.shape == (4,)
        out = np.empty((H, W, 4), np.float32)
        self._check(self._lib.pm_rof(self._h, int(view), _ptr(f), _ptr(lam),
                                     int(iters), _ptr(out)))
        return out

    def rof_device(self, view, d_f, lam, iters, d_out):
        """rof() on device pointers (ints) d_f and d_out, W x H x 4 each;
        enqueued on the context's stream, no synchronisation."""
        lam = np.ascontiguousarray(lam, np.float32)
        assert lam.shape == (4,)
        self._check(self._lib.pm_rof_device(
            self._h, int(view), ctypes.c_void_p(d_f), _ptr(lam), int(iters),
            ctypes.c_void_p(d_out)))


class PMHuberStereo:
    """cvt::PMHuberStereo (cvt/vision/PMHuberStereo.h:42-110).  By default
    its disparityMap pipeline: PatchMatch coupled to the diffusion-tensor
    ROF smoothing (PM_SMOOTH_HUBER); smoothing=PM_SMOOTH_NONE gives plain
    PatchMatch."""

    def __init__(self, device=0, smoothing=PM_SMOOTH_HUBER):
        self.device = device
        self.smoothing = smoothing
        self._ctx = None
        self._cfg = None

    def _context(self, patchsize, depthmax, smoothing=None):
        smoothing = self.smoothing if smoothing is None else smoothing
        cfg = (int(patchsize), float(depthmax), int(smoothing))
        if self._ctx is None or self._cfg != cfg:
            if self._ctx is not None:
                self._ctx.close()
            self._ctx = Context(radius=cfg[0], disparity_max=cfg[1],
                                smoothing=cfg[2], device=self.device)
            self._cfg = cfg
        return self._ctx

    def setup(self, left, right, patchsize, depthmax, initdisparity=None,
              initnormalmap=None, smoothing=None):
        ctx = self._context(patchsize, depthmax, smoothing)
        ctx.set_images(left, right)
        if initdisparity is not None:
            ctx.setup_prior(initdisparity, initnormalmap)
        else:
            ctx.setup()
        return ctx

    def iterate(self, n):
        self._ctx.iterate(n)

    def planes(self, view=0):
        return self._ctx.planes(view)

    def disparityMap(self, left, right, patchsize, depthmax, iterations,
                     pp=DISPARITY_PP_LRCHECK, normalmap=False,
                     initdisparity=None, initnormalmap=None):
        """Returns the left disparity map (and the normal map when
        ``normalmap`` is true).  Iterations are made odd like
        PMHuberStereo.cpp:490; LR check at 1 px / 5 degrees
        (PMHuberStereo.cpp:79-80)."""
        ctx = self.setup(left, right, patchsize, depthmax, initdisparity,
                         initnormalmap)
        ctx.iterate(int(iterations) | 1)
        d = ctx.disparity(pp, 1.0, 5.0)
        if normalmap:
            return d, ctx.normals(pp)
        return d

    def depthMap(self, left, right, patchsize, depthmax, iterations,
                 dscale=-1.0, pp=DISPARITY_PP_LRCHECK, normalmap=False,
                 initdisparity=None, initnormalmap=None):
        """PMHuberStereo::depthMap (PMHuberStereo.cpp:147-347): its own
        coupling (PM_SMOOTH_DEPTHMAP: Huber smoothing of filled fields,
        theta schedule from 1e-3), exactly ``iterations`` iterations, the
        result scaled by ``dscale`` (1 / depthmax when <= 0)."""
        if dscale <= 0.0:
            dscale = 1.0 / float(depthmax)
        ctx = self.setup(left, right, patchsize, depthmax, initdisparity,
                         initnormalmap, smoothing=PM_SMOOTH_DEPTHMAP)
        ctx.iterate(int(iterations))
        d = ctx.disparity(pp, 1.0, 5.0) * np.float32(dscale)
        if normalmap:
            return d, ctx.normals(pp)
        return d

    def close(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None
