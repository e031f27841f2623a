"""ctypes binding of the CPU oracle (``libpm_oracle.so``).

TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module; the product (``cvt_amd``) never does.

Arrays follow the layouts documented in ``pm_oracle.h``: images are
``(H, W)`` or ``(H, W, 4)`` float32, texel records ``(H, W, 8)``, plane
fields ``(H, W, 4)`` and view buffers ``(H, W)`` uint64.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libpm_oracle.so")
_lib = None

_f = ctypes.c_float
_i = ctypes.c_int
_p = ctypes.c_void_p
_u64 = ctypes.c_uint64

_SIGS = {
    "pmo_expf": (_f, [_f]),
    "pmo_sincosf": (None, [_f, _p, _p]),
    "pmo_acospif": (_f, [_f]),
    "pmo_mwc_state": (None, [_u64, _p, _p]),
    "pmo_draws": (None, [_u64, _i, _p]),
    "pmo_draws_stepped": (None, [_u64, _i, _p]),
    "pmo_prop_draw_start": (_u64, [_u64, _i]),
    "pmo_viewprop": (None, [_p, _p]),
    "pmo_transform": (_f, [_p, _f, _f]),
    "pmo_to_normal": (None, [_p, _p]),
    "pmo_from_disparity_normal": (None, [_f, _f, _f, _p, _p]),
    "pmo_state_init": (None, [_f, _f, _f, _f, _f, _i, _f, _p]),
    "pmo_state_refine": (None, [_f, _f, _f, _p, _f, _f, _f, _i, _p]),
    "pmo_to_ref_normal_depth": (None, [_p, _f, _f, _i, _p]),
    "pmo_prepare": (None, [_p, _i, _i, _i, _i, _i, _f, _p]),
    "pmo_cost_at": (_f, [_p, _p, _i, _i, _i, _f, _i, _f, _i, _i, _p, _f]),
    "pmo_cost_field": (None, [_p, _p, _i, _i, _i, _f, _i, _f, _p, _p]),
    "pmo_klt_at": (None, [_p, _p, _i, _i, _i, _f, _i, _i, _p, _p]),
    "pmo_init": (None, [_p, _p, _i, _i, _i, _f, _i, _f, _f, _i, _p]),
    "pmo_init_prior": (None, [_p, _p, _i, _i, _i, _f, _i, _f, _f, _i, _p, _p,
                              _p]),
    "pmo_viewbuf_clear": (None, [_p, _i, _i]),
    "pmo_propagate": (None, [_p, _p, _i, _i, _i, _f, _i, _f, _f, _f, _i, _i,
                             _p, _p, _p, _p, _p, _p, _i, _p]),
    "pmo_consistency": (None, [_p, _p, _i, _i, _f, _f, _p]),
    "pmo_to_disparity": (None, [_p, _i, _i, _f, _i, _p]),
    "pmo_run": (_i, [_p, _p, _i, _i, _i, _i, _i, _f, _i, _i, _f, _i, _f, _f,
                     _p, _p, _p, _p]),
    "pmo_scatter": (None, [_p, _i, _i, _i, _p]),
    "pmo_fill_depthmap": (None, [_p, _i, _i, _f, _p]),
    "pmo_fill_normalmap": (None, [_p, _i, _i, _p]),
    "pmo_disparity_to_normal": (None, [_p, _i, _i, _p]),
    "pmo_diffusion_tensor": (None, [_p, _i, _i, _i, _i, _f, _p]),
    "pmo_normal_depth": (None, [_p, _i, _i, _f, _i, _p]),
    "pmo_rof_step": (None, [_p, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p]),
    "pmo_rof": (_i, [_p, _p, _p, _i, _i, _i, _p]),
    "pmo_run_huber": (_i, [_p, _p, _i, _i, _i, _i, _i, _f, _i, _f, _f, _f,
                           _p, _p, _p, _p, _p]),
    "pmo_huber_step": (None, [_p, _p, _p, _p, _p, _p, _f, _i, _i, _p, _p,
                              _p]),
    "pmo_huber": (_i, [_p, _p, _p, _f, _i, _i, _i, _p]),
    "pmo_fill_state": (None, [_p, _i, _i, _f, _i, _p]),
    "pmo_run_depthmap": (_i, [_p, _p, _i, _i, _i, _i, _i, _f, _i, _f, _f,
                              _p, _p, _p, _p, _p]),
    "pmo_num_threads": (_i, []),
}

COLOR_WEIGHT = 12.0  # CENSUS_COLOR_WEIGHT, pmhstereo.cl:47
# candidate costs (the reference's `#define patch_eval`, pmhstereo.cl:753)
COST_CENSUS = 0       # patch_eval_census, pmhstereo.cl:661-736
COST_COLOR_GRAD = 1   # patch_eval_color_grad_weighted, pmhstereo.cl:510-565
CG_WEIGHT = 16.0      # its support-weight factor, pmhstereo.cl:531


def build(always=True):
    """Compile the oracle with its Makefile (gcc, a few seconds); with
    always=False only when the library is missing or older than its
    sources."""
    cmd = ["make", "-s", "-C", _HERE] + (["-B"] if always else [])
    subprocess.run(cmd, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build(always=False)
        L = ctypes.CDLL(_SO)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return a


def expf(x):
    return lib().pmo_expf(float(x))


def sincosf(x):
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    lib().pmo_sincosf(float(x), _ptr(s), _ptr(c))
    return float(s[0]), float(c[0])


def acospif(x):
    return lib().pmo_acospif(float(x))


def mwc_state(k):
    x = np.zeros(1, np.uint32)
    c = np.zeros(1, np.uint32)
    lib().pmo_mwc_state(int(k), _ptr(x), _ptr(c))
    return int(x[0]), int(c[0])


def draws(k, n, stepped=False):
    out = np.zeros(n, np.float32)
    fn = lib().pmo_draws_stepped if stepped else lib().pmo_draws
    fn(int(k), int(n), _ptr(out))
    return out


def prop_draw_start(p, it):
    """stream index of the first draw of the propagate at pixel p,
    iteration it (RNG_init at pmhstereo.cl:942)"""
    return int(lib().pmo_prop_draw_start(int(p), int(it)))


def viewprop(s):
    s = _f32(s, (4,))
    out = np.zeros(4, np.float32)
    lib().pmo_viewprop(_ptr(s), _ptr(out))
    return out


def transform(s, x, y):
    s = _f32(s, (4,))
    return lib().pmo_transform(_ptr(s), float(x), float(y))


def to_normal(s):
    s = _f32(s, (4,))
    out = np.zeros(3, np.float32)
    lib().pmo_to_normal(_ptr(s), _ptr(out))
    return out


def from_disparity_normal(x, y, d, n):
    n = _f32(n, (3,))
    out = np.zeros(4, np.float32)
    lib().pmo_from_disparity_normal(float(x), float(y), float(d), _ptr(n),
                                    _ptr(out))
    return out


def state_init(u, x, y, lr, dmax):
    out = np.zeros(4, np.float32)
    lib().pmo_state_init(float(u[0]), float(u[1]), float(u[2]), float(x),
                         float(y), int(lr), float(dmax), _ptr(out))
    return out


def state_refine(u, s, x, y, dmax, lr):
    s = _f32(s, (4,))
    out = np.zeros(4, np.float32)
    lib().pmo_state_refine(float(u[0]), float(u[1]), float(u[2]), _ptr(s),
                           float(x), float(y), float(dmax), int(lr),
                           _ptr(out))
    return out


def to_ref_normal_depth(s, x, y, lr):
    s = _f32(s, (4,))
    out = np.zeros(4, np.float32)
    lib().pmo_to_ref_normal_depth(_ptr(s), float(x), float(y), int(lr),
                                  _ptr(out))
    return out


def _img_args(img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W = img.shape[:2]
    C = 1 if img.ndim == 2 else img.shape[2]
    assert C in (1, 4)
    return img, W, H, C, W * C


def prepare(img, radius, color_weight=COLOR_WEIGHT):
    img, W, H, C, stride = _img_args(img)
    tex = np.zeros((H, W, 8), np.float32)
    lib().pmo_prepare(_ptr(img), W, H, C, stride, int(radius),
                      float(color_weight), _ptr(tex))
    return tex


def cost_at(tex1, tex2, radius, x, y, s, mincost=1e10,
            color_weight=COLOR_WEIGHT, cost_kind=COST_CENSUS,
            cost_weight=CG_WEIGHT):
    H, W = tex1.shape[:2]
    s = _f32(s, (4,))
    return lib().pmo_cost_at(_ptr(_f32(tex1)), _ptr(_f32(tex2)), W, H,
                             int(radius), float(color_weight), int(cost_kind),
                             float(cost_weight), int(x), int(y), _ptr(s),
                             float(mincost))


def cost_field(tex1, tex2, radius, planes, color_weight=COLOR_WEIGHT,
               cost_kind=COST_CENSUS, cost_weight=CG_WEIGHT):
    H, W = tex1.shape[:2]
    planes = _f32(planes, (H, W, 4))
    out = np.zeros((H, W), np.float32)
    lib().pmo_cost_field(_ptr(_f32(tex1)), _ptr(_f32(tex2)), W, H,
                         int(radius), float(color_weight), int(cost_kind),
                         float(cost_weight), _ptr(planes), _ptr(out))
    return out


def klt_at(tex1, tex2, radius, x, y, s, color_weight=COLOR_WEIGHT):
    H, W = tex1.shape[:2]
    s = _f32(s, (4,))
    out = np.zeros(4, np.float32)
    lib().pmo_klt_at(_ptr(_f32(tex1)), _ptr(_f32(tex2)), W, H, int(radius),
                     float(color_weight), int(x), int(y), _ptr(s), _ptr(out))
    return out


def init(tex1, tex2, radius, dmax, lr, color_weight=COLOR_WEIGHT,
         cost_kind=COST_CENSUS, cost_weight=CG_WEIGHT):
    H, W = tex1.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().pmo_init(_ptr(_f32(tex1)), _ptr(_f32(tex2)), W, H, int(radius),
                   float(color_weight), int(cost_kind), float(cost_weight),
                   float(dmax), int(lr), _ptr(out))
    return out


def init_prior(tex1, tex2, radius, dmax, lr, disparity, normals,
               color_weight=COLOR_WEIGHT, cost_kind=COST_CENSUS,
               cost_weight=CG_WEIGHT):
    H, W = tex1.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().pmo_init_prior(_ptr(_f32(tex1)), _ptr(_f32(tex2)), W, H,
                         int(radius), float(color_weight), int(cost_kind),
                         float(cost_weight), float(dmax), int(lr),
                         _ptr(_f32(disparity, (H, W))),
                         _ptr(_f32(normals, (H, W, 4))), _ptr(out))
    return out


def viewbuf_empty(W, H):
    return np.full((H, W), np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64)


def propagate(tex1, tex2, radius, dmax, lr, it, old, viewin=None,
              src_other=None, smooth=None, theta=0.0, scatter=True,
              color_weight=COLOR_WEIGHT, cost_kind=COST_CENSUS,
              cost_weight=CG_WEIGHT, color=-1, fresh=None, out=None,
              viewout=None):
    """One Jacobi propagate of one view, or with color 0 / 1 one pass of
    the red-black schedule (fresh: pass 0's output, read by pass 1);
    returns (new_states, viewout).  out / viewout: existing arrays to
    write into (a red-black pair of passes shares them)."""
    H, W = tex1.shape[:2]
    if out is None:
        out = np.zeros((H, W, 4), np.float32)
    vout = viewout
    if vout is None and scatter:
        vout = viewbuf_empty(W, H)
    if viewin is not None:
        viewin = np.ascontiguousarray(viewin, np.uint64)
        src_other = _f32(src_other, (H, W, 4))
    lib().pmo_propagate(_ptr(_f32(tex1)), _ptr(_f32(tex2)), W, H, int(radius),
                        float(color_weight), int(cost_kind),
                        float(cost_weight), float(dmax), float(theta),
                        int(lr), int(it), _ptr(_f32(old, (H, W, 4))),
                        _ptr(_f32(smooth) if smooth is not None else None),
                        _ptr(viewin), _ptr(src_other), _ptr(vout), _ptr(out),
                        int(color),
                        _ptr(_f32(fresh, (H, W, 4)) if fresh is not None
                             else None))
    return out, vout


def consistency(left, right, maxdisp=1.0, maxdeg=5.0):
    H, W = left.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().pmo_consistency(_ptr(_f32(left)), _ptr(_f32(right)), W, H,
                          float(maxdisp), float(maxdeg), _ptr(out))
    return out


def to_disparity(states, lr=1, scale=1.0):
    H, W = states.shape[:2]
    out = np.zeros((H, W), np.float32)
    lib().pmo_to_disparity(_ptr(_f32(states)), W, H, float(scale), int(lr),
                           _ptr(out))
    return out


SCHEDULE_JACOBI, SCHEDULE_RED_BLACK = 0, 1


def run(left, right, radius, dmax, iters, maxdisp=1.0, maxdeg=180.0,
        trace=False, cost_kind=COST_CENSUS, cost_weight=CG_WEIGHT,
        schedule=SCHEDULE_JACOBI):
    """Plain PatchMatch pipeline; returns dict of stateL, stateR, disp,
    and (if trace) cost_trace [iters, 2, H, W]."""
    left, W, H, C, stride = _img_args(left)
    right = np.ascontiguousarray(right, dtype=np.float32)
    assert right.shape == left.shape
    sL = np.zeros((H, W, 4), np.float32)
    sR = np.zeros((H, W, 4), np.float32)
    disp = np.zeros((H, W), np.float32)
    tr = np.zeros((iters, 2, H, W), np.float32) if trace else None
    rc = lib().pmo_run(_ptr(left), _ptr(right), W, H, C, stride, int(radius),
                       float(dmax), int(iters), int(cost_kind),
                       float(cost_weight), int(schedule), float(maxdisp),
                       float(maxdeg),
                       _ptr(sL), _ptr(sR), _ptr(disp), _ptr(tr))
    if rc != 0:
        raise MemoryError("pmo_run failed")
    out = {"stateL": sL, "stateR": sR, "disp": disp}
    if trace:
        out["cost_trace"] = tr
    return out


def scatter(states, lr):
    """View buffer the field's propagate leaves behind (fresh buffer)."""
    H, W = states.shape[:2]
    vb = viewbuf_empty(W, H)
    lib().pmo_scatter(_ptr(_f32(states, (H, W, 4))), W, H, int(lr), _ptr(vb))
    return vb


def fill_depthmap(states, scale=1.0):
    H, W = states.shape[:2]
    out = np.zeros((H, W), np.float32)
    lib().pmo_fill_depthmap(_ptr(_f32(states)), W, H, float(scale), _ptr(out))
    return out


def fill_normalmap(states):
    H, W = states.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().pmo_fill_normalmap(_ptr(_f32(states)), W, H, _ptr(out))
    return out


def disparity_to_normal(disp):
    disp = _f32(disp)
    H, W = disp.shape
    out = np.zeros((H, W, 4), np.float32)
    lib().pmo_disparity_to_normal(_ptr(disp), W, H, _ptr(out))
    return out


def num_threads():
    return lib().pmo_num_threads()


# ---- PM-Huber smoothing coupling -----------------------------------------
HUBER_THETA_MIN = 1e-4      # disparityMap's theta_min, PMHuberStereo.cpp:92
HUBER_ALPHA = 24.0          # PMHuberStereo.cpp:416


def diffusion_tensor(img, alpha=HUBER_ALPHA):
    img, W, H, C, stride = _img_args(img)
    out = np.zeros((H, W, 4), np.float32)
    lib().pmo_diffusion_tensor(_ptr(img), W, H, C, stride, float(alpha),
                               _ptr(out))
    return out


def normal_depth(states, dispscale, lr):
    H, W = states.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().pmo_normal_depth(_ptr(_f32(states)), W, H, float(dispscale),
                           int(lr), _ptr(out))
    return out


def rof_step(u, p, q, f, dt, lam):
    H, W = u.shape[:2]
    uo = np.zeros((H, W, 4), np.float32)
    po = np.zeros((H, W, 8), np.float32)
    qo = np.zeros((H, W, 4), np.float32)
    lam = np.asarray(lam, np.float32)
    lib().pmo_rof_step(_ptr(_f32(u)), _ptr(_f32(p)), _ptr(_f32(q)),
                       _ptr(_f32(f)), _ptr(_f32(dt)), _ptr(lam), W, H,
                       _ptr(uo), _ptr(po), _ptr(qo))
    return uo, po, qo


def rof(f, dt, lam, iters=40):
    H, W = f.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lam = np.asarray(lam, np.float32)
    if lib().pmo_rof(_ptr(_f32(f)), _ptr(_f32(dt)), _ptr(lam), int(iters), W,
                     H, _ptr(out)):
        raise ValueError("pmo_rof failed")
    return out


def run_huber(left, right, radius, dmax, iters, theta_min=HUBER_THETA_MIN,
              maxdisp=1.0, maxdeg=5.0):
    """PM-Huber pipeline; returns dict of stateL, stateR, disp, smoothL,
    smoothR."""
    left, W, H, C, stride = _img_args(left)
    right = np.ascontiguousarray(right, dtype=np.float32)
    assert right.shape == left.shape
    out = {k: np.zeros((H, W, 4), np.float32)
           for k in ("stateL", "stateR", "smoothL", "smoothR")}
    out["disp"] = np.zeros((H, W), np.float32)
    rc = lib().pmo_run_huber(_ptr(left), _ptr(right), W, H, C, stride,
                             int(radius), float(dmax), int(iters),
                             float(theta_min), float(maxdisp), float(maxdeg),
                             _ptr(out["stateL"]), _ptr(out["stateR"]),
                             _ptr(out["disp"]), _ptr(out["smoothL"]),
                             _ptr(out["smoothR"]))
    if rc != 0:
        raise MemoryError("pmo_run_huber failed")
    return out


def huber_step(u, p, q, f, dt, lam, eps):
    H, W = u.shape[:2]
    uo = np.zeros((H, W, 4), np.float32)
    po = np.zeros((H, W, 8), np.float32)
    qo = np.zeros((H, W, 4), np.float32)
    lam = np.asarray(lam, np.float32)
    lib().pmo_huber_step(_ptr(_f32(u)), _ptr(_f32(p)), _ptr(_f32(q)),
                         _ptr(_f32(f)), _ptr(_f32(dt)), _ptr(lam), float(eps),
                         W, H, _ptr(uo), _ptr(po), _ptr(qo))
    return uo, po, qo


def huber(f, dt, lam, eps, iters=40):
    H, W = f.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lam = np.asarray(lam, np.float32)
    if lib().pmo_huber(_ptr(_f32(f)), _ptr(_f32(dt)), _ptr(lam), float(eps),
                       int(iters), W, H, _ptr(out)):
        raise ValueError("pmo_huber failed")
    return out


def fill_state(states, dispscale, lr):
    H, W = states.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().pmo_fill_state(_ptr(_f32(states)), W, H, float(dispscale), int(lr),
                         _ptr(out))
    return out


def run_depthmap(left, right, radius, dmax, iters, maxdisp=1.0, maxdeg=5.0):
    """depthMap pipeline; returns dict of stateL, stateR, disp, smoothL,
    smoothR (disparity at scale 1)."""
    left, W, H, C, stride = _img_args(left)
    right = np.ascontiguousarray(right, dtype=np.float32)
    assert right.shape == left.shape
    out = {k: np.zeros((H, W, 4), np.float32)
           for k in ("stateL", "stateR", "smoothL", "smoothR")}
    out["disp"] = np.zeros((H, W), np.float32)
    rc = lib().pmo_run_depthmap(_ptr(left), _ptr(right), W, H, C, stride,
                                int(radius), float(dmax), int(iters),
                                float(maxdisp), float(maxdeg),
                                _ptr(out["stateL"]), _ptr(out["stateR"]),
                                _ptr(out["disp"]), _ptr(out["smoothL"]),
                                _ptr(out["smoothR"]))
    if rc != 0:
        raise MemoryError("pmo_run_depthmap failed")
    return out
