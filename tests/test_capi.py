"""The C ABI library without a GPU: it loads, exports every entry point
include/pm_stereo.h, include/pm_io.h, include/pm_net.h and
include/pm_batch.h declare, its headers compile as C and C++, parameters
are validated before any device is touched, and a host without a HIP
device gets PM_ERR_HIP (no CPU fallback)."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pm_stereo.h")
IO_HEADER = os.path.join(ROOT, "include", "pm_io.h")
NET_HEADER = os.path.join(ROOT, "include", "pm_net.h")
BATCH_HEADER = os.path.join(ROOT, "include", "pm_batch.h")
CXX_HEADER = os.path.join(ROOT, "include", "cvt_amd", "PMHuberStereo.h")


def declared_functions():
    names = set()
    for h in (HEADER, IO_HEADER, NET_HEADER, BATCH_HEADER):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(pm_[a-z_0-9]+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def native():
    from cvt_amd import _native
    return _native


def test_library_exports_every_declared_symbol(native):
    decl = declared_functions()
    assert len(decl) >= 27
    assert decl == set(native.SIGNATURES), decl ^ set(native.SIGNATURES)
    L = native.lib()
    for name in decl:
        assert getattr(L, name) is not None


def test_exported_dynamic_symbols(native):
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", native.SO_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert declared_functions() <= exported


def test_params_default_are_reference_constants(native):
    p = native.PmParams()
    native.lib().pm_params_default(ctypes.byref(p))
    assert p.radius == 5                 # cvt::PMHuberStereo patchsize
    assert p.disparity_max == 64.0
    assert p.color_weight == 12.0        # CENSUS_COLOR_WEIGHT
    assert p.cost_kind == native.PM_COST_CENSUS     # patch_eval_census
    assert p.cost_weight == 16.0         # patch_eval_color_grad_weighted
    assert p.schedule == native.PM_SCHEDULE_JACOBI  # the reference's sweep
    assert p.smoothing == native.PM_SMOOTH_NONE     # plain PatchMatch
    assert p.theta_min == np.float32(1e-4)          # PMHuberStereo.cpp:92
    assert p.schedule_iterations == 0
    assert p.device == 0 and not p.stream


@pytest.mark.parametrize("field,value", [
    ("radius", 0), ("radius", 17), ("disparity_max", 0.0),
    ("disparity_max", float("nan")), ("disparity_max", float("inf")),
    ("color_weight", -1.0), ("theta_min", -0.5), ("theta_min", float("nan")),
    ("smoothing", 3), ("schedule_iterations", -1), ("device", -1),
    ("cost_kind", 2), ("cost_kind", -1), ("cost_weight", -1.0),
    ("schedule", 2), ("schedule", -1),
    ("cost_weight", float("inf"))])
def test_create_rejects_bad_parameters(native, field, value):
    L = native.lib()
    p = native.PmParams()
    L.pm_params_default(ctypes.byref(p))
    setattr(p, field, value)
    ctx = ctypes.c_void_p()
    rc = L.pm_create(ctypes.byref(p), ctypes.byref(ctx))
    try:
        assert rc == native.PM_ERR_ARG
        assert L.pm_last_error(ctx)
    finally:
        L.pm_destroy(ctx)


def test_null_arguments(native):
    L = native.lib()
    assert L.pm_create(None, None) == native.PM_ERR_ARG
    assert L.pm_last_error(None) == b"context is NULL"
    L.pm_destroy(None)
    for name in ("pm_setup", "pm_sync"):
        assert getattr(L, name)(None) == native.PM_ERR_ARG
    assert L.pm_iterate(None, 1) == native.PM_ERR_ARG


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="host has a HIP device")
def test_no_device_is_an_error_not_a_fallback(native):
    L = native.lib()
    p = native.PmParams()
    L.pm_params_default(ctypes.byref(p))
    ctx = ctypes.c_void_p()
    rc = L.pm_create(ctypes.byref(p), ctypes.byref(ctx))
    try:
        assert rc == native.PM_ERR_HIP
        assert b"HIP device" in L.pm_last_error(ctx)
    finally:
        L.pm_destroy(ctx)
    import cvt_amd
    with pytest.raises(cvt_amd.PMError):
        cvt_amd.Context(radius=2, disparity_max=8.0)


def _batch_params(native, **kw):
    p = native.PmParams()
    native.lib().pm_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _batch_pair(native, **kw):
    q = native.PmBatchPair(width=32, height=16, channels=4, stride=128,
                           disparity_max=8.0, d_left=1, d_right=1, d_gt=None,
                           gt_stride=0)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


@pytest.mark.parametrize("params,pair", [
    ({"radius": 0}, {}), ({"smoothing": 1}, {}), ({"schedule": 1}, {}),
    ({"cost_kind": 2}, {}), ({}, {"width": 0}), ({}, {"channels": 3}),
    ({}, {"stride": 100}), ({}, {"disparity_max": 0.0}),
    ({}, {"disparity_max": float("nan")}), ({}, {"d_left": None}),
    ({}, {"d_gt": 1, "gt_stride": 8})])
def test_batch_create_rejects_bad_arguments(native, params, pair):
    L = native.lib()
    p = _batch_params(native, **params)
    q = (native.PmBatchPair * 2)(_batch_pair(native),
                                 _batch_pair(native, **pair))
    b = ctypes.c_void_p()
    rc = L.pm_batch_create(ctypes.byref(p), q, 2, ctypes.byref(b))
    try:
        assert rc == native.PM_ERR_ARG
        assert L.pm_batch_last_error(b)
    finally:
        L.pm_batch_destroy(b)


def test_batch_null_arguments(native):
    L = native.lib()
    assert L.pm_batch_create(None, None, 0, None) == native.PM_ERR_ARG
    assert L.pm_batch_last_error(None) == b"NULL batch"
    L.pm_batch_destroy(None)
    assert L.pm_batch_size(None) == -1
    assert L.pm_batch_sync(None) == native.PM_ERR_ARG
    assert L.pm_batch_run(None, 1, 1, 1.0, 5.0, -1.0, 0.0) == \
        native.PM_ERR_ARG


def test_missing_library_raises(native, monkeypatch):
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setattr(native, "SO_PATH", native.SO_PATH + ".absent")
    with pytest.raises(ImportError):
        native.lib()


C_PROGRAM = r"""
#include "pm_stereo.h"
#include "pm_io.h"
#include "pm_net.h"
#include "pm_batch.h"
#include <stdio.h>
int main(void) {
    pm_params p;
    pm_ctx* ctx = 0;
    pm_net* net = 0;
    int n = 0;
    pm_params_default(&p);
    p.radius = 0;
    int rc = pm_create(&p, &ctx);
    printf("%d %d %s\n", p.disparity_max == 64.0f, rc, pm_last_error(ctx));
    pm_destroy(ctx);
    if (pm_default_histogram(0, NULL, &n) != PM_OK || n != 15) return 2;
    /* no HIP device on this host: the net reports it, nothing runs */
    if (pm_net_create(0, NULL, 1, &net) != PM_ERR_HIP) return 3;
    pm_net_destroy(net);
    if (pm_setup_proposal(NULL, NULL) != PM_ERR_ARG) return 4;
    if (pm_setup_proposal_views(NULL, NULL, 1) != PM_ERR_ARG) return 4;
    if (pm_batch_set_proposal_views(NULL, NULL, 1) != PM_ERR_ARG) return 4;
    {   /* a batch checks its arguments before it looks for a device */
        pm_batch* b = 0;
        pm_batch_pair q = {16, 8, 4, 64, 4.0f, (const float*)&p,
                           (const float*)&p, NULL, 0};
        pm_params_default(&p);
        if (pm_batch_create(&p, &q, 0, &b) != PM_ERR_ARG) return 5;
        pm_batch_destroy(b);
        b = 0;
        if (pm_batch_create(&p, &q, 1, &b) != PM_ERR_HIP) return 6;
        pm_batch_destroy(b);
        if (pm_batch_run(NULL, 1, 1, 1.0f, 5.0f, -1.0f, 0.0f) != PM_ERR_ARG)
            return 7;
    }
    return rc == PM_ERR_ARG ? 0 : 1;
}
"""


def test_header_is_c99_and_links(native, tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "t.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "t"
    libdir = os.path.dirname(native.SO_PATH)
    subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic",
                    "-Werror", "-I", os.path.dirname(HEADER), str(src),
                    "-o", str(exe), "-L", libdir, "-lcvt_amd_pm",
                    "-Wl,-rpath," + libdir], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("1 1 ")


def test_cxx_header_compiles(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "t.cpp"
    src.write_text('#include "cvt_amd/PMHuberStereo.h"\n'
                   '#include "cvt_amd/Disparity.h"\n'
                   "int main() { cvt_amd::Image im(4, 3, 4);\n"
                   "  return im.stride() == 16 ? 0 : 1; }\n")
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror",
                    "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.dirname(HEADER), str(src)], check=True)
