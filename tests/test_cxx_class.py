"""The C++ drop-in class (include/cvt_amd/PMHuberStereo.h) driven from a
C++ program, as a cvt caller would drive cvt::PMHuberStereo::disparityMap
(cvt/vision/PMHuberStereo.h:47-48, PMHuberStereo.cpp:75-143)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "disparity_map.cpp")


def build_driver(tmp_path):
    from cvt_amd import _native
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    libdir = os.path.dirname(_native.SO_PATH)
    exe = str(tmp_path / "disparity_map")
    subprocess.run([gxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
                    "-L", libdir, "-lcvt_amd_pm", "-Wl,-rpath," + libdir],
                   check=True)
    return exe


def test_driver_builds(tmp_path):
    assert os.path.exists(build_driver(tmp_path))


def bits(a):
    a = np.ascontiguousarray(a, np.float32).copy()
    a[np.isnan(a)] = np.nan
    return a.view(np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("pp,mode", [(1, "huber"), (2, "huber"), (1, "plain"),
                                     (2, "depth")])
def test_disparity_map_matches_oracle(tmp_path, oracle, synth, pp, mode):
    O = oracle
    W, H, C, R, dmax, iters = 72, 40, 4, 3, 14.0, 2
    L, Rt, _ = synth.make_pair(W, H, cells=4, dmax=dmax, channels=C, seed=21)
    exe = build_driver(tmp_path)
    paths = [str(tmp_path / n) for n in ("l", "r", "d", "p", "n")]
    L.astype(np.float32).tofile(paths[0])
    Rt.astype(np.float32).tofile(paths[1])
    print("cxx driver: %dx%d C%d r=%d pp=%d %s" % (W, H, C, R, pp, mode),
          flush=True)
    r = subprocess.run([exe, str(W), str(H), str(C), str(R), str(dmax),
                        str(iters), str(pp)] + paths +
                       ([mode] if mode in ("plain", "depth") else []),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # disparityMap makes the iteration count odd (PMHuberStereo.cpp:490),
    # depthMap runs the count it is given
    n = iters if mode == "depth" else iters | 1
    assert "iterations run: %d" % n in r.stdout
    assert "mismatch:" in r.stdout
    d = np.fromfile(paths[2], np.float32).reshape(H, W)
    planes = np.fromfile(paths[3], np.float32).reshape(H, W, 4)
    normals = np.fromfile(paths[4], np.float32).reshape(H, W, 4)
    # the class's default is the reference's disparityMap pipeline
    # (PM-Huber, theta_min 1e-4)
    if mode == "huber":
        ref = O.run_huber(L, Rt, R, dmax, iters | 1, 1e-4, 1.0, 5.0)
    elif mode == "depth":
        ref = O.run_depthmap(L, Rt, R, dmax, iters, 1.0, 5.0)
    else:
        ref = O.run(L, Rt, R, dmax, iters | 1, 1.0, 5.0)
    np.testing.assert_array_equal(bits(planes), bits(ref["stateL"]))
    chk = O.consistency(ref["stateL"], ref["stateR"], 1.0, 5.0)
    want = O.to_disparity(chk) if pp == 1 else O.fill_depthmap(chk)
    if mode == "depth":   # the default dscale, 1 / depthmax
        want = want * (np.float32(1.0) / np.float32(dmax))
    np.testing.assert_array_equal(bits(d), bits(want))
    np.testing.assert_array_equal(bits(normals), bits(O.fill_normalmap(chk)))
