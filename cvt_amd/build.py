"""Builds the engine's shared library, ``cvt_amd/libcvt_amd_pm.so``.

One hipcc invocation compiles the gfx950 kernels, the C ABI and the C++
drop-in class.  Flags:
  -ffp-contract=off       only the fmaf() calls in the sources fuse (the
                          arithmetic contract of pm_device.h);
  -fhip-fp32-correctly-rounded-divide-sqrt
                          IEEE division and sqrt (hipcc's default, stated);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                          no packed-FP32 instructions.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SO = os.path.join(HERE, "libcvt_amd_pm.so")
SOURCES = ["csrc/pm_engine.hip", "csrc/pm_smooth.hip", "csrc/pm_net.hip",
           "csrc/pm_capi.cpp", "csrc/pm_net_capi.cpp",
           "csrc/pm_batch.cpp", "csrc/PMHuberStereo.cpp", "csrc/pm_io.cpp",
           "csrc/pm_noalloc.hip"]
ARCH = os.environ.get("PM_OFFLOAD_ARCH", "gfx950")


def hipcc():
    return os.environ.get("HIPCC") or shutil.which("hipcc") or \
        "/opt/rocm/bin/hipcc"


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
            "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-command-line-argument",
            "-I" + os.path.join(ROOT, "include"),
            "-I" + os.path.join(HERE, "csrc")]


def build(verbose=False):
    tmp = SO + ".tmp"
    cmd = [hipcc()] + flags() + ["-shared", "-o", tmp] + \
        [os.path.join(HERE, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if noise not in l]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, SO)
    return SO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
