"""The engine library's host-only PFM and metrics source
(cvt_amd/csrc/pm_io.cpp) under the host address and undefined-behaviour
sanitizers, on mutated PFM files (tests/cxx/pfm_fuzz.cpp).  Built with g++
from that source alone; no GPU code is compiled or run.  CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM_INCLUDE = "/opt/rocm/include"


def build(tmp_path, source):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    if not os.path.isdir(os.path.join(ROCM_INCLUDE, "hip")):
        pytest.skip("HIP headers not available")
    exe = str(tmp_path / "pfm_fuzz")
    r = subprocess.run(
        [gxx, "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Werror",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-D__HIP_PLATFORM_AMD__", "-I", ROCM_INCLUDE,
         "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(ROOT, "cvt_amd", "csrc"),
         source, os.path.join(ROOT, "tests", "cxx", "pfm_fuzz.cpp"),
         "-o", exe], capture_output=True, text=True)
    if r.returncode != 0 and "asan" in r.stderr.lower():
        pytest.skip("the address sanitizer's runtime is not available")
    assert r.returncode == 0, r.stderr
    return exe


def test_pfm_reader_on_mutated_files_under_sanitizers(tmp_path):
    exe = build(tmp_path, os.path.join(ROOT, "cvt_amd", "csrc", "pm_io.cpp"))
    run = tmp_path / "run"
    run.mkdir()
    r = subprocess.run([exe, str(run), "20000"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    assert words[0] == "cases" and words[1] == "20000"
    accepted, rejected = int(words[3]), int(words[5])
    assert accepted + rejected == 20000
    assert accepted > 1000 and rejected > 1000  # both sides exercised
