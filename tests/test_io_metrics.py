"""PFM I/O and disparity error metrics of the engine library (host code,
include/pm_io.h), against hand-written PFM fixtures, the oracle's numpy
restatement and the quirks of cvt's Disparity metrics.  CPU only."""
import json
import os

import numpy as np
import pytest

from cvt_amd import disparity as D
from cvt_amd import pfm
from oracle import metrics as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["gray_be.pfm", "rgb_le.pfm"])
def test_read_hand_written_fixtures(name):
    want = json.load(open(os.path.join(GOLDEN, "pfm_expected.json")))[name]
    got = pfm.load_pfm(os.path.join(GOLDEN, name))
    assert list(got.shape) == want["shape"]
    np.testing.assert_array_equal(got, np.asarray(want["data"], np.float32))
    np.testing.assert_array_equal(M.read_pfm(os.path.join(GOLDEN, name)),
                                  got)


@pytest.mark.parametrize("shape", [(5, 7), (4, 3, 4), (1, 1), (9, 1, 4)])
def test_write_matches_oracle_bytes_and_round_trips(tmp_path, shape):
    rng = np.random.default_rng(sum(shape))
    img = rng.normal(0, 10, shape).astype(np.float32)
    if len(shape) == 3:
        img[..., 3] = 1.0          # PF keeps RGB; alpha reads back as 1
    p_prod, p_orc = tmp_path / "prod.pfm", tmp_path / "orc.pfm"
    pfm.save_pfm(p_prod, img)
    M.write_pfm(p_orc, img)
    assert p_prod.read_bytes() == p_orc.read_bytes()
    np.testing.assert_array_equal(pfm.load_pfm(p_prod), img)


def test_big_endian_written_by_oracle(tmp_path):
    img = np.arange(12, dtype=np.float32).reshape(3, 4) - 5.5
    p = tmp_path / "be.pfm"
    M.write_pfm(p, img, little_endian=False, comment="big endian")
    np.testing.assert_array_equal(pfm.load_pfm(p), img)


def test_special_values_survive(tmp_path):
    img = np.array([[np.nan, np.inf, -np.inf, -0.0, 1e-40]], np.float32)
    p = tmp_path / "s.pfm"
    pfm.save_pfm(p, img)
    back = pfm.load_pfm(p)
    np.testing.assert_array_equal(back.view(np.uint32), img.view(np.uint32))


@pytest.mark.parametrize("content,what", [
    (b"P5\n2 2\n255\n", "unknown PFM type"),
    (b"Pf\n2\n-1.0\n" + b"\0" * 16, "expected <width>"),
    (b"Pf\n2 2\n-1.0\n" + b"\0" * 15, "not enough PFM data"),
    (b"Pf\n0 2\n-1.0\n", "invalid PFM size"),
    (b"", "header corrupt"),
    # W * H * 3 channels * 4 bytes wraps a 64-bit count to 11936
    (b"PF\n2147380029 715862424\n-1.0\n" + b"\0" * 11936,
     "not enough PFM data"),
])
def test_malformed_files_are_errors(tmp_path, content, what):
    p = tmp_path / "bad.pfm"
    p.write_bytes(content)
    with pytest.raises(D.IOError_) as e:
        pfm.load_pfm(p)
    assert e.value.code == 5 and what in str(e.value)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(D.IOError_) as e:
        pfm.load_pfm(tmp_path / "absent.pfm")
    assert "cannot open" in str(e.value)


@pytest.mark.parametrize("sizes,shape", [
    (b"0x3 02", (2, 3)),          # base 0: hexadecimal and octal
    (b"010 1", (1, 8)),
    (b"4294967299 1", (1, 3)),    # 2^32 + 3: the int keeps the low 32 bits
])
def test_header_sizes_read_as_the_reference_reads_them(tmp_path, sizes,
                                                       shape):
    """String::toInteger is strtol(s, NULL, 0) (cvt/util/String.h:510-513),
    stored into PFMHeader's int fields (PFMLoader.cpp:38-39,94-95)."""
    H, W = shape
    data = np.arange(W * H, dtype="<f4")
    p = tmp_path / "h.pfm"
    p.write_bytes(b"Pf\n" + sizes + b"\n-1.0\n" + data.tobytes())
    np.testing.assert_array_equal(pfm.load_pfm(p),
                                  data.reshape(H, W)[::-1])


def random_maps(seed, H=40, W=50):
    rng = np.random.default_rng(seed)
    gt = rng.uniform(0, 64, (H, W)).astype(np.float32)
    gt[rng.random((H, W)) < 0.2] = 0.0          # no ground truth
    est = (gt + rng.normal(0, 1.5, (H, W))).astype(np.float32)
    est[rng.random((H, W)) < 0.1] = 0.0         # LR-rejected
    est[rng.random((H, W)) < 0.02] = np.nan
    return est, gt


@pytest.mark.parametrize("abs_t,rel_t", [(1.0, 0.0), (3.0, 0.05), (0.5, 0.2)])
def test_evaluate_error_matches_oracle(abs_t, rel_t):
    est, gt = random_maps(int(abs_t * 10 + rel_t * 100))
    assert D.evaluate_error(est, gt, abs_t, rel_t) == \
        M.evaluate_error(est, gt, abs_t, rel_t)


def test_evaluate_error_closed_forms():
    gt = np.array([[0.0, 10.0, 10.0, 10.0, 10.0, -3.0]], np.float32)
    est = np.array([[5.0, 10.5, 12.0, 0.0, np.nan, 1.0]], np.float32)
    # gt > 0: columns 1-4; errors at 1 px: 12.0 (2 px), 0.0 (10 px);
    # a NaN estimate is not an error (the comparisons are false)
    assert D.evaluate_error(est, gt, 1.0, 0.0) == (2, 2, 4)
    # relative threshold 0.5 keeps only the 10 px miss
    assert D.evaluate_error(est, gt, 1.0, 0.5) == (1, 2, 4)
    e, dens = D.evaluate_error_rate(est, gt, 1.0, 0.0)
    assert (e, dens) == (0.5, 0.5)
    assert D.bad_pixel_rate(est, gt, 1.0) == 0.5


@pytest.mark.parametrize("relative", [False, True])
def test_histogram_matches_oracle(relative):
    est, gt = random_maps(7 + relative)
    bins, got = D.histogram_error(gt, est, relative=relative)
    np.testing.assert_array_equal(
        bins, np.float32(M.REL_BINS if relative else M.ABS_BINS))
    np.testing.assert_array_equal(got, M.histogram_error(gt, est, bins,
                                                         relative))


def test_histogram_quirks():
    gt = np.array([[2.0, 2.0, 2.0, 4.0, 0.0]], np.float32)
    est = np.array([[2.0, 2.5, 4.5, 3.0, 1.0]], np.float32)
    bins = np.array([0.5, 1.0, 1e10], np.float32)
    _, h = D.histogram_error(gt, est, bins)
    # 4 usable pixels; errors 0 (in no bin), 0.5, 2.5, 1.0
    np.testing.assert_array_equal(h, np.float32([0.25, 0.25, 0.25]))
    # no usable pixel: 0 / 0
    _, h = D.histogram_error(np.zeros((2, 2)), np.ones((2, 2)), bins)
    assert np.isnan(h).all()


def test_default_histograms_are_the_reference_bins():
    np.testing.assert_array_equal(D.default_histogram(False),
                                  np.float32(M.ABS_BINS))
    np.testing.assert_array_equal(D.default_histogram(True),
                                  np.float32(M.REL_BINS))


def test_cxx_mirror_of_pfm_and_disparity(tmp_path):
    """The C++ mirror (include/cvt_amd/Disparity.h) from a C++ program."""
    import shutil
    import subprocess
    from cvt_amd import _native
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    root = os.path.dirname(os.path.dirname(GOLDEN))
    libdir = os.path.dirname(_native.SO_PATH)
    exe = str(tmp_path / "pfm_metrics")
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cxx", "pfm_metrics.cpp"),
                    "-o", exe, "-L", libdir, "-lcvt_amd_pm",
                    "-Wl,-rpath," + libdir], check=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    # 5 of 6 columns carry ground truth on 4 rows; rows 2 and 3 are off by
    # 1.5 and 2.25 px
    assert lines[0] == "10 20 20"
    gt = M.read_pfm(tmp_path / "gt.pfm")
    est = M.read_pfm(tmp_path / "est.pfm")
    assert M.evaluate_error(est, gt, 1.0, 0.0) == (10, 20, 20)
    hist = dict(tok.split(":") for tok in lines[1].split())
    want = M.histogram_error(gt, est, M.ABS_BINS)
    for edge, v in zip(np.float32(M.ABS_BINS), want):
        assert abs(float(hist["%g" % edge]) - v) < 1e-6
    assert lines[2] == "missing: 5"
