"""numpy restatement of cvt's disparity error metrics and of the PFM file
format, the checker of the engine library's pm_io entry points.

TEST INFRASTRUCTURE ONLY: imported by tests/; the product never imports
it.

* evaluate_error: Disparity::evaluateError (cvt/vision/Disparity.cpp:
  201-233); histogram_error: histogramErrorAbsolute / histogramErrorRelative
  (Disparity.cpp:236-318), including the (previous edge, edge] bins that
  leave an error of exactly 0 out, and the float counts.
* read_pfm / write_pfm: the netpbm PFM layout as PFMLoader (plugins/
  iloader/PFM/PFMLoader.cpp:58-205) and PFMSaver (plugins/isaver/PFM/
  PFMSaver.cpp:117-181) use it: three header lines ("Pf"/"PF", width and
  height, scale whose sign gives the byte order), rows bottom to top.
"""
import numpy as np

ABS_BINS = [0.50, 0.75, 1.00, 2.00, 3.00, 4.00, 5.00, 6.00, 7.00, 8.00,
            9.00, 10.0, 20.0, 40.0, 1e10]
REL_BINS = [0.005, 0.01, 0.02, 0.03, 0.04, 0.05, 0.10, 0.20, 0.40, 1e10]


def evaluate_error(disp, gt, abs_thres=1.0, rel_thres=0.0):
    disp = np.asarray(disp, np.float32)
    gt = np.asarray(gt, np.float32)
    m = gt > 0
    with np.errstate(invalid="ignore", divide="ignore"):
        diff = np.abs(gt[m] - disp[m])
        err = (diff > np.float32(abs_thres)) & \
            ((diff / np.abs(gt[m])) > np.float32(rel_thres))
    return int(err.sum()), int((disp[m] > 0).sum()), int(m.sum())


def histogram_error(gt, test, bins, relative=False):
    gt = np.asarray(gt, np.float32)
    test = np.asarray(test, np.float32)
    bins = np.asarray(bins, np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        m = np.isfinite(gt) & (gt > 0) & np.isfinite(test) & (test > 0)
        diff = np.abs(test[m] - gt[m])
        if relative:
            diff = (diff / np.abs(gt[m])).astype(np.float32)
        out = np.zeros(bins.size, np.float32)
        last = np.float32(0.0)
        for i, b in enumerate(bins):
            out[i] = np.float32(((diff <= b) & (diff > last)).sum())
            last = b
        return out / np.float32(m.sum())


def write_pfm(path, img, little_endian=True, comment=None):
    img = np.asarray(img, np.float32)
    color = img.ndim == 3
    H, W = img.shape[:2]
    data = img[..., :3] if color else img
    head = "PF\n" if color else "Pf\n"
    if comment:
        head += "# %s\n" % comment
    head += "%d %d\n%s\n" % (W, H, "-1.0" if little_endian else "1.0")
    rows = np.ascontiguousarray(data[::-1])
    rows = rows.astype("<f4" if little_endian else ">f4")
    with open(path, "wb") as f:
        f.write(head.encode("ascii"))
        f.write(rows.tobytes())


def read_pfm(path):
    with open(path, "rb") as f:
        raw = f.read()
    lines, pos = [], 0
    while len(lines) < 3:
        end = raw.index(b"\n", pos)
        toks = raw[pos:end].replace(b"\t", b" ").split()
        pos = end + 1
        if toks and not toks[0].startswith(b"#"):
            lines.append(toks)
    color = lines[0][0] == b"PF"
    W, H = int(lines[1][0]), int(lines[1][1])
    scale = float(lines[2][0])
    C = 3 if color else 1
    dt = "<f4" if scale < 0 else ">f4"
    a = np.frombuffer(raw, dt, W * H * C, pos).astype(np.float32)
    a = a.reshape((H, W, C) if color else (H, W))[::-1]
    if color:
        a = np.concatenate([a, np.ones((H, W, 1), np.float32)], axis=2)
    return np.ascontiguousarray(a)
