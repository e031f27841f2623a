"""The proposal network's weights file on the host (cvt_amd/net.py): the
shipped file holds the parameter vector of the network's layout, and files
of another kind, count or length are refused.  CPU only; loading the
vector into a network (and the library's own count check) is covered by
tests/test_gpu_net.py."""
import struct

import numpy as np
import pytest

from cvt_amd import net


def test_shipped_weights_hold_the_layouts_parameter_vector():
    w = net.read_weights(net.WEIGHTS)
    assert w.dtype == np.float32 and w.size == net.param_count()
    assert np.isfinite(w).all()
    # every convolution carries trained weights: no block is all zero
    off = 0
    for cin, cout, _ in net.CONVS:
        size = net._up(cout, 32) * net._up(9 * cin, 32)
        assert np.abs(w[off:off + size]).max() > 0.0, (cin, cout)
        off = net._up(net._up(off + size, 64) + net._up(cout, 32), 64)
    assert off == w.size


def test_param_count_by_hand():
    # the first convolution, 8 -> 32: 32 x 96 weights (72 taps padded to
    # 96), then 32 biases padded to 64, so the second starts at 3136
    first = 32 * 96
    assert net._up(net._up(first, 64) + 32, 64) == 3136
    assert net.param_count() == 1291840


def write(path, magic, count, nfloats):
    with open(path, "wb") as f:
        f.write(magic + struct.pack("<Q", count))
        f.write(np.zeros(nfloats, "<f4").tobytes())


@pytest.mark.parametrize("magic,count,floats,what", [
    (b"PMNETW0\0", "n", "n", "not a proposal weights file"),
    (b"PMNETW1\0", 12, 12, "holds 12 parameters"),
    (b"PMNETW1\0", "n", "n-1", "truncated"),
    (b"PMNETW1\0", "n", 0, "truncated"),
])
def test_bad_weights_files_are_refused(tmp_path, magic, count, floats, what):
    n = net.param_count()
    size = {"n": n, "n-1": n - 1}
    p = tmp_path / "w.bin"
    write(p, magic, size.get(count, count), size.get(floats, floats))
    with pytest.raises(ValueError, match=what):
        net.read_weights(p)


def test_short_header_is_refused(tmp_path):
    p = tmp_path / "w.bin"
    p.write_bytes(b"PMNETW1\0\x01")
    with pytest.raises(ValueError, match="not a proposal weights file"):
        net.read_weights(p)
