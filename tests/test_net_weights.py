"""The proposal network's host side (cvt_amd/net.py): the shipped weights
file holds the parameter vector of the network's layout, files of another
kind, count or length are refused, and the layer table behind the
parameter and FLOP counts follows the GPU tests' topology.  CPU only;
loading the vector into a network (and the library's own count check) is
covered by tests/test_gpu_net.py."""
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


def test_layout_and_flops_follow_the_gpu_tests_topology():
    """net.CONVS (behind the parameter count and the FLOP count the
    network's roofline divides by) against the topology the GPU forward
    test builds its torch reference from; that test's head stores 8
    channels of which 4 are real outputs."""
    from test_gpu_net import CH, CONVS as TOPOLOGY, layout
    assert len(TOPOLOGY) == len(net.CONVS)
    for (in1, in2, o, _, _), (cin, cout, lev) in zip(TOPOLOGY, net.CONVS):
        assert cin == CH[in1][1] + (CH[in2][1] if in2 else 0)
        assert cout == (4 if o == "out" else CH[o][1])
        assert lev == CH[o][0]
    assert layout()[1] == net.param_count()
    # 2 FLOP per multiply-add, 9 taps, every output pixel of its level:
    # the 720p pair and its mirror (padded to 720 x 1280)
    macs = 0
    for in1, in2, o, _, _ in TOPOLOGY:
        lev = CH[o][0]
        cin = CH[in1][1] + (CH[in2][1] if in2 else 0)
        cout = 4 if o == "out" else CH[o][1]
        macs += 9 * cin * cout * (720 >> lev) * (1280 >> lev)
    assert net.forward_flops(720, 1280, N=2) == 2 * 2 * macs
    assert round(net.forward_flops(720, 1280, N=1) / 1e9, 1) == 261.7
