"""Python interface of the learned plane proposal (include/pm_net.h).

``ProposalNet`` owns one network on a device: its parameters (fp32 master
copy, loaded from or saved to a weights file), the forward pass, the
training step and the seeding of a ``Context`` through
``pm_setup_proposal``.  Every call runs the HIP kernels of the library;
device buffers are passed as raw pointers (torch tensors' ``data_ptr()``
work).
"""
import ctypes
import os
import struct

import numpy as np

from . import _native
from .stereo import PMError

DISP_SCALE = 128.0   # PM_NET_DISP_SCALE: disparity = 128 * out[0]
IN_CHANNELS = 8
OUT_CHANNELS = 4
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "weights", "proposal.bin")
_MAGIC = b"PMNETW1\0"

# the network's convolutions (pm_net_capi.cpp, kC): input channels, output
# channels and resolution level (2^-level of the padded input); the first
# convolutions of the second and third levels also read the 64-level fine
# and the 48-level coarse matching cost volume
CONVS = [(8, 32, 0), (32, 32, 0), (32 + 64, 64, 1), (64, 64, 1),
         (64 + 48, 96, 2),
         (96, 96, 2), (96, 128, 3), (128, 128, 3), (128, 128, 3),
         (128, 128, 3), (96 + 128, 96, 2), (96, 96, 2), (64 + 96, 64, 1),
         (64, 64, 1), (32 + 64, 32, 0), (32, 32, 0), (32, 4, 0)]


def padded(H, W):
    """the network's input size for an H x W pair: multiples of 8"""
    return (H + 7) // 8 * 8, (W + 7) // 8 * 8


def _up(v, m):
    return (v + m - 1) // m * m


def param_count():
    """floats of the network's parameter vector (pm_net_capi.cpp's layout):
    per convolution its weights, output channels padded to 32 times 9 x
    input channels padded to 32, then its 32-padded biases, each block
    starting on a multiple of 64"""
    off = 0
    for cin, cout, _ in CONVS:
        off = _up(off + _up(cout, 32) * _up(9 * cin, 32), 64)
        off = _up(off + _up(cout, 32), 64)
    return off


def read_weights(path):
    """The parameter vector of a weights file written by ``save``: the
    8-byte magic, the parameter count (little-endian u64), then that many
    little-endian floats.  ValueError for another file, another count or a
    short one."""
    with open(path, "rb") as f:
        head = f.read(16)
        if len(head) < 16 or head[:8] != _MAGIC:
            raise ValueError("%s is not a proposal weights file" % path)
        (n,) = struct.unpack("<Q", head[8:])
        if n != param_count():
            raise ValueError("%s holds %d parameters, the network %d"
                             % (path, n, param_count()))
        data = f.read(4 * n)
    if len(data) != 4 * n:
        raise ValueError("%s is truncated: %d of %d bytes of parameters"
                         % (path, len(data), 4 * n))
    return np.frombuffer(data, dtype="<f4").astype(np.float32)


def forward_flops(H, W, N=1):
    """multiply-add work of one forward pass over N padded H x W inputs
    (2 FLOP per multiply-add, 3 x 3 taps; the head's 4 real outputs)"""
    Hp, Wp = padded(H, W)
    total = 0
    for cin, cout, lev in CONVS:
        total += 2 * 9 * cin * cout * N * (Hp >> lev) * (Wp >> lev)
    return total


class ProposalNet:
    def __init__(self, device=0, stream=None, seed=0, weights=None):
        self._lib = _native.lib()
        h = ctypes.c_void_p()
        rc = self._lib.pm_net_create(int(device), stream, int(seed),
                                     ctypes.byref(h))
        self._h = h
        if rc != _native.PM_OK:
            msg = (self._lib.pm_net_last_error(h).decode() if h
                   else "pm_net_create")
            if h:
                self._lib.pm_net_destroy(h)
            self._h = None
            raise PMError(rc, msg)
        self.device = device
        if weights is not None:
            self.load(weights)

    def close(self):
        if self._h:
            self._lib.pm_net_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != _native.PM_OK:
            raise PMError(rc, self._lib.pm_net_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    @property
    def num_params(self):
        return int(self._lib.pm_net_num_params(self._h))

    def params(self):
        out = np.zeros(self.num_params, np.float32)
        self._check(self._lib.pm_net_get_params(
            self._h, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(self._lib.pm_net_set_params(
            self._h, w.ctypes.data_as(ctypes.c_void_p), w.size))

    def grads(self):
        out = np.zeros(self.num_params, np.float32)
        self._check(self._lib.pm_net_get_grads(
            self._h, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def save(self, path):
        w = self.params()
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "wb") as f:
            f.write(_MAGIC + struct.pack("<Q", w.size))
            f.write(w.astype("<f4").tobytes())

    def load(self, path):
        self.set_params(read_weights(path))

    def forward(self, d_x, N, H, W, d_out):
        """d_x: N x H x W x 8 fp32 device pointer; d_out: N x H x W x 4."""
        self._check(self._lib.pm_net_forward(self._h, ctypes.c_void_p(d_x),
                                             N, H, W,
                                             ctypes.c_void_p(d_out)))

    def train_step(self, d_x, d_target, N, H, W, lr, w_disp=1.0,
                   w_normal=0.05):
        loss = ctypes.c_float()
        self._check(self._lib.pm_net_train_step(
            self._h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_target), N, H, W,
            float(lr), float(w_disp), float(w_normal), ctypes.byref(loss)))
        return loss.value
