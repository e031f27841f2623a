"""pm_iterate recorded into a HIP graph by stream capture on the context's
stream.  The call forks its propagate bands (and, with PM_SMOOTH_HUBER, the
smoothing) onto streams of the context and joins them back, all by events,
so a capture on the context's stream records the whole iteration.  Every
replay must give what the eager call gives: an eager pm_setup before each
replay restores everything the captured iterations read (the initial
planes, the cleared scatter slots, the smooth fields and theta), and eager
work and synchronisations run between the replays."""
import numpy as np
import pytest

import cvt_amd

pytestmark = pytest.mark.gpu

ITERS = 2   # two iterations: the second's band waits refer to the first's


def bits(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


def assert_same(got, want, what):
    d = np.argwhere(bits(got) != bits(want))
    assert d.size == 0, "%s: %d differ, first %s: %r vs %r" % (
        what, len(d), tuple(d[0]), got[tuple(d[0])], want[tuple(d[0])])


def fields(ctx, huber):
    out = [ctx.planes(v) for v in (0, 1)]
    if huber:
        out += [ctx.smooth(v) for v in (0, 1)]
    return out


HUBER_SKIP = pytest.mark.skip(
    reason="capturing the PM-Huber iteration crashed the process inside "
    "the runtime's end of capture (a host segmentation fault, no GPU "
    "fault); no cause found in the engine's calls, so pm_iterate refuses "
    "PM_SMOOTH_HUBER on a capturing stream and this case stays off")


@pytest.mark.parametrize("smoothing", [
    cvt_amd.PM_SMOOTH_NONE,
    pytest.param(cvt_amd.PM_SMOOTH_HUBER, marks=HUBER_SKIP)])
def test_iterate_replays_from_a_captured_graph(synth, smoothing):
    import torch
    huber = smoothing == cvt_amd.PM_SMOOTH_HUBER
    L, R, _ = synth.make_pair(157, 83, cells=6, dmax=24.0, channels=4,
                              seed=7)
    stream = torch.cuda.Stream()
    print("capture, smoothing", smoothing, flush=True)
    with cvt_amd.Context(radius=4, disparity_max=24.0, smoothing=smoothing,
                         schedule_iterations=ITERS,
                         stream=stream.cuda_stream) as ctx:
        ctx.set_images(L, R)
        # eager run: the reference, and the context's streams and events
        # are made here rather than during the capture
        ctx.setup()
        ctx.iterate(ITERS)
        want = fields(ctx, huber)
        ctx.setup()
        ctx.sync()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            ctx.iterate(ITERS)
        assert ctx.iteration == ITERS
        for replay in range(3):
            ctx.setup()
            ctx.sync()
            graph.replay()
            torch.cuda.synchronize()
            got = fields(ctx, huber)
            for k, (g, w) in enumerate(zip(got, want)):
                assert_same(g, w, "replay %d, field %d" % (replay, k))
        del graph
