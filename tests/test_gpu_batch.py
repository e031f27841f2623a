"""Batched plain PatchMatch (include/pm_batch.h) against single-pair
contexts, through the C ABI.

A batch runs, for every pair, the sequence of a plain-PatchMatch context
(pm_setup, pm_iterate, pm_disparity), so each pair's plane fields and
disparity must equal a single-pair context's bit for bit, for every launch
shape the autotuner may choose, and on every replay of the recorded graph.
The single-pair context is itself bit-exact to the oracle (test_gpu_parity);
one pair is also compared with the oracle directly.
"""
import numpy as np
import pytest

import cvt_amd
from cvt_amd.batch import Batch
from test_gpu_parity import first_diff, same

pytestmark = pytest.mark.gpu

# uneven sizes: partial tile blocks at the right and bottom, one block row,
# one block column, gray and RGBA, disparity ranges 0.1 x width (at least 4)
SIZES = [(150, 70, 4), (64, 4, 4), (33, 41, 1), (211, 9, 4), (96, 120, 1),
         (17, 23, 4), (128, 64, 4)]
RADIUS = 3
ITERS = 3
COSTS = [(cvt_amd.PM_COST_CENSUS, 16.0), (cvt_amd.PM_COST_COLOR_GRAD, 25.5)]


def make_inputs(synth, sizes, seed0=100):
    import torch
    out = []
    for k, (W, H, C) in enumerate(sizes):
        dmax = max(4.0, 0.1 * W)
        L, R, gt = synth.make_pair(W, H, cells=4, dmax=dmax, channels=C,
                                   seed=seed0 + k)
        dev = {"L": torch.from_numpy(np.ascontiguousarray(L)).cuda(),
               "R": torch.from_numpy(np.ascontiguousarray(R)).cuda(),
               "gt": torch.from_numpy(np.ascontiguousarray(gt)).cuda()}
        out.append({"W": W, "H": H, "C": C, "dmax": dmax, "L": L, "R": R,
                    "gt": gt, "dev": dev})
    return out


def batch_pairs(inputs, with_gt=False):
    return [{"width": q["W"], "height": q["H"], "channels": q["C"],
             "disparity_max": q["dmax"], "left": q["dev"]["L"].data_ptr(),
             "right": q["dev"]["R"].data_ptr(),
             "gt": q["dev"]["gt"].data_ptr() if with_gt else None}
            for q in inputs]


def single(q, cost_kind, cost_weight, iters, pp=cvt_amd.DISPARITY_PP_LRCHECK):
    with cvt_amd.Context(radius=RADIUS, disparity_max=q["dmax"],
                         cost_kind=cost_kind, cost_weight=cost_weight) as ctx:
        ctx.set_images(q["L"], q["R"])
        ctx.setup()
        ctx.iterate(iters)
        disp = ctx.disparity(pp, 1.0, 5.0)
        return ctx.planes(0), ctx.planes(1), disp


def check_batch(b, inputs, want, what):
    for i in range(len(inputs)):
        for v in (0, 1):
            got = b.planes(i, v)
            assert same(got, want[i][v]), (what, i, v,
                                           first_diff(got, want[i][v]))
        got = b.disparity(i)
        assert same(got, want[i][2]), (what, i, first_diff(got, want[i][2]))


@pytest.fixture(scope="module")
def inputs(synth):
    return make_inputs(synth, SIZES)


@pytest.mark.parametrize("cost_kind,cost_weight", COSTS)
def test_batch_equals_single_pairs_for_every_shape(inputs, cost_kind,
                                                   cost_weight):
    want = [single(q, cost_kind, cost_weight, ITERS) for q in inputs]
    print("batch of %d pairs, cost %d" % (len(inputs), cost_kind),
          flush=True)
    with Batch(batch_pairs(inputs), radius=RADIUS, cost_kind=cost_kind,
               cost_weight=cost_weight) as b:
        for rows in (2, 4, 8):
            for wg in (8, 40, 0):
                print("  tile rows %d, work-groups %d" % (rows, wg),
                      flush=True)
                b.set_tuning(rows, wg)
                b.run(ITERS, cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)
                check_batch(b, inputs, want, (rows, wg))


def test_graph_replays_with_eager_work_between(inputs):
    import torch
    kind, weight = COSTS[1]
    want = [single(q, kind, weight, 2, cvt_amd.DISPARITY_PP_LRCHECK_FILL)
            for q in inputs]
    with Batch(batch_pairs(inputs), radius=RADIUS, cost_kind=kind,
               cost_weight=weight) as b:
        for replay in range(3):
            b.run(2, cvt_amd.DISPARITY_PP_LRCHECK_FILL, 1.0, 5.0)
            b.sync()
            check_batch(b, inputs, want, ("replay", replay))
            # eager work on the device between the replays
            x = torch.randn(1 << 20, device="cuda")
            torch.cuda.synchronize()
            assert torch.isfinite((x * 2).sum())
        # a run with other arguments records a second graph; the first
        # arguments then record again
        b.run(1, cvt_amd.DISPARITY_PP_NONE, 1.0, 5.0)
        b.run(2, cvt_amd.DISPARITY_PP_LRCHECK_FILL, 1.0, 5.0)
        check_batch(b, inputs, want, "re-recorded")


def test_zero_iterations_is_the_setup(inputs):
    kind, weight = COSTS[0]
    want = [single(q, kind, weight, 0) for q in inputs]
    with Batch(batch_pairs(inputs), radius=RADIUS, cost_kind=kind,
               cost_weight=weight) as b:
        b.run(0)
        check_batch(b, inputs, want, "setup")


def test_error_counts_match_evaluate_error(inputs):
    kind, weight = COSTS[1]
    with Batch(batch_pairs(inputs, with_gt=True), radius=RADIUS,
               cost_kind=kind, cost_weight=weight) as b:
        b.run(ITERS, cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0, 1.0, 0.0)
        counts = b.error_counts()
        for i, q in enumerate(inputs):
            d = b.disparity(i)
            e, v, a = cvt_amd.disparity.evaluate_error(d, q["gt"], 1.0, 0.0)
            assert tuple(int(c) for c in counts[i]) == (e, v, a), i
        with pytest.raises(cvt_amd.PMError):
            b.run(1, cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)
            b.error_counts()     # that run counted nothing


def test_one_pair_against_the_oracle(inputs, oracle):
    q = inputs[0]
    kind, weight = COSTS[1]
    with Batch(batch_pairs([q]), radius=RADIUS, cost_kind=kind,
               cost_weight=weight) as b:
        b.run(ITERS, cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)
        planes, disp = b.planes(0, 0), b.disparity(0)
    ref = oracle.run(q["L"], q["R"], RADIUS, q["dmax"], ITERS, maxdisp=1.0,
                     maxdeg=5.0, cost_kind=kind, cost_weight=weight)
    assert same(planes, ref["stateL"]), first_diff(planes, ref["stateL"])
    assert same(disp, ref["disp"]), first_diff(disp, ref["disp"])


def test_autotune_picks_a_shape_with_the_same_results(inputs):
    kind, weight = COSTS[1]
    want = [single(q, kind, weight, 2) for q in inputs]
    with Batch(batch_pairs(inputs), radius=RADIUS, cost_kind=kind,
               cost_weight=weight) as b:
        table = b.autotune(reps=1)
        assert len(table) >= 3
        assert all(t["ms"] > 0 for t in table)
        best = min(table, key=lambda t: t["ms"])
        chosen = b.tuning
        assert (chosen["tile_rows"], chosen["workgroups"]) == \
            (best["tile_rows"], best["workgroups"])
        b.run(2, cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)
        check_batch(b, inputs, want, "tuned")


def test_set_pair_swaps_images_without_recording(inputs, synth):
    import torch
    kind, weight = COSTS[0]
    q0 = inputs[0]
    L2, R2, _ = synth.make_pair(q0["W"], q0["H"], cells=4, dmax=q0["dmax"],
                                channels=q0["C"], seed=999)
    q2 = dict(q0, L=L2, R=R2,
              dev={"L": torch.from_numpy(L2).cuda(),
                   "R": torch.from_numpy(R2).cuda(), "gt": q0["dev"]["gt"]})
    with Batch(batch_pairs(inputs[:2]), radius=RADIUS, cost_kind=kind,
               cost_weight=weight) as b:
        b.run(1)
        b.set_pair(0, batch_pairs([q2])[0])
        b.run(1)
        want = single(q2, kind, weight, 1)
        assert same(b.planes(0, 0), want[0])
        assert same(b.disparity(0), want[2])
        with pytest.raises(cvt_amd.PMError):
            b.set_pair(0, dict(batch_pairs([q2])[0], width=q0["W"] + 1))


@pytest.mark.parametrize("views", [2, 1])
def test_proposal_seeded_batch_equals_single_pairs(inputs, views):
    """pm_batch_set_proposal_views: every pair seeded as
    pm_setup_proposal_views seeds a context (an untrained network: the
    seeding, not its quality, is under test), then two iterations, inside
    the recorded graph"""
    from cvt_amd.net import ProposalNet
    kind, weight = COSTS[1]
    with ProposalNet(seed=5) as net:
        want = []
        for q in inputs:
            with cvt_amd.Context(radius=RADIUS, disparity_max=q["dmax"],
                                 cost_kind=kind, cost_weight=weight) as ctx:
                ctx.set_images(q["L"], q["R"])
                ctx.setup_proposal(net, views)
                ctx.iterate(2)
                want.append((ctx.planes(0), ctx.planes(1),
                             ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0,
                                           5.0)))
        with Batch(batch_pairs(inputs), radius=RADIUS, cost_kind=kind,
                   cost_weight=weight) as b:
            b.set_proposal(net, views)
            for replay in range(2):
                b.run(2, cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)
                check_batch(b, inputs, want, ("proposal", replay))
            # back to the random init
            b.set_proposal(None)
            b.run(1)
            rnd = single(inputs[0], kind, weight, 1)
            assert same(b.planes(0, 0), rnd[0])
