"""Batched reconstructions of mixed box sizes in one HIP graph.

Every replay of the graph must give what the same reconstructions give when
run one after another eagerly, and the graph is replayed several times with
eager work in between (graph replays are checked one by one, not once).
"""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import to_device
from rexlib_amd import batched, em, synthetic

pytestmark = pytest.mark.gpu

BOXES = [32, 48, 64, 40, 60, 32]
COUNTS = [120, 300, 80, 500, 64, 33]


def make_recons(seed=0):
    recons = []
    for i, (n, count) in enumerate(zip(BOXES, COUNTS)):
        vol = synthetic.blob_volume(n, 10, n * 0.3, (1.0, 3.0), seed=seed + 10 * i)
        stack = synthetic.particle_stack(vol, count, n, seed=seed + 10 * i + 1, snr=0.3, device="cuda")
        recons.append(batched.Reconstruction(stack.images, to_device(stack.transforms), to_device(stack.ctf)))
    return recons


def snapshot(recons):
    return [(r.numerator.clone(), r.denominator.clone(), r.volume.clone()) for r in recons]


def test_graph_replays_equal_sequential_runs():
    recons = make_recons()
    batched.run_sequential(recons, tuning=False)
    torch.cuda.synchronize()
    expected = snapshot(recons)

    b = batched.BatchedReconstruction(recons, streams=3, tuning=False).capture()
    for replay in range(3):
        for r in recons:  # eager work between replays
            r.volume.fill_(float("nan"))
            r.numerator.mul_(2.0)
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        for (num, den, vol), r in zip(expected, recons):
            assert oracle.rel_l2(r.numerator.cpu().numpy(), num.cpu().numpy()) < 1e-5, replay
            assert oracle.rel_l2(r.denominator.cpu().numpy(), den.cpu().numpy()) < 1e-5, replay
            assert oracle.rel_l2(r.volume.cpu().numpy(), vol.cpu().numpy()) < 1e-4, replay


def test_one_batched_reconstruction_matches_the_oracle():
    recons = make_recons(seed=5)[:2]
    b = batched.BatchedReconstruction(recons, streams=2, tuning=False).capture()
    b.replay()
    torch.cuda.synchronize()
    r = recons[1]
    n = r.extent
    slices = np.fft.rfft2(r.images.cpu().numpy().astype(np.float64))
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    ref = oracle.insert_f64("ctf", [acc, acc.real.copy(), slices, r.transforms.cpu().numpy().astype(np.float64),
                                    r.ctf.cpu().numpy().astype(np.float64), np.ones(n // 2 + 1)],
                            oracle.Window.linear(), pixel_size=1.5)
    assert oracle.rel_l2(r.numerator.cpu().numpy(), ref[0]) <= 1e-4
    assert oracle.rel_l2(r.denominator.cpu().numpy(), ref[1]) <= 1e-4


def test_replay_after_a_larger_eager_call_on_the_capture_streams():
    # A graph holds the workspace addresses of its capture. A larger eager
    # insertion on the same streams afterwards grows their workspaces; the
    # areas the graph holds must stay valid (they are retired, never freed),
    # so a later replay still equals the sequential run.
    recons = make_recons(seed=11)
    batched.run_sequential(recons, tuning=False)
    torch.cuda.synchronize()
    expected = snapshot(recons)
    b = batched.BatchedReconstruction(recons, streams=2, tuning=False).capture()
    big = make_recons(seed=12)
    bigger = batched.Reconstruction(torch.cat([big[3].images] * 4), torch.cat([big[3].transforms] * 4),
                                    torch.cat([big[3].ctf] * 4))
    bigger.allocate()
    for stream in b.streams:
        with torch.cuda.stream(stream):
            batched.run_one(bigger, em.ExecutionContext(stream=stream))
    torch.cuda.synchronize()
    for r in recons:
        r.numerator.fill_(float("nan"))
    b.replay()
    torch.cuda.synchronize()
    for (num, den, vol), r in zip(expected, recons):
        assert oracle.rel_l2(r.numerator.cpu().numpy(), num.cpu().numpy()) < 1e-5
        assert oracle.rel_l2(r.denominator.cpu().numpy(), den.cpu().numpy()) < 1e-5
        assert oracle.rel_l2(r.volume.cpu().numpy(), vol.cpu().numpy()) < 1e-4
