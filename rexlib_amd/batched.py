"""Many independent reconstructions in one HIP graph launch.

Each reconstruction (its own box size n = N, particle stack, CTF and
accumulators) runs the whole path: zero the accumulators, rfft2 + CTF
insertion over launches of at most 4,096 images, weight quotient, irfft3,
gridding correction. ``BatchedReconstruction`` captures all of them into one
HIP graph, spread over a few streams (fork and join on the capture stream) so
that small boxes run beside large ones, and replays it with a single launch.
Before capture, one eager pass on the same streams sizes every workspace, so
the capture allocates nothing. The brick shape of each box size comes from
the autotuner's table (``rexlib_amd.tune``).

Reconstructions are assigned to streams largest first, each to the stream
with the least work so far.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from . import em, fourier, tune

LAUNCH = 4096


@dataclass
class Reconstruction:
    images: torch.Tensor        # (B, n, n) float32 on the device
    transforms: torch.Tensor    # (B, 3, 4) float32
    ctf: torch.Tensor           # (B, 4) float32
    pixel_size: float = 1.5
    numerator: torch.Tensor = field(default=None)
    denominator: torch.Tensor = field(default=None)
    volume: torch.Tensor = field(default=None)

    @property
    def extent(self) -> int:
        return int(self.images.shape[-1])

    def allocate(self):
        n = self.extent
        dev = self.images.device
        self.numerator = torch.zeros((n, n, n // 2 + 1), dtype=torch.complex64, device=dev)
        self.denominator = torch.zeros((n, n, n // 2 + 1), dtype=torch.float32, device=dev)
        self.volume = torch.zeros((n, n, n), dtype=torch.float32, device=dev)
        self.radial = torch.ones(n // 2 + 1, dtype=torch.float32, device=dev)
        self.quotient = torch.zeros_like(self.numerator)


WINDOW = em.GriddingWindow.linear()
OPTICS = em.CtfOptics(300.0, 2.7, 0.1)


def run_one(r: Reconstruction, context: em.ExecutionContext | None = None):
    """The whole path of one reconstruction, enqueued on the context's stream.
    The library picks the brick tuning of the reconstruction's extent itself
    (``tune.apply_all`` hands it the table once)."""
    r.numerator.zero_()
    r.denominator.zero_()
    count = r.images.shape[0]
    for s in range(0, count, LAUNCH):
        e = min(s + LAUNCH, count)
        em.insert_ctf_images(r.numerator, r.denominator, r.images[s:e], r.transforms[s:e], r.ctf[s:e], r.radial,
                             WINDOW, OPTICS, r.pixel_size, context=context)
    em.normalize(r.numerator, r.denominator, 1e-3, out=r.quotient, context=context)
    stream = context.stream if context is not None else None
    fourier.irfft3(r.quotient, out=r.volume, stream=stream)
    em.gridding_correct(r.volume, WINDOW, out=r.volume, context=context)


def assign_streams(recons, streams: int):
    """Largest work first, each to the least loaded stream."""
    loads = [0.0] * streams
    groups = [[] for _ in range(streams)]
    order = sorted(range(len(recons)), key=lambda i: -recons[i].images.shape[0] * recons[i].extent ** 2)
    for i in order:
        k = min(range(streams), key=lambda j: loads[j])
        groups[k].append(i)
        loads[k] += recons[i].images.shape[0] * recons[i].extent ** 2 + recons[i].extent ** 3
    return groups


class BatchedReconstruction:
    """All reconstructions captured in one HIP graph over `streams` streams."""

    def __init__(self, recons, streams: int = 4, tuning: bool = True):
        if tuning:
            tune.apply_all()
        self.recons = recons
        self.tuning = tuning
        for r in recons:
            if r.numerator is None:
                r.allocate()
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.groups = assign_streams(recons, streams)
        self.graph = None

    def _enqueue(self, origin: torch.cuda.Stream):
        for stream, group in zip(self.streams, self.groups):
            stream.wait_stream(origin)
            ctx = em.ExecutionContext(stream=stream)
            with torch.cuda.stream(stream):
                for i in group:
                    run_one(self.recons[i], ctx)
        for stream in self.streams:
            origin.wait_stream(stream)

    def capture(self):
        origin = torch.cuda.current_stream()
        # Eager pass on the same streams: sizes every workspace and cache.
        self._enqueue(origin)
        torch.cuda.synchronize()
        capture_stream = torch.cuda.Stream()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=capture_stream):
            self._enqueue(capture_stream)
        torch.cuda.synchronize()
        return self

    def replay(self):
        self.graph.replay()


def run_sequential(recons, tuning: bool = True):
    """The same reconstructions one after another on the current stream."""
    if tuning:
        tune.apply_all()
    for r in recons:
        if r.numerator is None:
            r.allocate()
        run_one(r, None)
