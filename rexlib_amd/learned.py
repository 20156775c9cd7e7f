"""Learned density compensation with a fallback to the iterative estimate.

The iterative estimate (``iterative_weights``, 10 passes of
W <- W / max(K3 * (W den), floor)) is the reference and the fallback. A small
perceptron (10 -> 64 -> 64 -> 64 -> 1, float16 weights, float32
accumulation on MFMA; ``CompensationModel``) predicts log(W den) -- the
correction that turns the plain quotient 1/den into the compensated weight --
from per-voxel features (``features``): the shell radius, the local sample
count (den against the mean of the positive weights), and an 8-bin
angular-coverage histogram of the slice normals whose planes pass within one
voxel (``coverage_table``, from the transforms); its output layer starts near
zero, so an untrained model gives the plain quotient. It is trained on the device by
the library's own forward, backward and Adam kernels, on the weights the
iterative estimate gives for many orientation distributions, at voxels
within the Nyquist sphere (|g| <= N/2); beyond it, in the corners of the
cube that only the corners of the slices reach, the predicted weights are the
plain quotient. A prediction is
used only while the reconstruction it gives keeps a Fourier shell correlation
of at least 0.99 with the iterative one, up to 0.9 Nyquist; otherwise the
iterative weights are used (``compensate``).

Reconstruction with weights W: V = gridding_correct(irfft3(numerator * W)).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N
from . import em, fourier

PASSES = 10
FLOOR = 1e-6
FEATURES = 16
SUPPORT_FSC = 0.99


def parameter_count() -> int:
    return int(N.library().rexlib_hip_mlp_parameter_count())


def _stream():
    return N.stream_handle()


def iterative_weights(den: torch.Tensor, passes: int = PASSES, floor: float = FLOOR) -> torch.Tensor:
    w = torch.empty_like(den)
    N.check(N.library().rexlib_hip_density_compensation(N.array(den), int(passes), float(floor), N.array(w),
                                                        _stream()))
    return w


def coverage_table(transforms: torch.Tensor, big: int) -> torch.Tensor:
    """(8 radii, 33 theta, 64 phi, 8 angle bins) float32 coverage fractions of
    the slice normals of `transforms` (..., 3, 4) for a volume of extent `big`."""
    out = torch.empty((8, 33, 64, 8), dtype=torch.float32, device=transforms.device)
    N.check(N.library().rexlib_hip_coverage_table(N.array(transforms.contiguous()), int(big), N.array(out), _stream()))
    return out


def density_scale(den: torch.Tensor) -> float:
    """1 / mean of the positive weights (1 when there are none)."""
    positive = den[den > 0]
    return 1.0 / float(positive.mean()) if positive.numel() else 1.0


def features(den: torch.Tensor, table: torch.Tensor, scale: float | None = None) -> torch.Tensor:
    scale = density_scale(den) if scale is None else scale
    out = torch.empty(den.shape + (FEATURES,), dtype=torch.float16, device=den.device)
    N.check(N.library().rexlib_hip_compensation_features(N.array(den), N.array(table), float(scale), N.array(out),
                                                         _stream()))
    return out


# Parameter layout (row-major W[k][n]).
SHAPES = [("w1", (16, 64)), ("w2", (64, 64)), ("w3", (64, 64)), ("w4", (64, 16)),
          ("b1", (64,)), ("b2", (64,)), ("b3", (64,)), ("b4", (16,))]


def unpack(flat: np.ndarray) -> dict:
    out = {}
    offset = 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = flat[offset:offset + size].reshape(shape)
        offset += size
    return out


def initial_parameters(seed: int = 0) -> np.ndarray:
    """He-uniform weights for the 10 real inputs and 64-wide layers, zero biases,
    only column 0 of W4 live."""
    rng = np.random.default_rng(seed)
    p = {name: np.zeros(shape, dtype=np.float64) for name, shape in SHAPES}
    p["w1"][:10] = rng.uniform(-1, 1, (10, 64)) * math.sqrt(6.0 / 10)
    p["w2"][:] = rng.uniform(-1, 1, (64, 64)) * math.sqrt(6.0 / 64)
    p["w3"][:] = rng.uniform(-1, 1, (64, 64)) * math.sqrt(6.0 / 64)
    p["w4"][:, 0] = rng.uniform(-1, 1, 64) * 0.01
    return np.concatenate([p[name].ravel() for name, _ in SHAPES]).astype(np.float32)


@dataclass
class TrainingLog:
    steps: int
    losses: list


class CompensationModel:
    """The perceptron's parameters on the device, with its Adam state."""

    def __init__(self, parameters: np.ndarray | None = None, device="cuda", seed: int = 0):
        p = initial_parameters(seed) if parameters is None else np.asarray(parameters, dtype=np.float32)
        assert p.size == parameter_count()
        self.f32 = torch.from_numpy(p.copy()).to(device)
        self.f16 = self.f32.to(torch.float16)
        self.grads = torch.zeros_like(self.f32)
        self.m = torch.zeros_like(self.f32)
        self.v = torch.zeros_like(self.f32)
        self.step = 0

    def parameters(self) -> np.ndarray:
        return self.f32.cpu().numpy()

    def predict(self, rows: torch.Tensor) -> torch.Tensor:
        """The correction log(W den) for feature rows (M, 16) float16."""
        out = torch.empty(rows.shape[0], dtype=torch.float32, device=rows.device)
        N.check(N.library().rexlib_hip_mlp_forward(N.array(rows), N.array(self.f16), N.array(self.f32), N.array(out),
                                                   _stream()))
        return out

    def gradients(self, rows: torch.Tensor, target: torch.Tensor, loss_scale: float = 1.0):
        """(float32 gradients / loss_scale, loss) without an update (the gradients are cleared first)."""
        self.grads.zero_()
        loss = torch.zeros(1, dtype=torch.float32, device=rows.device)
        N.check(N.library().rexlib_hip_mlp_backward(N.array(rows), N.array(target), N.array(self.f16),
                                                    N.array(self.f32), N.array(self.grads), N.array(loss),
                                                    float(loss_scale), _stream()))
        return self.grads / loss_scale, loss

    def train_step(self, rows, target, lr=1e-3, loss_scale=1024.0) -> torch.Tensor:
        loss = torch.zeros(1, dtype=torch.float32, device=rows.device)
        N.check(N.library().rexlib_hip_mlp_backward(N.array(rows), N.array(target), N.array(self.f16),
                                                    N.array(self.f32), N.array(self.grads), N.array(loss),
                                                    float(loss_scale), _stream()))
        self.step += 1
        N.check(N.library().rexlib_hip_mlp_adam(N.array(self.f32), N.array(self.f16), N.array(self.grads),
                                                N.array(self.m), N.array(self.v), float(lr), 0.9, 0.999, 1e-8,
                                                self.step, float(loss_scale), _stream()))
        return loss

    def train(self, rows: torch.Tensor, targets: torch.Tensor, steps: int, batch: int = 4096, lr: float = 1e-3,
              seed: int = 0) -> TrainingLog:
        g = torch.Generator(device=rows.device)
        g.manual_seed(seed)
        losses = []
        for i in range(steps):
            pick = torch.randint(0, rows.shape[0], (batch,), device=rows.device, generator=g)
            rate = lr * 0.5 * (1 + math.cos(math.pi * i / steps))
            loss = self.train_step(rows[pick].contiguous(), targets[pick].contiguous(), lr=rate)
            if i % max(1, steps // 20) == 0 or i == steps - 1:
                losses.append(float(loss.item()))
        return TrainingLog(steps, losses)

    def weights(self, den: torch.Tensor, table: torch.Tensor) -> torch.Tensor:
        """Predicted W = exp(correction) / den on the grid of `den` (0 where
        den <= 0; the plain quotient 1 / den beyond the Nyquist sphere)."""
        f = features(den, table).reshape(-1, FEATURES)
        correction = self.predict(f).reshape(den.shape)
        correction = torch.where(inside_nyquist(den.shape[0], den.device), correction, torch.zeros_like(correction))
        positive = den > 0
        return torch.where(positive, torch.exp(correction) / torch.where(positive, den, torch.ones_like(den)),
                           torch.zeros_like(den))

    def save(self, path: str):
        np.savez(path, parameters=self.parameters())

    @staticmethod
    def load(path: str, device="cuda") -> "CompensationModel":
        with np.load(path, allow_pickle=False) as f:
            return CompensationModel(f["parameters"], device)


def inside_nyquist(big: int, device) -> torch.Tensor:
    """(N, N, N/2+1) bool: |g| <= N/2 on the half grid (signed frequencies as numpy's fftfreq)."""
    key = (big, str(device))
    if key not in _NYQUIST:
        f = torch.fft.fftfreq(big, d=1.0 / big, device=device, dtype=torch.float64)
        x = torch.arange(big // 2 + 1, device=device, dtype=torch.float64)
        r2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + x[None, None, :] ** 2
        _NYQUIST.clear()
        _NYQUIST[key] = r2 <= (big // 2) ** 2
    return _NYQUIST[key]


_NYQUIST = {}


def shell_index(big: int, device) -> torch.Tensor:
    """(N, N, N/2+1) int64: the shell round(|g|) of each voxel of the half grid."""
    f = torch.fft.fftfreq(big, d=1.0 / big, device=device, dtype=torch.float64)
    x = torch.arange(big // 2 + 1, device=device, dtype=torch.float64)
    r2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + x[None, None, :] ** 2
    return torch.round(torch.sqrt(r2)).to(torch.int64)


def training_rows(den: torch.Tensor, table: torch.Tensor, samples: int, generator: torch.Generator | None = None):
    """(features, target log(W den)) at `samples` voxels with den > 0 within
    the Nyquist sphere, each shell round(|g|) >= 1 drawn with equal
    probability (as the Fourier shell correlation weighs the shells), then a
    voxel of the shell uniformly."""
    w = iterative_weights(den)
    f = features(den, table).reshape(-1, FEATURES)
    d = den.reshape(-1)
    mask = ((d > 0) & inside_nyquist(den.shape[0], den.device).reshape(-1)).nonzero().squeeze(1)
    shells = shell_index(den.shape[0], den.device).reshape(-1)[mask]
    per_shell = torch.bincount(shells).to(torch.float64)
    chance = torch.where(shells > 0, 1.0 / per_shell[shells], torch.zeros_like(per_shell[shells]))
    pick = mask[torch.multinomial(chance, samples, replacement=True, generator=generator)]
    return f[pick].contiguous(), torch.log(w.reshape(-1)[pick] * d[pick]).contiguous()


def reconstruct(numerator: torch.Tensor, weights: torch.Tensor, window=None) -> torch.Tensor:
    """V = gridding_correct(irfft3(numerator * W))."""
    window = window or em.GriddingWindow.linear()
    volume = fourier.irfft3((numerator * weights).contiguous())
    return em.gridding_correct(volume, window, out=volume)


MODEL_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "compensation_model.npz")


def shipped_model(device="cuda") -> CompensationModel | None:
    """The model trained by ``rexlib_amd.train_compensation`` and shipped with
    the package, or None when the file is absent."""
    return CompensationModel.load(MODEL_FILE, device) if os.path.exists(MODEL_FILE) else None


def judge(numerator, denominator, learned_weights, cutoff: float = 0.9):
    """The iterative weights and whether the learned ones keep FSC >= 0.99
    with them up to `cutoff` of Nyquist (the reconstruction numerator * W)."""
    from .precision import shell_report
    iterative = iterative_weights(denominator)
    report = shell_report((numerator * learned_weights).contiguous(), (numerator * iterative).contiguous(), cutoff)
    return iterative, report["min_fsc_to_cutoff"] >= SUPPORT_FSC, report["min_fsc_to_cutoff"]


def compensate(numerator, denominator, model: CompensationModel | None, table: torch.Tensor | None,
               cutoff: float = 0.9):
    """Weights for one reconstruction, judged: the model's when its result
    keeps FSC >= 0.99 with the iterative result up to `cutoff` of Nyquist,
    else the iterative weights. This always computes the iterative estimate;
    ``Compensator`` judges once for reconstructions that share a distribution.
    Returns (weights, report)."""
    if model is None or table is None:
        return iterative_weights(denominator), {"used": "iterative", "reason": "no model"}
    learned = model.weights(denominator, table)
    iterative, ok, fsc = judge(numerator, denominator, learned, cutoff)
    return (learned if ok else iterative), {"used": "learned" if ok else "iterative", "min_fsc": fsc}


class Compensator:
    """Weights for the reconstructions of one set of orientations (the two
    half-maps of a data set, the iterations of a refinement): the first call
    judges the model against the iterative estimate, later calls reuse the
    decision and skip the iterative estimate when the model was trusted.
    ``judge_every`` > 0 judges again every so many calls."""

    def __init__(self, model: CompensationModel | None, transforms: torch.Tensor, big: int, cutoff: float = 0.9,
                 judge_every: int = 0):
        self.model = model
        self.table = coverage_table(transforms, big) if model is not None else None
        self.cutoff = cutoff
        self.judge_every = judge_every
        self.calls = 0
        self.trusted = None
        self.min_fsc = None

    def __call__(self, numerator, denominator):
        self.calls += 1
        if self.model is None:
            return iterative_weights(denominator), {"used": "iterative", "reason": "no model"}
        due = self.trusted is None or (self.judge_every > 0 and (self.calls - 1) % self.judge_every == 0)
        if not due:
            if self.trusted:
                return self.model.weights(denominator, self.table), {"used": "learned", "judged": False,
                                                                     "min_fsc": self.min_fsc}
            return iterative_weights(denominator), {"used": "iterative", "judged": False, "min_fsc": self.min_fsc}
        learned = self.model.weights(denominator, self.table)
        iterative, self.trusted, self.min_fsc = judge(numerator, denominator, learned, self.cutoff)
        return (learned if self.trusted else iterative), {"used": "learned" if self.trusted else "iterative",
                                                          "judged": True, "min_fsc": self.min_fsc}
