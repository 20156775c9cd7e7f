"""Synthetic cryo-EM particle stacks for the insertion workloads.

A ground-truth volume is a sum of isotropic Gaussian blobs. A 3-D Gaussian
projects exactly onto a 2-D Gaussian of the same sigma with amplitude
a * sqrt(2 pi) * sigma at the rotated centre, so the projections carry no
projector error. Each image is separable in x and y per blob, which makes a
stack one batched matrix product: image = Gy^T diag(w) Gx.

Geometry follows the insertion contract
(include/rexlib/em/ops/reconstruction/insert_slices_operation.hpp:60-90): the
transform [A | t] maps the volume point p (voxels from the origin N // 2) to
the image point A p + t (pixels from n // 2), and the image is the
projection along the third coordinate. Orientations are uniform on SO(3)
from normalised Gaussian quaternions (the construction of
tests/integration/src/em/functional/test_cpu_reconstruction.cpp:299-316), or
von Mises-Fisher around the top view. The CTF is the contract's formula
(insert_ctf_slices_operation.hpp:43-58), applied in Fourier space; the noise is
white with variance var(signal) / SNR per image.

Parameters come from numpy's default_rng(seed); the images are computed with
torch on any device (float32 on the GPU), so a stack made on the GPU can be
copied to the host for the CPU oracle.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class BlobVolume:
    centres: np.ndarray   # (B, 3) offsets (x, y, z) from the origin, voxels
    sigmas: np.ndarray    # (B,)
    amplitudes: np.ndarray  # (B,)
    extent: int

    def render(self, device="cpu", dtype=torch.float32) -> torch.Tensor:
        """The volume (z, y, x), origin at extent // 2."""
        n = self.extent
        r = torch.arange(n, device=device, dtype=torch.float64) - n // 2
        out = torch.zeros((n, n, n), device=device, dtype=torch.float64)
        for c, s, a in zip(self.centres, self.sigmas, self.amplitudes):
            gx = torch.exp(-(r - c[0]) ** 2 / (2 * s * s))
            gy = torch.exp(-(r - c[1]) ** 2 / (2 * s * s))
            gz = torch.exp(-(r - c[2]) ** 2 / (2 * s * s))
            out += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        return out.to(dtype)


def blob_volume(extent: int, count: int, radius: float, sigma_range, seed: int) -> BlobVolume:
    rng = np.random.default_rng(seed)
    directions = rng.normal(size=(count, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    radii = radius * rng.uniform(0, 1, count) ** (1.0 / 3.0)
    centres = directions * radii[:, None]
    sigmas = rng.uniform(sigma_range[0], sigma_range[1], count)
    amplitudes = rng.uniform(0.5, 1.5, count)
    return BlobVolume(centres, sigmas, amplitudes, extent)


def quaternion_rotations(q: np.ndarray) -> np.ndarray:
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], axis=1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], axis=1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], axis=1),
    ], axis=1)


def uniform_rotations(count: int, rng: np.random.Generator) -> np.ndarray:
    return quaternion_rotations(rng.normal(size=(count, 4)))


def von_mises_fisher_rotations(count: int, kappa: float, rng: np.random.Generator) -> np.ndarray:
    """Rotations whose viewing direction (third row) follows a vMF law of
    concentration kappa around +z, with a uniform in-plane angle."""
    u = rng.uniform(0, 1, count)
    w = 1.0 + np.log(u + (1.0 - u) * np.exp(-2.0 * kappa)) / kappa  # cos of the polar angle
    phi = rng.uniform(0, 2 * math.pi, count)
    s = np.sqrt(np.clip(1 - w * w, 0, 1))
    view = np.stack([s * np.cos(phi), s * np.sin(phi), w], axis=1)
    psi = rng.uniform(0, 2 * math.pi, count)
    # An orthonormal frame with third row `view`, rotated in-plane by psi.
    helper = np.where(np.abs(view[:, 2:3]) < 0.9, np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]]))
    e1 = np.cross(helper, view)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(view, e1)
    r0 = np.cos(psi)[:, None] * e1 + np.sin(psi)[:, None] * e2
    r1 = -np.sin(psi)[:, None] * e1 + np.cos(psi)[:, None] * e2
    return np.stack([r0, r1, view], axis=1)


def transforms_of(rotations: np.ndarray) -> np.ndarray:
    t = np.zeros((rotations.shape[0], 3, 4))
    t[:, :, :3] = rotations
    return t


def ctf_parameters(count: int, rng: np.random.Generator, defocus=(10000.0, 30000.0),
                   astigmatism=1000.0) -> np.ndarray:
    """[D_u, D_v, angle, phase] per image: mean defocus uniform in `defocus` (A),
    astigmatism uniform up to `astigmatism` (A) at a uniform angle."""
    mean = rng.uniform(defocus[0], defocus[1], count)
    delta = rng.uniform(0, astigmatism, count)
    angle = rng.uniform(0, math.pi, count)
    return np.stack([mean + delta / 2, mean - delta / 2, angle, np.zeros(count)], axis=1)


def ctf_image(params: torch.Tensor, n: int, pixel_size: float, voltage=300.0, cs=2.7, q=0.1) -> torch.Tensor:
    """The CTF on the rfft2 grid (B, n, n//2+1), in float64, by the contract's formula."""
    volts = voltage * 1e3
    lam = 12.2643247 / math.sqrt(volts * (1.0 + 0.978466e-6 * volts))
    device = params.device
    ky = torch.fft.fftfreq(n, device=device, dtype=torch.float64) * n
    kx = torch.arange(n // 2 + 1, device=device, dtype=torch.float64)
    kx, ky = torch.meshgrid(kx, ky, indexing="xy")
    s2 = (kx * kx + ky * ky) / (n * pixel_size) ** 2
    theta = torch.atan2(ky, kx)
    p = params.to(torch.float64)
    du, dv, a, phase = (p[:, i][:, None, None] for i in range(4))
    defocus = 0.5 * (du + dv) + 0.5 * (du - dv) * torch.cos(2 * (theta[None] - a))
    chi = (math.pi * lam * defocus * s2[None] - 0.5 * math.pi * cs * 1e7 * lam ** 3 * s2[None] ** 2
           + phase + math.asin(q))
    return torch.sin(chi)


@dataclass
class ParticleStack:
    images: torch.Tensor      # (B, n, n) float32
    transforms: np.ndarray    # (B, 3, 4)
    ctf: np.ndarray           # (B, 4)
    volume: BlobVolume


def project(volume: BlobVolume, rotations: np.ndarray, n: int, device, chunk: int = 256) -> torch.Tensor:
    """Exact projections of the blob volume, (B, n, n) float32."""
    count = rotations.shape[0]
    out = torch.empty((count, n, n), dtype=torch.float32, device=device)
    r = torch.arange(n, device=device, dtype=torch.float32) - n // 2
    centres = torch.as_tensor(volume.centres, dtype=torch.float64, device=device)
    sig = torch.as_tensor(volume.sigmas, dtype=torch.float32, device=device)
    weight = torch.as_tensor(volume.amplitudes * math.sqrt(2 * math.pi) * volume.sigmas, dtype=torch.float32,
                             device=device)
    for s in range(0, count, chunk):
        rot = torch.as_tensor(rotations[s:s + chunk, :2, :], dtype=torch.float64, device=device)
        c2 = torch.matmul(centres[None], rot.transpose(1, 2).contiguous()).to(torch.float32)  # (b, blobs, 2)
        gx = torch.exp(-(r[None, None, :] - c2[..., 0:1]) ** 2 / (2 * sig[None, :, None] ** 2))
        gy = torch.exp(-(r[None, None, :] - c2[..., 1:2]) ** 2 / (2 * sig[None, :, None] ** 2))
        gx = gx * weight[None, :, None]
        out[s:s + chunk] = torch.matmul(gy.transpose(1, 2).contiguous(), gx.contiguous())
    return out


def particle_stack(volume: BlobVolume, count: int, n: int, seed: int, snr: float = 0.1, pixel_size: float = 1.5,
                   defocus=(10000.0, 30000.0), astigmatism=1000.0, rotations: np.ndarray | None = None,
                   device="cpu", chunk: int = 256) -> ParticleStack:
    rng = np.random.default_rng(seed)
    if rotations is None:
        rotations = uniform_rotations(count, rng)
    params = ctf_parameters(count, rng, defocus, astigmatism)
    images = project(volume, rotations, n, device, chunk)
    noise_seed = int(rng.integers(0, 2 ** 31 - 1))
    gen = torch.Generator(device=device)
    gen.manual_seed(noise_seed)
    for s in range(0, count, chunk):
        block = images[s:s + chunk]
        spec = torch.fft.rfft2(block.to(torch.float64))
        c = ctf_image(torch.as_tensor(params[s:s + chunk], device=device), n, pixel_size)
        signal = torch.fft.irfft2(spec * c, s=(n, n))
        var = signal.var(dim=(1, 2), keepdim=True)
        noise = torch.randn(signal.shape, generator=gen, device=device, dtype=torch.float64)
        images[s:s + chunk] = (signal + noise * torch.sqrt(var / snr)).to(torch.float32)
    return ParticleStack(images, transforms_of(rotations), params, volume)
