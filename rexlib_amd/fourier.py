"""rexlib's rfft2 / irfft2 / irfft3 on gfx950 (include/rexlib/functional/fourier.hpp:483,723).

Unscaled forward transforms and 1/n-per-axis inverse transforms, as pocketfft
computes them for rexlib (src/backends/cpu/kernels/pocketfft_fourier_transform.cpp:119-163,
src/ops/fourier/fourier_normalization.cpp:56-71).
"""
from __future__ import annotations

import torch

from . import _native as N


def _stream(stream):
    import ctypes
    if stream is not None:
        return ctypes.c_void_p(stream.cuda_stream)
    return N.stream_handle()


def rfft2(images: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Real 2-D DFT over the last two axes: (..., n, n) float32 -> (..., n, n//2+1) complex64."""
    n = images.shape[-1]
    if out is None:
        out = torch.empty(images.shape[:-1] + (n // 2 + 1,), dtype=torch.complex64, device=images.device)
    N.check(N.library().rexlib_hip_rfft2(N.array(images), N.array(out), _stream(stream)))
    return out


def irfft2(spectra: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Inverse of rfft2, scaled by 1/n^2: (..., n, n//2+1) -> (..., n, n)."""
    n = spectra.shape[-2]
    if out is None:
        out = torch.empty(spectra.shape[:-1] + (n,), dtype=torch.float32, device=spectra.device)
    N.check(N.library().rexlib_hip_irfft2(N.array(spectra), N.array(out), _stream(stream)))
    return out


def irfft3(spectrum: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Inverse real 3-D DFT, scaled by 1/N^3: (..., N, N, N//2+1) -> (..., N, N, N)."""
    big = spectrum.shape[-3]
    if out is None:
        out = torch.empty(spectrum.shape[:-3] + (big, big, big), dtype=torch.float32, device=spectrum.device)
    N.check(N.library().rexlib_hip_irfft3(N.array(spectrum), N.array(out), _stream(stream)))
    return out


NORMALIZATIONS = {"backward": 0, "ortho": 1, "forward": 2}


def _axes(axes, rank):
    import ctypes
    axes = tuple(range(rank)) if axes is None else tuple(int(a) for a in axes)
    return (ctypes.c_int32 * len(axes))(*axes), len(axes), tuple(a % rank for a in axes)


def fftn(x: torch.Tensor, axes=None, norm: str = "backward", out: torch.Tensor | None = None,
         stream=None) -> torch.Tensor:
    """Complex DFT over `axes` (all by default), as numpy.fft.fftn."""
    arr, count, _ = _axes(axes, x.dim())
    if out is None:
        out = torch.empty(x.shape, dtype=torch.complex64, device=x.device)
    N.check(N.library().rexlib_hip_fft(N.array(x), N.array(out), arr, count, 0, NORMALIZATIONS[norm],
                                       _stream(stream)))
    return out


def ifftn(x: torch.Tensor, axes=None, norm: str = "backward", out: torch.Tensor | None = None,
          stream=None) -> torch.Tensor:
    """Inverse complex DFT over `axes`, as numpy.fft.ifftn."""
    arr, count, _ = _axes(axes, x.dim())
    if out is None:
        out = torch.empty(x.shape, dtype=torch.complex64, device=x.device)
    N.check(N.library().rexlib_hip_fft(N.array(x), N.array(out), arr, count, 1, NORMALIZATIONS[norm],
                                       _stream(stream)))
    return out


def rfftn(x: torch.Tensor, axes=None, norm: str = "backward", out: torch.Tensor | None = None,
          stream=None) -> torch.Tensor:
    """Real-to-complex DFT over `axes`, the last one halved, as numpy.fft.rfftn."""
    arr, count, normal = _axes(axes, x.dim())
    if out is None:
        shape = list(x.shape)
        shape[normal[-1]] = shape[normal[-1]] // 2 + 1
        out = torch.empty(shape, dtype=torch.complex64, device=x.device)
    N.check(N.library().rexlib_hip_rfft(N.array(x), N.array(out), arr, count, NORMALIZATIONS[norm],
                                        _stream(stream)))
    return out


def irfftn(x: torch.Tensor, s=None, axes=None, norm: str = "backward", out: torch.Tensor | None = None,
           stream=None) -> torch.Tensor:
    """Complex-to-real inverse DFT over `axes`, as numpy.fft.irfftn; `s` the
    output extents along `axes` (default: 2 (m - 1) along the last one)."""
    arr, count, normal = _axes(axes, x.dim())
    if out is None:
        shape = list(x.shape)
        if s is None:
            shape[normal[-1]] = 2 * (shape[normal[-1]] - 1)
        else:
            for a, e in zip(normal, s):
                shape[a] = int(e)
        out = torch.empty(shape, dtype=torch.float32, device=x.device)
    N.check(N.library().rexlib_hip_irfft(N.array(x), N.array(out), arr, count, NORMALIZATIONS[norm],
                                         _stream(stream)))
    return out
