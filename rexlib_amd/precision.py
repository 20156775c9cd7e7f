"""Reduced-precision staging of particle images, and per-shell error reports.

Images are held in HBM as bfloat16, or as float16 scaled per image by its
largest magnitude (the scale, a float32 per image, multiplies the image back
when it is read); the transform, the CTF weighting and the accumulation run
in float32 (``rexlib_amd.em.insert_ctf_images``). The per-shell figures are
computed from float64 shell sums on the device (``em.shell_sums``).
"""
from __future__ import annotations

import torch

from . import em

# float16 holds magnitudes below 65504; images are scaled to at most this.
_FP16_TARGET = 32768.0


def stage_images(images: torch.Tensor, dtype: str):
    """(staged images, per-image float32 scales or None) for dtype 'bf16' or 'fp16'."""
    if dtype == "bf16":
        return images.to(torch.bfloat16), None
    if dtype == "fp16":
        peak = images.abs().amax(dim=(-2, -1)).clamp_min(1e-30)
        scales = (peak / _FP16_TARGET).to(torch.float32)
        staged = (images / scales[..., None, None]).to(torch.float16)
        return staged, scales
    if dtype == "f32":
        return images.to(torch.float32), None
    raise ValueError(f"unknown staging type {dtype}")


def shell_report(test: torch.Tensor, reference: torch.Tensor, cutoff: float = 0.9):
    """Per-shell FSC and relative error of two half spectra (N, N, N/2+1).

    Returns a dict with the shell index, the FSC, the relative error
    sqrt(sum |a-b|^2 / sum |b|^2) per shell, and their worst values over the
    shells up to `cutoff` of Nyquist (shell 0 excluded from the minimum FSC if
    it holds a single element), and the relative error of the whole band of
    shells 1 to that cutoff."""
    big = test.shape[-3]
    s = em.shell_sums(test, reference)
    cross, aa, bb, count = s[0], s[1], s[2], s[3]
    fsc = cross / torch.sqrt(torch.clamp(aa * bb, min=1e-300))
    # |a - b|^2 = |a|^2 + |b|^2 - 2 Re(a conj b)
    err = torch.sqrt(torch.clamp(aa + bb - 2 * cross, min=0.0) / torch.clamp(bb, min=1e-300))
    limit = int(cutoff * (big // 2))
    band = torch.clamp((aa + bb - 2 * cross)[1:limit + 1].sum(), min=0.0) / torch.clamp(bb[1:limit + 1].sum(),
                                                                                          min=1e-300)
    fsc_h = fsc.cpu()
    err_h = err.cpu()
    return {
        "shells": int(big // 2 + 1),
        "fsc": fsc_h.tolist(),
        "relative_error": err_h.tolist(),
        "min_fsc_to_cutoff": float(fsc_h[1:limit + 1].min()),
        "max_error_to_cutoff": float(err_h[1:limit + 1].max()),
        "cutoff_shell": limit,
        "relative_error_to_cutoff": float(torch.sqrt(band)),
        "count": count.cpu().tolist(),
    }
