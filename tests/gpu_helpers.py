"""Helpers of the GPU tests: run the HIP insertion on numpy operands."""
from __future__ import annotations

import numpy as np
import torch

import oracle
from rexlib_amd import em

DEVICE = "cuda"


def to_device(a: np.ndarray) -> torch.Tensor:
    if np.iscomplexobj(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(DEVICE)
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEVICE)


def hip_window(window: oracle.Window) -> em.GriddingWindow:
    if window.profile == "box":
        return em.GriddingWindow.nearest()
    if window.profile == "triangle":
        return em.GriddingWindow.linear()
    return em.GriddingWindow.kaiser_bessel(window.half_width, window.shape)


OPTICS = em.CtfOptics(300.0, 2.7, 0.1)


def hip_insert(variant: str, inputs, window: oracle.Window, pixel_size: float = 1.5, in_place: bool = True,
               tensors=None):
    """Insert with the HIP library; returns host numpy (complex64, float32)."""
    t = tensors if tensors is not None else [to_device(a) for a in inputs]
    w = hip_window(window)
    out = {}
    if in_place:
        out = dict(numerator_out=t[0], denominator_out=t[1])
    if variant == "plain":
        num, den = em.insert_slices(*t, w, **out)
    elif variant == "weighted":
        num, den = em.insert_weighted_slices(*t, w, **out)
    else:
        num, den = em.insert_ctf_slices(*t[:6], w, OPTICS, pixel_size, **out)
    torch.cuda.synchronize()
    return num.cpu().numpy(), den.cpu().numpy()


def max_abs(a, b) -> float:
    return float(np.max(np.abs(np.asarray(a, np.complex128) - np.asarray(b, np.complex128)))) if np.size(a) else 0.0
