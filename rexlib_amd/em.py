"""rexlib.em on gfx950: Fourier slice insertion and the reconstruction steps.

The three insertions mirror rexlib's Python API
(python/rexlib/em/_functional.py:151-310) over device tensors: same names,
operand order and meaning, the accumulators returned as a pair, and
``numerator_out`` / ``denominator_out`` to update in place. Invalid operands
raise ``ValueError`` (std::invalid_argument); operands no gfx950 program takes
raise ``UnsupportedOperationError``. The ``context`` argument selects the HIP
stream: None means torch's current stream.

Everything here calls librexlib_hip.so; there is no other implementation.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import _native as N
from ._native import UnsupportedOperationError

__all__ = [
    "GriddingWindow",
    "CtfOptics",
    "ExecutionContext",
    "UnsupportedOperationError",
    "insert_slices",
    "insert_weighted_slices",
    "insert_ctf_slices",
    "insert_ctf_images",
    "normalize",
    "gridding_correct",
    "standardize_background",
    "shell_sums",
    "fsc",
]


@dataclass(frozen=True)
class GriddingWindow:
    """A separable gridding window (include/rexlib/em/ops/parameters/gridding_window.hpp:23-90)."""

    profile: int
    half_width: float
    shape: float = 0.0

    @staticmethod
    def nearest() -> "GriddingWindow":
        return GriddingWindow(0, 0.5, 0.0)

    @staticmethod
    def linear() -> "GriddingWindow":
        return GriddingWindow(1, 1.0, 0.0)

    @staticmethod
    def kaiser_bessel(half_width: float, shape: float) -> "GriddingWindow":
        # Checks of src/em/ops/parameters/gridding_window.cpp:24-42.
        if not (math.isfinite(half_width) and half_width > 0.0):
            raise ValueError("gridding_window: the half-width of a Kaiser-Bessel window must be finite and "
                             f"positive, but it is {half_width}.")
        if not (math.isfinite(shape) and shape >= 0.0):
            raise ValueError("gridding_window: the shape of a Kaiser-Bessel window must be finite and 0 or more, "
                             f"but it is {shape}.")
        return GriddingWindow(2, float(half_width), float(shape))

    def native(self) -> N.Window:
        return N.Window(self.profile, self.half_width, self.shape)


@dataclass(frozen=True)
class CtfOptics:
    """Microscope optics (include/rexlib/em/ops/parameters/ctf_optics.hpp; checks of
    src/em/ops/parameters/ctf_optics.cpp:15-56)."""

    voltage: float
    spherical_aberration: float
    amplitude_contrast: float

    def __post_init__(self):
        if not (math.isfinite(self.voltage) and self.voltage > 0.0):
            raise ValueError(f"ctf_optics: the voltage must be finite and positive, but it is {self.voltage}.")
        if not math.isfinite(self.spherical_aberration):
            raise ValueError("ctf_optics: the spherical aberration must be finite, but it is "
                             f"{self.spherical_aberration}.")
        if not (0.0 <= self.amplitude_contrast < 1.0):
            raise ValueError("ctf_optics: the amplitude contrast must lie in [0, 1), but it is "
                             f"{self.amplitude_contrast}.")

    @property
    def wavelength(self) -> float:
        """12.2643247 / sqrt(V (1 + 0.978466e-6 V)), V in volts (ctf_optics.cpp:81-85)."""
        volts = self.voltage * 1e3
        return 12.2643247 / math.sqrt(volts * (1.0 + 0.978466e-6 * volts))

    def native(self) -> N.Optics:
        return N.Optics(self.voltage, self.spherical_aberration, self.amplitude_contrast)


@dataclass(frozen=True)
class ExecutionContext:
    """Where a call runs: a device and a HIP stream (None: torch's current stream)."""

    device: torch.device | None = None
    stream: torch.cuda.Stream | None = None

    def handle(self):
        import ctypes
        if self.stream is not None:
            return ctypes.c_void_p(self.stream.cuda_stream)
        return N.stream_handle(self.device)


def _context(context: ExecutionContext | None) -> ExecutionContext:
    return context if context is not None else ExecutionContext()


def _outputs(numerator, denominator, numerator_out, denominator_out):
    if numerator_out is None:
        numerator_out = torch.empty_like(numerator)
    if denominator_out is None:
        denominator_out = torch.empty_like(denominator)
    return numerator_out, denominator_out


def _device_of(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise UnsupportedOperationError(
                "the gfx950 insertion programs take device memory only (the CPU program takes host memory)")


def insert_slices(numerator, denominator, slices, transforms, radial_weights, window: GriddingWindow,
                  context: ExecutionContext | None = None, numerator_out=None, denominator_out=None):
    """Insert the spectra of projections into the spectrum of a volume.

    rexlib.em.insert_slices (python/rexlib/em/_functional.py:151-205;
    contract include/rexlib/em/ops/reconstruction/insert_slices_operation.hpp:31-112).
    """
    _device_of(numerator, denominator, slices, transforms, radial_weights)
    ctx = _context(context)
    num_out, den_out = _outputs(numerator, denominator, numerator_out, denominator_out)
    N.check(N.library().rexlib_hip_insert_slices(
        N.array(numerator), N.array(denominator), N.array(slices), N.array(transforms), N.array(radial_weights),
        window.native(), N.array(num_out), N.array(den_out), ctx.handle()))
    return num_out, den_out


def insert_weighted_slices(numerator, denominator, slices, weights, transforms, radial_weights,
                           window: GriddingWindow, context: ExecutionContext | None = None, numerator_out=None,
                           denominator_out=None):
    """rexlib.em.insert_weighted_slices (python/rexlib/em/_functional.py:207-254)."""
    _device_of(numerator, denominator, slices, weights, transforms, radial_weights)
    ctx = _context(context)
    num_out, den_out = _outputs(numerator, denominator, numerator_out, denominator_out)
    N.check(N.library().rexlib_hip_insert_weighted_slices(
        N.array(numerator), N.array(denominator), N.array(slices), N.array(weights), N.array(transforms),
        N.array(radial_weights), window.native(), N.array(num_out), N.array(den_out), ctx.handle()))
    return num_out, den_out


def insert_ctf_slices(numerator, denominator, slices, transforms, ctf, radial_weights, window: GriddingWindow,
                      optics: CtfOptics, pixel_size: float, context: ExecutionContext | None = None,
                      numerator_out=None, denominator_out=None):
    """rexlib.em.insert_ctf_slices (python/rexlib/em/_functional.py:256-310;
    contract insert_ctf_slices_operation.hpp:33-65)."""
    if not (math.isfinite(pixel_size) and pixel_size > 0.0):
        raise ValueError(f"insert_ctf_slices: the pixel size must be finite and positive, but it is {pixel_size}.")
    _device_of(numerator, denominator, slices, transforms, ctf, radial_weights)
    ctx = _context(context)
    num_out, den_out = _outputs(numerator, denominator, numerator_out, denominator_out)
    N.check(N.library().rexlib_hip_insert_ctf_slices(
        N.array(numerator), N.array(denominator), N.array(slices), N.array(transforms), N.array(ctf),
        N.array(radial_weights), window.native(), optics.native(), float(pixel_size), N.array(num_out),
        N.array(den_out), ctx.handle()))
    return num_out, den_out


def insert_ctf_images(numerator, denominator, images, transforms, ctf, radial_weights, window: GriddingWindow,
                      optics: CtfOptics, pixel_size: float, image_scales=None, padded: int = 0,
                      context: ExecutionContext | None = None):
    """rfft2 of real images (float32, bfloat16, or float16 with per-image
    scales), zero-padded on chip to ``padded``, then insert_ctf_slices into the
    accumulators in place. Arithmetic is float32 throughout."""
    if not (math.isfinite(pixel_size) and pixel_size > 0.0):
        raise ValueError(f"insert_ctf_slices: the pixel size must be finite and positive, but it is {pixel_size}.")
    _device_of(numerator, denominator, images, transforms, ctf, radial_weights)
    ctx = _context(context)
    N.check(N.library().rexlib_hip_insert_ctf_images(
        N.array(numerator), N.array(denominator), N.array(images), N.array(image_scales), int(padded),
        N.array(transforms), N.array(ctf), N.array(radial_weights), window.native(), optics.native(),
        float(pixel_size), ctx.handle()))
    return numerator, denominator


def normalize(numerator, denominator, threshold: float = 1e-3, out=None, context: ExecutionContext | None = None):
    """numerator / denominator where the denominator exceeds ``threshold``, else 0
    (the quotient of tests/integration/src/em/functional/test_cpu_reconstruction.cpp:457-464)."""
    ctx = _context(context)
    if out is None:
        out = torch.empty_like(numerator)
    N.check(N.library().rexlib_hip_normalize(N.array(numerator), N.array(denominator), float(threshold),
                                             N.array(out), ctx.handle()))
    return out


def gridding_correct(volume, window: GriddingWindow, out=None, crop: int | None = None,
                     context: ExecutionContext | None = None):
    """Divide a real volume (origin at N // 2) by the window's transform per axis.

    With ``crop`` = M < N (or an ``out`` of extent M), only the central M^3
    (from N//2 - M//2) is written: the crop of a reconstruction inserted with
    padding."""
    ctx = _context(context)
    if out is None:
        m = volume.shape[-1] if crop is None else int(crop)
        out = torch.empty(tuple(volume.shape[:-3]) + (m, m, m), dtype=volume.dtype, device=volume.device)
    N.check(N.library().rexlib_hip_gridding_correct(N.array(volume), window.native(), N.array(out),
                                                    ctx.handle()))
    return out


def standardize_background(value, key, radius: float, lower_bound: float, upper_bound: float, axes=(-2, -1),
                           out=None, context: ExecutionContext | None = None):
    """Standardize images against the statistics of their background
    (include/rexlib/em/ops/preprocessing/standardize_background_operation.hpp).

    An image is ``value``'s part along ``axes``; the other axes are batches.
    The background of an image is the elements whose squared distance from
    the centre (index n // 2 per axis) exceeds ``radius`` squared. Mean and
    population deviation of its finite elements are taken twice, the second
    time over the values that the first statistics standardize within
    [lower_bound, upper_bound]; the result is (value - mean) / deviation, and
    elements not finite or outside the bounds take standard normal draws of
    the Philox4x32-10 stream of their ``key`` (uint64, or int64 read as its
    bits, broadcasting against the batch axes)."""
    ctx = _context(context)
    rank = value.dim()
    mask = 0
    for a in axes:
        a = a + rank if a < 0 else a
        if not 0 <= a < rank or mask & (1 << a):
            raise ValueError("standardize_background: axes must be distinct axes of the value")
        mask |= 1 << a
    if key.dtype == torch.int64:
        key = key.view(torch.uint64)
    if out is None:
        out = torch.empty_like(value)
    N.check(N.library().rexlib_hip_standardize_background(N.array(value), N.array(key), mask, float(radius),
                                                          float(lower_bound), float(upper_bound), N.array(out),
                                                          ctx.handle()))
    return out


def shell_sums(a, b, context: ExecutionContext | None = None):
    """Per-shell float64 sums of two half spectra: (cross, |a|^2, |b|^2, count), each (N//2+1,)."""
    ctx = _context(context)
    big = a.shape[-3]
    shells = big // 2 + 1
    sums = torch.zeros(4 * shells, dtype=torch.float64, device=a.device)
    N.check(N.library().rexlib_hip_shell_sums(N.array(a), N.array(b), __import__("ctypes").c_void_p(sums.data_ptr()),
                                              ctx.handle()))
    return sums.view(4, shells)


def fsc(a, b, context: ExecutionContext | None = None):
    """Fourier shell correlation of two half spectra, from float64 shell sums."""
    s = shell_sums(a, b, context)
    return s[0] / torch.sqrt(torch.clamp(s[1] * s[2], min=1e-300))
