"""rexlib_amd — the MI355X (gfx950) backend of rexlib's Fourier slice insertion path.

The compute lives in librexlib_hip.so (C ABI: include/rexlib_hip.h), built by
rexlib_amd/build.py. This package is the host-side mirror of rexlib's Python
API for the path: ``rexlib_amd.em.insert_*_slices``, ``rexlib_amd.rfft2``,
``rexlib_amd.irfft3`` and the reconstruction steps.
"""
from . import em, fourier, functional
from ._native import UnsupportedOperationError, library
from .fourier import irfft2, irfft3, rfft2
from .functional import divide, fill_, full, greater, where

__all__ = ["em", "fourier", "functional", "rfft2", "irfft2", "irfft3", "divide", "fill_", "full", "greater", "where",
           "library", "UnsupportedOperationError"]
