"""rexlib's elementwise functions the reconstruction's quotient is written
with, on gfx950: ``full``/``fill_`` (ops::fill_operation), ``divide``,
``greater`` and ``where`` (include/rexlib/ops/arithmetic/divide_operation.hpp,
ops/compare/greater_operation.hpp, ops/selection/where_operation.hpp).
Operands broadcast against the result as numpy's rules say; the result's
shape is the broadcast shape.

With them the test recipe's quotient
(tests/integration/src/em/functional/test_cpu_reconstruction.cpp:457-464)
is dispatchable: ``where(greater(den, 1e-3), divide(num, den_c), 0)``.
"""
from __future__ import annotations

import torch

from . import _native as N
from .fourier import _stream


def _scalar(value, like: torch.Tensor) -> torch.Tensor:
    """A 0-d device tensor of `like`'s type holding `value`."""
    if isinstance(value, torch.Tensor):
        return value
    return full((), value, dtype=like.dtype, device=like.device)


def _shape(*tensors) -> tuple:
    return tuple(torch.broadcast_shapes(*(t.shape for t in tensors)))


def fill_(out: torch.Tensor, value, stream=None) -> torch.Tensor:
    """Set every element of `out` to `value` (complex for complex types)."""
    v = complex(value)
    N.check(N.library().rexlib_hip_fill(N.array(out), v.real, v.imag, _stream(stream)))
    return out


def full(shape, value, dtype=torch.float32, device="cuda", stream=None) -> torch.Tensor:
    return fill_(torch.empty(shape, dtype=dtype, device=device), value, stream)


def divide(dividend: torch.Tensor, divisor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    divisor = _scalar(divisor, dividend)
    if out is None:
        out = torch.empty(_shape(dividend, divisor), dtype=dividend.dtype, device=dividend.device)
    N.check(N.library().rexlib_hip_divide(N.array(dividend), N.array(divisor), N.array(out), _stream(stream)))
    return out


def greater(a: torch.Tensor, b, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    b = _scalar(b, a)
    if out is None:
        out = torch.empty(_shape(a, b), dtype=torch.bool, device=a.device)
    N.check(N.library().rexlib_hip_greater(N.array(a), N.array(b), N.array(out), _stream(stream)))
    return out


def where(condition: torch.Tensor, x, y, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    like = x if isinstance(x, torch.Tensor) else y
    x = _scalar(x, like)
    y = _scalar(y, like)
    if out is None:
        out = torch.empty(_shape(condition, x, y), dtype=like.dtype, device=like.device)
    N.check(N.library().rexlib_hip_where(N.array(condition), N.array(x), N.array(y), N.array(out),
                                         _stream(stream)))
    return out
