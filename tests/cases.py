"""Insertion cases shaped like the reference's fixture
(tests/unitary/src/backends/cpu/builders/em/reconstruction/fixtures/slice_insertion_cases.hpp:137-200,
594-625): uniform(-1, 1) complex accumulators, slices and transforms,
uniform(0, 2) denominators, radial weights and spectrum weights, CTF
defocus uniform(8000, 25000) A, angle uniform(0, pi), phase uniform(0, 0.5),
optics (300 kV, 2.7 mm, 0.1), pixel size 1.5 A."""
from __future__ import annotations

import math

import numpy as np

PIXEL_SIZE = 1.5
SIZES = [(6, 12), (5, 7), (6, 9), (5, 10)]


def make_case(variant: str, n: int, big: int, seed: int, slices: int = 3, accumulator_batch=(),
              transform_batch=None, zero_accumulators=False):
    """Operands in the operation's order, as numpy float64 / complex128."""
    rng = np.random.default_rng(seed)
    h = n // 2 + 1
    half = big // 2 + 1
    acc_shape = tuple(accumulator_batch) + (big, big, half)
    if zero_accumulators:
        num = np.zeros(acc_shape, np.complex128)
        den = np.zeros(acc_shape, np.float64)
    else:
        num = rng.uniform(-1, 1, acc_shape) + 1j * rng.uniform(-1, 1, acc_shape)
        den = rng.uniform(0, 2, acc_shape)
    sl = rng.uniform(-1, 1, (slices, n, h)) + 1j * rng.uniform(-1, 1, (slices, n, h))
    tshape = (slices, 3, 4) if transform_batch is None else tuple(transform_batch) + (3, 4)
    tr = rng.uniform(-1, 1, tshape)
    ra = rng.uniform(0, 2, (slices, h))
    if variant == "plain":
        return [num, den, sl, tr, ra]
    if variant == "weighted":
        wt = rng.uniform(0, 2, (slices, n, h))
        return [num, den, sl, wt, tr, ra]
    ctf = np.stack([rng.uniform(8000, 25000, slices), rng.uniform(8000, 25000, slices),
                    rng.uniform(0, math.pi, slices), rng.uniform(0, 0.5, slices)], axis=1)
    return [num, den, sl, tr, ctf, ra]


def identity_transform():
    return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], dtype=np.float64)


def spread_kat_case():
    """One sample landing at (1.25, 2.5, -0.75) with value 2 in an 8^3 volume
    (tests/unitary/src/backends/cpu/builders/em/reconstruction/test_row_insertion_kernel.cpp:47-112),
    posed as a weighted insertion: n = 3 so that only k = (1, 0) carries a
    weight, its shift phase is 1 (t_x = -1), and the -k copy lands at x < 0."""
    n, big = 3, 8
    h = n // 2 + 1
    scale = big / n
    num = np.zeros((big, big, big // 2 + 1), np.complex128)
    den = np.zeros_like(num, dtype=np.float64)
    sl = np.zeros((1, n, h), np.complex128)
    sl[0, 0, 1] = 2.0
    wt = np.zeros((1, n, h))
    wt[0, 0, 1] = 1.0
    tr = np.zeros((1, 3, 4))
    tr[0, 0, :3] = np.array([1.25, 2.5, -0.75]) / scale
    tr[0, 0, 3] = -1.0
    tr[0, 1, :3] = np.array([0.0, 0.0, 1.0]) / scale
    tr[0, 2, :3] = np.array([0.0, 1.0, 0.0])
    ra = np.array([[0.0, 1.0]])
    return [num, den, sl, wt, tr, ra]


def kat_index(z, y, x, big=8):
    return (z, y, x)
