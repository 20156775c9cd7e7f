"""TEST INFRASTRUCTURE — the oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker. The product path
(``rexlib_amd``) never imports it.

Contents:

* ``insert_f64``: the double-precision brute-force restatement of the
  insertion contract (``oracle/insertion_oracle.c``; reference
  ``include/rexlib/em/ops/reconstruction/insert_slices_operation.hpp:31-112``
  and the fixture oracle
  ``tests/unitary/src/backends/cpu/builders/em/reconstruction/fixtures/slice_insertion_cases.hpp:396-563``).
* ``ref_insert``: the reference's own CPU row kernel, compiled from its sources
  into ``oracle/_ref/librexlib_ref.so`` (float32 or float64, tiled and threaded
  like ``slice_insertion_functor.inl:105-156``).
* ``expand_batch``: the batch semantics of the insertion layout plan
  (``src/backends/cpu/plans/slice_insertion_layout_plan.cpp:81-185``):
  numpy broadcasting of the batch prefixes, the axes the accumulator lacks (or
  holds at extent 1) summed.
* FFTs through ``numpy.fft`` (pocketfft, the reference's FFT library; scale
  conventions of ``src/ops/fourier/fourier_normalization.cpp:56-71``).
* Closed forms for the steps the reference has no function for: the weight
  quotient (``tests/integration/src/em/functional/test_cpu_reconstruction.cpp:457-464``),
  gridding correction, per-shell sums and FSC.
* ``standardize_background``: the background standardization of
  ``include/rexlib/em/ops/preprocessing/standardize_background_operation.hpp:15-76``
  (``src/backends/cpu/builders/em/preprocessing/``), with its draws
  (``philox4x32_10``, ``standard_normal``: ``src/backends/cpu/kernels/philox.inl``,
  ``standard_normal.inl``), pinned by Random123's known answers and the
  reference kernel tests' own numbers.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

PROFILES = {"box": 0, "triangle": 1, "kaiser_bessel": 2}
VARIANTS = {"plain": 0, "weighted": 1, "ctf": 2}


@dataclass(frozen=True)
class Window:
    """A gridding window (include/rexlib/em/ops/parameters/gridding_window.hpp:23-90)."""

    profile: str
    half_width: float
    shape: float = 0.0

    @staticmethod
    def nearest() -> "Window":
        return Window("box", 0.5)

    @staticmethod
    def linear() -> "Window":
        return Window("triangle", 1.0)

    @staticmethod
    def kaiser_bessel(half_width: float, shape: float) -> "Window":
        return Window("kaiser_bessel", float(half_width), float(shape))


@dataclass(frozen=True)
class Optics:
    """CTF optics (include/rexlib/em/ops/parameters/ctf_optics.hpp)."""

    voltage: float = 300.0
    spherical_aberration: float = 2.7
    amplitude_contrast: float = 0.1


def _load(path):
    if not os.path.exists(path):
        return None
    return ctypes.CDLL(path)


_oracle = None
_ref = None


def _lib():
    global _oracle
    if _oracle is None:
        _oracle = _load(os.path.join(_HERE, "liboracle.so"))
        if _oracle is None:
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        d = ctypes.c_double
        i64 = ctypes.c_int64
        p = ctypes.c_void_p
        _oracle.oracle_insert_f64.argtypes = [ctypes.c_int, ctypes.c_int, d, d, i64, i64, i64, i64,
                                              p, p, p, p, p, p, d, d, d, d]
        _oracle.oracle_insert_f64.restype = None
        _oracle.oracle_bessel_i0.argtypes = [d]
        _oracle.oracle_bessel_i0.restype = d
        _oracle.oracle_wavelength.argtypes = [d]
        _oracle.oracle_wavelength.restype = d
        _oracle.oracle_window_weight.argtypes = [ctypes.c_int, d, d, d]
        _oracle.oracle_window_weight.restype = d
        _oracle.oracle_ctf.argtypes = [p, i64, d, d, d, d, d, d]
        _oracle.oracle_ctf.restype = d
    return _oracle


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "librexlib_ref.so"))


def _reflib():
    global _ref
    if _ref is None:
        _ref = _load(os.path.join(_HERE, "_ref", "librexlib_ref.so"))
        if _ref is None:
            raise RuntimeError("oracle/_ref/librexlib_ref.so is not built")
        d = ctypes.c_double
        i64 = ctypes.c_int64
        p = ctypes.c_void_p
        f = ctypes.c_float
        _ref.ref_insert.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, i64, i64, i64, i64,
                                    p, p, p, p, p, p, d, d, d, d, ctypes.c_int, i64]
        _ref.ref_insert.restype = ctypes.c_int
        _ref.ref_bessel_i0.argtypes = [d]
        _ref.ref_bessel_i0.restype = d
        _ref.ref_wavelength.argtypes = [d]
        _ref.ref_wavelength.restype = d
        _ref.ref_ctf_f32.argtypes = [d, d, d, d, i64, d, d, d, d, f, f]
        _ref.ref_ctf_f32.restype = f
        _ref.ref_kaiser_bessel_f32.argtypes = [d, d, f]
        _ref.ref_kaiser_bessel_f32.restype = f
        _ref.ref_axis_phase_f32.argtypes = [p, i64, i64, d]
        _ref.ref_axis_phase_f32.restype = None
        _ref.ref_choose_tile_extent.argtypes = [i64, i64, i64, i64, d]
        _ref.ref_choose_tile_extent.restype = i64
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# Scalar closed forms
# ---------------------------------------------------------------------------

def bessel_i0(x: float) -> float:
    return _lib().oracle_bessel_i0(float(x))


def wavelength(voltage_kv: float) -> float:
    return _lib().oracle_wavelength(float(voltage_kv))


def window_weight(window: Window, distance: float) -> float:
    return _lib().oracle_window_weight(PROFILES[window.profile], window.half_width, window.shape,
                                       float(distance))


def ctf(params, extent, pixel_size, optics: Optics, k_x, k_y) -> float:
    p = np.ascontiguousarray(params, dtype=np.float64)
    return _lib().oracle_ctf(_ptr(p), int(extent), float(pixel_size), optics.voltage,
                             optics.spherical_aberration, optics.amplitude_contrast,
                             float(k_x), float(k_y))


# ---------------------------------------------------------------------------
# Batch semantics of the insertion (slice_insertion_layout_plan.cpp:81-185)
# ---------------------------------------------------------------------------

CORE_RANKS = {
    "plain": (3, 3, 2, 2, 1),        # num, den, slices, transforms, radial
    "weighted": (3, 3, 2, 2, 2, 1),  # num, den, slices, weights, transforms, radial
    "ctf": (3, 3, 2, 2, 1, 1),       # num, den, slices, transforms, ctf, radial
}


def expand_batch(variant: str, inputs):
    """Broadcast the batch prefixes of the inputs and split kept from reduced.

    Returns (kept_shape, reduced_count, stacks) where stacks[i] is input i with
    its batch flattened to (kept, reduced, *core) for i >= 2 and
    (kept, *core) for the accumulators.
    """
    ranks = CORE_RANKS[variant]
    prefixes = [a.shape[: a.ndim - r] for a, r in zip(inputs, ranks)]
    frame = np.broadcast_shapes(*prefixes) if prefixes else ()
    acc_prefix = prefixes[0]
    padded = (1,) * (len(frame) - len(acc_prefix)) + tuple(acc_prefix)
    kept_axes = [ax for ax in range(len(frame)) if not (padded[ax] == 1 and frame[ax] > 1)]
    reduced_axes = [ax for ax in range(len(frame)) if padded[ax] == 1 and frame[ax] > 1]
    kept_shape = tuple(frame[ax] for ax in kept_axes)
    reduced_count = int(np.prod([frame[ax] for ax in reduced_axes], dtype=np.int64))
    order = kept_axes + reduced_axes
    stacks = []
    for i, (a, r) in enumerate(zip(inputs, ranks)):
        core = a.shape[a.ndim - r:]
        b = np.broadcast_to(a, tuple(frame) + tuple(core))
        b = np.transpose(b, order + list(range(len(frame), len(frame) + r)))
        kept_count = int(np.prod(kept_shape, dtype=np.int64))
        if i < 2:
            # Every reduced index of an accumulator is the same element.
            b = b.reshape((kept_count, reduced_count) + tuple(core))[:, 0]
        else:
            b = b.reshape((kept_count, reduced_count) + tuple(core))
        stacks.append(np.ascontiguousarray(b))
    return kept_shape, reduced_count, stacks


def _split(variant, inputs):
    if variant == "plain":
        num, den, slices, transforms, radial = inputs
        extra = None
    elif variant == "weighted":
        num, den, slices, extra, transforms, radial = inputs
    else:
        num, den, slices, transforms, extra, radial = inputs
    return num, den, slices, transforms, radial, extra


def _run_expanded(variant, inputs, fn):
    """Expand the batch, run fn over contiguous K x B stacks, restore shapes."""
    kept_shape, reduced, stacks = expand_batch(variant, inputs)
    num, den, slices, transforms, radial, extra = _split(variant, stacks)
    K = num.shape[0]
    out_num, out_den = fn(K, reduced, num, den, slices, transforms, radial, extra)
    # The kept axes are exactly the accumulator's own batch axes.
    return out_num.reshape(inputs[0].shape), out_den.reshape(inputs[1].shape)


def insert_f64(variant: str, inputs, window: Window, optics: Optics = Optics(), pixel_size: float = 1.0):
    """Brute-force FP64 insertion; ``inputs`` in the operation's operand order."""
    lib = _lib()

    def fn(K, B, num, den, slices, transforms, radial, extra):
        n = slices.shape[-2]
        big = num.shape[-3]
        num = np.ascontiguousarray(num, dtype=np.complex128).copy()
        den = np.ascontiguousarray(den, dtype=np.float64).copy()
        sl = np.ascontiguousarray(slices, dtype=np.complex128)
        tr = np.ascontiguousarray(transforms, dtype=np.float64)
        ra = np.ascontiguousarray(radial, dtype=np.float64)
        ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.float64)
        lib.oracle_insert_f64(VARIANTS[variant], PROFILES[window.profile], window.half_width,
                              window.shape, n, big, K, B, _ptr(num), _ptr(den), _ptr(sl), _ptr(tr),
                              _ptr(ra), _ptr(ex), optics.voltage, optics.spherical_aberration,
                              optics.amplitude_contrast, float(pixel_size))
        return num, den

    return _run_expanded(variant, inputs, fn)


def ref_insert(variant: str, inputs, window: Window, optics: Optics = Optics(), pixel_size: float = 1.0,
               dtype=np.float32, threads: int = 1, tile_budget: int = 1 << 20):
    """The reference's own row kernel, in float32 or float64."""
    lib = _reflib()
    real = np.dtype(dtype)
    cplx = np.complex64 if real == np.float32 else np.complex128

    def fn(K, B, num, den, slices, transforms, radial, extra):
        n = slices.shape[-2]
        big = num.shape[-3]
        num = np.ascontiguousarray(num, dtype=cplx).copy()
        den = np.ascontiguousarray(den, dtype=real).copy()
        sl = np.ascontiguousarray(slices, dtype=cplx)
        tr = np.ascontiguousarray(transforms, dtype=real)
        ra = np.ascontiguousarray(radial, dtype=real)
        ex = None if extra is None else np.ascontiguousarray(extra, dtype=real)
        rc = lib.ref_insert(0 if real == np.float32 else 1, VARIANTS[variant], PROFILES[window.profile],
                            window.half_width, window.shape, n, big, K, B, _ptr(num), _ptr(den), _ptr(sl),
                            _ptr(tr), _ptr(ra), _ptr(ex), optics.voltage, optics.spherical_aberration,
                            optics.amplitude_contrast, float(pixel_size), int(threads), int(tile_budget))
        if rc != 0:
            raise RuntimeError("ref_insert failed")
        return num, den

    return _run_expanded(variant, inputs, fn)


# ---------------------------------------------------------------------------
# FFTs (pocketfft through numpy), scale conventions of fourier_normalization.cpp:56-71
# ---------------------------------------------------------------------------

def rfft2(images):
    """Real 2-D DFT over the last two axes, unscaled (backward norm)."""
    return np.fft.rfft2(images, axes=(-2, -1))


def irfft3(spectrum, extent):
    """Inverse real 3-D DFT over the last three axes, scaled by 1/N^3."""
    return np.fft.irfftn(spectrum, s=(extent, extent, extent), axes=(-3, -2, -1))


# ---------------------------------------------------------------------------
# Closed forms downstream of the insertion
# ---------------------------------------------------------------------------

def normalize(num, den, threshold=1e-3):
    """num / den where den > threshold, else 0
    (tests/integration/src/em/functional/test_cpu_reconstruction.cpp:457-464)."""
    out = np.zeros_like(num)
    mask = den > threshold
    out[mask] = num[mask] / den[mask]
    return out


def window_transform(window: Window, extent: int):
    """The Fourier transform of the 1-D window sampled at the real-space
    positions x - C of a volume of extent N, normalised to 1 at 0.

    box: sinc(x / N); triangle: sinc^2(x / N); Kaiser-Bessel: the closed-form
    transform sinh(sqrt(b^2 - (2 pi w u)^2)) / sqrt(...) with u = x / N.
    """
    x = np.arange(extent, dtype=np.float64) - extent // 2
    u = x / extent
    if window.profile == "box":
        return np.sinc(u)
    if window.profile == "triangle":
        return np.sinc(u) ** 2
    w, b = window.half_width, window.shape
    z2 = b * b - (2 * math.pi * w * u) ** 2

    def kb(zz):
        zz = np.asarray(zz, dtype=np.float64)
        out = np.empty_like(zz)
        pos = zz > 1e-12
        neg = zz < -1e-12
        mid = ~(pos | neg)
        r = np.sqrt(np.abs(zz))
        out[pos] = np.sinh(r[pos]) / r[pos]
        out[neg] = np.sin(r[neg]) / r[neg]
        out[mid] = 1.0
        return out

    return kb(z2) / kb(np.array([b * b]))[0]


def gridding_correct(volume, window: Window, crop=None):
    """Divide a real-space volume (origin at C = N // 2) by the window's
    transform per axis; with ``crop`` = M, keep only the central M^3, from
    C - M // 2 (the crop of a reconstruction inserted with padding)."""
    extent = volume.shape[-1]
    c = window_transform(window, extent)
    out = volume / (c[:, None, None] * c[None, :, None] * c[None, None, :])
    if crop is None or crop == extent:
        return out
    o = extent // 2 - crop // 2
    return out[..., o:o + crop, o:o + crop, o:o + crop]


def shell_index(extent):
    """Shell of every element of a half spectrum (N, N, N/2+1): round-half-up of |g|."""
    f = np.fft.fftfreq(extent) * extent
    fx = np.arange(extent // 2 + 1)
    r = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + fx[None, None, :] ** 2)
    return np.floor(r + 0.5).astype(np.int64)


def fsc(a, b, extent):
    """Fourier shell correlation of two half spectra, per shell, in double."""
    shells = shell_index(extent).ravel()
    a = a.reshape(-1).astype(np.complex128)
    b = b.reshape(-1).astype(np.complex128)
    count = extent // 2 + 1
    keep = shells < count
    s = shells[keep]
    cross = np.bincount(s, weights=(a[keep] * np.conj(b[keep])).real, minlength=count)
    aa = np.bincount(s, weights=np.abs(a[keep]) ** 2, minlength=count)
    bb = np.bincount(s, weights=np.abs(b[keep]) ** 2, minlength=count)
    return cross / np.sqrt(np.maximum(aa * bb, 1e-300))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------------------
# Density compensation and its features (no reference function; restated on
# the full periodic spectrum, which the half spectrum stands for)
# ---------------------------------------------------------------------------

def full_from_half(half, extent):
    """The full (N, N, N) real field a half spectrum (N, N, N/2+1) of a real
    field stands for: element (z, y, x) with x beyond the half is (-z, -y, N-x)."""
    out = np.empty((extent, extent, extent), dtype=half.dtype)
    h = extent // 2 + 1
    out[:, :, :h] = half
    for x in range(h, extent):
        out[:, :, x] = np.roll(np.roll(half[::-1, ::-1, extent - x], 1, axis=0), 1, axis=1)
    return out


def density_compensation(den, passes=10, floor=1e-6):
    """From W = 1, W <- W / max(K3 * (W den), floor) with K3 = (1/6, 2/3, 1/6)
    per axis, periodic, in double; W = 0 where den <= 0.

    The fields live on the stored half grid: every pass expands W and den to
    the full grid through the Hermitian mirror of the stored half, so that a
    neighbour beyond the half always reads the mirrored stored value."""
    extent = den.shape[0]
    h = extent // 2 + 1
    d_half = np.asarray(den, dtype=np.float64)
    d = full_from_half(d_half, extent)
    w_half = np.ones_like(d_half)
    for _ in range(passes):
        c = convolve_k3(full_from_half(w_half, extent) * d)
        w_half = np.where(d_half > 0, w_half / np.maximum(c[:, :, :h], floor), 0.0)
    return w_half


K3 = {-1: 1 / 6, 0: 2 / 3, 1: 1 / 6}


def convolve_k3_direct(field):
    """The periodic 27-point sum of K3(dz) K3(dy) K3(dx) field[g + d]."""
    c = np.zeros_like(field)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                c += K3[dz] * K3[dy] * K3[dx] * np.roll(field, (-dz, -dy, -dx), axis=(0, 1, 2))
    return c


def convolve_k3(field):
    """convolve_k3_direct by its three axes in turn (the kernel is a product
    of one 1-D kernel per axis), 9 shifted sums instead of 27."""
    c = field
    for axis in (0, 1, 2):
        c = K3[0] * c + K3[1] * (np.roll(c, -1, axis=axis) + np.roll(c, 1, axis=axis))
    return c


COVERAGE_SHAPE = (8, 33, 64, 8)


def coverage_directions():
    """Unit directions u and the frame (e1 = du/dtheta, e2 = du/dphi) of the
    (theta, phi) grid of the coverage table, each (33, 64, 3): theta = k pi / 32
    from pole to pole, phi = k 2 pi / 64."""
    _, nt, npp, _ = COVERAGE_SHAPE
    theta = np.arange(nt) * (math.pi / (nt - 1))
    phi = np.arange(npp) * (2 * math.pi / npp)
    t, p = np.meshgrid(theta, phi, indexing="ij")
    u = np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], -1)
    e1 = np.stack([np.cos(t) * np.cos(p), np.cos(t) * np.sin(p), -np.sin(t)], -1)
    e2 = np.stack([-np.sin(p), np.cos(p), np.zeros_like(p)], -1)
    return u, e1, e2


def slice_normals(transforms):
    """Unit normals X x Y of the slice planes (rows 0 and 1 of each transform)."""
    t = np.asarray(transforms, np.float64).reshape(-1, 3, 4)
    n = np.cross(t[:, 0, :3], t[:, 1, :3])
    length = np.linalg.norm(n, axis=1, keepdims=True)
    return np.where(length > 0, n / np.where(length > 0, length, 1), 0.0)


def coverage_table(transforms, big):
    """(8, 33, 64, 8): for radius (r + 0.5) / 8 * big / 2 and each direction
    of the grid, the fractions of the slices with |n . u| < min(1, 1 / radius + 0.05)
    in each eighth of [0, pi) of the normal's angle around u (zeros where no
    plane passes)."""
    nr, nt, npp, nb = COVERAGE_SHAPE
    n = slice_normals(transforms)
    u, e1, e2 = coverage_directions()
    along = np.einsum("tpa,sa->tps", u, n)
    a1 = np.einsum("tpa,sa->tps", e1, n)
    a2 = np.einsum("tpa,sa->tps", e2, n)
    angle = np.arctan2(a2, a1)
    angle = np.where(angle < 0, angle + math.pi, angle)
    angle = np.where(angle >= math.pi, angle - math.pi, angle)
    bins = np.minimum(nb - 1, (angle * (nb / math.pi)).astype(np.int64))
    out = np.zeros(COVERAGE_SHAPE)
    for r in range(nr):
        radius = (r + 0.5) / nr * (0.5 * big)
        inside = (np.abs(along) < min(1.0, 1.0 / radius + 0.05)) & ((a1 != 0) | (a2 != 0))
        for b in range(nb):
            out[r, :, :, b] = (inside & (bins == b)).sum(-1)
        total = out[r].sum(-1, keepdims=True)
        out[r] = np.where(total > 0, out[r] / np.where(total > 0, total, 1), 0.0)
    return out


def compensation_features(den, table, scale):
    """(N, N, N/2+1, 10): |g|/(N/2), log1p(den * scale), and the 8 coverage
    fractions of the table entry nearest to the voxel's radius and direction
    (radius bin floor(8 |g| / (N/2)), the grid direction nearest in theta and
    in phi, phi wrapping around)."""
    extent = den.shape[0]
    h = extent // 2 + 1
    nr, nt, npp, nb = COVERAGE_SHAPE
    f = np.fft.fftfreq(extent) * extent
    z = f[:, None, None] * np.ones((1, extent, h))
    y = f[None, :, None] * np.ones((extent, 1, h))
    x = np.arange(h)[None, None, :] * np.ones((extent, extent, 1))
    length = np.sqrt(x * x + y * y + z * z)
    rho = length / (extent / 2)
    out = np.zeros((extent, extent, h, 10))
    out[..., 0] = rho
    out[..., 1] = np.log1p(np.maximum(np.asarray(den, np.float64), 0.0) * scale)
    r = np.minimum(nr - 1, (rho * nr).astype(np.int64))
    theta = np.where(length > 0, np.arccos(np.clip(z / np.where(length > 0, length, 1), -1, 1)), 0.0)
    phi = np.arctan2(y, x)
    phi = np.where(phi < 0, phi + 2 * math.pi, phi)
    it = np.minimum(nt - 1, (theta * ((nt - 1) / math.pi) + 0.5).astype(np.int64))
    ip = (phi * (npp / (2 * math.pi)) + 0.5).astype(np.int64) % npp
    out[..., 2:] = np.asarray(table)[r, it, ip]
    return out


# ---------------------------------------------------------------------------
# Background standardization
# ---------------------------------------------------------------------------

_PHILOX_M = (0xD2511F53, 0xCD9E8D57)
_PHILOX_W = (0x9E3779B9, 0xBB67AE85)
_MASK32 = 0xFFFFFFFF


def philox4x32_10(counter, key):
    """Philox4x32-10 of one block (src/backends/cpu/kernels/philox.inl:14-113;
    Salmon et al., SC 2011): (c0, c1, c2, c3), (k0, k1) -> four words."""
    c = [int(x) & _MASK32 for x in counter]
    k = [int(x) & _MASK32 for x in key]
    for r in range(10):
        if r:
            k = [(k[0] + _PHILOX_W[0]) & _MASK32, (k[1] + _PHILOX_W[1]) & _MASK32]
        p0 = _PHILOX_M[0] * c[0]
        p1 = _PHILOX_M[1] * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k[0]) & _MASK32, p1 & _MASK32, ((p0 >> 32) ^ c[3] ^ k[1]) & _MASK32, p0 & _MASK32]
    return tuple(c)


def uniform_open_closed(bits: int) -> float:
    """Top 53 bits plus one, times 2^-53: in (0, 1] (standard_normal.inl:20-25)."""
    return float((int(bits) >> 11) + 1) * (1.0 / 9007199254740992.0)


def standard_normal_pair(key: int, block: int):
    """Box-Muller of the block (block, 0, 0, 0) keyed by key (standard_normal.inl:27-44)."""
    key = int(key) & 0xFFFFFFFFFFFFFFFF
    block = int(block)
    r = philox4x32_10((block & _MASK32, block >> 32, 0, 0), (key & _MASK32, key >> 32))
    first = (r[1] << 32) | r[0]
    second = (r[3] << 32) | r[2]
    radius = math.sqrt(-2.0 * math.log(uniform_open_closed(first)))
    angle = 6.283185307179586476925 * uniform_open_closed(second)
    return radius * math.cos(angle), radius * math.sin(angle)


def standard_normal(key: int, counter: int) -> float:
    """Counters 2k and 2k + 1 take the pair of block k (standard_normal.inl:46-50)."""
    return standard_normal_pair(key, counter // 2)[counter % 2]


def background_moments(image, background, mean, scale, lower, upper):
    """(mean, 1 / deviation) of the finite background elements whose value,
    standardized with (mean, scale), lies within [lower, upper], in double;
    (0, 1) without any (background_moments_kernel.inl:117-135, 206-213)."""
    x = image.astype(np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        standardized = (x - mean) * scale
        valid = background & np.isfinite(x) & (standardized >= lower) & (standardized <= upper)
    count = int(valid.sum())
    if count == 0:
        return 0.0, 1.0
    kept = x[valid]
    m = float(kept.sum()) / count
    variance = float((kept * kept).sum()) / count - m * m
    return m, (1.0 / math.sqrt(variance) if variance > 0.0 else 1.0)


def background_mask(shape, radius):
    """Elements whose squared distance (in elements) from index n // 2 per
    axis exceeds radius^2 (background_geometry.inl:46-55)."""
    grids = np.meshgrid(*[np.arange(n) - n // 2 for n in shape], indexing="ij")
    squared = sum(g.astype(np.int64) ** 2 for g in grids)
    return squared.astype(np.float64) > float(radius) * float(radius)


def standardize_background(value, key, radius, lower, upper, image_rank=2):
    """The operation over trailing image axes, value (..., image), key
    broadcasting against the batch axes; the counter of an element is its
    row-major position among the elements sharing its key
    (standardize_background_operation.hpp:41-56)."""
    value = np.asarray(value)
    dtype = value.dtype.type
    batch_shape = value.shape[:value.ndim - image_rank]
    image_shape = value.shape[value.ndim - image_rank:]
    key = np.broadcast_to(np.asarray(key, dtype=np.uint64), batch_shape) if batch_shape else np.asarray(key, np.uint64)
    key_full = np.asarray(key, dtype=np.uint64)
    key_given = np.asarray(key, dtype=np.uint64)
    background = background_mask(image_shape, radius)
    image_elements = int(np.prod(image_shape)) if image_shape else 1
    out = np.empty_like(value)
    # Batch axes the key is shared along: its base (non-broadcast) strides are 0 there.
    shared = [key_full.strides[a] == 0 if key_full.ndim else True for a in range(len(batch_shape))]
    lower_t, upper_t = dtype(lower), dtype(upper)
    for b in np.ndindex(*batch_shape) if batch_shape else [()]:
        image = value[b]
        m1, s1 = background_moments(image, background, 0.0, 1.0, -math.inf, math.inf)
        m2, s2 = background_moments(image, background, m1, s1, lower, upper)
        standardized = (image.astype(dtype) - dtype(m2)) * dtype(s2)
        with np.errstate(invalid="ignore", over="ignore"):
            kept = (standardized >= lower_t) & (standardized <= upper_t) & np.isfinite(standardized)
        rank = 0
        for a, i in enumerate(b):
            if shared[a]:
                rank = rank * batch_shape[a] + i
        k = int(key_full[b]) if batch_shape else int(key_given)
        result = standardized.copy()
        flat_kept = kept.reshape(-1)
        flat = result.reshape(-1)
        for e in np.nonzero(~flat_kept)[0]:
            flat[e] = dtype(standard_normal(k, rank * image_elements + int(e)))
        out[b] = result
    return out
