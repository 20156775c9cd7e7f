"""General transforms on the GPU against numpy (pocketfft): complex, real and
inverse real DFTs over any axes, any strides, lengths with prime factors
beyond 5 (Bluestein), and rexlib's three normalisations, at the reference's
float32 tolerance (tests/python/functional/test_fourier.py:75-83: rtol and
atol 1e-4, here on the relative L2 and on the largest element)."""
import numpy as np
import pytest
import torch

import oracle
from rexlib_amd import fourier

pytestmark = pytest.mark.gpu

NORMS = ["backward", "ortho", "forward"]


def close(got, want, tol=1e-4):
    got = np.asarray(got, np.complex128)
    want = np.asarray(want, np.complex128)
    assert got.shape == want.shape
    assert oracle.rel_l2(got, want) <= tol
    assert np.max(np.abs(got - want)) <= tol * max(1.0, np.max(np.abs(want)))


def crandn(rng, shape):
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)


@pytest.mark.parametrize("shape,axes", [((12, 7, 10), (0, 2)), ((5, 97), (1,)), ((33, 17), (0, 1)),
                                         ((4, 6, 11, 3), (2, 0, 1)), ((1000,), (0,)), ((2048,), None),
                                         ((2, 3, 101), (-1,))])
@pytest.mark.parametrize("norm", NORMS)
def test_complex_transforms_match_numpy(shape, axes, norm):
    rng = np.random.default_rng(sum(shape))
    x = crandn(rng, shape)
    t = torch.from_numpy(x).cuda()
    close(fourier.fftn(t, axes, norm).cpu().numpy(), np.fft.fftn(x.astype(np.complex128), axes=axes, norm=norm))
    close(fourier.ifftn(t, axes, norm).cpu().numpy(), np.fft.ifftn(x.astype(np.complex128), axes=axes, norm=norm))


@pytest.mark.parametrize("shape,axes", [((9, 14), (0, 1)), ((6, 13, 8), (2, 1)), ((3, 7, 17), (0, 2)),
                                         ((40, 30, 22), None), ((5, 128), (1,))])
@pytest.mark.parametrize("norm", NORMS)
def test_real_transforms_match_numpy(shape, axes, norm):
    rng = np.random.default_rng(len(shape) * 100 + shape[-1])
    x = rng.normal(size=shape).astype(np.float32)
    t = torch.from_numpy(x).cuda()
    spec = fourier.rfftn(t, axes, norm)
    want = np.fft.rfftn(x.astype(np.float64), axes=axes, norm=norm)
    close(spec.cpu().numpy(), want)
    last = (axes[-1] if axes is not None else len(shape) - 1) % len(shape)
    ax = tuple(range(len(shape))) if axes is None else axes
    s = [shape[a % len(shape)] for a in ax]
    back = fourier.irfftn(spec, s=s, axes=axes, norm=norm)
    close(back.cpu().numpy(), x)
    assert spec.shape[last] == shape[last] // 2 + 1


def test_strided_operands_and_aliasing():
    rng = np.random.default_rng(5)
    base = crandn(rng, (10, 14, 6))
    t = torch.from_numpy(base).cuda()
    view = t[::2, 1:12, :].transpose(0, 2)  # (6, 11, 5), strided, non-contiguous
    want = np.fft.fftn(base[::2, 1:12, :].transpose(2, 1, 0).astype(np.complex128), axes=(0, 1))
    close(fourier.fftn(view, (0, 1)).cpu().numpy(), want)
    # In place: out aliases in.
    y = torch.from_numpy(crandn(rng, (7, 9))).cuda()
    want = np.fft.fftn(y.cpu().numpy().astype(np.complex128))
    fourier.fftn(y, out=y)
    close(y.cpu().numpy(), want)


# Dense operands over axes of 2-3-5 lengths take the direct path (axis passes
# from the input straight into the output, the normalisation in the last);
# the cases cover an image stack over its last two axes, an axis of extent 1
# among the transformed ones, and a transform in place.
@pytest.mark.parametrize("shape,axes", [((16, 256, 256), (-2, -1)), ((1, 8, 12), (0, 1)), ((3, 20, 45), (2, 0)),
                                         ((256, 4, 6), (0,))])
@pytest.mark.parametrize("norm", NORMS)
def test_direct_complex_transforms(shape, axes, norm):
    print("direct fft", shape, axes, norm, flush=True)
    rng = np.random.default_rng(sum(shape) + len(norm))
    x = crandn(rng, shape)
    t = torch.from_numpy(x).cuda()
    close(fourier.fftn(t, axes, norm).cpu().numpy(), np.fft.fftn(x.astype(np.complex128), axes=axes, norm=norm))
    close(fourier.ifftn(t, axes, norm).cpu().numpy(), np.fft.ifftn(x.astype(np.complex128), axes=axes, norm=norm))
    fourier.fftn(t, axes, norm, out=t)
    close(t.cpu().numpy(), np.fft.fftn(x.astype(np.complex128), axes=axes, norm=norm))


def test_irfft_ignores_the_imaginary_parts_of_self_conjugate_coefficients():
    rng = np.random.default_rng(7)
    spec = crandn(rng, (8, 5))  # n = 8 along the last axis
    t = torch.from_numpy(spec).cuda()
    got = fourier.irfftn(t, s=(8, 8), axes=(0, 1)).cpu().numpy()
    want = np.fft.irfftn(spec.astype(np.complex128), s=(8, 8), axes=(0, 1))
    close(got, want)


def test_float64_is_unsupported():
    from rexlib_amd import _native as N
    x = torch.zeros((4, 4), dtype=torch.complex128, device="cuda")
    with pytest.raises(N.UnsupportedOperationError):
        fourier.fftn(x, out=torch.zeros_like(x))
