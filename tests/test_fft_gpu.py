"""HIP FFTs against numpy.fft (pocketfft, the reference's FFT library).

The reference pins its float32 batch FFT to numpy with rtol/atol 1e-4
(tests/python/functional/test_fourier.py:75-83); the bar here is that, and a
size-independent relative error measured against the spectrum's norm.
"""
import numpy as np
import pytest
import torch

import rexlib_amd
from rexlib_amd import UnsupportedOperationError

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [4, 6, 8, 12, 15, 64, 96, 128, 192, 256, 320])
def test_rfft2_matches_numpy(n):
    print(f"rfft2 n={n}", flush=True)
    x = np.random.default_rng(n).normal(size=(3, n, n)).astype(np.float32)
    out = rexlib_amd.rfft2(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    ref = np.fft.rfft2(x.astype(np.float64))
    got = out.cpu().numpy()
    assert got.shape == ref.shape
    assert rel(got, ref) < 2e-6
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())


def test_rfft2_odd_batch_and_strides():
    x = np.random.default_rng(1).normal(size=(5, 16, 16)).astype(np.float32)
    t = torch.from_numpy(x).cuda().transpose(1, 2)  # strided rows
    out = torch.zeros((5, 16, 9), dtype=torch.complex64, device="cuda")
    rexlib_amd.rfft2(t, out=out)
    torch.cuda.synchronize()
    assert rel(out.cpu().numpy(), np.fft.rfft2(x.transpose(0, 2, 1).astype(np.float64))) < 2e-6


def test_rfft2_strided_rows_at_256():
    # Rows with an element stride of 256 take the four-step row kernel's
    # element-by-element loads rather than its 16-byte ones.
    x = np.random.default_rng(2).normal(size=(3, 256, 256)).astype(np.float32)
    t = torch.from_numpy(x).cuda().transpose(1, 2)
    out = rexlib_amd.rfft2(t)
    torch.cuda.synchronize()
    assert rel(out.cpu().numpy(), np.fft.rfft2(x.transpose(0, 2, 1).astype(np.float64))) < 2e-6


@pytest.mark.parametrize("n", [8, 12, 64, 128])
def test_irfft2_inverts_rfft2(n):
    x = np.random.default_rng(n).normal(size=(4, n, n)).astype(np.float32)
    spec = np.fft.rfft2(x.astype(np.float64)).astype(np.complex64)
    out = rexlib_amd.irfft2(torch.from_numpy(spec).cuda())
    torch.cuda.synchronize()
    ref = np.fft.irfft2(spec.astype(np.complex128), s=(n, n))
    assert rel(out.cpu().numpy(), ref) < 2e-6


@pytest.mark.parametrize("big", [8, 9, 12, 64, 96, 128, 256])
def test_irfft3_matches_numpy(big):
    print(f"irfft3 N={big}", flush=True)
    rng = np.random.default_rng(big)
    spec = (rng.normal(size=(big, big, big // 2 + 1)) + 1j * rng.normal(size=(big, big, big // 2 + 1)))
    spec = spec.astype(np.complex64)
    t = torch.from_numpy(spec).cuda()
    before = t.clone()
    out = rexlib_amd.irfft3(t)
    torch.cuda.synchronize()
    ref = np.fft.irfftn(spec.astype(np.complex128), s=(big, big, big))
    assert rel(out.cpu().numpy(), ref) < 3e-6
    assert torch.equal(t, before)  # the input is not modified


def test_irfft3_of_rfftn_is_identity_in_batches():
    big = 32
    x = np.random.default_rng(3).normal(size=(2, big, big, big)).astype(np.float32)
    spec = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3)).astype(np.complex64)
    out = rexlib_amd.irfft3(torch.from_numpy(spec).cuda())
    torch.cuda.synchronize()
    assert rel(out.cpu().numpy(), x) < 3e-6


def test_lengths_with_other_prime_factors_are_unsupported():
    with pytest.raises(UnsupportedOperationError):
        rexlib_amd.rfft2(torch.zeros((1, 7, 7), device="cuda"))
