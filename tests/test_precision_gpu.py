"""The fused image path and the reduced-precision staging forms.

insert_ctf_images (rfft2 of real images, zero-padded on chip, then the CTF
insertion) is checked against the double-precision oracle on the padded
float64 images: float32 staging to the 1e-4 bar; bfloat16 and per-image
scaled float16 staging by per-shell FSC >= 0.99 up to 0.9 Nyquist, with the
per-shell relative error reported by the float64 shell sums.
"""
import numpy as np
import pytest
import torch

import oracle
import rexlib_amd
from gpu_helpers import OPTICS, to_device
from rexlib_amd import em, precision, synthetic

pytestmark = pytest.mark.gpu


def padded(images, n):
    m = images.shape[-1]
    o = n // 2 - m // 2
    out = np.zeros(images.shape[:-2] + (n, n), dtype=np.float64)
    out[..., o:o + m, o:o + m] = images
    return out


def make_stack(count=48, m=32, n=64, seed=3):
    vol = synthetic.blob_volume(m, 12, m * 0.3, (1.0, 3.0), seed=seed)
    stack = synthetic.particle_stack(vol, count, m, seed=seed + 1, snr=0.5, pixel_size=1.5,
                                     defocus=(5000.0, 40000.0), device="cuda")
    return stack


def oracle_insert(stack, n, big):
    images = padded(stack.images.cpu().numpy().astype(np.float64), n)
    slices = np.fft.rfft2(images)
    acc = np.zeros((big, big, big // 2 + 1), np.complex128)
    return oracle.insert_f64("ctf", [acc, acc.real.copy(), slices, stack.transforms, stack.ctf,
                                     np.ones(n // 2 + 1)], oracle.Window.linear(), pixel_size=1.5)


def gpu_insert(stack, dtype, n, big):
    staged, scales = precision.stage_images(stack.images, dtype)
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    em.insert_ctf_images(num, den, staged, to_device(stack.transforms), to_device(stack.ctf),
                         torch.ones(n // 2 + 1, device="cuda"), em.GriddingWindow.linear(), OPTICS, 1.5,
                         image_scales=scales, padded=n)
    torch.cuda.synchronize()
    return num, den


def test_fused_images_equal_rfft2_then_insertion():
    stack = make_stack(m=64, n=64)
    n = big = 64
    num, den = gpu_insert(stack, "f32", n, big)
    slices = rexlib_amd.rfft2(stack.images)
    num2 = torch.zeros_like(num)
    den2 = torch.zeros_like(den)
    em.insert_ctf_slices(num2, den2, slices, to_device(stack.transforms), to_device(stack.ctf),
                         torch.ones(n // 2 + 1, device="cuda"), em.GriddingWindow.linear(), OPTICS, 1.5,
                         numerator_out=num2, denominator_out=den2)
    torch.cuda.synchronize()
    assert oracle.rel_l2(num.cpu().numpy(), num2.cpu().numpy()) < 1e-6
    assert oracle.rel_l2(den.cpu().numpy(), den2.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("schedule,m", [("auto", 128), ("scatter", 128), ("auto", 130)])
def test_fused_images_at_256_equal_rfft2_then_insertion(schedule, m, monkeypatch):
    # 256-point slices take the column transforms fused with the sample
    # preparation (brick schedule) or run them on their own (scatter). With
    # m = 130 the image starts 63 columns in, off the 4-element alignment of
    # the row kernel's 16-byte loads, so it loads element by element.
    monkeypatch.setenv("REXLIB_HIP_INSERTION_SCHEDULE", schedule)
    stack = make_stack(count=40, m=m, n=256)
    n, big = 256, 64
    num, den = gpu_insert(stack, "f32", n, big)
    images = torch.from_numpy(padded(stack.images.cpu().numpy(), n).astype(np.float32)).cuda()
    slices = rexlib_amd.rfft2(images)
    num2 = torch.zeros_like(num)
    den2 = torch.zeros_like(den)
    em.insert_ctf_slices(num2, den2, slices, to_device(stack.transforms), to_device(stack.ctf),
                         torch.ones(n // 2 + 1, device="cuda"), em.GriddingWindow.linear(), OPTICS, 1.5,
                         numerator_out=num2, denominator_out=den2)
    torch.cuda.synchronize()
    assert oracle.rel_l2(num.cpu().numpy(), num2.cpu().numpy()) < 1e-6
    assert oracle.rel_l2(den.cpu().numpy(), den2.cpu().numpy()) < 1e-6
    ref = oracle_insert(stack, n, big)
    assert oracle.rel_l2(num.cpu().numpy(), ref[0]) <= 1e-4
    assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_reduced_precision_staging_at_256_equals_the_dequantised_images(dtype):
    # The staged images, widened to float32 on chip (with the per-image
    # scale for fp16), must give what the same widened images give through
    # rfft2 and insert_ctf_slices in float32.
    stack = make_stack(count=24, m=128, n=256, seed=7)
    n, big = 256, 64
    staged, scales = precision.stage_images(stack.images, dtype)
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    em.insert_ctf_images(num, den, staged, to_device(stack.transforms), to_device(stack.ctf),
                         torch.ones(n // 2 + 1, device="cuda"), em.GriddingWindow.linear(), OPTICS, 1.5,
                         image_scales=scales, padded=n)
    widened = staged.float() if scales is None else staged.float() * scales[:, None, None]
    images = torch.zeros((widened.shape[0], n, n), device="cuda")
    o = n // 2 - 128 // 2
    images[:, o:o + 128, o:o + 128] = widened
    slices = rexlib_amd.rfft2(images)
    num2 = torch.zeros_like(num)
    den2 = torch.zeros_like(den)
    em.insert_ctf_slices(num2, den2, slices, to_device(stack.transforms), to_device(stack.ctf),
                         torch.ones(n // 2 + 1, device="cuda"), em.GriddingWindow.linear(), OPTICS, 1.5,
                         numerator_out=num2, denominator_out=den2)
    torch.cuda.synchronize()
    assert oracle.rel_l2(num.cpu().numpy(), num2.cpu().numpy()) < 1e-6
    assert oracle.rel_l2(den.cpu().numpy(), den2.cpu().numpy()) < 1e-6


def test_float32_staging_with_padding_matches_the_oracle():
    stack = make_stack()
    n = big = 64
    num, den = gpu_insert(stack, "f32", n, big)
    ref = oracle_insert(stack, n, big)
    assert oracle.rel_l2(num.cpu().numpy(), ref[0]) <= 1e-4
    assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_reduced_precision_staging_keeps_fsc(dtype):
    stack = make_stack()
    n = big = 64
    num, den = gpu_insert(stack, dtype, n, big)
    ref = oracle_insert(stack, n, big)
    ref_q = torch.from_numpy(oracle.normalize(ref[0], ref[1]).astype(np.complex64)).cuda()
    got_q = em.normalize(num, den)
    report = precision.shell_report(got_q, ref_q, cutoff=0.9)
    print(dtype, "min FSC", report["min_fsc_to_cutoff"], "max shell error", report["max_error_to_cutoff"])
    assert report["min_fsc_to_cutoff"] >= 0.99
    # The denominator does not see the images: it matches as in float32.
    assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4
    if dtype == "fp16":
        assert report["max_error_to_cutoff"] < 2e-3
    else:
        assert report["max_error_to_cutoff"] < 2e-2


def test_fp16_scales_restore_the_images():
    x = torch.randn((4, 16, 16), device="cuda") * torch.tensor([1e-3, 1.0, 1e3, 1e5], device="cuda")[:, None, None]
    staged, scales = precision.stage_images(x, "fp16")
    back = staged.float() * scales[:, None, None]
    rel = ((back - x).norm(dim=(1, 2)) / x.norm(dim=(1, 2))).max().item()
    assert rel < 1e-3
