"""Parity of the HIP insertion with the oracle and the reference's kernels.

The bar is the reference's own float32 tolerance: 1e-4 absolute against the
double-precision contract (test_slice_insertion_program_builder.cpp:168-191),
5e-4 for Kaiser-Bessel windows (slice_insertion_cases.hpp:648-654). The GPU
adds in an unspecified order (float atomics), so it is compared by tolerance,
never by bits.
"""
import os

import numpy as np
import pytest
import torch

import oracle
from cases import PIXEL_SIZE, SIZES, identity_transform, make_case, spread_kat_case
from gpu_helpers import hip_insert, hip_window, max_abs, to_device, OPTICS

from rexlib_amd import em

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "insertion_reference.npz")
WINDOWS = {
    "nearest": oracle.Window.nearest(),
    "linear": oracle.Window.linear(),
    "kb1.5_6": oracle.Window.kaiser_bessel(1.5, 6.0),
    "kb2_8": oracle.Window.kaiser_bessel(2.0, 8.0),
}
SCHEDULES = ["scatter", "brick"]


@pytest.fixture(params=SCHEDULES)
def schedule(request, monkeypatch):
    monkeypatch.setenv("REXLIB_HIP_INSERTION_SCHEDULE", request.param)
    return request.param


def tol32(window):
    return 5e-4 if window.profile == "kaiser_bessel" else 1e-4


def golden_names():
    with np.load(GOLDEN, allow_pickle=False) as g:
        return sorted({k.split("/")[0] for k in g.files})


@pytest.mark.parametrize("name", golden_names())
def test_matches_the_reference_golden_vectors(name, schedule):
    print(f"golden case {name} schedule {schedule}", flush=True)
    g = np.load(GOLDEN, allow_pickle=False)
    meta = g[f"{name}/meta"]
    variant = str(g[f"{name}/variant"])
    window = WINDOWS[str(g[f"{name}/window"])]
    inputs = [g[f"{name}/in{i}"] for i in range(int(meta[2]))]
    num, den = hip_insert(variant, inputs, window, PIXEL_SIZE)
    tol = tol32(window)
    assert max_abs(num, g[f"{name}/num32"]) <= tol
    assert max_abs(den, g[f"{name}/den32"]) <= tol
    assert max_abs(num, g[f"{name}/num64"]) <= tol
    assert max_abs(den, g[f"{name}/den64"]) <= tol


@pytest.mark.parametrize("n,big", SIZES)
@pytest.mark.parametrize("wname", ["nearest", "linear", "kb1.5_6"])
def test_plain_insertion_matches_the_oracle(n, big, wname, schedule):
    print(f"plain n={n} N={big} {wname} {schedule}", flush=True)
    inputs = make_case("plain", n, big, 21)
    w = WINDOWS[wname]
    num, den = hip_insert("plain", inputs, w)
    ref = oracle.insert_f64("plain", inputs, w)
    assert max_abs(num, ref[0]) <= tol32(w)
    assert max_abs(den, ref[1]) <= tol32(w)


@pytest.mark.parametrize("variant", ["weighted", "ctf"])
@pytest.mark.parametrize("n,big", [(6, 12), (5, 7), (8, 16)])
@pytest.mark.parametrize("wname", ["linear", "kb2_8", "nearest"])
def test_weighted_and_ctf_insertions_match_the_oracle(variant, n, big, wname, schedule):
    print(f"{variant} n={n} N={big} {wname} {schedule}", flush=True)
    inputs = make_case(variant, n, big, 23)
    w = WINDOWS[wname]
    num, den = hip_insert(variant, inputs, w, PIXEL_SIZE)
    ref = oracle.insert_f64(variant, inputs, w, pixel_size=PIXEL_SIZE)
    assert max_abs(num, ref[0]) <= tol32(w)
    assert max_abs(den, ref[1]) <= tol32(w)


def test_spread_of_one_sample_over_eight_voxels(schedule):
    num, den = hip_insert("weighted", spread_kat_case(), oracle.Window.linear())
    assert den[7, 2, 1] == pytest.approx(0.75 * 0.5 * 0.75, abs=1e-6)
    assert den[0, 3, 2] == pytest.approx(0.25 * 0.5 * 0.25, abs=1e-6)
    assert num[7, 2, 1].real == pytest.approx(2 * 0.75 * 0.5 * 0.75, abs=1e-6)
    assert num[0, 3, 2].real == pytest.approx(-2 * 0.25 * 0.5 * 0.25, abs=1e-6)
    assert float(den.sum()) == pytest.approx(1.0, abs=1e-6)


def test_ctf_at_zero_frequency_is_the_amplitude_contrast(schedule):
    n = 8
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    num, den = hip_insert("ctf", [acc, acc.real.copy(), np.ones((n, n // 2 + 1), np.complex128),
                                  identity_transform(), np.array([15000, 15000, 0, 0.0]), np.ones(n // 2 + 1)],
                          oracle.Window.nearest(), 1.5)
    assert den[0, 0, 0] == pytest.approx(0.01, rel=1e-5)
    assert num[0, 0, 0].real == pytest.approx(0.1, rel=1e-5)


@pytest.mark.parametrize("n", [7, 8])
def test_projection_lands_on_the_central_plane(n, schedule):
    volume = np.random.default_rng(1).normal(size=(n, n, n))
    slices = np.fft.rfft2(volume.sum(axis=0))
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    num, den = hip_insert("plain", [acc, acc.real.copy(), slices, identity_transform(), np.ones(n // 2 + 1)],
                          oracle.Window.nearest())
    inserted = den > 0
    assert np.all(den[inserted] == 1)
    assert not np.any(inserted[1:])
    expected = np.fft.rfftn(volume)[inserted]
    assert np.max(np.abs(num[inserted] - expected)) <= 1e-5 * np.max(np.abs(expected))


def test_nan_beyond_the_cutoff_is_left_out(schedule):
    inputs = make_case("plain", 8, 16, 4)
    inputs[4][:, 3] = 0.0
    inputs[4][:, 4] = 0.0
    inputs[2][:, 0, 4] = np.nan + 1j * np.nan
    num, den = hip_insert("plain", inputs, oracle.Window.linear())
    assert not np.isnan(num).any() and not np.isnan(den).any()
    ref = oracle.insert_f64("plain", inputs, oracle.Window.linear())
    assert max_abs(num, ref[0]) <= 1e-4


def test_batch_axes_the_accumulator_lacks_are_summed(schedule):
    inputs = make_case("plain", 6, 8, 3, accumulator_batch=(2, 1), transform_batch=(2, 3))
    num, den = hip_insert("plain", inputs, oracle.Window.linear())
    ref = oracle.insert_f64("plain", inputs, oracle.Window.linear())
    assert num.shape == (2, 1, 8, 8, 5)
    assert max_abs(num, ref[0]) <= 1e-4 and max_abs(den, ref[1]) <= 1e-4


def test_out_of_place_leaves_the_accumulators_and_matches_in_place(schedule):
    inputs = make_case("ctf", 8, 16, 8)
    t = [to_device(a) for a in inputs]
    before = t[0].clone(), t[1].clone()
    w = em.GriddingWindow.linear()
    num, den = em.insert_ctf_slices(*t, w, OPTICS, PIXEL_SIZE)
    torch.cuda.synchronize()
    assert torch.equal(t[0], before[0]) and torch.equal(t[1], before[1])
    ref = oracle.insert_f64("ctf", inputs, oracle.Window.linear(), pixel_size=PIXEL_SIZE)
    assert max_abs(num.cpu().numpy(), ref[0]) <= 1e-4
    em.insert_ctf_slices(*t, w, OPTICS, PIXEL_SIZE, numerator_out=t[0], denominator_out=t[1])
    torch.cuda.synchronize()
    assert max_abs(t[0].cpu().numpy(), ref[0]) <= 1e-4
    assert max_abs(t[1].cpu().numpy(), ref[1]) <= 1e-4


def test_split_batch_equals_whole(schedule):
    inputs = make_case("ctf", 8, 16, 7, slices=6, zero_accumulators=True)
    t = [to_device(a) for a in inputs]
    w = em.GriddingWindow.linear()
    whole = em.insert_ctf_slices(*t, w, OPTICS, PIXEL_SIZE)
    acc = t[0].clone(), t[1].clone()
    for part in (slice(0, 2), slice(2, 6)):
        em.insert_ctf_slices(acc[0], acc[1], *[x[part].contiguous() for x in t[2:]], w, OPTICS, PIXEL_SIZE,
                             numerator_out=acc[0], denominator_out=acc[1])
    torch.cuda.synchronize()
    assert torch.allclose(acc[0], whole[0], atol=1e-5)
    assert torch.allclose(acc[1], whole[1], atol=1e-5)


def test_strided_operands(schedule):
    inputs = make_case("weighted", 8, 12, 9)
    ref = oracle.insert_f64("weighted", inputs, oracle.Window.linear())
    t = [to_device(a) for a in inputs]
    # Slices and weights as transposed views, accumulators inside padded buffers.
    sl = t[2].transpose(1, 2).contiguous().transpose(1, 2)
    wt = t[3].permute(2, 0, 1).contiguous().permute(1, 2, 0)
    big_num = torch.zeros((12, 14, 9), dtype=torch.complex64, device="cuda")
    big_den = torch.zeros((12, 13, 8), dtype=torch.float32, device="cuda")
    num = big_num[:, :12, :7]
    den = big_den[:, :12, :7]
    num.copy_(t[0])
    den.copy_(t[1])
    em.insert_weighted_slices(num, den, sl, wt, t[4], t[5], em.GriddingWindow.linear(), numerator_out=num,
                              denominator_out=den)
    torch.cuda.synchronize()
    assert max_abs(num.cpu().numpy(), ref[0]) <= 1e-4
    assert max_abs(den.cpu().numpy(), ref[1]) <= 1e-4


@pytest.mark.parametrize("variant", ["plain", "ctf"])
def test_production_like_sizes_match_the_reference_kernels(variant, schedule):
    # 64 slices of n = 64 into N = 64 and n = 32 into N = 64 (2x padding).
    for n, big in [(64, 64), (32, 64)]:
        print(f"{variant} n={n} N={big} {schedule}", flush=True)
        rng = np.random.default_rng(n + big)
        inputs = make_case(variant, n, big, 31, slices=64, zero_accumulators=True)
        # Orthonormal frames, as real particles have.
        from rexlib_amd.synthetic import transforms_of, uniform_rotations
        t = transforms_of(uniform_rotations(64, rng))
        t[:, :2, 3] = rng.uniform(-3, 3, (64, 2))
        inputs[3] = t
        w = oracle.Window.linear()
        num, den = hip_insert(variant, inputs, w, PIXEL_SIZE)
        if oracle.ref_available():
            ref = oracle.ref_insert(variant, inputs, w, pixel_size=PIXEL_SIZE, dtype=np.float32, threads=8)
        else:
            ref = oracle.insert_f64(variant, inputs, w, pixel_size=PIXEL_SIZE)
        assert oracle.rel_l2(num, ref[0]) <= 1e-5
        assert oracle.rel_l2(den, ref[1]) <= 1e-5


def test_rows_longer_than_1023_samples(schedule):
    # n = 4096 into N = 8: every row of the slice reaches the one brick with
    # up to 2,049 samples, so the per-row counts no longer fit the 16-bit
    # halves of a shared scan and each class takes a scan of its own.
    # One sample in a thousand is nonzero (values and weights), which keeps
    # the float32 sums of the 320 voxels short.
    print(f"long rows n=4096 N=8 {schedule}", flush=True)
    inputs = make_case("weighted", 4096, 8, 41, slices=1, zero_accumulators=True)
    keep = np.random.default_rng(42).random(inputs[2].shape) < 1e-3
    inputs[2] = inputs[2] * keep
    inputs[3] = inputs[3] * keep
    w = oracle.Window.linear()
    num, den = hip_insert("weighted", inputs, w)
    ref = oracle.insert_f64("weighted", inputs, w)
    assert oracle.rel_l2(num, ref[0]) <= 1e-4
    assert oracle.rel_l2(den, ref[1]) <= 1e-4


def test_kernel_timing_records_each_launch(schedule):
    """Event timing brackets every launch of the schedule's insertion kernel
    and of nothing else; switching it off stops the records."""
    from rexlib_amd import _native as N
    inputs = make_case("ctf", 32, 32, 9, slices=40)
    kernel = N.KERNEL_BRICK if schedule == "brick" else N.KERNEL_SCATTER
    other = N.KERNEL_SCATTER if schedule == "brick" else N.KERNEL_BRICK
    N.set_kernel_timing(True)
    try:
        for _ in range(3):
            hip_insert("ctf", inputs, oracle.Window.linear(), PIXEL_SIZE)
        total, launches = N.kernel_timing(kernel)
        assert launches == 3 and total > 0
        assert N.kernel_timing(other) == (0.0, 0)
    finally:
        N.set_kernel_timing(False)
    hip_insert("ctf", inputs, oracle.Window.linear(), PIXEL_SIZE)
    assert N.kernel_timing(kernel) == (0.0, 0)
