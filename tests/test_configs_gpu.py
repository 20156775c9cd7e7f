"""The workloads at their own shapes, on the GPU, against the CPU references.

* 256^2 images into 256^3 in a full launch of 4,096 images plus a partial one
  of 848, through the fused image path and through rfft2 + insert_ctf_slices,
  against the reference's float32 kernels on numpy (pocketfft) spectra, and
  the finished volume against the same quotient and irfft3 on the host;
* every brick shape and task size the library holds, at every box size of
  the batched workload (64 to 320), against the reference's float32 kernels;
* 128^2 images staged as bfloat16 or scaled float16, zero-padded on chip to
  256^2 and inserted into 256^3: per-shell FSC against the float64 oracle,
  and the central 128^3 of the corrected volume;
* one graph holding reconstructions of all six box sizes with the tuning
  table applied, each against the reference and each replay against the
  sequential run;
* a preferred-orientation 256^3 denominator: the iterative compensation
  against its float64 restatement.

The bar is the reference's float32 tolerance, 1e-4 relative L2
(test_slice_insertion_program_builder.cpp:168-191).
"""
import os

import numpy as np
import pytest
import scipy.fft
import torch

import oracle
import rexlib_amd
from gpu_helpers import OPTICS, hip_window, to_device
from rexlib_amd import batched, em, learned, precision, synthetic, tune

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def reference_insert(inputs, window=oracle.Window.linear()):
    """The reference's float32 kernels where oracle/_ref is built, else the
    float64 contract."""
    if oracle.ref_available():
        return oracle.ref_insert("ctf", inputs, window, pixel_size=1.5, dtype=np.float32, threads=THREADS)
    return oracle.insert_f64("ctf", inputs, window, pixel_size=1.5)


def host_spectra(images):
    return scipy.fft.rfft2(np.asarray(images, np.float64), workers=THREADS).astype(np.complex64)


def zeros(big):
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    return num, den


@pytest.fixture
def tuning_table():
    tune.clear_tuning()
    yield tune.apply_all()
    tune.clear_tuning()


def test_full_launch_of_256_images_into_256_cubed(tuning_table):
    n = big = 256
    count = 4096 + 848
    print(f"256^2 -> 256^3, {count} images", flush=True)
    vol = synthetic.blob_volume(big, 400, 96.0, (2.0, 8.0), seed=2000)
    stack = synthetic.particle_stack(vol, count, n, seed=2001, snr=0.1, pixel_size=1.5, device="cuda")
    transforms = to_device(stack.transforms)
    ctf = to_device(stack.ctf)
    radial = torch.ones(n // 2 + 1, device="cuda")
    window = em.GriddingWindow.linear()

    num, den = zeros(big)
    for s in range(0, count, 4096):
        e = min(s + 4096, count)
        em.insert_ctf_images(num, den, stack.images[s:e], transforms[s:e], ctf[s:e], radial, window, OPTICS, 1.5)
    quotient = em.normalize(num, den, 1e-3)
    volume = rexlib_amd.irfft3(quotient)

    num2, den2 = zeros(big)
    for s in range(0, count, 4096):
        e = min(s + 4096, count)
        slices = rexlib_amd.rfft2(stack.images[s:e])
        em.insert_ctf_slices(num2, den2, slices, transforms[s:e], ctf[s:e], radial, window, OPTICS, 1.5,
                             numerator_out=num2, denominator_out=den2)
    torch.cuda.synchronize()

    acc = np.zeros((big, big, big // 2 + 1), np.complex64)
    inputs = [acc, acc.real.copy(), host_spectra(stack.images.cpu().numpy()), stack.transforms.astype(np.float32),
              stack.ctf.astype(np.float32), np.ones(n // 2 + 1, np.float32)]
    ref = reference_insert(inputs)
    for got_num, got_den in ((num, den), (num2, den2)):
        assert oracle.rel_l2(got_num.cpu().numpy(), ref[0]) <= 1e-4
        assert oracle.rel_l2(got_den.cpu().numpy(), ref[1]) <= 1e-4
    ref_volume = oracle.irfft3(oracle.normalize(ref[0].astype(np.complex128), ref[1].astype(np.float64)), big)
    assert oracle.rel_l2(volume.cpu().numpy(), ref_volume) <= 1e-4


BOXES = [64, 96, 128, 192, 256, 320]
SHARES = [32, 128, 1024]


def random_ctf_case(n, big, count, seed):
    rng = np.random.default_rng(seed)
    h = n // 2 + 1
    slices = (rng.uniform(-1, 1, (count, n, h)) + 1j * rng.uniform(-1, 1, (count, n, h))).astype(np.complex64)
    transforms = synthetic.transforms_of(synthetic.uniform_rotations(count, rng)).astype(np.float32)
    transforms[:, :2, 3] = rng.uniform(-3, 3, (count, 2))
    ctf = synthetic.ctf_parameters(count, rng).astype(np.float32)
    radial = rng.uniform(0, 2, h).astype(np.float32)
    acc = np.zeros((big, big, big // 2 + 1), np.complex64)
    return [acc, acc.real.copy(), slices, transforms, ctf, radial]


@pytest.mark.parametrize("big", BOXES)
def test_every_brick_shape_and_share(big, tuning_table):
    inputs = random_ctf_case(big, big, 160, seed=big)
    ref = reference_insert(inputs)
    t = [to_device(a) for a in inputs]
    try:
        for shape in tune.SHAPES:
            for share in SHARES:
                print(f"N={big} brick {shape} share={share}", flush=True)
                tune.set_tuning(big, shape, share)
                num, den = zeros(big)
                em.insert_ctf_slices(num, den, *t[2:], em.GriddingWindow.linear(), OPTICS, 1.5,
                                     numerator_out=num, denominator_out=den)
                torch.cuda.synchronize()
                assert oracle.rel_l2(num.cpu().numpy(), ref[0]) <= 1e-4, (shape, share)
                assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4, (shape, share)
    finally:
        tune.clear_tuning()


@pytest.mark.parametrize("n,big", [(96, 127), (128, 193), (160, 321)])
def test_odd_volume_extents(n, big, tuning_table):
    # Odd N: no +-N/2 aliasing, the volume phase is the tabulated product
    # (tabulated_phase.inl:12-37) rather than (-1)^(x+y+z), and the slices
    # are smaller than the volume, as in padded reconstructions.
    inputs = random_ctf_case(n, big, 200, seed=big)
    ref = reference_insert(inputs)
    t = [to_device(a) for a in inputs]
    num, den = zeros(big)
    em.insert_ctf_slices(num, den, *t[2:], em.GriddingWindow.linear(), OPTICS, 1.5,
                         numerator_out=num, denominator_out=den)
    torch.cuda.synchronize()
    assert oracle.rel_l2(num.cpu().numpy(), ref[0]) <= 1e-4
    assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4


@pytest.mark.parametrize("variant", ["plain", "weighted"])
def test_plain_and_weighted_insertion_into_256_cubed(variant, tuning_table):
    # The other two insertions of the path at the headline's volume, on the
    # tuned brick schedule, from accumulators that are not zero (in-place
    # accumulation onto earlier sums).
    rng = np.random.default_rng(77)
    n = big = 256
    count = 300
    h = n // 2 + 1
    slices = (rng.uniform(-1, 1, (count, n, h)) + 1j * rng.uniform(-1, 1, (count, n, h))).astype(np.complex64)
    transforms = synthetic.transforms_of(synthetic.uniform_rotations(count, rng)).astype(np.float32)
    radial = rng.uniform(0, 2, h).astype(np.float32)
    num0 = (rng.uniform(-1, 1, (big, big, h)) + 1j * rng.uniform(-1, 1, (big, big, h))).astype(np.complex64)
    den0 = rng.uniform(0, 2, (big, big, h)).astype(np.float32)
    if variant == "plain":
        inputs = [num0, den0, slices, transforms, radial]
    else:
        weights = rng.uniform(0, 2, (count, n, h)).astype(np.float32)
        inputs = [num0, den0, slices, weights, transforms, radial]
    if oracle.ref_available():
        ref = oracle.ref_insert(variant, inputs, oracle.Window.linear(), dtype=np.float32, threads=THREADS)
    else:
        ref = oracle.insert_f64(variant, inputs, oracle.Window.linear())
    t = [to_device(a) for a in inputs]
    if variant == "plain":
        num, den = em.insert_slices(*t, em.GriddingWindow.linear(), numerator_out=t[0], denominator_out=t[1])
    else:
        num, den = em.insert_weighted_slices(*t, em.GriddingWindow.linear(), numerator_out=t[0],
                                             denominator_out=t[1])
    torch.cuda.synchronize()
    assert oracle.rel_l2(num.cpu().numpy(), ref[0]) <= 1e-4
    assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4


@pytest.mark.parametrize("window", [oracle.Window.nearest(), oracle.Window.kaiser_bessel(2.0, 8.0)])
def test_every_brick_shape_with_one_and_four_taps(window, tuning_table):
    # Against the reference's float32 kernels: with tens of thousands of
    # samples some lie within a float32 rounding of a box window's half-voxel
    # boundary, where only the same float32 positions agree sample by sample.
    big = 96
    inputs = random_ctf_case(big, big, 80, seed=7)
    ref = reference_insert(inputs, window)
    t = [to_device(a) for a in inputs]
    try:
        for shape in tune.SHAPES:
            print(f"{window.profile} brick {shape}", flush=True)
            tune.set_tuning(big, shape, 64)
            num, den = zeros(big)
            em.insert_ctf_slices(num, den, *t[2:], hip_window(window), OPTICS, 1.5, numerator_out=num,
                                 denominator_out=den)
            torch.cuda.synchronize()
            assert oracle.rel_l2(num.cpu().numpy(), ref[0]) <= 1e-4, shape
            assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4, shape
    finally:
        tune.clear_tuning()


@pytest.mark.parametrize("window,taps", [(oracle.Window.nearest(), 1), (oracle.Window.kaiser_bessel(2.0, 8.0), 4),
                                         (oracle.Window.linear(), 2)])
def test_tuning_entry_for_one_window_family(window, taps, tuning_table):
    # A window family's own entry (here 8 x 6 x 6, share 1,024) beside a
    # different entry for the extent: the call takes its window's entry and
    # agrees with the reference's float32 kernels.
    big = 96
    inputs = random_ctf_case(big, big, 80, seed=11)
    ref = reference_insert(inputs, window)
    t = [to_device(a) for a in inputs]
    try:
        tune.set_tuning(big, (16, 16, 8), 64)
        tune.set_tuning(big, (8, 6, 6), 1024, taps=taps)
        num, den = zeros(big)
        em.insert_ctf_slices(num, den, *t[2:], hip_window(window), OPTICS, 1.5, numerator_out=num,
                             denominator_out=den)
        torch.cuda.synchronize()
        assert oracle.rel_l2(num.cpu().numpy(), ref[0]) <= 1e-4
        assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4
    finally:
        tune.clear_tuning()


def padded(images, n):
    m = images.shape[-1]
    o = n // 2 - m // 2
    out = np.zeros(images.shape[:-2] + (n, n), dtype=np.float64)
    out[..., o:o + m, o:o + m] = images
    return out


@pytest.fixture(scope="module")
def precision_case():
    """128^2 projections of a 128^3 blob volume (defocus 0.5 to 4 um),
    inserted with 2x padding into 256^3, and the float64 oracle's sums."""
    m, n = 128, 256
    count = 320
    vol = synthetic.blob_volume(m, 120, 48.0, (1.5, 5.0), seed=3000)
    stack = synthetic.particle_stack(vol, count, m, seed=3001, snr=0.1, pixel_size=1.5,
                                     defocus=(5000.0, 40000.0), device="cuda")
    images = padded(stack.images.cpu().numpy().astype(np.float64), n)
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    ref = oracle.insert_f64("ctf", [acc, acc.real.copy(), scipy.fft.rfft2(images, workers=THREADS), stack.transforms,
                                    stack.ctf, np.ones(n // 2 + 1)], oracle.Window.linear(), pixel_size=1.5)
    return stack, ref


@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp16"])
def test_128_images_padded_on_chip_into_256_cubed(dtype, precision_case):
    stack, ref = precision_case
    n = big = 256
    staged, scales = precision.stage_images(stack.images, dtype)
    num, den = zeros(big)
    em.insert_ctf_images(num, den, staged, to_device(stack.transforms), to_device(stack.ctf),
                         torch.ones(n // 2 + 1, device="cuda"), em.GriddingWindow.linear(), OPTICS, 1.5,
                         image_scales=scales, padded=n)
    quotient = em.normalize(num, den, 1e-3)
    ref_q = oracle.normalize(ref[0], ref[1])
    report = precision.shell_report(quotient, torch.from_numpy(ref_q.astype(np.complex64)).cuda(), cutoff=0.9)
    print(dtype, "min FSC", report["min_fsc_to_cutoff"], "max shell error", report["max_error_to_cutoff"], flush=True)
    assert report["min_fsc_to_cutoff"] >= 0.99
    assert oracle.rel_l2(den.cpu().numpy(), ref[1]) <= 1e-4
    # The central 128^3 of the corrected volume: the crop itself against the
    # host's irfft3 and correction of the same quotient, and the volume
    # against the float64 oracle's.
    window = em.GriddingWindow.linear()
    cropped = em.gridding_correct(rexlib_amd.irfft3(quotient), window, crop=128).cpu().numpy()
    assert cropped.shape == (128, 128, 128)
    own = oracle.gridding_correct(oracle.irfft3(quotient.cpu().numpy().astype(np.complex128), big),
                                  oracle.Window.linear(), crop=128)
    assert oracle.rel_l2(cropped, own) <= 1e-5
    ref_volume = oracle.gridding_correct(oracle.irfft3(ref_q, big), oracle.Window.linear(), crop=128)
    error = oracle.rel_l2(cropped, ref_volume)
    print(dtype, "cropped volume rel-L2 to the float64 oracle", error, flush=True)
    # float32 sums of 320 noisy images against float64: the outer shells,
    # whose weights are small, carry most of the difference.
    assert error <= (2e-3 if dtype == "f32" else 2e-2)


def test_one_graph_over_every_box_size(tuning_table):
    counts = [700, 500, 400, 300, 260, 220]
    recons = []
    for i, (n, count) in enumerate(zip(BOXES, counts)):
        vol = synthetic.blob_volume(n, 40, n * 0.3, (1.5, 4.0), seed=4000 + i)
        stack = synthetic.particle_stack(vol, count, n, seed=4100 + i, snr=0.05 + 0.09 * i, device="cuda")
        recons.append(batched.Reconstruction(stack.images, to_device(stack.transforms), to_device(stack.ctf)))
    batched.run_sequential(recons, tuning=True)
    torch.cuda.synchronize()
    expected = [(r.numerator.clone(), r.denominator.clone(), r.volume.clone()) for r in recons]
    for r in recons:
        n = r.extent
        print(f"box {n}: {r.images.shape[0]} images against the reference", flush=True)
        acc = np.zeros((n, n, n // 2 + 1), np.complex64)
        ref = reference_insert([acc, acc.real.copy(), host_spectra(r.images.cpu().numpy()),
                                r.transforms.cpu().numpy(), r.ctf.cpu().numpy(), np.ones(n // 2 + 1, np.float32)])
        assert oracle.rel_l2(r.numerator.cpu().numpy(), ref[0]) <= 1e-4, n
        assert oracle.rel_l2(r.denominator.cpu().numpy(), ref[1]) <= 1e-4, n
    graph = batched.BatchedReconstruction(recons, streams=4, tuning=True).capture()
    for replay in range(3):
        for r in recons:
            r.volume.fill_(float("nan"))
            r.numerator.mul_(2.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for (num, den, vol), r in zip(expected, recons):
            assert oracle.rel_l2(r.numerator.cpu().numpy(), num.cpu().numpy()) < 1e-5, (replay, r.extent)
            assert oracle.rel_l2(r.denominator.cpu().numpy(), den.cpu().numpy()) < 1e-5, (replay, r.extent)
            assert oracle.rel_l2(r.volume.cpu().numpy(), vol.cpu().numpy()) < 1e-4, (replay, r.extent)


def test_preferred_orientation_256_compensation_matches_the_oracle():
    big = 256
    count = 3000
    rng = np.random.default_rng(5000)
    k = int(0.8 * count)
    rot = np.concatenate([synthetic.von_mises_fisher_rotations(k, 6.0, rng),
                          synthetic.uniform_rotations(count - k, rng)])
    slices = torch.zeros((count, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    num, den = zeros(big)
    em.insert_slices(num, den, slices, to_device(synthetic.transforms_of(rot)), torch.ones(big // 2 + 1, device="cuda"),
                     em.GriddingWindow.linear(), numerator_out=num, denominator_out=den)
    w = learned.iterative_weights(den).cpu().numpy()
    ref = oracle.density_compensation(den.cpu().numpy(), passes=learned.PASSES, floor=learned.FLOOR)
    error = oracle.rel_l2(w, ref)
    print("256^3 iterative compensation rel-L2", error, flush=True)
    assert error < 1e-5
