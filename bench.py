"""Benchmark: images inserted per second, 256^2 images into a 256^3 half-complex grid.

Workload (the plain GPU configuration): a 256^3 volume of 400 Gaussian blobs
(sigma 2-8 voxels), 50,000 projections of 256 x 256 float32 at uniform
orientations, CTF defocus 1-3 um with up to 0.1 um astigmatism (300 kV,
Cs 2.7 mm, Q 0.1, 1.5 A pixels), white noise at SNR 0.1. One step is the whole
reconstruction: zero the accumulators, then for every launch of 4,096 images
rfft2 + CTF-weighted trilinear insertion into 256 x 256 x 129 complex64 +
float32 accumulators (the fused image entry point, rexlib_hip_insert_ctf_images,
on the float32 images), then the weight quotient, irfft3 and gridding
correction. Inputs are resident in HBM before the timed region. The same
reconstruction through the two calls of the rexlib API (rfft2, then
insert_ctf_slices) is timed beside it.

Printed: one JSON line with the throughput, the roofline of the dominant
kernel, the CPU baseline (the reference's own insertion kernels, compiled from
its sources, on the host cores) and the parity of a sample against it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("images inserted/s 256²→256³, plain FP32 vs BF16+batched+learned, "
          "at rel-L2≤1e-4")
HBM_PEAK_GBS = 8000.0


def log(*args):
    print(*args, file=sys.stderr, flush=True)


def algorithmic_bytes_per_image(n, big, launch, pixel_bytes=4):
    """n^2 s_pix + 64 + 24 N^2 (N/2+1) / L: the image, its transform and CTF,
    and the accumulators read and written once per launch."""
    return n * n * pixel_bytes + 64 + 24.0 * big * big * (big // 2 + 1) / launch


def traffic_per_launch(args, tuned):
    """HBM bytes per brick-kernel launch of this workload, from the committed
    rocprofv3 FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch, scaled by the
    factor a copy kernel of known size gives), or None when absent."""
    path = os.path.join(ROOT, "profiles", "brick_kernel_256_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        record = json.load(f)
    if record.get("n") != args.n or record.get("N") != args.big or record.get("launch") != args.launch:
        return None
    brick = record.get("brick", {})
    if tuned is None or ((brick.get("extent_x"), brick.get("extent_y"), brick.get("extent_z")),
                         brick.get("share")) != (tuple(tuned[0]), tuned[1]):
        return None  # measured for another brick shape
    return record["bytes_per_launch"]


def make_workload(args, torch):
    from rexlib_amd import synthetic
    t0 = time.time()
    vol = synthetic.blob_volume(args.big, args.blobs, args.big * 0.375, (2.0, 8.0), seed=1000)
    stack = synthetic.particle_stack(vol, args.images, args.n, seed=1001, snr=0.1, pixel_size=1.5,
                                     defocus=(10000.0, 30000.0), astigmatism=1000.0, device="cuda",
                                     chunk=512)
    torch.cuda.synchronize()
    log(f"generated {args.images} images of {args.n}^2 in {time.time() - t0:.1f} s")
    return stack


def host_cpu():
    """(logical cores, model name) of the host, from os.cpu_count() and lscpu."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count() or 1, model


def cpu_share():
    """Host threads for the CPU baseline: the cores this process may run on,
    at most OMP_NUM_THREADS (the GPU boxes allot 16 host cores per GPU and
    set it so)."""
    try:
        available = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        available = os.cpu_count() or 1
    limit = os.environ.get("OMP_NUM_THREADS")
    if limit and limit.isdigit() and int(limit) > 0:
        available = min(available, int(limit))
    return max(1, available)


def cpu_baseline(args, stack, threads):
    """The reference's CPU path on the host cores: pocketfft rfft2 (scipy.fft)
    and the reference's insertion kernels (oracle/_ref), over a bounded sample
    of the same images; then the weight quotient and the 3-D inverse FFT of
    the accumulators (numpy and scipy.fft), timed once, so that the
    end-to-end figure is the whole reconstruction of `args.images` images at
    the sampled insertion rate."""
    import oracle
    import scipy.fft
    if not oracle.ref_available():
        return None
    n, big = args.n, args.big
    h = n // 2 + 1
    acc = np.zeros((big, big, big // 2 + 1), np.complex64)
    den = np.zeros((big, big, big // 2 + 1), np.float32)
    chunk = 256
    done = 0
    elapsed = 0.0
    start = 0
    images = stack.images
    # Warm-up chunk, untimed.
    warm = images[:16].cpu().numpy()
    inputs = [acc, den, scipy.fft.rfft2(warm, workers=threads).astype(np.complex64),
              stack.transforms[:16].astype(np.float32), stack.ctf[:16].astype(np.float32), np.ones(h, np.float32)]
    oracle.ref_insert("ctf", inputs, oracle.Window.linear(), pixel_size=1.5, threads=threads)
    while elapsed < args.cpu_seconds and start + chunk <= images.shape[0]:
        host = images[start:start + chunk].cpu().numpy()
        t0 = time.perf_counter()
        spectra = scipy.fft.rfft2(host, workers=threads).astype(np.complex64)
        inputs = [acc, den, spectra, stack.transforms[start:start + chunk].astype(np.float32),
                  stack.ctf[start:start + chunk].astype(np.float32), np.ones(h, np.float32)]
        acc, den = oracle.ref_insert("ctf", inputs, oracle.Window.linear(), pixel_size=1.5, threads=threads)
        elapsed += time.perf_counter() - t0
        done += chunk
        start += chunk
    t0 = time.perf_counter()
    quotient = np.zeros_like(acc)
    mask = den > 1e-3
    quotient[mask] = acc[mask] / den[mask]
    scipy.fft.irfftn(quotient, s=(big, big, big), workers=threads)
    finish = time.perf_counter() - t0
    rate = done / elapsed
    cores, model = host_cpu()
    return {
        "value": rate,
        "unit": "images/s",
        "cores": threads,
        "kind": "reference",
        "host_cores": cores,
        "host_cpu": model,
        "thread_limit": "the cores this process may use, at most OMP_NUM_THREADS (16 host cores per GPU on the box)",
        "end_to_end_images_per_s": args.images / (args.images / rate + finish),
        "quotient_irfft3_s": finish,
        "sample": (f"{done} images of {n}^2 into one {big}^3 grid (rfft2 by scipy.fft/pocketfft, "
                   f"CTF trilinear insertion by the reference's row kernel, tile-owned over "
                   f"{threads} threads), {elapsed:.1f} s; then quotient + irfft3 of the {big}^3 grid "
                   f"({finish:.2f} s, scipy.fft, {threads} workers)"),
    }


def timed(torch, fn, repeats=1):
    """Wall time of `fn` (ms per call) between device synchronisations, after one untimed call."""
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        fn()
    torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t0) / repeats


def precision_forms(args, stack, transforms, ctf, radial, reference, torch, em, rexlib_amd):
    """The same reconstruction from images staged in HBM as float32, bfloat16 and
    per-image scaled float16 (fused rfft2 + insertion per launch), with the
    minimum FSC to the plain float32 result up to 0.9 Nyquist."""
    from rexlib_amd import precision
    n, big, L = args.n, args.big, args.launch
    window = em.GriddingWindow.linear()
    optics = em.CtfOptics(300.0, 2.7, 0.1)
    out = {}
    for dtype in ("f32", "bf16", "fp16"):
        staged, scales = precision.stage_images(stack.images, dtype)
        num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
        den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")

        def run():
            num.zero_()
            den.zero_()
            for s in range(0, args.images, L):
                e = min(s + L, args.images)
                em.insert_ctf_images(num, den, staged[s:e], transforms[s:e], ctf[s:e], radial, window, optics, 1.5,
                                     image_scales=None if scales is None else scales[s:e], padded=n)
            volume = rexlib_amd.irfft3(em.normalize(num, den, 1e-3))
            return em.gridding_correct(volume, window, out=volume)

        ms = timed(torch, run)
        report = precision.shell_report(num.contiguous(), reference, 0.9)
        out[dtype] = {"images_per_s": args.images / (ms * 1e-3), "ms": ms,
                      "min_fsc_to_fp32_to_0.9_nyquist": report["min_fsc_to_cutoff"],
                      "staged_bytes_per_image": n * n * staged.element_size() + (4 if scales is not None else 0)}
        del staged, scales, num, den
    log(f"precision forms: {out}")
    return out


def learned_form(den, transforms, big, torch):
    """Density compensation of the benchmark's weight volume: the 10-pass
    iterative estimate against the perceptron's weights (coverage table of the
    transforms, features, MFMA forward over every voxel, exp / den). The
    shipped model is used when present; the timing does not depend on the
    parameter values."""
    from rexlib_amd import learned
    model = learned.shipped_model() or learned.CompensationModel(seed=0)
    voxels = den.numel()
    iterative_ms = timed(torch, lambda: learned.iterative_weights(den), repeats=3)
    table_ms = timed(torch, lambda: learned.coverage_table(transforms, big), repeats=3)
    table = learned.coverage_table(transforms, big)
    learned_ms = timed(torch, lambda: model.weights(den, table), repeats=3)
    rows = learned.features(den, table).reshape(-1, learned.FEATURES)
    forward_ms = timed(torch, lambda: model.predict(rows), repeats=3)
    flops = 2.0 * (10 * 64 + 64 * 64 + 64 * 64 + 64 * 1) * voxels
    out = {"voxels": voxels, "iterative_ms": iterative_ms, "coverage_table_ms": table_ms, "learned_ms": learned_ms,
           "mlp_forward_ms": forward_ms, "mlp_forward_tflops": flops / (forward_ms * 1e-3) / 1e12,
           "mlp_forward_frac_of_fp16_dense_peak": flops / (forward_ms * 1e-3) / 2.5e15}
    log(f"learned form: {out}")
    return out


def event_ms(torch, fn, repeats=5):
    """Milliseconds per call from HIP events on torch's current stream (the
    stream the library's calls run on), after one untimed call."""
    fn()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(repeats):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / repeats


def widened_rows(torch, em):
    """The rows beside the hot path, each timed with HIP events against the
    HBM roofline of its compulsory bytes (every operand read once, every
    result written once):

    * background standardization of 4,096 float32 images of 256^2 (radius 100,
      bounds +-3, one Philox key per image);
    * complex FFTs over the last two axes of 1,024 complex64 256^2 images, and
      over the first axis (strided) of a 256 x 256 x 256 complex64 volume;
    * a real FFT of 4,096 float32 rows of length 1,000 (2^3 5^3, Stockham
      stages) and of length 1,009 (a prime: Bluestein);
    * the quotient where(den > 1e-3, num / den, 0) over the 256^3 accumulators
      by the library's elementwise divide, greater and where (the plugin's
      builders), against the fused normalize kernel;
    * the insertion builders with the other windows and weights (images/s)."""
    from rexlib_amd import fourier, functional
    out = {}

    def entry(name, ms, nbytes):
        out[name] = {"ms": ms, "bytes": nbytes, "gb_per_s": nbytes / (ms * 1e-3) / 1e9,
                     "frac_of_hbm": nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9)}

    g = torch.Generator(device="cuda")
    g.manual_seed(17)
    images = torch.randn((4096, 256, 256), device="cuda", generator=g)
    keys = torch.arange(4096, dtype=torch.int64, device="cuda")
    result = torch.empty_like(images)
    entry("standardize_background_4096x256^2",
          event_ms(torch, lambda: em.standardize_background(images, keys, 100.0, -3.0, 3.0, out=result)),
          2 * images.numel() * 4 + keys.numel() * 8)
    del images, result
    stack = torch.randn((1024, 256, 256), dtype=torch.complex64, device="cuda", generator=g)
    spectra = torch.empty_like(stack)
    entry("fft_last_two_axes_1024x256^2_c64", event_ms(torch, lambda: fourier.fftn(stack, axes=(-2, -1), out=spectra)),
          2 * stack.numel() * 8)
    del stack, spectra
    volume = torch.randn((256, 256, 256), dtype=torch.complex64, device="cuda", generator=g)
    transformed = torch.empty_like(volume)
    entry("fft_first_axis_256^3_c64", event_ms(torch, lambda: fourier.fftn(volume, axes=(0,), out=transformed)),
          2 * volume.numel() * 8)
    del volume, transformed
    for length in (1000, 1009):
        rows = torch.randn((4096, length), device="cuda", generator=g)
        half = torch.empty((4096, length // 2 + 1), dtype=torch.complex64, device="cuda")
        entry(f"rfft_4096_rows_of_{length}", event_ms(torch, lambda: fourier.rfftn(rows, axes=(-1,), out=half)),
              rows.numel() * 4 + half.numel() * 8)
        del rows, half
    num = torch.randn((256, 256, 129), dtype=torch.complex64, device="cuda", generator=g)
    den = torch.rand((256, 256, 129), device="cuda", generator=g)
    den_c = den.to(torch.complex64)
    zero = torch.zeros((), dtype=torch.complex64, device="cuda")
    quotient = torch.empty_like(num)
    ratio = torch.empty_like(num)
    mask = torch.empty(den.shape, dtype=torch.bool, device="cuda")

    def library_quotient():
        functional.greater(den, 1e-3, out=mask)
        functional.divide(num, den_c, out=ratio)
        functional.where(mask, ratio, zero, out=quotient)

    voxels = num.numel()
    entry("quotient_by_divide_greater_where_256^3", event_ms(torch, library_quotient),
          voxels * (8 + 8 + 4 + 8))
    entry("quotient_by_normalize_256^3", event_ms(torch, lambda: em.normalize(num, den, 1e-3, out=quotient)),
          voxels * (8 + 4 + 8))
    del num, den, den_c, quotient, ratio, mask

    # The insertion builders beside the fused CTF path: 4,096 slices of
    # 256 x 129 spectra into 256^3 accumulators per call, uniform views,
    # by insert_slices (trilinear), insert_weighted_slices (trilinear),
    # insert_ctf_slices with the nearest window and with a Kaiser-Bessel
    # window of half-width 2 (shape 7.4). Images/s per call.
    from rexlib_amd import synthetic
    count, n = 4096, 256
    rng = np.random.default_rng(29)
    transforms = torch.from_numpy(synthetic.transforms_of(synthetic.uniform_rotations(count, rng)).astype(
        np.float32)).cuda()
    ctf = torch.from_numpy(synthetic.ctf_parameters(count, rng).astype(np.float32)).cuda()
    slices = torch.randn((count, n, n // 2 + 1), dtype=torch.complex64, device="cuda", generator=g)
    weights = torch.rand((count, n, n // 2 + 1), device="cuda", generator=g)
    radial = torch.ones(n // 2 + 1, device="cuda")
    num = torch.zeros((n, n, n // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((n, n, n // 2 + 1), device="cuda")
    optics = em.CtfOptics(300.0, 2.7, 0.1)
    linear = em.GriddingWindow.linear()
    calls = {
        "insert_slices_linear": lambda: em.insert_slices(num, den, slices, transforms, radial, linear,
                                                         numerator_out=num, denominator_out=den),
        "insert_weighted_slices_linear": lambda: em.insert_weighted_slices(
            num, den, slices, weights, transforms, radial, linear, numerator_out=num, denominator_out=den),
        "insert_ctf_slices_nearest": lambda: em.insert_ctf_slices(
            num, den, slices, transforms, ctf, radial, em.GriddingWindow.nearest(), optics, 1.0,
            numerator_out=num, denominator_out=den),
        "insert_ctf_slices_kaiser_bessel_2": lambda: em.insert_ctf_slices(
            num, den, slices, transforms, ctf, radial, em.GriddingWindow.kaiser_bessel(2.0, 7.4), optics, 1.0,
            numerator_out=num, denominator_out=den),
    }
    for name, call in calls.items():
        ms = event_ms(torch, call, repeats=3)
        out[name] = {"ms": ms, "images_per_s": count / (ms * 1e-3)}
    del slices, weights, num, den
    torch.cuda.empty_cache()
    log(f"widened rows: {out}")
    return out


def mlp_profile(torch, big=256, rows=65536, seed=11):
    """The perceptron alone: features of a random (big, big, big/2+1) weight
    volume with a coverage table of 50,000 uniform orientations, the forward
    pass over every voxel three times, and three training steps (forward,
    backward, Adam) on `rows` of those features."""
    from rexlib_amd import learned, synthetic
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    den = torch.rand((big, big, big // 2 + 1), device="cuda", generator=g) + 0.5
    rot = synthetic.uniform_rotations(50000, np.random.default_rng(seed))
    table = learned.coverage_table(torch.from_numpy(synthetic.transforms_of(rot).astype(np.float32)).cuda(), big)
    f = learned.features(den, table).reshape(-1, learned.FEATURES)
    model = learned.CompensationModel(seed=seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        out = model.predict(f)
    torch.cuda.synchronize()
    forward_ms = 1e3 * (time.perf_counter() - t0) / 3
    target = torch.zeros(rows, device="cuda")
    batch = f[:rows].contiguous()
    for _ in range(3):
        model.train_step(batch, target)
    torch.cuda.synchronize()
    flops = 2.0 * (10 * 64 + 64 * 64 + 64 * 64 + 64 * 1) * f.shape[0]
    return {"mlp_profile": {"voxels": int(f.shape[0]), "forward_ms": forward_ms,
                            "forward_tflops": flops / (forward_ms * 1e-3) / 1e12, "train_rows": rows,
                            "finite": bool(torch.isfinite(out).all())}}


def learned_config(torch, em, rexlib_amd, images=80000, launch=4096, seed=7000):
    """The learned-weight form at its own shape: 256^3 from `images`
    projections of 256^2 of a 400-blob volume (CTF defocus 1-3 um, SNR 0.1),
    80% of the views von Mises-Fisher around the top view with kappa drawn
    uniform in [2, 20], 20% uniform on SO(3). Two paths over the same images:

    * plain: float32 images, one eager fused rfft2 + insertion call per
      launch, the iterative density compensation (10 passes), irfft3 of
      numerator * W, gridding correction;
    * combined: the images staged in HBM as bfloat16, the zeroing and every
      launch captured in one HIP graph and replayed, the shipped perceptron's
      weights (``learned.Compensator``: judged once against the iterative
      estimate, then reused), irfft3, gridding correction.

    Reports images/s of both (and of the plain quotient instead of the
    iterative compensation), the judge's FSC, and the FSC of the combined
    spectrum to the plain iterative one up to 0.9 Nyquist."""
    from rexlib_amd import learned, precision, synthetic
    n = big = 256
    rng = np.random.default_rng(seed)
    kappa = float(rng.uniform(2.0, 20.0))
    k = int(0.8 * images)
    rot = np.concatenate([synthetic.von_mises_fisher_rotations(k, kappa, rng),
                          synthetic.uniform_rotations(images - k, rng)])
    rot = rot[rng.permutation(images)]
    t0 = time.time()
    vol = synthetic.blob_volume(big, 400, big * 0.375, (2.0, 8.0), seed=seed)
    stack = synthetic.particle_stack(vol, images, n, seed=seed + 1, snr=0.1, rotations=rot, device="cuda")
    log(f"learned form: generated {images} images of {n}^2 (kappa {kappa:.2f}) in {time.time() - t0:.1f} s")
    transforms = torch.from_numpy(stack.transforms.astype(np.float32)).cuda()
    ctf = torch.from_numpy(stack.ctf.astype(np.float32)).cuda()
    radial = torch.ones(n // 2 + 1, device="cuda")
    window = em.GriddingWindow.linear()
    optics = em.CtfOptics(300.0, 2.7, 0.1)
    launches = [(s, min(s + launch, images)) for s in range(0, images, launch)]
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")

    def insert(staged, a, b):
        a.zero_()
        b.zero_()
        for s, e in launches:
            em.insert_ctf_images(a, b, staged[s:e], transforms[s:e], ctf[s:e], radial, window, optics, 1.5)

    def plain():
        insert(stack.images, num, den)
        w = learned.iterative_weights(den)
        return learned.reconstruct(num, w, window)

    def plain_quotient():
        insert(stack.images, num, den)
        volume = rexlib_amd.irfft3(em.normalize(num, den, 1e-3))
        return em.gridding_correct(volume, window, out=volume)

    plain_ms = timed(torch, plain)
    quotient_ms = timed(torch, plain_quotient)
    insert(stack.images, num, den)
    reference = num * learned.iterative_weights(den)
    staged, _ = precision.stage_images(stack.images, "bf16")
    del stack
    torch.cuda.empty_cache()
    num_c = torch.zeros_like(num)
    den_c = torch.zeros_like(den)
    capture = torch.cuda.Stream()
    capture.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(capture):
        insert(staged, num_c, den_c)  # eager pass on the capture stream: sizes its workspace
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=capture):
        insert(staged, num_c, den_c)
    torch.cuda.synchronize()
    model = learned.shipped_model()
    t1 = time.perf_counter()
    compensator = learned.Compensator(model, transforms, big)
    graph.replay()
    weights, judged = compensator(num_c, den_c)
    torch.cuda.synchronize()
    judged_ms = 1e3 * (time.perf_counter() - t1)

    def combined():
        graph.replay()
        w, _ = compensator(num_c, den_c)
        return learned.reconstruct(num_c, w, window)

    combined_ms = timed(torch, combined)
    graph.replay()
    weights, steady = compensator(num_c, den_c)
    agreement = precision.shell_report((num_c * weights).contiguous(), reference.contiguous(), 0.9)
    out = {"images": images, "n": n, "volume": f"{big}^3", "kappa": kappa, "preferred_fraction": 0.8,
           "launch": launch, "plain_fp32_iterative_ms": plain_ms, "plain_fp32_quotient_ms": quotient_ms,
           "combined_ms": combined_ms, "combined_first_call_with_judge_ms": judged_ms,
           "plain_images_per_s": images / (plain_ms * 1e-3),
           "plain_quotient_images_per_s": images / (quotient_ms * 1e-3),
           "combined_images_per_s": images / (combined_ms * 1e-3),
           "speedup_vs_plain_iterative": plain_ms / combined_ms,
           "speedup_vs_plain_quotient": quotient_ms / combined_ms,
           "weights_used": steady["used"], "judge": {k: v for k, v in judged.items()},
           "min_fsc_combined_to_plain_iterative_to_0.9_nyquist": agreement["min_fsc_to_cutoff"],
           "max_shell_error_combined_to_plain_iterative": agreement["max_error_to_cutoff"],
           "band_error_combined_to_plain_iterative": agreement["relative_error_to_cutoff"],
           "model": "shipped" if model is not None else "absent (iterative weights)"}
    log(f"learned form at its config: {out}")
    del graph, staged, num, den, num_c, den_c, reference, weights, compensator
    torch.cuda.empty_cache()
    return out


def precision_config(torch, em, rexlib_amd, images=200000, launch=4096, sample=256, seed=3000):
    """The precision form at its own shape: `images` projections of 128^2
    from a 128^3 blob volume (defocus 0.5-4 um, SNR 0.1), staged in HBM as
    float32, bfloat16 and per-image scaled float16, padded on chip to 256^2
    and inserted into 256^3 in launches of `launch`, then quotient, irfft3,
    gridding correction and the crop of the central 128^3. The per-shell FSC
    and relative error are taken on the first `sample` images against the
    float64 oracle (pocketfft rfft2 of the padded float64 images, double
    insertion), up to 0.9 Nyquist."""
    import oracle
    import scipy.fft
    from rexlib_amd import precision, synthetic
    m, n = 128, 256
    t0 = time.time()
    vol = synthetic.blob_volume(m, 120, 48.0, (1.5, 5.0), seed=seed)
    stack = synthetic.particle_stack(vol, images, m, seed=seed + 1, snr=0.1, pixel_size=1.5,
                                     defocus=(5000.0, 40000.0), device="cuda")
    transforms = torch.from_numpy(stack.transforms.astype(np.float32)).cuda()
    ctf = torch.from_numpy(stack.ctf.astype(np.float32)).cuda()
    log(f"precision form: generated {images} images of {m}^2 in {time.time() - t0:.1f} s")
    radial = torch.ones(n // 2 + 1, device="cuda")
    window = em.GriddingWindow.linear()
    optics = em.CtfOptics(300.0, 2.7, 0.1)
    # The float64 oracle on the sample.
    host = stack.images[:sample].cpu().numpy().astype(np.float64)
    padded = np.zeros((sample, n, n))
    o = n // 2 - m // 2
    padded[:, o:o + m, o:o + m] = host
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    ref = oracle.insert_f64("ctf", [acc, acc.real.copy(), scipy.fft.rfft2(padded, workers=min(16, os.cpu_count() or 1)),
                                    stack.transforms[:sample], stack.ctf[:sample], np.ones(n // 2 + 1)],
                            oracle.Window.linear(), pixel_size=1.5)
    ref_q = torch.from_numpy(oracle.normalize(ref[0], ref[1]).astype(np.complex64)).cuda()
    out = {"images": images, "image_extent": m, "padded_extent": n, "volume": f"{n}^3", "sample_for_fsc": sample}
    for dtype in ("f32", "bf16", "fp16"):
        staged, scales = precision.stage_images(stack.images, dtype)
        num = torch.zeros((n, n, n // 2 + 1), dtype=torch.complex64, device="cuda")
        den = torch.zeros((n, n, n // 2 + 1), dtype=torch.float32, device="cuda")

        def insert(count):
            num.zero_()
            den.zero_()
            for s in range(0, count, launch):
                e = min(s + launch, count)
                em.insert_ctf_images(num, den, staged[s:e], transforms[s:e], ctf[s:e], radial, window, optics, 1.5,
                                     image_scales=None if scales is None else scales[s:e], padded=n)

        def run():
            insert(images)
            volume = rexlib_amd.irfft3(em.normalize(num, den, 1e-3))
            return em.gridding_correct(volume, window, crop=m)

        ms = timed(torch, run)
        insert(sample)
        report = precision.shell_report(em.normalize(num, den, 1e-3), ref_q, 0.9)
        out[dtype] = {"images_per_s": images / (ms * 1e-3), "ms": ms,
                      "min_fsc_to_fp64_to_0.9_nyquist": report["min_fsc_to_cutoff"],
                      "max_shell_error_to_0.9_nyquist": report["max_error_to_cutoff"],
                      "staged_bytes_per_image": m * m * staged.element_size() + (4 if scales is not None else 0)}
        del staged, scales, num, den
    for dtype in ("bf16", "fp16"):
        out[dtype]["vs_f32"] = out[dtype]["images_per_s"] / out["f32"]["images_per_s"]
    log(f"precision form at its config: {out}")
    del stack, transforms, ctf
    torch.cuda.empty_cache()
    return out


def batched_config(torch, count=64, seed=5000, low=500, high=40000):
    """The batched form at its own shape: `count` reconstructions, box sizes
    drawn from {64, 96, 128, 192, 256, 320}, particle counts log-uniform in
    [low, high], each with its own blob volume, CTF and SNR (0.05-0.5), float32
    images; one HIP graph over four streams (autotuned brick per box size)
    against the same reconstructions run one after another."""
    from rexlib_amd import batched, synthetic
    rng = np.random.default_rng(seed)
    t0 = time.time()
    recons = []
    images = 0
    sizes = []
    for i in range(count):
        n = int(rng.choice([64, 96, 128, 192, 256, 320]))
        particles = int(np.exp(rng.uniform(np.log(low), np.log(high))))
        vol = synthetic.blob_volume(n, 40, n * 0.35, (1.0, 0.02 * n + 1.0), seed=seed + 10 * i)
        stack = synthetic.particle_stack(vol, particles, n, seed=seed + 10 * i + 1, snr=float(rng.uniform(0.05, 0.5)),
                                         device="cuda")
        recons.append(batched.Reconstruction(stack.images, torch.from_numpy(stack.transforms.astype(np.float32)).cuda(),
                                             torch.from_numpy(stack.ctf.astype(np.float32)).cuda()))
        images += particles
        sizes.append(n)
    log(f"batched form: generated {count} reconstructions, {images} images, in {time.time() - t0:.1f} s")
    sequential_ms = timed(torch, lambda: batched.run_sequential(recons))
    graph = batched.BatchedReconstruction(recons, streams=4).capture()
    graph_ms = timed(torch, graph.replay, repeats=2)
    out = {"reconstructions": count, "images": images, "box_sizes": {str(s): sizes.count(s) for s in sorted(set(sizes))},
           "sequential_ms": sequential_ms, "graph_ms": graph_ms, "speedup": sequential_ms / graph_ms,
           "sequential_images_per_s": images / (sequential_ms * 1e-3),
           "graph_images_per_s": images / (graph_ms * 1e-3)}
    log(f"batched form at its config: {out}")
    del recons, graph
    torch.cuda.empty_cache()
    return out


def plugin_path(args):
    """The rexlib call sequence through the plugin itself, as a rexlib caller
    would run it: tests/plugin_driver (a C++ rexlib caller) loads
    librexlib-hip.so and times rfft2 + insert_ctf_slices per launch, then
    irfft3, over the same image count and launch size (256^2 images into
    256^3; its own random images and orientations). None for other shapes
    or when the driver is not built."""
    import subprocess
    driver = os.path.join(ROOT, "tests", "plugin_driver")
    plugin = os.path.join(ROOT, "rexlib_amd", "rexlib-plugins", "librexlib-hip.so")
    if args.n != 256 or args.big != 256 or not (os.path.exists(driver) and os.path.exists(plugin)):
        return None
    proc = subprocess.run([driver, "bench", plugin, str(args.images), str(args.launch), "2"], capture_output=True,
                          text=True, timeout=600)
    lines = [line for line in proc.stdout.splitlines() if line.startswith("{")]
    if proc.returncode != 0 or not lines:
        log(f"plugin path failed ({proc.returncode}): {proc.stdout[-400:]} {proc.stderr[-400:]}")
        return None
    out = json.loads(lines[-1])
    out["calls"] = "through librexlib-hip.so: rfft2 + insert_ctf_slices per launch, then irfft3"
    log(f"plugin path: {out}")
    return out


def parity_sample(args, stack, torch, em, rexlib_amd, count=256):
    """rel-L2 of the GPU accumulators (fused rfft2 + insertion, as timed) on the
    first `count` images against the reference's CPU path: pocketfft rfft2 in
    float32 (scipy.fft) and the reference's float32 insertion kernels."""
    import oracle
    import scipy.fft
    n, big = args.n, args.big
    h = n // 2 + 1
    images = stack.images[:count]
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    em.insert_ctf_images(num, den, images, torch.from_numpy(stack.transforms[:count].astype(np.float32)).cuda(),
                         torch.from_numpy(stack.ctf[:count].astype(np.float32)).cuda(),
                         torch.ones(h, device="cuda"), em.GriddingWindow.linear(), em.CtfOptics(300, 2.7, 0.1), 1.5)
    torch.cuda.synchronize()
    acc = np.zeros((big, big, big // 2 + 1), np.complex64)
    threads = min(16, os.cpu_count() or 1)
    host_spectra = scipy.fft.rfft2(images.cpu().numpy(), workers=threads).astype(np.complex64)
    inputs = [acc, acc.real.copy(), host_spectra, stack.transforms[:count].astype(np.float32),
              stack.ctf[:count].astype(np.float32), np.ones(h, np.float32)]
    if oracle.ref_available():
        ref = oracle.ref_insert("ctf", inputs, oracle.Window.linear(), pixel_size=1.5, threads=threads)
        against = "pocketfft rfft2 (float32) + the reference's insertion kernels (float32)"
    else:
        ref = oracle.insert_f64("ctf", inputs, oracle.Window.linear(), pixel_size=1.5)
        against = "pocketfft rfft2 (float32) + the double-precision oracle"
    return {
        "images": count,
        "against": against,
        "rel_l2_numerator": oracle.rel_l2(num.cpu().numpy(), ref[0]),
        "rel_l2_denominator": oracle.rel_l2(den.cpu().numpy(), ref[1]),
    }


def main():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=3)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--images", type=int, default=50000)
    parser.add_argument("--launch", type=int, default=4096)
    parser.add_argument("--n", type=int, default=256)
    parser.add_argument("--big", type=int, default=256)
    parser.add_argument("--blobs", type=int, default=400)
    parser.add_argument("--cpu-seconds", type=float, default=12.0)
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-parity", action="store_true")
    parser.add_argument("--no-forms", action="store_true",
                        help="skip the precision, learned, batched and plugin-path legs reported beside the "
                             "headline value")
    parser.add_argument("--no-config-forms", action="store_true",
                        help="skip the precision (200,000 128^2 images), batched (64 reconstructions) and learned-weight "
                             "(80,000 256^2 images) legs")
    parser.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                        help="before the workload, copy a float32 buffer of MIB MiB three times with an elementwise "
                             "kernel (known traffic, to scale hardware byte counters)")
    parser.add_argument("--widened-rows", action="store_true",
                        help="only the rows beside the hot path (background standardization, general FFTs, the "
                             "elementwise quotient), timed against HBM, then exit")
    parser.add_argument("--mlp-profile", action="store_true",
                        help="only the perceptron: features of a random 256^3 weight volume, the forward pass over "
                             "every voxel three times and three training steps on 65,536 rows (the workload of the "
                             "MFMA profile), then exit")
    args = parser.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path would shard by "
              "images with one all-reduce of the accumulators; see the design notes).", file=sys.stderr)
        sys.exit(2)

    import torch

    import rexlib_amd
    from rexlib_amd import _native, em, tune

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.mlp_profile:
        print(json.dumps(mlp_profile(torch)), flush=True)
        return
    if args.widened_rows:
        tune.apply_all()
        print(json.dumps({"widened_rows": widened_rows(torch, em)}), flush=True)
        return
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration * (1 << 18), device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        log(f"copy calibration: 3 copies of {args.copy_calibration} MiB")
        del src, dst
    stack = make_workload(args, torch)
    n, big, L = args.n, args.big, args.launch
    h = n // 2 + 1
    transforms = torch.from_numpy(stack.transforms.astype(np.float32)).cuda()
    ctf = torch.from_numpy(stack.ctf.astype(np.float32)).cuda()
    radial = torch.ones(h, device="cuda")
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    spectra = torch.empty((L, n, h), dtype=torch.complex64, device="cuda")
    window = em.GriddingWindow.linear()
    optics = em.CtfOptics(300.0, 2.7, 0.1)
    launches = [(s, min(s + L, args.images)) for s in range(0, args.images, L)]
    # The autotuner's table (rexlib_amd/tuning.json); the library picks the
    # entry of this box size.
    tune.apply_all()
    tuned = tune.get_tuning(big)
    log(f"brick tuning ((extent x, extent y, extent z), share): {tuned}")

    def step():
        num.zero_()
        den.zero_()
        for s, e in launches:
            em.insert_ctf_images(num, den, stack.images[s:e], transforms[s:e], ctf[s:e], radial, window, optics, 1.5)
        volume = rexlib_amd.irfft3(em.normalize(num, den, 1e-3))
        return em.gridding_correct(volume, window, out=volume)

    def api_step():
        # The rexlib call sequence: rfft2 of each launch into a spectra
        # buffer, then insert_ctf_slices accumulating in place.
        num.zero_()
        den.zero_()
        for s, e in launches:
            out = spectra[: e - s]
            rexlib_amd.rfft2(stack.images[s:e], out=out)
            em.insert_ctf_slices(num, den, out, transforms[s:e], ctf[s:e], radial, window, optics, 1.5,
                                 numerator_out=num, denominator_out=den)
        volume = rexlib_amd.irfft3(em.normalize(num, den, 1e-3))
        return em.gridding_correct(volume, window, out=volume)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # The library brackets each launch of its insertion kernels with HIP
    # events on the stream they run on (torch's current stream here).
    _native.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    brick_ms, brick_launches = _native.kernel_timing(_native.KERNEL_BRICK)
    prepare_ms, prepare_launches = _native.kernel_timing(_native.KERNEL_PREPARE)
    _native.set_kernel_timing(False)
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.images * args.steps / elapsed
    log(f"{args.steps} steps of {args.images} images: {elapsed:.3f} s, {value:,.0f} images/s")
    api_ms = timed(torch, api_step)
    two_call = {"images_per_s": args.images / (api_ms * 1e-3), "ms_per_step": api_ms,
                "calls": "rfft2 + insert_ctf_slices per launch"}
    log(f"two-call API: {two_call}")
    plugin = None if args.no_forms else plugin_path(args)

    # Average over launches; the final partial launch holds images % L images.
    images_timed = args.images * args.steps
    b_img = algorithmic_bytes_per_image(n, big, L)
    avg_ms = brick_ms / brick_launches if brick_launches else float("nan")
    images_per_launch = images_timed / brick_launches if brick_launches else float("nan")
    achieved = b_img * images_timed / (brick_ms * 1e-3) / 1e9 if brick_launches else None
    roofline = {
        "bound": "hbm",
        "kernel": "brick_kernel (tile-owned LDS brick gather, CTF + trilinear)",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS if achieved else None,
        "traffic": traffic_per_launch(args, tuned),
        "launch_ms": avg_ms,
        "launches": brick_launches,
        "images_per_launch": images_per_launch,
        "bytes_per_image": b_img,
        "column_prepare_kernel_ms": prepare_ms / prepare_launches if prepare_launches else None,
        "pipeline_frac": value * b_img / (HBM_PEAK_GBS * 1e9),
    }

    forms = None
    if not args.no_forms:
        step()
        torch.cuda.synchronize()
        reference = num.clone()
        forms = {
            "precision_on_this_workload": precision_forms(args, stack, transforms, ctf, radial, reference, torch, em,
                                                          rexlib_amd),
            "learned": learned_form(den, transforms, big, torch),
        }
        forms["widened_rows"] = widened_rows(torch, em)
        if not args.no_config_forms:
            forms["precision"] = precision_config(torch, em, rexlib_amd)
            forms["batched"] = batched_config(torch)
            forms["learned_weights"] = learned_config(torch, em, rexlib_amd)
        del reference

    parity = None if args.no_parity else parity_sample(args, stack, torch, em, rexlib_amd)
    if parity:
        log(f"parity: {parity}")
    threads = cpu_share()
    baseline = None if args.no_cpu_baseline else cpu_baseline(args, stack, threads)
    if baseline:
        log(f"cpu baseline: {baseline}")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (blob volume, exact projections, CTF, SNR 0.1)",
        "config": {
            "workload": (f"plain FP32: {args.images} images {n}^2 -> {big}^3, launches of {L}, fused rfft2 + "
                         "CTF trilinear insertion, then quotient, irfft3, gridding correction"),
            "images": args.images, "n": n, "N": big, "launch": L, "window": "linear", "variant": "ctf",
            "brick": {"extent_x": tuned[0][0], "extent_y": tuned[0][1], "extent_z": tuned[0][2],
                      "share": tuned[1]} if tuned else None,
        },
        "roofline": roofline,
        "cpu_baseline": baseline,
        "parity": parity,
        "two_call_api": two_call,
        "plugin_path": plugin,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
