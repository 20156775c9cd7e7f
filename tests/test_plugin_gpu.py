"""The plugin boundary on the GPU: a C++ caller of the mirrored rexlib API
(tests/plugin_driver.cpp) loads librexlib-hip.so, opens hip:0, and runs
rfft2, insert_ctf_slices in two calls that accumulate in place, irfft3, and
the quotient as where(greater(den, 1e-3), divide(num, den), 0), with no HIP
code of its own. Results are checked against numpy's pocketfft
and the double-precision oracle, within the reference's float32 tolerance
(rel-L2 1e-4, tests/unitary/src/backends/cpu/builders/em/reconstruction/
test_slice_insertion_program_builder.cpp:190). Errors cross the boundary as
the reference's exceptions."""
import os

import numpy as np
import pytest

import oracle
from rexlib_amd import synthetic
from test_plugin import DRIVER, PLUGIN, driver

pytestmark = pytest.mark.gpu

WINDOWS = {0: oracle.Window.nearest(), 1: oracle.Window.linear(), 2: oracle.Window.kaiser_bessel(1.5, 6.0)}


@pytest.mark.parametrize("profile", [1, 0, 2])
def test_reconstruction_through_the_plugin(tmp_path, profile):
    rng = np.random.default_rng(40 + profile)
    count, n, big, split = 24, 32, 48, 10
    h = n // 2 + 1
    images = rng.normal(size=(count, n, n)).astype(np.float32)
    transforms = synthetic.transforms_of(synthetic.uniform_rotations(count, rng)).astype(np.float32)
    transforms[:, :2, 3] = rng.uniform(-2, 2, (count, 2))
    ctf = np.stack([rng.uniform(1e4, 3e4, count), rng.uniform(1e4, 3e4, count), rng.uniform(0, np.pi, count),
                    rng.uniform(0, 0.5, count)], axis=1).astype(np.float32)
    radial = rng.uniform(0.2, 2.0, h).astype(np.float32)
    for name, a in [("images", images), ("transforms", transforms), ("ctf", ctf), ("radial", radial)]:
        a.tofile(tmp_path / f"{name}.bin")
    w = WINDOWS[profile]
    meta = dict(count=count, n=n, big=big, split=split, window=profile, half_width=w.half_width, shape=w.shape,
                voltage=300, cs=2.7, q=0.1, pixel=1.5)
    (tmp_path / "meta.txt").write_text("".join(f"{k} {v}\n" for k, v in meta.items()))
    print(f"plugin run: window {profile}, {count} images {n}^2 -> {big}^3", flush=True)
    rc, lines, err = driver("run", PLUGIN, str(tmp_path))
    assert rc == 0, (lines, err)
    assert lines[-1]["run"] == "ok"

    spectra = np.fromfile(tmp_path / "spectra_out.bin", np.complex64).reshape(count, n, h)
    want = np.fft.rfft2(images.astype(np.float64))
    assert oracle.rel_l2(spectra, want) < 1e-5

    hb = big // 2 + 1
    num = np.fromfile(tmp_path / "num_out.bin", np.complex64).reshape(big, big, hb)
    den = np.fromfile(tmp_path / "den_out.bin", np.float32).reshape(big, big, hb)
    acc = np.zeros((big, big, hb), np.complex128)
    ref = oracle.insert_f64("ctf", [acc, acc.real.copy(), want, transforms.astype(np.float64),
                                    ctf.astype(np.float64), radial.astype(np.float64)], w, pixel_size=1.5)
    assert oracle.rel_l2(num, ref[0]) <= 1e-4
    assert oracle.rel_l2(den, ref[1]) <= 1e-4

    standardized = np.fromfile(tmp_path / "standardized_out.bin", np.float32).reshape(count, n, n)
    want_s = oracle.standardize_background(images, np.arange(1000, 1000 + count, dtype=np.uint64), n / 2 - 2,
                                           -3.0, 3.0)
    assert np.max(np.abs(standardized - want_s)) <= 1e-5

    quotient = np.fromfile(tmp_path / "quotient_out.bin", np.complex64).reshape(big, big, hb)
    want_q = oracle.normalize(num.astype(np.complex128), den.astype(np.float64), 1e-3)
    assert np.max(np.abs(quotient - want_q)) <= 1e-6 * np.max(np.abs(want_q))

    volume = np.fromfile(tmp_path / "volume_out.bin", np.float32).reshape(big, big, big)
    assert oracle.rel_l2(volume, np.fft.irfftn(num.astype(np.complex128), s=(big, big, big), axes=(0, 1, 2))) < 1e-4


def test_errors_cross_the_boundary_as_the_reference_exceptions():
    rc, lines, err = driver("errors", PLUGIN)
    assert rc == 0, (lines, err)
    e = lines[-1]
    assert e["valid"] == "ok"
    assert e["radial_extent"] == "invalid_argument"
    assert e["transform_shape"] == "invalid_argument"
    # float64 and host memory have no HIP program: the service finds no builder.
    assert e["float64"] == "unsupported_operation_error"
    assert e["host_operand"] == "unsupported_operation_error"
    assert e["new_outputs"] == "ok"


def test_benchmark_sequence_through_the_plugin():
    # The timed call sequence of bench.py's plugin-path leg, at a small size:
    # rfft2 + insert_ctf_slices per launch (a partial last launch included),
    # then irfft3, all dispatched through librexlib-hip.so.
    rc, lines, err = driver("bench", PLUGIN, "600", "256", "1")
    assert rc == 0, (lines, err)
    record = lines[-1]
    assert record["images"] == 600 and record["launch"] == 256 and record["steps"] == 1
    assert record["images_per_s"] > 0 and record["ms_per_step"] > 0
