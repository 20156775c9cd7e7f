"""The oracle, pinned before it is trusted (CPU only).

Anchors, each chosen so that a plausible slip (a dropped term, a wrong sign
or index, a transposed operand) fails one of them:
  * known answers copied as numbers from the reference's tests;
  * the reference's own CPU kernels, compiled from its sources into
    oracle/_ref (when built) and, everywhere, the golden vectors they wrote;
  * closed forms: the central plane of numpy.fft.rfftn, Hermitian symmetry.
"""
import math
import os

import numpy as np
import pytest

import oracle
from cases import PIXEL_SIZE, SIZES, identity_transform, make_case, spread_kat_case

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "insertion_reference.npz")
WINDOWS = {
    "nearest": oracle.Window.nearest(),
    "linear": oracle.Window.linear(),
    "kb1.5_6": oracle.Window.kaiser_bessel(1.5, 6.0),
    "kb2_8": oracle.Window.kaiser_bessel(2.0, 8.0),
}
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


def tolerance_for(window):
    # slice_insertion_cases.hpp:648-654: 1e-9 for box and triangle, 5e-4 for KB.
    return 5e-4 if window.profile == "kaiser_bessel" else 1e-9


# --- known answers from the reference's tests --------------------------------

def test_wavelength_known_values():
    # tests/unitary/src/em/ops/parameters/test_ctf_optics.cpp:40-53
    assert oracle.wavelength(300) == pytest.approx(0.019687615309, rel=1e-9)
    assert oracle.wavelength(200) == pytest.approx(0.025079493567, rel=1e-9)


def test_bessel_i0_known_values():
    # tests/unitary/src/backends/cpu/kernels/test_bessel_i0.cpp:11-24
    assert oracle.bessel_i0(0.0) == 1.0
    for x, v, rel in [(0.5, 1.0634833707413236, 1e-14), (1.0, 1.2660658777520082, 1e-14),
                      (2.0, 2.279585302336067, 1e-14), (5.0, 27.23987182360445, 1e-14),
                      (10.0, 2815.716628466255, 1e-13)]:
        assert oracle.bessel_i0(x) == pytest.approx(v, rel=rel)


def test_spread_of_one_sample_over_eight_voxels():
    # test_row_insertion_kernel.cpp:47-112: a sample at (1.25, 2.5, -0.75), value 2.
    num, den = oracle.insert_f64("weighted", spread_kat_case(), oracle.Window.linear())
    assert den[7, 2, 1] == pytest.approx(0.75 * 0.5 * 0.75, abs=1e-12)
    assert den[0, 3, 2] == pytest.approx(0.25 * 0.5 * 0.25, abs=1e-12)
    assert num[7, 2, 1].real == pytest.approx(2 * 0.75 * 0.5 * 0.75, abs=1e-12)
    assert num[0, 3, 2].real == pytest.approx(-2 * 0.25 * 0.5 * 0.25, abs=1e-12)
    assert num[2, 3, 0].real == pytest.approx(0.0, abs=1e-12)
    assert den.sum() == pytest.approx(1.0, abs=1e-12)
    assert np.abs(num.imag).max() < 1e-12


def test_ctf_at_zero_frequency_is_the_amplitude_contrast():
    # tests/python/em/functional/test_reconstruction.py:93-111
    n = 8
    slices = np.ones((n, n // 2 + 1), np.complex128)
    ctf = np.array([15000, 15000, 0, 0], np.float64)
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    num, den = oracle.insert_f64("ctf", [acc, acc.real.copy(), slices, identity_transform(), ctf,
                                         np.ones(n // 2 + 1)], oracle.Window.nearest(), pixel_size=1.5)
    assert den[0, 0, 0] == pytest.approx(0.01)
    assert num[0, 0, 0].real == pytest.approx(0.1)


@pytest.mark.parametrize("n", [7, 8])
def test_projection_lands_on_the_central_plane(n):
    # tests/python/em/functional/test_reconstruction.py:12-38 (and the C++
    # integration test test_cpu_reconstruction.cpp:320-346): nearest window,
    # identity transform, radial weight 1.
    volume = np.random.default_rng(1).normal(size=(n, n, n))
    slices = np.fft.rfft2(volume.sum(axis=0))
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    num, den = oracle.insert_f64("plain", [acc, acc.real.copy(), slices, identity_transform(),
                                           np.ones(n // 2 + 1)], oracle.Window.nearest())
    inserted = den > 0
    assert np.all(den[inserted] == 1)
    assert not np.any(inserted[1:])
    np.testing.assert_allclose(num[inserted], np.fft.rfftn(volume)[inserted], atol=1e-9)


def test_projection_along_x_lands_on_the_x_plane():
    # test_cpu_reconstruction.cpp:348-377: the x-projection under an axis swap.
    n = 8
    volume = np.random.default_rng(4).normal(size=(n, n, n))
    image = volume.sum(axis=2)  # axes (z, y): image x = volume z, image y = volume y
    slices = np.fft.rfft2(image.T)
    t = np.array([[0, 0, 1, 0], [0, 1, 0, 0], [1, 0, 0, 0]], np.float64)
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    num, den = oracle.insert_f64("plain", [acc, acc.real.copy(), slices, t, np.ones(n // 2 + 1)],
                                 oracle.Window.nearest())
    full = np.fft.rfftn(volume)
    inserted = den > 0
    assert np.all(inserted[..., 1:] == False)  # noqa: E712
    np.testing.assert_allclose(num[inserted], full[inserted], atol=1e-9)


def test_shift_is_undone_by_the_translation():
    # test_cpu_reconstruction.cpp:379-416: shift an image by (dx, dy) and pass
    # t = (dx, dy); the result equals the unshifted insertion.
    n = 8
    rng = np.random.default_rng(5)
    image = rng.normal(size=(n, n))
    shifted = np.roll(image, (1, 2), axis=(0, 1))
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    t = identity_transform()
    t2 = t.copy()
    t2[0, 3] = 2.0
    t2[1, 3] = 1.0
    w = oracle.Window.linear()
    a = oracle.insert_f64("plain", [acc, acc.real.copy(), np.fft.rfft2(image), t, np.ones(n // 2 + 1)], w)
    b = oracle.insert_f64("plain", [acc, acc.real.copy(), np.fft.rfft2(shifted), t2, np.ones(n // 2 + 1)], w)
    np.testing.assert_allclose(b[0], a[0], atol=1e-9)
    np.testing.assert_allclose(b[1], a[1], atol=1e-12)


def test_nan_beyond_the_cutoff_is_left_out():
    # test_slice_insertion_program_builder.cpp:221-246
    inputs = make_case("plain", 8, 16, 4)
    inputs[4][:, 3] = 0.0
    inputs[4][:, 4] = 0.0
    inputs[2][:, 0, 4] = np.nan + 1j * np.nan
    num, den = oracle.insert_f64("plain", inputs, oracle.Window.linear())
    assert not np.isnan(num).any()
    assert not np.isnan(den).any()


def test_batch_axes_the_accumulator_lacks_are_summed():
    # test_slice_insertion_program_builder.cpp:193-219: accumulators (2, 1),
    # transforms (2, 3): each accumulator receives three slices.
    inputs = make_case("plain", 6, 8, 3, accumulator_batch=(2, 1), transform_batch=(2, 3))
    whole = oracle.insert_f64("plain", inputs, oracle.Window.linear())
    assert whole[0].shape == (2, 1, 8, 8, 5)
    for k in range(2):
        single = [inputs[0][k, 0], inputs[1][k, 0], inputs[2], inputs[3][k], inputs[4]]
        part = oracle.insert_f64("plain", single, oracle.Window.linear())
        np.testing.assert_allclose(whole[0][k, 0], part[0], atol=1e-12)
        np.testing.assert_allclose(whole[1][k, 0], part[1], atol=1e-12)


def test_split_batch_equals_whole():
    # test_cpu_reconstruction.cpp:505-594 and test_reconstruction.py:40-74
    inputs = make_case("ctf", 8, 16, 7, slices=4, zero_accumulators=True)
    w = oracle.Window.linear()
    whole = oracle.insert_f64("ctf", inputs, w, pixel_size=PIXEL_SIZE)
    first = oracle.insert_f64("ctf", [inputs[0], inputs[1]] + [a[:2] for a in inputs[2:]], w,
                              pixel_size=PIXEL_SIZE)
    second = oracle.insert_f64("ctf", [first[0], first[1]] + [a[2:] for a in inputs[2:]], w,
                               pixel_size=PIXEL_SIZE)
    np.testing.assert_allclose(second[0], whole[0], atol=1e-12)
    np.testing.assert_allclose(second[1], whole[1], atol=1e-12)


def test_weighted_denominator_equals_the_weight():
    # tests/python/em/functional/test_reconstruction.py:76-91
    n = 8
    slices = np.fft.rfft2(np.random.default_rng(3).normal(size=(n, n)))
    weights = np.full(slices.shape, 0.25)
    acc = np.zeros((n, n, n // 2 + 1), np.complex128)
    _, den = oracle.insert_f64("weighted", [acc, acc.real.copy(), slices, weights, identity_transform(),
                                            np.ones(n // 2 + 1)], oracle.Window.nearest())
    inserted = den > 0
    assert np.all(den[inserted] == 0.25)


def test_ctf_formula_matches_the_reference_polynomial_form():
    # The oracle evaluates the contract's atan2/cos form in double; the
    # reference evaluates its float32 polynomial (contrast_transfer_function.inl:73-87).
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    lib = oracle._reflib()
    rng = np.random.default_rng(9)
    optics = oracle.Optics()
    for _ in range(200):
        p = [rng.uniform(8000, 25000), rng.uniform(8000, 25000), rng.uniform(0, math.pi), rng.uniform(0, 0.5)]
        kx, ky = float(rng.integers(0, 33)), float(rng.integers(-32, 33))
        a = oracle.ctf(p, 64, 1.5, optics, kx, ky)
        b = lib.ref_ctf_f32(300.0, 2.7, 0.1, 1.5, 64, *p, kx, ky)
        assert abs(a - b) < 2e-4


def test_kaiser_bessel_matches_the_reference_table():
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    lib = oracle._reflib()
    w = oracle.Window.kaiser_bessel(2.0, 8.0)
    for d in np.linspace(-2.5, 2.5, 101):
        assert abs(oracle.window_weight(w, d) - lib.ref_kaiser_bessel_f32(2.0, 8.0, d)) < 1e-5


# --- the reference's kernels --------------------------------------------------

def _golden():
    return np.load(GOLDEN, allow_pickle=False)


def golden_cases():
    with np.load(GOLDEN, allow_pickle=False) as g:
        names = sorted({k.split("/")[0] for k in g.files})
    return names


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_the_reference_golden_vectors(name):
    g = _golden()
    meta = g[f"{name}/meta"]
    variant = str(g[f"{name}/variant"])
    window = WINDOWS[str(g[f"{name}/window"])]
    inputs = [g[f"{name}/in{i}"] for i in range(int(meta[2]))]
    num, den = oracle.insert_f64(variant, inputs, window, pixel_size=PIXEL_SIZE)
    tol = tolerance_for(window)
    np.testing.assert_allclose(num, g[f"{name}/num64"], atol=tol, rtol=0)
    np.testing.assert_allclose(den, g[f"{name}/den64"], atol=tol, rtol=0)
    # The reference in float32 stays within the 1e-4 of
    # test_slice_insertion_program_builder.cpp:190 (KB: 5e-4).
    tol32 = max(tol, 1e-4)
    np.testing.assert_allclose(num, g[f"{name}/num32"], atol=tol32, rtol=0)
    np.testing.assert_allclose(den, g[f"{name}/den32"], atol=tol32, rtol=0)


@needs_ref
@pytest.mark.parametrize("n,big", SIZES)
@pytest.mark.parametrize("wname", ["nearest", "linear", "kb1.5_6"])
def test_oracle_matches_the_reference_kernels(n, big, wname):
    inputs = make_case("plain", n, big, 11)
    w = WINDOWS[wname]
    a = oracle.insert_f64("plain", inputs, w)
    b = oracle.ref_insert("plain", inputs, w, dtype=np.float64, threads=3, tile_budget=1)
    np.testing.assert_allclose(a[0], b[0], atol=tolerance_for(w), rtol=0)
    np.testing.assert_allclose(a[1], b[1], atol=tolerance_for(w), rtol=0)


@needs_ref
def test_reference_kernels_agree_across_tilings_and_threads():
    # The reference's own tile-independence (test_slice_insertion_functor.cpp:87-143).
    inputs = make_case("ctf", 8, 16, 2)
    w = oracle.Window.linear()
    a = oracle.ref_insert("ctf", inputs, w, pixel_size=PIXEL_SIZE, threads=1, tile_budget=1 << 30)
    b = oracle.ref_insert("ctf", inputs, w, pixel_size=PIXEL_SIZE, threads=4, tile_budget=1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# --- closed forms downstream of the insertion ----------------------------------

def test_window_transforms_at_known_points():
    assert oracle.window_transform(oracle.Window.linear(), 8)[4] == 1.0
    # triangle: sinc^2(1/4) at x - C = 2 of N = 8
    assert oracle.window_transform(oracle.Window.linear(), 8)[6] == pytest.approx(
        (math.sin(math.pi / 4) / (math.pi / 4)) ** 2)
    # box (nearest): sinc(1/4) = 2 sqrt(2) / pi at x - C = 2 of N = 8, and the
    # first zero of sinc at x - C = N (outside the volume), so sinc(-1/2) at
    # index 0 of N = 4 is 2 / pi.
    box = oracle.window_transform(oracle.Window.nearest(), 8)
    assert box[4] == 1.0
    assert box[6] == pytest.approx(2 * math.sqrt(2) / math.pi)
    assert box[2] == box[6]
    assert oracle.window_transform(oracle.Window.nearest(), 4)[0] == pytest.approx(2 / math.pi)
    # The box transform equals a direct numerical transform of the box of width 1.
    d = np.linspace(-0.5, 0.5, 2001)
    for j in (5, 7):
        u = (j - 4) / 8
        assert box[j] == pytest.approx(np.trapezoid(np.cos(2 * math.pi * u * d), d), rel=1e-6)
    kb = oracle.window_transform(oracle.Window.kaiser_bessel(2.0, 8.0), 64)
    assert kb[32] == pytest.approx(1.0)
    # The KB transform equals a direct numerical transform of the window.
    d = np.linspace(-2, 2, 4001)
    phi = np.array([oracle.window_weight(oracle.Window.kaiser_bessel(2.0, 8.0), x) for x in d])
    for j in (33, 40, 50):
        u = (j - 32) / 64
        direct = np.trapezoid(phi * np.cos(2 * math.pi * u * d), d) / np.trapezoid(phi, d)
        assert kb[j] == pytest.approx(direct, rel=1e-4, abs=1e-6)


def test_fsc_of_a_spectrum_with_itself_is_one():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(8, 8, 5)) + 1j * rng.normal(size=(8, 8, 5))
    assert np.allclose(oracle.fsc(a, a, 8), 1.0)
    assert np.allclose(oracle.fsc(a, -a, 8), -1.0)


def test_normalize_threshold():
    num = np.array([1 + 1j, 2, 3], np.complex128)
    den = np.array([2.0, 1e-4, 0.5])
    np.testing.assert_allclose(oracle.normalize(num, den), [0.5 + 0.5j, 0, 6])


def test_shell_index_counts_lattice_points_by_hand():
    # Shell s holds the points with s - 1/2 <= |g| < s + 1/2. On the half grid
    # (x >= 0) of N = 16: shell 0 is the origin; shell 1 is |g|^2 in {1, 2}:
    # 5 axis points and 8 face diagonals with x >= 0 (18 in the full grid);
    # shell 2 is |g|^2 in {3, 4, 5, 6}: 4 + 5 + 16 + 12 points with x >= 0.
    s = oracle.shell_index(16)
    assert [(s == k).sum() for k in range(3)] == [1, 13, 37]
    assert s[0, 0, 1] == 1 and s[1, 1, 1] == 2 and s[0, 2, 1] == 2  # |g| = 1, sqrt(3), sqrt(5)
    assert s[8, 0, 0] == 8 and s[8, 8, 8] == 14  # the Nyquist index is -8: |g| = 8, 8 sqrt(3)


def test_shell_is_round_half_up_of_the_radius_by_the_integer_walk():
    # The reference's walk (row_insertion_kernel.inl:180-220): the largest s
    # with 4 s^2 - 4 s + 1 <= 4 |k|^2; equal to floor(|k| + 1/2), ties up.
    for kx in range(0, 300, 7):
        for ky in range(-300, 300, 11):
            q4 = 4 * (kx * kx + ky * ky)
            s = 0
            while 4 * (s + 1) ** 2 - 4 * (s + 1) + 1 <= q4:
                s += 1
            assert s == math.floor(math.sqrt(kx * kx + ky * ky) + 0.5)
    # Exact half-integers round up: |k|^2 = (s + 1/2)^2 has no integer
    # solution, so ties never arise on the lattice; 4|k|^2 = (2s+1)^2 - 1 is
    # the last point of shell s.
    for s in range(1, 200):
        q4 = (2 * s + 1) ** 2 - 1
        assert 4 * s * s - 4 * s + 1 <= q4 < 4 * (s + 1) ** 2 - 4 * (s + 1) + 1


def test_one_select_each_way_makes_the_float_shell_estimate_exact():
    # The device's shell: s = int(sqrtf(q) + 0.5f) in float32, then one step
    # up and one down by the walk's own tests. Dense q up to 4e6 and random q
    # up to 2^29 (|k|^2 of slices up to 16384 points).
    rng = np.random.default_rng(0)
    q = np.concatenate([np.arange(0, 4_000_000, dtype=np.int64), rng.integers(0, 2 ** 29, 2_000_000)])
    s = (np.sqrt(q.astype(np.float32)) + np.float32(0.5)).astype(np.int64)
    q4 = 4 * q
    s = s + (4 * s * s + 4 * s + 1 <= q4)
    s = s - ((s > 0) & (4 * s * s - 4 * s + 1 > q4))
    exact = np.floor(np.sqrt(q.astype(np.float64)) + 0.5).astype(np.int64)
    assert np.array_equal(s, exact)


# ---------------------------------------------------------------------------
# Background standardization (the oracle's restatement, pinned)
# ---------------------------------------------------------------------------

def test_philox_reproduces_random123_known_answers():
    # tests/unitary/src/backends/cpu/kernels/test_philox.cpp:13-38.
    assert oracle.philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)
    assert oracle.philox4x32_10((0xffffffff,) * 4, (0xffffffff, 0xffffffff)) == (
        0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)
    assert oracle.philox4x32_10((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0)) == (
        0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)


def test_uniform_open_closed_spans_zero_to_one():
    assert oracle.uniform_open_closed(0) == 2.0 ** -53
    assert oracle.uniform_open_closed(0xFFFFFFFFFFFFFFFF) == 1.0


def test_draws_are_the_pairs_in_counter_order_and_standard_normal():
    key = 0x0123456789ABCDEF
    first, second = oracle.standard_normal_pair(key, 21)
    assert oracle.standard_normal(key, 42) == first
    assert oracle.standard_normal(key, 43) == second
    draws = np.array([oracle.standard_normal(key, c) for c in range(20000)])
    assert abs(draws.mean()) < 0.03 and abs(draws.std() - 1.0) < 0.03


def test_background_statistics_match_the_reference_kernel_test():
    # test_standardize_background_kernel.cpp:65-86: a row of nine, background
    # indices 0-2 and 6-8 (radius 1), bounds +-1. The first pass takes 23 in
    # (mean 5.33, deviation 7.95); the second leaves it out: mean 1.8.
    row = np.array([1, 3, 1, 0, 0, 0, 3, 1, 23], np.float64)
    background = oracle.background_mask((9,), 1)
    assert background.tolist() == [True, True, True, False, False, False, True, True, True]
    mean, scale = oracle.background_moments(row, background, 0.0, 1.0, -np.inf, np.inf)
    assert mean > 5 and abs(1.0 / scale - 7.951240294584376) < 1e-12
    mean2, _ = oracle.background_moments(row, background, mean, scale, -1, 1)
    assert mean2 == 1.8


def test_background_mask_counts_lattice_points_outside_the_disk():
    # 8 x 8 image, centre (4, 4), radius 2: 13 lattice points lie within
    # distance 2 (the disk x^2 + y^2 <= 4), the other 51 are background.
    mask = oracle.background_mask((8, 8), 2)
    assert int((~mask).sum()) == 13 and not mask[4, 4] and mask[0, 0]


def test_standardization_without_background_or_spread():
    # No background element: mean 0, deviation 1; a constant background:
    # the mean is subtracted and nothing scaled
    # (standardize_background_operation.hpp:33-35).
    image = np.arange(9, dtype=np.float64).reshape(3, 3)
    out = oracle.standardize_background(image, 7, 10.0, -1e9, 1e9)
    assert np.array_equal(out, image)
    image = np.full((5, 5), 4.0)
    image[2, 2] = 9.0
    out = oracle.standardize_background(image, 7, 0.5, -1e9, 1e9)
    assert out[2, 2] == 5.0 and out[0, 0] == 0.0
