"""The C ABI library on the CPU: it loads, exports what include/rexlib_hip.h
declares, and rejects invalid operands before touching a device."""
import ctypes
import os
import re

import pytest

from rexlib_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rexlib_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(rexlib_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = N.library()
    declared = declared_functions()
    assert "rexlib_hip_insert_ctf_slices" in declared
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(N.EXPORTS) == declared


def test_version_is_the_reference_contract():
    major, minor, patch = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    N.library().rexlib_hip_version(ctypes.byref(major), ctypes.byref(minor), ctypes.byref(patch))
    assert (major.value, minor.value) == (0, 1)


def test_no_device_is_not_an_error():
    assert N.library().rexlib_hip_device_count() >= 0


def fake(shape, dtype, address=0x1000):
    """A descriptor of memory that is never touched: validation fails first."""
    a = N.Array()
    a.data = address
    a.dtype = dtype
    a.rank = len(shape)
    stride = 1
    strides = []
    for e in reversed(shape):
        strides.append(stride)
        stride *= e
    for i, (e, s) in enumerate(zip(shape, reversed(strides))):
        a.extents[i] = e
        a.strides[i] = s
    return a


def call_ctf(num, den, sl, tr, ctf, ra, window=None, optics=None, pixel=1.5):
    lib = N.library()
    window = window or N.Window(1, 1.0, 0.0)
    optics = optics or N.Optics(300.0, 2.7, 0.1)
    return lib.rexlib_hip_insert_ctf_slices(ctypes.byref(num), ctypes.byref(den), ctypes.byref(sl),
                                            ctypes.byref(tr), ctypes.byref(ctf), ctypes.byref(ra),
                                            ctypes.byref(window), ctypes.byref(optics), pixel,
                                            None, None, None)


def good_operands(n=8, big=16, dtype_real=N.FLOAT32, dtype_complex=N.COMPLEX64):
    h = n // 2 + 1
    half = big // 2 + 1
    return [fake((big, big, half), dtype_complex), fake((big, big, half), dtype_real),
            fake((3, n, h), dtype_complex), fake((3, 3, 4), dtype_real), fake((3, 4), dtype_real),
            fake((3, h), dtype_real)]


def message():
    return N.library().rexlib_hip_last_error().decode()


def test_pixel_size_must_be_positive():
    # src/em/ops/reconstruction/insert_ctf_slices_operation.cpp:17-27
    assert call_ctf(*good_operands(), pixel=0.0) == N.INVALID_ARGUMENT
    assert "pixel size" in message()
    assert call_ctf(*good_operands(), pixel=float("nan")) == N.INVALID_ARGUMENT


def test_optics_are_checked():
    assert call_ctf(*good_operands(), optics=N.Optics(300.0, 2.7, 1.0)) == N.INVALID_ARGUMENT
    assert "amplitude contrast" in message()
    assert call_ctf(*good_operands(), optics=N.Optics(-1.0, 2.7, 0.1)) == N.INVALID_ARGUMENT


def test_kaiser_bessel_window_is_checked():
    assert call_ctf(*good_operands(), window=N.Window(2, -1.0, 2.0)) == N.INVALID_ARGUMENT
    assert "half-width" in message()
    assert call_ctf(*good_operands(), window=N.Window(2, 2.0, -1.0)) == N.INVALID_ARGUMENT
    assert call_ctf(*good_operands(), window=N.Window(2, 5.0, 2.0)) == N.UNSUPPORTED


def test_shapes_are_checked():
    ops = good_operands()
    ops[2] = fake((3, 8, 4), N.COMPLEX64)  # (n, n//2) instead of (n, n//2+1)
    assert call_ctf(*ops) == N.INVALID_ARGUMENT
    assert "slice" in message()
    ops = good_operands()
    ops[1] = fake((16, 16, 8), N.FLOAT32)
    assert call_ctf(*ops) == N.INVALID_ARGUMENT
    ops = good_operands()
    ops[4] = fake((3, 5), N.FLOAT32)
    assert call_ctf(*ops) == N.INVALID_ARGUMENT
    ops = good_operands()
    ops[3] = fake((2, 3, 4), N.FLOAT32)  # batch 2 against 3
    assert call_ctf(*ops) == N.INVALID_ARGUMENT
    assert "broadcast" in message()


def test_types_are_checked_and_float64_is_unsupported():
    ops = good_operands()
    ops[2] = fake((3, 8, 5), N.FLOAT32)  # real slices
    assert call_ctf(*ops) == N.INVALID_ARGUMENT
    assert call_ctf(*good_operands(dtype_real=N.FLOAT64, dtype_complex=N.COMPLEX128)) == N.UNSUPPORTED


def test_fft_lengths_with_other_prime_factors_are_unsupported():
    lib = N.library()
    images = fake((2, 7, 7), N.FLOAT32)
    spectra = fake((2, 7, 4), N.COMPLEX64)
    assert lib.rexlib_hip_rfft2(ctypes.byref(images), ctypes.byref(spectra), None) == N.UNSUPPORTED
    assert "prime factors" in message()


def test_python_api_rejects_host_tensors():
    import torch

    from rexlib_amd import em
    t = torch.zeros(4)
    with pytest.raises(N.UnsupportedOperationError):
        em.insert_slices(t, t, t, t, t, em.GriddingWindow.linear())
    with pytest.raises(ValueError):
        em.GriddingWindow.kaiser_bessel(0.0, 1.0)
    with pytest.raises(ValueError):
        em.CtfOptics(300, 2.7, 1.5)
    assert em.CtfOptics(300, 2.7, 0.1).wavelength == pytest.approx(0.019687615309, rel=1e-9)


def strided(shape, strides, dtype, address=0x1000):
    a = fake(shape, dtype, address)
    for i, s in enumerate(strides):
        a.strides[i] = s
    return a


def test_gridding_correction_checks_its_operands():
    lib = N.library()
    window = N.Window(1, 1.0, 0.0)

    def call(volume, out):
        return lib.rexlib_hip_gridding_correct(ctypes.byref(volume), ctypes.byref(window), ctypes.byref(out), None)

    # A (4, 4, 64) volume is not a cube: rejected before any memory is read.
    assert call(fake((4, 4, 64), N.FLOAT32), fake((4, 4, 64), N.FLOAT32)) == N.INVALID_ARGUMENT
    # A crop larger than the volume, and a crop aliasing its input.
    assert call(fake((8, 8, 8), N.FLOAT32), fake((16, 16, 16), N.FLOAT32, 0x2000)) == N.INVALID_ARGUMENT
    assert call(fake((16, 16, 16), N.FLOAT32), fake((8, 8, 8), N.FLOAT32)) == N.INVALID_ARGUMENT
    # Batch shapes differ.
    assert call(fake((2, 8, 8, 8), N.FLOAT32), fake((3, 8, 8, 8), N.FLOAT32, 0x2000)) == N.INVALID_ARGUMENT
    # A transposed view has no gfx950 program.
    assert call(strided((8, 8, 8), (1, 8, 64), N.FLOAT32), fake((8, 8, 8), N.FLOAT32, 0x2000)) == N.UNSUPPORTED
    assert "contiguous" in message()


def test_normalize_and_shell_sums_check_their_operands():
    lib = N.library()
    num = fake((8, 8, 5), N.COMPLEX64)
    den = fake((8, 8, 5), N.FLOAT32, 0x2000)
    out = strided((8, 8, 5), (1, 5, 40), N.COMPLEX64, 0x3000)
    assert lib.rexlib_hip_normalize(ctypes.byref(num), ctypes.byref(den), 1e-3, ctypes.byref(out), None) == \
        N.UNSUPPORTED
    bad = fake((8, 8, 8), N.COMPLEX64)
    assert lib.rexlib_hip_normalize(ctypes.byref(bad), ctypes.byref(den), 1e-3, ctypes.byref(bad), None) == \
        N.INVALID_ARGUMENT
    b = strided((8, 8, 5), (40, 1, 8), N.COMPLEX64, 0x2000)
    assert lib.rexlib_hip_shell_sums(ctypes.byref(num), ctypes.byref(b), None, None) == N.UNSUPPORTED
    batched = fake((2, 8, 8, 5), N.COMPLEX64)
    assert lib.rexlib_hip_shell_sums(ctypes.byref(batched), ctypes.byref(batched), None, None) == N.UNSUPPORTED


def test_elementwise_programs_check_their_operands():
    lib = N.library()
    a = fake((4, 5), N.FLOAT32)
    b = fake((3, 5), N.FLOAT32, 0x2000)
    out = fake((4, 5), N.FLOAT32, 0x3000)
    # (3, 5) does not broadcast to (4, 5).
    assert lib.rexlib_hip_divide(ctypes.byref(a), ctypes.byref(b), ctypes.byref(out), None) == N.INVALID_ARGUMENT
    assert "broadcast" in message()
    c = fake((4, 5), N.COMPLEX64, 0x2000)
    assert lib.rexlib_hip_divide(ctypes.byref(a), ctypes.byref(c), ctypes.byref(out), None) == N.INVALID_ARGUMENT
    mask = fake((4, 5), N.BOOL, 0x4000)
    assert lib.rexlib_hip_greater(ctypes.byref(a), ctypes.byref(a), ctypes.byref(out), None) == N.INVALID_ARGUMENT
    assert lib.rexlib_hip_where(ctypes.byref(a), ctypes.byref(a), ctypes.byref(a), ctypes.byref(out), None) == \
        N.INVALID_ARGUMENT
    d = fake((4, 5), N.FLOAT64, 0x2000)
    assert lib.rexlib_hip_greater(ctypes.byref(d), ctypes.byref(d), ctypes.byref(mask), None) == N.UNSUPPORTED


def test_insertion_tuning_is_a_table_per_volume_extent():
    """The brick tuning is kept per volume extent (0: every other extent),
    checked on entry, and read whole by each call; no device is needed."""
    from rexlib_amd import tune
    lib = N.library()
    tune.clear_tuning()
    try:
        assert tune.get_tuning(256) == ((16, 8, 6), 256)
        tune.set_tuning(256, (16, 16, 8), 128)
        assert tune.get_tuning(256) == ((16, 16, 8), 128)
        assert tune.get_tuning(64) == ((16, 8, 6), 256)
        tune.set_tuning(0, (8, 8, 8), 64)
        assert tune.get_tuning(64) == ((8, 8, 8), 64)
        assert tune.get_tuning(256) == ((16, 16, 8), 128)
        assert lib.rexlib_hip_set_insertion_tuning(96, 16, 8, 7, 128) == N.INVALID_ARGUMENT
        assert lib.rexlib_hip_set_insertion_tuning(96, 8, 16, 6, 128) == N.INVALID_ARGUMENT
        assert lib.rexlib_hip_set_insertion_tuning(96, 16, 8, 6, 8) == N.INVALID_ARGUMENT
        assert lib.rexlib_hip_set_insertion_tuning(-1, 16, 8, 6, 64) == N.INVALID_ARGUMENT
        assert tune.get_tuning(96) == ((8, 8, 8), 64)
    finally:
        tune.clear_tuning()
    assert tune.get_tuning(256) == ((16, 8, 6), 256)


def test_insertion_tuning_for_one_window_family():
    """Entries for one window family (taps per axis 1, 2 or 4) sit beside the
    extent's entry and leave it as it is; other taps are refused; the
    shipped table, which holds such an entry, applies whole."""
    from rexlib_amd import tune
    lib = N.library()
    tune.clear_tuning()
    try:
        tune.set_tuning(256, (8, 8, 8), 1024)
        tune.set_tuning(256, (8, 6, 6), 1024, taps=4)
        assert tune.get_tuning(256) == ((8, 8, 8), 1024)
        assert lib.rexlib_hip_set_insertion_tuning_for_window(256, 3, 8, 6, 6, 1024) == N.INVALID_ARGUMENT
        assert lib.rexlib_hip_set_insertion_tuning_for_window(256, 4, 8, 6, 5, 1024) == N.INVALID_ARGUMENT
        assert lib.rexlib_hip_set_insertion_tuning_for_window(1 << 20, 4, 8, 6, 6, 1024) == N.INVALID_ARGUMENT
        table = tune.apply_all()
        assert "256/4" in table
        assert tune.get_tuning(256) == ((table["256"]["extent_x"], table["256"]["extent_y"],
                                         table["256"]["extent_z"]), table["256"]["share"])
    finally:
        tune.clear_tuning()
