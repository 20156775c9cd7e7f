"""ctypes binding of librexlib_hip.so (include/rexlib_hip.h).

The library is built in-tree by rexlib_amd/build.py. Loading fails loudly when
it is missing: there is no other implementation of these operations here.
"""
from __future__ import annotations

import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "librexlib_hip.so")
MAX_RANK = 8

OK, INVALID_ARGUMENT, UNSUPPORTED, RUNTIME_ERROR = 0, 1, 2, 3
FLOAT32, FLOAT64, COMPLEX64, COMPLEX128, FLOAT16, BFLOAT16, BOOL, UINT64 = 0, 1, 2, 3, 4, 5, 6, 7

_DTYPES = {
    torch.float32: FLOAT32,
    torch.float64: FLOAT64,
    torch.complex64: COMPLEX64,
    torch.complex128: COMPLEX128,
    torch.float16: FLOAT16,
    torch.bfloat16: BFLOAT16,
    torch.bool: BOOL,
    torch.uint64: UINT64,
}


class UnsupportedOperationError(RuntimeError):
    """No gfx950 program takes these operands (rexlib's unsupported_operation_error,
    src/core/dispatch/program_service.cpp:84-89)."""


class Array(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("rank", ctypes.c_int32),
        ("extents", ctypes.c_int64 * MAX_RANK),
        ("strides", ctypes.c_int64 * MAX_RANK),
    ]


class Window(ctypes.Structure):
    _fields_ = [("profile", ctypes.c_int32), ("half_width", ctypes.c_double), ("shape", ctypes.c_double)]


class Optics(ctypes.Structure):
    _fields_ = [("voltage", ctypes.c_double), ("spherical_aberration", ctypes.c_double),
                ("amplitude_contrast", ctypes.c_double)]


EXPORTS = [
    "rexlib_hip_last_error",
    "rexlib_hip_version",
    "rexlib_hip_device_count",
    "rexlib_hip_set_insertion_tuning",
    "rexlib_hip_set_insertion_tuning_for_window",
    "rexlib_hip_get_insertion_tuning",
    "rexlib_hip_clear_insertion_tuning",
    "rexlib_hip_insert_slices",
    "rexlib_hip_insert_weighted_slices",
    "rexlib_hip_insert_ctf_slices",
    "rexlib_hip_insert_ctf_slices_no_alloc",
    "rexlib_hip_insert_ctf_images",
    "rexlib_hip_rfft2",
    "rexlib_hip_irfft2",
    "rexlib_hip_irfft3",
    "rexlib_hip_normalize",
    "rexlib_hip_gridding_correct",
    "rexlib_hip_shell_sums",
    "rexlib_hip_fft",
    "rexlib_hip_rfft",
    "rexlib_hip_irfft",
    "rexlib_hip_fill",
    "rexlib_hip_divide",
    "rexlib_hip_greater",
    "rexlib_hip_where",
    "rexlib_hip_standardize_background",
    "rexlib_hip_density_compensation",
    "rexlib_hip_coverage_table",
    "rexlib_hip_compensation_features",
    "rexlib_hip_mlp_parameter_count",
    "rexlib_hip_mlp_forward",
    "rexlib_hip_mlp_backward",
    "rexlib_hip_mlp_adam",
    "rexlib_hip_set_kernel_timing",
    "rexlib_hip_kernel_timing",
]

_lib = None


def library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        raise RuntimeError(
            f"{LIBRARY} is not built: run `python rexlib_amd/build.py` (the HIP library is the only implementation)")
    lib = ctypes.CDLL(LIBRARY)
    P = ctypes.POINTER
    A = P(Array)
    W = P(Window)
    O = P(Optics)
    vp = ctypes.c_void_p
    lib.rexlib_hip_last_error.restype = ctypes.c_char_p
    lib.rexlib_hip_last_error.argtypes = []
    lib.rexlib_hip_version.argtypes = [P(ctypes.c_int32)] * 3
    lib.rexlib_hip_version.restype = None
    lib.rexlib_hip_device_count.restype = ctypes.c_int32
    lib.rexlib_hip_device_count.argtypes = []
    lib.rexlib_hip_set_insertion_tuning.argtypes = [ctypes.c_int32] * 5
    lib.rexlib_hip_set_insertion_tuning.restype = ctypes.c_int
    lib.rexlib_hip_set_insertion_tuning_for_window.argtypes = [ctypes.c_int32] * 6
    lib.rexlib_hip_set_insertion_tuning_for_window.restype = ctypes.c_int
    lib.rexlib_hip_get_insertion_tuning.argtypes = [ctypes.c_int32] + [P(ctypes.c_int32)] * 4
    lib.rexlib_hip_get_insertion_tuning.restype = ctypes.c_int
    lib.rexlib_hip_clear_insertion_tuning.argtypes = []
    lib.rexlib_hip_clear_insertion_tuning.restype = None
    lib.rexlib_hip_insert_slices.argtypes = [A, A, A, A, A, W, A, A, vp]
    lib.rexlib_hip_insert_weighted_slices.argtypes = [A, A, A, A, A, A, W, A, A, vp]
    lib.rexlib_hip_insert_ctf_slices.argtypes = [A, A, A, A, A, A, W, O, ctypes.c_double, A, A, vp]
    lib.rexlib_hip_insert_ctf_slices_no_alloc.argtypes = [A, A, A, A, A, A, W, O, ctypes.c_double, A, A, vp]
    lib.rexlib_hip_insert_ctf_images.argtypes = [A, A, A, A, ctypes.c_int64, A, A, A, W, O, ctypes.c_double, vp]
    lib.rexlib_hip_rfft2.argtypes = [A, A, vp]
    lib.rexlib_hip_irfft2.argtypes = [A, A, vp]
    lib.rexlib_hip_irfft3.argtypes = [A, A, vp]
    lib.rexlib_hip_normalize.argtypes = [A, A, ctypes.c_double, A, vp]
    lib.rexlib_hip_gridding_correct.argtypes = [A, W, A, vp]
    lib.rexlib_hip_shell_sums.argtypes = [A, A, vp, vp]
    lib.rexlib_hip_fill.argtypes = [A, ctypes.c_double, ctypes.c_double, vp]
    I32 = ctypes.c_int32
    lib.rexlib_hip_fft.argtypes = [A, A, P(I32), I32, I32, I32, vp]
    lib.rexlib_hip_rfft.argtypes = [A, A, P(I32), I32, I32, vp]
    lib.rexlib_hip_irfft.argtypes = [A, A, P(I32), I32, I32, vp]
    lib.rexlib_hip_divide.argtypes = [A, A, A, vp]
    lib.rexlib_hip_greater.argtypes = [A, A, A, vp]
    lib.rexlib_hip_where.argtypes = [A, A, A, A, vp]
    d = ctypes.c_double
    lib.rexlib_hip_standardize_background.argtypes = [A, A, ctypes.c_uint32, d, d, d, A, vp]
    lib.rexlib_hip_density_compensation.argtypes = [A, ctypes.c_int32, d, A, vp]
    lib.rexlib_hip_coverage_table.argtypes = [A, ctypes.c_int32, A, vp]
    lib.rexlib_hip_compensation_features.argtypes = [A, A, d, A, vp]
    lib.rexlib_hip_mlp_parameter_count.argtypes = []
    lib.rexlib_hip_mlp_forward.argtypes = [A, A, A, A, vp]
    lib.rexlib_hip_mlp_backward.argtypes = [A, A, A, A, A, A, d, vp]
    lib.rexlib_hip_mlp_adam.argtypes = [A, A, A, A, A, d, d, d, d, ctypes.c_int32, d, vp]
    lib.rexlib_hip_set_kernel_timing.argtypes = [ctypes.c_int32]
    lib.rexlib_hip_kernel_timing.argtypes = [ctypes.c_int32, P(ctypes.c_double), P(ctypes.c_int64)]
    for name in EXPORTS[5:]:
        getattr(lib, name).restype = ctypes.c_int
    lib.rexlib_hip_mlp_parameter_count.restype = ctypes.c_int32
    _lib = lib
    return lib


def array(t: torch.Tensor | None):
    """Describe a device tensor; None passes a null operand."""
    if t is None:
        return None
    if t.dim() > MAX_RANK:
        raise ValueError(f"rank {t.dim()} exceeds {MAX_RANK}")
    if t.dtype not in _DTYPES:
        raise ValueError(f"unsupported element type {t.dtype}")
    a = Array()
    a.data = t.data_ptr()
    a.dtype = _DTYPES[t.dtype]
    a.rank = t.dim()
    for i, (e, s) in enumerate(zip(t.shape, t.stride())):
        a.extents[i] = e
        a.strides[i] = s
    return ctypes.byref(a)


def stream_handle(device: torch.device | None = None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def check(status: int) -> None:
    if status == OK:
        return
    message = library().rexlib_hip_last_error().decode()
    if status == INVALID_ARGUMENT:
        raise ValueError(message)
    if status == UNSUPPORTED:
        raise UnsupportedOperationError(message)
    raise RuntimeError(message)


KERNEL_BRICK, KERNEL_PREPARE, KERNEL_SCATTER = 0, 1, 2


def set_kernel_timing(enable: bool) -> None:
    """Bracket each launch of the insertion kernels with HIP events on its stream."""
    check(library().rexlib_hip_set_kernel_timing(1 if enable else 0))


def kernel_timing(kernel: int) -> tuple[float, int]:
    """(total milliseconds, launches) recorded for `kernel` since timing was switched on."""
    total = ctypes.c_double()
    launches = ctypes.c_int64()
    check(library().rexlib_hip_kernel_timing(int(kernel), ctypes.byref(total), ctypes.byref(launches)))
    return total.value, launches.value
