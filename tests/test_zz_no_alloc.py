"""The insertion entry point that allocates nothing
(rexlib_hip_insert_ctf_slices_no_alloc): through the C ABI, it gives what
rexlib_hip_insert_ctf_slices gives on the same operands, on two calls in a row
(the second one smaller, so it meets the scratch array's leftovers), and it
refuses a call whose workspace exceeds its fixed scratch array."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from rexlib_amd import _native as N
from rexlib_amd import synthetic

pytestmark = pytest.mark.gpu


def operands(n, big, count, seed):
    rng = np.random.default_rng(seed)
    h = n // 2 + 1
    slices = (rng.uniform(-1, 1, (count, n, h)) + 1j * rng.uniform(-1, 1, (count, n, h))).astype(np.complex64)
    transforms = synthetic.transforms_of(synthetic.uniform_rotations(count, rng)).astype(np.float32)
    ctf = synthetic.ctf_parameters(count, rng).astype(np.float32)
    radial = rng.uniform(0, 2, h).astype(np.float32)
    num = (rng.uniform(-1, 1, (big, big, big // 2 + 1)) + 1j * rng.uniform(-1, 1, (big, big, big // 2 + 1)))
    den = rng.uniform(0, 2, (big, big, big // 2 + 1))
    return [torch.from_numpy(a).cuda() for a in (num.astype(np.complex64), den.astype(np.float32), slices,
                                                  transforms, ctf, radial)]


def last_error():
    return N.library().rexlib_hip_last_error().decode()


def call(entry, ops):
    refs = [N.array(t) for t in ops]
    window = N.Window(1, 1.0, 0.0)
    optics = N.Optics(300.0, 2.7, 0.1)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return entry(refs[0], refs[1], refs[2], refs[3], refs[4], refs[5], ctypes.byref(window), ctypes.byref(optics),
                 1.5, refs[0], refs[1], stream)


# 256^2 slices into 64^3 take 396 KB of prepared samples each, so 600 of
# them go through the scratch array in two chunks.
@pytest.mark.parametrize("n,big,count", [(128, 128, 300), (64, 64, 40), (256, 64, 600)])
def test_no_alloc_insertion_equals_the_allocating_one(n, big, count):
    lib = N.library()
    for sizes in ((n, big, count), (n // 2, big // 2, count // 3)):
        print("no-alloc insertion", sizes, flush=True)
        ops = operands(*sizes, seed=sum(sizes))
        expected = [t.clone() for t in ops]
        assert call(lib.rexlib_hip_insert_ctf_slices, expected) == N.OK, last_error()
        assert call(lib.rexlib_hip_insert_ctf_slices_no_alloc, ops) == N.OK, last_error()
        torch.cuda.synchronize()
        assert oracle.rel_l2(ops[0].cpu().numpy(), expected[0].cpu().numpy()) <= 1e-5
        assert oracle.rel_l2(ops[1].cpu().numpy(), expected[1].cpu().numpy()) <= 1e-5


def test_no_alloc_insertion_refuses_what_its_scratch_cannot_hold():
    # Per slice of 8 x 5 the call keeps a 112-byte entry and 13 phases of
    # 8 bytes: 700,000 slices need 151 MB, beyond the 128 MiB scratch array.
    n, big, count = 8, 8, 700000
    h = n // 2 + 1
    ops = [torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda"),
           torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda"),
           torch.zeros((count, n, h), dtype=torch.complex64, device="cuda"),
           torch.zeros((count, 3, 4), dtype=torch.float32, device="cuda"),
           torch.zeros((count, 4), dtype=torch.float32, device="cuda"),
           torch.ones(h, dtype=torch.float32, device="cuda")]
    ops[3][:, 0, 0] = 1.0
    ops[3][:, 1, 1] = 1.0
    assert call(N.library().rexlib_hip_insert_ctf_slices_no_alloc, ops) == N.INVALID_ARGUMENT
    assert "scratch" in last_error()
