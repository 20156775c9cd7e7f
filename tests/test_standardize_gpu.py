"""Background standardization on the GPU against the oracle.

The statistics are sums in double (so the float32 mean and scale, rounded
from them, agree with the CPU's); the standardized values are computed in
the element type with no contraction, as the reference computes them
(standardization_map_kernel.inl:60-131). Draws come from double-precision
Box-Muller of the same Philox blocks; the GPU's log, sin and cos may differ
from the host's in the last bit, so draws are compared to 1e-6 after
rounding to the element type, and the set of replaced elements exactly.
"""
import numpy as np
import pytest
import torch

import oracle
from rexlib_amd import em

pytestmark = pytest.mark.gpu


def images(shape, seed, dtype=np.float32, hot=True, nonfinite=True):
    rng = np.random.default_rng(seed)
    x = rng.normal(3.0, 2.0, shape).astype(dtype)
    if hot:
        x.reshape(-1)[rng.integers(0, x.size, max(1, x.size // 200))] = 500.0
    if nonfinite:
        flat = x.reshape(-1)
        flat[rng.integers(0, x.size, 3)] = np.nan
        flat[rng.integers(0, x.size, 2)] = np.inf
    return x


def compare(got, want):
    replaced = ~np.isclose(got, want, rtol=0, atol=1e-5)
    assert not replaced.any(), (np.argwhere(replaced)[:5], got[replaced][:5], want[replaced][:5])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,radius", [((6, 32, 32), 10.0), ((3, 20, 27), 7.5), ((2, 64, 64), 0.0)])
def test_images_with_their_own_keys(shape, radius, dtype):
    print("standardize", shape, radius, dtype.__name__, flush=True)
    x = images(shape, seed=shape[1] + int(radius), dtype=dtype)
    keys = np.arange(100, 100 + shape[0], dtype=np.uint64)
    want = oracle.standardize_background(x, keys, radius, -3.0, 3.0)
    got = em.standardize_background(torch.from_numpy(x).cuda(), torch.from_numpy(keys.view(np.int64)).cuda(), radius,
                                    -3.0, 3.0).cpu().numpy()
    compare(got, want)


def test_one_key_for_every_image_gives_each_its_own_draws():
    x = images((4, 16, 16), seed=5, hot=False, nonfinite=False)
    x[:, 8, 8] = np.nan  # one draw per image, at the same place
    key = np.array(77, dtype=np.uint64)
    want = oracle.standardize_background(x, key, 4.0, -3.0, 3.0)
    got = em.standardize_background(torch.from_numpy(x).cuda(), torch.tensor(77, dtype=torch.int64).cuda(), 4.0,
                                    -3.0, 3.0).cpu().numpy()
    compare(got, want)
    assert len({float(v) for v in got[:, 8, 8]}) == 4


def test_an_image_draws_the_same_alone_and_in_a_batch():
    # With keys of their own, images draw the same however they are batched
    # (standardize_background_operation.hpp:51-56).
    x = images((5, 24, 24), seed=9)
    keys = torch.arange(10, 15, dtype=torch.int64).cuda()
    batch = em.standardize_background(torch.from_numpy(x).cuda(), keys, 6.0, -2.5, 2.5).cpu().numpy()
    for i in range(5):
        alone = em.standardize_background(torch.from_numpy(x[i:i + 1]).cuda(), keys[i:i + 1], 6.0, -2.5,
                                          2.5).cpu().numpy()
        assert np.array_equal(alone[0], batch[i])


def test_volumes_and_axes_that_are_not_the_last():
    x = images((10, 12, 2, 14), seed=11)
    keys = np.array([3, 4], dtype=np.uint64)
    moved = np.ascontiguousarray(np.moveaxis(x, 2, 0))  # (2, 10, 12, 14)
    want = np.moveaxis(oracle.standardize_background(moved, keys, 4.0, -3.0, 3.0, image_rank=3), 0, 2)
    got = em.standardize_background(torch.from_numpy(x).cuda(), torch.from_numpy(keys.view(np.int64)).cuda(), 4.0,
                                    -3.0, 3.0, axes=(0, 1, 3)).cpu().numpy()
    compare(got, want)


def test_a_hot_pixel_does_not_inflate_the_deviation():
    # test_cpu_preprocessing.cpp:503-525: a hot background pixel is left out
    # of the second statistics and replaced by a draw.
    rng = np.random.default_rng(3)
    x = rng.normal(0.0, 1.0, (1, 32, 32)).astype(np.float32)
    x[0, 0, 0] = 1.0e4
    got = em.standardize_background(torch.from_numpy(x).cuda(), torch.tensor([1], dtype=torch.int64).cuda(), 8.0,
                                    -4.0, 4.0).cpu().numpy()
    assert abs(got[0, 0, 0]) < 10 and 0.8 < got[0, :, 20:].std() < 1.2


def test_invalid_operands_are_refused():
    x = torch.zeros((2, 8, 8), device="cuda")
    k = torch.zeros(2, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError):
        em.standardize_background(x, k, -1.0, -3.0, 3.0)
    with pytest.raises(ValueError):
        em.standardize_background(x, k, 1.0, 1.0, 3.0)
    with pytest.raises(ValueError):
        em.standardize_background(x, torch.zeros(3, dtype=torch.int64, device="cuda"), 1.0, -3.0, 3.0)
    with pytest.raises(em.UnsupportedOperationError):
        em.standardize_background(x.half(), k, 1.0, -3.0, 3.0)


# Densely stored float32 images with a multiple of four columns run the
# register-resident kernel (one read and one write of each image); the
# shapes cover each of its register/LDS splits, an image that fills its
# workgroup's lanes in part (100 x 100, 20 x 8), a batch with a gap between
# images, a broadcast key and a radius beyond every element.
@pytest.mark.parametrize("shape,radius", [((3, 256, 256), 100.0), ((4, 128, 128), 50.0), ((5, 100, 100), 30.5),
                                          ((6, 64, 64), 20.0), ((7, 20, 8), 3.0), ((3, 176, 176), 70.0),
                                          ((2, 32, 32), 1000.0)])
def test_resident_kernel_shapes(shape, radius):
    print("standardize resident", shape, radius, flush=True)
    x = images(shape, seed=shape[1] + 3)
    keys = np.arange(7, 7 + shape[0], dtype=np.uint64)
    want = oracle.standardize_background(x, keys, radius, -3.0, 3.0)
    got = em.standardize_background(torch.from_numpy(x).cuda(), torch.from_numpy(keys.view(np.int64)).cuda(), radius,
                                    -3.0, 3.0).cpu().numpy()
    compare(got, want)


def test_resident_kernel_batch_with_gaps_and_a_shared_key():
    x = images((6, 64, 64), seed=21)
    strided = torch.from_numpy(x).cuda()[::2]  # images 0, 2, 4: a gap of one image between them
    want = oracle.standardize_background(x[::2], np.array(5, dtype=np.uint64), 12.0, -2.0, 2.0)
    got = em.standardize_background(strided, torch.tensor(5, dtype=torch.int64).cuda(), 12.0, -2.0, 2.0)
    compare(got.cpu().numpy(), want)


def test_resident_and_general_kernels_agree_on_an_offset_view():
    # An image view that starts one element into its buffer is not 16-byte
    # aligned and takes the general kernel; the aligned copy takes the
    # resident one. Both follow the same steps.
    x = images((3, 48, 48), seed=23)
    buffer = torch.zeros(x.size + 1, device="cuda")
    buffer[1:] = torch.from_numpy(x.reshape(-1)).cuda()
    offset = buffer[1:].view(3, 48, 48)
    keys = torch.arange(3, dtype=torch.int64).cuda()
    general = em.standardize_background(offset, keys, 15.0, -3.0, 3.0).cpu().numpy()
    resident = em.standardize_background(torch.from_numpy(x).cuda(), keys, 15.0, -3.0, 3.0).cpu().numpy()
    compare(resident, general)
    compare(resident, oracle.standardize_background(x, np.arange(3, dtype=np.uint64), 15.0, -3.0, 3.0))
