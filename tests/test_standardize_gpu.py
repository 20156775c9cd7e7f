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
