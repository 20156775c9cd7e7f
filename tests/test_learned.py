"""The double-precision oracle of the compensation perceptron and of the
density compensation, pinned on the CPU."""
import numpy as np
import pytest

import oracle
from oracle import mlp

P = sum(int(np.prod(s)) for _, s in mlp.SHAPES)


def random_parameters(seed):
    rng = np.random.default_rng(seed)
    p = {name: rng.normal(0, 0.3, shape) for name, shape in mlp.SHAPES}
    p["w1"][10:] = 0.0
    p["w4"][:, 1:] = 0.0
    p["b4"][1:] = 0.0
    return mlp.pack(p)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    flat = random_parameters(2)
    x = np.zeros((9, 16))
    x[:, :10] = rng.normal(size=(9, 10))
    y = rng.normal(size=9)
    _, g = mlp.gradients(flat, x, y)
    eps = 1e-6
    live = np.flatnonzero(np.abs(flat) > 0)
    for i in rng.choice(live, 60, replace=False):
        a = flat.copy()
        b = flat.copy()
        a[i] += eps
        b[i] -= eps
        fd = (mlp.loss(a, x, y) - mlp.loss(b, x, y)) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-8)


def test_one_unit_network_closed_form():
    # A single live path: out = w4 * relu(w3 * relu(w2 * relu(w1 x0 + b1) ...)).
    p = {name: np.zeros(shape) for name, shape in mlp.SHAPES}
    p["w1"][0, 0], p["w2"][0, 0], p["w3"][0, 0], p["w4"][0, 0] = 2.0, 3.0, 0.5, 4.0
    p["b4"][0] = 1.0
    flat = mlp.pack(p)
    x = np.zeros((1, 16))
    x[0, 0] = 1.5
    out, _ = mlp.forward(flat, x)
    assert out[0] == pytest.approx(4.0 * 0.5 * 3.0 * 2.0 * 1.5 + 1.0)
    _, g = mlp.gradients(flat, x, np.array([0.0]))
    gp = mlp.unpack(g)
    d = 2 * out[0]
    assert gp["b4"][0] == pytest.approx(d)
    assert gp["w1"][0, 0] == pytest.approx(d * 4.0 * 0.5 * 3.0 * 1.5)


def test_density_compensation_flattens_uniform_sampling():
    # A constant density converges to W = 1 / den after one pass.
    den = np.full((8, 8, 5), 4.0)
    w = oracle.density_compensation(den, passes=3)
    np.testing.assert_allclose(w, 0.25)


def test_density_compensation_fixed_point():
    rng = np.random.default_rng(0)
    den = rng.uniform(0.5, 2.0, (8, 8, 5))
    w = oracle.density_compensation(den, passes=60)
    full_w = oracle.full_from_half(w, 8)
    full_d = oracle.full_from_half(den, 8)
    prod = full_w * full_d
    t = {-1: 1 / 6, 0: 2 / 3, 1: 1 / 6}
    c = sum(t[a] * t[b] * t[cc] * np.roll(prod, (-a, -b, -cc), axis=(0, 1, 2))
            for a in (-1, 0, 1) for b in (-1, 0, 1) for cc in (-1, 0, 1))
    np.testing.assert_allclose(c, 1.0, atol=1e-3)


def test_full_from_half_is_hermitian_for_a_real_field():
    rng = np.random.default_rng(3)
    full = rng.normal(size=(6, 6, 6))
    half_spec = np.fft.rfftn(full)
    spec_full = np.fft.fftn(full)
    got = oracle.full_from_half(np.abs(half_spec), 6)
    np.testing.assert_allclose(got, np.abs(spec_full), atol=1e-9)


def test_features_at_the_origin_and_on_the_axes():
    den = np.zeros((8, 8, 5))
    den[0, 0, 0] = np.e - 1
    table = np.zeros(oracle.COVERAGE_SHAPE)
    table[0, 0, 0] = np.arange(8)
    table[3, 16, 0] = 10 + np.arange(8)  # theta = pi/2, phi = 0: the +x axis
    f = oracle.compensation_features(den, table, 1.0)
    assert f[0, 0, 0, 0] == 0.0
    assert f[0, 0, 0, 1] == pytest.approx(1.0)
    assert np.array_equal(f[0, 0, 0, 2:], np.arange(8))
    # g = (x, y, z) = (2, 0, 0): radius 2 / 4, bin floor(0.5 * 8) = 4? No:
    # (4, 0, 0) is radius 1, bin 7; (1, 0, 0) is radius 1/4, bin 2.
    assert f[0, 0, 1, 0] == 0.25
    table2 = np.zeros(oracle.COVERAGE_SHAPE)
    table2[2, 16, 0] = 1 + np.arange(8)
    f2 = oracle.compensation_features(den, table2, 1.0)
    assert np.array_equal(f2[0, 0, 1, 2:], 1 + np.arange(8))


def test_coverage_of_one_plane():
    # One slice in the xy plane (normal +z): the planes through the points of
    # the equator (u perpendicular to z) hold it, at the angle of -e1 = +z
    # around u, i.e. pi, which is 0 modulo pi: bin 0. Away from the equator
    # by more than the band it is not counted.
    t = np.zeros((1, 3, 4))
    t[0, 0, 0] = 1.0
    t[0, 1, 1] = 1.0
    table = oracle.coverage_table(t, 64)
    equator = table[:, 16]  # theta = pi/2
    assert np.all(equator[..., 0] == 1.0) and np.all(equator[..., 1:] == 0.0)
    assert np.all(table[7, 0] == 0.0)  # near the pole, far from the plane at radius ~30
    assert np.all(table[7, 16].sum(-1) == 1.0)
    # At radius 30 the band (0.083) is narrower than the grid step (pi / 32):
    # the neighbouring latitudes do not hold the plane.
    assert np.all(table[7, 15] == 0.0) and np.all(table[7, 17] == 0.0)


def test_coverage_fractions_sum_to_one_where_planes_pass():
    rng = np.random.default_rng(1)
    from rexlib_amd import synthetic
    t = synthetic.transforms_of(synthetic.uniform_rotations(500, rng))
    table = oracle.coverage_table(t, 32)
    totals = table.sum(-1)
    assert np.allclose(totals[totals > 0], 1.0)
    # Radius bin 0 (radius 1 voxel: band 1) holds every plane everywhere.
    assert np.all(totals[0] > 0.999)


def test_separable_k3_equals_the_27_point_sum():
    field = np.random.default_rng(3).normal(size=(6, 7, 5))
    np.testing.assert_allclose(oracle.convolve_k3(field), oracle.convolve_k3_direct(field), rtol=1e-13, atol=1e-13)
    # A delta spreads into the kernel itself: 2/3 at the centre, (1/6)^3 in the corners.
    delta = np.zeros((5, 5, 5))
    delta[2, 2, 2] = 1.0
    c = oracle.convolve_k3(delta)
    assert c[2, 2, 2] == pytest.approx(8 / 27)
    assert c[1, 1, 1] == pytest.approx(1 / 216)
    assert c.sum() == pytest.approx(1.0)


def test_features_take_the_nyquist_index_as_negative():
    # Index N/2 along y or z is the frequency -N/2 (numpy's fftfreq): the
    # voxel (z, y, x) = (N/2, 0, 0) lies on the -z axis (theta = pi, bin 32),
    # (0, N/2, 0) on the -y axis (phi = 3 pi / 2, bin 48).
    den = np.ones((16, 16, 9))
    table = np.zeros(oracle.COVERAGE_SHAPE)
    table[7, 32, 0] = np.arange(8)
    table[7, 16, 48] = 10 + np.arange(8)
    f = oracle.compensation_features(den, table, 1.0)
    assert np.array_equal(f[8, 0, 0, 2:], np.arange(8))
    assert np.array_equal(f[0, 8, 0, 2:], 10 + np.arange(8))
