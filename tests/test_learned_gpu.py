"""The learned density compensation on the GPU, against the oracle.

* the iterative compensation and the features match their double-precision
  restatements;
* the MFMA forward pass matches the oracle with the kernel's float16
  rounding points, and the MFMA backward pass matches the oracle's gradients
  within the error of float16 gradient operands;
* a model trained on the device by these kernels lowers its loss and, on a
  held-out orientation distribution, gives a reconstruction closer (in
  relative L2 of the weighted spectrum) to the iterative one than the plain
  quotient does: the gain of the method over no compensation.
"""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import to_device
from oracle import mlp
from rexlib_amd import em, learned, synthetic
from rexlib_amd.precision import shell_report

pytestmark = pytest.mark.gpu


def density(count, big, kappa, fraction, seed):
    """The denominator of `count` slices: vMF around the top view for a
    fraction of them, uniform for the rest (unit weights, no CTF)."""
    rng = np.random.default_rng(seed)
    k = int(fraction * count)
    rot = np.concatenate([synthetic.von_mises_fisher_rotations(k, kappa, rng),
                          synthetic.uniform_rotations(count - k, rng)])
    n = big
    slices = torch.zeros((count, n, n // 2 + 1), dtype=torch.complex64, device="cuda")
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    em.insert_slices(num, den, slices, to_device(synthetic.transforms_of(rot)), torch.ones(n // 2 + 1, device="cuda"),
                     em.GriddingWindow.linear(), numerator_out=num, denominator_out=den)
    torch.cuda.synchronize()
    return den, rot


def test_iterative_compensation_matches_the_oracle():
    den, _ = density(300, 16, 5.0, 0.8, 0)
    w = learned.iterative_weights(den).cpu().numpy()
    ref = oracle.density_compensation(den.cpu().numpy(), passes=learned.PASSES, floor=learned.FLOOR)
    assert oracle.rel_l2(w, ref) < 1e-5


def test_features_match_the_oracle():
    den, _ = density(200, 16, 8.0, 0.5, 1)
    f = learned.features(den).float().cpu().numpy()
    ref = oracle.compensation_features(den.cpu().numpy())
    np.testing.assert_allclose(f[..., :10], ref, rtol=2e-3, atol=2e-3)
    assert np.all(f[..., 10:] == 0)


def random_rows(count, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((count, 16))
    x[:, :10] = rng.normal(0, 1, (count, 10))
    x = x.astype(np.float16).astype(np.float64)
    return x, rng.normal(0, 1, count)


def test_forward_matches_the_oracle():
    model = learned.CompensationModel(seed=3)
    x, _ = random_rows(1000, 0)
    got = model.predict(torch.from_numpy(x.astype(np.float16)).cuda()).cpu().numpy()
    flat = model.f16.float().cpu().numpy().astype(np.float64)
    flat[learned.parameter_count() - 208:] = model.parameters()[learned.parameter_count() - 208:]  # biases are float32
    ref, _ = mlp.forward(flat, x, fp16_activations=True)
    assert np.max(np.abs(got - ref)) <= 2e-3 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("rows", [16, 37, 4096])
def test_gradients_match_the_oracle(rows):
    model = learned.CompensationModel(seed=4)
    x, y = random_rows(rows, rows)
    grads, loss = model.gradients(torch.from_numpy(x.astype(np.float16)).cuda(),
                                  torch.from_numpy(y.astype(np.float32)).cuda(), loss_scale=1024.0)
    flat = model.f16.float().cpu().numpy().astype(np.float64)
    flat[learned.parameter_count() - 208:] = model.parameters()[learned.parameter_count() - 208:]
    ref_loss, ref = mlp.gradients(flat, x, y, fp16_activations=True)
    got = grads.cpu().numpy().astype(np.float64)
    assert float(loss.item()) == pytest.approx(ref_loss, rel=1e-4)
    parts_got = mlp.unpack(got)
    parts_ref = mlp.unpack(ref)
    for name, _ in mlp.SHAPES:
        a, b = parts_got[name], parts_ref[name]
        if np.linalg.norm(b) == 0:
            assert np.linalg.norm(a) < 1e-6
            continue
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-2, name


def test_training_lowers_the_loss_and_beats_the_plain_quotient():
    big = 32
    rows, targets = [], []
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    for d in range(12):
        rng = np.random.default_rng(100 + d)
        den, _ = density(int(rng.integers(800, 2000)), big, float(rng.uniform(2, 20)), float(rng.uniform(0.5, 0.9)),
                         100 + d)
        f, t = learned.training_rows(den, 4096, g)
        rows.append(f)
        targets.append(t)
    rows = torch.cat(rows)
    targets = torch.cat(targets)
    model = learned.CompensationModel(seed=0)
    log = model.train(rows, targets, steps=400, batch=2048, lr=3e-3)
    print("losses", log.losses[0], "->", log.losses[-1])
    assert log.losses[-1] < 0.3 * log.losses[0]

    # Held out: a new distribution and real data.
    n = big
    vol = synthetic.blob_volume(n, 10, n * 0.3, (1.0, 3.0), seed=7)
    rng = np.random.default_rng(7)
    count = 1500
    k = int(0.8 * count)
    rot = np.concatenate([synthetic.von_mises_fisher_rotations(k, 6.0, rng),
                          synthetic.uniform_rotations(count - k, rng)])
    stack = synthetic.particle_stack(vol, count, n, seed=8, snr=10.0, rotations=rot, device="cuda")
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    import rexlib_amd
    em.insert_slices(num, den, rexlib_amd.rfft2(stack.images), to_device(stack.transforms),
                     torch.ones(n // 2 + 1, device="cuda"), em.GriddingWindow.linear(), numerator_out=num,
                     denominator_out=den)
    w_it = learned.iterative_weights(den)
    w_ml = model.weights(den)
    plain = torch.where(den > 0, 1.0 / den.clamp_min(1e-30), torch.zeros_like(den))
    reference = (num * w_it).contiguous()
    error_ml = oracle.rel_l2((num * w_ml).cpu().numpy(), reference.cpu().numpy())
    error_plain = oracle.rel_l2((num * plain).cpu().numpy(), reference.cpu().numpy())
    fsc_ml = shell_report((num * w_ml).contiguous(), reference, 0.9)["min_fsc_to_cutoff"]
    print("rel-L2 to the iterative reconstruction: learned", error_ml, "plain quotient", error_plain,
          "; learned min FSC", fsc_ml)
    assert error_ml < error_plain
    weights, report = learned.compensate(num, den, model)
    assert report["used"] in ("learned", "iterative")
    if report["used"] == "iterative":
        assert torch.equal(weights, w_it)
    else:
        assert report["min_fsc"] >= learned.SUPPORT_FSC
