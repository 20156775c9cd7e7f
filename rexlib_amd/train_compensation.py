"""Training of the density-compensation perceptron shipped as
``rexlib_amd/compensation_model.npz``.

    python -m rexlib_amd.train_compensation

For each of 200 random orientation distributions (a fraction in [0.5, 0.95]
of the views from a von Mises-Fisher law of concentration 2-20 around a
random axis, the rest uniform on SO(3)), slices are inserted into a grid of
extent 64, 96, 128, 192 or 256 with CTF weights (defocus 0.5-4 um, unit
radial weights); the slice count is 0.15-2.5 times the squared extent
(log-uniform, at most 120,000), which spans the sampling densities of
typical data sets. The targets are log(W den) of the iterative estimate at
8,192 random voxels with den > 0, the features those of ``learned.features``
with the distribution's coverage table. The network is trained on the
device by the library's own forward, backward and Adam kernels. Twenty
further distributions are held out: the report gives, for each, the minimum
FSC (to 0.9 Nyquist) between the reconstructions weighted by the model and
by the iterative estimate, and the relative error of the model's and of the
plain quotient's weighted spectra to the iterative one over the shells up
to 0.9 Nyquist (beyond it, in the corners of the cube, few slices pass and
the weights are not judged).
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np
import torch

from . import em, learned, synthetic
from .precision import shell_report

HERE = os.path.dirname(os.path.abspath(__file__))


def random_rotation(rng) -> np.ndarray:
    return synthetic.uniform_rotations(1, rng)[0]


def distribution(rng, count):
    """Rotations of one random orientation distribution."""
    fraction = rng.uniform(0.5, 0.95)
    kappa = rng.uniform(2.0, 20.0)
    k = int(fraction * count)
    axis = random_rotation(rng)
    preferred = synthetic.von_mises_fisher_rotations(k, kappa, rng) @ axis.T
    rot = np.concatenate([preferred, synthetic.uniform_rotations(count - k, rng)])
    return rot, {"fraction": fraction, "kappa": kappa, "count": count}


SIZES = (64, 96, 128, 192, 256)
CHUNK = 8192


def slice_count(rng, big) -> int:
    """0.15-2.5 slices per pixel of a central section (log-uniform), at most 120,000."""
    return int(min(120000, np.exp(rng.uniform(np.log(0.15), np.log(2.5))) * big * big))


def geometry(rot, big, rng, device="cuda"):
    """Numerator (of unit-random slices) and denominator of the rotations at
    extent `big`, with CTF weights; slices are drawn and inserted in chunks."""
    n = big
    count = rot.shape[0]
    g = torch.Generator(device=device)
    g.manual_seed(int(rng.integers(1 << 31)))
    transforms = torch.from_numpy(synthetic.transforms_of(rot).astype(np.float32)).to(device)
    ctf = torch.from_numpy(synthetic.ctf_parameters(count, rng, defocus=(5000.0, 40000.0)).astype(np.float32)).to(device)
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device=device)
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device=device)
    radial = torch.ones(n // 2 + 1, device=device)
    for s in range(0, count, CHUNK):
        e = min(s + CHUNK, count)
        slices = torch.randn((e - s, n, n // 2 + 1), dtype=torch.complex64, device=device, generator=g)
        em.insert_ctf_slices(num, den, slices, transforms[s:e], ctf[s:e], radial, em.GriddingWindow.linear(),
                             em.CtfOptics(300.0, 2.7, 0.1), 1.5, numerator_out=num, denominator_out=den)
    return num, den, transforms


def held_out(model, rng, count):
    """Judging report of the model on `count` new distributions."""
    out = []
    for _ in range(count):
        big = int(rng.choice(SIZES))
        rot, info = distribution(rng, slice_count(rng, big))
        num, den, transforms = geometry(rot, big, rng)
        table = learned.coverage_table(transforms, big)
        w_ml = model.weights(den, table)
        w_it, trusted, fsc = learned.judge(num, den, w_ml)
        plain = torch.where(den > 0, 1.0 / den.clamp_min(1e-30), torch.zeros_like(den))
        reference = (num * w_it).contiguous()
        plain_report = shell_report((num * plain).contiguous(), reference, 0.9)
        learned_report = shell_report((num * w_ml).contiguous(), reference, 0.9)
        out.append(dict(info, big=big, min_fsc=fsc, trusted=bool(trusted),
                        band_error_learned=learned_report["relative_error_to_cutoff"],
                        band_error_plain=plain_report["relative_error_to_cutoff"],
                        min_fsc_plain=plain_report["min_fsc_to_cutoff"]))
    return out


def main():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--distributions", type=int, default=200)
    parser.add_argument("--held-out", type=int, default=20)
    parser.add_argument("--samples", type=int, default=8192)
    parser.add_argument("--steps", type=int, default=6000)
    parser.add_argument("--batch", type=int, default=8192)
    parser.add_argument("--lr", type=float, default=3e-3)
    parser.add_argument("--seed", type=int, default=1)
    parser.add_argument("--out", default=os.path.join(HERE, "compensation_model.npz"))
    args = parser.parse_args()
    rng = np.random.default_rng(args.seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(args.seed)
    rows, targets = [], []
    t0 = time.time()
    for d in range(args.distributions):
        big = int(rng.choice(SIZES))
        rot, _ = distribution(rng, slice_count(rng, big))
        _, den, transforms = geometry(rot, big, rng)
        table = learned.coverage_table(transforms, big)
        f, t = learned.training_rows(den, table, args.samples, g)
        rows.append(f)
        targets.append(t)
        if d % 20 == 0:
            print(f"distribution {d}: {big}^3, {rot.shape[0]} slices ({time.time() - t0:.1f} s)", flush=True)
    rows = torch.cat(rows)
    targets = torch.cat(targets)
    model = learned.CompensationModel(seed=args.seed)
    log = model.train(rows, targets, steps=args.steps, batch=args.batch, lr=args.lr)
    print(f"loss {log.losses[0]:.4f} -> {log.losses[-1]:.4f} ({time.time() - t0:.1f} s)", flush=True)
    report = held_out(model, rng, args.held_out)
    trusted = sum(r["trusted"] for r in report)
    better = sum(r["band_error_learned"] < r["band_error_plain"] for r in report)
    print(f"held out: {trusted}/{len(report)} trusted; min FSC {min(r['min_fsc'] for r in report):.4f}; "
          f"closer to the iterative weighting than the plain quotient in {better}/{len(report)}", flush=True)
    model.save(args.out)
    with open(os.path.splitext(args.out)[0] + "_report.json", "w") as f:
        json.dump({"distributions": args.distributions, "rows": int(rows.shape[0]), "steps": args.steps,
                   "losses": log.losses, "held_out": report}, f, indent=1)


if __name__ == "__main__":
    main()
