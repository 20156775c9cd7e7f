"""Autotuner of the insertion's brick schedule.

For a workload (slice extent n, volume extent N, slices per launch L) it times
the insertion on the device for every brick shape the library holds (the
brick's first-tap cells along x, y and z) and several shares (expected plane
hits per task, which split busy bricks), checks every configuration's sums
against the first one's, and records the fastest per volume extent N in a
JSON table. ``apply_all()`` hands the
whole table to the library, which then picks each call's entry by its volume
extent, so reconstructions of different sizes need no switching between
calls.

    python -m rexlib_amd.tune --n 256 --N 256 --launch 4096
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

# (x, y, z) extents of the bricks the library is built for.
SHAPES = ((16, 8, 6), (16, 8, 8), (16, 16, 8), (8, 8, 8), (8, 8, 6), (8, 8, 4), (8, 6, 6))
SHARES = (32, 64, 128, 256, 512, 1024)
DEFAULT = {"extent_x": 16, "extent_y": 8, "extent_z": 6, "share": 256}


def load_table(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def set_tuning(extent: int, shape, share: int, taps: int = 0) -> None:
    """The library's entry for volumes of extent ``extent`` (0: every other
    extent); with ``taps`` (1 nearest, 2 trilinear, 4 Kaiser-Bessel) the
    entry for that window family only."""
    from . import _native as N
    if taps:
        N.check(N.library().rexlib_hip_set_insertion_tuning_for_window(
            int(extent), int(taps), int(shape[0]), int(shape[1]), int(shape[2]), int(share)))
        return
    N.check(N.library().rexlib_hip_set_insertion_tuning(int(extent), int(shape[0]), int(shape[1]), int(shape[2]),
                                                         int(share)))


def get_tuning(extent: int):
    """((extent_x, extent_y, extent_z), share) that calls with volume extent ``extent`` use."""
    from . import _native as N
    v = [ctypes.c_int32() for _ in range(4)]
    N.check(N.library().rexlib_hip_get_insertion_tuning(int(extent), *[ctypes.byref(x) for x in v]))
    return (v[0].value, v[1].value, v[2].value), v[3].value


def clear_tuning() -> None:
    from . import _native as N
    N.library().rexlib_hip_clear_insertion_tuning()


def apply_all(path: str = TABLE) -> dict:
    """Hand every recorded entry to the library; returns them. Keys are N
    ("256"), or N and the window's taps per axis ("256/4")."""
    table = load_table(path)
    for key, entry in table.items():
        extent, _, taps = key.partition("/")
        set_tuning(int(extent), (entry["extent_x"], entry["extent_y"], entry["extent_z"]), entry["share"],
                   int(taps) if taps else 0)
    return table


def workload(n: int, big: int, launch: int, seed: int = 0):
    import torch

    from .synthetic import ctf_parameters, transforms_of, uniform_rotations
    rng = np.random.default_rng(seed)
    h = n // 2 + 1
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    slices = torch.randn((launch, n, h), dtype=torch.complex64, device="cuda", generator=g)
    transforms = torch.from_numpy(transforms_of(uniform_rotations(launch, rng)).astype(np.float32)).cuda()
    ctf = torch.from_numpy(ctf_parameters(launch, rng).astype(np.float32)).cuda()
    return slices, transforms, ctf


def window_of(taps: int):
    """The window the tuner times for a family of `taps` per axis (0 and 2:
    trilinear; 1: nearest; 4: Kaiser-Bessel of half-width 2, shape 7.4)."""
    from . import em
    if taps == 1:
        return em.GriddingWindow.nearest()
    if taps == 4:
        return em.GriddingWindow.kaiser_bessel(2.0, 7.4)
    return em.GriddingWindow.linear()


def time_insertion(n, big, slices, transforms, ctf, repeats=3, taps=0):
    """Best time of a CTF insertion into zeroed accumulators, and its sums."""
    import torch

    from . import em
    h = n // 2 + 1
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    radial = torch.ones(h, device="cuda")
    window = window_of(taps)
    optics = em.CtfOptics(300.0, 2.7, 0.1)

    def run():
        em.insert_ctf_slices(num, den, slices, transforms, ctf, radial, window, optics, 1.5,
                             numerator_out=num, denominator_out=den)

    run()
    torch.cuda.synchronize()
    result = (num.clone(), den.clone())
    best = float("inf")
    for _ in range(repeats):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best, result


def rel_l2(a, b) -> float:
    import torch
    return float(torch.linalg.vector_norm(a - b) / torch.clamp(torch.linalg.vector_norm(b), min=1e-30))


def tune(n: int, big: int, launch: int, path: str = TABLE, verbose: bool = True, taps: int = 0) -> dict:
    """Time every brick shape and share; record the fastest under "N" (taps
    0: every window, timed with trilinear) or "N/taps" (one window family)."""
    slices, transforms, ctf = workload(n, big, launch)
    results = []
    reference = None
    previous = get_tuning(big)
    try:
        for shape in SHAPES:
            for share in SHARES:
                set_tuning(big, shape, share, taps)
                ms, sums = time_insertion(n, big, slices, transforms, ctf, taps=taps)
                if reference is None:
                    reference = sums
                error = max(rel_l2(sums[0], reference[0]), rel_l2(sums[1], reference[1]))
                if not error <= 1e-5:
                    raise RuntimeError(f"brick shape {shape}, share {share}: sums differ from "
                                       f"the first configuration's by rel-L2 {error:.3e}")
                results.append({"extent_x": shape[0], "extent_y": shape[1], "extent_z": shape[2], "share": share,
                                "ms": ms, "rel_l2": error})
                if verbose:
                    print(f"n={n} N={big} L={launch} brick {shape[0]}x{shape[1]}x{shape[2]} share={share}: "
                          f"{ms:.3f} ms (rel-L2 to the first {error:.1e})", flush=True)
    finally:
        set_tuning(big, *previous)
    best = min(results, key=lambda r: r["ms"])
    table = load_table(path)
    table[f"{big}/{taps}" if taps else str(big)] = {
        "extent_x": best["extent_x"], "extent_y": best["extent_y"], "extent_z": best["extent_z"],
        "share": best["share"], "n": n, "launch": launch, "ms": best["ms"],
        "images_per_s": launch / (best["ms"] * 1e-3), "tuned": time.strftime("%Y-%m-%d")}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    set_tuning(big, (best["extent_x"], best["extent_y"], best["extent_z"]), best["share"], taps)
    return {"best": best, "all": results}


def main():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--n", type=int, nargs="+", default=[256])
    parser.add_argument("--N", type=int, nargs="+", default=[256])
    parser.add_argument("--launch", type=int, default=4096)
    parser.add_argument("--table", default=TABLE)
    parser.add_argument("--taps", type=int, default=0, choices=(0, 1, 2, 4),
                        help="tune one window family (1 nearest, 2 trilinear, 4 Kaiser-Bessel); 0: every window")
    args = parser.parse_args()
    for n, big in zip(args.n, args.N):
        out = tune(n, big, args.launch, args.table, taps=args.taps)
        print(json.dumps(out["best"]), flush=True)


if __name__ == "__main__":
    main()
