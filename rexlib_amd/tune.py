"""Autotuner of the insertion's LDS brick schedule.

For a workload (slice extent n, volume extent N, slices per launch L) it times
the insertion on the device for every brick shape (extents along y and z;
the extent along x is 16; 6 along z goes with 8 along y) and share target (expected plane hits per task), and
records the fastest in a JSON table keyed by "n,N". ``apply(n, N)`` sets the
recorded choice in the library before an insertion.

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

EXTENTS_Y = (8, 16)
EXTENTS_Z = (6, 8, 16)
SHARES = (64, 96, 128, 192, 256, 512, 1024)


def load_table(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def set_tuning(extent_y: int, extent_z: int, share: int) -> None:
    from . import _native as N
    N.library().rexlib_hip_set_insertion_tuning(int(extent_y), int(extent_z), int(share))


def get_tuning():
    from . import _native as N
    w, d, s = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    N.library().rexlib_hip_get_insertion_tuning(ctypes.byref(w), ctypes.byref(d), ctypes.byref(s))
    return w.value, d.value, s.value


def apply(n: int, big: int, path: str = TABLE) -> tuple[int, int, int] | None:
    """Set the recorded choice for (n, N), or the nearest recorded N for n."""
    table = load_table(path)
    key = f"{n},{big}"
    entry = table.get(key)
    if entry is None:
        candidates = [(abs(int(k.split(",")[1]) - big), v) for k, v in table.items()]
        if not candidates:
            return None
        entry = min(candidates, key=lambda c: c[0])[1]
    set_tuning(entry["extent_y"], entry["extent_z"], entry["share"])
    return entry["extent_y"], entry["extent_z"], entry["share"]


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


def time_insertion(n, big, slices, transforms, ctf, repeats=3) -> float:
    import torch

    from . import em
    h = n // 2 + 1
    num = torch.zeros((big, big, big // 2 + 1), dtype=torch.complex64, device="cuda")
    den = torch.zeros((big, big, big // 2 + 1), dtype=torch.float32, device="cuda")
    radial = torch.ones(h, device="cuda")
    window = em.GriddingWindow.linear()
    optics = em.CtfOptics(300.0, 2.7, 0.1)

    def run():
        em.insert_ctf_slices(num, den, slices, transforms, ctf, radial, window, optics, 1.5,
                             numerator_out=num, denominator_out=den)

    run()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(repeats):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def tune(n: int, big: int, launch: int, path: str = TABLE, verbose: bool = True) -> dict:
    slices, transforms, ctf = workload(n, big, launch)
    results = []
    for ey in EXTENTS_Y:
        for ez in EXTENTS_Z:
            if ez == 6 and ey != 8:
                continue  # 6 along z is built for 8 along y only
            for share in SHARES:
                set_tuning(ey, ez, share)
                ms = time_insertion(n, big, slices, transforms, ctf)
                results.append({"extent_y": ey, "extent_z": ez, "share": share, "ms": ms})
                if verbose:
                    print(f"n={n} N={big} L={launch} extent_y={ey} extent_z={ez} share={share}: {ms:.3f} ms",
                          flush=True)
    best = min(results, key=lambda r: r["ms"])
    table = load_table(path)
    table[f"{n},{big}"] = {**best, "launch": launch, "images_per_s": launch / (best["ms"] * 1e-3),
                           "tuned": time.strftime("%Y-%m-%d")}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    set_tuning(best["extent_y"], best["extent_z"], best["share"])
    return {"best": best, "all": results}


def main():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--n", type=int, nargs="+", default=[256])
    parser.add_argument("--N", type=int, nargs="+", default=[256])
    parser.add_argument("--launch", type=int, default=4096)
    parser.add_argument("--table", default=TABLE)
    args = parser.parse_args()
    for n, big in zip(args.n, args.N):
        out = tune(n, big, args.launch, args.table)
        print(json.dumps(out["best"]), flush=True)


if __name__ == "__main__":
    main()
