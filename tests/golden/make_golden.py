"""Write the golden insertion vectors from the reference's own CPU kernels.

Run in a container that holds the reference tree, after `make -C oracle`:

    python tests/golden/make_golden.py

Each case of tests/cases.py (the fixture shapes of
tests/unitary/src/backends/cpu/builders/em/reconstruction/test_slice_insertion_program_builder.cpp:110-219)
is inserted by oracle/_ref/librexlib_ref.so — the reference's
row_insertion_kernel, windows, CTF and volume phases compiled from its
sources — in float32 and in float64. The inputs and both results are stored
as .npz (no pickles), so the GPU tests can compare against the reference's
results where the reference tree is absent.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from cases import PIXEL_SIZE, SIZES, make_case  # noqa: E402

WINDOWS = {
    "nearest": oracle.Window.nearest(),
    "linear": oracle.Window.linear(),
    "kb1.5_6": oracle.Window.kaiser_bessel(1.5, 6.0),
    "kb2_8": oracle.Window.kaiser_bessel(2.0, 8.0),
}


def cases():
    for n, big in SIZES:
        for w in ("nearest", "linear", "kb1.5_6"):
            yield f"plain_{n}_{big}_{w}", "plain", n, big, w, 1, {}
    for variant in ("weighted", "ctf"):
        for n, big in [(6, 12), (5, 7)]:
            for w in ("linear", "kb2_8"):
                yield f"{variant}_{n}_{big}_{w}", variant, n, big, w, 5, {}
    yield "batch_2x3", "plain", 6, 8, "linear", 3, dict(accumulator_batch=(2, 1), transform_batch=(2, 3))
    yield "float32_16_linear_ctf", "ctf", 8, 16, "linear", 2, {}


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/librexlib_ref.so is not built")
    out = {}
    for name, variant, n, big, w, seed, kw in cases():
        inputs = make_case(variant, n, big, seed, **kw)
        r32 = oracle.ref_insert(variant, inputs, WINDOWS[w], pixel_size=PIXEL_SIZE, dtype=np.float32)
        r64 = oracle.ref_insert(variant, inputs, WINDOWS[w], pixel_size=PIXEL_SIZE, dtype=np.float64)
        for i, a in enumerate(inputs):
            out[f"{name}/in{i}"] = a
        out[f"{name}/num32"] = r32[0]
        out[f"{name}/den32"] = r32[1]
        out[f"{name}/num64"] = r64[0]
        out[f"{name}/den64"] = r64[1]
        out[f"{name}/meta"] = np.array([n, big, len(inputs)])
        out[f"{name}/variant"] = np.array(variant)
        out[f"{name}/window"] = np.array(w)
    path = os.path.join(HERE, "insertion_reference.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
