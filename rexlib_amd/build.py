"""Build the gfx950 backend library in-tree: rexlib_amd/librexlib_hip.so.

One hipcc invocation compiles every HIP and C++ source of rexlib_amd/csrc for
gfx950 and links the shared object. Only the C ABI (rexlib_hip_*) and the
plugin entry points (rexlib_get_plugin*) are exported.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBRARY = os.path.join(HERE, "librexlib_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # Packed FP32 instructions stay off: ordinary float code only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and os.path.exists(candidate):
            return candidate
    raise RuntimeError("hipcc not found")


def sources():
    names = sorted(os.listdir(CSRC))
    return [os.path.join(CSRC, f) for f in names if f.endswith((".hip", ".cpp"))]


def build(verbose: bool = False) -> str:
    exports = os.path.join(CSRC, "exports.map")
    tmp = LIBRARY + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-shared", "-o", tmp, *sources(),
           f"-Wl,--version-script={exports}", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("hipcc failed to build librexlib_hip.so")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
