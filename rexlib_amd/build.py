"""Build the gfx950 backend in-tree.

- rexlib_amd/librexlib_hip.so: one hipcc invocation compiles every HIP and C++
  source of rexlib_amd/csrc for gfx950; only the C ABI (rexlib_hip_*) is
  exported.
- rexlib_amd/librexlib_host.so: the C++ core of the mirrored rexlib API
  (rexlib_amd/host/core.cpp): services, plugin loader, arrays, the eager
  dispatcher and the em / Fourier functions.
- rexlib_amd/rexlib-plugins/librexlib-hip.so: the plugin
  (rexlib_amd/host/hip_plugin.cpp), exporting rexlib_get_plugin_compatibility
  and rexlib_get_plugin; it runs librexlib_hip.so's kernels through the C ABI.
  It sits in a directory named rexlib-plugins inside the package, where
  rexlib's Python discovery looks for the plugins a distribution declares in
  the rexlib.plugins entry-point group (pyproject.toml).
- tests/plugin_driver: the C++ caller the plugin tests run.
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


ROOT = os.path.dirname(HERE)
HOST = os.path.join(HERE, "host")
HOST_LIBRARY = os.path.join(HERE, "librexlib_host.so")
PLUGIN_DIR = os.path.join(HERE, "rexlib-plugins")
PLUGIN = os.path.join(PLUGIN_DIR, "librexlib-hip.so")
DRIVER = os.path.join(ROOT, "tests", "plugin_driver")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
CXX_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"failed to build {what}")


def build_host(verbose: bool = False) -> list:
    """The mirrored rexlib core, the plugin and the test driver (host code, g++)."""
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, *CXX_FLAGS, "-fPIC", "-shared", "-fvisibility=hidden", "-I", HOST,
          os.path.join(HOST, "core.cpp"), "-ldl", "-o", HOST_LIBRARY + ".tmp"], "librexlib_host.so", verbose)
    os.replace(HOST_LIBRARY + ".tmp", HOST_LIBRARY)
    os.makedirs(PLUGIN_DIR, exist_ok=True)
    _run([cxx, *CXX_FLAGS, "-fPIC", "-shared", "-fvisibility=hidden", "-D__HIP_PLATFORM_AMD__",
          "-I", HOST, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROCM, "include"),
          os.path.join(HOST, "hip_plugin.cpp"), "-L", HERE, "-lrexlib_host", "-l:librexlib_hip.so",
          "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM}/lib",
          "-Wl,--no-undefined", "-o", PLUGIN + ".tmp"], "librexlib-hip.so", verbose)
    os.replace(PLUGIN + ".tmp", PLUGIN)
    _run([cxx, *CXX_FLAGS, "-I", HOST, os.path.join(ROOT, "tests", "plugin_driver.cpp"), "-L", HERE,
          "-lrexlib_host", "-ldl", "-Wl,-rpath,$ORIGIN/../rexlib_amd", "-o", DRIVER + ".tmp"], "plugin_driver",
         verbose)
    os.replace(DRIVER + ".tmp", DRIVER)
    return [HOST_LIBRARY, PLUGIN, DRIVER]


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
    build_host(verbose)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
