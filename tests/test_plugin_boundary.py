"""The plugin's source against rexlib's own headers.

rexlib_amd/host/hip_plugin.cpp is built here against restatements of the
declarations it uses (rexlib_amd/host/rexlib). This test type-checks the same
file against rexlib's installed header tree instead
(include/rexlib/core/plugin_definition.hpp, the hardware interfaces, the
operations), supplying only the version header that rexlib's CMake writes
from its VERSION file (cmake/config/version_macros.h.in). It runs where that
tree is present and is skipped elsewhere.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REXLIB_INCLUDE = "/root/reference/include"
VERSION_FILE = "/root/reference/VERSION"
PLUGIN_SOURCE = os.path.join(ROOT, "rexlib_amd", "host", "hip_plugin.cpp")
ROCM_INCLUDE = "/opt/rocm/include"


def version_header(directory):
    with open(VERSION_FILE) as f:
        major, minor, patch = f.read().strip().split(".")[:3]
    path = os.path.join(directory, "rexlib", "core")
    os.makedirs(path)
    with open(os.path.join(path, "version_macros.h"), "w") as f:
        f.write("#pragma once\n")
        f.write(f"#define REXLIB_VERSION_MAJOR {major}\n")
        f.write(f"#define REXLIB_VERSION_MINOR {minor}\n")
        f.write(f"#define REXLIB_VERSION_PATCH {patch}\n")


@pytest.mark.skipif(not os.path.isdir(REXLIB_INCLUDE) or shutil.which("g++") is None,
                    reason="rexlib's header tree (or g++) is not on this machine")
def test_plugin_source_compiles_against_rexlib_headers(tmp_path):
    version_header(str(tmp_path))
    cmd = ["g++", "-std=c++14", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__",
           "-I", str(tmp_path), "-I", REXLIB_INCLUDE, "-I", os.path.join(ROOT, "include"), "-I", ROCM_INCLUDE,
           PLUGIN_SOURCE]
    proc = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stderr[-4000:]


def test_plugin_source_names_only_rexlib_header_paths():
    """The plugin reaches rexlib through the installed paths only, so the
    check above sees rexlib's headers and not the restatements."""
    with open(PLUGIN_SOURCE) as f:
        includes = [line.strip() for line in f if line.startswith("#include")]
    rexlib = [i for i in includes if "rexlib/" in i and "rexlib_hip" not in i]
    assert rexlib and all(i.startswith("#include <rexlib/") for i in rexlib), rexlib
