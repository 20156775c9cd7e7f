"""The Python packaging of the plugin: setup.cfg declares the
``rexlib.plugins`` entry point, the package holds the plugin where rexlib's
discovery looks (a directory named ``rexlib-plugins``,
python/rexlib/_plugins.py:80-97), and a wheel built offline from the tree
carries both."""
import configparser
import os
import subprocess
import sys
import zipfile

import pytest

try:
    import tomllib
except ImportError:  # Python 3.10
    import tomli as tomllib

from rexlib_amd import plugins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_the_distribution_declares_the_entry_point():
    with open(os.path.join(ROOT, "pyproject.toml"), "rb") as f:
        assert tomllib.load(f)["build-system"]["build-backend"] == "setuptools.build_meta"
    cfg = configparser.ConfigParser()
    cfg.read(os.path.join(ROOT, "setup.cfg"))
    declared = [line.strip() for line in cfg["options.entry_points"][plugins.ENTRY_POINT_GROUP].splitlines()
                if line.strip()]
    assert declared == ["hip = rexlib_amd"]
    assert "rexlib-plugins/*.so" in cfg["options.package_data"]["rexlib_amd"]


def test_discovery_rules_find_the_plugin_and_only_it():
    files = plugins.plugin_files("rexlib_amd")
    assert [f.name for f in files] == ["librexlib-hip.so"]
    assert files[0].parent.name == plugins.PLUGIN_DIRECTORY_NAME
    # The C ABI library and the mirrored core sit beside the package, not in
    # the plugin directory: they are not plugins.
    assert not any(f.name in ("librexlib_hip.so", "librexlib_host.so") for f in files)
    with pytest.raises(ModuleNotFoundError):
        plugins.plugin_files("no_such_package_here")


def test_offline_wheel_carries_the_plugin_and_the_entry_point(tmp_path):
    proc = subprocess.run([sys.executable, "-m", "pip", "wheel", ROOT, "--no-deps", "--no-build-isolation",
                           "--no-index", "-w", str(tmp_path), "-q"], capture_output=True, text=True, timeout=600,
                          cwd=str(tmp_path))
    assert proc.returncode == 0, proc.stderr[-2000:]
    wheels = list(tmp_path.glob("rexlib_hip-0.1.0-*.whl"))
    assert len(wheels) == 1
    with zipfile.ZipFile(wheels[0]) as w:
        names = w.namelist()
        assert "rexlib_amd/rexlib-plugins/librexlib-hip.so" in names
        assert "rexlib_amd/librexlib_hip.so" in names and "rexlib_amd/librexlib_host.so" in names
        entry = next(n for n in names if n.endswith("entry_points.txt"))
        text = w.read(entry).decode()
    assert "[rexlib.plugins]" in text and "hip = rexlib_amd" in text
