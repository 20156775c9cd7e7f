"""Where rexlib finds this distribution's plugin.

rexlib's Python package loads the plugins that installed distributions
declare (python/rexlib/_plugins.py:80-97): every entry point in the group
``rexlib.plugins`` names a package, and every shared object directly in a
directory named ``rexlib-plugins``, at any depth of that package, is a
plugin. This distribution declares ``hip = "rexlib_amd"`` (pyproject.toml)
and its build writes ``rexlib_amd/rexlib-plugins/librexlib-hip.so``.

``plugin_files`` applies the same rules to one package; ``discover`` walks
the installed entry points as rexlib does, without loading anything.
"""
from __future__ import annotations

import importlib.metadata
import importlib.util
import pathlib
import sys

ENTRY_POINT_GROUP = "rexlib.plugins"
PLUGIN_DIRECTORY_NAME = "rexlib-plugins"


def shared_object_suffixes() -> tuple[str, ...]:
    return {"win32": (".dll",), "darwin": (".so", ".dylib")}.get(sys.platform, (".so",))


def package_directories(package: str) -> list[pathlib.Path]:
    """The directories a package spans, found without importing it."""
    spec = importlib.util.find_spec(package)
    if spec is None or spec.submodule_search_locations is None:
        raise ModuleNotFoundError(f"No package named {package!r} is installed", name=package)
    return [pathlib.Path(location) for location in spec.submodule_search_locations]


def plugin_files(package: str = "rexlib_amd") -> list[pathlib.Path]:
    """The shared objects directly in a ``rexlib-plugins`` directory, at any
    depth of the package, sorted."""
    suffixes = shared_object_suffixes()
    return sorted(path for directory in package_directories(package)
                  for plugins in directory.rglob(PLUGIN_DIRECTORY_NAME) if plugins.is_dir()
                  for path in plugins.iterdir() if path.suffix in suffixes and path.is_file())


def discover() -> dict[str, list[pathlib.Path]]:
    """Plugin files per installed entry point of the group (name -> files)."""
    found = {}
    for entry_point in importlib.metadata.entry_points(group=ENTRY_POINT_GROUP):
        try:
            found[entry_point.name] = plugin_files(entry_point.module)
        except ImportError:
            found[entry_point.name] = []
    return found
