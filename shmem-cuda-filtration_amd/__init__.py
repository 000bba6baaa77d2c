"""MI355X (gfx950) backend for the shmem-cuda-filtration time step.

The product is the C ABI in include/filtration_hip.h (lib/libfiltration_hip.so)
and the defines.h-compatible C++ shims over it (lib/libfiltration_defines.so).
This package adds Python access to that ABI (abi), host field bundles (grid),
and device-side helpers (device).
"""
import importlib
import os
import sys

from . import abi, device, grid, hiprt, workloads  # noqa: F401
from .abi import FiltParams, FiltFields, load_library, default_params_py  # noqa: F401
from .grid import HostGrid  # noqa: F401

PKG_DIR = abi.PKG_DIR


def import_self():
    """Return this package as a module (its directory name is not a Python
    identifier, so callers import it through importlib)."""
    return sys.modules[__name__]


def load(repo_root=None):
    root = repo_root or os.path.dirname(PKG_DIR)
    if root not in sys.path:
        sys.path.insert(0, root)
    return importlib.import_module(os.path.basename(PKG_DIR))
