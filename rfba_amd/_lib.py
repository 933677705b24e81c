"""ctypes loading of the in-tree native libraries.

The product path has no fallback: if a library is missing, loading raises
with the build command to run.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)

_cache: dict[str, ctypes.CDLL] = {}


def load(name: str) -> ctypes.CDLL:
    """Load rfba_amd/lib<name>.so (srrf, 3b, synth)."""
    if name in _cache:
        return _cache[name]
    path = os.path.join(PKG, "lib%s.so" % name)
    if not os.path.exists(path):
        raise ImportError("%s is not built: run `python -m rfba_amd.build` (or __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path)
    _cache[name] = lib
    return lib


def lib_path(name: str) -> str:
    return os.path.join(PKG, "lib%s.so" % name)
