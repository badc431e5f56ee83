"""Locate and load the in-tree native libraries of the package.

Libraries are built in place (``python -m cusingler_amd.build``) next to this
file.  A missing library is an error: there is no Python fallback for any
native component.
"""

from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
_cache = {}


class NativeLibraryMissing(RuntimeError):
    pass


def path(name: str) -> str:
    return os.path.join(PKG_DIR, name)


def load(name: str) -> ctypes.CDLL:
    if name in _cache:
        return _cache[name]
    p = path(name)
    if not os.path.exists(p):
        raise NativeLibraryMissing(
            f"{p} is not built; run `python -m cusingler_amd.build` (or __graft_entry__.build())")
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
    _cache[name] = lib
    return lib
