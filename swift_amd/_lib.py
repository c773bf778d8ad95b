"""Loader of libswsph.so (the HIP kernels behind the C ABI in include/swsph.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C swift_amd/csrc``).  There is no CPU fallback: when the library is
missing or cannot be loaded every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
# SWSPH_LIBRARY selects another build of the library (e.g. one compiled with
# other tuning knobs).
LIB_PATH = os.environ.get("SWSPH_LIBRARY") or os.path.join(_HERE, "libswsph.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "swsph.h")
_lib = None


class SwsphError(RuntimeError):
    """A non-zero status from libswsph (the reference would call error())."""


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SwsphError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(there is no CPU fallback for the SPH pass)")
        _lib = C.CDLL(LIB_PATH)
        _lib.swsph_last_error.restype = C.c_char_p
        _lib.swsph_server_last_error.restype = C.c_char_p
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().swsph_last_error().decode(errors="replace")
        raise SwsphError(f"{what}: status {rc}: {msg}")


def declared_symbols() -> list:
    """Function names declared in include/swsph.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(swsph_\w+)\s*\(",
                                 text, re.M)))
