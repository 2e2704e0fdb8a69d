"""Compile-time limits of the kernels, read from ``include/fpx_kernel_limits.h``.

Mirrors forcepho/kernel_limits.py:21-38: the header is the single source of
truth and every ``#define NAME value`` line becomes a module constant
(MAXBANDS, MAXSOURCES, NPARAMS, MAX_EXP_ARG, MAXRADII, FPX_WAVE, ...).
"""

import os
import re

_HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "include", "fpx_kernel_limits.h")

_regexp = re.compile(r"^\s*#define\s+(?P<name>\w+)\s+(?P<value>[-+0-9.eE]+)\s*$")

macros = {}
with open(_HEADER, "r") as _fp:
    for _line in _fp:
        _m = _regexp.match(_line)
        if _m:
            _v = _m.group("value")
            try:
                macros[_m.group("name")] = int(_v)
            except ValueError:
                macros[_m.group("name")] = float(_v)

locals().update(macros)

__all__ = list(macros)
