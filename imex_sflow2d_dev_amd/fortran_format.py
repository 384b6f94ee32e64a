"""Fortran edit descriptors of the reference's output files, for the host
writers: Fw.d, ESw.dEe, Ew.d and Aw output (ISO/IEC 1539 13.7.2).

- Fw.d: d digits after the point, right-justified in w; a value that does
  not fit prints w asterisks.  Negative values that round to zero keep the
  minus sign.
- ESw.dEe: one non-zero digit before the point (0.0 prints as 0.ddd),
  d after it, exponent with e digits; asterisks when the exponent needs
  more digits or the field is too narrow.
- Ew.d (no Ee): 0.ddd form with d significant digits; the exponent is
  written E+nn for |exp| <= 99 and +nnn (no E) for 99 < |exp| <= 999.
- Aw: right-justified with leading blanks when w exceeds the length, the
  first w characters otherwise.

Decimal conversion is Python's correctly rounded one (ties to even).
tests/test_fortran_format.py checks every function against the output of
the same descriptors from the Fortran runtime (tests/golden/fortran_formats.*).
"""
from __future__ import annotations

import math


def _special(x: float, w: int):
    if math.isnan(x):
        return "NaN".rjust(w) if w >= 3 else "*" * w
    if math.isinf(x):
        s = ("-" if x < 0 else "") + ("Infinity" if w >= 9 else "Inf")
        return s.rjust(w) if len(s) <= w else "*" * w
    return None


def f_fmt(x: float, w: int, d: int) -> str:
    """Fw.d"""
    sp = _special(x, w)
    if sp is not None:
        return sp
    s = f"{x:.{d}f}"
    if s.startswith("-") and float(s) == 0.0 and not math.copysign(1.0, x) < 0:
        s = s[1:]
    if len(s) > w and s.startswith("0.") and len(s) == w + 1:
        s = s[1:]
    if len(s) > w and s.startswith("-0.") and len(s) == w + 1:
        s = "-" + s[2:]
    return s.rjust(w) if len(s) <= w else "*" * w


def es_fmt(x: float, w: int, d: int, e: int = 2) -> str:
    """ESw.dEe"""
    sp = _special(x, w)
    if sp is not None:
        return sp
    s = f"{x:.{d}E}"
    mant, ex = s.split("E")
    n = int(ex)
    if len(str(abs(n))) > e:
        return "*" * w
    out = f"{mant}E{'-' if n < 0 else '+'}{abs(n):0{e}d}"
    return out.rjust(w) if len(out) <= w else "*" * w


def e_fmt(x: float, w: int, d: int) -> str:
    """Ew.d: 0.d1d2...dd x 10^n"""
    sp = _special(x, w)
    if sp is not None:
        return sp
    neg = math.copysign(1.0, x) < 0
    if x == 0.0:
        digits, n = "0" * d, 0
    else:
        s = f"{abs(x):.{d - 1}E}"
        mant, ex = s.split("E")
        digits = mant.replace(".", "")
        n = int(ex) + 1
    if abs(n) <= 99:
        exs = f"E{'-' if n < 0 else '+'}{abs(n):02d}"
    elif abs(n) <= 999:
        exs = f"{'-' if n < 0 else '+'}{abs(n):03d}"
    else:
        return "*" * w
    out = ("-" if neg else "") + "0." + digits + exs
    if len(out) > w and out.startswith("0."):
        out = out[1:]
    if len(out) > w and out.startswith("-0."):
        out = "-" + out[2:]
    return out.rjust(w) if len(out) <= w else "*" * w


def a_fmt(s: str, w: int) -> str:
    """Aw"""
    return s.rjust(w) if len(s) <= w else s[:w]
