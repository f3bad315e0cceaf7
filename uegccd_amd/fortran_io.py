"""Fortran list-directed output records, as the reference's `write(unit,*)`
lines in fort.59 are printed (gfortran): a record opens with one blank, items
are separated by one blank, REAL(8) items are edited as 1PG25.17E3 and default
INTEGER items in 11 columns, character items as they are."""
from __future__ import annotations

import math


def real8(x: float) -> str:
    """One REAL(8) item, 25 columns: F form (17 significant digits, right
    justified in 20, then 5 blanks) for 0.1 <= |x| < 1e17 and for zero, else
    E form with a 3-digit exponent, right justified in 25."""
    x = float(x)
    if x == 0.0:
        return f"{'-' if math.copysign(1.0, x) < 0 else ''}0.0000000000000000".rjust(20) + " " * 5
    mant, exp = f"{x:.16e}".split("e")
    k = int(exp) + 1                       # integer digits after rounding to 17 significant
    if 0 <= k <= 17:
        return f"{x:.{17 - k}f}".rjust(20) + " " * 5
    return f"{mant}E{int(exp):+04d}".rjust(25)


def integer(i: int) -> str:
    return f"{int(i):>11}"


def record(*items) -> str:
    """One list-directed record; arrays (lists, tuples, numpy arrays) are
    written element by element."""
    fields = []
    for it in items:
        if isinstance(it, str):
            fields.append(it)
        elif isinstance(it, (int,)) and not isinstance(it, bool):
            fields.append(integer(it))
        elif hasattr(it, "__len__"):
            for v in it:
                fields.append(integer(v) if isinstance(v, int) else real8(v))
        elif hasattr(it, "dtype") and it.dtype.kind in "iu":
            fields.append(integer(int(it)))
        else:
            fields.append(real8(it))
    return " " + " ".join(fields)
