"""Julia Base semantics the reference's integer outputs depend on.

The reference computes lag grids, bin edges, snapshot indices and the
coordinate grid with Julia's float ranges (``range(a, b, length=n)`` and
``a:s:b``), which use TwicePrecision arithmetic and an exact-rational path
(``Base.rat``).  Those choices decide the last bits of the range elements and,
through ``round``/``argmin``, the integer lag steps.  This module restates
them:

- ``rat``: Base.rat (continued fractions bounded by maxintfloat(Float32)).
- ``range_length``: range_start_stop_length for IEEE floats.  When both ends
  have exact rational forms the elements are the rationals rounded to Float64;
  otherwise ``_linspace``'s double-double elements, modelled as the exact value
  of start + (stop-start)(i-1)/(len-1) rounded once.
- ``colon``: ``(:)(start, step, stop)`` for IEEE floats, same two paths.
- ``round_int``: ``round(Int, x)``, ties to even.

The TwicePrecision results equal the once-rounded exact values except when
the exact value lies within 2^-106 relative of a rounding boundary; this is the
model's stated approximation.
"""
from __future__ import annotations

import math
from fractions import Fraction

MAXINTFLOAT_F32 = 16777216          # maxintfloat(Float32, Int): rat's bound for Float64
MAXINTFLOAT_F64 = 9007199254740992  # maxintfloat(Float64, Int)


def rat(x: float):
    """Base.rat(x) for Float64 x: (numerator, denominator) of a continued
    fraction approximation with both bounded by maxintfloat(Float32)."""
    y = float(x)
    a = d = 1
    b = c = 0
    m = MAXINTFLOAT_F32
    while abs(y) <= m:
        f = int(math.trunc(y))
        y -= f
        a, c = f * a + c, a
        b, d = f * b + d, b
        if max(abs(a), abs(b)) > m:
            return c, d
        if b != 0 and float(a) / float(b) == x:
            break
        if b == 0 and a != 0 and math.copysign(math.inf, a) == x:
            break
        if y == 0.0:
            # inv(0.0) == Inf ends Julia's loop on the next test
            break
        y = 1.0 / y
    return a, b


def _div_trunc(a: int, b: int) -> int:
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _is_between(a, x, b) -> bool:
    return (a <= x <= b) or (b <= x <= a)


def _to_float_div(n: int, d: int) -> float:
    """Julia T(n/d) with Int n, d: n/d is Float64 division of the Int values."""
    return float(n) / float(d)


def range_length(start: float, stop: float, length: int) -> list[float]:
    """collect(range(start, stop, length=length)) for Float64 ends."""
    start = float(start)
    stop = float(stop)
    n = int(length)
    if n < 0:
        raise ValueError("range length cannot be negative")
    if n == 0:
        return []
    if n == 1:
        if start != stop:
            raise ValueError("range(start, stop, length=1) needs start == stop")
        return [start]
    sn, sd = rat(start)
    en, ed = rat(stop)
    if sd != 0 and ed != 0:
        den = sd * ed // math.gcd(sd, ed)
        if den != 0 and abs(den * start) <= MAXINTFLOAT_F64 and abs(den * stop) <= MAXINTFLOAT_F64:
            s_n = round(den * start)
            e_n = round(den * stop)
            if _to_float_div(s_n, den) == start and _to_float_div(e_n, den) == stop:
                return [float(Fraction(s_n * (n - 1 - i) + e_n * i, (n - 1) * den)) for i in range(n)]
    fs, fe = Fraction(start), Fraction(stop)
    out = [float(fs + (fe - fs) * i / (n - 1)) for i in range(n)]
    out[0], out[-1] = start, stop
    return out


def range_int_length(start: int, stop: int, length: int) -> list[float]:
    """collect(range(start::Int, stop::Int, length=n)): Float64 elements, exact
    rationals rounded once."""
    n = int(length)
    if n == 1:
        return [float(start)]
    return [float(Fraction(start * (n - 1 - i) + stop * i, n - 1)) for i in range(n)]


def colon(start: float, step: float, stop: float) -> list[float]:
    """collect(start:step:stop) for Float64 arguments."""
    start, step, stop = float(start), float(step), float(stop)
    if step == 0.0:
        raise ValueError("range step cannot be zero")
    pn, pd = rat(step)
    if pd != 0 and _to_float_div(pn, pd) == step:
        sn, sd = rat(start)
        en, ed = rat(stop)
        if sd != 0 and ed != 0 and _to_float_div(sn, sd) == start and _to_float_div(en, ed) == stop:
            den = sd * pd // math.gcd(sd, pd)
            if (den != 0 and abs(start * den) <= MAXINTFLOAT_F64 and abs(step * den) <= MAXINTFLOAT_F64
                    and den % sd == 0 and den % pd == 0):
                s_n = round(start * den)
                p_n = round(step * den)
                ln = max(0, _div_trunc(den * en - ed * s_n + p_n * ed, p_n * ed))
                if (_is_between(start, start + (ln - 1) * step, stop + step / 2)
                        and not _is_between(start, start + ln * step, stop)):
                    return [float(Fraction(s_n + i * p_n, den)) for i in range(ln)]
    lf = (stop - start) / step
    if lf < 0:
        ln = 0
    elif lf == 0:
        ln = 1
    else:
        ln = round(lf) + 1
        stop2 = start + (ln - 1) * step
        ln -= int(start < stop < stop2) + int(start > stop > stop2)
    fs, fp = Fraction(start), Fraction(step)
    return [float(fs + fp * i) for i in range(ln)]


def round_int(x: float) -> int:
    """round(Int, x): round half to even, error on non-finite."""
    if not math.isfinite(x):
        raise ValueError(f"InexactError: round(Int, {x})")
    return int(round(float(x)))


def unique_stable(values):
    seen = set()
    out = []
    for v in values:
        if v not in seen:
            seen.add(v)
            out.append(v)
    return out
