"""Minimal forward-mode dual numbers over numpy arrays (value, d/dx).

TEST INFRASTRUCTURE (see ``oracle/__init__.py``).  Used by the oracle waveform
to obtain exact frequency derivatives (dA/du, dPsi/du at the interpolation
nodes) without finite differences.
"""

from __future__ import annotations

import numpy as np


class D:
    __slots__ = ("v", "d")
    __array_priority__ = 1000

    def __init__(self, v, d=0.0):
        self.v = np.asarray(v, dtype=np.float64)
        self.d = np.asarray(d, dtype=np.float64) * np.ones_like(self.v)

    # arithmetic -------------------------------------------------------------
    def __add__(self, o):
        o = lift(o)
        return D(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = lift(o)
        return D(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        return lift(o) - self

    def __neg__(self):
        return D(-self.v, -self.d)

    def __mul__(self, o):
        o = lift(o)
        return D(self.v * o.v, self.d * o.v + self.v * o.d)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = lift(o)
        return D(self.v / o.v, (self.d * o.v - self.v * o.d) / (o.v * o.v))

    def __rtruediv__(self, o):
        return lift(o) / self

    def __pow__(self, p):
        if isinstance(p, D):
            raise TypeError("dual exponent not supported")
        return D(self.v ** p, p * self.v ** (p - 1) * self.d)


def lift(x):
    return x if isinstance(x, D) else D(x, 0.0)


def val(x):
    return x.v if isinstance(x, D) else np.asarray(x, dtype=np.float64)


def der(x):
    return x.d if isinstance(x, D) else np.zeros_like(np.asarray(x, dtype=np.float64))


def log(x):
    if isinstance(x, D):
        return D(np.log(x.v), x.d / x.v)
    return np.log(x)


def exp(x):
    if isinstance(x, D):
        e = np.exp(x.v)
        return D(e, e * x.d)
    return np.exp(x)


def sqrt(x):
    if isinstance(x, D):
        s = np.sqrt(x.v)
        return D(s, 0.5 * x.d / s)
    return np.sqrt(x)


def cos(x):
    if isinstance(x, D):
        return D(np.cos(x.v), -np.sin(x.v) * x.d)
    return np.cos(x)


def atan(x):
    if isinstance(x, D):
        return D(np.arctan(x.v), x.d / (1.0 + x.v * x.v))
    return np.arctan(x)


def where(c, a, b):
    a, b = lift(a), lift(b)
    return D(np.where(c, a.v, b.v), np.where(c, a.d, b.d))
