"""Forward-mode dual numbers over NumPy arrays, for the oracle's tangent.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU baseline, never by the product package.

``D(v, d)`` carries a value and one tangent.  The non-smooth primitives
follow JAX's JVP rules, which define the reference's tangent
(``src/jaxrans/solvers/implicit.py:105-108`` linearizes the residual with
``jax.linearize``):

* ``maximum``/``minimum``: the derivative of the selected argument; at a tie,
  half of each (JAX's balanced rule for ``lax.max``/``lax.min``);
* ``abs``: ``+d`` where ``x >= 0``, ``-d`` where ``x < 0``;
* ``where``: the tangent of the selected branch only;
* ``x ** c`` (constant ``c``): ``c x^(c-1) d``; ``sqrt``: ``d * (0.5 / sqrt(x))``.

These tie conventions are what the device kernels implement as well; no
golden vector from the reference pins them (parity unpinned at exact ties).
"""

from __future__ import annotations

import numpy as np


class D:
    __slots__ = ("v", "d")
    __array_ufunc__ = None          # ndarray (op) D -> D.__r<op>__

    def __init__(self, v, d):
        self.v = np.asarray(v, dtype=np.float64)
        self.d = np.asarray(d, dtype=np.float64)

    # --- container protocol ---
    def __getitem__(self, k):
        return D(self.v[k], self.d[k])

    @property
    def shape(self):
        return self.v.shape

    @property
    def ndim(self):
        return self.v.ndim

    def reshape(self, *s):
        return D(self.v.reshape(*s), self.d.reshape(*s))

    def __len__(self):
        return len(self.v)

    # --- arithmetic ---
    def __add__(self, o):
        if isinstance(o, D):
            return D(self.v + o.v, self.d + o.d)
        return D(self.v + o, self.d + 0.0 * np.asarray(o))

    __radd__ = __add__

    def __sub__(self, o):
        if isinstance(o, D):
            return D(self.v - o.v, self.d - o.d)
        return D(self.v - o, self.d + 0.0 * np.asarray(o))

    def __rsub__(self, o):
        return D(o - self.v, -self.d + 0.0 * np.asarray(o))

    def __neg__(self):
        return D(-self.v, -self.d)

    def __mul__(self, o):
        if isinstance(o, D):
            return D(self.v * o.v, self.d * o.v + self.v * o.d)
        return D(self.v * o, self.d * o)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, D):
            q = self.v / o.v
            return D(q, (self.d - q * o.d) / o.v)
        return D(self.v / o, self.d / o)

    def __rtruediv__(self, o):
        q = o / self.v
        return D(q, -q * self.d / self.v)

    def __pow__(self, c):
        if isinstance(c, D):
            raise TypeError("only constant exponents")
        if isinstance(c, int) or (float(c).is_integer() and abs(c) < 16):
            n = int(c)
            return D(self.v ** n, n * self.v ** (n - 1) * self.d if n != 0 else 0.0 * self.d)
        return D(self.v ** c, c * self.v ** (c - 1.0) * self.d)

    # comparisons act on values (masks, no tangent)
    def __gt__(self, o):
        return self.v > _val(o)

    def __ge__(self, o):
        return self.v >= _val(o)

    def __lt__(self, o):
        return self.v < _val(o)

    def __le__(self, o):
        return self.v <= _val(o)


def _val(x):
    return x.v if isinstance(x, D) else x


def _tan(x, like):
    return x.d if isinstance(x, D) else np.zeros(np.broadcast_shapes(np.shape(x), np.shape(like)))


def is_dual(*xs) -> bool:
    return any(isinstance(x, D) for x in xs)


def value(x):
    return x.v if isinstance(x, D) else np.asarray(x)


def tangent(x):
    return x.d if isinstance(x, D) else np.zeros_like(np.asarray(x, dtype=np.float64))


def sqrt(x):
    if isinstance(x, D):
        r = np.sqrt(x.v)
        return D(r, x.d * (0.5 / r))
    return np.sqrt(x)


def fabs(x):
    if isinstance(x, D):
        return D(np.abs(x.v), np.where(x.v >= 0.0, x.d, -x.d))
    return np.abs(x)


def maximum(a, b):
    if not is_dual(a, b):
        return np.maximum(a, b)
    av, bv = _val(a), _val(b)
    v = np.maximum(av, bv)
    da, db = _tan(a, bv), _tan(b, av)
    wa = np.where(av == v, np.where(bv == v, 0.5, 1.0), 0.0)
    wb = np.where(bv == v, np.where(av == v, 0.5, 1.0), 0.0)
    return D(v, wa * da + wb * db)


def minimum(a, b):
    if not is_dual(a, b):
        return np.minimum(a, b)
    av, bv = _val(a), _val(b)
    v = np.minimum(av, bv)
    da, db = _tan(a, bv), _tan(b, av)
    wa = np.where(av == v, np.where(bv == v, 0.5, 1.0), 0.0)
    wb = np.where(bv == v, np.where(av == v, 0.5, 1.0), 0.0)
    return D(v, wa * da + wb * db)


def where(c, a, b):
    c = np.asarray(c)
    if not is_dual(a, b):
        return np.where(c, a, b)
    av, bv = _val(a), _val(b)
    return D(np.where(c, av, bv), np.where(c, _tan(a, bv), _tan(b, av)))


def stack(xs, axis=0):
    if not is_dual(*xs):
        return np.stack(xs, axis=axis)
    shape = np.broadcast_shapes(*(np.shape(_val(x)) for x in xs))
    return D(np.stack([np.broadcast_to(_val(x), shape) for x in xs], axis=axis),
             np.stack([np.broadcast_to(_tan(x, np.zeros(shape)), shape) for x in xs], axis=axis))


def concatenate(xs, axis=0):
    if not is_dual(*xs):
        return np.concatenate(xs, axis=axis)
    return D(np.concatenate([_val(x) for x in xs], axis=axis),
             np.concatenate([_tan(x, _val(x)) for x in xs], axis=axis))


def take(x, idx, axis=-1):
    if isinstance(x, D):
        return D(np.take(x.v, idx, axis=axis), np.take(x.d, idx, axis=axis))
    return np.take(x, idx, axis=axis)


def pad_edge(x, widths):
    if isinstance(x, D):
        return D(np.pad(x.v, widths, mode="edge"), np.pad(x.d, widths, mode="edge"))
    return np.pad(x, widths, mode="edge")


def isfinite(x):
    return np.isfinite(_val(x))


# --- the freestream as a function of the flow parameters (sensitivity) --------------

def sin(x):
    if isinstance(x, D):
        return D(np.sin(x.v), np.cos(x.v) * x.d)
    return np.sin(x)


def cos(x):
    if isinstance(x, D):
        return D(np.cos(x.v), -np.sin(x.v) * x.d)
    return np.cos(x)


def arctan2(y, x):
    """``atan2`` with the tangent ``(x dy - y dx) / (x^2 + y^2)``."""
    if not is_dual(y, x):
        return np.arctan2(y, x)
    yv, xv = _val(y), _val(x)
    return D(np.arctan2(yv, xv), (xv * _tan(y, xv) - yv * _tan(x, yv)) / (xv * xv + yv * yv))


def scalar(x):
    """A freestream component: a dual scalar kept, a number as ``np.float64``."""
    return x if isinstance(x, D) else np.float64(x)


def full(shape, x):
    """``x`` (number or dual scalar) broadcast to ``shape``."""
    if isinstance(x, D):
        return D(np.full(shape, x.v), np.full(shape, x.d))
    return np.full(shape, np.float64(x))
