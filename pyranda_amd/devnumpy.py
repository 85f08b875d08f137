"""The `numpy` and `scipy` names seen by EOM/IC strings.

pyranda translates `sqrt(` to `numpy.sqrt(` and friends (pyranda/pyranda.py:
812-869) and EOM strings call numpy directly (e.g. `numpy.minimum`). Here the
fields are torch tensors on the GPU, so these namespaces dispatch to torch
when any argument is a tensor and to numpy otherwise. Anything not listed
falls through to numpy, which fails loudly on device tensors rather than
silently copying them to the host.
"""
import math

import numpy as _np
import torch

_state = {"device": None, "dtype": torch.float64}


def set_default(device, dtype):
    _state["device"] = device
    _state["dtype"] = dtype


def _any_tensor(*args):
    return any(torch.is_tensor(a) for a in args)


def _unary(tname, nname):
    tfn = getattr(torch, tname)
    nfn = getattr(_np, nname)

    def f(x, *args, **kw):
        if torch.is_tensor(x):
            return tfn(x, *args, **kw)
        return nfn(x, *args, **kw)

    f.__name__ = nname
    return f


def _as_tensor_like(x, ref):
    if torch.is_tensor(x):
        return x
    return torch.as_tensor(x, dtype=ref.dtype, device=ref.device)


def _binary(tname, nname):
    tfn = getattr(torch, tname)
    nfn = getattr(_np, nname)

    def f(a, b):
        if _any_tensor(a, b):
            ref = a if torch.is_tensor(a) else b
            return tfn(_as_tensor_like(a, ref), _as_tensor_like(b, ref))
        return nfn(a, b)

    f.__name__ = nname
    return f


class _Random:
    def random(self, shape=None):
        if shape is None:
            return _np.random.random()
        # numpy's global generator, so seeded ICs draw the same numbers as
        # the reference; then moved to the device in Fortran layout
        arr = _np.random.random(shape)
        dev = _state["device"]
        if dev is None or len(shape) != 3:
            return arr
        t = torch.from_numpy(_np.ascontiguousarray(arr.transpose(2, 1, 0)))
        return t.to(device=dev, dtype=torch.float64).permute(2, 1, 0)

    def __getattr__(self, name):
        return getattr(_np.random, name)


class _NumpyNS:
    pi = math.pi
    e = math.e
    nan = float("nan")
    inf = float("inf")

    sqrt = staticmethod(_unary("sqrt", "sqrt"))
    abs = staticmethod(_unary("abs", "abs"))
    absolute = staticmethod(_unary("abs", "absolute"))
    sign = staticmethod(_unary("sign", "sign"))
    sin = staticmethod(_unary("sin", "sin"))
    cos = staticmethod(_unary("cos", "cos"))
    tan = staticmethod(_unary("tan", "tan"))
    tanh = staticmethod(_unary("tanh", "tanh"))
    exp = staticmethod(_unary("exp", "exp"))
    log = staticmethod(_unary("log", "log"))
    arctan = staticmethod(_unary("atan", "arctan"))
    arctanh = staticmethod(_unary("atanh", "arctanh"))
    floor = staticmethod(_unary("floor", "floor"))
    maximum = staticmethod(_binary("maximum", "maximum"))
    minimum = staticmethod(_binary("minimum", "minimum"))
    arctan2 = staticmethod(_binary("atan2", "arctan2"))
    random = _Random()

    @staticmethod
    def where(cond, a, b):
        if _any_tensor(cond, a, b):
            ref = next(t for t in (a, b, cond) if torch.is_tensor(t) and t.is_floating_point()) \
                if any(torch.is_tensor(t) and t.is_floating_point() for t in (a, b, cond)) else None
            if ref is None:
                ref = torch.zeros((), dtype=_state["dtype"], device=_state["device"])
            c = cond if torch.is_tensor(cond) else torch.as_tensor(cond, device=ref.device)
            return torch.where(c, _as_tensor_like(a, ref), _as_tensor_like(b, ref))
        return _np.where(cond, a, b)

    @staticmethod
    def dot(a, b):
        if _any_tensor(a, b):
            return torch.dot(a.reshape(-1), b.reshape(-1))
        return _np.dot(a, b)

    @staticmethod
    def isnan(x):
        if torch.is_tensor(x):
            return torch.isnan(x)
        return _np.isnan(x)

    @staticmethod
    def power(a, b):
        return a ** b

    @staticmethod
    def zeros_like(x):
        return torch.zeros_like(x) if torch.is_tensor(x) else _np.zeros_like(x)

    @staticmethod
    def ones_like(x):
        return torch.ones_like(x) if torch.is_tensor(x) else _np.ones_like(x)

    def __getattr__(self, name):
        return getattr(_np, name)


class _Special:
    @staticmethod
    def erf(x):
        if torch.is_tensor(x):
            return torch.special.erf(x)
        import scipy.special
        return scipy.special.erf(x)


class _ScipyNS:
    special = _Special()


numpy = _NumpyNS()
scipy = _ScipyNS()
