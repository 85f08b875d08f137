"""Boundary-plane updates for bounded directions (pyranda/pyrandaBC.py:15-160).

extrap: v[0] = 2 v[1] - v[2] (order 2) or v[1]; const: v[0] = val; on the
x1/xn/y1/yn/z1/zn planes that this block owns. Slip/exit/farfield/curvilinear
conditions are not implemented in pyranda_amd and raise.
"""
from .pyrandaPackage import pyrandaPackage

_PLANES = {
    "x1": (0, 0, 1, 2), "xn": (0, -1, -2, -3),
    "y1": (1, 0, 1, 2), "yn": (1, -1, -2, -3),
    "z1": (2, 0, 1, 2), "zn": (2, -1, -2, -3),
}


def _plane(data, axis, idx):
    sl = [slice(None)] * 3
    sl[axis] = idx
    return tuple(sl)


class pyrandaBC(pyrandaPackage):

    def __init__(self, pysim):
        pyrandaPackage.__init__(self, "BC", pysim)
        self.bcList = {}
        self.BCdata = {}

    def get_sMap(self):
        self.sMap = {
            "bc.extrap(": "self.packages['BC'].extrap(",
            "bc.const(": "self.packages['BC'].const(",
            "bc.exit(": "self.packages['BC'].exitbc(",
            "bc.slip(": "self.packages['BC'].slipbc(",
            "bc.field(": "self.packages['BC'].applyBC('field',",
            "bc.farfield(": "self.packages['BC'].farfieldbc(",
            "bc.symm(": "self.packages['BC'].symm(",
        }

    def _owns(self, direction):
        return getattr(self.pyranda.PyMPI, direction + "proc")

    def applyBC(self, func, var, direction, val=None):
        pFunc = getattr(self, func)
        var = var if isinstance(var, list) else [var]
        direction = direction if isinstance(direction, list) else [direction]
        for d in direction:
            for v in var:
                if val is not None:
                    pFunc(v, d, val)
                else:
                    pFunc(v, d)

    def extrap(self, var, direction, order=2):
        var = var if isinstance(var, list) else [var]
        direction = direction if isinstance(direction, list) else [direction]
        for d in direction:
            for v in var:
                self.extrapolate(v, d, order)

    def extrapolate(self, var, direction, order):
        if direction not in _PLANES or not self._owns(direction):
            return
        axis, i0, i1, i2 = _PLANES[direction]
        data = self.pyranda.variables[var].data
        if order == 2:
            data[_plane(data, axis, i0)] = 2 * data[_plane(data, axis, i1)] - data[_plane(data, axis, i2)]
        else:
            data[_plane(data, axis, i0)] = data[_plane(data, axis, i1)]

    def const(self, var, direction, val):
        var = var if isinstance(var, list) else [var]
        direction = direction if isinstance(direction, list) else [direction]
        for d in direction:
            for v in var:
                self.constant(v, d, val)

    def constant(self, var, direction, val):
        if direction not in _PLANES or not self._owns(direction):
            return
        axis, i0, _, _ = _PLANES[direction]
        data = self.pyranda.variables[var].data
        data[_plane(data, axis, i0)] = val

    def field(self, var, direction, field):
        if direction not in _PLANES or not self._owns(direction):
            return
        axis, i0, _, _ = _PLANES[direction]
        data = self.pyranda.variables[var].data
        data[_plane(data, axis, i0)] = field[_plane(field, axis, i0)]

    def _unsupported(self, *a, **k):
        raise NotImplementedError("this boundary condition is not implemented in pyranda_amd")

    exitbc = slipbc = farfieldbc = symm = _unsupported
