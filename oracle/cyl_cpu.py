"""Cylindrical (coordsys 1) and spherical (coordsys 2) operators restated on
the CPU (TEST INFRASTRUCTURE).

The compositions of pyranda/parcop/operators.f90 for coordsys 1 -- divV
CASE(1) :54-63, divT CASE(1) :124-147, gradS :193-197, LapS :527-544, ringS
:615-643 with d2 = radius*dy, filter 'spectral' :876-885 -- written out in
numpy over the C oracle's one-dimensional operators (pyoracle.Oracle, whose
apply_sym selects the SYMM-pair member as d1x(v, dv, bc) does,
compact_operators.f90:31-35). isymX**k is -1 for odd k when x has a SYMM
end, else +1 (patch.f90:89-91). Only tests/ import this module.

Arrays: float64, Fortran order (nx, ny, nz); r is the radius field.
"""
import math

import numpy as np

from . import pyoracle as po


class CylOracle:

    def __init__(self, P, r, dpatch, isym=(1, 1, 1)):
        self.P = P
        self.r = np.asfortranarray(np.broadcast_to(r, P.nn))
        self.dx, self.dy, self.dz = dpatch
        self.isym = isym           # (isymX, isymY, isymZ)

    def _d1(self, d, v, k=1):
        return self.P.apply_sym(po.D1, d, self.isym[d] ** k, v)

    def div(self, fx, fy, fz):
        r = self.r
        fA = self._d1(0, r * fx, 2)
        fB = self._d1(1, fy, 1)
        fC = self._d1(2, fz, 1)
        return (fA + fB) / r + fC

    def divT(self, fxx, fxy, fxz, fyx, fyy, fyz, fzx, fzy, fzz):
        r = self.r
        dfx = (self._d1(0, r * fxx, 3) + self._d1(1, fyx, 1) - fyy) / r + self._d1(2, fzx, 1)
        dfy = (self._d1(0, r ** 2 * fxy, 3) / r ** 2 + (self._d1(1, fyy, 2) + fyx - fxy) / r
               + self._d1(2, fzy, 1))
        dfz = (self._d1(0, r * fxz, 2) + self._d1(1, fyz, 1)) / r + self._d1(2, fzz, 2)
        return dfx, dfy, dfz

    def grad(self, v):
        P = self.P
        return P.ddx(v), P.ddy(v) / self.r, P.ddz(v)

    def laplacian(self, v):
        P, r = self.P, self.r
        lap = P.ddx(r * P.ddx(v)) / r
        lap = lap + P.d2(v, 1) / r ** 2
        return lap + P.d2(v, 2)

    def ring(self, v):
        P = self.P
        a = np.abs(P.d8(v, 0)) * self.dx ** 2
        b = np.abs(P.d8(v, 1)) * (self.r * self.dy) ** 2
        c = np.abs(P.d8(v, 2)) * self.dz ** 2
        return np.maximum(np.maximum(a, b), c)

    def cellvol(self):
        ny = self.P.nn[1]
        if ny == 1:
            return self.dx * 2.0 * math.pi * self.r * self.dz
        return self.dx * (self.r * self.dy) * self.dz

    def fbar(self, v):
        P, vol = self.P, self.cellvol()
        # a scalar's x parity (+1) flipped: the cell volume is odd in r
        t = P.apply_sym(po.SF, 0, -1, v * vol)
        t = P.apply(po.SF, 1, t)
        t = P.apply(po.SF, 2, t)
        return t / vol


class SphOracle:
    """Spherical (coordsys 2) compositions of pyranda/parcop/operators.f90
    restated in numpy over the C oracle's one-dimensional operators: divV
    CASE(2,4) :64-76, divT CASE(2,4) :148-180, gradS :198-202, LapS
    :545-559, ringS :615-643 with d2 = r dtheta and d3 = r sin(theta) dphi,
    filter 'spectral' :889-894 (the filter of f * CellVol divided by the
    filtered cell volume, mesh.f90:374-382); metrics mesh.f90:225-241.
    isym[d]**k selects the SYMM-pair member as in CylOracle. Only tests/
    import this module.

    r and theta are the radius and polar-angle fields (Fortran order)."""

    def __init__(self, P, r, theta, dpatch, isym=(1, 1, 1)):
        self.P = P
        self.r = np.asfortranarray(np.broadcast_to(r, P.nn))
        self.th = np.asfortranarray(np.broadcast_to(theta, P.nn))
        self.dx, self.dy, self.dz = dpatch
        self.isym = isym
        nx, ny, nz = P.nn
        self.d1 = self.dx
        self.d2 = self.r * self.dy
        self.d3 = self.r * np.sin(self.th) * self.dz
        if ny == 1 and nz == 1:
            self.vol = self.d1 * 4.0 * math.pi * self.r ** 2
        else:
            self.vol = self.d1 * self.d2 * self.d3
        t = P.apply(po.SF, 0, np.asfortranarray(self.vol))
        t = P.apply(po.SF, 1, t)
        self.vol_s = P.apply(po.SF, 2, t)

    def _d1(self, d, v, k=1):
        return self.P.apply_sym(po.D1, d, self.isym[d] ** k, np.asfortranarray(v))

    def div(self, fx, fy, fz):
        r, s = self.r, np.sin(self.th)
        fA = self._d1(0, r ** 2 * fx, 3) * self.dx / self.d1
        fB = self._d1(1, fy * s, 2)
        fC = self._d1(2, fz, 1)
        return fA / r ** 2 + (fB + fC) / (r * s)

    def divT(self, fxx, fxy, fxz, fyx, fyy, fyz, fzx, fzy, fzz):
        r, s = self.r, np.sin(self.th)
        iR, isinY, itanY = 1.0 / r, 1.0 / s, 1.0 / np.tan(self.th)
        fA = self._d1(0, r ** 2 * fxx, 4) * self.dx / self.d1
        fB = self._d1(1, fyx / isinY, 2)
        fC = self._d1(2, fzx, 1)
        dfx = fA * iR ** 2 + (fB + fC) * iR * isinY - (fyy + fzz) * iR
        fA = self._d1(0, r ** 3 * fxy, 4) * self.dx / self.d1
        fB = self._d1(1, fyy * s, 3)
        fC = self._d1(2, fzy, 1)
        dfy = fA * iR ** 3 + (fB + fC) * iR * isinY + (fyx - fxy - fzz * itanY) * iR
        fA = self._d1(0, r ** 3 * fxz, 4) * self.dx / self.d1
        fB = self._d1(1, fyz * s, 2)
        fC = self._d1(2, fzz, 2)
        dfz = fA * iR ** 3 + (fB + fC) * iR * isinY + (fzx - fxz + fzy * itanY) * iR
        return dfx, dfy, dfz

    def grad(self, v):
        P = self.P
        return P.ddx(v) * self.dx / self.d1, P.ddy(v) / self.r, P.ddz(v) / (self.r * np.sin(self.th))

    def laplacian(self, v):
        P, r, s = self.P, self.r, np.sin(self.th)
        lap = P.ddx(np.asfortranarray(r ** 2 * P.ddx(v))) / r ** 2
        lap = lap + P.ddy(np.asfortranarray(s * P.ddy(v))) / (r ** 2 * s)
        return lap + P.d2(v, 2) / (r * s) ** 2

    def ring(self, v):
        P = self.P
        a = np.abs(P.d8(v, 0)) * self.d1 ** 2
        b = np.abs(P.d8(v, 1)) * self.d2 ** 2
        c = np.abs(P.d8(v, 2)) * self.d3 ** 2
        return np.maximum(np.maximum(a, b), c)

    def fbar(self, v):
        P = self.P
        t = P.apply(po.SF, 0, np.asfortranarray(v * self.vol))
        t = P.apply(po.SF, 1, t)
        t = P.apply(po.SF, 2, t)
        return t / self.vol_s
