"""Cylindrical meshes (coordsys 1: x = radius, y = theta, z = axial).

The reference builds the operators of a cylindrical mesh from the same
one-dimensional compact operators as a Cartesian one, with the radius as a
pointwise metric (pyranda/parcop/operators.f90, CASE(1) of divV :54-63,
divT :124-147, gradS :193-197, LapS :527-544, filter :876-885; metrics
mesh.f90:213-224). Here the same compositions run on the GPU: each
directional operator is one launch through the C-ABI, the metric algebra
is elementwise on the device.

Symmetry: a direction with a SYMM end has a symmetric (default) and an
antisymmetric operator; the reference passes isymX**k (isymX = -1 when x
has a SYMM end, patch.f90:89-91) and takes the antisymmetric operator for
-1 (compact_operators.f90:31-35). `_op(..., odd=True)` asks for it; the
library falls back to the default operator where the direction has no
SYMM end, as the reference does.
"""
import math

import torch

from . import _lib


class CylindricalMetrics:
    """radius (the x coordinate), 1/radius, d1 = dx, d2 = radius*dy, d3 = dz,
    the cell volume and the minimum grid length (mesh.f90:213-224), as FP64
    device fields or scalars."""

    def __init__(self, ctx, nn, device):
        nx, ny, nz = nn
        r1 = torch.tensor(ctx.coords1d[0], dtype=torch.float64, device=device)
        self.radius = r1[None, None, :].expand(nz, ny, nx).contiguous().permute(2, 1, 0)
        dx, dy, dz = ctx.dpatch
        # a point meant at r = 0 lands within rounding of it (the grid is
        # x1 - dx/2 + (i - 1/2) dx): compare against the spacing
        if float(r1.abs().min()) <= 1e-9 * abs(dx):
            raise ValueError("cylindrical mesh: a point on the axis (radius 0)")
        self.iR = 1.0 / self.radius
        self.d1 = dx
        self.d2 = self.radius * dy
        self.d3 = dz
        if ny == 1:
            self.CellVol = dx * 2.0 * math.pi * self.radius * dz
            self.GridLen = min(dx, dz)
        else:
            self.CellVol = dx * self.d2 * dz
            self.GridLen = torch.clamp(self.d2, max=min(dx, dz))


def _op(ctx, op, d, v, odd=False):
    return ctx.apply(op | (_lib.OP_ANTISYM if odd else 0), d, v)


class CylindricalDer:
    """parcop_der on a cylindrical mesh (pyrandaMPI.py:664-712)."""

    def __init__(self, ctx, metrics):
        self.ctx = ctx
        self.m = metrics

    def ddx(self, v): return self.ctx.ddx(v)
    def ddy(self, v): return self.ctx.ddy(v)
    def ddz(self, v): return self.ctx.ddz(v)
    def dd8x(self, v): return self.ctx.dd8x(v)
    def dd8y(self, v): return self.ctx.dd8y(v)
    def dd8z(self, v): return self.ctx.dd8z(v)

    def dd4x(self, v):
        raise NotImplementedError("dd4x (explicit 4th derivative e4d4) is not implemented in pyranda_amd")

    dd4y = dd4z = dd4x

    def div(self, fx, fy, fz):
        """operators.f90:54-63: ((r fx)_r + (fy)_theta) / r + (fz)_z, with
        r fx even (isymX**2) and fy, fz of their own parity."""
        c, r = self.ctx, self.m.radius
        fA = _op(c, _lib.OP_D1, 0, r * fx)
        fB = _op(c, _lib.OP_D1, 1, fy, odd=True)
        fC = _op(c, _lib.OP_D1, 2, fz, odd=True)
        return (fA + fB) / r + fC

    def divT(self, fxx, fxy, fxz, fyx, fyy, fyz, fzx, fzy, fzz):
        """operators.f90:124-147 (tensor divergence, cylindrical)."""
        c, r = self.ctx, self.m.radius
        D1 = _lib.OP_D1
        fA = _op(c, D1, 0, r * fxx, odd=True)            # isymX**3
        fB = _op(c, D1, 1, fyx, odd=True)                # isymY
        fC = _op(c, D1, 2, fzx, odd=True)                # isymZ
        dfx = (fA + fB - fyy) / r + fC
        fA = _op(c, D1, 0, r * r * fxy, odd=True)        # isymX**3
        fB = _op(c, D1, 1, fyy)                          # isymY**2
        fC = _op(c, D1, 2, fzy, odd=True)                # isymZ
        dfy = fA / r ** 2 + (fB + fyx - fxy) / r + fC
        fA = _op(c, D1, 0, r * fxz)                      # isymX**2
        fB = _op(c, D1, 1, fyz, odd=True)                # isymY
        fC = _op(c, D1, 2, fzz)                          # isymZ**2
        dfz = (fA + fB) / r + fC
        return dfx, dfy, dfz

    def grad(self, v):
        """gradS (operators.f90:184-197): the theta derivative over r."""
        gx, gy, gz = self.ctx.grads(v)
        return gx, gy / self.m.radius, gz

    def laplacian(self, v):
        """LapS (operators.f90:527-544): (r f_r)_r / r + f_thth / r^2 + f_zz."""
        c, r = self.ctx, self.m.radius
        t = c.ddx(v)
        lap = c.ddx(r * t) / r
        lap = lap + c.apply(_lib.OP_D2, 1, v) / r ** 2
        return lap + c.apply(_lib.OP_D2, 2, v)

    def ring(self, v):
        """ringS with L=2 (operators.f90:615-643): the theta direction scaled
        by d2 = r dy at every point."""
        c, m = self.ctx, self.m
        a = c.apply(_lib.OP_D8, 0, v).abs() * (m.d1 * m.d1)
        b = c.apply(_lib.OP_D8, 1, v).abs() * (m.d2 * m.d2)
        e = c.apply(_lib.OP_D8, 2, v).abs() * (m.d3 * m.d3)
        return torch.maximum(torch.maximum(a, b.to(a.dtype)), e)

    def ringV(self, vx, vy, vz):
        raise NotImplementedError("ringV on cylindrical meshes is not implemented in pyranda_amd")


class CylindricalSFil:
    """fbar on a cylindrical mesh (filter 'spectral', operators.f90:876-885):
    the filter of f * CellVol, the x sweep with the parity flipped (the cell
    volume is odd in the radius), divided by CellVol."""

    def __init__(self, ctx, metrics):
        self.ctx = ctx
        self.m = metrics

    def filter(self, v):
        c, vol = self.ctx, self.m.CellVol
        t = (v * vol).to(v.dtype)
        t = _op(c, _lib.OP_SFILTER, 0, t, odd=True)
        t = c.apply(_lib.OP_SFILTER, 1, t)
        t = c.apply(_lib.OP_SFILTER, 2, t)
        return (t / vol).to(v.dtype)
