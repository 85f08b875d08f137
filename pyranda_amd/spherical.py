"""Spherical meshes (coordsys 2: x = radius, y = polar angle theta, z =
azimuth phi).

As for cylindrical meshes (cylindrical.py), the reference composes the
operators of a spherical mesh from the one-dimensional compact operators and
pointwise metrics (pyranda/parcop/operators.f90, CASE(2,4) of divV :64-76,
divT :148-180, gradS :198-202, LapS :545-559, ringS :615-643 with d2 = r dtheta
and d3 = r sin(theta) dphi, filter 'spectral' :889-894 divided by the filtered
cell volume; metrics mesh.f90:225-241 and :374-382). Here the same
compositions run on the GPU: each directional operator is one launch through
the C-ABI, the metric algebra is elementwise on the device, in the reference's
order of operations. The radially stretched variant (coordsys 4) is not
built.

Symmetry: as in cylindrical.py, `_op(..., odd=True)` asks for the
antisymmetric member where the reference passes isym**k = -1 for odd k.
"""
import math

import torch

from . import _lib
from .cylindrical import _op


class SphericalMetrics:
    """radius and theta fields, d1 = dx, d2 = radius*dy, d3 =
    radius*sin(theta)*dz, 1/radius, 1/sin(theta), 1/tan(theta), the cell
    volume and its sharp-filtered version, and the minimum grid length
    (mesh.f90:225-241, :374-382), as FP64 device fields or scalars."""

    def __init__(self, ctx, nn, device):
        nx, ny, nz = nn
        r1 = torch.tensor(ctx.coords1d[0], dtype=torch.float64, device=device)
        t1 = torch.tensor(ctx.coords1d[1], dtype=torch.float64, device=device)
        self.radius = r1[None, None, :].expand(nz, ny, nx).contiguous().permute(2, 1, 0)
        self.theta = t1[None, :, None].expand(nz, ny, nx).contiguous().permute(2, 1, 0)
        dx, dy, dz = ctx.dpatch
        # the grid is x1 - dx/2 + (i - 1/2) dx, so a point meant at 0 lands
        # within rounding of it: compare against the spacing
        if float(r1.abs().min()) <= 1e-9 * abs(dx):
            raise ValueError("spherical mesh: a point at the origin (radius 0)")
        sin_t = torch.sin(self.theta)
        if float(torch.sin(t1).abs().min()) <= 1e-9 * abs(dy):
            raise ValueError("spherical mesh: a point on the polar axis (sin(theta) = 0)")
        self.dx = dx
        self.d1 = dx
        self.d2 = self.radius * dy
        self.d3 = self.radius * sin_t * dz
        self.iR = 1.0 / self.radius
        self.isinY = 1.0 / sin_t
        self.itanY = 1.0 / torch.tan(self.theta)
        if ny == 1 and nz == 1:
            self.CellVol = self.d1 * 4.0 * math.pi * self.radius ** 2
            self.GridLen = self.d1
        else:
            self.CellVol = self.d1 * self.d2 * self.d3
            self.GridLen = torch.clamp(torch.minimum(self.d2, self.d3), max=self.d1)
        # the cell volume filtered x -> y -> z with the symmetric filters
        # (mesh.f90:377-379): fbar divides by it
        t = ctx.apply(_lib.OP_SFILTER, 0, self.CellVol)
        t = ctx.apply(_lib.OP_SFILTER, 1, t)
        self.CellVolS = ctx.apply(_lib.OP_SFILTER, 2, t)


class SphericalDer:
    """parcop_der on a spherical mesh (pyrandaMPI.py:664-712)."""

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
        """operators.f90:64-76: (r^2 fx)_r / r^2 + ((sin(theta) fy)_theta +
        (fz)_phi) / (r sin(theta)); r^2 fx odd in r (isymX**3)."""
        c, m = self.ctx, self.m
        r, th = m.radius, m.theta
        fA = _op(c, _lib.OP_D1, 0, r ** 2 * fx, odd=True)
        fA = fA * m.dx / m.d1
        fB = _op(c, _lib.OP_D1, 1, fy * torch.sin(th))          # isymY**2
        fC = _op(c, _lib.OP_D1, 2, fz, odd=True)                 # isymZ
        return fA / r ** 2 + (fB + fC) / (r * torch.sin(th))

    def divT(self, fxx, fxy, fxz, fyx, fyy, fyz, fzx, fzy, fzz):
        """operators.f90:148-180 (tensor divergence, spherical)."""
        c, m = self.ctx, self.m
        r, th = m.radius, m.theta
        D1 = _lib.OP_D1
        fA = _op(c, D1, 0, r ** 2 * fxx)                         # isymX**4
        fA = fA * m.dx / m.d1
        fB = _op(c, D1, 1, fyx / m.isinY)                        # isymY**2
        fC = _op(c, D1, 2, fzx, odd=True)                        # isymZ
        dfx = fA * m.iR ** 2 + (fB + fC) * m.iR * m.isinY - (fyy + fzz) * m.iR
        fA = _op(c, D1, 0, r ** 3 * fxy)                         # isymX**4
        fA = fA * m.dx / m.d1
        fB = _op(c, D1, 1, fyy * torch.sin(th), odd=True)        # isymY**3
        fC = _op(c, D1, 2, fzy, odd=True)                        # isymZ
        dfy = fA * m.iR ** 3 + (fB + fC) * m.iR * m.isinY + (fyx - fxy - fzz * m.itanY) * m.iR
        fA = _op(c, D1, 0, r ** 3 * fxz)                         # isymX**4
        fA = fA * m.dx / m.d1
        fB = _op(c, D1, 1, fyz * torch.sin(th))                  # isymY**2
        fC = _op(c, D1, 2, fzz)                                  # isymZ**2
        dfz = fA * m.iR ** 3 + (fB + fC) * m.iR * m.isinY + (fzx - fxz + fzy * m.itanY) * m.iR
        return dfx, dfy, dfz

    def grad(self, v):
        """gradS (operators.f90:198-202)."""
        gx, gy, gz = self.ctx.grads(v)
        m = self.m
        return gx * m.dx / m.d1, gy / m.radius, gz / (m.radius * torch.sin(m.theta))

    def laplacian(self, v):
        """LapS (operators.f90:545-559): (r^2 f_r)_r / r^2 + (sin(theta)
        f_theta)_theta / (r^2 sin(theta)) + f_phiphi / (r sin(theta))^2."""
        c, m = self.ctx, self.m
        r, sth = m.radius, torch.sin(m.theta)
        t = c.ddx(v)
        lap = c.ddx(r ** 2 * t) / r ** 2
        t = c.ddy(v)
        lap = lap + c.ddy(sth * t) / (r ** 2 * sth)
        return lap + c.apply(_lib.OP_D2, 2, v) / (r * sth) ** 2

    def ring(self, v):
        """ringS with L=2 (operators.f90:615-643) with d2 = r dtheta and
        d3 = r sin(theta) dphi at every point."""
        c, m = self.ctx, self.m
        a = c.apply(_lib.OP_D8, 0, v).abs() * (m.d1 * m.d1)
        b = c.apply(_lib.OP_D8, 1, v).abs() * (m.d2 * m.d2)
        e = c.apply(_lib.OP_D8, 2, v).abs() * (m.d3 * m.d3)
        return torch.maximum(torch.maximum(a, b.to(a.dtype)), e.to(a.dtype))

    def ringV(self, vx, vy, vz):
        raise NotImplementedError("ringV on spherical meshes is not implemented in pyranda_amd")


class SphericalSFil:
    """fbar on a spherical mesh (filter 'spectral', operators.f90:889-894):
    the filter of f * CellVol (symmetric members) divided by the filtered
    cell volume."""

    def __init__(self, ctx, metrics):
        self.ctx = ctx
        self.m = metrics

    def filter(self, v):
        c, m = self.ctx, self.m
        t = (v * m.CellVol).to(v.dtype)
        t = c.apply(_lib.OP_SFILTER, 0, t)
        t = c.apply(_lib.OP_SFILTER, 1, t)
        t = c.apply(_lib.OP_SFILTER, 2, t)
        return (t / m.CellVolS).to(v.dtype)
