"""Single-GPU runtime standing where pyranda's pyrandaMPI stands.

pyrandaMPI (pyranda/pyrandaMPI.py) owns the decomposition, the parcop setup
and the der/fil/gfil operator objects, plus global reductions. On one MI355X
the whole block is one rank's share, so this class keeps the same attribute
and method names (PyMPI.der.ddx, PyMPI.sum3D, emptyScalar, x1proc, ...) with
fields resident in HBM as torch tensors and operators going through
libpyranda_amd.so. Reductions (sum3D/max3D/min3D, pyrandaMPI.py:307-326)
return Python floats, as the allreduce does, or 0-d device tensors when
`device_scalars` is set (no host round trip).
"""
import torch

from .parcop import ParcopContext, empty_field


class _Der:
    """parcop_der (pyrandaMPI.py:664-712)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def ddx(self, v): return self.ctx.ddx(v)
    def ddy(self, v): return self.ctx.ddy(v)
    def ddz(self, v): return self.ctx.ddz(v)
    def dd8x(self, v): return self.ctx.dd8x(v)
    def dd8y(self, v): return self.ctx.dd8y(v)
    def dd8z(self, v): return self.ctx.dd8z(v)

    def dd4x(self, v):
        raise NotImplementedError("dd4x (explicit 4th derivative e4d4) is not implemented in pyranda_amd")

    dd4y = dd4z = dd4x

    def div(self, fx, fy, fz): return self.ctx.divergence(fx, fy, fz)

    def divT(self, fxx, fxy, fxz, fyx, fyy, fyz, fzx, fzy, fzz):
        """Tensor divergence (operators.f90:104-122): each diagonal term takes
        the symmetric operator (isym**2), each off-diagonal one the
        antisymmetric operator where its direction has a SYMM end."""
        from . import _lib
        c, D1, A = self.ctx, _lib.OP_D1, _lib.OP_D1 | _lib.OP_ANTISYM
        dfx = c.apply(D1, 0, fxx) + c.apply(A, 1, fyx) + c.apply(A, 2, fzx)
        dfy = c.apply(A, 0, fxy) + c.apply(D1, 1, fyy) + c.apply(A, 2, fzy)
        dfz = c.apply(A, 0, fxz) + c.apply(A, 1, fyz) + c.apply(D1, 2, fzz)
        return dfx, dfy, dfz

    def grad(self, v): return self.ctx.grads(v)
    def laplacian(self, v): return self.ctx.plaplacian(v)
    def ring(self, v): return self.ctx.pring(v)
    def ringV(self, vx, vy, vz): return self.ctx.pringv(vx, vy, vz)


class _GFil:
    """parcop_gfil (pyrandaMPI.py:721-730)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def filter(self, v): return self.ctx.gfilter(v)
    def filterDir(self, v, direction): return self.ctx.gfilterdir(v, direction)


class _SFil:
    """parcop_sfil (pyrandaMPI.py:734-740)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def filter(self, v): return self.ctx.sfilter(v)
    def filter_rounded(self, v): return self.ctx.sfilter_rounded(v)


class pyrandaMPI:

    def __init__(self, mesh, comm=None, dtype=torch.float64, device="cuda"):
        opts = mesh.options
        self.nx, self.ny, self.nz = (int(n) for n in opts["nn"])
        x1, xn = opts["x1"], opts["xn"]
        self.dx = (xn[0] - x1[0]) / max(self.nx - 1, 1)
        self.dy = (xn[1] - x1[1]) / max(self.ny - 1, 1)
        self.dz = (xn[2] - x1[2]) / max(self.nz - 1, 1)
        self.periodic = opts["periodic"]
        symmetric = opts["symmetric"]
        self.coordsys = opts["coordsys"]
        if self.coordsys not in (0, 1, 2):
            raise NotImplementedError("Cartesian (coordsys=0), cylindrical (1) and spherical (2) meshes are built")
        self.comm = comm
        self.master = True
        self.dtype = dtype
        self.device = torch.device(device)
        # BC strings as pyrandaMPI.py:101-131 builds them
        bcs = []
        for d in range(3):
            for side in range(2):
                b = "NONE"
                if self.periodic[d]:
                    b = "PERI"
                if symmetric[d][side]:
                    b = "SYMM"
                bcs.append(b)
        self.bcs = bcs
        self.ctx = ParcopContext((self.nx, self.ny, self.nz), (x1[0], x1[1], x1[2]),
                                 (xn[0], xn[1], xn[2]), bcs, device=self.device)
        self.ax, self.ay, self.az = self.nx, self.ny, self.nz
        self.chunk_3d_size = [self.nx, self.ny, self.nz]
        self.chunk_3d_lo = [0, 0, 0]
        self.chunk_3d_hi = [self.nx - 1, self.ny - 1, self.nz - 1]
        # physical-boundary ownership flags (mesh.f90:162-167)
        self.x1proc = (not self.periodic[0]) and self.nx > 1
        self.xnproc = self.x1proc
        self.y1proc = (not self.periodic[1]) and self.ny > 1
        self.ynproc = self.y1proc
        self.z1proc = (not self.periodic[2]) and self.nz > 1
        self.znproc = self.z1proc
        self.device_scalars = False
        self.der = _Der(self.ctx)
        self.gfil = _GFil(self.ctx)
        self.fil = _SFil(self.ctx)
        self.metrics = None
        if self.coordsys == 1:
            # radius metrics in div/divT/grad/lap/ring and the cell-volume
            # weighted fbar; gbar (filter 'smooth') stays Cartesian
            from .cylindrical import CylindricalDer, CylindricalMetrics, CylindricalSFil
            self.metrics = CylindricalMetrics(self.ctx, (self.nx, self.ny, self.nz), self.device)
            self.der = CylindricalDer(self.ctx, self.metrics)
            self.fil = CylindricalSFil(self.ctx, self.metrics)
        elif self.coordsys == 2:
            # radius and polar-angle metrics; fbar divides by the filtered
            # cell volume; gbar (filter 'smooth') stays Cartesian
            from .spherical import SphericalDer, SphericalMetrics, SphericalSFil
            self.metrics = SphericalMetrics(self.ctx, (self.nx, self.ny, self.nz), self.device)
            self.der = SphericalDer(self.ctx, self.metrics)
            self.fil = SphericalSFil(self.ctx, self.metrics)

    # -- allocation (pyrandaMPI.py:249-254) ---------------------------------
    def emptyScalar(self):
        return torch.zeros((self.nz, self.ny, self.nx), dtype=self.dtype, device=self.device).permute(2, 1, 0)

    def emptyVector(self, rank=3):
        return torch.zeros((rank, self.nz, self.ny, self.nx), dtype=self.dtype,
                           device=self.device).permute(3, 2, 1, 0)

    def setPatch(self):
        """Operator tables are per context; nothing global to switch."""

    # -- reductions -----------------------------------------------------------
    def _scalar(self, t):
        return t if self.device_scalars else t.item()

    def sum3D(self, data):
        if torch.is_tensor(data):
            return self._scalar(torch.sum(self._flat(data), dtype=torch.float64))
        import numpy as np
        return float(np.sum(data))

    @staticmethod
    def _flat(data):
        # the field layout (x fastest) is the contiguous order of the reversed view
        if data.dim() == 3 and data.permute(2, 1, 0).is_contiguous():
            return data.permute(2, 1, 0)
        return data

    def max3D(self, data):
        if torch.is_tensor(data):
            return self._scalar(torch.max(self._flat(data)))
        import numpy as np
        return float(np.max(data))

    def min3D(self, data):
        if torch.is_tensor(data):
            return self._scalar(torch.min(self._flat(data)))
        import numpy as np
        return float(np.min(data))

    # -- line/plane reductions used by example scripts ------------------------
    def xsum(self, data):
        return torch.sum(data, dim=0, dtype=torch.float64).cpu().numpy()

    def ysum(self, data):
        return torch.sum(data, dim=1, dtype=torch.float64).cpu().numpy()

    def zsum(self, data):
        return torch.sum(data, dim=2, dtype=torch.float64).cpu().numpy()

    xbar, ybar, zbar = xsum, ysum, zsum

    # reductions over two directions (pyrandaMPI.py:350-357): one value per
    # index of the remaining direction
    def xysum(self, data):
        return torch.sum(data, dim=(0, 1), dtype=torch.float64).cpu().numpy()

    def xzsum(self, data):
        return torch.sum(data, dim=(0, 2), dtype=torch.float64).cpu().numpy()

    def yzsum(self, data):
        return torch.sum(data, dim=(1, 2), dtype=torch.float64).cpu().numpy()

    def subsum3xz(self, data):
        return torch.sum(data, dim=(0, 2), dtype=torch.float64).cpu().numpy()

    def iprint(self, s):
        print(s, flush=True)
