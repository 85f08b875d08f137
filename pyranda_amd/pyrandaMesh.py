"""Mesh description and grid (pyranda/pyrandaMesh.py:16-233): Cartesian,
cylindrical (coordsys 1, radius/theta/z, see cylindrical.py) or spherical
(coordsys 2, radius/theta/phi, see spherical.py).

`xdom=(x1, xn, n, periodic=True)` lines are executed through the same
substitution map as the reference (pyrandaMesh.py:33-56). Coordinates follow
parcop's face-shifted construction (mesh.f90:173-177) and are returned by
the C-ABI, so operators and coordinates share one dx.
"""
import torch

from .pyrandaUtils import fortran3d, splitLines
from .pyrandaVar import pyrandaVar


def defaultMeshOptions():
    options = {}
    options["x1"] = [0.0, 0.0, 0.0]
    options["xn"] = [1, 1, 1]
    options["nn"] = [1, 1, 1]
    options["periodic"] = [False, False, False]
    options["periodicGrid"] = True
    options["symmetric"] = [[False, False], [False, False], [False, False]]
    options["dim"] = 1
    options["coordsys"] = 0
    options["function"] = None
    options["id"] = 0
    options["pn"] = [0, 0, 0]
    return options


class pyrandaMesh:

    def __init__(self):
        self.name = "base"
        self.options = {}
        self.dims = 0
        self.PyMPI = None
        self.coordsys = None
        self.shape = None

    def makeMeshStr(self, str_mesh):
        self.options = defaultMeshOptions()
        self.get_sMap()
        for msl in splitLines(str_mesh):
            exec(fortran3d(msl, self.sMap))

    def set_options(self, ind, x1, xn, nn, periodic=False):
        self.options["x1"][ind] = x1
        self.options["xn"][ind] = xn
        self.options["nn"][ind] = nn
        self.options["periodic"][ind] = periodic

    def get_sMap(self):
        self.sMap = {"xdom=(": "self.set_options(0,", "ydom=(": "self.set_options(1,",
                     "zdom=(": "self.set_options(2,"}

    def makeMesh(self):
        mpi = self.PyMPI
        ctx = mpi.ctx
        nx, ny, nz = mpi.nx, mpi.ny, mpi.nz
        dev = mpi.device
        self.nn = self.options["nn"]
        c1 = [torch.tensor(ctx.coords1d[a], dtype=torch.float64, device=dev) for a in range(3)]
        # (nx,ny,nz) Fortran-layout coordinate fields (always FP64: ICs are
        # evaluated in FP64 even when the run stores fields in FP32)
        X = c1[0][None, None, :].expand(nz, ny, nx).contiguous().permute(2, 1, 0)
        Y = c1[1][None, :, None].expand(nz, ny, nx).contiguous().permute(2, 1, 0)
        Z = c1[2][:, None, None].expand(nz, ny, nx).contiguous().permute(2, 1, 0)
        self.coords = [pyrandaVar("x", "mesh", "scalar"), pyrandaVar("y", "mesh", "scalar"),
                       pyrandaVar("z", "mesh", "scalar")]
        for cor, arr in zip(self.coords, (X, Y, Z)):
            cor.PyMPI = mpi
            cor.allocated = True
            cor.data = arr
        ii = torch.arange(nx, device=dev, dtype=torch.float64)
        jj = torch.arange(ny, device=dev, dtype=torch.float64)
        kk = torch.arange(nz, device=dev, dtype=torch.float64)
        self.indices = [pyrandaVar("iloc", "mesh", "scalar"), pyrandaVar("jloc", "mesh", "scalar"),
                        pyrandaVar("kloc", "mesh", "scalar")]
        for ind, arr in zip(self.indices, (ii[None, None, :].expand(nz, ny, nx), jj[None, :, None].expand(nz, ny, nx),
                                           kk[:, None, None].expand(nz, ny, nx))):
            ind.PyMPI = mpi
            ind.allocated = True
            ind.data = arr.contiguous().permute(2, 1, 0)
        self.shape = [nx, ny, nz]
        # uniform Cartesian metrics (mesh.f90:190-212) as scalars
        self.d1, self.d2, self.d3 = ctx.dmesh
        self.GridLen = ctx.gridlen
        self.CellVol = ctx.dpatch[0] * ctx.dpatch[1] * ctx.dpatch[2]
        if mpi.coordsys in (1, 2):
            # cylindrical (pyrandaMesh.py:152-155, mesh.f90:213-224): the
            # native (radius, theta, z) grid is kept, the mesh coordinates
            # become x = r cos(theta), y = r sin(theta), z; spherical
            # (pyrandaMesh.py:156-164, mesh.f90:225-241): native (radius,
            # theta, phi), x = r sin(theta) cos(phi), y = r sin(theta)
            # sin(phi), z = r cos(theta)
            m = mpi.metrics
            self.coordsNative = [pyrandaVar("x", "mesh", "scalar"), pyrandaVar("y", "mesh", "scalar"),
                                 pyrandaVar("z", "mesh", "scalar")]
            for cor, nat in zip(self.coordsNative, self.coords):
                cor.PyMPI = mpi
                cor.allocated = True
                cor.data = nat.data
            r, th = X, Y
            cx, cy = torch.empty_like(X), torch.empty_like(Y)     # same (x fastest) layout
            if mpi.coordsys == 1:
                torch.mul(r, torch.cos(th), out=cx)
                torch.mul(r, torch.sin(th), out=cy)
            else:
                ph = self.coords[2].data
                cz = torch.empty_like(X)
                rs = r * torch.sin(th)
                torch.mul(rs, torch.cos(ph), out=cx)
                torch.mul(rs, torch.sin(ph), out=cy)
                torch.mul(r, torch.cos(th), out=cz)
                self.coords[2].data = cz
            self.coords[0].data = cx
            self.coords[1].data = cy
            self.d1, self.d2, self.d3 = m.d1, m.d2, m.d3
            self.GridLen = m.GridLen
            self.CellVol = m.CellVol
