"""A named field (pyranda/pyrandaVar.py:13-136) holding a GPU tensor.

`data` is an (nx,ny,nz) tensor in the reference's Fortran layout (x fastest)
or, after an ALG update, whatever the right-hand side produced (pyranda
stores scalars such as dt the same way).
"""
import torch


class pyrandaVar:

    def __init__(self, name, kind, rank):
        self.name = name
        self.kind = kind
        self.rank = rank
        self.rhs = None
        self.dt = None
        self.data = None
        self.allocated = False
        self.PyMPI = None

    def __allocate__(self, pympi):
        if self.allocated:
            return
        self.PyMPI = pympi
        if self.rank == "scalar":
            self.data = pympi.emptyScalar()
        elif self.rank == "vector":
            self.data = pympi.emptyVector()
        else:
            raise ValueError("Error: rank: %s not recognized" % self.rank)
        self.allocated = True

    def sum(self, axis=None, idata=None):
        if idata is None or not isinstance(idata, type(self.data)):
            idata = self.data
        if self.rank == "vector":
            return [self.PyMPI.sum3D(idata[:, :, :, i]) for i in range(3)]
        if axis is None:
            return self.PyMPI.sum3D(idata)
        if isinstance(axis, (tuple, list)):
            if len(axis) != 2:
                self.PyMPI.iprint("Error: length of axis argument cant be larger than 2")
                return None
            return torch.sum(idata, dim=tuple(axis), dtype=torch.float64).cpu().numpy()
        if isinstance(axis, int):
            return torch.sum(idata, dim=axis, dtype=torch.float64).cpu().numpy()
        self.PyMPI.iprint("Error: unknown type given as axis")
        return None

    def mean(self, axis=None, idata=None):
        npts = float(self.PyMPI.nx * self.PyMPI.ny * self.PyMPI.nz)
        rdata = self.sum(axis=axis, idata=idata)
        size = getattr(rdata, "size", 1)
        return rdata * size / npts

    def var(self, axis=None):
        bar = self.mean(axis=axis)
        bar2 = self.mean(axis=axis, idata=self.data ** 2)
        return bar2 - bar * bar

    def __getitem__(self, given):
        if len(given) != 3:
            self.PyMPI.iprint("Error: 3 index ranges must be given")
            return None
        return self.data[given].squeeze().cpu().numpy()
