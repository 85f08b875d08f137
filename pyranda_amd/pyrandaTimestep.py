"""CFL and diffusive time-step limits (pyranda/pyrandaTimestep.py:15-88).

courant: 1/max(|u|/dx + |v|/dy + |w|/dz + |c|/GridLen) over the grid;
diff: min(rho*GridLen^2 / max(1e-12, bulk)). Both are device reductions
returning Python floats, as the reference's MPI allreduce does, or 0-d FP64
device tensors when the simulation keeps its scalars on the GPU.
"""
import torch

from . import fuse
from .devnumpy import numpy
from .pyrandaPackage import pyrandaPackage

# one-pass device reductions of the same expressions (same operation order)
_COURANT = fuse.FusedReduce("max", "(((fabs(x0) / x4 + fabs(x1) / x5) + fabs(x2) / x6) + fabs(x3) / x7)",
                            4, 4, "courant")
_DIFF = fuse.FusedReduce("min", "((x1 * x2) * x2) / pyr_max(1.0e-12, x0)", 2, 1, "diff")


def _uniform(*metrics):
    """True when every metric is one number (Cartesian meshes); curvilinear
    meshes carry d2, d3 or GridLen as fields and take the elementwise path."""
    return all(not isinstance(m, torch.Tensor) or m.numel() == 1 for m in metrics)


class pyrandaTimestep(pyrandaPackage):

    def __init__(self, pysim):
        pyrandaPackage.__init__(self, "Timestep", pysim)
        self.dx = pysim.mesh.d1
        self.dy = pysim.mesh.d2
        self.dz = pysim.mesh.d3
        self.pysim = pysim
        self.GridLen = pysim.mesh.GridLen

    def get_sMap(self):
        self.sMap = {
            "dt.courant(": "self.packages['Timestep'].courant(",
            "dt.diff(": "self.packages['Timestep'].diff(",
            "dt.diffDir(": "self.packages['Timestep'].diffDir(",
        }

    def courant(self, u, v, w, c):
        shape = self.pyranda.mesh.shape
        if fuse.fields_ok(shape, u, v, w, c) and _uniform(self.dx, self.dy, self.dz, self.GridLen):
            r = _COURANT(shape, [u, v, w, c], [self.dx, self.dy, self.dz, self.GridLen],
                         device=self.pyranda.PyMPI.device_scalars)
            # 1/x on the device is the correctly rounded quotient, as on the host
            return r.reciprocal() if isinstance(r, torch.Tensor) else 1.0 / r
        vrate = numpy.abs(u) / self.dx + numpy.abs(v) / self.dy + numpy.abs(w) / self.dz
        crate = numpy.abs(c) / self.GridLen
        return 1.0 / self.pyranda.PyMPI.max3D(vrate + crate)

    def diff(self, bulk, density):
        delta = self.GridLen
        shape = self.pyranda.mesh.shape
        if fuse.fields_ok(shape, bulk, density) and _uniform(delta):
            return _DIFF(shape, [bulk, density], [delta], device=self.pyranda.PyMPI.device_scalars)
        drate = density * delta * delta / numpy.maximum(1.0e-12, bulk)
        return self.pyranda.PyMPI.min3D(drate)

    def diffDir(self, bulk, density, delta):
        drate = density * delta * delta / numpy.maximum(1.0e-12, bulk)
        return self.pyranda.PyMPI.min3D(drate)
