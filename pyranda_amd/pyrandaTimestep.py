"""CFL and diffusive time-step limits (pyranda/pyrandaTimestep.py:15-88).

courant: 1/max(|u|/dx + |v|/dy + |w|/dz + |c|/GridLen) over the grid;
diff: min(rho*GridLen^2 / max(1e-12, bulk)). Both are device reductions
returning Python floats, as the reference's MPI allreduce does.
"""
from .devnumpy import numpy
from .pyrandaPackage import pyrandaPackage


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
        vrate = numpy.abs(u) / self.dx + numpy.abs(v) / self.dy + numpy.abs(w) / self.dz
        crate = numpy.abs(c) / self.GridLen
        return 1.0 / self.pyranda.PyMPI.max3D(vrate + crate)

    def diff(self, bulk, density):
        delta = self.GridLen
        drate = density * delta * delta / numpy.maximum(1.0e-12, bulk)
        return self.pyranda.PyMPI.min3D(drate)

    def diffDir(self, bulk, density, delta):
        drate = density * delta * delta / numpy.maximum(1.0e-12, bulk)
        return self.pyranda.PyMPI.min3D(drate)
