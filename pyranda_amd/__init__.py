"""MI355X-native compact finite-difference core with pyranda's simulation API.

    from pyranda_amd import pyrandaSim, pyrandaTimestep, pyrandaBC

Fields live in HBM as torch tensors; the compact operators (Pade derivatives,
compact/Gaussian filters, LAD ring) run as HIP kernels from
libpyranda_amd.so through the C-ABI in include/pyranda_amd.h.
"""
from .pyranda import pyrandaSim, pyrandaRestart
from .pyrandaTimestep import pyrandaTimestep
from .pyrandaBC import pyrandaBC
from .pyrandaIBM import pyrandaIBM
from .pyrandaPackage import pyrandaPackage
from .pyrandaMesh import defaultMeshOptions
from .group import pyrandaGroupSim

__all__ = ["pyrandaSim", "pyrandaRestart", "pyrandaTimestep", "pyrandaBC", "pyrandaIBM", "pyrandaPackage", "defaultMeshOptions",
           "pyrandaGroupSim"]
