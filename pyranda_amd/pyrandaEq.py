"""One equation of an EOM (pyranda/pyrandaEq.py:14-44).

`ddt(` marks a PDE, everything else is ALG; the right-hand side becomes a
`lambda self:` over device fields. The lambda's globals hold the device
`numpy`/`scipy` namespaces, so calls such as numpy.minimum act on GPU tensors.
"""
import os

from .pyrandaUtils import fortran3d, findVar
from .devnumpy import numpy, scipy  # noqa: F401  (names the lambdas resolve)
from .fuse import FusedRHS

# PYRANDA_AMD_FUSE=0 evaluates right-hand sides with one torch kernel per
# numpy-level operation instead of generated fused kernels
FUSE = os.environ.get("PYRANDA_AMD_FUSE", "1") != "0"


class pyrandaEq:

    def __init__(self, eqstr, sMap, pympi):
        self.eqstr = eqstr
        self.kind = "ALG"
        self.active = True
        self.rank = 1
        self.sRHS = ""
        if "=" in eqstr:
            # order-preserving: multiple return values map to LHS in order
            self.LHS = findVar(eqstr.split("=")[0], "scalar", unique=False)
            self.rank = len(self.LHS)
        else:
            self.LHS = None
        if self.LHS:
            Srhs = fortran3d(self.eqstr.split("=")[1], sMap)
        else:
            Srhs = fortran3d(self.eqstr, sMap)
        self.sRHS = Srhs
        self.RHS = FusedRHS(Srhs, globals()) if FUSE else eval("lambda self: " + Srhs)
        if "ddt(" in eqstr:
            self.kind = "PDE"
