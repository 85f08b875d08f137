"""Immersed boundaries by level-set extension (pyranda/pyrandaIBM.py:15-196).

The body is the region phi <= 0 of a signed-distance field phi with gradient
(gx, gy, gz). Inside it, values are replaced by an extension of the fluid
values along the level-set normal: two pseudo-time iterations of
    val <- val + 0.5 * GridLen * (grad(val) . grad(phi)),  then  gbar(val)
applied only where phi <= eps (pyrandaIBM.py:78-91). ibmV adds the slip
treatment of the normal velocity near the interface (pyrandaIBM.py:94-143),
ibmWall the no-slip one (:145-168).

Everything runs on the device: grad and gbar are the library's line sweeps,
the masks and updates elementwise tensor operations in the reference's order
of operations (so FP64 results equal the reference's to rounding of the
operators). The EOM tokens are those of the reference: ibmS(, ibmV(,
ibmWall(.
"""
import torch

from .pyrandaPackage import pyrandaPackage

# pseudo-time iterations, step factor and interface band (pyrandaIBM.py:4,15-16)
IMMERSED_ITER = 2
IMMERSED_CFL = 0.5
IMMERSED_EPS = 0.1


def _where(cond, a, b):
    """numpy.where for device tensors and Python scalars."""
    if not isinstance(a, torch.Tensor) and not isinstance(b, torch.Tensor):
        raise TypeError("ibm: a mask needs at least one field operand")
    return torch.where(cond, a, b)


class pyrandaIBM(pyrandaPackage):

    def __init__(self, pysim):
        pyrandaPackage.__init__(self, "IBM", pysim)

    def get_sMap(self):
        self.sMap = {
            "ibmWall(": "self.packages['IBM'].ibmWall(",
            "ibmV(": "self.packages['IBM'].ibmVel(",
            "ibmS(": "self.packages['IBM'].ibmS(",
        }

    # -- EOM entry points ----------------------------------------------------
    def ibmVel(self, vel, phi, gphi, phivar=None):
        """Slip velocity (pyrandaIBM.py:34-47)."""
        lens = self.pyranda.mesh.GridLen * IMMERSED_EPS
        return self.slip_velocity(phi, gphi[0], gphi[1], gphi[2], vel[0], vel[1], vel[2], lens,
                                  phivar=phivar, slip=True)

    def ibmWall(self, vel, phi, gphi, phivar=None):
        """No-slip wall (pyrandaIBM.py:49-62); the interface velocity is not
        used, as in the reference."""
        lens = self.pyranda.mesh.GridLen * IMMERSED_EPS
        return self.slip_velocity(phi, gphi[0], gphi[1], gphi[2], vel[0], vel[1], vel[2], lens,
                                  phivar=None, slip=False)

    def ibmS(self, scalar, phi, gphi):
        """Scalar extension into the body (pyrandaIBM.py:65-75)."""
        return self.smooth_terrain(phi, gphi[0], gphi[1], gphi[2], scalar, 0.0)

    # -- the extension -------------------------------------------------------
    def smooth_terrain(self, SDF, gDx, gDy, gDz, val_in, epsi, new=False):
        sim = self.pyranda
        step = IMMERSED_CFL * sim.mesh.GridLen
        inside = SDF <= epsi
        val = val_in * 1.0
        for _ in range(IMMERSED_ITER):
            tvx, tvy, tvz = sim.grad(val)
            term = tvx * gDx + tvy * gDy + tvz * gDz
            val = _where(inside, val + step * term, val)
            val = _where(inside, sim.gfilter(val), val)
        return val

    def slip_velocity(self, SDF, gDx, gDy, gDz, v1_in, v2_in, v3_in, lens, new=False, phivar=None,
                      slip=True):
        v1 = v1_in * 1.0
        v2 = v2_in * 1.0
        v3 = v3_in * 1.0
        if phivar:
            # relative to the interface velocity
            v1 = v1 - phivar[0]
            v2 = v2 - phivar[1]
            v3 = v3 - phivar[2]
        v1 = self.smooth_terrain(SDF, gDx, gDy, gDz, v1, 0.0, new=new)
        v2 = self.smooth_terrain(SDF, gDx, gDy, gDz, v2, 0.0, new=new)
        v3 = self.smooth_terrain(SDF, gDx, gDy, gDz, v3, 0.0, new=new)
        band = SDF < lens
        if slip:
            norm = v1 * gDx + v2 * gDy + v3 * gDz
            vn = _where(band, norm, 0.0)
            tmp = _where(band, 0.0, norm / SDF)
            # take the normal component out, then put back its linear profile
            # through the zero level
            v1 = v1 - vn * gDx
            v2 = v2 - vn * gDy
            v3 = v3 - vn * gDz
            tmp = self.smooth_terrain(SDF, gDx, gDy, gDz, tmp, lens, new=new)
            vn = _where(band, tmp * SDF, 0.0)
            v1 = v1 + vn * gDx
            v2 = v2 + vn * gDy
            v3 = v3 + vn * gDz
        else:
            norm = torch.sqrt(v1 * v1 + v2 * v2 + v3 * v3)
            vn = _where(band, norm, 0.0)
            tmp = _where(band, 0.0, norm / SDF)
            inorm = 1.0 / (norm + 1.0e-16)
            tx = v1 * inorm
            ty = v2 * inorm
            tz = v3 * inorm
            v1 = v1 - vn * tx
            v2 = v2 - vn * ty
            v3 = v3 - vn * tz
            tmp = self.smooth_terrain(SDF, gDx, gDy, gDz, tmp, lens, new=new)
            vn = _where(band, tmp * SDF, 0.0)
            v1 = v1 + vn * tx
            v2 = v2 + vn * ty
            v3 = v3 + vn * tz
        if phivar:
            v1 = v1 + phivar[0]
            v2 = v2 + phivar[1]
            v3 = v3 + phivar[2]
        return [v1, v2, v3]
