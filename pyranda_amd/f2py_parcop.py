"""Drop-in for pyranda's f2py module `parcop.parcop` over the C-ABI.

pyranda reaches its native operators only as `parcop.parcop.<name>(...)`
(pyranda/pyrandaMPI.py:151-155,247,664-740; pyranda/pyrandaMesh.py:85-208),
with host numpy arrays in Fortran order and a new array returned per call
(pyranda/parcop/parcop.f90:23-379). A maintainer who keeps pyranda's Python
replaces `from .parcop import parcop` (pyranda/parcop/__init__.py:1) with

    from pyranda_amd.f2py_parcop import parcop

and every call site runs on the GPU: each call stages its arguments into
device memory, runs the HIP operator and returns a new Fortran-order float64
array (the PCIe-inclusive path; pyranda_amd.pyrandaSim keeps fields on the
device instead). One Cartesian block per (patch, level), one rank.
"""
import ctypes

import numpy as np
import torch

from . import _lib

_BC = {"PERI": 0, "NONE": 1, "SYMM": 2}


class _Parcop:
    """The `parcop.parcop` names pyrandaMPI and pyrandaMesh call."""

    def __init__(self):
        self._ctx = {}          # (patch, level) -> (pyr_ctx*, (nx, ny, nz))
        self._cur = None
        self._nn = None
        self.device = None

    # -- setup (parcop.f90:23-61, 196-200) -------------------------------------
    def setup(self, patch, level, comm, nx, ny, nz, px, py, pz, coordsys, x1, xn, y1, yn, z1, zn,
              bx1, bxn, by1, byn, bz1, bzn):
        if (px, py, pz) != (1, 1, 1):
            raise ValueError("one rank per block: px = py = pz = 1")
        if coordsys != 0:
            raise ValueError("Cartesian meshes only (coordsys 0)")
        if self.device is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        L = _lib.lib()
        h = ctypes.c_void_p()
        bc = (ctypes.c_int * 6)(*[_BC[str(b).strip().upper()] for b in (bx1, bxn, by1, byn, bz1, bzn)])
        _lib.check(L.pyr_setup(ctypes.byref(h), int(nx), int(ny), int(nz), float(x1), float(xn), float(y1),
                               float(yn), float(z1), float(zn), bc))
        old = self._ctx.get((patch, level))
        if old is not None:
            L.pyr_destroy(old[0])
        self._ctx[(patch, level)] = (h, (int(nx), int(ny), int(nz)))
        self.set_patch(patch, level)

    def set_patch(self, patch, level):
        self._cur, self._nn = self._ctx[(patch, level)]

    # -- staging ---------------------------------------------------------------
    def _dev(self, val):
        a = np.asarray(val, dtype=np.float64)
        if a.shape != self._nn:
            a = np.broadcast_to(a, self._nn)
        # the device array (nz, ny, nx) C-order is the Fortran (nx, ny, nz) layout
        return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).to(self.device)

    @staticmethod
    def _host(t):
        return np.asfortranarray(t.cpu().numpy().transpose(2, 1, 0))

    def _call(self, fn, ins, nout=1, extra=()):
        d = [self._dev(v) for v in ins]
        o = [torch.empty(tuple(reversed(self._nn)), dtype=torch.float64, device=self.device) for _ in range(nout)]
        L = _lib.lib()
        L.pyr_set_stream(self._cur, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        args = [ctypes.c_void_p(t.data_ptr()) for t in d + o]
        _lib.check(getattr(L, fn)(self._cur, *args, *extra))
        out = [self._host(t) for t in o]
        return out[0] if nout == 1 else tuple(out)

    # -- operators (parcop.f90:202-379) ------------------------------------------
    def ddx(self, val, nx=None, ny=None, nz=None): return self._call("pyr_ddx_f64", [val])
    def ddy(self, val, nx=None, ny=None, nz=None): return self._call("pyr_ddy_f64", [val])
    def ddz(self, val, nx=None, ny=None, nz=None): return self._call("pyr_ddz_f64", [val])
    def dd8x(self, val, nx=None, ny=None, nz=None): return self._call("pyr_dd8x_f64", [val])
    def dd8y(self, val, nx=None, ny=None, nz=None): return self._call("pyr_dd8y_f64", [val])
    def dd8z(self, val, nx=None, ny=None, nz=None): return self._call("pyr_dd8z_f64", [val])
    def plaplacian(self, val, nx=None, ny=None, nz=None): return self._call("pyr_plaplacian_f64", [val])
    def pring(self, val, nx=None, ny=None, nz=None): return self._call("pyr_pring_f64", [val])
    def sfilter(self, val, nx=None, ny=None, nz=None): return self._call("pyr_sfilter_f64", [val])
    def gfilter(self, val, nx=None, ny=None, nz=None): return self._call("pyr_gfilter_f64", [val])

    def pringv(self, vx, vy, vz, nx=None, ny=None, nz=None):
        return self._call("pyr_pringv_f64", [vx, vy, vz])

    def gfilterdir(self, val, direction, nx=None, ny=None, nz=None):
        return self._call("pyr_gfilterdir_f64", [val], extra=(ctypes.c_int(int(direction)),))

    def grads(self, val, nx=None, ny=None, nz=None):
        return self._call("pyr_grads_f64", [val], nout=3)

    def divergence(self, fx, fy, fz, nx=None, ny=None, nz=None):
        return self._call("pyr_divergence_f64", [fx, fy, fz])

    # -- mesh getters (parcop.f90:130-191) ----------------------------------------
    def _axis(self, ax):
        n = self._nn[ax]
        buf = (ctypes.c_double * n)()
        _lib.check(_lib.lib().pyr_axis_coords(self._cur, ax, buf))
        return np.array(buf[:], dtype=np.float64)

    def _grid(self, ax):
        c = self._axis(ax)
        shape = [1, 1, 1]
        shape[ax] = len(c)
        return np.asfortranarray(np.broadcast_to(c.reshape(shape), self._nn))

    def xgrid(self, nx=None, ny=None, nz=None): return self._grid(0)
    def ygrid(self, nx=None, ny=None, nz=None): return self._grid(1)
    def zgrid(self, nx=None, ny=None, nz=None): return self._grid(2)

    def _spacing(self, ax):
        d = (ctypes.c_double * 3)()
        _lib.check(_lib.lib().pyr_mesh_scalars(self._cur, None, d, None))
        return np.asfortranarray(np.full(self._nn, d[ax]))

    def dxgrid(self, nx=None, ny=None, nz=None): return self._spacing(0)
    def dygrid(self, nx=None, ny=None, nz=None): return self._spacing(1)
    def dzgrid(self, nx=None, ny=None, nz=None): return self._spacing(2)

    def mesh_getcellvol(self, nx=None, ny=None, nz=None):
        d = (ctypes.c_double * 3)()
        _lib.check(_lib.lib().pyr_mesh_scalars(self._cur, d, None, None))
        return np.asfortranarray(np.full(self._nn, d[0] * d[1] * d[2]))

    def mesh_getgridlen(self, nx=None, ny=None, nz=None):
        g = ctypes.c_double()
        _lib.check(_lib.lib().pyr_mesh_scalars(self._cur, None, None, ctypes.byref(g)))
        return np.asfortranarray(np.full(self._nn, g.value))


parcop = _Parcop()
