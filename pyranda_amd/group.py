"""An ensemble of independent simulations as one pyrandaSim over one allocation.

Every field of the ensemble is one device array holding all members' fields
back to back (member m's (nx_m, ny_m, nz_m) block, x fastest, at offset off_m;
offsets are multiples of 16 elements). The EOM engine of pyrandaSim runs
unchanged on these (n, 1, 1) arrays, so every pointwise kernel, every RK
update and every operator is ONE launch for the whole ensemble:

* compact operators go through a group of the members' contexts
  (pyr_group_apply): one launch per operator and direction for all members;
* pointwise algebra is elementwise over the allocation;
* grid reductions (dt.courant, dt.diff, sum/max/min) are per member
  (segmented) and give one value per member, an (M,) FP64 device tensor;
  such values enter fused kernels per point through the member-index field,
  and rk4 takes dt as one value per member, so each member keeps its own
  time step, exactly as it would run alone.

All members share the EOM, the precision and the boundary kind of each
direction: periodic, or bounded (NONE ends, with pyrandaBC's extrap/const on
every member's boundary planes); SYMM ends are not grouped, and a direction
with a single point in every member is a null direction. Lines may differ in
length between members, up to 1024 points. No counterpart in the reference,
which advances one simulation per MPI communicator
(pyranda/pyranda.py:32-100).
"""
import ctypes

import torch

from . import _lib, fuse
from .devnumpy import set_default
from .parcop import ParcopContext
from .pyranda import pyrandaSim
from .pyrandaMesh import pyrandaMesh, defaultMeshOptions
from .pyrandaPackage import pyrandaPackage
from .pyrandaVar import pyrandaVar
from .pyrandaTimestep import pyrandaTimestep
from .pyrandaBC import pyrandaBC

ALIGN = 16


def _round_up(n, a):
    return (n + a - 1) // a * a


class GroupContext:
    """The operators of parcop over all members at once (pyr_group_*)."""

    def __init__(self, ctxs, offsets, shape, device):
        self.ctxs = ctxs
        self.shape = shape
        self.device = device
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(ctxs))(*[c._h.value for c in ctxs])
        offs = (ctypes.c_longlong * len(ctxs))(*offsets)
        _lib.check(self._L.pyr_group_create(ctypes.byref(h), len(ctxs), arr, offs))
        self._h = h
        # a direction is skipped when it is a null direction of every member
        self.null = [all(c.nn[d] <= 3 for c in ctxs) for d in range(3)]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.pyr_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _prep(self, v, dtype=None):
        if not torch.is_tensor(v):
            v = torch.full(self.shape, float(v), dtype=dtype or torch.float64, device=self.device)
        if dtype is not None and v.dtype != dtype:
            v = v.to(dtype)
        if v.dtype not in (torch.float64, torch.float32):
            v = v.to(torch.float64)
        return v.contiguous()

    def _out(self, like):
        return torch.empty(self.shape, dtype=like.dtype, device=self.device)

    def apply(self, op, d, v, out=None, epi=_lib.EPI_STORE, scale=0.0, power=0):
        v = self._prep(v)
        if out is None:
            out = self._out(v)
        self._L.pyr_group_set_stream(self._h, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        prec = _lib.F64 if v.dtype == torch.float64 else _lib.R32
        _lib.check(self._L.pyr_group_apply(self._h, op, d, prec, ctypes.c_void_p(v.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), epi, float(scale), int(power)))
        return out

    # -- parcop names -------------------------------------------------------------
    def ddx(self, v): return self.apply(_lib.OP_D1, 0, v)
    def ddy(self, v): return self.apply(_lib.OP_D1, 1, v)
    def ddz(self, v): return self.apply(_lib.OP_D1, 2, v)
    def dd8x(self, v): return self.apply(_lib.OP_D8, 0, v)
    def dd8y(self, v): return self.apply(_lib.OP_D8, 1, v)
    def dd8z(self, v): return self.apply(_lib.OP_D8, 2, v)

    def _filter3(self, op, v):
        v = self._prep(v)
        out = v
        for d in range(3):
            if not self.null[d]:            # a null filter direction is the identity
                out = self.apply(op, d, out)
        return out if out is not v else v.clone()

    def sfilter(self, v): return self._filter3(_lib.OP_SFILTER, v)
    def gfilter(self, v): return self._filter3(_lib.OP_GFILTER, v)

    def gfilterdir(self, v, direction):
        return self.apply(_lib.OP_GFILTER, int(direction) - 1, v)

    def pring(self, v):
        # ring L=2 (operators.f90:615-643): max_d |d8_d v| d_d^2 per member;
        # a null direction contributes |0| and leaves the maximum unchanged
        out = None
        for d in range(3):
            if self.null[d]:
                continue
            out = self.apply(_lib.OP_D8, d, v, out=out,
                             epi=_lib.EPI_ABS_SCALED if out is None else _lib.EPI_MAXABS_SCALED, scale=1.0, power=2)
        return out

    # a null derivative direction is zero: it is left out of sums and gives
    # a zero field on its own (as the per-member null operator would)
    def plaplacian(self, v):
        out = None
        for d in range(3):
            if not self.null[d]:
                out = self.apply(_lib.OP_D2, d, v, out=out, epi=_lib.EPI_STORE if out is None else _lib.EPI_ADD)
        return out

    def grads(self, v):
        v = self._prep(v)
        return [torch.zeros_like(v) if self.null[d] else self.apply(_lib.OP_D1, d, v) for d in range(3)]

    def divergence(self, fx, fy, fz):
        out = None
        for d, f in enumerate((fx, fy, fz)):
            if not self.null[d]:
                out = self.apply(_lib.OP_D1, d, f, out=out, epi=_lib.EPI_STORE if out is None else _lib.EPI_ADD)
        return out

    def pringv(self, vx, vy, vz):
        raise NotImplementedError("ringV over a group")


class _Der:
    def __init__(self, g): self.g = g
    def ddx(self, v): return self.g.ddx(v)
    def ddy(self, v): return self.g.ddy(v)
    def ddz(self, v): return self.g.ddz(v)
    def dd8x(self, v): return self.g.dd8x(v)
    def dd8y(self, v): return self.g.dd8y(v)
    def dd8z(self, v): return self.g.dd8z(v)
    def div(self, fx, fy, fz): return self.g.divergence(fx, fy, fz)
    def grad(self, v): return self.g.grads(v)
    def laplacian(self, v): return self.g.plaplacian(v)
    def ring(self, v): return self.g.pring(v)
    def ringV(self, vx, vy, vz): return self.g.pringv(vx, vy, vz)


class _GFil:
    def __init__(self, g): self.g = g
    def filter(self, v): return self.g.gfilter(v)
    def filterDir(self, v, direction): return self.g.gfilterdir(v, direction)


class _SFil:
    def __init__(self, g): self.g = g
    def filter(self, v): return self.g.sfilter(v)


class GroupMPI:
    """pyrandaMPI's role for a group: allocation, operators, per-member
    reductions (device values, one per member)."""

    def __init__(self, ctxs, offsets, sizes, shape, seg, dtype, device):
        self.ctxs = ctxs
        self.offsets = offsets
        self.sizes = sizes
        self.shape = shape
        self.seg = seg
        self.dtype = dtype
        self.device = device
        self.master = True
        self.device_scalars = True
        self.periodic = [all(c.periodic[d] for c in ctxs) for d in range(3)]
        self.group = GroupContext([c.ctx for c in ctxs], offsets, shape, device)
        self.der = _Der(self.group)
        self.gfil = _GFil(self.group)
        self.fil = _SFil(self.group)
        self.nx, self.ny, self.nz = shape

    def emptyScalar(self):
        return torch.zeros(self.shape, dtype=self.dtype, device=self.device)

    def emptyVector(self, rank=3):
        return torch.zeros(tuple(self.shape) + (rank,), dtype=self.dtype, device=self.device)

    def setPatch(self):
        pass

    def _seg_reduce(self, data, how):
        x = data.reshape(-1).double()
        mid = self.shape.mid.long()
        valid = self.shape.valid
        init = {"sum": 0.0, "amax": -float("inf"), "amin": float("inf")}[how]
        out = torch.full((self.seg.nmem,), init, dtype=torch.float64, device=self.device)
        return out.scatter_reduce(0, mid[valid], x[valid], reduce=how, include_self=True)

    def sum3D(self, data):
        return self._seg_reduce(data, "sum")

    def max3D(self, data):
        return self._seg_reduce(data, "amax")

    def min3D(self, data):
        return self._seg_reduce(data, "amin")

    def iprint(self, s):
        print(s, flush=True)


class _Member:
    """One member's mesh and context."""

    def __init__(self, mesh_options, device):
        self.mesh = pyrandaMesh()
        if isinstance(mesh_options, str):
            self.mesh.makeMeshStr(mesh_options)
            mesh_options = self.mesh.options
        opts = defaultMeshOptions()
        opts.update(mesh_options)
        if opts.get("coordsys", 0) != 0:
            raise NotImplementedError("only Cartesian members")
        self.opts = opts
        self.nn = tuple(int(n) for n in opts["nn"])
        self.periodic = [bool(p) for p in opts["periodic"]]
        sym = opts.get("symmetric") or [[False, False]] * 3
        if any(any(s) for s in sym):
            raise NotImplementedError("group members take periodic or bounded (NONE) directions, not SYMM")
        x1, xn = opts["x1"], opts["xn"]
        bcs = []
        for d in range(3):
            bcs += ["PERI" if self.periodic[d] else "NONE"] * 2
        self.ctx = ParcopContext(self.nn, (x1[0], x1[1], x1[2]), (xn[0], xn[1], xn[2]), bcs, device=device)
        self.npts = self.nn[0] * self.nn[1] * self.nn[2]


class _GroupBC(pyrandaBC):
    """pyrandaBC's extrap/const (pyrandaBC.py:58-160) on every member's
    boundary planes at once: the planes' points (offsets into the shared
    allocation) are gathered per direction once, and each update is one
    gather/scatter over all members, with the arithmetic of the single
    simulation (bit for bit its result)."""

    def __init__(self, pysim):
        pyrandaBC.__init__(self, pysim)
        self._idx = {}

    def _planes(self, direction):
        if direction not in self._idx:
            g = self.pyranda
            axis = {"x": 0, "y": 1, "z": 2}[direction[0]]
            hi = direction[1:] == "n"
            sets = ([], [], [])
            for mem, off in zip(g.members, g.offsets):
                nx, ny, nz = mem.nn
                n = mem.nn[axis]
                if mem.periodic[axis] or n < 3:
                    raise ValueError("bc on %s: member %s is periodic or too short along it" % (direction, mem.nn))
                ii, jj, kk = torch.arange(nx), torch.arange(ny), torch.arange(nz)
                for r, pos in enumerate([n - 1, n - 2, n - 3] if hi else [0, 1, 2]):
                    # the plane's points, in the same order for the three planes
                    if axis == 0:
                        flat = pos + nx * jj[None, :] + nx * ny * kk[:, None]
                    elif axis == 1:
                        flat = ii[None, :] + nx * pos + nx * ny * kk[:, None]
                    else:
                        flat = ii[None, :] + nx * jj[:, None] + nx * ny * pos
                    sets[r].append(off + flat.reshape(-1))
            dev = g.PyMPI.device
            self._idx[direction] = tuple(torch.cat(s_).to(dev) for s_ in sets)
        return self._idx[direction]

    def extrapolate(self, var, direction, order):
        i0, i1, i2 = self._planes(direction)
        d = self.pyranda.variables[var].data.view(-1)
        if order == 2:
            d[i0] = 2 * d[i1] - d[i2]
        else:
            d[i0] = d[i1]

    def constant(self, var, direction, val):
        i0, _, _ = self._planes(direction)
        # the value as a kernel argument (no host-to-device copy: capturable)
        self.pyranda.variables[var].data.view(-1).index_fill_(0, i0, float(val))

    def field(self, var, direction, field):
        raise NotImplementedError("bc.field over a group")


class _GroupTimestep(pyrandaTimestep):
    """dt.courant / dt.diff per member (pyrandaTimestep.py:42-76): segmented
    one-pass reductions with each member's spacings; one value per member."""

    def __init__(self, pysim):
        pyrandaPackage.__init__(self, "Timestep", pysim)
        self.pysim = pysim
        self._courant = fuse.SegmentedReduce("max", _courant_body(), 4, 4, "courant")
        self._diff = fuse.SegmentedReduce("min", _diff_body(), 2, 1, "diff")

    def courant(self, u, v, w, c):
        g = self.pysim
        fields = [g._field(x) for x in (u, v, w, c)]
        r = self._courant(g.seg, fields, g.member_spacings)
        return r.reciprocal()

    def diff(self, bulk, density):
        g = self.pysim
        return self._diff(g.seg, [g._field(bulk), g._field(density)], g.member_gridlen)

    def diffDir(self, bulk, density, delta):
        raise NotImplementedError("dt.diffDir over a group")


def _timestep_module():
    # the package exports the class under the module's name
    import importlib
    return importlib.import_module(__package__ + ".pyrandaTimestep")


def _courant_body():
    return _timestep_module()._COURANT.body


def _diff_body():
    return _timestep_module()._DIFF.body


class pyrandaGroupSim(pyrandaSim):
    """Independent simulations on their own grids advanced as one: same
    EOM/setIC/rk4 API as pyrandaSim; every launch covers all members.

    meshes: mesh strings or option dicts, one per member. Per-member values
    (dt, time) are (M,) FP64 device tensors; `member_field(name, m)` gives
    member m's (nx, ny, nz) view of a field."""

    def __init__(self, name, meshes, silent=True, precision="fp64", device="cuda", fp64_state=True):
        if precision not in ("fp64", "fp32", "fp32-fp64solve"):
            raise ValueError("precision must be 'fp64', 'fp32' or 'fp32-fp64solve'")
        if not torch.cuda.is_available():
            raise RuntimeError("pyranda_amd needs a HIP device (no GPU visible)")
        dev = torch.device(device)
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.name = name
        self.silent = silent
        # the update of a group reads one dt per member: F is stored first
        self.fuse_rk = False
        self.precision = precision
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        # every field of a group is stored in the run's precision (no FP64
        # storage of offset-dominated fields, pyrandaSim(wide_fields=None))
        self._storage = {}
        self.value_storage = self.named_storage = self.wide_if_wide_input = False
        self.members = [_Member(m, dev) for m in meshes]
        if precision == "fp32-fp64solve":
            for m in self.members:
                m.ctx.set_r32_solve("fp64")
        for d in range(3):
            flags = {m.nn[d] <= 1 for m in self.members}
            if len(flags) != 1:
                raise ValueError("members must agree on which directions have a single point")
        offs, o = [], 0
        for m in self.members:
            offs.append(o)
            o = _round_up(o + m.npts, ALIGN)
        ntot = o
        self.offsets = offs
        sizes = [m.npts for m in self.members]
        M = len(self.members)
        mid = torch.empty(ntot, dtype=torch.int32)
        valid = torch.zeros(ntot, dtype=torch.bool)
        for j, (off, n) in enumerate(zip(offs, sizes)):
            end = offs[j + 1] if j + 1 < M else ntot
            mid[off:end] = j                 # alignment padding belongs to the member before it
            valid[off:off + n] = True
        shape = fuse.GroupShape(ntot, mid.to(dev), M)
        shape.valid = valid.to(dev)
        self.seg = fuse.SegmentTable(sizes, offs, dev)
        self.PyMPI = GroupMPI(self.members, offs, sizes, shape, self.seg, self.dtype, dev)
        self.nx, self.ny, self.nz = ntot, 1, 1
        self.npts = ntot
        # per-member spacings: courant's dx, dy, dz, GridLen and diff's GridLen
        sp, gl = [], []
        for m in self.members:
            dp = m.ctx.dmesh
            sp.append([dp[0], dp[1], dp[2], m.ctx.gridlen])
            gl.append([m.ctx.gridlen])
        self.member_spacings = torch.tensor(sp, dtype=torch.float64, device=dev)
        self.member_gridlen = torch.tensor(gl, dtype=torch.float64, device=dev)
        # the group's mesh: coordinates of every member in the shared layout
        self.mesh = pyrandaMesh()
        self.mesh.options = defaultMeshOptions()
        self.mesh.coordsys = 0
        self.mesh.dims = 3 - sum(1 for d in range(3) if self.members[0].nn[d] == 1)
        self.mesh.PyMPI = self.PyMPI
        self.mesh.shape = shape
        self.mesh.GridLen = self.member_gridlen[:, 0].clone()
        # per-member spacings (pyrandaTimestep reads them before addPackage
        # swaps in the segmented version)
        self.mesh.d1, self.mesh.d2, self.mesh.d3 = (self.member_spacings[:, i].clone() for i in range(3))
        self.mesh.coords = []
        for ax, nm in enumerate(("meshx", "meshy", "meshz")):
            buf = torch.zeros(ntot, dtype=torch.float64)
            for m, off in zip(self.members, offs):
                nx, ny, nz = m.nn
                c = torch.tensor(m.ctx.coords1d[ax], dtype=torch.float64)
                if ax == 0:
                    f = c[None, None, :].expand(nz, ny, nx)
                elif ax == 1:
                    f = c[None, :, None].expand(nz, ny, nx)
                else:
                    f = c[:, None, None].expand(nz, ny, nx)
                buf[off:off + m.npts] = f.reshape(-1)
            var = pyrandaVar(nm, "mesh", "scalar")
            var.PyMPI = self.PyMPI
            var.allocated = True
            var.data = buf.to(dev).reshape(shape)
            self.mesh.coords.append(var)
        self.meshOptions = self.mesh.options
        self.equations = []
        self.conserved = []
        self.variables = {}
        self.initial_conditions = []
        self.nPDE = self.nALG = self.nELP = 0
        set_default(dev, self.dtype)
        self.zero = self.PyMPI.emptyScalar()
        self.packages = {}
        self.packagesRestart = []
        self.userDefined = {}
        self.get_sMap()
        self.time = 0.0
        self.cycle = 0
        self.deltat = 0.0
        self._shadow = {}
        self.fp64_state = fp64_state
        self._state64 = {}
        self._state_src = {}

    def checkForNan(self, names=None):
        """NaN check over the members' points (not the alignment padding)."""
        valid = self.mesh.shape.valid
        svars = ""
        for ivar in (names or list(self.variables)):
            d = self.variables[ivar].data if ivar in self.variables else None
            if not torch.is_tensor(d):
                continue
            x = d.reshape(-1)[valid] if tuple(d.shape) == tuple(self.mesh.shape) else d
            if torch.isnan(x).any().item():
                svars += ivar + " "
        return svars

    # -- helpers ---------------------------------------------------------------
    def _field(self, x):
        if torch.is_tensor(x) and tuple(x.shape) == tuple(self.mesh.shape):
            return x.contiguous()
        return torch.as_tensor(x, dtype=self.dtype, device=self.PyMPI.device).expand(self.mesh.shape).contiguous()

    def member_field(self, name, m):
        """Member m's field as an (nx, ny, nz) view (x fastest)."""
        mem = self.members[m]
        nx, ny, nz = mem.nn
        d = self.variables[name].data.reshape(-1)
        off = self.offsets[m]
        return d[off:off + mem.npts].reshape(nz, ny, nx).permute(2, 1, 0)

    def addPackage(self, package):
        if isinstance(package, pyrandaTimestep) and not isinstance(package, _GroupTimestep):
            package = _GroupTimestep(self)
        if isinstance(package, pyrandaBC) and not isinstance(package, _GroupBC):
            package = _GroupBC(self)
        pyrandaSim.addPackage(self, package)

    # -- operators (null directions of every member give 0) ---------------------
    def ddx(self, val):
        return 0.0 if self.PyMPI.group.null[0] else self.PyMPI.der.ddx(val)

    def ddy(self, val):
        return 0.0 if self.PyMPI.group.null[1] else self.PyMPI.der.ddy(val)

    def ddz(self, val):
        return 0.0 if self.PyMPI.group.null[2] else self.PyMPI.der.ddz(val)

    def div(self, f1, f2=None, f3=None):
        z = self.zero
        nul = self.PyMPI.group.null
        if f2 is None and f3 is None:
            for d in range(3):
                if not nul[d]:
                    args = [z, z, z]
                    args[d] = f1
                    return self.PyMPI.der.div(*args)
            return 0.0
        if f3 is None:
            live = [d for d in range(3) if not nul[d]]
            args = [z, z, z]
            for d, f in zip(live[:2], (f1, f2)):
                args[d] = f
            return self.PyMPI.der.div(*args)
        return self.PyMPI.der.div(f1, f2, f3)

    # -- time integration ---------------------------------------------------------
    def _rk_update(self, U, F, PHI, dt, A, B, fp32, store=True):
        u = self.variables[U].data
        if not (torch.is_tensor(dt) and dt.dim() == 1 and dt.shape[0] == len(self.members)
                and torch.is_tensor(F) and tuple(F.shape) == tuple(self.mesh.shape) and F.dtype == self.dtype):
            dtv = dt[self.mesh.shape.mid.long()].reshape(self.mesh.shape) if (
                torch.is_tensor(dt) and dt.dim() == 1) else dt
            return pyrandaSim._rk_update(self, U, F, PHI, dtv, A, B, fp32, store)
        out = torch.empty(self.mesh.shape, dtype=self.dtype, device=u.device)
        shadow = self._shadow[U].data_ptr() if fp32 and self._shadow[U] is not None else None
        dt64 = dt.to(torch.float64).contiguous()
        stream = ctypes.c_void_p(torch.cuda.current_stream(u.device).cuda_stream)
        _lib.check(_lib.lib().pyr_rk_update_member_dt(
            stream, _lib.R32 if fp32 else _lib.F64, self.npts, ctypes.c_void_p(F.contiguous().data_ptr()),
            ctypes.c_void_p(PHI[U].data_ptr()), ctypes.c_void_p(u.contiguous().data_ptr()),
            ctypes.c_void_p(out.data_ptr()), shadow, ctypes.c_void_p(dt64.data_ptr()),
            ctypes.c_void_p(self.mesh.shape.mid.data_ptr()), float(A), float(B)))
        return out
