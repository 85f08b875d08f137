"""pyrandaSim: pyranda's simulation object over GPU-resident fields.

Same public API as pyranda/pyranda.py:32-938 (EOM, setIC, rk4, eval, var,
addPackage, ...), with every field a torch tensor in HBM and every compact
operator a HIP kernel from libpyranda_amd.so. Nothing crosses to the host
inside a time step except the scalar reductions the EOM asks for (dt.courant,
sum(...)), exactly where the reference does an MPI allreduce.

Precision: `precision="fp64"` (default) keeps every field in FP64 like the
reference. `precision="fp32"` stores fields in FP32 and runs the operators as
FP32 sweeps plus one FP64-residual refinement; the RK4 state accumulation
(PHI and the conserved variables) stays in FP64 shadows.
`precision="fp32-fp64solve"` stores fields in FP32 like "fp32" but forms the
operators' right-hand sides and solves in FP64 registers (one rounding to
FP32 per result): on gfx950 FP32 and FP64 vector FMAs issue at the same rate,
so this keeps the halved memory traffic without the refinement's second solve.

Storage of FP32 runs (`wide_fields`): FP32 storage rounds a value to 6e-8 of
its magnitude, so a field whose values sit close to one large value (the
pressure and total energy of a low-Mach flow, stresses carrying -p, the
momentum fluxes) loses most of its variation, and a derivative turns that
rounding into noise. In a low-Mach flow the fluxes' derivatives also cancel
to O(M^2) of their size, so FP32 rounding of velocities and fluxes reaches
the right-hand sides amplified. wide_fields="auto" (the default of the fp32
modes) therefore stores in FP64: every field offset-dominated when first set
(fuse.offset_dominated), every field a pointwise line computes from FP64
values only (primitives from the FP64 conserved states), and every unnamed
pointwise result (flux argument) with an FP64 input; FP32 storage stays for
operator results (derivatives, the LAD ring and filters) and what is computed
from them. wide_fields="offset" keeps only the offset-dominated fields in
FP64 (faster, less accurate); None stores every field in FP32; a collection
of names stores exactly those variables in FP64. Decisions are kept for the
run.

`device_scalars=True` keeps the grid reductions (dt.courant, dt.diff, sum,
max, min) on the GPU as 0-d FP64 tensors instead of Python floats, and rk4
accepts such a tensor as dt: a time step then makes no host round trip and
can be captured in a HIP graph (pyranda_amd.ensemble). Values are the same
as with host scalars, bit for bit.
"""
import math
import os
import random
import re
import sys
import time as _time

import torch

from .devnumpy import numpy, scipy, set_default  # noqa: F401  (EOM/IC strings use these names)
from .pyrandaEq import pyrandaEq
from .pyrandaMesh import pyrandaMesh, defaultMeshOptions
from .pyrandaMPI import pyrandaMPI
from .pyrandaUtils import fortran3d, findVar, splitLines  # noqa: F401
from .pyrandaVar import pyrandaVar
from .pyrandaIO import pyrandaIO
from . import pyrandaIO as _io

# 5-stage, 2N-storage fourth-order Runge-Kutta coefficients (Carpenter &
# Kennedy 1994), as pyranda.py:760-779 uses them.
RK_A = [0.0, -6234157559845. / 12983515589748., -6194124222391. / 4410992767914.,
        -31623096876824. / 15682348800105., -12251185447671. / 11596622555746.]
RK_B = [494393426753. / 4806282396855., 4047970641027. / 5463924506627.,
        9795748752853. / 13190207949281., 4009051133189. / 8539092990294.,
        1348533437543. / 7166442652324.]
RK_ETA = [494393426753. / 4806282396855., 4702696611523. / 9636871101405.,
          3614488396635. / 5249666457482., 9766892798963. / 10823461281321., 1.0]


class pyrandaSim:

    def __init__(self, name, meshOptions, silent=False, precision="fp64", device="cuda",
                 device_scalars=False, fp64_state=True, wide_fields="auto"):
        self.name = name
        # fp32 modes: conserved variables rewritten only by the conservative
        # filter keep an FP64 state through it (False: the FP32 field is the state)
        self.fp64_state = fp64_state
        # each PDE right-hand side's last pointwise kernel runs inside the RK
        # update (PYRANDA_AMD_FUSE_RK=0: F is stored, then read by the update)
        self.fuse_rk = os.environ.get("PYRANDA_AMD_FUSE_RK", "1") != "0"
        self.silent = silent
        self.mesh = pyrandaMesh()
        if isinstance(meshOptions, str):
            self.mesh.makeMeshStr(meshOptions)
            meshOptions = self.mesh.options
        opts = defaultMeshOptions()
        opts.update(meshOptions)
        meshOptions = opts
        self.meshOptions = opts
        self.mesh.options = opts
        if "coordsys" in opts:
            self.mesh.coordsys = opts["coordsys"]
        else:
            raise ValueError("No suitable mesh type specified.")
        if precision not in ("fp64", "fp32", "fp32-fp64solve"):
            raise ValueError("precision must be 'fp64', 'fp32' or 'fp32-fp64solve'")
        self.precision = precision
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        if not torch.cuda.is_available():
            raise RuntimeError("pyranda_amd needs a HIP device (no GPU visible)")
        nx, ny, nz = opts["nn"]
        self.dx = (opts["xn"][0] - opts["x1"][0]) / max(nx - 1, 1)
        self.dy = (opts["xn"][1] - opts["x1"][1]) / max(ny - 1, 1)
        self.dz = (opts["xn"][2] - opts["x1"][2]) / max(nz - 1, 1)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.npts = nx * ny * nz
        self.equations = []
        self.conserved = []
        self.variables = {}
        self.initial_conditions = []
        self.nPDE = self.nALG = self.nELP = 0
        self.mesh.dims = opts["dim"]
        self.PyMPI = pyrandaMPI(self.mesh, comm=opts.get("comm"), dtype=self.dtype, device=device)
        self.PyMPI.device_scalars = bool(device_scalars)
        if precision == "fp32-fp64solve":
            self.PyMPI.ctx.set_r32_solve("fp64")
        self.mesh.PyMPI = self.PyMPI
        self.mesh.makeMesh()
        set_default(self.PyMPI.device, self.dtype)
        self.zero = self.PyMPI.emptyScalar()
        self.packages = {}
        self.packagesRestart = []
        self.userDefined = {}
        self.get_sMap()
        self.time = 0.0
        self.cycle = 0
        self.deltat = 0.0
        self.vizDumpHistory = []
        self.PyIO = pyrandaIO(self.name)
        self.eom = ""
        self.ics = ""
        self._shadow = {}  # FP64 RK state of conserved variables (fp32 mode)
        self._state64 = {}  # persistent FP64 state of self-filtered conserved variables
        self._state_src = {}  # the FP32 field last derived from it
        # storage of named fields in fp32 modes: name -> "f64"/"f32"
        self._storage = {}
        self.value_storage = False   # undecided fields are decided by value
        self.named_storage = False   # named fields have a storage of their own
        self.wide_if_wide_input = False
        if wide_fields not in ("auto", "offset", None) and isinstance(wide_fields, str):
            raise ValueError("wide_fields: 'auto', 'offset', None or a collection of names")
        if precision != "fp64":
            if wide_fields in ("auto", "offset"):
                self.value_storage = self.named_storage = True
                # auto: FP64 values stay FP64 through pointwise algebra
                self.wide_if_wide_input = wide_fields == "auto"
            elif wide_fields is not None:
                self._storage = {n: "f64" for n in wide_fields}
                self.named_storage = True

    def autotune(self, **kw):
        """Tune the line-operator launches of this grid and precision on the
        device (ParcopContext.autotune); returns the timing report
        {(op, dir): (choice, times)}. FP32 runs whose operators also see FP64
        fields (an FP64 state through the conservative filter, fields stored
        in FP64) tune the FP64 launches too, reported under (op, dir, "f64");
        the filter's z sweep is timed storing the FP32 rounding as well when
        the state is filtered in FP64."""
        from . import _lib
        ctx = self.PyMPI.ctx
        rep = ctx.autotune(self.dtype, **kw)
        if self.dtype != torch.float64:
            wide = any(v == "f64" for v in self._storage.values())
            filt = self._self_filtered()
            ops = kw.get("ops") or (_lib.OP_D1, _lib.OP_D2, _lib.OP_D8, _lib.OP_SFILTER, _lib.OP_GFILTER)
            passes = []
            if wide:
                passes.append(([o for o in ops if not (filt and o == _lib.OP_SFILTER)], False))
            if filt or (wide and _lib.OP_SFILTER in ops):
                passes.append(([_lib.OP_SFILTER], filt and self.dtype == torch.float32))
            for pops, rounded in passes:
                if not pops:
                    continue
                kw64 = dict(kw)
                kw64["ops"] = tuple(pops)
                kw64["rounded"] = rounded
                for (op, d), r in ctx.autotune(torch.float64, **kw64).items():
                    rep[(op, d, "f64")] = r
        return rep

    # -- bookkeeping -------------------------------------------------------------
    def iprint(self, sprnt, wpm=0):
        if self.PyMPI.master and not self.silent:
            print(sprnt)
            sys.stdout.flush()

    def addPackage(self, package):
        self.packages[package.name] = package
        self.packagesRestart.append(package.__module__)
        package.get_sMap()
        self.sMap.update(package.sMap)

    def addUserDefinedFunction(self, name, equation):
        if name in self.sMap:
            self.iprint("Error: Cant add user-defined equation '%s', it already exists" % name)
        self.userDefined[name] = equation
        self.sMap.update({"%s(" % name: "self.userDefined['%s'](self," % name})

    def allocate(self):
        for v in self.variables:
            self.variables[v].__allocate__(self.PyMPI)

    def var(self, name):
        if name in self.variables:
            return self.variables[name]
        raise ValueError("Error: variable name: %s not found in database" % name)

    def x(self):
        return self.variables["meshx"]

    def y(self):
        return self.variables["meshy"]

    def z(self):
        return self.variables["meshz"]

    def eval(self, expression):
        return eval(fortran3d(expression, self.sMap))

    def parse(self, expression, parseDict=None):
        if parseDict:
            expression = self.updateStringFromDictionary(expression, parseDict, "parse")
        exec(fortran3d(expression, self.sMap))

    def addVar(self, name, kind=None, rank="scalar"):
        if name not in self.variables:
            self.variables[name] = pyrandaVar(name, kind, rank)

    def addEqu(self, equation):
        peq = pyrandaEq(equation, self.sMap, self.PyMPI)
        self.equations.append(peq)
        if peq.kind == "PDE":
            self.nPDE += 1
            self.conserved.append(peq.LHS[0])
            if len(peq.LHS) > 1:
                raise ValueError("only single return values for PDEs allowed")
        elif peq.kind == "ALG":
            self.nALG += 1
        else:
            raise ValueError("No suitable equation type specified for %s" % peq.eqstr)

    def updateStringFromDictionary(self, string, dictionary, handle):
        for item in dictionary:
            if item in string:
                string = string.replace(item, str(dictionary[item]))
            else:
                self.iprint("Notice: Given %s dictionary value: '%s' does not exists in the %s string."
                            % (handle, item, handle))
        return string

    # -- EOM / IC (pyranda.py:231-369) ------------------------------------------
    def EOM(self, eom, eomDict=False):
        if eomDict:
            eom = self.updateStringFromDictionary(eom, eomDict, "EOM")
        self.eom = eom
        eom_lines = [el for el in eom.split("\n") if el.strip()]
        eom_lines = [el for el in eom_lines if el.strip()[0] != "#"]  # comments
        names = []
        for eq in eom_lines:
            names += findVar(eq, "scalar")
        for evar in sorted(set(names)):
            self.addVar(evar, kind="conserved")
        names = []
        for eq in eom_lines:
            names += findVar(eq, "vector")
        for evar in sorted(set(names)):
            self.addVar(evar, rank="vector", kind="conserved")
        for variable in self.variables:
            self.iprint("Adding variables: %s" % variable)
        self.iprint("Adding equations of motion: ")
        for n, eq in enumerate(eom_lines, 1):
            self.addEqu(eq)
            self.iprint("(%s)   %s" % (n, eq))
        self.allocate()
        self.variables["meshx"] = self.mesh.coords[0]
        self.variables["meshy"] = self.mesh.coords[1]
        self.variables["meshz"] = self.mesh.coords[2]

    def checkForNan(self, names=None):
        svars = ""
        for ivar in (names or list(self.variables)):
            if ivar not in self.variables:
                self.iprint("%s is not a variable" % ivar)
                continue
            d = self.variables[ivar].data
            bad = torch.isnan(d).any().item() if torch.is_tensor(d) else (d is not None and math.isnan(float(d)))
            if bad:
                svars += ivar + " "
        return svars

    def setIC(self, ics, icDict=False, addOnly=False, verbose=False):
        if icDict:
            ics = self.updateStringFromDictionary(ics, icDict, "ICs")
        self.ics = ics
        ic_lines = splitLines(ics)
        names = []
        for eq in ic_lines:
            names += findVar(eq, "scalar")
        for evar in sorted(set(names)):
            self.addVar(evar, kind="conserved")
        self.allocate()
        self.initial_conditions.extend(ic_lines)
        if addOnly:
            return
        # ICs are evaluated in FP64 (FP64 coordinates), then stored in the run's dtype
        scope = {"self": self}
        for ic in ic_lines:
            if verbose:
                self.iprint(ic)
            try:
                exec(fortran3d(ic, self.sMap), globals(), scope)
            except Exception as e:
                self.iprint("Error: cant parse following string from ICs")
                self.iprint(ic)
                self.iprint("Error message: %s" % e)
                raise
        if self.dtype != torch.float64:
            # the FP64 state of the conserved variables starts from the FP64
            # initial conditions
            self._state64 = {}
            for U in self._self_filtered():
                d = self.variables[U].data
                if torch.is_tensor(d) and tuple(d.shape) == tuple(self.mesh.shape):
                    self._state64[U] = d.to(torch.float64)
        # storage decided by value only for fields the ICs set (the others
        # hold their allocation's zeros until an ALG line writes them); with
        # wide_fields="auto" the fields pointwise ALG lines rewrite are
        # decided at their first evaluation, from its inputs
        deferred = self._pointwise_alg_lhs() if self.wide_if_wide_input else set()
        self._cast_fields(decide=set(names) - deferred)
        for U in self._state64:
            self._mark_state(U)
        snans = self.checkForNan(sorted(set(names)))
        if snans:
            raise FloatingPointError("Found some nans in inits: %s" % snans)
        self.updateVars(verbose=verbose)
        snans = self.checkForNan()
        if snans:
            raise FloatingPointError("Found some nans in Update after init: %s" % snans)

    def _cast_fields(self, decide=None):
        """Fields take the run's dtype (fp32 modes: their storage, see
        storage_of, decided by value for the names in `decide`, default all);
        device scalars (reductions kept on the GPU, one value or one per
        ensemble member) stay FP64 like the host floats they stand for."""
        for name, var in self.variables.items():
            if name.startswith("mesh"):
                continue
            d = var.data
            if torch.is_tensor(d) and tuple(d.shape) == tuple(self.mesh.shape) and d.is_floating_point():
                if decide is None or name in decide or self.storage_of(name) is not None:
                    want = self._storage_dtype(name, d)
                else:
                    want = self.dtype
                if name in self._state64 and want == torch.float64:
                    # an FP64 field with an FP64 state: the state is the field
                    var.data = self._state64[name]
                elif d.dtype != want:
                    var.data = d.to(want)

    # -- storage of fp32 runs' fields (see the module docstring) -----------------
    def _is_field(self, t):
        return torch.is_tensor(t) and tuple(t.shape) == tuple(self.mesh.shape) and t.is_floating_point()

    def storage_of(self, name):
        """Storage of a named field: "f64", "f32", or None while an fp32 run
        with storage by value has not seen the field yet."""
        if self.dtype == torch.float64:
            return "f64"
        st = self._storage.get(name)
        if st is None and not self.value_storage:
            return "f32"
        return st

    def _storage_dtype(self, name, t):
        """Storage dtype of `name`, deciding it from `t` if undecided."""
        st = self.storage_of(name)
        if st is None:
            from .fuse import offset_dominated
            st = self._storage[name] = "f64" if offset_dominated(t) else "f32"
        return torch.float64 if st == "f64" else torch.float32

    def decide_storage(self, name, st):
        """Record the storage ("f64"/"f32") of a field not decided yet."""
        if self.dtype != torch.float64 and self._storage.get(name) is None:
            self._storage[name] = st

    def _pointwise_alg_lhs(self):
        """Fields written by pointwise ALG lines (their storage is decided at
        their first evaluation from its inputs, not from an IC value)."""
        from .fuse import pure_pointwise
        out = set()
        for eq in self.equations:
            if eq.kind == "ALG" and eq.LHS and eq.rank == 1:
                pw = pure_pointwise(eq.sRHS)
                if pw is not None and pw is not False:
                    out.add(eq.LHS[0])
        return out

    def settle_storage(self, name):
        """Decide (once) and apply the storage of a named field from its current value."""
        var = self.variables[name]
        d = var.data
        if self.dtype == torch.float64 or not self._is_field(d):
            return
        want = self._storage_dtype(name, d)
        if d.dtype != want:
            var.data = d.to(want)

    def fp64_state_of(self, name):
        """The FP64 state of a conserved variable of an fp32 run whose field is
        its FP32 rounding (None otherwise)."""
        if self.dtype == torch.float64 or name not in self._state64 or self._wide(name):
            return None
        return self._state64[name] if self._state_current(name) else None

    def _wide(self, name):
        return self.dtype != torch.float64 and self._storage.get(name) == "f64"

    def emptyScalar(self, val=0.0):
        return self.PyMPI.emptyScalar() + val

    # -- time integration (pyranda.py:376-427, 758-810) --------------------------
    def _storage_mode(self):
        """Context of an evaluation: unnamed pointwise results are stored by
        value while this run stores its fields by value."""
        from . import fuse
        sim = self

        class _Mode:
            def __enter__(self):
                self.old = dict(fuse.STORAGE)
                fuse.STORAGE["value_storage"] = sim.value_storage
                fuse.STORAGE["wide_if_wide_input"] = getattr(sim, "wide_if_wide_input", False)

            def __exit__(self, *exc):
                fuse.STORAGE.update(self.old)

        return _Mode()

    def updateFlux(self, defer=False):
        """Right-hand sides of the PDEs. defer: a right-hand side whose last
        step is one fused pointwise kernel comes back as a fuse.DeferredF,
        which the RK update evaluates in its own kernel."""
        flux = {}
        with self._storage_mode():
            for eqo in self.equations:
                if eqo.kind == "PDE":
                    rhs = eqo.RHS
                    flux[eqo.LHS[0]] = rhs(self, defer=True) if defer and hasattr(rhs, "root") else rhs(self)
        return flux

    def _alg_plan(self):
        from . import fuse
        from .pyrandaEq import FUSE
        eqs = [eq for eq in self.equations if eq.kind == "ALG"]
        key = tuple((id(eq), eq.active) for eq in eqs)
        if getattr(self, "_plan_key", None) != key:
            active = [eq for eq in eqs if eq.active]
            self._plan = fuse.plan_alg(active) if FUSE else [("eq", eq) for eq in active]
            self._plan_key = key
        return self._plan

    def _eval_alg(self, eq, verbose):
        if verbose:
            self.iprint(eq.eqstr)
        U = eq.LHS[0] if eq.LHS and len(eq.LHS) == 1 else None
        if U is not None and U in self._state64 and U in self._self_filtered():
            # the conservative filter of an FP32 run acts on the FP64 state
            # (FP64 kernels) and the field is its FP32 rounding (or, for a
            # field stored in FP64, the state itself)
            if not self._state_current(U):
                self._state64[U] = self.variables[U].data.to(torch.float64)
            rounded = getattr(self.PyMPI.fil, "filter_rounded", None)
            if self._wide(U):
                out = out32 = self.filter(self._state64[U])
            elif rounded is not None and self.dtype == torch.float32:
                # the last sweep stores the rounding as well
                out, out32 = rounded(self._state64[U])
            else:
                out = self.filter(self._state64[U])
                out32 = out.to(self.dtype)
            self._state64[U] = out
            self.variables[U].data = out32
            self._mark_state(U)
            return
        if self.named_storage and eq.rank == 1 and self._storage.get(eq.LHS[0]) == "f32":
            # a derivative of an FP64 field stored in FP32: the FP64 sweep
            # stores only the rounding (the value settle_storage would keep)
            got = self._derivative_to32(eq)
            if got is not None:
                self.variables[eq.LHS[0]].data = got
                return
        root = getattr(eq.RHS, "root", None)
        if self.named_storage and root is not None and eq.rank == 1 and root.out_dtype is None:
            st = self.storage_of(eq.LHS[0])
            if st is not None:
                root.out_dtype = st
            else:
                root.named_all64 = True
        rhs = eq.RHS(self)
        if not eq.LHS:
            return
        if eq.rank == 1:
            self.variables[eq.LHS[0]].data = rhs
            if self.named_storage:
                if (self.wide_if_wide_input and root is not None and root.named_all64 and root.last_all64
                        and torch.is_tensor(rhs) and rhs.dtype == torch.float64):
                    self.decide_storage(eq.LHS[0], "f64")
                self.settle_storage(eq.LHS[0])
        else:
            for ii in range(len(rhs)):
                self.variables[eq.LHS[ii]].data = rhs[ii]
                if self.named_storage:
                    self.settle_storage(eq.LHS[ii])

    _D1_LINE = re.compile(r'^\s*self\.dd([xyz])\(self\.variables\["(\w+)"\]\.data\)\s*$')

    def _derivative_to32(self, eq):
        """FP32 result of a line `:a: = ddx(:b:)` (ddy, ddz) whose argument is
        an FP64 field, from one FP64 sweep that stores only its rounding;
        None where the line is not of that form."""
        m = getattr(eq, "_d1_to32", False)
        if m is False:
            m = eq._d1_to32 = self._D1_LINE.match(eq.sRHS or "")
        if m is None:
            return None
        d = "xyz".index(m.group(1))
        if (self.nx, self.ny, self.nz)[d] <= 1 or m.group(2) not in self.variables:
            return None
        v = self.variables[m.group(2)].data
        if not self._is_field(v) or v.dtype != torch.float64:
            return None
        from . import _lib
        return self.PyMPI.ctx.apply_to32(_lib.OP_D1, d, v)

    def updateVars(self, verbose=False):
        # pointwise ALG lines run as fused multi-output kernels where possible
        with self._storage_mode():
            for item in self._alg_plan():
                if item[0] == "eq":
                    self._eval_alg(item[1], verbose)
                elif verbose or not item[1](self):
                    for eq in item[2]:
                        self._eval_alg(eq, verbose)

    def rk4(self, time, dt):
        """5-stage 2N-storage RK4 (pyranda.py:758-810).

        Per stage and conserved variable: PHI = dt*F + A*PHI; U = U + B*PHI,
        as one fused kernel (pyr_rk_update) when F is a field. In fp32 mode
        PHI and an FP64 shadow of U carry the accumulation.
        """
        fp32 = self.dtype != torch.float64
        mpi = self.PyMPI
        PHI = {}
        # conserved variables whose only rewrite is the conservative filter keep
        # a persistent FP64 state (the filter acts on it, see _eval_alg); other
        # conserved variables that an ALG line overwrites after every stage
        # would have their FP64 shadow re-derived from the FP32 field each
        # stage, which is what the update does without one
        sf = self._self_filtered() if fp32 else {}
        # of those, the ones whose filter is the next reader of the field: the
        # update need not store their rounding
        skip = self._filtered_first(sf) if sf else set()
        rewritten = self._rewritten_conserved() if fp32 else set()
        for U in self.conserved:
            # the first stage has A = 0 and only writes PHI
            PHI[U] = torch.empty((self.nz, self.ny, self.nx), dtype=torch.float64,
                                 device=mpi.device).permute(2, 1, 0)
            if fp32 and U in sf:
                if not self._state_current(U):
                    self._state64[U] = self.variables[U].data.to(torch.float64)
                    if self._wide(U):
                        self.variables[U].data = self._state64[U]
                    self._mark_state(U)
                self._shadow[U] = self._state64[U]
            elif fp32:
                # a field stored in FP64 needs no FP64 shadow
                self._shadow[U] = None if (U in rewritten or self._wide(U)) else \
                    self.variables[U].data.to(torch.float64)
        time_i = time
        self.deltat = dt
        for ii in range(5):
            FLUX = self.updateFlux(defer=getattr(self, "fuse_rk", False))
            for U in self.conserved:
                self.variables[U].data = self._rk_update(U, FLUX[U], PHI, dt, RK_A[ii], RK_B[ii], fp32,
                                                         store=U not in skip)
                if U in sf:
                    # the update advanced the FP64 state in place and wrote its rounding
                    self._state64[U] = self._shadow[U]
                    self._mark_state(U)
            time = time_i + RK_ETA[ii] * dt
            self.time = time
            if fp32:
                before = {U: self.variables[U].data for U in self.conserved}
            self.updateVars()
            if fp32:
                for U in self.conserved:
                    if U in sf:
                        self._shadow[U] = self._state64[U]
                    elif (U not in rewritten and not self._wide(U)
                          and self.variables[U].data is not before[U]):
                        self._shadow[U] = self.variables[U].data.to(torch.float64)
        self.cycle += 1
        return time

    def _mark_state(self, U):
        d = self.variables[U].data
        self._state_src[U] = (d, getattr(d, "_version", None))

    def _state_current(self, U):
        """The FP64 state of U still matches its field: the field is the
        tensor derived from it and nothing wrote into it since."""
        src = self._state_src.get(U)
        d = self.variables[U].data
        return (U in self._state64 and src is not None and torch.is_tensor(d) and src[0] is d
                and d._version == src[1])

    def _self_filtered(self):
        """Conserved variables whose only ALG writer is the conservative
        filter `:U: = fbar(:U:)`, in an EOM without LHS-less ALG lines (BC
        calls and the like may change fields in place): their FP32 runs keep
        an FP64 state through the filter."""
        if not getattr(self, "fp64_state", True):
            return {}
        key = tuple((id(eq), eq.active) for eq in self.equations)
        if getattr(self, "_sf_key", None) == key:
            return self._sf
        writers = {}
        bare = False
        for eq in self.equations:
            if not eq.active or eq.kind != "ALG":
                continue
            if not eq.LHS:
                bare = True
                continue
            for U in eq.LHS:
                writers.setdefault(U, []).append(eq)
        out = {}
        if not bare:
            for U in self.conserved:
                eqs = writers.get(U, [])
                if len(eqs) == 1 and eqs[0].rank == 1 and eqs[0].sRHS.replace(" ", "").replace("'", '"') == \
                        'self.filter(self.variables["%s"].data)' % U:
                    out[U] = eqs[0]
        self._sf, self._sf_key = out, key
        return out

    def _filtered_first(self, sf):
        """Self-filtered variables whose filter line is among the ALG lines
        that run first (before any other line): after the RK update the
        filter of the FP64 state is the next reader of the field."""
        line = {id(eq): U for U, eq in sf.items()}
        out = set()
        for item in self._alg_plan():
            for eq in ([item[1]] if item[0] == "eq" else list(item[2])):
                U = line.get(id(eq))
                if U is None:
                    return out
                out.add(U)
        return out

    def _rewritten_conserved(self):
        out = set()
        for eq in self.equations:
            if eq.active and eq.kind == "ALG" and eq.LHS:
                out.update(U for U in eq.LHS if U in self.conserved)
        return out

    def _rk_update(self, U, F, PHI, dt, A, B, fp32, store=True):
        """One stage update of U; returns the new field. store=False (fp32
        with an FP64 state): only the state advances and the field returned
        is the current one, which the caller replaces before reading it."""
        from . import _lib
        from .fuse import DeferredF
        from .parcop import _contig_equiv
        u = self.variables[U].data
        nn = tuple(self.mesh.shape)
        floats = (torch.float32, torch.float64)
        # an FP64 field of an fp32 run whose FP64 state is the field itself
        state_is_field = fp32 and torch.is_tensor(u) and u.dtype == torch.float64 and self._shadow.get(U) is u
        if isinstance(F, DeferredF):
            got = self._rk_update_deferred(U, F, PHI, dt, A, B, fp32, store, u, state_is_field)
            if got is not None:
                return got
            F = F.materialize()
        ok = (torch.is_tensor(F) and tuple(F.shape) == nn and _contig_equiv(F)
              and F.dtype in floats and torch.is_tensor(u) and tuple(u.shape) == nn
              and _contig_equiv(u) and u.dtype in floats)
        if not ok:
            # scalar fluxes (e.g. ODE right-hand sides) and odd layouts: torch ops
            PHI[U] = dt * F + A * (PHI[U] if A != 0.0 else torch.zeros_like(PHI[U]))
            if state_is_field:
                self._shadow[U] = u + B * PHI[U]
                return self._shadow[U]
            if fp32 and u.dtype != torch.float64:
                base = self._shadow[U] if self._shadow[U] is not None else u.to(torch.float64)
                self._shadow[U] = base + B * PHI[U]
                return self._shadow[U].to(self.dtype)
            return u + B * PHI[U]
        if state_is_field:
            # the kernel advances the state in place; nothing else to write
            shadow, out, uin, optr = u.data_ptr(), u, None, None
        else:
            shadow = (self._shadow[U].data_ptr() if fp32 and u.dtype != torch.float64
                      and self._shadow[U] is not None else None)
            uin = u.data_ptr()
            if shadow is not None and not store:
                out, optr = u, None
            else:
                out = torch.empty((self.nz, self.ny, self.nx), dtype=u.dtype, device=u.device).permute(2, 1, 0)
                optr = out.data_ptr()
        ctx = self.PyMPI.ctx
        ctx._stream()
        f64, u64 = F.dtype == torch.float64, u.dtype == torch.float64
        prec = {(True, True): _lib.F64, (False, False): _lib.R32,
                (False, True): _lib.RK_F32_U64, (True, False): _lib.RK_F64_U32}[(f64, u64)]
        if torch.is_tensor(dt):
            # dt in device memory (device_scalars): read by the kernel
            dt64 = dt.to(device=u.device, dtype=torch.float64).reshape(())
            _lib.check(ctx._L.pyr_rk_update_dt(ctx._h, prec, self.npts, F.data_ptr(), PHI[U].data_ptr(),
                                               uin, optr, shadow, dt64.data_ptr(),
                                               float(A), float(B)))
        else:
            _lib.check(ctx._L.pyr_rk_update(ctx._h, prec, self.npts, F.data_ptr(), PHI[U].data_ptr(),
                                            uin, optr, shadow, float(dt), float(A),
                                            float(B)))
        return out

    def _rk_update_deferred(self, U, F, PHI, dt, A, B, fp32, store, u, state_is_field):
        """The stage update with the right-hand side's last pointwise kernel
        evaluated inside it (one kernel, F never stored); the same branches
        and arithmetic as the pyr_rk_update path. None where it does not
        apply (the caller materializes F)."""
        from .parcop import _contig_equiv
        nn = tuple(self.mesh.shape)
        if not (torch.is_tensor(u) and tuple(u.shape) == nn and _contig_equiv(u)
                and u.dtype in (torch.float32, torch.float64)):
            return None
        if torch.is_tensor(dt) and dt.numel() != 1:
            return None
        if torch.is_tensor(dt):
            dt = dt.to(device=u.device, dtype=torch.float64).reshape(())
        ts = "f64" if u.dtype == torch.float64 else "f32"
        if state_is_field:
            uin, out, u64, ret = None, None, u, u
        else:
            shadow = (self._shadow[U] if fp32 and u.dtype != torch.float64 and self._shadow[U] is not None
                      else None)
            uin, u64 = (None if shadow is not None else u), shadow
            if shadow is not None and not store:
                out, ret = None, u
            else:
                out = torch.empty((self.nz, self.ny, self.nx), dtype=u.dtype, device=u.device).permute(2, 1, 0)
                ret = out
        self.PyMPI.ctx._stream()
        if not F.rk_update(PHI[U], uin, out, u64, dt, A, B, ts):
            return None
        return ret

    # -- restart and viz files (pyranda.py:431-588) -------------------------------
    def _value_of(self, name):
        """FP64 value of a variable as a restart stores it: the FP64 state of
        a conserved variable of an fp32 run when it is current, else the
        field (upcast)."""
        if name in self._state64 and self._state_current(name):
            return self._state64[name]
        return self.variables[name].data

    def writeRestart(self, ivars=None, suffix=None, wVars=[]):
        """Write the simulation state to <rootname>/restart<suffix> (suffix
        default: _<cycle:07d>) as pyranda.py:475-588 does: serial.dat and
        proc-00000.bin with every variable (sorted by name) in FP64. ivars:
        local values to keep (strings, numbers, lists, dicts, booleans, None;
        numpy arrays go to <name>.npy)."""
        import os
        import numpy as np
        if not wVars:
            wVars = list(self.variables.keys())
        if not suffix:
            suffix = "_" + str(self.cycle).zfill(_io.RESTART_INT)
        self.PyIO.ensure_root()
        dump = os.path.join(self.PyIO.rootname, "restart" + suffix)
        os.makedirs(dump, exist_ok=True)
        serial = {}
        if ivars:
            serial["local_vars"] = {"numpyArrays": {}}
            for vv, val in ivars.items():
                if isinstance(val, np.ndarray):
                    fn = os.path.join(dump, "%s" % vv)
                    serial["local_vars"]["numpyArrays"][vv] = fn
                    np.save(fn, val, allow_pickle=False)
                elif val is None or isinstance(val, (str, int, float, bool, list, dict, np.floating)):
                    serial["local_vars"][vv] = val
        mesh = dict(self.meshOptions)
        if mesh.get("function"):
            mesh["function"] = None
        if "comm" in mesh:
            mesh["comm"] = "user specified"
        n = (self.nx, self.ny, self.nz)
        serial.update({
            "mesh": mesh, "EOM": self.eom, "ICs": self.ics, "decomp": [1, 1, 1],
            "procMap": {"0-g1": np.array([0, 0, 0], dtype=np.int32),
                        "0-gn": np.array([n[0] - 1, n[1] - 1, n[2] - 1], dtype=np.int32)},
            "packages": list(self.packagesRestart), "time": float(self.time),
            "deltat": float(self.deltat) if not torch.is_tensor(self.deltat) else float(self.deltat.item()),
            "cycle": int(self.cycle), "vizDumpHistory": list(self.vizDumpHistory), "sMap": dict(self.sMap),
        })
        names = sorted(wVars)
        serial["vars"] = {v: i for i, v in enumerate(names)}
        # pyranda_amd's own keys (other readers ignore them): the storage of an
        # fp32 run's fields
        serial["storage"] = dict(self._storage)
        _io.write_serial(os.path.join(dump, "serial.dat"), serial)
        cols = [_io._host_field(self, self._value_of(v), n) for v in names]
        _io.write_proc_bin(os.path.join(dump, "proc-%s.bin" % str(0).zfill(5)), cols, n)
        return dump

    def write(self, wVars=[]):
        """Viz dump of the conserved variables (or wVars): the legacy binary VTK
        file and pyranda.visit of pyranda.py:431-468 / pyrandaIO.py:115-217."""
        import os
        if not wVars:
            wVars = list(self.conserved)
        self.PyIO.ensure_root()
        dname = "vis" + str(self.cycle).zfill(_io.VIS_INT)
        os.makedirs(os.path.join(self.PyIO.rootname, dname), exist_ok=True)
        fname = os.path.join(self.PyIO.rootname, dname,
                             "proc-%s.%s.vtk" % (str(0).zfill(_io.PROC_INT), str(self.cycle).zfill(_io.VIS_INT)))
        n = (self.nx, self.ny, self.nz)
        coords = [_io._host_field(self, c.data, n) for c in self.mesh.coords]
        fields = [_io._host_field(self, self.variables[v].data, n) for v in wVars]
        fields = [f if not isinstance(f, float) else __import__("numpy").full(n, f) for f in fields]
        t = float(self.time.item()) if torch.is_tensor(self.time) else float(self.time)
        _io.write_vtk_binary(fname, coords, fields, list(wVars), int(self.cycle), t)
        self.vizDumpHistory.append([self.cycle, t])
        _io.write_visit(os.path.join(self.PyIO.rootname, "pyranda.visit"), self.vizDumpHistory)
        return fname

    def step(self, nsteps=1):
        for _ in range(nsteps):
            self.rk4(self.time, self.deltat)

    def euler(self, time, dt):
        FLUX = self.updateFlux()
        for U in self.conserved:
            self.variables[U].data = self.variables[U].data + FLUX[U] * dt
        time = time + dt
        self.time = time
        self.updateVars()
        return time

    # -- operators (pyranda.py:607-736) -------------------------------------------
    def ddx(self, val):
        if self.nx <= 1:
            return 0.0
        return self.PyMPI.der.ddx(val)

    def ddy(self, val):
        if self.ny <= 1:
            return 0.0
        return self.PyMPI.der.ddy(val)

    def ddz(self, val):
        if self.nz <= 1:
            return 0.0
        return self.PyMPI.der.ddz(val)

    def dd4x(self, val): return self.PyMPI.der.dd4x(val)
    def dd4y(self, val): return self.PyMPI.der.dd4y(val)
    def dd4z(self, val): return self.PyMPI.der.dd4z(val)
    def dd8x(self, val): return self.PyMPI.der.dd8x(val)
    def dd8y(self, val): return self.PyMPI.der.dd8y(val)
    def dd8z(self, val): return self.PyMPI.der.dd8z(val)

    def div(self, f1, f2=None, f3=None):
        z = self.zero
        if f2 is None and f3 is None:
            if self.nx > 1:
                return self.PyMPI.der.div(f1, z, z)
            if self.ny > 1:
                return self.PyMPI.der.div(z, f1, z)
            if self.nz > 1:
                return self.PyMPI.der.div(z, z, f1)
            return 0.0
        if f3 is None:
            if self.nx > 1 and self.ny > 1:
                return self.PyMPI.der.div(f1, f2, z)
            if self.nz > 1 and self.ny > 1:
                return self.PyMPI.der.div(z, f1, f2)
            if self.nz > 1 and self.nx > 1:
                return self.PyMPI.der.div(f1, z, f2)
            if self.nx > 1:
                return self.PyMPI.der.div(f1, z, z)
            if self.ny > 1:
                return self.PyMPI.der.div(z, f1, z)
            if self.nz > 1:
                return self.PyMPI.der.div(z, z, f1)
            return 0.0
        return self.PyMPI.der.div(f1, f2, f3)

    def divT(self, *args):
        return self.PyMPI.der.divT(*args)

    def grad(self, val):
        return self.PyMPI.der.grad(val)

    def laplacian(self, val):
        return self.PyMPI.der.laplacian(val)

    def ring(self, val):
        return self.PyMPI.der.ring(val)

    def ringV(self, vx, vy, vz):
        return self.PyMPI.der.ringV(vx, vy, vz)

    def filter(self, val):
        return self.PyMPI.fil.filter(val)

    def gfilterx(self, val):
        return self.PyMPI.gfil.filterDir(val, 1)

    def gfiltery(self, val):
        return self.PyMPI.gfil.filterDir(val, 2)

    def gfilterz(self, val):
        return self.PyMPI.gfil.filterDir(val, 3)

    def gfilter(self, val):
        return self.PyMPI.gfil.filter(val)

    # -- string map (pyranda.py:812-869) -------------------------------------------
    def get_sMap(self):
        sMap = {}
        sMap["div("] = "self.div("
        sMap["divT("] = "self.divT("
        sMap["ddx("] = "self.ddx("
        sMap["ddy("] = "self.ddy("
        sMap["ddz("] = "self.ddz("
        sMap["fbar("] = "self.filter("
        sMap["gbar("] = "self.gfilter("
        sMap["gbarx("] = "self.gfilterx("
        sMap["gbary("] = "self.gfiltery("
        sMap["gbarz("] = "self.gfilterz("
        sMap["grad("] = "self.grad("
        sMap["simtime"] = "self.time"
        sMap["deltat"] = "self.deltat"
        sMap["lap("] = "self.laplacian("
        sMap["ring("] = "self.ring("
        sMap["ringV("] = "self.ringV("
        sMap["dd4x("] = "self.dd4x("
        sMap["dd4y("] = "self.dd4y("
        sMap["dd4z("] = "self.dd4z("
        sMap["dd8x("] = "self.dd8x("
        sMap["dd8y("] = "self.dd8y("
        sMap["dd8z("] = "self.dd8z("
        sMap["sum("] = "self.PyMPI.sum3D("
        sMap["mean("] = "1.0/float(self.npts) * self.PyMPI.sum3D("
        sMap["mean3D("] = "self.PyMPI.mean3D("
        sMap["sign("] = "numpy.sign("
        sMap["dot("] = "numpy.dot("
        sMap["abs("] = "numpy.abs("
        sMap["sqrt("] = "numpy.sqrt("
        sMap["sin("] = "numpy.sin("
        sMap["cos("] = "numpy.cos("
        sMap["tanh("] = "numpy.tanh("
        sMap["exp("] = "numpy.exp("
        sMap["where("] = "numpy.where("
        sMap["max("] = "self.PyMPI.max3D("
        sMap["min("] = "self.PyMPI.min3D("
        sMap["3d("] = "self.emptyScalar("
        sMap["pi"] = "numpy.pi"
        sMap["meshVar("] = "self.PyMPI.getVar("
        sMap["gridLen"] = "self.mesh.GridLen"
        sMap["random3D()"] = "numpy.random.random( self.mesh.coords[0].data.shape) "
        sMap["erf("] = "scipy.special.erf("
        sMap["meshx"] = "self.mesh.coords[0].data"
        sMap["meshy"] = "self.mesh.coords[1].data"
        sMap["meshz"] = "self.mesh.coords[2].data"
        sMap["meshi"] = "self.mesh.indices[0].data"
        sMap["meshj"] = "self.mesh.indices[1].data"
        sMap["meshk"] = "self.mesh.indices[2].data"
        self.sMap = sMap


def pyrandaRestart(rootname, suffix=None, comm=None, **sim_options):
    """A pyrandaSim rebuilt from a restart directory (pyranda.py:979-1080):
    the latest restart* under rootname, or restart_<suffix>. Returns
    [pysim, local_vars]. sim_options (precision, device, ...) go to
    pyrandaSim; the fields are read back as they were written (FP64) and
    take the run's storage."""
    import glob
    import os
    import numpy as np
    if not suffix:
        dumps = sorted(glob.glob(os.path.join(rootname, "restart*")))
        if not dumps:
            raise FileNotFoundError("no restart directory under %s" % rootname)
        dump = dumps[-1]
    else:
        dump = os.path.join(rootname, "restart_%s" % suffix)
    if not os.path.isdir(dump):
        raise FileNotFoundError("Error: Cant read restart file %s" % dump)
    serial = _io.read_serial(os.path.join(dump, "serial.dat"))
    local_vars = {}
    if "local_vars" in serial:
        lv = serial.pop("local_vars")
        arrays = lv.pop("numpyArrays", {})
        local_vars.update(lv)
        for na, fn in arrays.items():
            local_vars[na] = np.load(fn + ".npy", allow_pickle=False)
    mesh = dict(serial["mesh"])
    if comm is not None:
        mesh["comm"] = comm
    elif mesh.get("comm") == "user specified":
        mesh.pop("comm")
    pysim = pyrandaSim(rootname, mesh, **sim_options)
    from .pyrandaBC import pyrandaBC as _BC
    from .pyrandaTimestep import pyrandaTimestep as _TS
    known = {"pyrandaTimestep": _TS, "pyrandaBC": _BC}
    for pack in serial["packages"]:
        name = pack.split(".")[1] if "." in pack else pack
        if name not in known:
            raise ValueError("restart names package %s, which pyranda_amd does not have" % pack)
        pysim.addPackage(known[name](pysim))
    pysim.sMap = dict(serial["sMap"])
    if "storage" in serial and pysim.dtype != torch.float64:
        pysim._storage.update(serial["storage"])
    pysim.EOM(serial["EOM"])
    pysim.setIC(serial["ICs"], addOnly=True)
    pysim.time = serial["time"]
    pysim.deltat = serial["deltat"]
    pysim.cycle = serial["cycle"]
    pysim.vizDumpHistory = list(serial["vizDumpHistory"])
    names = sorted(serial["vars"], key=lambda v: serial["vars"][v])
    n = (pysim.nx, pysim.ny, pysim.nz)
    decomp = [int(v) for v in serial["decomp"]]
    if decomp != [1, 1, 1]:
        raise ValueError("restart written by %s ranks; pyranda_amd reads single-rank restarts" % decomp)
    mm = _io.open_proc_bin(os.path.join(dump, "proc-%s.bin" % str(0).zfill(5)), n, len(names))
    dev = pysim.PyMPI.device
    for v in names:
        col = serial["vars"][v]
        if v not in pysim.variables:
            pysim.addVar(v, kind="conserved")
            pysim.variables[v].__allocate__(pysim.PyMPI)
        var = pysim.variables[v]
        if v.startswith("mesh"):
            continue  # the mesh comes from the mesh options (Cartesian)
        if torch.is_tensor(var.data) and var.data.dim() == 3:
            host = np.ascontiguousarray(np.asarray(mm[:, :, :, col]).transpose(2, 1, 0))
            var.data = torch.from_numpy(host).to(dev).permute(2, 1, 0)
        else:
            var.data = float(mm[0, 0, 0, col])
    del mm
    if pysim.dtype != torch.float64:
        # the FP64 state of self-filtered conserved variables is the stored
        # value; fields take their storage
        pysim._state64 = {}
        for U in pysim._self_filtered():
            d = pysim.variables[U].data
            if torch.is_tensor(d) and tuple(d.shape) == tuple(pysim.mesh.shape):
                pysim._state64[U] = d
        pysim._cast_fields()
        for U in pysim._state64:
            pysim._mark_state(U)
    for pack in pysim.packages.values():
        pack.__init__(pysim)
    pysim.iprint("Successful read of restart file: %s " % dump)
    return [pysim, local_vars]
