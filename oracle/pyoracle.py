"""ctypes front end for the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a plain-C FP64 restatement of pyranda's compact operators
(oracle/compact_oracle.c, each function citing its reference file:line).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker; the product package never does.

Arrays follow the reference layout: float64, Fortran order, shape (nx,ny,nz).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpyranda_ref.so")

D1, D2, D8, SF, GF = 0, 1, 2, 3, 4
BC_PERI, BC_NONE, BC_SYMM = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle not built: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.ora_create.restype = ctypes.c_void_p
    lib.ora_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.POINTER(ctypes.c_int)]
    lib.ora_destroy.argtypes = [ctypes.c_void_p]
    lib.ora_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.ora_apply_sym.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
    for name in ("ora_sfilter", "ora_gfilter", "ora_ring", "ora_laplacian"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.ora_gfilterdir.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    lib.ora_ringv.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
    lib.ora_divergence.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
    lib.ora_grads.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
    lib.ora_set_threads.argtypes = [ctypes.c_int]
    lib.ora_bpentLUD1.argtypes = [_dp, ctypes.c_int]
    lib.ora_ppentLUD1.argtypes = [_dp, ctypes.c_int]
    lib.ora_bpentLUS1.argtypes = [_dp, _dp, ctypes.c_int]
    lib.ora_ppentLUS1.argtypes = [_dp, _dp, ctypes.c_int]
    lib.ora_line_solve.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
    lib.ora_assemble.argtypes = [ctypes.c_int] * 5 + [_dp, _dp]
    lib.ora_assemble.restype = ctypes.c_int
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


class OraWeight(ctypes.Structure):
    _fields_ = [("nol", ctypes.c_int), ("nor", ctypes.c_int), ("ncl", ctypes.c_int),
                ("ncr", ctypes.c_int), ("nbc1", ctypes.c_int), ("nbc2", ctypes.c_int),
                ("null_option", ctypes.c_int), ("implicit_op", ctypes.c_int),
                ("ali", ctypes.c_double * 5), ("ari", ctypes.c_double * 9),
                ("alb1", ctypes.c_double * 60), ("alb2", ctypes.c_double * 60),
                ("arb1", ctypes.c_double * 108), ("arb2", ctypes.c_double * 108)]


def weights(op):
    """Weight tables of one scheme: dict of numpy arrays, variant axis -1..1."""
    w = OraWeight()
    lib().ora_get_weight.argtypes = [ctypes.c_int, ctypes.POINTER(OraWeight)]
    lib().ora_get_weight(op, ctypes.byref(w))
    return {
        "nol": w.nol, "nor": w.nor, "ncl": w.ncl, "ncr": w.ncr,
        "null_option": w.null_option, "implicit_op": w.implicit_op,
        "ali": np.array(w.ali[:]), "ari": np.array(w.ari[:]),
        "alb1": np.array(w.alb1[:]).reshape(3, 4, 5), "alb2": np.array(w.alb2[:]).reshape(3, 4, 5),
        "arb1": np.array(w.arb1[:]).reshape(3, 4, 9), "arb2": np.array(w.arb2[:]).reshape(3, 4, 9),
    }


def assemble(op, n, periodic, bc1=0, bc2=0):
    """Band + LU exactly as setup_compact_op1 (np=1). Returns (al(n,nal), ar(nr,n))."""
    al = np.zeros((9, n))  # column-major (n,9) == row-major (9,n)
    ar = np.zeros((n, 9))  # column-major (9,n) == row-major (n,9)
    nal = lib().ora_assemble(op, n, int(periodic), bc1, bc2, _ptr(al), _ptr(ar))
    return al[:nal].T.copy(), ar.T.copy()


def bc_codes(periodic, symmetric=None):
    codes = []
    for d in range(3):
        for side in range(2):
            if periodic[d]:
                codes.append(BC_PERI)
            elif symmetric is not None and symmetric[d][side]:
                codes.append(BC_SYMM)
            else:
                codes.append(BC_NONE)
    return codes


class Oracle:
    """Operator suite on one Cartesian block, as parcop.setup builds it."""

    def __init__(self, nn, dd, periodic, symmetric=None, threads=1):
        self.nn = tuple(int(n) for n in nn)
        self.dd = tuple(float(d) for d in dd)
        bc6 = (ctypes.c_int * 6)(*bc_codes(periodic, symmetric))
        self._lib = lib()
        self._lib.ora_set_threads(int(threads))
        self._h = self._lib.ora_create(*self.nn, *self.dd, bc6)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.ora_destroy(self._h)
            self._h = None

    def _in(self, v):
        v = np.asfortranarray(np.asarray(v, dtype=np.float64))
        if v.shape != self.nn:
            v = np.asfortranarray(np.broadcast_to(v, self.nn).astype(np.float64))
        return v

    def _out(self):
        return np.zeros(self.nn, dtype=np.float64, order="F")

    def apply(self, op, direction, v):
        v = self._in(v)
        out = self._out()
        self._lib.ora_apply(self._h, op, direction, _ptr(v), _ptr(out))
        return out

    def apply_sym(self, op, direction, isym, v):
        """The operator with the SYMM-pair member selected by isym (+1: the
        default, -1: the antisymmetric one where the direction has a SYMM
        end), as d1x/d8x(v, dv, bc) select it (compact_operators.f90:300-303)."""
        v = self._in(v)
        out = self._out()
        self._lib.ora_apply_sym(self._h, op, direction, int(isym), _ptr(v), _ptr(out))
        return out

    def ddx(self, v): return self.apply(D1, 0, v)
    def ddy(self, v): return self.apply(D1, 1, v)
    def ddz(self, v): return self.apply(D1, 2, v)
    def d2(self, v, d): return self.apply(D2, d, v)
    def d8(self, v, d): return self.apply(D8, d, v)

    def _unary(self, name, v):
        v = self._in(v)
        out = self._out()
        getattr(self._lib, name)(self._h, _ptr(v), _ptr(out))
        return out

    def sfilter(self, v): return self._unary("ora_sfilter", v)
    def gfilter(self, v): return self._unary("ora_gfilter", v)
    def ring(self, v): return self._unary("ora_ring", v)
    def laplacian(self, v): return self._unary("ora_laplacian", v)

    def gfilterdir(self, v, dir1):
        v = self._in(v)
        out = self._out()
        self._lib.ora_gfilterdir(self._h, _ptr(v), _ptr(out), int(dir1))
        return out

    def ringv(self, vx, vy, vz):
        vx, vy, vz = self._in(vx), self._in(vy), self._in(vz)
        out = self._out()
        self._lib.ora_ringv(self._h, _ptr(vx), _ptr(vy), _ptr(vz), _ptr(out))
        return out

    def divergence(self, fx, fy, fz):
        fx, fy, fz = self._in(fx), self._in(fy), self._in(fz)
        out = self._out()
        self._lib.ora_divergence(self._h, _ptr(fx), _ptr(fy), _ptr(fz), _ptr(out))
        return out

    def grads(self, v):
        v = self._in(v)
        g = [self._out() for _ in range(3)]
        self._lib.ora_grads(self._h, _ptr(v), *[_ptr(a) for a in g])
        return g


# ---------------------------------------------------------------------------
# The reference's own weights / LU (oracle/_ref), when it was built here.
# ---------------------------------------------------------------------------
def ref_available():
    return os.path.exists(REF_PATH)


_REF = None


def ref_lib():
    global _REF
    if _REF is None:
        r = ctypes.CDLL(REF_PATH)
        r.ref_get_weight.argtypes = [ctypes.c_int] + [_dp] * 6 + [ctypes.POINTER(ctypes.c_int)]
        r.ref_bpentLUD1.argtypes = [_dp, ctypes.c_int]
        r.ref_ppentLUD1.argtypes = [_dp, ctypes.c_int]
        r.ref_bpentLUS3.argtypes = [ctypes.c_int, _dp, _dp] + [ctypes.c_int] * 4
        r.ref_ppentLUS3.argtypes = [ctypes.c_int, _dp, _dp] + [ctypes.c_int] * 4
        _REF = r
    return _REF


def ref_weights(op):
    ali = np.zeros(5); ari = np.zeros(9)
    alb1 = np.zeros((3, 4, 5)); alb2 = np.zeros((3, 4, 5))
    arb1 = np.zeros((3, 4, 9)); arb2 = np.zeros((3, 4, 9))
    dims = (ctypes.c_int * 6)()
    ref_lib().ref_get_weight(op, _ptr(ali), _ptr(ari), _ptr(alb1), _ptr(alb2), _ptr(arb1), _ptr(arb2), dims)
    nol, nor, ncl, ncr, nullopt, impl = list(dims)
    return {"nol": nol, "nor": nor, "ncl": ncl, "ncr": ncr, "null_option": nullopt,
            "implicit_op": impl, "ali": ali, "ari": ari, "alb1": alb1, "alb2": alb2,
            "arb1": arb1, "arb2": arb2}

