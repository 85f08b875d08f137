"""Fusion of the pointwise algebra of EOM equations into generated HIP kernels.

An equation's right-hand side, after pyranda's string translation, is a
Python expression over fields (`self.variables["u"].data`), operator calls
(`self.ddx(...)`, `self.filter(...)`), reductions and numbers. numpy would
evaluate every + - * / ** sqrt ... with a fresh temporary; here every maximal
pointwise subtree becomes one elementwise kernel that reads each distinct
input once and writes its result once (compiled at first use with hiprtc,
cached per input-type signature). Operator calls and reductions stay calls;
their pointwise arguments are fused kernels of their own.

Evaluation order and rounding follow Python/numpy: operands are evaluated in
FP64 registers with the same expression tree, no FMA contraction, and the
result is stored in FP64 if any input field is FP64, else in FP32 (FP32
runs that store fields by value decide per result instead, see
offset_dominated). Subtrees
whose inputs are all scalars are evaluated in Python; inputs that are not
whole fields (slices, vectors) make that subtree evaluate with torch ops.
"""
import ast
import ctypes
import hashlib
import os
import re

import torch

from . import _lib
from .devnumpy import numpy as _dnp

_UNARY = {"sqrt": "sqrt", "abs": "fabs", "absolute": "fabs", "exp": "exp", "log": "log",
          "sin": "sin", "cos": "cos", "tan": "tan", "tanh": "tanh", "arctan": "atan",
          "floor": "floor"}
_BINARY = {"minimum": "pyr_min", "maximum": "pyr_max", "arctan2": "atan2"}
_CMP = {ast.Lt: "<", ast.Gt: ">", ast.LtE: "<=", ast.GtE: ">=", ast.Eq: "==", ast.NotEq: "!="}

# sibling subtrees of one right-hand side with plain-read leaves share one
# multi-output kernel (FusedGroup); PYRANDA_AMD_FUSE_GROUP=0 keeps one kernel each
GROUP_SIBLINGS = os.environ.get("PYRANDA_AMD_FUSE_GROUP", "1") != "0"

PRELUDE = r"""
struct PyrPwArgs { long long n; void* p[24]; double s[16]; int vec; };
__device__ __forceinline__ double pyr_min(double a, double b) { return (a < b || a != a) ? a : b; }
__device__ __forceinline__ double pyr_max(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ __forceinline__ double pyr_sign(double a) { return a != a ? a : (double)((a > 0.0) - (a < 0.0)); }
"""


# consecutive elements per thread and iteration of the elementwise loops
VEC = 4
_PTR_DECL = re.compile(r"^\s*((?:const )?\w+\* __restrict__ \w+) = (\((?:const )?\w+\*\))(a\.p\[\d+\]);$")
_LOOP_HEAD = ("  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;"
              " i += (long long)gridDim.x * blockDim.x) {")


def _vectorized(lines):
    """A generated elementwise kernel (its lines: prelude, signature,
    declarations, the grid-stride loop _LOOP_HEAD ... "  }", "}") with its
    loop split in two. With every pointer 16-byte aligned (a.vec, set at
    launch) each thread takes groups of VEC consecutive elements (group q:
    elements VEC q ... VEC q + VEC - 1) through a device function whose
    pointers are __restrict__ parameters: the compiler may then gather each
    stream's loads and stores of a group into 16-byte accesses (restrict on
    local pointers does not tell it that streams do not overlap); the
    elements past the last whole group, or every element without a.vec, go
    through the plain grid-stride loop."""
    h = lines.index(_LOOP_HEAD)
    assert lines[-2] == "  }" and lines[-1] == "}"
    prelude, sig, decls, body = lines[0], lines[1], lines[2:h], lines[h + 1:-2]
    kname = re.search(r"void (?:__launch_bounds__\(\d+\) )?(\w+)\(", sig).group(1)
    ptrs = [_PTR_DECL.match(d) for d in decls]
    rest = [d for d, m in zip(decls, ptrs) if m is None]
    ptrs = [m for m in ptrs if m is not None]
    fn = kname + "_q"
    params = ", ".join(["const PyrPwArgs& a"] + [m.group(1) for m in ptrs] + ["long long q_"])
    args = ", ".join(["a"] + ["%s__builtin_assume_aligned(%s, 16)" % (m.group(2), m.group(3)) for m in ptrs]
                     + ["q_"])
    ind = lambda ls, k: [" " * k + l.strip() for l in ls]
    F = ["static __device__ __forceinline__ void %s(%s) {" % (fn, params)]
    F += ind(rest, 2)
    F += ["#pragma unroll", "  for (int j_ = 0; j_ < %d; ++j_) {" % VEC,
          "    const long long i = q_ * %d + j_;" % VEC]
    F += ind(body, 4)
    F += ["  }", "}"]
    K = [sig] + decls
    K += ["  const long long nthr_ = (long long)gridDim.x * blockDim.x;",
          "  const long long t_ = (long long)blockIdx.x * blockDim.x + threadIdx.x;",
          "  long long i_ = t_;",
          "  if (a.vec) {",
          "    const long long nq_ = a.n / %d;" % VEC,
          "    for (long long q_ = t_; q_ < nq_; q_ += nthr_) %s(%s);" % (fn, args),
          "    i_ = nq_ * %d + t_;" % VEC,
          "  }",
          "  for (long long i = i_; i < a.n; i += nthr_) {"]
    K += ind(body, 4)
    K += ["  }", "}"]
    return [prelude] + F + K


def _is_numpy_call(node):
    return (isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute)
            and isinstance(node.func.value, ast.Name) and node.func.value.id == "numpy"
            and not node.keywords)


def _pointwise(node):
    """Is this node itself a pointwise operation (not a leaf)?"""
    if isinstance(node, ast.BinOp):
        return isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow))
    if isinstance(node, ast.UnaryOp):
        return isinstance(node.op, (ast.USub, ast.UAdd))
    if isinstance(node, ast.Compare):
        return len(node.ops) == 1 and type(node.ops[0]) in _CMP
    if _is_numpy_call(node):
        name = node.func.attr
        if name in _UNARY or name == "sign":
            return len(node.args) == 1
        if name in _BINARY:
            return len(node.args) == 2
        if name == "where":
            return len(node.args) == 3
    return False


def _const(node):
    return isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) \
        and not isinstance(node.value, bool)


class _Gen:
    """C++ expression of one pointwise subtree; leaves become x<i>."""

    def __init__(self):
        self.leaves = []       # ast nodes (non-pointwise children)
        self.leaf_src = {}     # unparsed source -> leaf index

    def leaf(self, node):
        src = ast.unparse(node)
        if src not in self.leaf_src:
            self.leaf_src[src] = len(self.leaves)
            self.leaves.append(node)
        return "x%d" % self.leaf_src[src]

    def expr(self, node):
        if _const(node):
            return "(%s)" % float(node.value).hex() if float(node.value) == float(node.value) else "NAN"
        if not _pointwise(node):
            return self.leaf(node)
        if isinstance(node, ast.BinOp):
            a, b = self.expr(node.left), self.expr(node.right)
            if isinstance(node.op, ast.Add):
                return "(%s + %s)" % (a, b)
            if isinstance(node.op, ast.Sub):
                return "(%s - %s)" % (a, b)
            if isinstance(node.op, ast.Mult):
                return "(%s * %s)" % (a, b)
            if isinstance(node.op, ast.Div):
                return "(%s / %s)" % (a, b)
            # numpy's fast paths for a scalar exponent (square, sqrt, reciprocal)
            if _const(node.right):
                e = float(node.right.value)
                if e == 2.0:
                    return "(%s * %s)" % (a, a)
                if e == 0.5:
                    return "sqrt(%s)" % a
                if e == -1.0:
                    return "(1.0 / %s)" % a
                if e == 1.0:
                    return a
            return "pow(%s, %s)" % (a, b)
        if isinstance(node, ast.UnaryOp):
            a = self.expr(node.operand)
            return "(-%s)" % a if isinstance(node.op, ast.USub) else a
        if isinstance(node, ast.Compare):
            return "(%s %s %s)" % (self.expr(node.left), _CMP[type(node.ops[0])],
                                   self.expr(node.comparators[0]))
        name = node.func.attr
        args = [self.expr(a) for a in node.args]
        if name in _UNARY:
            return "%s(%s)" % (_UNARY[name], args[0])
        if name == "sign":
            return "pyr_sign(%s)" % args[0]
        if name in _BINARY:
            return "%s(%s, %s)" % (_BINARY[name], args[0], args[1])
        if name == "where":
            return "((%s) ? %s : %s)" % (args[0], args[1], args[2])
        raise AssertionError(name)


# --- storage of FP32 runs' fields ---------------------------------------------
# FP32 storage rounds a value to 6e-8 of its magnitude. A field whose values
# all sit close to one large value (the pressure of a low-Mach flow, p ~ 71
# against variations ~ 0.4; the total energy; stresses carrying -p) loses
# most of its variation to that rounding, and a derivative of it turns the
# rounding into noise of ~1e-4 relative. In FP32 runs such fields are stored
# in FP64: a field is offset-dominated when its largest magnitude exceeds
# WIDE_RATIO times its range of values (max - min).
WIDE_RATIO = 8.0

# value_storage: True while an FP32 run with storage by value evaluates its
# equations. With wide_if_wide_input (the "auto" policy) a pointwise kernel's
# unnamed result (e.g. the flux argument of a derivative) is FP64 when an
# input field is FP64 and FP32 otherwise, a decision by types alone that a
# run resumed from a restart repeats; without it (the "offset" policy) the
# result is stored in FP64 when its first evaluation (in FP64) is
# offset-dominated, else in FP32 (kept per kernel and input signature).
# With value_storage False a result is FP64 when an input field is.
STORAGE = {"value_storage": False, "wide_if_wide_input": False}


def offset_dominated(t):
    """Is the field offset-dominated (max|t| > WIDE_RATIO * (max t - min t))?
    A constant nonzero field is; a zero or non-finite field is not."""
    lo, hi = torch.aminmax(t.detach())
    lo, hi = float(lo), float(hi)
    mag = max(abs(lo), abs(hi))
    if not (mag == mag) or mag == float("inf"):
        return False
    return mag > 0.0 and mag > WIDE_RATIO * (hi - lo)


def _device_scalar(v):
    return (torch.is_tensor(v) and v.dim() == 0 and v.is_cuda
            and v.dtype in (torch.float64, torch.float32))


class GroupShape(tuple):
    """Shape (n, 1, 1) of the fields of a group of simulations stored in one
    allocation, carrying `mid` (the member index of every point, int32 on the
    device) and `nmem`: a 1-D tensor of `nmem` values is then one scalar per
    member (dt, say), read per point as value[mid[i]]."""

    def __new__(cls, n, mid, nmem):
        t = super().__new__(cls, (n, 1, 1))
        t.mid = mid
        t.nmem = nmem
        return t


def _member_scalar(v, shape):
    return (isinstance(shape, GroupShape) and torch.is_tensor(v) and v.dim() == 1 and v.is_cuda
            and v.shape[0] == shape.nmem and v.dtype in (torch.float64, torch.float32))


def _expand_members(vals, shape):
    """Per-member values expanded to fields (for the torch path)."""
    if not isinstance(shape, GroupShape):
        return vals
    return [v[shape.mid.long()].reshape(shape) if _member_scalar(v, shape) else v for v in vals]


def _field_ok(t, shape):
    if not torch.is_tensor(t) or t.dtype not in (torch.float64, torch.float32) or not t.is_cuda:
        return False
    if tuple(t.shape) != shape:
        return False
    nx, ny, nz = shape
    want = (1, nx, nx * ny)
    return all(s == w or n == 1 for s, w, n in zip(t.stride(), want, shape))


class FusedSubtree:
    """One maximal pointwise subtree: a callable over its leaf values."""

    def __init__(self, node, src_tag):
        self.node = node
        g = _Gen()
        self.body = g.expr(node)
        self.leaves = g.leaves
        self.nleaves = len(g.leaves)
        self.is_bool = isinstance(node, ast.Compare)
        # torch/Python evaluation of the same subtree (for scalar-only or
        # non-field inputs)
        args = ", ".join("x%d" % i for i in range(self.nleaves))
        expr = ast.unparse(_rename_leaves(node, g))
        self._py = eval(compile("lambda %s: %s" % (args, expr), "<fused %s>" % src_tag, "eval"),
                        {"numpy": _dnp})
        self._kernels = {}
        self.tag = src_tag
        # storage of the result: out_dtype ('f64'/'f32', set by the owner of a
        # named result) wins; else decided per input signature when
        # STORAGE["value_storage"] is on; else by the input types
        self.out_dtype = None
        self._decided = {}
        # the result is a named field of a run storing by value: kept in FP64
        # when every field input is FP64 (last_all64 tells the owner)
        self.named_all64 = False
        self.last_all64 = False
        # set by FusedRHS(defer=True) around the call of a root subtree
        self._defer = False

    def _rk_source(self, sig, out_dt, variant):
        """The subtree's value F (rounded to its storage type) fused into the
        RK4 stage update of pyranda.py:800-804: PHI = dt F + A PHI, then the
        state U advanced by B PHI; the arithmetic of k_rk_update
        (pyr_core.hip), explicit FMAs included. variant = (ts, shadow,
        store, dtp, first): the field type, an FP64 state advanced in place,
        the field stored, dt read from device memory, A == 0 (the first
        stage: PHI is written, not read)."""
        ts, shadow, store, dtp, first = variant
        name = "pyr_rk_" + hashlib.sha1((self.body + repr(sig) + out_dt + repr(variant)).encode()).hexdigest()[:16]
        lines = [PRELUDE, 'extern "C" __global__ void %s(PyrPwArgs a) {' % name]
        ip = isc = 0
        for i, s in enumerate(sig):
            if s == "S":
                lines.append("  const double x%d = a.s[%d];" % (i, isc))
                isc += 1
            elif s in ("D64", "D32"):
                ct = "double" if s == "D64" else "float"
                lines.append("  const double x%d = (double)*(const %s*)a.p[%d];" % (i, ct, ip))
                ip += 1
            else:
                ct = "double" if s == "T64" else "float"
                lines.append("  const %s* __restrict__ p%d = (const %s*)a.p[%d];" % (ct, i, ct, ip))
                ip += 1
        tt = "double" if ts == "f64" else "float"
        ot = "double" if out_dt == "f64" else "float"
        lines.append("  double* __restrict__ phi_ = (double*)a.p[%d];" % ip)
        ip += 1
        if not shadow:
            lines.append("  const %s* __restrict__ uin = (const %s*)a.p[%d];" % (tt, tt, ip))
            ip += 1
        if store:
            lines.append("  %s* __restrict__ uout = (%s*)a.p[%d];" % (tt, tt, ip))
            ip += 1
        if shadow:
            lines.append("  double* __restrict__ u64 = (double*)a.p[%d];" % ip)
            ip += 1
        if dtp:
            lines.append("  const double dt = *(const double*)a.p[%d];" % ip)
        else:
            lines.append("  const double dt = a.s[%d];" % isc)
            isc += 1
        lines.append("  const double A = a.s[%d], B = a.s[%d];" % (isc, isc + 1))
        lines.append("  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;"
                     " i += (long long)gridDim.x * blockDim.x) {")
        for i, s in enumerate(sig):
            if s in ("T64", "T32"):
                lines.append("    const double x%d = (double)p%d[i];" % (i, i))
        lines.append("    const double f = (double)(%s)(%s);" % (ot, self.body))
        # the first stage (A == 0) neither reads PHI nor lets a stale value in
        lines.append("    const double tmp1 = %s;" % ("0.0" if first else "A * phi_[i]"))
        lines.append("    const double phi = fma(dt, f, tmp1);")
        lines.append("    phi_[i] = phi;")
        if shadow:
            lines.append("    const double u = fma(B, phi, u64[i]);")
            lines.append("    u64[i] = u;")
            if store:
                lines.append("    uout[i] = (%s)u;" % tt)
        else:
            lines.append("    uout[i] = (%s)fma(B, phi, (double)uin[i]);" % tt)
        lines.append("  }")
        lines.append("}")
        return name, "\n".join(_vectorized(lines))

    def rk_kernel(self, sig, out_dt, variant):
        key = ("rk", sig, out_dt, variant)
        fn = self._kernels.get(key)
        if fn is None:
            name, src = self._rk_source(sig, out_dt, variant)
            h = ctypes.c_void_p()
            L = _lib.lib()
            if L.pyr_pw_compile(src.encode(), name.encode(), ctypes.byref(h)) != 0:
                raise _lib.PyrandaAMDError(L.pyr_pw_last_error().decode())
            fn = h
            self._kernels[key] = fn
        return fn

    def _source(self, sig, out_dt):
        name = "pyr_pw_" + hashlib.sha1((self.body + repr(sig) + out_dt).encode()).hexdigest()[:16]
        lines = [PRELUDE, 'extern "C" __global__ void %s(PyrPwArgs a) {' % name]
        ip = isc = 0
        decl = []
        for i, s in enumerate(sig):
            if s == "S":
                decl.append("  const double x%d = a.s[%d];" % (i, isc))
                isc += 1
            elif s in ("D64", "D32"):
                ct = "double" if s == "D64" else "float"
                decl.append("  const double x%d = (double)*(const %s*)a.p[%d];" % (i, ct, ip))
                ip += 1
            else:
                ct = "double" if s in ("T64", "G64") else "float"
                decl.append("  const %s* __restrict__ p%d = (const %s*)a.p[%d];" % (ct, i, ct, ip))
                ip += 1
        if any(s[0] == "G" for s in sig):
            decl.append("  const int* __restrict__ pmid = (const int*)a.p[%d];" % ip)
            ip += 1
        ot = "double" if out_dt == "f64" else "float"
        lines += decl
        lines.append("  %s* __restrict__ out = (%s*)a.p[%d];" % (ot, ot, ip))
        lines.append("  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;"
                     " i += (long long)gridDim.x * blockDim.x) {")
        for i, s in enumerate(sig):
            if s in ("T64", "T32"):
                lines.append("    const double x%d = (double)p%d[i];" % (i, i))
            elif s in ("G64", "G32"):
                lines.append("    const double x%d = (double)p%d[pmid[i]];" % (i, i))
        lines.append("    out[i] = (%s)(%s);" % (ot, self.body))
        lines.append("  }")
        lines.append("}")
        return name, "\n".join(_vectorized(lines))

    def kernel(self, sig, out_dt):
        key = (sig, out_dt)
        fn = self._kernels.get(key)
        if fn is None:
            name, src = self._source(sig, out_dt)
            h = ctypes.c_void_p()
            L = _lib.lib()
            if L.pyr_pw_compile(src.encode(), name.encode(), ctypes.byref(h)) != 0:
                raise _lib.PyrandaAMDError(L.pyr_pw_last_error().decode())
            fn = h
            self._kernels[key] = fn
        return fn

    def __call__(self, shape, *vals):
        sig = []
        fields = []     # pointer operands: fields and device scalars, in order
        scalars = []
        any64 = False
        nfield = 0
        dev = None
        for v in vals:
            if torch.is_tensor(v) and v.dim() == 0 and not _device_scalar(v):
                v = v.item()
            if isinstance(v, (int, float)) or (hasattr(v, "dtype") and not torch.is_tensor(v)
                                                and getattr(v, "shape", None) == ()):
                sig.append("S")
                scalars.append(float(v))
            elif _device_scalar(v):
                # a 0-d device value (a reduction result kept on the GPU):
                # read once by the kernel, no host round trip; like a Python
                # scalar it does not decide the result type
                sig.append("D64" if v.dtype == torch.float64 else "D32")
                fields.append(v)
            elif _member_scalar(v, shape):
                # one value per member of a group (read per point by member)
                sig.append("G64" if v.dtype == torch.float64 else "G32")
                fields.append(v)
            elif _field_ok(v, shape):
                sig.append("T64" if v.dtype == torch.float64 else "T32")
                any64 |= v.dtype == torch.float64
                fields.append(v)
                nfield += 1
                dev = v.device
            else:
                return self._py(*_expand_members(vals, shape))
        if not nfield or self.is_bool:
            return self._py(*_expand_members(vals, shape) if nfield else vals)
        if any(k[0] == "G" for k in sig):
            fields.append(shape.mid)
        if len(fields) + 1 > _lib.PW_MAX_PTRS or len(scalars) > _lib.PW_MAX_SCALARS:
            return self._py(*_expand_members(vals, shape))
        sig = tuple(sig)
        decide = False
        all64 = all(k != "T32" for k in sig if k[0] == "T")
        self.last_all64 = all64
        if self.out_dtype is not None:
            out_dt = self.out_dtype
        elif STORAGE["value_storage"] and self.named_all64 and STORAGE["wide_if_wide_input"] and all64:
            out_dt = "f64"
        elif STORAGE["value_storage"] and not self.named_all64 and STORAGE["wide_if_wide_input"]:
            # an unnamed result: by its input types alone (no decision by
            # value, which a restarted run could not repeat)
            out_dt = "f64" if any64 else "f32"
        elif STORAGE["value_storage"]:
            out_dt = self._decided.get(sig)
            if out_dt is None:
                out_dt, decide = "f64", True
        else:
            out_dt = "f64" if any64 else "f32"
        if self._defer and not decide and not any(k[0] == "G" for k in sig):
            # the caller (an RK stage) evaluates the result inside its update
            return DeferredF(self, sig, out_dt, fields, scalars, tuple(shape), dev)
        fn = self.kernel(sig, out_dt)
        nx, ny, nz = shape
        out = torch.empty((nz, ny, nx), dtype=torch.float64 if out_dt == "f64" else torch.float32,
                          device=dev).permute(2, 1, 0)
        ptrs = (ctypes.c_void_p * (len(fields) + 1))(*[f.data_ptr() for f in fields], out.data_ptr())
        sc = (ctypes.c_double * max(1, len(scalars)))(*scalars)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check_pw(_lib.lib().pyr_pw_launch(fn, stream, nx * ny * nz, len(fields) + 1, ptrs,
                                               len(scalars), sc))
        if decide:
            # first evaluation in FP64: keep FP64 storage if offset-dominated
            self._decided[sig] = "f64" if offset_dominated(out) else "f32"
            if self._decided[sig] == "f32":
                out = out.to(torch.float32)
        return out


def _simple_load(node):
    """A leaf that is a plain read (names, attributes, constant subscripts; no
    call): evaluating it early has no effect and cannot fail differently."""
    if isinstance(node, ast.Name):
        return True
    if isinstance(node, ast.Attribute):
        return _simple_load(node.value)
    if isinstance(node, ast.Subscript):
        return isinstance(node.slice, ast.Constant) and _simple_load(node.value)
    return _const(node)


class FusedGroup:
    """Sibling pointwise subtrees of one right-hand side whose leaves are
    plain reads (the arguments of `-ddx(A) - ddy(B) - ddz(C)`): one kernel
    reads the union of their inputs once and writes one output per subtree.
    Each output is computed with its own subtree's expression and stored in
    its own type, so the results equal the separate kernels' bit for bit;
    where a separate evaluation would not take the kernel path the group
    falls back to the separate subtrees."""

    def __init__(self, subs, src_tag):
        self.subs = subs
        g = _Gen()
        self.bodies = [g.expr(sub.node) for sub in subs]
        self.leaves = g.leaves
        # each subtree's leaves as indices into the union
        self.maps = [[g.leaf_src[ast.unparse(l)] for l in sub.leaves] for sub in subs]
        self._kernels = {}
        self.tag = src_tag
        self._decided = {}  # per input signature: output types by value (FusedSubtree)

    def _source(self, sig, out_dts):
        key = "|".join(self.bodies) + repr(sig) + repr(out_dts)
        name = "pyr_pwg_" + hashlib.sha1(key.encode()).hexdigest()[:16]
        lines = [PRELUDE, 'extern "C" __global__ void %s(PyrPwArgs a) {' % name]
        ip = isc = 0
        for i, s in enumerate(sig):
            if s == "S":
                lines.append("  const double x%d = a.s[%d];" % (i, isc))
                isc += 1
            elif s in ("D64", "D32"):
                ct = "double" if s == "D64" else "float"
                lines.append("  const double x%d = (double)*(const %s*)a.p[%d];" % (i, ct, ip))
                ip += 1
            else:
                ct = "double" if s in ("T64", "G64") else "float"
                lines.append("  const %s* __restrict__ p%d = (const %s*)a.p[%d];" % (ct, i, ct, ip))
                ip += 1
        if any(s[0] == "G" for s in sig):
            lines.append("  const int* __restrict__ pmid = (const int*)a.p[%d];" % ip)
            ip += 1
        for k, dt in enumerate(out_dts):
            ot = "double" if dt == "f64" else "float"
            lines.append("  %s* __restrict__ out%d = (%s*)a.p[%d];" % (ot, k, ot, ip + k))
        lines.append("  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;"
                     " i += (long long)gridDim.x * blockDim.x) {")
        for i, s in enumerate(sig):
            if s in ("T64", "T32"):
                lines.append("    const double x%d = (double)p%d[i];" % (i, i))
            elif s in ("G64", "G32"):
                lines.append("    const double x%d = (double)p%d[pmid[i]];" % (i, i))
        for k, (body, dt) in enumerate(zip(self.bodies, out_dts)):
            lines.append("    out%d[i] = (%s)(%s);" % (k, "double" if dt == "f64" else "float", body))
        lines.append("  }")
        lines.append("}")
        return name, "\n".join(_vectorized(lines))

    def kernel(self, sig, out_dts):
        key = (sig, out_dts)
        fn = self._kernels.get(key)
        if fn is None:
            name, src = self._source(sig, out_dts)
            h = ctypes.c_void_p()
            L = _lib.lib()
            if L.pyr_pw_compile(src.encode(), name.encode(), ctypes.byref(h)) != 0:
                raise _lib.PyrandaAMDError(L.pyr_pw_last_error().decode())
            fn = h
            self._kernels[key] = fn
        return fn

    def _separate(self, shape, vals):
        return tuple(sub(shape, *[vals[j] for j in m]) for sub, m in zip(self.subs, self.maps))

    def __call__(self, shape, *vals):
        sig, fields, scalars = [], [], []
        is64 = []
        dev = None
        for v in vals:
            if torch.is_tensor(v) and v.dim() == 0 and not _device_scalar(v):
                v = v.item()
            if isinstance(v, (int, float)) or (hasattr(v, "dtype") and not torch.is_tensor(v)
                                                and getattr(v, "shape", None) == ()):
                sig.append("S")
                scalars.append(float(v))
            elif _device_scalar(v):
                sig.append("D64" if v.dtype == torch.float64 else "D32")
                fields.append(v)
            elif _member_scalar(v, shape):
                sig.append("G64" if v.dtype == torch.float64 else "G32")
                fields.append(v)
            elif _field_ok(v, shape):
                sig.append("T64" if v.dtype == torch.float64 else "T32")
                fields.append(v)
                dev = v.device
            else:
                return self._separate(shape, vals)
            is64.append(sig[-1] == "T64")
        out_dts = []
        for m in self.maps:
            # every subtree needs a field operand to take the kernel path
            if not any(sig[j][0] == "T" for j in m):
                return self._separate(shape, vals)
            out_dts.append("f64" if any(is64[j] for j in m) else "f32")
        decide = False
        if STORAGE["value_storage"] and STORAGE["wide_if_wide_input"]:
            # unnamed results by their input types alone (out_dts as set above)
            pass
        elif STORAGE["value_storage"]:
            known = self._decided.get(tuple(sig))
            if known is None:
                out_dts, decide = ["f64"] * len(self.maps), True
            else:
                out_dts = list(known)
        if any(k[0] == "G" for k in sig):
            fields.append(shape.mid)
        nout = len(self.subs)
        if len(fields) + nout > _lib.PW_MAX_PTRS or len(scalars) > _lib.PW_MAX_SCALARS:
            return self._separate(shape, vals)
        fn = self.kernel(tuple(sig), tuple(out_dts))
        nx, ny, nz = shape
        outs = [torch.empty((nz, ny, nx), dtype=torch.float64 if dt == "f64" else torch.float32,
                            device=dev).permute(2, 1, 0) for dt in out_dts]
        ptrs = (ctypes.c_void_p * (len(fields) + nout))(*[f.data_ptr() for f in fields],
                                                         *[o.data_ptr() for o in outs])
        sc = (ctypes.c_double * max(1, len(scalars)))(*scalars)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check_pw(_lib.lib().pyr_pw_launch(fn, stream, nx * ny * nz, len(fields) + nout, ptrs,
                                               len(scalars), sc))
        if decide:
            dts = tuple("f64" if offset_dominated(o) else "f32" for o in outs)
            self._decided[tuple(sig)] = dts
            outs = [o if dt == "f64" else o.to(torch.float32) for o, dt in zip(outs, dts)]
        return tuple(outs)


def _group_siblings(body, subtrees, tag):
    """Rewrite `body` so the fused subtrees whose leaves are all plain reads,
    outside conditional parts of the expression, are evaluated by one
    FusedGroup call ahead of the rest: `(lambda _g: body')(_pwg(_shape, ...))`
    with `_g[k]` in their places. Returns (new body, group or None)."""
    picks = []

    def walk(n, cond):
        if (isinstance(n, ast.Call) and isinstance(n.func, ast.Subscript)
                and isinstance(n.func.value, ast.Name) and n.func.value.id == "_pw"):
            sub = subtrees[n.func.slice.value]
            if not cond and not sub.is_bool and all(_simple_load(l) for l in sub.leaves):
                picks.append(n)
                return
        c = cond or isinstance(n, (ast.IfExp, ast.BoolOp, ast.Lambda, ast.ListComp, ast.GeneratorExp,
                                   ast.SetComp, ast.DictComp))
        for ch in ast.iter_child_nodes(n):
            walk(ch, c)

    walk(body, False)
    if len(picks) < 2:
        return body, None
    group = FusedGroup([subtrees[n.func.slice.value] for n in picks], tag)
    slot = {id(n): k for k, n in enumerate(picks)}

    class Put(ast.NodeTransformer):
        def visit_Call(self, n):
            if id(n) in slot:
                return ast.copy_location(ast.Subscript(value=ast.Name(id="_g", ctx=ast.Load()),
                                                       slice=ast.Constant(value=slot[id(n)]),
                                                       ctx=ast.Load()), n)
            return self.generic_visit(n)

    new = Put().visit(body)
    inner = ast.Lambda(args=ast.arguments(posonlyargs=[], args=[ast.arg(arg="_g")], kwonlyargs=[],
                                          kw_defaults=[], defaults=[]), body=new)
    call = ast.Call(func=ast.Name(id="_pwg", ctx=ast.Load()),
                    args=[ast.Name(id="_shape", ctx=ast.Load())] + [_copy(l) for l in group.leaves],
                    keywords=[])
    return ast.Call(func=inner, args=[call], keywords=[]), group


def _rename_leaves(node, gen):
    """Copy of the subtree with leaves replaced by their x<i> names."""

    class R(ast.NodeTransformer):
        def generic_visit(self, n):
            if n is not node and not _pointwise(n) and not _const(n) and isinstance(n, ast.expr):
                src = ast.unparse(n)
                if src in gen.leaf_src:
                    return ast.copy_location(ast.Name(id="x%d" % gen.leaf_src[src], ctx=ast.Load()), n)
            return super().generic_visit(n)

    import copy
    new = R().visit(copy.deepcopy(node))
    return ast.fix_missing_locations(new)


class _Fuser(ast.NodeTransformer):
    def __init__(self, tag):
        self.subtrees = []
        self.tag = tag

    def visit(self, node):
        if isinstance(node, ast.expr) and _pointwise(node) and _has_op(node):
            # fuse its non-pointwise children first (operator calls' arguments)
            sub = FusedSubtree(node, self.tag)
            idx = len(self.subtrees)
            self.subtrees.append(sub)
            args = [self.visit(_copy(l)) for l in sub.leaves]
            call = ast.Call(func=ast.Subscript(value=ast.Name(id="_pw", ctx=ast.Load()),
                                               slice=ast.Constant(value=idx), ctx=ast.Load()),
                            args=[ast.Name(id="_shape", ctx=ast.Load())] + args, keywords=[])
            return ast.copy_location(call, node)
        return self.generic_visit(node)


def _has_op(node):
    return _pointwise(node)


def _copy(n):
    import copy
    return copy.deepcopy(n)


class _DropAbsOfRing(ast.NodeTransformer):
    """numpy.abs(self.ring(...)) -> self.ring(...): ring and ringV are maxima of
    |d8|-scaled values (operators.f90:615-699), never negative, so the abs is
    the identity and would cost one more pass over the field."""

    def visit_Call(self, node):
        self.generic_visit(node)
        if (_is_numpy_call(node) and node.func.attr in ("abs", "absolute") and len(node.args) == 1
                and isinstance(node.args[0], ast.Call) and isinstance(node.args[0].func, ast.Attribute)
                and isinstance(node.args[0].func.value, ast.Name) and node.args[0].func.value.id == "self"
                and node.args[0].func.attr in ("ring", "ringV")):
            return node.args[0]
        return node


class DeferredF:
    """The value of a PDE right-hand side whose last pointwise kernel has not
    run: its operands are evaluated (operator results included) and the
    RK4 stage evaluates the kernel's expression inside its update
    (`rk_update`), saving the write and read of F. `materialize()` runs the
    kernel as FusedSubtree would."""

    def __init__(self, sub, sig, out_dt, fields, scalars, shape, dev):
        self.sub, self.sig, self.out_dt = sub, sig, out_dt
        self.fields, self.scalars, self.shape, self.device = fields, scalars, shape, dev
        self.dtype = torch.float64 if out_dt == "f64" else torch.float32

    def materialize(self):
        fn = self.sub.kernel(self.sig, self.out_dt)
        nx, ny, nz = self.shape
        out = torch.empty((nz, ny, nx), dtype=self.dtype, device=self.device).permute(2, 1, 0)
        ptrs = (ctypes.c_void_p * (len(self.fields) + 1))(*[f.data_ptr() for f in self.fields], out.data_ptr())
        sc = (ctypes.c_double * max(1, len(self.scalars)))(*self.scalars)
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check_pw(_lib.lib().pyr_pw_launch(fn, stream, nx * ny * nz, len(self.fields) + 1, ptrs,
                                               len(self.scalars), sc))
        return out

    def rk_update(self, phi, uin, uout, u64, dt, A, B, ts):
        """PHI, the FP64 state u64 (if given) and the field uout (if given)
        advanced by one stage with F evaluated in registers; dt a float or a
        0-d FP64 device tensor. Returns False when the operands do not fit one
        launch (the caller materializes F instead)."""
        shadow, store = u64 is not None, uout is not None
        dtp = torch.is_tensor(dt)
        ptrs = list(self.fields) + [phi] + ([] if shadow else [uin]) + ([uout] if store else []) \
            + ([u64] if shadow else []) + ([dt] if dtp else [])
        scal = list(self.scalars) + ([] if dtp else [float(dt)]) + [float(A), float(B)]
        if len(ptrs) > _lib.PW_MAX_PTRS or len(scal) > _lib.PW_MAX_SCALARS:
            return False
        fn = self.sub.rk_kernel(self.sig, self.out_dt, (ts, shadow, store, dtp, float(A) == 0.0))
        nx, ny, nz = self.shape
        pp = (ctypes.c_void_p * len(ptrs))(*[t.data_ptr() for t in ptrs])
        sc = (ctypes.c_double * len(scal))(*scal)
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check_pw(_lib.lib().pyr_pw_launch(fn, stream, nx * ny * nz, len(ptrs), pp, len(scal), sc))
        return True


class FusedRHS:
    """Callable replacing `lambda self: <rhs>` with fused pointwise kernels."""

    def __init__(self, src, globals_):
        tree = _DropAbsOfRing().visit(ast.parse(src, mode="eval"))
        f = _Fuser(src[:40])
        body = f.visit(tree.body)
        self.subtrees = f.subtrees
        self.src = src
        # the subtree that computes the whole right-hand side, if one does
        # (its result is the equation's value: its storage can be set)
        self.root = None
        if (isinstance(body, ast.Call) and isinstance(body.func, ast.Subscript)
                and isinstance(body.func.value, ast.Name) and body.func.value.id == "_pw"):
            self.root = self.subtrees[body.func.slice.value]
        body, self.group = _group_siblings(body, self.subtrees, src[:40]) if GROUP_SIBLINGS else (body, None)
        lam = ast.Expression(body=ast.Lambda(
            args=ast.arguments(posonlyargs=[], args=[ast.arg(arg="self"), ast.arg(arg="_shape"),
                                                     ast.arg(arg="_pw"), ast.arg(arg="_pwg")],
                               kwonlyargs=[], kw_defaults=[], defaults=[]),
            body=body))
        ast.fix_missing_locations(lam)
        self._fn = eval(compile(lam, "<eom %s>" % src[:40], "eval"), globals_)

    def __call__(self, sim, defer=False):
        """defer: a right-hand side whose last step is one pointwise kernel
        returns a DeferredF (its operands evaluated, the kernel not run)."""
        shape = sim.mesh.shape
        shape = shape if isinstance(shape, GroupShape) else tuple(shape)
        if not defer or self.root is None:
            return self._fn(sim, shape, self.subtrees, self.group)
        self.root._defer = True
        try:
            return self._fn(sim, shape, self.subtrees, self.group)
        finally:
            self.root._defer = False


_RED = {
    # identity, combine(acc, v); max/min propagate NaN as numpy's max/min do
    "max": ("-__builtin_huge_val()", "((v > acc || v != v || acc != acc) ? (acc != acc ? acc : v) : acc)"),
    "min": ("__builtin_huge_val()", "((v < acc || v != v || acc != acc) ? (acc != acc ? acc : v) : acc)"),
}
_RED_BLOCKS = 1024
_RED_UNROLL = int(os.environ.get("PYRANDA_AMD_RED_UNROLL", "4"))


class FusedReduce:
    """max/min over the grid of one pointwise expression of fields and
    scalars, in one pass: each workgroup of PYR_PW_BLOCK threads reduces its
    grid-stride share (wave shuffles, then LDS) to one FP64 partial, and the
    host reduces the partials. `body` is C over x<i> (inputs, fields first,
    then scalars) computed in FP64."""

    def __init__(self, op, body, nfields, nscalars, tag=""):
        self.op, self.body, self.nf, self.ns, self.tag = op, body, nfields, nscalars, tag
        self._kernels = {}

    def _source(self, sig):
        ident, comb = _RED[self.op]
        name = "pyr_red_" + hashlib.sha1((self.op + self.body + repr(sig) + "u%d" % _RED_UNROLL).encode()
                                         ).hexdigest()[:16]
        ptrs = ["const %s* __restrict__ p%d" % ("double" if t == "T64" else "float", i) for i, t in enumerate(sig)]
        scal = ["  const double x%d = a.s[%d];" % (self.nf + j, j) for j in range(self.ns)]
        body = ["    const double x%d = (double)p%d[i];" % (i, i) for i in range(len(sig))]
        body += ["    const double v = (%s);" % self.body, "    acc = %s;" % comb]
        # groups of VEC elements through restrict parameters (16-byte loads
        # when every pointer is 16-byte aligned, as for the elementwise
        # kernels); max/min do not depend on the order
        F = ["static __device__ __forceinline__ double %s_q(%s) {"
             % (name, ", ".join(["const PyrPwArgs& a"] + ptrs + ["long long q_", "double acc"]))]
        F += scal
        F += ["#pragma unroll", "  for (int j_ = 0; j_ < %d; ++j_) {" % VEC,
              "    const long long i = q_ * %d + j_;" % VEC] + body + ["  }", "  return acc;", "}"]
        args = ", ".join(["a"] + ["(const %s*)__builtin_assume_aligned(a.p[%d], 16)"
                                  % ("double" if t == "T64" else "float", i) for i, t in enumerate(sig)]
                         + ["q_", "acc"])
        L = [PRELUDE] + F + ['extern "C" __global__ void __launch_bounds__(256) %s(PyrPwArgs a) {' % name]
        L += ["  %s = (%s)a.p[%d];" % (d, d.split("*")[0] + "*", i) for i, d in enumerate(ptrs)]
        L += scal
        L.append("  double* __restrict__ part = (double*)a.p[%d];" % len(sig))
        L.append("  double acc = %s;" % ident)
        L += ["  const long long nthr_ = (long long)gridDim.x * blockDim.x;",
              "  const long long t_ = (long long)blockIdx.x * blockDim.x + threadIdx.x;",
              "  long long i_ = t_;",
              "  if (a.vec) {",
              "    const long long nq_ = a.n / %d;" % VEC,
              "    for (long long q_ = t_; q_ < nq_; q_ += nthr_) acc = %s_q(%s);" % (name, args),
              "    i_ = nq_ * %d + t_;" % VEC,
              "  }"]
        # several iterations' loads in flight
        L.append("#pragma unroll %d" % _RED_UNROLL)
        L.append("  for (long long i = i_; i < a.n; i += nthr_) {")
        L += body
        L.append("  }")
        L.append("  for (int off = 32; off > 0; off >>= 1) {")
        L.append("    const double v = __shfl_xor(acc, off, 64);")
        L.append("    acc = %s;" % comb)
        L.append("  }")
        L.append("  __shared__ double wsum[4];")
        L.append("  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;")
        L.append("  __syncthreads();")
        L.append("  if (threadIdx.x == 0) {")
        L.append("    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { const double v = wsum[w]; acc = %s; }" % comb)
        L.append("    part[blockIdx.x] = acc;")
        L.append("  }")
        L.append("}")
        return name, "\n".join(L)

    def _launch(self, n, fields, scalars, blocks):
        sig = tuple("T64" if f.dtype == torch.float64 else "T32" for f in fields)
        fn = self._kernels.get(sig)
        if fn is None:
            name, src = self._source(sig)
            h = ctypes.c_void_p()
            L = _lib.lib()
            if L.pyr_pw_compile(src.encode(), name.encode(), ctypes.byref(h)) != 0:
                raise _lib.PyrandaAMDError(L.pyr_pw_last_error().decode())
            fn = self._kernels[sig] = h
        dev = fields[0].device
        part = torch.empty(blocks, dtype=torch.float64, device=dev)
        ptrs = (ctypes.c_void_p * (len(fields) + 1))(*[f.data_ptr() for f in fields], part.data_ptr())
        sc = (ctypes.c_double * max(1, len(scalars)))(*[float(x) for x in scalars])
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check_pw(_lib.lib().pyr_pw_launch_blocks(fn, stream, n, blocks, len(fields) + 1, ptrs,
                                                      len(scalars), sc))
        return part

    def __call__(self, shape, fields, scalars, device=False):
        """The reduction as a Python float, or with device=True as a 0-d FP64
        device tensor (a second one-workgroup pass over the partials; no host
        round trip, so it can sit inside a captured HIP graph)."""
        n = shape[0] * shape[1] * shape[2]
        blocks = max(1, min(_RED_BLOCKS, (n + 1023) // 1024))
        part = self._launch(n, fields, scalars, blocks)
        if device:
            return _final(self.op)._launch(blocks, [part], [], 1).reshape(())
        import numpy as _np
        vals = part.cpu().numpy()
        return float(_np.max(vals) if self.op == "max" else _np.min(vals))


class SegmentedReduce:
    """max/min of one pointwise expression over each member of a group of
    simulations stored in one allocation (member m: points seg[m] ..
    seg[m+1]), in one launch: each workgroup reduces a share of one member's
    points (workgroups blk[m] .. blk[m+1] belong to member m, found by binary
    search) to a partial; a second launch (one workgroup per member) reduces
    each member's partials. Scalars are per member: x<nfields + j> reads
    sc[m * nscalars + j]. Returns one FP64 value per member (device)."""

    def __init__(self, op, body, nfields, nscalars, tag=""):
        self.op, self.body, self.nf, self.ns, self.tag = op, body, nfields, nscalars, tag
        self._kernels = {}

    def _source(self, sig):
        ident, comb = _RED[self.op]
        name = "pyr_segred_" + hashlib.sha1((self.op + self.body + repr(sig)).encode()).hexdigest()[:16]
        nf = len(sig)
        L = [PRELUDE, 'extern "C" __global__ void __launch_bounds__(256) %s(PyrPwArgs a) {' % name]
        for i, t in enumerate(sig):
            ct = "double" if t == "T64" else "float"
            L.append("  const %s* __restrict__ p%d = (const %s*)a.p[%d];" % (ct, i, ct, i))
        L.append("  double* __restrict__ part = (double*)a.p[%d];" % nf)
        L.append("  const long long* __restrict__ seg = (const long long*)a.p[%d];" % (nf + 1))
        L.append("  const int* __restrict__ blk = (const int*)a.p[%d];" % (nf + 2))
        L.append("  const double* __restrict__ sc = (const double*)a.p[%d];" % (nf + 3))
        L.append("  const int M = (int)a.n;")
        L.append("  int lo = 0, hi = M - 1;")
        L.append("  while (lo < hi) { const int md = (lo + hi + 1) >> 1; if (blk[md] <= (int)blockIdx.x) lo = md; else hi = md - 1; }")
        L.append("  const int m = lo;")
        for j in range(self.ns):
            L.append("  const double x%d = sc[m * %d + %d];" % (self.nf + j, max(self.ns, 1), j))
        L.append("  const long long s1 = seg[2 * m + 1];")
        L.append("  const long long nb = blk[m + 1] - blk[m];")
        L.append("  double acc = %s;" % ident)
        L.append("  for (long long i = seg[2 * m] + (long long)(blockIdx.x - blk[m]) * blockDim.x + threadIdx.x;"
                 " i < s1; i += nb * blockDim.x) {")
        for i in range(nf):
            L.append("    const double x%d = (double)p%d[i];" % (i, i))
        L.append("    const double v = (%s);" % self.body)
        L.append("    acc = %s;" % comb)
        L.append("  }")
        L.append("  for (int off = 32; off > 0; off >>= 1) {")
        L.append("    const double v = __shfl_xor(acc, off, 64);")
        L.append("    acc = %s;" % comb)
        L.append("  }")
        L.append("  __shared__ double wsum[4];")
        L.append("  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;")
        L.append("  __syncthreads();")
        L.append("  if (threadIdx.x == 0) {")
        L.append("    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { const double v = wsum[w]; acc = %s; }" % comb)
        L.append("    part[blockIdx.x] = acc;")
        L.append("  }")
        L.append("}")
        return name, "\n".join(L)

    def _fn(self, key, src_fn):
        fn = self._kernels.get(key)
        if fn is None:
            name, src = src_fn()
            h = ctypes.c_void_p()
            L = _lib.lib()
            if L.pyr_pw_compile(src.encode(), name.encode(), ctypes.byref(h)) != 0:
                raise _lib.PyrandaAMDError(L.pyr_pw_last_error().decode())
            fn = self._kernels[key] = h
        return fn

    def __call__(self, seg, fields, member_scalars):
        """seg: SegmentTable of the group; member_scalars: FP64 device tensor
        (M, nscalars) (or None without scalars)."""
        sig = tuple("T64" if f.dtype == torch.float64 else "T32" for f in fields)
        fn = self._fn(sig, lambda: self._source(sig))
        dev = fields[0].device
        part = torch.empty(seg.nblocks, dtype=torch.float64, device=dev)
        sc = member_scalars if member_scalars is not None else seg.zero
        ptrs = (ctypes.c_void_p * (len(fields) + 4))(*[f.data_ptr() for f in fields], part.data_ptr(),
                                                     seg.seg.data_ptr(), seg.blk.data_ptr(), sc.data_ptr())
        one = (ctypes.c_double * 1)(0.0)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        L = _lib.lib()
        _lib.check_pw(L.pyr_pw_launch_blocks(fn, stream, seg.nmem, seg.nblocks, len(fields) + 4, ptrs, 0, one))
        # second pass: member m reduces its partials blk[m] .. blk[m+1]
        fin = _segfinal(self.op)
        out = torch.empty(seg.nmem, dtype=torch.float64, device=dev)
        ptrs2 = (ctypes.c_void_p * 4)(part.data_ptr(), seg.pseg.data_ptr(), seg.blk.data_ptr(), out.data_ptr())
        _lib.check_pw(L.pyr_pw_launch_blocks(fin, stream, seg.nmem, seg.nmem, 4, ptrs2, 0, one))
        return out


_SEGFIN = {}


def _segfinal(op):
    """Per-member reduction of the partials: workgroup m reduces part[pblk[m]
    .. pblk[m+1]) (pseg/pblk: the partial ranges, one workgroup per member)."""
    if op not in _SEGFIN:
        ident, comb = _RED[op]
        name = "pyr_segfin_" + op
        L = [PRELUDE, 'extern "C" __global__ void __launch_bounds__(256) %s(PyrPwArgs a) {' % name,
             "  const double* __restrict__ part = (const double*)a.p[0];",
             "  const long long* __restrict__ pseg = (const long long*)a.p[1];",
             "  double* __restrict__ out = (double*)a.p[3];",
             "  const int m = blockIdx.x;",
             "  double acc = %s;" % ident,
             "  for (long long i = pseg[m] + threadIdx.x; i < pseg[m + 1]; i += blockDim.x) {",
             "    const double v = part[i];",
             "    acc = %s;" % comb,
             "  }",
             "  for (int off = 32; off > 0; off >>= 1) {",
             "    const double v = __shfl_xor(acc, off, 64);",
             "    acc = %s;" % comb,
             "  }",
             "  __shared__ double wsum[4];",
             "  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;",
             "  __syncthreads();",
             "  if (threadIdx.x == 0) {",
             "    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { const double v = wsum[w]; acc = %s; }" % comb,
             "    out[m] = acc;",
             "  }",
             "}"]
        h = ctypes.c_void_p()
        Lb = _lib.lib()
        if Lb.pyr_pw_compile("\n".join(L).encode(), name.encode(), ctypes.byref(h)) != 0:
            raise _lib.PyrandaAMDError(Lb.pyr_pw_last_error().decode())
        _SEGFIN[op] = h
    return _SEGFIN[op]


class SegmentTable:
    """Member segments of a group's fields and the workgroup partition of a
    segmented reduction: member m owns points seg[2m] .. seg[2m+1] (members
    may be separated by alignment padding) and workgroups blk[m] .. blk[m+1]
    (about four points per thread, at most 64 workgroups per member)."""

    def __init__(self, sizes, offsets, device):
        import numpy as _np
        self.nmem = len(sizes)
        blocks = [max(1, min(64, (n + 1023) // 1024)) for n in sizes]
        blk = _np.concatenate([[0], _np.cumsum(blocks)]).astype(_np.int32)
        self.nblocks = int(blk[-1])
        pairs = []
        for o, n in zip(offsets, sizes):
            pairs += [int(o), int(o) + int(n)]
        self.seg = torch.tensor(pairs, dtype=torch.int64, device=device)
        self.blk = torch.tensor(blk, dtype=torch.int32, device=device)
        self.pseg = torch.tensor(blk.astype(_np.int64), dtype=torch.int64, device=device)
        self.zero = torch.zeros(1, dtype=torch.float64, device=device)


_FINALS = {}


def _final(op):
    if op not in _FINALS:
        _FINALS[op] = FusedReduce(op, "x0", 1, 0, "final")
    return _FINALS[op]


def fields_ok(shape, *vals):
    """All values are FP32/FP64 device fields of `shape` in the field layout."""
    return all(_field_ok(v, tuple(shape)) for v in vals)


# --- runs of pointwise equations -------------------------------------------
_VAR = None


def _var_name(node):
    """'X' for the AST of self.variables["X"].data, else None."""
    if (isinstance(node, ast.Attribute) and node.attr == "data" and isinstance(node.value, ast.Subscript)
            and isinstance(node.value.value, ast.Attribute) and node.value.value.attr == "variables"
            and isinstance(node.value.value.value, ast.Name) and node.value.value.value.id == "self"
            and isinstance(node.value.slice, ast.Constant) and isinstance(node.value.slice.value, str)):
        return node.value.slice.value
    return None


def pure_pointwise(src):
    """Variables read by `src` if it is pointwise algebra over variables and
    constants only (no operator call, no reduction), else None."""
    try:
        node = ast.parse(src, mode="eval").body
    except SyntaxError:
        return None
    if not _pointwise(node):
        return None
    g = _Gen()
    g.expr(node)
    names = []
    for leaf in g.leaves:
        nm = _var_name(leaf)
        if nm is None:
            return None
        names.append(nm)
    return names


class PointwiseRun:
    """Consecutive pointwise equations `lhs_i = rhs_i` as one kernel: reads
    each outside variable once, writes every lhs. A value produced inside the
    run is rounded to its field type before later lines use it, so the results
    are those of evaluating the lines one after another; in an FP32 run a
    line whose result is stored in FP64 reads earlier lines' values unrounded
    (and conserved variables through their FP64 state), as an FP64 run
    would."""

    def __init__(self, lines, tag=""):
        self.lines = lines  # [(lhs, src)]
        self.tag = tag
        self.outputs = [l for l, _ in lines]
        self.inputs = []      # outside variables in first-use order
        self.bodies = []      # per line: expression over x<i> (inputs) and l<j> (earlier lines)
        self.deps = []        # per line: (input indices, earlier-line indices)
        produced = {}
        for j, (lhs, src) in enumerate(lines):
            node = ast.parse(src, mode="eval").body
            g = _Gen()
            body = g.expr(node)
            ins, ints = set(), set()
            # leaves are x<k> in g's numbering: map them to run inputs / earlier lines
            ren = {}
            for k, leaf in enumerate(g.leaves):
                nm = _var_name(leaf)
                if nm in produced:
                    ren["x%d" % k] = "l%d" % produced[nm]
                    ints.add(produced[nm])
                else:
                    if nm not in self.inputs:
                        self.inputs.append(nm)
                    ii = self.inputs.index(nm)
                    ren["x%d" % k] = "X%d" % ii
                    ins.add(ii)
            import re as _re
            body = _re.sub(r"\bx(\d+)\b", lambda m: ren["x" + m.group(1)], body)
            self.bodies.append(body.replace("X", "x"))
            self.deps.append((ins, ints))
            produced[lhs] = j
        self._kernels = {}

    def _types(self, sig):
        """Output type of every line ('f64'/'f32'), None if a line has no field input."""
        out = []
        for ins, ints in self.deps:
            kinds = [sig[i] for i in ins] + ["T64" if out[j] == "f64" else "T32" for j in ints]
            fields = [k for k in kinds if k[0] == "T"]
            if not fields:
                return None
            out.append("f64" if "T64" in fields else "f32")
        return out

    def _source(self, sig, types):
        key = repr(self.bodies) + repr(sig) + repr(types)
        name = "pyr_run_" + hashlib.sha1(key.encode()).hexdigest()[:16]
        L = [PRELUDE, 'extern "C" __global__ void %s(PyrPwArgs a) {' % name]
        ip = isc = 0
        for i, t in enumerate(sig):
            if t == "S":
                L.append("  const double x%d = a.s[%d];" % (i, isc))
                isc += 1
            elif t in ("D64", "D32"):
                ct = "double" if t == "D64" else "float"
                L.append("  const double x%d = (double)*(const %s*)a.p[%d];" % (i, ct, ip))
                ip += 1
            else:
                ct = "double" if t == "T64" else "float"
                L.append("  const %s* __restrict__ p%d = (const %s*)a.p[%d];" % (ct, i, ct, ip))
                ip += 1
        for j, t in enumerate(types):
            ot = "double" if t == "f64" else "float"
            L.append("  %s* __restrict__ o%d = (%s*)a.p[%d];" % (ot, j, ot, ip + j))
        L.append("  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;"
                 " i += (long long)gridDim.x * blockDim.x) {")
        for i, t in enumerate(sig):
            if t in ("T64", "T32"):
                L.append("    const double x%d = (double)p%d[i];" % (i, i))
        import re as _re
        for j, t in enumerate(types):
            ot = "double" if t == "f64" else "float"
            body = self.bodies[j]
            if t == "f64":
                # an FP64 result reads earlier lines' values unrounded (u<k>);
                # FP32 results read them rounded to their field type (l<k>)
                body = _re.sub(r"\bl(\d+)\b", r"u\1", body)
            L.append("    const double u%d = (%s);" % (j, body))
            L.append("    const %s r%d = (%s)u%d;" % (ot, j, ot, j))
            L.append("    o%d[i] = r%d;" % (j, j))
            L.append("    const double l%d = (double)r%d;" % (j, j))
        L.append("  }")
        L.append("}")
        return name, "\n".join(_vectorized(L))

    def __call__(self, sim):
        """Evaluate the run on `sim`; False if an input is not a field or scalar
        (the caller then evaluates the lines one by one)."""
        shape = tuple(sim.mesh.shape)
        sig, fields, scalars = [], [], []
        dev = None
        # a run with an FP64-stored result reads conserved variables through
        # their FP64 state (pyrandaSim.fp64_state_of) rather than the FP32 field
        named = getattr(sim, "named_storage", False)
        wide_out = named and any(sim.storage_of(l) != "f32" for l in self.outputs)
        state_of = getattr(sim, "fp64_state_of", None) if wide_out else None
        in64 = []   # per input: read as an FP64 field (or scalar)
        for nm in self.inputs:
            v = sim.variables[nm].data
            if state_of is not None:
                v64 = state_of(nm)
                if v64 is not None:
                    v = v64
            in64.append(not (torch.is_tensor(v) and v.dtype == torch.float32))
            if torch.is_tensor(v) and v.dim() == 0 and not _device_scalar(v):
                v = v.item()
            if isinstance(v, (int, float)) and not isinstance(v, bool):
                sig.append("S")
                scalars.append(float(v))
            elif _device_scalar(v):
                sig.append("D64" if v.dtype == torch.float64 else "D32")
                fields.append(v)
            elif _field_ok(v, shape):
                sig.append("T64" if v.dtype == torch.float64 else "T32")
                fields.append(v)
                dev = v.device
            else:
                return False
        sig = tuple(sig)
        types = self._types(sig)
        if types is None or dev is None:
            return False
        # an FP32 run's named fields have a storage type of their own
        # (pyrandaSim.storage_of: 'f64', 'f32', or None when not decided yet:
        # computed in FP64, then decided from the values)
        storage_of = getattr(sim, "storage_of", None) if getattr(sim, "named_storage", False) else None
        undecided = []
        if storage_of is not None:
            inputs_rule = STORAGE["wide_if_wide_input"]
            for j, lhs in enumerate(self.outputs):
                st = storage_of(lhs)
                if st is None:
                    ins, ints = self.deps[j]
                    # a line computed from FP64 values only keeps FP64 storage
                    if inputs_rule and all(sig[i] != "T32" or in64[i] for i in ins) \
                            and all(types[k] == "f64" for k in ints):
                        st = "f64"
                        sim.decide_storage(lhs, "f64")
                    else:
                        undecided.append(j)
                        st = "f64"
                types[j] = st
        if len(fields) + len(types) > _lib.PW_MAX_PTRS or len(scalars) > _lib.PW_MAX_SCALARS:
            return False
        key = (sig, tuple(types))
        fn = self._kernels.get(key)
        if fn is None:
            name, src = self._source(sig, types)
            h = ctypes.c_void_p()
            L = _lib.lib()
            if L.pyr_pw_compile(src.encode(), name.encode(), ctypes.byref(h)) != 0:
                raise _lib.PyrandaAMDError(L.pyr_pw_last_error().decode())
            fn = self._kernels[key] = h
        nx, ny, nz = shape
        outs = [torch.empty((nz, ny, nx), dtype=torch.float64 if t == "f64" else torch.float32,
                            device=dev).permute(2, 1, 0) for t in types]
        ptrs = (ctypes.c_void_p * (len(fields) + len(outs)))(*[f.data_ptr() for f in fields],
                                                             *[o.data_ptr() for o in outs])
        sc = (ctypes.c_double * max(1, len(scalars)))(*scalars)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check_pw(_lib.lib().pyr_pw_launch(fn, stream, nx * ny * nz, len(fields) + len(outs), ptrs,
                                               len(scalars), sc))
        for lhs, o in zip(self.outputs, outs):
            sim.variables[lhs].data = o
        for j in undecided:
            sim.settle_storage(self.outputs[j])
        return True


_SAFE_CALLS = {"ddx", "ddy", "ddz", "dd4x", "dd4y", "dd4z", "ring", "ringV", "gfilter", "gbar",
               "gbarx", "gbary", "gbarz", "fbar", "sfilter", "filter", "lap", "div", "grad",
               "laplacian", "d2x", "d2y", "d2z"}


def reads_of(src):
    """Variables an equation reads, or None if it calls anything whose reads
    are not visible in its text (packages, user functions)."""
    try:
        node = ast.parse(src, mode="eval").body
    except SyntaxError:
        return None
    names = set()
    for n in ast.walk(node):
        nm = _var_name(n)
        if nm is not None:
            names.add(nm)
        if isinstance(n, ast.Call):
            f = n.func
            if isinstance(f, ast.Attribute) and isinstance(f.value, ast.Name):
                if f.value.id == "self" and f.attr not in _SAFE_CALLS:
                    return None
                if f.value.id not in ("self", "numpy"):
                    return None
            elif not (isinstance(f, ast.Attribute)):
                return None
        if isinstance(n, ast.Name) and n.id not in ("self", "numpy"):
            return None
    return names


def plan_alg(eqs):
    """Execution plan of the ALG equations: a list of ('eq', eq) and
    ('run', PointwiseRun, [eqs]) items. Pointwise equations are grouped into
    runs; an equation between them may execute first when it neither reads a
    value the run produces nor writes anything the run reads or writes."""
    plan = []
    cur, cur_reads, cur_writes = [], set(), set()

    def flush():
        nonlocal cur, cur_reads, cur_writes
        if len(cur) == 1:
            plan.append(("eq", cur[0]))
        elif cur:
            plan.append(("run", PointwiseRun([(e.LHS[0], e.sRHS) for e in cur]), list(cur)))
        cur, cur_reads, cur_writes = [], set(), set()

    for eq in eqs:
        pw = (eq.rank == 1 and eq.LHS and pure_pointwise(eq.sRHS)) if eq.kind == "ALG" else None
        if pw is not None and pw is not False:
            cur.append(eq)
            cur_reads |= set(pw)
            cur_writes.add(eq.LHS[0])
            continue
        reads = reads_of(eq.sRHS) if eq.kind == "ALG" else None
        writes = set(eq.LHS or [])
        if cur and reads is not None and not (reads & cur_writes) and not (writes & (cur_reads | cur_writes)):
            plan.append(("eq", eq))
            continue
        flush()
        plan.append(("eq", eq))
    flush()
    return plan
