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
result is stored in FP64 if any input field is FP64, else in FP32. Subtrees
whose inputs are all scalars are evaluated in Python; inputs that are not
whole fields (slices, vectors) make that subtree evaluate with torch ops.
"""
import ast
import ctypes
import hashlib

import torch

from . import _lib
from .devnumpy import numpy as _dnp

_UNARY = {"sqrt": "sqrt", "abs": "fabs", "absolute": "fabs", "exp": "exp", "log": "log",
          "sin": "sin", "cos": "cos", "tan": "tan", "tanh": "tanh", "arctan": "atan",
          "floor": "floor"}
_BINARY = {"minimum": "pyr_min", "maximum": "pyr_max", "arctan2": "atan2"}
_CMP = {ast.Lt: "<", ast.Gt: ">", ast.LtE: "<=", ast.GtE: ">=", ast.Eq: "==", ast.NotEq: "!="}

PRELUDE = r"""
struct PyrPwArgs { long long n; void* p[24]; double s[16]; };
__device__ __forceinline__ double pyr_min(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double pyr_max(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double pyr_sign(double a) { return (double)((a > 0.0) - (a < 0.0)); }
"""


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

    def _source(self, sig, out_dt):
        name = "pyr_pw_" + hashlib.sha1((self.body + repr(sig) + out_dt).encode()).hexdigest()[:16]
        lines = [PRELUDE, 'extern "C" __global__ void %s(PyrPwArgs a) {' % name]
        ip = isc = 0
        decl = []
        for i, s in enumerate(sig):
            if s == "S":
                decl.append("  const double x%d = a.s[%d];" % (i, isc))
                isc += 1
            else:
                ct = "double" if s == "T64" else "float"
                decl.append("  const %s* __restrict__ p%d = (const %s*)a.p[%d];" % (ct, i, ct, ip))
                ip += 1
        ot = "double" if out_dt == "f64" else "float"
        lines += decl
        lines.append("  %s* __restrict__ out = (%s*)a.p[%d];" % (ot, ot, ip))
        lines.append("  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;"
                     " i += (long long)gridDim.x * blockDim.x) {")
        for i, s in enumerate(sig):
            if s != "S":
                lines.append("    const double x%d = (double)p%d[i];" % (i, i))
        lines.append("    out[i] = (%s)(%s);" % (ot, self.body))
        lines.append("  }")
        lines.append("}")
        return name, "\n".join(lines)

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
        fields = []
        scalars = []
        any64 = False
        for v in vals:
            if torch.is_tensor(v) and v.dim() == 0:
                v = v.item()
            if isinstance(v, (int, float)) or (hasattr(v, "dtype") and not torch.is_tensor(v)
                                                and getattr(v, "shape", None) == ()):
                sig.append("S")
                scalars.append(float(v))
            elif _field_ok(v, shape):
                sig.append("T64" if v.dtype == torch.float64 else "T32")
                any64 |= v.dtype == torch.float64
                fields.append(v)
            else:
                return self._py(*vals)
        if not fields or self.is_bool:
            return self._py(*vals)
        if len(fields) + 1 > _lib.PW_MAX_PTRS or len(scalars) > _lib.PW_MAX_SCALARS:
            return self._py(*vals)
        out_dt = "f64" if any64 else "f32"
        fn = self.kernel(tuple(sig), out_dt)
        dev = fields[0].device
        nx, ny, nz = shape
        out = torch.empty((nz, ny, nx), dtype=torch.float64 if any64 else torch.float32,
                          device=dev).permute(2, 1, 0)
        ptrs = (ctypes.c_void_p * (len(fields) + 1))(*[f.data_ptr() for f in fields], out.data_ptr())
        sc = (ctypes.c_double * max(1, len(scalars)))(*scalars)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check_pw(_lib.lib().pyr_pw_launch(fn, stream, nx * ny * nz, len(fields) + 1, ptrs,
                                               len(scalars), sc))
        return out


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


class FusedRHS:
    """Callable replacing `lambda self: <rhs>` with fused pointwise kernels."""

    def __init__(self, src, globals_):
        tree = ast.parse(src, mode="eval")
        f = _Fuser(src[:40])
        body = f.visit(tree.body)
        self.subtrees = f.subtrees
        self.src = src
        lam = ast.Expression(body=ast.Lambda(
            args=ast.arguments(posonlyargs=[], args=[ast.arg(arg="self"), ast.arg(arg="_shape"),
                                                     ast.arg(arg="_pw")],
                               kwonlyargs=[], kw_defaults=[], defaults=[]),
            body=body))
        ast.fix_missing_locations(lam)
        self._fn = eval(compile(lam, "<eom %s>" % src[:40], "eval"), globals_)

    def __call__(self, sim):
        return self._fn(sim, tuple(sim.mesh.shape), self.subtrees)
