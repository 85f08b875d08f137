"""Host checks of the pointwise-fusion pass (pyranda_amd/fuse.py).

* every generated kernel for the Taylor-Green EOM compiles with hiprtc for
  gfx950 (no device needed);
* the rewritten expression keeps the semantics of the original right-hand
  side: evaluated on CPU tensors (where every fused subtree takes its torch
  path) it equals the plain `lambda self: <rhs>` bit for bit;
* the generated C++ expression of a subtree is the expected one;
* the planner groups the Taylor-Green pointwise lines into runs, and running
  its plan gives the variables of the written order; the run and reduction
  kernels compile.
"""
import ast
import ctypes
import types

import pytest
import torch

from pyranda_amd import fuse
from pyranda_amd.pyrandaUtils import fortran3d
from pyranda_amd.workloads import TG_EOM


@pytest.fixture(scope="module")
def amd():
    import os
    from pyranda_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    _lib.lib()  # a built library that does not load is a failure, not a skip
    return _lib


def _tg_rhs():
    sm = {"ddx(": "self.ddx(", "ddy(": "self.ddy(", "ddz(": "self.ddz(", "fbar(": "self.filter(",
          "gbar(": "self.gfilter(", "ring(": "self.ring(", "abs(": "numpy.abs(",
          "sqrt(": "numpy.sqrt(", "dt.courant(": "self.packages['Timestep'].courant(",
          "dt.diff(": "self.packages['Timestep'].diff("}
    out = []
    for line in TG_EOM.split("\n"):
        if not line.strip() or line.strip()[0] == "#":
            continue
        out.append(fortran3d(line.split("=")[1], sm))
    return out


def test_expression_codegen():
    node = ast.parse("numpy.sqrt(self.variables['p'].data / self.variables['rho'].data * 1.4) "
                     "+ x**2 - numpy.minimum(y, 0.5)", mode="eval").body
    sub = fuse.FusedSubtree(node, "t")
    assert sub.nleaves == 4
    assert sub.body == ("((sqrt(((x0 / x1) * (0x1.6666666666666p+0))) + (x2 * x2)) - "
                        "pyr_min(x3, (0x1.0000000000000p-1)))")


def test_generated_kernels_compile(amd):
    L = amd.lib()
    n = 0
    for src in _tg_rhs():
        f = fuse.FusedRHS(src, {})
        for sub in f.subtrees:
            for sig in (tuple("T64" for _ in range(sub.nleaves)), tuple("T32" for _ in range(sub.nleaves))):
                name, code = sub._source(sig, "f64" if sig[0] == "T64" else "f32")
                rc = L.pyr_pw_check_source(code.encode(), name.encode())
                assert rc == 0, L.pyr_pw_last_error().decode()
                n += 1
    assert n > 20


def test_sibling_groups(amd):
    """The three flux arguments of every Taylor-Green PDE line form one
    multi-output group (their leaves are plain field reads); operator
    results never join a group; the group kernels compile with hiprtc."""
    L = amd.lib()
    groups = []
    for src in _tg_rhs():
        f = fuse.FusedRHS(src, {})
        if f.group is not None:
            groups.append(f.group)
            assert src.startswith("-self.ddx("), src
    assert [len(g.subs) for g in groups] == [3] * 5
    for g in groups:
        for t in ("T64", "T32"):
            sig = tuple(t for _ in g.leaves)
            dts = tuple("f64" if t == "T64" else "f32" for _ in g.subs)
            name, code = g._source(sig, dts)
            assert L.pyr_pw_check_source(code.encode(), name.encode()) == 0, L.pyr_pw_last_error().decode()


class _Sim:
    """CPU stand-in: fields are CPU tensors, operators are simple functions."""

    def __init__(self, names, shape=(6, 5, 4)):
        g = torch.Generator().manual_seed(0)
        self.variables = {k: types.SimpleNamespace(data=torch.rand(shape, generator=g, dtype=torch.float64) + 0.5)
                          for k in names}
        self.mesh = types.SimpleNamespace(shape=list(shape))
        self.packages = {"Timestep": types.SimpleNamespace(courant=lambda *a: 0.1, diff=lambda *a: 0.2)}

    def ddx(self, v): return torch.roll(v, 1, 0) - v
    def ddy(self, v): return torch.roll(v, 1, 1) * 0.5
    def ddz(self, v): return v.cumsum(2)
    def filter(self, v): return v * 0.9
    def gfilter(self, v): return v + 0.1
    def ring(self, v): return v.abs().sqrt()


def test_fused_rewrite_preserves_semantics():
    from pyranda_amd import devnumpy
    names = ["rho", "u", "v", "w", "rhou", "rhov", "rhow", "Et", "p", "gamma", "tauxx", "tauyy", "tauzz",
             "tauxy", "tauxz", "tauyz", "mu", "ux", "vy", "wz", "div", "uy", "uz", "vx", "vz", "wy", "wx",
             "beta", "taudia", "S", "cs", "dt", "enst", "tke"]
    sim = _Sim(names)
    g = {"numpy": devnumpy.numpy}
    for src in _tg_rhs():
        plain = eval("lambda self: " + src, g)
        fused = fuse.FusedRHS(src, g)
        a, b = plain(sim), fused(sim)
        if torch.is_tensor(a):
            assert torch.equal(a, b), src
        else:
            assert a == b, src


def _tg_eqs():
    from pyranda_amd import pyranda as P
    from pyranda_amd.pyrandaEq import pyrandaEq
    from pyranda_amd.pyrandaTimestep import pyrandaTimestep
    holder = types.SimpleNamespace(packages={})
    sMap = P.pyrandaSim.get_sMap(holder) or holder.sMap
    ts = pyrandaTimestep.__new__(pyrandaTimestep)
    ts.get_sMap()
    sMap.update(ts.sMap)
    return [pyrandaEq(line.strip(), sMap, None) for line in TG_EOM.split("\n")
            if line.strip() and not line.strip().startswith("#")]


def test_plan_groups_pointwise_runs():
    plan = fuse.plan_alg([e for e in _tg_eqs() if e.kind == "ALG"])
    runs = [[e.LHS[0] for e in item[2]] for item in plan if item[0] == "run"]
    assert runs == [["u", "v", "w", "p"], ["div", "enst", "tke", "S"],
                    ["taudia", "tauxx", "tauyy", "tauzz", "tauxy", "tauxz", "tauyz", "cs"]]
    # every ALG equation appears exactly once
    seen = _tg_eqs_alg_cache(plan)
    assert len(seen) == len(set(map(id, seen))) == len([e for e in _tg_eqs() if e.kind == "ALG"])


def _tg_eqs_alg_cache(plan):
    return [e for item in plan for e in ([item[1]] if item[0] == "eq" else item[2])]


def test_plan_order_preserves_results():
    """Executing the plan (runs line by line, as their CPU fallback does) gives
    the same variables as the equations in their written order."""
    from pyranda_amd import devnumpy
    alg = [e for e in _tg_eqs() if e.kind == "ALG"]
    names = set()
    for e in alg:
        names |= set(e.LHS)
        names |= set(fuse.reads_of(e.sRHS) or [])
    names |= {"rho", "rhou", "rhov", "rhow", "Et", "gamma", "dt"}
    g = {"numpy": devnumpy.numpy}

    def run(order):
        sim = _Sim(sorted(names))
        sim.variables["gamma"].data = 1.4
        for e in order:
            r = eval("lambda self: " + e.sRHS, g)(sim)
            sim.variables[e.LHS[0]].data = r
        return sim

    ref = run(alg)
    got = run(_tg_eqs_alg_cache(fuse.plan_alg(alg)))
    for k in names:
        a, b = ref.variables[k].data, got.variables[k].data
        if torch.is_tensor(a):
            torch.testing.assert_close(a, b, rtol=0, atol=0, equal_nan=True, msg=k)
        else:
            assert a == b, k


def test_run_and_reduction_kernels_compile(amd):
    L = amd.lib()
    plan = fuse.plan_alg([e for e in _tg_eqs() if e.kind == "ALG"])
    n = 0
    for item in plan:
        if item[0] != "run":
            continue
        run = item[1]
        for t in ("T32", "T64"):
            sig = tuple("S" if nm == "gamma" else t for nm in run.inputs)
            types_ = run._types(sig)
            name, code = run._source(sig, types_)
            assert L.pyr_pw_check_source(code.encode(), name.encode()) == 0, L.pyr_pw_last_error().decode()
            n += 1
    assert n == 6
    import importlib
    T = importlib.import_module("pyranda_amd.pyrandaTimestep")
    for red, nf in ((T._COURANT, 4), (T._DIFF, 2)):
        for t in ("T32", "T64"):
            name, code = red._source(tuple(t for _ in range(nf)))
            assert L.pyr_pw_check_source(code.encode(), name.encode()) == 0, L.pyr_pw_last_error().decode()


def test_device_scalar_operands_compile(amd):
    """0-d device values (reductions kept on the GPU) enter fused kernels as
    pointers read once; they do not decide the result type (like Python
    scalars), and the kernels compile for gfx950."""
    import ast
    L = amd.lib()
    src = ("self.ring(self.variables['T'].data) * self.variables['rho'].data * 1.5"
           " / (self.variables['T'].data * self.variables['dt'].data)")
    node = ast.parse(src, mode="eval").body
    sub = fuse.FusedSubtree(node, "kappa")
    assert sub.nleaves == 4
    sig = ("T32", "T32", "T32", "D64")
    name, code = sub._source(sig, "f32")
    assert "(double)*(const double*)a.p[3]" in code and "float* __restrict__ out = (float*)a.p[4]" in code
    assert L.pyr_pw_check_source(code.encode(), name.encode()) == 0, L.pyr_pw_last_error().decode()
    run = fuse.PointwiseRun([("a", "self.variables['x'].data * self.variables['dt'].data"),
                             ("b", "self.variables['a'].data + self.variables['x'].data")])
    sig = ("T32", "D64")
    types_ = run._types(sig)
    assert types_ == ["f32", "f32"]
    name, code = run._source(sig, types_)
    assert L.pyr_pw_check_source(code.encode(), name.encode()) == 0, L.pyr_pw_last_error().decode()
    # the one-workgroup final pass of a device reduction
    fin = fuse._final("max")
    name, code = fin._source(("T64",))
    assert L.pyr_pw_check_source(code.encode(), name.encode()) == 0, L.pyr_pw_last_error().decode()


def test_group_operands_and_segmented_reductions_compile(amd):
    """Per-member values of a grouped simulation enter fused kernels through
    the member-index field; the segmented dt reductions (one value per
    member) compile for gfx950, with member segments and workgroup ranges
    laid out as the kernels read them."""
    import ast
    import numpy as np
    L = amd.lib()
    src = "self.variables['T'].data * self.variables['dt'].data"
    sub = fuse.FusedSubtree(ast.parse(src, mode="eval").body, "kappa")
    name, code = sub._source(("T32", "G64"), "f32")
    assert "pmid[i]" in code and "const int* __restrict__ pmid = (const int*)a.p[2]" in code
    assert L.pyr_pw_check_source(code.encode(), name.encode()) == 0, L.pyr_pw_last_error().decode()
    import importlib
    T = importlib.import_module("pyranda_amd.pyrandaTimestep")
    red = fuse.SegmentedReduce("max", T._COURANT.body, 4, 4, "courant")
    name, code = red._source(("T32",) * 4)
    assert L.pyr_pw_check_source(code.encode(), name.encode()) == 0, L.pyr_pw_last_error().decode()
    fin = fuse._segfinal  # compiled at first use on a GPU; its source is built inline
    assert callable(fin)
    seg = fuse.SegmentTable([100, 5000, 70000], [0, 112, 5120], "cpu")
    assert seg.seg.tolist() == [0, 100, 112, 5112, 5120, 75120]
    blk = seg.blk.tolist()
    assert blk[0] == 0 and seg.nblocks == blk[-1] and all(b > a for a, b in zip(blk, blk[1:]))
    shape = fuse.GroupShape(75136, torch.zeros(75136, dtype=torch.int32), 3)
    assert tuple(shape) == (75136, 1, 1) and shape.nmem == 3


def test_offset_dominated_rule():
    """FP32 runs store a field in FP64 when max|v| > 8 (max v - min v)."""
    import torch
    from pyranda_amd.fuse import offset_dominated, WIDE_RATIO
    x = torch.linspace(-1.0, 1.0, 1000, dtype=torch.float64)
    p = 100.0 + 0.05 * torch.sin(7 * x)          # a low-Mach pressure
    assert offset_dominated(p)
    assert offset_dominated(torch.full((10,), 1.0))          # constant nonzero
    assert not offset_dominated(torch.zeros(10))              # zero field
    assert not offset_dominated(torch.sin(7 * x))             # sign-changing
    assert not offset_dominated(1.0 + x)                      # range [0, 2]
    assert not offset_dominated(torch.tensor([1.0, float("nan")]))
    assert not offset_dominated(torch.tensor([1.0, float("inf")]))
    # the threshold itself: max|v| = 9, range 1 is in; 8 and 1 is out
    assert offset_dominated(torch.tensor([8.0, 9.0])) == (9.0 > WIDE_RATIO * 1.0)
    assert not offset_dominated(torch.tensor([7.0, 8.0]))
    # FP32 and permuted (Fortran-order) views decide the same way
    f = (100.0 + torch.rand(4, 5, 6)).permute(2, 1, 0)
    assert offset_dominated(f) and offset_dominated(f.float())
