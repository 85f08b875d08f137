"""GPU checks of the fused pointwise kernels (pyranda_amd/fuse.py).

* a run of pointwise equations (one multi-output kernel) gives, bit for bit,
  the fields of evaluating its equations one after another with the
  per-equation fused kernels, for FP64 and FP32 fields;
* the one-pass max/min reductions of the time-step limits equal torch's
  max/min of the same expression evaluated in FP64.
"""
import importlib
import sys
import types

import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyranda_amd import _lib
    _lib.lib()
    return torch.device("cuda:0")


def _field(shape, dev, dtype, g, lo=0.5, hi=1.5):
    nx, ny, nz = shape
    t = torch.rand((nz, ny, nx), generator=g, dtype=torch.float64) * (hi - lo) + lo
    return t.to(dev, dtype).permute(2, 1, 0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_runs_equal_line_by_line(dev, dtype):
    sys.path.insert(0, __import__("os").path.dirname(__file__))
    from test_fuse_host import _tg_eqs
    from pyranda_amd import fuse
    shape = (40, 33, 17)
    alg = [e for e in _tg_eqs() if e.kind == "ALG"]
    plan = fuse.plan_alg(alg)
    runs = [item for item in plan if item[0] == "run"]
    assert len(runs) == 3
    g = torch.Generator().manual_seed(7)
    names = set()
    for _, run, eqs in runs:
        names |= set(run.inputs) | set(run.outputs)

    def make():
        sim = types.SimpleNamespace(mesh=types.SimpleNamespace(shape=list(shape)), variables={})
        gg = torch.Generator().manual_seed(7)
        for k in sorted(names):
            sim.variables[k] = types.SimpleNamespace(data=_field(shape, dev, dtype, gg))
        sim.variables["gamma"] = types.SimpleNamespace(data=1.4)
        return sim

    fused, ref = make(), make()
    for _, run, eqs in runs:
        _say("run", run.outputs)
        assert run(fused)
        for e in eqs:
            ref.variables[e.LHS[0]].data = e.RHS(ref)
        for k in run.outputs:
            a, b = fused.variables[k].data, ref.variables[k].data
            assert a.dtype == b.dtype == dtype, k
            torch.testing.assert_close(a, b, rtol=0, atol=0, equal_nan=True, msg=k)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_timestep_reductions(dev, dtype):
    T = importlib.import_module("pyranda_amd.pyrandaTimestep")
    shape = (64, 48, 20)
    g = torch.Generator().manual_seed(11)
    u, v, w, c = (_field(shape, dev, dtype, g, -2.0, 2.0) for _ in range(4))
    dx, dy, dz, gl = 0.1, 0.05, 0.2, 0.05
    got = T._COURANT(shape, [u, v, w, c], [dx, dy, dz, gl])
    d = [t.double() for t in (u, v, w, c)]
    ref = ((d[0].abs() / dx + d[1].abs() / dy) + d[2].abs() / dz + d[3].abs() / gl).max().item()
    assert got == ref
    bulk = _field(shape, dev, dtype, g, -1e-3, 1.0)
    rho = _field(shape, dev, dtype, g)
    got = T._DIFF(shape, [bulk, rho], [gl])
    ref = ((rho.double() * gl) * gl / torch.clamp(bulk.double(), min=1e-12)).min().item()
    assert got == ref


@pytest.mark.parametrize("mixed", [False, True])
def test_sibling_group_equals_separate_kernels(dev, mixed):
    """The flux arguments of each Taylor-Green PDE line, evaluated by one
    multi-output kernel (FusedGroup), give bit for bit the right-hand sides
    of one kernel per subtree; FP32 fields, or with one FP64 field mixed in
    (then only the outputs that read it are FP64)."""
    sys.path.insert(0, __import__("os").path.dirname(__file__))
    from test_fuse_host import _tg_rhs
    from pyranda_amd import devnumpy, fuse
    shape = (40, 33, 17)
    gg = torch.Generator().manual_seed(11)
    names = ["rho", "u", "v", "w", "rhou", "rhov", "rhow", "Et", "tauxx", "tauyy", "tauzz", "tauxy",
             "tauxz", "tauyz"]
    sim = types.SimpleNamespace(mesh=types.SimpleNamespace(shape=list(shape)), variables={})
    for k in names:
        dt = torch.float64 if mixed and k == "tauxy" else torch.float32
        sim.variables[k] = types.SimpleNamespace(data=_field(shape, dev, dt, gg))
    # operators stand-ins that keep the arguments (the pointwise results are compared)
    seen = []
    sim.ddx = sim.ddy = sim.ddz = lambda v: (seen.append(v), v)[1]
    g = {"numpy": devnumpy.numpy}
    pdes = [src for src in _tg_rhs() if src.startswith("-self.ddx(")]
    assert len(pdes) == 5
    for src in pdes:
        old = fuse.GROUP_SIBLINGS
        try:
            fuse.GROUP_SIBLINGS = False
            sep = fuse.FusedRHS(src, g)
            fuse.GROUP_SIBLINGS = True
            grp = fuse.FusedRHS(src, g)
        finally:
            fuse.GROUP_SIBLINGS = old
        assert sep.group is None and grp.group is not None and len(grp.group.subs) == 3
        seen.clear()
        a = sep(sim)
        args_a = list(seen)
        seen.clear()
        b = grp(sim)
        args_b = list(seen)
        torch.cuda.synchronize()
        assert torch.equal(a, b), src
        for x, y in zip(args_a, args_b):
            assert x.dtype == y.dtype and torch.equal(x, y), src
        if mixed:
            assert any(x.dtype == torch.float64 for x in args_b) == ("tauxy" in src), src
    _say("sibling groups equal", "mixed" if mixed else "fp32")


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_nan_propagates_like_numpy(dev, dtype):
    """minimum/maximum return NaN when either argument is NaN and sign(NaN)
    is NaN, as numpy does (a NaN in an EOM must reach the NaN checks)."""
    import ast
    import numpy as np
    from pyranda_amd import fuse
    shape = (64, 3, 2)
    n = shape[0] * shape[1] * shape[2]
    g = torch.Generator().manual_seed(5)
    a = torch.rand(n, generator=g, dtype=torch.float64) - 0.5
    b = torch.rand(n, generator=g, dtype=torch.float64) - 0.5
    a[::3] = float("nan")
    b[1::4] = float("nan")
    a[7] = b[7] = float("nan")
    a[8] = b[8] = 0.25
    an, bn = a.numpy().astype(np.float32 if dtype == torch.float32 else np.float64), b.numpy()
    bn = bn.astype(an.dtype)

    def field(x):
        return torch.from_numpy(x.copy()).reshape(shape[2], shape[1], shape[0]).to(dev).permute(2, 1, 0)

    for src, ref in [("numpy.minimum(a, b)", np.minimum(an, bn)), ("numpy.maximum(a, b)", np.maximum(an, bn)),
                     ("numpy.minimum(b, a)", np.minimum(bn, an)), ("numpy.maximum(b, a)", np.maximum(bn, an)),
                     ("numpy.sign(a)", np.sign(an)), ("numpy.sign(b)", np.sign(bn))]:
        node = ast.parse(src, mode="eval").body
        sub = fuse.FusedSubtree(node, "nan")
        names = [ast.unparse(leaf) for leaf in sub.leaves]
        got = sub(shape, *[field(an) if nm == "a" else field(bn) for nm in names])
        got = got.permute(2, 1, 0).reshape(-1).cpu().numpy()
        _say(src, int(np.isnan(ref).sum()), "NaNs")
        np.testing.assert_array_equal(np.isnan(got), np.isnan(ref), err_msg=src)
        ok = ~np.isnan(ref)
        np.testing.assert_array_equal(got[ok], ref[ok], err_msg=src)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_vector_and_scalar_paths_agree(dev, dtype):
    """The generated elementwise loop takes groups of four elements with
    16-byte accesses when every pointer is 16-byte aligned and goes element
    by element otherwise and past the last whole group: a pointwise
    expression over fields whose length is not a multiple of four, over
    16-byte aligned storage and over views one element into it, equals the
    same expression in torch (FP64 arithmetic, one rounding) bit for bit."""
    from pyranda_amd import devnumpy, fuse
    g = {"numpy": devnumpy.numpy}
    src = ('self.variables["a"].data*self.variables["b"].data-self.variables["c"].data'
           '/(1.0+self.variables["a"].data*self.variables["a"].data)')
    rhs = fuse.FusedRHS(src, g)
    gg = torch.Generator().manual_seed(5)
    for shape in ((37, 11, 5), (64, 16, 4)):
        n = shape[0] * shape[1] * shape[2]
        base = {k: torch.rand(n + 1, generator=gg, dtype=torch.float64).to(dev, dtype) for k in "abc"}
        for off in (0, 1):
            sim = types.SimpleNamespace(mesh=types.SimpleNamespace(shape=list(shape)), variables={})
            for k in "abc":
                t = base[k][off:off + n].view(shape[2], shape[1], shape[0]).permute(2, 1, 0)
                sim.variables[k] = types.SimpleNamespace(data=t)
            got = rhs(sim)
            a, b, c = (sim.variables[k].data.double() for k in "abc")
            want = (a * b - c / (1.0 + a * a)).to(dtype)
            torch.cuda.synchronize()
            _say("vector/scalar paths", dtype, shape, "offset", off, got.data_ptr() % 16)
            assert torch.equal(got, want), (shape, off)
