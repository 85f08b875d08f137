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
