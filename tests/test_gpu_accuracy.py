"""Accuracy of the FP32 Taylor-Green run against the FP64 run on the GPU.

The 3-D Taylor-Green Euler+LAD workload at 256^3 (gamma 1.4, Mach 0.1,
CFL 0.5 from the initial state, LAD on, the EOM of
examples/TaylorGreen.py:40-86), 100 RK4 steps (pyranda.py:758-810) in the FP32
mode (pyrandaSim wide_fields="auto": FP32 storage for operator results and the
LAD chain, FP64 for offset-dominated fields, values computed from FP64 values
only and flux arguments with FP64 inputs; FP32 sweeps plus one FP64-residual
refinement in every operator on FP32 fields),
against the same steps with FP64 fields: every conserved variable within
1e-6 of the FP64 result in both the max norm and the L2 norm (relative to
the FP64 field's max and L2 norms).
"""
import math
import sys

import pytest
import torch

pytestmark = pytest.mark.gpu

BAR = 1e-6  # rel_max and rel_L2 of every conserved variable after 100 steps


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _tg(pr, n, precision, p0):
    from pyranda_amd.workloads import TG_EOM, TG_IC
    mesh = """
    xdom = (0.0, 2*pi*FF,  Npts, periodic=True)
    ydom = (0.0, 2*pi*FF,  Npts, periodic=True)
    zdom = (0.0, 2*pi*FF,  Npts, periodic=True)
    """.replace("Npts", str(n)).replace("pi", str(math.pi)).replace("FF", str(float(n - 1) / n))
    ss = pr.pyrandaSim("TGvortex", mesh, silent=True, precision=precision)
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.EOM(TG_EOM)
    ss.setIC(TG_IC.replace("p0 = 100.0", "p0 = %r" % p0))
    return ss


@pytest.fixture(scope="module")
def pr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyranda_amd
    return pyranda_amd


def test_taylor_green_256_fp32_within_1e6_of_fp64_after_100_steps(pr):
    n, steps = 256, 100
    p0 = 1.0 / (1.4 * 0.1 ** 2)
    names = ("rho", "rhou", "rhov", "rhow", "Et")
    ref = _tg(pr, n, "fp64", p0)
    dt = float(ref.variables["dt"].data) * 0.5
    t = 0.0
    for _ in range(steps):
        t = ref.rk4(t, dt)
    want = {k: ref.var(k).data.clone() for k in names}
    del ref
    torch.cuda.empty_cache()
    ss = _tg(pr, n, "fp32", p0)
    _say("FP64 storage decided for", sorted(k for k, v in ss._storage.items() if v == "f64"))
    # offset-dominated (p, Et, the normal stresses) and computed from FP64
    # values only (the velocities) in FP64; operator results (the velocity
    # gradients) and the LAD chain built on them in FP32
    assert ss._storage.get("p") == "f64" and ss._storage.get("Et") == "f64"
    assert ss._storage.get("u") == "f64" and ss._storage.get("tauxx") == "f64"
    assert ss._storage.get("ux") == "f32" and ss._storage.get("mu") == "f32" and ss._storage.get("S") == "f32"
    t = 0.0
    for _ in range(steps):
        t = ss.rk4(t, dt)
    torch.cuda.synchronize()
    worst = 0.0
    for k in names:
        a, b = ss.var(k).data.double(), want[k]
        rmax = ((a - b).abs().max() / b.abs().max()).item()
        rl2 = (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item()
        _say("%-5s rel_max %.3e rel_l2 %.3e (max|fp64| %.3e)" % (k, rmax, rl2, b.abs().max().item()))
        worst = max(worst, rmax, rl2)
    assert worst <= BAR, worst
