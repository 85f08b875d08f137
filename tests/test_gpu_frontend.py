"""End-to-end parity of the pyrandaSim front end on the GPU.

The reference's own regression scripts, restated through pyranda_amd's
pyrandaSim (EOM/setIC/rk4/eval), against the values the reference holds:
unit-test-3D (tests/cases/testUnit.py:3), advec-1d-* (test1DAdvection.py),
integration-* (testIntegration.py), heat1D-analytic-* (testHeat1D.py) and
taylorgreen (testTaylorGreen.py:3), plus the 1-D advection + fbar workload
(100 points, 100 RK4 steps) against the oracle driven by the same RK4 on the
CPU (test_advection_fbar_100_vs_oracle).
"""
import json
import math
import os
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_scalars.json")


def _golden(name):
    with open(GOLDEN) as f:
        return json.load(f)[name]["value"]


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyranda_amd
    return pyranda_amd


def test_unit_test_3d(pr):
    """examples/unit_test.py through pysim.eval."""
    Nx = 32
    L = np.pi * 2.0 * (Nx - 1) / Nx
    domain = """
    xdom = (0.0 , L , Nx ,periodic=True)
    ydom = (0.0 , L , Nx ,periodic=True)
    zdom = (0.0 , L , Nx ,periodic=True)
    """.replace("L", str(L)).replace("Nx", str(Nx))
    pysim = pr.pyrandaSim("advection", domain, silent=True)
    E = 0.0
    expr = "sum( abs( ddx( sin(meshx) ) - cos(meshx) ))"
    for e in (expr, expr.replace("x", "y"), expr.replace("x", "z")):
        E += pysim.eval(e)
    domain = """
    xdom = (0.0 , 1.0 , Nx )
    ydom = (0.0 , 1.0 , Nx )
    zdom = (0.0 , 1.0 , Nx )
    """.replace("Nx", str(Nx))
    pysim2 = pr.pyrandaSim("nonperiodic", domain, silent=True)
    for expr in ("mean( abs(ddx(meshx*0.0)) )", "mean( abs(ddx(meshx*0.0 + 1.0)) )"):
        for e in (expr, expr.replace("x", "y"), expr.replace("x", "z")):
            E += pysim2.eval(e)
    expr = "mean( abs(ddx(meshx**power) - power*meshx**(power-1) )  )"
    for power in [1, 2, 3, 4]:
        ex = expr.replace("power", str(float(power)))
        for e in (ex, ex.replace("x", "y"), ex.replace("x", "z")):
            E += pysim2.eval(e)
    for f in ("fbar", "gbar"):
        E += pysim2.eval("mean( abs(%s(meshx*0.0)) )" % f)
        E += pysim2.eval("mean( abs(%s(meshx*0.0 + 1.0) - (meshx*0.0 + 1.0)  ) )" % f)
        E += pysim2.eval("mean( abs(%s(meshx) - (meshx)  ) )" % f)
        E += pysim.eval("mean( abs(%s(sin(meshx)) - sin(meshx)  ) )" % f)
        E += pysim.eval("mean( abs(%s( meshx*0.0 + 1.0) - 1.0  ) )" % f)
        E += pysim.eval("mean( abs(%s( meshx*0.0 )  ) )" % f)
    want = _golden("unit-test-3D")
    _say("unit-test-3D", E, want)
    assert abs(E - want) / want < 1e-9


def advection_error(pr, Npts):
    """examples/advection.py N 1: one period of periodic d1 + RK4."""
    L = np.pi * 2.0
    Lp = L * (Npts - 1.0) / Npts
    imesh = """
    Lp = %s
    Npts = %d
    xdom = (0.0, Lp,  Npts, periodic=True)
    """ % (Lp, Npts)
    ss = pr.pyrandaSim("advection", imesh, silent=True)
    ss.EOM(" ddt(:phi:)  =  -:c: * ddx(:phi:) ")
    ic = """
    r   = sqrt( (meshx-pi)**2  )
    :phi: = 1.0 + 0.1 * exp( -(r/(pi/4.0))**2 )
    :phi2: = 1.0*:phi:
    :c:   = 1.0
    """
    ss.setIC(ic)
    v = 1.0
    dt_max = v / ss.mesh.nn[0] * L * .90
    tt = L / v * 1.0
    dt = dt_max
    time = 0.0
    while tt > time:
        time = ss.rk4(time, dt)
        dt = min(dt_max, (tt - time))
    phi = ss.variables["phi"].data
    phi2 = ss.variables["phi2"].data
    return torch.sum((phi - phi2) ** 2).item()


@pytest.mark.parametrize("n", [50, 100, 200, 300, 400])
def test_advection_1d(pr, n):
    got = advection_error(pr, n)
    want = _golden("advec-1d-%d" % n)
    _say("advec", n, got, want)
    # the harness's own tolerance is 1e-4 (tests/run_tests.py:84)
    assert abs(got - want) / want < 1e-4


@pytest.mark.parametrize("n", [25, 50, 100, 200, 400])
def test_rk4_integration(pr, n):
    """examples/intTest.py N 1: the 5-stage RK4 on a'=a, b'=cos(6t)."""
    pysim = pr.pyrandaSim("integralTests", "xdom = (0.0 , 1.0 , 1 )", silent=True)
    pysim.EOM("""
    ddt(:a:)  =  :a:
    ddt(:b:)  =  cos( 6.0*simtime )
    """)
    pysim.setIC("""
    :a: = 1.0
    :b: = 0.0
    """)
    dt = 1.0 / float(n)
    time = 0.0
    tt, aa = [time], [1.0]
    while time < 1.0:
        time = pysim.rk4(time, dt)
        tt.append(time)
        aa.append(pysim.variables["a"].data[0, 0, 0].item())
    tt, aa = np.array(tt), np.array(aa)
    err = np.sum(np.abs(aa - (1.0 - 1.0 + np.exp(tt))))
    want = _golden("integration-%d" % n)
    _say("integration", n, err, want)
    assert abs(err - want) / want < 1e-4


def _host(t):
    return t.permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0)


def test_advection_fbar_100_vs_oracle(pr):
    """The small 1-D workload: xdom=(0,0.99,100,periodic), ddt(phi)=-ddx(phi)
    with the ALG line phi=fbar(phi) after every stage, dt=1e-3, `while
    t < 0.1` (examples/tutorials/Advection1D.py:15-30 plus the filter), against
    the oracle driven by the same RK4 (oracle/line1d_cpu.py), FP64 to 1e-12."""
    from oracle.line1d_cpu import advection_fbar_100
    ss = pr.pyrandaSim("advection", "xdom = (0.0 , 0.99 , 100 , periodic=True)", silent=True)
    ss.EOM("""
    ddt(:phi:)  =  - ddx(:phi:)
    :phi:       =  fbar( :phi: )
    """)
    ss.setIC(":phi: = 1.0 + 0.1 * exp( -(abs(meshx-.5)/.1 )**2 )")
    dt, time, steps = 1e-3, 0.0, 0
    while time < .1:
        time = ss.rk4(time, dt)
        steps += 1
    x, phi0, ref, nref = advection_fbar_100()
    got = _host(ss.var("phi").data)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    moved = np.abs(ref - phi0).max() / np.abs(phi0).max()
    _say("advection+fbar 100 pts", steps, "steps, rel err", err, "moved", moved)
    assert steps == nref == 100
    assert np.abs(_host(ss.mesh.coords[0].data)[:, 0, 0] - x).max() <= 1e-15
    assert moved > 1e-3  # the pulse moved (the check is not of a static field)
    assert err <= 1e-12


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256])
def test_heat1d_analytic(pr, n):
    """examples/heat1D.py N 1: bounded x, lap with bc.const at both ends
    after every stage; the reference-held heat1D-analytic-N is 0.0
    (tests/cases/testHeat1D.py:3-7; zero baselines compare absolutely to
    1e-4, tests/testObj.py:45-50)."""
    L = np.pi * 2.0
    Lp = L * (n - 1.0) / n
    mesh_options = {"coordsys": 0, "periodic": np.array([False, True, True]), "dim": 1,
                    "x1": [0.0, 0.0, 0.0], "xn": [Lp, Lp, Lp], "nn": [n, 1, 1]}
    ss = pr.pyrandaSim("heat_equation", mesh_options, silent=True)
    ss.addPackage(pr.pyrandaBC(ss))
    ss.EOM("""
    ddt(:phi:)  =  :c: * lap(:phi:)
    bc.const(['phi'],['x1'],2.0)
    bc.const(['phi'],['xn'],1.0)
    """)
    ss.setIC("""
    xnn = meshx[-1,0,0]
    :phi: = 1.0 + 1.0*(xnn - meshx)/xnn
    :c:   = 1.0
    """)
    x = _host(ss.mesh.coords[0].data)
    dt_max = L / ss.mesh.nn[0] * .005
    tt = dt_max * 500
    dt, time = dt_max, 0.0
    while tt > time:
        time = ss.rk4(time, dt)
        dt = min(dt_max, (tt - time))
    xnn = x[-1, 0, 0]
    anl = 1.0 + 1.0 * (xnn - x) / xnn
    err = float(np.sum(np.abs(anl - _host(ss.var("phi").data))[:, 0, 0]))
    _say("heat1D", n, err)
    assert err < 1e-4
    assert err < 1e-10  # FP64 keeps the linear steady state to rounding


from pyranda_amd.workloads import TG_EOM, TG_IC  # noqa: E402


def taylor_green_enstrophy(pr, Npts, tstop=0.1, precision="fp64"):
    imesh = """
    xdom = (0.0, 2*pi*FF,  Npts, periodic=True)
    ydom = (0.0, 2*pi*FF,  Npts, periodic=True)
    zdom = (0.0, 2*pi*FF,  Npts, periodic=True)
    """.replace("Npts", str(Npts)).replace("pi", str(np.pi)).replace("FF", str(float(Npts - 1) / Npts))
    ss = pr.pyrandaSim("TGvortex", imesh, silent=True, precision=precision)
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.EOM(TG_EOM)
    ss.setIC(TG_IC)
    CFL = 0.5
    dt = ss.variables["dt"].data * CFL
    enst0 = ss.var("enst").sum()
    time = 0.0
    while time < tstop:
        time = ss.rk4(time, dt)
        dt = ss.variables["dt"].data * CFL
    return ss.var("enst").sum() / enst0


def test_taylor_green_32(pr):
    got = taylor_green_enstrophy(pr, 32)
    want = _golden("taylorgreen")
    _say("taylorgreen", repr(got), want)
    assert abs(got - want) / want < 1e-6


@pytest.mark.parametrize("precision", ["fp32", "fp32-fp64solve"])
def test_taylor_green_32_fp32_fields(pr, precision):
    """FP32 fields: the enstrophy golden (11 digits) to FP32 storage accuracy."""
    got = taylor_green_enstrophy(pr, 32, precision=precision)
    want = _golden("taylorgreen")
    _say("taylorgreen", precision, repr(got), want)
    assert abs(got - want) / want < 1e-5


EULER_EOM = """
# Primary Equations of motion here
ddt(:rho:)  =  -ddx(:rho:*:u:)                  - ddy(:rho:*:v:)
ddt(:rhou:) =  -ddx(:rhou:*:u: + :p: - :tau:)   - ddy(:rhou:*:v:)
ddt(:rhov:) =  -ddx(:rhov:*:u:)                 - ddy(:rhov:*:v: + :p: - :tau:)
ddt(:Et:)   =  -ddx( (:Et: + :p: - :tau:)*:u: ) - ddy( (:Et: + :p: - :tau:)*:v: )
# Conservative filter of the EoM
:rho:       =  fbar( :rho:  )
:rhou:      =  fbar( :rhou: )
:rhov:      =  fbar( :rhov: )
:Et:        =  fbar( :Et:   )
# Update the primatives and enforce the EOS
:u:         =  :rhou: / :rho:
:v:         =  :rhov: / :rho:
:p:         =  ( :Et: - .5*:rho:*(:u:*:u: + :v:*:v:) ) * ( :gamma: - 1.0 )
# Artificial bulk viscosity (old school way)
:div:       =  ddx(:u:) + ddy(:v:)
:beta:      =  gbar(abs(ring(:div:))) * :rho: * 7.0e-2
:tau:       =  :beta:*:div:
# Apply constant BCs
bc.extrap(['rho','Et'],['x1','xn','y1','yn'])
bc.const(['u','v'],['x1','xn','y1','yn'],0.0)
"""


@pytest.mark.parametrize("npts", [64, 128])
def test_euler_2d_curve(pr, npts):
    """examples/euler.py (2-D Sod, bounded x and y, BC lines) against the
    reference-held curve tests/baselines/euler-2d-<n>.dat."""
    import numpy
    L = numpy.pi * 2.0
    Lp = L * (npts - 1.0) / npts
    mesh_options = {"coordsys": 0, "periodic": numpy.array([False, False, True]), "dim": 3,
                    "x1": [0.0, 0.0, 0.0], "xn": [Lp, Lp, Lp], "nn": [npts, npts, 1]}
    ss = pr.pyrandaSim("advection", mesh_options, silent=True)
    ss.addPackage(pr.pyrandaBC(ss))
    ss.EOM(EULER_EOM)
    ss.setIC("""rad = sqrt( (meshx-pi)**2  +  (meshy-pi)**2 )
    :gamma: = 1.4
    :Et:  = gbar( where( rad < pi/2.0, 1.0/(:gamma:-1.0) , .1 /(:gamma:-1.0) ) )
    :rho: = gbar( where( rad < pi/2.0, 1.0    , .125 ) )
    """)
    x = ss.mesh.coords[0].data
    xx = ss.PyMPI.zbar(x)
    ny = ss.PyMPI.ny
    dt_max = 1.0 / ss.mesh.nn[0] * 0.75
    tt = L / 1.0 * .125
    time, dt = 0.0, dt_max
    while tt > time:
        time = ss.rk4(time, dt)
        dt = min(dt_max, (tt - time))
    v = ss.PyMPI.zbar(ss.variables["rho"].data)
    base = numpy.loadtxt(os.path.join(os.path.dirname(GOLDEN), "euler-2d-%d.dat" % npts))
    assert numpy.allclose(xx[:, ny // 2], base[0], rtol=0, atol=1e-15)
    err = numpy.abs(v[:, ny // 2] - base[1]).max() / numpy.abs(base[1]).max()
    _say("euler-2d", npts, "max rel diff", err)
    # the reference harness accepts 1e-4 (tests/testObj.py:52-58)
    assert err <= 1e-9


RZ_EOM = """
ddt(:rho:)  =  -:mass:
ddt(:rhou:) =  -:xmom:
ddt(:rhow:) =  -:zmom:
ddt(:Et:)   =  -div( (:Et: + :p: - :tau:)*:u:,:v:, (:Et: + :p: - :tau:)*:w: )
:rho:       =  fbar( :rho:  )
:rhou:      =  fbar( :rhou: )
:rhow:      =  fbar( :rhow: )
:Et:        =  fbar( :Et:   )
:u:         =  :rhou: / :rho:
:w:         =  :rhow: / :rho:
:p:         =  ( :Et: - .5*:rho:*(:u:*:u: + :w:*:w:) ) * ( :gamma: - 1.0 )
:div:       =  div(:u:,:v:,:w:)
:beta:      =  gbar(abs(ring(:div:))) * :rho: * 7.0e-2
:tau:       =  :beta:*:div:
:v:         = :v: * 0.0
:mass: = div(:rho:*:u:               , :v:, :rho:*:w:)
[:fxx:,:fxy:,:fxz:] = [:rhou:*:u: + :p: - :tau:, :v:, :rhou:*:w:]
[:fyx:,:fyy:,:fyz:] = [      :v:               , :p:-:tau: , :v:       ]
[:fzx:,:fzy:,:fzz:] = [:rhow:*:u:,               :v:, :rhow:*:w: + :p: - :tau:]
[:xmom:,:ymom:,:zmom:] = divT(:fxx:,:fxy:,:fxz:,:fyx:,:fyy:,:fyz:,:fzx:,:fzy:,:fzz:)
"""


def test_euler_rz_curve(pr):
    """examples/eulerRZ.py at 128 (cylindrical coordsys 1: the radius-weighted
    div and divT, the cell-volume weighted fbar with the flipped parity at the
    axis, SYMM at r = 0) against the reference-held curve
    tests/baselines/euler-RZ-128.dat (pressure along z at the middle radius,
    t = 0.5)."""
    import numpy
    npts = 128
    L = numpy.pi * 2.0
    dx2 = L / (npts - 1) / 2.0
    opts = {"coordsys": 1, "periodic": numpy.array([False, False, True]), "x1": [dx2, -0.1, 0.0],
            "xn": [L, 0.1, L], "nn": [npts, 1, npts], "symmetric": [[True, False], [False, False], [False, False]]}
    ss = pr.pyrandaSim("sod2dRZ", opts, silent=True)
    ss.addPackage(pr.pyrandaBC(ss))
    ss.EOM(RZ_EOM)
    ss.setIC("""rad = sqrt( (meshx-pi)**2  +  (meshz-pi)**2 )
    :gamma: = 1.4
    wgt = .5*(1.0-tanh( (rad-pi/2.0)/.1) )
    :Et:  = 1.0/(:gamma:-1.0) * (wgt*.9 + .1)
    :rho: = 1.0*wgt + (1.0-wgt)*.125
    """)
    dt_max = 1.0 / ss.mesh.nn[0] * 0.75
    tt = 0.5
    time, dt = 0.0, dt_max
    steps = 0
    while tt > time:
        time = ss.rk4(time, dt)
        dt = min(dt_max, (tt - time))
        steps += 1
    half = npts // 2
    x = ss.mesh.coords[0].data[:, 0, 0].cpu().numpy()
    p = ss.var("p").data[half, 0, :].double().cpu().numpy()
    base = numpy.loadtxt(os.path.join(os.path.dirname(GOLDEN), "euler-RZ-128.dat"))
    assert numpy.allclose(x, base[0], rtol=0, atol=1e-14)
    err = numpy.abs(p - base[1]).max() / numpy.abs(base[1]).max()
    _say("euler-RZ-128", steps, "steps, max rel diff", err)
    # the reference harness accepts 1e-4 (tests/testObj.py:52-58)
    assert err <= 1e-9


def test_kh_2d_curve(pr):
    """examples/KH.py at 64^2 (doubly periodic shear layers, LAD with the
    kappa line reading the dt variable, dt from dt.courant/dt.diff in every
    stage) against the reference-held curve tests/baselines/
    KelvinHelmholtzKH-2d-64.dat: rho(x, 3ny/4) at t = 1.5 after 249 steps."""
    import numpy
    from pyranda_amd.workloads import kh_run, kh_setup
    ss = kh_setup(pr, 64, name="KelvinHelmholtzKH-2d-64")
    time, steps = kh_run(ss)
    ny = ss.PyMPI.ny
    xx = ss.PyMPI.zbar(ss.mesh.coords[0].data)
    v = ss.PyMPI.zbar(ss.variables["rho"].data)
    base = numpy.loadtxt(os.path.join(os.path.dirname(GOLDEN), "KH-2d-64.dat"))
    assert numpy.allclose(xx[:, 3 * ny // 4], base[0], rtol=0, atol=1e-15)
    err = numpy.abs(v[:, 3 * ny // 4] - base[1]).max() / numpy.abs(base[1]).max()
    _say("kh-2d-64 steps", steps, "max rel diff", err)
    assert steps == 249
    # the reference harness accepts 1e-4 (tests/testObj.py:52-58)
    assert err <= 1e-9


def test_fp32_fp64_state_through_filter(pr):
    """FP32 KH run: the conservatively filtered variables keep an FP64 state
    (the field is its rounding after every step) and end closer to the FP64
    run than with the state rounded to FP32 at every filter; a field written
    in place is taken as the new state."""
    import torch
    from pyranda_amd.workloads import kh_setup
    # every field in FP32 storage (wide_fields=None): the state mechanism alone
    runs = {"fp64": kh_setup(pr, 64, 48, precision="fp64"),
            "state": kh_setup(pr, 64, 48, precision="fp32", wide_fields=None),
            "rounded": kh_setup(pr, 64, 48, precision="fp32", wide_fields=None)}
    runs["rounded"].fp64_state = False
    st = runs["state"]
    assert sorted(st._self_filtered()) == ["Et", "rho", "rhou", "rhov"]
    dt = float(runs["fp64"].var("dt").data) * 0.5
    for name, s in runs.items():
        t = 0.0
        for _ in range(16):
            t = s.rk4(t, dt)
    torch.cuda.synchronize()
    for k in ("rho", "rhou", "rhov", "Et"):
        assert st._state64[k].dtype == torch.float64
        assert torch.equal(st.var(k).data, st._state64[k].float()), k
    ref = runs["fp64"].var("rhou").data
    err = {n: float((runs[n].var("rhou").data.double() - ref).norm() / ref.norm()) for n in ("state", "rounded")}
    _say("rhou rel l2 vs fp64", err)
    assert err["state"] < err["rounded"]
    # an in-place write to the field replaces the FP64 state
    assert st._state_current("rho")
    st.var("rho").data.mul_(1.0)
    assert not st._state_current("rho")
    st.rk4(t, dt)
    assert st._state_current("rho")
    assert torch.equal(st.var("rho").data, st._state64["rho"].float())


def test_fp32_offset_dominated_fields_stored_fp64(pr):
    """FP32 KH run with storage by value (the default): the fields set close
    to one large value (pressure 2.5, total energy) are stored in FP64, the
    FP64 state of a filtered FP64 field is the field itself, and the run ends
    closer to the FP64 run than with every field in FP32."""
    import torch
    from pyranda_amd.workloads import kh_setup
    runs = {"fp64": kh_setup(pr, 64, 48, precision="fp64"),
            "wide": kh_setup(pr, 64, 48, precision="fp32"),
            "offset": kh_setup(pr, 64, 48, precision="fp32", wide_fields="offset"),
            "f32": kh_setup(pr, 64, 48, precision="fp32", wide_fields=None)}
    w = runs["wide"]
    wide = sorted(k for k, v in w._storage.items() if v == "f64")
    off = sorted(k for k, v in runs["offset"]._storage.items() if v == "f64")
    _say("FP64 storage (auto):", wide, "(offset):", off)
    # auto: the offset-dominated fields and what is computed from FP64
    # values only (u = rhou / rho of the FP64 states); offset: only the former
    assert "p" in wide and "Et" in wide and "u" in wide
    assert "p" in off and "Et" in off and "u" not in off
    assert w.var("p").data.dtype == torch.float64 and w.var("u").data.dtype == torch.float64
    assert runs["offset"].var("u").data.dtype == torch.float32
    dt = float(runs["fp64"].var("dt").data) * 0.5
    for s in runs.values():
        t = 0.0
        for _ in range(16):
            t = s.rk4(t, dt)
    torch.cuda.synchronize()
    assert w.var("Et").data is w._state64["Et"]
    ref = runs["fp64"].var("rhou").data
    err = {n: float((runs[n].var("rhou").data.double() - ref).norm() / ref.norm()) for n in ("wide", "offset", "f32")}
    _say("rhou rel l2 vs fp64", err)
    assert err["wide"] < err["f32"] and err["offset"] < err["f32"]


def test_fp32_derivatives_of_fp64_fields_stored_by_the_sweep(pr, monkeypatch):
    """In an FP32 run, `:ux: = ddx(:u:)` with u stored in FP64 and ux in FP32
    runs one FP64 sweep that stores only the FP32 rounding
    (pyr_apply_to_r32): the run is bit for bit the one that sweeps into an
    FP64 temporary and rounds it afterwards."""
    import torch
    imesh = """
    xdom = (0.0, 2*pi*FF,  32, periodic=True)
    ydom = (0.0, 2*pi*FF,  32, periodic=True)
    zdom = (0.0, 2*pi*FF,  32, periodic=True)
    """.replace("pi", str(np.pi)).replace("FF", str(31.0 / 32))

    def run():
        ss = pr.pyrandaSim("TGvortex", imesh, silent=True, precision="fp32")
        ss.addPackage(pr.pyrandaTimestep(ss))
        ss.EOM(TG_EOM)
        ss.setIC(TG_IC)
        dt = float(ss.variables["dt"].data) * 0.5
        t = 0.0
        for _ in range(3):
            t = ss.rk4(t, dt)
        torch.cuda.synchronize()
        return ss

    calls = []
    orig = pr.pyrandaSim._derivative_to32

    def counted(self, eq):
        r = orig(self, eq)
        if r is not None:
            calls.append(eq.LHS[0])
        return r

    monkeypatch.setattr(pr.pyrandaSim, "_derivative_to32", counted)
    fast = run()
    assert fast._storage.get("u") == "f64" and fast._storage.get("ux") == "f32"
    assert "ux" in calls and "wy" in calls
    monkeypatch.setattr(pr.pyrandaSim, "_derivative_to32", lambda self, eq: None)
    slow = run()
    for k in ("rho", "rhou", "rhov", "rhow", "Et", "ux", "vz", "mu", "beta"):
        a, b = fast.var(k).data, slow.var(k).data
        assert a.dtype == b.dtype and torch.equal(a, b), k


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rhs_fused_into_rk_update(pr, monkeypatch, precision):
    """Each PDE right-hand side's last pointwise kernel (-a - b - c over the
    flux derivatives) evaluated inside the RK update (fuse.DeferredF): the
    run is bit for bit the one that stores F and updates from it."""
    import torch
    from pyranda_amd import fuse
    imesh = """
    xdom = (0.0, 2*pi*FF,  32, periodic=True)
    ydom = (0.0, 2*pi*FF,  32, periodic=True)
    zdom = (0.0, 2*pi*FF,  32, periodic=True)
    """.replace("pi", str(np.pi)).replace("FF", str(31.0 / 32))
    calls = []
    orig = fuse.DeferredF.rk_update

    def counted(self, *a):
        calls.append(1)
        return orig(self, *a)

    monkeypatch.setattr(fuse.DeferredF, "rk_update", counted)
    runs = {}
    for fused in (True, False):
        ss = pr.pyrandaSim("TGvortex", imesh, silent=True, precision=precision)
        ss.fuse_rk = fused
        ss.addPackage(pr.pyrandaTimestep(ss))
        ss.EOM(TG_EOM)
        ss.setIC(TG_IC)
        dt = float(ss.variables["dt"].data) * 0.5
        t = 0.0
        for _ in range(3):
            t = ss.rk4(t, dt)
        torch.cuda.synchronize()
        runs[fused] = ss
        if fused:
            _say("fused updates", len(calls))
            assert len(calls) == 3 * 5 * 5          # steps x stages x conserved variables
    for k in ("rho", "rhou", "rhov", "rhow", "Et", "p", "u"):
        a, b = runs[True].var(k).data, runs[False].var(k).data
        assert a.dtype == b.dtype and torch.equal(a, b), k
