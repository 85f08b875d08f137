"""Immersed boundaries (pyranda_amd/pyrandaIBM.py) on the GPU: examples/
cylinder.py (Mach-2 flow past a level-set cylinder: ibmV slip velocity and
ibmS extension of rho and p, LAD with conductivity, inflow/outflow BCs)
against the reference-held curves tests/baselines/cylinder-2d-32.dat and
cylinder-2d-64.dat (p along y = ny/2 at t = 1.5)."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

EOM = """
ddt(:rho:)  =  -ddx(:rho:*:u:)                  - ddy(:rho:*:v:)
ddt(:rhou:) =  -ddx(:rhou:*:u: + :p: - :tau:)   - ddy(:rhou:*:v:)
ddt(:rhov:) =  -ddx(:rhov:*:u:)                 - ddy(:rhov:*:v: + :p: - :tau:)
ddt(:Et:)   =  -ddx( (:Et: + :p: - :tau:)*:u: - :tx:*:kappa:) - ddy( (:Et: + :p: - :tau:)*:v: - :ty:*:kappa: )
:rho:       =  fbar( :rho:  )
:rhou:      =  fbar( :rhou: )
:rhov:      =  fbar( :rhov: )
:Et:        =  fbar( :Et:   )
:u:         =  :rhou: / :rho:
:v:         =  :rhov: / :rho:
:p:         =  ( :Et: - .5*:rho:*(:u:*:u: + :v:*:v:) ) * ( :gamma: - 1.0 )
:T:         = :p: / (:rho: * :R: )
:div:       =  ddx(:u:) + ddy(:v:)
:beta:      =  gbar( ring(:div:) * :rho: ) * 7.0e-2
:tau:       = :beta:*:div:
[:tx:,:ty:,:tz:] = grad(:T:)
:kappa:     = gbar( ring(:T:)* :rho:*:cv:/(:T: * :dt: ) ) * 1.0e-3
[:u:,:v:,:w:] = ibmV( [:u:,:v:,:w:], :phi:, [:gx:,:gy:,:gz:], [:u1:,:u2:,0.0] )
:rho: = ibmS( :rho: , :phi:, [:gx:,:gy:,:gz:] )
:p:   = ibmS( :p:   , :phi:, [:gx:,:gy:,:gz:] )
bc.extrap(['rho','p','u'],['xn'])
bc.const(['u'],['x1','y1','yn'],u0)
bc.const(['v'],['x1','xn','y1','yn'],0.0)
bc.const(['rho'],['x1','y1','yn'],rho0)
bc.const(['p'],['x1','y1','yn'],p0)
:Et:  = :p: / ( :gamma: - 1.0 )  + .5*:rho:*(:u:*:u: + :v:*:v:)
:rhou: = :rho:*:u:
:rhov: = :rho:*:v:
:cs:  = sqrt( :p: / :rho: * :gamma: )
:dt: = dt.courant(:u:,:v:,:w:,:cs:)
:dtB: = 0.2* dt.diff(:beta:,:rho:)
:dt: = numpy.minimum(:dt:,:dtB:)
:umag: = sqrt( :u:*:u: + :v:*:v: )
"""

IC = """
:gamma: = 1.4
:R: = 1.0
:cp: = :R: / (1.0 - 1.0/:gamma: )
:cv: = :cp: - :R:
rad = sqrt( (meshx-pi)**2  +  (meshy-pi)**2 )
:phi: = rad - pi/4.0
:rho: = 1.0 + 3d()
:p:  =  1.0 + 3d()
:u: = where( :phi:>0.5, mach * sqrt( :p: / :rho: * :gamma:) , 0.0 )
:u: = gbar( gbar( :u: ) )
:v: = 0.0 + 3d()
:Et: = :p:/( :gamma: - 1.0 ) + .5*:rho:*(:u:*:u: + :v:*:v:)
:rhou: = :rho:*:u:
:rhov: = :rho:*:v:
:cs:  = sqrt( :p: / :rho: * :gamma: )
:dt: = dt.courant(:u:,:v:,:w:,:cs:)*.1
[:gx:,:gy:,:gz:] = grad( :phi: )
:gx: = gbar( :gx: )
:gy: = gbar( :gy: )
"""


@pytest.fixture(scope="module")
def pr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyranda_amd
    return pyranda_amd


@pytest.mark.parametrize("npts", [32, 64])
def test_cylinder_curve(pr, npts):
    L = np.pi * 2.0
    lp = L * (npts - 1.0) / npts
    rho0, p0, gamma, mach = 1.0, 1.0, 1.4, 2.0
    u0 = np.sqrt(p0 / rho0 * gamma) * mach
    opts = {"coordsys": 0, "periodic": np.array([False, False, False]), "dim": 3,
            "x1": [0.0, 0.0, 0.0], "xn": [lp, lp, lp], "nn": [npts, npts, 1]}
    ss = pr.pyrandaSim("cylinder_test", opts, silent=True)
    ss.addPackage(pr.pyrandaBC(ss))
    ss.addPackage(pr.pyrandaIBM(ss))
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.EOM(EOM.replace("u0", str(u0)).replace("p0", str(p0)).replace("rho0", str(rho0)))
    ss.setIC(IC.replace("mach", str(mach)))
    time, cfl, tt = 0.0, 1.0, 1.5
    dt = float(ss.variables["dt"].data) * cfl * .1
    steps = 0
    while tt > time:
        time = ss.rk4(time, dt)
        dt = min(float(ss.variables["dt"].data) * cfl, 1.1 * dt)
        dt = min(dt, (tt - time))
        steps += 1
    ny = ss.PyMPI.ny
    xx = ss.PyMPI.zbar(ss.mesh.coords[0].data)
    v = ss.PyMPI.zbar(ss.variables["p"].data)
    base = np.loadtxt(os.path.join(GOLDEN, "cylinder-2d-%d.dat" % npts))
    assert np.allclose(xx[:, ny // 2], base[0], rtol=0, atol=1e-14)
    err = np.abs(v[:, ny // 2] - base[1]).max() / np.abs(base[1]).max()
    print("cylinder-2d-%d" % npts, steps, "steps, max rel diff", err, flush=True)
    # the reference harness accepts 1e-4 (tests/testObj.py:52-58). The bow
    # shock amplifies rounding: scaling the initial rho, rhou and Et by
    # (1 + 2^-52) moves this curve by 4e-11 (32) and 7e-11 (64) at t = 1.5,
    # and per-operation rounding differences enter at every stage; measured
    # against the reference 7.4e-10 (32) and 1.1e-9 (64).
    assert err <= 1e-8


MM_EOM = """
ddt(:rho:)  =  -ddx(:rho:*:u: - :tau:) - ddy(:rho:*:v: - :tau:)
ddt(:rhoA:)  =  -ddx(:rhoA:*:uA: - :tauA:) -ddy(:rhoA:*:vA: - :tauA:)
ddt(:phi:)  =  - :gx: * :u1: - :gy: * :v1:
:rho:       =  fbar( :rho:  )
:rhoA:      =  fbar( :rhoA:  )
:dr:        = ddx(:rho:)
:drA:       = ddx(:rhoA:)
:tau:       = 5.0e-7*gbar(abs(lap(lap(:rho:)))) * :dx4: * :u: * :dr: * :dx:
:tauA:      = 5.0e-7*gbar(abs(lap(lap(:rhoA:)))) * :dx4: * :uA: * :drA: * :dx:
[:gx:,:gy:,:gz:] = grad( :phi: )
:mgp:  =  sqrt( :gx:**2 + :gy:**2 + :gz:**2 )
:rho:         = ibmS( :rho: , :phi:, [:gx:,:gy:,:gz:] )
:rhoA:         = ibmS( :rhoA: , -:phi:, [-:gx:,-:gy:,-:gz:] )
:uT:        =  where( :phi: > 0.0, :u:, :uA:)
:rhoB:        =  where( :phi: > 0.0, :rho:, :rhoA:)
:rhoT:        =  where( abs(:phi:/:dx:*2.0) > 1.0, :rhoB:, .5*(:phi:/:dx:*2.0 + 1.0)*(:rho:-:rhoA:) + (:rhoA:)   )
"""


def _golden(name):
    import json
    with open(os.path.join(GOLDEN, "reference_scalars.json")) as f:
        return json.load(f)[name]["value"]


@pytest.mark.parametrize("diffusive", [False, True])
@pytest.mark.parametrize("npts", [50, 100, 200])
def test_interface_advection(pr, npts, diffusive):
    """examples/interface.py in its 1-D test form: two materials on either
    side of a level set advected at unit speed, each extended across the
    interface by ibmS (the second with the reversed level set), hyper-
    diffusion from lap(lap(.)); the minimum of the blended density over the
    right half at t = 0.2 L against MM-adv1d-<n>[-d]
    (tests/cases/testMM_simple.py:3-8)."""
    L = np.pi * 2.0
    lp = L * (npts - 1.0) / npts
    opts = {"coordsys": 0, "periodic": np.array([False, False, True]), "dim": 3,
            "x1": [0.0, 0.0, 0.0], "xn": [lp, lp, lp], "nn": [npts, 1, 1]}
    ss = pr.pyrandaSim("advection", opts, silent=True)
    ss.addPackage(pr.pyrandaIBM(ss))
    ss.EOM(MM_EOM)
    x = ss.mesh.coords[0].data
    amp = 1.0e6 if diffusive else 1.0
    V = ss.variables
    V["rho"].data += ss.gfilter(torch.where(x < np.pi * .25, amp, 1.0))
    V["rhoA"].data += 1.0e6 * (1.0 + .1 * torch.exp(-(x - np.pi * 1.5) ** 2 / .01))
    V["phi"].data = 3.14159 - x
    V["uA"].data += 1.0
    V["u"].data += 1.0
    V["u1"].data += 1.0
    V["u1"].data = 1.0
    V["dx"].data += (x[1, 0, 0] - x[0, 0, 0])
    V["dx4"].data += (x[1, 0, 0] - x[0, 0, 0]) ** 4
    ss.updateVars()
    time = 0.0
    dt_max = 1.0 / ss.mesh.nn[0] * L * .25
    tt = L / 1.0 * 1.0 * .2
    dt = dt_max
    steps = 0
    while tt > time:
        time = ss.rk4(time, dt)
        dt = min(dt_max, (tt - time))
        V["rho"].data[0, :, :] = amp
        steps += 1
    v2 = ss.PyMPI.zbar(V["rhoT"].data)
    got = np.min(v2[npts // 2:, 0])
    name = "MM-adv1d-%d%s" % (npts, "-d" if diffusive else "")
    want = _golden(name)
    err = abs(got - want) / abs(want)
    print(name, steps, "steps:", repr(got), "reference", want, "rel diff", err, flush=True)
    if diffusive and npts < 200:
        # the minimum sits at x = pi, the first point of the right half,
        # within reach of the 1e6 jump (the per-step reset of rho[0] alone
        # moves it by 1e-3 relative): measured 1.3e-6 (50) and 2.6e-5 (100)
        # relative, 4e-11 and 2e-11 of the 1e6 field scale, the same with
        # the right-hand sides fused into the RK update or not; held to the
        # reference harness's own bar (1e-4, tests/testObj.py:52-58)
        assert err <= 1e-4 and abs(got - want) <= 1e-10 * amp
    else:
        # the reference holds 12 significant digits
        assert err <= 1e-9
