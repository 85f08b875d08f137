"""examples/shu_osher.py on the GPU: the Shu-Osher shock/entropy-wave problem
(1-D Euler with the LAD bulk viscosity of examples/equation_library.py
euler_1d, inflow constants at x1, CFL and diffusive dt) at 100, 200, 400
and 800 points to t = 1.8; the convergence rate of rho(1.85) between the
resolutions against the reference-held shu-osher-CR
(tests/cases/testShuOsher.py:3)."""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_scalars.json")

EULER_1D = """
ddt(:rho:)  =  -ddx(:rho:*:u:)
ddt(:rhou:) =  -ddx(:rhou:*:u: + :p: - :tau:)
ddt(:Et:)   =  -ddx( (:Et: + :p: - :tau:)*:u: )
:rho:       =  fbar( :rho:  )
:rhou:      =  fbar( :rhou: )
:Et:        =  fbar( :Et:   )
:u:         =  :rhou: / :rho:
:p:         =  ( :Et: - .5*:rho:*(:u:*:u:) ) * ( :gamma: - 1.0 )
:div:       =  ddx(:u:)
:beta:      =  gbar(abs(ring(:div:))) * :rho: * 7.0e-2
:tau:       =  :beta:*:div:
bc.const(['u'],['xn'],0.0)
bc.const(['u'],['x1'],2.629369)
bc.const(['rho'],['x1'],3.857143)
bc.const(['p'],['x1'],10.33333)
bc.const(['Et'],['x1'],39.166585)
:rhou: = :rho:*:u:
:cs:  = sqrt( :p: / :rho: * :gamma: )
:dt: = dt.courant(:u:,0.0,0.0,:cs:)
:dt: = numpy.minimum(:dt:,0.2 * dt.diff(:beta:,:rho:))
"""

IC = """
:gamma: = 1.4
eps = 2.0e-1
:tmp: = (meshx+4.0)/%s
:dum: = tanh(:tmp:)
:dum: = (:dum:+1.0)/2.0
:rho: = 3.857143 + :dum:*(1.0+eps*sin(5.0*meshx) - 3.857143)
:u: = 2.629369*(1.0-:dum:)
:p: = 10.33333 + :dum:*(1.0-10.33333)
:rhou: = :rho: * :u:
:Et:  = :p:/(:gamma: -1.0) + .5*:rho:*:u:*:u:
"""


@pytest.fixture(scope="module")
def pr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyranda_amd
    return pyranda_amd


def _shu(pr, npts):
    """getShu of shu_osher.py:20-80: dt starts at 1% of the CFL limit and
    grows by at most 10% a step."""
    L = 10.0
    Lp = L * (npts - 1.0) / npts
    ss = pr.pyrandaSim("sod", "xdom = (-%s/2.0, %s/2.0, %s)" % (Lp, Lp, npts), silent=True)
    ss.addPackage(pr.pyrandaBC(ss))
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.EOM(EULER_1D)
    ss.setIC(IC % ss.dx)
    time, cfl = 0.0, 0.5
    dt = float(ss.variables["dt"].data) * cfl * .01
    xx = ss.PyMPI.zbar(ss.mesh.coords[0].data)
    tt = 1.8
    steps = 0
    while tt > time:
        time = ss.rk4(time, dt)
        dt = min(dt * 1.1, float(ss.variables["dt"].data) * cfl)
        dt = min(dt, (tt - time))
        steps += 1
    v = ss.PyMPI.zbar(ss.variables["rho"].data)
    return xx[:, 0], v[:, 0], steps


def test_shu_osher_convergence_rate(pr):
    sol = []
    for n in (100, 200, 400, 800):
        x, rho, steps = _shu(pr, n)
        sol.append(np.interp(1.85, x, rho))
        print("shu-osher", n, "points", steps, "steps, rho(1.85) =", repr(sol[-1]), flush=True)
    sol = np.array(sol)
    diff = np.abs(sol[1:] - sol[:-1])
    cr1 = np.log(diff[-1] / diff[-2]) / np.log(.5)
    cr2 = np.log(diff[-2] / diff[-3]) / np.log(.5)
    cr = (cr1 + cr2) / 2.0
    with open(GOLDEN) as f:
        want = json.load(f)["shu-osher-CR"]["value"]
    err = abs(cr - want) / abs(want)
    print("shu-osher-CR", repr(cr), "reference", want, "rel diff", err, flush=True)
    # the reference harness accepts 1e-4 (tests/testObj.py:52-58)
    assert err <= 1e-9
