"""Spherical meshes (coordsys 2) on the GPU: the device compositions of div,
divT, grad, lap, ring and fbar (pyranda_amd/spherical.py) against the
oracle's restatement (oracle/cyl_cpu.py SphOracle, pinned by closed forms in
tests/test_oracle.py) on the same fields, FP64, <= 1e-12 relative, and
examples/eulerSpherical1D.py against the reference-held curve."""
import os

import numpy as np
import pytest
import torch

from meshgen import mesh

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyranda_amd
    return pyranda_amd


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asfortranarray(a).transpose(2, 1, 0))).cuda().permute(2, 1, 0)


def _host(t):
    if not isinstance(t, torch.Tensor):
        return t
    return t.permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0)


@pytest.mark.parametrize("nn", [(48, 40, 32), (64, 1, 1)])
def test_spherical_operators_match_oracle(pr, oracle_mod, nn):
    """A 3-D (r, theta, phi) block away from the origin and the poles (phi
    periodic) and the one-dimensional radial mesh of the reference's
    spherical Sod problem (cell volume 4 pi r^2 dr)."""
    from oracle.cyl_cpu import SphOracle
    nx, ny, nz = nn
    if ny == 1:
        dom = [(1.0 / nx, 1.0, nx), (np.pi / 2 - .01, np.pi / 2 + .01, 1), (-np.pi / 16, np.pi / 16, 1)]
    else:
        dom = [(0.5, 1.5, nx), (0.3, 1.2, ny), (0.0, 2 * np.pi * (nz - 1) / nz, nz)]
    opts = {"coordsys": 2, "periodic": np.array([False, False, True]),
            "x1": [d[0] for d in dom], "xn": [d[1] for d in dom], "nn": list(nn)}
    ss = pr.pyrandaSim("sph", opts, silent=True)
    _say("spherical", nn)
    (R, T, F), dd = mesh(dom)
    m = ss.PyMPI.metrics
    assert np.allclose(_host(m.radius), R, rtol=0, atol=1e-14)
    assert np.allclose(_host(m.theta), T, rtol=0, atol=1e-14)
    P = oracle_mod.Oracle(nn, ss.PyMPI.ctx.dpatch, (False, False, True))
    S = SphOracle(P, R, T, ss.PyMPI.ctx.dpatch)
    assert np.abs(_host(m.CellVolS) - S.vol_s).max() <= 1e-12 * np.abs(S.vol_s).max()
    rng = np.random.default_rng(sum(nn))
    f = [np.asfortranarray(np.cos(3 * R) * np.sin(F + k) + 0.1 * np.cos(2 * T + k) + 0.01 * rng.standard_normal(nn))
         for k in range(9)]
    fd = [_dev(a) for a in f]
    der, fil = ss.PyMPI.der, ss.PyMPI.fil

    def close(got, want, what):
        err = np.abs(_host(got) - want).max() / max(np.abs(want).max(), 1e-300)
        _say("  ", what, err)
        assert err <= 1e-12, what

    close(der.div(fd[0], fd[1], fd[2]), S.div(f[0], f[1], f[2]), "div")
    for k, (g, w) in enumerate(zip(der.divT(*fd), S.divT(*f))):
        close(g, w, "divT %d" % k)
    for k, (g, w) in enumerate(zip(der.grad(fd[3]), S.grad(f[3]))):
        if np.abs(w).max() > 0:
            close(g, w, "grad %d" % k)
    close(der.laplacian(fd[4]), S.laplacian(f[4]), "lap")
    close(der.ring(fd[5]), S.ring(f[5]), "ring")
    close(fil.filter(fd[6]), S.fbar(f[6]), "fbar")
    # the mesh coordinates become Cartesian (pyrandaMesh.py:156-164)
    X = _host(ss.mesh.coords[0].data)
    Z = _host(ss.mesh.coords[2].data)
    assert np.allclose(X, R * np.sin(T) * np.cos(F), rtol=0, atol=1e-14)
    assert np.allclose(Z, R * np.cos(T), rtol=0, atol=1e-14)


SPH_EOM_2D = """
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
[:fxx:,:fxy:,:fxz:] = [:rhou:*:u: + :p: - :tau:, :v:       , :rhou:*:w:]
[:fyx:,:fyy:,:fyz:] = [      :v:               , :p:-:tau: , :v:       ]
[:fzx:,:fzy:,:fzz:] = [:rhow:*:u:,               :v:       , :rhow:*:w: + :p: - :tau:]
[:xmom:,:ymom:,:zmom:] = divT(:fxx:,:fxy:,:fxz:,:fyx:,:fyy:,:fyz:,:fzx:,:fzy:,:fzz:)
bc.const(['u'],['x1','xn'],0.0)
:cs:  = sqrt( :p: / :rho: * :gamma: )
:dt: = dt.courant(:u:,:v:,:w:,:cs:)
"""

SPH_IC = """rad = sqrt( (meshx-pi)**2  +  (meshy-0.0)**2 )
    :gamma: = 1.4
    wgt = .5*(1.0-tanh( (rad-pi/16.0)/.1) )
    :Et:  = 1.0/(:gamma:-1.0) * (wgt*.9 + .1)
    :rho: = 1.0*wgt + (1.0-wgt)*.125
    """


def _sph_run(ss, tt, dt_max):
    """The examples' time loop: dt from the dt variable, capped at dt_max
    and at the time left."""
    time, dt = 0.0, dt_max
    steps = 0
    while tt > time:
        time = ss.rk4(time, dt)
        dt = min(dt_max, float(ss.var("dt").data))
        dt = min(dt, (tt - time))
        steps += 1
    return steps


SPH_EOM = """
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
[:fxx:,:fxy:,:fxz:] = [:rhou:*:u: + :p: - :tau:, :v:       , :rhou:*:w:]
[:fyx:,:fyy:,:fyz:] = [      :v:               , :p:-:tau: , :v:       ]
[:fzx:,:fzy:,:fzz:] = [:rhow:*:u:,               :v:       , :rhow:*:w: + :p: - :tau:]
[:xmom:,:ymom:,:zmom:] = divT(:fxx:,:fxy:,:fxz:,:fyx:,:fyy:,:fyz:,:fzx:,:fzy:,:fzz:)
:cs:  = sqrt( :p: / :rho: * :gamma: )
:dt: = dt.courant(:u:,:v:,:w:,:cs:)
:dt: = numpy.minimum( :dt:, 0.1 * dt.diff(:beta:,:rho:) )
:rhow: = :rhow: * 0.0
bc.const(['u'],['x1'],0.0)
bc.extrap(['rho','p'],['x1'],order=1)
:rhou: = :rho:*:u:
:Et: = :p:/( :gamma: - 1.0 ) + .5*:rho:*(:u:*:u: + :w:*:w:)
"""


def test_euler_spherical_1d_curve(pr):
    """examples/eulerSpherical1D.py at 128 (512 radial points, coordsys 2:
    the r^2-weighted div and divT, the volume-weighted fbar, CFL and
    diffusive dt from the field metrics) against the reference-held curve
    tests/baselines/euler-SPH-1D.dat (p(r) at t = 0.1)."""
    import numpy
    npts = 128
    L = numpy.pi * 2.0
    mpts = npts * 4
    spike = numpy.pi / 16
    opts = {"coordsys": 2, "periodic": [False, False, True],
            "x1": [L / mpts, numpy.pi / 2 - .01, -spike], "xn": [L, numpy.pi / 2 + .01, spike],
            "nn": [mpts, 1, 1]}
    ss = pr.pyrandaSim("sod2dSpherical_1D", opts, silent=True)
    ss.addPackage(pr.pyrandaBC(ss))
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.EOM(SPH_EOM)
    ss.setIC(SPH_IC)
    steps = _sph_run(ss, 0.1, 1.0 / ss.mesh.nn[0] * 0.05)
    x = ss.mesh.coordsNative[0].data[:, 0, 0].cpu().numpy()
    p = ss.var("p").data[:, 0, 0].double().cpu().numpy()
    base = numpy.loadtxt(os.path.join(GOLDEN, "euler-SPH-1D.dat"))
    assert numpy.allclose(x, base[0], rtol=0, atol=1e-14)
    err = numpy.abs(p - base[1]).max() / numpy.abs(base[1]).max()
    _say("euler-SPH-1D", steps, "steps, max rel diff", err)
    # the reference harness accepts 1e-4 (tests/testObj.py:52-58)
    assert err <= 1e-9


def test_euler_spherical_2d_curve(pr):
    """examples/eulerSpherical.py at 32 (a 128 x 1 x 32 (r, theta, phi)
    mesh, phi periodic over pi/8: the azimuthal terms of div and divT and
    the three-direction filtered cell volume) against the reference-held
    curve tests/baselines/euler-SPH-2D.dat (rho averaged over phi at t =
    0.1)."""
    import numpy
    npts = 32
    L = numpy.pi * 2.0
    spike = numpy.pi / 16
    opts = {"coordsys": 2, "periodic": numpy.array([False, False, True]),
            "x1": [L / 10, numpy.pi / 2 - .01, -spike], "xn": [L, numpy.pi / 2 + .01, spike],
            "nn": [npts * 4, 1, npts]}
    ss = pr.pyrandaSim("2dSpherical", opts, silent=True)
    ss.addPackage(pr.pyrandaBC(ss))
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.EOM(SPH_EOM_2D)
    ss.setIC(SPH_IC)
    steps = _sph_run(ss, 0.1, 1.0 / ss.mesh.nn[0] * 0.05)
    x = ss.mesh.coordsNative[0].data[:, 0, npts // 2].cpu().numpy()
    rho = ss.var("rho").data[:, 0, :].double().cpu().numpy().mean(axis=1)
    base = numpy.loadtxt(os.path.join(GOLDEN, "euler-SPH-2D.dat"))
    assert numpy.allclose(x, base[0], rtol=0, atol=1e-14)
    err = numpy.abs(rho - base[1]).max() / numpy.abs(base[1]).max()
    _say("euler-SPH-2D", steps, "steps, max rel diff", err)
    # the reference harness accepts 1e-4 (tests/testObj.py:52-58)
    assert err <= 1e-9


@pytest.mark.parametrize("where", ["origin", "pole"])
def test_spherical_mesh_rejects_singular_points(pr, where):
    """A mesh point at r = 0 or on the polar axis (sin(theta) = 0) would
    divide by zero in the metrics; set-up raises instead."""
    if where == "origin":
        x1, xn = [0.0, 0.3, 0.0], [1.0, 1.2, 1.0]
    else:
        x1, xn = [0.5, 0.0, 0.0], [1.5, 1.2, 1.0]
    opts = {"coordsys": 2, "periodic": np.array([False, False, True]),
            "x1": x1, "xn": xn, "nn": [16, 12, 8]}
    with pytest.raises(ValueError):
        pr.pyrandaSim("sph_bad", opts, silent=True)
