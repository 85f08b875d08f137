"""CPU restatements of the reference's 1-D line workloads (TEST INFRASTRUCTURE).

Only tests/ use this module, as the checker of the GPU front end. FP64
throughout, the compact operators from the C oracle, and the 5-stage RK4 of
pyranda/pyranda.py:758-810 (PHI = dt*F + A*PHI; U += B*PHI; then the ALG
lines). Three workloads:

* advection: `ddt(:phi:) = -:c:*ddx(:phi:)` on a periodic line, optionally
  with the ALG line `:phi: = fbar(:phi:)` after every stage. With the mesh and
  loop of examples/advection.py it reproduces the reference-held
  advec-1d-* values (tests/cases/test1DAdvection.py:3-7); with
  xdom=(0,0.99,100,periodic), dt=1e-3, `while t < 0.1` and the filter it is
  the small 1-D workload of examples/tutorials/Advection1D.py:15-30 plus fbar.
* heat: `ddt(:phi:) = :c:*lap(:phi:)` on a bounded line with
  `bc.const(['phi'],['x1'],2.0)` and `bc.const(['phi'],['xn'],1.0)` after
  every stage (examples/heat1D.py:34-67; pyrandaBC.py:58-160), which keeps
  the linear steady state: the reference-held heat1D-analytic-* values are
  0.0 (tests/cases/testHeat1D.py:3-7).
"""
import numpy as np

from . import pyoracle as po
from .tg_cpu import RK_A, RK_B, RK_ETA, _axis


def rk4_step(U, rhs, algs, time, dt):
    """One 5-stage RK4 step of a single field U (pyranda.py:783-810)."""
    PHI = np.zeros_like(U)
    t0 = time
    for ii in range(5):
        F = rhs(U)
        PHI = dt * F + RK_A[ii] * PHI
        U = U + RK_B[ii] * PHI
        time = t0 + RK_ETA[ii] * dt
        for alg in algs:
            U = alg(U)
    return U, time


class Advection1DCPU:
    """Periodic 1-D advection at speed c, optionally filtered after every stage."""

    def __init__(self, n, x1, xn, c=1.0, filtered=False):
        x, d = _axis(x1, xn, n)
        self.x = x.reshape(n, 1, 1)
        self.d = d
        self.c = c
        self.P = po.Oracle((n, 1, 1), (d, 1.0, 1.0), (True, False, False))
        self.algs = [self.P.sfilter] if filtered else []

    def rhs(self, phi):
        return -self.c * self.P.ddx(phi)

    def run(self, phi, dt, until, dt_rule=None):
        """Step `while until > time` (examples/advection.py:66-72); dt_rule(time)
        gives the next dt (None keeps dt). Returns (phi, time, steps)."""
        time, steps = 0.0, 0
        while until > time:
            phi, time = rk4_step(phi, self.rhs, self.algs, time, dt)
            steps += 1
            if dt_rule is not None:
                dt = dt_rule(time)
        return phi, time, steps


def advection_error(npts):
    """examples/advection.py N 1: one period of the Gaussian pulse at speed 1
    on [0, 2*pi); returns sum((phi - phi0)**2) (the advec-1d-N value)."""
    L = np.pi * 2.0
    Lp = L * (npts - 1.0) / npts
    a = Advection1DCPU(npts, 0.0, Lp, c=1.0)
    r = np.sqrt((a.x - np.pi) ** 2)
    phi0 = 1.0 + 0.1 * np.exp(-(r / (np.pi / 4.0)) ** 2)
    dt_max = 1.0 / npts * L * .90
    tt = L
    phi, _, _ = a.run(phi0.copy(), dt_max, tt, dt_rule=lambda t: min(dt_max, tt - t))
    return float(np.sum((phi - phi0) ** 2))


def advection_fbar_100():
    """xdom=(0,0.99,100,periodic), phi0 = 1 + 0.1 exp(-((x-0.5)/0.1)^2),
    ddt(phi) = -ddx(phi) with phi = fbar(phi) after every stage, dt = 1e-3,
    `while t < 0.1`. setIC ends with the ALG lines (pyranda.py:365), so the
    initial field is filtered once before the first step. Returns (x, phi0,
    phi, steps), phi0 the initial field as set."""
    a = Advection1DCPU(100, 0.0, 0.99, c=1.0, filtered=True)
    phi0 = 1.0 + 0.1 * np.exp(-(np.abs(a.x - .5) / .1) ** 2)
    phi, _, steps = a.run(a.P.sfilter(phi0), 1e-3, 0.1)
    return a.x[:, 0, 0], phi0, phi, steps


def heat_error(npts):
    """examples/heat1D.py N 1 (bounded x, Lp = 2*pi*(N-1)/N, c = 1): the sum
    of |analytic - phi| after 500 steps of dt = L/N*0.005 (the
    heat1D-analytic-N value)."""
    L = np.pi * 2.0
    Lp = L * (npts - 1.0) / npts
    x, d = _axis(0.0, Lp, npts)
    x = x.reshape(npts, 1, 1)
    P = po.Oracle((npts, 1, 1), (d, 1.0, 1.0), (False, False, False))
    xnn = x[-1, 0, 0]
    phi = 1.0 + 1.0 * (xnn - x) / xnn

    def bc(u):
        u = u.copy()
        u[0, :, :] = 2.0
        u[-1, :, :] = 1.0
        return u

    dt_max = L / npts * .005
    tt = dt_max * 500
    dt, time = dt_max, 0.0
    while tt > time:
        phi, time = rk4_step(phi, lambda u: 1.0 * P.laplacian(u), [bc], time, dt)
        dt = min(dt_max, tt - time)
    anl = 1.0 + 1.0 * (xnn - x) / xnn
    return float(np.sum(np.abs(anl - phi)[:, 0, 0]))
