"""CPU restatement of the 2-D Euler (Sod) regression case (TEST INFRASTRUCTURE).

examples/euler.py with dim = 2 and problem = 'sod' (the script behind the
reference's euler-2d-64/128 baselines, tests/cases/test2deuler.py:14-22):
bounded x and y lines (NONE boundaries), LAD bulk viscosity through ring and
gbar, the conservative filter after every stage, and the pyrandaBC
extrapolation / constant boundary lines (pyranda/pyrandaBC.py:58-160), in
numpy with the C oracle's operators and the 5-stage RK4 of
pyranda/pyranda.py:758-810. FP64 throughout. Only tests import it.
"""
import numpy as np

from . import pyoracle as po
from .tg_cpu import RK_A, RK_B, _axis


class Euler2DCPU:

    def __init__(self, n):
        L = 2.0 * np.pi
        Lp = L * (n - 1.0) / n
        x, dx = _axis(0.0, Lp, n)
        y, dy = _axis(0.0, Lp, n)
        # the size-1 z axis: patch spacing of the face construction (n = 1)
        z, dz = _axis(0.0, Lp, 1)
        X, Y = np.meshgrid(x, y, indexing="ij")
        X, Y = X[:, :, None], Y[:, :, None]
        self.n = n
        self.x = x
        self.P = po.Oracle((n, n, 1), (dx, dy, dz), (False, False, True))
        self.gamma = 1.4
        F = np.asfortranarray
        P = self.P
        # initial conditions (euler.py:95-118): rad, then gbar of a step
        rad = np.sqrt((X - np.pi) ** 2 + (Y - np.pi) ** 2)
        g = self.gamma
        Et = P.gfilter(F(np.where(rad < np.pi / 2.0, 1.0 / (g - 1.0), .1 / (g - 1.0))))
        rho = P.gfilter(F(np.where(rad < np.pi / 2.0, 1.0, .125)))
        zero = F(np.zeros((n, n, 1)))
        self.U = {"rho": rho, "rhou": zero.copy(), "rhov": zero.copy(), "Et": Et}
        self.V = {"u": zero.copy(), "v": zero.copy(), "p": zero.copy(), "tau": zero.copy()}
        self.update_vars()

    @staticmethod
    def _extrap(a):
        # pyrandaBC.extrapolate, order 2, on x1, xn, y1, yn in that order
        a[0, :, :] = 2 * a[1, :, :] - a[2, :, :]
        a[-1, :, :] = 2 * a[-2, :, :] - a[-3, :, :]
        a[:, 0, :] = 2 * a[:, 1, :] - a[:, 2, :]
        a[:, -1, :] = 2 * a[:, -2, :] - a[:, -3, :]

    @staticmethod
    def _const(a, val):
        a[0, :, :] = val
        a[-1, :, :] = val
        a[:, 0, :] = val
        a[:, -1, :] = val

    def update_vars(self):
        P, U, V = self.P, self.U, self.V
        for k in ("rho", "rhou", "rhov", "Et"):
            U[k] = P.sfilter(U[k])
        rho = U["rho"]
        u = U["rhou"] / rho
        v = U["rhov"] / rho
        p = (U["Et"] - .5 * rho * (u * u + v * v)) * (self.gamma - 1.0)
        div = P.ddx(u) + P.ddy(v)
        beta = P.gfilter(np.abs(P.ring(div))) * rho * 7.0e-2
        tau = beta * div
        V.update(u=np.asfortranarray(u), v=np.asfortranarray(v), p=p, div=div, beta=beta, tau=tau)
        # bc.extrap(['rho','Et'], ['x1','xn','y1','yn']); bc.const(['u','v'], ..., 0.0)
        for k in ("rho", "Et"):
            self._extrap(U[k])
        for k in ("u", "v"):
            self._const(V[k], 0.0)

    def fluxes(self):
        P, U, V = self.P, self.U, self.V
        u, v, p, tau = V["u"], V["v"], V["p"], V["tau"]
        return {
            "rho": -P.ddx(U["rho"] * u) - P.ddy(U["rho"] * v),
            "rhou": -P.ddx(U["rhou"] * u + p - tau) - P.ddy(U["rhou"] * v),
            "rhov": -P.ddx(U["rhov"] * u) - P.ddy(U["rhov"] * v + p - tau),
            "Et": -P.ddx((U["Et"] + p - tau) * u) - P.ddy((U["Et"] + p - tau) * v),
        }

    def rk4(self, time, dt):
        PHI = {k: np.zeros_like(a) for k, a in self.U.items()}
        for ii in range(5):
            F = self.fluxes()
            for k in self.U:
                PHI[k] = dt * F[k] + RK_A[ii] * PHI[k]
                self.U[k] = np.asfortranarray(self.U[k] + RK_B[ii] * PHI[k])
            self.update_vars()
        return time + dt

    def run_case(self):
        """The euler.py test loop; returns (x, rho(x, ny/2)) at t = L/8."""
        n = self.n
        dt_max = 1.0 / n * 0.75
        tt = 2.0 * np.pi / 1.0 * .125
        time, dt = 0.0, dt_max
        while tt > time:
            time = self.rk4(time, dt)
            dt = min(dt_max, (tt - time))
        return self.x, self.U["rho"][:, n // 2, 0].copy()
