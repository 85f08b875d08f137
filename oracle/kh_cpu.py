"""CPU restatement of the 2-D Kelvin-Helmholtz regression case (TEST INFRASTRUCTURE).

examples/KH.py (McNally et al. shear layers, the script behind the
reference's KH-2d-64 baseline, tests/cases/test2deuler.py:42-45 with
arguments 64 0 KH-2d-64 1): doubly periodic 2-D grid, gamma = 5/3, LAD
shear and bulk viscosity and thermal diffusivity (ring + gbar), the
conservative filter after every stage, dt from dt.courant / dt.diff
(pyranda/pyrandaTimestep.py:42-76) re-evaluated in every stage, and the
kappa line reading the dt variable left by the previous evaluation. numpy
with the C oracle's operators and the 5-stage RK4 of
pyranda/pyranda.py:758-810; every expression keeps the operand order of the
EOM strings (KH.py:58-97, 101-118). FP64 throughout. Only tests import it.
"""
import numpy as np

from . import pyoracle as po
from .tg_cpu import RK_A, RK_B, _axis


class KelvinHelmholtzCPU:

    def __init__(self, n, vmean=0.0):
        gamma, R = 5. / 3., 1.0
        cp = R / (1.0 - 1.0 / gamma)
        self.gamma, self.R, self.cv = gamma, R, cp - R
        Lp = 1.0 * (n - 1.0) / n
        x, dx = _axis(0.0, Lp, n)
        y, dy = _axis(0.0, Lp, n)
        X, Y = np.meshgrid(x, y, indexing="ij")
        X, Y = X[:, :, None], Y[:, :, None]
        self.n, self.x = n, x
        # z is the size-1 axis of defaultMeshOptions (0..1, one point): a
        # null direction; pyrandaMesh's spacing rule for nz = 1 (mesh.f90:190-211)
        self.dx, self.dy = dx, dy
        self.dz = max(dx, dy)
        self.gridlen = min(dx, dy, self.dz)
        self.P = po.Oracle((n, n, 1), (dx, dy, 1.0), (True, True, False))
        F = np.asfortranarray
        u1, u2, p0, rho1, rho2, L = 0.5, -0.5, 2.5, 1.0, 2.0, 0.025
        # IC lines of KH.py:101-118, in order
        Um = (u1 - u2) / 2.0
        rhoM = (rho1 - rho2) / 2.0
        u = u1 - Um * np.exp(-(Y - .75) / L)
        u = np.where(Y < .75, u2 + Um * np.exp(-(.75 - Y) / L), u)
        u = np.where(Y < .50, u2 + Um * np.exp((-Y + .25) / L), u)
        u = np.where(Y < .25, u1 - Um * np.exp((Y - .25) / L), u)
        v = 0.01 * np.sin(4. * np.pi * X) + vmean
        rho = rho1 - rhoM * np.exp(-(Y - .75) / L)
        rho = np.where(Y < .75, rho2 + rhoM * np.exp(-(.75 - Y) / L), rho)
        rho = np.where(Y < .50, rho2 + rhoM * np.exp((-Y + .25) / L), rho)
        rho = np.where(Y < .25, rho1 - rhoM * np.exp((Y - .25) / L), rho)
        p = np.zeros_like(u)
        p += p0
        Et = p / (gamma - 1.0) + .5 * rho * (u * u + v * v)
        self.U = {"rho": F(rho), "rhou": F(rho * u), "rhov": F(rho * v), "Et": F(Et)}
        cs = np.sqrt(p / rho * gamma)
        w = np.zeros_like(u)
        self.V = {"dt": self.courant(u, v, w, cs) * .1}
        self.update_vars()

    def courant(self, u, v, w, c):
        vrate = np.abs(u) / self.dx + np.abs(v) / self.dy + np.abs(w) / self.dz
        crate = np.abs(c) / self.gridlen
        return 1.0 / np.max(vrate + crate)

    def diff(self, bulk, density):
        delta = self.gridlen
        return np.min(density * delta * delta / np.maximum(1.0e-12, bulk))

    def update_vars(self):
        P, U, V = self.P, self.U, self.V
        gamma, cv = self.gamma, self.cv
        for k in ("rho", "rhou", "rhov", "Et"):
            U[k] = P.sfilter(U[k])
        rho = U["rho"]
        u = U["rhou"] / rho
        v = U["rhov"] / rho
        w = np.zeros_like(u)
        p = (U["Et"] - .5 * rho * (u * u + v * v)) * (gamma - 1.0)
        ux = P.ddx(u)
        vy = P.ddy(v)
        div = ux + vy
        uy = P.ddy(u)
        vx = P.ddx(v)
        mu = P.gfilter(np.abs(P.ring(np.sqrt(ux * ux + vy * vy + .5 * ((uy + vx) ** 2))))) * rho * 1.0e-4
        beta = P.gfilter(np.abs(P.ring(div)) * rho) * 7.0e-3
        taudia = (beta - 2. / 3. * mu) * div - p
        T = p / (rho * self.R)
        tx, ty, _ = P.grads(T)
        kappa = P.gfilter(P.ring(T) * rho * cv / (T * V["dt"])) * 1.0e-3
        cs = np.sqrt(p / rho * gamma)
        dt = self.courant(u, v, w, cs) * 1.0
        dt = np.minimum(dt, 0.2 * self.diff(beta, rho))
        dt = np.minimum(dt, 0.2 * self.diff(mu, rho))
        V.update(u=u, v=v, p=p, kappa=kappa, tx=tx, ty=ty,
                 tauxx=2.0 * mu * ux + taudia, tauyy=2.0 * mu * vy + taudia,
                 tauxy=mu * (uy + vx), dt=dt)

    def fluxes(self):
        P, U, V = self.P, self.U, self.V
        u, v, kappa = V["u"], V["v"], V["kappa"]
        txx, tyy, txy = V["tauxx"], V["tauyy"], V["tauxy"]
        return {
            "rho": -P.ddx(U["rho"] * u) - P.ddy(U["rho"] * v),
            "rhou": -P.ddx(U["rhou"] * u - txx) - P.ddy(U["rhou"] * v - txy),
            "rhov": -P.ddx(U["rhov"] * u - txy) - P.ddy(U["rhov"] * v - tyy),
            "Et": -P.ddx((U["Et"] - txx) * u - txy * v - V["tx"] * kappa)
                  - P.ddy((U["Et"] - tyy) * v - txy * u - V["ty"] * kappa),
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

    def run_case(self, t_final=1.5, dt_max=1.0, max_steps=None):
        """The KH.py test loop; returns (x, rho(x, 3 ny/4)) at t_final."""
        time = 0.0
        dt = self.V["dt"]
        steps = 0
        while t_final > time:
            time = self.rk4(time, dt)
            dt = min(self.V["dt"], 1.1 * dt)
            dt = min(dt_max, dt)
            dt = min(dt, (t_final - time))
            steps += 1
            if max_steps is not None and steps >= max_steps:
                break
        self.steps = steps
        return self.x, self.U["rho"][:, 3 * self.n // 4, 0].copy()
