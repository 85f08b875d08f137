"""CPU restatement of the Taylor-Green Euler+LAD step (TEST INFRASTRUCTURE).

The EOM of examples/TaylorGreen.py:40-86 written out by hand in numpy, with
the compact operators from the C oracle and the 5-stage RK4 of
pyranda/pyranda.py:758-810. Used as the CPU baseline of bench.py and as an
independent check of the GPU front end's EOM path. FP64 throughout.
"""
import numpy as np

from . import pyoracle as po

RK_A = [0.0, -6234157559845. / 12983515589748., -6194124222391. / 4410992767914.,
        -31623096876824. / 15682348800105., -12251185447671. / 11596622555746.]
RK_B = [494393426753. / 4806282396855., 4047970641027. / 5463924506627.,
        9795748752853. / 13190207949281., 4009051133189. / 8539092990294.,
        1348533437543. / 7166442652324.]
RK_ETA = [494393426753. / 4806282396855., 4702696611523. / 9636871101405.,
          3614488396635. / 5249666457482., 9766892798963. / 10823461281321., 1.0]


def _axis(x1, xn, n):
    du = (xn - x1) / max(n - 1, 1)
    x1f, xnf = x1 - du / 2.0, xn + du / 2.0
    d = (xnf - x1f) / n
    return x1f + ((2.0 * np.arange(1, n + 1) - 1.0) * 0.5) * d, d


class TaylorGreenCPU:

    def __init__(self, n, p0=100.0, threads=1):
        L = 2 * np.pi * (n - 1) / n
        x, d = _axis(0.0, L, n)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        self.n = n
        self.d = d
        self.P = po.Oracle((n, n, n), (d, d, d), (True, True, True), threads=threads)
        self.gamma = 1.4
        F = lambda a: np.asfortranarray(a)
        u = F(np.sin(X) * np.cos(Y) * np.cos(Z))
        v = F(-np.cos(X) * np.sin(Y) * np.cos(Z))
        w = F(0.0 * u)
        p = F(p0 + 1.0 / 16.0 * ((np.cos(2. * X) + np.cos(2. * Y)) * (np.cos(2. * Z) + 2.0) - 2.0))
        rho = F(1.0 + 0.0 * u)
        self.U = {"rho": rho, "rhou": rho * u, "rhov": rho * v, "rhow": rho * w,
                  "Et": p / (self.gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)}
        self.V = {}
        self.update_vars()

    def courant(self, u, v, w, c):
        d = self.d
        vrate = np.abs(u) / d + np.abs(v) / d + np.abs(w) / d
        return 1.0 / np.max(vrate + np.abs(c) / d)

    def diff(self, bulk, rho):
        return np.min(rho * self.d * self.d / np.maximum(1.0e-12, bulk))

    def update_vars(self):
        P, U, V = self.P, self.U, self.V
        for k in ("rho", "rhou", "rhov", "rhow", "Et"):
            U[k] = P.sfilter(U[k])
        rho = U["rho"]
        u, v, w = U["rhou"] / rho, U["rhov"] / rho, U["rhow"] / rho
        p = (U["Et"] - .5 * rho * (u * u + v * v + w * w)) * (self.gamma - 1.0)
        ux, vy, wz = P.ddx(u), P.ddy(v), P.ddz(w)
        div = ux + vy + wz
        uy, uz, vx, vz, wy, wx = P.ddy(u), P.ddz(u), P.ddx(v), P.ddz(v), P.ddy(w), P.ddx(w)
        V["enst"] = np.sqrt((uy - vx) ** 2 + (uz - wx) ** 2 + (vz - wy) ** 2)
        V["tke"] = rho * (u * u + v * v + w * w)
        S = np.sqrt(ux * ux + vy * vy + wz * wz + .5 * ((uy + vx) ** 2 + (uz + wx) ** 2 + (vz + wy) ** 2))
        mu = P.gfilter(np.abs(P.ring(S))) * rho * 1.0e-4
        beta = P.gfilter(np.abs(P.ring(div)) * rho) * 7.0e-3
        taudia = (beta - 2. / 3. * mu) * div - p
        V.update(u=u, v=v, w=w, p=p, mu=mu, beta=beta,
                 tauxx=2.0 * mu * ux + taudia, tauyy=2.0 * mu * vy + taudia, tauzz=2.0 * mu * wz + taudia,
                 tauxy=mu * (uy + vx), tauxz=mu * (uz + wx), tauyz=mu * (vz + wz))
        cs = np.sqrt(p / rho * self.gamma)
        dt = self.courant(u, v, w, cs) * 1.0
        dt = np.minimum(dt, 0.2 * self.diff(beta, rho))
        dt = np.minimum(dt, 0.2 * self.diff(mu, rho))
        V["dt"] = float(dt)

    def fluxes(self):
        P, U, V = self.P, self.U, self.V
        u, v, w = V["u"], V["v"], V["w"]
        txx, tyy, tzz, txy, txz, tyz = (V[k] for k in ("tauxx", "tauyy", "tauzz", "tauxy", "tauxz", "tauyz"))
        return {
            "rho": -P.ddx(U["rho"] * u) - P.ddy(U["rho"] * v) - P.ddz(U["rho"] * w),
            "rhou": -P.ddx(U["rhou"] * u - txx) - P.ddy(U["rhou"] * v - txy) - P.ddz(U["rhou"] * w - txz),
            "rhov": -P.ddx(U["rhov"] * u - txy) - P.ddy(U["rhov"] * v - tyy) - P.ddz(U["rhov"] * w - tyz),
            "rhow": -P.ddx(U["rhow"] * u - txz) - P.ddy(U["rhow"] * v - tyz) - P.ddz(U["rhow"] * w - tzz),
            "Et": -P.ddx((U["Et"] - txx) * u - txy * v - txz * w)
                  - P.ddy((U["Et"] - tyy) * v - txy * u - tyz * w)
                  - P.ddz((U["Et"] - tzz) * w - txz * u - tyz * v),
        }

    def rk4(self, time, dt):
        PHI = {k: np.zeros_like(a) for k, a in self.U.items()}
        for ii in range(5):
            F = self.fluxes()
            for k in self.U:
                PHI[k] = dt * F[k] + RK_A[ii] * PHI[k]
                self.U[k] = self.U[k] + RK_B[ii] * PHI[k]
            self.update_vars()
        return time + dt
