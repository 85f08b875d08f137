"""CPU restatement of the Taylor-Green Euler+LAD step (TEST INFRASTRUCTURE).

The EOM of examples/TaylorGreen.py:40-86 written out by hand in numpy, with
the compact operators from the C oracle and the 5-stage RK4 of
pyranda/pyranda.py:758-810. Used as the CPU baseline of bench.py and as an
independent check of the GPU front end's EOM path (the reference-held
taylorgreen value, tests/test_oracle.py). FP64 throughout.

Threads: the oracle's operators run their lines on `threads` OpenMP threads;
the pointwise algebra between them runs on the same number of Python threads,
each over a slab of z-planes (numpy releases the GIL inside its loops). Every
value is computed by the same operations in the same order as a serial
evaluation, so the result does not depend on the thread count.
"""
from concurrent.futures import ThreadPoolExecutor

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
        self.threads = max(1, int(threads))
        self._pool = ThreadPoolExecutor(self.threads) if self.threads > 1 else None
        nz = n
        cuts = np.linspace(0, nz, min(self.threads, nz) + 1).astype(int)
        self._slabs = [slice(int(a), int(b)) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]
        self.P = po.Oracle((n, n, n), (d, d, d), (True, True, True), threads=self.threads)
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

    def close(self):
        if self._pool is not None:
            self._pool.shutdown()
            self._pool = None

    # -- pointwise algebra over z-slabs ---------------------------------------
    def pw(self, fn, *fields, nout=1):
        """Outputs of fn(*slabs) for every z-slab of the fields (Fortran order
        (x, y, z) arrays), assembled into new fields; fn returns one array or
        a tuple of nout arrays of the slab's shape."""
        shape = fields[0].shape
        outs = [np.empty(shape, order="F") for _ in range(nout)]

        def run(sl):
            r = fn(*[f[:, :, sl] for f in fields])
            r = r if nout > 1 else (r,)
            for o, a in zip(outs, r):
                o[:, :, sl] = a

        if self._pool is None:
            run(slice(None))
        else:
            list(self._pool.map(run, self._slabs))
        return tuple(outs) if nout > 1 else outs[0]

    def pmax(self, fn, *fields):
        """max over the grid of fn(*slabs) (slab maxima, then their max)."""
        if self._pool is None:
            return np.max(fn(*fields))
        return max(self._pool.map(lambda sl: np.max(fn(*[f[:, :, sl] for f in fields])), self._slabs))

    def pmin(self, fn, *fields):
        if self._pool is None:
            return np.min(fn(*fields))
        return min(self._pool.map(lambda sl: np.min(fn(*[f[:, :, sl] for f in fields])), self._slabs))

    def courant(self, u, v, w, c):
        d = self.d
        return 1.0 / self.pmax(lambda u, v, w, c: (np.abs(u) / d + np.abs(v) / d + np.abs(w) / d) + np.abs(c) / d,
                               u, v, w, c)

    def diff(self, bulk, rho):
        d = self.d
        return self.pmin(lambda b, r: r * d * d / np.maximum(1.0e-12, b), bulk, rho)

    def update_vars(self):
        P, U, V, g = self.P, self.U, self.V, self.gamma
        for k in ("rho", "rhou", "rhov", "rhow", "Et"):
            U[k] = P.sfilter(U[k])
        rho = U["rho"]

        def prim(rho, ru, rv, rw, Et):
            u, v, w = ru / rho, rv / rho, rw / rho
            p = (Et - .5 * rho * (u * u + v * v + w * w)) * (g - 1.0)
            return u, v, w, p

        u, v, w, p = self.pw(prim, rho, U["rhou"], U["rhov"], U["rhow"], U["Et"], nout=4)
        ux, vy, wz = P.ddx(u), P.ddy(v), P.ddz(w)
        div = self.pw(lambda a, b, c: a + b + c, ux, vy, wz)
        uy, uz, vx, vz, wy, wx = P.ddy(u), P.ddz(u), P.ddx(v), P.ddz(v), P.ddy(w), P.ddx(w)
        V["enst"] = self.pw(lambda uy, vx, uz, wx, vz, wy: np.sqrt((uy - vx) ** 2 + (uz - wx) ** 2 + (vz - wy) ** 2),
                            uy, vx, uz, wx, vz, wy)
        V["tke"] = self.pw(lambda r, u, v, w: r * (u * u + v * v + w * w), rho, u, v, w)
        S = self.pw(lambda ux, vy, wz, uy, vx, uz, wx, vz, wy:
                    np.sqrt(ux * ux + vy * vy + wz * wz + .5 * ((uy + vx) ** 2 + (uz + wx) ** 2 + (vz + wy) ** 2)),
                    ux, vy, wz, uy, vx, uz, wx, vz, wy)
        mu = self.pw(lambda a, r: a * r * 1.0e-4, P.gfilter(self.pw(np.abs, P.ring(S))), rho)
        beta = self.pw(lambda a: a * 7.0e-3,
                       P.gfilter(self.pw(lambda a, r: np.abs(a) * r, P.ring(div), rho)))

        def stress_rho(beta, mu, div, p, ux, vy, wz, uy, vx, uz, wx, vz, r):
            taudia = (beta - 2. / 3. * mu) * div - p
            return (2.0 * mu * ux + taudia, 2.0 * mu * vy + taudia, 2.0 * mu * wz + taudia,
                    mu * (uy + vx), mu * (uz + wx), mu * (vz + wz), np.sqrt(p / r * g))

        txx, tyy, tzz, txy, txz, tyz, cs = self.pw(stress_rho, beta, mu, div, p, ux, vy, wz, uy, vx, uz, wx,
                                                   vz, rho, nout=7)
        V.update(u=u, v=v, w=w, p=p, mu=mu, beta=beta, tauxx=txx, tauyy=tyy, tauzz=tzz,
                 tauxy=txy, tauxz=txz, tauyz=tyz)
        dt = self.courant(u, v, w, cs) * 1.0
        dt = np.minimum(dt, 0.2 * self.diff(beta, rho))
        dt = np.minimum(dt, 0.2 * self.diff(mu, rho))
        V["dt"] = float(dt)

    def fluxes(self):
        P, U, V = self.P, self.U, self.V
        u, v, w = V["u"], V["v"], V["w"]
        txx, tyy, tzz, txy, txz, tyz = (V[k] for k in ("tauxx", "tauyy", "tauzz", "tauxy", "tauxz", "tauyz"))
        pw = self.pw

        def rhs(a, b, c):
            return pw(lambda a, b, c: -a - b - c, a, b, c)

        return {
            "rho": rhs(P.ddx(pw(lambda r, u: r * u, U["rho"], u)), P.ddy(pw(lambda r, v: r * v, U["rho"], v)),
                       P.ddz(pw(lambda r, w: r * w, U["rho"], w))),
            "rhou": rhs(P.ddx(pw(lambda a, u, t: a * u - t, U["rhou"], u, txx)),
                        P.ddy(pw(lambda a, v, t: a * v - t, U["rhou"], v, txy)),
                        P.ddz(pw(lambda a, w, t: a * w - t, U["rhou"], w, txz))),
            "rhov": rhs(P.ddx(pw(lambda a, u, t: a * u - t, U["rhov"], u, txy)),
                        P.ddy(pw(lambda a, v, t: a * v - t, U["rhov"], v, tyy)),
                        P.ddz(pw(lambda a, w, t: a * w - t, U["rhov"], w, tyz))),
            "rhow": rhs(P.ddx(pw(lambda a, u, t: a * u - t, U["rhow"], u, txz)),
                        P.ddy(pw(lambda a, v, t: a * v - t, U["rhow"], v, tyz)),
                        P.ddz(pw(lambda a, w, t: a * w - t, U["rhow"], w, tzz))),
            "Et": rhs(P.ddx(pw(lambda E, txx, u, txy, v, txz, w: (E - txx) * u - txy * v - txz * w,
                               U["Et"], txx, u, txy, v, txz, w)),
                      P.ddy(pw(lambda E, tyy, v, txy, u, tyz, w: (E - tyy) * v - txy * u - tyz * w,
                               U["Et"], tyy, v, txy, u, tyz, w)),
                      P.ddz(pw(lambda E, tzz, w, txz, u, tyz, v: (E - tzz) * w - txz * u - tyz * v,
                               U["Et"], tzz, w, txz, u, tyz, v))),
        }

    def rk4(self, time, dt):
        PHI = {k: np.zeros_like(a) for k, a in self.U.items()}
        for ii in range(5):
            F = self.fluxes()
            A, B = RK_A[ii], RK_B[ii]
            for k in self.U:
                PHI[k], self.U[k] = self.pw(lambda f, ph, uu: ((lambda q: (q, uu + B * q))(dt * f + A * ph)),
                                            F[k], PHI[k], self.U[k], nout=2)
            self.update_vars()
        return time + dt
