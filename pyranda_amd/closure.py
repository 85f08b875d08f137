"""Learned stand-in for the LAD artificial bulk viscosity and species
diffusivity, on BF16 matrix cores, with the plain LAD chain as judge and
fallback.

The plain LAD chain (examples/RM3D.py:95-104) computes, per RK stage,
    beta  = 7e-3 * gbar(|ring(div)| * rho)
    adiff = gbar(rho * max(1e-4 ring(Yh), 1e2 (|Yh| - 1 + |1 - Yh|) gridLen^2) / dt)
with ring = max over directions of |d8|-scaled eighth derivatives: three
compact solves plus three Gaussian sweeps per quantity. The stand-in is a
pointwise MLP (pyr_mlp_infer, csrc/pyr_mlp.hip) over the 3x3 neighbourhoods
of dilatation, vorticity and |grad rho| on a 2-D grid, trained on fields the
plain chain generates (pyr_mlp_train_grad gives the gradient of the mean
squared error on the matrix cores; Adam on the FP32 master weights).

`pyrandaLearnedLAD` exposes it to EOM strings as
    [:beta:,:adiff:] = lad.closure(:div:,:vx:,:uy:,:rho:,:Yh:,:dt:)
with the plain chain as fallback, at two levels:
  * per point: the inference kernel flags every point whose normalised
    inputs reach beyond `sensor_factor` times the largest seen in training
    (inputs the network was never fitted on); when any point is flagged the
    plain chain is evaluated and its values replace the closure's at the
    flagged points (one host read of the flagged count per evaluation);
  * per run: the first evaluation (`judge_first`) and every `judge_every`-th
    one are judged: the closure is compared with the plain chain on the
    current state, and a relative L2 error above `judge_tol` in beta or
    adiff switches the closure off for the rest of the run (every later
    evaluation is the plain chain).
An untrained closure (no training range) falls back on every evaluation.
The network's second output is adiff * dt (the chain's value before its
division by the time step, which is not an input); the closure divides by
the current dt. `judge()` reports the closure error against the plain chain
on the current state.

No counterpart in the reference.
"""
import ctypes
import math

import torch

from . import _lib
from .pyrandaPackage import pyrandaPackage

NIN, NH, NO, NF = 32, 64, 16, 27
TRAIN_BLOCKS = 512


class MLPClosure:
    """Weights (FP32 masters), normalisation, the packed BF16 parameter block
    the kernels read, and Adam state."""

    def __init__(self, device, seed=0):
        self.device = torch.device(device)
        g = torch.Generator().manual_seed(seed)
        W1 = torch.randn(NIN, NH, generator=g) * math.sqrt(2.0 / NF)
        W1[NF:] = 0.0
        W2 = torch.randn(NH, NH, generator=g) * math.sqrt(2.0 / NH)
        W3 = torch.zeros(NH, NO)
        W3[:, :2] = torch.randn(NH, 2, generator=g) * math.sqrt(1.0 / NH)
        dev = self.device
        self.params = [t.to(dev) for t in (W1, W2, W3, torch.zeros(NH), torch.zeros(NH), torch.zeros(NO))]
        self.fshift = torch.zeros(4, device=dev)
        self.fscale = torch.ones(4, device=dev)
        self.oscale = torch.ones(4, device=dev)
        self.fmax = None          # largest normalised |feature| seen in training
        self.blob = torch.empty(_lib.MLP_PARAM_BYTES, dtype=torch.uint8, device=dev)
        self._m = [torch.zeros_like(p) for p in self.params]
        self._v = [torch.zeros_like(p) for p in self.params]
        self._t = 0
        self._partial = None
        self._grad = torch.empty(_lib.MLP_GRAD_FLOATS, dtype=torch.float32, device=dev)
        self.pack()

    # -- parameter block (layout of include/pyranda_amd.h) --------------------
    def pack(self):
        W1, W2, W3, b1, b2, b3 = self.params
        bf = torch.cat([W1.reshape(-1), W2.reshape(-1), W3.reshape(-1)]).to(torch.bfloat16)
        off = _lib.MLP_PARAM_FLOAT_OFFSET
        self.blob[:off].copy_(bf.view(torch.uint8))
        fl = torch.cat([b1, b2, b3, self.fshift, self.fscale, self.oscale]).to(torch.float32)
        self.blob[off:].copy_(fl.view(torch.uint8))

    def set_normalisation(self, feats, targets):
        """Feature shift/scale (mean, 1/std over the samples) and output
        scales (max |target|) from training samples."""
        for f in range(3):
            x = torch.cat([s[f].reshape(-1).double() for s in feats])
            self.fshift[f] = x.mean().float()
            sd = x.std().item()
            self.fscale[f] = 1.0 / sd if sd > 0 else 1.0
        for j in range(2):
            m = max(float(t[j].abs().max()) for t in targets)
            self.oscale[j] = m if m > 0 else 1.0
        self.fmax = max(float(((s[f].double() - float(self.fshift[f])) * float(self.fscale[f])).abs().max())
                        for s in feats for f in range(3))
        self.pack()

    # -- kernels -----------------------------------------------------------------
    @staticmethod
    def _prec(t):
        return _lib.F64 if t.dtype == torch.float64 else _lib.R32

    @staticmethod
    def _ptrs(ts):
        return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    @staticmethod
    def _contig(ts):
        """The fields in the layout and one precision the kernels read (FP64
        if any of them is FP64: an EOM can mix FP32 fields with FP64 ones
        derived from the mesh coordinates)."""
        dt = torch.float64 if any(t.dtype == torch.float64 for t in ts) else torch.float32
        out = []
        for t in ts:
            t = t.to(dt)
            out.append(t if t.permute(2, 1, 0).is_contiguous() else t.permute(2, 1, 0).contiguous().permute(2, 1, 0))
        return out

    def infer(self, feats, periodic, sensor=0.0):
        """(out0, out1) fields from three feature fields of shape (nx, ny, 1).
        With sensor > 0 also (mask, count): mask (uint8, the fields' shape)
        is 1 at points whose largest normalised |input| exceeds `sensor`,
        count (a 1-element int32 device tensor) the number of such points."""
        feats = self._contig(feats)
        nx, ny, nz = feats[0].shape
        if nz != 1:
            raise ValueError("the learned closure is built for 2-D grids (nz == 1)")
        outs = [torch.empty_like(feats[0]) for _ in range(2)]
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        args = (stream, self._prec(feats[0]), nx, ny, int(periodic[0]), int(periodic[1]), self._ptrs(feats),
                ctypes.c_void_p(self.blob.data_ptr()), self._ptrs(outs))
        if sensor <= 0.0:
            _lib.check(_lib.lib().pyr_mlp_infer(*args))
            return outs
        mask = torch.empty((1, ny, nx), dtype=torch.uint8, device=feats[0].device).permute(2, 1, 0)
        count = torch.zeros(1, dtype=torch.int32, device=feats[0].device)
        _lib.check(_lib.lib().pyr_mlp_infer_sensed(*args, ctypes.c_float(sensor), ctypes.c_void_p(mask.data_ptr()),
                                                   ctypes.c_void_p(count.data_ptr())))
        return outs, mask, count

    def gradient(self, feats, targets, periodic):
        """(gradient dict, mean squared error) of one sample on the MFMA kernels."""
        both = self._contig(list(feats) + list(targets))
        feats, targets = both[:3], both[3:]
        nx, ny, nz = feats[0].shape
        if self._partial is None:
            self._partial = torch.empty(TRAIN_BLOCKS * _lib.MLP_GRAD_FLOATS, dtype=torch.float32,
                                        device=self.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(_lib.lib().pyr_mlp_train_grad(
            stream, self._prec(feats[0]), nx, ny, int(periodic[0]), int(periodic[1]), self._ptrs(feats),
            self._ptrs(targets), ctypes.c_void_p(self.blob.data_ptr()), ctypes.c_void_p(self._partial.data_ptr()),
            TRAIN_BLOCKS, ctypes.c_void_p(self._grad.data_ptr())))
        return unpack_grad(self._grad), self._grad[-1] / (nx * ny)

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self._t += 1
        for p, g, m, v in zip(self.params, grads, self._m, self._v):
            m.mul_(b1).add_(g, alpha=1 - b1)
            v.mul_(b2).addcmul_(g, g, value=1 - b2)
            mh = m / (1 - b1 ** self._t)
            vh = v / (1 - b2 ** self._t)
            p.sub_(lr * mh / (vh.sqrt() + eps))
        self.pack()

    def train(self, samples, periodic, steps=200, lr=2e-3):
        """Adam over the samples [(feats, targets)], one sample per step in
        turn; returns the mean squared error of every step (device tensor)."""
        losses = []
        for it in range(steps):
            feats, targets = samples[it % len(samples)]
            grads, mse = self.gradient(feats, targets, periodic)
            losses.append(mse)
            self.adam_step(grads, lr=lr)
        return torch.stack(losses)


def unpack_grad(g):
    """Split the PYR_MLP_GRAD_FLOATS block into (gW1, gW2, gW3, gb1, gb2, gb3)."""
    sizes = [(NIN, NH), (NH, NH), (NH, NO), (NH,), (NH,), (NO,)]
    out, off = [], 0
    for s in sizes:
        n = int(torch.tensor(s).prod())
        out.append(g[off:off + n].reshape(s))
        off += n
    return out


# the plain chain of examples/RM3D.py:100-104, evaluated by the fallback and
# the judge exactly as the EOM evaluates it (same fused kernels)
RM_PLAIN_LAD = (":Dsgs:      =  ring(:Yh:) * 1.0e-4",
                ":Ysgs:      =  1.0e2*(abs(:Yh:) - 1.0 + abs(1.0-:Yh: ) )*gridLen**2",
                ":beta:      =  gbar( abs(ring(:div:)) * :rho: )  * 7.0e-3",
                ":adiff:     =  gbar( :rho:*numpy.maximum(:Dsgs:,:Ysgs:) / :dt: )")


class pyrandaLearnedLAD(pyrandaPackage):
    """`lad.closure(div, vx, uy, rho, Yh, dt)` for EOM strings: the learned
    stand-in for beta and adiff, with the plain chain as fallback. The
    arguments are the simulation's own fields; the plain chain reads them
    through its EOM lines (`plain_lines`, written to the same variables)."""

    def __init__(self, pysim, model, sensor_factor=1.5, plain_lines=RM_PLAIN_LAD, judge_every=0, judge_tol=0.5,
                 judge_first=True):
        pyrandaPackage.__init__(self, "LearnedLAD", pysim)
        self.pysim = pysim
        self.model = model
        self.sensor_factor = sensor_factor
        self.plain_lines = plain_lines
        self.judge_every = int(judge_every)
        self.judge_tol = float(judge_tol)
        self.judge_first = bool(judge_first)
        self._plain_eqs = None
        self.enabled = True
        self.calls = 0
        self.fallbacks = 0          # evaluations that were the plain chain throughout
        self.partial = 0            # evaluations with flagged points
        self.points_flagged = 0
        self.judged = []            # (call, {name: (rel_l2, rel_max)})
        p = pysim.PyMPI.periodic
        self.periodic = (bool(p[0]), bool(p[1]))

    def get_sMap(self):
        self.sMap = {"lad.closure(": "self.packages['LearnedLAD'].closure("}

    def features(self, div, vx, uy, rho):
        s = self.pysim
        gx, gy = s.ddx(rho), s.ddy(rho)
        return [div, vx - uy, torch.sqrt(gx * gx + gy * gy)]

    def sample(self):
        """One training sample from the simulation's current fields:
        (features, [beta, adiff * dt]). The network learns adiff * dt, the
        chain's value before its division by the time step (dt is not an
        input and changes along a run); closure() divides by the current dt."""
        s = self.pysim
        v = {k: s.var(k).data for k in ("div", "vx", "uy", "rho", "beta", "adiff", "dt")}
        feats = [f.clone() for f in self.features(v["div"], v["vx"], v["uy"], v["rho"])]
        return feats, [v["beta"].clone(), v["adiff"] * v["dt"]]

    def plain(self):
        """The plain LAD chain on the simulation's current fields: (beta, adiff)."""
        from .pyrandaEq import pyrandaEq
        s = self.pysim
        if self._plain_eqs is None:
            self._plain_eqs = [pyrandaEq(l, s.sMap, s.PyMPI) for l in self.plain_lines]
        out = {}
        for eq in self._plain_eqs:
            v = eq.RHS(s)
            name = eq.LHS[0]
            if name in ("beta", "adiff"):
                out[name] = v
            else:
                if name not in s.variables:     # lines the closure's EOM left out
                    s.addVar(name, kind="conserved")
                s.variables[name].data = v
        return out["beta"], out["adiff"]

    def closure(self, div, vx, uy, rho, Yh, dt):
        self.calls += 1
        m = self.model
        if not self.enabled or m is None or m.fmax is None:
            self.fallbacks += 1
            return list(self.plain())
        feats = self.features(div, vx, uy, rho)
        first = self.judge_first and not self.judged
        if first or (self.judge_every and self.calls % self.judge_every == 0):
            rep = self.judge(div, vx, uy, rho, Yh, dt, feats=feats)
            self.judged.append((self.calls, rep))
            if max(rep["beta"][0], rep["adiff"][0]) > self.judge_tol:
                self.enabled = False
                self.fallbacks += 1
                return list(self.plain())
        (beta, adiff_dt), mask, count = m.infer(feats, self.periodic, sensor=self.sensor_factor * m.fmax)
        adiff = adiff_dt / dt
        n = int(count.item())
        if n:
            self.partial += 1
            self.points_flagged += n
            pb, pa = self.plain()
            sel = mask.bool()
            beta = torch.where(sel, pb.to(beta.dtype), beta)
            adiff = torch.where(sel, pa.to(adiff.dtype), adiff)
        return [beta, adiff]

    def judge(self, div, vx, uy, rho, Yh, dt, feats=None, raw=False):
        """Closure error against the plain chain on the same state:
        {name: (rel_l2, rel_max)} for beta and adiff, of what closure()
        returns (the plain values at the points the sensor flags) or, with
        raw=True, of the network alone."""
        if feats is None:
            feats = self.features(div, vx, uy, rho)
        m = self.model
        if raw:
            got = m.infer(feats, self.periodic)
            sel = None
        else:
            got, mask, _ = m.infer(feats, self.periodic, sensor=self.sensor_factor * m.fmax)
            sel = mask.bool()
        got = [got[0], got[1] / dt]
        ref = self.plain()
        out = {}
        for name, g, r in zip(("beta", "adiff"), got, ref):
            r64, g64 = r.double(), g.double()
            if sel is not None:
                g64 = torch.where(sel, r64, g64)
            den = float(r64.norm()) or 1.0
            out[name] = (float((g64 - r64).norm()) / den,
                         float((g64 - r64).abs().max()) / (float(r64.abs().max()) or 1.0))
        return out
