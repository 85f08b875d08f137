"""Learned LAD closure on the GPU: BF16-MFMA kernels against a PyTorch FP32
reference of the same operation, training, and the closure's fallback.

The reference (below) evaluates the MLP with torch FP32 matrix products on
the same BF16-rounded weights, inputs and activations as the kernels; the
kernels accumulate in a different order, so an activation can round to the
neighbouring BF16 value now and then: outputs and gradients are compared
at a few parts in a thousand of their largest magnitude.
"""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu

NIN, NH, NO, NF = 32, 64, 16, 27


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyranda_amd
    return pyranda_amd


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def ref_features(fields, fshift, fscale, periodic):
    nx, ny = fields[0].shape[:2]
    i = torch.arange(nx, device=fields[0].device)
    j = torch.arange(ny, device=fields[0].device)
    cols = []
    for k in range(NIN):
        if k >= NF:
            cols.append(torch.zeros(nx * ny, device=fields[0].device))
            continue
        f, r = divmod(k, 9)
        dx, dy = r % 3 - 1, r // 3 - 1
        ii = (i + dx) % nx if periodic[0] else (i + dx).clamp(0, nx - 1)
        jj = (j + dy) % ny if periodic[1] else (j + dy).clamp(0, ny - 1)
        v = fields[f][:, :, 0].float()[ii][:, jj]
        v = (v - fshift[f]) * fscale[f]
        cols.append(v.t().reshape(-1))          # point p = j * nx + i
    return bf(torch.stack(cols, 1))


def ref_forward(m, X):
    W1, W2, W3, b1, b2, b3 = [p.float() for p in m.params]
    H1 = bf(torch.relu(X @ bf(W1) + b1))
    H2 = bf(torch.relu(H1 @ bf(W2) + b2))
    O = H2 @ bf(W3) + b3
    return H1, H2, O


def ref_grad(m, X, T):
    W1, W2, W3, b1, b2, b3 = [p.float() for p in m.params]
    H1, H2, O = ref_forward(m, X)
    n = X.shape[0]
    e = torch.zeros_like(O)
    e[:, :2] = O[:, :2] - T / m.oscale[:2]
    dO = bf(2.0 * e * (1.0 / n))
    dZ2 = bf((dO @ bf(W3).t()) * (H2 > 0))
    dZ1 = bf((dZ2 @ bf(W2).t()) * (H1 > 0))
    return [X.t() @ dZ1, H1.t() @ dZ2, H2.t() @ dO, dZ1.sum(0), dZ2.sum(0), dO.sum(0)], (e * e).sum()


def _model(pr, seed=3):
    from pyranda_amd.closure import MLPClosure
    m = MLPClosure("cuda", seed=seed)
    g = torch.Generator().manual_seed(seed + 1)
    m.params[3].copy_(0.1 * torch.randn(NH, generator=g))
    m.params[4].copy_(0.1 * torch.randn(NH, generator=g))
    m.params[5][:2].copy_(0.1 * torch.randn(2, generator=g))
    m.fshift[:3] = torch.tensor([0.1, -0.2, 0.3])
    m.fscale[:3] = torch.tensor([2.0, 0.5, 1.5])
    m.oscale[:2] = torch.tensor([3.0, 0.25])
    m.pack()
    return m


def _close(got, ref, frac=4e-3):
    scale = float(ref.abs().max()) or 1.0
    err = float((got.float() - ref.float()).abs().max()) / scale
    return err <= frac, err


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("shape,periodic", [((37, 29), (False, True)), ((64, 48), (True, True)),
                                            ((5, 3), (False, False))])
def test_mlp_infer_matches_torch(pr, dtype, shape, periodic):
    _say("mlp infer", dtype, shape, periodic)
    m = _model(pr)
    nx, ny = shape
    g = torch.Generator().manual_seed(nx * ny)
    fields = [torch.randn(ny, nx, generator=g, dtype=torch.float64).t().reshape(nx, ny, 1).to("cuda", dtype)
              for _ in range(3)]
    out = m.infer(fields, periodic)
    torch.cuda.synchronize()
    X = ref_features(fields, m.fshift, m.fscale, periodic)
    _, _, O = ref_forward(m, X)
    for j in range(2):
        want = (torch.relu(O[:, j]) * m.oscale[j]).reshape(ny, nx).t()
        got = out[j][:, :, 0]
        assert got.dtype == dtype
        ok, err = _close(got, want)
        _say("  output", j, "max err / max", err)
        assert ok, err


@pytest.mark.parametrize("shape,periodic", [((37, 29), (False, True)), ((128, 96), (True, True))])
def test_mlp_gradient_matches_torch(pr, shape, periodic):
    _say("mlp gradient", shape, periodic)
    m = _model(pr)
    nx, ny = shape
    g = torch.Generator().manual_seed(7 + nx)
    fields = [torch.randn(ny, nx, generator=g, dtype=torch.float64).t().reshape(nx, ny, 1).to("cuda", torch.float32)
              for _ in range(3)]
    targets = [torch.rand(ny, nx, generator=g, dtype=torch.float64).t().reshape(nx, ny, 1).to("cuda", torch.float32)
               * s for s in (3.0, 0.25)]
    grads, mse = m.gradient(fields, targets, periodic)
    torch.cuda.synchronize()
    X = ref_features(fields, m.fshift, m.fscale, periodic)
    T = torch.stack([t[:, :, 0].t().reshape(-1) for t in targets], 1)
    want, loss = ref_grad(m, X, T)
    names = ["gW1", "gW2", "gW3", "gb1", "gb2", "gb3"]
    for name, got, ref in zip(names, grads, want):
        ok, err = _close(got, ref, 1e-2)
        _say("  ", name, "max err / max", err)
        assert ok, (name, err)
    assert abs(float(mse) * nx * ny - float(loss)) <= 1e-3 * float(loss)
    # padded inputs and outputs get no gradient
    assert float(grads[0][NF:].abs().max()) == 0.0
    assert float(grads[2][:, 2:].abs().max()) == 0.0


@pytest.mark.parametrize("shape,periodic", [((37, 29), (False, True)), ((96, 64), (True, True))])
def test_mlp_kernels_match_cpu_oracle(pr, shape, periodic):
    """Outputs and gradients of the matrix-core kernels against the numpy
    restatement (oracle/mlp_cpu.py) on the same BF16-rounded operands: within
    a few parts in a thousand of each block's largest magnitude (the kernels
    accumulate in FP32 in another order, so an activation can round to the
    neighbouring BF16 value now and then); the loss to 1e-3."""
    import numpy as np
    from oracle import mlp_cpu as mc
    m = _model(pr)
    nx, ny = shape
    g = torch.Generator().manual_seed(31 + nx)
    fields = [torch.randn(ny, nx, generator=g, dtype=torch.float64).t().reshape(nx, ny, 1).to("cuda", torch.float32)
              for _ in range(3)]
    targets = [torch.rand(ny, nx, generator=g, dtype=torch.float64).t().reshape(nx, ny, 1).to("cuda", torch.float32)
               * s for s in (3.0, 0.25)]
    out = m.infer(fields, periodic)
    grads, mse = m.gradient(fields, targets, periodic)
    torch.cuda.synchronize()
    host = lambda t: t[:, :, 0].double().cpu().numpy()
    params = [p.float().cpu().numpy() for p in m.params]
    x = mc.features([host(f) for f in fields], m.fshift.cpu().numpy(), m.fscale.cpu().numpy(), periodic)
    osc = m.oscale.cpu().numpy()
    want_out = mc.infer(x, params, osc)
    for j in range(2):
        got = out[j][:, :, 0].t().reshape(-1).cpu().numpy()
        err = float(np.abs(got - want_out[j]).max()) / (float(np.abs(want_out[j]).max()) or 1.0)
        _say("  output", j, "max err / max", err)
        assert err <= 4e-3, (j, err)
    tg = [host(t).T.reshape(-1) for t in targets]
    want, loss = mc.gradient(x, tg, params, osc)
    for name, gk, ref in zip(["gW1", "gW2", "gW3", "gb1", "gb2", "gb3"], grads, want):
        err = float(np.abs(gk.cpu().numpy() - ref).max()) / (float(np.abs(ref).max()) or 1.0)
        _say("  ", name, "max err / max", err)
        assert err <= 1e-2, (name, err)
    assert abs(float(mse) * nx * ny - loss) <= 1e-3 * loss


def test_mlp_sensor_flags_points(pr):
    """pyr_mlp_infer_sensed: a point is flagged exactly when one of its 27
    normalised inputs (its 3x3 neighbourhood, wrapped or clamped) exceeds the
    threshold; the count is the number of flagged points and the outputs
    are those of the plain inference."""
    import numpy as np
    from oracle import mlp_cpu as mc
    m = _model(pr)
    nx, ny = 53, 40
    g = torch.Generator().manual_seed(5)
    fields = [torch.randn(ny, nx, generator=g, dtype=torch.float64).t().reshape(nx, ny, 1).to("cuda") for _ in range(3)]
    fields[1][7, 0, 0] = 40.0          # a spike on the wrapped y edge
    fields[2][52, 20, 0] = -60.0       # and on the clamped x edge
    thr = 8.0
    (o0, o1), mask, count = m.infer(fields, (False, True), sensor=thr)
    p0, p1 = m.infer(fields, (False, True))
    torch.cuda.synchronize()
    assert torch.equal(o0, p0) and torch.equal(o1, p1)
    host = [f[:, :, 0].cpu().numpy() for f in fields]
    x = mc.features(host, m.fshift.cpu().numpy(), m.fscale.cpu().numpy(), (False, True))
    # the kernel tests the normalised inputs before BF16 rounding
    xr = np.zeros_like(x)
    for k in range(mc.NF):
        f, r = divmod(k, 9)
        dx, dy = r % 3 - 1, r // 3 - 1
        ii = np.clip(np.arange(nx) + dx, 0, nx - 1)
        jj = (np.arange(ny) + dy) % ny
        v = host[f][np.ix_(ii, jj)].astype(np.float32)
        xr[:, k] = ((v - np.float32(m.fshift[f].item())) * np.float32(m.fscale[f].item())).T.reshape(-1)
    want = (np.abs(xr).max(1) > thr).astype(np.uint8)
    got = mask[:, :, 0].t().reshape(-1).cpu().numpy()
    _say("flagged", int(want.sum()), "count", int(count.item()))
    # 9 points around the wrapped spike, 6 around the clamped one
    assert want.sum() >= 15 and (got == want).all() and int(count.item()) == int(want.sum())


def test_closure_per_point_fallback_takes_the_plain_chain_where_flagged(pr):
    """With some points beyond the training range, the closure's beta and
    adiff are the plain chain's values (bit for bit) at the flagged points
    and the network's elsewhere."""
    from pyranda_amd.closure import MLPClosure, pyrandaLearnedLAD
    from pyranda_amd.workloads import rm_run, rm_setup
    ss = rm_setup(pr, 48)
    rm_run(ss, 3)
    lad = pyrandaLearnedLAD(ss, None, judge_first=False)
    args = [ss.var(k).data for k in ("div", "vx", "uy", "rho", "Yh")] + [ss.var("dt").data]
    feats = lad.features(*args[:4])
    model = MLPClosure("cuda")
    model.set_normalisation([feats], [[ss.var("beta").data, ss.var("adiff").data * args[5]]])
    model.fmax *= 0.3                  # as if trained on a narrower range
    lad.model = model
    beta, adiff = lad.closure(*args)
    (mb, ma), mask, count = model.infer(feats, lad.periodic, sensor=lad.sensor_factor * model.fmax)
    pb, pa = lad.plain()
    torch.cuda.synchronize()
    sel = mask.bool()
    n = int(count.item())
    _say("flagged points", n, "of", sel.numel(), "partial", lad.partial)
    assert 0 < n < sel.numel() and lad.partial == 1 and lad.fallbacks == 0
    assert torch.equal(beta[sel], pb[sel].to(beta.dtype)) and torch.equal(adiff[sel], pa[sel].to(adiff.dtype))
    # the network's second output is adiff * dt
    assert torch.equal(beta[~sel], mb[~sel]) and torch.equal(adiff[~sel], (ma / args[5])[~sel])


def test_closure_judge_switches_the_closure_off(pr):
    """judge_every / judge_tol: a judged evaluation whose closure error is
    above the tolerance switches the closure off; that evaluation and every
    later one are the plain chain."""
    from pyranda_amd.closure import MLPClosure, pyrandaLearnedLAD
    from pyranda_amd.workloads import rm_run, rm_setup
    ss = rm_setup(pr, 48)
    rm_run(ss, 2)
    lad = pyrandaLearnedLAD(ss, None, judge_every=2, judge_tol=1e-3, judge_first=False)
    args = [ss.var(k).data for k in ("div", "vx", "uy", "rho", "Yh")] + [ss.var("dt").data]
    model = MLPClosure("cuda")
    feats = lad.features(*args[:4])
    model.set_normalisation([feats], [[ss.var("beta").data, ss.var("adiff").data]])
    lad.model = model
    lad.closure(*args)                       # call 1: the (untrained) network
    assert lad.enabled and lad.fallbacks == 0
    b2, a2 = lad.closure(*args)              # call 2: judged, far off -> off
    assert not lad.enabled and lad.fallbacks == 1 and len(lad.judged) == 1
    pb, pa = lad.plain()
    b3, a3 = lad.closure(*args)
    torch.cuda.synchronize()
    assert lad.fallbacks == 2
    for got in ((b2, a2), (b3, a3)):
        assert torch.equal(got[0], pb) and torch.equal(got[1], pa)


def test_closure_first_evaluation_is_judged(pr):
    """judge_first (the default): the first evaluation of a trained closure
    is judged; above judge_tol the closure is off from that evaluation on,
    and it returns the plain chain's values."""
    from pyranda_amd.closure import MLPClosure, pyrandaLearnedLAD
    from pyranda_amd.workloads import rm_run, rm_setup
    ss = rm_setup(pr, 48)
    rm_run(ss, 2)
    args = [ss.var(k).data for k in ("div", "vx", "uy", "rho", "Yh")] + [ss.var("dt").data]
    for tol, on in ((1e-3, False), (1e30, True)):
        lad = pyrandaLearnedLAD(ss, None, judge_tol=tol)
        model = MLPClosure("cuda")
        model.set_normalisation([lad.features(*args[:4])], [[ss.var("beta").data, ss.var("adiff").data]])
        lad.model = model
        b1, a1 = lad.closure(*args)
        assert len(lad.judged) == 1 and lad.enabled == on, tol
        lad.closure(*args)
        assert len(lad.judged) == 1                 # judge_every = 0: the first only
        if not on:
            pb, pa = lad.plain()
            torch.cuda.synchronize()
            assert lad.fallbacks == 2 and torch.equal(b1, pb) and torch.equal(a1, pa)


def test_mlp_learns_a_smooth_map(pr):
    """Training on the matrix cores reduces the error of a map the network
    can represent (targets from another network of the same shape)."""
    teacher = _model(pr, seed=11)
    m = _model(pr, seed=5)
    nx, ny = 96, 64
    g = torch.Generator().manual_seed(2)
    samples = []
    for _ in range(4):
        fields = [torch.randn(ny, nx, generator=g, dtype=torch.float64).t().reshape(nx, ny, 1).to("cuda", torch.float32)
                  for _ in range(3)]
        samples.append((fields, teacher.infer(fields, (True, True))))
    before = float(m.gradient(*samples[0], (True, True))[1])
    losses = m.train(samples, (True, True), steps=300, lr=3e-3)
    after = float(m.gradient(*samples[0], (True, True))[1])
    _say("mse before", before, "after", after, "last losses", losses[-4:].tolist())
    assert math.isfinite(after) and after < 0.2 * before


def test_closure_fallback_is_the_plain_chain(pr):
    """An untrained closure flags every evaluation, and its fallback is the
    plain LAD chain: the RM run is bit for bit the plain run (FP64)."""
    from pyranda_amd.closure import MLPClosure
    from pyranda_amd.workloads import rm_run, rm_setup
    a = rm_setup(pr, 48)
    model = MLPClosure("cuda")
    b = rm_setup(pr, 48, closure=model)
    sa = rm_run(a, 4)
    sb = rm_run(b, 4)
    lad = b.packages["LearnedLAD"]
    assert lad.calls > 0 and lad.fallbacks == lad.calls
    assert sa == sb
    for k in ("rho", "Et", "beta", "adiff", "Yh"):
        assert torch.equal(a.var(k).data, b.var(k).data), k


def test_closure_fallback_is_the_plain_chain_at_full_size(pr):
    """The learned-closure workload at its full size, 1024 x 512 (x along the
    shock, y periodic, single-mode interface): an untrained closure falls
    back on every evaluation and three RM steps with it are bit for bit the
    plain run (FP64), which stays finite."""
    from pyranda_amd.closure import MLPClosure
    from pyranda_amd.workloads import rm_run, rm_setup
    a = rm_setup(pr, 512, nx=1024, single_mode=True)
    b = rm_setup(pr, 512, nx=1024, closure=MLPClosure("cuda"), single_mode=True)
    assert (a.nx, a.ny) == (1024, 512)
    sa = rm_run(a, 3)
    sb = rm_run(b, 3)
    lad = b.packages["LearnedLAD"]
    _say("calls", lad.calls, "fallbacks", lad.fallbacks, "state", sa)
    # setIC evaluates the ALG lines once, then 5 stages per step
    assert lad.calls == 16 and lad.fallbacks == lad.calls
    assert sa == sb and math.isfinite(sa["time"])
    for k in ("rho", "rhou", "Et", "beta", "adiff", "Yh"):
        assert torch.equal(a.var(k).data, b.var(k).data), k
    assert torch.isfinite(a.var("rho").data).all()


def test_closure_trained_on_rm(pr):
    """Train on snapshots of the plain RM run, then run the closure: it is
    used (no fallback) and the judge reports its error against the plain
    chain; the run stays finite."""
    from pyranda_amd.closure import MLPClosure
    from pyranda_amd.workloads import rm_run, rm_setup
    plain = rm_setup(pr, 64)
    lad_vars = ("div", "vx", "uy", "rho")
    from pyranda_amd.closure import pyrandaLearnedLAD
    helper = pyrandaLearnedLAD(plain, None)
    samples, st = [], None
    for _ in range(6):
        st = rm_run(plain, 5, state=st)
        samples.append(helper.sample())
    model = MLPClosure("cuda")
    model.set_normalisation([s[0] for s in samples], [s[1] for s in samples])
    losses = model.train(samples, (False, True), steps=200, lr=2e-3)
    _say("training mse first/last", float(losses[0]), float(losses[-1]))
    assert float(losses[-8:].mean()) < float(losses[:8].mean())
    run = rm_setup(pr, 64, closure=model, closure_kw={"judge_first": False})
    rm_run(run, 5)
    lad = run.packages["LearnedLAD"]
    rep = lad.judge(*[run.var(k).data for k in ("div", "vx", "uy", "rho", "Yh")], run.var("dt").data)
    _say("closure calls", lad.calls, "fallbacks", lad.fallbacks, "judge", rep)
    assert lad.calls > lad.fallbacks
    assert torch.isfinite(run.var("rho").data).all()
    for name, (l2, mx) in rep.items():
        assert math.isfinite(l2) and math.isfinite(mx), name


def test_mlp_mixed_precision_fields(pr):
    """Feature fields of mixed precision (an FP32 run whose EOM mixes in the
    FP64 mesh coordinates) are evaluated in FP64, as if all were FP64."""
    m = _model(pr)
    nx, ny = 40, 24
    g = torch.Generator().manual_seed(9)
    f64 = [torch.randn(ny, nx, generator=g, dtype=torch.float64).t().reshape(nx, ny, 1).to("cuda")
           for _ in range(3)]
    mixed = [f64[0].float(), f64[1], f64[2].float()]
    want = m.infer([f.float().double() if i != 1 else f for i, f in enumerate(mixed)], (False, True))
    got = m.infer(mixed, (False, True))
    torch.cuda.synchronize()
    for a, b in zip(got, want):
        assert a.dtype == torch.float64 and torch.equal(a, b)
