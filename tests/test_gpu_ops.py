"""GPU parity of the compact operators through the C-ABI (libpyranda_amd.so).

Bar: FP64 kernels within 1e-12 of the oracle (max error relative to max|ref|);
FP32-refined kernels within 1e-6 of the oracle applied to the same FP32 input.
Full 256^3 size: the seeded 64-mode Fourier field, whose exact result is known
per mode from the operators' transfer functions (size-independent check).
"""
import math
import sys

import numpy as np
import pytest
import torch

from meshgen import mesh

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyranda_amd import _lib
    _lib.lib()
    return torch.device("cuda:0")


def to_dev(a, dev, dtype=torch.float64):
    a = np.asfortranarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).to(dev, dtype)
    return t.permute(2, 1, 0)


def to_host(t):
    return t.permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0).astype(np.float64)


CASES = [
    ((32, 32, 32), ("PERI", "PERI", "PERI")),
    ((33, 40, 20), ("NONE", "NONE", "NONE")),
    ((64, 100, 16), ("PERI", "NONE", "PERI")),
    ((256, 8, 9), ("NONE", "PERI", "NONE")),
    ((24, 9, 12), ("SYMM", "SYMM", "NONE")),
    ((130, 70, 3), ("PERI", "NONE", "NONE")),
    ((300, 20, 10), ("PERI", "PERI", "NONE")),  # 10 chunks per x line
    # periodic y/z lines of whole 32-row chunks in whole line groups: the
    # pipelined y/z kernels (global->LDS copies of the next line group)
    ((64, 128, 128), ("PERI", "PERI", "PERI")),
    # periodic x lines of whole KiB: the pipelined persistent x kernel
    ((256, 24, 10), ("PERI", "NONE", "PERI")),
    # the longest lines (32 chunks; 16-line y/z workgroups; bounded LU tables
    # read from global memory) and an uneven 1000-point partition
    ((1024, 12, 8), ("PERI", "NONE", "NONE")),
    ((12, 1000, 8), ("NONE", "NONE", "PERI")),
    ((8, 9, 1024), ("NONE", "PERI", "NONE")),
    # bounded x lines of more than 16 chunks (x kernel, LU tables from global
    # memory): 1024 points, and 716 (uneven partition, 23 chunks)
    ((1024, 12, 8), ("NONE", "PERI", "NONE")),
    ((716, 10, 3), ("NONE", "PERI", "NONE")),
    # the 2-D RM grids: bounded x along the shock, periodic y of 512 points
    ((1024, 512, 1), ("NONE", "PERI", "PERI")),
    ((716, 512, 1), ("NONE", "PERI", "PERI")),
]


def _setup(nn, kinds, oracle_mod):
    from pyranda_amd.parcop import ParcopContext
    periodic = [k == "PERI" for k in kinds]
    sym = [[k == "SYMM", k == "SYMM"] for k in kinds]
    dom = [(0.0, 2 * np.pi * (n - 1) / n if p else 1.0, n) for n, p in zip(nn, periodic)]
    (X, Y, Z), dd = mesh(dom)
    bcs = []
    for k in kinds:
        bcs += [k, k]
    ctx = ParcopContext(nn, [d[0] for d in dom], [d[1] for d in dom], bcs)
    P = oracle_mod.Oracle(nn, dd, periodic, symmetric=sym)
    rng = np.random.default_rng(sum(nn))
    v = np.asfortranarray(np.sin(X) * np.cos(2 * Y) + 0.3 * np.sin(3 * Z + 1)
                          + 0.05 * rng.standard_normal(nn) + 2.0)
    return ctx, P, v


def _rel(got, ref):
    return np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)


@pytest.mark.parametrize("solver", ["circulant", "lu"])
@pytest.mark.parametrize("nn,kinds", CASES)
def test_directional_ops_f64_and_r32(dev, oracle_mod, nn, kinds, solver):
    if solver == "lu" and "PERI" not in kinds:
        pytest.skip("the periodic solver choice only affects periodic lines")
    _say("case", nn, kinds, solver)
    ctx, P, v = _setup(nn, kinds, oracle_mod)
    ctx.set_periodic_solver(solver)
    vd = to_dev(v, dev)
    v32 = v.astype(np.float32).astype(np.float64)
    vd32 = to_dev(v32, dev, torch.float32)
    for op in range(5):
        for d in range(3):
            ref = P.apply(op, d, v)
            got = to_host(ctx.apply(op, d, vd))
            assert _rel(got, ref) <= 1e-12, ("f64", op, d, _rel(got, ref))
            ref32 = P.apply(op, d, v32)
            got32 = to_host(ctx.apply(op, d, vd32))
            assert _rel(got32, ref32) <= 1e-6, ("r32", op, d, _rel(got32, ref32))
            # FP32 fields with the FP64 solve: one rounding of each result
            ctx.set_r32_solve("fp64")
            got32 = to_host(ctx.apply(op, d, vd32))
            ctx.set_r32_solve("refined")
            assert _rel(got32, ref32) <= 1.2e-7, ("r32-fp64solve", op, d, _rel(got32, ref32))
    ctx.close()


@pytest.mark.parametrize("nn,kinds", [((24, 9, 12), ("SYMM", "SYMM", "NONE")),
                                      ((40, 33, 64), ("NONE", "SYMM", "SYMM"))])
def test_antisymmetric_member(dev, oracle_mod, nn, kinds):
    """The antisymmetric member of the SYMM pair (variant -1 at SYMM ends,
    compact.f90:78-84) through pyr_apply(op | PYR_OP_ANTISYM): every operator
    and direction against the oracle's member (directions without a SYMM end
    keep the default operator)."""
    from pyranda_amd import _lib
    _say("antisymmetric member", nn, kinds)
    ctx, P, v = _setup(nn, kinds, oracle_mod)
    vd = to_dev(v, dev)
    v32 = v.astype(np.float32).astype(np.float64)
    vd32 = to_dev(v32, dev, torch.float32)
    for op in range(5):
        for d in range(3):
            ref = P.apply_sym(op, d, -1, v)
            got = to_host(ctx.apply(op | _lib.OP_ANTISYM, d, vd))
            assert _rel(got, ref) <= 1e-12, ("f64", op, d, _rel(got, ref))
            if kinds[d] == "SYMM":
                assert _rel(ref, P.apply(op, d, v)) > 1e-6, ("members differ", op, d)
            ref32 = P.apply_sym(op, d, -1, v32)
            got32 = to_host(ctx.apply(op | _lib.OP_ANTISYM, d, vd32))
            assert _rel(got32, ref32) <= 1e-6, ("r32", op, d, _rel(got32, ref32))
    ctx.close()


@pytest.mark.parametrize("nn,kinds", CASES[:3] + CASES[4:5] + [((40, 33, 64), ("NONE", "SYMM", "SYMM"))])
def test_composites(dev, oracle_mod, nn, kinds):
    _say("composites", nn, kinds)
    ctx, P, v = _setup(nn, kinds, oracle_mod)
    w = np.asfortranarray(np.cos(v) * 0.7)
    u = np.asfortranarray(v * v * 0.1)
    vd, wd, ud = to_dev(v, dev), to_dev(w, dev), to_dev(u, dev)
    checks = [
        ("sfilter", to_host(ctx.sfilter(vd)), P.sfilter(v)),
        ("gfilter", to_host(ctx.gfilter(vd)), P.gfilter(v)),
        ("pring", to_host(ctx.pring(vd)), P.ring(v)),
        ("pringv", to_host(ctx.pringv(vd, wd, ud)), P.ringv(v, w, u)),
        ("plaplacian", to_host(ctx.plaplacian(vd)), P.laplacian(v)),
        ("divergence", to_host(ctx.divergence(vd, wd, ud)), P.divergence(v, w, u)),
    ]
    for d in (1, 2, 3):
        checks.append(("gfilterdir%d" % d, to_host(ctx.gfilterdir(vd, d)), P.gfilterdir(v, d)))
    for g, r in zip(ctx.grads(vd), P.grads(v)):
        checks.append(("grads", to_host(g), r))
    for name, got, ref in checks:
        assert _rel(got, ref) <= 1e-12, (name, _rel(got, ref))
    # FP32-refined composites against the oracle on the rounded input
    v32 = v.astype(np.float32).astype(np.float64)
    vd32 = to_dev(v32, dev, torch.float32)
    for name, got, ref in [("sfilter", ctx.sfilter(vd32), P.sfilter(v32)),
                           ("gfilter", ctx.gfilter(vd32), P.gfilter(v32)),
                           ("pring", ctx.pring(vd32), P.ring(v32)),
                           ("plaplacian", ctx.plaplacian(vd32), P.laplacian(v32))]:
        assert _rel(to_host(got), ref) <= 1e-6, ("r32", name, _rel(to_host(got), ref))
    ctx.close()


def _symbols(theta):
    c, s = np.cos, np.sin
    d1 = 2 * (6.375 * s(theta) + 1.515 * s(2 * theta) + .015 * s(3 * theta)) / (9 + 9 * c(theta) + .9 * c(2 * theta))
    d2 = (-28670 + 19170 * c(theta) + 9342 * c(2 * theta) + 158 * c(3 * theta)) / (16182 + 12024 * c(theta) + 774 * c(2 * theta))
    sf = (.99965 + 1.33304 * c(theta) + .33348 * c(2 * theta) + 8e-5 * c(3 * theta) - 1e-5 * c(4 * theta)) / (1 + 1.33248 * c(theta) + .33376 * c(2 * theta))
    return d1, d2, sf


def test_fourier_field_256(dev):
    """Full-size 256^3 microbenchmark field: exact per-mode results."""
    from pyranda_amd.parcop import ParcopContext
    from pyranda_amd.workloads import fourier_modes, fourier_field
    n = 256
    _say("fourier 256^3")
    L = 2 * math.pi * (n - 1) / n
    ctx = ParcopContext((n, n, n), (0.0, 0.0, 0.0), (L, L, L), ["PERI"] * 6)
    h = ctx.dpatch[0]
    ks, amps, phis = fourier_modes()
    f = fourier_field(n, dev, (ks, amps, phis))
    for d in range(3):
        th = np.abs(ks[:, d]) * h
        s1, s2, ssf = _symbols(th)
        sgn = np.sign(ks[:, d])
        # d/dx cos(k.x+phi) -> -sigma/h sin(...) = sigma/h cos(... + pi/2)
        exact_d1 = fourier_field(n, dev, (ks, amps, phis + np.pi / 2), weights=sgn * s1 / h)
        got = ctx.apply(0, d, f)
        err = ((got - exact_d1).abs().max() / exact_d1.abs().max()).item()
        _say("ddx dir", d, err)
        assert err <= 1e-12
        exact_sf = fourier_field(n, dev, (ks, amps, phis), weights=ssf)
        got = ctx.apply(3, d, f)
        err = ((got - exact_sf).abs().max() / exact_sf.abs().max()).item()
        assert err <= 1e-12
        exact_d2 = fourier_field(n, dev, (ks, amps, phis), weights=s2 / h ** 2)
        got = ctx.apply(1, d, f)
        err = ((got - exact_d2).abs().max() / exact_d2.abs().max()).item()
        assert err <= 1e-12
        # refined FP32 path against the exact result of the FP32-rounded input
        f32 = f.float()
        got32 = ctx.apply(3, d, f32).double()
        err32 = ((got32 - exact_sf).abs().max() / exact_sf.abs().max()).item()
        _say("fbar r32 dir", d, err32)
        assert err32 <= 1e-6
    ctx.close()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_autotune_keeps_results(dev, dtype):
    """The autotuner only changes launch shapes (lines per x workgroup, the
    pipelined y/z kernel): every operator's result is bit for bit the same
    before and after tuning, and the chosen knobs are among the candidates."""
    from pyranda_amd import _lib
    from pyranda_amd.parcop import ParcopContext
    nn = (64, 96, 64)
    L = 2 * math.pi
    ctx = ParcopContext(nn, (0.0, 0.0, 0.0), tuple(L * (n - 1) / n for n in nn), ["PERI"] * 6, device=dev)
    g = torch.Generator(device=dev).manual_seed(3)
    v = torch.randn(tuple(reversed(nn)), generator=g, device=dev, dtype=dtype).permute(2, 1, 0)
    ops = (_lib.OP_D1, _lib.OP_D2, _lib.OP_D8, _lib.OP_SFILTER, _lib.OP_GFILTER)
    before = {(op, d): ctx.apply(op, d, v).clone() for op in ops for d in range(3)}
    rep = ctx.autotune(dtype, reps=2)
    assert set(rep) == set(before)
    for (op, d), (best, times) in rep.items():
        assert best in times
        assert best[0][0] == ("x_lds_kb" if d == 0 else "yz_pipe")
        after = ctx.apply(op, d, v)
        assert torch.equal(after, before[(op, d)]), (op, d, best)
    _say("autotune", dtype, {k: v_[0] for k, v_ in rep.items()})
    ctx.close()


@pytest.mark.parametrize("nn", [(256, 21, 9), (128, 24, 11)])
def test_x_kernel_variants_keep_results(dev, nn):
    """The one-pass x kernel at several LDS budgets, the pipelined kernel
    with a second LDS tile (yz_pipe 1) and, for the explicit filter, the one
    that prefetches the next lines into registers (yz_pipe 2) give the same
    result bit for bit, FP64 and FP32, also for an FP64 sweep that stores
    only its FP32 rounding; whole-KiB periodic lines (the pipelined kernels'
    shape), line counts that leave the last group of lines partly filled."""
    from pyranda_amd import _lib
    from pyranda_amd.parcop import ParcopContext
    L = 2 * math.pi
    ctx = ParcopContext(nn, (0.0, 0.0, 0.0), tuple(L * (n - 1) / n for n in nn), ["PERI"] * 6, device=dev)
    g = torch.Generator(device=dev).manual_seed(9)
    ops = (_lib.OP_D1, _lib.OP_D2, _lib.OP_D8, _lib.OP_SFILTER, _lib.OP_GFILTER)
    for dtype in (torch.float64, torch.float32):
        v = torch.randn(tuple(reversed(nn)), generator=g, device=dev, dtype=dtype).permute(2, 1, 0)
        for op in ops:
            ctx.set_tuning(op, 0, dtype, x_lds_kb=40, yz_pipe=0)
            base = ctx.apply(op, 0, v).clone()
            base32 = ctx.apply_to32(op, 0, v).clone() if dtype == torch.float64 else None
            shapes = [(kb, pp) for pp in (0, 1) for kb in (12, 24, 80)]
            if op == _lib.OP_GFILTER:
                shapes += [(kb, 2) for kb in (12, 24, 80)]
            for kb, pp in shapes:
                ctx.set_tuning(op, 0, dtype, x_lds_kb=kb, yz_pipe=pp)
                got = ctx.apply(op, 0, v)
                _say("x variants", nn, dtype, op, kb, pp)
                assert torch.equal(got, base), (dtype, op, kb, pp)
                if base32 is not None:
                    assert torch.equal(ctx.apply_to32(op, 0, v), base32), (op, kb, pp)
            ctx.set_tuning(op, 0, dtype)
    ctx.close()


@pytest.mark.parametrize("nn,kinds", [((64, 96, 64), ("PERI", "PERI", "PERI")),
                                      ((40, 72, 50), ("NONE", "NONE", "NONE"))])
def test_yz_lines_per_workgroup_keep_results(dev, nn, kinds):
    """16, 32 and 64 lines per y/z workgroup (pyr_set_tuning yz_lines, the
    plain kernel) and the plain kernel with runtime row counts (yz_pipe 3)
    give the same result bit for bit, FP64 and FP32, periodic (circulant) and
    bounded (chunked LU) lines, line counts that do not fill the last
    workgroup included."""
    from pyranda_amd import _lib
    from pyranda_amd.parcop import ParcopContext
    L = 2 * math.pi
    per = kinds[0] == "PERI"
    xn = tuple(L * (n - 1) / n for n in nn) if per else (L, L, L)
    ctx = ParcopContext(nn, (0.0, 0.0, 0.0), xn, [k for k in kinds for _ in (0, 1)], device=dev)
    g = torch.Generator(device=dev).manual_seed(5)
    ops = (_lib.OP_D1, _lib.OP_D2, _lib.OP_D8, _lib.OP_SFILTER, _lib.OP_GFILTER)
    for dtype in (torch.float64, torch.float32):
        v = torch.randn(tuple(reversed(nn)), generator=g, device=dev, dtype=dtype).permute(2, 1, 0)
        for op in ops:
            for d in (1, 2):
                got = {}
                for ly in (64, 32, 16):
                    ctx.set_tuning(op, d, dtype, yz_pipe=0, yz_lines=ly)
                    got[ly] = ctx.apply(op, d, v).clone()
                # the plain kernel with runtime row counts
                ctx.set_tuning(op, d, dtype, yz_pipe=3, yz_lines=64)
                got["rt"] = ctx.apply(op, d, v).clone()
                ctx.set_tuning(op, d, dtype)
                _say("yz lines", nn, kinds[0], dtype, op, d)
                assert torch.equal(got[16], got[64]) and torch.equal(got[32], got[64]), (dtype, op, d)
                assert torch.equal(got["rt"], got[64]), (dtype, op, d)
    ctx.close()


@pytest.mark.parametrize("nn,kinds", [((64, 96, 64), ("PERI", "PERI", "PERI")),
                                      ((33, 40, 20), ("NONE", "NONE", "NONE")),
                                      ((64, 48, 1), ("PERI", "PERI", "PERI"))])
def test_filter_with_rounded_copy(dev, nn, kinds):
    """fbar of an FP64 field with the FP32 rounding stored by the last sweep
    (pyr_sfilter_f64_rounded, and sweep by sweep through pyr_apply_rounded,
    the path the timed runs take): the FP64 result is bit for bit
    pyr_sfilter_f64's, the copy is its rounding; bounded lines and a null z
    direction (the copy kernel) included."""
    from pyranda_amd import _lib
    from pyranda_amd.parcop import ParcopContext
    _say("fbar with rounded copy", nn, kinds)
    x1 = (0.0, 0.0, 0.0)
    xn = tuple(2 * math.pi * (n - 1) / n if k == "PERI" and n > 1 else 1.0 for n, k in zip(nn, kinds))
    ctx = ParcopContext(nn, x1, xn, [k for k in kinds for _ in range(2)], device=dev)
    g = torch.Generator(device=dev).manual_seed(11)
    v = torch.randn(tuple(reversed(nn)), generator=g, device=dev, dtype=torch.float64).permute(2, 1, 0)
    ref = ctx.sfilter(v)
    out, out32 = ctx.sfilter_rounded(v)
    assert out.dtype == torch.float64 and out32.dtype == torch.float32
    assert torch.equal(out, ref) and torch.equal(out32, ref.float())
    ctx.timing = {}
    out_t, out32_t = ctx.sfilter_rounded(v)
    ctx.timing = None
    assert torch.equal(out_t, ref) and torch.equal(out32_t, ref.float())
    # an x sweep writes the copy too
    want = ctx.apply(_lib.OP_SFILTER, 0, v)
    o64 = torch.empty_like(want)
    _lib.check(ctx._L.pyr_apply_rounded(ctx._h, _lib.OP_SFILTER, 0, ctx._ptr(v), ctx._ptr(o64),
                                        ctx._ptr(out32)))
    torch.cuda.synchronize()
    assert torch.equal(o64, want) and torch.equal(out32, want.float())
    ctx.close()


@pytest.mark.parametrize("nn,kinds", CASES)
def test_fp64_sweep_storing_fp32(dev, nn, kinds):
    """pyr_apply_to_r32: an FP64 sweep that stores only the FP32 rounding of
    its result is bit for bit the rounding of the FP64 sweep's result, for
    every operator and direction (periodic, bounded, SYMM, null directions,
    long lines)."""
    from pyranda_amd import _lib
    import oracle.pyoracle as po
    _say("fp64 -> fp32 sweeps", nn, kinds)
    ctx, _, v = _setup(nn, kinds, po)
    vd = to_dev(v, dev)
    for op in (_lib.OP_D1, _lib.OP_D2, _lib.OP_D8, _lib.OP_SFILTER, _lib.OP_GFILTER):
        for d in range(3):
            want = ctx.apply(op, d, vd)
            got = ctx.apply_to32(op, d, vd)
            torch.cuda.synchronize()
            assert got.dtype == torch.float32 and torch.equal(got, want.float()), (op, d)
    ctx.close()


def test_autotune_solve_modes(dev):
    """With solve_modes the autotuner also picks, per operator and direction,
    refined-FP32 or FP64 solves for FP32 fields: every result is bit for bit
    the one of a context using the chosen arithmetic everywhere, and within
    the FP32-refined bar of the FP64 path on the same input."""
    from pyranda_amd import _lib
    from pyranda_amd.parcop import SOLVE_MODE_MARGIN, ParcopContext
    nn = (64, 96, 64)
    L = 2 * math.pi
    xn = tuple(L * (n - 1) / n for n in nn)
    ctx = ParcopContext(nn, (0.0, 0.0, 0.0), xn, ["PERI"] * 6, device=dev)
    ref = {m: ParcopContext(nn, (0.0, 0.0, 0.0), xn, ["PERI"] * 6, device=dev) for m in ("refined", "fp64")}
    ref["fp64"].set_r32_solve("fp64")
    g = torch.Generator(device=dev).manual_seed(5)
    v = torch.randn(tuple(reversed(nn)), generator=g, device=dev, dtype=torch.float64).permute(2, 1, 0)
    v32 = v.float()
    ops = (_lib.OP_D1, _lib.OP_D2, _lib.OP_D8, _lib.OP_SFILTER, _lib.OP_GFILTER)
    rep = ctx.autotune(torch.float32, reps=2, solve_modes=True)
    chosen = {}
    for (op, d), (best, times) in rep.items():
        assert best[0][0] == "r32_solve" and best in times
        mode = ("refined", "fp64")[best[0][1]]
        if mode == "fp64":
            # the FP64 solve only for a clear gain over the refined solve
            t_ref = min(t for c, t in times.items() if c[0] == ("r32_solve", 0))
            assert times[best] <= (1.0 - SOLVE_MODE_MARGIN) * t_ref, (op, d)
        chosen[(op, d)] = mode
        got = ctx.apply(op, d, v32)
        assert torch.equal(got, ref[mode].apply(op, d, v32)), (op, d, mode)
        exact = ref["refined"].apply(op, d, v32.double())
        err = float((got.double() - exact).abs().max() / exact.abs().max())
        assert err <= 1e-6, (op, d, mode, err)
    # composites follow the per-direction choice
    f = ctx.sfilter(v32)
    a = ref[chosen[(_lib.OP_SFILTER, 0)]].apply(_lib.OP_SFILTER, 0, v32)
    b = ref[chosen[(_lib.OP_SFILTER, 1)]].apply(_lib.OP_SFILTER, 1, a)
    c = ref[chosen[(_lib.OP_SFILTER, 2)]].apply(_lib.OP_SFILTER, 2, b)
    assert torch.equal(f, c)
    with pytest.raises(_lib.PyrandaAMDError):
        _lib.check(ctx._L.pyr_set_solve_tuning(ctx._h, 0, 0, 7))
    _say("solve modes", {k: m for k, m in chosen.items()})
    for c_ in [ctx] + list(ref.values()):
        c_.close()


def test_timing_only_selected_launches(dev):
    """Per-launch timing (ParcopContext.timing) of every launch, or with
    timing_only of the selected (op, dir, dtype) launches only; a composite
    is split into its sweeps only when its operator is timed. Results are
    the untimed ones bit for bit."""
    from pyranda_amd import _lib
    from pyranda_amd.parcop import ParcopContext
    nn = (64, 40, 24)
    L = 2 * math.pi
    ctx = ParcopContext(nn, (0.0, 0.0, 0.0), tuple(L * (n - 1) / n for n in nn), ["PERI"] * 6, device=dev)
    g = torch.Generator(device=dev).manual_seed(4)
    v = torch.randn(tuple(reversed(nn)), generator=g, device=dev, dtype=torch.float64).permute(2, 1, 0)
    want = (ctx.sfilter(v).clone(), ctx.pring(v).clone(), ctx.ddx(v).clone())
    for only in (None, {(_lib.OP_SFILTER, 0, "torch.float64")}):
        ctx.timing, ctx.timing_only = {}, only
        got = (ctx.sfilter(v), ctx.pring(v), ctx.ddx(v))
        summ = ctx.timing_summary()
        ctx.timing, ctx.timing_only = None, None
        for a, b in zip(got, want):
            assert torch.equal(a, b)
        _say("timed", only, sorted(summ))
        if only is None:
            assert summ[(_lib.OP_SFILTER, 2, "torch.float64")][0] == 1
            assert summ[(_lib.OP_D8, 1, "torch.float64")][0] == 1
            assert summ[(_lib.OP_D1, 0, "torch.float64")][0] == 1
        else:
            assert list(summ) == [(_lib.OP_SFILTER, 0, "torch.float64")] and summ[list(summ)[0]][0] == 1
    ctx.close()
