"""Benchmark: 3-D Taylor-Green Euler+LAD, 256^3 on one MI355X.

A step is one 5-stage RK4 step of the Taylor-Green EOM of
examples/TaylorGreen.py (gamma=1.4, LAD bulk viscosity via ring/gbar, fbar
on the conserved variables after every stage) with Mach 0.1
(p0 = 1/(gamma*0.1^2)), CFL 0.5 from the initial state, run through
pyranda_amd.pyrandaSim with every field resident in HBM.

The headline value is FP32-refined point-updates/s (FP32 fields and sweeps,
one FP64-residual refinement in every operator, FP64 RK4 state, FP64
storage of the offset-dominated fields). The FP64 path is timed beside it on
the same steps, and every conserved variable of the FP32 run is compared
with the FP64 run after the same steps (rel_max / rel_L2, bar 1e-6).

Roofline: every operator launch inside the timed region is bracketed by HIP
events on its stream; the operator kernel with the largest total time is the
dominant kernel, and its algorithmic bytes per launch (one read + one write
of every point, at the storage width) over its mean launch time give
`achieved` GB/s against the 8 TB/s HBM peak.

CPU baseline: the oracle's FP64 restatement of the same step
(oracle/tg_cpu.py, OpenMP operators + numpy pointwise) on the host cores, over
a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per dispatch of the line kernels at 256^3, measured with rocprofv3
# --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch), FETCH_SIZE scaled by 2.0 and
# WRITE_SIZE by 1.0 from a plain device copy of known size; the table and the
# commands are in profiles/line_ops_hbm_traffic_pmc.md. Key: (op, direction,
# field dtype) of the circulant solve.
MEASURED_TRAFFIC = {
    ("d1", "x", "torch.float64"): (131472 + 131072) * 1024,
    ("d1", "x", "torch.float32"): (65959 + 65536) * 1024,
    ("d1", "y", "torch.float64"): (131687 + 131072) * 1024,
    ("d1", "y", "torch.float32"): (65845 + 65536) * 1024,
    ("fbar", "y", "torch.float64"): (131510 + 131072) * 1024,
    ("fbar", "y", "torch.float32"): (65952 + 65536) * 1024,
    ("gbar", "y", "torch.float64"): (140248 + 131072) * 1024,
    ("gbar", "y", "torch.float32"): (65844 + 65536) * 1024,
    ("fbar", "x", "torch.float64"): (131496 + 131072) * 1024,
    # FP64 fbar z sweep that also stores the FP32 rounding (20 B/pt)
    ("fbar", "z", "torch.float64+float32"): (131502 + 196608) * 1024,
}
OPS = ["d1", "d2", "d8", "fbar", "gbar"]
CONSERVED = ("rho", "rhou", "rhov", "rhow", "Et")
DIRS = "xyz"


def tg_mesh(n):
    return """
    xdom = (0.0, 2*pi*FF,  Npts, periodic=True)
    ydom = (0.0, 2*pi*FF,  Npts, periodic=True)
    zdom = (0.0, 2*pi*FF,  Npts, periodic=True)
    """.replace("Npts", str(n)).replace("pi", str(math.pi)).replace("FF", str(float(n - 1) / n))


def make_tg(n, precision, p0, solver="circulant"):
    """precision: a pyrandaSim precision, "fp32-rounded-state" (fp32 with
    fp64_state=False: the conservative filter rounds the state to FP32) or
    "fp32-refined-only" (fp32 without the per-operator solve choice)."""
    import pyranda_amd as pr
    from pyranda_amd.workloads import TG_EOM, TG_IC
    rounded = precision == "fp32-rounded-state"
    base = "fp32" if precision in ("fp32-rounded-state", "fp32-offset", "fp32-all-f32") else precision
    wide = {"fp32-all-f32": None, "fp32-offset": "offset"}.get(precision, "auto")
    # grid reductions (dt.courant, dt.diff) stay on the device: no host round
    # trip inside a step (bit for bit the host-scalar values)
    ss = pr.pyrandaSim("TGvortex", tg_mesh(n), silent=True, precision=base, fp64_state=not rounded,
                       wide_fields=wide, device_scalars=True)
    ss.PyMPI.ctx.set_periodic_solver(solver)
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.EOM(TG_EOM)
    ss.setIC(TG_IC.replace("p0 = 100.0", "p0 = %r" % p0))
    return ss


def run_steps(ss, dt, nsteps, t0=0.0):
    t = t0
    for _ in range(nsteps):
        t = ss.rk4(t, dt)
    return t


def rel_err(a, b):
    import torch
    a = a.to(torch.float64)
    b = b.to(torch.float64)
    d = (a - b)
    rmax = (d.abs().max() / b.abs().max()).item()
    rl2 = (torch.linalg.vector_norm(d) / torch.linalg.vector_norm(b)).item()
    return rmax, rl2


def cpu_baseline(n, p0, dt, threads, sample_steps=1):
    """Point-updates/s of the oracle's FP64 Taylor-Green step (oracle/tg_cpu.py)
    on `threads` host threads: one untimed step, then `sample_steps` timed."""
    from oracle.tg_cpu import TaylorGreenCPU
    tg = TaylorGreenCPU(n, p0=p0, threads=threads)
    t = tg.rk4(0.0, dt)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        t = tg.rk4(t, dt)
    el = time.perf_counter() - t0
    tg.close()
    return n ** 3 * sample_steps / el, el


def host_info():
    """nproc, usable cores, OMP_NUM_THREADS and the CPU model of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cores": usable,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def _symbols(theta):
    import numpy as np
    c, s = np.cos, np.sin
    d1 = 2 * (6.375 * s(theta) + 1.515 * s(2 * theta) + .015 * s(3 * theta)) / (9 + 9 * c(theta) + .9 * c(2 * theta))
    d2 = (-28670 + 19170 * c(theta) + 9342 * c(2 * theta) + 158 * c(3 * theta)) / (16182 + 12024 * c(theta) + 774 * c(2 * theta))
    d8 = (4200 - 6720 * c(theta) + 3360 * c(2 * theta) - 960 * c(3 * theta) + 120 * c(4 * theta)) / (29 + 28 * c(theta) + 3 * c(2 * theta))
    sf = (.99965 + 1.33304 * c(theta) + .33348 * c(2 * theta) + 8e-5 * c(3 * theta) - 1e-5 * c(4 * theta)) / (1 + 1.33248 * c(theta) + .33376 * c(2 * theta))
    a, b, cc, d, e = 3565 / 10368, 3091 / 12960, 1997 / 25920, 149 / 12960, 107 / 103680
    gf = a + 2 * (b * c(theta) + cc * c(2 * theta) + d * c(3 * theta) + e * c(4 * theta))
    return [d1, d2, d8, sf, gf]


def ops_bench(n, steps, warmup, dev, only=None, solver="circulant", r32_solve="refined"):
    """Operator microbenchmark: every directional operator on the seeded
    64-mode Fourier field (|k|<=32, |k|^-5/3, seed 1234), FP64 and FP32-refined,
    timed with HIP events, error against the exact per-mode result."""
    import numpy as np
    import torch
    from pyranda_amd.parcop import ParcopContext, empty_field
    from pyranda_amd.workloads import fourier_modes, fourier_field
    L = 2 * math.pi * (n - 1) / n
    ctx = ParcopContext((n, n, n), (0.0, 0.0, 0.0), (L, L, L), ["PERI"] * 6, device=dev)
    ctx.set_periodic_solver(solver)
    ctx.set_r32_solve(r32_solve)
    h = ctx.dpatch[0]
    ks, amps, phis = fourier_modes()
    f64 = fourier_field(n, dev, (ks, amps, phis))
    f32 = f64.float()
    out = {}
    reps = max(steps, 1) * 4
    only = [t for t in (only or "").split(",") if t]

    def wanted(name):
        return not only or any(name.startswith(t) for t in only)

    # plain device copy of the same field: the practical bandwidth ceiling, and
    # the kernel whose FETCH_SIZE/WRITE_SIZE counters scale the others
    for name, f, bpp in (("copy_f64", f64, 16), ("copy_f32", f32, 8)):
        if not wanted(name):
            continue
        o = torch.empty_like(f)
        for _ in range(warmup + 1):
            o.copy_(f)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            o.copy_(f)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        gbs = n ** 3 * bpp / (us * 1e-6) / 1e9
        out[name] = {"us": round(us, 2), "GB_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 3)}
        print("[ops] %-10s %-4s %8.1f us %7.1f GB/s (%.1f%%)" % (name, "", us, gbs, 100 * gbs / HBM_PEAK_GBS),
              file=sys.stderr, flush=True)
    for op in range(5):
        for d in range(3):
            th = np.abs(ks[:, d]) * h
            sym = _symbols(th)[op]
            if op == 0:
                exact = fourier_field(n, dev, (ks, amps, phis + np.pi / 2), weights=np.sign(ks[:, d]) * sym / h)
            elif op == 1:
                exact = fourier_field(n, dev, (ks, amps, phis), weights=sym / h ** 2)
            else:
                exact = fourier_field(n, dev, (ks, amps, phis), weights=sym)
            variants = [("f64", f64, 16), ("r32", f32, 8)]
            if op == 3 and d > 0:
                # FP64 fbar y/z sweep that also stores the FP32 rounding (the
                # last sweep of the filter of an FP32 run's FP64 state)
                variants.append(("f64r", f64, 20))
            for name, f, bpp in variants:
                if not wanted("%s_%s_%s" % (OPS[op], DIRS[d], name)):
                    continue
                kw = {"out32": empty_field((n, n, n), torch.float32, dev)} if name == "f64r" else {}
                o = ctx.apply(op, d, f, **kw)
                for _ in range(warmup):
                    ctx.apply(op, d, f, out=o, **kw)
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    ctx.apply(op, d, f, out=o, **kw)
                e1.record()
                torch.cuda.synchronize()
                us = e0.elapsed_time(e1) * 1e3 / reps
                err = ((o.double() - exact).abs().max() / exact.abs().max()).item()
                gbs = n ** 3 * bpp / (us * 1e-6) / 1e9
                rec = {"us": round(us, 2), "GB_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 3),
                       "rel_max_err": err}
                if name == "r32":
                    # the refined solve against the FP64 path on the same (rounded) input
                    ref = ctx.apply(op, d, f.double())
                    rec["rel_max_err_same_input"] = ((o.double() - ref).abs().max() / ref.abs().max()).item()
                    del ref
                out["%s_%s_%s" % (OPS[op], DIRS[d], name)] = rec
                print("[ops] %-10s %-4s %8.1f us %7.1f GB/s (%.1f%%) err %.2e%s" %
                      (OPS[op] + "_" + DIRS[d], name, us, gbs, 100 * gbs / HBM_PEAK_GBS, err,
                       " same-input %.2e" % rec["rel_max_err_same_input"] if name == "r32" else ""),
                      file=sys.stderr, flush=True)
    # composites of the Taylor-Green step: ring (three d8 sweeps with |.| and
    # max fused into the stores), laplacian (d2 sweeps accumulated in place),
    # fbar (x, y, z filters). Bytes: one read and one write per directional sweep.
    exact = {}
    sym = [_symbols(np.abs(ks[:, d]) * h) for d in range(3)]
    exact["lap"] = sum(fourier_field(n, dev, (ks, amps, phis), weights=sym[d][1] / h ** 2) for d in range(3))
    exact["fbar3"] = fourier_field(n, dev, (ks, amps, phis), weights=sym[0][3] * sym[1][3] * sym[2][3])
    d8 = [fourier_field(n, dev, (ks, amps, phis), weights=sym[d][2]).abs() * h * h for d in range(3)]
    exact["ring"] = torch.maximum(torch.maximum(d8[0], d8[1]), d8[2])
    del d8
    comps = {"ring": ctx.pring, "lap": ctx.plaplacian, "fbar3": ctx.sfilter}
    for cname, fn in comps.items():
        for name, f, bpp in (("f64", f64, 16), ("r32", f32, 8)):
            key = "%s_%s" % (cname, name)
            if not wanted(key):
                continue
            o = fn(f)
            for _ in range(warmup):
                o = fn(f)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                o = fn(f)
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / reps
            ex = exact[cname]
            err = ((o.double() - ex).abs().max() / ex.abs().max()).item()
            gbs = 3 * n ** 3 * bpp / (us * 1e-6) / 1e9
            out[key] = {"us": round(us, 2), "GB_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 3),
                        "rel_max_err": err}
            print("[ops] %-10s %-4s %8.1f us %7.1f GB/s (%.1f%%) err %.2e" %
                  (cname, name, us, gbs, 100 * gbs / HBM_PEAK_GBS, err), file=sys.stderr, flush=True)
    del exact
    if wanted("host_d1_x_f64"):
        # the f2py-module stand-in: a host Fortran-order array in, a new host
        # array out (host -> device copy, the d1 x sweep, device -> host copy),
        # wall clock per call
        from pyranda_amd.f2py_parcop import parcop
        parcop.setup(0, 0, 0, n, n, n, 1, 1, 1, 0, 0.0, L, 0.0, L, 0.0, L,
                     "PERI", "PERI", "PERI", "PERI", "PERI", "PERI")
        host = np.asfortranarray(f64.permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0))
        hreps = 3
        r = parcop.ddx(host)
        t0 = time.perf_counter()
        for _ in range(hreps):
            r = parcop.ddx(host)
        us = (time.perf_counter() - t0) * 1e6 / hreps
        ref = ctx.apply(0, 0, f64).permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0)
        gbs = n ** 3 * 16 / (us * 1e-6) / 1e9
        out["host_d1_x_f64"] = {"us": round(us, 1), "GB_s": round(gbs, 1), "hbm_frac": None,
                                "equal_to_device_path": bool(np.array_equal(r, ref))}
        print("[ops] %-10s %-4s %8.1f us %7.1f GB/s (host arrays: PCIe-inclusive)" % ("host_d1_x", "f64", us, gbs),
              file=sys.stderr, flush=True)
        del host, r, ref
    best = max((kv for kv in out.items() if not kv[0].startswith(("copy", "host"))),
               key=lambda kv: kv[1]["GB_s"], default=(None, None))
    print(json.dumps({"metric": "compact operator throughput, 256^3 periodic Fourier field",
                      "unit": "GB/s (algorithmic: one read + one write per point)",
                      "n": n, "reps": reps, "solver": solver, "ops": out,
                      "best": best[0]}), flush=True)


def ensemble_bench(args, dev):
    """KH ensemble: independent 2-D shear layers on uneven grids. Default
    workload: 64 wall-bounded layers (periodic x, walls in y with
    bc.extrap/bc.const, density ratio 2, Mach 0.3 on the velocity jump,
    seeded six-mode perturbation of amplitude 1e-2), each axis log-uniform in
    [128, 1024]; --ens-workload periodic runs the doubly periodic
    examples/KH.py layer on [128, 512] instead. Timed four ways on the same
    members: member by member with eager launches; one captured HIP graph per
    step over all members; the ensemble as one pyrandaGroupSim (shared
    allocation, one launch per operator and direction for all members); and
    that group step captured in a HIP graph."""
    import torch
    import pyranda_amd as pr
    from pyranda_amd.ensemble import pyrandaEnsemble
    from pyranda_amd.workloads import (ensemble_sizes, kh_advance, kh_ensemble, kh_group, kh_wall_ensemble,
                                       kh_wall_group)
    wall = args.ens_workload == "wall"
    lo, hi = (128, 1024) if wall else (128, 512)
    sizes = ensemble_sizes(args.members, lo, hi)
    make_members, make_group = (kh_wall_ensemble, kh_wall_group) if wall else (kh_ensemble, kh_group)
    npts = sum(nx * ny for nx, ny in sizes)
    print("[ensemble] %d members, %d points, sizes %s" % (len(sizes), npts, sizes), file=sys.stderr, flush=True)
    out = {}

    def timed(step, label):
        step(max(1, args.warmup))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(args.steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[label] = {"ms_per_step": 1e3 * el / args.steps, "value": npts * args.steps / el}
        print("[ensemble] %-16s %8.2f ms/step" % (label, 1e3 * el / args.steps), file=sys.stderr, flush=True)

    modes = args.ens_modes.split(",")
    if "members" in modes:
        sims, state = make_members(pr, sizes, precision=args.ens_precision, device=dev)
        ens = pyrandaEnsemble(sims, kh_advance, state)
        timed(ens.step_eager, "members_eager")
        ens.capture(concurrent=True)
        timed(ens.step, "members_graph")
        ens.graph = None
        del ens, sims, state
    g, gst = make_group(pr, sizes, precision=args.ens_precision, device=dev)
    gens = pyrandaEnsemble([g], kh_advance, [gst])
    timed(gens.step_eager, "group_eager")
    gens.capture(concurrent=False)
    timed(gens.step, "group_graph")
    best = max(out, key=lambda k: out[k]["value"])
    line = {
        "metric": "KH ensemble point-updates/s (%d independent 2-D Euler+LAD layers, uneven grids%s)"
                  % (len(sizes), ", wall-bounded in y" if wall else ""),
        "value": out[best]["value"], "unit": "point-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": out[best]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp32": "f32 fields and sweeps, f64 residual refinement, f64 RK4 state",
                  "fp32-fp64solve": "f32 fields, f64 sweeps, f64 RK4 state",
                  "fp64": "f64"}[args.ens_precision],
        "data": ("synthetic (wall-bounded KH layer, seeded multimode perturbation, on each member grid)" if wall
                 else "synthetic (KH.py initial condition on each member grid)"),
        "config": {"workload": "kh_wall_ensemble" if wall else "kh_ensemble", "members": len(sizes),
                   "points": npts,
                   "sizes": "log-uniform per axis in [%d, %d], multiples of 8, seed 2024" % (lo, hi),
                   "mode": best}, "modes": out,
        "speedup_over_members_eager": (out[best]["value"] / out["members_eager"]["value"]
                                       if "members_eager" in out else None),
    }
    print(json.dumps(line))
    return 0


def closure_bench(args, dev):
    """Learned LAD closure on the 2-D Richtmyer-Meshkov case (examples/RM3D.py,
    2-D branch): snapshots of the plain run train the BF16-MFMA MLP; then the
    MLP's evaluation of beta and adiff is timed against the plain chain on the
    same state, RK4 steps with either are timed, and the judge reports the
    closure error and the solution drift after the same steps. The closure's
    first evaluation is judged (pyrandaLearnedLAD's judge_first): above its
    tolerance the closure is off for the run, and the line says so."""
    import torch
    import pyranda_amd as pr
    from pyranda_amd.closure import MLPClosure, pyrandaLearnedLAD
    from pyranda_amd.workloads import rm_run, rm_setup
    n, prec = args.closure_n, args.closure_precision
    nxr = args.closure_nx or None
    plain = rm_setup(pr, n, precision=prec, device=dev, nx=nxr, single_mode=True)
    nx, ny = plain.nx, plain.ny
    print("[closure] RM 2-D %dx%d %s" % (nx, ny, prec), file=sys.stderr, flush=True)
    helper = pyrandaLearnedLAD(plain, None)
    lad_in = ("div", "vx", "uy", "rho")
    # snapshots from the initial state on, so that the run the closure is
    # timed on stays inside the training range
    samples, st = [helper.sample()], None
    for _ in range(args.closure_snapshots - 1):
        st = rm_run(plain, args.closure_every, state=st)
        samples.append(helper.sample())
    model = MLPClosure(dev)
    model.set_normalisation([s_[0] for s_ in samples], [s_[1] for s_ in samples])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = model.train(samples, (False, True), steps=args.closure_train_steps, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    print("[closure] trained %d steps in %.2f s, mse %.3e -> %.3e" % (
        args.closure_train_steps, t_train, float(losses[:10].mean()), float(losses[-10:].mean())),
        file=sys.stderr, flush=True)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    feats = samples[-1][0]
    v = {k: plain.var(k).data for k in ("div", "vx", "uy", "rho", "Yh")}
    ms_mlp = timed(lambda: model.infer(feats, (False, True)), 50)
    ms_feat = timed(lambda: helper.features(v["div"], v["vx"], v["uy"], v["rho"]), 20)
    ms_plain = timed(lambda: helper.plain(), 20)
    npts = nx * ny
    mfma_flops = 2.0 * (32 * 64 + 64 * 64 + 64 * 16) * npts      # as issued (padded tiles)
    useful_flops = 2.0 * (27 * 64 + 64 * 64 + 64 * 2) * npts
    # steps of the RM run with either evaluation, from the same initial state
    runs = {}
    for mode in ("plain", "closure"):
        ss = rm_setup(pr, n, precision=prec, device=dev, nx=nxr, single_mode=True,
                      closure=model if mode == "closure" else None)
        stt = rm_run(ss, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stt = rm_run(ss, args.steps, state=stt)
        torch.cuda.synchronize()
        runs[mode] = (ss, (time.perf_counter() - t0) / args.steps)
    cl = runs["closure"][0]
    lad = cl.packages["LearnedLAD"]
    jargs = [cl.var(k).data for k in ("div", "vx", "uy", "rho", "Yh")] + [cl.var("dt").data]
    judge = lad.judge(*jargs)
    judge_raw = lad.judge(*jargs, raw=True)
    rp, rc = runs["plain"][0].var("rho").data.double(), cl.var("rho").data.double()
    drift = float((rc - rp).norm() / rp.norm())
    line = {
        "metric": "RM 2-D learned LAD closure: MLP evaluation points/s (BF16 MFMA, FP32 accumulate)",
        "value": npts / (ms_mlp * 1e-3), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": 1e3 * runs["closure"][1], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 weights/activations, f32 accumulate; %s fields" % prec,
        "data": "synthetic (RM3D.py 2-D initial condition, seeded perturbation)",
        "config": {"workload": "rm_2d_learned_closure", "grid": [nx, ny], "mlp": "27-64-64-2 relu",
                   "train_snapshots": args.closure_snapshots, "train_steps": args.closure_train_steps},
        "mlp_ms": ms_mlp, "features_ms": ms_feat, "plain_lad_ms": ms_plain,
        "closure_eval_speedup": ms_plain / (ms_mlp + ms_feat),
        "roofline": {"bound": "mfma", "achieved": mfma_flops / (ms_mlp * 1e-3) / 1e12, "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": mfma_flops / (ms_mlp * 1e-3) / 1e12 / 2500.0,
                     "useful_tflops": useful_flops / (ms_mlp * 1e-3) / 1e12, "traffic": None},
        "step_ms": {"plain": 1e3 * runs["plain"][1], "closure": 1e3 * runs["closure"][1]},
        "step_speedup": runs["plain"][1] / runs["closure"][1],
        "closure_enabled_at_end": lad.enabled,
        "judged": [{"call": c_, **{k: {"rel_l2": a_, "rel_max": b_} for k, (a_, b_) in r_.items()}}
                   for c_, r_ in lad.judged],
        "closure_calls": lad.calls, "fallbacks": lad.fallbacks, "partial_fallbacks": lad.partial,
        "points_flagged": lad.points_flagged,
        "judge": {k: {"rel_l2": a_, "rel_max": b_} for k, (a_, b_) in judge.items()},
        "judge_network_alone": {k: {"rel_l2": a_, "rel_max": b_} for k, (a_, b_) in judge_raw.items()},
        "drift_rho_rel_l2": drift, "train_s": t_train,
        "train_mse": [float(losses[:10].mean()), float(losses[-10:].mean())],
    }
    print(json.dumps(line))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=256,
                    help="grid of the CPU-baseline sample (one timed RK4 step on every usable core)")
    ap.add_argument("--cpu-n1", type=int, default=64,
                    help="grid of the one-thread CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--ops", action="store_true",
                    help="operator microbenchmark on the 256^3 seeded Fourier field instead")
    ap.add_argument("--solver", default="circulant", choices=["circulant", "lu"],
                    help="solver of periodic implicit lines")
    ap.add_argument("--r32-solve", default="refined", choices=["refined", "fp64"],
                    help="--ops: arithmetic of the solves on FP32 fields")
    ap.add_argument("--precision", default="fp64,fp32,fp32-offset",
                    help="comma-separated pyrandaSim precisions to time (the headline is fp32: FP32 "
                         "storage where it keeps 1e-6 (pyrandaSim wide_fields='auto'), FP32 sweeps plus "
                         "one FP64-residual refinement in every FP32 operator; fp32-offset stores only the "
                         "offset-dominated fields in FP64 (faster, less accurate); fp32-fp64solve solves "
                         "in FP64 on FP32 fields; fp32-rounded-state is fp32 without the FP64 state through "
                         "the filter; fp32-all-f32 stores every field in FP32)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (no per-operator autotuning before timing)")
    ap.add_argument("--ensemble", action="store_true",
                    help="KH ensemble (independent 2-D layers on uneven grids, one HIP graph per step)")
    ap.add_argument("--members", type=int, default=64, help="--ensemble: number of members")
    ap.add_argument("--ens-workload", default="wall", choices=["wall", "periodic"],
                    help="--ensemble: wall-bounded layers on [128,1024]^2 or doubly periodic KH.py on [128,512]^2")
    ap.add_argument("--ens-precision", default="fp32", choices=["fp64", "fp32", "fp32-fp64solve"],
                    help="--ensemble: member precision (default: FP32 fields with the refined FP32 solve)")
    ap.add_argument("--ens-modes", default="members,group",
                    help="--ensemble: 'members' (member by member, eager and graph) and/or 'group'")
    ap.add_argument("--closure", action="store_true",
                    help="learned LAD closure on the 2-D RM case (train, time, judge)")
    ap.add_argument("--closure-n", type=int, default=512, help="--closure: RM grid is nx x n (y periodic)")
    ap.add_argument("--closure-nx", type=int, default=1024,
                    help="--closure: points along the shock (x); 0: int(1.4 n) as examples/RM3D.py")
    ap.add_argument("--closure-precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--closure-snapshots", type=int, default=16)
    ap.add_argument("--closure-every", type=int, default=2)
    ap.add_argument("--closure-train-steps", type=int, default=1500)
    ap.add_argument("--only", default=None,
                    help="--ops: comma-separated name prefixes to run (e.g. d1_x_f64,copy_f64)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only (pyranda_amd has no multi-GPU path); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda:0")
    n = args.n
    if args.ops:
        return ops_bench(n, args.steps, args.warmup, dev, args.only, args.solver, args.r32_solve)
    if args.ensemble:
        return ensemble_bench(args, dev)
    if args.closure:
        return closure_bench(args, dev)
    gamma = 1.4
    p0 = 1.0 / (gamma * 0.1 ** 2)  # Mach 0.1 on the unit velocity

    print("[bench] setting up TG %d^3" % n, file=sys.stderr, flush=True)
    results = {}
    states = {}
    tuned = {}
    wide = None
    dt = None
    precs = [p for p in args.precision.split(",") if p and not (args.no_fp64 and p == "fp64")]
    if "fp32" not in precs:
        precs.append("fp32")
    for precision in precs:
        ss = make_tg(n, precision, p0, args.solver)
        if dt is None:
            dt = float(ss.variables["dt"].data) * 0.5  # CFL 0.5 from the initial state
        ctx = ss.PyMPI.ctx
        if not args.no_autotune:
            # launch knobs only: they leave results unchanged bit for bit (the
            # solve arithmetic is the precision's, not a timing decision)
            rep = ss.autotune()
            tuned[precision] = {"_".join([OPS[key[0]], DIRS[key[1]]] + list(key[2:])):
                                ",".join("%s=%d" % kv for kv in best)
                                for key, (best, _) in sorted(rep.items(), key=lambda kv: str(kv[0]))}
            print("[bench] %s autotune %s" % (precision, tuned[precision]), file=sys.stderr, flush=True)
        print("[bench] %s warmup %d" % (precision, args.warmup), file=sys.stderr, flush=True)
        # the warm-up steps time every operator launch (HIP events on the
        # library's stream) to find the dominant one; the timed steps record
        # events around that operator's launches only, so the rest of the
        # step runs as it does untimed (composites as one call)
        ctx.timing, ctx.timing_only = {}, None
        t = run_steps(ss, dt, args.warmup)
        torch.cuda.synchronize()
        wsumm = ctx.timing_summary()
        dom_key = max(wsumm, key=lambda k: wsumm[k][0] * wsumm[k][1]) if wsumm else None
        ctx.timing = {}
        ctx.timing_only = {dom_key} if dom_key is not None else None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t = run_steps(ss, dt, args.steps, t)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        summ = ctx.timing_summary()
        ctx.timing, ctx.timing_only = None, None
        if dom_key is None:
            dom_key = max(summ, key=lambda k: summ[k][0] * summ[k][1])
        results[precision] = {"elapsed_s": el, "ms_per_step": 1e3 * el / args.steps,
                              "pups": n ** 3 * args.steps / el, "kernels": summ, "dom": dom_key}
        states[precision] = {k: ss.variables[k].data.clone() for k in CONSERVED}
        if precision == "fp32":
            wide = {"f64": sorted(k for k, v in ss._storage.items() if v == "f64"),
                    "f32": sorted(k for k, v in ss._storage.items() if v == "f32")}
        print("[bench] %s %.3f ms/step, %.3e point-updates/s" %
              (precision, 1e3 * el / args.steps, n ** 3 * args.steps / el), file=sys.stderr, flush=True)
        del ss
        torch.cuda.empty_cache()

    head = results["fp32"]
    # dominant operator kernel of the headline run (from the warm-up) and
    # its launches in the timed steps
    dom = head["dom"]
    count, mean_ms = head["kernels"][dom]
    # one read and one write at the storage width; an FP64 sweep that also
    # stores the FP32 rounding writes 4 B/pt more
    bytes_per_pt = 20 if dom[2].endswith("+float32") else (8 if "float32" in dom[2] else 16)
    alg_bytes = n ** 3 * bytes_per_pt
    achieved = alg_bytes / (mean_ms * 1e-3) / 1e9
    # that kernel's share of the timed steps' wall time
    time_share = count * mean_ms / (head["elapsed_s"] * 1e3)

    # every variant against the FP64 run after the same warmup + timed steps
    acc = {}
    if "fp64" in states:
        for prec in states:
            if prec != "fp64":
                acc[prec] = {k: dict(zip(("rel_max", "rel_l2"), rel_err(states[prec][k], states["fp64"][k])))
                             for k in CONSERVED}
    worst = (max(max(v["rel_max"], v["rel_l2"]) for v in acc["fp32"].values()) if "fp32" in acc else None)

    cpu = None
    if not args.no_cpu:
        # every core this process may use; where OMP_NUM_THREADS is set (the
        # host's share of cores per GPU job) it bounds the count
        info = host_info()
        threads = int(info["usable_cores"] or 1)
        if info["omp_num_threads"]:
            threads = min(threads, max(1, int(info["omp_num_threads"])))
        print("[bench] cpu baseline %d^3, %d threads; %d^3, 1 thread" % (args.cpu_n, threads, args.cpu_n1),
              file=sys.stderr, flush=True)
        val, el = cpu_baseline(args.cpu_n, p0, dt, threads, sample_steps=1)
        val1, el1 = cpu_baseline(args.cpu_n1, p0, dt, 1, sample_steps=1)
        cpu = {"value": val, "unit": "point-updates/s", "cores": threads, "kind": "port",
               "sample": "one RK4 step (after one untimed step) of the same Taylor-Green Euler+LAD EOM at "
                         "%d^3 on %d threads: the oracle's FP64 operators (OpenMP over lines) and numpy "
                         "pointwise algebra over z-slabs in %d threads, %.1f s" % (args.cpu_n, threads, threads, el),
               "one_thread": {"value": val1, "sample": "one RK4 step at %d^3 on 1 thread, %.1f s"
                                                       % (args.cpu_n1, el1)},
               "host": info}

    line = {
        "metric": "256^3 Euler+LAD point-updates/s (FP32-refined), 3-D Taylor-Green",
        "value": head["pups"],
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage for operator results (derivatives, LAD ring and filters) and what is "
                 "computed from them, f64 storage for offset-dominated fields, for values computed from "
                 "f64 values only and for flux arguments with f64 inputs; f32 sweeps plus one "
                 "f64-residual refinement in every operator on f32 fields, f64 sweeps on f64 fields; "
                 "f64 RK4 state",
        "data": "synthetic (Taylor-Green initial condition)",
        "config": {"workload": "taylor_green_euler_lad_%d^3" % n, "grid": [n, n, n], "gamma": gamma,
                   "mach": 0.1, "cfl": 0.5, "dt": dt, "rk_stages": 5, "lad": True,
                   "periodic_solver": args.solver, "autotune": tuned.get("fp32"),
                   "storage": wide},
        "fp64": ({"value": results["fp64"]["pups"], "ms_per_step": results["fp64"]["ms_per_step"]}
                 if "fp64" in results else None),
        "speedup_fp32r_over_fp64": (head["pups"] / results["fp64"]["pups"]) if "fp64" in results else None,
        # the other variants: FP32 fields with FP64 solves ("fp32-fp64solve")
        # and any asked for with --precision
        "variants": {p.replace("-", "_"): {
            "value": results[p]["pups"], "ms_per_step": results[p]["ms_per_step"],
            "speedup_over_fp64": (results[p]["pups"] / results["fp64"]["pups"]) if "fp64" in results else None,
            "accuracy_vs_fp64": acc.get(p)} for p in results if p not in ("fp32", "fp64")},
        # the headline run against the FP64 run after warmup + steps RK4 steps
        "accuracy_fp32r_vs_fp64": acc.get("fp32"),
        "accuracy_steps": args.warmup + args.steps,
        "accuracy_worst": worst,
        "accuracy_bar": 1e-6,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (MEASURED_TRAFFIC.get((OPS[dom[0]], DIRS[dom[1]], dom[2])) if n == 256 else None),
                     "kernel": "%s_%s (%s)" % (OPS[dom[0]], DIRS[dom[1]],
                                              dom[2].replace("torch.", "").replace("+float32", ", FP32 copy")),
                     "launches": count, "mean_us": mean_ms * 1e3, "bytes_per_launch": alg_bytes,
                     "kernel_time_share": time_share},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
