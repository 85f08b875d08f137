"""Benchmark: 3-D Taylor-Green Euler+LAD, 256^3 on one MI355X.

A step is one 5-stage RK4 step of the Taylor-Green EOM of
examples/TaylorGreen.py (gamma=1.4, LAD bulk viscosity via ring/gbar, fbar
on the conserved variables after every stage) with Mach 0.1
(p0 = 1/(gamma*0.1^2)), CFL 0.5 from the initial state, run through
pyranda_amd.pyrandaSim with every field resident in HBM.

The headline value is FP32-refined point-updates/s (FP32 fields and sweeps,
one FP64-residual refinement per operator, FP64 RK4 state). The FP64 path is
timed beside it on the same steps, and the FP32-refined state is compared
with the FP64 state after the same steps (rel_max / rel_L2).

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
OPS = ["d1", "d2", "d8", "fbar", "gbar"]
DIRS = "xyz"


def tg_mesh(n):
    return """
    xdom = (0.0, 2*pi*FF,  Npts, periodic=True)
    ydom = (0.0, 2*pi*FF,  Npts, periodic=True)
    zdom = (0.0, 2*pi*FF,  Npts, periodic=True)
    """.replace("Npts", str(n)).replace("pi", str(math.pi)).replace("FF", str(float(n - 1) / n))


def make_tg(n, precision, p0):
    import pyranda_amd as pr
    from pyranda_amd.workloads import TG_EOM, TG_IC
    ss = pr.pyrandaSim("TGvortex", tg_mesh(n), silent=True, precision=precision)
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
    from oracle.tg_cpu import TaylorGreenCPU
    tg = TaylorGreenCPU(n, p0=p0, threads=threads)
    t0 = time.perf_counter()
    t = 0.0
    for _ in range(sample_steps):
        t = tg.rk4(t, dt)
    el = time.perf_counter() - t0
    return n ** 3 * sample_steps / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=128,
                    help="grid of the CPU-baseline sample (one RK4 step)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda:0")
    n = args.n
    gamma = 1.4
    p0 = 1.0 / (gamma * 0.1 ** 2)  # Mach 0.1 on the unit velocity

    print("[bench] setting up TG %d^3" % n, file=sys.stderr, flush=True)
    results = {}
    states = {}
    dt = None
    for precision in (["fp32"] if args.no_fp64 else ["fp64", "fp32"]):
        ss = make_tg(n, precision, p0)
        if dt is None:
            dt = float(ss.variables["dt"].data) * 0.5  # CFL 0.5 from the initial state
        ctx = ss.PyMPI.ctx
        print("[bench] %s warmup %d" % (precision, args.warmup), file=sys.stderr, flush=True)
        t = run_steps(ss, dt, args.warmup)
        torch.cuda.synchronize()
        ctx.timing = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t = run_steps(ss, dt, args.steps, t)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        summ = ctx.timing_summary()
        ctx.timing = None
        results[precision] = {"elapsed_s": el, "ms_per_step": 1e3 * el / args.steps,
                              "pups": n ** 3 * args.steps / el, "kernels": summ}
        states[precision] = {k: ss.variables[k].data.clone() for k in ("rho", "rhou", "Et")}
        print("[bench] %s %.3f ms/step, %.3e point-updates/s" %
              (precision, 1e3 * el / args.steps, n ** 3 * args.steps / el), file=sys.stderr, flush=True)
        del ss
        torch.cuda.empty_cache()

    head = results["fp32"]
    # dominant operator kernel of the headline run
    kern = head["kernels"]
    tot = {k: c * ms for k, (c, ms) in kern.items()}
    dom = max(tot, key=tot.get)
    count, mean_ms = kern[dom]
    bytes_per_pt = 8 if "float32" in dom[2] else 16
    alg_bytes = n ** 3 * bytes_per_pt
    achieved = alg_bytes / (mean_ms * 1e-3) / 1e9
    op_share = sum(tot.values()) / (head["elapsed_s"] * 1e3)

    acc = {}
    if "fp64" in states:
        for k in states["fp32"]:
            acc[k] = rel_err(states["fp32"][k], states["fp64"][k])

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        print("[bench] cpu baseline %d^3 x 2 steps, %d threads" % (args.cpu_n, threads),
              file=sys.stderr, flush=True)
        val, el = cpu_baseline(args.cpu_n, p0, dt, threads, sample_steps=2)
        cpu = {"value": val, "unit": "point-updates/s", "cores": threads, "kind": "port",
               "sample": "two RK4 steps of the same Taylor-Green Euler+LAD EOM at %d^3 (FP64 oracle "
                         "operators with OpenMP + numpy pointwise), %.1f s" % (args.cpu_n, el)}

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
        "dtype": "f32 fields and sweeps, f64 residual refinement, f64 RK4 state",
        "data": "synthetic (Taylor-Green initial condition)",
        "config": {"workload": "taylor_green_euler_lad_%d^3" % n, "grid": [n, n, n], "gamma": gamma,
                   "mach": 0.1, "cfl": 0.5, "dt": dt, "rk_stages": 5, "lad": True},
        "fp64": ({"value": results["fp64"]["pups"], "ms_per_step": results["fp64"]["ms_per_step"]}
                 if "fp64" in results else None),
        "speedup_fp32r_over_fp64": (head["pups"] / results["fp64"]["pups"]) if "fp64" in results else None,
        "accuracy_fp32r_vs_fp64": {k: {"rel_max": a, "rel_l2": b} for k, (a, b) in acc.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "%s_%s (%s)" % (OPS[dom[0]], DIRS[dom[1]], dom[2].replace("torch.", "")),
                     "launches": count, "mean_us": mean_ms * 1e3, "bytes_per_launch": alg_bytes,
                     "operator_time_share": op_share},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
