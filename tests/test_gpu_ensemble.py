"""Device-resident time steps and graph-captured ensembles on the GPU.

* A KH.py simulation with its scalars kept on the device (dt from the fused
  reductions as 0-d tensors, rk4 reading dt from device memory) is bit for
  bit the same as the host-scalar run.
* An ensemble of KH layers on uneven grids, one HIP graph replay per time
  step for all members, is bit for bit each member run alone, after every
  replay, with eager work and host synchronisation between the replays.
"""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyranda_amd
    return pyranda_amd


def _same(a, b):
    a = a if torch.is_tensor(a) else torch.tensor(float(a), dtype=torch.float64)
    b = b if torch.is_tensor(b) else torch.tensor(float(b), dtype=torch.float64)
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return a.shape == b.shape and torch.equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_scalars_match_host(pr, precision):
    from pyranda_amd.workloads import kh_advance, kh_run, kh_setup
    _say("device scalars vs host scalars, KH 64x48", precision)
    a = kh_setup(pr, 64, 48, precision=precision)
    b = kh_setup(pr, 64, 48, precision=precision, device_scalars=True)
    assert torch.is_tensor(b.var("dt").data) and b.var("dt").data.is_cuda
    assert _same(a.var("dt").data, b.var("dt").data)
    kh_run(a, max_steps=6)
    st = {"time": torch.zeros((), dtype=torch.float64, device="cuda"), "dt": b.var("dt").data.clone()}
    # the IC dt is the starting dt of the loop in both runs
    for _ in range(6):
        kh_advance(b, st)
    torch.cuda.synchronize()
    for k in ("rho", "rhou", "Et", "kappa", "mu", "beta", "dt"):
        assert _same(a.var(k).data, b.var(k).data), k


SIZES = [(64, 48), (40, 72), (96, 64), (48, 48)]


@pytest.mark.parametrize("precision,concurrent", [("fp64", True), ("fp32", True), ("fp32", False)])
def test_ensemble_graph_matches_members_alone(pr, precision, concurrent):
    from pyranda_amd.ensemble import pyrandaEnsemble
    from pyranda_amd.workloads import kh_advance, kh_ensemble
    _say("ensemble graph vs members alone", precision, "concurrent" if concurrent else "serial", SIZES)
    sims, state = kh_ensemble(pr, SIZES, precision=precision)
    solo, solo_state = kh_ensemble(pr, SIZES, precision=precision)
    ens = pyrandaEnsemble(sims, kh_advance, state)
    ens.step_eager(1)                     # compiles kernels, allocates work buffers
    for s, st in zip(solo, solo_state):
        kh_advance(s, st)
    ens.capture(concurrent=concurrent)
    junk = torch.randn(1 << 20, device="cuda")
    for rep in range(4):
        ens.step()
        # eager work and a host synchronisation between replays
        junk = junk * 0.5 + 1.0
        torch.cuda.synchronize()
        for s, st in zip(solo, solo_state):
            kh_advance(s, st)
        torch.cuda.synchronize()
        for i, (m, s) in enumerate(zip(ens.members, solo)):
            for k in ("rho", "rhou", "rhov", "Et", "kappa", "dt"):
                assert _same(m.var(k).data, s.var(k).data), (rep, i, k)
            assert _same(ens.state[i]["dt"], solo_state[i]["dt"]), (rep, i, "driver dt")
            assert _same(ens.state[i]["time"], solo_state[i]["time"]), (rep, i, "time")
        _say("replay", rep, "ok")
    # an eager ensemble step after replays continues the same trajectory
    ens.step_eager(1)
    for s, st in zip(solo, solo_state):
        kh_advance(s, st)
    ens.step()
    for s, st in zip(solo, solo_state):
        kh_advance(s, st)
    torch.cuda.synchronize()
    for i, (m, s) in enumerate(zip(ens.members, solo)):
        assert _same(m.var("rho").data, s.var("rho").data), i


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_group_sim_matches_members_alone(pr, precision):
    """The ensemble as ONE pyrandaGroupSim (shared allocation, grouped
    operator launches, per-member dt through segmented reductions) is bit for
    bit each member run alone, step after step."""
    from pyranda_amd.workloads import kh_advance, kh_ensemble, kh_group
    sizes = [(64, 48), (40, 72), (96, 64), (48, 48), (72, 40), (100, 36)]
    _say("group sim vs members alone", precision, sizes)
    g, gst = kh_group(pr, sizes, precision=precision)
    solo, solo_state = kh_ensemble(pr, sizes, precision=precision)
    for m in range(len(sizes)):
        assert _same(g.var("dt").data[m], solo[m].var("dt").data), m
    for step in range(4):
        kh_advance(g, gst)
        for s, st in zip(solo, solo_state):
            kh_advance(s, st)
        torch.cuda.synchronize()
        for m, s in enumerate(solo):
            for k in ("rho", "rhou", "rhov", "Et", "kappa", "mu", "beta"):
                assert _same(g.member_field(k, m), s.var(k).data), (step, m, k)
            assert _same(gst["dt"][m], solo_state[m]["dt"]), (step, m, "dt")
            assert _same(gst["time"][m], solo_state[m]["time"]), (step, m, "time")
        _say("step", step, "ok")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_wall_group_matches_members_alone(pr, precision):
    """Wall-bounded KH layers (periodic x, bounded y with bc.extrap/bc.const)
    on uneven grids up to 1024 points per line, as ONE pyrandaGroupSim
    (grouped launches split by chunk count, grouped bounded LU, one
    gather/scatter per boundary update): bit for bit each member alone."""
    from pyranda_amd.workloads import kh_advance, kh_wall_ensemble, kh_wall_group
    sizes = [(128, 96), (1024, 72), (64, 1000), (200, 130), (40, 48)]
    _say("wall group vs members alone", precision, sizes)
    g, gst = kh_wall_group(pr, sizes, precision=precision)
    solo, solo_state = kh_wall_ensemble(pr, sizes, precision=precision)
    for m in range(len(sizes)):
        assert _same(g.var("dt").data[m], solo[m].var("dt").data), m
        for k in ("rho", "u", "v", "Et"):
            assert _same(g.member_field(k, m), solo[m].var(k).data), ("ic", m, k)
    for step in range(3):
        kh_advance(g, gst)
        for s, st in zip(solo, solo_state):
            kh_advance(s, st)
        torch.cuda.synchronize()
        for m, s in enumerate(solo):
            for k in ("rho", "rhou", "rhov", "Et", "u", "v", "kappa", "mu", "beta"):
                assert _same(g.member_field(k, m), s.var(k).data), (step, m, k)
            assert _same(gst["dt"][m], solo_state[m]["dt"]), (step, m, "dt")
        _say("step", step, "ok")
    # the walls hold: v = 0 on both wall planes of every member
    for m in range(len(sizes)):
        v = g.member_field("v", m)
        assert float(v[:, 0, :].abs().max()) == 0.0 and float(v[:, -1, :].abs().max()) == 0.0


def test_wall_group_full_ensemble_matches_members_alone(pr):
    """The full wall-bounded KH ensemble: 64 members, each axis log-uniform in
    [128, 1024] (8.6M points), FP32 fields with the refined solve, as one
    pyrandaGroupSim: after each of two steps every member's conserved fields
    and dt equal that member run alone, bit for bit."""
    from pyranda_amd.workloads import ensemble_sizes, kh_advance, kh_wall_ensemble, kh_wall_group
    sizes = ensemble_sizes(64, 128, 1024)
    _say("full wall ensemble", len(sizes), "members,", sum(a * b for a, b in sizes), "points")
    g, gst = kh_wall_group(pr, sizes, precision="fp32")
    solo, solo_state = kh_wall_ensemble(pr, sizes, precision="fp32")
    for step in range(2):
        kh_advance(g, gst)
        for s, st in zip(solo, solo_state):
            kh_advance(s, st)
        torch.cuda.synchronize()
        for m, s in enumerate(solo):
            for k in ("rho", "rhou", "rhov", "Et"):
                assert _same(g.member_field(k, m), s.var(k).data), (step, m, sizes[m], k)
            assert _same(gst["dt"][m], solo_state[m]["dt"]), (step, m, "dt")
        _say("step", step, "ok")
    for m in range(len(sizes)):
        v = g.member_field("v", m)
        assert float(v[:, 0, :].abs().max()) == 0.0 and float(v[:, -1, :].abs().max()) == 0.0


def test_wall_group_graph_replays_match_eager(pr):
    """One step of a wall-bounded group (grouped bounded LU, BC gathers and
    scatters, segmented reductions) captured as a HIP graph: each replay, with
    eager work and host synchronisation between replays, is bit for bit the
    eager step of an identical group."""
    from pyranda_amd.ensemble import pyrandaEnsemble
    from pyranda_amd.workloads import kh_advance, kh_wall_group
    sizes = [(128, 96), (200, 130), (64, 48)]
    g, gst = kh_wall_group(pr, sizes, precision="fp32")
    ref, rst = kh_wall_group(pr, sizes, precision="fp32")
    ens = pyrandaEnsemble([g], kh_advance, [gst])
    ens.step_eager(1)
    kh_advance(ref, rst)
    ens.capture(concurrent=False)
    junk = torch.randn(1 << 20, device="cuda")
    for rep in range(3):
        ens.step()
        junk = junk * 0.5 + 1.0
        torch.cuda.synchronize()
        kh_advance(ref, rst)
        torch.cuda.synchronize()
        for k in ("rho", "rhou", "rhov", "Et", "v"):
            assert _same(g.var(k).data, ref.var(k).data), (rep, k)
        assert _same(ens.state[0]["dt"], rst["dt"]), (rep, "dt")
        _say("replay", rep, "ok")


def test_group_mode_switch_builds_new_plans(pr):
    """A grouped FP32 operator applied under the refined solve, then with the
    members switched to FP64 solves: each result equals every member's own
    apply in that mode (the group's plans are keyed by the solve mode)."""
    from pyranda_amd import _lib
    from pyranda_amd.workloads import kh_group
    sizes = [(64, 48), (40, 72)]
    g, _ = kh_group(pr, sizes, precision="fp32")
    grp = g.PyMPI.group
    f = g.var("rho").data
    for mode in ("refined", "fp64", "refined"):
        for mem in g.members:
            mem.ctx.set_r32_solve(mode)
        out = grp.apply(_lib.OP_D1, 0, f)
        torch.cuda.synchronize()
        for m, mem in enumerate(g.members):
            want = mem.ctx.apply(_lib.OP_D1, 0, g.member_field("rho", m).contiguous())
            assert _same(_member(g, out, m), want), (mode, m)
        _say("mode", mode, "ok")


def _member(g, field, m):
    mem = g.members[m]
    nx, ny, nz = mem.nn
    off = g.offsets[m]
    return field.reshape(-1)[off:off + mem.npts].reshape(nz, ny, nx).permute(2, 1, 0)
