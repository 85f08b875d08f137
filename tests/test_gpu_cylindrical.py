"""Cylindrical meshes (coordsys 1) on the GPU: the device compositions of
div, divT, grad, lap, ring and fbar (pyranda_amd/cylindrical.py) against the
oracle's restatement (oracle/cyl_cpu.py, pinned by closed forms in
tests/test_oracle.py) on the same fields, FP64, <= 1e-12 relative; the
reference-held RZ curve runs in tests/test_gpu_frontend.py."""
import numpy as np
import pytest
import torch

from meshgen import mesh

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pr():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pyranda_amd
    return pyranda_amd


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asfortranarray(a).transpose(2, 1, 0))).cuda().permute(2, 1, 0)


def _host(t):
    return t.permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0)


@pytest.mark.parametrize("nn,x1", [((64, 1, 48), None), ((40, 24, 32), 0.05)])
def test_cylindrical_operators_match_oracle(pr, oracle_mod, nn, x1):
    """RZ (one theta point, the reference script's grid with its SYMM face
    just off the axis) and a 3-D (r, theta, z) block away from the axis."""
    from oracle.cyl_cpu import CylOracle
    nx, ny, nz = nn
    L = 2 * np.pi
    r1 = L / (nx - 1) / 2.0 if x1 is None else x1
    th = (-0.1, 0.1) if ny == 1 else (0.0, 0.5)
    opts = {"coordsys": 1, "periodic": np.array([False, False, True]), "x1": [r1, th[0], 0.0],
            "xn": [L, th[1], L], "nn": list(nn),
            "symmetric": [[x1 is None, False], [False, False], [False, False]]}
    ss = pr.pyrandaSim("cyl", opts, silent=True)
    _say("cylindrical", nn, "first radius", float(ss.PyMPI.metrics.radius[0, 0, 0]))
    dom = [(r1, L, nx), (th[0], th[1], ny), (0.0, L, nz)]
    (R, T, Z), dd = mesh(dom)
    sym = [[x1 is None, False], [False, False], [False, False]]
    P = oracle_mod.Oracle(nn, ss.PyMPI.ctx.dpatch, (False, False, True), symmetric=sym)
    C = CylOracle(P, R, ss.PyMPI.ctx.dpatch, isym=(-1 if x1 is None else 1, 1, 1))
    assert np.allclose(_host(ss.PyMPI.metrics.radius), R, rtol=0, atol=1e-14)
    rng = np.random.default_rng(sum(nn))
    f = [np.asfortranarray(np.cos(R) * np.sin(Z + k) + 0.1 * np.cos(2 * T + k) + 0.01 * rng.standard_normal(nn))
         for k in range(9)]
    fd = [_dev(a) for a in f]
    der, fil = ss.PyMPI.der, ss.PyMPI.fil

    def close(got, want, what):
        err = np.abs(_host(got) - want).max() / max(np.abs(want).max(), 1e-300)
        _say("  ", what, err)
        assert err <= 1e-12, what

    close(der.div(fd[0], fd[1], fd[2]), C.div(f[0], f[1], f[2]), "div")
    for k, (g, w) in enumerate(zip(der.divT(*fd), C.divT(*f))):
        close(g, w, "divT %d" % k)
    for k, (g, w) in enumerate(zip(der.grad(fd[3]), C.grad(f[3]))):
        close(g, w, "grad %d" % k)
    close(der.laplacian(fd[4]), C.laplacian(f[4]), "lap")
    close(der.ring(fd[5]), C.ring(f[5]), "ring")
    close(fil.filter(fd[6]), C.fbar(f[6]), "fbar")


def test_cylindrical_mesh_rejects_axis_point(pr):
    """A mesh point on the axis (r = 0, placed within rounding of it by the
    grid formula) would divide by zero in the metrics; set-up raises."""
    opts = {"coordsys": 1, "periodic": np.array([False, False, True]), "x1": [0.0, 0.0, 0.0],
            "xn": [1.0, 0.5, 1.0], "nn": [16, 8, 8]}
    with pytest.raises(ValueError):
        pr.pyrandaSim("cyl_bad", opts, silent=True)
