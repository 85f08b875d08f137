"""The reference-side binding (pyranda_amd.f2py_parcop.parcop, the stand-in
for pyranda's f2py module parcop.parcop) called as pyrandaMPI.py and
pyrandaMesh.py call it: host numpy arrays in, new Fortran-order arrays out,
every operator against the oracle (FP64, 1e-12), on a block with periodic,
bounded and symmetric directions, and two patches switched with set_patch."""
import sys

import numpy as np
import pytest
import torch

from meshgen import mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyranda_amd import _lib
    _lib.lib()
    return torch.device("cuda:0")


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_parcop_stub_matches_oracle(dev, oracle_mod):
    from pyranda_amd.f2py_parcop import parcop
    nn = (24, 20, 16)
    L = 2 * np.pi * (nn[0] - 1) / nn[0]
    dom = [(0.0, L, nn[0]), (0.0, 1.0, nn[1]), (0.0, 2.0, nn[2])]
    (X, Y, Z), dd = mesh(dom)
    # parcop.setup(patch, level, comm, nx, ny, nz, px, py, pz, coordsys, x1..zn, bx1..bzn)
    parcop.setup(0, 0, None, *nn, 1, 1, 1, 0, 0.0, L, 0.0, 1.0, 0.0, 2.0,
                 "PERI", "PERI", "NONE", "NONE", "SYMM", "SYMM")
    P = oracle_mod.Oracle(nn, dd, (True, False, False), symmetric=[[False, False], [False, False], [True, True]])
    rng = np.random.default_rng(5)
    v = np.asfortranarray(np.sin(X) * np.cos(3 * Y) + Z ** 2 + 0.05 * rng.standard_normal(nn))
    w = np.asfortranarray(np.cos(X) * Y + 0.05 * rng.standard_normal(nn))
    u = np.asfortranarray(Z * np.sin(2 * X) + 0.05 * rng.standard_normal(nn))
    checks = [("ddx", parcop.ddx(v), P.ddx(v)), ("ddy", parcop.ddy(v), P.ddy(v)), ("ddz", parcop.ddz(v), P.ddz(v)),
              ("dd8x", parcop.dd8x(v), P.d8(v, 0)), ("dd8z", parcop.dd8z(v), P.d8(v, 2)),
              ("plaplacian", parcop.plaplacian(v), P.laplacian(v)), ("pring", parcop.pring(v), P.ring(v)),
              ("pringv", parcop.pringv(v, w, u), P.ringv(v, w, u)),
              ("sfilter", parcop.sfilter(v), P.sfilter(v)), ("gfilter", parcop.gfilter(v), P.gfilter(v)),
              ("gfilterdir2", parcop.gfilterdir(v, 2), P.gfilterdir(v, 2)),
              ("divergence", parcop.divergence(v, w, u), P.divergence(v, w, u))]
    for g, r in zip(parcop.grads(v), P.grads(v)):
        checks.append(("grads", g, r))
    for name, got, ref in checks:
        assert isinstance(got, np.ndarray) and got.flags.f_contiguous and got.shape == nn, name
        assert _rel(got, ref) <= 1e-12, (name, _rel(got, ref))
    assert np.abs(parcop.xgrid() - X).max() <= 1e-15 and np.abs(parcop.zgrid() - Z).max() <= 1e-15
    assert np.allclose(parcop.dygrid(), dd[1], rtol=1e-15, atol=0)
    # a second patch, then back to the first
    parcop.setup(1, 0, None, 16, 16, 1, 1, 1, 1, 0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0,
                 "NONE", "NONE", "PERI", "PERI", "NONE", "NONE")
    assert parcop.ddx(np.zeros((16, 16, 1))).shape == (16, 16, 1)
    parcop.set_patch(0, 0)
    assert _rel(parcop.ddy(v), P.ddy(v)) <= 1e-12
    print("parcop stub ok:", len(checks), "operator checks", flush=True)
