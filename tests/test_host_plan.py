"""CPU checks of the library's host side (no GPU needed).

* libpyranda_amd.so loads and exports every symbol include/pyranda_amd.h declares;
* its restated weight tables equal the oracle's (which are pinned to the
  reference build);
* its mesh scalars/coordinates follow parcop.setup/mesh.f90;
* the kernels' per-thread arithmetic (chunked LU and circulant line solves,
  run on the host by pyr_host_emulate) matches the oracle: <=1e-12 relative in
  FP64 and <=1e-6 in the FP32-refined mode, on periodic, bounded, symmetric
  and mixed grids.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from meshgen import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def amd():
    import os
    from pyranda_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    _lib.lib()  # a built library that does not load is a failure, not a skip
    return _lib


def test_exports_every_declared_symbol(amd):
    hdr = open(os.path.join(ROOT, "include", "pyranda_amd.h")).read()
    declared = set(re.findall(r"\b(pyr_[a-z0-9_]+)\s*\(", hdr))
    assert len(declared) > 30
    L = amd.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    # and the Python binding knows every one of them
    assert declared <= set(amd.SIGNATURES), declared - set(amd.SIGNATURES)


@pytest.mark.parametrize("op", range(5))
def test_library_weights_equal_oracle(amd, oracle_mod, op):
    L = amd.lib()
    ali, ari = np.zeros(5), np.zeros(9)
    alb1, alb2 = np.zeros((3, 4, 5)), np.zeros((3, 4, 5))
    arb1, arb2 = np.zeros((3, 4, 9)), np.zeros((3, 4, 9))
    dims = (ctypes.c_int * 6)()
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    amd.check(L.pyr_host_weights(op, p(ali), p(ari), p(alb1), p(alb2), p(arb1), p(arb2), dims))
    w = oracle_mod.weights(op)
    assert list(dims) == [w["nol"], w["nor"], w["ncl"], w["ncr"], w["null_option"], w["implicit_op"]]
    for k, a in (("ali", ali), ("ari", ari), ("alb1", alb1), ("alb2", alb2), ("arb1", arb1), ("arb2", arb2)):
        assert np.array_equal(a, w[k]), k


def _host_ctx(amd, dom, bcs):
    L = amd.lib()
    h = ctypes.c_void_p()
    bc6 = (ctypes.c_int * 6)(*[amd.BC[b] for b in bcs])
    args = []
    for x1, xn, _ in dom:
        args += [x1, xn]
    amd.check(L.pyr_setup_host(ctypes.byref(h), *[d[2] for d in dom], *args, bc6))
    return h


def test_mesh_matches_reference_construction(amd):
    dom = [(0.0, 2 * np.pi * 31 / 32, 32), (0.0, 1.0, 17), (-1.0, 3.0, 1)]
    h = _host_ctx(amd, dom, ["PERI", "PERI", "NONE", "NONE", "NONE", "NONE"])
    L = amd.lib()
    (X, Y, Z), dd = mesh(dom)
    dp, dm, gl = (ctypes.c_double * 3)(), (ctypes.c_double * 3)(), ctypes.c_double()
    amd.check(L.pyr_mesh_scalars(h, dp, dm, ctypes.byref(gl)))
    assert tuple(dp) == dd
    # d3 of a size-1 axis is max(dx,dy) (mesh.f90:204-209); GridLen = min
    assert tuple(dm) == (dd[0], dd[1], max(dd[0], dd[1]))
    assert gl.value == min(dd[0], dd[1])
    for ax, C in enumerate((X[:, 0, 0], Y[0, :, 0], Z[0, 0, :])):
        buf = (ctypes.c_double * len(C))()
        amd.check(L.pyr_axis_coords(h, ax, buf))
        assert np.array_equal(np.array(buf[:]), C)
    L.pyr_destroy(h)


def test_bad_inputs_return_errors(amd):
    L = amd.lib()
    h = ctypes.c_void_p()
    bc6 = (ctypes.c_int * 6)(1, 1, 1, 1, 1, 1)
    assert L.pyr_setup_host(ctypes.byref(h), 0, 4, 4, 0, 1, 0, 1, 0, 1, bc6) != 0
    assert b"grid" in L.pyr_last_error()
    assert L.pyr_setup_host(ctypes.byref(h), 1100, 16, 16, 0, 1, 0, 1, 0, 1, bc6) != 0
    assert b"exceeds 1024" in L.pyr_last_error()
    assert L.pyr_setup_host(ctypes.byref(h), 6, 4, 4, 0, 1, 0, 1, 0, 1, bc6) != 0
    assert b"below 8" in L.pyr_last_error()
    bad = (ctypes.c_int * 6)(1, 1, 7, 1, 1, 1)
    assert L.pyr_setup_host(ctypes.byref(h), 8, 8, 8, 0, 1, 0, 1, 0, 1, bad) != 0
    amd.check(L.pyr_setup_host(ctypes.byref(h), 8, 8, 8, 0, 1, 0, 1, 0, 1, bc6))
    # a host-only context refuses device work instead of touching memory
    assert L.pyr_apply(h, 0, 0, 0, ctypes.c_void_p(1), ctypes.c_void_p(2), 0, 0.0) != 0
    assert b"host-only" in L.pyr_last_error()
    L.pyr_destroy(h)


CASES = [
    ((33, 40, 20), ("PERI", "PERI", "PERI")),
    ((33, 40, 20), ("NONE", "NONE", "NONE")),
    ((64, 100, 16), ("PERI", "NONE", "PERI")),
    ((256, 8, 9), ("NONE", "PERI", "NONE")),
    ((24, 9, 12), ("SYMM", "SYMM", "NONE")),
    # shortest lines the reference schemes define (r4 boundary rows read up to
    # row 8, compact_r4.f90:117-120) and a null (n<4) direction
    ((8, 8, 3), ("PERI", "NONE", "NONE")),
    # P = 16 chunks per periodic line
    ((512, 9, 8), ("PERI", "PERI", "NONE")),
    # the longest lines, 32 chunks (periodic and bounded; LU tables read from
    # global memory on the device), and an uneven partition of 1000 points
    ((1024, 8, 3), ("PERI", "NONE", "NONE")),
    ((8, 1000, 3), ("NONE", "NONE", "NONE")),
    ((716, 8, 3), ("NONE", "PERI", "NONE")),
]


@pytest.mark.parametrize("solver", ["circulant", "lu"])
@pytest.mark.parametrize("nn,kinds", CASES)
def test_host_emulation_matches_oracle(amd, oracle_mod, nn, kinds, solver):
    if solver == "lu" and "PERI" not in kinds:
        pytest.skip("the periodic solver choice only affects periodic lines")
    po = oracle_mod
    periodic = [k == "PERI" for k in kinds]
    sym = [[k == "SYMM", k == "SYMM"] for k in kinds]
    dom = [(0.0, 2 * np.pi * (n - 1) / n if p else 1.0, n) for n, p in zip(nn, periodic)]
    (X, Y, Z), dd = mesh(dom)
    bcs = []
    for k in kinds:
        bcs += [k, k]
    h = _host_ctx(amd, dom, bcs)
    L = amd.lib()
    amd.check(L.pyr_set_periodic_solver(h, amd.PSOLVE[solver]))
    P = po.Oracle(nn, dd, periodic, symmetric=sym)
    rng = np.random.default_rng(sum(nn))
    v = np.asfortranarray(np.sin(X) * np.cos(2 * Y) + 0.3 * np.sin(3 * Z + 1)
                          + 0.05 * rng.standard_normal(nn) + 2.0)
    v32 = np.asfortranarray(v.astype(np.float32))
    L = amd.lib()
    for op in range(5):
        for d in range(3):
            ref = P.apply(op, d, v)
            out = np.zeros(nn, order="F")
            amd.check(L.pyr_host_emulate(h, op, d, amd.F64, v.ctypes.data, out.ctypes.data))
            scale = max(np.abs(ref).max(), 1e-300)
            assert np.abs(out - ref).max() / scale <= 1e-12, ("f64", op, d)
            ref32 = P.apply(op, d, v32.astype(np.float64))
            scale = max(np.abs(ref32).max(), 1e-300)
            # FP32 fields: refined FP32 solve (<=1e-6), or FP64 solve (one
            # rounding of each result, <= 2^-23 of max|ref|)
            for mode, tol in (("refined", 1e-6), ("fp64", 1.2e-7)):
                amd.check(L.pyr_set_r32_solve(h, amd.R32_SOLVE[mode]))
                o32 = np.zeros(nn, dtype=np.float32, order="F")
                amd.check(L.pyr_host_emulate(h, op, d, amd.R32, v32.ctypes.data, o32.ctypes.data))
                assert np.abs(o32.astype(np.float64) - ref32).max() / scale <= tol, ("r32", mode, op, d)
            amd.check(L.pyr_set_r32_solve(h, amd.R32_SOLVE["refined"]))
    L.pyr_destroy(h)
