"""Restart and viz file formats on the host (pyranda_amd/pyrandaIO.py).

* serial.dat is one line of str(dict) as pyranda.py:475-557 writes it; it is
  read back by a literal parser (no eval), which also takes numpy < 2's
  array(..., dtype=int32) form of the reference's procMap;
* proc-00000.bin is the C-order float64 (nx, ny, nz, nvars) array that
  readChunk reshapes (pyranda.py:1125-1126);
* the binary VTK file has the header lines, big-endian float32 TIME/CYCLE,
  points and scalars of pyrandaIO.makeDumpVTKbin (pyrandaIO.py:115-217),
  including its 2-D remapping.
"""
import os

import numpy as np

from pyranda_amd import pyrandaIO as io


def test_serial_roundtrip_and_reference_forms(tmp_path):
    data = {"mesh": {"x1": [0.0, 0.0, 0.0], "nn": [16, 8, 1], "periodic": np.array([True, False, False]),
                     "function": None, "dim": 2},
            "EOM": "ddt(:a:) = -ddx(:a:)\n", "decomp": [1, 1, 1], "time": 0.125, "cycle": 7,
            "procMap": {"0-g1": np.array([0, 0, 0], dtype=np.int32), "0-gn": np.array([15, 7, 0], dtype=np.int32)},
            "vizDumpHistory": [[0, 0.0], [5, 0.0625]], "vars": {"a": 0, "b": 1}, "deltat": np.float64(0.025)}
    p = os.path.join(tmp_path, "serial.dat")
    io.write_serial(p, data)
    text = open(p).read()
    assert "\n" not in text and "array([0, 0, 0], dtype=int32)" in text and "np.float64" not in text
    back = io.read_serial(p)
    assert back["procMap"]["0-gn"].dtype == np.int32 and list(back["procMap"]["0-gn"]) == [15, 7, 0]
    assert back["deltat"] == 0.025 and back["vars"] == {"a": 0, "b": 1} and back["mesh"]["function"] is None
    assert list(back["mesh"]["periodic"]) == [True, False, False]
    # forms other writers produce: numpy >= 2 scalar reprs, nan/inf, negative numbers
    with open(p, "w") as f:
        f.write("{'t': np.float64(1.5), 'n': nan, 'i': -inf, 'k': -3, 'a': array([1.0, 2.0])}")
    b = io.read_serial(p)
    assert b["t"] == 1.5 and b["n"] != b["n"] and b["i"] == float("-inf") and b["k"] == -3
    # anything that is not a literal is refused
    with open(p, "w") as f:
        f.write("{'x': __import__('os').getcwd()}")
    try:
        io.read_serial(p)
        raise AssertionError("a call was evaluated")
    except ValueError:
        pass


def test_proc_bin_layout_is_readchunk_layout(tmp_path):
    nn = (5, 4, 3)
    rng = np.random.default_rng(0)
    a = np.asfortranarray(rng.standard_normal(nn))
    b = np.asfortranarray(rng.standard_normal(nn))
    p = os.path.join(tmp_path, "proc-00000.bin")
    io.write_proc_bin(p, [a, 2.5, b], nn)
    # readChunk: numpy.reshape(numpy.fromfile(fid), (ax,ay,az,nvars), order='C')
    DATA = np.reshape(np.fromfile(p), nn + (3,), order="C")
    assert np.array_equal(DATA[..., 0], a) and np.all(DATA[..., 1] == 2.5) and np.array_equal(DATA[..., 2], b)
    mm = io.open_proc_bin(p, nn, 3)
    assert np.array_equal(mm[..., 2], b)


def test_vtk_binary_layout(tmp_path):
    nn = (6, 5, 4)
    x, y, z = np.meshgrid(np.arange(6) * 0.5, np.arange(5) * 0.25, np.arange(4) * 2.0, indexing="ij")
    f = np.asfortranarray(np.sin(x) + y * z)
    p = os.path.join(tmp_path, "d.vtk")
    io.write_vtk_binary(p, [x, y, z], [f], ["phi"], cycle=12, time=0.375)
    head = open(p, "rb").read()
    assert head.startswith(b"# vtk DataFile Version 3.0 \nvtk output \nBINARY \nDATASET STRUCTURED_GRID \n"
                           b"DIMENSIONS  6 5 4  \nFIELD FieldData 2 \nTIME 1 1 float \n")
    dims, t, cyc, pts, fields = io.read_vtk_binary(p)
    assert dims == nn and t == 0.375 and cyc == 12
    # points k, j, i with i fastest
    want = np.stack([x, y, z], -1).transpose(2, 1, 0, 3).reshape(-1, 3).astype(np.float32)
    assert np.array_equal(pts, want.astype(np.float64))
    assert np.array_equal(fields["phi"], f.transpose(2, 1, 0).reshape(-1).astype(np.float32).astype(np.float64))
    # a y-z plane dataset is written in the first two dimensions, z set to 0
    x2, y2, z2 = (a[:1] for a in (x, y, z))
    p2 = os.path.join(tmp_path, "e.vtk")
    io.write_vtk_binary(p2, [x2, y2, z2], [f[:1]], ["phi"], cycle=0, time=0.0)
    dims2, _, _, pts2, f2 = io.read_vtk_binary(p2)
    assert dims2 == (5, 4, 1)
    assert np.array_equal(pts2[:, 0], y2[0].T.reshape(-1).astype(np.float32))
    assert np.array_equal(pts2[:, 1], z2[0].T.reshape(-1).astype(np.float32)) and np.all(pts2[:, 2] == 0)


def test_visit_file(tmp_path):
    p = os.path.join(tmp_path, "pyranda.visit")
    io.write_visit(p, [[0, 0.0], [20, 0.5]])
    assert open(p).read() == ("!NBLOCKS 1 \n!TIME 0.0 \nvis0000000/proc-000000.0000000.vtk\n"
                              "!TIME 0.5 \nvis0000020/proc-000000.0000020.vtk\n")
