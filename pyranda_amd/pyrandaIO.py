"""Restart and visualisation files of a pyrandaSim (pyranda/pyranda.py:431-588,
979-1171; pyranda/pyrandaIO.py:16-230), on one GPU.

Restart (the reference's layout, one rank):
  <rootname>/restart_<cycle:07d>/serial.dat   one line: str() of a dict with
      mesh, EOM, ICs, decomp, procMap, packages, time, deltat, cycle,
      vizDumpHistory, sMap, vars (name -> column) and optionally local_vars
  <rootname>/restart_<cycle:07d>/proc-00000.bin   float64, C order of an
      (nx, ny, nz, nvars) array: every variable, sorted by name, one column each
  <rootname>/restart_<cycle:07d>/<name>.npy   numpy arrays passed as ivars
Fields move between the GPU and the file one variable at a time through a
memory map of proc-00000.bin, so a large grid never needs the interleaved
array in host memory.

Viz (pyrandaIO.py:115-217, BINARY): <rootname>/vis<cycle:07d>/proc-000000.<cycle:07d>.vtk,
a legacy VTK structured grid with big-endian float32 points and scalars and
TIME/CYCLE field data, plus <rootname>/pyranda.visit listing every dump.

serial.dat is read back with a literal parser (numbers, strings, lists,
dicts, tuples, and numpy's array(...)/float64(...) forms), never with eval.
"""
import ast
import os

import numpy as np
import torch

RESTART_INT = 7   # zero padding of the cycle in restart directory names
VIS_INT = 7       # and in viz directory/file names
PROC_INT = 6      # zero padding of the rank in viz file names


class pyrandaIO:
    """Root directory of a simulation's files (created on first write)."""

    def __init__(self, rootname):
        self.rootname = rootname

    def ensure_root(self):
        os.makedirs(self.rootname, exist_ok=True)


# --- serial.dat --------------------------------------------------------------
def _plain(obj):
    """Python builtins for numpy scalars and arrays (str() then prints them as
    numpy < 2 did: plain numbers, array([...]) lists)."""
    if isinstance(obj, dict):
        return {k: _plain(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        t = [_plain(v) for v in obj]
        return t if isinstance(obj, list) else tuple(t)
    if isinstance(obj, np.ndarray):
        return _NpArray(obj)
    if isinstance(obj, np.generic):
        return obj.item()
    if torch.is_tensor(obj):
        return obj.item() if obj.dim() == 0 else _NpArray(obj.detach().cpu().numpy())
    return obj


class _NpArray:
    """An array written as numpy < 2 prints it: array([...]) with a dtype for
    non-default integer types."""

    def __init__(self, a):
        self.a = np.asarray(a)

    def __repr__(self):
        body = repr(self.a.tolist())
        if self.a.dtype == np.int32:
            return "array(%s, dtype=int32)" % body
        return "array(%s)" % body


_NP_CALLS = {"array": lambda *a, dtype=None: np.array(*a, dtype=dtype),
             "float64": np.float64, "float32": np.float32, "int64": np.int64, "int32": np.int32,
             "bool_": np.bool_}
_NP_DTYPES = {"int32": np.int32, "int64": np.int64, "float64": np.float64, "float32": np.float32,
              "bool_": np.bool_, "bool": np.bool_}
_NAMES = {"nan": float("nan"), "inf": float("inf"), "True": True, "False": False, "None": None}


def _literal(node):
    """Value of a literal expression node; raises ValueError on anything else."""
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.Dict):
        return {_literal(k): _literal(v) for k, v in zip(node.keys, node.values)}
    if isinstance(node, ast.List):
        return [_literal(e) for e in node.elts]
    if isinstance(node, ast.Tuple):
        return tuple(_literal(e) for e in node.elts)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _literal(node.operand)
        if not isinstance(v, (int, float)):
            raise ValueError("unary sign on a non-number")
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.Name) and node.id in _NAMES:
        return _NAMES[node.id]
    if isinstance(node, ast.Call):
        f = node.func
        name = f.id if isinstance(f, ast.Name) else (f.attr if isinstance(f, ast.Attribute)
                                                     and isinstance(f.value, ast.Name)
                                                     and f.value.id in ("np", "numpy") else None)
        if name in _NP_CALLS:
            args = [_literal(a) for a in node.args]
            kw = {}
            for k in node.keywords:
                if k.arg != "dtype":
                    raise ValueError("unsupported keyword %r" % k.arg)
                d = k.value
                dn = d.id if isinstance(d, ast.Name) else (d.attr if isinstance(d, ast.Attribute) else None)
                if dn not in _NP_DTYPES:
                    raise ValueError("unsupported dtype")
                kw["dtype"] = _NP_DTYPES[dn]
            return _NP_CALLS[name](*args, **kw)
    raise ValueError("not a literal: %s" % ast.dump(node)[:80])


def read_serial(path):
    with open(path) as f:
        text = f.readline()
    return _literal(ast.parse(text.strip(), mode="eval").body)


def write_serial(path, data):
    with open(path, "w") as f:
        f.write(str(_plain(data)))


# --- proc-NNNNN.bin ------------------------------------------------------------
def _host_field(sim, data, nn):
    """FP64 host copy (Fortran order (nx, ny, nz)) of a variable's value."""
    if torch.is_tensor(data):
        if data.dim() == 0 or data.numel() == 1:
            return float(data.reshape(-1)[0].item())
        t = data.detach().to(torch.float64)
        if tuple(t.shape) != tuple(nn):
            t = t.expand(nn)
        return t.permute(2, 1, 0).contiguous().cpu().numpy().transpose(2, 1, 0)
    if isinstance(data, np.ndarray):
        return np.broadcast_to(data.astype(np.float64), nn)
    return float(data)


def write_proc_bin(path, columns, nn):
    """columns: list of FP64 host fields or floats, in variable order."""
    nx, ny, nz = nn
    mm = np.memmap(path, dtype=np.float64, mode="w+", shape=(nx, ny, nz, len(columns)), order="C")
    for i, c in enumerate(columns):
        mm[:, :, :, i] = c
    mm.flush()
    del mm


def open_proc_bin(path, nn, nvars):
    nx, ny, nz = nn
    want = nx * ny * nz * nvars * 8
    if os.path.getsize(path) != want:
        raise ValueError("%s holds %d bytes, expected %d for (%d,%d,%d,%d) float64"
                         % (path, os.path.getsize(path), want, nx, ny, nz, nvars))
    return np.memmap(path, dtype=np.float64, mode="r", shape=(nx, ny, nz, nvars), order="C")


# --- VTK ------------------------------------------------------------------------
def write_vtk_binary(path, coords, fields, names, cycle, time, make2d=True):
    """Legacy VTK structured grid as pyrandaIO.makeDumpVTKbin writes it for one
    rank (the ghost layer of the reference is clipped again on a single
    rank, pyrandaMPI.py:514-558). coords: three FP64 host (ax, ay, az)
    arrays; fields: host arrays in `names` order."""
    fx, fy, fz = (np.array(c, dtype=np.float64, order="F") for c in coords)
    ax, ay, az = fx.shape
    reshape = None
    if make2d and az > 1 and (ax == 1 or ay == 1):
        # a 2-D dataset in the y-z or x-z plane is written in the first two dimensions
        if ax == 1:
            ax, ay, az = ay, az, 1
            fx[:, :, :] = fy
            fy[:, :, :] = fz
            fz[:, :, :] = 0.0
        else:
            ay, az = az, 1
            fy[:, :, :] = fz
            fz[:, :, :] = 0.0
        reshape = (ax, ay, 1)
        fx, fy, fz = (a.reshape(reshape, order="F") for a in (fx, fy, fz))
    if az == 1 and fz.max() == fz.min():
        fz[:, :, :] = 0.0
    npts = ax * ay * az
    with open(path, "wb") as f:
        f.write(b"# vtk DataFile Version 3.0 \n")
        f.write(b"vtk output \n")
        f.write(b"BINARY \n")
        f.write(b"DATASET STRUCTURED_GRID \n")
        f.write(("DIMENSIONS  %s %s %s  \n" % (ax, ay, az)).encode())
        f.write(b"FIELD FieldData 2 \n")
        f.write(b"TIME 1 1 float \n")
        f.write(np.array([time], dtype=">f4").tobytes())
        f.write(b"CYCLE 1 1 int \n")
        f.write(np.array([cycle], dtype=">i4").tobytes())
        f.write(("POINTS %s float  \n" % npts).encode())
        # point k,j,i (i fastest) as x y z
        pts = np.stack([fx, fy, fz], axis=-1).transpose(2, 1, 0, 3)
        f.write(np.ascontiguousarray(pts, dtype=">f4").tobytes())
        f.write(("\nPOINT_DATA %s  " % npts).encode())
        for name, g in zip(names, fields):
            f.write(("\nSCALARS %s float \n" % name).encode())
            f.write(b"LOOKUP_TABLE default \n")
            g = np.asarray(g, dtype=np.float64)
            if reshape:
                g = g.reshape(reshape, order="F")
            f.write(np.ascontiguousarray(g.transpose(2, 1, 0), dtype=">f4").tobytes())


def write_visit(path, history, nblocks=1):
    """pyranda.visit: the block count, then per dump its time and files
    (pyranda.py:461-468)."""
    with open(path, "w") as vid:
        vid.write("!NBLOCKS %s \n" % nblocks)
        for iv, t in history:
            vid.write("!TIME %s \n" % t)
            for p in range(nblocks):
                vid.write("%s\n" % os.path.join("vis" + str(iv).zfill(VIS_INT),
                                                "proc-%s.%s.%s" % (str(p).zfill(PROC_INT), str(iv).zfill(VIS_INT),
                                                                   "vtk")))


def read_vtk_binary(path):
    """Parse a file written by write_vtk_binary: (dims, time, cycle, points
    (n,3), {name: values (n,)}), all float32 as stored (test helper)."""
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0

    def line():
        nonlocal pos
        end = buf.index(b"\n", pos)
        s = buf[pos:end].decode()
        pos = end + 1
        return s

    for _ in range(4):
        line()
    dims = tuple(int(v) for v in line().split()[1:4])
    line()  # FIELD
    line()  # TIME meta
    time = float(np.frombuffer(buf, ">f4", 1, pos)[0])
    pos += 4
    line()  # CYCLE meta
    cycle = int(np.frombuffer(buf, ">i4", 1, pos)[0])
    pos += 4
    n = int(line().split()[1])
    pts = np.frombuffer(buf, ">f4", 3 * n, pos).reshape(n, 3)
    pos += 12 * n
    fields = {}
    rest = buf[pos:]
    p = rest.index(b"POINT_DATA")
    pos = pos + p
    line()
    while pos < len(buf):
        if buf[pos:pos + 1] == b"\n":
            pos += 1
        hdr = line()
        if not hdr.startswith("SCALARS"):
            break
        name = hdr.split()[1]
        line()  # LOOKUP_TABLE
        fields[name] = np.frombuffer(buf, ">f4", n, pos).astype(np.float64)
        pos += 4 * n
    return dims, time, cycle, pts.astype(np.float64), fields
