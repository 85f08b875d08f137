"""Host logic of the pyrandaSim front end (no GPU).

The EOM translator must reproduce pyranda's string semantics exactly
(pyranda/pyrandaUtils.py:22-141): token discovery, map insertion order,
whitespace removal, `#arg#` substitutions. The expected strings below are
worked out by hand from those rules.
"""
import pytest

from pyranda_amd.pyrandaUtils import fortran3d, findVar, splitLines
from pyranda_amd.pyrandaMesh import pyrandaMesh, defaultMeshOptions


def base_map():
    # the head of pyrandaSim.get_sMap, in the reference's insertion order
    return {"ddx(": "self.ddx(", "fbar(": "self.filter(", "sqrt(": "numpy.sqrt(",
            "pi": "numpy.pi", "meshx": "self.mesh.coords[0].data"}


def test_scalar_and_vector_tokens():
    s = ":rho:*:u: + :[vel]: - a[:,0]"
    assert sorted(findVar(s, "scalar")) == ["rho", "u"]
    assert findVar(s, "vector") == ["vel"]
    # order preserved when unique=False (multi-value LHS)
    assert findVar("[:tx:,:ty:,:tz:]", "scalar", unique=False) == ["tx", "ty", "tz"]


def test_fortran3d_rewrites_in_map_order():
    sm = base_map()
    out = fortran3d("ddx( :rho: * sqrt(:p:) ) + pi*meshx", sm)
    assert out == ('self.ddx(self.variables["rho"].data*numpy.sqrt(self.variables["p"].data))'
                   '+numpy.pi*self.mesh.coords[0].data')
    # variable tokens were appended to the shared map after the base keys
    assert list(sm)[-2:] in ([":rho:", ":p:"], [":p:", ":rho:"])


def test_base_keys_apply_before_variable_tokens():
    # the reference rewrites 'pi' before it substitutes ':pi:' (bug-compatible)
    sm = base_map()
    out = fortran3d(":pi: + 1", sm)
    assert out == ":numpy.pi:+1"


def test_arg_substitution():
    sm = {"dt.diff(#arg#)": "self.packages['Timestep'].diff(#arg#)"}
    out = fortran3d("0.2*dt.diff(:beta:,:rho:)", sm)
    assert "self.packages['Timestep'].diff(" in out


def test_split_lines_drops_comments_and_blanks():
    text = """
    # comment
    :a: = 1.0

    :b: = :a: + 2
    """
    assert splitLines(text) == [":a:=1.0", ":b:=:a:+2"]


def test_mesh_string():
    m = pyrandaMesh()
    m.makeMeshStr("""
    xdom = (0.0, 2.0, 16, periodic=True)
    ydom = (-1.0, 1.0, 8)
    """)
    o = m.options
    assert o["x1"][:2] == [0.0, -1.0] and o["xn"][:2] == [2.0, 1.0]
    assert o["nn"] == [16, 8, 1]
    assert o["periodic"] == [True, False, False]
    d = defaultMeshOptions()
    assert d["coordsys"] == 0 and d["symmetric"] == [[False, False]] * 3


def test_device_numpy_namespace_on_host_values():
    from pyranda_amd.devnumpy import numpy
    assert numpy.minimum(1.0, 2.0) == 1.0
    assert numpy.sqrt(4.0) == 2.0
    assert numpy.pi == pytest.approx(3.141592653589793)
