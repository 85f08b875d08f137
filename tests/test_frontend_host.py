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


def _sim_with_eom(eom, fp64_state=True):
    """A pyrandaSim shell holding only what the FP64-state rule reads: its
    equations (parsed like EOM() parses them) and conserved variables."""
    from pyranda_amd.pyranda import pyrandaSim
    from pyranda_amd.pyrandaEq import pyrandaEq
    sm = dict(base_map())
    sm.update({"fbar(": "self.filter(", "bc.extrap(": "self.packages['BC'].extrap("})
    s = object.__new__(pyrandaSim)
    s.fp64_state = fp64_state
    s.equations = [pyrandaEq(l, sm, None) for l in splitLines(eom)]
    s.conserved = [eq.LHS[0] for eq in s.equations if eq.kind == "PDE"]
    return s


FILTERED_EOM = """
ddt(:rho:)  = -ddx(:rho:*:u:)
ddt(:Et:)   = -ddx(:Et:*:u:)
:rho:       =  fbar( :rho: )
:Et:        =  fbar( :Et:  )
:u:         =  :Et: / :rho:
"""


def test_fp64_state_rule():
    """FP32 runs keep an FP64 state for conserved variables whose only ALG
    writer is their own conservative filter, in EOMs without LHS-less lines."""
    assert sorted(_sim_with_eom(FILTERED_EOM)._self_filtered()) == ["Et", "rho"]
    assert _sim_with_eom(FILTERED_EOM, fp64_state=False)._self_filtered() == {}
    # a second writer of rho, or a filter of another field, takes rho out
    two = FILTERED_EOM + ":rho: = :rho: * 1.0\n"
    assert sorted(_sim_with_eom(two)._self_filtered()) == ["Et"]
    other = FILTERED_EOM.replace("fbar( :rho: )", "fbar( :Et: )")
    assert sorted(_sim_with_eom(other)._self_filtered()) == ["Et"]
    # an LHS-less line (a boundary condition call) may write any field in place
    bc = FILTERED_EOM + "bc.extrap(['rho'],['x1'])\n"
    assert _sim_with_eom(bc)._self_filtered() == {}
    # inactive lines do not count
    s = _sim_with_eom(bc)
    s.equations[-1].active = False
    assert sorted(s._self_filtered()) == ["Et", "rho"]
    # the RK update skips storing the rounding only where the filter lines
    # run before any other ALG line (they are then the field's next reader)
    s = _sim_with_eom(FILTERED_EOM)
    assert sorted(s._filtered_first(s._self_filtered())) == ["Et", "rho"]
    late = FILTERED_EOM.replace(":Et:        =  fbar( :Et:  )\n", "") + ":Et: = fbar( :Et: )\n"
    s = _sim_with_eom(late)
    assert sorted(s._self_filtered()) == ["Et", "rho"]
    assert sorted(s._filtered_first(s._self_filtered())) == ["rho"]
