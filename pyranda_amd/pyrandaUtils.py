"""EOM string translation, with pyranda's semantics.

fortran3d (pyranda/pyrandaUtils.py:22-66) rewrites an equation string into a
Python expression: whitespace is dropped, every `:v:` (scalar) and `:[v]:`
(vector) token is registered in the shared substitution map, then each map
key is replaced in the map's insertion order. The insertion order is part of
the contract (e.g. `pi` is rewritten before any variable token), so the same
ordered str.replace pass is kept.
"""
import re

SCALAR_KEY = ":"
VECTOR_KEY = (":[", "]:")
TENSOR_KEY = (":{", "}:")
VAR_DATA = 'self.variables["%s"].data'


def fortran3d(form, sMap, latex=False):
    form = form.replace(" ", "").strip()
    var_data = "%s" if latex else VAR_DATA
    for ivar in findVar(form, "scalar"):
        kvar = "%s%s%s" % (SCALAR_KEY, ivar, SCALAR_KEY)
        if kvar not in sMap:
            sMap[kvar] = var_data % ivar
    for ivar in findVar(form, "vector"):
        kvar = "%s%s%s" % (VECTOR_KEY[0], ivar, VECTOR_KEY[1])
        if kvar not in sMap:
            sMap[kvar] = var_data % ivar
    for mvar in list(sMap):
        if "#arg#" in mvar:
            eMap = sMap[mvar]
            prefix = mvar.split("(")[0]
            try:
                found = re.findall(prefix + r"\(.*?\)", form)[0]
                args = found.replace("(", "").replace(")", "").replace(prefix, "")
                form = form.replace(mvar.replace("#arg#", args), eMap.replace("#arg#", args))
            except Exception:  # the reference ignores any failure here
                pass
        else:
            form = form.replace(mvar, sMap[mvar])
    return form


def splitLines(aLongString):
    lines = [el.replace(" ", "") for el in aLongString.split("\n") if el]
    lines = [el for el in lines if el]
    return [el for el in lines if el.strip()[0] != "#"]


def findVar(string, kind, unique=True):
    if kind == "scalar":
        found = _find_scalar(string)
    elif kind == "vector":
        found = _find_delimited(string, VECTOR_KEY)
    elif kind == "tensor":
        found = _find_delimited(string, TENSOR_KEY)
    else:
        found = []
    if unique:
        found = list(set(found))
    return found


def _find_scalar(string):
    excluded = "[]{}#%*@"
    out = []
    for tok in re.findall("%s.*?%s" % (SCALAR_KEY, SCALAR_KEY), string):
        if not any(c in tok for c in excluded):
            out.append(tok.replace(SCALAR_KEY, ""))
    return out


def _find_delimited(string, keys):
    pat = re.escape(keys[0]) + ".*?" + re.escape(keys[1])
    return [s.replace(keys[0], "").replace(keys[1], "") for s in re.findall(pat, string)]
