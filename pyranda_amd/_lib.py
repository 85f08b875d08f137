"""ctypes binding of libpyranda_amd.so (the C-ABI in include/pyranda_amd.h).

The library is built in-tree by pyranda_amd._build.build(). Loading fails
loudly if it is missing: there is no CPU fallback for the operators.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PYRANDA_AMD_LIB selects another in-tree build of the same library (tuning
# variants built by _build.build(out=..., defines=...))
LIB_PATH = os.environ.get("PYRANDA_AMD_LIB") or os.path.join(HERE, "libpyranda_amd.so")

BC = {"PERI": 0, "NONE": 1, "SYMM": 2}
OP_D1, OP_D2, OP_D8, OP_SFILTER, OP_GFILTER = 0, 1, 2, 3, 4
OP_ANTISYM = 0x10  # or-ed into an op: antisymmetric member of the SYMM pair
EPI_STORE, EPI_ADD, EPI_ABS_SCALED, EPI_MAXABS_SCALED = 0, 1, 2, 3
F64, R32 = 0, 1
RK_F32_U64, RK_F64_U32 = 2, 3  # RK update with F and U of different widths
PSOLVE = {"circulant": 0, "lu": 1}
R32_SOLVE = {"refined": 0, "fp64": 1}

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# every symbol the header declares, with its ctypes signature
SIGNATURES = {
    "pyr_setup": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [_ip]),
    "pyr_setup_host": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [_ip]),
    "pyr_destroy": (ctypes.c_int, [_vp]),
    "pyr_set_stream": (ctypes.c_int, [_vp, _vp]),
    "pyr_set_periodic_solver": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pyr_set_tuning": (ctypes.c_int, [_vp] + [ctypes.c_int] * 6),
    "pyr_last_error": (ctypes.c_char_p, []),
    "pyr_axis_coords": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pyr_mesh_scalars": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "pyr_apply": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_double]),
    "pyr_host_weights": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
    "pyr_host_emulate": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "pyr_rk_update": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_longlong, _vp, _vp, _vp, _vp, _vp,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "pyr_rk_update_dt": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_longlong, _vp, _vp, _vp, _vp, _vp,
                                        _vp, ctypes.c_double, ctypes.c_double]),
    "pyr_group_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_vp),
                                        ctypes.POINTER(ctypes.c_longlong)]),
    "pyr_group_destroy": (ctypes.c_int, [_vp]),
    "pyr_group_set_stream": (ctypes.c_int, [_vp, _vp]),
    "pyr_group_apply": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int]),
    "pyr_rk_update_member_dt": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_longlong, _vp, _vp, _vp, _vp, _vp,
                                               _vp, _vp, ctypes.c_double, ctypes.c_double]),
    "pyr_mlp_infer": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.POINTER(_vp), _vp, ctypes.POINTER(_vp)]),
    "pyr_mlp_infer_sensed": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.POINTER(_vp), _vp, ctypes.POINTER(_vp),
                                            ctypes.c_float, _vp, _vp]),
    "pyr_mlp_train_grad": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp,
                                          _vp, ctypes.c_int, _vp]),
    "pyr_pw_compile": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "pyr_pw_check_source": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p]),
    "pyr_pw_launch": (ctypes.c_int, [_vp, _vp, ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(_vp),
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "pyr_pw_launch_blocks": (ctypes.c_int, [_vp, _vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "pyr_pw_last_error": (ctypes.c_char_p, []),
    "pyr_set_r32_solve": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pyr_set_solve_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pyr_sfilter_f64_rounded": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "pyr_apply_rounded": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
    "pyr_apply_to_r32": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "pyr_sfilter_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
}
PW_MAX_PTRS, PW_MAX_SCALARS = 24, 16
MLP_PARAM_FLOAT_OFFSET, MLP_PARAM_BYTES, MLP_GRAD_FLOATS = 14336, 14960, 7313
for _prec in ("f64", "r32"):
    for _name in ("ddx", "ddy", "ddz", "plaplacian", "pring", "sfilter", "gfilter"):
        SIGNATURES["pyr_%s_%s" % (_name, _prec)] = (ctypes.c_int, [_vp, _vp, _vp])
    SIGNATURES["pyr_gfilterdir_%s" % _prec] = (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int])
    SIGNATURES["pyr_grads_%s" % _prec] = (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp])
    SIGNATURES["pyr_divergence_%s" % _prec] = (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp])
for _name in ("dd8x", "dd8y", "dd8z", "d2x", "d2y", "d2z"):
    SIGNATURES["pyr_%s_f64" % _name] = (ctypes.c_int, [_vp, _vp, _vp])
SIGNATURES["pyr_pringv_f64"] = (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp])

_LIB = None


class PyrandaAMDError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise PyrandaAMDError(
                "libpyranda_amd.so is not built; run pyranda_amd._build.build() "
                "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status):
    if status != 0:
        raise PyrandaAMDError(lib().pyr_last_error().decode())
    return status


def check_pw(status):
    if status != 0:
        raise PyrandaAMDError(lib().pyr_pw_last_error().decode())
    return status
