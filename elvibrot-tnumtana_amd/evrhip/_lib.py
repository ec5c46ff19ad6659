"""ctypes binding of libevrhip.so (the HIP engine's C ABI, include/evr_hip.h)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libevrhip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)


class BasisDesc(ctypes.Structure):
    _fields_ = [
        ("D", ctypes.c_int), ("nlev", ctypes.c_int), ("nterms", ctypes.c_int64),
        ("term_levels", _ip), ("weights", _dp), ("nb", _ip), ("nq", _ip),
        ("x", _dp), ("B", _dp), ("W", _dp), ("D1", _dp), ("D2", _dp),
        ("nb_packed", ctypes.c_int64), ("map", _ip), ("grid_last_fastest", ctypes.c_int),
    ]


class OpDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("V", _dp), ("G0", _dp), ("Ga", _dp), ("F2", _dp),
                ("F1", _dp), ("hermitian", ctypes.c_int)]


class DavidsonOpts(ctypes.Structure):
    _fields_ = [("nev", ctypes.c_int), ("block", ctypes.c_int), ("max_sub", ctypes.c_int),
                ("max_it", ctypes.c_int), ("conv_resi", ctypes.c_double), ("conv_ene", ctypes.c_double),
                ("precond", ctypes.c_int), ("diag", _dp), ("guess", _dp), ("nguess", ctypes.c_int),
                ("precision", ctypes.c_int), ("refine_resi", ctypes.c_double),
                ("max_nvec_hint", ctypes.c_int), ("verbose", ctypes.c_int),
                ("basis", ctypes.c_void_p), ("guess_file", ctypes.c_char_p), ("save_file", ctypes.c_char_p),
                ("graph", ctypes.c_int)]


class DavidsonInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("nconv", ctypes.c_int),
                ("hpsi_fp64", ctypes.c_int64), ("hpsi_fp32", ctypes.c_int64),
                ("max_resi", ctypes.c_double), ("seconds", ctypes.c_double)]


class PropaOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("dt", ctypes.c_double), ("nsteps", ctypes.c_int),
                ("npoly", ctypes.c_int), ("poly_tol", ctypes.c_double), ("Hmin", ctypes.c_double),
                ("Hmax", ctypes.c_double), ("field", _dp), ("mu", ctypes.c_void_p), ("psi0", ctypes.c_void_p),
                ("precision", ctypes.c_int), ("device_ptrs", ctypes.c_int), ("graph", ctypes.c_int),
                ("field_sub", _dp), ("nsub", ctypes.c_int)]


class PropaInfo(ctypes.Structure):
    _fields_ = [("hpsi", ctypes.c_int64), ("max_order", ctypes.c_int), ("retries", ctypes.c_int),
                ("Hmin", ctypes.c_double), ("Hmax", ctypes.c_double), ("seconds", ctypes.c_double)]


class ControlOpts(ctypes.Structure):
    _fields_ = [("T", ctypes.c_double), ("nsteps", ctypes.c_int), ("nk", ctypes.c_int),
                ("checkpoint", ctypes.c_int), ("npoly", ctypes.c_int), ("poly_tol", ctypes.c_double),
                ("Hmin", ctypes.c_double), ("Hmax", ctypes.c_double), ("precision", ctypes.c_int),
                ("graph", ctypes.c_int)]


class ControlInfo(ctypes.Structure):
    _fields_ = [("hpsi", ctypes.c_int64), ("max_order", ctypes.c_int), ("retries", ctypes.c_int),
                ("Hmin", ctypes.c_double), ("Hmax", ctypes.c_double), ("seconds_forward", ctypes.c_double),
                ("seconds_backward", ctypes.c_double)]


# name: (restype, argtypes)
_SIGS = {
    "evr_last_error": (ctypes.c_char_p, []),
    "evr_sg4_hermite_desc": (ctypes.c_int, [ctypes.c_int] * 7 + [ctypes.POINTER(ctypes.POINTER(BasisDesc))]),
    "evr_sg4_hermite_desc_modes": (ctypes.c_int, [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int)] * 4 +
                                   [ctypes.POINTER(ctypes.POINTER(BasisDesc))]),
    "evr_dp_hermite_desc": (ctypes.c_int, [ctypes.c_int, _ip, _ip, ctypes.POINTER(ctypes.POINTER(BasisDesc))]),
    "evr_basis_desc_free": (None, [ctypes.POINTER(BasisDesc)]),
    "evr_basis_desc_sizes": (ctypes.c_int, [ctypes.POINTER(BasisDesc), _i64p]),
    "evr_basis_desc_grid": (ctypes.c_int, [ctypes.POINTER(BasisDesc), _dp]),
    "evr_propa_save_restart": (ctypes.c_int, [ctypes.c_double, ctypes.c_int, ctypes.c_int64, _dp,
                                              ctypes.c_char_p]),
    "evr_propa_read_restart": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, _dp, _dp,
                                              ctypes.POINTER(ctypes.c_int)]),
    "evr_psi_save": (ctypes.c_int, [ctypes.POINTER(BasisDesc), ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
    "evr_psi_read": (ctypes.c_int, [ctypes.POINTER(BasisDesc), ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                    _ip, ctypes.c_char_p]),
    "evr_plan_create": (ctypes.c_int, [ctypes.POINTER(BasisDesc), ctypes.POINTER(OpDesc), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "evr_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "evr_plan_set_potential": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "evr_plan_set_keo_grid": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "evr_plan_set_type1_grid": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "evr_plan_gen_potential_poly": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "evr_plan_sizes": (ctypes.c_int, [ctypes.c_void_p, _i64p]),
    "evr_plan_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "evr_plan_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "evr_plan_kernel_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, _dp, _dp, _i64p,
                                             ctypes.POINTER(ctypes.c_int)]),
    "evr_oppsi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "evr_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), _dp, _dp]),
    "evr_basis_to_grid_n": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int]),
    "evr_grid_to_basis_n": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int]),
    "evr_grid_deriv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "evr_oppsi_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int]),
    "evr_basis_to_grid": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "evr_plan_diagonal": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int]),
    "evr_davidson": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DavidsonOpts), _dp, _dp,
                                    ctypes.POINTER(DavidsonInfo)]),
    "evr_grid_to_basis": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "evr_plan_max_nvec": (ctypes.c_int, [ctypes.c_void_p]),
    "evr_propagate": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PropaOpts), ctypes.c_int, ctypes.c_void_p,
                                     _dp, ctypes.POINTER(PropaInfo)]),
    "evr_plan_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "evr_plan_assembled": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "evr_plan_release_assembled": (ctypes.c_int, [ctypes.c_void_p]),
    "evr_plan_bounds": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp]),
    "evr_control_field": (ctypes.c_int, [ctypes.POINTER(ControlOpts), ctypes.c_int, _dp, _dp]),
    "evr_control_build_field": (ctypes.c_int, [ctypes.POINTER(ControlOpts), ctypes.c_int, _dp, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int, _dp, _dp]),
    "evr_control_field_sub": (ctypes.c_int, [ctypes.POINTER(ControlOpts), ctypes.c_int, _dp, ctypes.c_int,
                                             ctypes.c_int, _dp]),
    "evr_control_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ControlOpts),
                                            ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                            ctypes.POINTER(ControlInfo)]),
    "evr_bessel_j": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _dp]),
    "evr_multi_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "evr_multi_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "evr_multi_offsets": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "evr_multi_count": (ctypes.c_int, [ctypes.c_void_p]),
    "evr_multi_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "evr_multi_oppsi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]),
    "evr_multi_propagate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, _dp, _dp, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.POINTER(PropaInfo)]),
    "evr_coupled_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "evr_coupled_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "evr_coupled_sizes": (ctypes.c_int, [ctypes.c_void_p, _i64p, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "evr_coupled_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "evr_coupled_oppsi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]),
    "evr_coupled_propagate": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PropaOpts), ctypes.c_int,
                                             ctypes.c_void_p, _dp, ctypes.POINTER(PropaInfo)]),
    "evr_coupled_diagonal": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int]),
    "evr_coupled_davidson": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DavidsonOpts), _dp, _dp,
                                            ctypes.POINTER(DavidsonInfo)]),
    "evr_coupled_propagate_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(PropaOpts),
                                                   ctypes.c_int, ctypes.c_void_p, _dp,
                                                   ctypes.POINTER(PropaInfo)]),
}

_LIB = None


def load():
    """Load libevrhip.so; raises if it is not built (there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libevrhip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def declared_symbols():
    return sorted(_SIGS)


def check(rc):
    if rc != 0:
        raise RuntimeError("evrhip error %d: %s" % (rc, load().evr_last_error().decode()))
    return rc
