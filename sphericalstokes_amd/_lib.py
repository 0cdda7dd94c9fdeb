"""ctypes binding of libsspt.so (the C ABI of include/sspt.h, sspt_cart.h and sspt_cyl.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or cannot be loaded, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsspt.so"

SS_OK, SS_ERR_ARG, SS_ERR_NONFINITE, SS_NOT_CONVERGED, SS_ERR_STATE, SS_ERR_NOMEM = 0, -1, -2, -3, -4, -5
STATUS_NAMES = {0: "SS_OK", -1: "SS_ERR_ARG", -2: "SS_ERR_NONFINITE", -3: "SS_NOT_CONVERGED",
                -4: "SS_ERR_STATE", -5: "SS_ERR_NOMEM"}


class ss_params(C.Structure):
    _fields_ = [
        ("_dr", C.c_double), ("_dth", C.c_double), ("_dph", C.c_double),
        ("dmp", C.c_double), ("vpdt", C.c_double), ("lr", C.c_double),
        ("tau_C", C.c_double), ("n_exp", C.c_double), ("g", C.c_double),
        ("r", C.c_double), ("sh", C.c_double), ("relax", C.c_double),
        ("eta_min", C.c_double), ("eta_max", C.c_double),
        ("bc_mode", C.c_int), ("err_drop_nan_vph", C.c_int), ("etasm_spread", C.c_int),
    ]

    @classmethod
    def from_params(cls, p) -> "ss_params":
        return cls(p._dr, p._dth, p._dph, p.dmp, p.vpdt, p.lr, p.tau_C, p.n_exp, p.g, p.r, p.sh,
                   p.relax, p.eta_min, p.eta_max if p.eta_max is not None else math.inf,
                   int(p.bc_mode), int(p.err_drop_nan_vph), int(getattr(p, "etasm_spread", 0)))


class ss_result(C.Structure):
    _fields_ = [("ittot", C.c_int), ("iter_end", C.c_int), ("err", C.c_double),
                ("err_comp", C.c_double * 3), ("seconds", C.c_double)]


MIXED_MAX_OUTER = 64


class ss_mixed_opts(C.Structure):
    _fields_ = [("max_outer", C.c_int), ("max_inner", C.c_int), ("nout_inner", C.c_int),
                ("tol_inner", C.c_double), ("tol", C.c_double)]


class ss_mixed_result(C.Structure):
    _fields_ = [("outer", C.c_int), ("inner_total", C.c_int), ("converged", C.c_int),
                ("inner_hist", C.c_int * MIXED_MAX_OUTER), ("r_hist", C.c_double * (MIXED_MAX_OUTER + 1)),
                ("err_inner_hist", C.c_double * MIXED_MAX_OUTER), ("r", C.c_double), ("seconds", C.c_double),
                ("inner_seconds", C.c_double)]


class ss_adjoint_result(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("J", C.c_double), ("err", C.c_double),
                ("seconds", C.c_double)]


class ss_adjoint_mixed_result(C.Structure):
    _fields_ = [("outer", C.c_int), ("inner_total", C.c_int), ("converged", C.c_int), ("J", C.c_double),
                ("err", C.c_double), ("seconds", C.c_double), ("inner_seconds", C.c_double)]


# adjoint field ids (enum ss_adj_field)
ADJ_FIELDS = ["VR", "Vth", "Vph", "P", "tau_RR", "tau_thth", "tau_phph", "tau_Rth", "tau_Rph", "tau_thph", "eta",
              "eta_PL", "grad_rho", "grad_eta_ini"]

# name -> (restype, argtypes); every symbol declared in include/*.h
P = C.c_void_p
D = C.c_double
I = C.c_int
DP = C.POINTER(C.c_double)
SIGNATURES = {
    "ss_create": (I, [I, I, I, I, C.POINTER(P)]),
    "ss_destroy": (I, [P]),
    "ss_last_error": (C.c_char_p, [P]),
    "ss_field_dims": (I, [P, I, C.POINTER(C.c_int)]),
    "ss_set_grid": (I, [P, DP, DP, DP]),
    "ss_upload": (I, [P, I, DP, I]),
    "ss_download": (I, [P, I, DP, I]),
    "ss_fill": (I, [P, I, D]),
    "ss_copy_field": (I, [P, I, I]),
    "ss_initialize_inclusion": (I, [P, I, D, D, D]),
    "ss_initialize_velocity": (I, [P, D, D]),
    "ss_equal3": (I, [P]),
    "ss_maxloc": (I, [P]),
    "ss_spread_etasm": (I, [P, I]),
    "ss_copy_bc": (I, [P, I]),
    "ss_compute_P": (I, [P, D, D, D, D, D, D]),
    "ss_compute_TII": (I, [P]),
    "ss_power_law": (I, [P, D, D]),
    "ss_power_law_ex": (I, [P, D, D, D, D, D]),
    "ss_compute_V": (I, [P, D, D, D, D]),
    "ss_bc_R": (I, [P]),
    "ss_bc_th": (I, [P, I, D, D]),
    "ss_bc_ph": (I, [P, I, D, D]),
    "ss_check_err": (I, [P, I, DP, DP]),
    "ss_check_err_noalloc": (I, [P, P, P, P, P, P, I, I, I, I, P, P]),
    "ss_pt_iterate": (I, [P, C.POINTER(ss_params), I]),
    "ss_pt_solve": (I, [P, C.POINTER(ss_params), I, I, D, C.POINTER(ss_result)]),
    "ss_pt_solve_unfused": (I, [P, C.POINTER(ss_params), I, I, D, C.POINTER(ss_result)]),
    "ss_residual": (I, [P, C.POINTER(ss_params), DP]),
    "ss_pt_solve_residual": (I, [P, C.POINTER(ss_params), I, I, D, C.POINTER(ss_result)]),
    "ss_pt_solve_mixed": (I, [P, C.POINTER(ss_params), C.POINTER(ss_mixed_opts), C.POINTER(ss_mixed_result)]),
    "ss_adjoint_prepare": (I, [P, C.POINTER(ss_params)]),
    "ss_adjoint_zero": (I, [P]),
    "ss_adjoint_objective": (I, [P, D, D, DP]),
    "ss_adjoint_upload": (I, [P, I, DP, I]),
    "ss_adjoint_download": (I, [P, I, DP, I]),
    "ss_adjoint_step": (I, [P, C.POINTER(ss_params), I]),
    "ss_adjoint_solve": (I, [P, C.POINTER(ss_params), D, D, I, I, D, C.POINTER(ss_adjoint_result)]),
    "ss_adjoint_solve_mixed": (I, [P, C.POINTER(ss_params), D, D, C.POINTER(ss_mixed_opts),
                                   C.POINTER(ss_adjoint_mixed_result)]),
    "ss_last_kernel_ms": (I, [P, DP]),
    "ss_pt_iterate_timed": (I, [P, C.POINTER(ss_params), I, DP, DP]),
}


class ss_cart_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("_dx", "_dy", "_dz", "dmp", "vpdt", "lx", "tau_C", "n_exp", "relax",
                                         "g", "sh")]


class ss_cart_result(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("err", C.c_double), ("seconds", C.c_double)]


# the Cartesian twin, include/sspt_cart.h
SIGNATURES.update({
    "ss_cart_create": (I, [I, I, I, I, C.POINTER(P)]),
    "ss_cart_destroy": (None, [P]),
    "ss_cart_last_error": (C.c_char_p, [P]),
    "ss_cart_set_y": (I, [P, DP]),
    "ss_cart_upload": (I, [P, I, DP, I]),
    "ss_cart_download": (I, [P, I, DP, I]),
    "ss_cart_equal": (I, [P, I, I]),
    "ss_cart_maxloc": (I, [P]),
    "ss_cart_copy_boundary": (I, [P, I]),
    "ss_cart_timesteps": (I, [P, D, D, D]),
    "ss_cart_compute_P": (I, [P, D, D, D, D, D, D]),
    "ss_cart_compute_TII": (I, [P]),
    "ss_cart_power_law": (I, [P, D, D, D]),
    "ss_cart_compute_V": (I, [P, D, D, D, D]),
    "ss_cart_bc_x_lin": (I, [P, I, D]),
    "ss_cart_bc_x_0": (I, [P, I]),
    "ss_cart_copy_bc_x": (I, [P, I]),
    "ss_cart_bc_y_fact": (I, [P, I, D]),
    "ss_cart_bc_y_0": (I, [P, I]),
    "ss_cart_bc_z_0": (I, [P, I]),
    "ss_cart_copy_bc_z": (I, [P, I]),
    "ss_cart_check_err": (I, [P, DP, DP]),
    "ss_cart_pt_solve": (I, [P, C.POINTER(ss_cart_params), I, I, D, C.POINTER(ss_cart_result)]),
    "ss_cart_pt_iterate_timed": (I, [P, C.POINTER(ss_cart_params), I, DP, DP]),
})


class ss_cyl_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dr", "dth", "dz", "dmp", "vpdt", "lr", "tau_C", "n_exp", "n_exp_PL",
                                         "relax", "g", "sh", "r")]


class ss_cyl_result(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("err", C.c_double), ("n_exp", C.c_double),
                ("seconds", C.c_double)]


# the cylindrical twin, include/sspt_cyl.h
SIGNATURES.update({
    "ss_cyl_create": (I, [I, I, I, I, C.POINTER(P)]),
    "ss_cyl_destroy": (None, [P]),
    "ss_cyl_last_error": (C.c_char_p, [P]),
    "ss_cyl_set_grid": (I, [P, DP, DP]),
    "ss_cyl_upload": (I, [P, I, DP, I]),
    "ss_cyl_download": (I, [P, I, DP, I]),
    "ss_cyl_equal": (I, [P, I, I]),
    "ss_cyl_multiply": (I, [P, I, I, D]),
    "ss_cyl_maxloc": (I, [P]),
    "ss_cyl_timesteps": (I, [P, D, D, D]),
    "ss_cyl_timesteps2": (I, [P, D]),
    "ss_cyl_compute_P": (I, [P, D, D, D]),
    "ss_cyl_compute_sigma": (I, [P]),
    "ss_cyl_compute_TII": (I, [P]),
    "ss_cyl_power_law": (I, [P, D, D, D]),
    "ss_cyl_compute_dV": (I, [P, D, D, D]),
    "ss_cyl_compute_V": (I, [P]),
    "ss_cyl_bc": (I, [P, I, I, D, D]),
    "ss_cyl_err_rth": (I, [P]),
    "ss_cyl_exit_error": (I, [P, DP, DP]),
    "ss_cyl_pt_solve": (I, [P, C.POINTER(ss_cyl_params), I, I, D, C.POINTER(ss_cyl_result)]),
    "ss_cyl_pt_iterate_timed": (I, [P, C.POINTER(ss_cyl_params), I, DP, DP]),
})

_LIB = None


def lib() -> C.CDLL:
    """Loads libsspt.so once; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class SSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        self.code = code
        name = STATUS_NAMES.get(code, f"hipError {code}" if code > 0 else str(code))
        super().__init__(f"{name}: {msg}")
