"""Host-side mirror of the reference's kernel interface on top of libsspt.so.

A :class:`Solver` owns one device context.  Its methods carry the names and
arguments of the ParallelStencil kernels the reference scripts call
(SphericalStokes_diapir.jl:413-428), so a script's solver loop maps one to
one::

    @parallel compute_P!(P, DT_R, τ..., VR, Vθ, Vφ, R, θ, η, ηSM, _dr, _dθ, _dφ, dmp, vpdt, lr)
    ->  s.compute_P(_dr, _dth, _dph, dmp, vpdt, lr)

The arrays are the context's device buffers; they are moved with
:meth:`upload` / :meth:`download` (numpy, Julia column-major layout by
default, or C row-major).  Every call blocks until the GPU work is done,
as each ``@parallel`` call does.  Errors raise :class:`SSError`; nothing
falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import SSError, ss_mixed_opts, ss_mixed_result, ss_params, ss_result
from .fields import FIELD_ID, FIELDS, LAYOUT_C, LAYOUT_JULIA, canon, shape
from .setups import Params, Setup


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    def __init__(self, nr: int, nth: int, nph: int, device: int = 0):
        self._L = _lib.lib()
        self.nr, self.nth, self.nph = nr, nth, nph
        h = C.c_void_p()
        rc = self._L.ss_create(nr, nth, nph, device, C.byref(h))
        if rc != 0:
            raise SSError(rc, f"ss_create({nr},{nth},{nph},device={device})")
        self._h = h

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            self._L.ss_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc: int, what: str):
        if rc != 0:
            msg = self._L.ss_last_error(self._h)
            raise SSError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    # -- grid and data movement
    def set_grid(self, R, th, ph):
        R, th, ph = (np.ascontiguousarray(x, dtype=np.float64) for x in (R, th, ph))
        if R.shape != (self.nr,) or th.shape != (self.nth,) or ph.shape != (self.nph,):
            raise ValueError("coordinate arrays must have lengths nr, nth, nph")
        self._chk(self._L.ss_set_grid(self._h, _dptr(R), _dptr(th), _dptr(ph)), "set_grid")

    def field_shape(self, name: str):
        return shape(name, self.nr, self.nth, self.nph)

    def upload(self, name: str, a: np.ndarray):
        name = canon(name)
        shp = self.field_shape(name)
        a = np.asarray(a, dtype=np.float64)
        if a.shape != shp:
            raise ValueError(f"{name}: expected shape {shp}, got {a.shape}")
        if a.flags.f_contiguous:
            layout = LAYOUT_JULIA
        else:
            a = np.ascontiguousarray(a)
            layout = LAYOUT_C
        self._chk(self._L.ss_upload(self._h, FIELD_ID[name], _dptr(a), layout), f"upload {name}")

    def download(self, name: str, order: str = "F") -> np.ndarray:
        name = canon(name)
        shp = self.field_shape(name)
        out = np.empty(shp, dtype=np.float64, order=order)
        layout = LAYOUT_JULIA if order == "F" else LAYOUT_C
        self._chk(self._L.ss_download(self._h, FIELD_ID[name], _dptr(out), layout), f"download {name}")
        return out

    def fill(self, name: str, value: float):
        name = canon(name)
        self._chk(self._L.ss_fill(self._h, FIELD_ID[name], float(value)), f"fill {name}")

    def copy_field(self, dst: str, src: str):
        self._chk(self._L.ss_copy_field(self._h, FIELD_ID[canon(dst)], FIELD_ID[canon(src)]), "copy_field")

    def load_setup(self, s: Setup):
        """Grid + every field of a setup."""
        self.set_grid(s.R, s.th, s.ph)
        for n in FIELDS:
            self.upload(n, s.fields[n])

    def download_all(self) -> dict:
        return {n: self.download(n) for n in FIELDS}

    # -- setup kernels (SphericalStokes_diapir.jl:127-146)
    def initialize_inclusion(self, name: str, r: float, radius: float, val: float):
        self._chk(self._L.ss_initialize_inclusion(self._h, FIELD_ID[canon(name)], r, radius, val),
                  "initialize_inclusion")

    def initialize_velocity(self, fact: float, r: float):
        self._chk(self._L.ss_initialize_velocity(self._h, fact, r), "initialize_velocity")

    # -- the reference kernels
    def equal3(self):
        self._chk(self._L.ss_equal3(self._h), "equal3")

    def maxloc(self):
        self._chk(self._L.ss_maxloc(self._h), "maxloc")

    def spread_etasm(self, m: int):
        """The etaSM spread of the etasm_spread = m option (no reference
        counterpart), applied after maxloc + copy_BC(etaSM)."""
        self._chk(self._L.ss_spread_etasm(self._h, int(m)), "spread_etasm")

    def copy_BC(self, name: str):
        self._chk(self._L.ss_copy_bc(self._h, FIELD_ID[canon(name)]), "copy_BC")

    def compute_P(self, _dr, _dth, _dph, dmp, vpdt, lr):
        self._chk(self._L.ss_compute_P(self._h, _dr, _dth, _dph, dmp, vpdt, lr), "compute_P")

    def compute_TII(self):
        self._chk(self._L.ss_compute_TII(self._h), "compute_TII")

    def power_law(self, tau_C, n_exp, relax=1.0, eta_min=0.0, eta_max=math.inf):
        if relax == 1.0 and eta_min == 0.0 and eta_max == math.inf:
            self._chk(self._L.ss_power_law(self._h, tau_C, n_exp), "power_law")
        else:
            self._chk(self._L.ss_power_law_ex(self._h, tau_C, n_exp, relax, eta_min, eta_max), "power_law")

    def compute_V(self, _dr, _dth, _dph, g):
        self._chk(self._L.ss_compute_V(self._h, _dr, _dth, _dph, g), "compute_V")

    def BC_R(self):
        self._chk(self._L.ss_bc_R(self._h), "BC_R")

    def BC_th(self, bc_mode, r=0.0, sh=0.0):
        self._chk(self._L.ss_bc_th(self._h, int(bc_mode), r, sh), "BC_th")

    def BC_ph(self, bc_mode, r=0.0, sh=0.0):
        self._chk(self._L.ss_bc_ph(self._h, int(bc_mode), r, sh), "BC_ph")

    def check_err(self, drop_nan_vph=False):
        comp = (C.c_double * 3)()
        err = C.c_double()
        self._chk(self._L.ss_check_err(self._h, int(drop_nan_vph), comp, C.byref(err)), "check_err")
        return err.value, tuple(comp)

    def pt_step_unfused(self, p: Params, equal3=True):
        """One PT iteration as the reference's launch sequence (SphericalStokes_diapir.jl:413-425)."""
        if equal3:
            self.equal3()
        self.maxloc()
        self.copy_BC("etaSM")
        if getattr(p, "etasm_spread", 0) > 0:
            self.spread_etasm(p.etasm_spread)
        self.compute_P(p._dr, p._dth, p._dph, p.dmp, p.vpdt, p.lr)
        self.compute_TII()
        self.copy_BC("tII")
        self.power_law(p.tau_C, p.n_exp, p.relax, p.eta_min, p.eta_max)
        self.compute_V(p._dr, p._dth, p._dph, p.g)
        self.BC_R()
        self.BC_th(p.bc_mode, p.r, p.sh)
        self.BC_ph(p.bc_mode, p.r, p.sh)

    # -- fused PT loops
    def pt_iterate(self, p: Params, niter: int):
        sp = ss_params.from_params(p)
        self._chk(self._L.ss_pt_iterate(self._h, C.byref(sp), int(niter)), "pt_iterate")

    def pt_solve(self, p: Params, maxiter: int, nout: int, eps: float, fused: bool = True, raise_on=("nonfinite",)):
        """The reference PT loop with its exit test.  Returns a dict with
        ittot, iter_end, err, err_comp, seconds and status."""
        sp = ss_params.from_params(p)
        res = ss_result()
        fn = self._L.ss_pt_solve if fused else self._L.ss_pt_solve_unfused
        rc = fn(self._h, C.byref(sp), int(maxiter), int(nout), float(eps), C.byref(res))
        status = _lib.STATUS_NAMES.get(rc, str(rc))
        if rc not in (0, _lib.SS_NOT_CONVERGED) and not (rc == _lib.SS_ERR_NONFINITE and "nonfinite" not in raise_on):
            self._chk(rc, "pt_solve")
        return {"ittot": res.ittot, "iter_end": res.iter_end, "err": res.err, "err_comp": tuple(res.err_comp),
                "seconds": res.seconds, "status": status}

    def residual(self, p: Params):
        """(r, r_div, r_mom) of the current state with tau replaced by the
        stress consistent with V and eta (the state is not modified)."""
        sp = ss_params.from_params(p)
        out = (C.c_double * 3)()
        self._chk(self._L.ss_residual(self._h, C.byref(sp), out), "residual")
        return tuple(out)

    def pt_solve_residual(self, p: Params, maxiter: int, ncheck: int, tol: float):
        """Plain FP64 fused iterations until residual <= tol (checked every ncheck)."""
        sp = ss_params.from_params(p)
        res = ss_result()
        rc = self._L.ss_pt_solve_residual(self._h, C.byref(sp), int(maxiter), int(ncheck), float(tol), C.byref(res))
        if rc not in (0, _lib.SS_NOT_CONVERGED):
            self._chk(rc, "pt_solve_residual")
        return {"iters": res.ittot, "iter_end": res.iter_end, "r": res.err, "r_div": res.err_comp[0],
                "r_mom": res.err_comp[1], "seconds": res.seconds, "converged": rc == 0}

    def pt_solve_mixed(self, p: Params, *, max_outer=20, max_inner=2000, nout_inner=100, tol_inner=1e-4,
                       tol=1e-8):
        """FP32 sweeps inside an FP64 defect-correction loop (ss_pt_solve_mixed)."""
        sp = ss_params.from_params(p)
        o = ss_mixed_opts(int(max_outer), int(max_inner), int(nout_inner), float(tol_inner), float(tol))
        r = ss_mixed_result()
        rc = self._L.ss_pt_solve_mixed(self._h, C.byref(sp), C.byref(o), C.byref(r))
        if rc not in (0, _lib.SS_NOT_CONVERGED):
            self._chk(rc, "pt_solve_mixed")
        n = r.outer
        return {"outer": n, "inner_total": r.inner_total, "converged": bool(r.converged), "r": r.r,
                "r_hist": list(r.r_hist[:n + 1]), "inner_hist": list(r.inner_hist[:n]),
                "err_inner_hist": list(r.err_inner_hist[:n]), "seconds": r.seconds,
                "inner_seconds": r.inner_seconds}

    def pt_iterate_timed(self, p: Params, niter: int):
        """pt_iterate with HIP events around each main-kernel launch; returns
        (main kernel ms per launch, whole-iteration ms)."""
        sp = ss_params.from_params(p)
        a, b = C.c_double(), C.c_double()
        self._chk(self._L.ss_pt_iterate_timed(self._h, C.byref(sp), int(niter), C.byref(a), C.byref(b)),
                  "pt_iterate_timed")
        return a.value, b.value

    def last_kernel_ms(self) -> float:
        v = C.c_double()
        self._chk(self._L.ss_last_kernel_ms(self._h, C.byref(v)), "last_kernel_ms")
        return v.value

    # ---- adjoint of the converged solve (include/sspt.h, ss_adjoint_*)
    _ADJ_SHAPE = {"grad_rho": "rho", "grad_eta_ini": "eta_ini"}

    def _adj_id(self, name: str) -> int:
        if name not in _lib.ADJ_FIELDS:
            raise ValueError(f"unknown adjoint field {name}")
        return _lib.ADJ_FIELDS.index(name)

    def _adj_shape(self, name: str):
        return self.field_shape(self._ADJ_SHAPE.get(name, name))

    def adjoint_prepare(self, p: Params):
        sp = ss_params.from_params(p)
        self._chk(self._L.ss_adjoint_prepare(self._h, C.byref(sp)), "adjoint_prepare")

    def adjoint_zero(self):
        self._chk(self._L.ss_adjoint_zero(self._h), "adjoint_zero")

    def adjoint_objective(self, r: float, radius: float) -> float:
        J = C.c_double()
        self._chk(self._L.ss_adjoint_objective(self._h, float(r), float(radius), C.byref(J)), "adjoint_objective")
        return J.value

    def adjoint_upload(self, name: str, a: np.ndarray):
        a = np.asarray(a, dtype=np.float64)
        if a.shape != self._adj_shape(name):
            raise ValueError(f"{name}: expected shape {self._adj_shape(name)}, got {a.shape}")
        layout = LAYOUT_JULIA if a.flags.f_contiguous else LAYOUT_C
        if layout == LAYOUT_C:
            a = np.ascontiguousarray(a)
        self._chk(self._L.ss_adjoint_upload(self._h, self._adj_id(name), _dptr(a), layout), f"adjoint_upload {name}")

    def adjoint_download(self, name: str) -> np.ndarray:
        out = np.empty(self._adj_shape(name), dtype=np.float64, order="F")
        self._chk(self._L.ss_adjoint_download(self._h, self._adj_id(name), _dptr(out), LAYOUT_JULIA),
                  f"adjoint_download {name}")
        return out

    def adjoint_step(self, p: Params, gradients: bool = False):
        sp = ss_params.from_params(p)
        self._chk(self._L.ss_adjoint_step(self._h, C.byref(sp), int(bool(gradients))), "adjoint_step")

    def adjoint_solve(self, p: Params, r: float, radius: float, maxiter: int = 4_000_000, ncheck: int = 100,
                      tol: float = 1e-12):
        """Gradient of J (volume-weighted mean VR in the inclusion) at the
        current converged state: dict with J, iterations, and the gradients
        with respect to rho and log10(eta_ini) (centre arrays)."""
        sp = ss_params.from_params(p)
        res = _lib.ss_adjoint_result()
        rc = self._L.ss_adjoint_solve(self._h, C.byref(sp), float(r), float(radius), int(maxiter), int(ncheck),
                                      float(tol), C.byref(res))
        if rc not in (0, _lib.SS_NOT_CONVERGED):
            self._chk(rc, "adjoint_solve")
        g_ini = self.adjoint_download("grad_eta_ini")
        return {"J": res.J, "iters": res.iters, "converged": bool(res.converged), "err": res.err,
                "seconds": res.seconds, "grad_rho": self.adjoint_download("grad_rho"),
                "grad_log10_eta_ini": g_ini * self.download("eta_ini") * math.log(10.0)}

    def adjoint_solve_mixed(self, p: Params, r: float, radius: float, *, max_outer=20, max_inner=200_000,
                            nout_inner=100, tol_inner=1e-4, tol=1e-12, gradients=True):
        """adjoint_solve with FP32 sweeps inside an FP64 defect correction
        (ss_adjoint_solve_mixed): dict with J, outer steps, FP32 sweeps and,
        with gradients, the gradient arrays as adjoint_solve returns them."""
        sp = ss_params.from_params(p)
        o = ss_mixed_opts(int(max_outer), int(max_inner), int(nout_inner), float(tol_inner), float(tol))
        res = _lib.ss_adjoint_mixed_result()
        rc = self._L.ss_adjoint_solve_mixed(self._h, C.byref(sp), float(r), float(radius), C.byref(o),
                                            C.byref(res))
        if rc not in (0, _lib.SS_NOT_CONVERGED):
            self._chk(rc, "adjoint_solve_mixed")
        out = {"J": res.J, "outer": res.outer, "inner_total": res.inner_total, "converged": bool(res.converged),
               "err": res.err, "seconds": res.seconds, "inner_seconds": res.inner_seconds}
        if gradients:
            g_ini = self.adjoint_download("grad_eta_ini")
            out["grad_rho"] = self.adjoint_download("grad_rho")
            out["grad_log10_eta_ini"] = g_ini * self.download("eta_ini") * math.log(10.0)
        return out
