"""The cylindrical twin of the diapir solver on libsspt.so
(scripts_diapir/CylindricalStokes_diapir.jl, CYL_3D_Inclusion).

:class:`CylSolver` owns one device context (include/sspt_cyl.h) whose
methods carry the script's kernel names and arguments (``compute_P!(D_RR,
..., KBDT, GSDT, η, dr, dθ, dz)`` -> ``s.compute_P(dr, dth, dz)``), plus the
script's loop with its exponent switch and exit test (:329-384) as
:meth:`CylSolver.pt_solve`.  :func:`cyl_diapir` is the script's setup
(:222-327) with every constant a keyword.  With r = 1000 and lθ = lr/r
the sector is a thin cylindrical shell that approaches the Cartesian box.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import SSError
from .fields import LAYOUT_C, LAYOUT_JULIA

CYL_FIELDS = ["P", "DIVV", "VR", "Vth", "VZ", "dVR", "dVth", "dVZ", "VR_OLD", "Vth_OLD", "VZ_OLD", "D_RR", "D_thth",
              "D_ZZ", "D_Rth", "D_RZ", "D_thZ", "tau_RR", "tau_thth", "tau_ZZ", "tau_Rth", "tau_RZ", "tau_thZ",
              "sig_RR", "sig_thth", "sig_ZZ", "rho", "rhoG", "eta", "eta_INI", "eta_PL", "eta_PL_OLD", "etaSM",
              "KBDT", "GSDT", "DT_R", "tII", "ERR_Rth"]
CYL_ID = {n: i for i, n in enumerate(CYL_FIELDS)}
_SHIFT = {"VR": (1, 0, 0), "VR_OLD": (1, 0, 0), "Vth": (0, 1, 0), "Vth_OLD": (0, 1, 0), "VZ": (0, 0, 1),
          "VZ_OLD": (0, 0, 1), "dVR": (-1, -2, -2), "dVth": (-2, -1, -2), "dVZ": (-2, -2, -1),
          "D_Rth": (-1, -1, -2), "tau_Rth": (-1, -1, -2), "ERR_Rth": (-1, -1, -2), "D_RZ": (-1, -2, -1),
          "tau_RZ": (-1, -2, -1), "D_thZ": (-2, -1, -1), "tau_thZ": (-2, -1, -1)}


def cyl_shape(name: str, nr: int, nth: int, nz: int) -> tuple:
    """Reference array sizes (CylindricalStokes_diapir.jl:278-317)."""
    if name not in CYL_ID:
        raise KeyError(name)
    a, b, c = _SHIFT.get(name, (0, 0, 0))
    return (nr + a, nth + b, nz + c)


@dataclass
class CylParams:
    """Scalars (names of CylindricalStokes_diapir.jl:224-256): grid spacings
    dr, dθ, dz (the kernels divide by them), the PT step, the power law
    (n_exp until the exit-test error falls below 1e-5, then n_exp_PL, :342-344)."""
    dr: float
    dth: float
    dz: float
    dmp: float
    vpdt: float
    lr: float
    tau_C: float = 1.0
    n_exp: float = 1.0
    n_exp_PL: float = 5.0
    relax: float = 1e-3
    g: float = 1.0
    sh: float = 1.0
    r: float = 1000.0


@dataclass
class CylSetup:
    nr: int
    nth: int
    nz: int
    R: np.ndarray
    th: np.ndarray
    z: np.ndarray
    params: CylParams
    fields: dict = field(default_factory=dict)


def cyl_diapir(nr, nth=None, nz=None, *, eta0=1.0, vr=1e2, rho0=0.0, drho=-10.0, g=1.0, n_exp=1.0, n_exp_PL=5.0,
               tau_C=1.0, relax=1e-3, r=1000.0, lr=6.0, lth=None, lz=None, radius=1.0, sh=1.0,
               dmp=4.5) -> CylSetup:
    """CYL_3D_Inclusion's setup (CylindricalStokes_diapir.jl:222-327):
    inclusion (R-r)^2 + (th r)^2 + Z^2 < radius (squared distance against
    `radius`), rhoG = rho g, VZ = th r sh on the th and z boundary faces only
    (initialize_velocity!), eta_INI = eta_PL = eta, vpdt = dr CFL."""
    nth = nr if nth is None else nth
    nz = nr if nz is None else nz
    lth = lr / r if lth is None else lth
    lz = lr if lz is None else lz
    dr, dth, dz = lr / (nr - 1), lth / (nth - 1), lz / (nz - 1)
    CFL = 1.0 / (2.0 + 4.5 * math.log10(vr))
    R = np.linspace(r - lr / 2.0, r + lr / 2.0, nr)
    th = np.linspace(-lth / 2.0, lth / 2.0, nth)
    z = np.linspace(-lz / 2.0, lz / 2.0, nz)
    f = {n: np.zeros(cyl_shape(n, nr, nth, nz), order="F") for n in CYL_FIELDS}
    f["rho"][...] = rho0
    f["eta"][...] = eta0
    a = (R - r)[:, None, None]
    b = (th * r)[None, :, None]
    c = z[None, None, :]
    inside = ((a * a + b * b) + c * c) < radius
    f["rho"][inside] = rho0 + drho
    f["eta"][inside] = eta0 / vr
    f["rhoG"][...] = f["rho"] * g
    VZ = f["VZ"]
    prof = ((th * r) * sh)[None, :, None]
    VZ[:, 0, :] = prof[:, 0, :]
    VZ[:, -1, :] = prof[:, -1, :]
    VZ[:, :, 0] = np.broadcast_to(prof[:, :, 0], VZ[:, :, 0].shape)
    VZ[:, :, -1] = np.broadcast_to(prof[:, :, 0], VZ[:, :, -1].shape)
    f["eta_INI"][...] = f["eta"]
    f["eta_PL"][...] = f["eta"]
    p = CylParams(dr=dr, dth=dth, dz=dz, dmp=dmp, vpdt=dr * CFL, lr=lr, tau_C=tau_C, n_exp=n_exp, n_exp_PL=n_exp_PL,
                  relax=relax, g=g, sh=sh, r=r)
    return CylSetup(nr, nth, nz, R, th, z, p, f)


CYL_BC = {"copy_bc_r": 0, "copy_bc_th": 1, "copy_bc_z": 2, "bc_r_0": 3, "bc_th_fact": 4, "bc_th_0": 5, "bc_z_0": 6,
          "bc_z_lin": 7}


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _cparams(p: CylParams):
    return _lib.ss_cyl_params(p.dr, p.dth, p.dz, p.dmp, p.vpdt, p.lr, p.tau_C, p.n_exp, p.n_exp_PL, p.relax, p.g,
                              p.sh, p.r)


class CylSolver:
    """One device context of the cylindrical twin (include/sspt_cyl.h)."""

    def __init__(self, nr: int, nth: int, nz: int, device: int = 0):
        self._L = _lib.lib()
        self.nr, self.nth, self.nz = nr, nth, nz
        h = C.c_void_p()
        rc = self._L.ss_cyl_create(nr, nth, nz, device, C.byref(h))
        if rc != 0:
            raise SSError(rc, f"ss_cyl_create({nr},{nth},{nz},device={device})")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.ss_cyl_destroy(self._h)
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

    def _chk(self, rc, what):
        if rc != 0:
            raise SSError(rc, f"{what}: {self._L.ss_cyl_last_error(self._h).decode(errors='replace')}")

    def set_grid(self, R, th):
        r = np.ascontiguousarray(R, dtype=np.float64)
        t = np.ascontiguousarray(th, dtype=np.float64)
        assert r.shape == (self.nr,) and t.shape == (self.nth,)
        self._chk(self._L.ss_cyl_set_grid(self._h, _dptr(r), _dptr(t)), "ss_cyl_set_grid")

    def upload(self, name: str, a: np.ndarray):
        shp = cyl_shape(name, self.nr, self.nth, self.nz)
        a = np.asarray(a, dtype=np.float64)
        if a.shape != shp:
            raise ValueError(f"{name}: shape {a.shape}, expected {shp}")
        if a.flags.f_contiguous:
            self._chk(self._L.ss_cyl_upload(self._h, CYL_ID[name], _dptr(a), LAYOUT_JULIA), "ss_cyl_upload")
        else:
            c = np.ascontiguousarray(a)
            self._chk(self._L.ss_cyl_upload(self._h, CYL_ID[name], _dptr(c), LAYOUT_C), "ss_cyl_upload")

    def download(self, name: str) -> np.ndarray:
        out = np.empty(cyl_shape(name, self.nr, self.nth, self.nz), order="F")
        self._chk(self._L.ss_cyl_download(self._h, CYL_ID[name], _dptr(out), LAYOUT_JULIA), "ss_cyl_download")
        return out

    def load_setup(self, s: CylSetup, nonzero_only: bool = False):
        """Grid and every field of the setup (nonzero_only: skip all-zero
        fields, for a fresh context, which starts at zero)."""
        self.set_grid(s.R, s.th)
        for n in CYL_FIELDS:
            if not nonzero_only or np.any(s.fields[n]):
                self.upload(n, s.fields[n])

    # -- the reference's kernels
    def equal(self, dst, src):
        self._chk(self._L.ss_cyl_equal(self._h, CYL_ID[dst], CYL_ID[src]), "equal!")

    def multiply(self, dst, src, fact):
        self._chk(self._L.ss_cyl_multiply(self._h, CYL_ID[dst], CYL_ID[src], float(fact)), "multiply!")

    def maxloc(self):
        self._chk(self._L.ss_cyl_maxloc(self._h), "maxloc!")

    def timesteps(self, dmp, vpdt, lr):
        self._chk(self._L.ss_cyl_timesteps(self._h, dmp, vpdt, lr), "timesteps!")

    def timesteps2(self, vpdt):
        self._chk(self._L.ss_cyl_timesteps2(self._h, vpdt), "timesteps2!")

    def compute_P(self, dr, dth, dz):
        self._chk(self._L.ss_cyl_compute_P(self._h, dr, dth, dz), "compute_P!")

    def compute_sigma(self):
        self._chk(self._L.ss_cyl_compute_sigma(self._h), "compute_sigma!")

    def compute_TII(self):
        self._chk(self._L.ss_cyl_compute_TII(self._h), "compute_TII!")

    def power_law(self, tau_C, n_exp, relax):
        self._chk(self._L.ss_cyl_power_law(self._h, tau_C, n_exp, relax), "power_law!")

    def compute_dV(self, dr, dth, dz):
        self._chk(self._L.ss_cyl_compute_dV(self._h, dr, dth, dz), "compute_dV!")

    def compute_V(self):
        self._chk(self._L.ss_cyl_compute_V(self._h), "compute_V!")

    def bc(self, op, name, fact=0.0, r=0.0):
        self._chk(self._L.ss_cyl_bc(self._h, CYL_ID[name], CYL_BC[op], float(fact), float(r)), op + "!")

    def err_rth(self):
        self._chk(self._L.ss_cyl_err_rth(self._h), "err_rth!")

    def exit_error(self):
        err = C.c_double()
        comp = (C.c_double * 3)()
        self._chk(self._L.ss_cyl_exit_error(self._h, C.byref(err), comp), "exit error")
        return err.value, tuple(comp)

    def pt_solve(self, p: CylParams, maxiter: int, nout: int, eps: float):
        """The reference loop with its exponent switch and exit test."""
        r = _lib.ss_cyl_result()
        rc = self._L.ss_cyl_pt_solve(self._h, C.byref(_cparams(p)), int(maxiter), int(nout), float(eps), C.byref(r))
        if rc not in (0, _lib.SS_NOT_CONVERGED):
            self._chk(rc, "ss_cyl_pt_solve")
        return {"iters": r.iters, "converged": bool(r.converged), "err": r.err, "n_exp": r.n_exp,
                "seconds": r.seconds}

    def pt_iterate_timed(self, p: CylParams, niter: int):
        it, k = C.c_double(), C.c_double()
        self._chk(self._L.ss_cyl_pt_iterate_timed(self._h, C.byref(_cparams(p)), int(niter), C.byref(it),
                                                  C.byref(k)), "ss_cyl_pt_iterate_timed")
        return it.value, k.value
