"""The Cartesian twin of the diapir solver on libsspt.so
(scripts_diapir/CartesianStokes_diapir.jl, CART_3D_Inclusion_Perf).

:class:`CartSolver` owns one device context (include/sspt_cart.h); its
methods carry the names and arguments of the script's kernels
(``compute_P!(P, τ..., VX, VY, VZ, η, ηSM, dmp, vpdt, lx, _dx, _dy, _dz)``
-> ``s.compute_P(dmp, vpdt, lx, _dx, _dy, _dz)``), plus the script's loop
with its exit test (:349-391) as :meth:`CartSolver.pt_solve`.
:func:`cart_diapir` is the script's setup (:263-346) with every constant a
keyword.  It serves the cross-coordinate check of the spherical solver: a
thin spherical sector (r = 1000, lr = 6, the diapir script's own setup)
approaches this Cartesian box (lx = 6).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import SSError, ss_cart_params, ss_cart_result
from .fields import LAYOUT_C, LAYOUT_JULIA

CART_FIELDS = ["P", "tau_XX", "tau_YY", "tau_ZZ", "tau_XY", "tau_XZ", "tau_YZ", "VX", "VY", "VZ",
               "ERR_VX", "ERR_VY", "ERR_VZ", "rho", "eta", "eta_INI", "eta_PL", "eta_PL_OLD", "etaSM",
               "DT_R", "tII"]
CART_ID = {n: i for i, n in enumerate(CART_FIELDS)}
VELOCITIES = ("VX", "VY", "VZ")


def cart_shape(name: str, nx: int, ny: int, nz: int) -> tuple:
    """Reference array sizes (CartesianStokes_diapir.jl:317-337)."""
    if name in ("VX", "ERR_VX"):
        return (nx + 1, ny, nz)
    if name in ("VY", "ERR_VY"):
        return (nx, ny + 1, nz)
    if name in ("VZ", "ERR_VZ"):
        return (nx, ny, nz + 1)
    if name == "tau_XY":
        return (nx - 1, ny - 1, nz - 2)
    if name == "tau_XZ":
        return (nx - 1, ny - 2, nz - 1)
    if name == "tau_YZ":
        return (nx - 2, ny - 1, nz - 1)
    if name not in CART_ID:
        raise KeyError(name)
    return (nx, ny, nz)


@dataclass
class CartParams:
    """Scalars of one iteration (names of CartesianStokes_diapir.jl:267-293)."""
    _dx: float
    _dy: float
    _dz: float
    dmp: float
    vpdt: float
    lx: float
    tau_C: float = 1.0
    n_exp: float = 5.0
    relax: float = 1e-3
    g: float = 1.0
    sh: float = 1.0

    def c(self) -> ss_cart_params:
        return ss_cart_params(self._dx, self._dy, self._dz, self.dmp, self.vpdt, self.lx, self.tau_C, self.n_exp,
                              self.relax, self.g, self.sh)


@dataclass
class CartSetup:
    nx: int
    ny: int
    nz: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    params: CartParams
    fields: dict = field(default_factory=dict)

    def copy_fields(self) -> dict:
        return {k: np.array(v, order="F", copy=True) for k, v in self.fields.items()}


def cart_diapir(nx, ny=None, nz=None, *, eta0=1.0, vr=1e2, rho0=0.0, drho=-10.0, g=1.0, n_exp=5.0, tau_C=1.0,
                relax=1e-3, lx=6.0, ly=None, lz=None, radius=1.0, sh=1.0, dmp=4.5) -> CartSetup:
    """CART_3D_Inclusion_Perf's setup (CartesianStokes_diapir.jl:263-346):
    the inclusion test x^2 + y^2 + z^2 < radius (squared distance against
    `radius`, :112), VX = y*sh everywhere (initialize_velocity!, :117-123),
    eta_INI = eta_PL = eta, vpdt = dx*CFL."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    ly = lx if ly is None else ly
    lz = lx if lz is None else lz
    dx, dy, dz = lx / (nx - 1), ly / (ny - 1), lz / (nz - 1)
    CFL = 1.0 / (2.0 + 4.5 * math.log10(vr))
    x = np.linspace(-lx / 2.0, lx / 2.0, nx)
    y = np.linspace(-ly / 2.0, ly / 2.0, ny)
    z = np.linspace(-lz / 2.0, lz / 2.0, nz)
    f = {n: np.zeros(cart_shape(n, nx, ny, nz), order="F") for n in CART_FIELDS}
    for n in ("ERR_VX", "ERR_VY", "ERR_VZ"):
        f[n][...] = 1.0                                  # @ones (:320-322)
    f["rho"][...] = rho0
    f["eta"][...] = eta0
    d = (x[:, None, None] * x[:, None, None] + y[None, :, None] * y[None, :, None]) + z[None, None, :] * z[None, None, :]
    inside = d < radius
    f["rho"][inside] = rho0 + drho
    f["eta"][inside] = eta0 / vr
    f["VX"][...] = (y * sh)[None, :, None]               # COORD[ix,iy,iz]*fact, the last x face from ix-1
    f["eta_INI"][...] = f["eta"]
    f["eta_PL"][...] = f["eta"]
    p = CartParams(_dx=1.0 / dx, _dy=1.0 / dy, _dz=1.0 / dz, dmp=dmp, vpdt=dx * CFL, lx=lx, tau_C=tau_C,
                   n_exp=n_exp, relax=relax, g=g, sh=sh)
    return CartSetup(nx, ny, nz, x, y, z, p, f)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class CartSolver:
    def __init__(self, nx: int, ny: int, nz: int, device: int = 0):
        self._L = _lib.lib()
        self.nx, self.ny, self.nz = nx, ny, nz
        h = C.c_void_p()
        rc = self._L.ss_cart_create(nx, ny, nz, device, C.byref(h))
        if rc != 0:
            raise SSError(rc, f"ss_cart_create({nx},{ny},{nz},device={device})")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.ss_cart_destroy(self._h)
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
            raise SSError(rc, f"{what}: {self._L.ss_cart_last_error(self._h).decode(errors='replace')}")

    # -- data
    def set_y(self, y):
        a = np.ascontiguousarray(y, dtype=np.float64)
        assert a.shape == (self.ny,)
        self._chk(self._L.ss_cart_set_y(self._h, _dptr(a)), "ss_cart_set_y")

    def upload(self, name: str, a: np.ndarray):
        shp = cart_shape(name, self.nx, self.ny, self.nz)
        a = np.asarray(a, dtype=np.float64)
        if a.shape != shp:
            raise ValueError(f"{name}: shape {a.shape}, expected {shp}")
        if a.flags.f_contiguous:
            self._chk(self._L.ss_cart_upload(self._h, CART_ID[name], _dptr(a), LAYOUT_JULIA), "ss_cart_upload")
        else:
            c = np.ascontiguousarray(a)
            self._chk(self._L.ss_cart_upload(self._h, CART_ID[name], _dptr(c), LAYOUT_C), "ss_cart_upload")

    def download(self, name: str) -> np.ndarray:
        out = np.empty(cart_shape(name, self.nx, self.ny, self.nz), order="F")
        self._chk(self._L.ss_cart_download(self._h, CART_ID[name], _dptr(out), LAYOUT_JULIA), "ss_cart_download")
        return out

    def load_setup(self, s: CartSetup):
        self.set_y(s.y)
        for n in CART_FIELDS:
            self.upload(n, s.fields[n])

    def download_all(self) -> dict:
        return {n: self.download(n) for n in CART_FIELDS}

    # -- the reference's kernels
    def equal(self, dst: str, src: str):
        self._chk(self._L.ss_cart_equal(self._h, CART_ID[dst], CART_ID[src]), "equal!")

    def maxloc(self):
        self._chk(self._L.ss_cart_maxloc(self._h), "maxloc!")

    def copy_boundary(self, name: str):
        self._chk(self._L.ss_cart_copy_boundary(self._h, CART_ID[name]), "copy_boundary!")

    def timesteps(self, vpdt, dmp, lx):
        self._chk(self._L.ss_cart_timesteps(self._h, vpdt, dmp, lx), "timesteps!")

    def compute_P(self, dmp, vpdt, lx, _dx, _dy, _dz):
        self._chk(self._L.ss_cart_compute_P(self._h, dmp, vpdt, lx, _dx, _dy, _dz), "compute_P!")

    def compute_TII(self):
        self._chk(self._L.ss_cart_compute_TII(self._h), "compute_TII!")

    def power_law(self, tau_C, n_exp, relax):
        self._chk(self._L.ss_cart_power_law(self._h, tau_C, n_exp, relax), "power_law!")

    def compute_V(self, _dx, _dy, _dz, g):
        self._chk(self._L.ss_cart_compute_V(self._h, _dx, _dy, _dz, g), "compute_V!")

    def bc_x_lin(self, name, fact):
        self._chk(self._L.ss_cart_bc_x_lin(self._h, CART_ID[name], fact), "bc_x_lin!")

    def bc_x_0(self, name):
        self._chk(self._L.ss_cart_bc_x_0(self._h, CART_ID[name]), "bc_x_0!")

    def copy_bc_x(self, name):
        self._chk(self._L.ss_cart_copy_bc_x(self._h, CART_ID[name]), "copy_bc_x!")

    def bc_y_fact(self, name, fact):
        self._chk(self._L.ss_cart_bc_y_fact(self._h, CART_ID[name], fact), "bc_y_fact!")

    def bc_y_0(self, name):
        self._chk(self._L.ss_cart_bc_y_0(self._h, CART_ID[name]), "bc_y_0!")

    def bc_z_0(self, name):
        self._chk(self._L.ss_cart_bc_z_0(self._h, CART_ID[name]), "bc_z_0!")

    def copy_bc_z(self, name):
        self._chk(self._L.ss_cart_copy_bc_z(self._h, CART_ID[name]), "copy_bc_z!")

    def check_err(self):
        """error!(ERR_V, V) for each component and the exit-test error."""
        err = C.c_double()
        comp = (C.c_double * 3)()
        self._chk(self._L.ss_cart_check_err(self._h, C.byref(err), comp), "check_err")
        return err.value, tuple(comp)

    def iteration(self, p: CartParams, check: bool = False):
        """One iteration of the reference loop as separate kernel calls
        (CartesianStokes_diapir.jl:351-379)."""
        self.equal("etaSM", "eta")
        self.maxloc()
        self.copy_boundary("etaSM")
        self.timesteps(p.vpdt, p.dmp, p.lx)
        self.compute_P(p.dmp, p.vpdt, p.lx, p._dx, p._dy, p._dz)
        if p.n_exp > 1:
            self.compute_TII()
            self.copy_boundary("tII")
            self.equal("eta_PL_OLD", "eta_PL")
            self.power_law(p.tau_C, p.n_exp, p.relax)
        if check:
            for v in VELOCITIES:
                self.equal("ERR_" + v, v)
        self.compute_V(p._dx, p._dy, p._dz, p.g)
        self.bc_x_lin("VX", p.sh)
        self.bc_x_0("VY")
        self.bc_x_0("VZ")
        self.bc_y_fact("VX", p.sh)
        self.bc_y_0("VY")
        self.bc_y_fact("VZ", 0.0)
        self.copy_bc_z("VX")
        self.copy_bc_z("VY")
        self.bc_z_0("VZ")

    def pt_solve(self, p: CartParams, maxiter: int, nout: int, eps: float):
        """The reference loop with its exit test in one call."""
        r = ss_cart_result()
        rc = self._L.ss_cart_pt_solve(self._h, C.byref(p.c()), int(maxiter), int(nout), float(eps), C.byref(r))
        if rc not in (0, _lib.SS_NOT_CONVERGED):
            self._chk(rc, "ss_cart_pt_solve")
        return {"iters": r.iters, "converged": bool(r.converged), "err": r.err, "seconds": r.seconds}

    def pt_iterate_timed(self, p: CartParams, niter: int):
        it, k = C.c_double(), C.c_double()
        self._chk(self._L.ss_cart_pt_iterate_timed(self._h, C.byref(p.c()), int(niter), C.byref(it), C.byref(k)),
                  "ss_cart_pt_iterate_timed")
        return it.value, k.value
