"""Model setups: the diapir and plateau configurations as plain numpy state.

A setup returns a :class:`Setup` holding the 1D coordinates, every field in
the reference's Julia layout (Fortran-ordered float64 arrays, r fastest) and
the scalar parameters of the PT iteration, so that the same input can be
given to the HIP library and to the CPU oracle.

* :func:`diapir` follows ``SPH_3D_Inclusion`` (SphericalStokes_diapir.jl:347-404)
  with every constant a keyword; its defaults are the script's values.
* :func:`diapir_sector` is the diapir in a wide sector r in [1,2],
  th in [pi/3, 2pi/3], phi in [-pi/6, pi/6] (rising sphere, radius 0.1).
* :func:`plateau_sector` is a thin-shell plateau load: r in [0.9, 1.0],
  a power-law crust of thickness 0.01 carrying a tanh-edged 0.005 bump over
  the central 60 % of a 30 deg x 40 deg sector, a linear mantle below.

Coordinates come from ``numpy.linspace``; Julia's ``range`` may round the
last bit of some points differently.  The library takes the caller's
coordinates, so a Julia caller passes its own.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .fields import BC_FREE_SLIP, BC_PURE_SHEAR, BC_SHEAR, FIELDS, shape


@dataclass
class Params:
    """Scalars of one PT iteration (names as in the reference scripts)."""
    _dr: float
    _dth: float
    _dph: float
    dmp: float
    vpdt: float
    lr: float
    tau_C: float = 1.0
    n_exp: float = 1.0
    g: float = 1.0
    r: float = 1.0
    sh: float = 0.0
    relax: float = 1.0
    eta_min: float = 0.0
    eta_max: float = math.inf
    bc_mode: int = BC_SHEAR
    err_drop_nan_vph: int = 0
    # 0 = the reference's maxloc!; m > 0 = etaSM spread over m cells with a
    # decay of 10 per cell (build option for sharp viscosity contrasts above ~1e3)
    etasm_spread: int = 0


@dataclass
class Setup:
    nr: int
    nth: int
    nph: int
    R: np.ndarray
    th: np.ndarray
    ph: np.ndarray
    params: Params
    fields: dict = field(default_factory=dict)
    name: str = ""

    def copy_fields(self) -> dict:
        return {k: np.array(v, order="F", copy=True) for k, v in self.fields.items()}


def _zeros(nr, nth, nph):
    return {n: np.zeros(shape(n, nr, nth, nph), dtype=np.float64, order="F") for n in FIELDS}


def _inclusion(A, R, th, ph, r, radius, val):
    """initialize_inclusion! (SphericalStokes_diapir.jl:127-137): squared
    distance compared with `radius`."""
    a = (R - r)[:, None, None]
    b = ((th - math.pi / 2.0) * r)[None, :, None]
    c = (ph * r)[None, None, :]
    d = (a * a + b * b) + c * c
    A[d < radius] = val


def _velocity_shear(Vph, th, r, fact):
    """initialize_velocity!(Vph, th, sh, r) (SphericalStokes_diapir.jl:139-146)."""
    Vph[...] = (((th - math.pi / 2.0) * r) * fact)[None, :, None]


def _faces(x):
    """Face positions of a uniform centre grid x: x[0]-dx/2, midpoints, x[-1]+dx/2."""
    dx = x[1] - x[0]
    return np.concatenate([[x[0] - 0.5 * dx], 0.5 * (x[:-1] + x[1:]), [x[-1] + 0.5 * dx]])


def pure_shear_velocity(th, ph, r, sh):
    """The far-field pure shear V = (0, -sh*r*(th-pi/2)/sin(th), sh*r*ph) at
    the staggered Vth and Vph positions.  It is divergence-free, and its
    normal components on the th and phi boundaries are what the pure-shear
    BC prescribes there (a build extension; the reference has simple shear
    only).  Returns (Vth[nth+1], Vph[nph+1]) as 1D profiles."""
    thf, phf = _faces(th), _faces(ph)
    return -(((thf - math.pi / 2.0) * r) * sh) / np.sin(thf), (phf * r) * sh


def _velocity_pure_shear(Vth, Vph, th, ph, r, sh):
    vt, vp = pure_shear_velocity(th, ph, r, sh)
    Vth[...] = vt[None, :, None]
    Vph[...] = vp[None, None, :]


def diapir(nr, nth, nph, *, eta0=1.0, vr=1e2, rho0=0.0, drho=-10.0, g=1.0, n_exp=5.0, tau_C=1.0,
           r=1000.0, lr=6.0, lth=None, lph=None, radius=1.0, sh=1.0, dmp=4.5, bc_mode=BC_SHEAR,
           relax=1.0, eta_min=0.0, eta_max=math.inf, init_shear=None) -> Setup:
    """SPH_3D_Inclusion setup, SphericalStokes_diapir.jl:350-404."""
    lth = lr / r if lth is None else lth
    lph = lr / r if lph is None else lph
    CFL = 1.0 / (2.0 + 4.5 * math.log10(vr))
    dr = lr / (nr - 1)
    R = np.linspace(r - lr / 2.0, r + lr / 2.0, nr)
    th = np.linspace(math.pi / 2.0 - lth / 2.0, math.pi / 2.0 + lth / 2.0, nth)
    ph = np.linspace(-lph / 2.0, lph / 2.0, nph)
    f = _zeros(nr, nth, nph)
    f["rho"][...] = rho0
    f["eta"][...] = eta0
    _inclusion(f["rho"], R, th, ph, r, radius, rho0 + drho)
    _inclusion(f["eta"], R, th, ph, r, radius, eta0 / vr)
    # far-field flow as the initial guess (initialize_velocity!, :403), for
    # both kinematic BCs; free slip starts at rest
    if init_shear if init_shear is not None else bc_mode in (BC_SHEAR, BC_PURE_SHEAR):
        if bc_mode == BC_PURE_SHEAR:
            _velocity_pure_shear(f["Vth"], f["Vph"], th, ph, r, sh)
        else:
            _velocity_shear(f["Vph"], th, r, sh)
    f["eta_ini"][...] = f["eta"]
    f["eta_PL"][...] = f["eta"]
    p = Params(_dr=(nr - 1) / lr, _dth=(nth - 1) / lth, _dph=(nph - 1) / lph, dmp=dmp, vpdt=dr * CFL,
               lr=lr, tau_C=tau_C, n_exp=n_exp, g=g, r=r, sh=sh, relax=relax, eta_min=eta_min,
               eta_max=eta_max, bc_mode=bc_mode)
    return Setup(nr, nth, nph, R, th, ph, p, f, name="diapir")


def diapir_sector(n, *, n_exp=1.0, vr=1e3, bc_mode=BC_FREE_SLIP, sh=0.0, relax=1.0, eta_min=0.0,
                  eta_max=math.inf, nth=None, nph=None) -> Setup:
    """Rising sphere in the sector r in [1,2], th in [pi/3,2pi/3], phi in [-pi/6,pi/6]:
    sphere of radius 0.1 at (1.5, pi/2, 0), drho = -1, g = 1, eta_in/eta_m = 1/vr."""
    s = diapir(n, nth or n, nph or n, eta0=1.0, vr=vr, rho0=0.0, drho=-1.0, g=1.0, n_exp=n_exp, tau_C=1.0,
               r=1.5, lr=1.0, lth=math.pi / 3.0, lph=math.pi / 3.0, radius=0.1 ** 2, sh=sh, dmp=4.5,
               bc_mode=bc_mode, relax=relax, eta_min=eta_min, eta_max=eta_max)
    s.name = "diapir_sector"
    return s


def lognormal_log10(shape, *, seed=0, std=0.5, sigma=8.0, clip=2.0) -> np.ndarray:
    """Smooth log-normal field in log10: independent N(0, 1) draws from
    numpy's PCG64 generator (np.random.default_rng(seed).standard_normal,
    C order over `shape`), smoothed by a separable Gaussian of `sigma` cells
    truncated at 3 sigma with mirrored walls (scipy.ndimage.gaussian_filter1d,
    mode='mirror', along each axis in turn), rescaled to standard deviation
    `std` and clipped to +-clip (a total contrast of 10^(2 clip))."""
    from scipy.ndimage import gaussian_filter1d
    x = np.random.default_rng(seed).standard_normal(shape)
    for ax in range(len(shape)):
        x = gaussian_filter1d(x, sigma, axis=ax, mode="mirror", truncate=3.0)
    x *= std / x.std()
    return np.clip(x, -clip, clip)


def diapir_lognormal(n, *, n_exp=3.0, seed=0, std=0.5, sigma=8.0, clip=2.0, eta_in=1e-3,
                     bc_mode=BC_FREE_SLIP, cfl_factor=1.0, etasm_spread=5) -> Setup:
    """diapir_sector (sphere of radius 0.1 at (1.5, pi/2, 0), drho = -1,
    power law n_exp) with a random matrix viscosity: log10 eta_matrix from
    lognormal_log10 (seed 0: N(0, 0.5^2), smoothed over 8 cells, contrast
    clipped to 1e4), inclusion viscosity eta_in.  vpdt follows the total
    viscosity contrast as in the reference (CFL = 1/(2 + 4.5 log10 vr)),
    times cfl_factor.  The inclusion's edge is a sharp jump against the
    matrix next to it (1.2e3 at most at 63^3, 1.4e4 at 95^3, 6.9e3 at 127^3:
    finer grids resolve more of the random field's high values): with the
    reference's etaSM any sharp jump above ~3e3 diverges within a hundred
    iterations, also at half the step, so this setup spreads etaSM over 5
    cells (etasm_spread = 5) by default."""
    s = diapir_sector(n, n_exp=n_exp, bc_mode=bc_mode)
    f = s.fields
    inc = f["rho"] != 0.0
    eta = np.asfortranarray(10.0 ** lognormal_log10((s.nr, s.nth, s.nph), seed=seed, std=std, sigma=sigma,
                                                    clip=clip))
    eta[inc] = eta_in
    for k in ("eta", "eta_ini", "eta_PL"):
        f[k][...] = eta
    vr = float(eta.max() / eta.min())
    CFL = cfl_factor / (2.0 + 4.5 * math.log10(vr))
    s.params.vpdt = (1.0 / (s.nr - 1)) * CFL
    s.params.etasm_spread = etasm_spread
    s.name = "diapir_lognormal"
    return s


def plateau_sector(n, *, nth=None, nph=None, n_crust=3.0, eta_mantle=1e-2, rho_crust=1.0, rho_mantle=1.2,
                   eta_air=1e-3, dmp=4.5) -> Setup:
    """Thin-shell plateau: r in [0.9, 1.0] (planet radius 1), th span 30 deg,
    phi span 40 deg, a power-law crust 0.01 thick under a surface raised by a
    tanh-edged 0.005 bump over the central 60 % of the sector; linear mantle
    (viscosity ratio eta_mantle) below, weak low-density air above the surface.
    Free slip on all faces."""
    nth = nth or n
    nph = nph or n
    r_top, r_bot = 1.0, 0.9
    lr = r_top - r_bot
    lth, lph = math.radians(30.0), math.radians(40.0)
    r_c = 0.5 * (r_top + r_bot)
    R = np.linspace(r_bot, r_top, n)
    th = np.linspace(math.pi / 2.0 - lth / 2.0, math.pi / 2.0 + lth / 2.0, nth)
    ph = np.linspace(-lph / 2.0, lph / 2.0, nph)
    w_th, w_ph = 0.03 * lth, 0.03 * lph
    bth = 0.25 * (1 + np.tanh((th - (math.pi / 2 - 0.3 * lth)) / w_th)) * (1 - np.tanh((th - (math.pi / 2 + 0.3 * lth)) / w_th))
    bph = 0.25 * (1 + np.tanh((ph + 0.3 * lph) / w_ph)) * (1 - np.tanh((ph - 0.3 * lph) / w_ph))
    surf = 0.99 + 0.005 * bth[:, None] * bph[None, :]         # surface radius over (th, phi)
    moho = surf - 0.01                                        # base of the crust
    Rg = R[:, None, None]
    crust = (Rg <= surf[None]) & (Rg > moho[None])
    air = Rg > surf[None]
    f = _zeros(n, nth, nph)
    f["rho"][...] = rho_mantle
    f["eta"][...] = eta_mantle
    f["rho"][crust] = rho_crust
    f["eta"][crust] = 1.0
    f["rho"][air] = 0.0
    f["eta"][air] = eta_air
    f["eta_ini"][...] = f["eta"]
    f["eta_PL"][...] = f["eta"]
    vr = float(f["eta"].max() / f["eta"].min())
    CFL = 1.0 / (2.0 + 4.5 * math.log10(vr))
    dr = lr / (n - 1)
    # The reference power law has one exponent for the whole domain.  Where
    # tII << tau_C it tends to eta = 2*eta_ini (linear), which is the regime
    # of the mantle and air here; the crust carries the stresses that reach
    # tau_C and so the n_crust rheology.
    p = Params(_dr=(n - 1) / lr, _dth=(nth - 1) / lth, _dph=(nph - 1) / lph, dmp=dmp, vpdt=dr * CFL, lr=lr,
               tau_C=1.0, n_exp=n_crust, g=1.0, r=r_c, sh=0.0, bc_mode=BC_FREE_SLIP)
    return Setup(n, nth, nph, R, th, ph, p, f, name="plateau_sector")


def random_state(s: Setup, seed: int = 0, scale: float = 1e-2) -> Setup:
    """Fills V, P and stresses with seeded random values (parity tests on a
    non-trivial state without running a solve first).  The auxiliary fields
    that single kernels read (etaSM for compute_P!, DT_R for compute_V!, tII
    for power_law!) get positive values of a realistic size, so that every
    per-kernel call changes its outputs: etaSM the 7-point maximum of eta
    times a random factor in [1, 2), DT_R of the size of vpdt^2/etaSM, tII
    of the size of the stresses."""
    rng = np.random.default_rng(seed)
    for n in ("VR", "Vth", "Vph", "P", "tau_RR", "tau_thth", "tau_phph", "tau_Rth", "tau_Rph", "tau_thph"):
        a = s.fields[n]
        a[...] = scale * rng.standard_normal(a.shape)
    e = s.fields["eta"]
    e[...] = np.exp(rng.standard_normal(e.shape))
    s.fields["eta_ini"][...] = np.exp(rng.standard_normal(e.shape))
    s.fields["eta_PL"][...] = s.fields["eta"]
    s.fields["rho"][...] = rng.standard_normal(e.shape)
    m = e.copy()
    for ax in range(3):
        for sh in (1, -1):
            m = np.maximum(m, np.roll(e, sh, axis=ax))
    s.fields["etaSM"][...] = m * (1.0 + rng.random(e.shape))
    p = s.params
    s.fields["DT_R"][...] = (p.vpdt * p.vpdt) / s.fields["etaSM"] * (0.5 + rng.random(e.shape))
    s.fields["tII"][...] = scale * np.exp(rng.standard_normal(e.shape))
    return s
