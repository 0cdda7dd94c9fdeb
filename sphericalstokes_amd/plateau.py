"""The plateau study's continuation workflow on the GPU solver
(scripts_plateau/SphericalStokes_plateau.jl, SPH_3D_Plateau :368-595).

The reference builds a thin spherical shell under a raised plateau from
dimensional parameters (:370-398), scales them (:400-439), fills density and
viscosity with ``initialize_geometry!`` (:120-167: mantle, lithospheric
mantle, crust with a root under the highland and a tapered transition zone,
sticky air above the surface) and then runs eight convergence steps
(:485-559), each a PT loop of at most ``maxiter`` iterations with the exit
test every ``nout``:

* steps 2-6 raise the crustal viscosity (eta_c -> eta_c_2 ... eta_c_6): eta
  and eta_ini are re-initialised from the geometry and vr, CFL and vpdt are
  recomputed; V, P and tau carry over;
* step 7 switches the power-law exponent to n_exp_PL (3), step 8 to
  n_exp_PL2 (6).

``plateau_reference(nr, nth, nph)`` gives the setup (grid, scaled
parameters, fields) and ``run_continuation(solver, setup)`` the eight steps
on a ``Solver`` holding that setup, optionally writing each step's output
folder in the reference's format (``out_plateau_<it>``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace
from pathlib import Path

import numpy as np

from .fields import BC_FREE_SLIP
from .setups import Params, Setup, _zeros


@dataclass
class PlateauParams:
    """Dimensional parameters of SphericalStokes_plateau.jl:370-398."""
    eta_m: float = 1e20      # Pa s, mantle
    eta_l: float = 1e22      # lithospheric mantle
    eta_c: float = 1e20      # crust, step 1
    eta_c_steps: tuple = (5e20, 1e21, 3e21, 6e21, 1e22)   # crust, steps 2-6
    eta_a: float = 5e18      # sticky air
    rho_m: float = 3300.0
    rho_l: float = 3300.0
    rho_c: float = 2800.0
    rho_a: float = 0.0
    hm: float = 360e3        # m, mantle thickness
    hl: float = 87e3         # lithospheric mantle
    hr: float = 28e3         # root
    hc: float = 35e3         # crust
    he: float = 5e3          # elevation
    ha: float = 50e3         # sticky air
    dl: float = 700e3        # lowland distance
    dt: float = 100e3        # transition zone
    dh: float = 600e3        # highland distance
    g: float = 9.81
    n_exp: float = 1.0       # steps 1-6
    n_exp_PL: float = 3.0    # step 7
    n_exp_PL2: float = 6.0   # step 8
    tau_C: float = 24e6      # Pa, reference stress of the power law
    r: float = 6371e3        # m, planet radius
    dmp: float = 4.5


@dataclass
class Scaled:
    """The scaled (non-dimensional) values the reference computes (:406-439)."""
    sl: float
    s_eta: float
    s_rho: float
    lr: float
    lth: float
    lph: float
    rho: dict
    eta: dict
    eta_c_steps: tuple
    r: float
    hm: float
    hl: float
    hr: float
    hc: float
    he: float
    ha: float
    dl: float
    dt: float
    dh: float
    g: float
    tau_C: float


def scale(pp: PlateauParams) -> Scaled:
    lr_d = pp.hm + pp.hl + pp.hr + pp.hc + pp.he + pp.ha
    lth_d = (pp.dl + pp.dt + pp.dh) / pp.r
    sl, se, sr = lr_d, pp.eta_m, pp.rho_m
    return Scaled(sl=sl, s_eta=se, s_rho=sr, lr=lr_d / sl, lth=lth_d, lph=lth_d,
                  rho={"m": pp.rho_m / sr, "l": pp.rho_l / sr, "c": pp.rho_c / sr, "a": pp.rho_a / sr},
                  eta={"m": pp.eta_m / se, "l": pp.eta_l / se, "c": pp.eta_c / se, "a": pp.eta_a / se},
                  eta_c_steps=tuple(x / se for x in pp.eta_c_steps),
                  r=pp.r / sl, hm=pp.hm / sl, hl=pp.hl / sl, hr=pp.hr / sl, hc=pp.hc / sl, he=pp.he / sl,
                  ha=pp.ha / sl, dl=pp.dl / sl, dt=pp.dt / sl, dh=pp.dh / sl,
                  g=pp.g / se ** 2 * sr ** 2 * sl ** 3, tau_C=pp.tau_C / se ** 2 * sr * sl ** 2)


def initialize_geometry(A, R, th, ph, r, he, hc, hr, hl, dt, gl, gc, ga):
    """initialize_geometry! (SphericalStokes_plateau.jl:120-167) on the 3D
    array A over the 1D axes R, th, ph: the same thirteen conditions in the
    same order, later ones overriding earlier ones."""
    x = (np.asarray(R) - r)[:, None, None]
    th3 = np.asarray(th)[None, :, None]
    ph3 = np.asarray(ph)[None, None, :]
    y = (th3 - math.pi / 2.0) * r          # (theta - pi/2) r
    z = ph3 * r                            # phi r
    tr = (hr / dt) * (math.pi / 2.0 - th3) * r
    pr = (hr / dt) * ph3 * r
    te = -((he / dt) * (math.pi / 2.0 - th3) * r)
    pe = -((he / dt) * ph3 * r)
    shp = A.shape
    b = lambda c: np.broadcast_to(c, shp)
    A[b(x > -hc - hr - hl)] = gl
    A[b((x > -hc) & (y < 0))] = gc
    A[b((x > -hc) & (z > 0))] = gc
    A[b((x > -hc - hr) & (y >= dt) & (z <= -dt))] = gc
    A[b((x > -hc + pr) & (y >= dt) & (z <= 0) & (z >= -dt))] = gc
    A[b((x > -hc + tr) & (y >= 0) & (y <= dt) & (z <= -dt))] = gc
    A[b((x > np.maximum(-hc + tr, -hc + pr)) & (y >= 0) & (y <= dt) & (z <= 0) & (z >= -dt))] = gc
    A[b((x > 0) & (y < 0))] = ga
    A[b((x > 0) & (z > 0))] = ga
    A[b((x > he) & (y >= dt) & (z <= -dt))] = ga
    A[b((x > pe) & (y >= dt) & (z <= 0) & (z >= -dt))] = ga
    A[b((x > te) & (y >= 0) & (y <= dt) & (z <= -dt))] = ga
    A[b((x > np.minimum(te, pe)) & (y >= 0) & (y <= dt) & (z <= 0) & (z >= -dt))] = ga
    return A


def _vpdt(setup_nr, lr, etas):
    vr = max(etas) / min(etas)
    CFL = 1.0 / (2.0 + 4.5 * math.log10(vr))
    return (lr / (setup_nr - 1)) * CFL, vr


def plateau_reference(nr=70, nth=5, nph=5, pp: PlateauParams | None = None) -> Setup:
    """SPH_3D_Plateau's grid, scaled parameters and initial fields
    (SphericalStokes_plateau.jl:400-481); the defaults are the script's
    (NR, NTH, NPH = 70, 5, 5)."""
    pp = pp or PlateauParams()
    s = scale(pp)
    base = s.hc + s.hr + s.hl + s.hm
    R = np.linspace(s.r - base, s.r + s.lr - base, nr)
    th = np.linspace(math.pi / 2.0 - s.dl / s.r, math.pi / 2.0 + (s.dt + s.dh) / s.r, nth)
    ph = np.linspace(-s.dl / s.r, (s.dt + s.dh) / s.r, nph)
    f = _zeros(nr, nth, nph)
    f["rho"][...] = s.rho["m"]
    f["eta"][...] = s.eta["m"]
    initialize_geometry(f["rho"], R, th, ph, s.r, s.he, s.hc, s.hr, s.hl, s.dt, s.rho["l"], s.rho["c"], s.rho["a"])
    initialize_geometry(f["eta"], R, th, ph, s.r, s.he, s.hc, s.hr, s.hl, s.dt, s.eta["l"], s.eta["c"], s.eta["a"])
    f["eta_ini"][...] = f["eta"]
    f["eta_PL"][...] = f["eta"]
    vpdt, _ = _vpdt(nr, s.lr, (s.eta["m"], s.eta["l"], s.eta["c"], s.eta["a"]))
    p = Params(_dr=(nr - 1) / s.lr, _dth=(nth - 1) / s.lth, _dph=(nph - 1) / s.lph, dmp=pp.dmp, vpdt=vpdt,
               lr=s.lr, tau_C=s.tau_C, n_exp=pp.n_exp, g=s.g, r=s.r, sh=0.0, bc_mode=BC_FREE_SLIP,
               err_drop_nan_vph=1)
    st = Setup(nr, nth, nph, R, th, ph, p, f, name="plateau_reference")
    st.plateau = (pp, s)
    return st


def step_update(setup: Setup, p: Params, it: int):
    """Changes at the start of convergence step `it` (1-based,
    SphericalStokes_plateau.jl:490-521) to the parameters p of the step
    before: returns (params, eta or None); eta is the re-initialised
    viscosity (also the new eta_ini) for steps 2-6."""
    pp, s = setup.plateau
    if 2 <= it <= 6:
        ec = s.eta_c_steps[it - 2]
        eta = np.full((setup.nr, setup.nth, setup.nph), s.eta["m"], order="F")
        initialize_geometry(eta, setup.R, setup.th, setup.ph, s.r, s.he, s.hc, s.hr, s.hl, s.dt, s.eta["l"], ec,
                            s.eta["a"])
        vpdt, _ = _vpdt(setup.nr, s.lr, (s.eta["m"], s.eta["l"], ec, s.eta["a"]))
        return replace(p, vpdt=vpdt), eta
    if it == 7:
        return replace(p, n_exp=pp.n_exp_PL), None
    if it == 8:
        return replace(p, n_exp=pp.n_exp_PL2), None
    return p, None


def pt_loop(solver, p: Params, maxiter: int, nout: int, eps: float):
    """The reference's PT loop (SphericalStokes_plateau.jl:525-559) on the
    fused iteration: nout iterations at a time, the exit test
    (err < eps and iter > 20) after each, err dropping a NaN Vph component.
    Returns (ittot, err_evo, iter_evo); ittot keeps the reference's value on
    a break (the iteration before it)."""
    err_evo, iter_evo = [], []
    it = ittot = 0
    while it < maxiter:
        n = min(nout - it % nout, maxiter - it)
        solver.pt_iterate(p, n)
        it += n
        if it % nout == 0:
            err, _ = solver.check_err(drop_nan_vph=True)
            err_evo.append(err)
            iter_evo.append(it)
            if err < eps and it > 20:
                break
        ittot = it
    return ittot, err_evo, iter_evo


def run_continuation(solver, setup: Setup, *, nt: int = 8, maxiter=200, nout: int = 10, eps: float = 5e-7,
                     save_dir=None, log=None):
    """The eight convergence steps on `solver` (which holds `setup`'s fields).
    maxiter is the PT iteration limit of every step, or a sequence with one
    per step.  Returns one dict per step (ittot, err_evo, iter_evo, n_exp,
    vpdt).  With save_dir, step it's fields go to save_dir/out_plateau_<it>
    in the reference's format (its plateau _phys.inf line)."""
    from . import io

    pp, s = setup.plateau
    out = []
    p = setup.params
    for it in range(1, nt + 1):
        p, eta = step_update(setup, p, it)
        if eta is not None:
            solver.upload("eta", eta)
            solver.upload("eta_ini", eta)
            solver.upload("eta_PL", eta)
        mi = maxiter[it - 1] if hasattr(maxiter, "__len__") else maxiter
        ittot, err_evo, iter_evo = pt_loop(solver, p, int(mi), nout, eps)
        out.append({"it": it, "ittot": ittot, "err_evo": err_evo, "iter_evo": iter_evo, "n_exp": p.n_exp,
                    "vpdt": p.vpdt})
        if log:
            log(f"step {it}: {ittot} iterations, err {err_evo[-1] if err_evo else float('nan'):.3e}")
        if save_dir is not None:
            ec = s.eta_c_steps[min(it, 6) - 2] if it >= 2 else s.eta["c"]
            etas = (s.eta["m"], s.eta["l"], ec, s.eta["a"])
            vr = max(etas) / min(etas)
            phys = ("%d %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f"
                    % (1, s.sl, s.s_eta, s.s_rho, s.eta["m"], s.eta["l"], ec, s.eta["a"], s.rho["m"], s.rho["l"],
                       s.rho["c"], s.rho["a"], s.hm, s.hl, s.hr, s.hc, s.he, p.n_exp, s.tau_C, vr, s.g * 1e15,
                       s.r, p.dmp))
            io.save_reference_format(Path(save_dir) / f"out_plateau_{it}", solver, replace(setup, params=p),
                                     num=1, eps=eps * 1e7, runtime=0.0, err_evo=err_evo, iter_evo=iter_evo,
                                     phys_line=phys)
    return out
