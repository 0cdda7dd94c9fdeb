"""Pins the CPU reference adjoint (oracle/sspt_adjoint.c), which is the
checker of the HIP adjoint.

* Dot-product test of one adjoint step at a converged state:
  <(Phi(x*+hu) - Phi(x*-hu)) / 2h, w> = <u, T^T w> with Phi the oracle's own
  PT step (so the transposed stencils are checked against the forward code,
  not against a restatement of it).  A dropped term, a wrong index or a sign
  error in any reversed statement breaks the equality for random u, w.
* Gradient of J (volume-weighted mean VR in the inclusion) with respect to
  rho and log10 eta_ini against central differences of full forward solves
  along seeded random directions.
"""
import numpy as np
import pytest

from oracle.binding import ADJ_FIELDS, Oracle
from sphericalstokes_amd import setups
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_SHEAR

R0, RAD = 1.5, 0.1 ** 2       # inclusion of diapir_sector: centre r = 1.5, squared radius


def _converged(st, iters):
    o = Oracle(st)
    o.step(iters)
    return o


CASES = {
    "linear_free_slip": (lambda: setups.diapir_sector(9, n_exp=1.0, bc_mode=BC_FREE_SLIP), False),
    "powerlaw_free_slip": (lambda: setups.diapir_sector(9, n_exp=3.0, bc_mode=BC_FREE_SLIP), False),
    "linear_simple_shear": (lambda: setups.diapir_sector(9, n_exp=1.0, bc_mode=BC_SHEAR, sh=0.5), False),
    "powerlaw_relax_cutoffs": (lambda: setups.diapir_sector(9, n_exp=3.0, bc_mode=BC_FREE_SLIP, relax=0.5,
                                                            eta_min=1e-4, eta_max=1e4), True),
}


@pytest.mark.parametrize("case", list(CASES))
def test_adjoint_step_dot_product(case):
    make, relax = CASES[case]
    st = make()
    o = _converged(st, 30000)
    base = {n: o.f[n].copy(order="F") for n in o.f}
    rng = np.random.default_rng(5)
    u, w = {}, {}
    for n in ADJ_FIELDS:
        a = base[n]
        if n in ("eta", "eta_PL"):        # positive fields: relative perturbations
            du = a * rng.standard_normal(a.shape)
            du[~np.isfinite(du)] = 0.0
        else:
            du = np.abs(a).max() * rng.standard_normal(a.shape)
        u[n] = du
        w[n] = np.where(np.isfinite(a), rng.standard_normal(a.shape), 0.0)   # eta_PL is +inf where tII = 0
    if not relax:                         # eta_PL is not part of the state without relaxation
        u["eta_PL"][:] = 0.0
        w["eta_PL"][:] = 0.0
    h = 1e-6

    def phi(sign):
        f = {n: base[n].copy(order="F") for n in base}
        for n in ADJ_FIELDS:
            f[n] = np.asfortranarray(np.where(u[n] != 0.0, f[n] + sign * h * u[n], f[n]))
        ob = Oracle(st, f)
        ob.step(1)
        return ob.f

    fp, fm = phi(+1), phi(-1)
    with np.errstate(invalid="ignore"):
        lhs = sum(float(np.sum(np.where(w[n] != 0, (fp[n] - fm[n]) / (2 * h) * w[n], 0.0))) for n in ADJ_FIELDS)
    tw = o.adj_step(w)
    rhs = sum(float(np.sum(np.where(u[n] != 0, u[n] * tw[n], 0.0))) for n in ADJ_FIELDS)
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs), (lhs, rhs)


@pytest.mark.parametrize("case", ["linear_free_slip", "powerlaw_free_slip"])
def test_adjoint_gradient_matches_central_differences(case):
    st = {"linear_free_slip": lambda: setups.diapir_sector(7, n_exp=1.0, bc_mode=BC_FREE_SLIP),
          "powerlaw_free_slip": lambda: setups.diapir_sector(7, n_exp=3.0, bc_mode=BC_FREE_SLIP, vr=10.0)}[case]()
    iters = 20000
    o = _converged(st, iters)
    it, _, g_rho, g_ini = o.adj_solve(R0, RAD, 200000, 100, 1e-15)
    assert it > 0
    g_log = g_ini * o.f["eta_ini"] * np.log(10.0)
    base = {n: o.f[n].copy(order="F") for n in o.f}
    rng = np.random.default_rng(1)
    h = 1e-5
    iters_fd = 14000
    # 16 Gaussian directions per parameter from seed 1 (the adjoint
    # configuration's check); each perturbed solve starts from the converged
    # state
    for what in ("rho", "log10_eta_ini"):
        for d in range(16):
            dm = rng.standard_normal(base["rho"].shape)
            J = []
            for sgn in (+1, -1):
                f = {n: base[n].copy(order="F") for n in base}
                if what == "rho":
                    f["rho"] = np.asfortranarray(f["rho"] + sgn * h * dm)
                else:
                    f["eta_ini"] = np.asfortranarray(f["eta_ini"] * 10.0 ** (sgn * h * dm))
                ob = Oracle(st, f)
                ob.step(iters_fd)   # from the converged state: the 1e-5 change converges in fewer steps
                J.append(ob.objective(R0, RAD)[0])
            fd = (J[0] - J[1]) / (2 * h)
            ad = float(np.sum((g_rho if what == "rho" else g_log) * dm))
            assert abs(fd - ad) <= 1e-6 * abs(fd), (what, d, fd, ad)
