"""Pins of the CPU oracle against mathematics (the reference ships no golden
vectors or tests, SURVEY.md has none to offer).

* compute_P!: strain rates and divergence of a smooth analytic velocity field
  converge at second order to the textbook spherical formulas (sympy); with
  eta -> inf and tau = 0 the stress update returns G*D, so D is read off.
* compute_V!: the momentum residual of analytic stress / pressure / density
  fields converges at second order to div(sigma) - rho g e_r in spherical
  coordinates.
* maxloc!: brute-force 7-point max.  copy_BC! / BC_*!: independent
  "last BC wins" formulation and the free-slip interpolation property.
* compute_TII! / power_law!: closed forms and limits.
* The whole loop: V = 0 is a fixed point; a uniform-density column relaxes
  to hydrostatic pressure dP/dr = -rho g.
A dropped metric term, a sign or an index slip in the oracle fails one of them.
"""
import math

import numpy as np
import pytest
import sympy as sp

from oracle.binding import Oracle
from sphericalstokes_amd import setups
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_PURE_SHEAR, BC_SHEAR

r_, t_, f_ = sp.symbols("r theta phi", real=True)

# analytic velocity and stress fields (smooth, all components non-trivial)
VR_E = sp.sin(2 * r_) * sp.cos(t_) * sp.sin(f_ + 0.3) + 0.2 * r_ ** 2
VT_E = sp.cos(r_ + t_) * sp.cos(2 * f_) + 0.1 * r_ * t_
VP_E = sp.exp(-r_) * sp.sin(3 * t_) * sp.cos(f_ - 0.2) + 0.3 * sp.sin(r_ * f_)
P_E = sp.cos(r_ * t_) + f_ ** 2
TRR_E = sp.sin(r_ + f_) * t_
TTT_E = sp.cos(2 * r_) * sp.sin(t_ + f_)
TPP_E = r_ * sp.cos(t_ - f_)
TRT_E = sp.sin(r_ * t_) + f_
TRP_E = sp.cos(r_) * sp.sin(2 * f_ + t_)
TTP_E = sp.exp(0.3 * r_) * sp.cos(t_) * sp.sin(f_)
RHO_E = sp.cos(r_ + f_) + t_


def _lam(e):
    return sp.lambdify((r_, t_, f_), e, "numpy")


def _strain_exprs():
    vr, vt, vp = VR_E, VT_E, VP_E
    err = sp.diff(vr, r_)
    ett = sp.diff(vt, t_) / r_ + vr / r_
    epp = sp.diff(vp, f_) / (r_ * sp.sin(t_)) + vr / r_ + vt * sp.cot(t_) / r_
    div = err + ett + epp
    d_rt = sp.diff(vt, r_) - vt / r_ + sp.diff(vr, t_) / r_          # 2 e_rth
    d_rp = sp.diff(vp, r_) - vp / r_ + sp.diff(vr, f_) / (r_ * sp.sin(t_))
    d_tp = sp.diff(vp, t_) / r_ - vp * sp.cot(t_) / r_ + sp.diff(vt, f_) / (r_ * sp.sin(t_))
    return {k: _lam(v) for k, v in dict(div=div, Drr=err - div / 3, Dtt=ett - div / 3, Dpp=epp - div / 3,
                                         Drt=d_rt, Drp=d_rp, Dtp=d_tp).items()}


def _divsig_exprs():
    srr, stt, spp = -P_E + TRR_E, -P_E + TTT_E, -P_E + TPP_E
    srt, srp, stp = TRT_E, TRP_E, TTP_E
    s, c = sp.sin(t_), sp.cot(t_)
    fr = sp.diff(srr, r_) + sp.diff(srt, t_) / r_ + sp.diff(srp, f_) / (r_ * s) + (2 * srr - stt - spp + srt * c) / r_
    ft = sp.diff(srt, r_) + sp.diff(stt, t_) / r_ + sp.diff(stp, f_) / (r_ * s) + ((stt - spp) * c + 3 * srt) / r_
    fp = sp.diff(srp, r_) + sp.diff(stp, t_) / r_ + sp.diff(spp, f_) / (r_ * s) + (3 * srp + 2 * stp * c) / r_
    return _lam(fr - RHO_E), _lam(ft), _lam(fp)


def _grid(n):
    st = setups.diapir_sector(n, n_exp=1.0, bc_mode=BC_FREE_SLIP)
    R, th, ph = st.R, st.th, st.ph
    dr, dt, dp = R[1] - R[0], th[1] - th[0], ph[1] - ph[0]
    return st, R, th, ph, dr, dt, dp


def _mesh(a, b, c):
    return np.meshgrid(a, b, c, indexing="ij")


def _positions(R, th, ph, dr, dt, dp):
    Rf = np.concatenate([R - dr / 2, [R[-1] + dr / 2]])
    Tf = np.concatenate([th - dt / 2, [th[-1] + dt / 2]])
    Pf = np.concatenate([ph - dp / 2, [ph[-1] + dp / 2]])
    return Rf, Tf, Pf


def _p_errors(n):
    st, R, th, ph, dr, dt, dp = _grid(n)
    Rf, Tf, Pf = _positions(R, th, ph, dr, dt, dp)
    f = st.fields
    for k in f:
        f[k][...] = 0.0
    f["VR"][...] = _lam(VR_E)(*_mesh(Rf, th, ph))
    f["Vth"][...] = _lam(VT_E)(*_mesh(R, Tf, ph))
    f["Vph"][...] = _lam(VP_E)(*_mesh(R, th, Pf))
    f["etaSM"][...] = 1.0
    f["eta"][...] = 1e300
    o = Oracle(st)
    o.compute_P()
    p = st.params
    gs = ((4.0 * math.pi) * p.vpdt) / p.lr
    kb = (((p.dmp * 2.0) * math.pi) * p.vpdt) / p.lr
    ex = _strain_exprs()
    C = _mesh(R, th, ph)
    errs = {}
    errs["div"] = np.abs(-o.f["P"] / kb - ex["div"](*C)).max()
    errs["Drr"] = np.abs(o.f["tau_RR"] / (2 * gs) - ex["Drr"](*C)).max()
    errs["Dtt"] = np.abs(o.f["tau_thth"] / (2 * gs) - ex["Dtt"](*C)).max()
    errs["Dpp"] = np.abs(o.f["tau_phph"] / (2 * gs) - ex["Dpp"](*C)).max()
    E_rt = _mesh(R[:-1] + dr / 2, th[:-1] + dt / 2, ph[1:-1])
    E_rp = _mesh(R[:-1] + dr / 2, th[1:-1], ph[:-1] + dp / 2)
    E_tp = _mesh(R[1:-1], th[:-1] + dt / 2, ph[:-1] + dp / 2)
    errs["Drt"] = np.abs(o.f["tau_Rth"] / gs - ex["Drt"](*E_rt)).max()
    errs["Drp"] = np.abs(o.f["tau_Rph"] / gs - ex["Drp"](*E_rp)).max()
    errs["Dtp"] = np.abs(o.f["tau_thph"] / gs - ex["Dtp"](*E_tp)).max()
    return errs


def test_compute_P_strain_rates_second_order():
    e1, e2, e3 = _p_errors(11), _p_errors(21), _p_errors(41)
    for k in e1:
        assert e3[k] < 2e-3, (k, e3[k])
        assert e1[k] / e2[k] > 3.3 and e2[k] / e3[k] > 3.6, (k, e1[k], e2[k], e3[k])


def _v_errors(n):
    st, R, th, ph, dr, dt, dp = _grid(n)
    f = st.fields
    for k in f:
        f[k][...] = 0.0
    C = _mesh(R, th, ph)
    f["P"][...] = _lam(P_E)(*C)
    f["tau_RR"][...] = _lam(TRR_E)(*C)
    f["tau_thth"][...] = _lam(TTT_E)(*C)
    f["tau_phph"][...] = _lam(TPP_E)(*C)
    f["rho"][...] = _lam(RHO_E)(*C)
    f["DT_R"][...] = 1.0
    f["tau_Rth"][...] = _lam(TRT_E)(*_mesh(R[:-1] + dr / 2, th[:-1] + dt / 2, ph[1:-1]))
    f["tau_Rph"][...] = _lam(TRP_E)(*_mesh(R[:-1] + dr / 2, th[1:-1], ph[:-1] + dp / 2))
    f["tau_thph"][...] = _lam(TTP_E)(*_mesh(R[1:-1], th[:-1] + dt / 2, ph[:-1] + dp / 2))
    st.params.g = 1.0
    o = Oracle(st)
    o.compute_V()
    fr, ft, fp = _divsig_exprs()
    # interior faces written by compute_V!
    eR = np.abs(o.f["VR"][1:-1, 1:-1, 1:-1] - fr(*_mesh(R[:-1] + dr / 2, th[1:-1], ph[1:-1]))).max()
    eT = np.abs(o.f["Vth"][1:-1, 1:-1, 1:-1] - ft(*_mesh(R[1:-1], th[:-1] + dt / 2, ph[1:-1]))).max()
    eP = np.abs(o.f["Vph"][1:-1, 1:-1, 1:-1] - fp(*_mesh(R[1:-1], th[1:-1], ph[:-1] + dp / 2))).max()
    return {"VR": eR, "Vth": eT, "Vph": eP}


def test_compute_V_momentum_second_order():
    e1, e2, e3 = _v_errors(11), _v_errors(21), _v_errors(41)
    for k in e1:
        assert e3[k] < 5e-3, (k, e3[k])
        assert e1[k] / e2[k] > 3.3 and e2[k] / e3[k] > 3.6, (k, e1[k], e2[k], e3[k])


def test_maxloc_brute_force():
    st = setups.random_state(setups.diapir(6, 7, 5), seed=3)
    o = Oracle(st)
    eta = o.f["eta"].copy()
    o.f["etaSM"][...] = -1.0
    o.maxloc()
    sm = o.f["etaSM"]
    for i in range(1, 5):
        for j in range(1, 6):
            for k in range(1, 4):
                nb = [eta[i, j, k], eta[i - 1, j, k], eta[i + 1, j, k], eta[i, j - 1, k], eta[i, j + 1, k],
                      eta[i, j, k - 1], eta[i, j, k + 1]]
                assert sm[i, j, k] == max(nb)
    # boundary untouched by maxloc!, then copy_BC! takes the nearest interior value
    assert (sm[0] == -1.0).all()
    o.copy_bc("etaSM")
    I = np.clip(np.arange(6), 1, 4)[:, None, None]
    J = np.clip(np.arange(7), 1, 5)[None, :, None]
    K = np.clip(np.arange(5), 1, 3)[None, None, :]
    assert np.array_equal(o.f["etaSM"], sm[I, J, K])


def _last_bc_wins(V0, comp, mode, th, ph, r, sh):
    """Final ghost values of the BC_R! -> BC_th! -> BC_ph! sequence written as
    'the last BC among the ghost directions decides, from the final inward
    value' (independent of the sequential oracle code)."""
    V = V0.copy()
    n = V.shape
    out = V.copy()

    def ghost(d, idx):
        return idx == 0 or idx == n[d] - 1

    def inward(d, idx):
        return 1 if idx == 0 else n[d] - 2

    def final(i, j, k):
        pos = [i, j, k]
        for d in (2, 1, 0):
            if ghost(d, pos[d]):
                q = list(pos)
                q[d] = inward(d, pos[d])
                side = 0 if pos[d] == 0 else 1
                if d == 0:
                    return (-1.0 if comp == 0 else 1.0) * final(*q)
                if d == 1:
                    if comp == 1:
                        if mode == BC_PURE_SHEAR:
                            tb = th[0 if side == 0 else -1]
                            vb = -((((tb - math.pi / 2) * r) * sh) / math.sin(tb))
                            return 2.0 * vb - final(*q)
                        return -final(*q)
                    if mode == BC_SHEAR:
                        return 0.0 if comp == 0 else ((th[0 if side == 0 else -1] - math.pi / 2) * r) * sh
                    return final(*q)
                if comp == 2:
                    if mode == BC_SHEAR:
                        return ((2.0 * (th[j] - math.pi / 2)) * r) * sh - final(*q)
                    if mode == BC_FREE_SLIP:
                        return -final(*q)
                    return 2.0 * ((ph[0 if side == 0 else -1] * r) * sh) - final(*q)
                return 0.0 if mode == BC_SHEAR else final(*q)
        return V[i, j, k]

    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                out[i, j, k] = final(i, j, k)
    return out


@pytest.mark.parametrize("mode", [BC_SHEAR, BC_FREE_SLIP, BC_PURE_SHEAR])
def test_bc_sequence_last_bc_wins(mode):
    st = setups.diapir(5, 6, 7, bc_mode=mode, sh=0.7)
    rng = np.random.default_rng(11)
    for k in ("VR", "Vth", "Vph"):
        st.fields[k][...] = rng.standard_normal(st.fields[k].shape)
    o = Oracle(st)
    V0 = {k: o.f[k].copy() for k in ("VR", "Vth", "Vph")}
    o.bc_R()
    o.bc_th()
    o.bc_ph()
    for c, k in enumerate(("VR", "Vth", "Vph")):
        exp = _last_bc_wins(V0[k], c, mode, st.th, st.ph, st.params.r, st.params.sh)
        assert np.array_equal(o.f[k], exp), k


def test_free_slip_interpolates_zero_normal_velocity():
    st = setups.diapir(6, 7, 8, bc_mode=BC_FREE_SLIP)
    rng = np.random.default_rng(2)
    for k in ("VR", "Vth", "Vph"):
        st.fields[k][...] = rng.standard_normal(st.fields[k].shape)
    o = Oracle(st)
    o.bc_R(); o.bc_th(); o.bc_ph()
    VR, VT, VP = o.f["VR"], o.f["Vth"], o.f["Vph"]
    assert np.allclose(VR[0] + VR[1], 0) and np.allclose(VR[-1] + VR[-2], 0)
    assert np.allclose(VT[:, 0] + VT[:, 1], 0) and np.allclose(VT[:, -1] + VT[:, -2], 0)
    assert np.allclose(VP[:, :, 0] + VP[:, :, 1], 0) and np.allclose(VP[:, :, -1] + VP[:, :, -2], 0)
    # tangential components: zero normal derivative (ghost = neighbour) away from edges
    assert np.array_equal(VT[0, 1:-1, 1:-1], VT[1, 1:-1, 1:-1])
    assert np.array_equal(VR[1:-1, 0, 1:-1], VR[1:-1, 1, 1:-1])


def test_compute_TII_closed_form():
    st = setups.random_state(setups.diapir(7, 6, 8), seed=4, scale=1.0)
    o = Oracle(st)
    o.compute_TII()
    f = o.f
    a, b, c = f["tau_RR"][1:-1, 1:-1, 1:-1], f["tau_thth"][1:-1, 1:-1, 1:-1], f["tau_phph"][1:-1, 1:-1, 1:-1]
    RT, RP, TP = f["tau_Rth"], f["tau_Rph"], f["tau_thph"]
    x = (RT[:-1, :-1] + RT[:-1, 1:] + RT[1:, :-1] + RT[1:, 1:]) / 4      # r-th edges around each centre
    y = (RP[:-1, :, :-1] + RP[:-1, :, 1:] + RP[1:, :, :-1] + RP[1:, :, 1:]) / 4
    z = (TP[:, :-1, :-1] + TP[:, :-1, 1:] + TP[:, 1:, :-1] + TP[:, 1:, 1:]) / 4
    exp = np.sqrt(0.5 * (a ** 2 + b ** 2 + c ** 2) + x ** 2 + y ** 2 + z ** 2)
    assert np.allclose(f["tII"][1:-1, 1:-1, 1:-1], exp, rtol=1e-14, atol=0)


def test_power_law_limits():
    st = setups.diapir(5, 5, 5, n_exp=3.0)
    f = st.fields
    rng = np.random.default_rng(5)
    f["eta_ini"][...] = np.exp(rng.standard_normal(f["eta"].shape))
    ini = f["eta_ini"].copy()
    o = Oracle(st)
    o.f["tII"][...] = 1.0                      # tII = tau_C -> eta = eta_ini
    o.power_law()
    assert np.allclose(o.f["eta"], ini, rtol=2e-16 * 4)
    o.f["tII"][...] = 0.0                      # tII -> 0 with n > 1 -> 2 eta_ini
    o.power_law()
    assert np.allclose(o.f["eta"], 2 * ini, rtol=4e-16, atol=0)
    o.f["tII"][...] = 2.0                      # closed form
    o.power_law()
    pl = ini * 2.0 ** (-2.0)
    assert np.allclose(o.f["eta"], 2 / (1 / ini + 1 / pl), rtol=1e-15)
    st.params.n_exp = 1.0
    o.set_params(st.params)
    o.f["tII"][...] = rng.random(ini.shape) * 10
    o.power_law()
    assert np.allclose(o.f["eta"], ini, rtol=1e-15)


def test_power_law_relaxation_and_cutoffs():
    st = setups.diapir(5, 5, 5, n_exp=3.0, relax=0.5, eta_min=0.3, eta_max=3.0)
    f = st.fields
    rng = np.random.default_rng(6)
    ini = np.exp(rng.standard_normal(f["eta"].shape))
    f["eta_ini"][...] = ini
    old = np.exp(rng.standard_normal(f["eta"].shape))
    f["eta_PL"][...] = old
    f["tII"][...] = rng.random(ini.shape) * 3 + 0.1
    o = Oracle(st)
    o.power_law()
    pl = ini * (f["tII"] / 1.0) ** (-2.0)
    relaxed = np.sqrt(pl * old)                      # relax = 0.5: geometric mean
    assert np.allclose(o.f["eta_PL"], relaxed, rtol=1e-13)
    assert np.allclose(o.f["eta"], np.clip(2 / (1 / ini + 1 / relaxed), 0.3, 3.0), rtol=1e-13)


def test_zero_state_is_fixed_point():
    st = setups.diapir_sector(9, n_exp=3.0, bc_mode=BC_FREE_SLIP)
    st.fields["rho"][...] = 0.0
    o = Oracle(st)
    o.step(3)
    for k in ("VR", "Vth", "Vph", "P", "tau_RR", "tau_Rth", "tau_thph"):
        assert not o.f[k].any(), k


def test_hydrostatic_column():
    """Uniform density, free slip: V -> 0 and dP/dr -> -rho g (the radial
    momentum residual with tau = 0 is -dP/dr + (2 s_rr - s_thth - s_phph)/r - rho g)."""
    st = setups.diapir_sector(9, n_exp=1.0, bc_mode=BC_FREE_SLIP, vr=1.0)
    st.fields["rho"][...] = 1.0
    st.fields["eta"][...] = 1.0
    st.fields["eta_ini"][...] = 1.0
    o = Oracle(st)
    o.step(4000)      # the relative err of the exit test is undefined as V -> 0; run a fixed count
    P = o.f["P"]
    dPdr = (P[1:] - P[:-1]) * st.params._dr
    assert np.allclose(dPdr[:, 1:-1, 1:-1], -1.0, rtol=0, atol=1e-12)
    assert np.abs(o.f["VR"]).max() < 1e-14


def test_pure_shear_boundary_fluxes_cancel():
    """The pure-shear BC prescribes normal velocities whose fluxes through
    the th and phi boundaries cancel (free slip closes the r faces), so a
    divergence-free solution can exist.  The prescribed value on a face is
    read back from the oracle's ghost/inner pair, (ghost + inner)/2, at every
    tangential position, and integrated with the face area element
    (R sin(th_b) dR dphi on a th face, R dR dth on a phi face)."""
    st = setups.diapir_sector(9, nth=11, nph=13, n_exp=1.0, bc_mode=BC_PURE_SHEAR, sh=0.7)
    rng = np.random.default_rng(5)
    for k in ("VR", "Vth", "Vph"):
        st.fields[k][...] = rng.standard_normal(st.fields[k].shape)
    o = Oracle(st)
    o.bc_R(); o.bc_th(); o.bc_ph()
    VR, VT, VP = o.f["VR"], o.f["Vth"], o.f["Vph"]
    R, th, ph = st.R, st.th, st.ph
    def trapz(x):
        w = np.gradient(x)
        w[[0, -1]] *= 0.5
        return w
    wR, wT, wP = trapz(R) * R, trapz(th), trapz(ph)   # trapezoid rule over the boundary nodes
    vr_lo, vr_hi = 0.5 * (VR[0] + VR[1]), 0.5 * (VR[-1] + VR[-2])
    vt_lo, vt_hi = 0.5 * (VT[:, 0] + VT[:, 1]), 0.5 * (VT[:, -1] + VT[:, -2])      # (nr, nph)
    vp_lo, vp_hi = 0.5 * (VP[:, :, 0] + VP[:, :, 1]), 0.5 * (VP[:, :, -1] + VP[:, :, -2])  # (nr, nth)
    assert np.abs(vr_lo).max() < 1e-15 and np.abs(vr_hi).max() < 1e-15
    # a uniform normal velocity on each face (the BC prescribes a constant)
    for v in (vt_lo, vt_hi, vp_lo, vp_hi):
        assert np.ptp(v) <= 1e-14 * np.abs(v).max()
    flux_th = math.sin(th[-1]) * np.einsum("i,k,ik->", wR, wP, vt_hi) - math.sin(th[0]) * np.einsum("i,k,ik->", wR, wP, vt_lo)
    flux_ph = np.einsum("i,j,ij->", wR, wT, vp_hi) - np.einsum("i,j,ij->", wR, wT, vp_lo)
    assert abs(flux_th) > 0.1 and abs(flux_ph) > 0.1        # both faces carry flow
    assert abs(flux_th + flux_ph) <= 1e-13 * abs(flux_ph), (flux_th, flux_ph)


def test_pure_shear_seed_is_divergence_free():
    """The seeded far-field pure shear (setups.pure_shear_velocity) is
    divergence-free in the continuum, so compute_P!'s discrete divV of it is
    O(h^2) and shrinks by about 4x per grid doubling."""
    errs = []
    for n in (9, 17, 33):
        st = setups.diapir_sector(n, n_exp=1.0, bc_mode=BC_PURE_SHEAR, sh=1.0)
        f = st.fields
        for k in ("P", "tau_RR", "tau_thth", "tau_phph", "tau_Rth", "tau_Rph", "tau_thph"):
            f[k][...] = 0.0
        f["etaSM"][...] = 1.0
        f["eta"][...] = 1e300
        o = Oracle(st)
        o.compute_P()
        p = st.params
        kb = (((p.dmp * 2.0) * math.pi) * p.vpdt) / p.lr
        errs.append(np.abs(o.f["P"] / kb).max())
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5, errs
    assert errs[-1] < 1e-3, errs


def _check_err_brute(Vs, Eprev, drop_nan_vph):
    comp = []
    for V, E in zip(Vs, Eprev):
        d = np.abs(V.ravel(order="F") - E.ravel(order="F"))
        a = np.abs(V.ravel(order="F"))
        comp.append(float("nan") if (np.isnan(d).any() or np.isnan(a).any()) else float(d.max() / a.max()))
    cs = comp[:2] if (drop_nan_vph and math.isnan(comp[2])) else comp
    return (float("nan") if any(math.isnan(c) for c in cs) else max(cs)), comp


@pytest.mark.parametrize("case", ["plain", "nan_vph", "nan_vph_drop", "nan_vr"])
def test_check_err_brute_force(case):
    """check_err! + the host reduction (err = max over components of
    max|V - V_prev| / max|V|, NaN-propagating; the plateau variant drops a
    NaN Vph component) against a numpy brute force."""
    st = setups.diapir(6, 7, 8, bc_mode=BC_FREE_SLIP)
    rng = np.random.default_rng(3)
    for k in ("VR", "Vth", "Vph"):
        st.fields[k][...] = rng.standard_normal(st.fields[k].shape) * (1 + 3 * ("VR" == k))
        st.fields[k + "_err"][...] = st.fields[k] + 1e-3 * rng.standard_normal(st.fields[k].shape)
    if case.startswith("nan_vph"):
        st.fields["Vph"][2, 3, 4] = np.nan
    if case == "nan_vr":
        st.fields["VR_err"][1, 1, 1] = np.nan
    st.params.err_drop_nan_vph = int(case == "nan_vph_drop")
    Vs = [st.fields[k].copy() for k in ("VR", "Vth", "Vph")]
    Es = [st.fields[k + "_err"].copy() for k in ("VR", "Vth", "Vph")]
    o = Oracle(st)
    err, comp = o.check_err()
    exp_err, exp_comp = _check_err_brute(Vs, Es, st.params.err_drop_nan_vph)
    for a, b in zip(comp, exp_comp):
        assert (math.isnan(a) and math.isnan(b)) or a == b
    assert (math.isnan(err) and math.isnan(exp_err)) or err == exp_err
    if case == "plain":
        assert err == exp_comp[int(np.argmax(exp_comp))] and err > 0
    # V_err <- V - V_prev
    for k, V, E in zip(("VR", "Vth", "Vph"), Vs, Es):
        assert np.array_equal(o.f[k + "_err"], V - E, equal_nan=True)


def test_initialize_inclusion_brute_force():
    """initialize_inclusion! tests the squared distance against `radius`
    (SphericalStokes_diapir.jl:132): a point at distance d from the centre
    is inside iff d^2 < radius, i.e. d < sqrt(radius).  Brute force over a
    tiny grid with explicit Python loops, plus the quirk itself: radius 0.01
    marks a sphere of radius 0.1."""
    from oracle.binding import initialize_inclusion
    st = setups.diapir_sector(11, nth=9, nph=13)
    R, th, ph, r = st.R, st.th, st.ph, 1.5
    for radius in (0.01, 0.05, 0.3):
        A = np.full((11, 9, 13), 7.0, order="F")
        initialize_inclusion(A, R, th, ph, r, radius, -2.5)
        exp = np.full_like(A, 7.0)
        for i in range(11):
            for j in range(9):
                for k in range(13):
                    d = math.sqrt((R[i] - r) ** 2 + ((th[j] - math.pi / 2) * r) ** 2 + (ph[k] * r) ** 2)
                    if d < math.sqrt(radius):
                        exp[i, j, k] = -2.5
        assert np.array_equal(A, exp), radius
        # the numpy setup does the same
        B = np.full_like(A, 7.0)
        setups._inclusion(B, R, th, ph, r, radius, -2.5)
        assert np.array_equal(B, exp), radius
    A = np.zeros((11, 9, 13), order="F")
    initialize_inclusion(A, R, th, ph, r, 0.01, 1.0)
    inside = np.argwhere(A == 1.0)
    dist = [math.sqrt((R[i] - r) ** 2 + ((th[j] - math.pi / 2) * r) ** 2 + (ph[k] * r) ** 2) for i, j, k in inside]
    assert len(dist) > 0 and max(dist) < 0.1


def test_initialize_velocity_brute_force():
    """initialize_velocity!(Vph, th, sh, r) sets Vph = (th - pi/2) r sh on
    every (nr, nth, nph+1) point (SphericalStokes_diapir.jl:139-146): a
    simple shear, linear in th, zero on the equator."""
    from oracle.binding import initialize_velocity
    st = setups.diapir_sector(7, nth=9, nph=5)
    th = st.th
    V = np.full((7, 9, 6), 3.0, order="F")
    initialize_velocity(V, th, 0.8, 1.5)
    for j in range(9):
        assert np.all(V[:, j, :] == (th[j] - math.pi / 2) * 1.5 * 0.8)
    assert V[:, 4, :].max() == 0.0 and V[:, 4, :].min() == 0.0        # th = pi/2
    slope = np.diff(V[0, :, 0]) / np.diff(th)
    assert np.allclose(slope, 1.5 * 0.8, rtol=1e-12)
    # the numpy setup's seed for simple shear is the same field
    W = np.zeros_like(V)
    setups._velocity_shear(W, th, 1.5, 0.8)
    assert np.array_equal(W, V)
