"""GPU: the etaSM spread option (ss_params.etasm_spread = m) against the
oracle's maxloc! + copy_BC! + orc_spread_etasm -- fused (pre-passes +
F_SMBUF) and per-kernel paths bit for bit at n_exp = 1, within 1e-10 with
the power law; the log-normal mixed-precision configuration's field at
127^3 iterating stably with it (the reference rule blows up there); and the
FP32-refined solve with it against an oracle-converged state."""
import sys

import numpy as np
import pytest

from helpers import compare_fields, rel_l2
from oracle.binding import Oracle
from sphericalstokes_amd import setups
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_SHEAR, FIELDS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _solver(st, fields=None):
    from sphericalstokes_amd.solver import Solver
    s = Solver(st.nr, st.nth, st.nph)
    s.set_grid(st.R, st.th, st.ph)
    f = fields if fields is not None else st.fields
    for n in FIELDS:
        s.upload(n, f[n])
    return s


def _case(shape, seed, spread=3, **kw):
    st = setups.diapir_sector(shape[0], nth=shape[1], nph=shape[2], **kw)
    setups.random_state(st, seed=seed)
    st.params.etasm_spread = spread
    return st


@pytest.mark.parametrize("shape", [(9, 7, 70), (40, 33, 5), (17, 70, 130)])
def test_spread_pass_every_m_bitwise(shape):
    """ss_spread_etasm (the r pass, then the th and phi passes through LDS
    tiles) against orc_spread_etasm for every m, bit for bit, on shapes with
    dimensions below 2m (clamping at both ends of a line) and above one
    tile (64 columns, 32 rows)."""
    from oracle.binding import _ptr, lib
    st = setups.diapir_sector(shape[0], nth=shape[1], nph=shape[2], n_exp=1.0)
    rng = np.random.default_rng(sum(shape))
    with _solver(st) as s:
        for m in range(1, 9):
            A = np.asfortranarray(np.exp(3 * rng.standard_normal(shape)))
            s.upload("etaSM", A)
            s.spread_etasm(m)
            got = s.download("etaSM")
            exp = A.copy(order="F")
            tmp = np.zeros_like(exp)
            lib().orc_spread_etasm(_ptr(exp), _ptr(tmp), *shape, m)
            say("spread", shape, "m", m, "differing", int(np.count_nonzero(got != exp)))
            assert np.array_equal(got, exp), m


@pytest.mark.parametrize("fused", [True, False])
def test_spread_etasm_bitwise_linear(fused):
    st = _case((21, 40, 70), 11, n_exp=1.0, bc_mode=BC_FREE_SLIP)
    say("spread etaSM, linear, fused" if fused else "spread etaSM, linear, per-kernel")
    o = Oracle(st)
    o.step(6)
    with _solver(st) as s:
        if fused:
            s.pt_iterate(st.params, 6)
        else:
            for _ in range(6):
                s.pt_step_unfused(st.params)
        got = s.download_all()
    res, bad = compare_fields(got, o.f)
    assert not bad, {k: res[k] for k in bad}


def test_spread_etasm_powerlaw_tolerance():
    st = _case((21, 40, 70), 12, spread=5, n_exp=3.0, bc_mode=BC_SHEAR, sh=1.0, relax=0.5, eta_min=1e-4,
               eta_max=1e4)
    o = Oracle(st)
    o.step(10)
    with _solver(st) as s:
        s.pt_iterate(st.params, 10)
        got = s.download_all()
    for n in FIELDS:
        assert rel_l2(got[n], o.f[n]) <= 1e-10, n


def test_lognormal_127_spread_iterates_stably():
    """The mixed-precision configuration's field (log-normal matrix, inclusion
    1e-3; jumps of up to 6.9e3 at the inclusion's edge at this size, where
    the reference step rule blows up) at 127^3 with n_exp = 3: the first 20
    fused iterations match the oracle within 1e-10, and 4000 more stay
    finite with the reference's err falling."""
    st = setups.diapir_lognormal(127)
    assert st.params.etasm_spread == 5
    o = Oracle(st)
    o.step(20)
    with _solver(st) as s:
        s.pt_iterate(st.params, 20)
        for n in FIELDS:
            assert rel_l2(s.download(n), o.f[n]) <= 1e-10, n
        errs = []
        for k in range(4):
            r = s.pt_solve(st.params, 1000, 1000, 0.0, raise_on=())
            errs.append(r["err"])
            say(f"  {20 + 1000 * (k + 1)} iterations: err {r['err']:.3e}, "
                f"max|VR| {np.abs(s.download('VR')).max():.3e}")
            assert np.isfinite(r["err"])
    assert errs[-1] < errs[0]


def test_mixed_spread_reaches_oracle_converged_state():
    """FP32-refined solve (spread etaSM, n_exp = 3 log-normal field) at 15^3
    against the oracle's own converged state: V and interior zero-mean P
    within 1e-8."""
    st = setups.diapir_lognormal(15)
    o = Oracle(st)
    r = o.solve(400000, 500, 1e-15)
    say("oracle", r)
    with _solver(st) as s:
        rm = s.pt_solve_mixed(st.params, max_outer=60, max_inner=100000, nout_inner=200, tol_inner=1e-5, tol=1e-12)
        say("mixed", {k: v for k, v in rm.items() if k not in ("r_hist", "inner_hist", "err_inner_hist")})
        got = s.download_all()
    assert rm["converged"]
    for n in ("VR", "Vth", "Vph"):
        e = rel_l2(got[n], o.f[n])
        say(n, "rel error vs oracle", e)
        assert e < 1e-8, (n, e)
    dp = (got["P"] - o.f["P"])[1:-1, 1:-1, 1:-1]
    pc = o.f["P"][1:-1, 1:-1, 1:-1]
    e = float(np.linalg.norm(dp - dp.mean()) / np.linalg.norm(pc - pc.mean()))
    say("P (interior, zero-mean) rel error vs oracle", e)
    assert e < 1e-8
