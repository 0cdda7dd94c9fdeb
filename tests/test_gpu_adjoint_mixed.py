"""GPU: the adjoint solve with FP32 sweeps inside an FP64 defect correction
(ss_adjoint_solve_mixed) against the FP64 adjoint solve (ss_adjoint_solve,
itself checked against the CPU adjoint and central differences) on the same
converged states.

Both stop on a relative increment below tol (the FP64 solve on lam, the
mixed one on the BC-folded lam, whose fixed point is the same); at tol 1e-13
both gradients lie within about 1e-10 of the fixed point's, so they agree to
rel. 1e-8 of each gradient's max.  The FP32 sweeps' rounding (state, tables
and correction rounded to FP32, ~1e-7 relative per sweep) only slows the
outer convergence; it does not enter the FP64 residual that stops it.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from sphericalstokes_amd import _lib, setups
from sphericalstokes_amd._lib import SS_ERR_ARG
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_PURE_SHEAR, BC_SHEAR

pytestmark = pytest.mark.gpu

R0, RAD = 1.5, 0.1 ** 2

CASES = {
    "linear_free_slip": dict(n_exp=1.0, bc_mode=BC_FREE_SLIP),
    "powerlaw_simple_shear": dict(n_exp=3.0, bc_mode=BC_SHEAR, sh=0.7),
    "powerlaw_relax_cutoffs": dict(n_exp=3.0, bc_mode=BC_FREE_SLIP, relax=0.5, eta_min=0.3, eta_max=3.0),
    "powerlaw_pure_shear": dict(n_exp=3.0, bc_mode=BC_PURE_SHEAR, sh=1.0),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _solver(st):
    from sphericalstokes_amd.solver import Solver
    s = Solver(st.nr, st.nth, st.nph)
    s.load_setup(st)
    return s


def _compare(a, b, what):
    for k in ("grad_rho", "grad_log10_eta_ini"):
        scale = np.abs(a[k]).max()
        err = np.abs(a[k] - b[k]).max() / scale
        say(f"  {what} {k}: rel diff {err:.2e}")
        assert err <= 1e-8, (what, k, err)
    assert abs(a["J"] - b["J"]) <= 1e-14 * abs(a["J"])


@pytest.mark.parametrize("path", ["marching", "per_family"])
@pytest.mark.parametrize("case", list(CASES))
def test_mixed_adjoint_matches_fp64_adjoint(case, path, monkeypatch):
    """Every BC mode, relaxation and cutoffs at 15^3; the FP64 outer step
    through either adjoint path (the FP32 sweeps always use the r-marching
    kernel)."""
    monkeypatch.setenv("SSPT_ADJ_FUSED", "1" if path == "marching" else "0")
    st = setups.diapir_sector(15, vr=10.0, **CASES[case])
    say("mixed adjoint", case, path)
    with _solver(st) as s:
        f = s.pt_solve(st.params, 2_000_000, 100, 1e-14)
        assert f["status"] == "SS_OK", f
        ref = s.adjoint_solve(st.params, R0, RAD, maxiter=2_000_000, ncheck=100, tol=1e-13)
        got = s.adjoint_solve_mixed(st.params, R0, RAD, tol=1e-13)
    say("  fp64", {k: v for k, v in ref.items() if not k.startswith("grad")})
    say("  mixed", {k: v for k, v in got.items() if not k.startswith("grad")})
    assert ref["converged"] and got["converged"]
    _compare(got, ref, case)


def test_mixed_adjoint_63():
    """The 63^3 power-law rising sphere of the central-difference check
    (free slip, inclusion viscosity 0.1, cutoffs): mixed and FP64 gradients
    agree; the FP32 sweeps take about as many sweeps as the FP64 iteration."""
    st = setups.diapir_sector(63, n_exp=3.0, bc_mode=BC_FREE_SLIP, vr=10.0, eta_min=1e-4, eta_max=1e4)
    with _solver(st) as s:
        f = s.pt_solve(st.params, 2_000_000, 500, 1e-14)
        assert f["status"] == "SS_OK", f
        ref = s.adjoint_solve(st.params, R0, RAD, maxiter=2_000_000, ncheck=100, tol=1e-13)
        got = s.adjoint_solve_mixed(st.params, R0, RAD, tol=1e-13)
    say("63^3 fp64", {k: v for k, v in ref.items() if not k.startswith("grad")})
    say("63^3 mixed", {k: v for k, v in got.items() if not k.startswith("grad")})
    assert ref["converged"] and got["converged"]
    assert got["inner_total"] <= 2 * ref["iters"]
    _compare(got, ref, "63^3")


def test_mixed_adjoint_option_errors():
    st = setups.diapir_sector(9, n_exp=1.0, bc_mode=BC_FREE_SLIP)
    L = _lib.lib()
    with _solver(st) as s:
        sp = _lib.ss_params.from_params(st.params)
        res = _lib.ss_adjoint_mixed_result()
        for bad in ((0, 10, 10, 1e-4, 1e-8), (_lib.MIXED_MAX_OUTER + 1, 10, 10, 1e-4, 1e-8),
                    (4, 0, 10, 1e-4, 1e-8), (4, 10, 0, 1e-4, 1e-8)):
            o = _lib.ss_mixed_opts(*bad)
            assert L.ss_adjoint_solve_mixed(s._h, C.byref(sp), R0, RAD, C.byref(o), C.byref(res)) == SS_ERR_ARG
        o = _lib.ss_mixed_opts(4, 10, 10, 1e-4, 1e-8)
        assert L.ss_adjoint_solve_mixed(s._h, None, R0, RAD, C.byref(o), C.byref(res)) == SS_ERR_ARG
        assert L.ss_adjoint_solve_mixed(s._h, C.byref(sp), R0, RAD, C.byref(o), None) == SS_ERR_ARG


def test_autograd_function_mixed_adjoint():
    """The torch.autograd.Function with the mixed-precision backward
    (SolveOptions.adj_mixed) gives the FP64 backward's gradients."""
    import torch
    from sphericalstokes_amd.adjoint import SolveOptions, ascent_velocity

    st = setups.diapir_sector(13, n_exp=3.0, bc_mode=BC_FREE_SLIP, vr=10.0)
    rho0 = np.ascontiguousarray(st.fields["rho"])
    le0 = np.log10(np.ascontiguousarray(st.fields["eta_ini"]))
    grads = {}
    for mixed in (False, True):
        opts = SolveOptions(eps=1e-13, nout=500, adj_mixed=mixed, adj_ncheck=100)
        rho = torch.tensor(rho0, requires_grad=True)
        le = torch.tensor(le0, requires_grad=True)
        ascent_velocity(rho, le, st, opts).backward()
        say("autograd, mixed" if mixed else "autograd, fp64", opts.info.get("adjoint"))
        grads[mixed] = (rho.grad.numpy(), le.grad.numpy())
    for a, b in zip(grads[True], grads[False]):
        assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()


def test_mixed_adjoint_nonfinite_state_is_an_error():
    """A NaN in the converged state (here one viscosity) reaches the adjoint
    fields in the first FP64 step; the solve stops there with
    SS_ERR_NONFINITE instead of sweeping on."""
    from sphericalstokes_amd._lib import SS_ERR_NONFINITE, SSError
    st = setups.diapir_sector(15, n_exp=1.0, bc_mode=BC_FREE_SLIP, vr=10.0)
    with _solver(st) as s:
        f = s.pt_solve(st.params, 2_000_000, 100, 1e-10)
        assert f["status"] == "SS_OK", f
        eta = s.download("eta")
        eta[7, 7, 7] = np.nan
        s.upload("eta", eta)
        with pytest.raises(SSError) as e:
            s.adjoint_solve_mixed(st.params, R0, RAD, tol=1e-13)
        assert e.value.code == SS_ERR_NONFINITE
