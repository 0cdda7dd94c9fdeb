"""GPU: the HIP adjoint (gather form, sspt_adjoint.hip: the r-marching kernel
and the per-family kernels) against the CPU reference adjoint (scatter form,
oracle/sspt_adjoint.c, itself pinned by tests/test_adjoint_oracle.py), and
its gradients against central differences of GPU forward solves.

Floating point: the two adjoints sum the same terms in different orders, so
they agree to rounding (rel. 1e-12 of each field's max); the gradients of
the converged adjoint to rel. 1e-8; finite differences to rel. 1e-5 (the
north_star bound for the adjoint gradient).
"""
import sys

import numpy as np
import pytest

from oracle.binding import ADJ_FIELDS as ORC_ADJ, Oracle
from sphericalstokes_amd import setups
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_PURE_SHEAR, BC_SHEAR, FIELDS

pytestmark = pytest.mark.gpu

R0, RAD = 1.5, 0.1 ** 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def Solver():
    from sphericalstokes_amd.solver import Solver as S
    return S


def _load(S, st, fields):
    s = S(st.nr, st.nth, st.nph)
    s.set_grid(st.R, st.th, st.ph)
    for n in FIELDS:
        s.upload(n, fields[n])
    return s


STEP_CASES = {
    "linear_free_slip": dict(n_exp=1.0, bc_mode=BC_FREE_SLIP),
    "powerlaw_simple_shear": dict(n_exp=3.0, bc_mode=BC_SHEAR, sh=0.7),
    "powerlaw_relax_cutoffs": dict(n_exp=3.0, bc_mode=BC_FREE_SLIP, relax=0.5, eta_min=0.3, eta_max=3.0),
    "powerlaw_pure_shear": dict(n_exp=3.0, bc_mode=BC_PURE_SHEAR, sh=1.0),
}


@pytest.mark.parametrize("path", ["marching", "per_family"])
@pytest.mark.parametrize("case", list(STEP_CASES))
def test_adjoint_step_matches_reference(Solver, case, path, monkeypatch):
    """One application of T^T (and of (dPhi/dm)^T) on random adjoint fields at
    a random state (T^T is defined at any state with its step sizes held),
    by the r-marching kernel and by the per-family kernels."""
    monkeypatch.setenv("SSPT_ADJ_FUSED", "1" if path == "marching" else "0")
    st = setups.diapir_sector(11, nth=13, nph=70, **STEP_CASES[case])
    setups.random_state(st, seed=11)
    o = Oracle(st)
    o.step(2)                                   # etaSM, DT_R, tII consistent with the state
    rng = np.random.default_rng(3)
    lam = {n: rng.standard_normal(o.f[n].shape) for n in ORC_ADJ}
    if st.params.relax == 1.0:
        lam["eta_PL"][:] = 0.0
    ref, g_rho, g_ini = o.adj_step(lam, grads=True)
    say("adjoint step", case, path)
    with _load(Solver, st, o.f) as s:
        s.adjoint_prepare(st.params)
        s.adjoint_zero()
        for n in ORC_ADJ:
            s.adjoint_upload(n, lam[n])
        s.adjoint_step(st.params, gradients=True)
        got = {n: s.adjoint_download(n) for n in ORC_ADJ}
        got_rho, got_ini = s.adjoint_download("grad_rho"), s.adjoint_download("grad_eta_ini")
    for n in ORC_ADJ:
        scale = max(np.abs(ref[n]).max(), 1e-300)
        err = np.abs(got[n] - ref[n]).max() / scale
        assert err <= 1e-12, (n, err)
    for a, b, n in ((got_rho, g_rho, "grad_rho"), (got_ini, g_ini, "grad_eta_ini")):
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300), n


@pytest.mark.parametrize("path", ["marching", "per_family"])
def test_adjoint_solve_matches_reference(Solver, path, monkeypatch):
    monkeypatch.setenv("SSPT_ADJ_FUSED", "1" if path == "marching" else "0")
    st = setups.diapir_sector(9, n_exp=3.0, bc_mode=BC_FREE_SLIP)
    o = Oracle(st)
    o.step(30000)
    it, _, g_rho, g_ini = o.adj_solve(R0, RAD, 200000, 100, 1e-14)
    assert it > 0
    say("reference adjoint iterations", it)
    with _load(Solver, st, o.f) as s:
        r = s.adjoint_solve(st.params, R0, RAD, maxiter=200000, ncheck=100, tol=1e-14)
    say("gpu adjoint", {k: v for k, v in r.items() if not k.startswith("grad")})
    assert r["converged"]
    J_ref, _ = o.objective(R0, RAD)
    assert abs(r["J"] - J_ref) <= 1e-13 * abs(J_ref)
    assert np.abs(r["grad_rho"] - g_rho).max() <= 1e-8 * np.abs(g_rho).max()
    g_log = g_ini * o.f["eta_ini"] * np.log(10.0)
    assert np.abs(r["grad_log10_eta_ini"] - g_log).max() <= 1e-8 * np.abs(g_log).max()


@pytest.mark.parametrize("n_exp", [1.0, 3.0])
def test_adjoint_gradient_central_differences(Solver, n_exp):
    """Gradient of J at 17^3 on the GPU against central differences (h = 1e-6)
    of GPU forward solves along 4 Gaussian directions per parameter from
    seed 1 (16 per parameter run on the CPU reference adjoint at 7^3,
    tests/test_adjoint_oracle.py; this test stays under a minute)."""
    st = setups.diapir_sector(17, n_exp=n_exp, bc_mode=BC_FREE_SLIP, vr=10.0)
    p = st.params
    iters = 30000
    with _load(Solver, st, st.fields) as s:
        s.pt_iterate(p, iters)
        base = s.download_all()
        r = s.adjoint_solve(p, R0, RAD, maxiter=400000, ncheck=200, tol=1e-13)
    say("adjoint", {k: v for k, v in r.items() if not k.startswith("grad")})
    assert r["converged"]
    rng = np.random.default_rng(1)
    h = 1e-6
    worst = 0.0
    for what in ("rho", "log10_eta_ini"):
        for d in range(4):
            dm = rng.standard_normal(base["rho"].shape)
            J = []
            for sgn in (+1, -1):
                f = {k: v.copy(order="F") for k, v in base.items()}
                if what == "rho":
                    f["rho"] = np.asfortranarray(f["rho"] + sgn * h * dm)
                else:
                    f["eta_ini"] = np.asfortranarray(f["eta_ini"] * 10.0 ** (sgn * h * dm))
                with _load(Solver, st, f) as s2:
                    s2.pt_iterate(p, iters)
                    s2.adjoint_prepare(p)
                    J.append(s2.adjoint_objective(R0, RAD))
            fd = (J[0] - J[1]) / (2 * h)
            ad = float(np.sum((r["grad_rho"] if what == "rho" else r["grad_log10_eta_ini"]) * dm))
            worst = max(worst, abs(fd - ad) / abs(fd))
            assert abs(fd - ad) <= 1e-5 * abs(fd), (what, d, fd, ad)
    say("worst relative difference over 8 directions", worst)


def test_autograd_function_matches_central_differences():
    """J.backward() of the torch.autograd.Function (forward PT solve, adjoint
    backward) against central differences of the same forward."""
    import torch
    from sphericalstokes_amd.adjoint import SolveOptions, ascent_velocity

    st = setups.diapir_sector(13, n_exp=3.0, bc_mode=BC_FREE_SLIP, vr=10.0)
    rho0 = np.ascontiguousarray(st.fields["rho"])
    le0 = np.log10(np.ascontiguousarray(st.fields["eta_ini"]))
    opts = SolveOptions(eps=1e-13, nout=500)
    rho = torch.tensor(rho0, requires_grad=True)
    le = torch.tensor(le0, requires_grad=True)
    J = ascent_velocity(rho, le, st, opts)
    J.backward()
    say("autograd", float(J), opts.info)
    rng = np.random.default_rng(7)
    h = 1e-6
    for name, grad, x0 in (("rho", rho.grad.numpy(), rho0), ("log10_eta_ini", le.grad.numpy(), le0)):
        dm = rng.standard_normal(x0.shape)
        Js = []
        for sgn in (+1, -1):
            args = [torch.tensor(rho0), torch.tensor(le0)]
            k = 0 if name == "rho" else 1
            args[k] = torch.tensor(x0 + sgn * h * dm)
            with torch.no_grad():
                Js.append(float(ascent_velocity(args[0], args[1], st, SolveOptions(eps=1e-13, nout=500))))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float(np.sum(grad * dm))
        say(name, "fd", fd, "autograd", ad)
        assert abs(fd - ad) <= 1e-5 * abs(fd), (name, fd, ad)


def test_adjoint_gradient_63_sixteen_directions(Solver):
    """The adjoint check at 63^3 with 16 directions per parameter: the
    power-law (n_exp = 3) rising sphere, free slip, inclusion viscosity 0.1,
    J = mean VR in the inclusion; the gradient with respect to rho and
    log10(eta_ini) against central differences (h = 1e-6) along 16 Gaussian
    directions per parameter from seed 1, within 1e-5.  Each perturbed solve
    restarts from the converged state and runs a fixed 40 000 iterations: the
    part of J's error that does not cancel between the two signs is about
    h * |dx/dm| * rho^K (rho the PT contraction), and this setup contracts
    by 1e-10 in about 30 000 iterations.  With the adjoint configuration's
    pure shear and relaxation the iteration contracts 5-8 times more slowly
    (150 000-250 000 iterations to an err of 1e-13): there 40 000 and
    110 000 iterations per perturbed solve left 1.5e-5 and 1.0e-5, so its
    16-direction check does not fit the test budget; its gradients are
    checked at small size against the CPU adjoint (test_adjoint_solve_*,
    test_adjoint_step_*).  The 64 solves run four at a time on separate
    contexts and streams."""
    from concurrent.futures import ThreadPoolExecutor
    st = setups.diapir_sector(63, n_exp=3.0, bc_mode=BC_FREE_SLIP, vr=10.0, eta_min=1e-4, eta_max=1e4)
    p = st.params
    with _load(Solver, st, st.fields) as s:
        f = s.pt_solve(p, 2_000_000, 500, 1e-14)
        say("forward", f)
        assert f["status"] == "SS_OK"
        base = s.download_all()
        r = s.adjoint_solve(p, R0, RAD, maxiter=2_000_000, ncheck=500, tol=1e-14)
    say("adjoint", {k: v for k, v in r.items() if not k.startswith("grad")})
    assert r["converged"]
    rng = np.random.default_rng(1)
    dirs = [(what, rng.standard_normal(base["rho"].shape)) for what in ("rho", "log10_eta_ini") for _ in range(16)]
    h, iters = 1e-6, 40_000

    def J_of(what, dm, sgn):
        g = {k: v.copy(order="F") for k, v in base.items()}
        if what == "rho":
            g["rho"] = np.asfortranarray(g["rho"] + sgn * h * dm)
        else:
            g["eta_ini"] = np.asfortranarray(g["eta_ini"] * 10.0 ** (sgn * h * dm))
        with _load(Solver, st, g) as s2:
            s2.pt_iterate(p, iters)
            s2.adjoint_prepare(p)
            return s2.adjoint_objective(R0, RAD)

    with ThreadPoolExecutor(4) as ex:
        futs = [(what, dm, ex.submit(J_of, what, dm, +1), ex.submit(J_of, what, dm, -1)) for what, dm in dirs]
        worst = 0.0
        for q, (what, dm, fp, fm) in enumerate(futs):
            fd = (fp.result() - fm.result()) / (2 * h)
            ad = float(np.sum((r["grad_rho"] if what == "rho" else r["grad_log10_eta_ini"]) * dm))
            rel = abs(fd - ad) / abs(fd)
            worst = max(worst, rel)
            say(f"  {what:14s} direction {q % 16:2d}: fd {fd:+.10e} adjoint {ad:+.10e} rel {rel:.2e}")
            assert rel <= 1e-5, (what, q, fd, ad)
    say("worst relative difference over 32 directions", worst)
