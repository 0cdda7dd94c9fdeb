"""GPU: the command-line driver (sphericalstokes_amd/sspt_run, a C++ caller
of include/sspt.h) runs the reference loop on the setups of setups.py; its
result equals the Python-driven solve of the same setup through ctypes bit
for bit (iteration count, iteration of the break, err)."""
import json
import subprocess
from pathlib import Path

import pytest

from sphericalstokes_amd import setups
from sphericalstokes_amd.solver import Solver

pytestmark = pytest.mark.gpu
CLI = Path(__file__).resolve().parent.parent / "sphericalstokes_amd" / "sspt_run"


def _cli(*args):
    r = subprocess.run([str(CLI), *args, "--json"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode in (0, 3), r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1]), r.stdout


def _py(st, maxiter, nout, eps):
    with Solver(st.nr, st.nth, st.nph) as s:
        s.load_setup(st)
        return s.pt_solve(st.params, maxiter, nout, eps)


def test_cli_sector_solve_equals_python():
    got, out = _cli("--n", "23", "--maxiter", "200000", "--nout", "100", "--eps", "1e-6")
    ref = _py(setups.diapir_sector(23), 200_000, 100, 1e-6)
    assert got["converged"] and (got["ittot"], got["iter_end"]) == (ref["ittot"], ref["iter_end"])
    assert got["err"] == ref["err"]
    assert "Total iters = %d" % ref["ittot"] in out


def test_cli_reference_setup_equals_python():
    """The script's own setup (r = 1000, lr = 6, n = 5, simple shear) at 15^3."""
    got, _ = _cli("--setup", "reference", "--n", "15", "--maxiter", "3000", "--nout", "100", "--eps", "1e-12")
    ref = _py(setups.diapir(15, 15, 15), 3000, 100, 1e-12)
    assert (got["ittot"], got["iter_end"]) == (ref["ittot"], ref["iter_end"])
    assert got["err"] == ref["err"]


def test_cli_timed_reports_t_eff():
    got, _ = _cli("--n", "63", "--timed", "20")
    assert got["iterations"] == 20 and got["iter_ms"] > 0 and got["T_eff_GBps"] > 0


@pytest.mark.parametrize("mode", ["--gradient", "--mixed-adjoint"])
def test_cli_gradient_equals_python(mode):
    """--gradient / --mixed-adjoint after the sector solve at 15^3: J and the
    adjoint's iteration counts and err equal the Python-driven solves of the
    same setup (ss_adjoint_solve / ss_adjoint_solve_mixed through ctypes)."""
    got, out = _cli("--n", "15", "--maxiter", "400000", "--nout", "100", "--eps", "1e-12", mode, "--adj-tol", "1e-12")
    st = setups.diapir_sector(15)
    with Solver(15, 15, 15) as s:
        s.load_setup(st)
        f = s.pt_solve(st.params, 400_000, 100, 1e-12)
        if mode == "--gradient":
            a = s.adjoint_solve(st.params, 1.5, 0.1 ** 2, maxiter=800_000, ncheck=100, tol=1e-12)
            assert got["adjoint_iters"] == a["iters"]
        else:
            a = s.adjoint_solve_mixed(st.params, 1.5, 0.1 ** 2, max_outer=20, max_inner=200_000, nout_inner=100,
                                      tol_inner=1e-4, tol=1e-12)
            assert (got["adjoint_outer"], got["adjoint_sweeps"]) == (a["outer"], a["inner_total"])
    assert got["ittot"] == f["ittot"] and got["adjoint_converged"] and a["converged"]
    assert got["J"] == a["J"] and got["adjoint_err"] == a["err"]
    assert "J = " in out
