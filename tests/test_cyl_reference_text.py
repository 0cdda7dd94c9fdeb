"""The cylindrical twin's oracle (oracle/sspt_cyl_oracle.c) and the package's
cylindrical setup against vectors evaluated from the text of
scripts_diapir/CylindricalStokes_diapir.jl (tests/golden/reference_text_vectors.npz,
cases cyl_*).  That script's kernels are ParallelStencil @parallel kernels;
the evaluator expands their FiniteDifferences3D macros as
tests/golden/jl_subset.py:FD3D_TEXT restates them (the package is not part
of the reference, so the term order of its averages is pinned by that
restatement, not by the reference).  Bit for bit.
"""
import numpy as np
import pytest

from oracle.binding import CYL, CylOracle
from sphericalstokes_amd.cylindrical import CYL_FIELDS, CylParams, CylSetup, cyl_diapir
from test_reference_text import FIX, META, bits_differ

CY = META["cyl"]


def cyl_text_setup():
    nr, nt, nz = CY["grid"]
    sc = CY["scalars"]
    p = CylParams(dr=sc["dr"], dth=sc["dth"], dz=sc["dz"], dmp=sc["dmp"], vpdt=sc["vpdt"], lr=sc["lr"],
                  tau_C=sc["tau_C"], n_exp=sc["n_exp"], n_exp_PL=sc["n_exp"], relax=sc["relax"], g=sc["g"],
                  sh=sc["sh"], r=sc["r"])
    f = {n: np.array(FIX[f"cyl_in__state__{n}"], order="F", copy=True) for n in CYL_FIELDS}
    return CylSetup(nr, nt, nz, FIX["cyl__R"].copy(), FIX["cyl__th"].copy(), FIX["cyl__z"].copy(), p, f)


def expect(o, case):
    keys = [k for k in FIX.files if k.startswith(case + "__out__")]
    assert keys, case
    bad = {}
    for k in keys:
        n = k.split("__out__")[1]
        d = bits_differ(o.f[n], FIX[k])
        if d:
            bad[n] = d
    assert not bad, f"{case}: values differing from the reference text per field: {bad}"


def cyl_kernels(p):
    return {
        "timesteps": lambda o: o.timesteps(p.dmp, p.vpdt, p.lr),
        "timesteps2": lambda o: o.timesteps2(p.vpdt),
        "compute_P": lambda o: o.compute_P(p.dr, p.dth, p.dz),
        "compute_sigma": lambda o: o.compute_sigma(),
        "compute_TII": lambda o: o.compute_TII(),
        "power_law": lambda o: o.power_law(p.tau_C, p.n_exp, p.relax),
        "compute_dV": lambda o: o.compute_dV(p.dr, p.dth, p.dz),
        "compute_V": lambda o: o.compute_V(),
        "err_rth": lambda o: o.err_rth(),
        "maxloc": lambda o: o.maxloc(),
        "copy_bc_r_VZ": lambda o: o.bc("copy_bc_r", "VZ"),
        "copy_bc_th_etaSM": lambda o: o.bc("copy_bc_th", "etaSM"),
        "copy_bc_z_tII": lambda o: o.bc("copy_bc_z", "tII"),
        "bc_r_0_VR": lambda o: o.bc("bc_r_0", "VR"),
        "bc_th_fact_VZ": lambda o: o.bc("bc_th_fact", "VZ", p.sh, p.r),
        "bc_th_fact_VR": lambda o: o.bc("bc_th_fact", "VR", 0.0, p.r),
        "bc_th_0_Vth": lambda o: o.bc("bc_th_0", "Vth"),
        "bc_z_0_VR": lambda o: o.bc("bc_z_0", "VR"),
        "bc_z_lin_VZ": lambda o: o.bc("bc_z_lin", "VZ", p.sh, p.r),
    }


CYL_KERNEL_NAMES = sorted(cyl_kernels(CylParams(1, 1, 1, 1, 1, 1)))


@pytest.mark.parametrize("kernel", CYL_KERNEL_NAMES)
def test_cyl_oracle_kernel_equals_reference_text(kernel):
    st = cyl_text_setup()
    o = CylOracle(st)
    cyl_kernels(st.params)[kernel](o)
    expect(o, "cyl_" + kernel)


def test_cyl_three_iterations_equal_reference_loop_text():
    """Three iterations of the loop body (:331-366) with n_exp = 3, err_rth!
    after the third, and the exit test's error of the third iteration."""
    st = cyl_text_setup()
    o = CylOracle(st)
    o.step(st.params, 3)
    o.err_rth()
    expect(o, "cyl_loop")
    err, comp = o.exit_error()
    assert np.array_equal(np.array(comp), FIX["cyl_loop__err"]) and err == float(np.max(FIX["cyl_loop__err"]))


def test_cyl_setup_equals_reference_text():
    """initialize_inclusion! ((R-r)^2 + (th r)^2 + Z^2 < radius) and
    initialize_velocity! (VZ = th r sh on the th and z boundary faces)
    against cylindrical.cyl_diapir's numpy setup."""
    R, th, z = FIX["cyl_init__R"], FIX["cyl_init__th"], FIX["cyl_init__z"]
    n = len(R)
    s = cyl_diapir(n, r=1000.0, lr=6.0, radius=1.0, vr=1e2)
    assert np.array_equal(s.R, R) and np.array_equal(s.th, th) and np.array_equal(s.z, z)
    assert np.array_equal(s.fields["eta"] != 1.0, FIX["cyl_init__out__inclusion"] != 1.0)
    assert bits_differ(s.fields["VZ"], FIX["cyl_init__out__velocity"]) == 0


def test_cyl_oracle_solve_switches_exponent_and_converges():
    """The loop on the script's setup at 15^3: linear until the exit-test
    error falls below 1e-5, then n_exp_PL = 5, and err < eps at a multiple
    of nout."""
    s = cyl_diapir(15)
    o = CylOracle(s)
    r = o.solve(s.params, 100_000, 100, 5e-7)
    assert r["converged"] and r["iters"] % 100 == 0 and r["err"] < 5e-7
    assert np.all(np.isfinite(o.f["VZ"]))
    # the power law was active: eta departs from its initial values
    assert np.any(o.f["eta"] != s.fields["eta"])
