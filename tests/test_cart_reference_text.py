"""The Cartesian twin's oracle (oracle/sspt_cart_oracle.c) and the package's
Cartesian setup against golden vectors evaluated from the text of
scripts_diapir/CartesianStokes_diapir.jl (tests/golden/reference_text_vectors.npz,
cases cart_*; made by tests/golden/make_reference_text_fixture.py).  Bit for
bit: the vectors and the oracle use the same C library for pow, exp, log.
"""
import numpy as np
import pytest

from oracle.binding import CartOracle
from sphericalstokes_amd.cartesian import CART_FIELDS, CartParams, CartSetup, cart_diapir
from test_reference_text import FIX, META, bits_differ

CM = META["cart"]


def cart_text_setup():
    nx, ny, nz = CM["grid"]
    sc = CM["scalars"]
    p = CartParams(_dx=sc["_dx"], _dy=sc["_dy"], _dz=sc["_dz"], dmp=sc["dmp"], vpdt=sc["vpdt"], lx=sc["lx"],
                   tau_C=sc["tau_C"], n_exp=sc["n_exp"], relax=sc["relax"], g=sc["g"], sh=sc["sh"])
    f = {n: np.array(FIX[f"cart_in__state__{n}"], order="F", copy=True) for n in CART_FIELDS}
    return CartSetup(nx, ny, nz, FIX["cart__x"].copy(), FIX["cart__y"].copy(), FIX["cart__z"].copy(), p, f)


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


def _kernels(p):
    return {
        "timesteps": lambda o: o.timesteps(p.vpdt, p.dmp, p.lx),
        "compute_P": lambda o: o.compute_P(p.dmp, p.vpdt, p.lx, p._dx, p._dy, p._dz),
        "compute_TII": lambda o: o.compute_TII(),
        "power_law": lambda o: o.power_law(p.tau_C, p.n_exp, p.relax),
        "compute_V": lambda o: o.compute_V(p._dx, p._dy, p._dz, p.g),
        "copy_boundary_tII": lambda o: o.copy_boundary("tII"),
        "bc_x_lin_VX": lambda o: o.bc("bc_x_lin", "VX", p.sh),
        "bc_x_0_VY": lambda o: o.bc("bc_x_0", "VY"),
        "copy_bc_x_VZ": lambda o: o.bc("copy_bc_x", "VZ"),
        "bc_y_fact_VX": lambda o: o.bc("bc_y_fact", "VX", p.sh),
        "bc_y_fact_VZ": lambda o: o.bc("bc_y_fact", "VZ", 0.0),
        "bc_y_0_VY": lambda o: o.bc("bc_y_0", "VY"),
        "bc_z_0_VZ": lambda o: o.bc("bc_z_0", "VZ"),
        "copy_bc_z_VX": lambda o: o.bc("copy_bc_z", "VX"),
    }


KERNEL_NAMES = sorted(_kernels(CartParams(1, 1, 1, 1, 1, 1)))


@pytest.mark.parametrize("kernel", KERNEL_NAMES)
def test_cart_oracle_kernel_equals_reference_text(kernel):
    st = cart_text_setup()
    o = CartOracle(st)
    _kernels(st.params)[kernel](o)
    expect(o, "cart_" + kernel)


def test_cart_error_equals_reference_text():
    st = cart_text_setup()
    o = CartOracle(st)
    err, comp = o.check_err()
    expect(o, "cart_error")
    ref = [float(np.max(FIX[f"cart_error__out__ERR_{v}"])) / float(np.max(np.abs(FIX[f"cart_in__state__{v}"])))
           for v in ("VX", "VY", "VZ")]
    assert comp == tuple(ref) and err == max(ref)


def test_cart_three_iterations_equal_reference_loop_text():
    """Three iterations of the loop body (:351-379), the exit-test copies
    of V before compute_V! and error! after the BCs on the third."""
    st = cart_text_setup()
    o = CartOracle(st)
    o.step(st.params, 2)
    o.step(st.params, 1, check=True)
    o.check_err()
    expect(o, "cart_loop")


def test_cart_setup_equals_reference_text():
    """initialize_inclusion! (x^2+y^2+z^2 < radius) and initialize_velocity!
    (VX = y sh) against cartesian.cart_diapir's numpy setup."""
    xs = FIX["cart_init__x"]
    n = len(xs)
    s = cart_diapir(n, lx=6.0, radius=1.0, vr=1e2, eta0=1.0)
    eta_ref = FIX["cart_init__out__inclusion"]        # 1 outside, 0.01 inside (radius 1)
    inside = eta_ref != 1.0
    assert np.array_equal(s.fields["eta"] != 1.0, inside)
    assert np.array_equal(s.x, xs)
    assert bits_differ(s.fields["VX"], FIX["cart_init__out__velocity"]) == 0


def test_cart_oracle_solve_reaches_the_exit_test():
    """The loop with its exit test on the script's setup at 15^3 (linear
    viscosity for speed): err falls below eps at a multiple of nout."""
    s = cart_diapir(15, n_exp=1.0)
    o = CartOracle(s)
    r = o.solve(s.params, 100_000, 100, 5e-7)
    assert r["converged"] and r["iters"] % 100 == 0 and r["err"] < 5e-7
    # the light sphere (rho_in < rho0) moves along +z: compute_V! adds
    # -g*rho to the z momentum, so VZ at its centre is positive
    VZ = o.f["VZ"]
    c = s.nx // 2
    assert VZ[c, c, c] > 0.0
