"""GPU: error behaviour of the Cartesian and cylindrical twins' C ABI
(include/sspt_cart.h, include/sspt_cyl.h): argument and state checks that
return a status without launching anything, and a NaN in the state that the
loop's exit test turns into SS_ERR_NONFINITE (the reference's err is NaN then
and its loop never breaks; the library stops at that check)."""
import ctypes as C

import numpy as np
import pytest

from sphericalstokes_amd import _lib
from sphericalstokes_amd._lib import SS_ERR_ARG, SS_ERR_NONFINITE, SS_ERR_STATE, SSError
from sphericalstokes_amd.cartesian import CART_ID, CartSolver, cart_diapir
from sphericalstokes_amd.cylindrical import CYL_BC, CYL_ID, CylSolver, cyl_diapir

pytestmark = pytest.mark.gpu


def _code(excinfo):
    return excinfo.value.code


def test_create_rejects_small_or_huge_boxes():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.ss_cart_create(2, 9, 9, 0, C.byref(h)) == SS_ERR_ARG and not h.value
    assert L.ss_cart_create(9, 9, 5000, 0, C.byref(h)) == SS_ERR_ARG and not h.value
    assert L.ss_cyl_create(3, 9, 9, 0, C.byref(h)) == SS_ERR_ARG and not h.value
    assert L.ss_cyl_create(9, 9, 9, 7, C.byref(h)) == SS_ERR_ARG and not h.value   # no such device
    assert L.ss_cart_create(9, 9, 9, 0, None) == SS_ERR_ARG


def test_cart_argument_and_state_errors():
    with CartSolver(9, 9, 9) as s:
        with pytest.raises(SSError) as e:          # y was never set
            s.bc_x_lin("VX", 1.0)
        assert _code(e) == SS_ERR_STATE
        s.set_y(np.linspace(-0.5, 0.5, 9))
        with pytest.raises(SSError) as e:          # VY has y at faces
            s.bc_y_fact("VY", 1.0)
        assert _code(e) == SS_ERR_ARG
        with pytest.raises(SSError) as e:          # not a velocity
            s.bc_x_0("P")
        assert _code(e) == SS_ERR_ARG
        L = _lib.lib()
        assert L.ss_cart_equal(s._h, 99, CART_ID["P"]) == SS_ERR_ARG
        assert L.ss_cart_equal(s._h, CART_ID["VX"], CART_ID["P"]) == SS_ERR_ARG   # shapes differ
        assert L.ss_cart_last_error(s._h)                                          # a message is left
        with pytest.raises(ValueError):
            s.upload("VX", np.zeros((9, 9, 9)))


def test_cyl_argument_and_state_errors():
    with CylSolver(9, 9, 9) as s:
        p = cyl_diapir(9).params
        with pytest.raises(SSError) as e:          # R, th never set
            s.compute_P(p.dr, p.dth, p.dz)
        assert _code(e) == SS_ERR_STATE
        st = cyl_diapir(9)
        s.set_grid(st.R, st.th)
        with pytest.raises(SSError) as e:          # theta at faces for Vth
            s.bc("bc_th_fact", "Vth", 1.0, 1000.0)
        assert _code(e) == SS_ERR_ARG
        L = _lib.lib()
        assert L.ss_cyl_bc(s._h, CYL_ID["VZ"], 99, 0.0, 0.0) == SS_ERR_ARG
        assert L.ss_cyl_equal(s._h, CYL_ID["VR"], CYL_ID["P"]) == SS_ERR_ARG
        assert L.ss_cyl_upload(s._h, CYL_ID["P"], None, 0) == SS_ERR_ARG
        assert L.ss_cyl_exit_error(s._h, None, None) == SS_ERR_ARG
        assert CYL_BC["bc_z_lin"] == 7


def test_cart_nan_stops_the_loop_nonfinite():
    st = cart_diapir(11, n_exp=1.0)
    st.fields["VX"][5, 5, 5] = np.nan
    with CartSolver(11, 11, 11) as s:
        s.load_setup(st)
        with pytest.raises(SSError) as e:
            s.pt_solve(st.params, 1000, 100, 1e-8)
        assert _code(e) == SS_ERR_NONFINITE


def test_cyl_nan_stops_the_loop_nonfinite():
    st = cyl_diapir(11, n_exp_PL=1.0)
    st.fields["VR"][5, 5, 5] = np.nan
    with CylSolver(11, 11, 11) as s:
        s.load_setup(st)
        with pytest.raises(SSError) as e:
            s.pt_solve(st.params, 1000, 100, 1e-8)
        assert _code(e) == SS_ERR_NONFINITE
