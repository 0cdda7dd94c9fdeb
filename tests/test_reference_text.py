"""The oracle against golden vectors evaluated from the reference's own
kernel text (tests/golden/reference_text_vectors.npz, made by
tests/golden/make_reference_text_fixture.py with the Julia-subset evaluator
tests/golden/jl_subset.py from SphericalStokes_diapir.jl,
SphericalStokes_plateau.jl, CartesianStokes_diapir.jl:210-220 and
SphericalStokes_mxpu.jl:295-297).

Every comparison is bit for bit: the vectors and the oracle use the same C
library for sin, tan, pow, exp and log, so any difference is a difference
in an expression -- a pseudo-time step coefficient (KBDT = dmp*2pi*vpdt/lr
*etaSM, GSDT = 4pi*vpdt/lr*etaSM, DT_R = vpdt^2/(KBDT + GSDT/(1+GSDT/eta))),
a metric term, an index, a guard range or an association order.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle.binding import Oracle, initialize_inclusion, initialize_velocity
from sphericalstokes_amd import plateau as PL
from sphericalstokes_amd import setups
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_SHEAR, FIELDS
from sphericalstokes_amd.setups import Params, Setup

GOLDEN = Path(__file__).resolve().parent / "golden"
FIX = np.load(GOLDEN / "reference_text_vectors.npz")
META = json.loads((GOLDEN / "reference_text_vectors.json").read_text())
STATE = [n for n in FIELDS if n != "eta_PL"]


def text_setup(bc_mode):
    """The fixture's input state and scalars as a Setup."""
    nr, nth, nph = META["grid"]
    sc = META["scalars"]
    p = Params(_dr=sc["_dr"], _dth=sc["_dth"], _dph=sc["_dph"], dmp=sc["dmp"], vpdt=sc["vpdt"], lr=sc["lr"],
               tau_C=sc["tau_C"], n_exp=sc["n_exp"], g=sc["g"], r=sc["r"], sh=sc["sh"], bc_mode=bc_mode)
    f = {n: np.array(FIX[f"in__state__{n}"], order="F", copy=True) for n in STATE}
    f["eta_PL"] = np.array(f["eta"], order="F", copy=True)
    return Setup(nr, nth, nph, FIX["R1"].copy(), FIX["th1"].copy(), FIX["ph1"].copy(), p, f, name="reference_text")


def bits_differ(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    return int(np.count_nonzero(a.view(np.uint64) != b.view(np.uint64)))


def expect(o, case, names=None):
    keys = [k for k in FIX.files if k.startswith(case + "__out__")]
    assert keys, case
    bad = {}
    for k in keys:
        n = k.split("__out__")[1]
        if names is not None and n not in names:
            continue
        d = bits_differ(o.f[n], FIX[k])
        if d:
            bad[n] = d
    assert not bad, f"{case}: values differing from the reference text per field: {bad}"


KERNELS = {
    "compute_P": lambda o: o.compute_P(),
    "compute_TII": lambda o: o.compute_TII(),
    "power_law": lambda o: o.power_law(),
    "compute_V": lambda o: o.compute_V(),
    "BC_R": lambda o: o.bc_R(),
    "BC_th": lambda o: o.bc_th(),
    "BC_ph": lambda o: o.bc_ph(),
    "equal3": lambda o: o.equal3(),
    "copy_BC_tII": lambda o: o.copy_bc("tII"),
}


@pytest.mark.parametrize("kernel", sorted(KERNELS))
def test_oracle_kernel_equals_reference_text(kernel):
    o = Oracle(text_setup(BC_SHEAR))
    KERNELS[kernel](o)
    expect(o, "diapir_" + kernel)


def test_pseudo_time_steps_are_those_of_the_reference_text():
    """compute_P! alone: DT_R carries KBDT and GSDT; P carries KBDT; the
    normal and shear stresses carry GSDT and its 4-point averages."""
    o = Oracle(text_setup(BC_SHEAR))
    o.compute_P()
    for n in ("DT_R", "P", "tau_RR", "tau_Rth", "tau_Rph", "tau_thph"):
        assert bits_differ(o.f[n], FIX[f"diapir_compute_P__out__{n}"]) == 0, n
    # and the vectors do depend on the coefficients: a 2pi <-> 4pi slip in
    # KBDT changes P and DT_R everywhere compute_P! writes
    s = text_setup(BC_SHEAR)
    s.params.dmp *= 2.0
    o2 = Oracle(s)
    o2.compute_P()
    assert bits_differ(o2.f["P"], FIX["diapir_compute_P__out__P"]) > 0.9 * o2.f["P"].size
    assert bits_differ(o2.f["DT_R"], FIX["diapir_compute_P__out__DT_R"]) > 0.9 * o2.f["DT_R"].size


def test_check_err_equals_reference_text():
    """check_err! (V_err = V - V_err) and the host's max-of-ratios
    (SphericalStokes_diapir.jl:428-431) on the kernel's output."""
    o = Oracle(text_setup(BC_SHEAR))
    err, comp = o.check_err()
    expect(o, "diapir_check_err")
    ref = []
    for v, e in (("VR", "VR_err"), ("Vth", "Vth_err"), ("Vph", "Vph_err")):
        E = FIX[f"diapir_check_err__out__{e}"]
        V = FIX[f"in__state__{v}"]
        ref.append(float(np.max(np.abs(E))) / float(np.max(np.abs(V))))
    assert comp == tuple(ref)
    assert err == max(ref)


def test_maxloc_equals_reference_text():
    o = Oracle(text_setup(BC_SHEAR))
    o.maxloc()
    expect(o, "maxloc")


@pytest.mark.parametrize("kernel", ["BC_R", "BC_th", "BC_ph"])
def test_free_slip_bc_equals_plateau_text(kernel):
    o = Oracle(text_setup(BC_FREE_SLIP))
    KERNELS[kernel](o)
    expect(o, "plateau_" + kernel)


@pytest.mark.parametrize("case,bc_mode", [("diapir_loop", BC_SHEAR), ("plateau_loop", BC_FREE_SLIP)])
def test_three_iterations_equal_reference_loop_text(case, bc_mode):
    """Three iterations of the scripts' loop body (diapir :413-425, plateau
    :526-538) with n_exp = 3: every field bit for bit."""
    o = Oracle(text_setup(bc_mode))
    o.step(3)
    expect(o, case)


def test_relaxed_power_law_equals_cartesian_text():
    """power_law! with relaxation (CartesianStokes_diapir.jl:210-220):
    eta_PL relaxed in log space against the previous eta_PL, then the
    harmonic combination."""
    nr, nth, nph = META["grid"]
    s = text_setup(BC_SHEAR)
    s.params.relax = META["relax"]["relax"]
    s.params.tau_C = META["relax"]["tau_C"]
    s.params.n_exp = META["relax"]["n_exp"]
    s.fields["tII"] = np.array(FIX["relax__in__tII"], order="F", copy=True)
    s.fields["eta_ini"] = np.array(FIX["relax__in__eta_ini"], order="F", copy=True)
    s.fields["eta_PL"] = np.array(FIX["relax__in__eta_PL_old"], order="F", copy=True)
    o = Oracle(s)
    o.power_law()
    assert bits_differ(o.f["eta_PL"], FIX["relax__out__eta_PL"]) == 0
    assert bits_differ(o.f["eta"], FIX["relax__out__eta"]) == 0


def test_setup_kernels_equal_reference_text():
    """initialize_inclusion! (squared distance against radius) and
    initialize_velocity! of the diapir script: the oracle's and the
    package's numpy setups."""
    m = META["init"]
    R1, th1, ph1 = FIX["init__R1"], FIX["init__th1"], FIX["init__ph1"]
    n = len(R1)
    A = np.zeros((n, n, n), order="F")
    initialize_inclusion(A, R1, th1, ph1, m["r"], m["radius"], m["in"])
    assert bits_differ(A, FIX["init__out__inclusion"]) == 0
    B = np.zeros((n, n, n), order="F")
    setups._inclusion(B, R1, th1, ph1, m["r"], m["radius"], m["in"])
    assert bits_differ(B, FIX["init__out__inclusion"]) == 0
    V = np.zeros((n, n, n + 1), order="F")
    initialize_velocity(V, th1, m["fact"], m["r"])
    assert bits_differ(V, FIX["init__out__velocity"]) == 0


@pytest.mark.parametrize("tag", ["geo70", "geo31"])
def test_plateau_geometry_equals_reference_text(tag):
    """initialize_geometry! of the plateau script (thirteen ordered
    conditions) against plateau.plateau_reference's density and viscosity."""
    gnr, gnt, gnp = META[tag]["grid"]
    su = PL.plateau_reference(gnr, gnt, gnp)
    assert np.array_equal(su.R, FIX[f"{tag}__R1"]) and np.array_equal(su.th, FIX[f"{tag}__th1"])
    assert np.array_equal(su.ph, FIX[f"{tag}__ph1"])
    for name in ("rho", "eta"):
        assert bits_differ(su.fields[name], FIX[f"{tag}__out__{name}"]) == 0, name
    assert len(np.unique(FIX[f"{tag}__out__rho"])) >= 3


def test_vectors_are_not_trivial():
    """Each kernel's vector differs from its input (the random state makes
    every output move)."""
    for k in FIX.files:
        if "__out__" not in k or not k.startswith(("diapir_", "plateau_BC")):
            continue
        n = k.split("__out__")[1]
        src = f"in__state__{n}"
        if src in FIX.files and not k.startswith("diapir_loop"):
            assert bits_differ(FIX[k], FIX[src]) > 0, k
    assert math.isfinite(float(np.sum(FIX["diapir_loop__out__VR"])))
