"""GPU: libsspt.so against the golden vectors evaluated from the reference's
own kernel text (tests/golden/reference_text_vectors.npz; see
tests/test_reference_text.py for how they are made and what pins them).

The per-kernel entry points must give the vectors bit for bit, power_law
within 4 ulp (x^(1-n_exp) is the correctly rounded power on the GPU, the C
library's pow in the vectors); three iterations of the unfused and of the
fused loop within rel-L2 1e-12 per field (n_exp = 3, so the power law's
last bits enter), the FP64 drop-in tolerance being 1e-10.
"""
import sys

import numpy as np
import pytest

from helpers import max_ulp, rel_l2
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_SHEAR, FIELDS
from test_reference_text import FIX, META, text_setup

pytestmark = pytest.mark.gpu
POW_ULP = 4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _solver(st):
    from sphericalstokes_amd.solver import Solver
    s = Solver(st.nr, st.nth, st.nph)
    s.set_grid(st.R, st.th, st.ph)
    for n in FIELDS:
        s.upload(n, st.fields[n])
    return s


def _check(s, case, ulp_fields=()):
    bad = {}
    for k in FIX.files:
        if not k.startswith(case + "__out__"):
            continue
        n = k.split("__out__")[1]
        got, ref = s.download(n), FIX[k]
        if n in ulp_fields:
            u = max_ulp(got, ref)
            if u > POW_ULP:
                bad[n] = f"{u} ulp"
        elif not np.array_equal(got, ref):
            bad[n] = int(np.count_nonzero(got != ref))
    assert not bad, f"{case}: {bad}"


KERNELS = {
    "compute_P": lambda s, p: s.compute_P(p._dr, p._dth, p._dph, p.dmp, p.vpdt, p.lr),
    "compute_TII": lambda s, p: s.compute_TII(),
    "power_law": lambda s, p: s.power_law(p.tau_C, p.n_exp),
    "compute_V": lambda s, p: s.compute_V(p._dr, p._dth, p._dph, p.g),
    "BC_R": lambda s, p: s.BC_R(),
    "BC_th": lambda s, p: s.BC_th(p.bc_mode, p.r, p.sh),
    "BC_ph": lambda s, p: s.BC_ph(p.bc_mode, p.r, p.sh),
    "equal3": lambda s, p: s.equal3(),
    "copy_BC_tII": lambda s, p: s.copy_BC("tII"),
}


@pytest.mark.parametrize("kernel", sorted(KERNELS))
def test_kernel_equals_reference_text(kernel):
    st = text_setup(BC_SHEAR)
    say("reference text:", kernel, tuple(META["grid"]))
    with _solver(st) as s:
        KERNELS[kernel](s, st.params)
        _check(s, "diapir_" + kernel, ulp_fields=("eta",) if kernel == "power_law" else ())


def test_maxloc_equals_reference_text():
    st = text_setup(BC_SHEAR)
    say("reference text: maxloc")
    with _solver(st) as s:
        s.maxloc()
        _check(s, "maxloc")


@pytest.mark.parametrize("kernel", ["BC_R", "BC_th", "BC_ph"])
def test_free_slip_bc_equals_plateau_text(kernel):
    st = text_setup(BC_FREE_SLIP)
    say("reference text: plateau", kernel)
    with _solver(st) as s:
        KERNELS[kernel](s, st.params)
        _check(s, "plateau_" + kernel)


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("case,bc_mode", [("diapir_loop", BC_SHEAR), ("plateau_loop", BC_FREE_SLIP)])
def test_three_iterations_match_reference_loop_text(case, bc_mode, fused):
    st = text_setup(bc_mode)
    say("reference text:", case, "fused" if fused else "unfused")
    with _solver(st) as s:
        if fused:
            s.pt_iterate(st.params, 3)
        else:
            for _ in range(3):
                s.pt_step_unfused(st.params)
        worst = 0.0
        for k in FIX.files:
            if not k.startswith(case + "__out__"):
                continue
            n = k.split("__out__")[1]
            e = rel_l2(s.download(n), FIX[k])
            worst = max(worst, e)
            assert e <= 1e-12, (n, e)
        say("  worst field rel-L2", worst)
