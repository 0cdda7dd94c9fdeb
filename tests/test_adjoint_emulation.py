"""CPU: the r-marching adjoint kernel (sphericalstokes_amd/csrc/
sspt_adj_fused_body.h), run by the host emulation of its block schedule,
against the CPU reference adjoint step (oracle/sspt_adjoint.c, pinned by
tests/test_adjoint_oracle.py).

One application of T^T on random adjoint fields at a random state, every BC
mode, relaxation and cutoffs, several tilings (blocks of 62 x 10 owned
points, r-chunks down to one plane).  The two sum the same terms in different
orders: rel. 1e-12 of each field's max.
"""
import numpy as np
import pytest

from emu import emu
from oracle.binding import ADJ_FIELDS as ORC_ADJ, Oracle
from sphericalstokes_amd import setups
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_PURE_SHEAR, BC_SHEAR

CASES = {
    "linear_free_slip": dict(n_exp=1.0, bc_mode=BC_FREE_SLIP),
    "powerlaw_simple_shear": dict(n_exp=3.0, bc_mode=BC_SHEAR, sh=0.7),
    "powerlaw_relax_cutoffs": dict(n_exp=3.0, bc_mode=BC_FREE_SLIP, relax=0.5, eta_min=0.3, eta_max=3.0),
    "powerlaw_pure_shear": dict(n_exp=3.0, bc_mode=BC_PURE_SHEAR, sh=1.0),
}
# (nr, nth, nph, r-chunk): several tiles in th and phi, chunk edges, the
# smallest box
SHAPES = [(11, 13, 70, 0), (11, 13, 70, 3), (9, 30, 7, 2), (5, 5, 130, 1), (3, 3, 3, 1)]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{a}x{b}x{c}_r{d}" for a, b, c, d in SHAPES])
@pytest.mark.parametrize("case", list(CASES))
def test_marching_adjoint_matches_reference(case, shape):
    nr, nth, nph, rch = shape
    st = setups.diapir_sector(nr, nth=nth, nph=nph, **CASES[case])
    setups.random_state(st, seed=11)
    o = Oracle(st)
    o.step(2)
    rng = np.random.default_rng(3)
    lam = {n: rng.standard_normal(o.f[n].shape) for n in ORC_ADJ}
    if st.params.relax == 1.0:
        lam["eta_PL"][:] = 0.0
    ref = o.adj_step(lam)
    got = emu.adj_step(st, o.f, lam, st.params, rchunk=rch)
    for n in ORC_ADJ:
        scale = max(np.abs(ref[n]).max(), 1e-300)
        err = np.abs(got[n] - ref[n]).max() / scale
        assert err <= 1e-12, (n, err)


@pytest.mark.parametrize("shape", [(11, 13, 70, 0), (11, 13, 70, 3), (9, 30, 7, 2)],
                         ids=["11x13x70_r0", "11x13x70_r3", "9x30x7_r2"])
def test_marching_adjoint_objective_matches_reference(shape):
    """The step with the objective: T^T lam + dJ/dVR, dJ the oracle's
    objective weights (centre r = 1.5, squared radius 0.1^2), which are zero
    outside the inclusion; the kernel adds them where its own inclusion test
    holds."""
    nr, nth, nph, rch = shape
    st = setups.diapir_sector(nr, nth=nth, nph=nph, n_exp=3.0, bc_mode=BC_FREE_SLIP)
    setups.random_state(st, seed=11)
    o = Oracle(st)
    o.step(2)
    _, dJ = o.objective(1.5, 0.1 ** 2)
    assert np.count_nonzero(dJ) > 0
    rng = np.random.default_rng(4)
    lam = {n: rng.standard_normal(o.f[n].shape) for n in ORC_ADJ}
    lam["eta_PL"][:] = 0.0
    ref = o.adj_step(lam)
    ref["VR"] = ref["VR"] + dJ
    got = emu.adj_step(st, o.f, lam, st.params, rchunk=rch, objective=(1.5, 0.1 ** 2, dJ))
    for n in ORC_ADJ:
        scale = max(np.abs(ref[n]).max(), 1e-300)
        err = np.abs(got[n] - ref[n]).max() / scale
        assert err <= 1e-12, (n, err)


@pytest.mark.parametrize("shape", [(11, 13, 70, 0), (9, 30, 7, 2), (5, 5, 130, 1)],
                         ids=["11x13x70_r0", "9x30x7_r2", "5x5x130_r1"])
@pytest.mark.parametrize("case", list(CASES))
def test_marching_adjoint_rhs_instances(case, shape):
    """The right-hand-side instances of the kernel (the correction sweeps
    d = T^T d + r of the mixed-precision adjoint solve): in FP64 the output
    is the plain step's plus r at every point of every field, bit for bit
    (the kernel writes each point of each field once, and adding r is its
    last operation); in FP32 (state, lam, r and tables rounded, FP32
    arithmetic) within 2e-6 of each field's max of the oracle's T^T lam + r
    (FP32 rounding of a few dozen terms: 8e-8 to 3.3e-7 measured)."""
    nr, nth, nph, rch = shape
    st = setups.diapir_sector(nr, nth=nth, nph=nph, **CASES[case])
    setups.random_state(st, seed=11)
    o = Oracle(st)
    o.step(2)
    rng = np.random.default_rng(5)
    lam = {n: rng.standard_normal(o.f[n].shape) for n in ORC_ADJ}
    rhs = {n: rng.standard_normal(o.f[n].shape) for n in ORC_ADJ}
    if st.params.relax == 1.0:
        lam["eta_PL"][:] = 0.0
        rhs["eta_PL"][:] = 0.0   # eta_PL is not an output then
    copy = lambda d: {k: v.copy() for k, v in d.items()}
    base = emu.adj_step(st, o.f, copy(lam), st.params, rchunk=rch)
    got = emu.adj_step_rhs(st, o.f, copy(lam), st.params, rhs, rchunk=rch)
    g32 = emu.adj_step_rhs(st, o.f, copy(lam), st.params, rhs, f32=True, rchunk=rch)
    ref = o.adj_step(copy(lam))
    for n in ORC_ADJ:
        assert np.array_equal(got[n], base[n] + rhs[n]), n
        e = ref[n] + rhs[n]
        rel = np.abs(g32[n] - e).max() / max(np.abs(e).max(), 1e-300)
        assert rel <= 2e-6, (n, rel)
