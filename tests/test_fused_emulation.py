"""The fused PT schedule (sspt_fused_body.h), run by the host emulation, must
reproduce the reference kernel sequence (the CPU oracle) bit for bit: the
same expressions in the same order, only re-scheduled into one r-marching
pass plus a boundary pass.  Divisions by metric coefficients and by the
shared stress denominator use a correctly rounded reciprocal plus one FMA
correction, which returns the IEEE quotient, so they stay bitwise.  The one
exception is x^(1-n_exp) for integer n_exp > 1, the correctly rounded power
instead of pow (which misses it in about 1 case of 1 400 by 1 ulp): those
cases are compared to rel-L2 1e-12."""
import math

import numpy as np
import pytest

from emu import emu
from helpers import compare_fields
from oracle.binding import Oracle
from sphericalstokes_amd import setups
from sphericalstokes_amd.fields import BC_FREE_SLIP, BC_PURE_SHEAR, BC_SHEAR


def _case(kind, n=(13, 11, 17), seed=1, **kw):
    nr, nt, np_ = n
    if kind == "diapir":
        s = setups.diapir(nr, nt, np_, vr=1e2, **kw)
    else:
        s = setups.diapir_sector(nr, nth=nt, nph=np_, **kw)
    setups.random_state(s, seed=seed)
    return s


CASES = [
    ("diapir", dict(n_exp=3.0, bc_mode=BC_SHEAR)),
    ("diapir", dict(n_exp=1.0, bc_mode=BC_SHEAR)),
    ("sector", dict(n_exp=1.0, bc_mode=BC_FREE_SLIP)),
    ("sector", dict(n_exp=3.0, bc_mode=BC_PURE_SHEAR, sh=0.7)),
    ("sector", dict(n_exp=3.0, bc_mode=BC_SHEAR, sh=1.0, relax=0.5, eta_min=0.5, eta_max=2.0)),
]


@pytest.mark.parametrize("kind,kw", CASES)
@pytest.mark.parametrize("niter", [1, 3])
def test_fused_matches_oracle_bitwise(kind, kw, niter):
    s = _case(kind, **kw)
    o = Oracle(s)
    o.step(niter)
    got, _ = emu.run(s, s.fields, s.params, niter)
    exact = kw.get("n_exp", 1.0) == 1.0
    res, bad = compare_fields(got, o.f, exact=exact, tol=1e-12)
    assert not bad, {k: res[k] for k in bad}


@pytest.mark.parametrize("shape", [(4, 4, 4), (7, 70, 5), (20, 9, 130), (9, 13, 63), (9, 13, 64), (6, 10, 7),
                                   (6, 11, 9), (5, 21, 66), (5, 31, 3)])
def test_fused_tiling_edges(shape):
    """Sizes around the 62 x 10 tile edges and the smallest grids."""
    s = _case("sector", n=shape, n_exp=1.0, bc_mode=BC_FREE_SLIP, seed=3)
    o = Oracle(s)
    o.step(2)
    got, _ = emu.run(s, s.fields, s.params, 2)
    res, bad = compare_fields(got, o.f)
    assert not bad, {k: res[k] for k in bad}


@pytest.mark.parametrize("rchunk", [1, 2, 3, 5])
def test_fused_r_chunks(rchunk):
    """Blocks that start mid-column warm up over two planes; any chunk length
    must give the same result as one march over all of r."""
    s = _case("diapir", n=(17, 9, 11), n_exp=1.0, seed=4)
    o = Oracle(s)
    o.step(2)
    got, _ = emu.run(s, s.fields, s.params, 2, rchunk=rchunk)
    res, bad = compare_fields(got, o.f)
    assert not bad, {k: res[k] for k in bad}


def test_fused_check_reduction_matches_check_err():
    s = _case("diapir", n_exp=1.0, seed=5)
    o = Oracle(s)
    o.step(1, equal3=True)
    err_ref, comp_ref = o.check_err()
    got, red = emu.run(s, s.fields, s.params, 1, check_last=True)
    comp = red[:3] / red[3:]
    assert np.array_equal(comp, np.array(comp_ref))
    res, bad = compare_fields(got, o.f)
    assert not bad, {k: res[k] for k in bad}


def test_powerlaw_fused_vs_emulated_chunks_bitwise():
    """With n_exp = 3 the fused schedule is bitwise independent of tiling and
    r-chunking (same arithmetic everywhere)."""
    s = _case("diapir", n=(17, 9, 70), n_exp=3.0, seed=6, relax=0.5, eta_min=0.5, eta_max=2.0)
    a, _ = emu.run(s, s.fields, s.params, 3, rchunk=0)
    b, _ = emu.run(s, s.fields, s.params, 3, rchunk=4)
    res, bad = compare_fields(b, a)
    assert not bad, {k: res[k] for k in bad}


@pytest.mark.parametrize("kind,kw", CASES)
@pytest.mark.parametrize("niter", [1, 2])
def test_paired_buffers_give_the_same_iterates(kind, kw, niter):
    """With each state field's in/out buffers the two halves of one array,
    the main kernel addresses both through one base and byte offsets
    (F_PAIRED); the iterates must not change by a bit (an odd number of
    iterations leaves the state in the upper halves)."""
    s = _case(kind, **kw)
    a, _ = emu.run(s, s.fields, s.params, niter)
    b, _ = emu.run(s, s.fields, s.params, niter, paired=True)
    res, bad = compare_fields(b, a)
    assert not bad, {k: res[k] for k in bad}


@pytest.mark.parametrize("shape,rchunk", [((13, 11, 17), 0), ((9, 13, 64), 3), ((20, 9, 130), 5), ((6, 21, 66), 0)])
@pytest.mark.parametrize("paired", [False, True])
def test_spread_etasm_matches_oracle_bitwise(shape, rchunk, paired):
    """etasm_spread = 3: the spread etaSM, filled before the launch and read
    by the main kernel (F_SMBUF), against the oracle's maxloc! + copy_BC! +
    orc_spread_etasm."""
    s = _case("sector", n=shape, n_exp=1.0, bc_mode=BC_FREE_SLIP, seed=7)
    s.params.etasm_spread = 3
    o = Oracle(s)
    o.step(3)
    got, _ = emu.run(s, s.fields, s.params, 3, rchunk=rchunk, paired=paired)
    res, bad = compare_fields(got, o.f)
    assert not bad, {k: res[k] for k in bad}
    # the spread etaSM differs from the reference one on this random viscosity
    s1 = _case("sector", n=shape, n_exp=1.0, bc_mode=BC_FREE_SLIP, seed=7)
    o1 = Oracle(s1)
    o1.step(1)
    assert not np.array_equal(o1.f["etaSM"], o.f["etaSM"])
