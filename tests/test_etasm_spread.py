"""The etaSM spread option (ss_params.etasm_spread) and why it exists, on
the CPU oracle.

With the reference's pseudo-time steps (maxloc! over the 7-point star,
SphericalStokes_diapir.jl:149-156, CFL = 1/(2 + 4.5 log10 vr)) a sharp
viscosity jump of 1e4 between neighbouring cells diverges within a few
hundred iterations in the diapir sector, also at half the step: next to the
jump a face's step averages DT_R over a cell whose etaSM saw the stiff side
and a cell whose etaSM did not, and the stiff pressure/stress update of the
first is then driven with the second's large step.  Taking etaSM over two
cells only moves that interface.  Spreading etaSM with a decay of 10 per
cell bounds the ratio of neighbouring cells' steps, which keeps the
iteration stable at any contrast; the converged solution does not depend on
the step sizes."""
import math

import numpy as np

from oracle.binding import Oracle
from sphericalstokes_amd import setups


def _max_v(o):
    return max(float(np.abs(o.f[k]).max()) for k in ("VR", "Vth", "Vph"))


def _run(vr, spread, iters, cfl=1.0, n=31):
    st = setups.diapir_sector(n, n_exp=1.0, vr=vr)
    st.params.vpdt *= cfl
    st.params.etasm_spread = spread
    o = Oracle(st)
    o.step(iters)
    return o


def test_reference_rule_diverges_at_contrast_1e4():
    for n, cfl, iters in ((31, 1.0, 300), (63, 0.5, 100)):
        o = _run(1e4, 0, iters, cfl, n)
        m = _max_v(o)
        assert not (m < 1e3), (n, cfl, m)         # blown up (or non-finite)


def test_reference_rule_stable_at_contrast_1e3():
    o = _run(1e3, 0, 600)
    assert _max_v(o) < 1e-1


def test_spread_is_stable_at_contrast_1e5():
    """Bounded where the reference rule blows up within 100 iterations (at
    63^3 the damped pseudo-time waves reach max|V| of order 1 in the first
    ~1500 iterations and then decay)."""
    o = _run(1e5, 5, 600, n=31)
    assert _max_v(o) < 1e-1
    o = _run(1e5, 5, 300, n=63)
    assert np.isfinite(_max_v(o)) and _max_v(o) < 5.0


def test_lognormal_95_spread_stable_where_reference_diverges():
    """The mixed-precision configuration's field at 95^3 (jump of 1.4e4 at the
    inclusion's edge): the reference rule blows up within 100 iterations, the
    spread etaSM keeps V bounded and falling."""
    ref = setups.diapir_lognormal(95, etasm_spread=0)
    o = Oracle(ref)
    o.step(100)
    assert not (_max_v(o) < 1e3)
    st = setups.diapir_lognormal(95)
    assert st.params.etasm_spread == 5
    o = Oracle(st)
    o.step(100)
    m1 = _max_v(o)
    o.step(400)
    m2 = _max_v(o)
    assert np.isfinite(m2) and m2 < m1 and m2 < 0.2, (m1, m2)


def test_spread_definition_brute_force():
    """orc_spread_etasm against its definition written as one box maximum:
    etaSM(i) = max over |a|,|b|,|c| <= m of in(clamp(i + (a,b,c))) *
    0.1^|a| * 0.1^|b| * 0.1^|c| (the three separable passes reach the same
    candidates; products may round differently, hence the 4e-16)."""
    from oracle.binding import lib, _ptr
    rng = np.random.default_rng(4)
    nr, nt, np_ = 7, 9, 8
    A = np.asfortranarray(np.exp(3 * rng.standard_normal((nr, nt, np_))))
    for m in (1, 2, 5):
        got = A.copy(order="F")
        tmp = np.zeros_like(got)
        lib().orc_spread_etasm(_ptr(got), _ptr(tmp), nr, nt, np_, m)
        w = [1.0]
        for _ in range(8):
            w.append(w[-1] * 0.1)
        exp = np.zeros_like(A)
        for i in range(nr):
            for j in range(nt):
                for k in range(np_):
                    best = 0.0
                    for a in range(-m, m + 1):
                        for b in range(-m, m + 1):
                            for c in range(-m, m + 1):
                                v = A[min(max(i + a, 0), nr - 1), min(max(j + b, 0), nt - 1),
                                      min(max(k + c, 0), np_ - 1)]
                                best = max(best, v * w[abs(a)] * w[abs(b)] * w[abs(c)])
                    exp[i, j, k] = best
        assert np.allclose(got, exp, rtol=4e-16, atol=0), m
        assert np.all(got >= A)


def test_spread_bounds_the_neighbour_ratio():
    """A single stiff cell (1e5) in a field of 1e-3: after the spread the
    etaSM of neighbouring cells differs by at most a factor 10 (to rounding)
    wherever the spread reaches, and the stiff cell keeps its value."""
    from oracle.binding import lib, _ptr
    nr = nt = np_ = 15
    A = np.full((nr, nt, np_), 1e-3, order="F")
    A[7, 7, 7] = 1e5
    got = A.copy(order="F")
    tmp = np.zeros_like(got)
    lib().orc_spread_etasm(_ptr(got), _ptr(tmp), nr, nt, np_, 5)
    assert got[7, 7, 7] == 1e5
    for ax in range(3):
        r = np.maximum(got, np.roll(got, 1, axis=ax)) / np.minimum(got, np.roll(got, 1, axis=ax))
        inner = [slice(3, 12)] * 3
        assert r[tuple(inner)].max() <= 10 * (1 + 1e-12), ax
    assert math.isclose(got[7, 7, 2], 1.0, rel_tol=1e-12)      # 5 cells away: 1e5 * 0.1^5
