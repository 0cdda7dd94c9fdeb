"""The T_eff accounting against the reference's own single-GPU perf logs.

tests/golden/reference_perf_logs.json holds every line of
out_perf/out_SphericalStokes_mtp_{volta,daint}.txt (nx ny nz niter wtime A_eff
t_it T_eff, written by SphericalStokes_mxpu_mtp.jl:436-443).  The reference
computes A_eff = (4*2 + 6*2 + 2 + 2*2)/1e9*nr*nth*nph*8 GB,
wtime_it = wtime/(ittot-10) and T_eff = A_eff/wtime_it
(SphericalStokes_mxpu_mtp.jl:436-438; SphericalStokes_diapir.jl:447-450).
bench.py reports T_eff from a_eff_bytes(n) = 26 n^3 8 B, and the fused solve
times iterations 11..ittot (ss_result.seconds).  These tests check that
accounting, and the 511^3 baseline figure bench.py divides by, against the
reference's numbers.
"""
import json
from pathlib import Path

import pytest

import bench

ROWS = json.loads((Path(__file__).resolve().parent / "golden" / "reference_perf_logs.json").read_text())


def test_fixture_covers_the_published_sizes():
    sizes = {(r["gpu"], r["nx"]) for r in ROWS}
    assert {("volta", n) for n in (31, 63, 127, 255, 511)} <= sizes
    assert {("daint", n) for n in (31, 63, 127, 255, 383)} <= sizes
    assert all(r["nx"] == r["ny"] == r["nz"] for r in ROWS)


@pytest.mark.parametrize("row", ROWS, ids=lambda r: r["source"])
def test_a_eff_and_t_eff_match_the_reference_logs(row):
    n = row["nx"]
    # the reference's own expression, in its evaluation order: bit for bit
    assert (4 * 2 + 6 * 2 + 2 + 2 * 2) / 1e9 * n * n * n * 8 == row["A_eff_GB"]
    # bench.py's integer byte count gives the same GB figure (to rounding)
    assert bench.a_eff_bytes(n) / 1e9 == pytest.approx(row["A_eff_GB"], rel=1e-15, abs=0)
    # 10 warm-up iterations excluded from the timed region
    t_it = row["wtime"] / (row["niter"] - 10)
    assert t_it == pytest.approx(row["t_it"], rel=1e-15, abs=0)
    assert row["A_eff_GB"] / t_it == pytest.approx(row["T_eff"], rel=1e-14, abs=0)


def test_baseline_figure_is_the_best_511_run():
    """bench.py's vs_baseline divides by the best of the five 511^3 runs
    (the max over repeats, as the reference's plots take it)."""
    t511 = [r["T_eff"] for r in ROWS if r["gpu"] == "volta" and r["nx"] == 511]
    assert len(t511) == 5
    assert max(t511) == pytest.approx(bench.REF_TEFF_511, abs=0.005)
