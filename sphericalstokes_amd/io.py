"""Field dump and restore in the reference's output format.

The reference scripts write, at the end of a run (SphericalStokes_diapir.jl
:455-481 with the writers at :23-43), into an output folder:

* ``{num}_infos.inf``: ``num lr lth lph nr nth nph eps runtime``;
* ``{num}_phys.inf``: ``num eta0 rho0 rho_in n_exp tau_C vr g, sh, r, radius , dmp``
  (the reference's printf format, commas included);
* ``A_<NAME>.bin``: each array as raw FP64 in Julia's column-major order
  (r fastest): R, TH, PH (the 3D coordinate arrays), ETAS, RHO, P, VR, VTH,
  VPH, TAU_RR, TAU_THTH, TAU_PHPH, TAU_RTH, TAU_RPH, TAU_THPH, TII, and the
  convergence history err_evo / iter_evo.

``save_reference_format`` writes that set from a Solver (or a dict of host
fields), so that a script which swaps its loop for this solver still feeds
the reference's post-processing (visu_diapir/visu_diapir.m reads the .inf
files with ``load`` and the arrays with ``fread``/``reshape``).  Two arrays
the reference does not write, A_ETA_INI.bin and A_ETA_PL.bin, are added so
that ``load_reference_format`` + ``restore`` can resume a run (the power law
needs eta_ini; the relaxation needs eta_PL).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .fields import shape

# reference file name -> field name of this package
ARRAYS = {
    "ETAS": "eta", "RHO": "rho", "P": "P", "VR": "VR", "VTH": "Vth", "VPH": "Vph",
    "TAU_RR": "tau_RR", "TAU_THTH": "tau_thth", "TAU_PHPH": "tau_phph",
    "TAU_RTH": "tau_Rth", "TAU_RPH": "tau_Rph", "TAU_THPH": "tau_thph", "TII": "tII",
}
# written in addition to the reference's set, for resuming
EXTRA = {"ETA_INI": "eta_ini", "ETA_PL": "eta_PL"}


def _write(path: Path, a) -> None:
    np.asarray(a, dtype=np.float64).ravel(order="F").tofile(path)


def _read(path: Path, shp) -> np.ndarray:
    a = np.fromfile(path, dtype=np.float64)
    n = int(np.prod(shp))
    if a.size != n:
        raise ValueError(f"{path}: {a.size} values, expected {n} for shape {shp}")
    return np.asfortranarray(a.reshape(shp, order="F"))


def save_reference_format(out, source, setup, *, num: int = 1, eps: float = 1e-8, runtime: float = 0.0,
                          err_evo=(), iter_evo=(), eta0: float = 1.0, rho0: float = 0.0, rho_in: float = -1.0,
                          vr: float = 1e3, radius: float = 0.01, phys_line: str | None = None) -> Path:
    """Writes the reference's output set for `setup` (grid, parameters) with
    the fields of `source`: a Solver (downloaded) or a dict name -> array.
    The physical constants of the diapir's _phys.inf line that the solver
    does not hold (eta0, rho0, rho_in, vr, radius) are taken from the
    keywords; phys_line replaces that line (the plateau script writes its
    own)."""
    out = Path(out)
    out.mkdir(parents=True, exist_ok=True)
    p = setup.params
    nr, nth, nph = setup.nr, setup.nth, setup.nph
    lr = (nr - 1) / p._dr
    lth = (nth - 1) / p._dth
    lph = (nph - 1) / p._dph
    get = source.download if hasattr(source, "download") else source.__getitem__
    # SphericalStokes_diapir.jl:24-25, :29-31 (format strings as written there;
    # eps and runtime go through %d, i.e. as integers)
    (out / f"{num}_infos.inf").write_text(
        "%d %f %f %f %d %d %d %d %d" % (num, lr, lth, lph, nr, nth, nph, round(eps), round(runtime)))
    (out / f"{num}_phys.inf").write_text(
        phys_line if phys_line is not None else
        "%d   %f  %f  %f    %f     %f   %f  %f, %f, %f, %f    , %f"
        % (num, eta0, rho0, rho_in, p.n_exp, p.tau_C, vr, p.g, p.sh, p.r, radius, p.dmp))
    _write(out / "A_err_evo.bin", np.asarray(err_evo, dtype=np.float64))
    _write(out / "A_iter_evo.bin", np.asarray(iter_evo, dtype=np.float64))
    R3 = np.broadcast_to(np.asarray(setup.R)[:, None, None], (nr, nth, nph))
    T3 = np.broadcast_to(np.asarray(setup.th)[None, :, None], (nr, nth, nph))
    P3 = np.broadcast_to(np.asarray(setup.ph)[None, None, :], (nr, nth, nph))
    _write(out / "A_R.bin", R3)
    _write(out / "A_TH.bin", T3)
    _write(out / "A_PH.bin", P3)
    for fname, field in {**ARRAYS, **EXTRA}.items():
        _write(out / f"A_{fname}.bin", get(field))
    return out


def load_reference_format(out, *, num: int = 1):
    """Reads a folder written by the reference (or by save_reference_format):
    returns (dims, coords, fields) with dims = (nr, nth, nph), coords = the 1D
    R, th, ph axes and fields = name -> array for every array present."""
    out = Path(out)
    info = (out / f"{num}_infos.inf").read_text().split()
    nr, nth, nph = int(info[4]), int(info[5]), int(info[6])
    R = _read(out / "A_R.bin", (nr, nth, nph))[:, 0, 0].copy()
    th = _read(out / "A_TH.bin", (nr, nth, nph))[0, :, 0].copy()
    ph = _read(out / "A_PH.bin", (nr, nth, nph))[0, 0, :].copy()
    fields = {}
    for fname, field in {**ARRAYS, **EXTRA}.items():
        path = out / f"A_{fname}.bin"
        if path.exists():
            fields[field] = _read(path, shape(field, nr, nth, nph))
    return (nr, nth, nph), (R, th, ph), fields


def restore(solver, fields) -> None:
    """Uploads the fields of a dump into a Solver of the same size (resume).
    A dump from the reference itself has no eta_ini: eta is used for it, which
    is exact for a linear viscosity and for a run saved before any power-law
    update."""
    f = dict(fields)
    if "eta_ini" not in f and "eta" in f:
        f["eta_ini"] = f["eta"]
    if "eta_PL" not in f and "eta_ini" in f:
        f["eta_PL"] = f["eta_ini"]
    for name, a in f.items():
        solver.upload(name, a)
