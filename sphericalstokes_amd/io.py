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


# ---------------------------------------------------------------- XDMF output
# The reference's visualisation output (scripts_perf/SphericalStokes_mxpu.jl
# :443-464, writers in scripts_perf/data_io2.jl:1-132): per save, the inner
# cells' Cartesian coordinates X, Y, Z (sph2cart!, SphericalStokes_mxpu.jl
# :32-38) and fields averaged onto the inner cell centres, described by an
# XDMF 3.0 file as a temporal collection of 3DSMesh grids with cell-centred
# attributes.  The reference stores the arrays in one HDF5 file per save;
# this build has no HDF5 library, so each save is one raw little-endian FP64
# file holding the datasets back to back, and every XDMF DataItem points
# into it with Format="Binary" and a byte Seek offset.  Dataset names,
# order of dimensions and the XML layout are the reference's.
XDMF_FIELDS = ["eta", "P", "tII", "VR", "VT", "VF", "t_RR", "t_TT", "t_FF", "t_RT", "τ_RF", "τ_TF"]


def sph2cart(R, th, ph):
    """X, Y, Z of the points (R_i, th_j, ph_k) as the reference maps them
    (sph2cart!, SphericalStokes_mxpu.jl:32-38; the pi/2 - th is specific to
    its sector about the equator): X = R cos(pi/2-th) cos(ph),
    Y = R sin(pi/2-th) cos(ph), Z = R sin(ph).  3D arrays (Julia order)."""
    R3 = np.asarray(R, dtype=np.float64)[:, None, None]
    T3 = np.asarray(th, dtype=np.float64)[None, :, None]
    P3 = np.asarray(ph, dtype=np.float64)[None, None, :]
    X = R3 * np.cos(np.pi / 2 - T3) * np.cos(P3)
    Y = R3 * np.sin(np.pi / 2 - T3) * np.cos(P3)
    Z = np.broadcast_to(R3 * np.sin(P3), X.shape)
    return X, Y, np.ascontiguousarray(Z)


def _inn(A):
    return A[1:-1, 1:-1, 1:-1]


def xdmf_datasets(source, setup) -> dict:
    """The reference's save set (SphericalStokes_mxpu.jl:450-452) on the
    (nr-2, nth-2, nph-2) inner cells: X, Y, Z, then eta, P, tII, the
    velocities averaged from their faces (av_x2i/av_y2i/av_z2i) and the
    stresses (inner centres; shear stresses averaged from their 4 edges,
    av_xya/av_xza/av_yza), as Fortran-ordered float64 arrays."""
    get = source.download if hasattr(source, "download") else source.__getitem__
    X, Y, Z = sph2cart(setup.R[1:-1], setup.th[1:-1], setup.ph[1:-1])
    VR, VT, VP = get("VR"), get("Vth"), get("Vph")
    RT, RP, TP = get("tau_Rth"), get("tau_Rph"), get("tau_thph")
    d = {"X": X, "Y": Y, "Z": Z,
         "eta": _inn(get("eta")), "P": _inn(get("P")), "tII": _inn(get("tII")),
         "VR": (VR[2:-1, 1:-1, 1:-1] + VR[1:-2, 1:-1, 1:-1]) * 0.5,
         "VT": (VT[1:-1, 2:-1, 1:-1] + VT[1:-1, 1:-2, 1:-1]) * 0.5,
         "VF": (VP[1:-1, 1:-1, 2:-1] + VP[1:-1, 1:-1, 1:-2]) * 0.5,
         "t_RR": _inn(get("tau_RR")), "t_TT": _inn(get("tau_thth")), "t_FF": _inn(get("tau_phph")),
         "t_RT": (((RT[:-1, :-1, :] + RT[1:, :-1, :]) + RT[:-1, 1:, :]) + RT[1:, 1:, :]) * 0.25,
         "τ_RF": (((RP[:-1, :, :-1] + RP[1:, :, :-1]) + RP[:-1, :, 1:]) + RP[1:, :, 1:]) * 0.25,
         "τ_TF": (((TP[:, :-1, :-1] + TP[:, 1:, :-1]) + TP[:, :-1, 1:]) + TP[:, 1:, 1:]) * 0.25}
    return {k: np.asfortranarray(v, dtype=np.float64) for k, v in d.items()}


class XdmfSeries:
    """A time series of saves: ``add(source, setup, t)`` writes
    ``<name>_<isave>.bin`` and rewrites ``<name>.xdmf`` over all saves so
    far (the reference's write_xdmf with h5_names and timesteps)."""

    def __init__(self, out, name: str = "results3D"):
        self.out = Path(out)
        self.out.mkdir(parents=True, exist_ok=True)
        self.name = name
        self.saves = []   # (file name, time, dims, [(dataset, seek)])

    def add(self, source, setup, t: float) -> Path:
        d = xdmf_datasets(source, setup)
        dims = d["X"].shape
        fname = f"{self.name}_{len(self.saves) + 1}.bin"
        items, seek = [], 0
        with open(self.out / fname, "wb") as fh:
            for key in ["X", "Y", "Z"] + XDMF_FIELDS:
                a = d[key]
                assert a.shape == dims, (key, a.shape, dims)
                fh.write(a.astype("<f8").tobytes(order="F"))
                items.append((key, seek))
                seek += a.size * 8
        self.saves.append((fname, float(t), dims, items))
        return self.write_xdmf()

    def write_xdmf(self) -> Path:
        import xml.etree.ElementTree as ET
        root = ET.Element("Xdmf", Version="3.0")
        dom = ET.SubElement(root, "Domain")
        coll = ET.SubElement(dom, "Grid", GridType="Collection", CollectionType="Temporal")

        def item(parent, fname, dims, seek, **kw):
            di = ET.SubElement(parent, "DataItem", **kw, Format="Binary", NumberType="Float", Precision="8",
                               Endian="Little", Seek=str(seek), Dimensions=" ".join(str(n) for n in reversed(dims)))
            di.text = fname
            return di

        for fname, t, dims, items in self.saves:
            g = ET.SubElement(coll, "Grid", GridType="Uniform")
            ET.SubElement(g, "Topology", TopologyType="3DSMesh", Dimensions=" ".join(str(n) for n in reversed(dims)))
            ET.SubElement(g, "Time", Value=repr(t))
            geo = ET.SubElement(g, "Geometry", GeometryType="X_Y_Z")
            seeks = dict(items)
            for c in ("X", "Y", "Z"):
                item(geo, fname, dims, seeks[c], Name=c)
            for key in XDMF_FIELDS:
                att = ET.SubElement(g, "Attribute", Name=key, Center="Cell")
                item(att, fname, dims, seeks[key])
        path = self.out / f"{self.name}.xdmf"
        ET.ElementTree(root).write(path, encoding="utf-8", xml_declaration=True)
        return path


def read_xdmf_dataitem(xdmf_path, grid_index: int, name: str) -> np.ndarray:
    """Reads one dataset of a save back through its XDMF description
    (file, Seek, Dimensions): the geometry's X/Y/Z or an attribute."""
    import xml.etree.ElementTree as ET
    xdmf_path = Path(xdmf_path)
    grids = ET.parse(xdmf_path).getroot().find("Domain").find("Grid").findall("Grid")
    g = grids[grid_index]
    di = None
    for d in g.find("Geometry").findall("DataItem"):
        if d.get("Name") == name:
            di = d
    for a in g.findall("Attribute"):
        if a.get("Name") == name:
            di = a.find("DataItem")
    if di is None:
        raise KeyError(name)
    dims = tuple(int(x) for x in di.get("Dimensions").split())[::-1]
    n = int(np.prod(dims))
    with open(xdmf_path.parent / di.text.strip(), "rb") as fh:
        fh.seek(int(di.get("Seek")))
        a = np.frombuffer(fh.read(8 * n), dtype="<f8")
    return np.asfortranarray(a.reshape(dims, order="F"))
