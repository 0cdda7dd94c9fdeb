"""XDMF output (the reference's visualisation save, SphericalStokes_mxpu.jl
:443-464 with data_io2.jl:14-118): XML layout, byte offsets and the data
read back through the description, on the CPU."""
import math
import xml.etree.ElementTree as ET

import numpy as np

from sphericalstokes_amd import io, setups


def _state(seed=0):
    st = setups.diapir_sector(9, nth=7, nph=11, n_exp=1.0)
    setups.random_state(st, seed=seed)
    return st


def test_sph2cart_maps_radius_and_equator():
    R = np.array([1.0, 1.5, 2.0])
    th = np.linspace(math.pi / 3, 2 * math.pi / 3, 5)
    ph = np.linspace(-math.pi / 6, math.pi / 6, 7)
    X, Y, Z = io.sph2cart(R, th, ph)
    # |x| = R at every point
    assert np.allclose(X * X + Y * Y + Z * Z, (R[:, None, None] * np.ones(X.shape)) ** 2, rtol=1e-14)
    # th = pi/2, ph = 0 lies on the X axis; ph > 0 gives Z > 0; th > pi/2 gives Y < 0
    assert np.allclose([X[1, 2, 3], Y[1, 2, 3], Z[1, 2, 3]], [1.5, 0.0, 0.0], atol=1e-15)
    assert np.all(Z[:, :, 4:] > 0) and np.all(Y[:, 3:, :] < 0)


def test_xdmf_series_layout_and_readback(tmp_path):
    sts = [_state(0), _state(1)]
    series = io.XdmfSeries(tmp_path, "results3D")
    for q, st in enumerate(sts):
        path = series.add(st.fields, st, t=float(q + 1))
    assert path.name == "results3D.xdmf"
    root = ET.parse(path).getroot()
    assert root.tag == "Xdmf" and root.get("Version") == "3.0"
    coll = root.find("Domain").find("Grid")
    assert coll.get("GridType") == "Collection" and coll.get("CollectionType") == "Temporal"
    grids = coll.findall("Grid")
    assert len(grids) == 2
    nr, nt, np_ = 9, 7, 11
    dims = f"{np_ - 2} {nt - 2} {nr - 2}"          # slowest first: the Julia order reversed
    for q, g in enumerate(grids):
        assert g.find("Topology").get("TopologyType") == "3DSMesh"
        assert g.find("Topology").get("Dimensions") == dims
        assert float(g.find("Time").get("Value")) == q + 1
        geo = g.find("Geometry")
        assert geo.get("GeometryType") == "X_Y_Z"
        assert [d.get("Name") for d in geo.findall("DataItem")] == ["X", "Y", "Z"]
        assert [a.get("Name") for a in g.findall("Attribute")] == io.XDMF_FIELDS
        seeks = [int(d.get("Seek")) for d in geo.findall("DataItem")] + \
                [int(a.find("DataItem").get("Seek")) for a in g.findall("Attribute")]
        n = (nr - 2) * (nt - 2) * (np_ - 2)
        assert seeks == [8 * n * k for k in range(15)]
        fname = geo.find("DataItem").text
        assert fname == f"results3D_{q + 1}.bin"
        assert (tmp_path / fname).stat().st_size == 15 * 8 * n
    # read back through the description and compare with the definitions
    for q, st in enumerate(sts):
        f = st.fields
        X, Y, Z = io.sph2cart(st.R[1:-1], st.th[1:-1], st.ph[1:-1])
        exp = {"X": X, "Y": Y, "Z": Z, "eta": f["eta"][1:-1, 1:-1, 1:-1], "P": f["P"][1:-1, 1:-1, 1:-1],
               "tII": f["tII"][1:-1, 1:-1, 1:-1], "t_RR": f["tau_RR"][1:-1, 1:-1, 1:-1]}
        VR = f["VR"]
        i, j, k = 3, 2, 4     # inner cell (i+1, j+1, k+1) of the full grid
        exp_vr = 0.5 * (VR[i + 2, j + 1, k + 1] + VR[i + 1, j + 1, k + 1])
        RT = f["tau_Rth"]
        exp_rt = (((RT[i, j, k] + RT[i + 1, j, k]) + RT[i, j + 1, k]) + RT[i + 1, j + 1, k]) * 0.25
        for name, a in exp.items():
            assert np.array_equal(io.read_xdmf_dataitem(path, q, name), a), (q, name)
        assert io.read_xdmf_dataitem(path, q, "VR")[i, j, k] == exp_vr
        assert io.read_xdmf_dataitem(path, q, "t_RT")[i, j, k] == exp_rt
        for name in io.XDMF_FIELDS:
            assert io.read_xdmf_dataitem(path, q, name).shape == (nr - 2, nt - 2, np_ - 2)
