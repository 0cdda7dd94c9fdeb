"""CPU: the field dump in the reference's output format
(SphericalStokes_diapir.jl:23-43, :455-481) and its restore."""
import numpy as np

from sphericalstokes_amd import io, setups
from sphericalstokes_amd.fields import shape


def test_dump_layout_and_round_trip(tmp_path):
    st = setups.diapir_sector(9, nth=7, nph=11, n_exp=3.0)
    setups.random_state(st, seed=4)
    f = st.fields
    io.save_reference_format(tmp_path, f, st, num=1, eps=1e-8, runtime=12.7, err_evo=[1.0, 0.5],
                             iter_evo=[100, 200])
    # file set and sizes: raw FP64, column-major (r fastest)
    for name in list(io.ARRAYS) + list(io.EXTRA) + ["R", "TH", "PH"]:
        assert (tmp_path / f"A_{name}.bin").exists(), name
    assert (tmp_path / "A_VR.bin").stat().st_size == 10 * 7 * 11 * 8
    assert (tmp_path / "A_TAU_RTH.bin").stat().st_size == int(np.prod(shape("tau_Rth", 9, 7, 11))) * 8
    vr = np.fromfile(tmp_path / "A_VR.bin").reshape((10, 7, 11), order="F")
    assert np.array_equal(vr, f["VR"])
    r3 = np.fromfile(tmp_path / "A_R.bin").reshape((9, 7, 11), order="F")
    assert np.array_equal(r3[:, 3, 5], st.R) and np.all(r3 == st.R[:, None, None])
    assert np.array_equal(np.fromfile(tmp_path / "A_iter_evo.bin"), [100.0, 200.0])
    # the .inf lines, read the way visu_diapir.m reads them (numbers split on
    # blanks and commas)
    info = [float(x) for x in (tmp_path / "1_infos.inf").read_text().split()]
    assert info[0] == 1 and abs(info[1] - 1.0) < 1e-6 and info[4:7] == [9, 7, 11]
    phys = [float(x) for x in (tmp_path / "1_phys.inf").read_text().replace(",", " ").split()]
    assert len(phys) == 12 and phys[4] == 3.0 and abs(phys[10] - 0.01) < 1e-6 and phys[11] == 4.5
    # read back
    dims, (R, th, ph), g = io.load_reference_format(tmp_path)
    assert dims == (9, 7, 11)
    assert np.array_equal(R, st.R) and np.array_equal(th, st.th) and np.array_equal(ph, st.ph)
    for name in list(io.ARRAYS.values()) + list(io.EXTRA.values()):
        assert np.array_equal(g[name], f[name]), name


class _FakeSolver:
    def __init__(self):
        self.up = {}

    def upload(self, name, a):
        self.up[name] = np.array(a)


def test_restore_from_a_reference_dump_without_eta_ini(tmp_path):
    st = setups.diapir_sector(7, n_exp=1.0)
    io.save_reference_format(tmp_path, st.fields, st)
    (tmp_path / "A_ETA_INI.bin").unlink()
    (tmp_path / "A_ETA_PL.bin").unlink()
    _, _, g = io.load_reference_format(tmp_path)
    s = _FakeSolver()
    io.restore(s, g)
    assert np.array_equal(s.up["eta_ini"], st.fields["eta"]) and np.array_equal(s.up["eta_PL"], st.fields["eta"])
    assert set(io.ARRAYS.values()) <= set(s.up)
