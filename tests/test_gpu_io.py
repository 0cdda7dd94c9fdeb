"""GPU: resume through the reference-format dump -- iterate, dump, restore
into a new context, iterate on: the same fields, bit for bit, as iterating
on without the round trip."""
import sys

import numpy as np
import pytest

from sphericalstokes_amd import io, setups
from sphericalstokes_amd.fields import FIELDS, STATE

pytestmark = pytest.mark.gpu


def test_dump_restore_resumes_bit_for_bit(tmp_path):
    from sphericalstokes_amd.solver import Solver
    st = setups.diapir_sector(15, n_exp=3.0, relax=0.5)
    p = st.params
    s = Solver(15, 15, 15)
    s.set_grid(st.R, st.th, st.ph)
    for n in FIELDS:
        s.upload(n, st.fields[n])
    s.pt_iterate(p, 40)
    io.save_reference_format(tmp_path, s, st)
    s.pt_iterate(p, 40)
    ref = s.download_all()
    s.close()
    dims, (R, th, ph), g = io.load_reference_format(tmp_path)
    s2 = Solver(*dims)
    s2.set_grid(R, th, ph)
    s2.upload("rho", g["rho"])
    io.restore(s2, g)
    s2.pt_iterate(p, 40)
    got = s2.download_all()
    s2.close()
    print("resumed", flush=True)
    sys.stdout.flush()
    for n in STATE + ["eta_PL"]:
        assert np.array_equal(got[n], ref[n]), n


def test_dump_matches_oracle_dump_byte_for_byte(tmp_path):
    """The reference-format folder written from the GPU state and from the
    oracle's state after the same 50 linear iterations (n_exp = 1, where the
    iteration is bitwise) are byte-identical, file by file
    (SphericalStokes_diapir.jl:23-43, :455-481)."""
    from oracle.binding import Oracle
    from sphericalstokes_amd.solver import Solver
    st = setups.diapir_sector(23, nth=19, nph=29, n_exp=1.0)
    p = st.params
    o = Oracle(st)
    o.step(50)
    with Solver(st.nr, st.nth, st.nph) as s:
        s.load_setup(st)
        s.pt_iterate(p, 50)
        io.save_reference_format(tmp_path / "gpu", s, st, err_evo=[1e-3, 2e-4], iter_evo=[100, 200])
    io.save_reference_format(tmp_path / "cpu", o.f, st, err_evo=[1e-3, 2e-4], iter_evo=[100, 200])
    names = sorted(q.name for q in (tmp_path / "gpu").iterdir())
    assert names == sorted(q.name for q in (tmp_path / "cpu").iterdir())
    assert len(names) >= 20
    for name in names:
        a, b = (tmp_path / "gpu" / name).read_bytes(), (tmp_path / "cpu" / name).read_bytes()
        assert a == b, name
    print("dumps identical:", len(names), "files", flush=True)
