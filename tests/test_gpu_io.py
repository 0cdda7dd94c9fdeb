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
