"""GPU: ss_check_err_noalloc (no allocation, static device scratch) against
ss_check_err (context scratch) on the same data: the three component
errors, err and the V - V_prev arrays bit for bit, on calls in a row with
shrinking inputs and a NaN case, through the C ABI.  Device arrays are
allocated here with the HIP runtime that libsspt.so itself uses (looked up
through its handle), also when torch's copy of the runtime is mapped too."""
import ctypes as C
import sys

import numpy as np
import pytest

from sphericalstokes_amd import setups
from sphericalstokes_amd.fields import FIELDS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    H2D, D2H = 1, 2

    def __init__(self):
        # the HIP runtime libsspt.so itself uses: symbols looked up through
        # libsspt.so's handle resolve in its own dependency tree, whatever
        # other copy of libamdhip64 (torch's) the process also maps
        from sphericalstokes_amd import _lib
        self.L = _lib.lib()
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def ok(self, rc, what):
        assert rc == 0, (what, rc)

    def put(self, a):
        a = np.ascontiguousarray(a.ravel(order="F"))
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.ok(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy H2D")
        return p, a.size

    def get(self, p, n, shape=None):
        out = np.empty(n, dtype=np.float64)
        self.ok(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H), "hipMemcpy D2H")
        return out if shape is None else np.asfortranarray(out.reshape(shape, order="F"))

    def free(self, p):
        self.ok(self.L.hipFree(p), "hipFree")


def _case(shape, seed, nan=False):
    st = setups.diapir_sector(shape[0], nth=shape[1], nph=shape[2], n_exp=1.0)
    rng = np.random.default_rng(seed)
    f = st.fields
    for k in ("VR", "Vth", "Vph"):
        f[k][...] = rng.standard_normal(f[k].shape) * (1.0 + 2.0 * (k == "Vth"))
        f[k + "_err"][...] = f[k] + 1e-3 * rng.standard_normal(f[k].shape)
    if nan:
        f["Vph"][1, 2, 3] = np.nan
    return st


@pytest.mark.parametrize("drop", [0, 1])
def test_noalloc_check_err_matches_context_check(drop):
    from sphericalstokes_amd import _lib
    from sphericalstokes_amd.solver import Solver
    L = _lib.lib()
    hip = Hip()
    for shape, seed, nan in (((21, 40, 70), 1, False), ((13, 17, 33), 2, False), ((9, 11, 66), 3, True)):
        say("noalloc check_err", shape, "nan" if nan else "", "drop", drop)
        st = _case(shape, seed, nan)
        with Solver(*shape) as s:
            s.load_setup(st)
            err, comp = s.check_err(drop_nan_vph=bool(drop))
            ref = {k: s.download(k) for k in ("VR_err", "Vth_err", "Vph_err")}
        names = ("VR", "Vth", "Vph", "VR_err", "Vth_err", "Vph_err")
        dev = {k: hip.put(st.fields[k]) for k in names}
        out, _ = hip.put(np.full(4, 7.0))
        rc = L.ss_check_err_noalloc(*[dev[k][0] for k in names], *shape, drop, out, None)
        assert rc == 0, rc
        hip.ok(hip.L.hipDeviceSynchronize(), "sync")
        got = hip.get(out, 4)
        say("  context", comp, err, " noalloc", got)
        for q in range(3):
            assert (np.isnan(got[q]) and np.isnan(comp[q])) or got[q] == comp[q], q
        assert (np.isnan(got[3]) and np.isnan(err)) or got[3] == err
        for k in ("VR_err", "Vth_err", "Vph_err"):
            g = hip.get(dev[k][0], dev[k][1], st.fields[k].shape)
            assert np.array_equal(g, ref[k], equal_nan=True), k
        for p, _ in dev.values():
            hip.free(p)
        hip.free(out)
