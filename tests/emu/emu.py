"""ctypes binding of the host emulation of the fused kernel (test support)."""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

from sphericalstokes_amd._lib import ss_params
from sphericalstokes_amd.fields import FIELDS, STATE

EMU_LIB = Path(__file__).resolve().parent / "libsspt_emu.so"
SHEAR = ("tau_Rth", "tau_Rph", "tau_thph")
PINGPONG = set(STATE)

_L = None


def lib():
    global _L
    if _L is None:
        if not EMU_LIB.exists():
            raise RuntimeError(f"{EMU_LIB} not built")
        L = C.CDLL(str(EMU_LIB))
        DP = C.POINTER(C.c_double)
        L.emu_fused_iterations.restype = C.c_int
        L.emu_fused_iterations.argtypes = [C.c_int] * 4 + [DP] * 3 + [C.POINTER(ss_params), C.POINTER(DP),
                                           C.POINTER(DP), C.c_int, C.c_int, C.c_int,
                                           C.POINTER(C.c_ulonglong), C.c_longlong]
        L.emu_pow_ex.restype = None
        L.emu_pow_ex.argtypes = [DP, C.c_int, C.c_double, DP]
        _L = L
    return _L


def pitch_of(nph):
    return ((nph + 1) + 63) // 64 * 64


def to_box(setup, fields):
    nr, nt, np_ = setup.nr, setup.nth, setup.nph
    P = pitch_of(np_)
    out = {}
    for n in FIELDS:
        b = np.zeros((nr + 1, nt + 1, P))
        a = fields[n]
        o = 1 if n in SHEAR else 0
        b[o:o + a.shape[0], o:o + a.shape[1], o:o + a.shape[2]] = a
        out[n] = b
    return out


def from_box(setup, box, shapes):
    out = {}
    for n in FIELDS:
        o = 1 if n in SHEAR else 0
        s = shapes[n]
        out[n] = np.asfortranarray(box[n][o:o + s[0], o:o + s[1], o:o + s[2]])
    return out


def run(setup, fields, params, niter, check_last=False, rchunk=0, paired=False):
    """Runs the fused schedule on the host; returns (fields, red[6]).  With
    paired, each state field's two buffers are the halves of one array and
    the main kernel runs with F_PAIRED."""
    L = lib()
    cur = to_box(setup, fields)
    if paired:
        pairs = {n: np.zeros((2,) + cur[n].shape) for n in PINGPONG}
        for n in PINGPONG:
            pairs[n][0] = cur[n]
        alt = {n: (pairs[n][1] if n in PINGPONG else cur[n]) for n in FIELDS}
        cur = {n: (pairs[n][0] if n in PINGPONG else cur[n]) for n in FIELDS}
        pair_bytes = cur["P"].nbytes
    else:
        alt = {n: (np.zeros_like(cur[n]) if n in PINGPONG else cur[n]) for n in FIELDS}
        pair_bytes = 0
    DP = C.POINTER(C.c_double)
    pc = (DP * len(FIELDS))(*[cur[n].ctypes.data_as(DP) for n in FIELDS])
    pa = (DP * len(FIELDS))(*[alt[n].ctypes.data_as(DP) for n in FIELDS])
    red = (C.c_ulonglong * 6)()
    sp = ss_params.from_params(params)
    R, th, ph = (np.ascontiguousarray(x, dtype=np.float64) for x in (setup.R, setup.th, setup.ph))
    which = L.emu_fused_iterations(setup.nr, setup.nth, setup.nph, pitch_of(setup.nph), R.ctypes.data_as(DP),
                                   th.ctypes.data_as(DP), ph.ctypes.data_as(DP), C.byref(sp), pc, pa,
                                   int(niter), int(check_last), int(rchunk), red, pair_bytes)
    if which < 0:
        raise RuntimeError("state buffers are not paired")
    box = cur if which == 0 else alt
    shapes = {n: fields[n].shape for n in FIELDS}
    reds = np.array(list(red), dtype=np.uint64).view(np.float64)
    return from_box(setup, box, shapes), reds


ADJ = ["VR", "Vth", "Vph", "P", "tau_RR", "tau_thth", "tau_phph", "tau_Rth", "tau_Rph", "tau_thph",
       "eta", "eta_PL"]
ADJ_STATE = ["tau_RR", "tau_thth", "tau_phph", "tau_Rth", "tau_Rph", "tau_thph", "eta", "eta_ini", "eta_PL", "tII",
             "DT_R", "etaSM"]


def _box1(setup, name, a):
    nr, nt, np_ = setup.nr, setup.nth, setup.nph
    b = np.zeros((nr + 1, nt + 1, pitch_of(np_)))
    o = 1 if name in SHEAR else 0
    b[o:o + a.shape[0], o:o + a.shape[1], o:o + a.shape[2]] = a
    return b


def adj_step(setup, state, lam, params, rchunk=0, objective=None):
    """One application of the r-marching adjoint kernel on the host: state
    (forward fields by name), lam (adjoint fields by name) -> new lam; with
    objective = (r, squared radius, weights of VR's shape) the step also adds
    the weights on the r-faces inside the inclusion."""
    L = lib()
    DP = C.POINTER(C.c_double)
    if not hasattr(L, "_adj_sig"):
        L.emu_adj_step.restype = C.c_int
        L.emu_adj_step.argtypes = [C.c_int] * 4 + [DP] * 3 + [C.POINTER(ss_params), C.POINTER(DP), C.POINTER(DP),
                                                              C.POINTER(DP), C.c_int, DP, C.c_double, C.c_double]
        L._adj_sig = True
    sb = [_box1(setup, n, state[n]) for n in ADJ_STATE]
    li = [_box1(setup, n, lam[n]) for n in ADJ]
    lo = [np.zeros_like(x) for x in li]
    ps = (DP * len(sb))(*[x.ctypes.data_as(DP) for x in sb])
    pi = (DP * len(li))(*[x.ctypes.data_as(DP) for x in li])
    po = (DP * len(lo))(*[x.ctypes.data_as(DP) for x in lo])
    sp = ss_params.from_params(params)
    R, th, ph = (np.ascontiguousarray(x, dtype=np.float64) for x in (setup.R, setup.th, setup.ph))
    bj, r_obj, rad = None, 0.0, 0.0
    if objective is not None:
        r_obj, rad, w = objective
        bj = _box1(setup, "VR", w)
    L.emu_adj_step(setup.nr, setup.nth, setup.nph, pitch_of(setup.nph), R.ctypes.data_as(DP), th.ctypes.data_as(DP),
                   ph.ctypes.data_as(DP), C.byref(sp), ps, pi, po, int(rchunk),
                   bj.ctypes.data_as(DP) if bj is not None else None, float(r_obj), float(rad))
    out = {}
    for n, x in zip(ADJ, lo):
        o = 1 if n in SHEAR else 0
        s = lam[n].shape
        out[n] = np.asfortranarray(x[o:o + s[0], o:o + s[1], o:o + s[2]])
    return out



def adj_step_rhs(setup, state, lam, params, rhs, f32=False, rchunk=0):
    """lam_out = T^T lam + rhs (every output) with the RHS instances of the
    r-marching adjoint kernel on the host, in FP64 or (f32) with state, lam,
    rhs and the tables rounded to FP32 and FP32 arithmetic."""
    L = lib()
    VP = C.c_void_p
    DP = C.POINTER(C.c_double)
    if not hasattr(L, "_adj_rhs_sig"):
        L.emu_adj_step_rhs.restype = C.c_int
        L.emu_adj_step_rhs.argtypes = [C.c_int] * 5 + [DP] * 3 + [C.POINTER(ss_params), C.POINTER(VP), C.POINTER(VP),
                                                                  C.POINTER(VP), C.c_int, C.POINTER(VP)]
        L._adj_rhs_sig = True
    dt = np.float32 if f32 else np.float64
    sb = [_box1(setup, n, state[n]).astype(dt) for n in ADJ_STATE]
    li = [_box1(setup, n, lam[n]).astype(dt) for n in ADJ]
    rb = [_box1(setup, n, rhs[n]).astype(dt) for n in ADJ]
    lo = [np.zeros_like(x) for x in li]
    arr = lambda xs: (VP * len(xs))(*[x.ctypes.data for x in xs])
    sp = ss_params.from_params(params)
    R, th, ph = (np.ascontiguousarray(x, dtype=np.float64) for x in (setup.R, setup.th, setup.ph))
    L.emu_adj_step_rhs(int(f32), setup.nr, setup.nth, setup.nph, pitch_of(setup.nph), R.ctypes.data_as(DP),
                       th.ctypes.data_as(DP), ph.ctypes.data_as(DP), C.byref(sp), arr(sb), arr(li), arr(lo),
                       int(rchunk), arr(rb))
    out = {}
    for n, x in zip(ADJ, lo):
        o = 1 if n in SHEAR else 0
        s = lam[n].shape
        out[n] = np.asfortranarray(x[o:o + s[0], o:o + s[1], o:o + s[2]].astype(np.float64))
    return out

def pow_ex(x, ex):
    """The fused kernels' x^ex (sspt_fused_body.h pow_ex), run on the host."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    DP = C.POINTER(C.c_double)
    lib().emu_pow_ex(x.ctypes.data_as(DP), x.size, float(ex), out.ctypes.data_as(DP))
    return out
