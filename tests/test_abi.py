"""The C ABI: libsspt.so loads and exports every symbol include/*.h
declares; argument checks that need no GPU return SS_ERR_ARG."""
import ctypes as C
import re
from pathlib import Path

import pytest

from sphericalstokes_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    syms = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        txt = h.read_text()
        syms |= set(re.findall(r"^\s*(?:int|void|const char \*)\s*(ss_\w+)\s*\(", txt, flags=re.M))
    return sorted(syms)


def test_header_symbols_exported():
    L = C.CDLL(str(_lib.LIB_PATH))
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_create_rejects_bad_sizes_without_gpu():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.ss_create(2, 10, 10, 0, C.byref(h)) == _lib.SS_ERR_ARG
    assert L.ss_create(10, 10, 10, 0, None) == _lib.SS_ERR_ARG
    assert L.ss_destroy(None) == _lib.SS_ERR_ARG
    assert L.ss_last_error(None) == b"null context"
    assert L.ss_cart_create(2, 10, 10, 0, C.byref(h)) == _lib.SS_ERR_ARG
    assert L.ss_cart_create(10, 10, 10, 0, None) == _lib.SS_ERR_ARG
    assert L.ss_cart_last_error(None) == b"null context"
    L.ss_cart_destroy(None)
    assert L.ss_cart_maxloc(None) == _lib.SS_ERR_ARG
    assert L.ss_cyl_create(3, 10, 10, 0, C.byref(h)) == _lib.SS_ERR_ARG
    assert L.ss_cyl_create(10, 10, 10, 0, None) == _lib.SS_ERR_ARG
    assert L.ss_cyl_last_error(None) == b"null context"
    L.ss_cyl_destroy(None)
    assert L.ss_cyl_compute_V(None) == _lib.SS_ERR_ARG


def test_params_struct_layout():
    # ss_params: 14 doubles + 3 ints, padded to 8 bytes
    assert C.sizeof(_lib.ss_params) == 14 * 8 + 4 * 4
    assert C.sizeof(_lib.ss_result) == 4 + 4 + 8 + 24 + 8
    assert C.sizeof(_lib.ss_cart_params) == 11 * 8
    assert C.sizeof(_lib.ss_cart_result) == 4 + 4 + 8 + 8
    assert C.sizeof(_lib.ss_cyl_params) == 13 * 8
    assert C.sizeof(_lib.ss_cyl_result) == 4 + 4 + 8 + 8 + 8


def test_check_err_noalloc_rejects_without_gpu():
    """Argument checks of the allocation-free check run before any HIP call."""
    L = _lib.lib()
    dummy = C.c_void_p(8)
    args = [dummy] * 6
    assert L.ss_check_err_noalloc(None, *args[1:], 10, 10, 10, 0, dummy, None) == _lib.SS_ERR_ARG
    assert L.ss_check_err_noalloc(*args, 10, 10, 10, 0, None, None) == _lib.SS_ERR_ARG
    assert L.ss_check_err_noalloc(*args, 0, 10, 10, 0, dummy, None) == _lib.SS_ERR_ARG
    # 700^3 cells: beyond the scratch's 268 435 456 elements per component
    assert L.ss_check_err_noalloc(*args, 700, 700, 700, 0, dummy, None) == _lib.SS_ERR_ARG
