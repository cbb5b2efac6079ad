"""C ABI checks that need no GPU: the library loads and exports every declared symbol."""

import ctypes as C
import os
import re

import pytest

from mhdflows_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mhdflows_amd.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mhd_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    decl = declared_symbols()
    assert len(decl) >= 30
    assert set(decl) == set(_lib.SIGNATURES), set(decl) ^ set(_lib.SIGNATURES)


def test_library_exports_all_symbols():
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("library not built")
    L = C.CDLL(_lib.LIBPATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_pure_entry_points_without_gpu():
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("library not built")
    L = _lib.lib()
    assert L.mhd_abi_version() == 1
    for n in (8, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024):
        assert L.mhd_supported_length(n) == 1
    for n in (0, 10, 33, 2048):
        assert L.mhd_supported_length(n) == 0
    cfg = _lib.Config(nx=256, ny=256, nz=256, Lx=6.283185307179586, Ly=6.283185307179586,
                      Lz=6.283185307179586, nu=1e-3, eta=1e-3, precision=0, physics=1, stepper=0)
    b = C.c_uint64()
    assert L.mhd_query_bytes(C.byref(cfg), C.byref(b)) == 0
    Vc = 8 * 129 * 256 * 256
    Vr = 4 * 256 ** 3
    assert b.value >= 27 * Vc + 6 * Vr
    assert abs(L.mhd_step_bytes_model(C.byref(cfg)) - (438 * Vc + 120 * Vr)) < 1
    bad = _lib.Config(nx=33, ny=32, nz=32, Lx=1, Ly=1, Lz=1)
    assert L.mhd_query_bytes(C.byref(bad), C.byref(b)) == -1
    assert b"even" in L.mhd_last_error()
    big = _lib.Config(nx=2048, ny=32, nz=32, Lx=1, Ly=1, Lz=1)
    assert L.mhd_query_bytes(C.byref(big), C.byref(b)) == -2
