"""Serving-loop admission rules that need no GPU: requests that can never be admitted are
refused on arrival (recorded as failed results) instead of waiting forever."""

import os

import numpy as np
import pytest

from mhdflows_amd import _lib, serve


@pytest.mark.timeout(60)
def test_server_refuses_what_can_never_fit():
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("library not built")
    reqs = [serve.Request(0, 64, np.float32, 10, 1e-3, 1e-3, 1, 0.0),   # pool larger than the budget
            serve.Request(1, 10, np.float32, 10, 1e-3, 1e-3, 2, 0.0),   # no compiled transform
            serve.Request(2, 24, np.float64, 10, 1e-3, 1e-3, 3, 0.0)]   # disabled row kernel
    srv = serve.Server(hbm_budget=1 << 20, slots={64: 1})
    res = srv.run(reqs, realtime=False)      # returns: no pool is ever created
    srv.close()
    assert sorted(res) == [0, 1, 2]
    assert "budget" in res[0].error
    assert "compiled transform" in res[1].error
    assert "nx = 24" in res[2].error
    s = serve.summarize(res, 1.0)
    assert s["requests"] == 0 and s["refused"] == 3


def test_disabled_row_kernel_is_refused_at_validation():
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("library not built")
    import ctypes as C
    L = _lib.lib()
    cfg = _lib.Config(nx=24, ny=24, nz=24, Lx=1, Ly=1, Lz=1, precision=1, physics=1)
    b = C.c_uint64()
    assert L.mhd_query_bytes(C.byref(cfg), C.byref(b)) == -2
    for prec, phys, ny in ((0, 1, 24), (1, 0, 24), (1, 1, 48)):
        ok = _lib.Config(nx=24 if ny == 24 else 48, ny=ny, nz=24, Lx=1, Ly=1, Lz=1, precision=prec, physics=phys)
        assert L.mhd_query_bytes(C.byref(ok), C.byref(b)) == 0
