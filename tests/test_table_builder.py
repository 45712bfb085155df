"""The library's table-grid search (mcd_grid_search: host C++, FP64 adaptive
Gauss-Kronrod reference integral) against the search restated in Python with
scipy's quad (mcd.tables, probTables.py:165-411): the same float32 grids, bit
for bit, for the test and benchmark configurations, and the grid shapes the
reference's search gives for C1 and C2.  CPU only: the search uses no device."""
import ctypes as C

import numpy as np
import pytest

from mcd import _abi
from mcd._lib import MCDError, check, lib
from mcd.material import Material
from mcd.tables import gryz_grid, moller_grid

CASES = [("Si", 10000.0, 1000.0, 100.0, 1e-2, 1e-2),      # C1
         ("Cu", 20000.0, 1000.0, 50.0, 1e-3, 1e-3),       # C2 and the differentiation workload
         ("Al", 20000.0, 15000.0, 10.0, 1e-4, 1e-7),      # the reference's test input
         ("Cu", 20000.0, 1000.0, 20.0, 1e-3, 1e-3),       # Wc range ends
         ("Cu", 20000.0, 1000.0, 200.0, 1e-3, 1e-3),
         ("Cu", 20000.0, 1000.0, 202.0, 1e-3, 1e-3)]      # E grid 2000 -> 1000 between 200 and 202 eV
CASES += [(m, E0, 500.0, 100.0, 1e-3, 1e-3) for m in ("Al", "Si", "Cu", "Au")
          for E0 in (5000.0, 10000.0, 20000.0, 30000.0)]  # the batched workload's DS tables


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]:g}-{c[2]:g}-Wc{c[3]:g}-{c[4]:g}-{c[5]:g}")
def test_library_search_equals_python_search(case):
    mat, E0, Emin, Wc, tE, tW = case
    m = Material(mat)
    for make in (lambda: moller_grid(m, E0, Emin, Wc), lambda: gryz_grid(m, E0, Emin)):
        py = make().build_grids(tE, tW)
        cc = make().build_grids_library(tE, tW)
        assert py.shape == cc.shape, (py.kind, py.shape, cc.shape)
        assert np.array_equal(py.Ws, cc.Ws) and np.array_equal(py.Es, cc.Es)


def test_library_search_gives_the_reference_shapes():
    """(nE, nW - 1) of the CDF tables: Si 10 keV tol 1e-2 Moller 300 x 29,
    Gryzinski[2p] 400 x 39; Cu 20 keV tol 1e-3 2000 x 199 for both."""
    si, cu = Material("Si"), Material("Cu")
    assert moller_grid(si, 10000.0, 1000.0, 100.0).build_grids_library(1e-2, 1e-2).shape == (300, 29)
    assert gryz_grid(si, 10000.0, 1000.0).build_grids_library(1e-2, 1e-2).shape == (400, 39)
    assert moller_grid(cu, 20000.0, 1000.0, 50.0).build_grids_library(1e-3, 1e-3).shape == (2000, 199)
    assert gryz_grid(cu, 20000.0, 1000.0).build_grids_library(1e-3, 1e-3).shape == (2000, 199)


def test_grid_search_argument_errors():
    nE, nW = C.c_int32(0), C.c_int32(0)
    L = lib()
    with pytest.raises(MCDError) as e:
        check(L.mcd_grid_search(0, 50.0, 1.0, 0.0, 7.0, 1000.0, 20000.0, 0.0, 1e-3, None, C.byref(nE), None,
                                C.byref(nW)))
    assert e.value.code == _abi.ERR_INVALID
    with pytest.raises(MCDError):
        check(L.mcd_grid_search(5, 50.0, 1.0, 0.0, 7.0, 1000.0, 20000.0, 1e-3, 1e-3, None, C.byref(nE), None,
                                C.byref(nW)))
    # sizes first, then a buffer too small for them
    check(L.mcd_grid_search(0, 50.0, 1.0, 0.0, 7.0, 1000.0, 20000.0, 1e-3, 1e-3, None, C.byref(nE), None,
                            C.byref(nW)))
    assert nE.value > 2 and nW.value > 2
    Es = np.zeros(nE.value, np.float32)
    Ws = np.zeros(2, np.float32)
    nW2 = C.c_int32(2)
    with pytest.raises(MCDError) as e:
        check(L.mcd_grid_search(0, 50.0, 1.0, 0.0, 7.0, 1000.0, 20000.0, 1e-3, 1e-3, Es.ctypes.data, C.byref(nE),
                                Ws.ctypes.data, C.byref(nW2)))
    assert e.value.code == _abi.ERR_INVALID and nW2.value == nW.value
