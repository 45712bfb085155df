"""mcd_forward_noalloc, the entry point that allocates nothing (its deferred
guard parks electrons in the library's static device scratch), against
mcd_forward through the C-ABI: two calls in a row, the second with fewer
electrons, give mcd_forward's records and counters byte for byte."""
import ctypes as C

import numpy as np
import pytest
import torch

from mcd import _abi, engine
from mcd._lib import MCDError, check, lib

pytestmark = pytest.mark.gpu


def _call(fn, tr, n, first, seed, flags=0, scats=None):
    dev = torch.device("cuda", 0)
    els = torch.zeros(max(n, 1) * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
    la = _abi.Launch(0, 0, _abi.PREC_FP32, flags)
    sb = None
    rows = None
    if scats:
        rows = torch.zeros(scats * _abi.SCATTER_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        sb = _abi.ScatterBuf(rows.data_ptr(), scats, first + 8)
    ctx = engine.context(0)
    s = torch.cuda.current_stream(0).cuda_stream
    check(fn(ctx, C.byref(tr.cfg), tr._tables, seed, 0, first, n, els.data_ptr(),
             C.byref(sb) if sb is not None else None, cnt.data_ptr(), C.byref(la), s), ctx)
    torch.cuda.synchronize(0)
    c = cnt.cpu().numpy().view(np.uint64).copy()
    r = None
    if rows is not None:
        used = int(c[_abi.CNT_SCAT_ROWS])
        r = rows.cpu().numpy()[: used * _abi.SCATTER_DTYPE.itemsize].view(_abi.SCATTER_DTYPE)
        r = np.sort(r, order=["electron", "step"])
    return els.cpu().numpy()[: n * _abi.ELECTRON_DTYPE.itemsize].tobytes(), c, r


# counters that describe the run (the internal work counters may differ between forms)
PUBLIC = [_abi.CNT_EVENTS, _abi.CNT_ELECTRONS, _abi.CNT_BKS, _abi.CNT_ABS, _abi.CNT_TRSM, _abi.CNT_LONG,
          _abi.CNT_FP64_PICKS, _abi.CNT_SCAT_ROWS, _abi.CNT_GUARD_TIMEOUTS]


@pytest.mark.parametrize("material,E0,Emin,mode", [("Cu", 20000.0, 1000.0, "DS"), ("Au", 30000.0, 1000.0, "DS"),
                                                    ("Au", 30000.0, 1000.0, "cont")])
def test_noalloc_forward_equals_forward_on_two_calls(material, E0, Emin, mode):
    tr = engine.Transport(material, E0, Emin, Wc=50.0, s_tilt=70.0, mode=mode, bethe_model="JL", tol_E=1e-2,
                          tol_W=1e-2)
    L = lib()
    for n, first, seed in (((1 << 18) + 3, 5, 11), (20000, 1 << 33, 12)):  # the second call is smaller
        print(f"noalloc {material} {mode} n={n} first={first}", flush=True)
        a_els, a_cnt, _ = _call(L.mcd_forward_noalloc, tr, n, first, seed)
        b_els, b_cnt, _ = _call(L.mcd_forward, tr, n, first, seed, flags=_abi.LAUNCH_DEFER_GUARD)
        assert a_els == b_els
        assert np.array_equal(a_cnt[PUBLIC], b_cnt[PUBLIC])
        assert a_cnt[_abi.CNT_GUARD_TIMEOUTS] == 0
        assert int(a_cnt[_abi.CNT_ELECTRONS]) == n


def test_noalloc_forward_scatter_rows_and_size_limit():
    tr = engine.Transport("Si", 10000.0, 1000.0, Wc=100.0, s_tilt=70.0, tol_E=1e-2, tol_W=1e-2)
    L = lib()
    for n in (4096, 300):
        a_els, a_cnt, a_rows = _call(L.mcd_forward_noalloc, tr, n, 0, 3, scats=1 << 16)
        b_els, b_cnt, b_rows = _call(L.mcd_forward, tr, n, 0, 3, flags=_abi.LAUNCH_DEFER_GUARD, scats=1 << 16)
        assert a_els == b_els and a_rows.tobytes() == b_rows.tobytes()
    with pytest.raises(MCDError, match="static scratch"):
        _call(L.mcd_forward_noalloc, tr, _abi.NOALLOC_MAX_ELECTRONS + 1, 0, 3)
