"""Host-side drop-in interface on the CPU: input parsing, output columns, the
store name and the C-ABI symbol table (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

from mcd import _abi
from mcd.fileTools import els_columns, els_rows, readInput, zipDict

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_read_reference_input_files():
    p = readInput(os.path.join(HERE, "golden", "input.file"))
    assert p["num_el"] == 10 and p["material"] == "Si" and p["mode"] == "DS" and p["Wc"] == 10.0
    assert p["tol_E"] == 1e-4 and p["tol_W"] == 1e-7 and p["maxScatt"] == 1000 and p["thickness"] == 3e6
    assert p["electron_output"] == ["outcome", "final_E", "final_dir", "last_pos"]
    assert p["scatter_output"] == ["type", "position", "energy"]
    assert p["diffMFP"] is False and p["gen_tables"] is False
    t = readInput(os.path.join(HERE, "golden", "testInput.file"))
    assert t["material"] == "Al" and t["Emin"] == 15000.0 and t["scatter_output"][-1] == "az_angle"


def test_read_seed_and_maxscatt_off(tmp_path):
    f = tmp_path / "in.file"
    f.write_text("num_el : 8\nmaxScatt : off\nseed : 42  # Philox key\nE0 : 20000\n")
    p = readInput(str(f))
    assert p == {"num_el": 8, "maxScatt": 0, "seed": 42, "E0": 20000.0}


def test_els_rows_expand_exit_then_absorb():
    """An exit step that also absorbs gives two els rows (electron.py:52-54,68-70)."""
    e = np.zeros(3, dtype=_abi.ELECTRON_DTYPE)
    e["outcome"] = [_abi.OUT_BKS, _abi.OUT_ABS, _abi.OUT_BKS]
    e["outcome2"] = [0, 0, _abi.OUT_ABS]
    e["final_E"] = [1.0, 2.0, 3.0]
    rows, idx = els_rows(e)
    assert list(idx) == [0, 1, 2, 2]
    cols = els_columns(e, ["outcome", "final_E", "final_dir", "last_pos"])
    assert list(cols["outcome"]) == ["bks", "abs", "bks", "abs"]
    assert list(cols["final_E"]) == [1.0, 2.0, 3.0, 3.0]
    assert list(cols) == ["outcome", "final_E", "dx", "dy", "dz", "last_x", "last_y", "last_z"]


def test_zipdict_flattens_nested_columns():
    a = {}
    zipDict(a, {"outcome": ["bks"], "final_dir": {"dx": [1.0], "dy": [0.0], "dz": [0.0]}})
    zipDict(a, {"outcome": ["abs"], "final_dir": {"dx": [0.5], "dy": [0.5], "dz": [0.0]}})
    assert a == {"outcome": ["bks", "abs"], "dx": [1.0, 0.5], "dy": [0.0, 0.5], "dz": [0.0, 0.0]}


def test_store_name_follows_reference():
    import doScatter
    p = readInput(os.path.join(HERE, "golden", "input.file"))
    assert doScatter.store_name(p) == ("TRSM_diff:False_thick:3000000.0_mat:Si_mode:DS_elastic:Ruth_vanilla"
                                       "_tilt:40.0_Emin:10000.0_E0:20000.0_tolE:0.0001_tolW:1e-07.h5")


def test_library_exports_every_declared_symbol():
    """libmcd.so loads (no GPU call) and exports each function of include/mcd.h."""
    from mcd import _lib
    hdr = open(os.path.join(ROOT, "include", "mcd.h")).read()
    names = set(re.findall(r"^int (mcd_\w+)\(", hdr, re.M))
    assert {"mcd_forward", "mcd_tables_create", "mcd_ctx_create"} <= names
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n
    assert L.mcd_version() == 1


def test_python_constants_match_the_header():
    """The ctypes mirror (_abi) uses the header's values for launch flags,
    record flags and counter slots."""
    from mcd import _abi
    hdr = open(os.path.join(ROOT, "include", "mcd.h")).read()
    vals = {k: int(v) for k, v in re.findall(r"#define (MCD_LAUNCH_\w+) (\d+)", hdr)}
    vals.update({k: int(v) for k, v in re.findall(r"\b(MCD_(?:FLAG|CNT)_\w+) = (\d+)", hdr)})
    want = {"MCD_LAUNCH_INLINE_GUARD": _abi.LAUNCH_INLINE_GUARD, "MCD_LAUNCH_DEFER_GUARD": _abi.LAUNCH_DEFER_GUARD,
            "MCD_LAUNCH_SERVICE_GUARD": _abi.LAUNCH_SERVICE_GUARD, "MCD_FLAG_FP64_PICK": _abi.FLAG_FP64_PICK,
            "MCD_FLAG_GUARD_TIMEOUT": _abi.FLAG_GUARD_TIMEOUT, "MCD_CNT_EVENTS": _abi.CNT_EVENTS,
            "MCD_CNT_FP64_PICKS": _abi.CNT_FP64_PICKS, "MCD_CNT_QUEUE": _abi.CNT_QUEUE,
            "MCD_CNT_GUARD_TIMEOUTS": _abi.CNT_GUARD_TIMEOUTS, "MCD_CNT_PARKED": _abi.CNT_PARKED,
            "MCD_CNT_RESUMED": _abi.CNT_RESUMED, "MCD_LAUNCH_REFILL_SHIFT": _abi.LAUNCH_REFILL_SHIFT}
    for k, v in want.items():
        assert vals[k] == v, k


def test_refill_period_flags():
    from mcd import _abi
    from mcd.engine import refill_flags
    assert refill_flags(0) == 0 and refill_flags(1) == 0
    assert refill_flags(4) == 2 << _abi.LAUNCH_REFILL_SHIFT
    assert refill_flags(128) & ~_abi.LAUNCH_REFILL_MASK == 0
    for bad in (3, 256):
        with pytest.raises(ValueError):
            refill_flags(bad)


def test_unsupported_reference_options_raise():
    from mcd.multiScatter import genTables
    with pytest.raises(NotImplementedError):
        genTables({"material": "Si", "mode": "DS", "elastic": "Mott", "E0": 1e4, "Emin": 1e3, "Wc": 100})
    with pytest.raises(NotImplementedError):
        genTables({"material": "Si", "mode": "DS", "diffMFP": True, "E0": 1e4, "Emin": 1e3, "Wc": 100})
