"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY (the checker): imported by tests/, the smoke check and
bench.py's CPU-baseline leg, never by the product.  The product's failure to
load its own HIP library is never answered by calling this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "mc-discrete_amd"))

from mcd import _abi  # noqa: E402  (the interface structs, not the product's algorithm)

LIB_PATH = os.path.join(HERE, "liboracle.so")


class OracleTable(C.Structure):
    _fields_ = [("nE", C.c_int32), ("nW", C.c_int32), ("Es", C.POINTER(C.c_float)),
                ("Ws", C.POINTER(C.c_float)), ("cdf", C.POINTER(C.c_float))]


ORACLE_ELS_DTYPE = np.dtype([
    ("final_E", "<f8"), ("dir", "<f8", (3,)), ("pos", "<f8", (3,)), ("n_scatt", "<u4"),
    ("outcome", "u1"), ("outcome2", "u1"), ("flags", "u1"), ("pad", "u1"), ("path", "<f8"),
])
assert ORACLE_ELS_DTYPE.itemsize == 72

ORACLE_STEP_DTYPE = np.dtype([
    ("electron", "<u4"), ("step", "<u4"), ("channel", "<i4"), ("outcome", "<i4"),
    ("u", "<f8", (4,)), ("E", "<f8"), ("E_loss", "<f8"), ("pol", "<f8"), ("az", "<f8"), ("pathl", "<f8"),
    ("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("energy", "<f8"),
    ("dx", "<f8"), ("dy", "<f8"), ("dz", "<f8"), ("yx", "<f8"), ("yy", "<f8"), ("yz", "<f8"),
])

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i, u32, u64 = C.c_double, C.c_int, C.c_uint32, C.c_uint64
        P = C.POINTER
        L.oracle_sigma_ruth.restype = d; L.oracle_sigma_ruth.argtypes = [d, i]
        L.oracle_sigma_ruth_model.restype = d; L.oracle_sigma_ruth_model.argtypes = [d, i, i]
        L.oracle_sigma_moller.restype = d; L.oracle_sigma_moller.argtypes = [d, d, d]
        L.oracle_sigma_gryz.restype = d; L.oracle_sigma_gryz.argtypes = [d, d, d]
        L.oracle_sigma_quinn.restype = d; L.oracle_sigma_quinn.argtypes = [d, d, d, d]
        L.oracle_sigma_set.restype = i; L.oracle_sigma_set.argtypes = [P(_abi.Config), d, P(d), P(i)]
        L.oracle_pick.restype = i; L.oracle_pick.argtypes = [P(d), i, d]
        L.oracle_table_row.restype = i; L.oracle_table_row.argtypes = [P(OracleTable), d]
        L.oracle_table_col.restype = i; L.oracle_table_col.argtypes = [P(OracleTable), i, d]
        L.oracle_rotate.restype = None; L.oracle_rotate.argtypes = [d, d, P(d), P(d)]
        L.oracle_philox4x32_10.restype = None
        L.oracle_philox4x32_10.argtypes = [P(u32), P(u32), P(u32)]
        L.oracle_u01.restype = d; L.oracle_u01.argtypes = [u32]
        L.oracle_bethe_cl.restype = d; L.oracle_bethe_cl.argtypes = [i, d, d, d]
        L.oracle_bethe_jl.restype = d; L.oracle_bethe_jl.argtypes = [i, d, d, d, d]
        L.oracle_bethe_explicit.restype = d
        L.oracle_bethe_explicit.argtypes = [d, d, P(d), P(d), i, d, d]
        L.oracle_run.restype = i
        L.oracle_run.argtypes = [P(_abi.Config), P(OracleTable), P(OracleTable), u64, u32, u64, u64,
                                 C.c_void_p, C.c_void_p, u64, P(u64), P(u64), i]
        _lib = L
    return _lib


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return [int(x) for x in o]


class Table:
    """An inverse-CDF table in the reference layout (Es, Ws, cdf with masked = 0)."""

    def __init__(self, Es, Ws, cdf):
        self.Es = np.ascontiguousarray(Es, dtype=np.float32)
        self.Ws = np.ascontiguousarray(Ws, dtype=np.float32)
        self.cdf = np.ascontiguousarray(cdf, dtype=np.float32)
        assert self.cdf.shape == (self.Es.size, self.Ws.size - 1)
        self.c = OracleTable(self.Es.size, self.Ws.size,
                             self.Es.ctypes.data_as(C.POINTER(C.c_float)),
                             self.Ws.ctypes.data_as(C.POINTER(C.c_float)),
                             self.cdf.ctypes.data_as(C.POINTER(C.c_float)))

    def row(self, E):
        return lib().oracle_table_row(C.byref(self.c), float(E))

    def col(self, row, u):
        return lib().oracle_table_col(C.byref(self.c), int(row), float(u))


def sigma_set(cfg, E):
    sig = (C.c_double * _abi.MAX_CHANNELS)()
    lab = (C.c_int * _abi.MAX_CHANNELS)()
    n = lib().oracle_sigma_set(C.byref(cfg), float(E), sig, lab)
    return np.array(sig[:n]), np.array(lab[:n])


def pick(sig, u):
    arr = (C.c_double * len(sig))(*[float(s) for s in sig])
    return lib().oracle_pick(arr, len(sig), float(u))


def rotate(c2, half_phi, d, y):
    dd = (C.c_double * 3)(*d)
    yy = (C.c_double * 3)(*y)
    lib().oracle_rotate(float(c2), float(half_phi), dd, yy)
    return np.array(dd[:]), np.array(yy[:])


def run(cfg, moller, gryz, seed, first, n, config_id=0, trace_cap=0, nthreads=1):
    """Run electrons [first, first+n); returns (els, trace, events)."""
    els = np.zeros(n, dtype=ORACLE_ELS_DTYPE)
    trace = np.zeros(trace_cap, dtype=ORACLE_STEP_DTYPE) if trace_cap else None
    tn = C.c_uint64(0)
    ev = C.c_uint64(0)
    rc = lib().oracle_run(C.byref(cfg), C.byref(moller.c) if moller else None,
                          C.byref(gryz.c) if gryz else None, int(seed), int(config_id), int(first), int(n),
                          els.ctypes.data, trace.ctypes.data if trace is not None else None, int(trace_cap),
                          C.byref(tn), C.byref(ev), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed with status {rc}")
    if trace is not None:
        if tn.value > trace_cap:
            raise RuntimeError(f"trace buffer too small: {tn.value} rows > {trace_cap}")
        trace = trace[: tn.value]
    return els, trace, ev.value
