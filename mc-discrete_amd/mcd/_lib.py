"""Loader for libmcd.so, the gfx950 transport library.

torch is imported first so that the process holds exactly one HIP runtime
(torch's): libmcd.so then binds to it and torch tensors, torch streams and the
library share device pointers and queues.  There is no CPU fallback: a missing
library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (loads the HIP runtime libmcd.so binds to)

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcd.so")

_lib = None


class MCDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmcd error {code}: {msg}")
        self.code = code


def lib():
    """Load libmcd.so and declare its prototypes (include/mcd.h)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C mc-discrete_amd/csrc` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    P, vp, i32, u32, u64, sz = C.POINTER, C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_size_t
    L.mcd_version.restype = i32
    L.mcd_version.argtypes = []
    L.mcd_ctx_create.restype = i32
    L.mcd_ctx_create.argtypes = [i32, P(vp)]
    L.mcd_ctx_destroy.restype = i32
    L.mcd_ctx_destroy.argtypes = [vp]
    L.mcd_last_error.restype = i32
    L.mcd_last_error.argtypes = [vp, C.c_char_p, sz]
    L.mcd_tables_create.restype = i32
    L.mcd_tables_create.argtypes = [vp, P(_abi.Config), P(_abi.TableGridC), P(_abi.TableGridC), i32, P(vp)]
    L.mcd_tables_upload.restype = i32
    L.mcd_tables_upload.argtypes = [vp, P(_abi.Config), P(_abi.TableGridC), vp, P(_abi.TableGridC), vp, P(vp)]
    L.mcd_tables_download.restype = i32
    L.mcd_tables_download.argtypes = [vp, vp, i32, vp, sz]
    L.mcd_tables_destroy.restype = i32
    L.mcd_tables_destroy.argtypes = [vp]
    L.mcd_tables_info.restype = i32
    L.mcd_tables_info.argtypes = [vp, P(i32), P(i32), P(i32), P(C.c_float)]
    L.mcd_tables_set_pick_nodes.restype = i32
    L.mcd_tables_set_pick_nodes.argtypes = [vp, vp, i32]
    f64 = C.c_double
    L.mcd_grid_search.restype = i32
    L.mcd_grid_search.argtypes = [i32, f64, f64, f64, f64, f64, f64, f64, f64, vp, P(i32), vp, P(i32)]
    L.mcd_tables_build.restype = i32
    L.mcd_tables_build.argtypes = [vp, P(_abi.Config), f64, f64, i32, P(vp)]
    L.mcd_tables_build_batch.restype = i32
    L.mcd_scatter_pack.restype = i32
    L.mcd_scatter_pack.argtypes = [vp, vp, C.c_uint64, C.c_uint32, vp, C.c_size_t, P(C.c_size_t), vp]
    L.mcd_tables_build_batch.argtypes = [vp, P(_abi.Config), i32, f64, f64, i32, P(vp)]
    L.mcd_grid_search_batch.restype = i32
    L.mcd_grid_search_batch.argtypes = [vp, P(_abi.GridSpec), i32, f64, f64, P(i32), P(i32), P(i32)]
    L.mcd_tables_grid.restype = i32
    L.mcd_tables_grid.argtypes = [vp, vp, i32, P(i32), P(i32), vp, vp]
    L.mcd_tables_save.restype = i32
    L.mcd_tables_save.argtypes = [vp, vp, C.c_char_p]
    L.mcd_tables_load.restype = i32
    L.mcd_tables_load.argtypes = [vp, P(_abi.Config), f64, f64, i32, C.c_char_p, P(vp)]
    L.mcd_forward.restype = i32
    L.mcd_forward.argtypes = [vp, P(_abi.Config), vp, u64, u32, u64, u64, vp, P(_abi.ScatterBuf), vp,
                              P(_abi.Launch), vp]
    L.mcd_forward_noalloc.restype = i32
    L.mcd_forward_noalloc.argtypes = [vp, P(_abi.Config), vp, u64, u32, u64, u64, vp, P(_abi.ScatterBuf), vp,
                                      P(_abi.Launch), vp]
    L.mcd_batch_create.restype = i32
    L.mcd_batch_create.argtypes = [vp, P(_abi.BatchItem), i32, i32, P(vp)]
    L.mcd_batch_run.restype = i32
    L.mcd_batch_run.argtypes = [vp, vp, u64, vp, P(_abi.Launch), i32, vp]
    L.mcd_batch_destroy.restype = i32
    L.mcd_batch_destroy.argtypes = [vp]
    L.mcd_backward.restype = i32
    L.mcd_backward.argtypes = [vp, P(_abi.Config), vp, u64, u32, u64, u64, vp, vp, vp, P(_abi.BwdOpts), vp]
    if L.mcd_version() != 1:
        raise ImportError(f"libmcd.so ABI version {L.mcd_version()} does not match the bindings (1)")
    _lib = L
    return L


def last_error(ctx=None) -> str:
    buf = C.create_string_buffer(512)
    lib().mcd_last_error(ctx, buf, 512)
    return buf.value.decode(errors="replace")


def check(rc, ctx=None):
    if rc != 0:
        raise MCDError(rc, last_error(ctx))
