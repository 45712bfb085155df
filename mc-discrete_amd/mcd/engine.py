"""Device-side transport engine: contexts, tables and runs on one GPU.

A ``Transport`` holds one configuration (material, E0, Emin, Wc, tilt, mode,
Bethe model) with its device tables and runs electron ranges through
``mcd_forward``.  Output buffers are torch tensors on the GPU; results come back
as numpy structured arrays (``_abi.ELECTRON_DTYPE`` / ``_abi.SCATTER_DTYPE``).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from ._lib import check, lib
from .material import Material
from .tables import TableGrid, gryz_grid, moller_grid

_CTX = {}


def context(device: int = 0):
    """Process-wide libmcd context for a device (created on first use)."""
    if device not in _CTX:
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the transport runs only on the GPU (there is no CPU fallback)")
        h = C.c_void_p()
        check(lib().mcd_ctx_create(device, C.byref(h)))
        _CTX[device] = h
    return _CTX[device]


@dataclass
class RunResult:
    els: np.ndarray              # ELECTRON_DTYPE, one record per electron
    counters: np.ndarray         # uint64[NCOUNTERS]
    scats: np.ndarray | None     # SCATTER_DTYPE rows sorted by (electron, step), or None
    seconds: float | None = None

    @property
    def events(self) -> int:
        return int(self.counters[_abi.CNT_EVENTS])


class GuardError(RuntimeError):
    """A launch left guard picks unsettled (counter CNT_GUARD_TIMEOUTS > 0):
    its records are not valid and must not be used."""


def check_guard(counters, what="launch"):
    """Raise GuardError when a counters block (uint64[NCOUNTERS], host or
    device) reports unsettled guard picks."""
    v = counters[_abi.CNT_GUARD_TIMEOUTS]
    v = int(v.item()) if hasattr(v, "item") else int(v)
    if v:
        raise GuardError(f"{what}: {v} guard picks were not settled (no reply from the service wavefront "
                         "within the wait bound, or a parked electron outside the launch); records are invalid")


class ScatterOverflow(RuntimeError):
    """The scatter rows of a run exceeded the row buffer (no rows are lost on
    the device side beyond the capacity being full: rerunning the same
    electrons with more room, or in smaller ranges, gives the same rows)."""


def guard_flags(guard):
    """mcd_launch.flags for a guard placement: None, "inline", "deferred" or
    "service" (the last for single-configuration launches only)."""
    if guard is None:
        return 0
    if guard == "inline":
        return _abi.LAUNCH_INLINE_GUARD
    if guard == "deferred":
        return _abi.LAUNCH_DEFER_GUARD
    if guard == "service":
        return _abi.LAUNCH_SERVICE_GUARD
    raise ValueError(f"guard must be None, 'inline', 'deferred' or 'service', not {guard!r}")


def refill_flags(period):
    """mcd_launch.flags bits for a compaction period of 1, 2, 4, ... 128 steps
    (0 or None: every step)."""
    if not period or period == 1:
        return 0
    k = int(period).bit_length() - 1
    if period != 1 << k or k > 7:
        raise ValueError(f"refill_period must be a power of two up to 128, not {period!r}")
    return k << _abi.LAUNCH_REFILL_SHIFT


class Transport:
    """One configuration on one GPU."""

    def __init__(self, material: str | Material, E0: float, Emin: float, Wc: float = 50.0, s_tilt: float = 0.0,
                 mode: str = "DS", bethe_model: str = "classical", tol_E: float = 1e-3, tol_W: float = 1e-3,
                 thickness: float = float("inf"), max_scatt: int = 0, J: float | None = None,
                 table_precision: int = _abi.PREC_FP64, device: int = 0, grids=None, cdfs=None,
                 pick_nodes: int = 0, grid_search: str = "host", table_file: str | None = None,
                 elastic: str | int = "Ruth_vanilla", tables: C.c_void_p | None = None):
        """DS tables: from `grids` (and `cdfs`) when given; else read from
        `table_file` (mcd_tables_load: refused when the file was built for
        other parameters); else built with the grid search run in Python
        (grid_search="host", scipy's quad as the reference integral) or by the
        library (grid_search="library", mcd_tables_build).  `tables`: a handle
        from mcd_tables_build_batch for this configuration, which the
        Transport then owns (see build_transports_batch)."""
        self.mat = material if isinstance(material, Material) else Material(material)
        self.device = device
        self.cfg = _abi.make_config(self.mat, E0, Emin, Wc, s_tilt, thickness, mode, bethe_model, max_scatt, J,
                                    elastic=elastic)
        self.mode = self.cfg.mode
        self.tol_E, self.tol_W = tol_E, tol_W
        self.grids = grids
        # launch shape used when a launch leaves block / waves_per_simd / guard
        # at their defaults; set by tune()
        self.tuned = None
        self._tables = C.c_void_p()
        ctx = context(device)
        self.table_precision = int(table_precision)
        if tables is not None:
            self._tables = tables
            if self.mode == _abi.MODE_DS:
                self._lazy_grids = (E0, Emin, Wc)  # copied from the tables on first use of .grids
                if pick_nodes:
                    check(lib().mcd_tables_set_pick_nodes(ctx, self._tables, int(pick_nodes)), ctx)
        elif self.mode == _abi.MODE_DS and self.grids is None and cdfs is None and (
                table_file is not None or grid_search == "library"):
            if table_file is not None:
                check(lib().mcd_tables_load(ctx, C.byref(self.cfg), float(tol_E), float(tol_W), int(table_precision),
                                            os.fsencode(table_file), C.byref(self._tables)), ctx)
            else:
                check(lib().mcd_tables_build(ctx, C.byref(self.cfg), float(tol_E), float(tol_W), int(table_precision),
                                             C.byref(self._tables)), ctx)
            self.grids = (self._library_grid(_abi.TABLE_MOLLER, moller_grid(self.mat, E0, Emin, Wc)),
                          self._library_grid(_abi.TABLE_GRYZ, gryz_grid(self.mat, E0, Emin)))
            if pick_nodes:
                check(lib().mcd_tables_set_pick_nodes(ctx, self._tables, int(pick_nodes)), ctx)
        elif self.mode == _abi.MODE_DS:
            if self.grids is None:
                self.grids = (moller_grid(self.mat, E0, Emin, Wc).build_grids(tol_E, tol_W),
                              gryz_grid(self.mat, E0, Emin).build_grids(tol_E, tol_W))
            gm = _abi.grid_struct(self.grids[0], _abi.TABLE_MOLLER)
            gg = _abi.grid_struct(self.grids[1], _abi.TABLE_GRYZ)
            if cdfs is None:
                check(lib().mcd_tables_create(ctx, C.byref(self.cfg), C.byref(gm), C.byref(gg), int(table_precision),
                                              C.byref(self._tables)), ctx)
            else:
                cm = np.ascontiguousarray(cdfs[0], dtype=np.float32)
                cg = np.ascontiguousarray(cdfs[1], dtype=np.float32)
                check(lib().mcd_tables_upload(ctx, C.byref(self.cfg), C.byref(gm), cm.ctypes.data, C.byref(gg),
                                              cg.ctypes.data, C.byref(self._tables)), ctx)
            if pick_nodes:
                # channel-pick table of a given node count (a batch item's, to reproduce its records)
                check(lib().mcd_tables_set_pick_nodes(ctx, self._tables, int(pick_nodes)), ctx)
        else:
            check(lib().mcd_tables_create(ctx, C.byref(self.cfg), None, None, int(table_precision),
                                          C.byref(self._tables)), ctx)

    @classmethod
    def _config_only(cls, material, E0, Emin, Wc=50.0, s_tilt=0.0, mode="DS", bethe_model="classical",
                     thickness=float("inf"), max_scatt=0, J=None, elastic="Ruth_vanilla", device=0):
        """An object holding only the mcd_config of these parameters (no tables)."""
        self = cls.__new__(cls)
        self.mat = material if isinstance(material, Material) else Material(material)
        self.cfg = _abi.make_config(self.mat, E0, Emin, Wc, s_tilt, thickness, mode, bethe_model, max_scatt, J,
                                    elastic=elastic)
        self._tables = C.c_void_p()
        return self

    def __del__(self):
        try:
            if self._tables:
                lib().mcd_tables_destroy(self._tables)
                self._tables = C.c_void_p()
        except Exception:
            pass

    # -- tables ---------------------------------------------------------------
    @property
    def grids(self):
        """(Moller, outer-shell Gryzinski) TableGrid of a DS configuration, or
        None; for tables from a batched build, read from the device tables on
        first use."""
        lazy = getattr(self, "_lazy_grids", None)
        if lazy is not None:
            E0, Emin, Wc = lazy
            self._lazy_grids = None
            self._grids = (self._library_grid(_abi.TABLE_MOLLER, moller_grid(self.mat, E0, Emin, Wc)),
                           self._library_grid(_abi.TABLE_GRYZ, gryz_grid(self.mat, E0, Emin)))
        return getattr(self, "_grids", None)

    @grids.setter
    def grids(self, value):
        self._lazy_grids = None
        self._grids = value

    def _library_grid(self, which, grid):
        ctx = context(self.device)
        nE, nW = C.c_int32(0), C.c_int32(0)
        check(lib().mcd_tables_grid(ctx, self._tables, int(which), C.byref(nE), C.byref(nW), None, None), ctx)
        grid.Es = np.zeros(nE.value, dtype=np.float32)
        grid.Ws = np.zeros(nW.value, dtype=np.float32)
        check(lib().mcd_tables_grid(ctx, self._tables, int(which), C.byref(nE), C.byref(nW), grid.Es.ctypes.data,
                                    grid.Ws.ctypes.data), ctx)
        return grid

    def save_tables(self, path: str):
        """Write the DS tables to a parameter-keyed table file (mcd_tables_save;
        tables from grid_search="library" or a table file only)."""
        ctx = context(self.device)
        check(lib().mcd_tables_save(ctx, self._tables, os.fsencode(path)), ctx)

    def cdf(self, which: int) -> np.ndarray:
        """Device-built CDF in the reference layout (masked = 0)."""
        g = self.grids[which]
        out = np.zeros((g.Es.size, g.Ws.size - 1), dtype=np.float32)
        ctx = context(self.device)
        check(lib().mcd_tables_download(ctx, self._tables, int(which), out.ctypes.data, out.size), ctx)
        return out

    def pick_info(self):
        n, s, e, d = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        check(lib().mcd_tables_info(self._tables, C.byref(n), C.byref(s), C.byref(e), C.byref(d)))
        return dict(n_nodes=n.value, swap_cells=s.value, exact_cells=e.value, delta=d.value)

    # -- runs -----------------------------------------------------------------
    def launch(self, n, els, counters, seed=1, first=0, config_id=0, scat_rows=None, scat_max_electron=0,
               precision=_abi.PREC_FP32, block=0, waves_per_simd=0, guard=None, stream=None, refill_period=0):
        """Asynchronous launch on `stream` into caller-owned torch buffers.
        guard: None (library default: served for DS launches of at least 2^21
        electrons with no shape given, inline otherwise), "inline", "deferred"
        or "service" -- where the DS FP64 guard pick runs (same records in
        every form).  refill_period: the compaction period, idle lanes take
        new electrons every refill_period steps (0: every step)."""
        if self.tuned is not None and precision == _abi.PREC_FP32:
            block = block or self.tuned["block"]
            waves_per_simd = waves_per_simd or self.tuned["waves_per_simd"]
            guard = guard or self.tuned["guard"]
            refill_period = refill_period or self.tuned.get("refill_period", 0)
        ctx = context(self.device)
        sb = None
        if scat_rows is not None:
            sb = _abi.ScatterBuf(scat_rows.data_ptr(), scat_rows.numel() // _abi.SCATTER_DTYPE.itemsize,
                                 int(scat_max_electron))
        la = _abi.Launch(int(block), int(waves_per_simd), int(precision),
                         guard_flags(guard) | refill_flags(refill_period))
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        check(lib().mcd_forward(ctx, C.byref(self.cfg), self._tables, int(seed), int(config_id), int(first), int(n),
                                els.data_ptr(), C.byref(sb) if sb is not None else None, counters.data_ptr(),
                                C.byref(la), s), ctx)

    def tune(self, n=1 << 22, **kw):
        """Time the launch shapes of mcd.autotune.TRANSPORT_CANDIDATES on n
        electrons of this configuration, keep the fastest for later launches
        (records are the same for every shape) and return the tuner's report."""
        from .autotune import tune_transport
        res = tune_transport(self, n, **kw)
        self.tuned = dict(block=res["block"], waves_per_simd=res["waves_per_simd"], guard=res["guard"],
                          refill_period=res.get("refill_period", 1))
        return res

    def run(self, n, seed=1, first=0, config_id=0, scats=False, scat_max_electron=None, scat_capacity=None,
            precision=_abi.PREC_FP32, block=0, waves_per_simd=0, guard=None, refill_period=0,
            scat_fields=None) -> RunResult:
        """Run electrons [first, first+n) and return host results.  scat_fields:
        the scatter-row fields to bring to the host (mcd_scatter_pack packs them
        into columns on the device); None brings whole rows.  Either way the
        rows come back sorted by (electron, step)."""
        dev = torch.device("cuda", self.device)
        els = torch.empty(max(n, 1) * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
        rows = None
        if scats:
            if scat_max_electron is None:
                scat_max_electron = first + n
            if scat_capacity is None:
                # room for 4096 rows per recorded electron, at most 2^26 rows (4 GiB);
                # a run that needs more raises ScatterOverflow (multiScatter.run_rows
                # splits the range instead)
                scat_capacity = min(1 << 26, max(1024, 4096 * min(n, max(0, scat_max_electron - first))))
            rows = torch.empty(scat_capacity * _abi.SCATTER_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        self.launch(n, els, cnt, seed=seed, first=first, config_id=config_id, scat_rows=rows,
                    scat_max_electron=scat_max_electron or 0, precision=precision, block=block,
                    waves_per_simd=waves_per_simd, guard=guard, refill_period=refill_period)
        torch.cuda.synchronize(dev)
        counters = cnt.cpu().numpy().view(np.uint64)
        check_guard(counters)
        e = els.cpu().numpy()[: n * _abi.ELECTRON_DTYPE.itemsize].view(_abi.ELECTRON_DTYPE)
        srows = None
        if rows is not None:
            used = int(counters[_abi.CNT_SCAT_ROWS])
            if used > scat_capacity:
                raise ScatterOverflow(f"scatter-row buffer overflow: {used} rows requested, capacity {scat_capacity}")
            if scat_fields is None:
                srows = rows[: used * _abi.SCATTER_DTYPE.itemsize].cpu().numpy().view(_abi.SCATTER_DTYPE)
            else:
                srows = self._packed_rows(rows, used, scat_fields)
            srows = srows[np.lexsort((srows["step"], _abi.row_offsets(srows, first)))]
        return RunResult(els=e, counters=counters, scats=srows)


    def _packed_rows(self, rows, used, fields):
        """The given fields of the first `used` scatter rows (device), packed
        into columns on the device and copied as such: a structured array of
        those fields (electron and step always among them)."""
        names = [f for f in _abi.SCATTER_DTYPE.names if f in set(fields) | {"electron", "step"}]
        mask = 0
        for f in names:
            mask |= _abi.SCAT_FIELD_BITS[f]
        ctx = context(self.device)
        need = C.c_size_t(0)
        check(lib().mcd_scatter_pack(ctx, rows.data_ptr(), used, mask, None, 0, C.byref(need), None), ctx)
        out = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=rows.device)
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().mcd_scatter_pack(ctx, rows.data_ptr(), used, mask, out.data_ptr(), out.numel(), C.byref(need), s),
              ctx)
        host = out.cpu().numpy()
        dt = np.dtype([(f, _abi.SCATTER_DTYPE.fields[f][0]) for f in names])
        res = np.empty(used, dtype=dt)
        at = 0
        for f in names:
            t = _abi.SCATTER_DTYPE.fields[f][0]
            res[f] = host[at: at + used * t.itemsize].view(t)
            at += (used * t.itemsize + 255) // 256 * 256
        return res

    # -- differentiation ------------------------------------------------------
    def run_chunks(self, n, chunk, seed=1, first=0, streams=2, guard=None):
        """Electrons [first, first+n) in launches of `chunk` electrons, yielded
        in order as (first electron, RunResult without scatter rows).  The
        launches alternate between `streams` HIP streams (default priority),
        each with its own record and counter buffers, so the device starts the
        next launch while the previous one's last trajectories finish and while
        the host handles the previous results.  Records equal those of one
        launch over the whole range (they are keyed by electron id).  guard:
        as for launch (every form may overlap on separate streams: the
        deferred form's park buffer is per stream)."""
        dev = torch.device("cuda", self.device)
        esz = _abi.ELECTRON_DTYPE.itemsize
        chunk = max(1, int(chunk))
        ss = [torch.cuda.Stream(device=dev) for _ in range(max(1, int(streams)))]
        bufs = [(torch.empty(chunk * esz, dtype=torch.uint8, device=dev),
                 torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)) for _ in ss]
        starts = list(range(int(first), int(first) + int(n), chunk))

        def launch(k):
            s, (els, cnt) = ss[k % len(ss)], bufs[k % len(ss)]
            m = min(chunk, int(first) + int(n) - starts[k])
            with torch.cuda.stream(s):
                cnt.zero_()
                self.launch(m, els, cnt, seed=seed, first=starts[k], stream=s, guard=guard)
            return k, m

        pending = [launch(k) for k in range(min(len(ss), len(starts)))]
        nxt = len(pending)
        while pending:
            k, m = pending.pop(0)
            s, (els, cnt) = ss[k % len(ss)], bufs[k % len(ss)]
            with torch.cuda.stream(s):  # copies on the finished launch's own stream
                counters = cnt.cpu().numpy().view(np.uint64).copy()
                e = els[: m * esz].cpu().numpy().view(_abi.ELECTRON_DTYPE).copy()
            if nxt < len(starts):
                pending.append(launch(nxt))
                nxt += 1
            check_guard(counters)
            yield starts[k], RunResult(els=e, counters=counters, scats=None)

    def _bwd(self, n, seed, first, config_id, grad_out, out_vjp, out_jac, n_bins, rel_step, graze, block,
             stream):
        ctx = context(self.device)
        o = _abi.BwdOpts(int(n_bins), float(rel_step), float(graze), int(block))
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        check(lib().mcd_backward(ctx, C.byref(self.cfg), self._tables, int(seed), int(config_id), int(first), int(n),
                                 ptr(grad_out), ptr(out_vjp), ptr(out_jac), C.byref(o), s), ctx)

    def jacobian(self, n, seed=1, first=0, config_id=0, n_bins=64, rel_step=0.0, graze=0.0, block=0,
                 out=None, stream=None):
        """Replay estimate of per-electron Y and H_b with their derivatives:
        array (1 + n_bins, 1 + NGRAD), row 0 = yield, row 1+b = bin b; column 0 =
        value, column 1+g = d/d(parameter g) in mcd.h's units (g = GRAD_*).
        With `out` (a zeroed float64 device tensor of that shape) the launch is
        asynchronous and accumulates raw sums into it; returns None."""
        dev = torch.device("cuda", self.device)
        acc = out if out is not None else torch.zeros((1 + n_bins, 1 + _abi.NGRAD), dtype=torch.float64, device=dev)
        self._bwd(n, seed, first, config_id, None, None, acc, n_bins, rel_step, graze, block, stream)
        if out is not None:
            return None
        return acc.cpu().numpy() / float(n)

    def vjp(self, n, grad_out, seed=1, first=0, config_id=0, rel_step=0.0, graze=0.0, block=0, out=None,
            stream=None):
        """L = dL/dY * Y + sum_b dL/dH_b * H_b (per electron) and dL/dparam.
        grad_out: float32 device tensor (1 + n_bins).  Returns float64 array
        (1 + NGRAD), or accumulates raw sums into `out` asynchronously."""
        dev = torch.device("cuda", self.device)
        g = grad_out.to(device=dev, dtype=torch.float32).contiguous()
        acc = out if out is not None else torch.zeros(1 + _abi.NGRAD, dtype=torch.float64, device=dev)
        self._bwd(n, seed, first, config_id, g, acc, None, g.numel() - 1, rel_step, graze, block, stream)
        if out is not None:
            return None
        return acc.cpu().numpy() / float(n)


def backscatter_stats(els: torch.Tensor, n: int, E0: float, n_bins: int = 64):
    """Backscatter count and exit-energy histogram (n_bins equal bins over
    [0, E0], last bin closed) of n electron records held in a device byte
    tensor, computed on the device.  Returns (count, counts[n_bins]) tensors."""
    rec = els[: n * _abi.ELECTRON_DTYPE.itemsize].view(n, _abi.ELECTRON_DTYPE.itemsize)
    bks = rec[:, _abi.ELECTRON_DTYPE.fields["outcome"][1]] == _abi.OUT_BKS
    E = rec[:, 0:4].contiguous().view(torch.float32).reshape(n)
    b = torch.clamp((E[bks].double() * (n_bins / E0)).long(), 0, n_bins - 1)
    return bks.sum(), torch.bincount(b, minlength=n_bins)


class Batch:
    """Many configurations in one persistent launch (mcd_batch_*).

    items: list of (Transport, n_electrons).  Item i draws from Philox counter
    word config = i, electrons 0..n-1, so results equal item-by-item runs with
    config_id = i."""

    def __init__(self, items, pick_nodes=0, device=0):
        self.device = device
        self.items = list(items)
        dev = torch.device("cuda", device)
        self.els = [torch.empty(max(n, 1) * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
                    for _, n in self.items]
        self.counters = torch.zeros(len(self.items) * _abi.NCOUNTERS, dtype=torch.int64, device=dev)
        arr = (_abi.BatchItem * len(self.items))()
        for i, ((tr, n), e) in enumerate(zip(self.items, self.els)):
            arr[i].cfg = C.pointer(tr.cfg)
            arr[i].tables = tr._tables
            arr[i].n = int(n)
            arr[i].els = e.data_ptr()
        self._arr = arr
        self._h = C.c_void_p()
        ctx = context(device)
        check(lib().mcd_batch_create(ctx, arr, len(self.items), int(pick_nodes), C.byref(self._h)), ctx)

    def __del__(self):
        try:
            if self._h:
                lib().mcd_batch_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def launch(self, seed=3, use_graph=True, block=0, guard=None, stream=None, refill_period=0):
        """guard: None (library default: inline), "inline", "deferred" or "service".
        refill_period: the compaction period, idle lanes take new electrons
        every refill_period steps (0: every step)."""
        ctx = context(self.device)
        la = _abi.Launch(int(block), 0, _abi.PREC_FP32, guard_flags(guard) | refill_flags(refill_period))
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        check(lib().mcd_batch_run(ctx, self._h, int(seed), self.counters.data_ptr(), C.byref(la), int(bool(use_graph)),
                                  s), ctx)

    def results(self):
        torch.cuda.synchronize(self.device)
        cnt = self.counters.cpu().numpy().view(np.uint64).reshape(len(self.items), _abi.NCOUNTERS)
        out = []
        for i, ((tr, n), e, c) in enumerate(zip(self.items, self.els, cnt)):
            check_guard(c, f"batch item {i}")
            rec = e.cpu().numpy()[: n * _abi.ELECTRON_DTYPE.itemsize].view(_abi.ELECTRON_DTYPE)
            out.append(RunResult(els=rec, counters=c.copy(), scats=None))
        return out


def grid_search_batch(grids, tol_E, tol_W, device=0, statuses=False):
    """Grid sizes of many tables from one device launch (mcd_grid_search_batch,
    findWseries / findEseries, probTables.py:280-382).  grids: TableGrid
    objects (mcd.tables.moller_grid / gryz_grid).  Returns int arrays nE, nW
    and, with statuses=True, the per-table status (0 or ERR_TABLE) instead of
    raising on the first table no grid can serve."""
    n = len(grids)
    arr = (_abi.GridSpec * max(n, 1))()
    for i, g in enumerate(grids):
        arr[i] = _abi.GridSpec(0 if g.kind == "Moller" else 1, 0, float(g.Wmin), float(g.n_el), float(g.Eb),
                               float(g.Ef), float(g.Emin), float(g.Emax))
    nE = np.zeros(max(n, 1), np.int32)
    nW = np.zeros(max(n, 1), np.int32)
    st = np.zeros(max(n, 1), np.int32)
    P = C.POINTER(C.c_int32)
    ctx = context(device)
    check(lib().mcd_grid_search_batch(ctx, arr, n, float(tol_E), float(tol_W), nE.ctypes.data_as(P),
                                      nW.ctypes.data_as(P), st.ctypes.data_as(P) if statuses else None), ctx)
    return (nE[:n], nW[:n], st[:n]) if statuses else (nE[:n], nW[:n])


def build_transports_batch(specs, tol_E=1e-3, tol_W=1e-3, table_precision=_abi.PREC_FP64, device=0, pick_nodes=0):
    """Transports for many configurations with their tables built together
    (mcd_tables_build_batch: the grid searches, the CDF rows and the row
    records of every table in one device launch each).  specs: list of
    dicts of Transport's keyword arguments (material, E0, Emin, Wc, s_tilt,
    mode, bethe_model, ...); the tables equal those of
    Transport(..., grid_search="library") bit for bit."""
    built = [Transport._config_only(device=device, **sp) for sp in specs]
    cfgs = (_abi.Config * len(built))(*[b.cfg for b in built])
    handles = (C.c_void_p * len(built))()
    ctx = context(device)
    check(lib().mcd_tables_build_batch(ctx, cfgs, len(built), float(tol_E), float(tol_W), int(table_precision),
                                       handles), ctx)
    out = []
    try:
        for i, sp in enumerate(specs):
            h = C.c_void_p(handles[i])
            handles[i] = None  # the Transport owns it from here
            out.append(Transport(tol_E=tol_E, tol_W=tol_W, table_precision=table_precision, device=device,
                                 pick_nodes=pick_nodes if sp.get("mode", "DS") == "DS" else 0, tables=h, **sp))
    finally:
        for h in handles:  # handles no Transport took (a constructor raised)
            if h:
                lib().mcd_tables_destroy(C.c_void_p(h))
    return out


def grids_for(mat: Material, E0, Emin, Wc, tol_E, tol_W):
    """The (Moller, outer-shell Gryzinski) grids of a DS configuration."""
    return (moller_grid(mat, E0, Emin, Wc).build_grids(tol_E, tol_W),
            gryz_grid(mat, E0, Emin).build_grids(tol_E, tol_W))


__all__ = ["Transport", "Batch", "RunResult", "context", "grids_for", "TableGrid", "backscatter_stats",
           "build_transports_batch", "grid_search_batch",
           "GuardError", "check_guard"]
