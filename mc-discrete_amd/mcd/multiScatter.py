"""Drop-in batch workers of the reference (MC/tools/multiScatter.py) running
the trajectories on the GPU.

* ``genTables(inputPar)``   replaces probTables.genTables (probTables.py:541-614):
  grids on the host, CDF rows on the GPU; only the Moller table and the
  outermost-shell Gryzinski table are built, the ones the transport reads
  (scattering.py:603,625).
* ``multiTraj_DS(inputPar, numTraj, material, tables, thingsToSave)`` and
  ``multiTraj_cont(inputPar, numTraj, material, thingsToSave, tables)`` keep the
  reference's signatures (multiScatter.py:66,171) and return the same list of
  per-electron ``{'els': {...}, 'scats': {...}}`` dicts.
* ``scatterMultiEl_DS`` / ``scatterMultiEl_cont`` keep the legacy names
  (multiScatter.py:15,118) and fill the given output mapping.
* ``run_columns`` is the columnar fast path used by doScatter.py.

Successive calls on one tables object continue the electron numbering, so the
Philox streams (key = seed, counter = (electron, config, step, 0)) never repeat
within a run; the reference instead re-seeded per job with pid*time
(multiScatter.py:94), which repeats streams within a second.
"""
from __future__ import annotations

import os

import numpy as np

from . import _abi
from .engine import Transport
from .fileTools import els_columns, per_electron_dicts, scats_columns, zipDict
from .material import Material
from .tables import grid_key, gryz_grid, moller_grid


class DeviceTables(dict):
    """genTables() result: {'Moller': grid, 'Gryz': [outer-shell grid]} plus
    the device transports built from them and the electron counter."""

    def __init__(self, inputPar, material):
        super().__init__()
        self.inputPar = dict(inputPar)
        self.material = material
        self.next_electron = 0
        self._transports = {}
        self.table_file = None   # the parameter-keyed table file the DS tables were read from or saved to

    def claim(self, n):
        first = self.next_electron
        self.next_electron += int(n)
        return first

    def transport(self, inputPar, material=None, mode=None):
        p = dict(self.inputPar)
        p.update(inputPar)
        mode = mode or p.get("mode", "DS")
        key = (mode, p.get("Bethe_model", "classical"), p.get("elastic", "Ruth_vanilla"), float(p.get("s_tilt", 0.0)),
               float(p.get("thickness", float("inf"))), int(p.get("maxScatt", 0) or 0))
        if key not in self._transports:
            mat = _material(material or self.material, p)
            ds = "DS" in mode
            from_file = ds and self.table_file is not None
            grids = (self["Moller"], self["Gryz"][0]) if ("Moller" in self and not from_file) else None
            self._transports[key] = Transport(
                mat, p["E0"], p["Emin"], Wc=p.get("Wc", 50.0), s_tilt=p.get("s_tilt", 0.0),
                mode="DS" if ds else "cont", bethe_model=p.get("Bethe_model", "classical"),
                tol_E=p.get("tol_E", 1e-3), tol_W=p.get("tol_W", 1e-3),
                thickness=p.get("thickness", float("inf")), max_scatt=int(p.get("maxScatt", 0) or 0),
                grids=grids, table_file=self.table_file if from_file else None,
                elastic=p.get("elastic", "Ruth_vanilla"))
        return self._transports[key]


def _material(material, inputPar):
    if isinstance(material, Material):
        return material
    if isinstance(material, str):
        return Material(material)
    return Material(inputPar["material"])


def _check_elastic(inputPar):
    """The Rutherford models (Ruth_vanilla, Ruth_vanilla_wDefl, Ruth_nigram,
    Ruth_nigram_wDefl) resolve as the reference resolves them in each mode
    (_abi.elastic_model); Mott needs the Ioffe tables, which the reference
    does not ship.  diffMFP: the reference's branch reads self.xip_g, whose
    assignment is commented out (scattering.py:511-513), so it raises."""
    _abi.elastic_model(inputPar.get("elastic", "Ruth_vanilla"), inputPar.get("mode", "DS"))
    if inputPar.get("diffMFP", False):
        raise NotImplementedError("diffMFP: the reference's diffraction branch reads self.xip_g, which is never set "
                                  "(scattering.py:511-513); not implemented")


def table_file(inputPar, table_dir, material=None):
    """The parameter-keyed table file of a DS run under table_dir: its name
    holds every parameter the tables depend on (the reference's
    '{type}_{shell}.table' names hold none, probTables.py:160)."""
    mat = _material(material, inputPar)
    key = grid_key("tables", mat, inputPar["E0"], inputPar["Emin"], inputPar["Wc"], inputPar.get("tol_E", 1e-3),
                   inputPar.get("tol_W", 1e-3))
    return os.path.join(table_dir, key + "_fp64.mcdt")


def genTables(inputPar, table_dir=None):
    """Tables for a run (probTables.genTables, probTables.py:541-614).

    With a table directory (the argument or inputPar['table_dir']) the DS
    tables go through a parameter-keyed table file there: gen_tables true
    builds them (library grid search, device CDF fill) and writes the file;
    gen_tables false reads the file, and builds and writes it when there is
    none.  A file built for other parameters is refused, never reused.
    Without a table directory the tables are built in memory."""
    _check_elastic(inputPar)
    mat = Material(inputPar["material"])
    t = DeviceTables(inputPar, mat)
    if "DS" in inputPar.get("mode", "DS"):
        tol_E, tol_W = inputPar.get("tol_E", 1e-3), inputPar.get("tol_W", 1e-3)
        table_dir = table_dir or inputPar.get("table_dir")
        if table_dir is None:
            t["Moller"] = moller_grid(mat, inputPar["E0"], inputPar["Emin"], inputPar["Wc"]).build_grids(tol_E, tol_W)
            t["Gryz"] = [gryz_grid(mat, inputPar["E0"], inputPar["Emin"]).build_grids(tol_E, tol_W)]
            return t
        path = table_file(inputPar, table_dir, mat)
        kw = dict(Wc=inputPar["Wc"], tol_E=tol_E, tol_W=tol_W)
        if inputPar.get("gen_tables", True) or not os.path.exists(path):
            if not inputPar.get("gen_tables", True):
                print(f" - no table file {path}: generating the tables")
            tr = Transport(mat, inputPar["E0"], inputPar["Emin"], grid_search="library", **kw)
            os.makedirs(table_dir, exist_ok=True)
            tr.save_tables(path)
            print(f" - tables written to {path}")
        else:
            tr = Transport(mat, inputPar["E0"], inputPar["Emin"], table_file=path, **kw)
            print(f" - tables read from {path}")
        t["Moller"], t["Gryz"] = tr.grids[0], [tr.grids[1]]
        t.table_file = path
    return t


def _tables_for(inputPar, tables):
    if isinstance(tables, DeviceTables):
        return tables
    return genTables(inputPar)


def _run(inputPar, numTraj, material, tables, thingsToSave, mode):
    _check_elastic(inputPar)
    tables = _tables_for(inputPar, tables)
    tr = tables.transport(inputPar, material, mode)
    first = tables.claim(numTraj)
    want_scats = bool(thingsToSave.get("scat_output"))
    seed = int(inputPar.get("seed", 0))
    r = (run_rows(tr, int(numTraj), seed, first, scat_fields=_abi.scat_fields_for(thingsToSave["scat_output"]))
         if want_scats else tr.run(int(numTraj), seed=seed, first=first))
    return per_electron_dicts(r.els, r.scats, tr.mat, thingsToSave, first)


def multiTraj_DS(inputPar, numTraj, material, tables, thingsToSave):
    """numTraj DS trajectories (multiScatter.py:66-116)."""
    return _run(inputPar, numTraj, material, tables, thingsToSave, "DS")


def multiTraj_cont(inputPar, numTraj, material, thingsToSave, tables):
    """numTraj continuous-slowing-down trajectories (multiScatter.py:171-229),
    with the Bethe law of inputPar['Bethe_model'] (classical, JL or explicit)."""
    return _run(inputPar, numTraj, material, tables, thingsToSave, "cont")


def scatterMultiEl_DS(inputPar, tables, listParToSave, outputQ, count=0):
    """Legacy entry point (multiScatter.py:15-61): inputPar['num_el']
    trajectories; outputQ['els'] / outputQ['scats'] receive the zipped columns."""
    res = multiTraj_DS(inputPar, inputPar["num_el"], inputPar.get("material"), tables, listParToSave)
    els, scats = {}, {}
    for r in res:
        zipDict(els, r["els"])
        zipDict(scats, r["scats"])
    outputQ["els"], outputQ["scats"] = els, scats
    return outputQ


def scatterMultiEl_cont(inputPar, thingsToSave, output, num, count=0):
    """Legacy entry point (multiScatter.py:118-168): `num` continuous-mode
    trajectories; output['electrons'] / output['scatterings'] receive the
    zipped columns."""
    res = multiTraj_cont(inputPar, num, inputPar.get("material"), thingsToSave, None)
    els, scats = {}, {}
    for r in res:
        zipDict(els, r["els"])
        zipDict(scats, r["scats"])
    output["electrons"], output["scatterings"] = els, scats
    return output


# Scatter rows per device buffer when the caller gives no capacity: 2^26 rows
# of 64 bytes (4 GiB); a range whose rows do not fit is split (run_rows).
SCAT_ROWS_DEFAULT = 1 << 26


def run_rows(tr, n, seed, first, scat_max_electron=None, scat_capacity=None, scat_fields=None):
    """Electrons [first, first+n) with their scatter rows, in as many device
    runs as the row buffer needs: when a range's rows overflow the buffer it is
    run again as two halves (records and rows are keyed by electron id and the
    Philox counter, so the halves give exactly the rows of the whole).  Returns
    one RunResult with the records, rows (sorted by electron, step) and summed
    counters of the range."""
    from .engine import RunResult, ScatterOverflow
    cap = int(scat_capacity) if scat_capacity else min(SCAT_ROWS_DEFAULT, max(1024, 4096 * int(n)))
    try:
        return tr.run(int(n), seed=seed, first=first, scats=True, scat_max_electron=scat_max_electron,
                      scat_capacity=cap, scat_fields=scat_fields)
    except ScatterOverflow:
        if n <= 1:
            raise
    h = int(n) // 2
    a = run_rows(tr, h, seed, first, scat_max_electron, cap, scat_fields)
    b = run_rows(tr, int(n) - h, seed, first + h, scat_max_electron, cap, scat_fields)
    return RunResult(els=np.concatenate([a.els, b.els]), counters=a.counters + b.counters,
                     scats=np.concatenate([a.scats, b.scats]))


def iter_columns(inputPar, n, chunk, tables=None, first=None, streams=2):
    """Columnar runs of n electrons in launches of `chunk`, pipelined over
    `streams` HIP streams (Transport.run_chunks); yields (els_columns,
    RunResult) per launch in electron order.  Scatter rows are not recorded
    here (run_columns records them)."""
    _check_elastic(inputPar)
    tables = _tables_for(inputPar, tables)
    tr = tables.transport(inputPar, None, inputPar.get("mode", "DS"))
    first = tables.claim(n) if first is None else first
    el_want = list(inputPar.get("electron_output", ("outcome", "final_E", "final_dir", "last_pos")))
    for _, r in tr.run_chunks(int(n), int(chunk), seed=int(inputPar.get("seed", 0)), first=first, streams=streams):
        yield els_columns(r.els, el_want), r


def run_columns(inputPar, n, tables=None, first=None, scat_max_electron=None, scat_capacity=None):
    """Columnar run of n electrons: returns (els_columns, scats_columns or
    None, RunResult).  scats are recorded for electrons with id <
    scat_max_electron (default: all of them)."""
    _check_elastic(inputPar)
    tables = _tables_for(inputPar, tables)
    tr = tables.transport(inputPar, None, inputPar.get("mode", "DS"))
    first = tables.claim(n) if first is None else first
    sc_want = list(inputPar.get("scatter_output", ()))
    el_want = list(inputPar.get("electron_output", ("outcome", "final_E", "final_dir", "last_pos")))
    seed = int(inputPar.get("seed", 0))
    if sc_want:
        r = run_rows(tr, int(n), seed, first, scat_max_electron, scat_capacity, _abi.scat_fields_for(sc_want))
    else:
        r = tr.run(int(n), seed=seed, first=first)
    els = els_columns(r.els, el_want)
    scats = scats_columns(r.scats, tr.mat, sc_want) if sc_want else None
    return els, scats, r


__all__ = ["genTables", "multiTraj_DS", "multiTraj_cont", "scatterMultiEl_DS", "scatterMultiEl_cont",
           "run_columns", "iter_columns", "run_rows", "DeviceTables"]
