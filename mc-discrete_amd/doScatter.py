#!/usr/bin/env python3
"""Command-line driver (MC/doScatter.py): ``python doScatter.py -i input.file``.

Reads the input file (fileTools.readInput), builds the tables, runs
inputPar['num_el'] trajectories on the GPU in chunks and writes the ``input``,
``els`` and ``scats`` tables.  The store is named as the reference names it
(doScatter.py:175-186) under --out (default ./data); DS tables are kept in a
parameter-keyed table file under <out>/tables (or the input key table_dir)
and read back when the input says gen_tables: no.  The store is an HDF5
file: written through pandas.HDFStore(format='table', data_columns=True) like
the reference's listener (doScatter.py:112-117) when PyTables is importable,
otherwise by mcd.hdf5 with the same groups, tables and attributes.  -u
(units) is accepted and ignored, as it is unused in the reference.
"""
from __future__ import annotations

import os
import sys
import time
from getopt import GetoptError, getopt

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from mcd.fileTools import readInput  # noqa: E402


def usage():
    print("doScatter.py -i <input file> [-o <output dir>] [-c <electrons per chunk>]")


def dealWithInput(argv):
    """getopt handling of doScatter.dealWithInput (doScatter.py:121-151)."""
    inputfile, outdir, chunk = "../inputFile/input.file", "data", 1 << 20
    try:
        opts, _ = getopt(argv, "uhi:o:c:", ["ifile=", "out=", "chunk="])
    except GetoptError as err:
        print(err)
        usage()
        sys.exit(2)
    for opt, arg in opts:
        if opt == "-h":
            usage()
            sys.exit()
        elif opt in ("-i", "--ifile"):
            inputfile = arg
        elif opt in ("-o", "--out"):
            outdir = arg
        elif opt in ("-c", "--chunk"):
            chunk = int(arg)
    return inputfile, outdir, chunk


def store_name(p):
    """doScatter.py:175-186."""
    return ("TRSM_diff:" + str(p.get("diffMFP", False)) + "_thick:" + str(p.get("thickness", "inf")) +
            "_mat:" + str(p["material"]) + "_mode:" + str(p["mode"]) + "_elastic:" +
            str(p.get("elastic", "Ruth_vanilla")) + "_tilt:" + str(p["s_tilt"]) + "_Emin:" + str(p["Emin"]) +
            "_E0:" + str(p["E0"]) + "_tolE:" + str(p.get("tol_E")) + "_tolW:" + str(p.get("tol_W")) + ".h5")


class Sink:
    """Appends column dicts to the run's HDF5 store: through pandas.HDFStore
    (format='table', data_columns=True) when PyTables is importable, as the
    reference's listener does (doScatter.py:112-117); otherwise through
    mcd.hdf5.H5Store, which writes the same groups and tables itself."""

    MIN_ITEMSIZE = {"els": {"outcome": 6}, "scats": {"type": 12}}

    def __init__(self, path, inputPar):
        import pandas as pd

        self.pd = pd
        self.path = path
        try:
            import tables  # noqa: F401

            self.hdf = True
        except ImportError:
            self.hdf = False
        if os.path.exists(path):
            os.remove(path)
        if self.hdf:
            with pd.HDFStore(path, mode="w") as st:
                st["input"] = pd.DataFrame(pd.Series(inputPar)).T   # doScatter.py:66
            self.store = None
        else:
            from mcd.hdf5 import H5Store

            self.store = H5Store(path)
            self.store.put_input(inputPar)

    def append(self, key, cols):
        if cols is None or not cols:
            return
        if self.hdf:
            df = self.pd.DataFrame(cols)
            with self.pd.HDFStore(self.path) as st:
                st.put(key, df, format="table", data_columns=True, append=True,
                       min_itemsize=self.MIN_ITEMSIZE.get(key))
        else:
            self.store.append(key, cols, min_itemsize=self.MIN_ITEMSIZE.get(key))

    def close(self):
        if self.store is not None:
            self.store.close()


def main(argv=None):
    inputfile, outdir, chunk = dealWithInput(sys.argv[1:] if argv is None else argv)
    from mcd.multiScatter import genTables, iter_columns, run_columns

    p = readInput(inputfile)
    p.setdefault("thickness", float("inf"))
    p.setdefault("seed", 0)
    print(f" - Number of incident electrons: {p['num_el']}\n - Material is: {p['material']}\n"
          f" - Scattering mode is: {p['mode']}\n - Elastic scattering mode is: {p.get('elastic', 'Ruth_vanilla')}")
    t0 = time.time()
    os.makedirs(outdir, exist_ok=True)
    # DS tables through a parameter-keyed table file: input key table_dir, by
    # default <output dir>/tables (gen_tables: no reads it back)
    tables = genTables(p, table_dir=p.get("table_dir") or os.path.join(outdir, "tables"))
    sink = Sink(os.path.join(outdir, store_name(p)), p)
    done, events = 0, 0
    if not p.get("scatter_output"):
        # electron records only: launches pipelined over two streams, the host
        # writing one chunk while the device runs the next
        for els, r in iter_columns(p, p["num_el"], chunk, tables=tables, first=0):
            sink.append("els", els)
            done += len(r.els)
            events += r.events
            print(f"   {done}/{p['num_el']} trajectories, {events} scatter events", flush=True)
    while done < p["num_el"]:
        n = min(chunk, p["num_el"] - done)
        els, scats, r = run_columns(p, n, tables=tables, first=done)
        sink.append("els", els)
        sink.append("scats", scats)
        done += n
        events += r.events
        print(f"   {done}/{p['num_el']} trajectories, {events} scatter events", flush=True)
    sink.close()
    print(f" BSE data has been written to {sink.path} ({time.time() - t0:.1f} s)")
    return sink.path


if __name__ == "__main__":
    main()
