#!/usr/bin/env python3
"""Command-line driver (MC/doScatter.py): ``python doScatter.py -i input.file``.

Reads the input file (fileTools.readInput), builds the tables, runs
inputPar['num_el'] trajectories on the GPU in chunks and writes the ``input``,
``els`` and ``scats`` tables.  The store is named as the reference names it
(doScatter.py:175-186) under --out (default ./data).  With PyTables importable
the store is an HDF5 file written through pandas.HDFStore(format='table',
data_columns=True) like the reference's listener (doScatter.py:112-117);
otherwise the same tables are written as Parquet files in a directory of that
name.  -u (units) is accepted and ignored, as it is unused in the reference.
"""
from __future__ import annotations

import json
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
    """Appends column dicts to an HDF5 store (PyTables present) or Parquet."""

    def __init__(self, path, inputPar):
        import pandas as pd

        self.pd = pd
        self.path = path
        try:
            import tables  # noqa: F401

            self.hdf = True
        except ImportError:
            self.hdf = False
        if self.hdf:
            if os.path.exists(path):
                os.remove(path)
            with pd.HDFStore(path, mode="w") as st:
                st["input"] = pd.Series(list(inputPar.values()), index=list(inputPar.keys()), dtype=str)
        else:
            os.makedirs(path, exist_ok=True)
            with open(os.path.join(path, "input.json"), "w") as f:
                json.dump({k: (v if isinstance(v, (int, float, str, list, bool)) else str(v))
                           for k, v in inputPar.items()}, f, indent=1)
            self.parts = {"els": 0, "scats": 0}

    def append(self, key, cols):
        if cols is None or not cols:
            return
        df = self.pd.DataFrame(cols)
        if self.hdf:
            with self.pd.HDFStore(self.path) as st:
                st.put(key, df, format="table", data_columns=True, append=True,
                       min_itemsize={"outcome": 6} if key == "els" else {"type": 12})
        else:
            df.to_parquet(os.path.join(self.path, f"{key}.{self.parts[key]:05d}.parquet"))
            self.parts[key] += 1


def main(argv=None):
    inputfile, outdir, chunk = dealWithInput(sys.argv[1:] if argv is None else argv)
    from mcd.multiScatter import genTables, iter_columns, run_columns

    p = readInput(inputfile)
    p.setdefault("thickness", float("inf"))
    p.setdefault("seed", 0)
    print(f" - Number of incident electrons: {p['num_el']}\n - Material is: {p['material']}\n"
          f" - Scattering mode is: {p['mode']}\n - Elastic scattering mode is: {p.get('elastic', 'Ruth_vanilla')}")
    t0 = time.time()
    tables = genTables(p)
    os.makedirs(outdir, exist_ok=True)
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
    print(f" BSE data has been written to {sink.path} ({time.time() - t0:.1f} s)")
    return sink.path


if __name__ == "__main__":
    main()
