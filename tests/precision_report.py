#!/usr/bin/env python3
"""Per-output error of the FP32 transport against the all-FP64 GPU run of the
same Philox streams (needs the GPU).

Workloads: DS Cu 20 keV -> 1 keV, 70 deg, Wc 50 eV, tol 1e-3 (seed 1), and the
continuous-mode precision study, Au 30 keV -> 1 keV at 0 and 70 deg with the
classical, Joy-Luo and explicit Bethe laws (seed 2); 2^22 electrons each by
default.  For every output column of the electron records it prints the share
of trajectories that are identical end to end (same outcome and scatter count),
and over those the largest and median relative error of the exit energy and
the largest absolute error of the direction and of the last position relative
to the largest |position|; plus the backscatter yields of both runs.

    python tests/precision_report.py [--electrons N] [--markdown]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "mc-discrete_amd"))


def compare(r32, r64):
    from mcd import _abi
    a, b = r32.els, r64.els
    same = (a["outcome"] == b["outcome"]) & (a["outcome2"] == b["outcome2"]) & (a["n_scatt"] == b["n_scatt"])
    s32, s64 = a[same], b[same]
    relE = np.abs(s32["final_E"].astype(np.float64) - s64["final_E"]) / s64["final_E"]
    ddir = np.abs(s32["dir"].astype(np.float64) - s64["dir"]).max(axis=1)
    scale = max(np.abs(s64["pos"]).max(), 1.0)
    dpos = np.abs(s32["pos"].astype(np.float64) - s64["pos"]).max(axis=1) / scale
    y32 = float((a["outcome"] == _abi.OUT_BKS).mean())
    y64 = float((b["outcome"] == _abi.OUT_BKS).mean())
    return dict(identical=float(same.mean()), relE_max=float(relE.max()), relE_med=float(np.median(relE)),
                dir_max=float(ddir.max()), pos_max=float(dpos.max()), y32=y32, y64=y64,
                ev32=int(r32.events), ev64=int(r64.events))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--electrons", type=int, default=1 << 22)
    ap.add_argument("--markdown", action="store_true")
    a = ap.parse_args()
    from mcd import _abi
    from mcd.engine import Transport
    n = a.electrons
    cases = [("DS Cu 20 keV 70 deg (seed 1)", dict(material="Cu", E0=20000.0, Wc=50.0, s_tilt=70.0, mode="DS"), 1)]
    for tilt in (0.0, 70.0):
        for law in ("classical", "JL", "explicit"):
            cases.append((f"cont Au 30 keV {tilt:.0f} deg {law} (seed 2)",
                          dict(material="Au", E0=30000.0, s_tilt=tilt, mode="cont", bethe_model=law), 2))
    rows = []
    for name, kw, seed in cases:
        kw = dict(kw)
        mat, E0 = kw.pop("material"), kw.pop("E0")
        tr = Transport(mat, E0, 1000.0, **kw)
        r32 = tr.run(n, seed=seed)
        r64 = tr.run(n, seed=seed, precision=_abi.PREC_FP64)
        c = compare(r32, r64)
        rows.append((name, c))
        print(name, c, flush=True)
    if a.markdown:
        print("\n| workload | identical trajectories | exit energy rel. err. max / median | direction abs. err. max "
              "| last position err. max (rel. to max abs pos) | yield FP32 / FP64 |")
        print("|---|---|---|---|---|---|")
        for name, c in rows:
            print(f"| {name} | {100 * c['identical']:.4f} % | {c['relE_max']:.2e} / {c['relE_med']:.1e} "
                  f"| {c['dir_max']:.2e} | {c['pos_max']:.2e} | {c['y32']:.6f} / {c['y64']:.6f} |")


if __name__ == "__main__":
    main()
