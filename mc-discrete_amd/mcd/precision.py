"""Per-output error of the FP32 transport against the all-FP64 GPU run of the
same Philox streams (the FP64 run is the device's own FP64 restatement of the
reference step, k_transport_f64).

The precision study: DS Cu 20 keV -> 1 keV, 70 deg, Wc 50 eV, tol 1e-3 (seed 1)
and continuous mode, Au 30 keV -> 1 keV at 0 and 70 deg with the classical,
Joy-Luo and explicit Bethe laws (seed 2).  For each workload: the share of
trajectories identical end to end (same outcome and scatter count), over
those the largest and median relative error of the exit energy, the largest
absolute error of the exit direction and of the last position (relative to
the largest |position|), and both runs' backscatter yields with the binomial
standard error.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi


def compare(r32, r64) -> dict:
    """Errors of FP32 records r32 against FP64 records r64 of the same electrons."""
    a, b = r32.els, r64.els
    same = (a["outcome"] == b["outcome"]) & (a["outcome2"] == b["outcome2"]) & (a["n_scatt"] == b["n_scatt"])
    s32, s64 = a[same], b[same]
    relE = np.abs(s32["final_E"].astype(np.float64) - s64["final_E"]) / s64["final_E"]
    ddir = np.abs(s32["dir"].astype(np.float64) - s64["dir"]).max(axis=1)
    scale = max(float(np.abs(s64["pos"]).max()), 1.0)
    dpos = np.abs(s32["pos"].astype(np.float64) - s64["pos"]).max(axis=1) / scale
    n = a.size
    y32 = float((a["outcome"] == _abi.OUT_BKS).mean())
    y64 = float((b["outcome"] == _abi.OUT_BKS).mean())
    return dict(electrons=int(n), identical=float(same.mean()), relE_max=float(relE.max()),
                relE_med=float(np.median(relE)), dir_max=float(ddir.max()), pos_max=float(dpos.max()),
                y32=y32, y64=y64, y_sigma=math.sqrt(max(y64 * (1 - y64), 1e-300) / n),
                ev32=int(r32.events), ev64=int(r64.events))


def study_cases():
    """(name, Transport keyword arguments, seed) of the precision study."""
    cases = [("DS Cu 20 keV 70 deg (seed 1)",
              dict(material="Cu", E0=20000.0, Emin=1000.0, Wc=50.0, s_tilt=70.0, mode="DS"), 1)]
    for tilt in (0.0, 70.0):
        for law in ("classical", "JL", "explicit"):
            cases.append((f"cont Au 30 keV {tilt:.0f} deg {law} (seed 2)",
                          dict(material="Au", E0=30000.0, Emin=1000.0, s_tilt=tilt, mode="cont", bethe_model=law), 2))
    return cases


def run_study(n=1 << 22, cases=None, device=0):
    """Run every case at n electrons in FP32 and FP64; returns [(name, compare dict)]."""
    from .engine import Transport
    out = []
    for name, kw, seed in (cases or study_cases()):
        tr = Transport(device=device, **kw)
        r32 = tr.run(n, seed=seed)
        r64 = tr.run(n, seed=seed, precision=_abi.PREC_FP64)
        out.append((name, compare(r32, r64)))
    return out


def markdown(rows) -> str:
    lines = ["| workload | identical trajectories | exit energy rel. err. max / median | direction abs. err. max "
             "| last position err. max (rel. to max abs pos) | yield FP32 / FP64 |", "|---|---|---|---|---|---|"]
    for name, c in rows:
        lines.append(f"| {name} | {100 * c['identical']:.4f} % | {c['relE_max']:.2e} / {c['relE_med']:.1e} "
                     f"| {c['dir_max']:.2e} | {c['pos_max']:.2e} | {c['y32']:.6f} / {c['y64']:.6f} |")
    return "\n".join(lines)


__all__ = ["compare", "study_cases", "run_study", "markdown"]
