"""Named workloads of the benchmark and test suites.

* headline DS run: Cu, E0 20 keV -> 1 keV, 70 deg, Wc 50 eV, tol 1e-3, 2^22 electrons, seed 1
* batched uneven workload: {Al, Si, Cu, Au} x E0 {5, 10, 20, 30} keV x s_tilt {0, 70} deg
  x {DS, continuous Joy-Luo}; Emin 0.5 keV, Wc 100 eV, tol 1e-3; electrons per
  configuration log-uniform in [2^14, 2^20] drawn from seed 3.
"""
from __future__ import annotations

import numpy as np


def batched_configs(seed=3, scale=1.0):
    """The 64 configurations: list of dicts (material, E0, s_tilt, mode, n)."""
    rng = np.random.default_rng(seed)
    u = rng.random(64)
    out = []
    k = 0
    for mat in ("Al", "Si", "Cu", "Au"):
        for E0 in (5000.0, 10000.0, 20000.0, 30000.0):
            for tilt in (0.0, 70.0):
                for mode in ("DS", "cont"):
                    n = int(2 ** (14 + 6 * u[k]))
                    out.append(dict(material=mat, E0=E0, Emin=500.0, Wc=100.0, s_tilt=tilt, mode=mode,
                                    bethe_model="JL", tol=1e-3, n=max(1, int(n * scale))))
                    k += 1
    return out


def build_transports(configs, device=0, pick_nodes=0, batch_tables=False):
    """One Transport per configuration; pick_nodes > 0 sets the node count of
    the DS channel-pick tables (the count a batch's items use, so that
    per-configuration launches reproduce the batch's records).
    batch_tables: build every configuration's tables in one library call
    (mcd_tables_build_batch; the configurations must share one tolerance),
    instead of one grid search in Python per configuration."""
    from .engine import Transport, build_transports_batch

    if batch_tables:
        tols = {c["tol"] for c in configs}
        if len(tols) != 1:
            raise ValueError("batch_tables needs one tolerance for every configuration")
        tol = tols.pop()
        specs = [dict(material=c["material"], E0=c["E0"], Emin=c["Emin"], Wc=c["Wc"], s_tilt=c["s_tilt"],
                      mode=c["mode"], bethe_model=c["bethe_model"]) for c in configs]
        return build_transports_batch(specs, tol_E=tol, tol_W=tol, device=device, pick_nodes=pick_nodes)
    return [Transport(c["material"], c["E0"], c["Emin"], Wc=c["Wc"], s_tilt=c["s_tilt"], mode=c["mode"],
                      bethe_model=c["bethe_model"], tol_E=c["tol"], tol_W=c["tol"], device=device,
                      pick_nodes=pick_nodes if c["mode"] == "DS" else 0)
            for c in configs]
