"""Writes tests/golden/c1_si10_seed0.npz: golden vectors of the CPU-runnable DS
case (Si 10 keV -> 1 keV, 70 deg, Wc 100 eV, tol_E = tol_W = 1e-2), made with
the oracle (the FP64 C restatement of the reference's step functions) and the
float32 restatement of the reference's table build:

* the grids (Es, Ws) of the Moller and outer-shell Gryzinski tables from the
  tolerance search, and their float32 CDF tables (nE x (nW - 1), masked = 0);
* the per-step trace (channel, W, angles, path, positions, energies,
  directions) and the final records of electrons 0..3 under Philox seed 0.

Run from the repository root: python tests/golden/make_c1_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "mc-discrete_amd")]

N_ELECTRONS, SEED = 4, 0
TRACE_FIELDS = ("electron", "step", "channel", "outcome", "E", "E_loss", "pol", "az", "pathl", "x", "y", "z",
                "energy", "dx", "dy", "dz")
ELS_FIELDS = ("final_E", "dir", "pos", "n_scatt", "outcome", "outcome2", "flags")


def golden():
    """The golden arrays, computed now (dict of numpy arrays)."""
    import oracle as O
    import tables_ref as TR
    from conftest import SI10, build_grids
    from mcd import _abi
    m, gm, gg = build_grids(SI10)
    cdf_m = TR.cdf_table(gm.Es, gm.Ws, m.fermi_e, "Moller", gm.n_el)
    cdf_g = TR.cdf_table(gg.Es, gg.Ws, m.fermi_e, "Gryzinski", gg.n_el, gg.Eb)
    cfg = _abi.make_config(m, SI10["E0"], SI10["Emin"], SI10["Wc"], SI10["s_tilt"], float("inf"), "DS", "classical",
                           0, None)
    els, trace, _ = O.run(cfg, O.Table(gm.Es, gm.Ws, cdf_m), O.Table(gg.Es, gg.Ws, cdf_g), SEED, 0, N_ELECTRONS,
                          trace_cap=100000)
    out = {"Es_m": gm.Es, "Ws_m": gm.Ws, "cdf_m": cdf_m, "Es_g": gg.Es, "Ws_g": gg.Ws, "cdf_g": cdf_g}
    out.update({"trace_" + k: trace[k] for k in TRACE_FIELDS})
    out.update({"els_" + k: els[k] for k in ELS_FIELDS})
    return out


if __name__ == "__main__":
    g = golden()
    np.savez_compressed(os.path.join(HERE, "c1_si10_seed0.npz"), **g)
    print({k: v.shape for k, v in g.items()})
