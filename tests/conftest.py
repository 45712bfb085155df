"""Shared test setup: import paths, the `gpu` marker and common fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "mc-discrete_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through libmcd.so)")


# Configurations used across the tests (names follow the workloads they model).
SI10 = dict(material="Si", E0=10000.0, Emin=1000.0, Wc=100.0, tol=1e-2, s_tilt=70.0)   # CPU-runnable DS case
CU20 = dict(material="Cu", E0=20000.0, Emin=1000.0, Wc=50.0, tol=1e-3, s_tilt=70.0)   # headline DS case


def build_grids(c):
    from mcd.material import Material
    from mcd.tables import gryz_grid, moller_grid

    m = Material(c["material"])
    gm = moller_grid(m, c["E0"], c["Emin"], c["Wc"]).build_grids(c["tol"], c["tol"])
    gg = gryz_grid(m, c["E0"], c["Emin"]).build_grids(c["tol"], c["tol"])
    return m, gm, gg


def oracle_tables(m, gm, gg, cdf_m=None, cdf_g=None):
    """Oracle tables from the float32 restatement (or given CDFs)."""
    import oracle as O
    import tables_ref as TR

    if cdf_m is None:
        cdf_m = TR.cdf_table(gm.Es, gm.Ws, m.fermi_e, "Moller", gm.n_el)
    if cdf_g is None:
        cdf_g = TR.cdf_table(gg.Es, gg.Ws, m.fermi_e, "Gryzinski", gg.n_el, gg.Eb)
    return O.Table(gm.Es, gm.Ws, cdf_m), O.Table(gg.Es, gg.Ws, cdf_g)


@pytest.fixture(scope="session")
def si10():
    m, gm, gg = build_grids(SI10)
    tm, tg = oracle_tables(m, gm, gg)
    return dict(mat=m, gm=gm, gg=gg, tm=tm, tg=tg, **SI10)


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(12345)


def identical(g, o):
    """Trajectories with the same outcomes and scatter count (records of the
    HIP kernels against oracle records)."""
    return (g["outcome"] == o["outcome"]) & (g["outcome2"] == o["outcome2"]) & (g["n_scatt"] == o["n_scatt"])


def mfp_at(cfg, E):
    """Mean free path [A] of a configuration at energy E from the oracle's
    sigma set (DS: every channel; continuous: Rutherford only)."""
    import oracle as O
    if cfg.mode == 0:
        sig, _ = O.sigma_set(cfg, float(E))
        tot = float(np.sum(sig))
    else:
        tot = O.lib().oracle_sigma_ruth_model(float(E), int(cfg.mat.Z), int(cfg.elastic))
    return 1.0 / (cfg.mat.atnd * tot) * 1e14


def check_values(g, o, same, rel_E=1e-5, dir_abs=1e-5, pos_rel=1e-5, length=40.0):
    """Energies and directions of identical trajectories within tolerance; the
    position error is measured against the trajectory's length scale, since a
    coordinate can sit near zero: its total path length when the records
    carry it (oracle records do), else `length` (a mean free path) x
    sqrt(events)."""
    if not np.any(same):
        return 0.0, 0.0, 0.0
    E_err = np.abs(g["final_E"][same] - o["final_E"][same]) / o["final_E"][same]
    assert E_err.max() <= rel_E, E_err.max()
    d_err = np.abs(g["dir"][same] - o["dir"][same]).max()
    assert d_err <= dir_abs, d_err
    if "path" in o.dtype.names:
        ref = o["path"][same]
    else:
        ref = length * np.sqrt(o["n_scatt"][same] + 1.0)
    scale = np.maximum(np.linalg.norm(o["pos"][same], axis=1), ref)
    p_err = (np.linalg.norm(g["pos"][same] - o["pos"][same], axis=1) / scale).max()
    assert p_err <= pos_rel, p_err
    return E_err.max(), d_err, p_err
