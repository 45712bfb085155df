"""GPU parity of the HIP transport against the FP64 oracle (same Philox stream,
same table bytes), through the C-ABI (libmcd.so)."""
import math

import numpy as np
import pytest

import oracle as O
import tables_ref as TR
from conftest import CU20, SI10, check_values, identical
from divergence import classify_divergence
from mcd import _abi

pytestmark = pytest.mark.gpu


def transport(c, **kw):
    from mcd import engine
    return engine.Transport(c["material"], c["E0"], c["Emin"], Wc=c["Wc"], s_tilt=c["s_tilt"], tol_E=c["tol"],
                            tol_W=c["tol"], **kw)


def oracle_for(tr):
    """Oracle tables holding the device-built CDF bytes."""
    gm, gg = tr.grids
    return O.Table(gm.Es, gm.Ws, tr.cdf(0)), O.Table(gg.Es, gg.Ws, tr.cdf(1))




@pytest.fixture(scope="module")
def si10():
    tr = transport(SI10)
    tm, tg = oracle_for(tr)
    return tr, tm, tg


@pytest.fixture(scope="module")
def cu20():
    tr = transport(CU20)
    tm, tg = oracle_for(tr)
    return tr, tm, tg


def test_device_tables_match_float32_restatement(si10, cu20):
    """Device FP64 CDF rows (rounded once to float32) against the reference's
    float32 arithmetic (probTables.py:415-462): same mask, values within the
    float32 cumulative-sum error."""
    for tr, _, _ in (si10, cu20):
        m = tr.mat
        gm, gg = tr.grids
        ref = [TR.cdf_table(gm.Es, gm.Ws, m.fermi_e, "Moller", gm.n_el),
               TR.cdf_table(gg.Es, gg.Ws, m.fermi_e, "Gryzinski", gg.n_el, gg.Eb)]
        for which in (0, 1):
            dev = tr.cdf(which)
            assert np.array_equal(dev == 0, ref[which] == 0)
            assert np.abs(dev - ref[which]).max() < 5e-6


def test_device_fp32_table_build_matches_restatement(si10):
    from mcd import engine
    tr32 = engine.Transport("Si", 10000.0, 1000.0, Wc=100.0, s_tilt=70.0, tol_E=1e-2, tol_W=1e-2,
                            table_precision=_abi.PREC_FP32, grids=si10[0].grids)
    m = tr32.mat
    gm, _ = tr32.grids
    ref = TR.cdf_table(gm.Es, gm.Ws, m.fermi_e, "Moller", gm.n_el)
    dev = tr32.cdf(0)
    assert np.array_equal(dev == 0, ref == 0)
    assert np.abs(dev - ref).max() < 1e-6


@pytest.mark.parametrize("seed", [0, 9])
def test_fp64_transport_matches_oracle_exactly(si10, seed):
    tr, tm, tg = si10
    n = 2000
    o, _, ev = O.run(tr.cfg, tm, tg, seed, 0, n, nthreads=16)
    r = tr.run(n, seed=seed, precision=_abi.PREC_FP64)
    same = identical(r.els, o)
    assert same.all(), np.flatnonzero(~same)[:10]
    assert r.events == ev
    check_values(r.els, o, same, rel_E=1e-7, dir_abs=1e-7, pos_rel=1e-7)


@pytest.mark.parametrize("seed", [0, 9])
def test_fp32_transport_matches_oracle(si10, seed):
    tr, tm, tg = si10
    n = 2000
    o, _, ev = O.run(tr.cfg, tm, tg, seed, 0, n, nthreads=16)
    r = tr.run(n, seed=seed)
    same = identical(r.els, o)
    assert same.mean() >= 0.999, same.mean()
    check_values(r.els, o, same)
    c = r.counters
    assert int(c[_abi.CNT_ELECTRONS]) == n
    assert int(c[_abi.CNT_BKS]) == int((r.els["outcome"] == _abi.OUT_BKS).sum())


def test_step_rows_match_oracle_trace(si10):
    """Per-step scatter rows (type, W, angles, path, positions, energies,
    directions) of the first 16 electrons against the oracle's trace."""
    tr, tm, tg = si10
    n, rec = 64, 16
    _, tro, _ = O.run(tr.cfg, tm, tg, 4, 0, rec, trace_cap=100000)
    r = tr.run(n, seed=4, scats=True, scat_max_electron=rec)
    s = r.scats
    assert len(s) == len(tro)
    assert np.array_equal(s["electron"], tro["electron"]) and np.array_equal(s["step"], tro["step"])
    assert np.array_equal(s["channel"], tro["channel"])
    assert np.array_equal(s["E_loss"], tro["E_loss"].astype(np.float32))
    assert np.allclose(s["E"], tro["E"], rtol=1e-6)
    assert np.allclose(s["energy"], tro["energy"], rtol=1e-6)
    assert np.allclose(s["pol"], tro["pol"], rtol=1e-5, atol=1e-6)
    assert np.allclose(s["az"], tro["az"], rtol=1e-6)
    assert np.allclose(s["pathl"], tro["pathl"], rtol=2e-5)
    for k in ("dx", "dy", "dz"):
        assert np.abs(s[k] - tro[k]).max() < 2e-5
    assert np.array_equal(s["outcome"], tro["outcome"].astype(np.uint8))


def test_step_rows_match_committed_golden_trace():
    """The device, given the committed golden float32 tables of the
    CPU-runnable DS case (tests/golden/c1_si10_seed0.npz), reproduces the
    committed per-step trace and records of its electrons 0..3 (seed 0): same
    grids, channels, W values, outcomes and step counts; energies, angles,
    paths and directions to FP32 accuracy."""
    import os
    from conftest import SI10
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_si10_seed0.npz"))
    tr = transport(SI10, cdfs=(g["cdf_m"], g["cdf_g"]))
    gm, gg = tr.grids
    for a, b in ((gm.Es, g["Es_m"]), (gm.Ws, g["Ws_m"]), (gg.Es, g["Es_g"]), (gg.Ws, g["Ws_g"])):
        assert np.array_equal(a, b)
    assert np.array_equal(tr.cdf(0), g["cdf_m"]) and np.array_equal(tr.cdf(1), g["cdf_g"])
    n = len(g["els_final_E"])
    r = tr.run(n, seed=0, scats=True)
    s = r.scats
    assert len(s) == len(g["trace_step"])
    assert np.array_equal(s["electron"], g["trace_electron"]) and np.array_equal(s["step"], g["trace_step"])
    assert np.array_equal(s["channel"], g["trace_channel"])
    assert np.array_equal(s["outcome"], g["trace_outcome"].astype(np.uint8))
    assert np.array_equal(s["E_loss"], g["trace_E_loss"].astype(np.float32))
    for k, rtol in (("E", 1e-6), ("energy", 1e-6), ("az", 1e-6), ("pathl", 2e-5)):
        assert np.allclose(s[k], g["trace_" + k], rtol=rtol), k
    assert np.allclose(s["pol"], g["trace_pol"], rtol=1e-5, atol=1e-6)
    for k in ("dx", "dy", "dz"):
        assert np.abs(s[k] - g["trace_" + k]).max() < 2e-5, k
    assert np.array_equal(r.els["outcome"], g["els_outcome"]) and np.array_equal(r.els["n_scatt"], g["els_n_scatt"])
    assert np.allclose(r.els["final_E"], g["els_final_E"], rtol=1e-6)


def test_headline_config_first_1e4_electrons(cu20):
    """Cu 20 keV, 70 deg, Wc 50, tol 1e-3: the first 10^4 electrons of the
    headline workload trajectory by trajectory."""
    tr, tm, tg = cu20
    n = 10000
    o, _, ev = O.run(tr.cfg, tm, tg, 1, 0, n, nthreads=16)
    r = tr.run(n, seed=1)
    same = identical(r.els, o)
    assert same.mean() >= 0.999, same.mean()
    check_values(r.els, o, same)
    r64 = tr.run(n, seed=1, precision=_abi.PREC_FP64)
    assert identical(r64.els, o).all()


def test_headline_first_1e4_electrons_per_step(cu20):
    """The headline workload (Cu 20 keV, 70 deg, Wc 50 eV, tol 1e-3, seed 1),
    its first 10^4 electrons step by step against the oracle trace
    (scattering.py:39-59, singleScatter.py:10-41): the FP64 kernel's scatter
    rows equal the trace (channel, loss, outcome and energies bit for bit as
    float32); the FP32 kernel's rows leave the trace only where an FP32
    position or direction decides an exit, or where u ties a boundary, never
    through a missed guard-band pick or a different table lookup."""
    tr, tm, tg = cu20
    n, chunk = 10000, 2500
    g64, g32, otr = [], [], []
    for first in range(0, n, chunk):
        _, t, _ = O.run(tr.cfg, tm, tg, 1, first, chunk, trace_cap=chunk * 2500)
        otr.append(t)
        g64.append(tr.run(chunk, seed=1, first=first, scats=True, precision=_abi.PREC_FP64).scats)
        g32.append(tr.run(chunk, seed=1, first=first, scats=True).scats)
    o = np.concatenate(otr)
    r64 = np.concatenate(g64)
    r32 = np.concatenate(g32)
    print(f"{len(o)} oracle rows, {len(r64)} FP64 rows, {len(r32)} FP32 rows", flush=True)
    assert len(r64) == len(o)
    assert np.array_equal(r64["electron"], o["electron"]) and np.array_equal(r64["step"], o["step"])
    assert np.array_equal(r64["channel"], o["channel"])
    assert np.array_equal(r64["E_loss"], o["E_loss"].astype(np.float32))
    assert np.array_equal(r64["outcome"], o["outcome"].astype(np.uint8))
    assert np.array_equal(r64["E"], o["E"].astype(np.float32))
    assert np.array_equal(r64["energy"], o["energy"].astype(np.float32))
    cls = classify_divergence(tr.cfg, (tm, tg), r32, o)
    counts = {}
    for _, _, c in cls:
        counts[c] = counts.get(c, 0) + 1
    print(f"FP32 electrons leaving the trace: {len(cls)} of {n}; first divergence classes: {counts}")
    for e, st, c in cls:
        print(f"  electron {e} step {st}: {c}")
    assert counts.get("pick_miss", 0) == 0 and counts.get("table_miss", 0) == 0, counts
    assert len(cls) <= 0.001 * n


def test_headline_config_full_size_statistics(cu20):
    """2^22 electrons: FP32 vs the all-FP64 GPU run of the same stream
    (identical trajectories), and vs an independent FP64 sample (yield within
    3 sigma, two-sample KS on BSE energy and exit angle, p > 0.01)."""
    from scipy import stats
    tr, _, _ = cu20
    n = 1 << 22
    r32 = tr.run(n, seed=1)
    r64 = tr.run(n, seed=1, precision=_abi.PREC_FP64)
    same = identical(r32.els, r64.els)
    assert same.mean() >= 0.999, same.mean()
    e32 = r32.els[same]
    e64 = r64.els[same]
    assert (np.abs(e32["final_E"] - e64["final_E"]) / e64["final_E"]).max() < 1e-5
    ind = tr.run(n, seed=2, precision=_abi.PREC_FP64)
    b32 = r32.els["outcome"] == _abi.OUT_BKS
    bind = ind.els["outcome"] == _abi.OUT_BKS
    y32, yind = b32.mean(), bind.mean()
    sig = math.sqrt(y32 * (1 - y32) / n + yind * (1 - yind) / n)
    assert abs(y32 - yind) < 3 * sig, (y32, yind, sig)
    assert stats.ks_2samp(r32.els["final_E"][b32], ind.els["final_E"][bind]).pvalue > 0.01
    ang32 = np.degrees(np.arccos(-r32.els["dir"][b32][:, 2]))
    angind = np.degrees(np.arccos(-ind.els["dir"][bind][:, 2]))
    assert stats.ks_2samp(ang32, angind).pvalue > 0.01


@pytest.mark.parametrize("tilt", [0.0, 70.0])
@pytest.mark.parametrize("bethe", ["classical", "JL", "explicit"])
def test_continuous_mode_matches_oracle(tilt, bethe):
    """Au 30 keV continuous mode with each Bethe law (trajectory_cont_cl,
    singleScatter.py:205-239; stoppingPowers.py:112-215)."""
    from mcd import engine
    tr = engine.Transport("Au", 30000.0, 1000.0, s_tilt=tilt, mode="cont", bethe_model=bethe)
    n = 2000
    o, _, ev = O.run(tr.cfg, None, None, 2, 0, n, nthreads=16)
    r64 = tr.run(n, seed=2, precision=_abi.PREC_FP64)
    assert identical(r64.els, o).all()
    r = tr.run(n, seed=2)
    same = identical(r.els, o)
    assert same.mean() >= 0.999, same.mean()
    check_values(r.els, o, same)


@pytest.mark.parametrize("elastic", ["Ruth_vanilla_wDefl", "Ruth_nigram", "Ruth_nigram_wDefl"])
def test_ds_elastic_models_match_oracle(si10, elastic):
    """DS Si 10 keV with each Rutherford model of the input key 'elastic'
    (scattering.py:483-493; crossSections.py:62-147): the FP64 kernel equals
    the oracle trajectory by trajectory, the FP32 kernel to the FP32 bar."""
    tr0, tm, tg = si10
    tr = transport(SI10, elastic=elastic, grids=tr0.grids)
    assert tr.cfg.elastic == _abi.elastic_model(elastic, "DS")
    n = 2000
    o, _, ev = O.run(tr.cfg, tm, tg, 0, 0, n, nthreads=16)
    r64 = tr.run(n, seed=0, precision=_abi.PREC_FP64)
    assert identical(r64.els, o).all() and r64.events == ev
    r = tr.run(n, seed=0)
    same = identical(r.els, o)
    assert same.mean() >= 0.999, same.mean()
    check_values(r.els, o, same)
    # a different elastic cross section gives different trajectories
    base = tr0.run(n, seed=0)
    assert not np.array_equal(base.els["n_scatt"], r.els["n_scatt"])


@pytest.mark.parametrize("elastic", ["Ruth_nigram", "Ruth_nigram_wDefl", "Ruth_vanilla_wDefl"])
def test_continuous_elastic_models_match_oracle(elastic):
    """Continuous mode resolves the names as the reference does
    (scattering.py:780-789: the wDefl names fall back to the base models)."""
    from mcd import engine
    tr = engine.Transport("Au", 30000.0, 1000.0, s_tilt=70.0, mode="cont", bethe_model="JL", elastic=elastic)
    assert tr.cfg.elastic == (_abi.ELASTIC_NIGRAM if "nigram" in elastic else _abi.ELASTIC_RUTH)
    n = 2000
    o, _, ev = O.run(tr.cfg, None, None, 2, 0, n, nthreads=16)
    r64 = tr.run(n, seed=2, precision=_abi.PREC_FP64)
    assert identical(r64.els, o).all()
    r = tr.run(n, seed=2)
    same = identical(r.els, o)
    assert same.mean() >= 0.999, same.mean()
    check_values(r.els, o, same)


@pytest.mark.parametrize("material,E0,Emin", [("Au", 30000.0, 1000.0), ("Al", 5000.0, 500.0)])
def test_ds_other_materials_match_oracle(material, E0, Emin):
    """DS parity at 11 channels (Au: Rutherford, Moller, Quinn and eight
    Gryzinski shells in the pick) and for Al at 5 keV: the FP64 kernel equals
    the oracle trajectory by trajectory, the FP32 kernel within the FP32
    bar, and the per-step channel sequence of the first electrons matches."""
    from mcd import engine
    tr = engine.Transport(material, E0, Emin, Wc=100.0, s_tilt=70.0, tol_E=1e-2, tol_W=1e-2)
    tm, tg = oracle_for(tr)
    n = 2000
    o, _, ev = O.run(tr.cfg, tm, tg, 3, 0, n, nthreads=16)
    r64 = tr.run(n, seed=3, precision=_abi.PREC_FP64)
    assert identical(r64.els, o).all() and r64.events == ev
    check_values(r64.els, o, identical(r64.els, o), rel_E=1e-7, dir_abs=1e-7, pos_rel=1e-7)
    r = tr.run(n, seed=3)
    same = identical(r.els, o)
    assert same.mean() >= 0.999, same.mean()
    check_values(r.els, o, same)
    _, tro, _ = O.run(tr.cfg, tm, tg, 3, 0, 8, trace_cap=400000)
    s = tr.run(8, seed=3, scats=True, precision=_abi.PREC_FP64).scats
    assert len(s) == len(tro) and np.array_equal(s["channel"], tro["channel"])
    assert np.array_equal(s["E_loss"], tro["E_loss"].astype(np.float32))
    assert len(np.unique(s["channel"])) >= 4


@pytest.mark.parametrize("bethe", ["classical", "JL", "explicit"])
def test_continuous_step_rows_match_oracle_trace(bethe):
    """Per-step rows of continuous mode (path, Bethe loss W = l * S(E_pre),
    Rutherford angle, azimuth, energies, directions) for the first 12
    electrons of Au 30 keV at 70 deg, from the FP64 and the FP32 kernels,
    against the oracle's trace."""
    from mcd import engine
    tr = engine.Transport("Au", 30000.0, 1000.0, s_tilt=70.0, mode="cont", bethe_model=bethe)
    rec = 12
    _, tro, _ = O.run(tr.cfg, None, None, 5, 0, rec, trace_cap=400000)
    for prec, tol in ((_abi.PREC_FP64, 1e-6), (_abi.PREC_FP32, 2e-5)):
        s = tr.run(32, seed=5, scats=True, scat_max_electron=rec, precision=prec).scats
        if prec == _abi.PREC_FP32:
            # FP32 trajectories may part from the FP64 ones; compare the
            # electrons whose step counts agree
            keep = np.isin(s["electron"], [e for e in range(rec)
                                           if (s["electron"] == e).sum() == (tro["electron"] == e).sum()])
            assert keep.mean() >= 0.9, keep.mean()
            s = s[keep]
            t = tro[np.isin(tro["electron"], np.unique(s["electron"]))]
        else:
            t = tro
        assert len(s) == len(t)
        assert np.array_equal(s["electron"], t["electron"]) and np.array_equal(s["step"], t["step"])
        assert np.array_equal(s["channel"], t["channel"])
        assert np.array_equal(s["outcome"], t["outcome"].astype(np.uint8))
        for k in ("E", "energy", "pathl", "az"):
            assert np.allclose(s[k], t[k], rtol=tol, atol=0), (prec, k)
        assert np.allclose(s["E_loss"], t["E_loss"], rtol=10 * tol, atol=1e-3), prec
        assert np.allclose(s["pol"], t["pol"], rtol=10 * tol, atol=1e-6), prec
        for k in ("dx", "dy", "dz"):
            assert np.abs(s[k] - t[k]).max() < 20 * tol, (prec, k)


def test_split_launches_equal_one_launch(cu20):
    """Philox keys use the global electron id: two half-range launches give
    bit-identical records to one launch (the sharding invariance a multi-GPU
    split relies on)."""
    tr = cu20[0]
    n = 1 << 16
    full = tr.run(n, seed=5)
    a = tr.run(n // 2, seed=5)
    b = tr.run(n // 2, seed=5, first=n // 2)
    both = np.concatenate([a.els, b.els])
    assert both.tobytes() == full.els.tobytes()


def test_electron_ids_across_2_32_match_oracle(si10):
    """Electrons 2^32 - 24 .. 2^32 + 23: the FP64 kernel equals the oracle bit
    for bit (bits 32..47 of the id in Philox counter word 1), every guard form
    gives the same FP32 records, the scatter rows (which hold the low 32 bits
    of the id) come back in electron order across the wrap and equal the
    oracle's trace, and a range past 2^48 is refused before launch."""
    tr, tm, tg = si10
    first, n = (1 << 32) - 24, 48
    o, tro, ev = O.run(tr.cfg, tm, tg, 6, first, n, trace_cap=200000)
    r = tr.run(n, seed=6, first=first, precision=_abi.PREC_FP64)
    assert identical(r.els, o).all() and r.events == ev
    check_values(r.els, o, identical(r.els, o), rel_E=1e-7, dir_abs=1e-7, pos_rel=1e-7)
    base = tr.run(n, seed=6, first=first, guard="inline")
    for guard in ("deferred", "service"):
        assert tr.run(n, seed=6, first=first, guard=guard).els.tobytes() == base.els.tobytes(), guard
    s = tr.run(n, seed=6, first=first, precision=_abi.PREC_FP64, scats=True).scats
    assert len(s) == len(tro)
    assert np.array_equal(s["electron"], tro["electron"]) and np.array_equal(s["step"], tro["step"])
    assert np.array_equal(s["channel"], tro["channel"])
    with pytest.raises(RuntimeError, match="2\\^48"):
        tr.run(8, first=(1 << 48) - 4)


@pytest.mark.parametrize("n", [0, 1, 15, 17, 1000, (1 << 20) + 3])
def test_every_electron_runs_once_at_any_count(cu20, n):
    """Wavefronts take electron indices in blocks of 16 per atomic: for counts
    that are not multiples of the block (or of a wavefront) every electron is
    transported exactly once -- its record is written and equals the record of
    the same electron in a larger launch (records are keyed by electron id).
    An empty launch returns at once and leaves the counters untouched."""
    tr = cu20[0]
    for guard in ("inline", "deferred", "service"):
        r = tr.run(n, seed=7, first=3, guard=guard)
        assert int(r.counters[_abi.CNT_ELECTRONS]) == n
        outs = r.els["outcome"]
        assert np.isin(outs, [_abi.OUT_BKS, _abi.OUT_ABS]).all(), guard
        # an exit step that also ends below Emin records a second 'abs' outcome
        # (saveOutcomes runs twice, electron.py:52-54,105-113)
        o2 = r.els["outcome2"]
        assert int(r.counters[_abi.CNT_BKS]) == int((outs == _abi.OUT_BKS).sum())
        assert int(r.counters[_abi.CNT_ABS]) == int((outs == _abi.OUT_ABS).sum() + (o2 == _abi.OUT_ABS).sum())
        big = tr.run(n + 40, seed=7, first=3, guard=guard)
        assert big.els[:n].tobytes() == r.els.tobytes(), guard
        assert int(r.counters[_abi.CNT_EVENTS]) == int(r.els["n_scatt"].astype(np.int64).sum()), guard
        # compaction period: lanes refilled every 8 steps give the same records
        assert tr.run(n, seed=7, first=3, guard=guard, refill_period=8).els.tobytes() == r.els.tobytes(), guard


def test_max_scatt_and_thin_film_outcomes():
    """maxScatt -> 'long' (singleScatter.py:19-23) and a finite thickness ->
    'trsm' (electron.py:72-81) against the oracle."""
    from mcd import engine
    tr = engine.Transport("Si", 10000.0, 1000.0, Wc=100.0, s_tilt=0.0, tol_E=1e-2, tol_W=1e-2, thickness=3000.0,
                          max_scatt=80)
    tm, tg = oracle_for(tr)
    n = 2000
    o, _, _ = O.run(tr.cfg, tm, tg, 6, 0, n, nthreads=16)
    r = tr.run(n, seed=6)
    for out in (_abi.OUT_LONG, _abi.OUT_TRSM, _abi.OUT_BKS):
        assert (o["outcome"] == out).sum() > 50
    same = identical(r.els, o)
    assert same.mean() >= 0.999
    check_values(r.els, o, same)
    # the 'long' record carries the compensated position, like every other outcome
    check_values(r.els, o, same & (o["outcome"] == _abi.OUT_LONG))
    for guard in ("service", "deferred"):  # the 'long' and 'trsm' paths in the other guard forms
        assert tr.run(n, seed=6, guard=guard).els.tobytes() == r.els.tobytes(), guard


def test_continuous_max_scatt_and_thin_film_outcomes():
    """Continuous mode (Au 30 keV, JL) with maxScatt 150 and a 3000 A film:
    'long', 'trsm' and 'bks' outcomes against the oracle, FP64 kernel exactly
    and the FP32 kernel within its bar."""
    from mcd import engine
    tr = engine.Transport("Au", 30000.0, 1000.0, s_tilt=0.0, mode="cont", bethe_model="JL", thickness=3000.0,
                          max_scatt=150)
    n = 2000
    o, _, ev = O.run(tr.cfg, None, None, 6, 0, n, nthreads=16)
    for out in (_abi.OUT_LONG, _abi.OUT_TRSM, _abi.OUT_BKS):
        assert (o["outcome"] == out).sum() > 200
    r64 = tr.run(n, seed=6, precision=_abi.PREC_FP64)
    assert identical(r64.els, o).all() and r64.events == ev
    r = tr.run(n, seed=6)
    same = identical(r.els, o)
    assert same.mean() >= 0.999, same.mean()
    check_values(r.els, o, same)
    check_values(r.els, o, same & (o["outcome"] == _abi.OUT_LONG))
    assert int(r.counters[_abi.CNT_ELECTRONS]) == n


def test_invalid_configuration_is_rejected_before_launch():
    from mcd import engine
    from mcd._lib import MCDError
    with pytest.raises((MCDError, ValueError)):
        engine.Transport("Au", 30000.0, 400.0, mode="cont", bethe_model="classical")  # S(Emin) < 0
    with pytest.raises((MCDError, ValueError)):
        engine.Transport("Cu", 1000.0, 2000.0, Wc=50.0)  # Emin > E0


def test_scatter_buffer_overflow_is_reported(si10):
    from mcd.engine import ScatterOverflow
    tr = si10[0]
    with pytest.raises(ScatterOverflow, match="overflow"):
        tr.run(64, seed=1, scats=True, scat_capacity=100)


@pytest.mark.parametrize("guard", [None, "deferred", "service"])
def test_pipelined_chunks_on_two_streams_equal_one_launch(cu20, guard):
    """Launches alternating between two streams, each with its own counters
    block (which holds its work counters) and, in the deferred form, the
    stream's own park buffer, run concurrently on the device and give the
    records of one launch over the whole range."""
    tr = cu20[0]
    n = (1 << 18) + 5
    whole = tr.run(n, seed=3, first=11)
    got, events = [], 0
    print(f"run_chunks guard={guard}", flush=True)
    for start, r in tr.run_chunks(n, 1 << 16, seed=3, first=11, streams=2, guard=guard):
        assert start == 11 + sum(len(x) for x in got)
        got.append(r.els)
        events += int(r.counters[_abi.CNT_EVENTS])
    assert np.concatenate(got).tobytes() == whole.els.tobytes()
    assert events == int(whole.counters[_abi.CNT_EVENTS])


def test_scatter_rows_split_into_ranges_equal_one_run(si10):
    """run_rows splits an electron range whose scatter rows overflow the row
    buffer into halves until each fits; the records, rows and counters equal
    one run with room for every row."""
    from mcd.multiScatter import run_rows
    tr = si10[0]
    n = 512
    whole = tr.run(n, seed=0, first=5, scats=True)
    per_el = len(whole.scats) / n
    parts = run_rows(tr, n, 0, 5, scat_capacity=int(40 * per_el))  # about 1/13 of the rows per device run
    assert parts.els.tobytes() == whole.els.tobytes()
    assert parts.scats.tobytes() == whole.scats.tobytes()
    assert (parts.counters[:_abi.CNT_SCAT_ROWS] == whole.counters[:_abi.CNT_SCAT_ROWS]).all()


def test_scatter_rows_packed_on_the_device_equal_whole_rows(si10):
    """mcd_scatter_pack: the requested fields of the scatter rows, packed into
    columns on the device, equal those fields of the whole rows bit for bit
    (after the same (electron, step) ordering), for the field sets the
    scatter_output names need, every field, and ranges split by run_rows;
    an empty mask, an unknown field and a short buffer are refused."""
    import ctypes as C

    import torch
    from mcd import engine
    from mcd._lib import MCDError, check, lib
    from mcd.multiScatter import run_rows
    tr = si10[0]
    n = 300
    whole = tr.run(n, seed=2, first=11, scats=True)
    sets = [_abi.scat_fields_for(["type", "E_loss"]), _abi.scat_fields_for(["position", "energy", "direction"]),
            list(_abi.SCATTER_DTYPE.names), ["electron", "step"]]
    for fields in sets:
        part = tr.run(n, seed=2, first=11, scats=True, scat_fields=fields)
        assert part.els.tobytes() == whole.els.tobytes()
        assert set(part.scats.dtype.names) == set(fields) | {"electron", "step"}
        for f in part.scats.dtype.names:
            assert part.scats[f].tobytes() == whole.scats[f].tobytes(), (fields, f)
    split = run_rows(tr, n, 2, 11, scat_capacity=len(whole.scats) // 5, scat_fields=["channel", "E"])
    for f in split.scats.dtype.names:
        assert split.scats[f].tobytes() == whole.scats[f].tobytes(), f
    ctx = engine.context(0)
    rows = torch.empty(64 * 8, dtype=torch.uint8, device="cuda")
    need = C.c_size_t(0)
    for mask in (0, 1 << 18):
        with pytest.raises(MCDError) as e:
            check(lib().mcd_scatter_pack(ctx, rows.data_ptr(), 8, mask, None, 0, C.byref(need), None), ctx)
        assert e.value.code == _abi.ERR_INVALID
    check(lib().mcd_scatter_pack(ctx, rows.data_ptr(), 8, 0b11, None, 0, C.byref(need), None), ctx)
    assert need.value == 2 * 256
    out = torch.empty(need.value - 1, dtype=torch.uint8, device="cuda")
    with pytest.raises(MCDError) as e:
        check(lib().mcd_scatter_pack(ctx, rows.data_ptr(), 8, 0b11, out.data_ptr(), out.numel(), C.byref(need), None),
              ctx)
    assert e.value.code == _abi.ERR_INVALID


def test_reference_statistical_checks_on_gpu_output(si10):
    """The reference's chi-square channel test and Watson U^2 polar-angle test
    (tests/testThings.py:321-434) applied to GPU scatter rows."""
    from scipy import stats
    from test_oracle_algorithms import watson_u2
    tr = si10[0]
    r = tr.run(4000, seed=13, scats=True, scat_capacity=4000 * 600)
    first = r.scats[r.scats["step"] == 0]
    sig, lab = O.sigma_set(tr.cfg, 10000.0)
    p = sig / sig.sum()
    obs = np.array([(first["channel"] == l).sum() for l in lab], dtype=float)
    exp = p * len(first)
    keep = exp >= 5
    chi2 = (((obs - exp) ** 2 / exp)[keep]).sum()
    assert chi2 < stats.chi2.ppf(0.999, keep.sum() - 1)
    ru = first[first["channel"] == _abi.CH_RUTH]
    mc = np.degrees(np.arccos(2 * ru["pol"].astype(np.float64) - 1))[:200]
    a = 3.4 * 14 ** 0.666667 / 10000.0
    u = np.random.default_rng(3).random(200)
    theory = np.degrees(np.arccos(1 - (2 * a * u) / (1 + a - u)))
    assert watson_u2(theory, mc) < 0.152


@pytest.mark.parametrize("guard,shape", [
    ("deferred", {}),
    ("service", {}),
    ("service", dict(block=128, waves_per_simd=8)),
    ("service", dict(block=1024, waves_per_simd=8)),
    ("service", dict(block=512, waves_per_simd=6)),
])
def test_guard_forms_equal_inline_guard(cu20, si10, guard, shape):
    """The two-kernel DS form (FP64 guard picks parked and resumed in a second
    kernel) and the served form (guard picks posted through LDS to wavefront 0
    of the workgroup, at any workgroup size: one transport wavefront up to
    fifteen) write the same records, scatter rows and counters as the
    one-kernel form with the guard inline."""
    tr = cu20[0]
    n = 1 << 20
    a = tr.run(n, seed=1, guard=guard, **shape)
    b = tr.run(n, seed=1)
    assert a.els.tobytes() == b.els.tobytes()
    for k in (_abi.CNT_EVENTS, _abi.CNT_BKS, _abi.CNT_ABS, _abi.CNT_FP64_PICKS):
        assert int(a.counters[k]) == int(b.counters[k]), k
    assert int(a.counters[_abi.CNT_FP64_PICKS]) > 0  # guard picks were settled in FP64
    assert int(a.counters[_abi.CNT_GUARD_TIMEOUTS]) == 0
    trs = si10[0]
    ra = trs.run(512, seed=0, scats=True, guard=guard, **shape)
    rb = trs.run(512, seed=0, scats=True, guard="inline")
    assert ra.els.tobytes() == rb.els.tobytes()
    assert ra.scats.tobytes() == rb.scats.tobytes()


def test_library_default_from_2_21_electrons_is_served_with_the_same_records(cu20):
    """With no guard placement and no launch shape, a DS launch of 2^21
    electrons runs the served form and one of 2^21 - 1 the inline form; both
    write the inline form's records."""
    tr = cu20[0]
    n = 1 << 21
    a = tr.run(n, seed=2)
    b = tr.run(n, seed=2, guard="inline")
    assert a.els.tobytes() == b.els.tobytes()
    c = tr.run(n - 1, seed=2)
    assert c.els.tobytes() == b.els[: n - 1].tobytes()


@pytest.mark.parametrize("material,E0,Emin", [("Au", 30000.0, 1000.0), ("Al", 5000.0, 500.0)])
def test_service_guard_other_channel_counts(material, E0, Emin):
    """The served guard at 11 channels (Au, four Gryzinski shells listed plus
    Moller, Rutherford, Quinn) and at 6 (Al) writes the inline form's records."""
    from mcd import engine
    tr = engine.Transport(material, E0, Emin, Wc=100.0, s_tilt=70.0, tol_E=1e-2, tol_W=1e-2)
    a = tr.run(1 << 17, seed=4, guard="service")
    b = tr.run(1 << 17, seed=4, guard="inline")
    assert a.els.tobytes() == b.els.tobytes()
    assert int(a.counters[_abi.CNT_FP64_PICKS]) == int(b.counters[_abi.CNT_FP64_PICKS])


def test_served_guard_timeout_is_reported(cu20, monkeypatch):
    """The served guard's wait for a reply is bounded; driven past the bound
    from the host (MCD_GUARD_SPIN_CAP=1: a transport wavefront polls its
    mailbox once), the launch still ends normally, the steps that gave up are
    flagged and counted, and the host refuses the records (GuardError).  With
    the default bound the same launch has no timeouts and the inline form's
    records."""
    import torch
    from mcd.engine import GuardError
    tr = cu20[0]
    n = 1 << 16
    monkeypatch.setenv("MCD_GUARD_SPIN_CAP", "1")
    with pytest.raises(GuardError):
        tr.run(n, seed=1, guard="service")
    dev = torch.device("cuda", 0)
    els = torch.empty(n * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
    tr.launch(n, els, cnt, seed=1, guard="service")
    torch.cuda.synchronize(dev)
    c = cnt.cpu().numpy().view(np.uint64)
    rec = els.cpu().numpy().view(_abi.ELECTRON_DTYPE)
    assert int(c[_abi.CNT_GUARD_TIMEOUTS]) > 0 and int(c[_abi.CNT_ELECTRONS]) == n
    assert ((rec["flags"] & _abi.FLAG_GUARD_TIMEOUT) != 0).any()
    monkeypatch.delenv("MCD_GUARD_SPIN_CAP")
    ok = tr.run(n, seed=1, guard="service")
    assert int(ok.counters[_abi.CNT_GUARD_TIMEOUTS]) == 0
    assert ok.els.tobytes() == tr.run(n, seed=1, guard="inline").els.tobytes()


def test_launch_shape_errors_are_reported(cu20):
    """The C-ABI rejects two guard placements at once and out-of-range shapes
    with MCD_ERR_INVALID before launching anything."""
    import ctypes as C
    import torch
    from mcd import engine
    from mcd._lib import MCDError, check, lib
    tr = cu20[0]
    dev = torch.device("cuda", 0)
    els = torch.empty(64 * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
    ctx = engine.context(0)
    s = torch.cuda.current_stream(0).cuda_stream
    bad = [(0, 0, _abi.LAUNCH_INLINE_GUARD | _abi.LAUNCH_SERVICE_GUARD),
           (0, 0, _abi.LAUNCH_DEFER_GUARD | _abi.LAUNCH_SERVICE_GUARD),
           (1088, 0, _abi.LAUNCH_SERVICE_GUARD),
           (64, 0, _abi.LAUNCH_SERVICE_GUARD),
           (512, 5, _abi.LAUNCH_SERVICE_GUARD),
           (512, 0, _abi.LAUNCH_INLINE_GUARD)]
    for block, w, flags in bad:
        la = _abi.Launch(block, w, _abi.PREC_FP32, flags)
        rc = lib().mcd_forward(ctx, C.byref(tr.cfg), tr._tables, 1, 0, 0, 64, els.data_ptr(), None,
                               cnt.data_ptr(), C.byref(la), s)
        with pytest.raises(MCDError):
            check(rc, ctx)
    torch.cuda.synchronize(dev)


def test_transport_autotuner_picks_a_shape_and_keeps_records(cu20, tmp_path):
    """The single-configuration tuner times every candidate shape, caches the
    fastest by configuration and electron count, and launches with the tuned
    shape write the same records as the library default."""
    from mcd import autotune
    tr = cu20[0]
    n = 1 << 16
    ref = tr.run(n, seed=1)
    path = str(tmp_path / "tt.json")
    first = autotune.tune_transport(tr, n, reps=1, cache_path=path)
    names = {autotune._tname(c) for c in autotune.TRANSPORT_CANDIDATES}
    assert not first["cached"] and set(first["ms"]) == names
    assert autotune._tname(first) in names
    again = autotune.tune_transport(tr, n, reps=1, cache_path=path)
    assert again["cached"] and autotune._tname(again) == autotune._tname(first)
    assert autotune.transport_key(tr, n) != autotune.transport_key(tr, 2 * n)
    for cand in autotune.TRANSPORT_CANDIDATES:
        r = tr.run(n, seed=1, block=cand["block"], waves_per_simd=cand["waves_per_simd"], guard=cand["guard"])
        assert r.els.tobytes() == ref.els.tobytes(), cand
    try:
        tr.tuned = dict(block=first["block"], waves_per_simd=first["waves_per_simd"], guard=first["guard"])
        assert tr.run(n, seed=1).els.tobytes() == ref.els.tobytes()
    finally:
        tr.tuned = None


def test_served_guard_matches_inline_across_the_batch_configurations():
    """The served-guard kernel writes the inline form's records for each of
    the 32 DS configurations of the batched workload ({Al, Si, Cu, Au} x E0
    {5, 10, 20, 30} keV x tilt {0, 70} deg: 6, 7 and 11 channels), 2^14
    electrons each, and settles the same number of FP64 guard picks."""
    from mcd import engine, workloads
    picks = 0
    for c in workloads.batched_configs():
        if c["mode"] != "DS":
            continue
        tr = engine.Transport(c["material"], c["E0"], c["Emin"], Wc=c["Wc"], s_tilt=c["s_tilt"], tol_E=c["tol"],
                              tol_W=c["tol"])
        a = tr.run(1 << 14, seed=9, guard="service")
        b = tr.run(1 << 14, seed=9, guard="inline")
        assert a.els.tobytes() == b.els.tobytes(), c
        assert int(a.counters[_abi.CNT_FP64_PICKS]) == int(b.counters[_abi.CNT_FP64_PICKS]), c
        picks += int(a.counters[_abi.CNT_FP64_PICKS])
    assert picks > 0
