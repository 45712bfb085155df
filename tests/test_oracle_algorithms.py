"""Pins of the oracle's discrete algorithms and of whole trajectories.

* the channel pick against the reference's own idiom (dict sorted by value,
  np.cumsum / sum, bisect.bisect_left) on random and tied inputs;
* the table lookup against bisect over numpy.ma rows (masked = 0), the
  semantics the reference reads its tables with;
* the rotation against an independent closed form and scipy's Rotation;
* the float32 CDF restatement against the analytic Moller CDF;
* whole trajectories against an independent Python restatement (baseline_b1);
* the reference's own statistical checks (chi-square channel frequencies,
  Watson U^2 polar angles) on oracle output.
"""
import bisect
import os
import math
import operator

import numpy as np
import numpy.ma as ma
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
import tables_ref as TR
import baseline_b1 as B1
from mcd import _abi
from mcd.material import Material


def ref_pick(sig, u):
    """pickFromSigmas' idiom (scattering.py:48-56) on a dict in insertion order."""
    d = {i: s for i, s in enumerate(sig)}
    srt = dict(sorted(d.items(), key=operator.itemgetter(1), reverse=True))
    probs = np.cumsum(list(srt.values())) / sum(srt.values())
    return list(srt.keys())[bisect.bisect_left(probs, u)]


def test_pick_matches_reference_idiom(rng):
    for trial in range(3000):
        n = rng.integers(2, 12)
        sig = 0.01 + rng.random(n) * 10.0 ** rng.integers(-3, 2, n)
        if trial % 3 == 0:  # zeros (channels below threshold) and exact ties
            sig[rng.integers(0, n, 2)] = 0.0
            sig[rng.integers(0, n)] = sig[0]
            if sig.sum() == 0:
                sig[-1] = 1.0
        u = rng.random()
        if trial % 5 == 0:  # u exactly on a boundary
            srt = sorted(sig, reverse=True)
            u = float(np.cumsum(srt)[rng.integers(0, n)] / sum(srt))
            if u >= 1.0:
                u = float(np.nextafter(1.0, 0.0))
        assert O.pick(sig, u) == ref_pick(sig, u), (sig, u)


def test_table_lookup_matches_masked_bisect(si10, rng):
    for t in (si10["tm"], si10["tg"]):
        masked = ma.masked_array(t.cdf, t.cdf == 0)
        for _ in range(2000):
            E = float(rng.uniform(t.Es[0] * 0.999, t.Es[-1]))
            if rng.random() < 0.2:
                E = float(t.Es[rng.integers(0, t.Es.size)])  # on a grid energy
            u = float(rng.random())
            row = bisect.bisect_left(t.Es, E)
            assert t.row(E) == row
            assert t.col(row, u) == bisect.bisect_left(masked[row], u)


def test_rotation_against_closed_form_and_scipy(rng):
    for _ in range(500):
        d = rng.normal(size=3); d /= np.linalg.norm(d)
        y = np.cross(d, rng.normal(size=3)); y /= np.linalg.norm(y)
        c2, hp = rng.random(), math.pi * rng.random()
        nd, ny = O.rotate(c2, hp, d, y)
        # closed form: w = y x d; d' = cos t d + sin t (cos p w + sin p y); y' = cos p y - sin p w
        ct, st = 2 * c2 - 1, 2 * math.sqrt(c2 * (1 - c2))
        s_hp = math.sin(hp); c_hp = math.sqrt(1 - s_hp * s_hp)
        cp, sp = c_hp * c_hp - s_hp * s_hp, 2 * s_hp * c_hp
        w = np.cross(y, d)
        assert np.allclose(nd, ct * d + st * (cp * w + sp * y), atol=1e-13)
        assert np.allclose(ny, cp * y - sp * w, atol=1e-13)
        # scipy: q_az * q_polar (scalar-last for scipy)
        q_pol = np.r_[math.sin(math.acos(math.sqrt(c2))) * y, math.sqrt(c2)]
        q_az = np.r_[s_hp * d, c_hp]
        R = Rotation.from_quat(q_az) * Rotation.from_quat(q_pol)
        assert np.allclose(nd, R.apply(d), atol=1e-12)
        assert np.allclose(ny, R.apply(y), atol=1e-12)
        assert abs(np.dot(nd, d) - ct) < 1e-12 and abs(np.dot(nd, ny)) < 1e-12


def test_moller_cdf_against_analytic(si10):
    """The float32 trapezoid CDF of the Moller table approaches the analytic
    CDF of moller_dCS, F(W) = [g(Wc) - g(W)] / [g(Wc) - g(Wmax)] with
    g(W) = 1/eps - 1/(1-eps) + ln(eps/(1-eps)), eps = W/E."""
    t, m = si10["tm"], si10["mat"]
    for i in (0, t.Es.size // 2, t.Es.size - 1):
        E = float(t.Es[i])
        row = t.cdf[i]
        nv = int((row > 0).sum())
        W = t.Ws.astype(np.float64)[1:nv + 1]
        g = lambda w: 1 / (w / E) - 1 / (1 - w / E) + np.log((w / E) / (1 - w / E))
        Wc = float(t.Ws[0])
        F = (g(Wc) - g(W)) / (g(Wc) - g(W[-1]))
        assert np.all(np.diff(row[:nv]) >= 0) and row[nv - 1] == 1.0
        assert np.abs(row[:nv] - F).max() < 0.02  # trapezoid on a 29-point geometric grid
        # masked pattern: upper bin edge above (E - E_F)/2 in float32
        wmax = (np.float32(E) - np.float32(m.fermi_e)) * np.float32(0.5)
        assert np.array_equal(row > 0, ~(t.Ws[1:] > wmax))


def test_table_grid_shapes_match_independent_restatement():
    """Grid sizes of the tolerance search (probTables.py:280-382) equal the
    independently restated values: Si 10 keV tol 1e-2 gives 300x29
    (Moller) and 400x39 (Gryz 2p); Cu 20 keV tol 1e-3 gives 2000x199 twice."""
    from conftest import CU20, SI10, build_grids
    _, gm, gg = build_grids(SI10)
    assert gm.shape == (300, 29) and gg.shape == (400, 39)   # (nE, nW - 1) CDF shapes
    m, gm, gg = build_grids(CU20)
    assert gm.shape == (2000, 199) and gg.shape == (2000, 199)
    assert gm.Ws[0] == 50.0 and gg.Ws[0] == 74.0
    assert gm.Ws[-1] == np.float32((20000.0 - m.fermi_e) * 0.5)
    assert gm.Es[0] == 1000.0 and gm.Es[-1] == 20000.0


@pytest.mark.parametrize("mode,bethe", [("DS", "classical"), ("cont", "classical"), ("cont", "JL"),
                                        ("cont", "explicit")])
def test_oracle_matches_independent_python_restatement(si10, mode, bethe):
    cfg = _abi.make_config(si10["mat"], 10000, 1000, 100, 70, mode=mode, bethe_model=bethe)
    els, _, ev = O.run(cfg, si10["tm"], si10["tg"], 3, 100, 24)
    P = B1.params_from(cfg, si10["tm"], si10["tg"], seed=3)
    ev1, res = B1.run_serial(P, 100, 24)
    assert ev1 == ev
    names = {1: "abs", 2: "bks", 3: "trsm", 5: "long"}
    for r, e in zip(res, els):
        assert r["els"]["outcome"][0] == names[e["outcome"]]
        assert r["els"]["final_E"][0] == pytest.approx(e["final_E"], rel=1e-12)
        assert np.allclose(r["els"]["final_dir"][0], e["dir"], atol=1e-11)
        assert np.allclose(r["els"]["last_pos"][0], e["pos"], rtol=1e-10, atol=1e-8)


def test_baseline_pool_forms_do_the_same_work(si10):
    """The CPU-baseline drivers (serial, pool with tables loaded once per
    worker, pool with the tables shipped in every job) transport the same
    electrons: equal event counts, equal to the oracle's."""
    cfg = _abi.make_config(si10["mat"], 10000, 1000, 100, 70)
    _, _, ev = O.run(cfg, si10["tm"], si10["tg"], 2, 0, 12)
    P = B1.params_from(cfg, si10["tm"], si10["tg"], seed=2)
    assert B1.run_serial(P, 0, 12)[0] == ev
    assert B1.run_pool(P, 12, workers=2)[0] == ev
    assert B1.run_pool_shipped(P, 12, workers=2, n_gryz_tables=3)[0] == ev


def test_electron_ids_above_2_32_have_streams_of_their_own(si10):
    """Electron ids are 48 bits: bits 32..47 go into Philox counter word 1
    beside the configuration id.  Electron 2^32 + k must not replay electron
    k's stream, the two implementations agree across the 2^32 boundary, and
    ranges that reach past 2^48 are refused."""
    cfg = _abi.make_config(si10["mat"], 10000, 1000, 100, 70)
    lo, _, _ = O.run(cfg, si10["tm"], si10["tg"], 3, 0, 32)
    hi, _, _ = O.run(cfg, si10["tm"], si10["tg"], 3, 1 << 32, 32)
    same = (lo["n_scatt"] == hi["n_scatt"]) & (lo["final_E"] == hi["final_E"])
    assert same.sum() <= 1, same.sum()
    first = (1 << 32) - 4
    els, _, ev = O.run(cfg, si10["tm"], si10["tg"], 3, first, 8)
    ev1, res = B1.run_serial(B1.params_from(cfg, si10["tm"], si10["tg"], seed=3), first, 8)
    assert ev1 == ev
    for r, e in zip(res, els):
        assert r["els"]["final_E"][0] == pytest.approx(e["final_E"], rel=1e-12)
    assert np.array_equal(els[4:], hi[:4])
    O.run(cfg, si10["tm"], si10["tg"], 3, (1 << 48) - 2, 2)
    with pytest.raises(RuntimeError):
        O.run(cfg, si10["tm"], si10["tg"], 3, (1 << 48) - 2, 3)


def test_trace_invariants(si10):
    cfg = _abi.make_config(si10["mat"], 10000, 1000, 100, 70)
    els, tr, ev = O.run(cfg, si10["tm"], si10["tg"], 0, 0, 64, trace_cap=200000)
    assert len(tr) == ev
    d = np.stack([tr["dx"], tr["dy"], tr["dz"]], 1)
    y = np.stack([tr["yx"], tr["yy"], tr["yz"]], 1)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-12)
    assert np.allclose(np.linalg.norm(y, axis=1), 1, atol=1e-12)
    assert np.abs((d * y).sum(1)).max() < 1e-9
    assert np.all(tr["energy"] == tr["E"] - tr["E_loss"])
    assert np.all(tr["az"] < math.pi) and np.all(tr["az"] > 0)
    # every table loss is a grid value and energy never increases
    ins = np.isin(tr["channel"], [_abi.CH_MOLLER] + [_abi.CH_GRYZ0 + k for k in range(3)])
    grid = np.concatenate([si10["tm"].Ws, si10["tg"].Ws]).astype(np.float64)
    assert np.all(np.isin(tr["E_loss"][ins], grid))
    q = tr["channel"] == _abi.CH_QUINN
    assert np.all(tr["E_loss"][q] == si10["mat"].plasmon_e)
    assert np.all(tr["E_loss"][tr["channel"] == _abi.CH_RUTH] == 0)


def test_reference_chi_square_channel_test(si10):
    """Test_crossSection (tests/testThings.py:321-387): channel frequencies of
    events at E0 against the cross-section ratios, chi-square at q = 0.90."""
    from scipy import stats
    cfg = _abi.make_config(si10["mat"], 10000, 1000, 100, 70)
    _, tr, _ = O.run(cfg, si10["tm"], si10["tg"], 5, 0, 4000, trace_cap=2_000_000)
    first = tr[tr["step"] == 0]  # every electron's first event is at E0
    sig, lab = O.sigma_set(cfg, 10000.0)
    p = sig / sig.sum()
    obs = np.array([(first["channel"] == l).sum() for l in lab], dtype=float)
    keep = p * len(first) >= 5
    exp = p * len(first)
    chi2 = (((obs - exp) ** 2 / exp)[keep]).sum() + ((obs[~keep].sum() - exp[~keep].sum()) ** 2 /
                                                   max(exp[~keep].sum(), 1e-9) if (~keep).any() else 0)
    dof = keep.sum() - 1 + (1 if (~keep).any() else 0)
    assert chi2 < stats.chi2.ppf(0.999, dof), (obs, exp)


def watson_u2(x, y):
    """Two-sample Watson U^2 (Mardia & Jupp eq. 8.3.7), the statistic of
    watson_two_test (tests/testThings.py:255-316)."""
    x, y = np.sort(x), np.sort(y)
    nx, ny, N = len(x), len(y), len(x) + len(y)
    allv = np.sort(np.unique(np.r_[x, y]))
    Fx = np.searchsorted(x, allv, side="right") / nx
    Fy = np.searchsorted(y, allv, side="right") / ny
    dk = Fx - Fy
    return nx * ny * ((dk ** 2).sum() - dk.sum() ** 2 / N) / N ** 2


def test_reference_watson_polar_angle_test(si10):
    """TestPolarProb_Ruth_E0 (testThings.py:393-434): Rutherford polar angles
    at E0 against draws from the analytic inverse CDF (Ruth_diffCS_E._ppf,
    crossSections.py:410-421), Watson U^2 < 0.152."""
    cfg = _abi.make_config(si10["mat"], 10000, 1000, 100, 0)
    _, tr, _ = O.run(cfg, si10["tm"], si10["tg"], 11, 0, 2000, trace_cap=2_000_000)
    r = tr[(tr["step"] == 0) & (tr["channel"] == _abi.CH_RUTH)]
    mc = np.degrees(np.arccos(2 * r["pol"] - 1))[:200]
    a = 3.4 * 14 ** 0.666667 / 10000.0
    u = np.random.default_rng(7).random(200)
    theory = np.degrees(np.arccos(1 - (2 * a * u) / (1 + a - u)))
    assert watson_u2(theory, mc) < 0.152
    # azimuth: 2*halfPhi uniform on [0, 360) (TestAzimProb_Ruth, testThings.py:491-522)
    az = np.degrees(2 * r["az"])[:200]
    assert watson_u2(np.random.default_rng(8).random(200) * 360, az) < 0.152


def test_cfg_struct_layout_matches_header(tmp_path):
    """ctypes structs agree with include/mcd.h (sizes and offsets, via gcc)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "lay.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "mcd.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %d %d %d\\n", sizeof(mcd_config), '
                   'offsetof(mcd_config,E0), offsetof(mcd_config,J), sizeof(mcd_material), sizeof(mcd_table_grid), '
                   'sizeof(mcd_electron_rec), sizeof(mcd_scatter_rec), sizeof(mcd_scatter_buf), sizeof(mcd_launch), '
                   'sizeof(mcd_batch_item), sizeof(mcd_bwd_opts), offsetof(mcd_bwd_opts,graze), MCD_NGRAD, '
                   'MCD_GRAD_LNRHO, MCD_MAX_BINS);'
                   'return 0;}\n')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    import ctypes as C
    want = [C.sizeof(_abi.Config), _abi.Config.E0.offset, _abi.Config.J.offset, C.sizeof(_abi.Material),
            C.sizeof(_abi.TableGridC), _abi.ELECTRON_DTYPE.itemsize, _abi.SCATTER_DTYPE.itemsize,
            C.sizeof(_abi.ScatterBuf), C.sizeof(_abi.Launch), C.sizeof(_abi.BatchItem), C.sizeof(_abi.BwdOpts),
            _abi.BwdOpts.graze.offset, _abi.NGRAD, _abi.GRAD_LNRHO, _abi.MAX_BINS]
    assert got == want


def _golden_module():
    import importlib.util
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "make_c1_golden.py")
    spec = importlib.util.spec_from_file_location("make_c1_golden", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_oracle_and_table_restatement_reproduce_the_committed_golden_vectors():
    """tests/golden/c1_si10_seed0.npz freezes the grids, the float32 CDF tables
    and the per-step trace of 4 electrons of the CPU-runnable DS case; the
    grid search, the table restatement and the oracle reproduce every array
    bit for bit."""
    mod = _golden_module()
    ref = np.load(os.path.join(os.path.dirname(mod.__file__), "c1_si10_seed0.npz"))
    now = mod.golden()
    assert set(ref.files) == set(now)
    for k in ref.files:
        assert np.array_equal(ref[k], now[k]), k
    assert len(ref["trace_step"]) > 1000 and set(ref["els_outcome"]) <= {1, 2}


def test_divergence_classifier_on_planted_rows(si10):
    """The per-step parity test's classifier, on kernel-like rows made from
    an oracle trace with one planted difference per electron: a channel
    changed where u is far from every boundary (a missed guard pick), a
    channel changed where u sits on a boundary (a tie), a different loss
    (a table miss), a different outcome (an exit flip) and a shortened
    trajectory; unchanged electrons are not reported."""
    import oracle as O
    from divergence import classify_divergence
    from mcd import _abi
    from mcd.material import Material
    m = Material("Si")
    cfg = _abi.make_config(m, si10["E0"], si10["Emin"], si10["Wc"], si10["s_tilt"])
    _, tr, _ = O.run(cfg, si10["tm"], si10["tg"], 4, 0, 8, trace_cap=100000)
    g = np.zeros(len(tr), dtype=_abi.SCATTER_DTYPE)
    for k in ("electron", "step", "E", "E_loss", "pol", "az", "pathl", "x", "y", "z", "energy", "dx", "dy", "dz"):
        g[k] = tr[k]
    g["channel"] = tr["channel"]
    g["outcome"] = tr["outcome"]
    o = tr.copy()
    assert classify_divergence(cfg, (si10["tm"], si10["tg"]), g, o) == []

    def row(e, st):
        return int(np.flatnonzero((tr["electron"] == e) & (tr["step"] == st))[0])

    # electron 1, step 3: a channel swap far from any boundary -> pick_miss
    i = row(1, 3)
    g["channel"][i] = _abi.CH_QUINN if g["channel"][i] != _abi.CH_QUINN else _abi.CH_RUTH
    # electron 2, step 4: the oracle's u put exactly on a boundary, channel swapped -> pick_tie
    i = row(2, 4)
    sig, _ = O.sigma_set(cfg, float(o["E"][i]))
    srt = sorted(range(len(sig)), key=lambda q: -sig[q])
    o["u"][i, 0] = float(sig[srt[0]]) / sum(float(sig[q]) for q in srt)
    g["channel"][i] = _abi.CH_QUINN if g["channel"][i] != _abi.CH_QUINN else _abi.CH_RUTH
    # the first other electron with a Moller or Gryzinski step: another loss there -> table_miss
    lossy = (tr["channel"] == _abi.CH_MOLLER) | (tr["channel"] >= _abi.CH_GRYZ0)
    e3 = next(e for e in (3, 4, 7, 0) if np.any(lossy & (tr["electron"] == e)))
    i = row(e3, int(tr["step"][lossy & (tr["electron"] == e3)][0]))
    g["E_loss"][i] = g["E_loss"][i] * 2 + 1
    # electron 5: the last step without its outcome -> exit
    i = row(5, int(tr["step"][tr["electron"] == 5].max()))
    g["outcome"][i] = 0
    # electron 6: its last row missing -> length
    keep = np.ones(len(g), bool)
    keep[row(6, int(tr["step"][tr["electron"] == 6].max()))] = False
    got = classify_divergence(cfg, (si10["tm"], si10["tg"]), g[keep], o)
    classes = {e: c for e, _, c in got}
    assert classes == {1: "pick_miss", 2: "pick_tie", e3: "table_miss", 5: "exit", 6: "length"}, got
