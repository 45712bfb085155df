"""The 64-configuration batched workload: every configuration against the
FP64 oracle (same Philox streams, item i uses counter word config = i), the
batched launch against per-configuration launches, and the full-size batch
through properties that do not depend on its size."""
import math

import numpy as np
import pytest

import oracle as O
from conftest import check_values, identical, mfp_at

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    """The 64 configurations at 1/64 of their electron counts (2^8 .. 2^14
    each), with per-configuration transports whose pick tables have the node
    count the batch uses, so both launch paths read identical tables."""
    from mcd import workloads
    cfgs = workloads.batched_configs(scale=1 / 64)
    trs = workloads.build_transports(cfgs, pick_nodes=128)
    return cfgs, trs


def oracle_tables(tr):
    """Oracle tables holding the device-built CDF bytes (DS), else None."""
    if tr.grids is None or tr.cfg.mode != 0:
        return None, None
    gm, gg = tr.grids
    return O.Table(gm.Es, gm.Ws, tr.cdf(0)), O.Table(gg.Es, gg.Ws, tr.cdf(1))


@pytest.fixture(scope="module")
def batch_oracle(small_batch):
    """FP64 oracle records of every item of the 1/64-scale batch (seed 3,
    counter word config = item index)."""
    cfgs, trs = small_batch
    out = []
    for i, (tr, c) in enumerate(zip(trs, cfgs)):
        tm, tg = oracle_tables(tr)
        o, _, ev = O.run(tr.cfg, tm, tg, 3, 0, c["n"], config_id=i, nthreads=16)
        out.append((o, ev))
    return out


def _name(c):
    return f"{c['material']} {c['E0'] / 1000:g} keV {c['s_tilt']:g} deg {c['mode']}"


def test_every_batch_configuration_fp64_kernel_equals_oracle(small_batch, batch_oracle):
    """All 64 configurations (Emin 0.5 keV, Wc 100 eV, tol 1e-3; DS items with
    6, 7 and 11 channels, the Cu 5 keV and Au 5 and 10 keV DS items whose inner
    shells lie above E0, continuous Joy-Luo items): the all-FP64 kernel gives
    the oracle's records trajectory by trajectory and the same event count
    (multiScatter.py:66-116, singleScatter.py:10-41,205-239)."""
    from mcd import _abi
    cfgs, trs = small_batch
    for i, (tr, c, (o, ev)) in enumerate(zip(trs, cfgs, batch_oracle)):
        print(f"item {i}: {_name(c)}, {c['n']} electrons", flush=True)
        r64 = tr.run(c["n"], seed=3, config_id=i, precision=_abi.PREC_FP64)
        same = identical(r64.els, o)
        assert same.all(), (i, _name(c), np.flatnonzero(~same)[:5])
        assert r64.events == ev, (i, _name(c))
        check_values(r64.els, o, same, rel_E=1e-7, dir_abs=1e-7, pos_rel=1e-7, length=mfp_at(tr.cfg, c["E0"]))


def test_every_batch_configuration_fp32_batched_launch_matches_oracle(small_batch, batch_oracle):
    """One batched launch of all 64 configurations (FP32) against the oracle:
    at least 99.9 % of all trajectories identical end to end and no item
    with more than max(1, 0.1 %) differing ones; energies, directions and
    positions of identical trajectories within the FP32 bar."""
    from mcd import engine
    cfgs, trs = small_batch
    b = engine.Batch([(tr, c["n"]) for tr, c in zip(trs, cfgs)], pick_nodes=128)
    b.launch(seed=3, use_graph=True)
    n_same = n_all = 0
    for i, (tr, c, r, (o, ev)) in enumerate(zip(trs, cfgs, b.results(), batch_oracle)):
        same = identical(r.els, o)
        bad = int((~same).sum())
        assert bad <= max(1, math.ceil(1e-3 * c["n"])), (i, _name(c), bad, c["n"])
        errs = check_values(r.els, o, same, length=mfp_at(tr.cfg, c["E0"]))
        print(f"item {i}: {_name(c)} n={c['n']} differing {bad}; max errors E {errs[0]:.2e} dir {errs[1]:.2e} "
              f"pos {errs[2]:.2e}", flush=True)
        n_same += int(same.sum())
        n_all += c["n"]
    print(f"identical trajectories: {n_same}/{n_all} = {n_same / n_all:.6f}")
    assert n_same / n_all >= 0.999


@pytest.mark.parametrize("use_graph,guard,period", [(False, None, 1), (True, None, 1), (True, "deferred", 1),
                                                    (False, "service", 1), (True, "service", 1), (True, None, 4),
                                                    (False, "deferred", 2), (True, "service", 2)])
def test_batched_equals_per_configuration_launches(small_batch, use_graph, guard, period):
    """Every guard placement and compaction period (idle lanes refilled every
    `period` steps) gives the per-configuration records."""
    from mcd import _abi, engine
    cfgs, trs = small_batch
    b = engine.Batch([(tr, c["n"]) for tr, c in zip(trs, cfgs)], pick_nodes=128)
    for rep in range(2):  # graph replays must repeat the eager result
        b.counters.zero_()
        b.launch(seed=3, use_graph=use_graph, guard=guard, refill_period=period)
        res = b.results()
        for i, (tr, c, r) in enumerate(zip(trs, cfgs, res)):
            ref = tr.run(c["n"], seed=3, config_id=i)
            assert r.els.tobytes() == ref.els.tobytes(), (i, c)
            assert int(r.counters[_abi.CNT_EVENTS]) == ref.events
            assert int(r.counters[_abi.CNT_ELECTRONS]) == c["n"]


def test_served_batch_at_several_workgroup_sizes(small_batch):
    """The served guard in the batched kernel (wavefront 0 of each workgroup
    settles the guard picks of the other wavefronts, whatever configuration
    each is on) gives the per-configuration records at 2 to 16 wavefronts per
    workgroup."""
    from mcd import engine
    cfgs, trs = small_batch
    b = engine.Batch([(tr, c["n"]) for tr, c in zip(trs, cfgs)], pick_nodes=128)
    refs = [tr.run(c["n"], seed=3, config_id=i).els.tobytes() for i, (tr, c) in enumerate(zip(trs, cfgs))]
    for block in (128, 640, 1024):
        b.counters.zero_()
        b.launch(seed=3, use_graph=False, block=block, guard="service")
        for i, r in enumerate(b.results()):
            assert r.els.tobytes() == refs[i], (block, i)


def test_served_batch_guard_timeout_is_reported(small_batch, monkeypatch):
    """The batched served guard's wait is bounded as well: a batch created
    with MCD_GUARD_SPIN_CAP=1 still ends normally, counts the steps that gave
    up and the host refuses its records (GuardError); a batch created with the
    default bound has no timeouts and the per-configuration records."""
    from mcd import _abi, engine
    from mcd.engine import GuardError
    cfgs, trs = small_batch
    items = [(tr, c["n"]) for tr, c in zip(trs, cfgs)]
    monkeypatch.setenv("MCD_GUARD_SPIN_CAP", "1")
    b = engine.Batch(items, pick_nodes=128)
    b.launch(seed=3, use_graph=False, guard="service")
    with pytest.raises(GuardError):
        b.results()
    cnt = b.counters.cpu().numpy().view(np.uint64).reshape(len(items), _abi.NCOUNTERS)
    assert int(cnt[:, _abi.CNT_GUARD_TIMEOUTS].sum()) > 0
    assert [int(c) for c in cnt[:, _abi.CNT_ELECTRONS]] == [n for _, n in items]
    monkeypatch.delenv("MCD_GUARD_SPIN_CAP")
    b = engine.Batch(items, pick_nodes=128)
    b.launch(seed=3, use_graph=False, guard="service")
    res = b.results()
    assert sum(int(r.counters[_abi.CNT_FP64_PICKS]) for r in res) > 0
    for i, ((tr, n), r) in enumerate(zip(items, res)):
        assert int(r.counters[_abi.CNT_GUARD_TIMEOUTS]) == 0
        assert r.els.tobytes() == tr.run(n, seed=3, config_id=i).els.tobytes(), i


def test_autotuner_picks_and_caches_a_launch_shape(small_batch, tmp_path):
    """The tuner times every candidate, returns the fastest, caches it by the
    batch's configurations, and the tuned batch reproduces per-configuration
    launches built with the same pick-table node count."""
    import os
    from mcd import autotune, engine, workloads
    cfgs, trs = small_batch
    items = [(tr, c["n"]) for tr, c in zip(trs, cfgs)]
    cands = (dict(block=256, pick_nodes=128, guard="deferred"), dict(block=128, pick_nodes=128, guard="deferred"),
             dict(block=256, pick_nodes=128, guard="inline"), dict(block=256, pick_nodes=128, guard="inline",
                                                                   refill_period=2))
    path = str(tmp_path / "tune.json")
    first = autotune.tune_batch(items, candidates=cands, reps=1, cache_path=path)
    assert not first["cached"]
    assert (first["block"], first["pick_nodes"], first["guard"], first["refill_period"]) in {
        (256, 128, "deferred", 1), (128, 128, "deferred", 1), (256, 128, "inline", 1), (256, 128, "inline", 2)}
    assert set(first["ms"]) == {"block=256,pick_nodes=128,guard=deferred,refill_period=1",
                                "block=128,pick_nodes=128,guard=deferred,refill_period=1",
                                "block=256,pick_nodes=128,guard=inline,refill_period=1",
                                "block=256,pick_nodes=128,guard=inline,refill_period=2"}
    again = autotune.tune_batch(items, candidates=cands, reps=1, cache_path=path)
    assert again["cached"] and again["block"] == first["block"] and again["guard"] == first["guard"]
    assert again["refill_period"] == first["refill_period"]
    other = workloads.batched_configs(scale=1 / 128)
    assert autotune.batch_key(items) != autotune.batch_key([(tr, c["n"]) for tr, c in zip(trs, other)])
    b = engine.Batch(items, pick_nodes=first["pick_nodes"])
    b.launch(seed=3, use_graph=True, block=first["block"], guard=first["guard"], refill_period=first["refill_period"])
    res = b.results()
    for i, (tr, c, r) in enumerate(zip(trs, cfgs, res)):
        assert r.els.tobytes() == tr.run(c["n"], seed=3, config_id=i).els.tobytes(), i
    assert os.path.exists(path)


@pytest.mark.parametrize("guard", [None, "deferred", "service"])
def test_ragged_batch_with_empty_and_single_electron_items(small_batch, guard):
    """Ragged items (0, 1 and 17 electrons beside larger ones, DS and
    continuous) give the per-configuration records and counts; empty items
    are never assigned a wavefront and count nothing, and a batch whose items
    are all empty ends with zero counters, eagerly and as a graph replay."""
    from mcd import _abi, engine
    cfgs, trs = small_batch
    sizes = [0, 1, 17, 0, 300, 5, 0, 1000]
    picks = [0, 9, 18, 27, 36, 45, 54, 63]  # DS and continuous items of all four materials
    modes = {cfgs[k]["mode"] for k in picks}
    assert len(picks) == len(sizes) and len(modes) == 2, modes
    items = [(trs[k], n) for k, n in zip(picks, sizes)]
    b = engine.Batch(items, pick_nodes=128)
    for use_graph in (False, True):
        b.counters.zero_()
        b.launch(seed=8, use_graph=use_graph, guard=guard, block=640 if guard == "service" else 0)
        for i, ((tr, n), r) in enumerate(zip(items, b.results())):
            ref = tr.run(n, seed=8, config_id=i)
            assert r.els.tobytes() == ref.els.tobytes(), (i, n)
            assert int(r.counters[_abi.CNT_ELECTRONS]) == n, (i, n)
            assert int(r.counters[_abi.CNT_EVENTS]) == (ref.events if n else 0), (i, n)
    empty = engine.Batch([(trs[k], 0) for k in picks[:3]], pick_nodes=128)
    for use_graph in (False, True):
        empty.counters.zero_()
        empty.launch(seed=8, use_graph=use_graph, guard=guard, block=640 if guard == "service" else 0)
        assert all(not r.counters.any() for r in empty.results())


def test_batch_with_thin_film_and_max_scatt_items():
    """Items with a finite thickness and maxScatt (DS and continuous) in one
    batch give the per-configuration records ('long' and 'trsm' outcomes
    included) in every guard form."""
    from mcd import _abi, engine
    trs = [engine.Transport("Si", 10000.0, 1000.0, Wc=100.0, s_tilt=0.0, tol_E=1e-2, tol_W=1e-2,
                            thickness=3000.0, max_scatt=80, pick_nodes=128),
           engine.Transport("Au", 30000.0, 1000.0, s_tilt=0.0, mode="cont", bethe_model="JL", thickness=3000.0,
                            max_scatt=150),
           engine.Transport("Cu", 20000.0, 1000.0, Wc=100.0, s_tilt=70.0, tol_E=1e-2, tol_W=1e-2, pick_nodes=128)]
    sizes = [3000, 2000, 4000]
    refs = [tr.run(n, seed=6, config_id=i) for i, (tr, n) in enumerate(zip(trs, sizes))]
    for out in (_abi.OUT_LONG, _abi.OUT_TRSM):
        assert (refs[0].els["outcome"] == out).sum() > 50 and (refs[1].els["outcome"] == out).sum() > 50
    b = engine.Batch(list(zip(trs, sizes)), pick_nodes=128)
    for guard in (None, "deferred", "service"):
        b.counters.zero_()
        b.launch(seed=6, use_graph=False, guard=guard, block=512 if guard == "service" else 0)
        for i, (r, ref) in enumerate(zip(b.results(), refs)):
            assert r.els.tobytes() == ref.els.tobytes(), (guard, i)
            assert int(r.counters[_abi.CNT_EVENTS]) == ref.events, (guard, i)


def test_full_size_batch_properties():
    """The full 64-configuration workload (14.67 M electrons, log-uniform
    2^14..2^20 per configuration, seed 3) in one graph-replayed launch:
    every item finishes each of its electrons once (counters and records
    agree), its event count is the sum of its records' scatter counts, every
    outcome is 'abs' or 'bks' on a semi-infinite target; each item's
    backscatter yield is within 3 sigma of the all-FP64 per-configuration run
    of the same streams, and the yields against independent FP32 streams
    (seed 103) pass a chi-square test over the 64 items (p > 0.001)."""
    from scipy import stats
    from mcd import _abi, engine, workloads
    cfgs = workloads.batched_configs()
    trs = workloads.build_transports(cfgs, pick_nodes=128)
    b = engine.Batch([(tr, c["n"]) for tr, c in zip(trs, cfgs)], pick_nodes=128)
    b.launch(seed=3, use_graph=True)
    res = b.results()
    chi2 = 0.0
    for i, (tr, c, r) in enumerate(zip(trs, cfgs, res)):
        n = c["n"]
        cnt = r.counters
        assert int(cnt[_abi.CNT_ELECTRONS]) == n, (i, _name(c))
        assert int(cnt[_abi.CNT_EVENTS]) == int(r.els["n_scatt"].astype(np.int64).sum()), (i, _name(c))
        out = r.els["outcome"]
        assert np.isin(out, (_abi.OUT_ABS, _abi.OUT_BKS)).all(), (i, _name(c))
        nb = int((out == _abi.OUT_BKS).sum())
        assert int(cnt[_abi.CNT_BKS]) == nb, (i, _name(c))
        y = nb / n
        r64 = tr.run(n, seed=3, config_id=i, precision=_abi.PREC_FP64)
        y64 = float((r64.els["outcome"] == _abi.OUT_BKS).mean())
        sig = math.sqrt(max(y64 * (1 - y64), 1e-12) / n)
        assert abs(y - y64) < 3 * sig, (i, _name(c), y, y64, sig)
        ind = tr.run(n, seed=103, config_id=i)
        yi = float((ind.els["outcome"] == _abi.OUT_BKS).mean())
        chi2 += (y - yi) ** 2 / max(y * (1 - y) / n + yi * (1 - yi) / n, 1e-300)
        print(f"item {i}: {_name(c)} n={n} yield {y:.5f} fp64 {y64:.5f} independent {yi:.5f}", flush=True)
    p = stats.chi2.sf(chi2, len(cfgs))
    print(f"chi2 = {chi2:.1f} over {len(cfgs)} items, p = {p:.3g}")
    assert p > 1e-3


@pytest.mark.parametrize("guard", [None, "service"])
def test_batched_elastic_models_equal_per_configuration_launches(guard):
    """Items with every Rutherford elastic model (each item's constants,
    elastic ones included, are copied into its wavefronts' LDS slices), their
    tables from one mcd_tables_build_batch call, give the per-configuration
    records."""
    from mcd import _abi, engine
    specs = [dict(material="Cu", E0=20000.0, Emin=1000.0, Wc=50.0, s_tilt=70.0, mode="DS", elastic=el)
             for el in ("Ruth_vanilla", "Ruth_vanilla_wDefl", "Ruth_nigram", "Ruth_nigram_wDefl")]
    specs += [dict(material="Au", E0=10000.0, Emin=1000.0, s_tilt=0.0, mode="cont", bethe_model="JL", elastic=el)
              for el in ("Ruth_vanilla", "Ruth_nigram")]
    trs = engine.build_transports_batch(specs, tol_E=1e-3, tol_W=1e-3, pick_nodes=128)
    ns = [3000, 2048, 1500, 2500, 1024, 2000]
    b = engine.Batch(list(zip(trs, ns)), pick_nodes=128)
    b.launch(seed=7, use_graph=False, guard=guard, block=512 if guard == "service" else 0)
    res = b.results()
    yields = []
    for i, (tr, n, r) in enumerate(zip(trs, ns, res)):
        ref = tr.run(n, seed=7, config_id=i)
        assert r.els.tobytes() == ref.els.tobytes(), (i, specs[i]["elastic"])
        yields.append(float(np.mean(ref.els["outcome"] == _abi.OUT_BKS)))
    # the models differ: the four DS items do not share one yield
    assert len({round(y, 6) for y in yields[:4]}) == 4
