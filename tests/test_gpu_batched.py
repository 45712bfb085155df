"""Batched multi-configuration launch against per-configuration launches of
the same Philox streams (item i uses counter word config = i)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    """Per-configuration transports whose pick tables have the node count the
    batch uses (given explicitly), so both launch paths read identical tables."""
    import os
    from mcd import workloads
    old = os.environ.get("MCD_PICK_NODES")
    os.environ["MCD_PICK_NODES"] = "128"
    try:
        cfgs = workloads.batched_configs(scale=1 / 64)  # 2^8 .. 2^14 electrons per configuration
        trs = workloads.build_transports(cfgs)
    finally:
        if old is None:
            del os.environ["MCD_PICK_NODES"]
        else:
            os.environ["MCD_PICK_NODES"] = old
    return cfgs, trs


@pytest.mark.parametrize("use_graph,guard", [(False, None), (True, None), (True, "deferred"), (False, "service"),
                                             (True, "service")])
def test_batched_equals_per_configuration_launches(small_batch, use_graph, guard):
    from mcd import _abi, engine
    cfgs, trs = small_batch
    b = engine.Batch([(tr, c["n"]) for tr, c in zip(trs, cfgs)], pick_nodes=128)
    for rep in range(2):  # graph replays must repeat the eager result
        b.counters.zero_()
        b.launch(seed=3, use_graph=use_graph, guard=guard)
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


def test_autotuner_picks_and_caches_a_launch_shape(small_batch, tmp_path):
    """The tuner times every candidate, returns the fastest, caches it by the
    batch's configurations, and the tuned batch reproduces per-configuration
    launches built with the same pick-table node count."""
    import os
    from mcd import autotune, engine, workloads
    cfgs, trs = small_batch
    items = [(tr, c["n"]) for tr, c in zip(trs, cfgs)]
    cands = (dict(block=256, pick_nodes=128, guard="deferred"), dict(block=128, pick_nodes=128, guard="deferred"),
             dict(block=256, pick_nodes=128, guard="inline"))
    path = str(tmp_path / "tune.json")
    first = autotune.tune_batch(items, candidates=cands, reps=1, cache_path=path)
    assert not first["cached"]
    assert (first["block"], first["pick_nodes"], first["guard"]) in {(256, 128, "deferred"), (128, 128, "deferred"),
                                                                     (256, 128, "inline")}
    assert set(first["ms"]) == {"block=256,pick_nodes=128,guard=deferred", "block=128,pick_nodes=128,guard=deferred",
                                "block=256,pick_nodes=128,guard=inline"}
    again = autotune.tune_batch(items, candidates=cands, reps=1, cache_path=path)
    assert again["cached"] and again["block"] == first["block"] and again["guard"] == first["guard"]
    other = workloads.batched_configs(scale=1 / 128)
    assert autotune.batch_key(items) != autotune.batch_key([(tr, c["n"]) for tr, c in zip(trs, other)])
    b = engine.Batch(items, pick_nodes=first["pick_nodes"])
    b.launch(seed=3, use_graph=True, block=first["block"], guard=first["guard"])
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
    import os
    from mcd import _abi, engine
    old = os.environ.get("MCD_PICK_NODES")
    os.environ["MCD_PICK_NODES"] = "128"
    try:
        trs = [engine.Transport("Si", 10000.0, 1000.0, Wc=100.0, s_tilt=0.0, tol_E=1e-2, tol_W=1e-2,
                                thickness=3000.0, max_scatt=80),
               engine.Transport("Au", 30000.0, 1000.0, s_tilt=0.0, mode="cont", bethe_model="JL", thickness=3000.0,
                                max_scatt=150),
               engine.Transport("Cu", 20000.0, 1000.0, Wc=100.0, s_tilt=70.0, tol_E=1e-2, tol_W=1e-2)]
    finally:
        if old is None:
            del os.environ["MCD_PICK_NODES"]
        else:
            os.environ["MCD_PICK_NODES"] = old
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
