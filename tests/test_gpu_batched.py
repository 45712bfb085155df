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


@pytest.mark.parametrize("use_graph", [False, True])
def test_batched_equals_per_configuration_launches(small_batch, use_graph):
    from mcd import _abi, engine
    cfgs, trs = small_batch
    b = engine.Batch([(tr, c["n"]) for tr, c in zip(trs, cfgs)], pick_nodes=128)
    for rep in range(2):  # graph replays must repeat the eager result
        b.counters.zero_()
        b.launch(seed=3, use_graph=use_graph)
        res = b.results()
        for i, (tr, c, r) in enumerate(zip(trs, cfgs, res)):
            ref = tr.run(c["n"], seed=3, config_id=i)
            assert r.els.tobytes() == ref.els.tobytes(), (i, c)
            assert int(r.counters[_abi.CNT_EVENTS]) == ref.events
            assert int(r.counters[_abi.CNT_ELECTRONS]) == c["n"]
