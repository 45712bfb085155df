"""The reference-shaped host interface on the GPU: multiTraj_DS /
multiTraj_cont return the per-electron dicts of multiScatter.py:66-116, and
doScatter.py writes the els / scats tables."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def si_input(**kw):
    p = dict(num_el=64, E0=10000.0, Emin=1000.0, material="Si", s_tilt=70.0, thickness=float("inf"),
             elastic="Ruth_vanilla", mode="DS", Bethe_model="classical", diffMFP=False, gen_tables=True,
             tol_E=1e-2, tol_W=1e-2, maxScatt=0, Wc=100.0, seed=0,
             electron_output=["outcome", "final_E", "final_dir", "last_pos"],
             scatter_output=["type", "E", "E_loss", "pol_angle", "az_angle", "pathl", "position", "energy"])
    p.update(kw)
    return p


def test_multitraj_ds_matches_oracle_electrons():
    from mcd.material import Material
    from mcd.multiScatter import genTables, multiTraj_DS
    p = si_input()
    tables = genTables(p)
    mat = Material("Si")
    things = {"el_output": p["electron_output"], "scat_output": p["scatter_output"]}
    res = multiTraj_DS(p, 4, mat, tables, things) + multiTraj_DS(p, 4, mat, tables, things)
    assert len(res) == 8
    tr = tables.transport(p)
    gm, gg = tr.grids
    o, trace, _ = O.run(tr.cfg, O.Table(gm.Es, gm.Ws, tr.cdf(0)), O.Table(gg.Es, gg.Ws, tr.cdf(1)), 0, 0, 8,
                        trace_cap=100000)
    names = {1: "abs", 2: "bks", 3: "trsm"}
    for i, r in enumerate(res):
        assert r["els"]["outcome"][0] == names[o["outcome"][i]]
        assert r["els"]["final_E"][0] == pytest.approx(o["final_E"][i], rel=1e-5)
        assert set(r["els"]["final_dir"]) == {"dx", "dy", "dz"}
        assert set(r["els"]["last_pos"]) == {"last_x", "last_y", "last_z"}
        t = trace[trace["electron"] == i]
        assert len(r["scats"]["type"]) == len(t)
        assert r["scats"]["E_loss"] == pytest.approx(list(t["E_loss"]), rel=1e-6, abs=1e-6)
        assert set(r["scats"]["position"]) == {"x", "y", "z"}


def test_multitraj_cont_signature_and_dicts():
    from mcd.multiScatter import multiTraj_cont
    p = si_input(mode="cont", Bethe_model="JL")
    res = multiTraj_cont(p, 4, "Si", {"el_output": ["outcome", "final_E"], "scat_output": []}, None)
    assert len(res) == 4 and all(r["els"]["outcome"][0] in ("abs", "bks", "trsm") for r in res)


def test_doscatter_cli_writes_tables(tmp_path):
    import pandas as pd
    import doScatter
    f = tmp_path / "in.file"
    f.write_text("num_el : 3000\nE0 : 10000\nEmin : 1000\nmaterial : Si\ns_tilt : 70\nelastic : Ruth_vanilla\n"
                 "mode : DS\nBethe_model : classical\ndiffMFP : no\ngen_tables : yes\ntol_E : 1e-2\ntol_W : 1e-2\n"
                 "maxScatt : off\nWc : 100\nseed : 0\nelectron_output : outcome, final_E, final_dir, last_pos\n"
                 "scatter_output : type, E, energy\n")
    path = doScatter.main(["-i", str(f), "-o", str(tmp_path / "out"), "-c", "1000"])
    if path.endswith(".h5") and os.path.isfile(path):
        els = pd.read_hdf(path, "els")
        scats = pd.read_hdf(path, "scats")
    else:
        els = pd.concat([pd.read_parquet(os.path.join(path, f)) for f in sorted(os.listdir(path))
                         if f.startswith("els")])
        scats = pd.concat([pd.read_parquet(os.path.join(path, f)) for f in sorted(os.listdir(path))
                           if f.startswith("scats")])
    assert len(els) >= 3000 and list(els.columns[:2]) == ["outcome", "final_E"]
    assert set(els["outcome"]) <= {"abs", "bks", "trsm"}
    assert list(scats.columns) == ["type", "E", "energy"]
    assert np.all(scats["E"] >= scats["energy"])
    assert set(scats["type"]) <= {"Ruth", "Moller", "Quinn", "Gryz1s", "Gryz2s", "Gryz2p"}


def test_doscatter_without_scatter_rows_pipelines_launches(tmp_path):
    """With no scatter_output, doScatter runs its chunks as launches pipelined
    over two streams; the els table equals one launch over all electrons."""
    import pandas as pd
    import doScatter
    from mcd.multiScatter import genTables, run_columns
    f = tmp_path / "in.file"
    text = ("num_el : 2500\nE0 : 10000\nEmin : 1000\nmaterial : Si\ns_tilt : 70\nelastic : Ruth_vanilla\n"
            "mode : DS\nBethe_model : classical\ndiffMFP : no\ngen_tables : yes\ntol_E : 1e-2\ntol_W : 1e-2\n"
            "maxScatt : off\nWc : 100\nseed : 0\nelectron_output : outcome, final_E, final_dir, last_pos\n")
    f.write_text(text)
    path = doScatter.main(["-i", str(f), "-o", str(tmp_path / "out"), "-c", "1000"])
    if path.endswith(".h5") and os.path.isfile(path):
        els = pd.read_hdf(path, "els")
    else:
        els = pd.concat([pd.read_parquet(os.path.join(path, f)) for f in sorted(os.listdir(path))
                         if f.startswith("els")], ignore_index=True)
    from mcd.fileTools import readInput
    p = readInput(str(f))
    one, _, _ = run_columns(p, 2500, tables=genTables(p), first=0)
    assert len(els) == len(one["final_E"])
    assert np.array_equal(els["final_E"].to_numpy(), np.asarray(one["final_E"]))
    assert list(els["outcome"]) == list(one["outcome"])
