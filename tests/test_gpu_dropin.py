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


NAMES = {1: "abs", 2: "bks", 3: "trsm", 5: "long"}


def oracle_columns(o, trace, material):
    """The zipped els / scats columns the reference's listener would append for
    oracle records and trace (thingsToSave + zipDict, fileTools.py:67-99,
    193-242): an electron that exits and is absorbed in one step has two els
    rows (electron.py:52-54)."""
    from mcd.fileTools import channel_names
    els = {k: [] for k in ("outcome", "final_E", "dx", "dy", "dz", "last_x", "last_y", "last_z")}
    for e in o:
        outs = [e["outcome"]] + ([e["outcome2"]] if e["outcome2"] else [])
        for oc in outs:
            els["outcome"].append(NAMES[int(oc)])
            els["final_E"].append(float(e["final_E"]))
            for j, k in enumerate(("dx", "dy", "dz")):
                els[k].append(float(e["dir"][j]))
            for j, k in enumerate(("last_x", "last_y", "last_z")):
                els[k].append(float(e["pos"][j]))
    names = channel_names(material)
    scats = {"type": [names[int(c)] for c in trace["channel"]], "E": list(trace["E"]),
             "E_loss": list(trace["E_loss"]), "pol_angle": list(trace["pol"]), "az_angle": list(trace["az"]),
             "pathl": list(trace["pathl"]), "energy": list(trace["energy"]),
             "x": list(trace["x"]), "y": list(trace["y"]), "z": list(trace["z"])}
    return els, scats


def compare_columns(got_els, got_scats, want_els, want_scats, table_losses=True):
    """table_losses: DS losses are table values or constants, equal as float32;
    continuous-mode losses are l * S evaluated in FP32 (relative 2e-5, as the
    path length)."""
    assert list(got_els["outcome"]) == want_els["outcome"]
    assert np.allclose(got_els["final_E"], want_els["final_E"], rtol=1e-5)
    for k in ("dx", "dy", "dz"):
        assert np.abs(np.asarray(got_els[k]) - want_els[k]).max() < 1e-5, k
    for k in ("last_x", "last_y", "last_z"):
        assert np.allclose(got_els[k], want_els[k], rtol=1e-5, atol=1e-2), k
    assert list(got_scats["type"]) == want_scats["type"]
    if table_losses:
        assert np.array_equal(np.asarray(got_scats["E_loss"], np.float32),
                              np.asarray(want_scats["E_loss"], np.float32))
    else:
        assert np.allclose(got_scats["E_loss"], want_scats["E_loss"], rtol=2e-5)
    for k, rtol in (("E", 1e-6), ("energy", 1e-6), ("az_angle", 1e-6), ("pathl", 2e-5)):
        assert np.allclose(got_scats[k], want_scats[k], rtol=rtol), k
    assert np.allclose(got_scats["pol_angle"], want_scats["pol_angle"], rtol=1e-5, atol=1e-6)
    for k in ("x", "y", "z"):
        assert np.allclose(got_scats[k], want_scats[k], rtol=1e-5, atol=1e-2), k


ALL_SCATS = ["type", "E", "E_loss", "pol_angle", "az_angle", "pathl", "position", "energy"]


def test_scatter_multi_el_ds_matches_oracle():
    """The legacy DS entry point scatterMultiEl_DS (multiScatter.py:15-61):
    inputPar['num_el'] trajectories, zipped els / scats columns in outputQ,
    against the oracle's records and per-step trace of the same electrons."""
    from mcd.material import Material
    from mcd.multiScatter import genTables, scatterMultiEl_DS
    p = si_input(num_el=48, seed=5)
    tables = genTables(p)
    q = scatterMultiEl_DS(p, tables, {"el_output": p["electron_output"], "scat_output": ALL_SCATS}, {})
    tr = tables.transport(p)
    gm, gg = tr.grids
    o, trace, _ = O.run(tr.cfg, O.Table(gm.Es, gm.Ws, tr.cdf(0)), O.Table(gg.Es, gg.Ws, tr.cdf(1)), 5, 0, 48,
                        trace_cap=200000)
    want_els, want_scats = oracle_columns(o, trace, Material("Si"))
    compare_columns(q["els"], q["scats"], want_els, want_scats)


@pytest.mark.parametrize("bethe", ["classical", "JL", "explicit"])
def test_scatter_multi_el_cont_matches_oracle(bethe):
    """The legacy continuous entry point scatterMultiEl_cont
    (multiScatter.py:118-168; its reference body fails with a NameError at
    :131) with each Bethe law, against the oracle's continuous trajectories."""
    from mcd.material import Material
    from mcd.multiScatter import scatterMultiEl_cont
    p = si_input(mode="cont", Bethe_model=bethe, seed=7)
    things = {"el_output": p["electron_output"], "scat_output": ALL_SCATS}
    out = scatterMultiEl_cont(p, things, {}, 40)
    from mcd import _abi
    cfg = _abi.make_config(Material("Si"), p["E0"], p["Emin"], p["Wc"], p["s_tilt"], mode="cont",
                           bethe_model=bethe)
    o, trace, _ = O.run(cfg, None, None, 7, 0, 40, trace_cap=400000)
    want_els, want_scats = oracle_columns(o, trace, Material("Si"))
    compare_columns(out["electrons"], out["scatterings"], want_els, want_scats, table_losses=False)


def test_multitraj_cont_matches_oracle():
    """multiTraj_cont (multiScatter.py:171-229, unreachable in the reference:
    its call at :206-207 passes 5 arguments to a 6-parameter function) returns
    per-electron dicts equal to the oracle's continuous trajectories; two
    calls on one tables object continue the electron numbering."""
    from mcd.material import Material
    from mcd.multiScatter import genTables, multiTraj_cont
    from mcd.fileTools import zipDict
    p = si_input(mode="cont", Bethe_model="JL", seed=9)
    tables = genTables(p)
    things = {"el_output": p["electron_output"], "scat_output": ALL_SCATS}
    res = multiTraj_cont(p, 4, Material("Si"), things, tables) + multiTraj_cont(p, 4, Material("Si"), things, tables)
    assert len(res) == 8
    els, scats = {}, {}
    for r in res:
        zipDict(els, r["els"])
        zipDict(scats, r["scats"])
    from mcd import _abi
    cfg = _abi.make_config(Material("Si"), p["E0"], p["Emin"], p["Wc"], p["s_tilt"], mode="cont", bethe_model="JL")
    o, trace, _ = O.run(cfg, None, None, 9, 0, 8, trace_cap=200000)
    want_els, want_scats = oracle_columns(o, trace, Material("Si"))
    compare_columns(els, scats, want_els, want_scats, table_losses=False)


def read_store(path, key):
    """A table of the CLI's HDF5 store: pandas.read_hdf where PyTables is
    importable, else the store module's own reader."""
    import pandas as pd
    try:
        import tables  # noqa: F401
        return pd.read_hdf(path, key)
    except ImportError:
        from mcd.hdf5 import read_frame
        return read_frame(path, key)


def test_doscatter_cli_output_matches_oracle(tmp_path):
    """doScatter.py end to end (input file -> tables -> chunked GPU launches
    -> HDF5 store) against the oracle's records and per-step trace of the same
    electrons: Si 10 keV -> 1 keV, 70 deg, Wc 100 eV, tol 1e-2, seed 0,
    200 electrons in chunks of 64, every els and scats column the reference's
    listener writes (doScatter.py:100-117, fileTools.py:67-99, 193-242)."""
    import doScatter
    from mcd.fileTools import readInput
    from mcd.material import Material
    from mcd.multiScatter import genTables
    f = tmp_path / "in.file"
    f.write_text("num_el : 200\nE0 : 10000\nEmin : 1000\nmaterial : Si\ns_tilt : 70\nelastic : Ruth_vanilla\n"
                 "mode : DS\nBethe_model : classical\ndiffMFP : no\ngen_tables : yes\ntol_E : 1e-2\ntol_W : 1e-2\n"
                 "maxScatt : off\nWc : 100\nseed : 0\nelectron_output : outcome, final_E, final_dir, last_pos\n"
                 "scatter_output : type, E, E_loss, pol_angle, az_angle, pathl, position, energy\n")
    path = doScatter.main(["-i", str(f), "-o", str(tmp_path / "out"), "-c", "64"])
    els, scats = read_store(path, "els"), read_store(path, "scats")
    inp = read_store(path, "input")
    assert inp.shape[0] == 1 and str(inp["material"].iloc[0]) == "Si" and int(float(inp["num_el"].iloc[0])) == 200
    p = readInput(str(f))
    tr = genTables(p).transport(p)
    gm, gg = tr.grids
    o, trace, _ = O.run(tr.cfg, O.Table(gm.Es, gm.Ws, tr.cdf(0)), O.Table(gg.Es, gg.Ws, tr.cdf(1)), 0, 0, 200,
                        trace_cap=400000)
    want_els, want_scats = oracle_columns(o, trace, Material("Si"))
    compare_columns({k: els[k].to_numpy() for k in els.columns}, {k: scats[k].to_numpy() for k in scats.columns},
                    want_els, want_scats)


def test_doscatter_cli_writes_tables(tmp_path):
    import doScatter
    f = tmp_path / "in.file"
    f.write_text("num_el : 3000\nE0 : 10000\nEmin : 1000\nmaterial : Si\ns_tilt : 70\nelastic : Ruth_vanilla\n"
                 "mode : DS\nBethe_model : classical\ndiffMFP : no\ngen_tables : yes\ntol_E : 1e-2\ntol_W : 1e-2\n"
                 "maxScatt : off\nWc : 100\nseed : 0\nelectron_output : outcome, final_E, final_dir, last_pos\n"
                 "scatter_output : type, E, energy\n")
    path = doScatter.main(["-i", str(f), "-o", str(tmp_path / "out"), "-c", "1000"])
    assert path.endswith(".h5") and os.path.isfile(path)
    els, scats = read_store(path, "els"), read_store(path, "scats")
    assert len(els) >= 3000 and list(els.columns[:2]) == ["outcome", "final_E"]
    assert set(els["outcome"]) <= {"abs", "bks", "trsm"}
    assert list(scats.columns) == ["type", "E", "energy"]
    assert np.all(scats["E"] >= scats["energy"])
    assert set(scats["type"]) <= {"Ruth", "Moller", "Quinn", "Gryz1s", "Gryz2s", "Gryz2p"}


def test_doscatter_without_scatter_rows_pipelines_launches(tmp_path):
    """With no scatter_output, doScatter runs its chunks as launches pipelined
    over two streams; the els table equals one launch over all electrons."""
    import doScatter
    from mcd.multiScatter import genTables, run_columns
    f = tmp_path / "in.file"
    text = ("num_el : 2500\nE0 : 10000\nEmin : 1000\nmaterial : Si\ns_tilt : 70\nelastic : Ruth_vanilla\n"
            "mode : DS\nBethe_model : classical\ndiffMFP : no\ngen_tables : yes\ntol_E : 1e-2\ntol_W : 1e-2\n"
            "maxScatt : off\nWc : 100\nseed : 0\nelectron_output : outcome, final_E, final_dir, last_pos\n")
    f.write_text(text)
    path = doScatter.main(["-i", str(f), "-o", str(tmp_path / "out"), "-c", "1000"])
    els = read_store(path, "els")
    from mcd.fileTools import readInput
    p = readInput(str(f))
    one, _, _ = run_columns(p, 2500, tables=genTables(p), first=0)
    assert len(els) == len(one["final_E"])
    assert np.array_equal(els["final_E"].to_numpy(), np.asarray(one["final_E"]))
    assert list(els["outcome"]) == list(one["outcome"])
