"""Tables built by the library from input-file parameters (mcd_tables_build:
grid search plus the device CDF fill) and parameter-keyed table files
(mcd_tables_save / mcd_tables_load), the counterparts of
probTables.genTables / mapToMemory / readFromMemory (probTables.py:489-614).

Pins: the grids equal the Python search's (scipy quad) and the CDFs equal the
tables built from those grids bit for bit, hence the float32 restatement of
computeBlock within 1e-5; a table file gives back the same tables and the same
electron records; a file built for other parameters is refused (the
reference silently reuses a stale '{type}_{shell}.table', Q12)."""
import os

import numpy as np
import pytest

import tables_ref as TR
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no GPU")]

C1 = dict(material="Si", E0=10000.0, Emin=1000.0, Wc=100.0, tol_E=1e-2, tol_W=1e-2, s_tilt=70.0)
C2 = dict(material="Cu", E0=20000.0, Emin=1000.0, Wc=50.0, tol_E=1e-3, tol_W=1e-3, s_tilt=70.0)


def _tr(cfg, **kw):
    from mcd.engine import Transport
    c = dict(cfg)
    return Transport(c.pop("material"), c.pop("E0"), c.pop("Emin"), **c, **kw)


@pytest.mark.parametrize("cfg", [C1, C2], ids=["Si10", "Cu20"])
def test_library_tables_equal_python_built_tables(cfg):
    lib_tr, py_tr = _tr(cfg, grid_search="library"), _tr(cfg)
    for q in (0, 1):
        a, b = lib_tr.grids[q], py_tr.grids[q]
        assert np.array_equal(a.Es, b.Es) and np.array_equal(a.Ws, b.Ws)
        ca, cb = lib_tr.cdf(q), py_tr.cdf(q)
        assert np.array_equal(ca, cb)
        mat = lib_tr.mat
        ref = TR.cdf_table(a.Es, a.Ws, mat.fermi_e, "Moller" if q == 0 else "Gryzinski", a.n_el, a.Eb)
        assert np.abs(ca - ref).max() <= 1e-5


def test_table_file_round_trip_and_records(tmp_path):
    import torch
    from mcd import _abi
    built = _tr(C2, grid_search="library")
    path = str(tmp_path / "cu20.mcdt")
    built.save_tables(path)
    loaded = _tr(C2, table_file=path)
    for q in (0, 1):
        assert np.array_equal(built.grids[q].Es, loaded.grids[q].Es)
        assert np.array_equal(built.grids[q].Ws, loaded.grids[q].Ws)
        assert np.array_equal(built.cdf(q), loaded.cdf(q))
    n = 1 << 14
    dev = torch.device("cuda", 0)
    outs = []
    for tr in (built, loaded):
        els = torch.empty(n * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
        tr.launch(n, els, cnt, seed=1)
        torch.cuda.synchronize(dev)
        outs.append(els.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_table_file_for_other_parameters_is_refused(tmp_path):
    from mcd import _abi
    from mcd._lib import MCDError
    from mcd.material import Material
    path = str(tmp_path / "si10.mcdt")
    _tr(C1, grid_search="library").save_tables(path)
    for change in (dict(Wc=101.0), dict(tol_W=2e-2), dict(Emin=900.0), dict(material=Material("Si", 1.01))):
        cfg = dict(C1)
        cfg.update(change)
        with pytest.raises(MCDError) as e:
            _tr(cfg, table_file=path)
        assert e.value.code == _abi.ERR_TABLE and "other parameters" in str(e.value)
    raw = open(path, "rb").read()
    bad = str(tmp_path / "short.mcdt")
    open(bad, "wb").write(raw[:-100])
    with pytest.raises(MCDError) as e:
        _tr(C1, table_file=bad)
    assert e.value.code == _abi.ERR_TABLE
    open(bad, "wb").write(b"not a table file")
    with pytest.raises(MCDError) as e:
        _tr(C1, table_file=bad)
    assert e.value.code == _abi.ERR_TABLE
    # tables built from given grids carry no tolerances: nothing to key a file with
    with pytest.raises(MCDError):
        _tr(C1).save_tables(str(tmp_path / "x.mcdt"))


def test_gen_tables_no_reads_the_keyed_file(tmp_path):
    """doScatter.py with gen_tables: yes writes the keyed table file; the same
    input with gen_tables: no reads it back and gives the same store; a run
    with another Wc and gen_tables: no gets its own file, not the stale one."""
    import doScatter
    from test_gpu_dropin import read_store
    text = ("num_el : 600\nE0 : 10000\nEmin : 1000\nmaterial : Si\ns_tilt : 70\nelastic : Ruth_vanilla\n"
            "mode : DS\nBethe_model : classical\ndiffMFP : no\ngen_tables : {gen}\ntol_E : 1e-2\ntol_W : 1e-2\n"
            "maxScatt : off\nWc : {wc}\nseed : 0\nelectron_output : outcome, final_E\nscatter_output : type, E_loss\n")
    out = tmp_path / "out"
    f = tmp_path / "in.file"
    f.write_text(text.format(gen="yes", wc=100))
    p1 = doScatter.main(["-i", str(f), "-o", str(out), "-c", "256"])
    files = sorted(os.listdir(out / "tables"))
    assert len(files) == 1 and files[0].endswith(".mcdt")
    mtime = os.path.getmtime(out / "tables" / files[0])
    a_els, a_scats = read_store(p1, "els"), read_store(p1, "scats")
    f.write_text(text.format(gen="no", wc=100))
    p2 = doScatter.main(["-i", str(f), "-o", str(out), "-c", "256"])
    assert os.path.getmtime(out / "tables" / files[0]) == mtime
    b_els, b_scats = read_store(p2, "els"), read_store(p2, "scats")
    assert a_els.equals(b_els) and a_scats.equals(b_scats)
    f.write_text(text.format(gen="no", wc=120))
    doScatter.main(["-i", str(f), "-o", str(out), "-c", "256"])
    assert len(os.listdir(out / "tables")) == 2


def test_device_grid_search_equals_host_search():
    """mcd_grid_search_batch (one device launch, a workgroup per table) picks
    the grid sizes the host search (mcd_grid_search, itself equal bit for bit
    to the scipy-based restatement) picks, for the 44 tables of the host
    search's test cases: C1, C2, the reference's test input at tol_W 1e-7, the
    Wc range ends and the batched workload's DS tables, one launch per
    tolerance pair."""
    from mcd.engine import grid_search_batch
    from mcd.material import Material
    from mcd.tables import gryz_grid, moller_grid
    from test_table_builder import CASES
    groups = {}
    for mat, E0, Emin, Wc, tE, tW in CASES:
        m = Material(mat)
        groups.setdefault((tE, tW), []).extend([moller_grid(m, E0, Emin, Wc), gryz_grid(m, E0, Emin)])
    for (tE, tW), grids in groups.items():
        nE, nW, st = grid_search_batch(grids, tE, tW, statuses=True)
        assert not st.any()
        for g, e, w in zip(grids, nE, nW):
            h = g.build_grids_library(tE, tW)
            assert (int(e), int(w)) == (h.Es.size, h.Ws.size), (g.kind, g.Emax, g.Wmin, tE, tW)


@pytest.mark.parametrize("tol", [1e-2, 1e-3, 1e-4])
def test_device_grid_search_random_configurations(tol):
    """Device and host searches pick the same sizes (or both fail) for 48
    random DS configurations per tolerance (seeded: material, E0 in 3-30 keV,
    Emin from 300 eV to 0.7 E0, Wc in 10-300 eV), 96 tables each."""
    from mcd.engine import grid_search_batch
    from mcd.material import Material
    from mcd.tables import gryz_grid, moller_grid
    rng = np.random.default_rng(int(-np.log10(tol)) * 7919)
    mats = {m: Material(m) for m in ("Al", "Si", "Cu", "Au")}
    grids = []
    while len(grids) < 96:
        m = mats[rng.choice(list(mats))]
        E0 = float(np.round(rng.uniform(3000.0, 30000.0), 1))
        Emin = float(np.round(max(300.0, E0 * rng.uniform(0.05, 0.7)), 1))
        Wc = float(np.round(rng.uniform(10.0, 300.0), 2))
        grids += [moller_grid(m, E0, Emin, Wc), gryz_grid(m, E0, Emin)]
    nE, nW, st = grid_search_batch(grids, tol, tol, statuses=True)
    for g, e, w, s_ in zip(grids, nE, nW, st):
        try:
            h = g.build_grids_library(tol, tol)
        except Exception:
            assert s_ != 0, (g.kind, g.Emax, g.Emin, g.Wmin)
            continue
        assert s_ == 0 and (int(e), int(w)) == (h.Es.size, h.Ws.size), (g.kind, g.Emax, g.Emin, g.Wmin, tol)


def test_device_grid_search_many_tables_in_chunks():
    """More tables than one launch takes (256 per launch): five copies of the
    batched workload's 64 DS tables give five copies of the 64 tables' sizes."""
    from mcd import workloads
    from mcd.engine import grid_search_batch
    from mcd.material import Material
    from mcd.tables import gryz_grid, moller_grid
    grids = []
    for c in workloads.batched_configs():
        if c["mode"] == "DS":
            m = Material(c["material"])
            grids += [moller_grid(m, c["E0"], c["Emin"], c["Wc"]), gryz_grid(m, c["E0"], c["Emin"])]
    nE1, nW1 = grid_search_batch(grids, 1e-3, 1e-3)
    nE, nW = grid_search_batch(grids * 5, 1e-3, 1e-3)
    assert len(grids) * 5 > 256
    assert np.array_equal(nE, np.tile(nE1, 5)) and np.array_equal(nW, np.tile(nW1, 5))


def test_device_grid_search_reports_each_table():
    """A table no grid can serve (a Moller table whose loss range is empty:
    Wc above W_max(E0)) gets ERR_TABLE in its own status slot while the others
    are searched; without the status array the call fails naming it."""
    from mcd import _abi
    from mcd._lib import MCDError
    from mcd.engine import grid_search_batch
    from mcd.material import Material
    from mcd.tables import gryz_grid, moller_grid
    si = Material("Si")
    grids = [moller_grid(si, 10000.0, 1000.0, 100.0), moller_grid(si, 1000.0, 500.0, 600.0),
             gryz_grid(si, 10000.0, 1000.0)]
    nE, nW, st = grid_search_batch(grids, 1e-2, 1e-2, statuses=True)
    assert list(st) == [0, _abi.ERR_TABLE, 0]
    assert (int(nE[0]), int(nW[0]) - 1) == (300, 29) and (int(nE[2]), int(nW[2]) - 1) == (400, 39)
    with pytest.raises(MCDError) as e:
        grid_search_batch(grids, 1e-2, 1e-2)
    assert e.value.code == _abi.ERR_TABLE and "table 1" in str(e.value)


def test_batched_build_equals_per_configuration_builds():
    """mcd_tables_build_batch (every table's grid search in one device launch,
    one launch for the CDF rows of every table and one for their row records)
    gives each configuration the tables mcd_tables_build (host grid search)
    gives it: grids, CDFs and channel-pick tables bit for bit, and the same
    records; continuous configurations in the batch get their tables too.  The
    64-configuration workload's DS tables at tol 1e-3, in FP64 and in the
    reference's float32 arithmetic; the same with the searches on host threads
    (MCD_GRID_SEARCH=host)."""
    import torch
    from mcd import _abi, workloads
    from mcd.engine import Transport, build_transports_batch
    cfgs = workloads.batched_configs()
    ds = [c for c in cfgs if c["mode"] == "DS" and c["s_tilt"] == 70.0]
    specs = [dict(material=c["material"], E0=c["E0"], Emin=c["Emin"], Wc=c["Wc"], s_tilt=c["s_tilt"], mode="DS")
             for c in ds]
    specs.append(dict(material="Au", E0=30000.0, Emin=1000.0, s_tilt=0.0, mode="cont", bethe_model="JL"))
    dev = torch.device("cuda", 0)
    for prec, where in ((_abi.PREC_FP64, None), (_abi.PREC_FP32, None), (_abi.PREC_FP64, "host")):
        if where:
            os.environ["MCD_GRID_SEARCH"] = where
        try:
            batch = build_transports_batch(specs, tol_E=1e-3, tol_W=1e-3, table_precision=prec)
        finally:
            os.environ.pop("MCD_GRID_SEARCH", None)
        assert len(batch) == len(specs)
        for sp, b in zip(specs, batch):
            one = Transport(tol_E=1e-3, tol_W=1e-3, table_precision=prec, grid_search="library", **sp)
            if sp["mode"] == "DS":
                for q in (0, 1):
                    assert np.array_equal(b.grids[q].Es, one.grids[q].Es) and np.array_equal(b.grids[q].Ws,
                                                                                             one.grids[q].Ws)
                    assert np.array_equal(b.cdf(q), one.cdf(q)), (sp, q, prec)
                assert b.pick_info() == one.pick_info()
            if prec == _abi.PREC_FP64:
                n = 1 << 12
                outs = []
                for tr in (b, one):
                    els = torch.empty(n * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
                    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
                    tr.launch(n, els, cnt, seed=5)
                    torch.cuda.synchronize(dev)
                    outs.append(els.cpu().numpy())
                assert np.array_equal(outs[0], outs[1]), sp
        del batch


def test_batched_build_names_the_failing_configuration(tmp_path):
    from mcd import _abi
    from mcd._lib import MCDError
    from mcd.engine import build_transports_batch
    specs = [dict(material="Si", E0=10000.0, Emin=1000.0, Wc=100.0, s_tilt=70.0, mode="DS"),
             dict(material="Cu", E0=20000.0, Emin=1000.0, Wc=50.0, s_tilt=70.0, mode="DS")]
    good = build_transports_batch(specs, tol_E=1e-2, tol_W=1e-2)
    good[1].save_tables(str(tmp_path / "cu.mcdt"))  # keyed tables, as from mcd_tables_build
    with pytest.raises(MCDError) as e:
        build_transports_batch(specs, tol_E=0.0, tol_W=1e-2)
    assert e.value.code == _abi.ERR_INVALID
    bad = specs + [dict(material="Al", E0=5000.0, Emin=6000.0, Wc=100.0, mode="DS")]  # Emin above E0
    with pytest.raises(MCDError) as e:
        build_transports_batch(bad, tol_E=1e-2, tol_W=1e-2)
    assert "configuration 2" in str(e.value)
