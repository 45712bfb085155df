"""The HDF5 store writer (mcd.hdf5) used for the ``input``, ``els`` and
``scats`` tables of doScatter.py (MC/doScatter.py:63-67, 112-117).

Pins: the lookup3 checksum against the published test vectors of lookup3.c
(Jenkins 2006, driver5); the files against an independent reader, the HDF5
library itself (libhdf5 through ctypes, when the environment has one), which
verifies every metadata checksum and decodes the compound datatypes; and a
round trip through this module's own reader.  Whether pandas.read_hdf (with
PyTables, absent here) reads these files as it reads the reference's is not
checked: that part of the layout is unpinned."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from mcd import hdf5


def test_lookup3_published_vectors():
    assert hdf5.lookup3(b"") == 0xDEADBEEF
    assert hdf5.lookup3(b"", 0xDEADBEEF) == 0xBD5B7DDE
    assert hdf5.lookup3(b"Four score and seven years ago") == 0x17770551
    assert hdf5.lookup3(b"Four score and seven years ago", 1) == 0xCD628161


def _frames():
    rng = np.random.default_rng(0)
    n1, n2 = 1000, 37
    els = [{"outcome": rng.choice(["bks", "abs", "trans", "long"], n),
            "final_E": rng.random(n).astype(np.float32) * 2e4,
            "n_scatt": rng.integers(0, 5000, n).astype(np.int32),
            "x": rng.standard_normal(n)} for n in (n1, n2)]
    return els


def _write(path):
    with hdf5.H5Store(path) as st:
        st.put_input({"num_el": 1037, "E0": 20000.0, "material": "Cu", "mode": "DS", "s_tilt": 70.0})
        for chunk in _frames():
            st.append("els", chunk, min_itemsize={"outcome": 6})
        st.append("scats", {"type": np.array(["Ruth", "Moller"]), "E": np.array([2e4, 1.9e4])},
                  min_itemsize={"type": 12})


def test_round_trip(tmp_path):
    path = str(tmp_path / "store.h5")
    _write(path)
    assert hdf5.keys(path) == ["els", "input", "scats"]
    df = hdf5.read_frame(path, "els")
    want = {k: np.concatenate([c[k] for c in _frames()]) for k in _frames()[0]}
    assert list(df.columns) == list(want)
    assert (df.index.values == np.arange(1037)).all()
    assert (df["outcome"].values == want["outcome"]).all()
    assert (df["final_E"].values == want["final_E"]).all() and df["final_E"].dtype == np.float32
    assert (df["n_scatt"].values == want["n_scatt"]).all() and df["n_scatt"].dtype == np.int32
    assert (df["x"].values == want["x"]).all()
    inp = hdf5.read_frame(path, "input")
    assert inp.shape == (1, 5) and inp["material"][0] == "Cu" and inp["E0"][0] == "20000.0"
    arr, tattrs, gattrs = hdf5.read_table(path, "els")
    assert tattrs["CLASS"] == "TABLE" and tattrs["NROWS"] == 1037
    assert [tattrs[f"FIELD_{i}_NAME"] for i in range(5)] == ["index", "outcome", "final_E", "n_scatt", "x"]
    assert gattrs["pandas_type"] == "frame_table" and gattrs["table_type"] == "appendable_frame"
    assert gattrs["index_cols"] == [(0, "index")] and gattrs["data_columns"] == list(want)
    assert arr.dtype["outcome"].itemsize == 6
    # the per-column attributes pandas reads back for a frame table
    # (pandas/io/pytables.py Table.indexables: <c>_kind, <c>_dtype, <c>_meta)
    for c in want:
        assert tattrs[f"{c}_kind"] == [c] and tattrs[f"{c}_meta"] is None
        assert tattrs[f"{c}_dtype"] == arr.dtype[c].name
    assert tattrs["outcome_dtype"] == "bytes48" and tattrs["final_E_dtype"] == "float32"
    import pandas.io.pytables as pt
    for c in want:  # the kind pandas derives from the stored dtype name
        pt._dtype_to_kind(tattrs[f"{c}_dtype"])
    assert not glob.glob(path + ".*.part")


def test_string_columns_ascii_and_not(tmp_path):
    """String columns are stored as fixed-width bytes: ASCII text (object or
    unicode arrays) converted in numpy, other text encoded one by one as UTF-8
    (the width then counts bytes); both read back as the same strings."""
    path = str(tmp_path / "s.h5")
    names = np.array(["Ruth", "Moller", "Gryz3p"], dtype=object)
    with hdf5.H5Store(path) as st:
        st.append("a", {"type": names[[0, 1, 2, 1]], "u": np.array(["x", "yz", "", "w"])})
        st.append("b", {"type": np.array(["Møller", "Ruth"], dtype=object)})
    arr, _, _ = hdf5.read_table(path, "a")
    assert arr.dtype["type"].itemsize == 6 and arr.dtype["u"].itemsize == 2
    assert list(hdf5.read_frame(path, "a")["type"]) == ["Ruth", "Moller", "Gryz3p", "Moller"]
    assert list(hdf5.read_frame(path, "a")["u"]) == ["x", "yz", "", "w"]
    arr, _, _ = hdf5.read_table(path, "b")
    assert arr.dtype["type"].itemsize == len("Møller".encode())
    assert list(hdf5.read_frame(path, "b")["type"]) == ["Møller", "Ruth"]


def test_column_width_and_schema_are_enforced(tmp_path):
    with pytest.raises(ValueError, match="min_itemsize"):
        with hdf5.H5Store(str(tmp_path / "a.h5")) as st:
            st.append("els", {"outcome": np.array(["bks"])}, min_itemsize={"outcome": 3})
            st.append("els", {"outcome": np.array(["trans"])})
    with pytest.raises(ValueError, match="columns"):
        with hdf5.H5Store(str(tmp_path / "b.h5")) as st:
            st.append("els", {"a": np.zeros(2)})
            st.append("els", {"b": np.zeros(2)})


def test_corrupted_metadata_is_detected(tmp_path):
    path = str(tmp_path / "store.h5")
    _write(path)
    raw = bytearray(open(path, "rb").read())
    i = raw.rfind(b"OHDR")          # the root group's object header
    raw[i + 12] ^= 0x01
    bad = str(tmp_path / "bad.h5")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        hdf5.keys(bad)


def test_attribute_pickles_name_no_code(tmp_path):
    """Object-valued attributes are read back only as plain data: a pickle
    that names a class or function is left as text."""
    assert hdf5._plain_loads(b"(lp0\nI1\na.") == [1]
    with pytest.raises(Exception):
        hdf5._plain_loads(b"cos\nsystem\n(S'true'\ntR.")


def _libhdf5():
    for p in sorted(glob.glob("/opt/conda/lib/libhdf5.so*")) + sorted(glob.glob("/usr/lib/x86_64-linux-gnu/libhdf5*.so*")):
        if "_hl" in p or "fortran" in p or "cpp" in p:
            continue
        try:
            return C.CDLL(p)
        except OSError:
            continue
    return None


def test_files_read_by_the_hdf5_library(tmp_path):
    """libhdf5 opens the file (superblock and object-header checksums
    verified by the library), finds the datasets by path, and reads the
    compound rows and attributes back byte for byte."""
    L = _libhdf5()
    if L is None:
        pytest.skip("no libhdf5 in this environment")
    path = str(tmp_path / "store.h5")
    _write(path)
    hid = C.c_int64
    L.H5open.restype = C.c_int
    assert L.H5open() >= 0
    for f in ("H5Fopen", "H5Dopen2", "H5Dget_type", "H5Dget_space", "H5Aopen_by_name", "H5Aget_type"):
        getattr(L, f).restype = hid
    L.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid]
    L.H5Dopen2.argtypes = [hid, C.c_char_p, hid]
    L.H5Dget_type.argtypes = [hid]
    L.H5Dget_space.argtypes = [hid]
    L.H5Sget_simple_extent_dims.argtypes = [hid, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.H5Tget_nmembers.argtypes = [hid]
    L.H5Tget_size.argtypes = [hid]
    L.H5Tget_size.restype = C.c_size_t
    L.H5Tget_member_name.argtypes = [hid, C.c_uint]
    L.H5Tget_member_name.restype = C.c_void_p
    L.H5Tget_member_offset.argtypes = [hid, C.c_uint]
    L.H5Tget_member_offset.restype = C.c_size_t
    L.H5Dread.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    L.H5Aopen_by_name.argtypes = [hid, C.c_char_p, C.c_char_p, hid, hid]
    L.H5Aget_type.argtypes = [hid]
    L.H5Aread.argtypes = [hid, hid, C.c_void_p]
    for f in ("H5Aclose", "H5Tclose", "H5Sclose", "H5Dclose", "H5Fclose"):
        getattr(L, f).argtypes = [hid]
    fid = L.H5Fopen(path.encode(), 0, 0)   # H5F_ACC_RDONLY, H5P_DEFAULT
    assert fid >= 0
    try:
        for key in ("els", "scats", "input"):
            arr, tattrs, _ = hdf5.read_table(path, key)
            d = L.H5Dopen2(fid, f"/{key}/table".encode(), 0)
            assert d >= 0
            t, sp = L.H5Dget_type(d), L.H5Dget_space(d)
            dims = (C.c_uint64 * 1)()
            assert L.H5Sget_simple_extent_dims(sp, dims, None) == 1 and dims[0] == len(arr)
            assert L.H5Tget_size(t) == arr.dtype.itemsize
            names = [C.string_at(L.H5Tget_member_name(t, i)).decode() for i in range(L.H5Tget_nmembers(t))]
            assert names == list(arr.dtype.names)
            assert [L.H5Tget_member_offset(t, i) for i in range(len(names))] == \
                [arr.dtype.fields[n][1] for n in names]
            buf = np.zeros(len(arr), dtype=arr.dtype)
            assert L.H5Dread(d, t, 0, 0, 0, buf.ctypes.data) >= 0   # H5S_ALL, H5P_DEFAULT
            assert buf.tobytes() == arr.tobytes()
            a = L.H5Aopen_by_name(fid, f"/{key}/table".encode(), b"NROWS", 0, 0)
            v = np.zeros(1, np.int64)
            at = L.H5Aget_type(a)
            assert L.H5Aread(a, at, v.ctypes.data) >= 0 and v[0] == tattrs["NROWS"]
            L.H5Tclose(at)
            L.H5Aclose(a)
            a = L.H5Aopen_by_name(fid, f"/{key}".encode(), b"pandas_type", 0, 0)
            at = L.H5Aget_type(a)
            s = C.create_string_buffer(L.H5Tget_size(at))
            assert L.H5Aread(a, at, s) >= 0 and s.raw.rstrip(b"\0") == b"frame_table"
            L.H5Tclose(at)
            L.H5Aclose(a)
            L.H5Tclose(t)
            L.H5Sclose(sp)
            L.H5Dclose(d)
    finally:
        L.H5Fclose(fid)
    os.remove(path)
