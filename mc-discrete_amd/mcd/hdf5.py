"""A small HDF5 writer and reader for the transport's output store.

The reference writes its results with ``pandas.HDFStore`` (PyTables):
``store.put(key, df, format='table', data_columns=True, append=True)`` for
``els`` and ``scats`` and ``store['input'] = ...`` for the run parameters
(MC/doScatter.py:63-67, 112-117, 198-205).  PyTables is not part of this
environment, so this module writes the same store itself, from the HDF5 file
format specification (version 3.0 structures):

* superblock version 2, 8-byte offsets and lengths, every metadata block
  carrying its Jenkins lookup3 checksum;
* groups as version-2 object headers holding Link Info, Group Info and one
  Link message per child (compact link storage);
* each frame as the group ``/<key>`` with the dataset ``/<key>/table``: a
  one-dimensional compound dataset (``index`` int64, then one field per
  column, as ``data_columns=True`` gives), stored contiguously;
* the attributes PyTables and pandas put on those objects (CLASS, VERSION,
  TITLE, FIELD_<i>_NAME, NROWS, index_kind and, per data column, <c>_kind,
  <c>_meta and <c>_dtype; pandas_type 'frame_table', table_type
  'appendable_frame', index_cols, values_cols, non_index_axes, data_columns,
  nan_rep, encoding, levels, info: the attributes pandas' HDFStore reads back
  for a frame table, pandas/io/pytables.py Table.get_attrs / indexables).
  Object-valued pandas attributes are stored as protocol-0 pickles in string
  attributes.

The HDF5 structure is checked against the HDF5 library itself where one is
installed (the test suite reads these files through libhdf5); that
``pandas.read_hdf`` reads them as the reference's files is not checked here
(no PyTables), so the pandas layout is a restatement, not a pinned format.
``read_frame`` reads the files this module writes.

Rows are appended in chunks (each chunk spills to a side file) and the file
is assembled on ``close``: datasets are contiguous, so their sizes must be
final before the metadata is written.
"""
from __future__ import annotations

import os
import pickle
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"


# ---------------------------------------------------------------------------
# Jenkins lookup3 hashlittle (public domain, B. Jenkins 2006), the checksum of
# HDF5 metadata blocks (initial value 0).
def _rot(x, k):
    return ((x << k) | (x >> (32 - k))) & 0xFFFFFFFF


def lookup3(data: bytes, initval: int = 0) -> int:
    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & 0xFFFFFFFF
    i = 0
    M = 0xFFFFFFFF
    while n > 12:
        a = (a + int.from_bytes(data[i:i + 4], "little")) & M
        b = (b + int.from_bytes(data[i + 4:i + 8], "little")) & M
        c = (c + int.from_bytes(data[i + 8:i + 12], "little")) & M
        a = (a - c) & M; a ^= _rot(c, 4); c = (c + b) & M
        b = (b - a) & M; b ^= _rot(a, 6); a = (a + c) & M
        c = (c - b) & M; c ^= _rot(b, 8); b = (b + a) & M
        a = (a - c) & M; a ^= _rot(c, 16); c = (c + b) & M
        b = (b - a) & M; b ^= _rot(a, 19); a = (a + c) & M
        c = (c - b) & M; c ^= _rot(b, 4); b = (b + a) & M
        n -= 12
        i += 12
    if n == 0:
        return c
    tail = data[i:i + n] + b"\0" * (12 - n)
    a = (a + int.from_bytes(tail[0:4], "little")) & M
    b = (b + int.from_bytes(tail[4:8], "little")) & M
    c = (c + int.from_bytes(tail[8:12], "little")) & M
    c ^= b; c = (c - _rot(b, 14)) & M
    a ^= c; a = (a - _rot(c, 11)) & M
    b ^= a; b = (b - _rot(a, 25)) & M
    c ^= b; c = (c - _rot(b, 16)) & M
    a ^= c; a = (a - _rot(c, 4)) & M
    b ^= a; b = (b - _rot(a, 14)) & M
    c ^= b; c = (c - _rot(b, 24)) & M
    return c


# ---------------------------------------------------------------------------
# Datatype and dataspace messages.
def _dtype_msg(dt: np.dtype) -> bytes:
    """HDF5 datatype encoding of a little-endian numpy dtype."""
    dt = np.dtype(dt)
    if dt.names:
        size = dt.itemsize
        width = 1 if size < 256 else 2 if size < 65536 else 3 if size < (1 << 24) else 4
        props = b""
        for name in dt.names:
            sub, off = dt.fields[name][0], dt.fields[name][1]
            props += name.encode() + b"\0" + off.to_bytes(width, "little") + _dtype_msg(sub)
        n = len(dt.names)
        return bytes([0x36]) + struct.pack("<HB", n, 0) + struct.pack("<I", size) + props
    if dt.kind in "iu":
        bits = 0x08 if dt.kind == "i" else 0
        return bytes([0x10, bits, 0, 0]) + struct.pack("<I", dt.itemsize) + struct.pack("<HH", 0, 8 * dt.itemsize)
    if dt.kind == "f":
        if dt.itemsize == 4:
            sign, eloc, esz, msz, bias = 31, 23, 8, 23, 127
        elif dt.itemsize == 8:
            sign, eloc, esz, msz, bias = 63, 52, 11, 52, 1023
        else:
            raise TypeError(f"unsupported float size {dt.itemsize}")
        return (bytes([0x11, 0x20, sign, 0]) + struct.pack("<I", dt.itemsize)
                + struct.pack("<HHBBBBI", 0, 8 * dt.itemsize, eloc, esz, 0, msz, bias))
    if dt.kind == "S":
        return bytes([0x13, 0x01, 0, 0]) + struct.pack("<I", dt.itemsize)  # null-padded ASCII
    raise TypeError(f"unsupported dtype {dt}")


def _space_msg(shape) -> bytes:
    if shape == ():
        return bytes([2, 0, 0, 0])
    return bytes([2, len(shape), 0, 1]) + b"".join(struct.pack("<Q", int(d)) for d in shape)


def _attr_msg(name: str, value) -> bytes:
    """Attribute message (version 3) for a scalar or 1-D value."""
    if isinstance(value, str):
        value = value.encode()
    if isinstance(value, bytes):
        arr = np.array(value if value else b"\0", dtype=f"S{max(1, len(value))}")
    elif isinstance(value, (int, np.integer)):
        arr = np.array(value, dtype="<i8")
    elif isinstance(value, (float, np.floating)):
        arr = np.array(value, dtype="<f8")
    elif isinstance(value, np.ndarray):
        arr = value
    else:  # other Python objects: a protocol-0 pickle, as PyTables stores them
        blob = pickle.dumps(value, 0)
        arr = np.array(blob, dtype=f"S{len(blob)}")
    dtm, spm = _dtype_msg(arr.dtype), _space_msg(arr.shape)
    nm = name.encode() + b"\0"
    return (struct.pack("<BBHHHB", 3, 0, len(nm), len(dtm), len(spm), 0) + nm + dtm + spm
            + np.ascontiguousarray(arr).tobytes())


def _msg(mtype: int, data: bytes) -> bytes:
    if len(data) > 0xFFFF:
        raise ValueError("object header message above 64 KiB")
    return struct.pack("<BHB", mtype, len(data), 0) + data


def _ohdr(messages) -> bytes:
    body = b"".join(messages)
    head = b"OHDR" + bytes([2, 0x02]) + struct.pack("<I", len(body))
    blk = head + body
    return blk + struct.pack("<I", lookup3(blk))


def _group_ohdr(links, attrs) -> bytes:
    msgs = [_msg(0x02, bytes([0, 0]) + struct.pack("<QQ", UNDEF, UNDEF)),  # Link Info: compact storage
            _msg(0x0A, bytes([0, 0]))]                                      # Group Info: defaults
    for name, addr in links:
        nm = name.encode()
        if len(nm) > 255:
            raise ValueError("link name above 255 bytes")
        msgs.append(_msg(0x06, bytes([1, 0x00, len(nm)]) + nm + struct.pack("<Q", addr)))
    msgs += [_msg(0x0C, _attr_msg(k, v)) for k, v in attrs.items()]
    return _ohdr(msgs)


def _dataset_ohdr(dtype, nrows, addr, nbytes, attrs) -> bytes:
    msgs = [_msg(0x01, _space_msg((nrows,))),
            _msg(0x03, _dtype_msg(dtype)),
            _msg(0x05, bytes([3, 0x0A])),                                   # fill value: late, if set
            _msg(0x08, bytes([3, 1]) + struct.pack("<QQ", addr if nbytes else UNDEF, nbytes))]
    msgs += [_msg(0x0C, _attr_msg(k, v)) for k, v in attrs.items()]
    return _ohdr(msgs)


# ---------------------------------------------------------------------------
class H5Store:
    """Append-only writer of frames in the pandas table layout.

    ``append(key, columns)`` adds rows (a dict of equal-length 1-D arrays; str
    values become fixed-width byte strings, at least ``min_itemsize[col]``
    wide, as HDFStore.put(min_itemsize=...)); ``put_input(params)`` stores the
    run parameters as a one-row frame of strings; ``close()`` writes the file.
    """

    def __init__(self, path: str):
        self.path = path
        self.frames = {}   # key -> dict(dtype, nrows, spill path, spill file, columns)
        self._closed = False

    @staticmethod
    def _ascii(a):
        """A column of str as fixed-width bytes, converted in numpy (None when
        some string is not ASCII: the caller then encodes one by one)."""
        try:
            return np.asarray(a).astype("S")
        except (UnicodeEncodeError, ValueError, TypeError):
            return None

    def _dtype_for(self, key, cols, min_itemsize):
        fields = [("index", "<i8")]
        for name, v in cols.items():
            a = np.asarray(v)
            if a.dtype.kind in "USO":
                b = self._ascii(a) if a.dtype.kind != "S" else a
                if b is not None:
                    longest = int(np.strings.str_len(b).max()) if b.size else 0
                else:
                    longest = max([len(str(x).encode()) for x in a.tolist()] + [0])
                w = max(longest, 1, int((min_itemsize or {}).get(name, 0)))
                fields.append((name, f"S{w}"))
            elif a.dtype.kind == "b":
                fields.append((name, "<i1"))
            else:
                fields.append((name, a.dtype.newbyteorder("<").str))
        return np.dtype(fields)

    def append(self, key: str, columns: dict, min_itemsize=None):
        if self._closed:
            raise ValueError("store is closed")
        if not columns:
            return
        n = len(next(iter(columns.values())))
        f = self.frames.get(key)
        if f is None:
            dt = self._dtype_for(key, columns, min_itemsize)
            spill = f"{self.path}.{key}.part"
            f = dict(dtype=dt, nrows=0, spill=spill, fh=open(spill, "wb"), columns=list(columns))
            self.frames[key] = f
        elif list(columns) != f["columns"]:
            raise ValueError(f"{key}: columns {list(columns)} differ from the stored {f['columns']}")
        dt = f["dtype"]
        rec = np.zeros(n, dtype=dt)
        rec["index"] = np.arange(f["nrows"], f["nrows"] + n, dtype=np.int64)
        for name in f["columns"]:
            a = np.asarray(columns[name])
            if dt[name].kind == "S":
                b = self._ascii(a) if a.dtype.kind != "S" else a
                if b is not None:
                    width = int(np.strings.str_len(b).max()) if b.size else 0
                else:
                    b = np.array([str(x).encode() for x in a.tolist()], dtype=object)
                    width = max((len(x) for x in b), default=0)
                if width > dt[name].itemsize:
                    raise ValueError(f"{key}.{name}: string of {width} bytes above the column width "
                                     f"{dt[name].itemsize} (raise min_itemsize)")
                rec[name] = b.astype(dt[name])
            else:
                rec[name] = a
        rec.tofile(f["fh"])
        f["nrows"] += n

    def put_input(self, params: dict):
        """The run parameters as a one-row frame of string columns (the
        reference's final 'input' entry is pd.DataFrame(pd.Series(inputPar)).T,
        doScatter.py:66)."""
        self.append("input", {k: [str(v)] for k, v in params.items()})

    def close(self):
        if self._closed:
            return
        self._closed = True
        for f in self.frames.values():
            f["fh"].close()
        with open(self.path, "wb") as out:
            out.write(b"\0" * 48)
            data_addr = {}
            for key, f in self.frames.items():
                addr = out.tell()
                with open(f["spill"], "rb") as src:
                    while True:
                        buf = src.read(1 << 24)
                        if not buf:
                            break
                        out.write(buf)
                data_addr[key] = (addr, out.tell() - addr)
            links = []
            for key, f in self.frames.items():
                dt, cols = f["dtype"], f["columns"]
                tattrs = {"CLASS": "TABLE", "VERSION": "2.7", "TITLE": ""}
                for i, name in enumerate(dt.names):
                    tattrs[f"FIELD_{i}_NAME"] = name
                tattrs["NROWS"] = int(f["nrows"])
                tattrs["index_kind"] = "integer"
                # per data column, as pandas' DataCol.set_attr writes them: the
                # block's column list, the numpy dtype name (bytes<8n> for an
                # S<n> string column) and the metadata kind (none)
                for c in cols:
                    tattrs[f"{c}_kind"] = [c]
                    tattrs[f"{c}_meta"] = None
                    tattrs[f"{c}_dtype"] = dt.fields[c][0].name
                t_addr = out.tell()
                out.write(_dataset_ohdr(dt, f["nrows"], data_addr[key][0], data_addr[key][1], tattrs))
                gattrs = {"CLASS": "GROUP", "TITLE": "", "VERSION": "1.0",
                          "pandas_type": "frame_table", "pandas_version": "0.15.2",
                          "table_type": "appendable_frame", "encoding": "UTF-8", "errors": "strict",
                          "levels": 1, "nan_rep": "nan",
                          "index_cols": [(0, "index")], "values_cols": list(cols),
                          "non_index_axes": [(1, list(cols))], "data_columns": list(cols),
                          "info": {1: {"names": [None], "type": "Index"}, "index": {}}}
                g_addr = out.tell()
                out.write(_group_ohdr([("table", t_addr)], gattrs))
                links.append((key, g_addr))
            root = out.tell()
            out.write(_group_ohdr(links, {"CLASS": "GROUP", "TITLE": "", "VERSION": "1.0",
                                          "PYTABLES_FORMAT_VERSION": "2.1"}))
            eof = out.tell()
            sb = SIGNATURE + bytes([2, 8, 8, 0]) + struct.pack("<QQQQ", 0, UNDEF, eof, root)
            out.seek(0)
            out.write(sb + struct.pack("<I", lookup3(sb)))
        for f in self.frames.values():
            os.remove(f["spill"])

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# Reader of the structures written above.
class _PlainUnpickler(pickle.Unpickler):
    """Unpickles lists, tuples, dicts, strings and numbers only: a pickle
    that names any class or function is refused (a file is data, not code)."""

    def find_class(self, module, name):
        raise pickle.UnpicklingError(f"refusing {module}.{name} in an attribute pickle")


def _plain_loads(blob: bytes):
    import io
    return _PlainUnpickler(io.BytesIO(blob)).load()


class _Reader:
    def __init__(self, path):
        with open(path, "rb") as f:
            self.buf = f.read()
        b = self.buf
        if b[:8] != SIGNATURE or b[8] not in (2, 3):
            raise ValueError("not an HDF5 file with a version 2/3 superblock")
        if b[9] != 8 or b[10] != 8:
            raise ValueError("only 8-byte offsets and lengths are supported")
        if lookup3(b[:44]) != struct.unpack_from("<I", b, 44)[0]:
            raise ValueError("superblock checksum mismatch")
        self.root = struct.unpack_from("<Q", b, 36)[0]

    def messages(self, addr):
        b = self.buf
        if b[addr:addr + 4] != b"OHDR" or b[addr + 4] != 2:
            raise ValueError(f"no version 2 object header at {addr}")
        flags = b[addr + 5]
        p = addr + 6
        if flags & 0x20:
            p += 16
        if flags & 0x10:
            p += 4
        w = 1 << (flags & 3)
        size = int.from_bytes(b[p:p + w], "little")
        p += w
        end = p + size
        if lookup3(b[addr:end]) != struct.unpack_from("<I", b, end)[0]:
            raise ValueError(f"object header checksum mismatch at {addr}")
        out = []
        while p + 4 <= end:
            t, n, mf = struct.unpack_from("<BHB", b, p)
            p += 4
            if flags & 0x04:
                p += 2
            out.append((t, b[p:p + n]))
            p += n
        return out

    @staticmethod
    def dtype(m, off=0):
        """numpy dtype of a datatype message at m[off:], and its encoded length."""
        cls, ver = m[off] & 0x0F, m[off] >> 4
        bits = m[off + 1:off + 4]
        size = struct.unpack_from("<I", m, off + 4)[0]
        p = off + 8
        if cls == 0:
            if bits[0] & 1:
                raise ValueError("big-endian integers are not supported")
            return np.dtype(("<i" if bits[0] & 0x08 else "<u") + str(size)), p + 4 - off
        if cls == 1:
            return np.dtype("<f" + str(size)), p + 12 - off
        if cls == 3:
            return np.dtype(f"S{size}"), p - off
        if cls == 6 and ver == 3:
            n = bits[0] | (bits[1] << 8)
            width = 1 if size < 256 else 2 if size < 65536 else 3 if size < (1 << 24) else 4
            names, formats, offsets = [], [], []
            for _ in range(n):
                z = m.index(b"\0", p)
                names.append(m[p:z].decode())
                p = z + 1
                offsets.append(int.from_bytes(m[p:p + width], "little"))
                p += width
                sub, ln = _Reader.dtype(m, p)
                formats.append(sub)
                p += ln
            return np.dtype(dict(names=names, formats=formats, offsets=offsets, itemsize=size)), p - off
        raise ValueError(f"unsupported datatype class {cls} version {ver}")

    @staticmethod
    def space(m, off=0):
        ver, rank, flags = m[off], m[off + 1], m[off + 2]
        if ver != 2:
            raise ValueError("only version 2 dataspaces are supported")
        p = off + 4
        dims = tuple(struct.unpack_from("<Q", m, p + 8 * i)[0] for i in range(rank))
        p += 8 * rank * (2 if flags & 1 else 1)
        return dims, p - off

    def attrs(self, msgs):
        out = {}
        for t, m in msgs:
            if t != 0x0C:
                continue
            ver, _, nl, dl, sl, _ = struct.unpack_from("<BBHHHB", m, 0)
            p = 9
            name = m[p:p + nl - 1].decode()
            p += nl
            dt, _ = self.dtype(m, p)
            p += dl
            dims, _ = self.space(m, p)
            p += sl
            cnt = int(np.prod(dims)) if dims else 1
            val = np.frombuffer(m, dtype=dt, count=cnt, offset=p)
            val = val.reshape(dims) if dims else val[0]
            if isinstance(val, bytes):
                s = val.rstrip(b"\0")
                if s == b"N.":  # a protocol-0 pickle of None
                    val = None
                elif s.startswith(b"(") and s.endswith(b"."):
                    try:
                        val = _plain_loads(s)
                    except Exception:
                        val = s.decode()
                else:
                    val = s.decode()
            elif isinstance(val, np.generic):
                val = val.item()
            out[name] = val
        return out

    def links(self, msgs):
        out = {}
        for t, m in msgs:
            if t == 0x06:
                flags = m[1]
                p = 2
                if flags & 0x08:
                    p += 1
                if flags & 0x04:
                    p += 8
                if flags & 0x10:
                    p += 1
                w = 1 << (flags & 3)
                n = int.from_bytes(m[p:p + w], "little")
                p += w
                name = m[p:p + n].decode()
                out[name] = struct.unpack_from("<Q", m, p + n)[0]
        return out

    def dataset(self, addr):
        msgs = self.messages(addr)
        dims = dt = lay = None
        for t, m in msgs:
            if t == 0x01:
                dims, _ = self.space(m)
            elif t == 0x03:
                dt, _ = self.dtype(m)
            elif t == 0x08:
                if m[0] != 3 or m[1] != 1:
                    raise ValueError("only contiguous layouts are supported")
                lay = struct.unpack_from("<QQ", m, 2)
        n = int(np.prod(dims))
        if n == 0:
            return np.zeros(dims, dtype=dt), self.attrs(msgs)
        return np.frombuffer(self.buf, dtype=dt, count=n, offset=lay[0]).reshape(dims).copy(), self.attrs(msgs)


def keys(path):
    """Names of the frames in a store written by H5Store."""
    r = _Reader(path)
    return sorted(r.links(r.messages(r.root)))


def read_table(path, key):
    """(structured array of /key/table, table attributes, group attributes)."""
    r = _Reader(path)
    groups = r.links(r.messages(r.root))
    if key not in groups:
        raise KeyError(key)
    gm = r.messages(groups[key])
    arr, tattrs = r.dataset(r.links(gm)["table"])
    return arr, tattrs, r.attrs(gm)


def _strings(a):
    """A fixed-width bytes column as str: numpy's conversion for ASCII, UTF-8
    (the store's encoding attribute) one by one otherwise."""
    try:
        return a.astype(str)
    except UnicodeDecodeError:
        return np.array([x.decode("utf-8") for x in a.tolist()], dtype=object)


def read_frame(path, key):
    """The frame stored under key as a pandas DataFrame (byte-string columns
    decoded to str), indexed by the stored index column."""
    import pandas as pd

    arr, _, gattrs = read_table(path, key)
    cols = {}
    for name in arr.dtype.names[1:]:
        a = arr[name]
        cols[name] = _strings(a) if a.dtype.kind == "S" else a
    return pd.DataFrame(cols, index=pd.Index(arr["index"], name=None))


__all__ = ["H5Store", "read_frame", "read_table", "keys", "lookup3"]
