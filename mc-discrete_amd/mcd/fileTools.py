"""Input parsing and output-column assembly (MC/tools/fileTools.py).

* ``readInput`` follows fileTools.readInput (:8-46): one ``key : value`` per
  line, ``#`` comments, typed by key.  ``seed`` (the Philox key, an addition of
  this implementation) is read as an int; ``maxScatt : off`` means no limit.
* ``els_columns`` / ``scats_columns`` / ``per_electron_dicts`` produce the column
  layout of thingsToSave (:193-242) after zipDict's flattening (:67-99):
  ``final_dir`` -> dx, dy, dz; ``last_pos`` -> last_x, last_y, last_z;
  ``position`` -> x, y, z.
"""
from __future__ import annotations

import numpy as np

from . import _abi

STRING_KEYS = ("mode", "material", "elastic", "Bethe_model")
INT_KEYS = ("num_el", "maxScatt", "seed")
BOOL_KEYS = ("gen_tables", "diffMFP")

ELECTRON_OUTPUTS = ("outcome", "final_E", "final_dir", "last_pos")
SCATTER_OUTPUTS = ("type", "E", "E_loss", "pol_angle", "az_angle", "pathl", "position", "energy", "direction")


def readInput(fileName="input.file"):
    """Parse an MC-discrete input file into a parameter dict."""
    with open(fileName) as f:
        raw = f.read().split("\n")
    data = {}
    for line in raw:
        if not line.strip() or line.startswith("#"):
            continue
        line = line.partition("#")[0]
        parts = [p.strip(" ") for p in line.split(":")]
        key = parts[0].strip()
        val = parts[1].strip() if len(parts) > 1 else ""
        if key in STRING_KEYS:
            data[key] = val
        elif key in INT_KEYS:
            data[key] = 0 if (key == "maxScatt" and val.lower() in ("off", "none", "")) else int(val)
        elif "output" in key:
            data[key] = val.replace(" ", "").split(",")
        elif key in BOOL_KEYS:
            data[key] = "es" in val
        else:
            data[key] = float(val)
    return data


def _outcome_names(codes):
    names = np.array(["", "abs", "bks", "trsm", "trdf", "long"], dtype=object)
    return names[codes]


def channel_names(material):
    """Channel label -> scats 'type' string (Ruth, Moller, Gryz<shell>, Quinn)."""
    names = {_abi.CH_RUTH: "Ruth", _abi.CH_MOLLER: "Moller", _abi.CH_QUINN: "Quinn"}
    for i, s in enumerate(material.params["name_s"]):
        names[_abi.CH_GRYZ0 + i] = "Gryz" + s
    return names


def els_rows(els: np.ndarray) -> np.ndarray:
    """Expand electron records into `els` rows: an electron that exits and is
    absorbed in the same step has a second row with outcome 'abs'
    (electron.py:44-59,64-84)."""
    two = els["outcome2"] != 0
    idx = np.repeat(np.arange(els.size), 1 + two.astype(np.int64))
    rows = els[idx].copy()
    second = np.zeros(idx.size, dtype=bool)
    second[1:] = idx[1:] == idx[:-1]
    rows["outcome"][second] = rows["outcome2"][second]
    return rows, idx


def els_columns(els: np.ndarray, wanted=ELECTRON_OUTPUTS) -> dict:
    """Columnar `els` table in the reference's column order."""
    rows, idx = els_rows(els)
    out = {}
    for w in wanted:
        if w == "outcome":
            out["outcome"] = _outcome_names(rows["outcome"])
        elif w == "final_E":
            out["final_E"] = rows["final_E"].astype(np.float64)
    for w in wanted:
        if w == "final_dir":
            out.update(dx=rows["dir"][:, 0], dy=rows["dir"][:, 1], dz=rows["dir"][:, 2])
        elif w == "last_pos":
            out.update(last_x=rows["pos"][:, 0], last_y=rows["pos"][:, 1], last_z=rows["pos"][:, 2])
    return out


def scats_columns(rows: np.ndarray, material, wanted=("type", "E", "energy")) -> dict:
    """Columnar `scats` table (rows sorted by electron, step)."""
    names = channel_names(material)
    out = {}
    for w in wanted:
        if w == "type":
            lut = np.array([names.get(i, "") for i in range(_abi.MAX_CHANNELS + 3)], dtype=object)
            out["type"] = lut[rows["channel"]]
        elif w == "E":
            out["E"] = rows["E"].astype(np.float64)
        elif w == "E_loss":
            out["E_loss"] = rows["E_loss"].astype(np.float64)
        elif w == "pol_angle":
            out["pol_angle"] = rows["pol"].astype(np.float64)
        elif w == "az_angle":
            out["az_angle"] = rows["az"].astype(np.float64)
        elif w == "pathl":
            out["pathl"] = rows["pathl"].astype(np.float64)
        elif w == "energy":
            out["energy"] = rows["energy"].astype(np.float64)
        elif w == "direction":
            out.update(dir_x=rows["dx"], dir_y=rows["dy"], dir_z=rows["dz"])
    if "position" in wanted:
        out.update(x=rows["x"], y=rows["y"], z=rows["z"])
    return out


def per_electron_dicts(els: np.ndarray, rows, material, things, first=0):
    """The list of {'els': {...}, 'scats': {...}} dicts multiTraj_DS returns
    (multiScatter.py:105-116), one per electron, with nested final_dir /
    last_pos / position dicts as thingsToSave builds them.  `rows` are the
    scatter rows of electrons first.. sorted by (electron, step)."""
    el_want = list(things.get("el_output", ()))
    sc_want = list(things.get("scat_output", ()))
    names = channel_names(material)
    if rows is not None and sc_want:
        ids = _abi.row_offsets(rows, first)
        lo = np.searchsorted(ids, np.arange(els.size), side="left")
        hi = np.searchsorted(ids, np.arange(els.size), side="right")
    out = []
    for i, e in enumerate(els):
        ed = {}
        n_rows = 2 if e["outcome2"] else 1
        outs = [e["outcome"], e["outcome2"]][:n_rows]
        for w in el_want:
            if w == "outcome":
                ed["outcome"] = [_abi.OUTCOME_NAMES[o] for o in outs]
            elif w == "final_E":
                ed["final_E"] = [float(e["final_E"])] * n_rows
            elif w == "final_dir":
                ed["final_dir"] = {k: [float(e["dir"][j])] * n_rows for j, k in enumerate(("dx", "dy", "dz"))}
            elif w == "last_pos":
                ed["last_pos"] = {k: [float(e["pos"][j])] * n_rows
                                  for j, k in enumerate(("last_x", "last_y", "last_z"))}
            else:
                ed[w] = []
        sd = {}
        if rows is not None and sc_want:
            sel = rows[lo[i]:hi[i]]
            for w in sc_want:
                if w == "type":
                    sd["type"] = [names[c] for c in sel["channel"]]
                elif w == "position":
                    sd["position"] = {"x": sel["x"].tolist(), "y": sel["y"].tolist(), "z": sel["z"].tolist()}
                elif w == "pol_angle":
                    sd[w] = sel["pol"].tolist()
                elif w == "az_angle":
                    sd[w] = sel["az"].tolist()
                elif w in ("E", "E_loss", "pathl", "energy"):
                    sd[w] = sel[w].tolist()
                elif w == "direction":
                    sd[w] = [np.array([a, b, c]) for a, b, c in zip(sel["dx"], sel["dy"], sel["dz"])]
        out.append({"els": ed, "scats": sd})
    return out


def zipDict(dictA, dictB):
    """Merge per-electron column dicts, flattening final_dir / position /
    last_pos (fileTools.zipDict, :67-99)."""
    for nested in ("final_dir", "position", "last_pos"):
        if nested in dictB:
            for k, v in dictB[nested].items():
                dictB[k] = v
            del dictB[nested]
    for k, v in dictB.items():
        if k in dictA:
            dictA[k] += v
        else:
            dictA[k] = v
    return dictA
