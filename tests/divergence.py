"""First step at which a kernel's per-step rows leave the oracle trace, and
what decided it (test helper: the per-step parity test and its CPU check)."""
import numpy as np

import oracle as O
from mcd import _abi


def _row_keys(electron, step):
    return electron.astype(np.uint64) << np.uint64(32) | step.astype(np.uint64)


def _pick_margin(cfg, E, u):
    """Distance of u from the nearest boundary of the FP64 sorted cumulative
    channel probabilities at energy E (pickFromSigmas, scattering.py:39-59)."""
    sig, _ = O.sigma_set(cfg, float(E))
    order = sorted(range(len(sig)), key=lambda i: -sig[i])  # stable: ties keep insertion order
    tot = 0.0
    for i in order:
        tot += float(sig[i])
    acc, margin = 0.0, 1.0
    for i in order[:-1]:
        acc += float(sig[i])
        margin = min(margin, abs(acc / tot - u))
    return margin


def classify_divergence(cfg, tables, g_rows, o_rows):
    """First step at which the FP32 kernel's rows leave the oracle trace, for
    every electron whose rows differ, classified by what decided it:
    'pick_tie' / 'pick_miss' (channel differs; u within / beyond 1e-9 of a
    boundary of the FP64 pick at the oracle's energy), 'table_tie' /
    'table_miss' (same channel, different loss: E on a grid energy or u on a
    CDF value / neither), 'exit' (same channel and loss, a different outcome:
    an exit or absorption decided by an FP32 position or direction), 'length'
    (one trajectory ends earlier).  Returns a list of (electron, step, class)."""
    kg = _row_keys(g_rows["electron"], g_rows["step"])
    ko = _row_keys(o_rows["electron"], o_rows["step"])
    common, ig, io = np.intersect1d(kg, ko, assume_unique=True, return_indices=True)
    g, o = g_rows[ig], o_rows[io]
    diff = ((g["channel"] != o["channel"]) | (g["E_loss"] != o["E_loss"].astype(np.float32)) |
            (g["outcome"] != o["outcome"].astype(np.uint8)))
    first = {}
    for k in common[diff]:
        e, st = int(k >> np.uint64(32)), int(k & np.uint64(0xFFFFFFFF))
        first[e] = min(first.get(e, st), st)
    only = np.setxor1d(kg, ko, assume_unique=True)
    for k in only:
        e, st = int(k >> np.uint64(32)), int(k & np.uint64(0xFFFFFFFF))
        if st < first.get(e, 1 << 40):
            first[e] = st
    out = []
    assert np.all(np.diff(kg.astype(np.int64)) > 0) and np.all(np.diff(ko.astype(np.int64)) > 0)

    def find(keys, key):  # index of key in sorted keys, or -1
        i = int(np.searchsorted(keys, np.uint64(key)))
        return i if i < len(keys) and int(keys[i]) == key else -1

    for e, st in sorted(first.items()):
        key = (e << 32) | st
        jg, jo = find(kg, key), find(ko, key)
        if jg < 0 or jo < 0:
            out.append((e, st, "length"))
            continue
        a, b = g_rows[jg], o_rows[jo]
        if a["channel"] != b["channel"]:
            m = _pick_margin(cfg, b["E"], b["u"][0])
            out.append((e, st, "pick_tie" if m <= 1e-9 else "pick_miss"))
        elif a["E_loss"] != np.float32(b["E_loss"]):
            t = tables[0] if b["channel"] == _abi.CH_MOLLER else tables[1]
            row = t.row(b["E"])
            on_grid = float(t.Es[min(row, len(t.Es) - 1)]) == b["E"]
            col = t.col(row, b["u"][2])
            on_cdf = col < t.cdf.shape[1] and abs(float(t.cdf[row, col]) - b["u"][2]) <= 1e-9
            out.append((e, st, "table_tie" if (on_grid or on_cdf) else "table_miss"))
        else:
            out.append((e, st, "exit"))
    return out
