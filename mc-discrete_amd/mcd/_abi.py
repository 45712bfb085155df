"""ctypes mirror of include/mcd.h (structs, enums and prototypes)."""
from __future__ import annotations

import ctypes as C

import numpy as np

MAX_SHELLS = 8
MAX_CHANNELS = MAX_SHELLS + 3

OK, ERR_INVALID, ERR_DEVICE, ERR_TABLE, ERR_OVERFLOW, ERR_NOMEM, ERR_UNSUPPORTED = range(7)
MODE_DS, MODE_CONT = 0, 1
BETHE_CLASSICAL, BETHE_JL, BETHE_EXPLICIT = 0, 1, 2
OUT_NONE, OUT_ABS, OUT_BKS, OUT_TRSM, OUT_TRDF, OUT_LONG = range(6)
OUTCOME_NAMES = {OUT_ABS: "abs", OUT_BKS: "bks", OUT_TRSM: "trsm", OUT_TRDF: "trdf", OUT_LONG: "long"}
CH_RUTH, CH_MOLLER, CH_QUINN, CH_GRYZ0 = 0, 1, 2, 3
FLAG_NAN_DIR, FLAG_LOSS_GE_E, FLAG_LONG_PATH, FLAG_FP64_PICK, FLAG_GUARD_TIMEOUT = 1, 2, 4, 8, 16
TABLE_MOLLER, TABLE_GRYZ = 0, 1
PREC_FP64, PREC_FP32 = 0, 1

(CNT_EVENTS, CNT_ELECTRONS, CNT_BKS, CNT_ABS, CNT_TRSM, CNT_LONG, CNT_FP64_PICKS, CNT_SCAT_ROWS, CNT_QUEUE,
 CNT_GUARD_TIMEOUTS, CNT_PARKED, CNT_RESUMED) = range(12)
NCOUNTERS = 16
GRAD_WC, GRAD_TILT, GRAD_J, GRAD_LNRHO = range(4)
NGRAD = 4
GRAD_NAMES = ("Wc", "s_tilt", "J", "density_scale")
MAX_BINS = 256


class Material(C.Structure):
    _fields_ = [
        ("Z", C.c_int32), ("n_shells", C.c_int32),
        ("atwt", C.c_double), ("density", C.c_double), ("n_val", C.c_double), ("E_val", C.c_double),
        ("bethe_k", C.c_double),
        ("shell_Eb", C.c_double * MAX_SHELLS), ("shell_n", C.c_double * MAX_SHELLS),
        ("atnd", C.c_double), ("fermi_e", C.c_double), ("plasmon_e", C.c_double),
    ]


class Config(C.Structure):
    _fields_ = [
        ("mat", Material),
        ("E0", C.c_double), ("Emin", C.c_double), ("Wc", C.c_double), ("s_tilt", C.c_double),
        ("thickness", C.c_double),
        ("mode", C.c_int32), ("bethe_model", C.c_int32),
        ("max_scatt", C.c_int64),
        ("J", C.c_double), ("explicit_lnEstar", C.c_double), ("explicit_Zeff", C.c_double),
        ("elastic", C.c_int32),
    ]


class TableGridC(C.Structure):
    _fields_ = [
        ("nE", C.c_int32), ("nW", C.c_int32),
        ("Es", C.POINTER(C.c_float)), ("Ws", C.POINTER(C.c_float)),
        ("Wmin", C.c_double), ("n_el", C.c_double), ("Eb", C.c_double), ("kind", C.c_int32),
    ]


class GridSpec(C.Structure):
    """mcd_grid_spec: one table's grid-search parameters."""
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("Wmin", C.c_double), ("n_el", C.c_double),
                ("Eb", C.c_double), ("Ef", C.c_double), ("Emin", C.c_double), ("Emax", C.c_double)]


class ScatterBuf(C.Structure):
    _fields_ = [("rows", C.c_void_p), ("capacity", C.c_uint64), ("max_electron", C.c_uint64)]


class BatchItem(C.Structure):
    _fields_ = [("cfg", C.POINTER(Config)), ("tables", C.c_void_p), ("n", C.c_uint64), ("els", C.c_void_p)]


class Launch(C.Structure):
    _fields_ = [("block", C.c_int32), ("waves_per_simd", C.c_int32), ("precision", C.c_int32),
                ("flags", C.c_int32)]


NOALLOC_MAX_ELECTRONS = 1 << 20

LAUNCH_INLINE_GUARD = 1
LAUNCH_DEFER_GUARD = 2
LAUNCH_SERVICE_GUARD = 4
LAUNCH_REFILL_SHIFT = 8
LAUNCH_REFILL_MASK = 7 << LAUNCH_REFILL_SHIFT


class BwdOpts(C.Structure):
    _fields_ = [("n_bins", C.c_int32), ("rel_step", C.c_float), ("graze", C.c_float), ("block", C.c_int32)]


# mcd_electron_rec (36 bytes)
ELECTRON_DTYPE = np.dtype([
    ("final_E", "<f4"), ("dir", "<f4", (3,)), ("pos", "<f4", (3,)), ("n_scatt", "<u4"),
    ("outcome", "u1"), ("outcome2", "u1"), ("flags", "u1"), ("pad", "u1"),
])
assert ELECTRON_DTYPE.itemsize == 36

# mcd_scatter_rec (64 bytes)
SCATTER_DTYPE = np.dtype([
    ("electron", "<u4"), ("step", "<u4"), ("config", "<u2"), ("channel", "u1"), ("outcome", "u1"),
    ("E", "<f4"), ("E_loss", "<f4"), ("pol", "<f4"), ("az", "<f4"), ("pathl", "<f4"),
    ("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("energy", "<f4"),
    ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("step_flags", "<u4"),
])
assert SCATTER_DTYPE.itemsize == 64
# mcd_scatter_pack field bits (MCD_SCAT_*): the record's fields in order
SCAT_FIELD_BITS = {name: 1 << i for i, name in enumerate(SCATTER_DTYPE.names)}
# scatter_output names (fileTools.thingsToSave) -> the row fields they read
SCAT_OUTPUT_FIELDS = {"type": ("channel",), "E": ("E",), "E_loss": ("E_loss",), "pol_angle": ("pol",),
                      "az_angle": ("az",), "pathl": ("pathl",), "energy": ("energy",),
                      "position": ("x", "y", "z"), "direction": ("dx", "dy", "dz")}


def scat_fields_for(names):
    """Row fields a scatter_output list needs, plus electron and step (the
    rows are ordered by them on the host)."""
    out = {"electron", "step"}
    for n in names:
        out.update(SCAT_OUTPUT_FIELDS.get(n, ()))
    return [f for f in SCATTER_DTYPE.names if f in out]


def row_offsets(rows, first):
    """Electron index of each scatter row within a launch that began at
    electron `first` (int64).  A row holds the low 32 bits of the electron id,
    so the offset is their difference modulo 2^32 (a launch spans fewer than
    2^32 electrons)."""
    return (rows["electron"] - np.uint32(first & 0xFFFFFFFF)).astype(np.int64)


ELASTIC_RUTH, ELASTIC_RUTH_WDEFL, ELASTIC_NIGRAM, ELASTIC_NIGRAM_WDEFL = range(4)


def elastic_model(elastic, mode) -> int:
    """The elastic cross section the reference uses for the input key
    'elastic' in each mode, with its string tests as written:
    DS (scattering.py:483-493): == 'Ruth_vanilla', then 'vanilla_wDefl' in,
    then == 'Ruth_nigram', then 'nigram_wDefl' in;
    continuous (scattering.py:780-789): 'vanilla' in, then 'nigram' in -- so
    in continuous mode Ruth_vanilla_wDefl gets ruther_sigma and
    Ruth_nigram_wDefl gets ruther_N_sigma, their wDefl branches unreachable."""
    if isinstance(elastic, int):
        return elastic
    cont = mode in ("cont", "continuous", MODE_CONT)
    if "Ruth" in elastic:
        if cont:
            if "vanilla" in elastic:
                return ELASTIC_RUTH
            if "nigram" in elastic:
                return ELASTIC_NIGRAM
        else:
            if elastic == "Ruth_vanilla":
                return ELASTIC_RUTH
            if "vanilla_wDefl" in elastic:
                return ELASTIC_RUTH_WDEFL
            if elastic == "Ruth_nigram":
                return ELASTIC_NIGRAM
            if "nigram_wDefl" in elastic:
                return ELASTIC_NIGRAM_WDEFL
        raise ValueError(f"elastic model {elastic!r}: the reference sets no elastic cross section for it in "
                         f"{'continuous' if cont else 'DS'} mode")
    raise NotImplementedError(f"elastic model {elastic!r}: only the Rutherford models are implemented "
                              "(the Mott tables are not part of the reference)")


def make_config(mat, E0, Emin, Wc=50.0, s_tilt=0.0, thickness=float("inf"), mode="DS",
                bethe_model="classical", max_scatt=0, J=None, elastic="Ruth_vanilla") -> Config:
    """Fill an mcd_config from a mcd.material.Material and run parameters."""
    from .material import bethe_J

    c = Config()
    m = c.mat
    p = mat.params
    m.Z = int(p["Z"])
    shells = mat.shells
    if len(shells) > MAX_SHELLS:
        raise ValueError("too many shells")
    m.n_shells = len(shells)
    m.atwt = float(p["atwt"])
    m.density = float(p["density"]) * mat.density_scale
    m.n_val = float(p["n_val"])
    m.E_val = float(p["E_val"])
    m.bethe_k = float(p["bethe_k"])
    for i, (_, Eb, ns) in enumerate(shells):
        m.shell_Eb[i] = Eb
        m.shell_n[i] = ns
    m.atnd = mat.atnd
    m.fermi_e = mat.fermi_e
    m.plasmon_e = mat.plasmon_e
    c.E0, c.Emin, c.Wc, c.s_tilt, c.thickness = float(E0), float(Emin), float(Wc), float(s_tilt), float(thickness)
    modes = {"DS": MODE_DS, "cont": MODE_CONT, "continuous": MODE_CONT}
    if isinstance(mode, str):
        if mode not in modes:
            raise ValueError(f"unknown mode {mode!r}")
        mode = modes[mode]
    c.mode = int(mode)
    c.elastic = elastic_model(elastic, c.mode)
    models = {"classical": BETHE_CLASSICAL, "JL": BETHE_JL, "explicit": BETHE_EXPLICIT}
    if isinstance(bethe_model, str):
        if bethe_model not in models:
            raise ValueError(f"unknown Bethe_model {bethe_model!r}")
        bethe_model = models[bethe_model]
    c.bethe_model = int(bethe_model)
    c.max_scatt = int(max_scatt) if max_scatt else 0
    c.J = float(J) if J is not None else bethe_J(m.Z)
    zeff, lnE = mat.explicit_bethe_constants()
    c.explicit_Zeff, c.explicit_lnEstar = zeff, lnE
    return c


def grid_struct(grid, kind):
    """Build an mcd_table_grid pointing into the (kept-alive) float32 arrays of a TableGrid."""
    Es = np.ascontiguousarray(grid.Es, dtype=np.float32)
    Ws = np.ascontiguousarray(grid.Ws, dtype=np.float32)
    g = TableGridC()
    g.nE, g.nW = Es.size, Ws.size
    g.Es = Es.ctypes.data_as(C.POINTER(C.c_float))
    g.Ws = Ws.ctypes.data_as(C.POINTER(C.c_float))
    g.Wmin, g.n_el, g.Eb, g.kind = float(grid.Wmin), float(grid.n_el), float(grid.Eb), int(kind)
    g._keep = (Es, Ws)
    return g
