"""Constants of the hand-written FP32 device arithmetic, checked on the CPU
against the mathematics they implement (no GPU needed): the literals are read
from mc-discrete_amd/csrc/mcd_device.h, so a mistyped coefficient fails here.

* u01f: (1 + m 2^-23) - (1 - 2^-24) must equal (2m + 1) 2^-24 exactly for all
  2^23 values of m = x >> 9.
* cos_abs_sin_2pi: cos(2 pi u) and |sin(2 pi u)| from the quadrant reduction
  and the two fitted polynomials, evaluated with single-rounding FMAs in
  float32, within 2 ulp of the FP64 values over the generator's u grid.
"""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEVICE_H = os.path.join(ROOT, "mc-discrete_amd", "csrc", "mcd_device.h")


def _source():
    with open(DEVICE_H) as f:
        return f.read()


def _function(src, name):
    i = src.index(name)
    return src[i:src.index("\n}\n", i)]


def _floats(text):
    return [np.float32(float(x)) for x in re.findall(r"(-?\d+\.\d+)f", text)]


def test_u01_mapping_constant_is_exact_for_every_word():
    body = _function(_source(), "float u01f(uint32_t x)")
    assert "__builtin_amdgcn_alignbit(0x7Fu, x, 9u)" in body
    (c,) = _floats(body)
    assert c == np.float32(1.0 - 2.0 ** -24)
    m = np.arange(1 << 23, dtype=np.uint32)
    # {0x7F, x} >> 9 keeps the low 23 bits of the shifted word and sets the exponent of 1.0
    f = (np.uint32(0x3F800000) | m).view(np.float32)
    u = f - c
    want = ((2 * m.astype(np.float64) + 1) * 2.0 ** -24)
    assert np.array_equal(u.astype(np.float64), want)
    assert u.min() > 0 and u.max() < 1


def _fma32(a, b, c):
    # single rounding: the product of two float32 values is exact in float64
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def test_folded_azimuth_polynomials_within_two_ulp():
    body = _function(_source(), "void cos_abs_sin_2pi(float u, float& cp, float& sp)")
    consts = _floats(body)
    # the literals in source order: -0.25 (quadrant step), then the sine and
    # cosine coefficients, highest power first
    assert consts[0] == np.float32(-0.25)
    coeffs = consts[1:]
    assert len(coeffs) == 8, coeffs
    s3, s2, s1, s0, c3, c2, c1, c0 = coeffs
    m = np.arange(0, 1 << 23, 7, dtype=np.int64)
    u = ((2 * m + 1) * 2.0 ** -24).astype(np.float32)
    a = np.abs(u - np.rint(u)).astype(np.float32)
    q = np.rint(np.float32(4) * a).astype(np.float32)
    z = _fma32(np.full_like(q, -0.25), q, a)
    w = (z * z).astype(np.float32)
    ps = _fma32(w, np.full_like(w, s3), np.full_like(w, s2))
    ps = _fma32(w, ps, np.full_like(w, s1))
    ps = _fma32(w, ps, np.full_like(w, s0))
    sn = (z * ps).astype(np.float32)
    pc = _fma32(w, np.full_like(w, c3), np.full_like(w, c2))
    pc = _fma32(w, pc, np.full_like(w, c1))
    pc = _fma32(w, pc, np.full_like(w, c0))
    cs = _fma32(w, pc, np.ones_like(w))
    cp = np.where(q == 0, cs, np.where(q == 1, -sn, -cs))
    sp = np.where(q == 0, sn, np.where(q == 1, cs, -sn))
    ref_c = np.cos(2 * np.pi * u.astype(np.float64))
    ref_s = np.abs(np.sin(2 * np.pi * u.astype(np.float64)))
    ulp_c = np.spacing(np.abs(ref_c).astype(np.float32)).astype(np.float64)
    ulp_s = np.spacing(np.abs(ref_s).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(cp - ref_c) / ulp_c) <= 2.0
    assert np.max(np.abs(sp - ref_s) / ulp_s) <= 2.0
    assert (sp >= 0).all()
