"""Independent pins of the oracle's Joy-Luo and explicit Bethe stopping powers,
of its Quinn plasmon cross section, of the Nigram and wDefl elastic models and
of the Gryzinski differential cross section of the tables.

The reference's JL and explicit laws (MC/tools/stoppingPowers.py:145-173,
177-215) are never reached by its own code and none of its tests evaluates
them, and its Quinn cross section (MC/tools/crossSections.py:247-284) has no
numeric test either.  Each law is therefore tied here to something that is not
its own formula:

* JL: with k = 0 it is the classical law, which the textbook engineering form
  pins (test_oracle_rng_physics); for E >> kJ its excess over the classical law
  is the first-order term 2 pi e^4 n Z / E * kJ / E; and it equals Joy and
  Luo's engineering form 78500 rho Z / (A E) ln(1.166 (E + kJ) / J) keV/cm
  (Scanning 11, 176 (1989)).
* explicit: with one valence term (n = Z, E_val) and no inner shells it is the
  classical law with J = 1.166 E_val; splitting a shell in two halves at the
  same binding energy leaves it unchanged; each shell adds
  2 pi e^4 n n_i / E ln(E / E_i).
* Quinn: the inverse mean free path E_pl / (2 a0 E) ln(q_c / q_min) rebuilt
  from electron wave vectors (q_min = k - k' for a loss E_pl, q_c =
  k_F (sqrt(1 + E_pl / E_F) - 1)) with scipy's CODATA constants; its
  E >> E_F asymptote with q_min -> E_pl / (2 sqrt(E E_F)) k_F; zero at and
  below the plasmon threshold.
* Gryzinski's differential form (extFunctions.py:189-219): its integral over
  [E_b, E] against closed forms at both ends of the energy range, the
  threshold limit n pi e^4 2^-3/2 (8/9) ln(2.7) (U - 1)^3/2 / E_b^2 and the
  high-energy limit (Gryzinski's total, crossSections.py:213-242); together
  they fix every factor of the form.

Every pin is written as a predicate and also applied to deliberately wrong
variants of the law (a dropped factor, a swapped argument, a wrong sign) to
show that the pin rejects them.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy.constants import e as E_CHARGE, hbar, m_e
from scipy.integrate import quad

import oracle as O
from mcd import physics
from mcd.material import Material, bethe_J

PI_EFOUR = physics.PI_EFOUR
BOHR_R = 5.29177210903e-11  # m, parameters.py:8 (scipy CODATA 2018)


def S_cl(Z, E, n, J):
    return O.lib().oracle_bethe_cl(int(Z), float(E), float(n), float(J))


def S_jl(Z, E, n, k, J):
    return O.lib().oracle_bethe_jl(int(Z), float(E), float(n), float(k), float(J))


def S_expl(E, n, shells, Zval, Eval):
    Zi = (ctypes.c_double * max(1, len(shells)))(*[s[0] for s in shells])
    Ei = (ctypes.c_double * max(1, len(shells)))(*[s[1] for s in shells])
    return O.lib().oracle_bethe_explicit(float(E), float(n), Zi, Ei, len(shells), float(Zval), float(Eval))


def sig_quinn(E, Epl, Ef, n):
    return O.lib().oracle_sigma_quinn(float(E), float(Epl), float(Ef), float(n))


# ---------------------------------------------------------------------------
# Joy-Luo Bethe law

def pin_jl(jl, cl=S_cl):
    """True when `jl(Z, E, n, k, J)` passes every JL pin."""
    for sp in ("Al", "Cu", "Au"):
        m = Material(sp)
        Z, n, k, J = m.Z, m.atnd, m.params["bethe_k"], bethe_J(m.Z)
        for E in (1500.0, 5000.0, 20000.0):
            if jl(Z, E, n, 0.0, J) != cl(Z, E, n, J):  # k = 0: the classical law, bit for bit
                return False
        # E >> kJ: S_JL - S_cl = 2 pi e^4 n Z / E ln(1 + kJ/E) ~ 2 pi e^4 n Z / E * kJ / E
        E = 1e8
        first = 2 * PI_EFOUR * n * Z / E * (k * J / E) * 1e-14
        if not math.isclose(jl(Z, E, n, k, J) - cl(Z, E, n, J), first, rel_tol=1e-4):
            return False
        # Joy & Luo's engineering form in keV/cm, E and J in keV (1 keV/cm = 1e-5 eV/A)
        p = m.params
        for E in (1500.0, 5000.0, 20000.0):
            Ek, Jk = E / 1000.0, J / 1000.0
            S_kev_cm = 78500.0 * p["density"] * Z / (p["atwt"] * Ek) * math.log(1.166 * (Ek + k * Jk) / Jk)
            if not math.isclose(jl(Z, E, n, k, J), S_kev_cm * 1e-5, rel_tol=2e-3):
                return False
    return True


def test_joy_luo_law_pins():
    assert pin_jl(S_jl)


@pytest.mark.parametrize("name,mutant", [
    ("k*J -> k", lambda Z, E, n, k, J: 2 * PI_EFOUR * n * (Z / E) * math.log(1.166 * (E + k) / J) * 1e-14),
    ("E + kJ -> E - kJ", lambda Z, E, n, k, J: 2 * PI_EFOUR * n * (Z / E) * math.log(1.166 * (E - k * J) / J) * 1e-14),
    ("dropped 2", lambda Z, E, n, k, J: PI_EFOUR * n * (Z / E) * math.log(1.166 * (E + k * J) / J) * 1e-14),
    ("1.166 on kJ only", lambda Z, E, n, k, J: 2 * PI_EFOUR * n * (Z / E) * math.log((E + 1.166 * k * J) / J) * 1e-14),
    ("E and J swapped in the log", lambda Z, E, n, k, J: 2 * PI_EFOUR * n * (Z / E) * math.log(1.166 * (J + k * E) / E)
     * 1e-14),
])
def test_joy_luo_pins_reject_planted_mistakes(name, mutant):
    assert not pin_jl(mutant), name


# ---------------------------------------------------------------------------
# explicit Bethe law (shell sum)

def pin_explicit(expl, cl=S_cl):
    """True when `expl(E, n, shells, Zval, Eval)` passes every explicit-law pin
    (shells: list of (occupancy, binding energy))."""
    for sp in ("Si", "Cu", "Au"):
        m = Material(sp)
        Z, n = m.Z, m.atnd
        J = bethe_J(Z)
        for E in (1000.0, 5000.0, 30000.0):
            # one valence term holding all Z electrons, no shells: the classical law at J = 1.166 E_val
            if not math.isclose(expl(E, n, [], Z, J / 1.166), cl(Z, E, n, J), rel_tol=1e-12):
                return False
            shells = [(s[2], s[1]) for s in m.shells]
            full = expl(E, n, shells, m.params["n_val"], m.params["E_val"])
            # a shell split into two halves at the same binding energy changes nothing
            split = [(shells[0][0] / 2, shells[0][1]), (shells[0][0] / 2, shells[0][1])] + shells[1:]
            if not math.isclose(expl(E, n, split, m.params["n_val"], m.params["E_val"]), full, rel_tol=1e-12):
                return False
            # each shell adds 2 pi e^4 n n_i / E ln(E / E_i)
            for i, (ni, Ei) in enumerate(shells):
                less = expl(E, n, shells[:i] + shells[i + 1:], m.params["n_val"], m.params["E_val"])
                add = 2 * PI_EFOUR * n / E * 1e-14 * ni * math.log(E / Ei)
                if not math.isclose(full - less, add, rel_tol=1e-9, abs_tol=1e-12 * abs(full)):
                    return False
    return True


def _expl(E, n, shells, Zval, Eval, *, f=lambda E, Ei: math.log(E / Ei), pref=2.0, lead=True):
    s = Zval * f(E, Eval) if lead else 0.0
    for ni, Ei in shells:
        s += ni * f(E, Ei)
    return pref * PI_EFOUR * n / E * 1e-14 * s


def test_explicit_law_pins():
    assert pin_explicit(S_expl)


@pytest.mark.parametrize("name,mutant", [
    ("valence term dropped", lambda E, n, sh, Zv, Ev: _expl(E, n, sh, Zv, Ev, lead=False)),
    ("ratio inverted", lambda E, n, sh, Zv, Ev: _expl(E, n, sh, Zv, Ev, f=lambda E, Ei: math.log(Ei / E))),
    ("dropped 2", lambda E, n, sh, Zv, Ev: _expl(E, n, sh, Zv, Ev, pref=1.0)),
    ("1.166 inside the shell logs", lambda E, n, sh, Zv, Ev: _expl(E, n, sh, Zv, Ev,
                                                                   f=lambda E, Ei: math.log(1.166 * E / Ei))),
    ("occupancy and energy swapped", lambda E, n, sh, Zv, Ev: _expl(E, n, [(b, a) for a, b in sh], Zv, Ev)),
])
def test_explicit_pins_reject_planted_mistakes(name, mutant):
    assert not pin_explicit(mutant), name


# ---------------------------------------------------------------------------
# Quinn plasmon cross section

def quinn_from_wave_vectors(E, Epl, Ef, n):
    """sigma = 1 / (lambda n) with 1/lambda = E_pl / (2 a0 E) ln(q_c / q_min),
    the wave vectors built in SI units from the energies [eV]; returns cm^2."""
    k = math.sqrt(2 * m_e * E * E_CHARGE) / hbar
    kp = math.sqrt(2 * m_e * (E - Epl) * E_CHARGE) / hbar
    kF = math.sqrt(2 * m_e * Ef * E_CHARGE) / hbar
    q_min = k - kp
    q_c = kF * (math.sqrt(1 + Epl / Ef) - 1)
    inv_lambda = Epl / (2 * BOHR_R * E) * math.log(q_c / q_min)  # 1/m
    return inv_lambda / n * 1e4


def pin_quinn(sig):
    """True when `sig(E, Epl, Ef, n)` passes every Quinn pin."""
    for sp in ("Al", "Si", "Cu", "Au"):
        m = Material(sp)
        Epl, Ef, n = m.plasmon_e, m.fermi_e, m.atnd
        for E in (5 * Epl, 1000.0, 20000.0, 1e6):
            if not math.isclose(sig(E, Epl, Ef, n), quinn_from_wave_vectors(E, Epl, Ef, n), rel_tol=1e-9):
                return False
        # E >> E_F: q_min / k_F -> E_pl / (2 sqrt(E E_F)), so the log argument is
        # (sqrt(1 + E_pl/E_F) - 1) 2 sqrt(E E_F) / E_pl up to O(E_pl / E)
        E = 1e6
        asym = Epl / (2 * BOHR_R * E) * math.log((math.sqrt(1 + Epl / Ef) - 1) * 2 * math.sqrt(E * Ef) / Epl) / n * 1e4
        if not math.isclose(sig(E, Epl, Ef, n), asym, rel_tol=Epl / E):
            return False
        # no plasmon below threshold: zero at and below E_pl, positive well above it
        if sig(Epl, Epl, Ef, n) != 0.0 or sig(0.5 * Epl, Epl, Ef, n) != 0.0:
            return False
        if not sig(10 * Epl, Epl, Ef, n) > 0.0:
            return False
    return True


def _quinn(E, Epl, Ef, n, *, two=2.0, swap=False, cm2=1e4):
    if E <= Epl:
        return 0.0
    a, b = (Ef, Epl) if swap else (Epl, Ef)
    x, p = E / b, a / b
    return a * math.log((math.sqrt(1 + p) - 1) / (math.sqrt(x) - math.sqrt(x - p))) / (two * BOHR_R * E) / n * cm2


def test_quinn_pins():
    assert pin_quinn(sig_quinn)


@pytest.mark.parametrize("name,mutant", [
    ("E_pl and E_F swapped", lambda E, Epl, Ef, n: _quinn(E, Epl, Ef, n, swap=True)),
    ("dropped 2", lambda E, Epl, Ef, n: _quinn(E, Epl, Ef, n, two=1.0)),
    ("m^2 -> cm^2 factor missing", lambda E, Epl, Ef, n: _quinn(E, Epl, Ef, n, cm2=1.0)),
    ("no threshold", lambda E, Epl, Ef, n: _quinn(E, Epl * 0.999999, Ef, n) if E > 0.5 * Epl else 0.0),
])
def test_quinn_pins_reject_planted_mistakes(name, mutant):
    assert not pin_quinn(mutant), name


def test_quinn_mean_free_path_scale():
    """Plasmon mean free paths of the order of 10^2 A at 10 keV, growing
    about linearly with E (the E_pl / E prefactor with a slowly varying
    log), as measured for Al and Si (Egerton, EELS in the Electron
    Microscope, 3rd ed., Sec. 3.3: ~100-200 nm at 100 keV)."""
    for sp in ("Al", "Si"):
        m = Material(sp)
        lam = lambda E: 1.0 / (m.atnd * 1e-6 * sig_quinn(E, m.plasmon_e, m.fermi_e, m.atnd)) * 1e8  # noqa: E731
        assert 50.0 < lam(10000.0) < 400.0
        assert 1000.0 < lam(100000.0) < 3000.0
        r = lam(20000.0) / lam(10000.0)
        assert 1.7 < r < 2.0


# ---------------------------------------------------------------------------
# Elastic cross-section models (crossSections.py:37-147, scattering.py:483-493,
# 780-789).  The Nigram model is the total of the screened Rutherford law
# dsigma/dOmega = (Z e^2 / 4E)^2 / (sin^2(theta/2) + beta)^2 with beta* =
# 0.48 * 5.43 Z^(2/3) / E: its integral over the sphere is
# pi Z^2 e^4 / (4 E^2 beta (1 + beta)), checked here by quadrature of the
# differential law, not by the closed form.  The wDefl models replace Z^2 by
# Z (Z + 1): their ratio to the base model is (Z + 1) / Z exactly.

def _nigram_quadrature(E, Z, wdefl=False):
    from scipy.integrate import quad
    beta = 0.48 * 5.43 * Z ** (2 / 3) / E
    e4 = PI_EFOUR / math.pi                       # e^4 in cm^2 eV^2
    zz = Z * (Z + 1) if wdefl else Z * Z
    pref = zz * e4 / (16.0 * E * E)

    def f(th):
        s = math.sin(0.5 * th) ** 2
        return 2 * math.pi * math.sin(th) * pref / (s + beta) ** 2
    # the law is sharply forward-peaked: split at the screening angle
    th_s = 2 * math.asin(min(1.0, math.sqrt(beta)))
    pts = sorted({th_s / 10, th_s, min(math.pi, 10 * th_s)})
    tot, last = 0.0, 0.0
    for p in pts + [math.pi]:
        if p > last:
            tot += quad(f, last, p, epsabs=0, epsrel=1e-12, limit=200)[0]
            last = p
    return tot


@pytest.mark.parametrize("Z", [13, 14, 29, 79])
@pytest.mark.parametrize("E", [500.0, 5000.0, 30000.0])
def test_nigram_sigma_is_the_screened_rutherford_total(E, Z):
    from mcd import _abi
    L = O.lib()
    got = L.oracle_sigma_ruth_model(E, Z, _abi.ELASTIC_NIGRAM)
    assert got == pytest.approx(_nigram_quadrature(E, Z), rel=1e-9, abs=0)
    got_w = L.oracle_sigma_ruth_model(E, Z, _abi.ELASTIC_NIGRAM_WDEFL)
    assert got_w == pytest.approx(_nigram_quadrature(E, Z, wdefl=True), rel=1e-9, abs=0)
    # a dropped factor 4, a missing (1 + beta) or an unscreened beta fail the pin
    beta = 0.48 * 5.43 * Z ** (2 / 3) / E
    for wrong in (got * 4, got * (1 + beta), got * 0.48):
        assert wrong != pytest.approx(_nigram_quadrature(E, Z), rel=1e-6, abs=0)


@pytest.mark.parametrize("Z", [13, 29, 79])
@pytest.mark.parametrize("E", [1000.0, 20000.0])
def test_wdefl_models_scale_by_z_plus_one_over_z(E, Z):
    from mcd import _abi
    L = O.lib()
    for base, wd in ((_abi.ELASTIC_RUTH, _abi.ELASTIC_RUTH_WDEFL), (_abi.ELASTIC_NIGRAM, _abi.ELASTIC_NIGRAM_WDEFL)):
        r = L.oracle_sigma_ruth_model(E, Z, wd) / L.oracle_sigma_ruth_model(E, Z, base)
        assert r == pytest.approx((Z + 1) / Z, rel=1e-14, abs=0)
    assert L.oracle_sigma_ruth_model(E, Z, _abi.ELASTIC_RUTH) == L.oracle_sigma_ruth(E, Z)


def test_elastic_model_resolution_follows_the_reference_string_tests():
    """DS tests the names in the order ==, in, ==, in (scattering.py:483-493);
    continuous mode tests 'vanilla' in, then 'nigram' in (:780-789), so its
    wDefl names fall back to the base models."""
    from mcd import _abi
    em = _abi.elastic_model
    assert [em(n, "DS") for n in ("Ruth_vanilla", "Ruth_vanilla_wDefl", "Ruth_nigram", "Ruth_nigram_wDefl")] == \
        [_abi.ELASTIC_RUTH, _abi.ELASTIC_RUTH_WDEFL, _abi.ELASTIC_NIGRAM, _abi.ELASTIC_NIGRAM_WDEFL]
    assert [em(n, "cont") for n in ("Ruth_vanilla", "Ruth_vanilla_wDefl", "Ruth_nigram", "Ruth_nigram_wDefl")] == \
        [_abi.ELASTIC_RUTH, _abi.ELASTIC_RUTH, _abi.ELASTIC_NIGRAM, _abi.ELASTIC_NIGRAM]
    with pytest.raises(ValueError):
        em("Ruth_nigram_wDelf", "DS")          # the reference's test input spells it so: no elastic sigma in DS
    assert em("Ruth_nigram_wDelf", "cont") == _abi.ELASTIC_NIGRAM
    with pytest.raises(NotImplementedError):
        em("Mott", "DS")


# ---------------------------------------------------------------------------
# Gryzinski differential cross section (extFunctions.py:189-219)
#
# The differential form and Gryzinski's closed-form total (crossSections.py:
# 213-242, the oracle's oracle_sigma_gryz, pinned by its threshold and
# high-energy asymptote in test_oracle_rng_physics) describe one cross
# section; their integral relation is fixed at both ends of the energy range.
# With U = E / E_b:
# * U -> 1+: eps, eps_B -> 1, so (1 + eps_B)^-3/2 -> 2^-3/2, the tail factor
#   (1 - eps)^(eps_B / (eps_B + eps)) -> (1 - eps)^1/2 and the bracket ->
#   (4/3) ln 2.7; with dW = E d(1 - eps) over [0, U - 1],
#   integral ~ n pi e^4 2^-3/2 (4/3) ln(2.7) (2/3) (U - 1)^3/2 / E_b^2.
# * U -> inf: the integral over [E_b, E] tends to the closed-form total (the
#   leading terms 1/(E E_b) and (2/3) ln sqrt(U) / (E E_b) are the same).

def _gryz_integral(dcs, U, Eb=50.0, n=2.0):
    import warnings

    from scipy.integrate import IntegrationWarning
    warnings.simplefilter("ignore", IntegrationWarning)  # quad's roundoff note at 1e-13 near threshold
    E = U * Eb
    if U < 1.5:  # near threshold: t = 1 - W/E = s^2 removes the square-root endpoint
        smax = math.sqrt(1.0 - 1.0 / U)
        f = lambda s: float(dcs(E, E * (1.0 - s * s), n, Eb)) * E * 2.0 * s  # noqa: E731
        v, _ = quad(f, 0.0, smax, limit=2000, epsrel=1e-13, epsabs=0.0)
    else:  # in ln W: the integrand falls like 1/W^2
        f = lambda t: float(dcs(E, math.exp(t), n, Eb)) * math.exp(t)  # noqa: E731
        v, _ = quad(f, math.log(Eb), math.log(E), limit=2000, epsrel=1e-13, epsabs=0.0)
    return v


def pin_gryz_dcs(dcs):
    """True when `dcs(E, W, n_s, E_b)` (float64) has both integral limits."""
    Eb, n = 50.0, 2.0
    try:
        d = 1e-8
        lead = n * PI_EFOUR * 2.0 ** -1.5 * (4.0 / 3.0) * math.log(2.7) * (2.0 / 3.0) * d ** 1.5 / Eb ** 2
        if not math.isclose(_gryz_integral(dcs, 1.0 + d, Eb, n), lead, rel_tol=1e-4):
            return False
        for U in (1e5, 1e6):
            tot = O.lib().oracle_sigma_gryz(U * Eb, Eb, n)
            if not math.isclose(_gryz_integral(dcs, U, Eb, n), tot, rel_tol=1e-4 if U < 1e6 else 1e-5):
                return False
    except (ZeroDivisionError, OverflowError, ValueError):
        return False
    return True


def _gryz_variant(E, W, n, Eb, part):
    eps, eB = W / E, Eb / E
    if not (eps < 1.0 and eB < 1.0):
        return 0.0
    a = (1.0 + eB) ** -1.5
    b = (1.0 - eps) ** (eB / (eB + eps))
    c = (1.0 - eB) + 4.0 * eB * math.log(2.7 + ((1.0 - eps) / eB) ** 0.5) / (3.0 * eps)
    if part == "no tail factor":
        b = 1.0
    elif part == "tail exponent eB/(eB - eps)":
        b = (1.0 - eps) ** (eB / (eB - eps))
    elif part == "tail exponent inverted":
        b = (1.0 - eps) ** ((eB + eps) / eB)
    elif part == "(1 + eB)^+3/2":
        a = (1.0 + eB) ** 1.5
    elif part == "bracket 1 + eB":
        c = (1.0 + eB) + 4.0 * eB * math.log(2.7 + ((1.0 - eps) / eB) ** 0.5) / (3.0 * eps)
    elif part == "log coefficient 2/3":
        c = (1.0 - eB) + 2.0 * eB * math.log(2.7 + ((1.0 - eps) / eB) ** 0.5) / (3.0 * eps)
    elif part == "ln(e + ...) for ln(2.7 + ...)":
        c = (1.0 - eB) + 4.0 * eB * math.log(math.e + ((1.0 - eps) / eB) ** 0.5) / (3.0 * eps)
    elif part == "no square root in the log":
        c = (1.0 - eB) + 4.0 * eB * math.log(2.7 + (1.0 - eps) / eB) / (3.0 * eps)
    elif part == "eps^2 under the log term":
        c = (1.0 - eB) + 4.0 * eB * math.log(2.7 + ((1.0 - eps) / eB) ** 0.5) / (3.0 * eps * eps)
    elif part == "W^2 for W^3":
        return n * PI_EFOUR * eps * a * b * c / W ** 2
    return n * PI_EFOUR * eps * a * b * c / W ** 3


def test_gryzinski_dcs_integral_limits():
    assert pin_gryz_dcs(physics.gryz_dCS)
    assert pin_gryz_dcs(lambda E, W, n, Eb: _gryz_variant(E, W, n, Eb, None))


@pytest.mark.parametrize("part", ["no tail factor", "tail exponent eB/(eB - eps)", "tail exponent inverted",
                                  "(1 + eB)^+3/2", "bracket 1 + eB", "log coefficient 2/3",
                                  "ln(e + ...) for ln(2.7 + ...)", "no square root in the log",
                                  "eps^2 under the log term", "W^2 for W^3"])
def test_gryzinski_dcs_pins_reject_planted_mistakes(part):
    assert not pin_gryz_dcs(lambda E, W, n, Eb: _gryz_variant(E, W, n, Eb, part)), part


def test_gryzinski_f32_restatement_matches_the_pinned_form():
    """The oracle's float32 restatement of the table build evaluates the
    pinned float64 form to float32 rounding (Al 2p, Cu 3p, Au 5p outer shells,
    losses across the whole W range, energies from above threshold to 30 keV)."""
    import tables_ref as TR
    for Eb, ns in ((72.9, 6.0), (74.0, 6.0), (57.0, 6.0)):
        for E in (1.2 * Eb, 3.0 * Eb, 1000.0, 5000.0, 20000.0, 30000.0):
            W = np.geomspace(Eb, 0.999 * E, 257).astype(np.float32)
            got = TR.gryz_dcs_f32(np.float32(E), W, ns, Eb).astype(np.float64)
            ref = physics.gryz_dCS(float(np.float32(E)), W.astype(np.float64), ns, Eb)
            assert np.allclose(got, ref, rtol=2e-5, atol=0), (Eb, E)
