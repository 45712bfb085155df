"""Pins of the FP64 oracle's random stream and physics formulas.

Each check compares the oracle with something that is not the oracle's own
formula: published known-answer vectors, an independent implementation, a
scipy quadrature of the differential cross section, a textbook form of the
same law, or tabulated free-electron values.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import oracle as O
import baseline_b1 as B1
from mcd import physics
from mcd.material import Material, bethe_J

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key, output)
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(O.philox(ctr, key)) == out


def test_philox_two_implementations_agree():
    """The C oracle's Philox and the vectorised numpy one (baseline_b1) agree
    on the transport's counter layout (electron, config, step, 0)."""
    seed = 0x1234_5678_9ABC_DEF0
    u = B1.philox_block(77, 3, range(0, 40), seed)
    for s in (0, 1, 17, 39):
        words = O.philox([77, 3, s, 0], [seed & 0xFFFFFFFF, seed >> 32])
        assert np.array_equal(u[s], [O.lib().oracle_u01(w) for w in words])


def test_uniform_mapping_exact_and_open():
    lib = O.lib()
    for x, want in [(0, 2.0 ** -24), (0xFFFFFFFF, 1 - 2.0 ** -24), (1 << 9, 3 * 2.0 ** -24)]:
        assert lib.oracle_u01(x) == want
    xs = np.random.default_rng(0).integers(0, 2 ** 32, 1000, dtype=np.uint64)
    us = np.array([lib.oracle_u01(int(x)) for x in xs])
    assert np.all((us > 0) & (us < 1))
    assert np.all(us.astype(np.float32).astype(np.float64) == us)  # exact in FP32


# ---------------------------------------------------------------------------
# total cross sections

@pytest.mark.parametrize("E,Wc,n", [(20000.0, 50.0, 11), (5000.0, 100.0, 4), (1200.0, 10.0, 3)])
def test_moller_sigma_is_integral_of_dcs(E, Wc, n):
    """moller_sigma (crossSections.py:183-208) is the integral of moller_dCS
    (extFunctions.py:159-184) over W in [Wc, E/2]."""
    val, _ = quad(lambda w: float(physics.moller_dCS(E, w, n)), Wc, E / 2, limit=400, epsabs=0, epsrel=1e-12)
    assert O.lib().oracle_sigma_moller(E, Wc, n) == pytest.approx(val, rel=1e-9)


@pytest.mark.parametrize("E,Z", [(20000.0, 29), (1000.0, 14), (30000.0, 79), (5000.0, 13)])
def test_rutherford_sigma_is_sphere_integral_of_screened_dcs(E, Z):
    """ruther_sigma's 4pi/(alpha(1+alpha)) is the sphere integral of the
    screened Rutherford shape 1/(sin^2(theta/2) + alpha)^2 (ruther_sigma_p,
    crossSections.py:359-385) with the same prefactor."""
    a = 3.4 * Z ** 0.666667 / E
    Ek = E * 1e-3
    pref = 5.21e-21 * (Z * Z / (Ek * Ek)) * ((Ek + 511.) / (Ek + 1024.)) ** 2
    ang, _ = quad(lambda t: 2 * math.pi * math.sin(t) / (math.sin(t / 2) ** 2 + a) ** 2, 0, math.pi,
                  points=[min(1.0, 20 * math.sqrt(a))], limit=400, epsabs=0, epsrel=1e-11)
    assert O.lib().oracle_sigma_ruth(E, Z) == pytest.approx(pref * ang, rel=1e-8)


def test_rutherford_angle_sampler_inverts_reference_cdf():
    """The polar-angle sampler c2 = 1 - alpha u/(1 + alpha - u)
    (Rutherford_halfPol, scattering.py:125-152) inverts the CDF the reference's
    tests hold for it (Ruth_diffCS_E._cdf, crossSections.py:404-408); the
    oracle's Rutherford trace rows are checked against it."""
    import tables_ref  # noqa: F401
    Z, E = 29, 20000.0
    a = 3.4 * Z ** 0.666667 / E
    for u in (1e-6, 0.01, 0.3, 0.5, 0.9, 0.999999):
        c2 = 1.0 - (a * u / (1.0 + a - u))
        cos_t = 2 * c2 - 1
        cdf = (-a - 1) * (1 - cos_t) / (-2 * a + cos_t - 1)
        assert cdf == pytest.approx(u, rel=1e-9)
    # also against the numerically integrated shape
    def cdf_num(theta):
        f = lambda t: math.sin(t) / (math.sin(t / 2) ** 2 + a) ** 2
        return quad(f, 0, theta, limit=400, epsrel=1e-12)[0] / quad(f, 0, math.pi, limit=400, epsrel=1e-12)[0]
    for u in (0.2, 0.7, 0.95):
        c2 = 1.0 - (a * u / (1.0 + a - u))
        assert cdf_num(math.acos(2 * c2 - 1)) == pytest.approx(u, rel=1e-7)


@pytest.mark.parametrize("E", [1000.5, 2000.0, 9000.0, 20000.0])
def test_gryzinski_sigma_close_to_integral_of_dcs(E):
    """Gryzinski's closed-form total (crossSections.py:213-242) and the
    integral of his differential form (extFunctions.py:189-219) over [E_b, E]
    are two approximations of one cross section; they agree to a few percent,
    which a dropped term in either would break."""
    for Eb, ns in [(74.0, 6.0), (977.0, 8.0), (120.0, 2.0)]:
        if E <= Eb:
            continue
        s = O.lib().oracle_sigma_gryz(E, Eb, ns)
        integ, _ = quad(lambda w: float(physics.gryz_dCS(E, w, ns, Eb)), Eb, E, limit=400, epsrel=1e-10)
        assert s == pytest.approx(integ, rel=0.25), (E, Eb)


def test_gryzinski_threshold_and_asymptote():
    lib = O.lib()
    assert lib.oracle_sigma_gryz(977.0, 977.0, 8) == 0.0
    assert lib.oracle_sigma_gryz(977.0 * (1 + 1e-9), 977.0, 8) > 0.0
    # sigma ~ (ns pi e^4 / (E Eb)) (1 + (2/3) ln(2.7 + sqrt(U))) for U >> 1
    E, Eb, ns = 1e9, 50.0, 2.0
    U = E / Eb
    approx = ns * physics.PI_EFOUR / (E * Eb) * (1 + 2 / 3 * math.log(2.7 + math.sqrt(U)))
    assert lib.oracle_sigma_gryz(E, Eb, ns) == pytest.approx(approx, rel=1e-3)


@pytest.mark.parametrize("sp", ["Al", "Si", "Cu", "Au"])
def test_quinn_sigma_two_algebraic_forms(sp):
    """The oracle's literal Quinn form (crossSections.py:278-281) and the
    cancellation-free form used by the host agree."""
    m = Material(sp)
    for E in (m.plasmon_e * 1.5, 1000.0, 20000.0):
        a = O.lib().oracle_sigma_quinn(E, m.plasmon_e, m.fermi_e, m.atnd)
        b = float(physics.quinn_sigma(E, m.plasmon_e, m.fermi_e, m.atnd))
        assert a == pytest.approx(b, rel=1e-9)


def test_mfp_magnitudes():
    """Mean free paths of the sigma set stay in the Angstrom range the
    reference's tables were tuned for (tens of A at 20 keV in Cu)."""
    m = Material("Cu")
    from mcd import _abi
    cfg = _abi.make_config(m, 20000, 1000, 50, 70)
    sig, lab = O.sigma_set(cfg, 20000.0)
    mfp = 1.0 / (m.atnd * sig.sum()) * 1e14
    assert 30 < mfp < 50
    p = sig / sig.sum()
    assert p[0] == pytest.approx(0.744, abs=0.01)          # Rutherford
    assert p[list(lab).index(2)] == pytest.approx(0.200, abs=0.01)  # Quinn


# ---------------------------------------------------------------------------
# material constants against free-electron values (Ashcroft & Mermin, Table 2.1)

def test_free_electron_constants():
    from mcd.material import fermi_energy, plasmon_energy, at_num_dens
    al = Material("Al")
    assert al.fermi_e == pytest.approx(11.7, rel=0.01)
    assert al.plasmon_e == pytest.approx(15.8, rel=0.01)
    cu1 = fermi_energy(at_num_dens(8.96, 63.546), 1)
    assert cu1 == pytest.approx(7.00, rel=0.01)
    assert plasmon_energy(at_num_dens(8.96, 63.546), 1) == pytest.approx(10.8, rel=0.02)
    # values restated independently from the reference's formulas with scipy's CODATA constants
    cu = Material("Cu")
    assert cu.atnd == pytest.approx(8.4912e28, rel=1e-4)
    assert cu.fermi_e == pytest.approx(34.843, rel=1e-4)
    assert cu.plasmon_e == pytest.approx(35.887, rel=1e-4)


# ---------------------------------------------------------------------------
# stopping powers

@pytest.mark.parametrize("sp,E", [("Cu", 20000.0), ("Au", 5000.0), ("Al", 1500.0)])
def test_bethe_classical_matches_textbook_units(sp, E):
    """S = 2 pi e^4 n Z / E ln(1.166 E / J) in eV/A equals the engineering form
    78500 rho Z / (A E[keV]) ln(1.166 E/J) keV/cm (Joy, Monte Carlo Modeling,
    eq. 3.21) to the precision of its rounded constant."""
    m = Material(sp)
    p = m.params
    J = bethe_J(p["Z"])
    S = O.lib().oracle_bethe_cl(p["Z"], E, m.atnd, J)
    S_keV_cm = 78500 * p["density"] * p["Z"] / (p["atwt"] * E / 1000) * math.log(1.166 * E / J)
    assert S == pytest.approx(S_keV_cm * 1e-5, rel=2e-3)


def test_mean_ionisation_fit():
    """J = 9.76 Z + 58.5 Z^-0.19 (Berger-Seltzer fit, stoppingPowers.py:134-137)."""
    assert bethe_J(29) == pytest.approx(313.9, abs=0.2)
    assert bethe_J(79) == pytest.approx(796.54, abs=0.02)  # 771.04 + 58.5/2.2937
    assert bethe_J(12) == pytest.approx(138.0)


def test_explicit_bethe_folded_form():
    """The kernel's folded explicit law Z_eff ln(E/E*) equals the reference's
    shell sum (stoppingPowers.py:208-215) as evaluated by the oracle."""
    for sp in ("Au", "Cu", "Si"):
        m = Material(sp)
        zeff, lnE = m.explicit_bethe_constants()
        shells = m.shells
        Zi = (C_double := __import__("ctypes").c_double * len(shells))(*[s[2] for s in shells])
        Ei = C_double(*[s[1] for s in shells])
        for E in (1000.0, 5000.0, 30000.0):
            a = O.lib().oracle_bethe_explicit(E, m.atnd, Zi, Ei, len(shells), m.params["n_val"], m.params["E_val"])
            b = 2 * physics.PI_EFOUR * m.atnd / E * 1e-14 * zeff * (math.log(E) - lnE)
            assert a == pytest.approx(b, rel=1e-9, abs=1e-12)
