"""Host-side FP64 physics formulas.

These are used on the host for table grids, launch constants and the per-config
channel-order analysis; the transport itself runs on the GPU.

* differential cross sections moller_dCS / gryz_dCS: MC/tools/extFunctions.py:159-219
* total cross sections: MC/tools/crossSections.py:14-58 (Rutherford), :183-208
  (Moller), :213-242 (Gryzinski), :247-284 (Quinn)
* stopping powers: MC/tools/stoppingPowers.py:112-215
* maximum energy losses: MC/tools/probTables.py:17-41
"""
from __future__ import annotations

import numpy as np

PI_EFOUR = 6.51408491409531e-14     # cm^2 eV^2, parameters.py:7
BOHR_R = 5.29177210903e-11          # m, parameters.py:8 (scipy CODATA 2018)
PI = np.pi


def maxW_moller(E, Ef):
    """Largest Moller energy loss (E - E_F)/2, probTables.py:17-25."""
    return (E - Ef) * 0.5


maxW_gryz = maxW_moller  # probTables.py:27-41 uses the same bound


def moller_dCS(E, W, nfree):
    """Moller differential cross section [cm^2/eV], extFunctions.py:159-184."""
    eps = W / E
    with np.errstate(divide="ignore", invalid="ignore"):
        val = nfree * PI_EFOUR * (1.0 / (eps ** 2) + (1.0 / ((1.0 - eps) ** 2)) - (1.0 / (eps * (1.0 - eps)))) / E ** 3
    return np.where(eps < 1.0, val, 0)


def gryz_dCS(E, W, nsi, Ebi):
    """Gryzinski differential cross section [cm^2/eV], extFunctions.py:189-219."""
    eps = W / E
    epsB = Ebi / E
    with np.errstate(divide="ignore", invalid="ignore"):
        val = (nsi * PI_EFOUR * eps * (1.0 + epsB) ** (-1.5) * (1.0 - eps) ** (epsB / (epsB + eps))
               * ((1.0 - epsB) + 4.0 * epsB * np.log(2.7 + ((1.0 - eps) / epsB) ** 0.5) / (3.0 * eps)) / (W ** 3))
    return np.where((eps < 1.0) & (epsB < 1.0), val, 0)


def alpha(E, Z):
    """Rutherford screening parameter, crossSections.py:14-27."""
    return 3.4 * (Z ** 0.666667) / E


def ruther_sigma(E, Z):
    """Rutherford total cross section [cm^2], crossSections.py:37-58."""
    a = alpha(E, Z)
    Ek = E * 1e-3
    return 5.21e-21 * (Z * Z / (Ek * Ek)) * ((Ek + 511.0) / (Ek + 1024.0)) ** 2 * (4.0 * PI) / (a * (1.0 + a))


def moller_sigma(E, Wc, nfree):
    """Moller total cross section [cm^2], crossSections.py:183-208."""
    E = np.asarray(E, dtype=np.float64)
    eps = Wc / E
    with np.errstate(divide="ignore", invalid="ignore"):
        val = nfree * PI_EFOUR * (1.0 / (E * E)) * (1.0 / eps - 1.0 / (1.0 - eps) + np.log(eps / (1.0 - eps)))
    return np.where(E <= Wc, 0.0, val)


def gryz_sigma(E, Esi, nsi):
    """Gryzinski total cross section [cm^2], crossSections.py:213-242."""
    E = np.asarray(E, dtype=np.float64)
    U = E / Esi
    with np.errstate(divide="ignore", invalid="ignore"):
        val = (nsi * PI_EFOUR * ((U - 1.0) / (U + 1.0)) ** 1.5
               * (1.0 + 2.0 * (1.0 - (1.0 / (2.0 * U))) / 3.0 * np.log(2.7 + np.sqrt(np.maximum(U - 1.0, 0.0))))
               / (E * Esi))
    return np.where(E <= Esi, 0.0, val)


def quinn_sigma(E, Epl, Ef, n):
    """Quinn plasmon total cross section [cm^2], crossSections.py:247-284,
    with sqrt(a) - sqrt(a - x) evaluated as x / (sqrt(a) + sqrt(a - x))."""
    E = np.asarray(E, dtype=np.float64)
    a = E / Ef
    x = Epl / Ef
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = (np.sqrt(1.0 + x) - 1.0) * (np.sqrt(a) + np.sqrt(np.maximum(a - x, 0.0))) / x
        val = Epl * np.log(ratio) / (2.0 * BOHR_R * E) / n * 1e4
    return np.where(E <= Epl, 0.0, val)


def sigma_set(mat, Wc, E):
    """Insertion-ordered channel cross sections (scattering.py:480-508):
    rows Ruth, Moller, Gryz<shell> (name_s order), Quinn."""
    E = np.asarray(E, dtype=np.float64)
    rows = [ruther_sigma(E, mat.Z), moller_sigma(E, Wc, mat.params["n_val"])]
    for _, Eb, ns in mat.shells:
        rows.append(gryz_sigma(E, Eb, ns))
    rows.append(quinn_sigma(E, mat.plasmon_e, mat.fermi_e, mat.atnd))
    return np.stack(rows)


def bethe_cl_sp(Z, E, n, J):
    """Classical Bethe stopping power [eV/A], stoppingPowers.py:112-140."""
    return 2.0 * PI_EFOUR * n * (Z / E) * np.log(1.166 * E / J) * 1e-14


def bethe_mod_sp_k(Z, E, n, k, J):
    """Joy-Luo Bethe stopping power [eV/A], stoppingPowers.py:145-173."""
    return 2.0 * PI_EFOUR * n * (Z / E) * np.log(1.166 * (E + k * J) / J) * 1e-14


def bethe_mod_sp(E, n, Zi, Ei, Zval, Eval):
    """Explicit-shell Bethe stopping power [eV/A], stoppingPowers.py:177-215."""
    s = Zval * np.log(E / Eval)
    for z, eb in zip(Zi, Ei):
        s = s + z * np.log(E / eb)
    return (2.0 * PI_EFOUR * n / E) * 1e-14 * s
