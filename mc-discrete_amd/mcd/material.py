"""Material database and derived electron-gas constants.

Values follow MC/tools/material.py:151-302 (the reference's scatter_params) and
the derived quantities of material.__init__ (material.py:117-134):

* atomic number density   n    = rho * N_A / A * 1e6            [m^-3]  (:15-30)
* Fermi energy            E_F  = hbar^2 (3 pi^2 n_val n)^(2/3) / (2 m_e) / e  [eV] (:34-56)
* plasmon energy          E_pl = hbar e sqrt(n_val n / (eps0 m_e)) / e        [eV] (:61-88)

The reference takes its physical constants from scipy.constants; so does this
module.  The conversion J -> eV uses scipy's elementary charge.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

from scipy.constants import Avogadro, e, epsilon_0, hbar, m_e, pi

MAX_SHELLS = 8

# scatter_params, material.py:151-302.  name_s order matters: it is the
# insertion order of the Gryzinski channels (scattering.py:500-503) and its
# reverse is the table build order (probTables.py:579).
_DB = {
    "Al": dict(Z=13, n_val=3, E_val=72.55, name_val="3s2-3p1",
               name_s=["1s", "2s", "2p"], Es={"1s": 1559.0, "2s": 118.0, "2p": 73.5},
               ns={"1s": 2, "2s": 2, "2p": 6}, density=2.70, atwt=26.98154,
               bethe_k=0.815, penn_b=-2.16),
    "Si": dict(Z=14, n_val=4, E_val=99.42, name_val="3s2-3p2",
               name_s=["1s", "2s", "2p"], Es={"1s": 1189, "2s": 149, "2p": 100},
               ns={"1s": 2, "2s": 2, "2p": 6}, density=2.33, atwt=28.085,
               bethe_k=0.822, penn_b=-2.19, xip_g=1360),
    "Cu": dict(Z=29, n_val=11, E_val=75.1, name_val="3d10-4s2",
               name_s=["1s", "2s2p", "3s", "3p"], Es={"1s": 8980, "2s2p": 977, "3s": 120, "3p": 74},
               ns={"1s": 2, "2s2p": 8, "3s": 2, "3p": 6}, density=8.96, atwt=63.546,
               bethe_k=0.83, penn_b=-3.21),
    "Au": dict(Z=79, n_val=11, E_val=75.1, name_val="5d10-6s1",
               name_s=["2s2p", "3s3p3d", "4s4p", "4d", "5s", "4f", "5p2", "5p4"],
               Es={"2s2p": 12980, "3s3p3d": 2584, "4s4p": 624, "4d": 341, "5s": 178, "4f": 85,
                   "5p2": 72, "5p4": 54},
               ns={"2s2p": 8, "3s3p3d": 18, "4s4p": 8, "4d": 10, "5s": 2, "4f": 14, "5p2": 2, "5p4": 4},
               density=19.30, atwt=196.967, bethe_k=0.851, penn_b=-2.16),
}

SPECIES = tuple(_DB)


def at_num_dens(dens: float, atom_mass: float) -> float:
    """Atomic number density [m^-3] (material.py:15-30)."""
    return dens * Avogadro / atom_mass * 1e6


def fermi_energy(atnd: float, nvalence: float) -> float:
    """Free-electron Fermi energy [eV] (material.py:34-56)."""
    n = nvalence * atnd
    return hbar ** 2 * (3.0 * (pi ** 2) * n) ** (2.0 / 3.0) / (2.0 * m_e) / e


def plasmon_energy(atnd: float, nvalence: float) -> float:
    """Free-electron plasmon energy [eV] (material.py:61-88)."""
    n = nvalence * atnd
    return hbar * e * (n / (epsilon_0 * m_e)) ** 0.5 / e


def bethe_J(Z: int) -> float:
    """Mean ionisation energy of the Bethe laws [eV] (stoppingPowers.py:134-137)."""
    return 9.76 * Z + 58.5 / (Z ** 0.19) if Z >= 13 else 11.5 * Z


@dataclass
class Material:
    """A target material: tabulated constants plus derived n, E_F, E_pl.

    ``density_scale`` multiplies the mass density before the derived quantities
    are formed, so it moves n, E_F and E_pl together the way the reference's
    ``density`` parameter does (material.py:127-131).
    """

    species: str
    density_scale: float = 1.0
    params: dict = field(init=False)
    atnd: float = field(init=False)
    fermi_e: float = field(init=False)
    plasmon_e: float = field(init=False)

    def __post_init__(self):
        if self.species not in _DB:
            raise ValueError(f"unknown material {self.species!r}; known: {', '.join(SPECIES)}")
        p = dict(_DB[self.species])
        p["Es"] = dict(p["Es"])
        p["ns"] = dict(p["ns"])
        p["name_s"] = list(p["name_s"])
        self.params = p
        self.atnd = at_num_dens(p["density"] * self.density_scale, p["atwt"])
        self.plasmon_e = plasmon_energy(self.atnd, p["n_val"])
        self.fermi_e = fermi_energy(self.atnd, p["n_val"])

    @property
    def Z(self) -> int:
        return self.params["Z"]

    @property
    def shells(self):
        """(name, E_b, n) in name_s order."""
        p = self.params
        return [(s, float(p["Es"][s]), float(p["ns"][s])) for s in p["name_s"]]

    @property
    def outer_shell(self) -> str:
        """The shell whose table serves every Gryzinski event: tables are built
        over name_s[::-1] and events read tables['Gryz'][0]
        (probTables.py:579, scattering.py:625)."""
        return self.params["name_s"][-1]

    def explicit_bethe_constants(self):
        """(Z_eff, ln E*) with sum_i n_i ln(E/E_i) = Z_eff ln(E/E*), the folded
        shell sum of bethe_mod_sp (stoppingPowers.py:208-215) including the
        valence term."""
        p = self.params
        terms = [(float(p["n_val"]), float(p["E_val"]))] + [(n, Eb) for _, Eb, n in self.shells]
        zeff = sum(n for n, _ in terms)
        ln_estar = sum(n * math.log(Eb) for n, Eb in terms) / zeff
        return zeff, ln_estar
