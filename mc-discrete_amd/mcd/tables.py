"""Inverse-CDF energy-loss tables (Moller and outer-shell Gryzinski).

The grids follow the reference's tolerance-driven search (MC/tools/probTables.py):

* ``find_W_series``  (probTable.findWseries, :280-328): float32 geomspace W grid on
  [Wmin, W_max(E0)], its size the first n of 100, 1000, ... (ordersOfMag, :54-58)
  and then the first 10 % step (percents, :61-64) whose trapezoid integral of
  dsigma/dW agrees with scipy's quad within tol_W (check_intTol, :165-188).
* ``find_E_series`` (findEseries, :332-382): float32 geomspace, then linspace, E grid
  on [Emin', E0] chosen with check_CDFTol (:191-232) at tolerance tol_E.

Arithmetic types follow the numpy 1.x promotion rules the reference ran under:
a numpy float32 scalar combined with a Python float or another scalar is
computed in float64, while an array combined with a scalar of the same kind
stays float32 (the scalar is rounded to float32 first).

The CDF rows themselves (computeBlock, :415-462) are filled on the GPU by
``mcd_tables_create`` (FP64 arithmetic by default, rounded to float32).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.integrate import quad

from . import physics
from .material import Material


def orders_of_mag(max_size=1000000, num_iter=5):
    """probTables.ordersOfMag (:54-58)."""
    for n in np.geomspace(100, max_size, num=num_iter, dtype=int):
        yield int(n)


def percents(max_size, num_iter=10):
    """probTables.percents (:61-64)."""
    for n in np.linspace(max_size / 10, max_size, num_iter, dtype=np.float32):
        yield int(n)


@dataclass
class TableGrid:
    """One table's grids and integrand parameters (probTable.__init__, :89-162)."""

    kind: str            # 'Moller' or 'Gryzinski'
    shell: str
    Wmin: float
    n_el: float
    Eb: float            # Gryzinski binding energy (0 for Moller)
    Ef: float
    Emin: float
    Emax: float
    Es: np.ndarray = None  # float32 [nE]
    Ws: np.ndarray = None  # float32 [nW]

    # -- integrands ---------------------------------------------------------
    def f64(self, E, W):
        """dsigma/dW in float64 (scalar-scalar numpy 1.x promotion)."""
        if self.kind == "Moller":
            return physics.moller_dCS(np.float64(E), np.asarray(W, dtype=np.float64), self.n_el)
        return physics.gryz_dCS(np.float64(E), np.asarray(W, dtype=np.float64), self.n_el, self.Eb)

    def f32_array(self, E32, W32):
        """dsigma/dW with a float32 energy scalar and a float32 loss array, in the
        float32 array arithmetic of the reference (float64 scalars rounded to
        float32 where they meet the array)."""
        f32 = np.float32
        E32 = f32(E32)
        W32 = np.asarray(W32, dtype=np.float32)
        eps = W32 / E32
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            if self.kind == "Moller":
                k = f32(self.n_el * physics.PI_EFOUR)
                br = (f32(1.0) / (eps ** 2) + (f32(1.0) / ((f32(1.0) - eps) ** 2))
                      - (f32(1.0) / (eps * (f32(1.0) - eps))))
                e3 = f32(float(E32) ** 3)
                val = k * br / e3
                return np.where(eps < 1.0, val, f32(0)).astype(np.float32)
            epsB = float(self.Eb) / float(E32)
            a = f32((1.0 + epsB) ** (-1.5))
            expo = f32(epsB) / (f32(epsB) + eps)
            inner = ((f32(1.0) - eps) / f32(epsB)) ** f32(0.5)
            c = f32(1.0 - epsB) + f32(4.0 * epsB) * np.log(f32(2.7) + inner) / (f32(3.0) * eps)
            val = f32(self.n_el * physics.PI_EFOUR) * eps * a * (f32(1.0) - eps) ** expo * c / (W32 ** 3)
            return np.where((eps < 1.0) & (epsB < 1.0), val, f32(0)).astype(np.float32)

    def Wmax(self, E):
        return (E - self.Ef) * 0.5

    # -- tolerance tests -----------------------------------------------------
    def _check_int_tol(self, series, ref, E, tol_W):
        """check_intTol (:165-188): float64 integrand on the float32 grid,
        float32 grid differences, float64 trapezoid."""
        y = self.f64(E, series.astype(np.float64))
        d = np.diff(series).astype(np.float64)
        trap = float((d * (y[1:] + y[:-1]) / 2.0).sum())
        return abs(trap - ref) / ref < tol_W

    def _check_cdf_tol(self, Eseries, ref, tol_E):
        """check_CDFTol (:191-232): float32 trapezoid of dsigma/dW at the
        second-largest grid energy against the quad integral at E0."""
        E = Eseries[-2]
        wmax = np.float32((float(E) - self.Ef) * 0.5)
        sel = self.Ws[self.Ws <= wmax]
        y = self.f32_array(E, sel)
        d = np.diff(sel)
        tot = np.float32((d * (y[1:] + y[:-1]) / np.float32(2.0)).sum(dtype=np.float32))
        return abs(float(tot) - ref) / ref < tol_E

    # -- searches ------------------------------------------------------------
    def find_W_series(self, tol_W):
        """findWseries (:280-328) at E = Emax."""
        E = self.Emax
        W_max = self.Wmax(E)
        ref, err = quad(lambda w: float(self.f64(E, w)), self.Wmin, W_max, limit=300, epsabs=1e-34)
        if not err < tol_W * ref:
            raise ValueError(f"{self.kind} {self.shell}: quad error {err:g} exceeds tol_W*integral "
                             f"(probTables.py:305)")
        n1 = None
        for n in orders_of_mag(1000000):
            s = np.geomspace(self.Wmin, W_max, num=n, dtype=np.float32)
            if self._check_int_tol(s, ref, E, tol_W):
                n1 = n
                break
        if n1 is None:
            raise ValueError(f"{self.kind} {self.shell}: no W grid up to 1e6 points meets tol_W")
        for n in percents(n1):
            s = np.geomspace(self.Wmin, W_max, num=n, dtype=np.float32)
            if self._check_int_tol(s, ref, E, tol_W):
                return s
        raise ValueError(f"{self.kind} {self.shell}: W refinement failed")

    def find_E_series(self, tol_E):
        """findEseries (:332-382)."""
        Eref = self.Emax
        ref, _ = quad(lambda w: float(self.f64(Eref, w)), self.Wmin, self.Wmax(Eref), limit=300, epsabs=1e-34)
        n1 = None
        for n in orders_of_mag(1000000):
            s = np.geomspace(self.Emin, self.Emax, num=n, dtype=np.float32)
            if self._check_cdf_tol(s, ref, tol_E):
                n1 = n
                break
        if n1 is None:
            raise ValueError(f"{self.kind} {self.shell}: no E grid up to 1e6 points meets tol_E")
        for n in percents(n1):
            s = np.linspace(self.Emin, self.Emax, num=n, dtype=np.float32)
            if self._check_cdf_tol(s, ref, tol_E):
                return s
        raise ValueError(f"{self.kind} {self.shell}: E refinement failed")

    def build_grids(self, tol_E, tol_W):
        """set_Ws then set_Es (:385-411)."""
        self.Ws = self.find_W_series(tol_W)
        if not np.all(self.Ws < self.Emax):
            raise ValueError("energy-loss grid reaches the incident energy (probTables.py:396)")
        self.Es = self.find_E_series(tol_E)
        return self

    def build_grids_library(self, tol_E, tol_W):
        """The same search run by the library (mcd_grid_search: host C++, an
        FP64 Gauss-Kronrod integral in place of scipy's quad)."""
        import ctypes as C

        from ._lib import check, lib

        kind = 0 if self.kind == "Moller" else 1
        args = (kind, float(self.Wmin), float(self.n_el), float(self.Eb), float(self.Ef), float(self.Emin),
                float(self.Emax), float(tol_E), float(tol_W))
        nE, nW = C.c_int32(0), C.c_int32(0)
        check(lib().mcd_grid_search(*args, None, C.byref(nE), None, C.byref(nW)))
        Es = np.zeros(nE.value, dtype=np.float32)
        Ws = np.zeros(nW.value, dtype=np.float32)
        check(lib().mcd_grid_search(*args, Es.ctypes.data, C.byref(nE), Ws.ctypes.data, C.byref(nW)))
        self.Es, self.Ws = Es, Ws
        return self

    @property
    def shape(self):
        return (int(self.Es.size), int(self.Ws.size) - 1)


def moller_grid(mat: Material, E0: float, Emin: float, Wc: float) -> TableGrid:
    """The Moller table of genTables (:557-575)."""
    return TableGrid(kind="Moller", shell=mat.params["name_val"], Wmin=float(Wc),
                     n_el=float(mat.params["n_val"]), Eb=0.0, Ef=mat.fermi_e,
                     Emin=float(Emin), Emax=float(E0))


def gryz_grid(mat: Material, E0: float, Emin: float, shell: str | None = None) -> TableGrid:
    """A Gryzinski table (:577-599); by default the outermost shell, the only
    one the transport reads (scattering.py:625)."""
    shell = shell or mat.outer_shell
    Eb = float(mat.params["Es"][shell])
    lo = Eb if Emin < Eb else float(Emin)
    return TableGrid(kind="Gryzinski", shell=shell, Wmin=Eb, n_el=float(mat.params["ns"][shell]),
                     Eb=Eb, Ef=mat.fermi_e, Emin=lo, Emax=float(E0))



def grid_key(kind, mat: Material, E0, Emin, Wc, tol_E, tol_W):
    """File key naming every parameter the table depends on (the reference's
    '{type}_{shell}.table' omits them, probTables.py:160)."""
    return (f"{kind}_{mat.species}_E0{E0:g}_Emin{Emin:g}_Wc{Wc:g}_tolE{tol_E:g}_tolW{tol_W:g}"
            f"_rho{mat.density_scale:g}")


def fmt_shape(g: TableGrid) -> str:
    return f"{g.kind}[{g.shell}] {g.shape[0]}x{g.shape[1]}"


__all__ = ["TableGrid", "moller_grid", "gryz_grid", "orders_of_mag", "percents", "grid_key"]
