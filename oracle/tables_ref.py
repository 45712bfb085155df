"""Float32 restatement of the reference's CDF table fill.

TEST INFRASTRUCTURE ONLY (the checker): imported by tests/, the smoke check and
bench.py's CPU-baseline leg, never by the product.

probTable.computeBlock (MC/tools/probTables.py:415-462) with numpy.ma semantics:
a trapezoid bin j of row i is masked when its upper edge W[j+1] exceeds
W_max(E_i) = (E_i - E_F)/2 (:429-439); the masked bins count as zero in the
float32 cumulative sum (MaskedArray.cumsum fills with 0) and each row is
divided by its last cumulative value (:457-460, np.broadcast_to drops the
mask); masked entries are stored as 0 (mapToMemory, :507).  All arithmetic is
float32, as the float32 Es/Ws arrays force under numpy 1.x promotion.

Pinning (tests/test_oracle_pins.py, tests/test_oracle_algorithms.py): the
Moller rows against the analytic Moller CDF; the Gryzinski differential form
through the limits of its integral at threshold and at high energy (against
Gryzinski's closed-form total), with planted mistakes rejected, and this
float32 evaluation against the pinned float64 form to float32 rounding.
"""
import numpy as np

PI_EFOUR = 6.51408491409531e-14  # parameters.py:7


def moller_dcs_f32(E, W, nfree):
    """extFunctions.moller_dCS (:159-184) on float32 arrays."""
    f = np.float32
    eps = W / E
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        val = f(nfree * PI_EFOUR) * (f(1.) / np.square(eps) + (f(1.) / np.square(f(1.) - eps))
                                     - (f(1.) / (eps * (f(1.) - eps)))) / np.power(E, f(3))
    return np.where(eps < f(1.), val, f(0)).astype(np.float32)


def gryz_dcs_f32(E, W, nsi, Ebi):
    """extFunctions.gryz_dCS (:189-219) on float32 arrays."""
    f = np.float32
    eps = W / E
    epsB = f(Ebi) / E
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        val = (f(nsi * PI_EFOUR) * eps * np.power(f(1.) + epsB, f(-1.5))
               * np.power(f(1.) - eps, epsB / (epsB + eps))
               * ((f(1.) - epsB) + f(4.) * epsB * np.log(f(2.7) + np.sqrt((f(1.) - eps) / epsB)) / (f(3.) * eps))
               / np.power(W, f(3)))
    return np.where((eps < f(1.)) & (epsB < f(1.)), val, f(0)).astype(np.float32)


def cdf_table(Es, Ws, Ef, kind, n_el, Eb=0.0):
    """Return the (nE, nW-1) float32 CDF with masked entries stored as 0."""
    f = np.float32
    Es = np.asarray(Es, dtype=np.float32)
    Ws = np.asarray(Ws, dtype=np.float32)
    wmax = (Es - f(Ef)) * f(0.5)
    Em, Wm = np.meshgrid(Es, Ws, indexing="ij")
    if kind == "Moller":
        fv = moller_dcs_f32(Em, Wm, n_el)
    else:
        fv = gryz_dcs_f32(Em, Wm, n_el, Eb)
    valid = ~(Wm[:, 1:] > wmax[:, None])           # upper edge inside W_max(E)
    d = Wm[:, 1:] - Wm[:, :-1]
    step = (d * (fv[:, 1:] + fv[:, :-1]) / f(2.)).astype(np.float32)
    step = np.where(valid, step, f(0))
    cum = np.cumsum(step, axis=1, dtype=np.float32)
    tot = cum[:, -1:]
    with np.errstate(divide="ignore", invalid="ignore"):
        cdf = (cum / tot).astype(np.float32)
    return np.where(valid, cdf, f(0)).astype(np.float32)
