"""The reference's distribution tests (tests/testThings.py:200-830,
TestScatterAnglesforDS) applied to GPU scatter rows, and on the CPU to the
oracle's step trace.

Workload: the reference's test input (tests/testInput.file): Al, E0 20 keV,
Emin 15 keV, Wc 10 eV, tilt 0, DS, tables at tol_E = 1e-4, tol_W = 1e-7
(3000 x 20000 Moller table).  The table draw returns the lower W node of the
bin that u falls in, so the fine W grid matters: at tol_W = 1e-4 (1.4% node
spacing) the angles of 2000 draws already differ from the continuous law
(Watson p = 0.004 on the oracle's rows) while matching the discrete law of
the table exactly.

Each sampled quantity is checked twice:
* as the reference does: 200 MC angles against 200 angles drawn from the
  theoretical distribution, two-sample Watson U^2 (watson_two_test,
  testThings.py:255-316).  The reference asserts U^2 < 0.152, the critical
  value at p = 0.10: a correct sampler fails that one time in ten, and about
  1.6 times over the 16 cases here.  The criterion here is p > 1e-3 from the
  asymptotic law of U^2, P(U^2 > x) = 2 sum_k (-1)^(k-1) exp(-2 k^2 pi^2 x)
  (Watson 1961), the level of the KS checks below;
* over every row of the run: the row's value mapped through the theoretical
  CDF at the row's own pre-step energy (probability integral transform) must
  be uniform on (0, 1), Kolmogorov-Smirnov p > 1e-3.  This sees errors of
  about 1/sqrt(n) in the CDF, far below what 200 draws can.  Table-sampled
  energy losses take grid values only (the lower node of the CDF bin that
  u falls in), so for them the transform is the randomised one of a discrete
  law, with the exact CDF evaluated at the table row's energy and nodes.

Theoretical CDFs:
* Rutherford polar angle: the analytic CDF of Ruth_diffCS_E
  (crossSections.py:391-421);
* Moller energy loss on [Wc, (E - E_F)/2] (Moller_W_E, distributions.py:20-48):
  the closed-form antiderivative of moller_dCS (extFunctions.py:159-184),
  G(e) = -1/e + 1/(1 - e) - ln e + ln(1 - e), e = W/E;
* Gryzinski energy loss of one shell on [E_b, (E - E_F)/2] (Gryz_Last_W_E,
  distributions.py:80-125): 64-point Gauss-Legendre quadrature of gryz_dCS
  (extFunctions.py:189-219) in ln W, itself checked against the Moller closed
  form on the Moller integrand;
* polar angle of Moller and Gryzinski events from the energy loss by the
  binary collision model, cos^2(theta/2) = (sqrt(1 - W/E) + 1)/2
  (scattering.py:156-180);
* azimuth: phi/2 uniform on [0, pi) (testThings.py:492-524).
"""
import numpy as np
import pytest
from scipy import stats

from conftest import gpu_available
from test_oracle_algorithms import watson_u2

E0, EMIN, WC = 20000.0, 15000.0, 10.0
PI_EFOUR = 6.51e-14  # cm^2 eV^2; the constant cancels in every normalised CDF
U2_CRIT_REF = 0.152  # testThings.py:310-313 (p = 0.10)

_GL_X, _GL_W = np.polynomial.legendre.leggauss(64)


@pytest.fixture(scope="module", params=[
    "oracle",
    pytest.param("gpu", marks=[pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no GPU")])])
def al20(request):
    """Scatter rows of the workload: 16384 electrons on the GPU (mcd_forward
    scatter rows, about 220 per electron); 4096 electrons of the oracle's
    step trace on the CPU (the same checks then pin the oracle's sampling,
    with less power)."""
    from mcd import _abi
    from mcd.material import Material
    from mcd.tables import gryz_grid, moller_grid
    mat = Material("Al")
    assert mat.outer_shell == min(mat.params["Es"], key=mat.params["Es"].get)
    if request.param == "gpu":
        from mcd import engine
        tr = engine.Transport(mat, E0, EMIN, Wc=WC, s_tilt=0.0, tol_E=1e-4, tol_W=1e-7)
        n = 16384
        r = tr.run(n, seed=21, scats=True, scat_capacity=n * 400)
        return dict(mat=mat, cfg=tr.cfg, grids=tr.grids, rows=r.scats, big=True)
    import oracle as O
    import tables_ref as TR
    gm = moller_grid(mat, E0, EMIN, WC).build_grids(1e-4, 1e-7)
    gg = gryz_grid(mat, E0, EMIN).build_grids(1e-4, 1e-7)
    tm = O.Table(gm.Es, gm.Ws, TR.cdf_table(gm.Es, gm.Ws, mat.fermi_e, "Moller", gm.n_el))
    tg = O.Table(gg.Es, gg.Ws, TR.cdf_table(gg.Es, gg.Ws, mat.fermi_e, "Gryzinski", gg.n_el, gg.Eb))
    cfg = _abi.make_config(mat, E0, EMIN, WC, 0.0)
    _, trace, _ = O.run(cfg, tm, tg, 21, 0, 4096, trace_cap=4096 * 1500, nthreads=8)
    return dict(mat=mat, cfg=cfg, grids=(gm, gg), rows=trace, big=False)


def moller_G(eps):
    return -1.0 / eps + 1.0 / (1.0 - eps) - np.log(eps) + np.log(1.0 - eps)


def moller_cdf(E, W, Ef):
    """Moller energy-loss CDF at incident energy E (closed form)."""
    lo, hi = WC / E, 0.5 * (E - Ef) / E
    return (moller_G(W / E) - moller_G(lo)) / (moller_G(hi) - moller_G(lo))


def moller_dcs(E, W, nfree):
    eps = W / E
    return nfree * PI_EFOUR * (1.0 / eps ** 2 + 1.0 / (1.0 - eps) ** 2 - 1.0 / (eps * (1.0 - eps))) / E ** 3


def gryz_dcs(E, W, ns, Eb):
    eps, epsB = W / E, Eb / E
    return (ns * PI_EFOUR * eps * (1.0 + epsB) ** -1.5 * (1.0 - eps) ** (epsB / (epsB + eps))
            * ((1.0 - epsB) + 4.0 * epsB * np.log(2.7 + ((1.0 - eps) / epsB) ** 0.5) / (3.0 * eps)) / W ** 3)


def gl_integral(f, E, lo, hi):
    """Integral of f(E, W) over W in [lo, hi] per row (arrays), Gauss-Legendre
    in t = ln W (the integrands fall like W^-2)."""
    E, lo, hi = np.broadcast_arrays(np.asarray(E, float), np.asarray(lo, float), np.asarray(hi, float))
    a, b = np.log(lo)[:, None], np.log(hi)[:, None]
    t = 0.5 * (b - a) * _GL_X[None, :] + 0.5 * (b + a)
    W = np.exp(t)
    return 0.5 * (b - a)[:, 0] * (f(E[:, None], W) * W * _GL_W[None, :]).sum(axis=1)


def gryz_cdf(E, W, ns, Eb, Ef):
    f = lambda e, w: gryz_dcs(e, w, ns, Eb)  # noqa: E731
    hi = 0.5 * (E - Ef)
    return gl_integral(f, E, np.full_like(E, Eb), W) / gl_integral(f, E, np.full_like(E, Eb), hi)


def invert(cdf, E, lo, hi, u, iters=60):
    """W with cdf(E, W) = u by bisection (vectorised)."""
    a, b = np.array(lo, float), np.array(hi, float)
    for _ in range(iters):
        m = 0.5 * (a + b)
        low = cdf(E, m) < u
        a, b = np.where(low, m, a), np.where(low, b, m)
    return 0.5 * (a + b)


def theta_deg(c2):
    """thetaFromCos (testThings.py:153-163): degrees from cos^2(theta/2)."""
    return np.degrees(2.0 * np.arccos(np.sqrt(np.clip(c2, 0.0, 1.0))))


def binary_coll(E, W):
    return 0.5 * (np.sqrt(1.0 - W / E) + 1.0)


def watson_p(u2):
    """Asymptotic upper-tail probability of the two-sample Watson U^2."""
    k = np.arange(1, 50)
    return float(np.clip(2.0 * np.sum((-1.0) ** (k - 1) * np.exp(-2.0 * k * k * np.pi ** 2 * u2)), 0.0, 1.0))


def assert_watson(theory, mc):
    u2 = watson_u2(theory, mc)
    print(f"Watson U^2 = {u2:.4f} (reference critical value {U2_CRIT_REF}), p = {watson_p(u2):.3g}")
    assert watson_p(u2) > 1e-3


def test_watson_tail_matches_critical_values():
    """The series gives the tabulated critical values (Mardia & Jupp,
    Appendix 2.14): 0.152 at p = 0.10, 0.187 at 0.05, 0.268 at 0.01."""
    for x, p in ((0.152, 0.10), (0.187, 0.05), (0.268, 0.01)):
        assert abs(watson_p(x) - p) < 0.003 * max(1.0, p / 0.01)


def ks_uniform(u):
    return stats.kstest(np.asarray(u, float), "uniform").pvalue


def mc_sample(rows, n=200, seed=0):
    rng = np.random.default_rng(seed)
    return rows[rng.choice(len(rows), size=min(n, len(rows)), replace=False)]


def channel_rows(rows, ch, at_E0=False):
    r = rows[rows["channel"] == ch]
    return r[r["step"] == 0] if at_E0 else r


def last_shell(mat):
    """The channel and (E_b, n_s) of the shell with the smallest binding
    energy: the shell the reference tests (testThings.py:686-689) and, for
    Al, the outermost shell, whose table serves every Gryzinski event
    (scattering.py:625)."""
    from mcd import _abi
    shells = mat.shells
    i = min(range(len(shells)), key=lambda k: shells[k][1])
    return _abi.CH_GRYZ0 + i, shells[i][1], shells[i][2]


def test_gauss_legendre_cdf_matches_moller_closed_form():
    """The quadrature used for the Gryzinski CDF reproduces the closed-form
    Moller CDF (pins the integrator, not the GPU)."""
    E = np.array([15000.0, 17500.0, 20000.0])
    Ef = 11.6587
    W = np.array([12.0, 150.0, 4000.0])
    f = lambda e, w: moller_dcs(e, w, 3)  # noqa: E731
    gl = gl_integral(f, E, np.full(3, WC), W) / gl_integral(f, E, np.full(3, WC), 0.5 * (E - Ef))
    assert np.allclose(gl, moller_cdf(E, W, Ef), rtol=0, atol=1e-10)


def test_rutherford_polar_angle_all_energies(al20):
    """TestPolarProb_Ruth (testThings.py:437-488)."""
    from mcd import _abi
    rows = channel_rows(al20["rows"], _abi.CH_RUTH)
    Z = al20["mat"].params["Z"]
    E = rows["E"].astype(np.float64)
    a = 3.4 * Z ** 0.666667 / E
    pol = rows["pol"].astype(np.float64)
    u = (1.0 + a) * (1.0 - pol) / (a + 1.0 - pol)   # Ruth_diffCS_E._cdf with cos = 2 pol - 1
    assert len(u) > (100_000 if al20["big"] else 50_000)
    assert ks_uniform(u) > 1e-3
    mc = mc_sample(rows, seed=1)
    rng = np.random.default_rng(2)
    Et = mc_sample(rows, seed=3)["E"].astype(np.float64)
    at = 3.4 * Z ** 0.666667 / Et
    x = rng.random(len(Et))
    theory = np.degrees(np.arccos(1.0 - 2.0 * at * x / (1.0 + at - x)))  # Ruth_diffCS_E._ppf
    assert_watson(theory, theta_deg(mc["pol"].astype(np.float64)))


@pytest.mark.parametrize("kind", ["Ruth", "Moller", "Gryz"])
def test_azimuth_is_uniform(al20, kind):
    """TestAzimProb_Ruth / _Moller / _Gryz (testThings.py:492-524, 640-662):
    phi/2 uniform on [0, pi)."""
    from mcd import _abi
    ch = {"Ruth": _abi.CH_RUTH, "Moller": _abi.CH_MOLLER, "Gryz": last_shell(al20["mat"])[0]}[kind]
    rows = channel_rows(al20["rows"], ch)
    az = rows["az"].astype(np.float64)
    assert az.min() >= 0.0 and az.max() < np.pi
    assert ks_uniform(az / np.pi) > 1e-3
    mc = np.degrees(2.0 * mc_sample(rows, seed=4)["az"].astype(np.float64))
    theory = np.random.default_rng(5).random(len(mc)) * 360.0
    assert_watson(theory, mc)


@pytest.mark.parametrize("kind", ["Moller", "Gryz"])
def test_binary_collision_polar_angle(al20, kind):
    """Polar angle of every Moller / Gryzinski row from its energy loss
    (binaryCollModel, scattering.py:156-180)."""
    from mcd import _abi
    ch = _abi.CH_MOLLER if kind == "Moller" else last_shell(al20["mat"])[0]
    rows = channel_rows(al20["rows"], ch)
    c2 = binary_coll(rows["E"].astype(np.float64), rows["E_loss"].astype(np.float64))
    assert np.abs(rows["pol"] - c2).max() < 2e-6


def _loss_integral(kind, mat, E, a, b):
    """Integral of the energy-loss dCS over [a, b] at energies E (arrays)."""
    if kind == "Moller":
        return (moller_G(b / E) - moller_G(a / E)) * 3 * PI_EFOUR / E ** 2
    _, Eb, ns = last_shell(mat)
    return gl_integral(lambda e, w: gryz_dcs(e, w, ns, Eb), E, a, b)


def _loss_cdf(kind, mat, E, W):
    """Continuous energy-loss CDF at E on [W_min, (E - E_F)/2]."""
    lo, hi = _loss_range(kind, mat, E)
    return _loss_integral(kind, mat, E, lo, W) / _loss_integral(kind, mat, E, lo, hi)


def _loss_range(kind, mat, E):
    lo = WC if kind == "Moller" else last_shell(mat)[1]
    return np.full_like(E, lo), 0.5 * (E - mat.fermi_e)


def table_pit(kind, mat, grid, E, W, rng, tail=0.0):
    """Randomised probability integral transform of table-sampled losses.

    The transport takes row i = the first grid energy >= E (oracle_table_row)
    and returns W = Ws[j], the lower node of the bin whose cumulative value
    first reaches u (pickTable, scattering.py:39-59); bins whose upper node
    exceeds (E_i - E_F)/2 are masked and the row renormalised
    (probTables.py:429-460).  So P(W = Ws[j]) = F_i(Ws[j+1]) - F_i(Ws[j]), F_i
    the exact CDF at E_i truncated at the last valid node, and
    u = F_i(Ws[j]) + V (F_i(Ws[j+1]) - F_i(Ws[j])), V uniform, is uniform.
    tail > 0: only the draws at nodes >= tail * E_i, through the conditional
    law (F - F_c) / (1 - F_c), F_c = F_i at the first such node; returns the
    transforms of those draws."""
    Es = np.asarray(grid.Es, np.float32)
    Ws = np.asarray(grid.Ws, np.float32)
    i = np.searchsorted(Es.astype(np.float64), E, side="left")
    j = np.searchsorted(Ws, W.astype(np.float32))
    assert (i < len(Es)).all() and (Ws[np.minimum(j, len(Ws) - 1)] == W.astype(np.float32)).all()
    Ei = Es[i].astype(np.float64)
    wmax = (Es[i] - np.float32(mat.fermi_e)) * np.float32(0.5)
    n_valid = np.searchsorted(Ws, wmax, side="right") - 1   # bins with upper node <= wmax
    assert (j < n_valid).all()
    W64 = Ws.astype(np.float64)
    lo = np.full_like(Ei, W64[0])
    tot = _loss_integral(kind, mat, Ei, lo, W64[n_valid])
    F0 = _loss_integral(kind, mat, Ei, lo, np.maximum(W64[j], W64[0] * (1 + 1e-15))) / tot
    F1 = _loss_integral(kind, mat, Ei, lo, W64[j + 1]) / tot
    F0 = np.where(j == 0, 0.0, F0)
    u = F0 + rng.random(len(Ei)) * (F1 - F0)
    if tail <= 0.0:
        return u
    jc = np.searchsorted(W64, tail * Ei, side="left")
    keep = (j >= jc) & (jc > 0) & (jc < n_valid)
    Fc = _loss_integral(kind, mat, Ei[keep], lo[keep], W64[jc[keep]]) / tot[keep]
    return (u[keep] - Fc) / (1.0 - Fc)


@pytest.mark.parametrize("kind", ["Moller", "Gryz"])
@pytest.mark.parametrize("at_E0", [True, False], ids=["E0", "all_E"])
def test_energy_loss_and_polar_angle(al20, kind, at_E0):
    """TestPolarProb_Moller_E0 / _Moller / _Gryz_E0 / _Gryz
    (testThings.py:529-639, 664-800): energy losses through the theoretical
    distribution of the table draw at each row's energy; polar angles
    against the binary-collision angles of theoretical energy losses."""
    from mcd import _abi
    mat = al20["mat"]
    ch = _abi.CH_MOLLER if kind == "Moller" else last_shell(mat)[0]
    grid = al20["grids"][0 if kind == "Moller" else 1]
    rows = channel_rows(al20["rows"], ch, at_E0=at_E0)
    if at_E0:
        assert (rows["E"] == np.float32(E0)).all()
    E = rows["E"].astype(np.float64)
    W = rows["E_loss"].astype(np.float64)
    lo, hi = _loss_range(kind, mat, E)
    assert (W >= lo * (1 - 1e-6)).all() and (W <= hi * (1 + 1e-6)).all()
    u = table_pit(kind, mat, grid, E, W, np.random.default_rng(10))
    assert len(u) > (300 if at_E0 else 20_000) // (1 if al20["big"] else 2)
    assert ks_uniform(u) > 1e-3
    mc = mc_sample(rows, seed=6)
    Et = mc_sample(rows, seed=7)["E"].astype(np.float64)
    x = np.random.default_rng(8).random(len(Et))
    lo_t, hi_t = _loss_range(kind, mat, Et)
    Wt = invert(lambda e, w: _loss_cdf(kind, mat, e, w), Et, lo_t, hi_t, x)
    theory = theta_deg(binary_coll(Et, Wt))
    assert_watson(theory, theta_deg(mc["pol"].astype(np.float64)))


@pytest.mark.parametrize("kind", ["Moller", "Gryz"])
def test_energy_loss_tail(al20, kind):
    """The large-loss tail, W >= 0.1 E, through its conditional law.  Near
    W_min the 1/W^2 term dominates every row of the full-range check; in the
    tail the (1 - e)^-2 and interference terms of moller_dCS shape the density
    (on the oracle's rows, dropping either gives p < 1e-4 at 1000 tail draws)
    and so does the log term of gryz_dCS.  Its (1 - e)^(e_B/(e_B + e)) factor
    moves the tail density by under 3% here and is not resolved by this
    test."""
    from mcd import _abi
    mat = al20["mat"]
    ch = _abi.CH_MOLLER if kind == "Moller" else last_shell(mat)[0]
    grid = al20["grids"][0 if kind == "Moller" else 1]
    rows = channel_rows(al20["rows"], ch)
    u = table_pit(kind, mat, grid, rows["E"].astype(np.float64), rows["E_loss"].astype(np.float64),
                  np.random.default_rng(11), tail=0.1)
    assert len(u) > (800 if al20["big"] else 200)
    assert ks_uniform(u) > 1e-3


def test_channel_frequencies_at_all_energies(al20):
    """Test_crossSection (testThings.py:321-390) extended from E0 to every
    energy: channel counts of 20,000 rows against the sum of their per-row
    channel probabilities from the oracle's sigma set, chi-square p > 1e-3."""
    import oracle as O
    rows = mc_sample(al20["rows"], n=20_000, seed=9)
    cfg = al20["cfg"]
    expect = {}
    for E in rows["E"].astype(np.float64):
        sig, lab = O.sigma_set(cfg, float(E))
        for s, l in zip(sig / sig.sum(), lab):
            expect[int(l)] = expect.get(int(l), 0.0) + s
    labs = sorted(expect)
    obs = np.array([(rows["channel"] == l).sum() for l in labs], float)
    exp = np.array([expect[l] for l in labs])
    keep = exp >= 5
    assert obs[~keep].sum() <= 5
    chi2 = (((obs - exp) ** 2 / exp)[keep]).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3
