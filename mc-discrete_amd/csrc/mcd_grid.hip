// mcd_grid.hip — the tolerance-driven search for the energy-loss table grids
// (host code of libmcd.so; the CDF rows are filled on the device by
// launch_build_cdf).
//
// The search of the reference (MC/tools/probTables.py):
//   findWseries (:280-328): W grid = float32 geomspace(Wmin, W_max(E0), n) with
//     n the first of 100, 1000, ... 1e6 (ordersOfMag, :54-58), refined over
//     10 %-steps of that n (percents, :61-64), such that the trapezoid integral
//     of dsigma/dW on the grid (float64 integrand, float32 grid differences,
//     check_intTol :165-188) is within tol_W of the reference integral;
//   findEseries (:332-382): E grid = float32 geomspace, then linspace, over
//     [Emin', E0] (Emin' = max(Emin, E_b) for Gryzinski), the first size for
//     which the float32 trapezoid integral at the second-largest grid energy
//     over the W nodes below W_max(E) is within tol_E of the reference
//     integral at E0 (check_CDFTol :191-232).
// The reference integral is scipy.integrate.quad (QUADPACK); here it is an
// adaptive Gauss-Kronrod (7/15-point) integration in FP64 to 1e-13 relative.
// Grids are built as numpy builds them: geomspace and linspace evaluated in
// float64 (log10, step = (stop - start)/(n - 1), y_i = i*step + start,
// pow(10, y_i), exact end points) and rounded to float32; sums in numpy's
// pairwise order (8 accumulators over blocks of up to 128 elements).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mcd_internal.h"

#pragma clang fp contract(off)

namespace mcd {

namespace {

constexpr double kPiEFour = 6.51408491409531e-14;  // parameters.py:7 (cm^2 eV^2)

template <class T>
T pairwise_sum(const T* a, size_t n) {
  if (n < 8) {
    T res = 0;
    for (size_t i = 0; i < n; ++i) res += a[i];
    return res;
  }
  if (n <= 128) {
    T r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    size_t i;
    for (i = 8; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    T res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  size_t n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
}

std::vector<double> linspace64(double start, double stop, int n) {
  std::vector<double> y((size_t)std::max(n, 0));
  if (n <= 0) return y;
  const double step = (stop - start) / (double)(n - 1);
  for (int i = 0; i < n; ++i) {
    double v = (double)i;
    v *= step;
    v += start;
    y[i] = v;
  }
  if (n > 1) y[n - 1] = stop;
  return y;
}

std::vector<float> geomspace32(double start, double stop, int n) {
  std::vector<double> y = linspace64(log10(start), log10(stop), n);
  std::vector<float> out(y.size());
  for (size_t i = 0; i < y.size(); ++i) out[i] = (float)pow(10.0, y[i]);
  if (n > 0) out[0] = (float)start;
  if (n > 1) out[n - 1] = (float)stop;
  return out;
}

std::vector<float> linspace32(double start, double stop, int n) {
  std::vector<double> y = linspace64(start, stop, n);
  return std::vector<float>(y.begin(), y.end());
}

// ordersOfMag (:54-58): geomspace(100, max_size, 5) truncated to int
std::vector<int> orders_of_mag(int max_size) {
  std::vector<double> y = linspace64(log10(100.0), log10((double)max_size), 5);
  std::vector<int> out;
  for (size_t i = 0; i < y.size(); ++i) {
    double v = pow(10.0, y[i]);
    if (i == 0) v = 100.0;
    if (i + 1 == y.size()) v = (double)max_size;
    out.push_back((int)v);
  }
  return out;
}

// percents (:61-64): linspace(max_size/10, max_size, 10, dtype=float32) truncated to int
std::vector<int> percents(int max_size) {
  std::vector<float> y = linspace32((double)max_size / 10.0, (double)max_size, 10);
  std::vector<int> out;
  for (float v : y) out.push_back((int)v);
  return out;
}

struct Spec {
  int kind;  // MCD_TABLE_MOLLER or MCD_TABLE_GRYZ
  double Wmin, n_el, Eb, Ef, Emin, Emax;
  double Wmax(double E) const { return (E - Ef) * 0.5; }
};

// extFunctions.moller_dCS (:159-184) / gryz_dCS (:189-219) in float64, in
// the order numpy evaluates the reference expressions
double f64(const Spec& s, double E, double W) {
  const double eps = W / E;
  if (s.kind == MCD_TABLE_MOLLER) {
    if (!(eps < 1.0)) return 0.0;
    const double om = 1.0 - eps;
    const double br = (1.0 / (eps * eps) + 1.0 / (om * om)) - 1.0 / (eps * om);
    return s.n_el * kPiEFour * br / pow(E, 3.0);
  }
  const double epsB = s.Eb / E;
  if (!(eps < 1.0 && epsB < 1.0)) return 0.0;
  const double a = pow(1.0 + epsB, -1.5);
  const double b = pow(1.0 - eps, epsB / (epsB + eps));
  const double c = (1.0 - epsB) + 4.0 * epsB * log(2.7 + sqrt((1.0 - eps) / epsB)) / (3.0 * eps);
  double v = s.n_el * kPiEFour * eps;
  v *= a;
  v *= b;
  v *= c;
  return v / pow(W, 3.0);
}

// probTable.f32_array semantics: float32 energy scalar, float32 loss array,
// float32 arithmetic with float64 scalars rounded to float32 where they meet it
void f32_array(const Spec& s, float E32, const float* W, size_t n, float* out) {
  if (s.kind == MCD_TABLE_MOLLER) {
    const float k = (float)(s.n_el * kPiEFour);
    const float e3 = (float)pow((double)E32, 3.0);
    for (size_t i = 0; i < n; ++i) {
      const float eps = W[i] / E32;
      const float om = 1.0f - eps;
      const float br = (1.0f / (eps * eps) + 1.0f / (om * om)) - 1.0f / (eps * om);
      const float val = k * br / e3;
      out[i] = eps < 1.0f ? val : 0.0f;
    }
    return;
  }
  const double epsB = s.Eb / (double)E32;
  const float a = (float)pow(1.0 + epsB, -1.5);
  const float epsBf = (float)epsB;
  const float c0 = (float)(1.0 - epsB), c1 = (float)(4.0 * epsB);
  const float k = (float)(s.n_el * kPiEFour);
  for (size_t i = 0; i < n; ++i) {
    const float eps = W[i] / E32;
    const float expo = epsBf / (epsBf + eps);
    const float inner = sqrtf((1.0f - eps) / epsBf);
    const float c = c0 + c1 * logf(2.7f + inner) / (3.0f * eps);
    float v = k * eps;
    v *= a;
    v *= powf(1.0f - eps, expo);
    v *= c;
    const float w3 = powf(W[i], 3.0f);
    v /= w3;
    out[i] = (eps < 1.0f && epsB < 1.0) ? v : 0.0f;
  }
}

// adaptive Gauss-Kronrod 7/15 integration of f64(s, E, .) over [a, b]
double gk15(const Spec& s, double E, double a, double b, double* err) {
  static const double xk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
  static const double wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                               0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                               0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                               0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
  static const double wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                               0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  const double fc = f64(s, E, c);
  double rk = fc * wk[7], rg = fc * wg[3];
  for (int j = 0; j < 7; ++j) {
    const double x = h * xk[j];
    const double f = f64(s, E, c - x) + f64(s, E, c + x);
    rk += wk[j] * f;
    if (j % 2 == 1) rg += wg[j / 2] * f;
  }
  *err = fabs((rk - rg) * h);
  return rk * h;
}

double integrate(const Spec& s, double E, double a, double b) {
  struct Seg { double a, b, v, e; };
  std::vector<Seg> segs;
  double e0;
  // start from 8 panels uniform in ln W: the integrands fall like W^-2
  const int n0 = 8;
  const double la = log(a), lb = log(b);
  for (int i = 0; i < n0; ++i) {
    const double x0 = i == 0 ? a : exp(la + (lb - la) * i / n0);
    const double x1 = i == n0 - 1 ? b : exp(la + (lb - la) * (i + 1) / n0);
    const double v = gk15(s, E, x0, x1, &e0);
    segs.push_back({x0, x1, v, e0});
  }
  for (int it = 0; it < 4000; ++it) {
    double tot = 0., err = 0.;
    size_t worst = 0;
    for (size_t i = 0; i < segs.size(); ++i) {
      tot += segs[i].v;
      err += segs[i].e;
      if (segs[i].e > segs[worst].e) worst = i;
    }
    if (err <= 1e-13 * fabs(tot)) break;
    const Seg w = segs[worst];
    const double m = 0.5 * (w.a + w.b);
    double e1, e2;
    const double v1 = gk15(s, E, w.a, m, &e1), v2 = gk15(s, E, m, w.b, &e2);
    segs[worst] = {w.a, m, v1, e1};
    segs.push_back({m, w.b, v2, e2});
  }
  std::vector<double> v;
  for (const Seg& g : segs) v.push_back(g.v);
  std::sort(v.begin(), v.end());
  double tot = 0.;
  for (double x : v) tot += x;
  return tot;
}

// check_intTol (:165-188)
bool check_int_tol(const Spec& s, const std::vector<float>& series, double ref, double E, double tol_W) {
  const size_t n = series.size();
  std::vector<double> t(n - 1);
  double ylo = f64(s, E, (double)series[0]);
  for (size_t i = 0; i + 1 < n; ++i) {
    const double yhi = f64(s, E, (double)series[i + 1]);
    const double d = (double)(float)(series[i + 1] - series[i]);
    t[i] = d * (yhi + ylo) / 2.0;
    ylo = yhi;
  }
  const double trap = pairwise_sum(t.data(), t.size());
  return fabs(trap - ref) / ref < tol_W;
}

// check_CDFTol (:191-232)
bool check_cdf_tol(const Spec& s, const std::vector<float>& Es, const std::vector<float>& Ws, double ref,
                   double tol_E) {
  const float E = Es[Es.size() - 2];
  const float wmax = (float)(((double)E - s.Ef) * 0.5);
  std::vector<float> sel;
  for (float w : Ws)
    if (w <= wmax) sel.push_back(w);
  if (sel.size() < 2) return false;
  std::vector<float> y(sel.size());
  f32_array(s, E, sel.data(), sel.size(), y.data());
  std::vector<float> t(sel.size() - 1);
  for (size_t i = 0; i + 1 < sel.size(); ++i) {
    const float d = sel[i + 1] - sel[i];
    t[i] = d * (y[i + 1] + y[i]) / 2.0f;
  }
  const float tot = pairwise_sum(t.data(), t.size());
  return fabs((double)tot - ref) / ref < tol_E;
}

}  // namespace

int search_grid(int kind, double Wmin, double n_el, double Eb, double Ef, double Emin, double Emax, double tol_E,
                double tol_W, std::vector<float>& Es, std::vector<float>& Ws, char* err, size_t err_len) {
  Spec s{kind, Wmin, n_el, Eb, Ef, Emin, Emax};
  const double E = Emax, W_max = s.Wmax(E);
  if (!(W_max > Wmin)) {
    snprintf(err, err_len, "table %d: W range [%g, %g] is empty", kind, Wmin, W_max);
    return MCD_ERR_TABLE;
  }
  const double ref = integrate(s, E, Wmin, W_max);
  int n1 = 0;
  for (int n : orders_of_mag(1000000))
    if (check_int_tol(s, geomspace32(Wmin, W_max, n), ref, E, tol_W)) {
      n1 = n;
      break;
    }
  if (!n1) {
    snprintf(err, err_len, "table %d: no W grid up to 1e6 points meets tol_W = %g", kind, tol_W);
    return MCD_ERR_TABLE;
  }
  Ws.clear();
  for (int n : percents(n1)) {
    std::vector<float> g = geomspace32(Wmin, W_max, n);
    if (check_int_tol(s, g, ref, E, tol_W)) {
      Ws = g;
      break;
    }
  }
  if (Ws.empty()) {
    snprintf(err, err_len, "table %d: W refinement failed", kind);
    return MCD_ERR_TABLE;
  }
  for (float w : Ws)
    if (!(w < Emax)) {
      snprintf(err, err_len, "table %d: energy-loss grid reaches the incident energy (probTables.py:396)", kind);
      return MCD_ERR_TABLE;
    }
  // findEseries: the reference integral at E0 over [Wmin, W_max(E0)]
  n1 = 0;
  for (int n : orders_of_mag(1000000))
    if (check_cdf_tol(s, geomspace32(Emin, Emax, n), Ws, ref, tol_E)) {
      n1 = n;
      break;
    }
  if (!n1) {
    snprintf(err, err_len, "table %d: no E grid up to 1e6 points meets tol_E = %g", kind, tol_E);
    return MCD_ERR_TABLE;
  }
  Es.clear();
  for (int n : percents(n1)) {
    std::vector<float> g = linspace32(Emin, Emax, n);
    if (check_cdf_tol(s, g, Ws, ref, tol_E)) {
      Es = g;
      break;
    }
  }
  if (Es.empty()) {
    snprintf(err, err_len, "table %d: E refinement failed", kind);
    return MCD_ERR_TABLE;
  }
  return MCD_OK;
}

double grid_reference_integral(int kind, double Wmin, double n_el, double Eb, double Ef, double E) {
  Spec s{kind, Wmin, n_el, Eb, Ef, 0., E};
  return integrate(s, E, Wmin, s.Wmax(E));
}

}  // namespace mcd
