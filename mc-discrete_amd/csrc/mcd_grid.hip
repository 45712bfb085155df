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
  __host__ __device__ double Wmax(double E) const { return (E - Ef) * 0.5; }
};

// extFunctions.moller_dCS (:159-184) / gryz_dCS (:189-219) in float64, in
// the order numpy evaluates the reference expressions (host and device search)
__host__ __device__ double f64(const Spec& s, double E, double W) {
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

// probTable.f32_array semantics: float32 energy scalar, float32 loss values,
// float32 arithmetic with float64 scalars rounded to float32 where they meet
// it.  The per-energy constants once, then one loss value at a time.
struct F32Row {
  bool moller, epsB_lt1;
  float E32, k, e3, a, epsBf, c0, c1;
  __host__ __device__ F32Row(const Spec& s, float E) : moller(s.kind == MCD_TABLE_MOLLER), E32(E) {
    k = (float)(s.n_el * kPiEFour);
    if (moller) {
      e3 = (float)pow((double)E32, 3.0);
      epsB_lt1 = true;
      a = epsBf = c0 = c1 = 0.f;
      return;
    }
    const double epsB = s.Eb / (double)E32;
    epsB_lt1 = epsB < 1.0;
    a = (float)pow(1.0 + epsB, -1.5);
    epsBf = (float)epsB;
    c0 = (float)(1.0 - epsB);
    c1 = (float)(4.0 * epsB);
    e3 = 0.f;
  }
  __host__ __device__ float operator()(float W) const {
    const float eps = W / E32;
    if (moller) {
      const float om = 1.0f - eps;
      const float br = (1.0f / (eps * eps) + 1.0f / (om * om)) - 1.0f / (eps * om);
      const float val = k * br / e3;
      return eps < 1.0f ? val : 0.0f;
    }
    const float expo = epsBf / (epsBf + eps);
    const float inner = sqrtf((1.0f - eps) / epsBf);
    const float c = c0 + c1 * logf(2.7f + inner) / (3.0f * eps);
    float v = k * eps;
    v *= a;
    v *= powf(1.0f - eps, expo);
    v *= c;
    const float w3 = powf(W, 3.0f);
    v /= w3;
    return (eps < 1.0f && epsB_lt1) ? v : 0.0f;
  }
};

void f32_array(const Spec& s, float E32, const float* W, size_t n, float* out) {
  const F32Row row(s, E32);
  for (size_t i = 0; i < n; ++i) out[i] = row(W[i]);
}

// Gauss-Kronrod 7/15 nodes and weights (QUADPACK qk15)
constexpr double kGkXk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                             0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                             0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                             0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
constexpr double kGkWk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                             0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                             0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                             0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
constexpr double kGkWg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                             0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
__host__ __device__ inline double gk_xk(int j) { return kGkXk[j]; }
__host__ __device__ inline double gk_wk(int j) { return kGkWk[j]; }
__host__ __device__ inline double gk_wg(int j) { return kGkWg[j]; }
// The 15 abscissae of a panel: q = 0 the centre, 1..7 c - h x_j, 8..14 c + h x_j.
__host__ __device__ inline double gk_point(double a, double b, int q) {
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  if (q == 0) return c;
  const double x = h * gk_xk(q <= 7 ? q - 1 : q - 8);
  return q <= 7 ? c - x : c + x;
}
// The panel's Kronrod value and error estimate from f at its 15 points.
__host__ __device__ inline double gk_combine(const double* f, double a, double b, double* err) {
  const double h = 0.5 * (b - a);
  const double fc = f[0];
  double rk = fc * gk_wk(7), rg = fc * gk_wg(3);
  for (int j = 0; j < 7; ++j) {
    const double fs = f[1 + j] + f[8 + j];
    rk += gk_wk(j) * fs;
    if (j % 2 == 1) rg += gk_wg(j / 2) * fs;
  }
  *err = fabs((rk - rg) * h);
  return rk * h;
}

// adaptive Gauss-Kronrod 7/15 integration of f64(s, E, .) over [a, b]
double gk15(const Spec& s, double E, double a, double b, double* err) {
  double f[15];
  for (int q = 0; q < 15; ++q) f[q] = f64(s, E, gk_point(a, b, q));
  return gk_combine(f, a, b, err);
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

// ---------------------------------------------------------------------------
// The same search on the device, for many tables in one launch: one
// workgroup per table runs search_grid's steps in order (reference integral,
// W sizes by orders of magnitude then by tenths, E sizes likewise) without
// returning to the host.  Threads share the work of each step: the 15 points
// of each Gauss-Kronrod panel, and the leaf blocks of numpy's pairwise sum of
// a candidate grid's trapezoids.  Every quantity the decisions depend on is
// accumulated in the host search's order (one thread walks the segments, the
// sort, and the pairwise tree over the leaf sums), so the two searches pick
// the same sizes; the device's libm (pow, log, exp) may differ from the
// host's in the last bit, which moves a tolerance test only at an exact tie.
// The grids themselves are materialised on the host from the sizes.

namespace {

constexpr int kGsBlock = 1024;
constexpr int kGsMaxSeg = 8 + 4000 + 8;
constexpr int kGsMaxTerms = 1000000;                 // a candidate grid has at most 1e6 points
constexpr int kGsMaxLeaves = kGsMaxTerms / 64 + 16;  // pairwise leaves hold >= 64 terms past the first split
constexpr int kGsChunk = 16384;                      // grid points evaluated per pass

struct GsScratch {
  double seg[kGsMaxSeg][4];  // a, b, value, error
  double vals[kGsMaxSeg];
  uint32_t leaf_off[kGsMaxLeaves + 1];
  double leaf_sum[kGsMaxLeaves];
  float x[kGsChunk];         // grid values of the current pass
  double y[kGsChunk];        // integrand at them (float32 values in the first half for a float sum)
};

constexpr int kPwDepth = 32;  // the pairwise recursion over < 2^20 terms is < 20 levels deep

struct GsShared {
  double f[128];       // Gauss-Kronrod point values
  double pan[2][2];    // the two halves of the segment being split
  double dres;         // broadcast value
  float fres;
  int flag, n_leaves, l1;
  // thread 0's walks of the pairwise recursion
  uint32_t so[kPwDepth], sn[kPwDepth];
  double fl[kPwDepth];
  uint8_t st[kPwDepth];
};

// numpy's pairwise summation blocks (pairwise_sum above): the leaves of the
// recursion, left to right, as offsets; off[L] = n.  One thread.
__device__ int pw_leaves(uint32_t n, uint32_t* off, GsShared* sh) {
  uint32_t* so = sh->so;
  uint32_t* sn = sh->sn;
  int sp = 1, L = 0;
  so[0] = 0;
  sn[0] = n;
  while (sp) {
    --sp;
    const uint32_t o = so[sp], m = sn[sp];
    if (m <= 128) {
      off[L++] = o;
      continue;
    }
    uint32_t n2 = m / 2;
    n2 -= n2 % 8;
    so[sp] = o + n2;
    sn[sp] = m - n2;
    ++sp;
    so[sp] = o;
    sn[sp] = n2;
    ++sp;
  }
  off[L] = n;
  return L;
}

// The pairwise tree over the leaf sums, in the recursion's order.  One thread.
template <class T>
__device__ T pw_combine(uint32_t n, const double* leaf, GsShared* sh) {
  uint32_t* fm = sh->so;
  double* fl = sh->fl;  // a float partial sum is exact in a double slot
  uint8_t* st = sh->st;
  int sp = 1, k = 0;
  fm[0] = n;
  st[0] = 0;
  bool have = false;
  T ret = 0;
  while (true) {
    const int f = sp - 1;
    if (have) {
      if (st[f] == 1) {  // left value in: descend right
        fl[f] = (double)ret;
        st[f] = 2;
        have = false;
        uint32_t n2 = fm[f] / 2;
        n2 -= n2 % 8;
        fm[sp] = fm[f] - n2;
        st[sp] = 0;
        ++sp;
        continue;
      }
      ret = (T)fl[f] + ret;  // pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2)
      if (--sp == 0) return ret;
      continue;
    }
    if (fm[f] <= 128) {
      ret = (T)leaf[k++];
      if (--sp == 0) return ret;
      have = true;
      continue;
    }
    st[f] = 1;
    uint32_t n2 = fm[f] / 2;
    n2 -= n2 % 8;
    fm[sp] = n2;
    st[sp] = 0;
    ++sp;
  }
}

// Trapezoid k of a pass: T-typed width (check_intTol: the float32 width in
// float64; check_CDFTol: float32) times the mean of the two values.
template <class T>
__device__ __forceinline__ T trapezoid(const float* x, const T* y, int k) {
  const T d = (T)(x[k + 1] - x[k]);
  return d * (y[k + 1] + y[k]) / (T)2;
}

// A leaf's sum as pairwise_sum forms it (sequential below 8 terms, else 8
// accumulators), its terms k0 .. k0 + m - 1 of the pass.
template <class T>
__device__ T leaf_sum(const float* x, const T* y, int k0, int m) {
  if (m < 8) {
    T res = 0;
    for (int q = 0; q < m; ++q) res += trapezoid(x, y, k0 + q);
    return res;
  }
  T r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = trapezoid(x, y, k0 + j);
  const int full = m - m % 8;
  int q = 8;
  for (; q < full; q += 8)
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] += trapezoid(x, y, k0 + q + j);
  T res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; q < m; ++q) res += trapezoid(x, y, k0 + q);
  return res;
}

// Element i of geomspace32 / linspace32 (n points), as those build it.
struct GeomGrid {
  double start, stop, lo, step;
  int n;
  __device__ GeomGrid(double a, double b, int n_) : start(a), stop(b), n(n_) {
    lo = log10(a);
    step = (log10(b) - lo) / (double)(n - 1);
  }
  __device__ float at(int i) const {
    if (i == 0) return (float)start;
    if (i == n - 1) return (float)stop;
    double v = (double)i;
    v *= step;
    v += lo;
    return (float)pow(10.0, v);
  }
};
__device__ float lin_at(double start, double stop, int n, int i) {
  if (i == n - 1) return (float)stop;
  const double step = (stop - start) / (double)(n - 1);
  double v = (double)i;
  v *= step;
  v += start;
  return (float)v;
}

// pairwise_sum of the n_pts - 1 trapezoids of fn over the grid points
// g.at(0 .. n_pts - 1), by the whole workgroup; every thread gets the value.
// Points are evaluated one per thread in passes of whole leaves.
template <class T, class Fn>
__device__ T block_trapezoids(int n_pts, const GeomGrid& g, const Fn& fn, GsScratch* sc, GsShared* sh) {
  const int t = threadIdx.x;
  const uint32_t n = (uint32_t)(n_pts - 1);
  if (t == 0) sh->n_leaves = pw_leaves(n, sc->leaf_off, sh);
  __syncthreads();
  const int L = sh->n_leaves;
  T* y = reinterpret_cast<T*>(sc->y);
  for (int l0 = 0; l0 < L;) {
    if (t == 0) {
      int l1 = l0 + 1;  // a leaf has at most 129 points
      while (l1 < L && (int)(sc->leaf_off[l1 + 1] - sc->leaf_off[l0]) + 1 <= kGsChunk) ++l1;
      sh->l1 = l1;
    }
    __syncthreads();
    const int l1 = sh->l1;
    const int p0 = (int)sc->leaf_off[l0], np = (int)sc->leaf_off[l1] - p0 + 1;
    for (int k = t; k < np; k += blockDim.x) {
      const float x = g.at(p0 + k);
      sc->x[k] = x;
      y[k] = fn(x);
    }
    __syncthreads();
    for (int l = l0 + t; l < l1; l += blockDim.x) {
      const int k0 = (int)sc->leaf_off[l] - p0;
      sc->leaf_sum[l] = (double)leaf_sum<T>(sc->x, y, k0, (int)(sc->leaf_off[l + 1] - sc->leaf_off[l]));
    }
    __syncthreads();
    l0 = l1;
  }
  if (t == 0) {
    const T v = pw_combine<T>(n, sc->leaf_sum, sh);
    if constexpr (sizeof(T) == 8) sh->dres = (double)v; else sh->fres = (float)v;
  }
  __syncthreads();
  T v;
  if constexpr (sizeof(T) == 8) v = (T)sh->dres; else v = (T)sh->fres;
  __syncthreads();  // the broadcast slots are free again
  return v;
}

// integrate() by the workgroup: the same panels, splits and final sort
__device__ double integrate_dev(const Spec& s, double E, double a, double b, GsScratch* sc, GsShared* sh) {
  const int t = threadIdx.x;
  const int n0 = 8;
  const double la = log(a), lb = log(b);
  auto edge = [&](int i) { return i == 0 ? a : (i == n0 ? b : exp(la + (lb - la) * i / n0)); };
  if (t < 15 * n0) {
    const int p = t / 15;
    sh->f[t] = f64(s, E, gk_point(edge(p), edge(p + 1), t % 15));
  }
  __syncthreads();
  int nseg = n0;
  if (t == 0)
    for (int p = 0; p < n0; ++p) {
      double e0;
      const double x0 = edge(p), x1 = edge(p + 1);
      const double v = gk_combine(sh->f + 15 * p, x0, x1, &e0);
      sc->seg[p][0] = x0; sc->seg[p][1] = x1; sc->seg[p][2] = v; sc->seg[p][3] = e0;
    }
  int worst = 0;
  for (int it = 0; it < 4000; ++it) {
    if (t == 0) {
      double tot = 0., err = 0.;
      worst = 0;
      for (int i = 0; i < nseg; ++i) {
        tot += sc->seg[i][2];
        err += sc->seg[i][3];
        if (sc->seg[i][3] > sc->seg[worst][3]) worst = i;
      }
      sh->flag = err <= 1e-13 * fabs(tot);
      const double wa = sc->seg[worst][0], wb = sc->seg[worst][1], m = 0.5 * (wa + wb);
      sh->pan[0][0] = wa; sh->pan[0][1] = m;
      sh->pan[1][0] = m; sh->pan[1][1] = wb;
    }
    __syncthreads();
    if (sh->flag) break;
    if (t < 30) sh->f[t] = f64(s, E, gk_point(sh->pan[t / 15][0], sh->pan[t / 15][1], t % 15));
    __syncthreads();
    if (t == 0) {
      double e1, e2;
      const double v1 = gk_combine(sh->f, sh->pan[0][0], sh->pan[0][1], &e1);
      const double v2 = gk_combine(sh->f + 15, sh->pan[1][0], sh->pan[1][1], &e2);
      sc->seg[worst][0] = sh->pan[0][0]; sc->seg[worst][1] = sh->pan[0][1];
      sc->seg[worst][2] = v1; sc->seg[worst][3] = e1;
      sc->seg[nseg][0] = sh->pan[1][0]; sc->seg[nseg][1] = sh->pan[1][1];
      sc->seg[nseg][2] = v2; sc->seg[nseg][3] = e2;
    }
    ++nseg;
  }
  if (t == 0) {
    // the values in ascending order (std::sort), summed in that order
    double* v = sc->vals;
    for (int i = 0; i < nseg; ++i) {
      const double x = sc->seg[i][2];
      int j = i;
      while (j > 0 && v[j - 1] > x) { v[j] = v[j - 1]; --j; }
      v[j] = x;
    }
    double tot = 0.;
    for (int i = 0; i < nseg; ++i) tot += v[i];
    sh->dres = tot;
  }
  __syncthreads();
  const double ref = sh->dres;
  __syncthreads();
  return ref;
}

__device__ bool check_int_dev(const Spec& s, double ref, double E, double tol_W, int n, GsScratch* sc, GsShared* sh) {
  auto fn = [&](float w) { return f64(s, E, (double)w); };
  const double trap = block_trapezoids<double>(n, GeomGrid(s.Wmin, s.Wmax(E), n), fn, sc, sh);
  return fabs(trap - ref) / ref < tol_W;
}

__device__ bool check_cdf_dev(const Spec& s, double ref, double tol_E, float E, const GeomGrid& Wg, GsScratch* sc,
                              GsShared* sh) {
  const float wmax = (float)(((double)E - s.Ef) * 0.5);
  int lo = 0, hi = Wg.n;  // W values at or below wmax: a prefix of the nondecreasing grid
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (Wg.at(mid) <= wmax) lo = mid + 1; else hi = mid;
  }
  if (lo < 2) return false;
  const F32Row row(s, E);
  auto fn = [&](float w) { return row(w); };
  GeomGrid sel = Wg;
  const float tot = block_trapezoids<float>(lo, sel, fn, sc, sh);
  return fabs((double)tot - ref) / ref < tol_E;
}

}  // namespace

__global__ __launch_bounds__(kGsBlock) void k_grid_search(const GridSearchJob* __restrict__ jobs, int n_jobs,
                                                         GridCandidates cand, double tol_E, double tol_W,
                                                         void* scratch, GridSearchResult* __restrict__ res) {
  __shared__ GsShared sh;
  const int j = blockIdx.x;
  if (j >= n_jobs) return;
  const GridSearchJob& job = jobs[j];
  GsScratch* sc = reinterpret_cast<GsScratch*>(scratch) + j;
  const Spec s{job.kind, job.Wmin, job.n_el, job.Eb, job.Ef, job.Emin, job.Emax};
  GridSearchResult r{0, 0, 0, 0, 0.};
  const double E = s.Emax, W_max = s.Wmax(E);
  if (!(W_max > s.Wmin)) {
    r.status = kGridEmptyW;
  } else {
    r.ref = integrate_dev(s, E, s.Wmin, W_max, sc, &sh);
    int oi = -1;
    for (int q = 0; q < kGridOrders && oi < 0; ++q)
      if (check_int_dev(s, r.ref, E, tol_W, cand.orders[q], sc, &sh)) oi = q;
    if (oi < 0) r.status = kGridNoW;
    for (int q = 0; q < kGridPercents && oi >= 0 && r.nW == 0; ++q)
      if (check_int_dev(s, r.ref, E, tol_W, cand.pct[oi][q], sc, &sh)) r.nW = cand.pct[oi][q];
    if (oi >= 0 && r.nW == 0) r.status = kGridWRefine;
    if (r.status == 0 && !((float)W_max < s.Emax)) r.status = kGridWReachesE;
    if (r.status == 0) {
      const GeomGrid Wg(s.Wmin, W_max, r.nW);
      int oe = -1;
      for (int q = 0; q < kGridOrders && oe < 0; ++q) {
        const GeomGrid Eg(s.Emin, s.Emax, cand.orders[q]);
        if (check_cdf_dev(s, r.ref, tol_E, Eg.at(cand.orders[q] - 2), Wg, sc, &sh)) oe = q;
      }
      if (oe < 0) r.status = kGridNoE;
      for (int q = 0; q < kGridPercents && oe >= 0 && r.nE == 0; ++q) {
        const int n = cand.pct[oe][q];
        if (check_cdf_dev(s, r.ref, tol_E, lin_at(s.Emin, s.Emax, n, n - 2), Wg, sc, &sh)) r.nE = n;
      }
      if (oe >= 0 && r.nE == 0) r.status = kGridERefine;
    }
  }
  if (threadIdx.x == 0) res[j] = r;
}

GridCandidates grid_candidates() {
  GridCandidates c;
  const std::vector<int> om = orders_of_mag(kGsMaxTerms);
  for (int q = 0; q < kGridOrders; ++q) {
    c.orders[q] = om[q];
    const std::vector<int> pc = percents(om[q]);
    for (int p = 0; p < kGridPercents; ++p) c.pct[q][p] = pc[p];
  }
  return c;
}

size_t grid_search_scratch_bytes(int n_jobs) { return (size_t)n_jobs * sizeof(GsScratch); }

hipError_t launch_grid_search(const GridSearchJob* jobs, int n_jobs, double tol_E, double tol_W, void* scratch,
                              GridSearchResult* res, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_grid_search, dim3(n_jobs), dim3(kGsBlock), 0, s, jobs, n_jobs, grid_candidates(), tol_E,
                     tol_W, scratch, res);
  return hipGetLastError();
}

void grid_materialize(const GridSearchJob& j, int nE, int nW, std::vector<float>& Es, std::vector<float>& Ws) {
  Ws = geomspace32(j.Wmin, (j.Emax - j.Ef) * 0.5, nW);
  Es = linspace32(j.Emin, j.Emax, nE);
}

const char* grid_status_text(int status) {
  switch (status) {
    case kGridEmptyW: return "the W range is empty";
    case kGridNoW: return "no W grid up to 1e6 points meets tol_W";
    case kGridWRefine: return "W refinement failed";
    case kGridWReachesE: return "energy-loss grid reaches the incident energy (probTables.py:396)";
    case kGridNoE: return "no E grid up to 1e6 points meets tol_E";
    case kGridERefine: return "E refinement failed";
    default: return "ok";
  }
}

}  // namespace mcd
