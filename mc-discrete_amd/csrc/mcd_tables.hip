// mcd_tables.hip — device-side table construction.
//
// * k_build_cdf: the inverse-CDF energy-loss rows of probTable.computeBlock
//   (MC/tools/probTables.py:415-462).  One thread per energy row; FP64
//   trapezoids and cumulative sums (rounded to float32 once), or the float32
//   arithmetic of the reference.  The mask W[j+1] > (E - E_F)/2 is evaluated in
//   float32 exactly as the reference's float32 arrays do (:429-439).
// * k_pick_nodes / k_pick_check: the LDS pick table (sorted cumulative channel
//   probabilities and ln mfp on a ln E grid) and its interpolation-error check
//   against the exact FP64 pick, which sets the guard band.
#include <hip/hip_runtime.h>

#include "mcd_device.h"
#include "mcd_internal.h"

namespace mcd {

#pragma clang fp contract(off)
__device__ __forceinline__ double d_moller_dcs(double E, double W, double n) {  // extFunctions.py:159-184
  double eps = W / E;
  if (!(eps < 1.)) return 0.;
  return n * kPiEFour * (1. / (eps * eps) + (1. / ((1. - eps) * (1. - eps))) - (1. / (eps * (1. - eps)))) /
         (E * E * E);
}
__device__ __forceinline__ double d_gryz_dcs(double E, double W, double n, double Eb) {  // :189-219
  double eps = W / E, epsB = Eb / E;
  if (!(eps < 1. && epsB < 1.)) return 0.;
  return n * kPiEFour * eps * pow(1. + epsB, -1.5) * pow(1. - eps, epsB / (epsB + eps)) *
         ((1. - epsB) + 4. * epsB * log(2.7 + sqrt((1. - eps) / epsB)) / (3. * eps)) / (W * W * W);
}
#pragma clang fp contract(on)

__device__ __forceinline__ float f_moller_dcs(float E, float W, float kn) {
  float eps = W / E;
  float om = 1.f - eps;
  float v = kn * (1.f / (eps * eps) + (1.f / (om * om)) - (1.f / (eps * om))) / powf(E, 3.f);
  return eps < 1.f ? v : 0.f;
}
__device__ __forceinline__ float f_gryz_dcs(float E, float W, float kn, float Eb) {
  float eps = W / E, epsB = Eb / E;
  float v = kn * eps * powf(1.f + epsB, -1.5f) * powf(1.f - eps, epsB / (epsB + eps)) *
            ((1.f - epsB) + 4.f * epsB * logf(2.7f + sqrtf((1.f - eps) / epsB)) / (3.f * eps)) / powf(W, 3.f);
  return (eps < 1.f && epsB < 1.f) ? v : 0.f;
}

// Energy row i of a table's CDF (computeBlock, probTables.py:415-462).
__device__ __forceinline__ void cdf_row(const float* __restrict__ Es, const float* __restrict__ Ws, int i, int nW,
                                        double Ef, int kind, double n_el, double Eb, int precision,
                                        float* __restrict__ cdf) {
  const int nb = nW - 1;
  float* row = cdf + (size_t)i * nb;
  const float E32 = Es[i];
  const float wmax32 = (E32 - (float)Ef) * 0.5f;  // float32 mask, probTables.py:429-439
  // number of valid bins: W[j+1] <= wmax32 (Ws increasing)
  int nvalid = 0;
  while (nvalid < nb && !(Ws[nvalid + 1] > wmax32)) ++nvalid;
  if (precision == MCD_PREC_FP64) {
    const double E = (double)E32;
    double tot = 0.;
    double f0 = kind == MCD_TABLE_MOLLER ? d_moller_dcs(E, Ws[0], n_el) : d_gryz_dcs(E, Ws[0], n_el, Eb);
    for (int j = 0; j < nvalid; ++j) {
      double f1 = kind == MCD_TABLE_MOLLER ? d_moller_dcs(E, Ws[j + 1], n_el) : d_gryz_dcs(E, Ws[j + 1], n_el, Eb);
      tot += ((double)Ws[j + 1] - (double)Ws[j]) * (f0 + f1) / 2.0;
      f0 = f1;
    }
    double acc = 0.;
    f0 = kind == MCD_TABLE_MOLLER ? d_moller_dcs(E, Ws[0], n_el) : d_gryz_dcs(E, Ws[0], n_el, Eb);
    for (int j = 0; j < nvalid; ++j) {
      double f1 = kind == MCD_TABLE_MOLLER ? d_moller_dcs(E, Ws[j + 1], n_el) : d_gryz_dcs(E, Ws[j + 1], n_el, Eb);
      acc += ((double)Ws[j + 1] - (double)Ws[j]) * (f0 + f1) / 2.0;
      row[j] = (float)(acc / tot);
      f0 = f1;
    }
  } else {
    const float kn = (float)(n_el * kPiEFour);
    const float Ebf = (float)Eb;
    float acc = 0.f;
    float f0 = kind == MCD_TABLE_MOLLER ? f_moller_dcs(E32, Ws[0], kn) : f_gryz_dcs(E32, Ws[0], kn, Ebf);
    for (int j = 0; j < nvalid; ++j) {
      float f1 = kind == MCD_TABLE_MOLLER ? f_moller_dcs(E32, Ws[j + 1], kn) : f_gryz_dcs(E32, Ws[j + 1], kn, Ebf);
      float st = (Ws[j + 1] - Ws[j]) * (f0 + f1) / 2.f;
      acc += st;
      row[j] = acc;
      f0 = f1;
    }
    for (int j = 0; j < nvalid; ++j) row[j] = row[j] / acc;
  }
  for (int j = nvalid; j < nb; ++j) row[j] = kMasked;
}

__global__ void k_build_cdf(const float* __restrict__ Es, const float* __restrict__ Ws, int nE, int nW, double Ef,
                            int kind, double n_el, double Eb, int precision, float* __restrict__ cdf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nE) return;
  cdf_row(Es, Ws, i, nW, Ef, kind, n_el, Eb, precision, cdf);
}

// The job holding global row g of a batch: jobs cover consecutive row ranges
// [row0, row0 + nE) in order.
__device__ __forceinline__ int job_of_row(const CdfJob* __restrict__ jobs, int n_jobs, long long g) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].row0 <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The CDF rows of many tables in one launch: one thread per energy row of
// any table, the same arithmetic as k_build_cdf.
__global__ void k_build_cdf_batched(const CdfJob* __restrict__ jobs, int n_jobs, long long total_rows,
                                    int precision) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_rows) return;
  const CdfJob& j = jobs[job_of_row(jobs, n_jobs, g)];
  cdf_row(j.Es, j.Ws, (int)(g - j.row0), j.nW, j.Ef, j.kind, j.n_el, j.Eb, precision, j.cdf);
}

hipError_t launch_build_cdf(const float* Es, const float* Ws, int nE, int nW, double Ef, int kind, double n_el,
                            double Eb, int precision, float* cdf, hipStream_t s) {
  const int block = 64;
  hipLaunchKernelGGL(k_build_cdf, dim3((nE + block - 1) / block), dim3(block), 0, s, Es, Ws, nE, nW, Ef, kind, n_el,
                     Eb, precision, cdf);
  return hipGetLastError();
}

// Likelihood-ratio scores of loss-table column picks for the derivative
// replay, from the discrete tables themselves (their float32 CDFs, masks
// included).  p_j is the mass of column j of a row (CDF difference; 0 where
// masked).
__device__ __forceinline__ double col_mass(const float* row, int j) {
  const float c = row[j];
  if (c == kMasked) return 0.;
  return (double)c - (j > 0 ? (double)row[j - 1] : 0.);
}

// Parameter score: (p_j(+) - p_j(-)) / (2 delta p_j) from the tables rebuilt
// at the parameter +- delta (the central difference of the column
// probabilities, masks included).
__global__ void k_ratio_score(const float* __restrict__ c0, const float* __restrict__ cp,
                              const float* __restrict__ cm, int nE, int nb, double inv_2delta,
                              float* __restrict__ out) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)nE * nb) return;
  const size_t i = k / nb;
  const int j = (int)(k - i * nb);
  const float* r0 = c0 + i * nb;
  const double p0 = col_mass(r0, j);
  out[k] = p0 > 0. ? (float)((col_mass(cp + i * nb, j) - col_mass(cm + i * nb, j)) * inv_2delta / p0) : 0.f;
}

// Energy score: rows are piecewise constant in energy (row i serves
// (Es[i-1], Es[i]]); the score of a pick from row i is the row-to-row change
// (p_j(i) - p_j(i-1)) / ((Es[i] - Es[i-1]) p_j(i)).
__global__ void k_row_score(const float* __restrict__ cdf, const float* __restrict__ Es, int nE, int nb,
                            float* __restrict__ out) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)nE * nb) return;
  const size_t i = k / nb;
  const int j = (int)(k - i * nb);
  const double p = col_mass(cdf + i * nb, j);
  if (i == 0 || !(p > 0.)) {
    out[k] = 0.f;
    return;
  }
  const double q = col_mass(cdf + (i - 1) * nb, j);
  out[k] = (float)((p - q) / (((double)Es[i] - (double)Es[i - 1]) * p));
}

hipError_t launch_ratio_score(const float* c0, const float* cp, const float* cm, int nE, int nW, double inv_2delta,
                              float* out, hipStream_t s) {
  const size_t n = (size_t)nE * (nW - 1);
  hipLaunchKernelGGL(k_ratio_score, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c0, cp, cm, nE, nW - 1,
                     inv_2delta, out);
  return hipGetLastError();
}

hipError_t launch_row_score(const float* cdf, const float* Es, int nE, int nW, float* out, hipStream_t s) {
  const size_t n = (size_t)nE * (nW - 1);
  hipLaunchKernelGGL(k_row_score, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cdf, Es, nE, nW - 1, out);
  return hipGetLastError();
}

// Row records (layout in mcd_device.h): bracket energies and the quantile
// guide q -> bisect_left(cdf_row, q / G) (masked entries hold kMasked, above
// every q / G).  One workgroup per row.
__device__ __forceinline__ void row_record(const float* __restrict__ cdf, const float* __restrict__ Es, int row,
                                           int nW, int narrow, uint8_t* __restrict__ rows) {
  const int nb = nW - 1;
  const float* r = cdf + (size_t)row * nb;
  uint8_t* rec = rows + (size_t)row * kRowBytes;
  if (threadIdx.x == 0) {
    float2 br;
    br.x = Es[row > 0 ? row - 1 : 0];
    br.y = Es[row];
    *reinterpret_cast<float2*>(rec) = br;
  }
  const int G = narrow ? kGuideNarrow : kGuideWide;
  const int used = kRowHeader + (narrow ? 1 : 2) * (G + 1);
  for (int b = used + threadIdx.x; b < kRowBytes; b += blockDim.x) rec[b] = 0;
  for (int q = threadIdx.x; q <= G; q += blockDim.x) {
    const float v = (float)q / (float)G;
    int lo = 0, hi = nb;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (r[mid] < v) lo = mid + 1; else hi = mid;
    }
    if (narrow) rec[kRowHeader + q] = (uint8_t)lo;
    else reinterpret_cast<uint16_t*>(rec + kRowHeader)[q] = (uint16_t)lo;
  }
}

__global__ void k_build_rows(const float* __restrict__ cdf, const float* __restrict__ Es, int nE, int nW, int narrow,
                             uint8_t* __restrict__ rows) {
  row_record(cdf, Es, blockIdx.x, nW, narrow, rows);
}

// Row records of many tables in one launch: one workgroup per energy row of
// any table.
__global__ void k_build_rows_batched(const CdfJob* __restrict__ jobs, int n_jobs) {
  const long long g = blockIdx.x;
  const CdfJob& j = jobs[job_of_row(jobs, n_jobs, g)];
  row_record(j.cdf, j.Es, (int)(g - j.row0), j.nW, j.nW - 1 <= 255, j.rows);
}

hipError_t launch_build_tables_batched(const CdfJob* jobs, int n_jobs, long long total_rows, int precision,
                                       hipStream_t s) {
  if (n_jobs < 1 || total_rows < 1 || total_rows > 0x7FFFFFFFLL) return hipErrorInvalidValue;
  const int block = 64;
  hipLaunchKernelGGL(k_build_cdf_batched, dim3((unsigned)((total_rows + block - 1) / block)), dim3(block), 0, s, jobs,
                     n_jobs, total_rows, precision);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_build_rows_batched, dim3((unsigned)total_rows), dim3(128), 0, s, jobs, n_jobs);
  return hipGetLastError();
}

hipError_t launch_build_rows(const float* cdf, const float* Es, int nE, int nW, uint8_t* rows, hipStream_t s) {
  const int narrow = nW - 1 <= 255;
  hipLaunchKernelGGL(k_build_rows, dim3(nE), dim3(128), 0, s, cdf, Es, nE, nW, narrow, rows);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pick-table nodes: FP64 sorted cumulative probabilities, ln mfp, labels.
__global__ void k_pick_nodes(DevCfg c, int n_nodes, double lnE0, double dln, float* nodes, int stride) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_nodes) return;
  double lnE = k == n_nodes - 1 ? log(c.E0) : lnE0 + k * dln;
  double E = k == n_nodes - 1 ? c.E0 : (k == 0 ? c.Emin : exp(lnE));
  double sig[MCD_MAX_CHANNELS], probs[MCD_MAX_CHANNELS];
  int lab[MCD_MAX_CHANNELS], order[MCD_MAX_CHANNELS];
  const int n = d_sigma_set(E, c, sig, lab);
  double s = 0.;
  for (int i = 0; i < n; ++i) s += sig[i];
  d_sorted_probs(sig, n, order, probs);
  // Zero-probability channels are never picked, so their labels are free:
  // order them by rising threshold so that a channel switching on at its
  // threshold keeps its position and the cell needs no reordering.
  int nz = 0;
  while (nz < n && sig[order[nz]] > 0.) ++nz;
  double thr[MCD_MAX_CHANNELS];
  for (int i = 0; i < n; ++i) {
    const int l = lab[i];
    thr[i] = l == MCD_CH_MOLLER ? c.Wc : (l == MCD_CH_QUINN ? c.Epl : (l >= MCD_CH_GRYZ0 ? c.Eb[l - MCD_CH_GRYZ0] : 0.));
  }
  for (int i = nz + 1; i < n; ++i) {
    int v = order[i], j = i - 1;
    while (j >= nz && thr[order[j]] > thr[v]) { order[j + 1] = order[j]; --j; }
    order[j + 1] = v;
  }
  float* o = nodes + (size_t)k * stride;
  for (int i = 0; i < n - 1; ++i) o[i] = (float)probs[i];
  o[n - 1] = (float)log2(1. / (c.atnd * s) * 1e14 * M_LN2);  // log2(mfp ln 2)
  uint32_t lo = 0, hi = 0;
  for (int i = 0; i < n; ++i) {
    uint32_t l = (uint32_t)lab[order[i]];
    if (i < 8) lo |= l << (4 * i); else hi |= l << (4 * (i - 8));
  }
  o[n] = __uint_as_float(lo);
  o[n + 1] = __uint_as_float(hi | (kMetaNoSwap << 12));
  for (int i = n + 2; i < stride; ++i) o[i] = 0.f;
}

hipError_t launch_pick_nodes(const DevCfg& c, int n_nodes, double lnE0, double dln, float* nodes, int stride,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_pick_nodes, dim3((n_nodes + 63) / 64), dim3(64), 0, s, c, n_nodes, lnE0, dln, nodes, stride);
  return hipGetLastError();
}

// Check points t = (m + 0.5)/M inside every cell: compare the FP32
// interpolation (the transport's own f_pick_eval) with the exact FP64 order and
// cumulative probabilities.  err[2*i] = max boundary error, err[2*i+1] = error
// of pA - pB in swap cells; bad[i] = order mismatch not explained by the swap.
template <int NCH>
__global__ void k_pick_check(DevCfg c, int M, float* err, int* bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncell = c.pick.n_nodes - 1;
  if (i >= ncell * M) return;
  const int k = i / M, m = i % M;
  const double lnE0 = log(c.Emin);
  const double lnE = lnE0 + ((double)k + ((double)m + 0.5) / (double)M) * (log(c.E0) - lnE0) / (double)ncell;
  const double E = exp(lnE);
  PickEval<NCH> o;
  f_pick_eval<NCH>(c.pick.nodes, c.pick, (float)E, o);
  err[2 * i] = 0.f;
  err[2 * i + 1] = 0.f;
  bad[i] = 0;
  if (o.exact) return;
  double sig[MCD_MAX_CHANNELS], probs[MCD_MAX_CHANNELS];
  int lab[MCD_MAX_CHANNELS], order[MCD_MAX_CHANNELS];
  const int n = d_sigma_set(E, c, sig, lab);
  d_sorted_probs(sig, n, order, probs);
  // last position with non-zero probability
  int last = 0;
  for (int q = 0; q < n; ++q)
    if (sig[order[q]] > 0.) last = q;
  float eb = 0.f;
  int mism = 0;
  for (int q = 0; q <= last; ++q) {
    if (perm_label(o.perm_lo, o.perm_hi, q) != lab[order[q]]) mism = 1;
  }
  for (int q = 0; q < NCH - 1; ++q) eb = fmaxf(eb, fabsf((float)((double)o.B[q] - probs[q])));
  if (o.swap != (int)kMetaNoSwap) {
    const int j = o.swap;
    // true probabilities of the node-k position-j (A) and j+1 (B) channels
    uint32_t la = (uint32_t)perm_label(__float_as_uint(c.pick.nodes[(size_t)k * c.pick.stride + NCH]),
                                       __float_as_uint(c.pick.nodes[(size_t)k * c.pick.stride + NCH + 1]) & 0xFFFu, j);
    uint32_t lb = (uint32_t)perm_label(__float_as_uint(c.pick.nodes[(size_t)k * c.pick.stride + NCH]),
                                       __float_as_uint(c.pick.nodes[(size_t)k * c.pick.stride + NCH + 1]) & 0xFFFu,
                                       j + 1);
    double s = 0., sa = 0., sb = 0.;
    for (int q = 0; q < n; ++q) {
      s += sig[q];
      if ((uint32_t)lab[q] == la) sa = sig[q];
      if ((uint32_t)lab[q] == lb) sb = sig[q];
    }
    const double dtrue = (sa - sb) / s;
    err[2 * i + 1] = fabsf((float)((double)(o.pA - o.pB) - dtrue));
    if (mism) {
      // allowed only when the two orders differ by the swap itself (then the
      // guard must catch it: the boundary error at j is not meaningful)
      int other = 0;
      for (int q = 0; q <= last; ++q)
        if (q != j && q != j + 1 && perm_label(o.perm_lo, o.perm_hi, q) != lab[order[q]]) other = 1;
      mism = other;
      float eb2 = 0.f;
      for (int q = 0; q < NCH - 1; ++q)
        if (q != j) eb2 = fmaxf(eb2, fabsf((float)((double)o.B[q] - probs[q])));
      eb = eb2;
    }
  }
  err[2 * i] = eb;
  bad[i] = mism;
}

hipError_t launch_pick_check(const DevCfg& c, int M, float* err, int* bad, hipStream_t s) {
  const int total = (c.pick.n_nodes - 1) * M;
  const dim3 g((total + 63) / 64), b(64);
  switch (c.n_shells + 3) {
    case 4: hipLaunchKernelGGL(k_pick_check<4>, g, b, 0, s, c, M, err, bad); break;
    case 5: hipLaunchKernelGGL(k_pick_check<5>, g, b, 0, s, c, M, err, bad); break;
    case 6: hipLaunchKernelGGL(k_pick_check<6>, g, b, 0, s, c, M, err, bad); break;
    case 7: hipLaunchKernelGGL(k_pick_check<7>, g, b, 0, s, c, M, err, bad); break;
    case 8: hipLaunchKernelGGL(k_pick_check<8>, g, b, 0, s, c, M, err, bad); break;
    case 9: hipLaunchKernelGGL(k_pick_check<9>, g, b, 0, s, c, M, err, bad); break;
    case 10: hipLaunchKernelGGL(k_pick_check<10>, g, b, 0, s, c, M, err, bad); break;
    case 11: hipLaunchKernelGGL(k_pick_check<11>, g, b, 0, s, c, M, err, bad); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 63) / 64), dim3(64), 0, s, p, v, n);
  return hipGetLastError();
}

// masked sentinel -> 0 (the reference's stored layout), or 0 -> sentinel
__global__ void k_unmask(const float* __restrict__ src, float* __restrict__ dst, size_t n, int to_reference) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = src[i];
  if (to_reference) dst[i] = v == kMasked ? 0.f : v;
  else dst[i] = v == 0.f ? kMasked : v;
}

hipError_t launch_unmask(const float* src, float* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_unmask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n, 1);
  return hipGetLastError();
}

hipError_t launch_mask(const float* src, float* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_unmask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n, 0);
  return hipGetLastError();
}

}  // namespace mcd
