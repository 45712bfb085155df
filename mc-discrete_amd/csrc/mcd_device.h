// mcd_device.h — device-side building blocks of the gfx950 transport kernels.
//
// One electron per lane.  The FP32 transport keeps the energy as an unevaluated
// float pair (hi + lo, compensated subtraction) so every energy threshold
// (absorption, channel thresholds, table rows, the binary-collision limit) is
// decided on an energy that agrees with an FP64 run to ~48 bits.  The channel
// pick reads an LDS-staged table of cumulative channel probabilities
// interpolated in ln E and falls back to the exact FP64 pick whenever the
// random number lies within the table's guard band of a channel boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mcd.h"

namespace mcd {

constexpr double kPiEFour = 6.51408491409531e-14;  // cm^2 eV^2, parameters.py:7
constexpr double kBohrR = 5.29177210903e-11;       // m, parameters.py:8
constexpr double kPi = 3.141592653589793;
constexpr float kPiF = 3.14159265358979f;
constexpr float kMasked = 2.0f;              // device sentinel for masked CDF entries (> any u)
constexpr int64_t kMaxStepsOff = 1ll << 24;  // step cap when maxScatt is off

// Pick-table node layout: [B_0 .. B_{n-2}, log2(mfp ln 2), perm_lo, perm_hi|meta, pad..]
// perm: 4-bit channel labels in sorted (pick) order, positions 0..7 in perm_lo,
// 8..10 in bits 0..11 of perm_hi.  meta (bits 12..): cell [k, k+1] swap position
// (bits 12..15, 15 = no swap) and exact-only flag (bit 16).
// Floats per pick-table node: n_ch - 1 cumulative probabilities, log2 of
// (mfp ln 2) and two label words, padded to an odd count so that the nodes
// read by the 64 lanes of a wavefront (one node each) start in different LDS
// banks.
__host__ __device__ constexpr int pick_stride(int n_ch) { return (n_ch + 2) | 1; }
// The wide node record of an LDS copy: the n_ch + 2 floats of the table's
// record, then the differences to the next node of its n_ch - 1 cumulative
// probabilities and of its log2 mean free path (the interpolation's slopes,
// formed once when the copy is staged instead of at every step).
__host__ __device__ constexpr int pick_stride_wide(int n_ch) { return (2 * n_ch + 2) | 1; }

constexpr uint32_t kMetaNoSwap = 15u;
constexpr uint32_t kMetaExact = 1u << 16;

// Row records of a loss table.  Record i (kRowBytes bytes) holds the grid
// energies Es[i-1], Es[i] that bracket row i (record 0: Es[0] twice) and the
// quantile guide of row i's CDF: entry q = bisect_left(cdf_row, q / G), the
// first column whose CDF >= q / G, so the column of u lies in
// [guide[q], guide[q+1]] with q = floor(u G).  Tables with at most 255 loss
// bins keep a uint8 guide with G = 256, wider ones a uint16 guide with G = 128.
// A lookup reads the guessed row's bracket and guide pair in one round trip
// and, when the pair differs, up to four CDF entries in a second.  G = 256
// keeps the two tables of a configuration with their records (~4.3 MB at
// 2000 x 199) close to one XCD's 4 MB L2; G = 512 settled more lookups from
// the guide alone but sent seven times the L2 misses to the fabric.
constexpr int kGuideNarrow = 256;
constexpr int kGuideWide = 128;
constexpr int kRowHeader = 8;
constexpr int kRowBytes = 272;
static_assert(kRowHeader + kGuideNarrow + 1 <= kRowBytes && kRowHeader + 2 * (kGuideWide + 1) <= kRowBytes,
              "row record too small");

struct DevTable {
  const float* Es;   // nE energies (float32 grid)
  const float* Ws;   // nW losses
  const float* cdf;  // nE x (nW-1), masked = kMasked
  const uint8_t* rows;  // nE row records of kRowBytes
  int nE, nW;
  int narrow;           // 1: uint8 guide (nW - 1 <= 255), 0: uint16 guide
  float E_first, inv_h;  // linspace index guess
};

struct DevPick {
  const float* nodes;  // n_nodes x stride floats (global; staged into LDS)
  int n_nodes, stride, n_ch;
  float log2E0, inv_dlog2E;  // node k sits at log2 E = log2E0 + k / inv_dlog2E
  float delta;           // guard band on the cumulative probabilities
};

// Per-configuration constants, prepared once on the host in FP64.
struct DevCfg {
  int mode, bethe, n_shells, Z;
  uint32_t config_id;
  int64_t max_scatt;
  // FP64 copies (FP64 transport and the exact pick)
  double E0, Emin, Wc, Ef, Epl, atnd, thickness, c_alpha, ZZ, n_val;
  double J, bethe_kk, zeff, lnEstar, E_val;
  double expl_k;  // explicit Bethe law: 2 pi e^4 n Z_eff 1e-14 (S = expl_k (ln E - lnEstar) / E)
  int elastic;    // MCD_ELASTIC_*
  double nig_zp, nig_num;  // Nigram sigma: 5.43 Z^(2/3) and pi e^4 Z^2 (or pi e^4 (Z+1) Z)
  double Eb[MCD_MAX_SHELLS], ns[MCD_MAX_SHELLS];
  double dir0[3];
  // FP32 constants
  float E0_hi, E0_lo;
  float Emin_hi, Emin_lo, Ef2_hi, Ef2_lo;  // Ef2 = E_F (pair)
  float Epl_hi, Epl_lo;
  float thickness_f, c_alpha_f, ruth_k_f, mfp_k_f;
  float bethe_k1_f, bethe_c1_f, bethe_kJ_f, expl_k_f, expl_inv_Estar_f;
  float nig_zp_f, nig_num_f;
  float dir0_f[3];
  DevTable tm, tg;
  DevPick pick;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and the (0,1) mapping.
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                       uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    // three-input XOR in one v_bitop3_b32 (truth table 0x96)
    uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
    uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return {c0, c1, c2, c3};
}

// Loads from global memory through a pointer whose address space the compiler
// cannot see (table pointers arrive through LDS descriptors or by reference):
// without the cast they become flat loads, which also wait on the LDS counter.
template <class T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// Value of a uniform-indexed lane, as a wave-uniform (scalar) value.
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// ---------------------------------------------------------------------------
// Work distribution of the persistent kernels.  A wavefront refills its idle
// lanes from a reservation [next, end) of indices it has taken from a shared
// counter; when the reservation is short it takes a block of at least
// kRefillBlock indices with one atomic.  Every index is handed out exactly once
// (blocks from the counter are disjoint, a reservation is only consumed by its
// wavefront), and the wavefront keeps refilling until its reservation passes n,
// so no reserved index is left behind.  Taking one index per idle lane instead
// makes thousands of resident wavefronts queue on the counter's cache line for
// every refill; with blocks the counter sees one atomic per kRefillBlock
// electrons.
constexpr uint64_t kRefillBlock = 16;

struct WaveQueue {
  uint64_t next = 0, end = 0;
};

// Index for this lane if it is idle (caller checks idx < n); clears `open`
// once the reservation has passed n.  idle = ballot of idle lanes, nonzero.
__device__ __forceinline__ uint64_t refill_index(WaveQueue& q, unsigned long long* counter, uint64_t idle, int lane,
                                                 uint64_t n, bool& open) {
  const uint64_t nidle = (uint64_t)__popcll(idle);
  const uint64_t pfx = (uint64_t)__popcll(idle & ((1ull << lane) - 1ull));
  const uint64_t avail = q.end - q.next;
  uint64_t idx;
  if (nidle > avail) {
    const uint64_t fetch = nidle - avail > kRefillBlock ? nidle - avail : kRefillBlock;
    const int leader = __ffsll((unsigned long long)idle) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)fetch);
    const uint64_t blk = lane_u64(base, leader);
    idx = pfx < avail ? q.next + pfx : blk + (pfx - avail);
    q.next = blk + (nidle - avail);
    q.end = blk + fetch;
  } else {
    idx = q.next + pfx;
    q.next += nidle;
  }
  if (q.next >= n) open = false;
  return idx;
}

// u = (2*(x>>9)+1) * 2^-24, exact in FP32: 1 + (x>>9) 2^-23 built from the
// bits ({0x7F, x} >> 9 = 0x3F800000 | x >> 9, one v_alignbit_b32), minus
// 1 - 2^-24 (the difference is representable, so exact).
__device__ __forceinline__ float u01f(uint32_t x) {
  return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, x, 9u)) - 0.99999994039535522461f;
}

// ln and exp from the hardware base-2 instructions (v_log_f32, v_exp_f32).
// Every argument on the transport path is a normal, finite float (energies,
// probabilities in [2^-24, 1), bounded exponents), so the subnormal scaling
// and range handling of the library functions are not needed.
constexpr float kLn2F = 0.69314718055994530942f;
constexpr float kLog2eF = 1.44269504088896340736f;
__device__ __forceinline__ float log2_f(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float exp2_f(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float ln_f(float x) { return __builtin_amdgcn_logf(x) * kLn2F; }
__device__ __forceinline__ float exp_f(float x) { return __builtin_amdgcn_exp2f(x * kLog2eF); }
__device__ __forceinline__ double u01d(uint32_t x) { return (double)(2u * (x >> 9) + 1u) * 5.9604644775390625e-08; }

// ---------------------------------------------------------------------------
// Float-pair energy arithmetic.
struct F2 { float hi, lo; };

__device__ __forceinline__ F2 f2_from_double(double v) {
  float h = (float)v;
  return {h, (float)(v - (double)h)};
}

// a - b, TwoSum on the high parts then Fast2Sum renormalisation.
__device__ __forceinline__ F2 f2_sub(F2 a, float bh, float bl) {
  float s = a.hi - bh;
  float bb = s - a.hi;
  float err = (a.hi - (s - bb)) + (-bh - bb);
  float lo = (a.lo - bl) + err;
  float hi = s + lo;
  lo = lo - (hi - s);
  return {hi, lo};
}

// a <= b for normalised pairs
__device__ __forceinline__ bool f2_le(F2 a, float bh, float bl) { return a.hi < bh || (a.hi == bh && a.lo <= bl); }
// v < a for a float v
__device__ __forceinline__ bool f_lt_f2(float v, F2 a) { return v < a.hi || (v == a.hi && a.lo > 0.f); }

// ---------------------------------------------------------------------------
// FP64 cross sections, mirroring crossSections.py operation order.
#pragma clang fp contract(off)
__device__ __forceinline__ double d_sigma_ruth(double E, const DevCfg& c) {  // crossSections.py:37-147
  if (c.elastic >= MCD_ELASTIC_NIGRAM) {  // ruther_N_sigma(_wDefl), :91-147
    const double b = 0.48 * (c.nig_zp / E);
    return c.nig_num / (4. * b * (1. + b) * (E * E));
  }
  // ruther_sigma (:37-58), ruther_sigma_wDefl (:62-88): ZZ = Z^2 or Z (Z + 1)
  double a = c.c_alpha / E;
  double Ek = E * 1e-3;
  double r = (Ek + 511.) / (Ek + 1024.);
  return 5.21e-21 * (c.ZZ / (Ek * Ek)) * (r * r) * (4. * kPi) / (a * (1. + a));
}
__device__ __forceinline__ double d_sigma_moller(double E, const DevCfg& c) {  // :183-208
  if (E <= c.Wc) return 0.;
  double eps = c.Wc * 1. / E;
  return c.n_val * kPiEFour * (1. / (E * E)) * (1. / eps - 1. / (1. - eps) + log(eps / (1. - eps)));
}
__device__ __forceinline__ double d_sigma_gryz(double E, double Esi, double nsi) {  // :213-242
  if (E <= Esi) return 0.;
  double U = E * 1. / Esi;
  const double q = (U - 1.) / (U + 1.);  // q^1.5 as q*sqrt(q): equal to pow to an ulp
  return nsi * kPiEFour * (q * sqrt(q)) *
         (1. + (2. * (1. - (1. / (2. * U)))) / 3. * log(2.7 + sqrt(U - 1.))) / (E * Esi);
}
__device__ __forceinline__ double d_sigma_quinn(double E, const DevCfg& c) {  // :247-284
  if (E <= c.Epl) return 0.;
  double E_Ef = E * 1. / c.Ef;
  double Epl_Ef = c.Epl * 1. / c.Ef;
  double tot = c.Epl * log((sqrt(1. + Epl_Ef) - 1.) / (sqrt(E_Ef) - sqrt(E_Ef - Epl_Ef))) / (2. * kBohrR * E);
  return (tot / c.atnd) * 1e4;
}

// Insertion-ordered sigma set (scattering.py:480-508); returns the count.
__device__ __forceinline__ int d_sigma_set(double E, const DevCfg& c, double* sig, int* lab) {
  int k = 0;
  sig[k] = d_sigma_ruth(E, c); lab[k++] = MCD_CH_RUTH;
  sig[k] = d_sigma_moller(E, c); lab[k++] = MCD_CH_MOLLER;
  for (int s = 0; s < c.n_shells; ++s) { sig[k] = d_sigma_gryz(E, c.Eb[s], c.ns[s]); lab[k++] = MCD_CH_GRYZ0 + s; }
  sig[k] = d_sigma_quinn(E, c); lab[k++] = MCD_CH_QUINN;
  return k;
}

// Sorted (pick-order) cumulative probabilities: stable descending sort,
// cumulative sum / sequential total (pickFromSigmas, scattering.py:39-59).
// Writes order[] (insertion indices) and probs[] and returns the total sigma.
__device__ __forceinline__ double d_sorted_probs(const double* sig, int n, int* order, double* probs) {
  for (int i = 0; i < n; ++i) order[i] = i;
  for (int i = 1; i < n; ++i) {
    int v = order[i], j = i - 1;
    while (j >= 0 && sig[order[j]] < sig[v]) { order[j + 1] = order[j]; --j; }
    order[j + 1] = v;
  }
  double acc = 0., tot = 0.;
  for (int i = 0; i < n; ++i) { acc += sig[order[i]]; probs[i] = acc; }
  for (int i = 0; i < n; ++i) tot += sig[order[i]];
  for (int i = 0; i < n; ++i) probs[i] = probs[i] / tot;
  return tot;
}

// Exact FP64 pick: returns the channel label; also the FP64 mean free path.
__device__ __noinline__ int d_pick(double E, double u, const DevCfg& c, double* mfp_out) {
  double sig[MCD_MAX_CHANNELS], probs[MCD_MAX_CHANNELS];
  int lab[MCD_MAX_CHANNELS], order[MCD_MAX_CHANNELS];
  int n = d_sigma_set(E, c, sig, lab);
  double s = 0.;
  for (int i = 0; i < n; ++i) s += sig[i];
  *mfp_out = 1. / (c.atnd * s) * 1e14;  // mfp_from_sigma, scattering.py:19-36
  d_sorted_probs(sig, n, order, probs);
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (probs[mid] < u) lo = mid + 1; else hi = mid;
  }
  if (lo >= n) lo = n - 1;
  return lab[order[lo]];
}
#pragma clang fp contract(off)
// Sigma of one channel label in FP64 (one formula per call keeps the guard
// path's live registers to a handful of doubles).
__device__ __forceinline__ double d_sigma_label(int l, double E, const DevCfg& c) {
  if (l == MCD_CH_RUTH) return d_sigma_ruth(E, c);
  if (l == MCD_CH_MOLLER) return d_sigma_moller(E, c);
  if (l == MCD_CH_QUINN) return d_sigma_quinn(E, c);
  const int s = l - MCD_CH_GRYZ0;
  return d_sigma_gryz(E, c.Eb[s], c.ns[s]);
}

__device__ __forceinline__ int insertion_index(int l, int n) {
  return l == MCD_CH_RUTH ? 0 : (l == MCD_CH_MOLLER ? 1 : (l == MCD_CH_QUINN ? n - 1 : 2 + l - MCD_CH_GRYZ0));
}
#pragma clang fp contract(on)
// ---------------------------------------------------------------------------
// Interpolated pick (FP32).  Node k holds the FP64 sorted cumulative
// probabilities at E_k; within cell [k, k+1] the order is constant except in
// cells flagged with an adjacent swap (a sigma crossing), where the two
// swapping channels' probabilities are interpolated separately and compared.
template <int NCH>
struct PickEval {
  float B[NCH - 1];      // cumulative probabilities in pick order (last = 1 implied)
  uint32_t perm_lo, perm_hi;
  float pA, pB;          // swap cells: interpolated probabilities of the two channels
  int swap;              // swap position or kMetaNoSwap
  bool exact;            // exact-only cell
  float mfp_ln2;         // mean free path times ln 2: the step is -mfp_ln2 log2(u)
};

template <int NCH, bool WIDE = false>
__device__ __forceinline__ void f_pick_eval(const float* __restrict__ nodes, const DevPick& P, float E,
                                            PickEval<NCH>& o) {
  constexpr int NB = NCH - 1;
  constexpr int S = WIDE ? pick_stride_wide(NCH) : pick_stride(NCH);
  constexpr int D = NB + 3;  // WIDE: the slopes follow the record
  float x = (log2_f(E) - P.log2E0) * P.inv_dlog2E;
  x = fminf(fmaxf(x, 0.f), (float)(P.n_nodes - 1));
  const int k = min((int)x, P.n_nodes - 2);
  const float f = x - (float)k;
  const float* a = nodes + k * S;
  const float* b = a + S;
  o.mfp_ln2 = exp2_f(a[NB] + f * (WIDE ? a[D + NB] : b[NB] - a[NB]));
  o.perm_lo = __float_as_uint(a[NB + 1]);
  const uint32_t meta = __float_as_uint(a[NB + 2]);
  o.perm_hi = meta & 0xFFFu;
  o.exact = (meta & kMetaExact) != 0;
  o.swap = (int)((meta >> 12) & 15u);
#pragma unroll
  for (int i = 0; i < NB; ++i) o.B[i] = a[i] + f * (WIDE ? a[D + i] : b[i] - a[i]);
  o.pA = o.pB = 0.f;
  if (o.swap != (int)kMetaNoSwap) {
    // node k: channel A at position j, B at j+1; node k+1: B at j, A at j+1
    const int j = o.swap;
    const float am = j > 0 ? a[j - 1] : 0.f, bm = j > 0 ? b[j - 1] : 0.f;
    const float ap = j + 1 < NB ? a[j + 1] : 1.f, bp = j + 1 < NB ? b[j + 1] : 1.f;
    const float pA_k = a[j] - am, pB_k = ap - a[j];
    const float pB_k1 = b[j] - bm, pA_k1 = bp - b[j];
    o.pA = pA_k + f * (pA_k1 - pA_k);
    o.pB = pB_k + f * (pB_k1 - pB_k);
    const bool bfirst = o.pB > o.pA;
    const float pj = bfirst ? o.pB : o.pA;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const float prev = i > 0 ? o.B[i - 1] : 0.f;
      o.B[i] = (i == j) ? prev + pj : o.B[i];
    }
    if (bfirst) {
      uint64_t perm = ((uint64_t)o.perm_hi << 32) | o.perm_lo;
      const uint64_t la = (perm >> (4 * j)) & 15u, lb = (perm >> (4 * (j + 1))) & 15u;
      perm &= ~((uint64_t)0xFFu << (4 * j));
      perm |= (lb << (4 * j)) | (la << (4 * (j + 1)));
      o.perm_lo = (uint32_t)perm;
      o.perm_hi = (uint32_t)(perm >> 32);
    }
  }
}

__device__ __forceinline__ int perm_label(uint32_t lo, uint32_t hi, int pos) {
  return (int)((pos < 8 ? (lo >> (4 * pos)) : (hi >> (4 * (pos - 8)))) & 15u);
}

// Returns the channel label, or -1 when the guard band (or an exact-only cell)
// requires the FP64 pick.  mfp_ln2 is always set.
template <int NCH, bool WIDE = false>
__device__ __forceinline__ int f_pick_table(const float* __restrict__ nodes, const DevPick& P, float E, float u,
                                            float& mfp_ln2, PickEval<NCH>& o) {
  f_pick_eval<NCH, WIDE>(nodes, P, E, o);
  mfp_ln2 = o.mfp_ln2;
  if (o.exact) return -1;
  if (o.swap != (int)kMetaNoSwap && fabsf(o.pA - o.pB) <= P.delta) {
    // the order of the crossing pair (positions j, j+1) is uncertain, but the
    // pick depends on it only when u falls in the pair's cumulative range
    // [B[j-1], B[j+1]]; outside it (by more than delta) the count below
    // decides, whichever way the pair is ordered
    const int j = o.swap;
    float lo = 0.f, hi = 1.f;
#pragma unroll
    for (int i = 0; i < NCH - 1; ++i) {
      lo = (i == j - 1) ? o.B[i] : lo;
      hi = (i == j + 1) ? o.B[i] : hi;
    }
    if (u >= lo - P.delta && u <= hi + P.delta) return -1;
  }
  // count of boundaries below u from the sign of B[i] - u: exact (the
  // difference is +0 only when they are equal, and a flushed subnormal keeps
  // its sign)
  int cnt = 0;
  float gap = 1.f;
#pragma unroll
  for (int i = 0; i < NCH - 1; ++i) {
    const float d = o.B[i] - u;
    cnt += (int)(__float_as_uint(d) >> 31);
    gap = fminf(gap, fabsf(d));
  }
  if (gap <= P.delta) return -1;
  return perm_label(o.perm_lo, o.perm_hi, cnt);
}

// Exact FP64 pick for a lane whose interpolated pick fell inside the guard
// band.  The descending-sigma order is the cell's order (node labels), with
// the swap pair of a crossing cell ordered by comparing its two sigmas in
// FP64 (ties keep insertion order); in exact-only cells the order is found by
// repeated selection of the largest remaining sigma.  Cumulative sums run in
// that order, so probs[k] = cum_k / total equals pickFromSigmas' FP64 values
// (scattering.py:48-56) and the first k with probs[k] >= u is its bisect_left.
#pragma clang fp contract(off)
template <int NCH>
__device__ __forceinline__ int d_pick_ordered(double E, double u, const DevCfg& c, uint32_t perm_lo, uint32_t perm_hi,
                                           int swap, bool exact) {
  uint64_t perm = ((uint64_t)(perm_hi & 0xFFFu) << 32) | perm_lo;
  if (exact) {
    // selection order: position k gets the largest remaining sigma (lowest insertion index on ties)
    uint32_t used = 0;
    uint64_t out = 0;
#pragma unroll 1
    for (int k = 0; k < NCH; ++k) {
      int best = -1;
      double bv = -1.0;
#pragma unroll 1
      for (int q = 0; q < NCH; ++q) {
        if (used & (1u << q)) continue;
        const int l = q == 0 ? MCD_CH_RUTH : (q == 1 ? MCD_CH_MOLLER : (q == NCH - 1 ? MCD_CH_QUINN : MCD_CH_GRYZ0 + q - 2));
        const double v = d_sigma_label(l, E, c);
        if (v > bv) { bv = v; best = q; }
      }
      used |= 1u << best;
      const int lb = best == 0 ? MCD_CH_RUTH : (best == 1 ? MCD_CH_MOLLER : (best == NCH - 1 ? MCD_CH_QUINN : MCD_CH_GRYZ0 + best - 2));
      out |= (uint64_t)lb << (4 * k);
    }
    perm = out;
  } else if (swap != (int)kMetaNoSwap) {
    const int la = (int)((perm >> (4 * swap)) & 15u), lb = (int)((perm >> (4 * (swap + 1))) & 15u);
    const double sa = d_sigma_label(la, E, c), sb = d_sigma_label(lb, E, c);
    const bool b_first = sb > sa || (sb == sa && insertion_index(lb, NCH) < insertion_index(la, NCH));
    if (b_first) {
      perm &= ~((uint64_t)0xFFu << (4 * swap));
      perm |= ((uint64_t)lb << (4 * swap)) | ((uint64_t)la << (4 * (swap + 1)));
    }
  }
  double tot = 0.;
#pragma unroll 1
  for (int k = 0; k < NCH; ++k) tot += d_sigma_label((int)((perm >> (4 * k)) & 15u), E, c);
  double acc = 0.;
  int last = (int)((perm >> (4 * (NCH - 1))) & 15u);
#pragma unroll 1
  for (int k = 0; k < NCH; ++k) {
    const int l = (int)((perm >> (4 * k)) & 15u);
    acc += d_sigma_label(l, E, c);
    if (acc / tot >= u) return l;
  }
  return last;
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Inverse-CDF table lookup (pickTable, scattering.py:63-94).
// Row = first grid energy >= E (bisect_left), column = first CDF entry >= u.
__device__ __forceinline__ void guide_pair(const DevTable& t, const uint8_t* rec, float u, int& g0, int& g1) {
  if (t.narrow) {
    const int q = min((int)(u * (float)kGuideNarrow), kGuideNarrow - 1);
    g0 = ldg(rec + kRowHeader + q);
    g1 = ldg(rec + kRowHeader + q + 1);
  } else {
    const int q = min((int)(u * (float)kGuideWide), kGuideWide - 1);
    const uint16_t* g = reinterpret_cast<const uint16_t*>(rec + kRowHeader) + q;
    g0 = ldg(g);
    g1 = ldg(g + 1);
  }
}

// Loss-table lookup (scattering.py:603-608,625-630): row = bisect_left(Es, E)
// with es_lt_E(v) deciding v < E, col = bisect_left(cdf_row, u).  The
// linspace guess of the row is checked against its bracket; the guide pair of
// the guessed row is loaded in the same round trip.
template <class EnergyLess>
__device__ __forceinline__ int table_lookup(const DevTable& t, float Eguess, EnergyLess es_lt_E, float u, int& row) {
  int i = (int)ceilf((Eguess - t.E_first) * t.inv_h);
  i = max(1, min(i, t.nE - 1));
  const uint8_t* rec = t.rows + (size_t)i * kRowBytes;
  float2 br;
  br.x = ldg(reinterpret_cast<const float*>(rec));
  br.y = ldg(reinterpret_cast<const float*>(rec) + 1);
  int g0, g1;
  guide_pair(t, rec, u, g0, g1);
  if (!(es_lt_E(br.x) && !es_lt_E(br.y))) {
    int a = 0, b = t.nE;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (es_lt_E(ldg(t.Es + mid))) a = mid + 1; else b = mid;
    }
    i = min(a, t.nE - 1);
    rec = t.rows + (size_t)i * kRowBytes;
    guide_pair(t, rec, u, g0, g1);
  }
  row = i;
  const float* r = t.cdf + (size_t)i * (size_t)(t.nW - 1);
  int lo = g0;
  const int span = g1 - g0;
  if (span > 0) {
    if (span <= 4) {
      // the CDF row is nondecreasing: col = g0 + #{j in [g0, g1) : cdf[j] < u};
      // four independent loads, clamped into [g0, g1)
      const int l = span - 1;
      const float a0 = ldg(r + lo), a1 = ldg(r + lo + min(1, l)), a2 = ldg(r + lo + min(2, l)),
                  a3 = ldg(r + lo + min(3, l));
      lo += (int)(a0 < u) + (int)(1 < span && a1 < u) + (int)(2 < span && a2 < u) + (int)(3 < span && a3 < u);
    } else {
      int hi = g1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ldg(r + mid) < u) lo = mid + 1; else hi = mid;
      }
    }
  }
  return lo;
}

// cos(2 pi u) and |sin(2 pi u)| for u in (0, 1): the folded azimuth of the
// direction update.  The reductions a = |u - rint(u)| in [0, 1/2],
// q = rint(4a) in {0, 1, 2} and z = a - q/4 in [-1/8, 1/8] are exact; then
// sin(2 pi z) = z P(z^2) and cos(2 pi z) = 1 + z^2 Q(z^2) with coefficients
// fitted by weighted least squares on [0, 1/8] (within 1.7 ulp of the true
// values at every u = (2m+1) 2^-24), and the quadrant q rotates the pair.
__device__ __forceinline__ void cos_abs_sin_2pi(float u, float& cp, float& sp) {
  const float a = fabsf(u - __builtin_rintf(u));
  const float q = __builtin_rintf(4.f * a);
  const float z = __builtin_fmaf(-0.25f, q, a);
  const float w = z * z;
  float ps = __builtin_fmaf(w, -75.43057250976562f, 81.59331512451172f);
  ps = __builtin_fmaf(w, ps, -41.34166717529297f);
  ps = __builtin_fmaf(w, ps, 6.2831854820251465f);
  const float sn = z * ps;
  float pc = __builtin_fmaf(w, 59.25483322143555f, -85.44401550292969f);
  pc = __builtin_fmaf(w, pc, 64.9393310546875f);
  pc = __builtin_fmaf(w, pc, -19.739208221435547f);
  const float cs = __builtin_fmaf(w, pc, 1.f);
  cp = q == 0.f ? cs : (q == 1.f ? -sn : -cs);
  sp = q == 0.f ? sn : (q == 1.f ? cs : -sn);
}

// ---------------------------------------------------------------------------
// Direction update, closed form of q_az * q_polar acting on (d, y_local)
// (electron.py:87-99, rotation.py:41-53): with w = y x d,
//   d' = cos(th) d + sin(th) (cos(ph) w + sin(ph) y),  y' = cos(ph) y - sin(ph) w,
// cos(th) = c2 - s2, sin(th) = 2 sqrt(c2 s2); the azimuth fold
// c_hPhi = sqrt(1 - sin^2(pi u)) gives cos(ph) = cos(2 pi u), sin(ph) = |sin(2 pi u)|.
__device__ __forceinline__ void f_rotate(float c2, float s2, float u_az, float* d, float* y) {
  float ct = c2 - s2;
  float st = 2.f * __builtin_amdgcn_sqrtf(c2 * s2);
  float sp, cp;
  cos_abs_sin_2pi(u_az, cp, sp);
  float w0 = y[1] * d[2] - y[2] * d[1];
  float w1 = y[2] * d[0] - y[0] * d[2];
  float w2 = y[0] * d[1] - y[1] * d[0];
  float nd0 = ct * d[0] + st * (cp * w0 + sp * y[0]);
  float nd1 = ct * d[1] + st * (cp * w1 + sp * y[1]);
  float nd2 = ct * d[2] + st * (cp * w2 + sp * y[2]);
  float ny0 = cp * y[0] - sp * w0;
  float ny1 = cp * y[1] - sp * w1;
  float ny2 = cp * y[2] - sp * w2;
  float rd = __builtin_amdgcn_rsqf(nd0 * nd0 + nd1 * nd1 + nd2 * nd2);
  float ry = __builtin_amdgcn_rsqf(ny0 * ny0 + ny1 * ny1 + ny2 * ny2);
  d[0] = nd0 * rd; d[1] = nd1 * rd; d[2] = nd2 * rd;
  y[0] = ny0 * ry; y[1] = ny1 * ry; y[2] = ny2 * ry;
}

#pragma clang fp contract(off)
// FP64 rotation as the reference composes it: Hamilton product, rotation
// matrix normalised by |q|^2, renormalisation of both vectors.
__device__ __forceinline__ void d_rotate(double c2, double halfPhi, double* d, double* y) {
  double s_ht = pow(1. - c2, 0.5), c_ht = pow(c2, 0.5);
  double s_hp = sin(halfPhi), c_hp = pow(1. - s_hp * s_hp, 0.5);
  double a1 = c_hp, b1 = d[0] * s_hp, c1 = d[1] * s_hp, d1 = d[2] * s_hp;
  double a2 = c_ht, b2 = y[0] * s_ht, c2q = y[1] * s_ht, d2 = y[2] * s_ht;
  double w = a1 * a2 - b1 * b2 - c1 * c2q - d1 * d2;
  double x = a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2q;
  double yy = a1 * c2q - b1 * d2 + c1 * a2 + d1 * b2;
  double z = a1 * d2 + b1 * c2q - c1 * b2 + d1 * a2;
  double n = w * w + x * x + yy * yy + z * z;
  double M00 = 1.0 - 2 * (yy * yy + z * z) / n, M01 = 2 * (x * yy - z * w) / n, M02 = 2 * (x * z + yy * w) / n;
  double M10 = 2 * (x * yy + z * w) / n, M11 = 1.0 - 2 * (x * x + z * z) / n, M12 = 2 * (yy * z - x * w) / n;
  double M20 = 2 * (x * z - yy * w) / n, M21 = 2 * (yy * z + x * w) / n, M22 = 1.0 - 2 * (x * x + yy * yy) / n;
  double nd0 = M00 * d[0] + M01 * d[1] + M02 * d[2];
  double nd1 = M10 * d[0] + M11 * d[1] + M12 * d[2];
  double nd2 = M20 * d[0] + M21 * d[1] + M22 * d[2];
  double ny0 = M00 * y[0] + M01 * y[1] + M02 * y[2];
  double ny1 = M10 * y[0] + M11 * y[1] + M12 * y[2];
  double ny2 = M20 * y[0] + M21 * y[1] + M22 * y[2];
  double ld = sqrt(nd0 * nd0 + nd1 * nd1 + nd2 * nd2);
  double ly = sqrt(ny0 * ny0 + ny1 * ny1 + ny2 * ny2);
  d[0] = nd0 / ld; d[1] = nd1 / ld; d[2] = nd2 / ld;
  y[0] = ny0 / ly; y[1] = ny1 / ly; y[2] = ny2 / ly;
}
#pragma clang fp contract(on)

}  // namespace mcd
