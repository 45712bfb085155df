// mcd_backward.hip — derivative replay of electron trajectories on gfx950.
//
// Quantities differentiated (per electron, summed over electrons):
//   Y   = 1 if the electron backscatters (outcome 'bks'), else 0
//   H_b = Y * [exit energy in bin b of n_bins equal bins over [0, E0]]
// with respect to three parameters per mode (slot -> output column):
//   DS:         slot 0 = Wc, slot 1 = s_tilt [deg], slot 2 = ln(density)
//   continuous: slot 0 = J,  slot 1 = s_tilt [deg], slot 2 = ln(density)
//               (identically zero on a semi-infinite target: not carried)
//
// The replay regenerates every draw from the electron's Philox counters
// (electron, config, step) instead of storing trajectories, and estimates
// the derivative of the expectation with three ingredients:
//
// 1. Survival biasing at the surface.  At a step that flies towards the
//    surface (d_z < 0) from depth z, the exit probability is
//    q = exp(-z / (|d_z| mfp)); the step adds w*q to Y (and to the bin of the
//    pre-step energy, which is the exit energy the reference records), the
//    weight becomes w*(1-q), and the free path is drawn from the exponential
//    truncated to stay inside: l = -mfp ln(q + u (1-q)).  The trajectory never
//    leaves, so Y is a smooth function of positions and directions and the
//    geometry is differentiated pathwise (forward-mode tangents of z, d, y_local,
//    the energy and w).
//    Flights with d_z in (-graze, 0), where the direction derivative of q
//    grows like 1/|d_z|, exit as in the forward run instead; their exit
//    indicator is not differentiated (a share of order graze^2 of the yield
//    for a cosine-like exit distribution; graze defaults to 0.005).
// 2. Likelihood-ratio scores for the discrete channel picks (DS): a pick of
//    channel c adds d ln p_c / d theta (through the energy tangent and the
//    explicit parameter dependence) to the score S, and each later
//    contribution w*q carries w*q*S.  The node table holds p_c and the
//    logarithmic derivatives, built in FP64 on the device.
// 3. Threshold terms.  An electron is alive at step j iff E_j > Emin, and a
//    contribution lands in the bin of E_j; both indicators jump where E_j
//    crosses a threshold.  Each contribution adds the central difference of
//    the two indicators along the energy tangent,
//      C * ([E_j + dE_j delta] - [E_j - dE_j delta]) / (2 delta),
//    with delta the parameter step (default 1% of the parameter).  The replay
//    therefore continues past Emin while the electron is alive on one side
//    of that stencil (E + |dE_i delta_i| > Emin): those steps add the
//    threshold terms of the contributions that exist at theta -+ delta only.
//    For a smooth exit-energy density this is the threshold (boundary) term
//    of the derivative up to O(delta^2); DS exit energies are a discrete comb
//    (E0 - k E_pl minus table grid values), for which it is the
//    finite-difference quotient at step delta of the edge crossings, the
//    quantity a finite-difference check at that step measures.
//
// Energy-loss sampling: a Moller loss is W = Ws[col], the lower edge of the
// picked column of the table built on the grid geomspace(Wc, W_top, nW); W
// moves with Wc pathwise (dWs[j]/dWc = Ws[j] (1 - j/(nW-1)) / Wc), and the
// column pick adds a likelihood-ratio score for the change of the column
// probabilities, taken from the discrete tables themselves so that columns
// appearing or vanishing under the W_max = (E - E_F)/2 mask count:
//   Wc:     (p_j(Wc + delta) - p_j(Wc - delta)) / (2 delta p_j), tables rebuilt
//           at Wc +- delta on the device;
//   energy: (p_j(row i) - p_j(row i-1)) / ((Es[i] - Es[i-1]) p_j(row i)) times
//           dE/dtheta (rows are piecewise constant in energy), for the Moller
//           and the Gryzinski table.
// A Gryzinski loss does not depend on the parameters otherwise; a Quinn loss
// is E_pl, with dE_pl/dln(rho) = E_pl/2.  Left out: the tables' dependence on
// E_F (the W grid's upper end and the W_max mask move by 0.23 eV per percent
// of density).
#include <hip/hip_runtime.h>

#include "mcd_device.h"
#include "mcd_internal.h"

namespace mcd {

// ---------------------------------------------------------------------------
// Node table: n_nodes nodes on a uniform ln E grid over [ln Emin, ln E0].
// Node layout (floats): [ln mfp, d ln mfp/d ln E, d ln mfp/d Wc, d ln mfp/d ln rho,
//   P_0 .. P_{n_ch-2} (cumulative pick probabilities, insertion order),
//   then per channel c: (d ln p_c/d ln E, d ln p_c/d Wc, d ln p_c/d ln rho)].
__global__ void k_bwd_nodes(DevCfg c, int n_nodes, double lnE0, double dln, float* nodes, int stride) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_nodes) return;
  const double E = k == 0 ? c.Emin : (k == n_nodes - 1 ? c.E0 : exp(lnE0 + k * dln));
  float* o = nodes + (size_t)k * stride;
  for (int i = 0; i < stride; ++i) o[i] = 0.f;
  const double h = 1e-4;
  if (c.mode == MCD_MODE_CONT) {
    // Rutherford-only mean free path (scattering.py:780-797): mfp = 1e14 / (n sigma_R)
    const double s0 = d_sigma_ruth(E, c), sp = d_sigma_ruth(E * exp(h), c), sm = d_sigma_ruth(E * exp(-h), c);
    o[0] = (float)log(1e14 / (c.atnd * s0));
    o[1] = (float)(-(log(sp) - log(sm)) / (2. * h));
    o[2] = 0.f;
    o[3] = -1.f;
    return;
  }
  // 7 evaluations of the insertion-ordered sigma set: base, E*e^{+-h},
  // Wc*(1+-h), density*e^{+-h} (n, E_F ~ n^{2/3}, E_pl ~ n^{1/2})
  double sg[7][MCD_MAX_CHANNELS];
  int lab[MCD_MAX_CHANNELS];
  int n = 0;
  for (int v = 0; v < 7; ++v) {
    DevCfg cc = c;
    double Ev = E;
    if (v == 1) Ev = E * exp(h);
    if (v == 2) Ev = E * exp(-h);
    if (v == 3) cc.Wc = c.Wc * (1. + h);
    if (v == 4) cc.Wc = c.Wc * (1. - h);
    if (v == 5 || v == 6) {
      const double s = v == 5 ? h : -h;
      cc.atnd = c.atnd * exp(s);
      cc.Ef = c.Ef * exp(2. * s / 3.);
      cc.Epl = c.Epl * exp(0.5 * s);
    }
    n = d_sigma_set(Ev, cc, sg[v], lab);
  }
  double tot[7];
  for (int v = 0; v < 7; ++v) {
    tot[v] = 0.;
    for (int i = 0; i < n; ++i) tot[v] += sg[v][i];
  }
  const double dWc = 2. * h * c.Wc;
  const double gE = (log(tot[1]) - log(tot[2])) / (2. * h);
  const double gW = (log(tot[3]) - log(tot[4])) / dWc;
  const double gR = (log(tot[5]) - log(tot[6])) / (2. * h);
  o[0] = (float)log(1e14 / (c.atnd * tot[0]));
  o[1] = (float)(-gE);
  o[2] = (float)(-gW);
  o[3] = (float)(-1. - gR);
  double acc = 0.;
  for (int i = 0; i < n - 1; ++i) {
    acc += sg[0][i];
    o[4 + i] = (float)(acc / tot[0]);
  }
  float* g = o + 4 + (n - 1);
  for (int i = 0; i < n; ++i) {
    if (!(sg[0][i] > 0.)) continue;  // never picked here
    // one-sided difference where the channel opens within the stencil
    const double le = sg[0][i], lp = sg[1][i], lm = sg[2][i];
    const double dE = lm > 0. ? (log(lp) - log(lm)) / (2. * h) : (log(lp) - log(le)) / h;
    g[3 * i + 0] = (float)(dE - gE);
    g[3 * i + 1] = (float)((log(sg[3][i]) - log(sg[4][i])) / dWc - gW);
    g[3 * i + 2] = (float)((log(sg[5][i]) - log(sg[6][i])) / (2. * h) - gR);
  }
}

hipError_t launch_bwd_nodes(const DevCfg& c, int n_nodes, double lnE0, double dln, float* nodes, int stride,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_bwd_nodes, dim3((n_nodes + 63) / 64), dim3(64), 0, s, c, n_nodes, lnE0, dln, nodes, stride);
  return hipGetLastError();
}

// odd, so that the nodes read by the lanes of a wavefront start in different LDS banks
int bwd_stride(int mode, int n_ch) { return (mode == MCD_MODE_CONT ? 4 : 4 + (n_ch - 1) + 3 * n_ch) | 1; }

// ---------------------------------------------------------------------------
// Replay kernel.
constexpr float kDeg = 0.017453292519943295f;  // d/ds_tilt of the tilt angle [rad/deg]
constexpr float kXCut = 30.f;                   // exp(-30) ~ 1e-13: no exit contribution beyond
constexpr float kWMin = 1e-12f;                 // weight below which a replay ends
constexpr uint32_t kBwdStepCap = 1u << 22;

template <int MODE, bool JAC>
__global__ __launch_bounds__(256) void k_backward(DevCfg c, BwdArgs A) {
  extern __shared__ float lds[];
  float* nodes = lds;
  float* hist = lds + A.n_nodes * A.stride;  // JAC: n_bins x 4; VJP: 1 + n_bins upstream gradients
  {
    const int nf = A.n_nodes * A.stride;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) nodes[i] = A.nodes[i];
    if (JAC) {
      for (int i = threadIdx.x; i < 4 * A.n_bins; i += blockDim.x) hist[i] = 0.f;
    } else {
      for (int i = threadIdx.x; i < 1 + A.n_bins; i += blockDim.x) hist[i] = A.grad_out[i];
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
  const int stride = A.stride, n_ch = A.n_ch;
  const float Emin = c.Emin_hi;
  const float Epl = c.Epl_hi, Ef = c.Ef2_hi;
  const float Jf = (float)c.J, kJ = (float)(c.bethe_kk * c.J);
  const float inv_J = 1.f / Jf;
  const float* gup = hist;  // VJP: [dL/dY, dL/dH_0 ..]
  // Slots carried.  Continuous mode: free paths scale as 1/n and stopping
  // powers as n, so on a semi-infinite target (the only kind replayed) no exit
  // probability, energy or direction depends on the density scale: its slot
  // (2) is identically zero and is not carried, and its outputs stay 0.
  constexpr int NS = MODE == MCD_MODE_CONT ? 2 : 3;

  // per-lane accumulators: value and 3 slot derivatives
  float accv = 0.f, accg[3] = {0.f, 0.f, 0.f};

  bool active = false, open = true;
  WaveQueue wq;
  uint64_t e = 0;
  uint32_t step = 0;
  float E = 0.f, z = 0.f, w = 0.f;
  float d[3], y[3];
  float dE[3], dz[3], dw[3], S[3];
  float dd[3][3], dy[3][3];  // [slot][component]; DS keeps slot 1 implicit (rigid tilt)

  while (true) {
    // key words as loop-carried SGPR values, set where all lanes are converged
    // (keeps the twenty round keys from being hoisted and spilled)
    asm volatile("" : "+s"(k0), "+s"(k1));
    if (open) {
      const uint64_t idle = __ballot(!active);
      if (idle) {
        const uint64_t idx = refill_index(wq, A.queue, idle, lane, A.n, open);
        if (!active) {
          if (idx < A.n) {
            active = true;
            e = A.first + idx;
            step = 0;
            E = c.E0_hi;
            z = 0.f;
            w = 1.f;
            d[0] = c.dir0_f[0]; d[1] = c.dir0_f[1]; d[2] = c.dir0_f[2];
            y[0] = 0.f; y[1] = 1.f; y[2] = 0.f;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              dE[i] = dz[i] = dw[i] = S[i] = 0.f;
#pragma unroll
              for (int j = 0; j < 3; ++j) dd[i][j] = dy[i][j] = 0.f;
            }
            // d0 = (-sin s, 0, cos s), s in degrees (multiScatter.py:97-100)
            dd[1][0] = -kDeg * d[2];
            dd[1][2] = kDeg * d[0];
          }
        }
      }
    }
    if (!__any(active)) break;
    if (!active) continue;

    const U4 r = philox10((uint32_t)e, c.config_id | (uint32_t)(e >> 32) << 16, step, 0u, k0, k1);
    const float iE = __builtin_amdgcn_rcpf(E);
    float dlnE[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < NS; ++i) dlnE[i] = dE[i] * iE;
    if (MODE == MCD_MODE_DS) dlnE[1] = 0.f;  // DS energies do not depend on the tilt

    // node interpolation
    float tn = (ln_f(E) - A.lnE0) * A.inv_dln;
    tn = fminf(fmaxf(tn, 0.f), (float)(A.n_nodes - 1) - 1e-3f);
    const int kn = (int)tn;
    const float fn = tn - (float)kn;
    const float* n0 = nodes + kn * stride;
    const float* n1 = n0 + stride;
    auto lerp = [&](int o) { return n0[o] + fn * (n1[o] - n0[o]); };
    const float lnl = lerp(0), aE = lerp(1);
    const float lam = exp_f(lnl);
    float dlnl[3];
    dlnl[0] = aE * dlnE[0] + lerp(2);
    dlnl[1] = aE * dlnE[1];
    dlnl[2] = NS > 2 ? aE * dlnE[2] + lerp(3) : 0.f;

    float u_path, u_ang, u_az;
    int ch = 0;
    float sc[3] = {0.f, 0.f, 0.f};
    if (MODE == MCD_MODE_DS) {
      const float u0 = u01f(r.x);
      u_path = u01f(r.y);
      u_ang = u01f(r.z);
      u_az = u01f(r.w);
      // bisect_left on the cumulative probabilities: count of P_c < u
      for (int i = 0; i < n_ch - 1; ++i) ch += (lerp(4 + i) < u0) ? 1 : 0;
      const int go = 4 + (n_ch - 1) + 3 * ch;
      const float G = lerp(go), HW = lerp(go + 1), HR = lerp(go + 2);
      sc[0] = G * dlnE[0] + HW;
      sc[2] = G * dlnE[2] + HR;
    } else {
      u_path = u01f(r.x);
      u_ang = u01f(r.y);
      u_az = u01f(r.z);
    }

    // ---- exit towards the surface, survival-biased free path ----
    float ddz[3];
#pragma unroll
    for (int i = 0; i < NS; ++i) ddz[i] = dd[i][2];
    if (MODE == MCD_MODE_DS) ddz[1] = kDeg * d[0];  // rigid rotation about y
    // Accumulate an exit contribution C (value) with its pathwise/score
    // derivative dC into the bin of E, plus the threshold terms: the alive
    // indicator [E > Emin] and the bin indicator, differenced along the energy
    // tangent with step delta_i (central difference of the indicator).
    // Past Emin (base indicator 0) a contribution adds only its threshold
    // terms: the steps where the electron is alive at theta + delta or
    // theta - delta but not at theta.
    auto contribute = [&](float C, const float* dC) {
      const bool base = E > Emin;
      const int b = min((int)(E * A.inv_bin_w), A.n_bins - 1);
      float tv[3] = {0.f, 0.f, 0.f};   // threshold term of the yield
      int bp[3] = {-1, -1, -1}, bm[3] = {-1, -1, -1};
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        const float s = dE[i] * A.delta[i];
        const float Ep = E + s, Em = E - s;
        bp[i] = Ep > Emin ? min((int)(Ep * A.inv_bin_w), A.n_bins - 1) : -1;
        bm[i] = Em > Emin ? min((int)(Em * A.inv_bin_w), A.n_bins - 1) : -1;
        tv[i] = C * A.inv_2delta[i] * (float)((bp[i] >= 0) - (bm[i] >= 0));
      }
      if (JAC) {
        if (base) accv += C;
#pragma unroll
        for (int i = 0; i < NS; ++i) accg[i] += (base ? dC[i] : 0.f) + tv[i];
        float* hb = hist + 4 * b;
        if (base) atomicAdd(&hb[0], C);
#pragma unroll
        for (int i = 0; i < NS; ++i) {
          if (base) atomicAdd(&hb[1 + i], dC[i]);
          if (bp[i] != bm[i]) {
            const float f = C * A.inv_2delta[i];
            if (bp[i] >= 0) atomicAdd(&hist[4 * bp[i] + 1 + i], f);
            if (bm[i] >= 0) atomicAdd(&hist[4 * bm[i] + 1 + i], -f);
          }
        }
      } else {
        const float g = base ? gup[0] + gup[1 + b] : 0.f;
        accv += C * g;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
          const float gp = bp[i] >= 0 ? gup[0] + gup[1 + bp[i]] : 0.f;
          const float gm = bm[i] >= 0 ? gup[0] + gup[1 + bm[i]] : 0.f;
          accg[i] += dC[i] * g + C * A.inv_2delta[i] * (gp - gm);
        }
      }
    };
    float l, dl[3];
    float x = kXCut;
    const float rr = d[2] < -A.graze ? -__builtin_amdgcn_rcpf(d[2] * lam) : 0.f;
    if (d[2] < -A.graze) x = z * rr;
    if (x < kXCut) {
      const float q = exp_f(-x);
      const float idz = __builtin_amdgcn_rcpf(d[2]);
      float dx[3], dC[3];
#pragma unroll
      for (int i = 0; i < NS; ++i) dx[i] = rr * dz[i] - x * (ddz[i] * idz + dlnl[i]);
      const float C = w * q;
#pragma unroll
      for (int i = 0; i < NS; ++i) dC[i] = q * dw[i] + C * (S[i] - dx[i]);
      contribute(C, dC);
      // survival weight and truncated free path
      const float U = q + u_path * (1.f - q);
      l = -lam * ln_f(U);
      const float kq = lam * (1.f - u_path) * q * __builtin_amdgcn_rcpf(U);
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        dl[i] = l * dlnl[i] + kq * dx[i];
        dw[i] = dw[i] * (1.f - q) + w * q * dx[i];
      }
      w *= 1.f - q;
    } else {
      l = -lam * ln_f(u_path);
#pragma unroll
      for (int i = 0; i < NS; ++i) dl[i] = l * dlnl[i];
      if (d[2] < 0.f && z + l * d[2] <= 0.f) {
        // grazing flight (|d_z| < graze) that leaves: counted as it happens,
        // the exit indicator itself is not differentiated
        float dC[3];
#pragma unroll
        for (int i = 0; i < NS; ++i) dC[i] = dw[i] + w * S[i];
        contribute(w, dC);
        active = false;
        continue;
      }
    }
    z += l * d[2];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      dz[i] += dl[i] * d[2] + l * ddz[i];
      S[i] += sc[i];
    }

    // ---- energy loss and polar angle ----
    float W = 0.f, dW[3] = {0.f, 0.f, 0.f};
    float c2, s2, ds2[3] = {0.f, 0.f, 0.f};
    bool lost = false;
    auto ruth = [&](float u) {  // Rutherford_halfPol (scattering.py:125-152)
      const float a = c.c_alpha_f * iE;
      const float om = 1.f - u;
      const float inv = __builtin_amdgcn_rcpf(om + a);
      s2 = a * u * inv;
      c2 = om * (1.f + a) * inv;
      const float k = -u * om * inv * inv * a;  // ds2/dlnE
#pragma unroll
      for (int i = 0; i < NS; ++i) ds2[i] = k * dlnE[i];
    };
    if (MODE == MCD_MODE_DS) {
      if (ch == 0) {
        ruth(u_ang);
      } else if (ch == n_ch - 1) {  // Quinn: W = E_pl, no deflection (y_local still turns by the azimuth)
        W = Epl;
        dW[2] = 0.5f * Epl;
        c2 = 1.f; s2 = 0.f;
      } else {
        const DevTable& t = (ch == 1) ? c.tm : c.tg;
        const float Ec = E;
        int row;
        const int col = table_lookup(t, E, [Ec](float v) { return v < Ec; }, u_ang, row);
        W = ldg(t.Ws + col);
        const size_t cell = (size_t)row * (size_t)(t.nW - 1) + col;
        // the column pick's probability depends on the energy (likelihood ratio)
        const float gE = (ch == 1 ? A.m_score_e : A.g_score_e)[cell];
        S[0] += gE * dE[0];
        S[2] += gE * dE[2];
        if (ch == 1) {
          // W = Ws[col] on the grid geomspace(Wc, W_top, nW) moves with Wc,
          // and the column probabilities change with Wc
          dW[0] = W * (1.f - (float)col * A.inv_nw1) * A.inv_wc;
          S[0] += A.m_score_wc[cell];
        }
        if (2.f * W > E - Ef) {  // binaryCollModel NaN branch: direction undefined
          lost = true;
        } else {
          const float v = W * iE;
          const float sq = __builtin_amdgcn_sqrtf(1.f - v);
          c2 = 0.5f * (1.f + sq);
          s2 = 0.5f * v * __builtin_amdgcn_rcpf(1.f + sq);
          const float k = 0.25f * iE * __builtin_amdgcn_rcpf(sq);
#pragma unroll
          for (int i = 0; i < NS; ++i) ds2[i] = k * (dW[i] - v * dE[i]);
        }
      }
    } else {
      // Bethe stopping power at the pre-step energy and its derivatives
      float Sv, dSdE, dSdJ;
      if (c.bethe == MCD_BETHE_CLASSICAL) {
        const float L = ln_f(c.bethe_c1_f * E);
        Sv = c.bethe_k1_f * iE * L;
        dSdE = -Sv * iE + c.bethe_k1_f * iE * iE;
        dSdJ = -c.bethe_k1_f * iE * inv_J;
      } else if (c.bethe == MCD_BETHE_JL) {
        const float L = ln_f(c.bethe_c1_f * (E + kJ));
        const float ie2 = __builtin_amdgcn_rcpf(E + kJ);
        Sv = c.bethe_k1_f * iE * L;
        dSdE = -Sv * iE + c.bethe_k1_f * iE * ie2;
        dSdJ = c.bethe_k1_f * iE * ((float)c.bethe_kk * ie2 - inv_J);
      } else {
        Sv = c.expl_k_f * iE * ln_f(E * c.expl_inv_Estar_f);
        dSdE = -Sv * iE + c.expl_k_f * iE * iE;
        dSdJ = 0.f;
      }
      W = l * Sv;
      const float th[3] = {dSdJ, 0.f, Sv};
#pragma unroll
      for (int i = 0; i < NS; ++i) dW[i] = dl[i] * Sv + l * (dSdE * dE[i] + th[i]);
      ruth(u_ang);
    }
    E -= W;
#pragma unroll
    for (int i = 0; i < NS; ++i) dE[i] -= dW[i];
    step++;
    // alive at theta or at one side of the threshold stencil theta +- delta
    float Ereach = E;
#pragma unroll
    for (int i = 0; i < NS; ++i) Ereach = fmaxf(Ereach, E + fabsf(dE[i] * A.delta[i]));
    if (lost || !(Ereach > Emin) || !(E > 0.f) || !(z > 0.f) || w < kWMin || step >= kBwdStepCap) {
      active = false;
      continue;
    }

    // ---- rotation of (d, y_local) and its tangents ----
    {
      const float ct = c2 - s2;
      const float st = 2.f * __builtin_amdgcn_sqrtf(c2 * s2);
      const float kst = st > 0.f ? 2.f * ct * __builtin_amdgcn_rcpf(st) : 0.f;
      float sp, cp;
      cos_abs_sin_2pi(u_az, cp, sp);
      const float w0 = y[1] * d[2] - y[2] * d[1];
      const float w1 = y[2] * d[0] - y[0] * d[2];
      const float w2 = y[0] * d[1] - y[1] * d[0];
      const float m0 = cp * w0 + sp * y[0], m1 = cp * w1 + sp * y[1], m2 = cp * w2 + sp * y[2];
      float nd[3] = {ct * d[0] + st * m0, ct * d[1] + st * m1, ct * d[2] + st * m2};
      float ny[3] = {cp * y[0] - sp * w0, cp * y[1] - sp * w1, cp * y[2] - sp * w2};
      const float rd = __builtin_amdgcn_rsqf(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
      const float ry = __builtin_amdgcn_rsqf(ny[0] * ny[0] + ny[1] * ny[1] + ny[2] * ny[2]);
#pragma unroll
      for (int j = 0; j < 3; ++j) { nd[j] *= rd; ny[j] *= ry; }
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        if (MODE == MCD_MODE_DS && i == 1) continue;
        const float dct = -2.f * ds2[i], dst = kst * ds2[i];
        const float* a = dd[i];
        const float* b = dy[i];
        const float v0 = b[1] * d[2] - b[2] * d[1] + y[1] * a[2] - y[2] * a[1];
        const float v1 = b[2] * d[0] - b[0] * d[2] + y[2] * a[0] - y[0] * a[2];
        const float v2 = b[0] * d[1] - b[1] * d[0] + y[0] * a[1] - y[1] * a[0];
        float na[3] = {(dct * d[0] + ct * a[0] + dst * m0 + st * (cp * v0 + sp * b[0])) * rd,
                       (dct * d[1] + ct * a[1] + dst * m1 + st * (cp * v1 + sp * b[1])) * rd,
                       (dct * d[2] + ct * a[2] + dst * m2 + st * (cp * v2 + sp * b[2])) * rd};
        float nb[3] = {(cp * b[0] - sp * v0) * ry, (cp * b[1] - sp * v1) * ry, (cp * b[2] - sp * v2) * ry};
        // Keep the tangents on the tangent space of orthonormal frames
        // (d.dd = 0, y.dy = 0, d.dy + y.dd = 0): the exact derivative of the
        // renormalisation.  Components off it are amplified by the rotation
        // map (a y-radial part by sqrt(1 + sin^2 theta) per step), so FP32
        // rounding would otherwise grow without bound.
        const float pa = nd[0] * na[0] + nd[1] * na[1] + nd[2] * na[2];
        const float pb = ny[0] * nb[0] + ny[1] * nb[1] + ny[2] * nb[2];
#pragma unroll
        for (int j = 0; j < 3; ++j) { na[j] -= pa * nd[j]; nb[j] -= pb * ny[j]; }
        const float pc = nd[0] * nb[0] + nd[1] * nb[1] + nd[2] * nb[2] + ny[0] * na[0] + ny[1] * na[1] + ny[2] * na[2];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          dd[i][j] = na[j];
          dy[i][j] = nb[j] - pc * nd[j];
        }
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) { d[j] = nd[j]; y[j] = ny[j]; }
    }
  }

  // ---- reduction ----
  float v[4] = {accv, accg[0], accg[1], accg[2]};
#pragma unroll
  for (int i = 0; i < 4; ++i)
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
  const int* col = A.col_of_slot;
  if (lane == 0) {
    atomicAdd(&A.out[0], (double)v[0]);
#pragma unroll
    for (int i = 0; i < 3; ++i) atomicAdd(&A.out[col[i]], (double)v[1 + i]);
  }
  if (JAC) {
    __syncthreads();
    const int ncol = 1 + MCD_NGRAD;
    for (int i = threadIdx.x; i < 4 * A.n_bins; i += blockDim.x) {
      const int b = i >> 2, q = i & 3;
      const float x = hist[i];
      if (x != 0.f) atomicAdd(&A.out[(1 + b) * ncol + (q == 0 ? 0 : col[q - 1])], (double)x);
    }
  }
}

static size_t bwd_lds_bytes(const BwdArgs& A, bool jac) {
  return sizeof(float) * ((size_t)A.n_nodes * A.stride + (jac ? 4 * (size_t)A.n_bins : 1 + (size_t)A.n_bins));
}

template <int MODE, bool JAC>
static const void* bwd_kernel() {
  return reinterpret_cast<const void*>(&k_backward<MODE, JAC>);
}

static const void* pick_bwd(int mode, bool jac) {
  if (mode == MCD_MODE_DS) return jac ? bwd_kernel<MCD_MODE_DS, true>() : bwd_kernel<MCD_MODE_DS, false>();
  return jac ? bwd_kernel<MCD_MODE_CONT, true>() : bwd_kernel<MCD_MODE_CONT, false>();
}

int bwd_blocks_per_cu(const DevCfg& c, const BwdArgs& A, bool jac, int block) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pick_bwd(c.mode, jac), block, bwd_lds_bytes(A, jac)) !=
      hipSuccess)
    return 1;
  return nb < 1 ? 1 : nb;
}

hipError_t launch_backward(const DevCfg& c, const BwdArgs& A, bool jac, int grid, int block, hipStream_t s) {
  const size_t lds = bwd_lds_bytes(A, jac);
  if (c.mode == MCD_MODE_DS) {
    if (jac) hipLaunchKernelGGL((k_backward<MCD_MODE_DS, true>), dim3(grid), dim3(block), lds, s, c, A);
    else hipLaunchKernelGGL((k_backward<MCD_MODE_DS, false>), dim3(grid), dim3(block), lds, s, c, A);
  } else {
    if (jac) hipLaunchKernelGGL((k_backward<MCD_MODE_CONT, true>), dim3(grid), dim3(block), lds, s, c, A);
    else hipLaunchKernelGGL((k_backward<MCD_MODE_CONT, false>), dim3(grid), dim3(block), lds, s, c, A);
  }
  return hipGetLastError();
}

}  // namespace mcd
