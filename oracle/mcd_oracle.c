/*
 * mcd_oracle.c — FP64 CPU restatement of MC-discrete's electron transport loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * check and bench.py's CPU-baseline leg load it.  The product (libmcd.so) never
 * links or calls it.
 *
 * It restates, in plain C and FP64, the reference's per-step semantics with the
 * reference's random draws replaced by Philox4x32-10 (key = seed, counter =
 * (electron low 32 bits, config | electron bits 32..47 << 16, step, 0)), one
 * call per step:
 *   DS   slots [0]=channel pick, [1]=free path, [2]=W or Rutherford angle, [3]=azimuth
 *   cont slots [0]=free path, [1]=Rutherford angle, [2]=azimuth
 * Every function cites the reference lines it follows (paths relative to the
 * MC-discrete repository root).
 *
 * Pinning (see tests/test_oracle_*.py): Philox against the Random123 known-answer
 * vectors; sigma formulas against scipy quad of the differential cross sections,
 * closed-form limits and hand-evaluated values; the Joy-Luo and explicit Bethe
 * laws and the Quinn sigma against independent forms and limits, with planted
 * mistakes rejected; the rotation against an independent closed form and scipy's
 * Rotation; the pick and table lookup against brute force.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mcd.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* MC/tools/parameters.py:7-8 (bohr_r = scipy CODATA 2018 Bohr radius) */
#define OR_PI_EFOUR 6.51408491409531e-14
#define OR_BOHR_R 5.29177210903e-11
#define OR_PI 3.141592653589793

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), written from the paper. */
static void or_philox(const uint32_t ctr_in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  or_philox(ctr, key[0], key[1], out);
}

/* u = (2*(x>>9)+1) * 2^-24: exact in FP32 and FP64, strictly inside (0,1). */
static double or_u01(uint32_t x) { return (double)(2u * (x >> 9) + 1u) * 5.9604644775390625e-08; }

double oracle_u01(uint32_t x) { return or_u01(x); }

static void or_draw(uint64_t seed, uint64_t electron, uint32_t config, uint32_t step, double u[4]) {
  uint32_t ctr[4] = {(uint32_t)electron, config | (uint32_t)(electron >> 32) << 16, step, 0u};
  uint32_t o[4];
  or_philox(ctr, (uint32_t)seed, (uint32_t)(seed >> 32), o);
  for (int i = 0; i < 4; ++i) u[i] = or_u01(o[i]);
}

/* ------------------------------------------------------------------------- */
/* Cross sections, cm^2 per atom.                                            */

/* alpha: MC/tools/crossSections.py:14-27 */
static double or_alpha(double E, int Z) { return 3.4 * pow((double)Z, 0.666667) / E; }

/* ruther_sigma: crossSections.py:37-58 (E converted to keV, :55) */
double oracle_sigma_ruth(double E, int Z) {
  double a = or_alpha(E, Z);
  double Ek = E * 1e-3;
  double r = (Ek + 511.) / (Ek + 1024.);
  return 5.21e-21 * ((double)(Z * Z) / (Ek * Ek)) * (r * r) * (4. * OR_PI) / (a * (1. + a));
}

/* The elastic models of the input key 'elastic' (scattering.py:483-493):
 * ruther_sigma_wDefl (crossSections.py:62-88) replaces Z^2 by Z (Z + 1);
 * ruther_N_sigma / ruther_N_sigma_wDefl (:91-147) use the Nigram screening
 * beta* = 0.48 * 5.43 Z^(2/3) / E (E in eV) in
 * sigma = pi e^4 Z^2 / (4 beta* (1 + beta*) E^2). */
double oracle_sigma_ruth_model(double E, int Z, int elastic) {
  if (elastic == MCD_ELASTIC_NIGRAM || elastic == MCD_ELASTIC_NIGRAM_WDEFL) {
    double beta_N = 5.43 * pow((double)Z, 2. / 3.) / E;
    double bs = 0.48 * beta_N;
    double num = elastic == MCD_ELASTIC_NIGRAM_WDEFL ? OR_PI_EFOUR * (double)(Z + 1) * (double)Z
                                                     : OR_PI_EFOUR * (double)(Z * Z);
    return num / (4. * bs * (1. + bs) * (E * E));
  }
  double a = or_alpha(E, Z);
  double Ek = E * 1e-3;
  double r = (Ek + 511.) / (Ek + 1024.);
  double zz = elastic == MCD_ELASTIC_RUTH_WDEFL ? (double)(Z * (Z + 1)) : (double)(Z * Z);
  return 5.21e-21 * (zz / (Ek * Ek)) * (r * r) * (4. * OR_PI) / (a * (1. + a));
}

/* moller_sigma: crossSections.py:183-208 */
double oracle_sigma_moller(double E, double Wc, double nfree) {
  if (E <= Wc) return 0.;
  double eps = Wc * 1. / E;
  return nfree * OR_PI_EFOUR * (1. / (E * E)) * (1. / eps - 1. / (1. - eps) + log(eps / (1. - eps)));
}

/* gryz_sigma: crossSections.py:213-242 */
double oracle_sigma_gryz(double E, double Esi, double nsi) {
  if (E <= Esi) return 0.;
  double U = E * 1. / Esi;
  return nsi * OR_PI_EFOUR * pow((U - 1.) / (U + 1.), 1.5) *
         (1. + (2. * (1. - (1. / (2. * U)))) / 3. * log(2.7 + pow(U - 1., 0.5))) / (E * Esi);
}

/* quinn_sigma: crossSections.py:247-284 (m^2 -> cm^2 factor 1e4, :281).
 * The reference raises NameError at E <= Epl; unreachable for E > Emin > Epl. */
double oracle_sigma_quinn(double E, double Epl, double Ef, double n) {
  if (E <= Epl) return 0.;
  double E_Ef = E * 1. / Ef;
  double Epl_Ef = Epl * 1. / Ef;
  double tot = Epl * log((pow(1. + Epl_Ef, 0.5) - 1.) / (pow(E_Ef, 0.5) - pow(E_Ef - Epl_Ef, 0.5))) /
               (2. * OR_BOHR_R * E);
  return (tot / n) * 1e4;
}

/* Insertion-ordered sigma set of scatter_discrete.__init__ (scattering.py:480-508):
 * Ruth, Moller, Gryz<shell> in name_s order, Quinn.  Returns the channel count
 * and writes labels (MCD_CH_*). */
int oracle_sigma_set(const mcd_config* c, double E, double* sig, int* label) {
  const mcd_material* m = &c->mat;
  int k = 0;
  sig[k] = oracle_sigma_ruth_model(E, m->Z, c->elastic); label[k++] = MCD_CH_RUTH;
  sig[k] = oracle_sigma_moller(E, c->Wc, m->n_val); label[k++] = MCD_CH_MOLLER;
  for (int s = 0; s < m->n_shells; ++s) {
    sig[k] = oracle_sigma_gryz(E, m->shell_Eb[s], m->shell_n[s]);
    label[k++] = MCD_CH_GRYZ0 + s;
  }
  sig[k] = oracle_sigma_quinn(E, m->plasmon_e, m->fermi_e, m->atnd); label[k++] = MCD_CH_QUINN;
  return k;
}

/* mfp_from_sigma: scattering.py:19-36; the total is Python's sum() over the
 * dict in insertion order (scattering.py:519). */
static double or_mfp(const double* sig, int n, double atnd) {
  double s = 0.;
  for (int i = 0; i < n; ++i) s += sig[i];
  return 1. / (atnd * s) * 1e14;
}

/* pickFromSigmas: scattering.py:39-59.  Stable descending sort, cumulative sum
 * over the sorted values divided by their (sequential) sum, bisect_left(u).
 * Returns the insertion index of the chosen channel. */
int oracle_pick(const double* sig, int n, double u) {
  int order[MCD_MAX_CHANNELS];
  for (int i = 0; i < n; ++i) order[i] = i;
  /* stable insertion sort, descending (sorted(..., reverse=True) keeps ties in order) */
  for (int i = 1; i < n; ++i) {
    int v = order[i], j = i - 1;
    while (j >= 0 && sig[order[j]] < sig[v]) { order[j + 1] = order[j]; --j; }
    order[j + 1] = v;
  }
  double cum[MCD_MAX_CHANNELS];
  double acc = 0.;
  for (int i = 0; i < n; ++i) { acc += sig[order[i]]; cum[i] = acc; }
  double total = 0.;
  for (int i = 0; i < n; ++i) total += sig[order[i]];
  /* bisect_left on probs = cum/total */
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (cum[mid] / total < u) lo = mid + 1; else hi = mid;
  }
  if (lo >= n) lo = n - 1; /* unreachable: probs[-1] == 1.0 > u */
  return order[lo];
}

/* ------------------------------------------------------------------------- */
/* Table lookup: pickTable, scattering.py:63-94.  Row = bisect_left(Es, E), the
 * first grid energy >= E; column = bisect_left over the CDF row where masked
 * (zero) entries compare as "not less than u" (numpy.ma semantics); W = Ws[col]. */
typedef struct oracle_table {
  int32_t nE, nW;
  const float* Es;
  const float* Ws;
  const float* cdf; /* nE x (nW-1), masked = 0 */
} oracle_table;

int oracle_table_row(const oracle_table* t, double E) {
  int lo = 0, hi = t->nE;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if ((double)t->Es[mid] < E) lo = mid + 1; else hi = mid;
  }
  return lo;
}

int oracle_table_col(const oracle_table* t, int row, double u) {
  const float* r = t->cdf + (size_t)row * (size_t)(t->nW - 1);
  int lo = 0, hi = t->nW - 1;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    float v = r[mid];
    int less = (v != 0.0f) && ((double)v < u);
    if (less) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static double or_pick_table(const oracle_table* t, double E, double u) {
  int row = oracle_table_row(t, E);
  if (row >= t->nE) row = t->nE - 1; /* unreachable: E <= E0 == Es[-1] */
  int col = oracle_table_col(t, row, u);
  return (double)t->Ws[col];
}

/* ------------------------------------------------------------------------- */
/* Direction update: electron.update_direction (electron.py:87-99) and newdir
 * (rotation.py:41-53): q = q_az * q_polar, rotate d and y_local by q's rotation
 * matrix (numpy-quaternion as_rotation_matrix normalises by |q|^2), then
 * renormalise both. */
void oracle_rotate(double c2, double halfPhi, double d[3], double y[3]) {
  double s_ht = pow(1. - c2, 0.5);
  double c_ht = pow(c2, 0.5);
  double s_hp = sin(halfPhi);
  double c_hp = pow(1. - s_hp * s_hp, 0.5);
  /* q_polar = (c_ht, y*s_ht); q_az = (c_hp, d*s_hp) */
  double a1 = c_hp, b1 = d[0] * s_hp, c1 = d[1] * s_hp, d1 = d[2] * s_hp;
  double a2 = c_ht, b2 = y[0] * s_ht, c2q = y[1] * s_ht, d2 = y[2] * s_ht;
  double w = a1 * a2 - b1 * b2 - c1 * c2q - d1 * d2;
  double x = a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2q;
  double yy = a1 * c2q - b1 * d2 + c1 * a2 + d1 * b2;
  double z = a1 * d2 + b1 * c2q - c1 * b2 + d1 * a2;
  double n = w * w + x * x + yy * yy + z * z;
  double M[3][3];
  M[0][0] = 1.0 - 2 * (yy * yy + z * z) / n;
  M[0][1] = 2 * (x * yy - z * w) / n;
  M[0][2] = 2 * (x * z + yy * w) / n;
  M[1][0] = 2 * (x * yy + z * w) / n;
  M[1][1] = 1.0 - 2 * (x * x + z * z) / n;
  M[1][2] = 2 * (yy * z - x * w) / n;
  M[2][0] = 2 * (x * z - yy * w) / n;
  M[2][1] = 2 * (yy * z + x * w) / n;
  M[2][2] = 1.0 - 2 * (x * x + yy * yy) / n;
  double nd[3], ny[3];
  for (int i = 0; i < 3; ++i) {
    nd[i] = M[i][0] * d[0] + M[i][1] * d[1] + M[i][2] * d[2];
    ny[i] = M[i][0] * y[0] + M[i][1] * y[1] + M[i][2] * y[2];
  }
  double ld = sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
  double ly = sqrt(ny[0] * ny[0] + ny[1] * ny[1] + ny[2] * ny[2]);
  for (int i = 0; i < 3; ++i) { d[i] = nd[i] / ld; y[i] = ny[i] / ly; }
}

/* ------------------------------------------------------------------------- */
/* Stopping powers, eV/Angstrom (MC/tools/stoppingPowers.py).  J is an explicit
 * input (the reference derives it from Z at :134-137). */
double oracle_bethe_cl(int Z, double E, double n, double J) { /* :112-140 */
  return 2. * OR_PI_EFOUR * n * ((double)Z / E) * log(1.166 * (E) / J) * 1e-14;
}
double oracle_bethe_jl(int Z, double E, double n, double k, double J) { /* :145-173 */
  return 2. * OR_PI_EFOUR * n * ((double)Z / E) * log(1.166 * (E + k * J) / J) * 1e-14;
}
double oracle_bethe_explicit(double E, double n, const double* Zi, const double* Ei, int ns, double Zval,
                             double Eval) { /* :177-215 */
  double pref = (2. * OR_PI_EFOUR * n / E) * 1e-14;
  double sumi = Zval * log(E * 1. / Eval);
  for (int i = 0; i < ns; ++i) sumi = sumi + Zi[i] * log(E * 1. / Ei[i]);
  return pref * sumi;
}

/* ------------------------------------------------------------------------- */
/* Trajectory loops.                                                         */

typedef struct oracle_els {
  double final_E;
  double dir[3];
  double pos[3];
  uint32_t n_scatt;
  uint8_t outcome;
  uint8_t outcome2;
  uint8_t flags;
  uint8_t pad;
  double path;       /* total free path of the trajectory (sum of l over its steps), a length scale */
} oracle_els;

typedef struct oracle_step {
  uint32_t electron;
  uint32_t step;
  int32_t channel;   /* MCD_CH_* (continuous: MCD_CH_RUTH) */
  int32_t outcome;   /* outcome set during this step (last one), 0 if none */
  double u[4];
  double E;          /* pre-step energy */
  double E_loss;
  double pol;        /* c2 */
  double az;         /* halfPhi */
  double pathl;
  double x, y, z;    /* pre-step position */
  double energy;     /* post-step energy */
  double dx, dy, dz; /* post-step direction */
  double yx, yy, yz; /* post-step y_local */
} oracle_step;

typedef struct or_state {
  double E, xyz[3], dir[3], y[3];
  int outcome, outcome2;
  int nsaves;
  uint8_t flags;
  oracle_els* rec;
} or_state;

/* electron.saveOutcomes (electron.py:105-113): first call fills the record,
 * a second call (absorbed after an exit in the same step) only sets outcome2
 * since outcome, E, dir and position are unchanged in between. */
static void or_save(or_state* s, int outcome) {
  if (s->nsaves == 0) {
    s->rec->outcome = (uint8_t)outcome;
    s->rec->final_E = s->E;
    for (int i = 0; i < 3; ++i) { s->rec->dir[i] = s->dir[i]; s->rec->pos[i] = s->xyz[i]; }
  } else {
    s->rec->outcome2 = (uint8_t)outcome;
  }
  s->nsaves++;
}

/* electron.update_xyz (electron.py:64-84); the diffracting flag is always off
 * (diffMFP is not supported), so a transmission is 'trsm'. */
static void or_update_xyz(or_state* s, double l, double zmax) {
  double np_[3];
  for (int i = 0; i < 3; ++i) np_[i] = s->xyz[i] + l * s->dir[i];
  if (np_[2] <= 0.) { s->outcome = MCD_OUT_BKS; or_save(s, MCD_OUT_BKS); }
  else if (np_[2] >= zmax) { s->outcome = MCD_OUT_TRSM; or_save(s, MCD_OUT_TRSM); }
  else { for (int i = 0; i < 3; ++i) s->xyz[i] = np_[i]; }
}

/* electron.update_energy (electron.py:44-59): saveOutcomes runs before the
 * energy is replaced, so an absorbed electron records its pre-loss energy. */
static void or_update_energy(or_state* s, double W, double Emin) {
  double e = s->E - W;
  if (e <= Emin) { s->outcome = MCD_OUT_ABS; or_save(s, MCD_OUT_ABS); }
  s->E = e;
}

static int or_done(int outcome) { return outcome != MCD_OUT_NONE; }

/* One trajectory.  trajectory_DS (singleScatter.py:10-41) for DS and
 * trajectory_cont_cl (singleScatter.py:205-239, with the Bethe law chosen by
 * the configuration) for continuous mode. */
static void or_trajectory(const mcd_config* c, const oracle_table* tm, const oracle_table* tg, uint64_t seed,
                          uint32_t config_id, uint64_t electron, oracle_els* rec, oracle_step* trace,
                          uint64_t trace_cap, uint64_t* trace_n, uint64_t* events, uint64_t* fp_picks) {
  (void)fp_picks;
  const mcd_material* m = &c->mat;
  or_state s;
  memset(&s, 0, sizeof(s));
  memset(rec, 0, sizeof(*rec));
  s.rec = rec;
  s.E = c->E0;
  /* multiScatter.py:97-100 */
  double t = c->s_tilt * (OR_PI / 180.);
  s.dir[0] = -sin(t); s.dir[1] = 0.; s.dir[2] = cos(t);
  s.y[0] = 0.; s.y[1] = 1.; s.y[2] = 0.; /* electron.py:38 */
  uint64_t nsc = 0;
  uint32_t step = 0;
  double sig[MCD_MAX_CHANNELS];
  int lab[MCD_MAX_CHANNELS];
  while (!or_done(s.outcome)) {
    nsc += 1;
    if (c->max_scatt > 0 && nsc > (uint64_t)c->max_scatt) {
      s.outcome = MCD_OUT_LONG;
      or_save(&s, MCD_OUT_LONG);
      break;
    }
    double u[4];
    or_draw(seed, electron, config_id, step, u);
    double E = s.E;
    double pre[3] = {s.xyz[0], s.xyz[1], s.xyz[2]};
    int ch;
    double l, W, c2, hp;
    if (c->mode == MCD_MODE_DS) {
      int n = oracle_sigma_set(c, E, sig, lab);
      double mfp = or_mfp(sig, n, m->atnd);
      ch = lab[oracle_pick(sig, n, u[0])];                 /* det_type, scattering.py:530-549 */
      l = -mfp * log(u[1]);                                 /* compute_pathl, :551-565 */
      or_update_xyz(&s, l, c->thickness);
      if (ch == MCD_CH_RUTH) {                              /* compute_Eloss_and_angles, :568-679 */
        W = 0.;
        double a = or_alpha(E, m->Z);
        c2 = 1. - (a * u[2] / (1. + a - u[2]));             /* Rutherford_halfPol, :125-152 */
      } else if (ch == MCD_CH_QUINN) {
        W = m->plasmon_e;
        c2 = 1.;
      } else {
        /* Moller table, or the outermost-shell Gryzinski table for every shell (:625) */
        W = or_pick_table(ch == MCD_CH_MOLLER ? tm : tg, E, u[2]);
        if (!(W < E)) s.flags |= MCD_FLAG_LOSS_GE_E;        /* assert, :605,627 */
        /* binaryCollModel, :156-179 */
        if (W <= (E - m->fermi_e) * 0.5) c2 = 0.5 * (sqrt(1 - (W / E)) + 1);
        else { c2 = NAN; s.flags |= MCD_FLAG_NAN_DIR; }
      }
      hp = OR_PI * u[3];
      or_update_energy(&s, W, c->Emin);
    } else {
      double sr = oracle_sigma_ruth_model(E, m->Z, c->elastic);
      double mfp = 1. / (m->atnd * sr) * 1e14;              /* scattering.py:792 */
      ch = MCD_CH_RUTH;
      l = -mfp * log(u[0]);                                 /* :802-813 */
      if (!(l < 1e4)) s.flags |= MCD_FLAG_LONG_PATH;
      or_update_xyz(&s, l, c->thickness);
      double S;
      if (c->bethe_model == MCD_BETHE_CLASSICAL) S = oracle_bethe_cl(m->Z, E, m->atnd, c->J);
      else if (c->bethe_model == MCD_BETHE_JL) S = oracle_bethe_jl(m->Z, E, m->atnd, m->bethe_k, c->J);
      else S = oracle_bethe_explicit(E, m->atnd, m->shell_n, m->shell_Eb, m->n_shells, m->n_val, m->E_val);
      W = l * S;                                            /* compute_Eloss, :815-826 */
      if (!(W < E)) s.flags |= MCD_FLAG_LOSS_GE_E;
      or_update_energy(&s, W, c->Emin);
      double a = or_alpha(E, m->Z);                         /* compute_sAngles, :828-840 */
      c2 = 1. - (a * u[1] / (1. + a - u[1]));
      hp = OR_PI * u[2];
    }
    if (c2 != c2) s.flags |= MCD_FLAG_NAN_DIR;
    rec->path += l;
    oracle_rotate(c2, hp, s.dir, s.y);
    if (trace && *trace_n < trace_cap) {
      oracle_step* r = &trace[*trace_n];
      r->electron = (uint32_t)electron; r->step = step; r->channel = ch; r->outcome = s.outcome;
      for (int i = 0; i < 4; ++i) r->u[i] = u[i];
      r->E = E; r->E_loss = W; r->pol = c2; r->az = hp; r->pathl = l;
      r->x = pre[0]; r->y = pre[1]; r->z = pre[2];
      r->energy = s.E;
      r->dx = s.dir[0]; r->dy = s.dir[1]; r->dz = s.dir[2];
      r->yx = s.y[0]; r->yy = s.y[1]; r->yz = s.y[2];
    }
    if (trace) (*trace_n)++;
    step++;
  }
  rec->n_scatt = (uint32_t)(nsc > 0 && s.outcome == MCD_OUT_LONG ? nsc - 1 : nsc);
  rec->flags = s.flags;
  *events += rec->n_scatt;
}

/* Run electrons [first, first+n).  With a trace buffer the run is serial and
 * per-step rows are appended (trace_n counts rows requested). */
int oracle_run(const mcd_config* c, const oracle_table* tm, const oracle_table* tg, uint64_t seed,
               uint32_t config_id, uint64_t first, uint64_t n, oracle_els* els, oracle_step* trace,
               uint64_t trace_cap, uint64_t* trace_n, uint64_t* events, int nthreads) {
  if (!c || !els || !events) return MCD_ERR_INVALID;
  if (c->mode == MCD_MODE_DS && (!tm || !tg)) return MCD_ERR_INVALID;
  /* the counter holds 48 electron bits beside a 16-bit configuration id */
  if (config_id > 0xFFFFu || first > (1ull << 48) || n > (1ull << 48) - first) return MCD_ERR_INVALID;
  uint64_t ev = 0;
  if (trace || nthreads <= 1) {
    uint64_t tn = trace_n ? *trace_n : 0;
    for (uint64_t i = 0; i < n; ++i)
      or_trajectory(c, tm, tg, seed, config_id, first + i, &els[i], trace, trace_cap, trace ? &tn : NULL, &ev, NULL);
    if (trace_n) *trace_n = tn;
  } else {
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16) reduction(+ : ev)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      uint64_t e = 0;
      or_trajectory(c, tm, tg, seed, config_id, first + (uint64_t)i, &els[i], NULL, 0, NULL, &e, NULL);
      ev += e;
    }
  }
  *events += ev;
  return MCD_OK;
}
