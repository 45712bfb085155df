// mcd_api.hip — the C-ABI of libmcd.so (declared in include/mcd.h).
#include <hip/hip_runtime.h>

#include <math.h>

#include <cmath>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <string>
#include <thread>
#include <vector>

#include "mcd_device.h"
#include "mcd_internal.h"

using namespace mcd;

// Park buffer of the deferred guard for launches on one stream: launches on
// one stream run one after the other, so they can share it; launches on
// different streams get different buffers and may overlap.
struct ParkBuf {
  hipStream_t stream;
  ParkRec* p;
  uint64_t cap;  // electrons
};

// Device scratch of mcd_forward_noalloc: the deferred guard's park buffer for
// up to MCD_NOALLOC_MAX_ELECTRONS electrons, in a file-scope device array that
// each GPU the library runs on holds until the process ends.  Host code never
// uses the array's name as an address: the context keeps the device address
// from hipGetSymbolAddress (ctx->scratch) and passes that to kernels.
constexpr size_t kScratchBytes = (size_t)MCD_NOALLOC_MAX_ELECTRONS * sizeof(ParkRec);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct mcd_ctx {
  int device;
  int n_cu;
  int n_xcc;                    // XCDs of the device (each with its own L2)
  unsigned long long* d_queue;  // [0] backward work counter
  std::vector<ParkBuf> parks;   // deferred-guard park buffers, one per stream
  ParkRec* scratch;             // this device's address of `scratch` (mcd_forward_noalloc)
  char err[512];
};

struct TableDev {
  float* Es = nullptr;
  float* Ws = nullptr;
  float* cdf = nullptr;
  uint8_t* rows = nullptr;
  int nE = 0, nW = 0;
  float E_first = 0.f, inv_h = 0.f;
};

struct mcd_tables {
  int device;
  mcd_config cfg;
  bool has_w;
  // search tolerances and CDF precision, known for tables from
  // mcd_tables_build / mcd_tables_load (keyed = true): what a table file records
  bool keyed;
  double tol_E, tol_W;
  int precision;
  TableDev m, g;
  float* nodes;
  int n_nodes, stride, n_ch;
  double lnE0, dln;
  float delta;
  int n_swap_cells, n_exact_cells;
  // derivative node table, built on the first mcd_backward with these tables
  double moller_n_el;  // electrons per atom of the Moller table
  mutable float* bwd_nodes = nullptr;
  mutable float* bwd_m_swc = nullptr;   // DS: Wc score per Moller table column pick
  mutable float* bwd_m_se = nullptr;    // DS: energy score per Moller table column pick
  mutable float* bwd_g_se = nullptr;    // DS: energy score per Gryzinski table column pick
  mutable double bwd_rel = 0.;          // parameter step the Wc scores were built with
  mutable int bwd_n = 0, bwd_stride = 0, bwd_nch = 0;
  mutable double bwd_lnE0 = 0., bwd_dln = 0.;
};

static thread_local char g_err[512];

static int set_err(mcd_ctx* ctx, int code, const char* fmt, ...) {
  char* buf = ctx ? ctx->err : g_err;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, 512, fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(ctx, expr)                                                                        \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) return set_err(ctx, MCD_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

extern "C" int mcd_version(void) { return MCD_ABI_VERSION; }

extern "C" int mcd_last_error(const mcd_ctx* ctx, char* buf, size_t len) {
  if (!buf || !len) return MCD_ERR_INVALID;
  snprintf(buf, len, "%s", ctx ? ctx->err : g_err);
  return MCD_OK;
}

extern "C" int mcd_ctx_create(int device, mcd_ctx** out) {
  if (!out) return set_err(nullptr, MCD_ERR_INVALID, "mcd_ctx_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(nullptr, hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return set_err(nullptr, MCD_ERR_INVALID, "mcd_ctx_create: device %d not present (%d devices)", device, ndev);
  HIP_TRY(nullptr, hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(nullptr, hipGetDeviceProperties(&prop, device));
  mcd_ctx* c = new mcd_ctx();
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  c->n_xcc = 1;
  if (hipDeviceGetAttribute(&c->n_xcc, hipDeviceAttributeNumberOfXccs, device) != hipSuccess || c->n_xcc < 1)
    c->n_xcc = 1;
  c->err[0] = 0;
  if (hipMalloc(&c->d_queue, 64 * sizeof(unsigned long long)) != hipSuccess) {
    delete c;
    return set_err(nullptr, MCD_ERR_NOMEM, "mcd_ctx_create: hipMalloc of the work counters failed");
  }
  // the device address of the static scratch array on this device (current above)
  void* sp = nullptr;
  const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  if (se != hipSuccess || !sp) {
    (void)hipFree(c->d_queue);
    delete c;
    return set_err(nullptr, MCD_ERR_DEVICE, "mcd_ctx_create: hipGetSymbolAddress(scratch): %s", hipGetErrorString(se));
  }
  c->scratch = static_cast<ParkRec*>(sp);
  *out = c;
  return MCD_OK;
}

extern "C" int mcd_ctx_destroy(mcd_ctx* ctx) {
  if (!ctx) return MCD_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipFree(ctx->d_queue);
  for (ParkBuf& b : ctx->parks)
    if (b.p) (void)hipFree(b.p);
  delete ctx;
  return MCD_OK;
}

// ---------------------------------------------------------------------------
static void pair_of(double v, float& hi, float& lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}

static int check_config(mcd_ctx* ctx, const mcd_config* cfg) {
  if (!cfg) return set_err(ctx, MCD_ERR_INVALID, "config is NULL");
  const mcd_material& m = cfg->mat;
  if (m.n_shells < 1 || m.n_shells > MCD_MAX_SHELLS)
    return set_err(ctx, MCD_ERR_INVALID, "n_shells %d outside [1, %d]", m.n_shells, MCD_MAX_SHELLS);
  if (!(cfg->E0 > cfg->Emin) || !(cfg->Emin > 0.))
    return set_err(ctx, MCD_ERR_INVALID, "need E0 > Emin > 0 (E0=%g, Emin=%g)", cfg->E0, cfg->Emin);
  if (!(m.atnd > 0.) || !(m.fermi_e > 0.) || !(m.plasmon_e > 0.))
    return set_err(ctx, MCD_ERR_INVALID, "material derived constants (atnd, E_F, E_pl) must be positive");
  if (cfg->mode != MCD_MODE_DS && cfg->mode != MCD_MODE_CONT)
    return set_err(ctx, MCD_ERR_INVALID, "unknown mode %d", cfg->mode);
  if (cfg->mode == MCD_MODE_DS && !(cfg->Wc > 0.))
    return set_err(ctx, MCD_ERR_INVALID, "DS mode needs Wc > 0");
  if (cfg->mode == MCD_MODE_CONT && (cfg->bethe_model < 0 || cfg->bethe_model > 2))
    return set_err(ctx, MCD_ERR_INVALID, "unknown Bethe model %d", cfg->bethe_model);
  if (!(cfg->thickness > 0.)) return set_err(ctx, MCD_ERR_INVALID, "thickness must be > 0");
  if (cfg->elastic < MCD_ELASTIC_RUTH || cfg->elastic > MCD_ELASTIC_NIGRAM_WDEFL)
    return set_err(ctx, MCD_ERR_INVALID, "unknown elastic model %d", cfg->elastic);
  if (cfg->mode == MCD_MODE_CONT) {
    // the Bethe laws must remove energy down to Emin, or a trajectory never ends
    const double E = cfg->Emin, J = cfg->J;
    double arg = cfg->bethe_model == MCD_BETHE_CLASSICAL ? 1.166 * E / J
                 : cfg->bethe_model == MCD_BETHE_JL        ? 1.166 * (E + m.bethe_k * J) / J
                                                           : exp(log(E) - cfg->explicit_lnEstar);
    if (!(J > 0.) || !(arg > 1.))
      return set_err(ctx, MCD_ERR_INVALID, "stopping power is not positive at Emin=%g eV for this Bethe model", E);
  }
  return MCD_OK;
}

static DevCfg make_devcfg(const mcd_config& cfg, const mcd_tables* t, uint32_t config_id) {
  DevCfg c;
  memset(&c, 0, sizeof(c));
  const mcd_material& m = cfg.mat;
  c.mode = cfg.mode;
  c.bethe = cfg.bethe_model;
  c.n_shells = m.n_shells;
  c.Z = m.Z;
  c.config_id = config_id;
  // maxScatt off: a cap of 2^24 steps keeps every wavefront finite (a
  // trajectory is ~10^3 steps; the cap only binds on a runaway random walk)
  c.max_scatt = cfg.max_scatt > 0 ? cfg.max_scatt : (int64_t)kMaxStepsOff;
  c.E0 = cfg.E0; c.Emin = cfg.Emin; c.Wc = cfg.Wc;
  c.Ef = m.fermi_e; c.Epl = m.plasmon_e; c.atnd = m.atnd;
  c.thickness = cfg.thickness;
  c.c_alpha = 3.4 * pow((double)m.Z, 0.666667);       // crossSections.py:27
  c.elastic = cfg.elastic;
  c.ZZ = cfg.elastic == MCD_ELASTIC_RUTH_WDEFL ? (double)(m.Z * (m.Z + 1)) : (double)(m.Z * m.Z);
  // Nigram screening (crossSections.py:113-116, 140-145): beta* = 0.48 * 5.43 Z^(2/3) / E
  c.nig_zp = 5.43 * pow((double)m.Z, 2. / 3.);
  c.nig_num = cfg.elastic == MCD_ELASTIC_NIGRAM_WDEFL ? kPiEFour * (double)(m.Z + 1) * (double)m.Z
                                                      : kPiEFour * (double)(m.Z * m.Z);
  c.n_val = m.n_val;
  c.J = cfg.J; c.bethe_kk = m.bethe_k;
  c.zeff = cfg.explicit_Zeff; c.lnEstar = cfg.explicit_lnEstar; c.E_val = m.E_val;
  for (int i = 0; i < MCD_MAX_SHELLS; ++i) { c.Eb[i] = m.shell_Eb[i]; c.ns[i] = m.shell_n[i]; }
  const double tr = cfg.s_tilt * (3.141592653589793 / 180.);  // np.radians
  c.dir0[0] = -sin(tr); c.dir0[1] = 0.; c.dir0[2] = cos(tr);
  pair_of(cfg.E0, c.E0_hi, c.E0_lo);
  pair_of(cfg.Emin, c.Emin_hi, c.Emin_lo);
  pair_of(m.fermi_e, c.Ef2_hi, c.Ef2_lo);
  pair_of(m.plasmon_e, c.Epl_hi, c.Epl_lo);
  c.thickness_f = (float)cfg.thickness;
  c.c_alpha_f = (float)c.c_alpha;
  c.ruth_k_f = (float)(5.21e-21 * c.ZZ * 4. * 3.141592653589793);
  c.mfp_k_f = (float)(1e14 / m.atnd);
  c.nig_zp_f = (float)c.nig_zp;
  c.nig_num_f = (float)c.nig_num;
  c.bethe_k1_f = (float)(2. * kPiEFour * m.atnd * (double)m.Z * 1e-14);
  c.bethe_c1_f = (float)(1.166 / cfg.J);
  c.bethe_kJ_f = (float)(m.bethe_k * cfg.J);
  c.expl_k = 2. * kPiEFour * m.atnd * 1e-14 * cfg.explicit_Zeff;
  c.expl_k_f = (float)c.expl_k;
  c.expl_inv_Estar_f = (float)exp(-cfg.explicit_lnEstar);
  for (int i = 0; i < 3; ++i) c.dir0_f[i] = (float)c.dir0[i];
  if (t && t->has_w) {
    const TableDev* src[2] = {&t->m, &t->g};
    DevTable* dst[2] = {&c.tm, &c.tg};
    for (int q = 0; q < 2; ++q) {
      dst[q]->Es = src[q]->Es; dst[q]->Ws = src[q]->Ws; dst[q]->cdf = src[q]->cdf; dst[q]->rows = src[q]->rows;
      dst[q]->nE = src[q]->nE; dst[q]->nW = src[q]->nW;
      dst[q]->narrow = src[q]->nW - 1 <= 255;
      dst[q]->E_first = src[q]->E_first; dst[q]->inv_h = src[q]->inv_h;
    }
  }
  if (t && t->nodes) {
    c.pick.nodes = t->nodes;
    c.pick.n_nodes = t->n_nodes;
    c.pick.stride = t->stride;
    c.pick.n_ch = t->n_ch;
    c.pick.log2E0 = (float)(t->lnE0 / M_LN2);
    c.pick.inv_dlog2E = (float)(M_LN2 / t->dln);
    c.pick.delta = t->delta;
  }
  return c;
}

static void free_table(TableDev& t) {
  if (t.Es) (void)hipFree(t.Es);
  if (t.Ws) (void)hipFree(t.Ws);
  if (t.cdf) (void)hipFree(t.cdf);
  if (t.rows) (void)hipFree(t.rows);
  t = TableDev();
}

extern "C" int mcd_tables_destroy(mcd_tables* t) {
  if (!t) return MCD_OK;
  (void)hipSetDevice(t->device);
  free_table(t->m);
  free_table(t->g);
  if (t->nodes) (void)hipFree(t->nodes);
  if (t->bwd_nodes) (void)hipFree(t->bwd_nodes);
  if (t->bwd_m_swc) (void)hipFree(t->bwd_m_swc);
  if (t->bwd_m_se) (void)hipFree(t->bwd_m_se);
  if (t->bwd_g_se) (void)hipFree(t->bwd_g_se);
  delete t;
  return MCD_OK;
}

static int upload_grid(mcd_ctx* ctx, const mcd_table_grid* gr, TableDev& td) {
  if (!gr || !gr->Es || !gr->Ws || gr->nE < 2 || gr->nW < 2)
    return set_err(ctx, MCD_ERR_TABLE, "table grid missing or smaller than 2 x 2");
  for (int i = 1; i < gr->nE; ++i)
    if (!(gr->Es[i] > gr->Es[i - 1])) return set_err(ctx, MCD_ERR_TABLE, "table energies not increasing at %d", i);
  for (int i = 1; i < gr->nW; ++i)
    if (!(gr->Ws[i] > gr->Ws[i - 1])) return set_err(ctx, MCD_ERR_TABLE, "table losses not increasing at %d", i);
  td.nE = gr->nE;
  td.nW = gr->nW;
  td.E_first = gr->Es[0];
  td.inv_h = (float)((double)(gr->nE - 1) / ((double)gr->Es[gr->nE - 1] - (double)gr->Es[0]));
  const size_t ncdf = (size_t)gr->nE * (size_t)(gr->nW - 1);
  if (gr->nW - 1 > 65535) return set_err(ctx, MCD_ERR_TABLE, "tables wider than 65535 loss bins are not supported");
  if (hipMalloc(&td.Es, gr->nE * sizeof(float)) != hipSuccess || hipMalloc(&td.Ws, gr->nW * sizeof(float)) != hipSuccess ||
      hipMalloc(&td.cdf, ncdf * sizeof(float)) != hipSuccess ||
      hipMalloc(&td.rows, (size_t)gr->nE * kRowBytes) != hipSuccess)
    return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for a %d x %d table failed", gr->nE, gr->nW - 1);
  HIP_TRY(ctx, hipMemcpy(td.Es, gr->Es, gr->nE * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(td.Ws, gr->Ws, gr->nW * sizeof(float), hipMemcpyHostToDevice));
  return MCD_OK;
}

// Served-guard wait bound of a launch: kGuardSpinCap, or MCD_GUARD_SPIN_CAP
// when set to a positive integer (for tests of the timeout path).
static uint32_t guard_spin_cap() {
  const char* s = getenv("MCD_GUARD_SPIN_CAP");
  const long v = s ? atol(s) : 0;
  return v > 0 && v <= (long)kGuardSpinCap ? (uint32_t)v : kGuardSpinCap;
}

static int pick_nodes_default() {
  const char* s = getenv("MCD_PICK_NODES");
  int n = s ? atoi(s) : 512;
  return n < 16 ? 16 : (n > 8192 ? 8192 : n);
}

static uint32_t label_at(const float* node, int n_ch, int pos) {
  uint32_t lo, hi;
  memcpy(&lo, &node[n_ch], 4);
  memcpy(&hi, &node[n_ch + 1], 4);
  return pos < 8 ? (lo >> (4 * pos)) & 15u : (hi >> (4 * (pos - 8))) & 15u;
}

static void set_meta(float* node, int n_ch, uint32_t swap, bool exact) {
  uint32_t hi;
  memcpy(&hi, &node[n_ch + 1], 4);
  hi = (hi & 0xFFFu) | (swap << 12) | (exact ? kMetaExact : 0u);
  memcpy(&node[n_ch + 1], &hi, 4);
}

// Build the LDS pick tables of n DS configurations (see mcd_device.h): every
// table's nodes in one pass of launches, the cell classification on the host,
// then every table's interpolation check into one buffer read back once.
static int build_pick_tables(mcd_ctx* ctx, mcd_tables* const* ts, const int* n_nodes, int n) {
  std::vector<std::vector<float>> h(n);
  for (int i = 0; i < n; ++i) {
    mcd_tables* t = ts[i];
    const mcd_config& cfg = t->cfg;
    t->n_ch = cfg.mat.n_shells + 3;
    t->stride = pick_stride(t->n_ch);
    t->n_nodes = n_nodes[i];
    if (t->nodes) { (void)hipFree(t->nodes); t->nodes = nullptr; }
    t->lnE0 = log(cfg.Emin);
    t->dln = (log(cfg.E0) - t->lnE0) / (double)(t->n_nodes - 1);
    t->delta = 0.f;
    const size_t nf = (size_t)t->n_nodes * t->stride;
    if (hipMalloc(&t->nodes, nf * sizeof(float)) != hipSuccess)
      return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for the pick table failed");
    DevCfg c = make_devcfg(cfg, t, 0);
    HIP_TRY(ctx, launch_pick_nodes(c, t->n_nodes, t->lnE0, t->dln, t->nodes, t->stride, 0));
  }
  // cell classification from the node orders
  std::vector<std::vector<char>> exact(n);
  for (int i = 0; i < n; ++i) {
    mcd_tables* t = ts[i];
    const size_t nf = (size_t)t->n_nodes * t->stride;
    h[i].resize(nf);
    HIP_TRY(ctx, hipMemcpy(h[i].data(), t->nodes, nf * sizeof(float), hipMemcpyDeviceToHost));
    t->n_swap_cells = t->n_exact_cells = 0;
    exact[i].assign(t->n_nodes, 0);
    for (int k = 0; k + 1 < t->n_nodes; ++k) {
      float* a = &h[i][(size_t)k * t->stride];
      float* b = &h[i][(size_t)(k + 1) * t->stride];
      int ndiff = 0, first = -1;
      for (int q = 0; q < t->n_ch; ++q)
        if (label_at(a, t->n_ch, q) != label_at(b, t->n_ch, q)) { if (first < 0) first = q; ++ndiff; }
      uint32_t swap = kMetaNoSwap;
      bool ex = false;
      if (ndiff == 2 && first + 1 < t->n_ch && label_at(a, t->n_ch, first) == label_at(b, t->n_ch, first + 1) &&
          label_at(a, t->n_ch, first + 1) == label_at(b, t->n_ch, first))
        swap = (uint32_t)first;
      else if (ndiff != 0)
        ex = true;
      for (int q = 0; q < t->stride; ++q)
        if (!(h[i][(size_t)k * t->stride + q] == h[i][(size_t)k * t->stride + q])) ex = true;  // NaN node
      if (swap != kMetaNoSwap) t->n_swap_cells++;
      exact[i][k] = ex;
      set_meta(a, t->n_ch, swap, ex);
    }
    HIP_TRY(ctx, hipMemcpy(t->nodes, h[i].data(), nf * sizeof(float), hipMemcpyHostToDevice));
  }
  // interpolation check against the exact FP64 pick, M points per cell
  const int M = 16;
  std::vector<size_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + (size_t)(ts[i]->n_nodes - 1) * M;
  const size_t npts = off[n];
  float* d_err = nullptr;
  int* d_bad = nullptr;
  if (hipMalloc(&d_err, 2 * npts * sizeof(float)) != hipSuccess || hipMalloc(&d_bad, npts * sizeof(int)) != hipSuccess) {
    (void)hipFree(d_err);
    return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for the pick-table check failed");
  }
  hipError_t e1 = hipSuccess;
  for (int i = 0; i < n && e1 == hipSuccess; ++i) {
    DevCfg c = make_devcfg(ts[i]->cfg, ts[i], 0);
    e1 = launch_pick_check(c, M, d_err + 2 * off[i], d_bad + off[i], 0);
  }
  std::vector<float> herr(2 * npts);
  std::vector<int> hbad(npts);
  hipError_t e2 = e1 == hipSuccess ? hipMemcpy(herr.data(), d_err, 2 * npts * sizeof(float), hipMemcpyDeviceToHost)
                                   : hipSuccess;
  hipError_t e3 = e1 == hipSuccess ? hipMemcpy(hbad.data(), d_bad, npts * sizeof(int), hipMemcpyDeviceToHost)
                                   : hipSuccess;
  (void)hipFree(d_err);
  (void)hipFree(d_bad);
  HIP_TRY(ctx, e1);
  HIP_TRY(ctx, e2);
  HIP_TRY(ctx, e3);
  for (int i = 0; i < n; ++i) {
    mcd_tables* t = ts[i];
    const float* err = herr.data() + 2 * off[i];
    const int* bad_pt = hbad.data() + off[i];
    double maxerr = 0.;
    for (int k = 0; k + 1 < t->n_nodes; ++k) {
      bool bad = exact[i][k];
      double ek = 0.;
      for (int m = 0; m < M; ++m) {
        const size_t j = (size_t)k * M + m;
        if (bad_pt[j]) bad = true;
        ek = fmax(ek, fmax((double)err[2 * j], (double)err[2 * j + 1]));
        if (!(err[2 * j] == err[2 * j]) || !(err[2 * j + 1] == err[2 * j + 1])) bad = true;
      }
      if (ek > 1e-4) bad = true;  // badly resolved cell: always exact
      if (bad) {
        if (!exact[i][k]) {
          uint32_t hi;
          memcpy(&hi, &h[i][(size_t)k * t->stride + t->n_ch + 1], 4);
          set_meta(&h[i][(size_t)k * t->stride], t->n_ch, (hi >> 12) & 15u, true);
        }
        exact[i][k] = 1;
        t->n_exact_cells++;
      } else {
        maxerr = fmax(maxerr, ek);
      }
    }
    t->delta = (float)(2.0 * maxerr + 2e-7);
    HIP_TRY(ctx, hipMemcpy(t->nodes, h[i].data(), h[i].size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return MCD_OK;
}

static int build_pick_table(mcd_ctx* ctx, mcd_tables* t, int n_nodes) { return build_pick_tables(ctx, &t, &n_nodes, 1); }

// New tables for a checked configuration; DS: the grids uploaded and the CDF
// and row-record buffers allocated, not yet filled.
static int tables_new(mcd_ctx* ctx, const mcd_config* cfg, const mcd_table_grid* moller, const mcd_table_grid* gryz,
                      mcd_tables** out) {
  *out = nullptr;
  mcd_tables* t = new mcd_tables();
  memset(t, 0, sizeof(*t));
  t->device = ctx->device;
  t->cfg = *cfg;
  if (cfg->mode == MCD_MODE_DS) {
    if (!moller || !gryz) {
      delete t;
      return set_err(ctx, MCD_ERR_TABLE, "DS mode needs the Moller and Gryzinski table grids");
    }
    int rc;
    if ((rc = upload_grid(ctx, moller, t->m)) || (rc = upload_grid(ctx, gryz, t->g))) {
      mcd_tables_destroy(t);
      return rc;
    }
    t->has_w = true;
    t->moller_n_el = moller->n_el;
  }
  *out = t;
  return MCD_OK;
}

// The channel-pick table of DS tables whose CDFs and row records are built
// (or queued on the null stream), then the wait for the whole build.
static int tables_finish(mcd_ctx* ctx, mcd_tables* t) {
  if (!t->has_w) return MCD_OK;
  int rc = build_pick_table(ctx, t, pick_nodes_default());
  if (rc) return rc;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return set_err(ctx, MCD_ERR_DEVICE, "table build: %s", hipGetErrorString(e));
  return MCD_OK;
}

extern "C" int mcd_tables_create(mcd_ctx* ctx, const mcd_config* cfg, const mcd_table_grid* moller,
                                 const mcd_table_grid* gryz, int precision, mcd_tables** out) {
  if (!ctx || !out) return set_err(ctx, MCD_ERR_INVALID, "mcd_tables_create: NULL argument");
  *out = nullptr;
  int rc = check_config(ctx, cfg);
  if (rc) return rc;
  if (precision != MCD_PREC_FP64 && precision != MCD_PREC_FP32)
    return set_err(ctx, MCD_ERR_INVALID, "unknown table precision %d", precision);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  mcd_tables* t = nullptr;
  if ((rc = tables_new(ctx, cfg, moller, gryz, &t))) return rc;
  if (t->has_w) {
    const mcd_table_grid* gr[2] = {moller, gryz};
    TableDev* td[2] = {&t->m, &t->g};
    for (int q = 0; q < 2; ++q) {
      hipError_t e = launch_build_cdf(td[q]->Es, td[q]->Ws, td[q]->nE, td[q]->nW, cfg->mat.fermi_e, gr[q]->kind,
                                      gr[q]->n_el, gr[q]->Eb, precision, td[q]->cdf, 0);
      if (e != hipSuccess) {
        mcd_tables_destroy(t);
        return set_err(ctx, MCD_ERR_DEVICE, "CDF build: %s", hipGetErrorString(e));
      }
    }
    for (int q = 0; q < 2; ++q) {
      hipError_t e = launch_build_rows(td[q]->cdf, td[q]->Es, td[q]->nE, td[q]->nW, td[q]->rows, 0);
      if (e != hipSuccess) {
        mcd_tables_destroy(t);
        return set_err(ctx, MCD_ERR_DEVICE, "row-record build: %s", hipGetErrorString(e));
      }
    }
  }
  if ((rc = tables_finish(ctx, t))) {
    mcd_tables_destroy(t);
    return rc;
  }
  *out = t;
  return MCD_OK;
}

// Same as mcd_tables_create, then replace the device CDFs with host CDFs in the
// reference layout (masked = 0; probTable.readFromMemory, probTables.py:513-535).
extern "C" int mcd_tables_upload(mcd_ctx* ctx, const mcd_config* cfg, const mcd_table_grid* moller,
                                     const float* cdf_moller, const mcd_table_grid* gryz, const float* cdf_gryz,
                                     mcd_tables** out) {
  int rc = mcd_tables_create(ctx, cfg, moller, gryz, MCD_PREC_FP64, out);
  if (rc) return rc;
  mcd_tables* t = *out;
  if (!t->has_w) return MCD_OK;
  const float* src[2] = {cdf_moller, cdf_gryz};
  TableDev* td[2] = {&t->m, &t->g};
  for (int q = 0; q < 2; ++q) {
    if (!src[q]) continue;
    const size_t n = (size_t)td[q]->nE * (size_t)(td[q]->nW - 1);
    float* tmp = nullptr;
    if (hipMalloc(&tmp, n * sizeof(float)) != hipSuccess) {
      mcd_tables_destroy(t);
      *out = nullptr;
      return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for the CDF upload failed");
    }
    hipError_t e = hipMemcpy(tmp, src[q], n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_mask(tmp, td[q]->cdf, n, 0);
    if (e == hipSuccess) e = launch_build_rows(td[q]->cdf, td[q]->Es, td[q]->nE, td[q]->nW, td[q]->rows, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(tmp);
    if (e != hipSuccess) {
      mcd_tables_destroy(t);
      *out = nullptr;
      return set_err(ctx, MCD_ERR_DEVICE, "CDF upload: %s", hipGetErrorString(e));
    }
  }
  return MCD_OK;
}

extern "C" int mcd_tables_download(mcd_ctx* ctx, const mcd_tables* t, int which, float* cdf_host, size_t n_floats) {
  if (!ctx || !t || !cdf_host) return set_err(ctx, MCD_ERR_INVALID, "mcd_tables_download: NULL argument");
  if (!t->has_w) return set_err(ctx, MCD_ERR_INVALID, "tables hold no energy-loss tables (continuous mode)");
  const TableDev& td = which == MCD_TABLE_MOLLER ? t->m : t->g;
  const size_t n = (size_t)td.nE * (size_t)(td.nW - 1);
  if (n_floats < n) return set_err(ctx, MCD_ERR_INVALID, "buffer holds %zu floats, table has %zu", n_floats, n);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  float* tmp = nullptr;
  if (hipMalloc(&tmp, n * sizeof(float)) != hipSuccess) return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc failed");
  hipError_t e = launch_unmask(td.cdf, tmp, n, 0);
  if (e == hipSuccess) e = hipMemcpy(cdf_host, tmp, n * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(tmp);
  HIP_TRY(ctx, e);
  return MCD_OK;
}

// ---------------------------------------------------------------------------
// Tables from input-file parameters: the tolerance-driven grid search
// (mcd_grid.hip) and the device CDF fill (probTables.genTables, :541-614).
// The two grids of a DS configuration; no context, so searches of several
// configurations may run on separate threads.  err receives the reason of a
// failure.
// The two table searches of a DS configuration: Moller, losses over
// [Wc, W_max(E0)] and energies over [Emin, E0] (probTables.py:557-575); the
// outermost shell's Gryzinski table, the one every Gryzinski event reads
// (probTables.py:579, scattering.py:625), energies from max(Emin, E_b).
static void config_grid_jobs(const mcd_config* cfg, GridSearchJob j[2]) {
  const mcd_material& m = cfg->mat;
  j[0] = {MCD_TABLE_MOLLER, cfg->Wc, m.n_val, 0.0, m.fermi_e, cfg->Emin, cfg->E0};
  const int sh = m.n_shells - 1;
  const double Eb = m.shell_Eb[sh], ns = m.shell_n[sh];
  j[1] = {MCD_TABLE_GRYZ, Eb, ns, Eb, m.fermi_e, cfg->Emin < Eb ? Eb : cfg->Emin, cfg->E0};
}

static void grid_of_job(const GridSearchJob& j, std::vector<float>& E, std::vector<float>& W, mcd_table_grid* g) {
  *g = {(int32_t)E.size(), (int32_t)W.size(), E.data(), W.data(), j.Wmin, j.n_el, j.Eb, j.kind};
}

static int config_grids(const mcd_config* cfg, double tol_E, double tol_W, std::vector<float>* E,
                        std::vector<float>* W, mcd_table_grid* g, char* err, size_t err_len) {
  GridSearchJob j[2];
  config_grid_jobs(cfg, j);
  char why[200];
  for (int q = 0; q < 2; ++q) {
    const int rc = search_grid(j[q].kind, j[q].Wmin, j[q].n_el, j[q].Eb, j[q].Ef, j[q].Emin, j[q].Emax, tol_E, tol_W,
                               E[q], W[q], why, sizeof why);
    if (rc) {
      snprintf(err, err_len, "%s grid search: %s", q == 0 ? "Moller" : "Gryzinski", why);
      return rc;
    }
  }
  for (int q = 0; q < 2; ++q) grid_of_job(j[q], E[q], W[q], &g[q]);
  return MCD_OK;
}

// The device search of n tables (one launch per 256 tables: the scratch is
// about 0.55 MB per table); res[i] per job.
static int device_grid_search(mcd_ctx* ctx, const std::vector<GridSearchJob>& jobs, double tol_E, double tol_W,
                              std::vector<GridSearchResult>& res) {
  const int n = (int)jobs.size();
  res.assign(n, GridSearchResult{0, 0, 0, 0, 0.});
  if (n == 0) return MCD_OK;
  constexpr int kChunk = 256;
  const int m = std::min(n, kChunk);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  GridSearchJob* d_jobs = nullptr;
  GridSearchResult* d_res = nullptr;
  void* d_scr = nullptr;
  hipError_t e = hipMalloc(&d_jobs, n * sizeof(GridSearchJob));
  if (e == hipSuccess) e = hipMalloc(&d_res, n * sizeof(GridSearchResult));
  if (e == hipSuccess) e = hipMalloc(&d_scr, grid_search_scratch_bytes(m));
  if (e == hipSuccess) e = hipMemcpy(d_jobs, jobs.data(), n * sizeof(GridSearchJob), hipMemcpyHostToDevice);
  for (int i = 0; i < n && e == hipSuccess; i += kChunk)
    e = launch_grid_search(d_jobs + i, std::min(kChunk, n - i), tol_E, tol_W, d_scr, d_res + i, 0);
  if (e == hipSuccess) e = hipMemcpy(res.data(), d_res, n * sizeof(GridSearchResult), hipMemcpyDeviceToHost);
  (void)hipFree(d_jobs);
  (void)hipFree(d_res);
  (void)hipFree(d_scr);
  if (e != hipSuccess) return set_err(ctx, MCD_ERR_DEVICE, "device grid search: %s", hipGetErrorString(e));
  return MCD_OK;
}

static int build_grids(mcd_ctx* ctx, const mcd_config* cfg, double tol_E, double tol_W, std::vector<float>* E,
                       std::vector<float>* W, mcd_table_grid* g) {
  char err[256];
  int rc = config_grids(cfg, tol_E, tol_W, E, W, g, err, sizeof err);
  if (rc) return set_err(ctx, rc, "%s", err);
  return MCD_OK;
}

extern "C" int mcd_grid_search(int kind, double Wmin, double n_el, double Eb, double Ef, double Emin, double Emax,
                               double tol_E, double tol_W, float* Es, int32_t* nE, float* Ws, int32_t* nW) {
  if (!nE || !nW) return set_err(nullptr, MCD_ERR_INVALID, "mcd_grid_search: NULL size argument");
  if (kind != MCD_TABLE_MOLLER && kind != MCD_TABLE_GRYZ)
    return set_err(nullptr, MCD_ERR_INVALID, "unknown table kind %d", kind);
  if (!(tol_E > 0.0 && tol_E < 1.0) || !(tol_W > 0.0 && tol_W < 1.0) || !(Wmin > 0.0) || !(Emax > Emin) ||
      !(Emin > 0.0))
    return set_err(nullptr, MCD_ERR_INVALID, "grid search: invalid range or tolerance");
  std::vector<float> E, W;
  char err[256];
  int rc = search_grid(kind, Wmin, n_el, Eb, Ef, Emin, Emax, tol_E, tol_W, E, W, err, sizeof err);
  if (rc) return set_err(nullptr, rc, "%s", err);
  const int32_t capE = *nE, capW = *nW;
  *nE = (int32_t)E.size();
  *nW = (int32_t)W.size();
  if (!Es && !Ws) return MCD_OK;
  if (!Es || !Ws || capE < *nE || capW < *nW)
    return set_err(nullptr, MCD_ERR_INVALID, "grid buffers hold %d and %d values, the grids need %d and %d", capE,
                   capW, *nE, *nW);
  memcpy(Es, E.data(), E.size() * sizeof(float));
  memcpy(Ws, W.data(), W.size() * sizeof(float));
  return MCD_OK;
}

extern "C" int mcd_grid_search_batch(mcd_ctx* ctx, const mcd_grid_spec* specs, int n, double tol_E, double tol_W,
                                     int32_t* nE, int32_t* nW, int32_t* status) {
  if (!ctx || (n > 0 && (!specs || !nE || !nW)) || n < 0)
    return set_err(ctx, MCD_ERR_INVALID, "mcd_grid_search_batch: NULL argument");
  if (!(tol_E > 0.0 && tol_E < 1.0) || !(tol_W > 0.0 && tol_W < 1.0))
    return set_err(ctx, MCD_ERR_INVALID, "tolerances tol_E = %g, tol_W = %g outside (0, 1)", tol_E, tol_W);
  std::vector<GridSearchJob> jobs(n);
  for (int i = 0; i < n; ++i) {
    const mcd_grid_spec& p = specs[i];
    if (p.kind != MCD_TABLE_MOLLER && p.kind != MCD_TABLE_GRYZ)
      return set_err(ctx, MCD_ERR_INVALID, "table %d: unknown table kind %d", i, p.kind);
    if (!(p.Wmin > 0.0) || !(p.Emax > p.Emin) || !(p.Emin > 0.0))
      return set_err(ctx, MCD_ERR_INVALID, "table %d: invalid range", i);
    jobs[i] = {p.kind, p.Wmin, p.n_el, p.Eb, p.Ef, p.Emin, p.Emax};
  }
  std::vector<GridSearchResult> res;
  int rc = device_grid_search(ctx, jobs, tol_E, tol_W, res);
  if (rc) return rc;
  int first_bad = -1;
  for (int i = 0; i < n; ++i) {
    nE[i] = res[i].nE;
    nW[i] = res[i].nW;
    if (status) status[i] = res[i].status ? MCD_ERR_TABLE : MCD_OK;
    if (res[i].status && first_bad < 0) first_bad = i;
  }
  // with a status array the per-table codes are the report and the call succeeds
  if (first_bad >= 0 && !status)
    return set_err(ctx, MCD_ERR_TABLE, "table %d: %s", first_bad, grid_status_text(res[first_bad].status));
  return MCD_OK;
}

extern "C" int mcd_tables_build(mcd_ctx* ctx, const mcd_config* cfg, double tol_E, double tol_W, int precision,
                                mcd_tables** out) {
  if (!ctx || !cfg || !out) return set_err(ctx, MCD_ERR_INVALID, "mcd_tables_build: NULL argument");
  *out = nullptr;
  if (!(tol_E > 0.0 && tol_E < 1.0) || !(tol_W > 0.0 && tol_W < 1.0))
    return set_err(ctx, MCD_ERR_INVALID, "tolerances tol_E = %g, tol_W = %g outside (0, 1)", tol_E, tol_W);
  int rc = check_config(ctx, cfg);
  if (rc) return rc;
  if (cfg->mode != MCD_MODE_DS) {
    rc = mcd_tables_create(ctx, cfg, nullptr, nullptr, precision, out);
  } else {
    std::vector<float> E[2], W[2];
    mcd_table_grid g[2];
    if ((rc = build_grids(ctx, cfg, tol_E, tol_W, E, W, g))) return rc;
    rc = mcd_tables_create(ctx, cfg, &g[0], &g[1], precision, out);
  }
  if (rc) return rc;
  (*out)->keyed = true;
  (*out)->tol_E = tol_E;
  (*out)->tol_W = tol_W;
  (*out)->precision = precision;
  return MCD_OK;
}

// Tables of many configurations: the grid searches of every table in one
// device launch (on host threads with MCD_GRID_SEARCH=host), then the CDF rows
// of every table in one launch and their row records in another.
extern "C" int mcd_tables_build_batch(mcd_ctx* ctx, const mcd_config* cfgs, int n, double tol_E, double tol_W,
                                      int precision, mcd_tables** out) {
  if (!ctx || !cfgs || !out || n < 0) return set_err(ctx, MCD_ERR_INVALID, "mcd_tables_build_batch: NULL argument");
  for (int i = 0; i < n; ++i) out[i] = nullptr;
  if (!(tol_E > 0.0 && tol_E < 1.0) || !(tol_W > 0.0 && tol_W < 1.0))
    return set_err(ctx, MCD_ERR_INVALID, "tolerances tol_E = %g, tol_W = %g outside (0, 1)", tol_E, tol_W);
  if (precision != MCD_PREC_FP64 && precision != MCD_PREC_FP32)
    return set_err(ctx, MCD_ERR_INVALID, "unknown table precision %d", precision);
  for (int i = 0; i < n; ++i) {
    int rc = check_config(ctx, &cfgs[i]);
    if (rc) {
      char msg[256];
      snprintf(msg, sizeof msg, "%s", ctx->err);
      return set_err(ctx, rc, "configuration %d: %s", i, msg);
    }
  }
  if (n == 0) return MCD_OK;
  std::vector<std::array<std::vector<float>, 2>> E(n), W(n);
  std::vector<std::array<mcd_table_grid, 2>> g(n);
  const char* where = getenv("MCD_GRID_SEARCH");
  if (where && !strcmp(where, "host")) {
    // grid searches on host threads, configurations dealt round-robin
    std::vector<int> rcs(n, MCD_OK);
    std::vector<std::array<char, 256>> errs(n);
    unsigned hw = std::thread::hardware_concurrency();
    const int n_threads = (int)std::max(1u, std::min({hw ? hw : 1u, 16u, (unsigned)n}));
    auto work = [&](int tid) {
      for (int i = tid; i < n; i += n_threads)
        if (cfgs[i].mode == MCD_MODE_DS)
          rcs[i] = config_grids(&cfgs[i], tol_E, tol_W, E[i].data(), W[i].data(), g[i].data(), errs[i].data(),
                                errs[i].size());
    };
    {
      std::vector<std::thread> pool;
      for (int t = 1; t < n_threads; ++t) pool.emplace_back(work, t);
      work(0);
      for (std::thread& th : pool) th.join();
    }
    for (int i = 0; i < n; ++i)
      if (rcs[i]) return set_err(ctx, rcs[i], "configuration %d: %s", i, errs[i].data());
  } else {
    // every table's search in one device launch, the grids from their sizes
    std::vector<GridSearchJob> jobs;
    std::vector<int> of;
    for (int i = 0; i < n; ++i)
      if (cfgs[i].mode == MCD_MODE_DS) {
        GridSearchJob j[2];
        config_grid_jobs(&cfgs[i], j);
        jobs.push_back(j[0]);
        jobs.push_back(j[1]);
        of.push_back(i);
      }
    std::vector<GridSearchResult> res;
    int rc = device_grid_search(ctx, jobs, tol_E, tol_W, res);
    if (rc) return rc;
    for (size_t k = 0; k < of.size(); ++k) {
      const int i = of[k];
      for (int q = 0; q < 2; ++q) {
        const GridSearchResult& r = res[2 * k + q];
        if (r.status)
          return set_err(ctx, MCD_ERR_TABLE, "configuration %d: %s grid search: %s", i,
                         q == 0 ? "Moller" : "Gryzinski", grid_status_text(r.status));
        grid_materialize(jobs[2 * k + q], r.nE, r.nW, E[i][q], W[i][q]);
        grid_of_job(jobs[2 * k + q], E[i][q], W[i][q], &g[i][q]);
      }
    }
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  auto fail = [&](int rc) {
    for (int i = 0; i < n; ++i)
      if (out[i]) { mcd_tables_destroy(out[i]); out[i] = nullptr; }
    return rc;
  };
  std::vector<CdfJob> jobs;
  long long rows = 0;
  for (int i = 0; i < n; ++i) {
    const bool ds = cfgs[i].mode == MCD_MODE_DS;
    int rc = tables_new(ctx, &cfgs[i], ds ? &g[i][0] : nullptr, ds ? &g[i][1] : nullptr, &out[i]);
    if (rc) return fail(rc);
    if (!ds) continue;
    const TableDev* td[2] = {&out[i]->m, &out[i]->g};
    for (int q = 0; q < 2; ++q) {
      CdfJob j;
      j.Es = td[q]->Es;
      j.Ws = td[q]->Ws;
      j.cdf = td[q]->cdf;
      j.rows = td[q]->rows;
      j.nE = td[q]->nE;
      j.nW = td[q]->nW;
      j.kind = g[i][q].kind;
      j.Ef = cfgs[i].mat.fermi_e;
      j.n_el = g[i][q].n_el;
      j.Eb = g[i][q].Eb;
      j.row0 = rows;
      rows += j.nE;
      jobs.push_back(j);
    }
  }
  if (!jobs.empty()) {
    CdfJob* d_jobs = nullptr;
    if (hipMalloc(&d_jobs, jobs.size() * sizeof(CdfJob)) != hipSuccess)
      return fail(set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for %zu table jobs failed", jobs.size()));
    hipError_t e = hipMemcpy(d_jobs, jobs.data(), jobs.size() * sizeof(CdfJob), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_build_tables_batched(d_jobs, (int)jobs.size(), rows, precision, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(d_jobs);
    if (e != hipSuccess) return fail(set_err(ctx, MCD_ERR_DEVICE, "batched table build: %s", hipGetErrorString(e)));
  }
  {
    // the channel-pick tables of every DS configuration together
    std::vector<mcd_tables*> ds;
    for (int i = 0; i < n; ++i)
      if (out[i]->has_w) ds.push_back(out[i]);
    const std::vector<int> nodes(ds.size(), pick_nodes_default());
    int rc = ds.empty() ? MCD_OK : build_pick_tables(ctx, ds.data(), nodes.data(), (int)ds.size());
    if (rc) return fail(rc);
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(set_err(ctx, MCD_ERR_DEVICE, "table build: %s", hipGetErrorString(e)));
  }
  for (int i = 0; i < n; ++i) {
    out[i]->keyed = true;
    out[i]->tol_E = tol_E;
    out[i]->tol_W = tol_W;
    out[i]->precision = precision;
  }
  return MCD_OK;
}

extern "C" int mcd_tables_grid(mcd_ctx* ctx, const mcd_tables* t, int which, int32_t* nE, int32_t* nW, float* Es,
                               float* Ws) {
  if (!ctx || !t || !nE || !nW) return set_err(ctx, MCD_ERR_INVALID, "mcd_tables_grid: NULL argument");
  if (!t->has_w) return set_err(ctx, MCD_ERR_INVALID, "tables hold no energy-loss tables (continuous mode)");
  if (which != MCD_TABLE_MOLLER && which != MCD_TABLE_GRYZ)
    return set_err(ctx, MCD_ERR_INVALID, "unknown table %d", which);
  const TableDev& td = which == MCD_TABLE_MOLLER ? t->m : t->g;
  *nE = td.nE;
  *nW = td.nW;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (Es) HIP_TRY(ctx, hipMemcpy(Es, td.Es, td.nE * sizeof(float), hipMemcpyDeviceToHost));
  if (Ws) HIP_TRY(ctx, hipMemcpy(Ws, td.Ws, td.nW * sizeof(float), hipMemcpyDeviceToHost));
  return MCD_OK;
}

// Table files.  Little-endian: the 8-byte magic, the key (every parameter the
// tables depend on), then per table nE, nW, Wmin, n_el, Eb, Es, Ws and the
// (nE x (nW-1)) float32 CDF in the reference layout (masked = 0).
namespace {
const char kTableMagic[8] = {'M', 'C', 'D', 'T', 'A', 'B', '0', '1'};
struct TableKey {
  int32_t Z, precision;
  double E0, Emin, Wc, n_val, fermi_e, Eb_outer, ns_outer, tol_E, tol_W;
};
TableKey table_key(const mcd_config* c, double tol_E, double tol_W, int precision) {
  TableKey k;
  memset(&k, 0, sizeof k);
  const mcd_material& m = c->mat;
  k.Z = m.Z;
  k.precision = precision;
  k.E0 = c->E0;
  k.Emin = c->Emin;
  k.Wc = c->Wc;
  k.n_val = m.n_val;
  k.fermi_e = m.fermi_e;
  k.Eb_outer = m.n_shells > 0 ? m.shell_Eb[m.n_shells - 1] : 0.0;
  k.ns_outer = m.n_shells > 0 ? m.shell_n[m.n_shells - 1] : 0.0;
  k.tol_E = tol_E;
  k.tol_W = tol_W;
  return k;
}
}  // namespace

extern "C" int mcd_tables_save(mcd_ctx* ctx, const mcd_tables* t, const char* path) {
  if (!ctx || !t || !path) return set_err(ctx, MCD_ERR_INVALID, "mcd_tables_save: NULL argument");
  if (!t->has_w) return set_err(ctx, MCD_ERR_INVALID, "tables hold no energy-loss tables (continuous mode)");
  if (!t->keyed)
    return set_err(ctx, MCD_ERR_INVALID, "only tables from mcd_tables_build or mcd_tables_load carry the "
                                         "tolerances a table file records");
  const TableKey k = table_key(&t->cfg, t->tol_E, t->tol_W, t->precision);
  std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return set_err(ctx, MCD_ERR_INVALID, "cannot write %s", tmp.c_str());
  bool ok = fwrite(kTableMagic, 1, 8, f) == 8 && fwrite(&k, sizeof k, 1, f) == 1;
  const TableDev* td[2] = {&t->m, &t->g};
  const double gparam[2][3] = {{t->cfg.Wc, t->moller_n_el, 0.0}, {k.Eb_outer, k.ns_outer, k.Eb_outer}};
  for (int q = 0; q < 2 && ok; ++q) {
    const int32_t dims[2] = {td[q]->nE, td[q]->nW};
    std::vector<float> Es(td[q]->nE), Ws(td[q]->nW), cdf((size_t)td[q]->nE * (td[q]->nW - 1));
    int32_t a, b;
    int rc = mcd_tables_grid(ctx, t, q, &a, &b, Es.data(), Ws.data());
    if (!rc) rc = mcd_tables_download(ctx, t, q, cdf.data(), cdf.size());
    if (rc) {
      fclose(f);
      remove(tmp.c_str());
      return rc;
    }
    ok = fwrite(dims, sizeof dims, 1, f) == 1 && fwrite(gparam[q], sizeof gparam[q], 1, f) == 1 &&
         fwrite(Es.data(), sizeof(float), Es.size(), f) == Es.size() &&
         fwrite(Ws.data(), sizeof(float), Ws.size(), f) == Ws.size() &&
         fwrite(cdf.data(), sizeof(float), cdf.size(), f) == cdf.size();
  }
  ok = (fclose(f) == 0) && ok;
  if (!ok || rename(tmp.c_str(), path) != 0) {
    remove(tmp.c_str());
    return set_err(ctx, MCD_ERR_INVALID, "writing the table file %s failed", path);
  }
  return MCD_OK;
}

extern "C" int mcd_tables_load(mcd_ctx* ctx, const mcd_config* cfg, double tol_E, double tol_W, int precision,
                               const char* path, mcd_tables** out) {
  if (!ctx || !cfg || !path || !out) return set_err(ctx, MCD_ERR_INVALID, "mcd_tables_load: NULL argument");
  *out = nullptr;
  int rc = check_config(ctx, cfg);
  if (rc) return rc;
  if (cfg->mode != MCD_MODE_DS) return set_err(ctx, MCD_ERR_INVALID, "continuous mode reads no table file");
  FILE* f = fopen(path, "rb");
  if (!f) return set_err(ctx, MCD_ERR_TABLE, "no table file %s", path);
  char magic[8];
  TableKey k;
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kTableMagic, 8) == 0 && fread(&k, sizeof k, 1, f) == 1;
  if (!ok) {
    fclose(f);
    return set_err(ctx, MCD_ERR_TABLE, "%s is not a table file", path);
  }
  const TableKey want = table_key(cfg, tol_E, tol_W, precision);
  if (memcmp(&k, &want, sizeof k) != 0) {
    fclose(f);
    return set_err(ctx, MCD_ERR_TABLE,
                   "table file %s was built for other parameters (Z %d E0 %g Emin %g Wc %g tol_E %g tol_W %g "
                   "precision %d; asked Z %d E0 %g Emin %g Wc %g tol_E %g tol_W %g precision %d)",
                   path, k.Z, k.E0, k.Emin, k.Wc, k.tol_E, k.tol_W, k.precision, want.Z, want.E0, want.Emin, want.Wc,
                   want.tol_E, want.tol_W, want.precision);
  }
  std::vector<float> E[2], W[2], cdf[2];
  mcd_table_grid g[2];
  for (int q = 0; q < 2 && ok; ++q) {
    int32_t dims[2];
    double gp[3];
    ok = fread(dims, sizeof dims, 1, f) == 1 && fread(gp, sizeof gp, 1, f) == 1 && dims[0] >= 2 && dims[1] >= 2 &&
         dims[0] <= (1 << 24) && dims[1] <= 65536;
    if (!ok) break;
    E[q].resize(dims[0]);
    W[q].resize(dims[1]);
    cdf[q].resize((size_t)dims[0] * (dims[1] - 1));
    ok = fread(E[q].data(), sizeof(float), E[q].size(), f) == E[q].size() &&
         fread(W[q].data(), sizeof(float), W[q].size(), f) == W[q].size() &&
         fread(cdf[q].data(), sizeof(float), cdf[q].size(), f) == cdf[q].size();
    g[q] = {dims[0], dims[1], E[q].data(), W[q].data(), gp[0], gp[1], gp[2], q};
  }
  ok = ok && fgetc(f) == EOF;
  fclose(f);
  if (!ok) return set_err(ctx, MCD_ERR_TABLE, "table file %s is truncated or malformed", path);
  rc = mcd_tables_upload(ctx, cfg, &g[0], cdf[0].data(), &g[1], cdf[1].data(), out);
  if (rc) return rc;
  (*out)->keyed = true;
  (*out)->tol_E = tol_E;
  (*out)->tol_W = tol_W;
  (*out)->precision = precision;
  return MCD_OK;
}

// Pick-table diagnostics: n_nodes, stride, swap cells, exact cells, delta.
extern "C" int mcd_tables_info(const mcd_tables* t, int* n_nodes, int* n_swap, int* n_exact, float* delta) {
  if (!t) return MCD_ERR_INVALID;
  if (n_nodes) *n_nodes = t->n_nodes;
  if (n_swap) *n_swap = t->n_swap_cells;
  if (n_exact) *n_exact = t->n_exact_cells;
  if (delta) *delta = t->delta;
  return MCD_OK;
}

extern "C" int mcd_tables_set_pick_nodes(mcd_ctx* ctx, mcd_tables* t, int n_nodes) {
  if (!ctx || !t) return set_err(ctx, MCD_ERR_INVALID, "mcd_tables_set_pick_nodes: NULL argument");
  if (n_nodes < 16 || n_nodes > 8192)
    return set_err(ctx, MCD_ERR_INVALID, "pick-table node count %d outside [16, 8192]", n_nodes);
  if (t->cfg.mode != MCD_MODE_DS || !t->has_w)
    return set_err(ctx, MCD_ERR_INVALID, "continuous-mode tables have no channel-pick table");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipDeviceSynchronize());  // launches may still read the old table
  int rc = build_pick_table(ctx, t, n_nodes);
  if (rc) return rc;
  HIP_TRY(ctx, hipDeviceSynchronize());
  return MCD_OK;
}

// occupancy variant used when the caller does not choose one
// Occupancy of the one-kernel FP32 transport: 4 resident wavefronts per SIMD
// (128 VGPRs, no spills) hides more of the table and LDS latency than the
// compiler's own allocation (3 per SIMD) at every configuration measured;
// 6 and 8 spill.  MCD_MINW overrides it.
// Smallest DS FP32 launch that takes the served guard when the caller names
// neither a guard placement nor a launch shape.
static const uint64_t kServedDefaultMin = 1ull << 21;
// Electron ids are 48 bits: the low 32 are Philox counter word 0, the high 16
// share word 1 with the 16-bit configuration id.
static const uint64_t kMaxElectronId = 1ull << 48;

static int default_minw(const mcd_config&) {
  const char* s = getenv("MCD_MINW");
  if (s) return atoi(s);
  return 4;
}

static bool same_config(const mcd_config& a, const mcd_config& b) {
  return a.mat.Z == b.mat.Z && a.mat.atnd == b.mat.atnd && a.mat.fermi_e == b.mat.fermi_e && a.E0 == b.E0 &&
         a.Emin == b.Emin && a.Wc == b.Wc && a.mat.n_shells == b.mat.n_shells && a.elastic == b.elastic;
}

// static_park: nullptr for mcd_forward (the deferred guard parks in the
// stream's buffer, allocated on demand); the context's scratch address for
// mcd_forward_noalloc (DS FP32 launches then default to the deferred guard
// and park there; nothing is allocated).
static int forward_impl(mcd_ctx* ctx, const mcd_config* cfg, const mcd_tables* tables, uint64_t seed,
                        uint32_t config_id, uint64_t first, uint64_t n, mcd_electron_rec* els,
                        const mcd_scatter_buf* scats, uint64_t* counters, const mcd_launch* launch, void* stream,
                        ParkRec* static_park) {
  if (!ctx) return set_err(nullptr, MCD_ERR_INVALID, "mcd_forward: ctx is NULL");
  int rc = check_config(ctx, cfg);
  if (rc) return rc;
  if (!counters) return set_err(ctx, MCD_ERR_INVALID, "mcd_forward: counters is NULL");
  if (n == 0) return MCD_OK;
  if (!els) return set_err(ctx, MCD_ERR_INVALID, "mcd_forward: els is NULL");
  if (config_id > 0xFFFFu) return set_err(ctx, MCD_ERR_INVALID, "config_id must fit 16 bits");
  if (first > kMaxElectronId || n > kMaxElectronId - first)
    return set_err(ctx, MCD_ERR_INVALID, "electrons [first, first+n) must lie below 2^48");
  const int precision = launch ? launch->precision : MCD_PREC_FP32;
  if (cfg->mode == MCD_MODE_DS) {
    if (!tables || !tables->has_w) return set_err(ctx, MCD_ERR_TABLE, "DS mode needs tables from mcd_tables_create");
    if (!same_config(*cfg, tables->cfg))
      return set_err(ctx, MCD_ERR_TABLE, "tables were built for a different configuration");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  DevCfg c = make_devcfg(*cfg, tables, config_id);
  RunArgs R;
  memset(&R, 0, sizeof(R));
  R.seed = seed;
  R.first = first;
  R.n = n;
  R.els = els;
  R.counters = (unsigned long long*)counters;
  // the launch's work counter is slot MCD_CNT_QUEUE of the caller's counters
  // block, so launches with separate blocks share no device counters (the
  // deferred form's park buffer is per stream)
  unsigned long long* q = (unsigned long long*)counters + MCD_CNT_QUEUE;
  R.queue = q;
  if (scats && scats->rows && scats->capacity) {
    R.rows = scats->rows;
    R.row_cap = scats->capacity;
    R.row_max_electron = scats->max_electron;
    R.row_cursor = (unsigned long long*)&counters[MCD_CNT_SCAT_ROWS];
  }
  const int flags = launch ? launch->flags : 0;
  R.refill_mask = (1u << ((flags & MCD_LAUNCH_REFILL_MASK) >> MCD_LAUNCH_REFILL_SHIFT)) - 1u;
  R.guard_spins = guard_spin_cap();
  const int guard_flags = flags & (MCD_LAUNCH_INLINE_GUARD | MCD_LAUNCH_DEFER_GUARD | MCD_LAUNCH_SERVICE_GUARD);
  if (guard_flags & (guard_flags - 1))
    return set_err(ctx, MCD_ERR_INVALID, "at most one of MCD_LAUNCH_INLINE_GUARD, _DEFER_GUARD, _SERVICE_GUARD");
  // the FP64 guard pick runs inline unless another placement is asked for
  // (DS, FP32 only; continuous mode has no guard)
  const bool ds32 = cfg->mode == MCD_MODE_DS && precision == MCD_PREC_FP32;
  const bool deferred = ds32 && ((flags & MCD_LAUNCH_DEFER_GUARD) || (static_park && !guard_flags));
  // with no placement and no shape asked for, DS launches of at least
  // kServedDefaultMin electrons take the served guard (on MI355X the fastest
  // form from 2^21 electrons up; below it the inline form is)
  const bool shape_given = launch && (launch->block > 0 || launch->waves_per_simd > 0);
  const bool service = ds32 && ((flags & MCD_LAUNCH_SERVICE_GUARD) ||
                                (!guard_flags && !shape_given && !static_park && n >= kServedDefaultMin));
  int block = launch && launch->block > 0 ? launch->block : (service ? 1024 : 256);
  if (service) {
    if (block % 64 || block < 128 || block > 1024)
      return set_err(ctx, MCD_ERR_INVALID, "service guard: block %d must be a multiple of 64 in [128, 1024]", block);
  } else if (block % 64 || block > 256) {
    return set_err(ctx, MCD_ERR_INVALID, "block %d must be a multiple of 64, at most 256", block);
  }
  int minw = launch && launch->waves_per_simd > 0 ? launch->waves_per_simd : (service ? 8 : default_minw(*cfg));
  if (service ? !valid_svc_minw(minw) : !valid_minw(minw))
    return set_err(ctx, MCD_ERR_INVALID, service ? "service guard: waves_per_simd %d not 6, 7 or 8"
                                                 : "waves_per_simd %d not one of 1, 4, 5, 6, 8", minw);
  const uint64_t want = (n + block - 1) / block;
  if (service) {
    // wavefront 0 of each workgroup serves the guard; the others carry electrons
    const uint64_t want_svc = (n + (uint64_t)(block - 64) - 1) / (uint64_t)(block - 64);
    uint64_t grid = (uint64_t)ctx->n_cu * (uint64_t)svc_blocks_per_cu(c, minw, block);
    if (grid > want_svc) grid = want_svc;
    if (grid < 1) grid = 1;
    HIP_TRY(ctx, launch_fill_u64(q, 0ull, 1, s));
    HIP_TRY(ctx, launch_transport_svc(c, R, minw, (int)grid, block, s));
  } else if (deferred) {
    if (static_park) {
      R.park = static_park;  // capacity MCD_NOALLOC_MAX_ELECTRONS >= n, checked by the caller
    } else {
      // the stream's park buffer (launches on other streams use their own)
      ParkBuf* pb = nullptr;
      for (ParkBuf& b : ctx->parks)
        if (b.stream == s) pb = &b;
      if (!pb) {
        ctx->parks.push_back(ParkBuf{s, nullptr, 0});
        pb = &ctx->parks.back();
      }
      if (pb->cap < n) {
        // earlier launches on this stream may still read the buffer
        HIP_TRY(ctx, hipStreamSynchronize(s));
        if (pb->p) { (void)hipFree(pb->p); pb->p = nullptr; pb->cap = 0; }
        if (hipMalloc(&pb->p, n * sizeof(ParkRec)) != hipSuccess) {
          pb->p = nullptr;
          return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc of the park buffer (%llu electrons) failed",
                         (unsigned long long)n);
        }
        pb->cap = n;
      }
      R.park = pb->p;
    }
    // parked count and resume counter: slots of the caller's counters block
    R.park_count = (unsigned long long*)counters + MCD_CNT_PARKED;
    R.resume_queue = (unsigned long long*)counters + MCD_CNT_RESUMED;
    uint64_t g1 = (uint64_t)ctx->n_cu * (uint64_t)deferred_blocks_per_cu(c, 1, block);
    uint64_t g2 = (uint64_t)ctx->n_cu * (uint64_t)deferred_blocks_per_cu(c, 2, block);
    if (g1 > want) g1 = want;
    if (g2 > want) g2 = want;
    HIP_TRY(ctx, launch_fill_u64(q, 0ull, 1, s));
    HIP_TRY(ctx, launch_fill_u64(R.park_count, 0ull, 2, s));
    HIP_TRY(ctx, launch_transport_deferred(c, R, 1, (int)(g1 < 1 ? 1 : g1), block, s));
    HIP_TRY(ctx, launch_transport_deferred(c, R, 2, (int)(g2 < 1 ? 1 : g2), block, s));
  } else {
    int per_cu = max_blocks_per_cu(c, precision, minw, block);
    uint64_t grid = (uint64_t)ctx->n_cu * (uint64_t)per_cu;
    if (grid > want) grid = want;
    if (grid < 1) grid = 1;
    HIP_TRY(ctx, launch_fill_u64(q, 0ull, 1, s));
    HIP_TRY(ctx, launch_transport(c, R, precision, minw, (int)grid, block, s));
  }
  HIP_TRY(ctx, launch_count_outcomes(els, n, (unsigned long long*)counters, s));
  return MCD_OK;
}

extern "C" int mcd_forward(mcd_ctx* ctx, const mcd_config* cfg, const mcd_tables* tables, uint64_t seed,
                           uint32_t config_id, uint64_t first, uint64_t n, mcd_electron_rec* els,
                           const mcd_scatter_buf* scats, uint64_t* counters, const mcd_launch* launch, void* stream) {
  return forward_impl(ctx, cfg, tables, seed, config_id, first, n, els, scats, counters, launch, stream, nullptr);
}

extern "C" int mcd_forward_noalloc(mcd_ctx* ctx, const mcd_config* cfg, const mcd_tables* tables, uint64_t seed,
                                   uint32_t config_id, uint64_t first, uint64_t n, mcd_electron_rec* els,
                                   const mcd_scatter_buf* scats, uint64_t* counters, const mcd_launch* launch,
                                   void* stream) {
  if (!ctx) return set_err(nullptr, MCD_ERR_INVALID, "mcd_forward_noalloc: ctx is NULL");
  if (n > (uint64_t)MCD_NOALLOC_MAX_ELECTRONS)
    return set_err(ctx, MCD_ERR_INVALID,
                   "mcd_forward_noalloc: %llu electrons exceed the static scratch's %d; split the range "
                   "(records are keyed by electron id)", (unsigned long long)n, MCD_NOALLOC_MAX_ELECTRONS);
  return forward_impl(ctx, cfg, tables, seed, config_id, first, n, els, scats, counters, launch, stream,
                      ctx->scratch);
}

// ---------------------------------------------------------------------------
// Batched launches.
struct mcd_batch {
  int device;
  int n_items;
  int slice_floats;
  std::vector<mcd_tables*> own_tables;  // per-item pick tables sized for the LDS slices
  BatchItemDev* d_items;                // grouped by XCD, followed by the group bounds (n_groups + 1 ints)
  int* d_group_start;
  int n_groups;
  unsigned long long* d_work;           // queues [n], active [n], parked [n], resumed [n]
  ParkRec* d_park;                      // deferred guard: parked electrons, n_i per DS item
  bool has_ds;
  hipStream_t capture_stream;
  hipGraphExec_t graph;
  uint64_t graph_seed;
  uint64_t* graph_counters;
  hipStream_t graph_stream;
  int graph_block, graph_grid, graph_deferred;  // graph_deferred: the guard form the graph was captured with
  uint32_t graph_refill;                        // the compaction period less one it was captured with
};

static double work_per_electron(const mcd_config& c) {
  // events per electron, a rough model for load balancing only
  const double dE = c.E0 - c.Emin;
  if (c.mode == MCD_MODE_DS) return dE / 25.0;
  const double E = c.E0, Z = c.mat.Z;
  const double a = 3.4 * pow(Z, 0.666667) / E, Ek = E * 1e-3, r = (Ek + 511.) / (Ek + 1024.);
  const double sig = 5.21e-21 * (Z * Z / (Ek * Ek)) * r * r * 4. * 3.141592653589793 / (a * (1. + a));
  const double mfp = 1e14 / (c.mat.atnd * sig);
  const double S = 2. * kPiEFour * c.mat.atnd * Z / E * log(1.166 * (E + c.mat.bethe_k * c.J) / c.J) * 1e-14;
  return fmax(1.0, dE / fmax(mfp * S, 1e-3) * 0.6);
}

extern "C" int mcd_batch_destroy(mcd_batch* b) {
  if (!b) return MCD_OK;
  (void)hipSetDevice(b->device);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  for (mcd_tables* t : b->own_tables) mcd_tables_destroy(t);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_work) (void)hipFree(b->d_work);
  if (b->d_park) (void)hipFree(b->d_park);
  delete b;
  return MCD_OK;
}

extern "C" int mcd_batch_create(mcd_ctx* ctx, const mcd_batch_item* items, int n_items, int pick_nodes,
                                mcd_batch** out) {
  if (!ctx || !items || !out || n_items < 1 || n_items > 0xFFFF)
    return set_err(ctx, MCD_ERR_INVALID, "mcd_batch_create: bad arguments (n_items=%d)", n_items);
  *out = nullptr;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  // per-wavefront LDS slice: configuration copy + pick table within ~10 KB
  const int pick_budget = 2048;  // floats
  mcd_batch* b = new mcd_batch();
  b->device = ctx->device;
  b->n_items = n_items;
  std::vector<BatchItemDev> h(n_items);
  std::vector<double> work(n_items);
  int slice = 0;
  for (int i = 0; i < n_items; ++i) {
    const mcd_batch_item& it = items[i];
    int rc = check_config(ctx, it.cfg);
    if (rc) { mcd_batch_destroy(b); return rc; }
    if (!it.els && it.n) { mcd_batch_destroy(b); return set_err(ctx, MCD_ERR_INVALID, "item %d: els is NULL", i); }
    if (it.n > kMaxElectronId) {
      mcd_batch_destroy(b);
      return set_err(ctx, MCD_ERR_INVALID, "item %d: electron count above 2^48", i);
    }
    mcd_tables* own = nullptr;
    if (it.cfg->mode == MCD_MODE_DS) {
      if (!it.tables || !it.tables->has_w || !same_config(*it.cfg, it.tables->cfg)) {
        mcd_batch_destroy(b);
        return set_err(ctx, MCD_ERR_TABLE, "item %d: DS tables missing or built for another configuration", i);
      }
      // a private pick table sized for the per-wavefront LDS slice, sharing the W tables
      own = new mcd_tables();
      memset((void*)own, 0, sizeof(*own));
      own->device = ctx->device;
      own->cfg = *it.cfg;
      own->has_w = false;
      own->m = it.tables->m;
      own->g = it.tables->g;
      own->has_w = true;
      const int stride = pick_stride(it.cfg->mat.n_shells + 3);
      int nodes = pick_nodes > 0 ? pick_nodes : pick_budget / stride;
      if (nodes > 512) nodes = 512;
      rc = build_pick_table(ctx, own, nodes);
      own->m = TableDev();  // W tables stay owned by the caller's object
      own->g = TableDev();
      own->has_w = false;
      if (rc) { mcd_tables_destroy(own); mcd_batch_destroy(b); return rc; }
      b->own_tables.push_back(own);
      slice = slice > own->n_nodes * own->stride ? slice : own->n_nodes * own->stride;
    }
    DevCfg c = make_devcfg(*it.cfg, it.tables, (uint32_t)i);
    if (own) {
      c.pick.nodes = own->nodes;
      c.pick.n_nodes = own->n_nodes;
      c.pick.stride = own->stride;
      c.pick.n_ch = own->n_ch;
      c.pick.log2E0 = (float)(own->lnE0 / M_LN2);
      c.pick.inv_dlog2E = (float)(M_LN2 / own->dln);
      c.pick.delta = own->delta;
    }
    memset(&h[i], 0, sizeof(BatchItemDev));
    h[i].c = c;
    h[i].R.first = 0;
    h[i].R.n = it.n;
    h[i].R.els = it.els;
    h[i].R.guard_spins = guard_spin_cap();
    work[i] = (double)it.n * work_per_electron(*it.cfg);
    h[i].cost = (float)work_per_electron(*it.cfg);
  }
  // One group of items per XCD: largest estimated work first, each to the
  // group with the least work so far (longest-processing-time partition).
  b->n_groups = std::max(1, std::min(std::min(ctx->n_xcc, kMaxXcdGroups), n_items));
  std::vector<int> by_work(n_items);
  for (int i = 0; i < n_items; ++i) by_work[i] = i;
  std::stable_sort(by_work.begin(), by_work.end(), [&](int a, int c) { return work[a] > work[c]; });
  std::vector<std::vector<int>> groups(b->n_groups);
  std::vector<double> load(b->n_groups, 0.);
  for (int i : by_work) {
    const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    groups[g].push_back(i);
    load[g] += work[i];
  }
  std::vector<int> order, group_start(1, 0);
  for (int g = 0; g < b->n_groups; ++g) {
    double acc = 0.;
    for (int i : groups[g]) {
      h[i].cum_lo = (float)(load[g] > 0 ? acc / load[g] : 0.);
      acc += work[i];
      h[i].cum_hi = (float)(load[g] > 0 ? acc / load[g] : 0.);
    }
    if (!groups[g].empty()) h[groups[g].back()].cum_hi = 1.01f;
    order.insert(order.end(), groups[g].begin(), groups[g].end());
    group_start.push_back((int)order.size());
  }
  b->slice_floats = batched_cfg_floats() + (slice > 0 ? slice : 4);
  const size_t items_bytes = ((size_t)n_items * sizeof(BatchItemDev) + 255) & ~(size_t)255;
  if (hipMalloc(&b->d_items, items_bytes + (kMaxXcdGroups + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc(&b->d_work, 4 * (size_t)n_items * sizeof(unsigned long long)) != hipSuccess) {
    mcd_batch_destroy(b);
    return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for the batch failed");
  }
  uint64_t n_park = 0;
  for (int i = 0; i < n_items; ++i)
    if (items[i].cfg->mode == MCD_MODE_DS) n_park += items[i].n;
  b->has_ds = n_park > 0;
  if (n_park && hipMalloc(&b->d_park, n_park * sizeof(ParkRec)) != hipSuccess) {
    b->d_park = nullptr;
    mcd_batch_destroy(b);
    return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc of the batch park buffer (%llu electrons) failed",
                   (unsigned long long)n_park);
  }
  uint64_t off = 0;
  for (int i = 0; i < n_items; ++i) {
    h[i].R.queue = b->d_work + i;
    h[i].active = b->d_work + n_items + i;
    h[i].R.park_count = b->d_work + 2 * n_items + i;
    h[i].R.resume_queue = b->d_work + 3 * n_items + i;
    if (items[i].cfg->mode == MCD_MODE_DS) {
      h[i].R.park = b->d_park + off;
      off += items[i].n;
    }
  }
  std::vector<BatchItemDev> hs(n_items);
  for (int k = 0; k < n_items; ++k) {
    hs[k] = h[order[k]];
    hs[k].slot = order[k];
  }
  b->d_group_start = reinterpret_cast<int*>(reinterpret_cast<char*>(b->d_items) + items_bytes);
  hipError_t e = hipMemcpy(b->d_items, hs.data(), n_items * sizeof(BatchItemDev), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_group_start, group_start.data(), group_start.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) { mcd_batch_destroy(b); return set_err(ctx, MCD_ERR_DEVICE, "batch upload: %s", hipGetErrorString(e)); }
  *out = b;
  return MCD_OK;
}

// guard placement of a batch run: 0 inline, 1 deferred (phases 1 + 2), 3 served;
// refill_mask: compaction period less one
static hipError_t batch_enqueue(mcd_ctx* ctx, mcd_batch* b, uint64_t seed, uint64_t* counters, int form,
                                int block, uint32_t refill_mask, hipStream_t s) {
  hipError_t e = launch_fill_u64(b->d_work, 0ull, 4 * b->n_items, s);
  unsigned long long* cnt = (unsigned long long*)counters;
  const bool deferred = form == 1;
  auto run = [&](int phase, int grid) {
    return launch_batched(b->d_items, b->n_items, b->d_group_start, b->n_groups, b->slice_floats, seed,
                          refill_mask, cnt, phase, grid, block, s);
  };
  if (form == 3) {
    const int g3 = ctx->n_cu * batched_blocks_per_cu(3, block, b->slice_floats, refill_mask != 0u);
    if (e == hipSuccess) e = run(3, g3);
  } else if (deferred) {
    const int g1 = ctx->n_cu * batched_blocks_per_cu(1, block, b->slice_floats, refill_mask != 0u);
    const int g2 = ctx->n_cu * batched_blocks_per_cu(2, block, b->slice_floats, false);
    if (e == hipSuccess) e = run(1, g1);
    if (e == hipSuccess) e = run(2, g2);
  } else {
    const int g0 = ctx->n_cu * batched_blocks_per_cu(0, block, b->slice_floats, refill_mask != 0u);
    if (e == hipSuccess) e = run(0, g0);
  }
  if (e == hipSuccess) e = launch_count_outcomes_batched(b->d_items, b->n_items, cnt, s);
  return e;
}

extern "C" int mcd_batch_run(mcd_ctx* ctx, mcd_batch* b, uint64_t seed, uint64_t* counters,
                             const mcd_launch* launch, int use_graph, void* stream) {
  if (!ctx || !b || !counters) return set_err(ctx, MCD_ERR_INVALID, "mcd_batch_run: NULL argument");
  if (launch && launch->precision == MCD_PREC_FP64)
    return set_err(ctx, MCD_ERR_UNSUPPORTED, "batched launches run the FP32 transport");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int flags = launch ? launch->flags : 0;
  const int guard_flags = flags & (MCD_LAUNCH_INLINE_GUARD | MCD_LAUNCH_DEFER_GUARD | MCD_LAUNCH_SERVICE_GUARD);
  if (guard_flags & (guard_flags - 1))
    return set_err(ctx, MCD_ERR_INVALID, "at most one of MCD_LAUNCH_INLINE_GUARD, _DEFER_GUARD, _SERVICE_GUARD");
  // DS items take the FP64 guard pick inline unless another placement is asked
  // for; the served form keeps its workgroup shape when no DS item has
  // electrons (its service wavefronts then stay idle), the deferred form then
  // runs as the inline one
  const int form = (flags & MCD_LAUNCH_SERVICE_GUARD) ? 3 : ((b->has_ds && (flags & MCD_LAUNCH_DEFER_GUARD)) ? 1 : 0);
  const int block = launch && launch->block > 0 ? launch->block : 256;
  if (form == 3) {
    if (block % 64 || block < 128 || block > 1024)
      return set_err(ctx, MCD_ERR_INVALID, "served guard: block %d must be a multiple of 64 in [128, 1024]", block);
  } else if (block % 64 || block > 256) {
    return set_err(ctx, MCD_ERR_INVALID, "block %d invalid", block);
  }
  const uint32_t refill = (1u << ((flags & MCD_LAUNCH_REFILL_MASK) >> MCD_LAUNCH_REFILL_SHIFT)) - 1u;
  hipStream_t s = (hipStream_t)stream;
  if (!use_graph) {
    HIP_TRY(ctx, batch_enqueue(ctx, b, seed, counters, form, block, refill, s));
    return MCD_OK;
  }
  if (!b->graph || b->graph_seed != seed || b->graph_counters != counters || b->graph_block != block ||
      b->graph_deferred != form || b->graph_refill != refill) {
    if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    if (!b->capture_stream) HIP_TRY(ctx, hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    HIP_TRY(ctx, hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = batch_enqueue(ctx, b, seed, counters, form, block, refill, b->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->capture_stream, &g);
    HIP_TRY(ctx, e);
    HIP_TRY(ctx, e2);
    hipError_t e3 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(ctx, e3);
    b->graph_seed = seed;
    b->graph_counters = counters;
    b->graph_block = block;
    b->graph_deferred = form;
    b->graph_refill = refill;
  }
  HIP_TRY(ctx, hipGraphLaunch(b->graph, s));
  return MCD_OK;
}

// ---------------------------------------------------------------------------
// Scatter rows to columns on the device.
extern "C" int mcd_scatter_pack(mcd_ctx* ctx, const mcd_scatter_rec* rows, uint64_t n_rows, uint32_t fields,
                                void* out, size_t out_bytes, size_t* needed, void* stream) {
  if (!ctx || !needed) return set_err(ctx, MCD_ERR_INVALID, "mcd_scatter_pack: NULL argument");
  if (fields == 0u || (fields >> kScatFields) != 0u)
    return set_err(ctx, MCD_ERR_INVALID, "field mask 0x%x selects no field or an unknown one", fields);
  typedef mcd_scatter_rec R;
  static const int off[kScatFields] = {
      (int)offsetof(R, electron), (int)offsetof(R, step), (int)offsetof(R, config), (int)offsetof(R, channel),
      (int)offsetof(R, outcome), (int)offsetof(R, E), (int)offsetof(R, E_loss), (int)offsetof(R, pol),
      (int)offsetof(R, az), (int)offsetof(R, pathl), (int)offsetof(R, x), (int)offsetof(R, y),
      (int)offsetof(R, z), (int)offsetof(R, energy), (int)offsetof(R, dx), (int)offsetof(R, dy),
      (int)offsetof(R, dz), (int)offsetof(R, step_flags)};
  static const int sz[kScatFields] = {
      (int)sizeof(R::electron), (int)sizeof(R::step), (int)sizeof(R::config), (int)sizeof(R::channel),
      (int)sizeof(R::outcome), (int)sizeof(R::E), (int)sizeof(R::E_loss), (int)sizeof(R::pol),
      (int)sizeof(R::az), (int)sizeof(R::pathl), (int)sizeof(R::x), (int)sizeof(R::y),
      (int)sizeof(R::z), (int)sizeof(R::energy), (int)sizeof(R::dx), (int)sizeof(R::dy),
      (int)sizeof(R::dz), (int)sizeof(R::step_flags)};
  ScatPackPlan plan{};
  uint64_t at = 0;
  for (int f = 0; f < kScatFields; ++f) {
    if (!((fields >> f) & 1u)) continue;
    plan.src[plan.n] = off[f];
    plan.size[plan.n] = sz[f];
    plan.dst[plan.n] = at;
    at += (n_rows * (uint64_t)sz[f] + 255u) / 256u * 256u;
    plan.n++;
  }
  *needed = (size_t)at;
  if (!out) return MCD_OK;
  if (n_rows && !rows) return set_err(ctx, MCD_ERR_INVALID, "mcd_scatter_pack: NULL rows");
  if (out_bytes < at)
    return set_err(ctx, MCD_ERR_INVALID, "output holds %zu bytes, the columns need %llu", out_bytes,
                   (unsigned long long)at);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, launch_pack_scatter(rows, n_rows, plan, out, (hipStream_t)stream));
  return MCD_OK;
}

// ---------------------------------------------------------------------------
// Differentiation.
// Node count of the replay's node table.  192: a 256-node table (31.7 KB of
// LDS per workgroup in DS) fits fewer workgroups per CU and made the DS
// replay 12 % slower on MI355X; 48 to 192 nodes all ran at the same speed.
static int bwd_nodes_default() {
  const char* s = getenv("MCD_BWD_NODES");
  int n = s ? atoi(s) : 192;
  return n < 16 ? 16 : (n > 4096 ? 4096 : n);
}

// Moller W grid at a shifted cut-off: geomspace(Wc, W_top, nW) in float32
// (probTables.py:251), the grid the reference would build for that Wc.
static void moller_grid_at(double Wc, float Wtop, int nW, std::vector<float>& out) {
  out.resize(nW);
  const double l0 = log(Wc), l1 = log((double)Wtop);
  for (int j = 0; j < nW; ++j) out[j] = (float)exp(l0 + (double)j / (double)(nW - 1) * (l1 - l0));
  out[0] = (float)Wc;
  out[nW - 1] = Wtop;
}

// Wc scores of the Moller column picks: tables rebuilt at Wc +- delta with
// the same grid sizes, central difference of the column probabilities.
static int build_wc_scores(mcd_ctx* ctx, const mcd_tables* t, double rel, hipStream_t s) {
  const TableDev& m = t->m;
  const size_t ncell = (size_t)m.nE * (size_t)(m.nW - 1);
  float wtop = 0.f;
  HIP_TRY(ctx, hipMemcpy(&wtop, m.Ws + (m.nW - 1), sizeof(float), hipMemcpyDeviceToHost));
  const double delta = rel * t->cfg.Wc;
  float *wp = nullptr, *wm = nullptr, *cp = nullptr, *cm = nullptr;
  int rc = MCD_OK;
  if (hipMalloc(&wp, m.nW * sizeof(float)) != hipSuccess || hipMalloc(&wm, m.nW * sizeof(float)) != hipSuccess ||
      hipMalloc(&cp, ncell * sizeof(float)) != hipSuccess || hipMalloc(&cm, ncell * sizeof(float)) != hipSuccess) {
    rc = set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for the Wc column scores failed");
  } else {
    std::vector<float> gp, gm;
    moller_grid_at(t->cfg.Wc + delta, wtop, m.nW, gp);
    moller_grid_at(t->cfg.Wc - delta, wtop, m.nW, gm);
    hipError_t e = hipMemcpy(wp, gp.data(), m.nW * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(wm, gm.data(), m.nW * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = launch_build_cdf(m.Es, wp, m.nE, m.nW, t->cfg.mat.fermi_e, MCD_TABLE_MOLLER, t->moller_n_el, 0., MCD_PREC_FP64,
                           cp, s);
    if (e == hipSuccess)
      e = launch_build_cdf(m.Es, wm, m.nE, m.nW, t->cfg.mat.fermi_e, MCD_TABLE_MOLLER, t->moller_n_el, 0., MCD_PREC_FP64,
                           cm, s);
    if (e == hipSuccess) e = launch_ratio_score(m.cdf, cp, cm, m.nE, m.nW, 0.5 / delta, t->bwd_m_swc, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = set_err(ctx, MCD_ERR_DEVICE, "Wc column scores: %s", hipGetErrorString(e));
  }
  if (wp) (void)hipFree(wp);
  if (wm) (void)hipFree(wm);
  if (cp) (void)hipFree(cp);
  if (cm) (void)hipFree(cm);
  if (rc == MCD_OK) t->bwd_rel = rel;
  return rc;
}

static int ensure_bwd_nodes(mcd_ctx* ctx, const mcd_tables* t, const DevCfg& c, double rel, hipStream_t s) {
  const int n_ch = t->cfg.mode == MCD_MODE_DS ? t->cfg.mat.n_shells + 3 : 1;
  const int nn = bwd_nodes_default();
  if (!t->bwd_nodes || t->bwd_n != nn) {
    if (t->bwd_nodes) { (void)hipFree(t->bwd_nodes); t->bwd_nodes = nullptr; }
    const int stride = bwd_stride(t->cfg.mode, n_ch);
    if (hipMalloc(&t->bwd_nodes, (size_t)nn * stride * sizeof(float)) != hipSuccess) {
      t->bwd_nodes = nullptr;
      return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for the derivative node table failed");
    }
    t->bwd_n = nn;
    t->bwd_stride = stride;
    t->bwd_nch = n_ch;
    t->bwd_lnE0 = log(t->cfg.Emin);
    t->bwd_dln = (log(t->cfg.E0) - t->bwd_lnE0) / (double)(nn - 1);
    HIP_TRY(ctx, launch_bwd_nodes(c, nn, t->bwd_lnE0, t->bwd_dln, t->bwd_nodes, stride, s));
  }
  if (t->cfg.mode != MCD_MODE_DS) return MCD_OK;
  const TableDev& m = t->m;
  const TableDev& g = t->g;
  if (!t->bwd_m_swc) {
    const size_t nm = (size_t)m.nE * (size_t)(m.nW - 1), ng = (size_t)g.nE * (size_t)(g.nW - 1);
    if (hipMalloc(&t->bwd_m_swc, nm * sizeof(float)) != hipSuccess ||
        hipMalloc(&t->bwd_m_se, nm * sizeof(float)) != hipSuccess ||
        hipMalloc(&t->bwd_g_se, ng * sizeof(float)) != hipSuccess)
      return set_err(ctx, MCD_ERR_NOMEM, "hipMalloc for the loss-table scores failed");
    HIP_TRY(ctx, launch_row_score(m.cdf, m.Es, m.nE, m.nW, t->bwd_m_se, s));
    HIP_TRY(ctx, launch_row_score(g.cdf, g.Es, g.nE, g.nW, t->bwd_g_se, s));
    t->bwd_rel = 0.;
  }
  if (t->bwd_rel != rel) return build_wc_scores(ctx, t, rel, s);
  return MCD_OK;
}

extern "C" int mcd_backward(mcd_ctx* ctx, const mcd_config* cfg, const mcd_tables* tables, uint64_t seed,
                            uint32_t config_id, uint64_t first, uint64_t n, const float* grad_out, double* out_vjp,
                            double* out_jac, const mcd_bwd_opts* opts, void* stream) {
  if (!ctx) return set_err(nullptr, MCD_ERR_INVALID, "mcd_backward: ctx is NULL");
  int rc = check_config(ctx, cfg);
  if (rc) return rc;
  if ((out_jac == nullptr) == (out_vjp == nullptr))
    return set_err(ctx, MCD_ERR_INVALID, "mcd_backward: pass exactly one of out_jac and out_vjp");
  if (out_vjp && !grad_out) return set_err(ctx, MCD_ERR_INVALID, "mcd_backward: out_vjp needs grad_out");
  if (config_id > 0xFFFFu) return set_err(ctx, MCD_ERR_INVALID, "config_id must fit 16 bits");
  if (first > kMaxElectronId || n > kMaxElectronId - first)
    return set_err(ctx, MCD_ERR_INVALID, "electrons [first, first+n) must lie below 2^48");
  if (!tables || !same_config(*cfg, tables->cfg) || tables->cfg.mode != cfg->mode)
    return set_err(ctx, MCD_ERR_TABLE, "mcd_backward needs tables built for this configuration");
  if (cfg->mode == MCD_MODE_DS && !tables->has_w)
    return set_err(ctx, MCD_ERR_TABLE, "DS mode needs tables from mcd_tables_create");
  if (std::isfinite(cfg->thickness))
    return set_err(ctx, MCD_ERR_UNSUPPORTED, "mcd_backward supports semi-infinite targets only (thickness = inf)");
  if (cfg->max_scatt > 0) return set_err(ctx, MCD_ERR_UNSUPPORTED, "mcd_backward needs maxScatt off");
  const int n_bins = opts && opts->n_bins > 0 ? opts->n_bins : 64;
  if (n_bins > MCD_MAX_BINS) return set_err(ctx, MCD_ERR_INVALID, "n_bins %d above %d", n_bins, MCD_MAX_BINS);
  const double bin_w = cfg->E0 / n_bins;
  const double rel = opts && opts->rel_step > 0.f ? opts->rel_step : 1e-2;
  const double graze = opts && opts->graze > 0.f ? opts->graze : 0.005;
  if (!(rel < 0.5)) return set_err(ctx, MCD_ERR_INVALID, "rel_step must be below 0.5");
  if (!(graze < 1.0)) return set_err(ctx, MCD_ERR_INVALID, "graze must be below 1");
  const int block = opts && opts->block > 0 ? opts->block : 256;
  if (block % 64 || block > 256) return set_err(ctx, MCD_ERR_INVALID, "block %d must be a multiple of 64, at most 256", block);
  if (n == 0) return MCD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  DevCfg c = make_devcfg(*cfg, tables, config_id);
  if ((rc = ensure_bwd_nodes(ctx, tables, c, rel, s))) return rc;
  BwdArgs A;
  memset(&A, 0, sizeof(A));
  A.seed = seed;
  A.first = first;
  A.n = n;
  A.nodes = tables->bwd_nodes;
  A.n_nodes = tables->bwd_n;
  A.stride = tables->bwd_stride;
  A.n_ch = tables->bwd_nch;
  A.lnE0 = (float)tables->bwd_lnE0;
  A.inv_dln = (float)(1.0 / tables->bwd_dln);
  A.n_bins = n_bins;
  A.bin_w = (float)bin_w;
  A.inv_bin_w = (float)(1.0 / bin_w);
  // parameter steps of the threshold terms, per tangent slot
  const double p0 = cfg->mode == MCD_MODE_DS ? cfg->Wc : cfg->J;
  const double steps[3] = {rel * fabs(p0), rel * fmax(fabs(cfg->s_tilt), 1.0), rel};
  for (int i = 0; i < 3; ++i) {
    A.delta[i] = (float)steps[i];
    A.inv_2delta[i] = (float)(0.5 / steps[i]);
  }
  A.graze = (float)graze;
  if (cfg->mode == MCD_MODE_DS) {
    A.m_score_wc = tables->bwd_m_swc;
    A.m_score_e = tables->bwd_m_se;
    A.g_score_e = tables->bwd_g_se;
    A.inv_wc = (float)(1.0 / cfg->Wc);
    A.inv_nw1 = (float)(1.0 / (double)(tables->m.nW - 1));
  }
  A.grad_out = grad_out;
  A.out = out_jac ? out_jac : out_vjp;
  A.queue = ctx->d_queue;
  A.col_of_slot[0] = 1 + (cfg->mode == MCD_MODE_DS ? MCD_GRAD_WC : MCD_GRAD_J);
  A.col_of_slot[1] = 1 + MCD_GRAD_TILT;
  A.col_of_slot[2] = 1 + MCD_GRAD_LNRHO;
  const bool jac = out_jac != nullptr;
  const int per_cu = bwd_blocks_per_cu(c, A, jac, block);
  uint64_t want = (n + block - 1) / block;
  uint64_t grid = (uint64_t)ctx->n_cu * (uint64_t)per_cu;
  if (grid > want) grid = want;
  if (grid < 1) grid = 1;
  HIP_TRY(ctx, launch_fill_u64(ctx->d_queue, 0ull, 1, s));
  HIP_TRY(ctx, launch_backward(c, A, jac, (int)grid, block, s));
  return MCD_OK;
}
