/*
 * mcd.h — C-ABI of the MI355X electron-trajectory transport library (libmcd.so).
 *
 * The library replaces the per-electron scattering loop of MC-discrete:
 *   multiTraj_DS   MC/tools/multiScatter.py:66-116   -> mcd_forward (mode MCD_MODE_DS)
 *   multiTraj_cont MC/tools/multiScatter.py:171-229  -> mcd_forward (mode MCD_MODE_CONT)
 *   trajectory_DS  MC/tools/singleScatter.py:10-41   (the device step loop)
 *   trajectory_cont_cl MC/tools/singleScatter.py:205-239
 *   probTable.computeBlock MC/tools/probTables.py:415-462 -> mcd_tables_create
 * The Python side (ctypes) owns every buffer; the library owns only the
 * opaque handles below.  Every entry point returns an int status (MCD_OK = 0)
 * and records a message retrievable with mcd_last_error().
 *
 * Device pointers passed in are plain HIP device addresses; streams are
 * hipStream_t passed as void* (NULL = the null stream).  One context is used
 * by one host thread at a time.  mcd_forward keeps a launch's work counters in
 * the caller's counters block (MCD_CNT_QUEUE, and MCD_CNT_PARKED /
 * MCD_CNT_RESUMED for the deferred guard, whose park buffer the context keeps
 * per stream), so launches with separate counters blocks may run concurrently
 * on different streams of one context.  mcd_backward uses a work counter held
 * by the context, so backward calls on one context must not overlap in time.
 */
#ifndef MCD_H
#define MCD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCD_ABI_VERSION 1

#define MCD_MAX_SHELLS 8
#define MCD_MAX_CHANNELS (MCD_MAX_SHELLS + 3)

/* status codes */
enum {
  MCD_OK = 0,
  MCD_ERR_INVALID = 1,     /* bad argument or configuration */
  MCD_ERR_DEVICE = 2,      /* HIP runtime error */
  MCD_ERR_TABLE = 3,       /* table grid inconsistent with the configuration */
  MCD_ERR_OVERFLOW = 4,    /* scatter-row buffer full; see mcd_counters.scat_rows */
  MCD_ERR_NOMEM = 5,
  MCD_ERR_UNSUPPORTED = 6
};

/* transport modes (input.file `mode`, `Bethe_model`) */
enum { MCD_MODE_DS = 0, MCD_MODE_CONT = 1 };
enum { MCD_BETHE_CLASSICAL = 0, MCD_BETHE_JL = 1, MCD_BETHE_EXPLICIT = 2 };
/* elastic cross section (the input key 'elastic', MC/tools/scattering.py:483-493,
 * 780-789; crossSections.py:37-147): Rutherford with Joy's screening, with
 * Z^2 -> Z(Z+1) for inelastic deflections (wDefl), or with the Nigram
 * screening parameter.  The polar angle is Rutherford_halfPol in every case. */
enum {
  MCD_ELASTIC_RUTH = 0,          /* Ruth_vanilla:       ruther_sigma */
  MCD_ELASTIC_RUTH_WDEFL = 1,    /* Ruth_vanilla_wDefl: ruther_sigma_wDefl */
  MCD_ELASTIC_NIGRAM = 2,        /* Ruth_nigram:        ruther_N_sigma */
  MCD_ELASTIC_NIGRAM_WDEFL = 3   /* Ruth_nigram_wDefl:  ruther_N_sigma_wDefl */
};

/* electron outcomes (MC/tools/electron.py:44-84, singleScatter.py:15-23) */
enum {
  MCD_OUT_NONE = 0,
  MCD_OUT_ABS = 1,
  MCD_OUT_BKS = 2,
  MCD_OUT_TRSM = 3,
  MCD_OUT_TRDF = 4,
  MCD_OUT_LONG = 5
};

/* scattering channels; label values, not the pick order.  The pick order is
 * the descending-sigma order with ties kept in insertion order Ruth, Moller,
 * Gryz<shell> (name_s order), Quinn (MC/tools/scattering.py:39-59,483-508). */
enum {
  MCD_CH_RUTH = 0,
  MCD_CH_MOLLER = 1,
  MCD_CH_QUINN = 2,
  MCD_CH_GRYZ0 = 3 /* + shell index in name_s order */
};

/* per-electron anomaly flags */
enum {
  MCD_FLAG_NAN_DIR = 1,    /* binaryCollModel returned NaN (W > (E-E_F)/2) */
  MCD_FLAG_LOSS_GE_E = 2,  /* energy loss >= energy (reference asserts, scattering.py:605,627,822) */
  MCD_FLAG_LONG_PATH = 4,  /* continuous-mode path >= 1e4 A (reference asserts, scattering.py:809) */
  MCD_FLAG_FP64_PICK = 8,  /* at least one channel pick was settled by the FP64 guard path */
  MCD_FLAG_GUARD_TIMEOUT = 16 /* service guard: no reply within the wait bound; the pick is not
                                 settled (counted in MCD_CNT_GUARD_TIMEOUTS; an error) */
};

/* Material constants (MC/tools/material.py:151-302) plus the derived
 * quantities the host computes once (material.py:117-134). */
typedef struct mcd_material {
  int32_t Z;
  int32_t n_shells;
  double atwt;       /* g/mol */
  double density;    /* g/cm^3 */
  double n_val;      /* valence electrons */
  double E_val;      /* eV */
  double bethe_k;    /* Joy-Luo k */
  double shell_Eb[MCD_MAX_SHELLS]; /* binding energies, eV, name_s order */
  double shell_n[MCD_MAX_SHELLS];  /* occupancies, name_s order */
  double atnd;       /* atomic number density, m^-3 */
  double fermi_e;    /* eV */
  double plasmon_e;  /* eV */
} mcd_material;

/* One transport configuration (input.file keys, MC/tools/fileTools.py:8-46). */
typedef struct mcd_config {
  mcd_material mat;
  double E0;         /* eV */
  double Emin;       /* eV */
  double Wc;         /* eV, Moller cut-off (DS) */
  double s_tilt;     /* degrees; dir0 = (-sin t, 0, cos t), multiScatter.py:98-100 */
  double thickness;  /* Angstrom; +inf = semi-infinite */
  int32_t mode;      /* MCD_MODE_* */
  int32_t bethe_model; /* MCD_BETHE_* */
  int64_t max_scatt; /* <= 0: no limit */
  double J;          /* mean ionisation energy for the Bethe laws, eV (stoppingPowers.py:134-137) */
  double explicit_lnEstar; /* explicit Bethe: sum_i n_i ln E_i / Z_eff (host-computed, FP64) */
  double explicit_Zeff;    /* explicit Bethe: n_val + sum_i n_i */
  int32_t elastic;         /* MCD_ELASTIC_* */
} mcd_config;

/* Inverse-CDF energy-loss table grid (probTables.py:70-535).  Es has nE
 * float32 energies, Ws has nW float32 losses; the CDF is nE x (nW-1) float32,
 * row-major, entries with W > W_max(E) stored as 0 (probTables.py:500-507). */
typedef struct mcd_table_grid {
  int32_t nE;
  int32_t nW;
  const float* Es;
  const float* Ws;
  double Wmin;       /* lower integration limit: Wc (Moller) or E_b (Gryzinski) */
  double n_el;       /* electrons in the shell: n_val (Moller) or ns (Gryzinski) */
  double Eb;         /* Gryzinski binding energy; 0 for Moller */
  int32_t kind;      /* 0 Moller, 1 Gryzinski */
} mcd_table_grid;

enum { MCD_TABLE_MOLLER = 0, MCD_TABLE_GRYZ = 1 };
enum { MCD_PREC_FP64 = 0, MCD_PREC_FP32 = 1 };

/* Per-electron record written by mcd_forward (the `els` row(s) of one electron,
 * electron.py:105-113).  outcome2 != 0 marks the second saveOutcomes row that
 * the reference appends when an exit step also absorbs (electron.py:52-54). */
typedef struct mcd_electron_rec {
  float final_E;
  float dir[3];
  float pos[3];
  uint32_t n_scatt;  /* scatter events (loop iterations that reached det_type) */
  uint8_t outcome;
  uint8_t outcome2;
  uint8_t flags;
  uint8_t pad;
} mcd_electron_rec;

/* One `scats` row (scattering.py:527-679, electron.py:59,99), appended to a
 * device buffer in arbitrary order; sort by (config, electron, step). */
typedef struct mcd_scatter_rec {
  uint32_t electron; /* low 32 bits of the electron id */
  uint32_t step;
  uint16_t config;
  uint8_t channel;   /* MCD_CH_* */
  uint8_t outcome;   /* outcome set during this step, 0 if none */
  float E;           /* pre-step energy */
  float E_loss;
  float pol;         /* cos^2(theta/2) */
  float az;          /* halfPhi */
  float pathl;
  float x, y, z;     /* pre-step position */
  float energy;      /* post-step energy */
  float dx, dy, dz;  /* post-step direction */
  uint32_t step_flags; /* MCD_FLAG_* raised during this step */
} mcd_scatter_rec;

/* Device-side append buffer for scatter rows. */
typedef struct mcd_scatter_buf {
  mcd_scatter_rec* rows;     /* device */
  uint64_t capacity;         /* rows */
  uint64_t max_electron;     /* record only electrons with id < max_electron */
} mcd_scatter_buf;

/* Fields of mcd_scatter_rec, in record order, for mcd_scatter_pack. */
enum {
  MCD_SCAT_ELECTRON = 1u << 0,
  MCD_SCAT_STEP = 1u << 1,
  MCD_SCAT_CONFIG = 1u << 2,
  MCD_SCAT_CHANNEL = 1u << 3,
  MCD_SCAT_OUTCOME = 1u << 4,
  MCD_SCAT_E = 1u << 5,
  MCD_SCAT_E_LOSS = 1u << 6,
  MCD_SCAT_POL = 1u << 7,
  MCD_SCAT_AZ = 1u << 8,
  MCD_SCAT_PATHL = 1u << 9,
  MCD_SCAT_X = 1u << 10,
  MCD_SCAT_Y = 1u << 11,
  MCD_SCAT_Z = 1u << 12,
  MCD_SCAT_ENERGY = 1u << 13,
  MCD_SCAT_DX = 1u << 14,
  MCD_SCAT_DY = 1u << 15,
  MCD_SCAT_DZ = 1u << 16,
  MCD_SCAT_FLAGS = 1u << 17
};

/* Counters written by mcd_forward into caller-owned device memory. */
enum {
  MCD_CNT_EVENTS = 0,     /* scatter events */
  MCD_CNT_ELECTRONS = 1,  /* electrons finished */
  MCD_CNT_BKS = 2,
  MCD_CNT_ABS = 3,
  MCD_CNT_TRSM = 4,
  MCD_CNT_LONG = 5,
  MCD_CNT_FP64_PICKS = 6, /* picks settled by the FP64 guard */
  MCD_CNT_SCAT_ROWS = 7,  /* scatter rows requested (may exceed capacity) */
  MCD_CNT_QUEUE = 8,      /* internal: the launch's work counter (mcd_forward resets it) */
  MCD_CNT_GUARD_TIMEOUTS = 9, /* guard picks left unsettled: a served-guard wait that timed out, or
                                 a parked electron the resume kernel could not place (must be 0;
                                 otherwise the launch's records are not valid) */
  MCD_CNT_PARKED = 10,    /* internal, deferred guard: electrons parked by the first kernel
                             (mcd_forward resets it) */
  MCD_CNT_RESUMED = 11,   /* internal, deferred guard: the resume kernel's work counter */
  MCD_NCOUNTERS = 16
};

/* Launch tuning (0 = library default). */
typedef struct mcd_launch {
  int32_t block;          /* threads per workgroup, multiple of 64, at most 256 (served
                             guard: 128 to 1024, default 1024) */
  int32_t waves_per_simd; /* occupancy variant of the one-kernel FP32 transport: resident
                             wavefronts per SIMD its register budget guarantees (1 = compiler's
                             choice, 4 = default, 5, 6, 8; served guard: 6, 7 or 8 = default) */
  int32_t precision;      /* MCD_PREC_FP32 transport (default) or MCD_PREC_FP64 (all-FP64 run) */
  int32_t flags;          /* MCD_LAUNCH_* */
} mcd_launch;

/* DS in FP32 can run the FP64 guard pick inline (one kernel), deferred (two
 * kernels: the first has no FP64 code and parks an electron whose channel pick
 * falls in the guard band, the second resumes the parked electrons with the
 * FP64 pick inline) or served (one kernel in which wavefront 0 of each
 * workgroup settles the guard-band picks that the other wavefronts post to it
 * through LDS, so the transport code carries no FP64; for mcd_batch_run the
 * block may then be 128 to 1024).  Records are identical in every form.
 * Without a flag (and without block or waves_per_simd), mcd_forward runs DS
 * launches of at least 2^21 electrons served and smaller ones inline;
 * mcd_batch_run runs inline unless another form is asked for (the faster form
 * depends on the configuration; mcd/autotune.py measures it).  Setting more
 * than one flag is an error. */
#define MCD_LAUNCH_INLINE_GUARD 1
#define MCD_LAUNCH_DEFER_GUARD 2
#define MCD_LAUNCH_SERVICE_GUARD 4

/* Compaction period of mcd_forward and mcd_batch_run: a wavefront refills its
 * idle lanes with new electrons every 2^k steps, k = (flags &
 * MCD_LAUNCH_REFILL_MASK) >> MCD_LAUNCH_REFILL_SHIFT (0 = every step, the
 * default; at most 2^7), and at once when all its lanes are idle.  Records do
 * not depend on it. */
#define MCD_LAUNCH_REFILL_SHIFT 8
#define MCD_LAUNCH_REFILL_MASK (7 << MCD_LAUNCH_REFILL_SHIFT)

typedef struct mcd_ctx mcd_ctx;
typedef struct mcd_tables mcd_tables;

/* library version */
int mcd_version(void);

/* context */
int mcd_ctx_create(int device, mcd_ctx** out);
int mcd_ctx_destroy(mcd_ctx* ctx);
/* copies the last error message of ctx (or of the library if ctx is NULL) */
int mcd_last_error(const mcd_ctx* ctx, char* buf, size_t len);

/* Build the inverse-CDF tables on the device from host grids (FP64 or FP32
 * arithmetic), replacing probTable.computeBlock (probTables.py:415-462).
 * gryz may be NULL for continuous mode.  The grids are copied. */
int mcd_tables_create(mcd_ctx* ctx, const mcd_config* cfg, const mcd_table_grid* moller,
                      const mcd_table_grid* gryz, int precision, mcd_tables** out);
/* As mcd_tables_create, then replace the device CDFs with host CDFs in the
 * reference layout (nE x (nW-1), masked = 0), the tables the reference reads
 * back from its memmap files (probTable.readFromMemory, probTables.py:513-535).
 * A NULL cdf keeps the device-built table. */
int mcd_tables_upload(mcd_ctx* ctx, const mcd_config* cfg, const mcd_table_grid* moller, const float* cdf_moller,
                      const mcd_table_grid* gryz, const float* cdf_gryz, mcd_tables** out);
/* Copy one table's CDF back to the host in the reference layout. */
int mcd_tables_download(mcd_ctx* ctx, const mcd_tables* t, int which, float* cdf_host, size_t n_floats);
int mcd_tables_destroy(mcd_tables* t);
/* Channel-pick table diagnostics (no reference counterpart): node count, cells
 * with a cross-section crossing, cells always settled in FP64, guard band. */
int mcd_tables_info(const mcd_tables* t, int* n_nodes, int* n_swap_cells, int* n_exact_cells, float* guard_delta);
/* Rebuild the DS channel-pick table with n_nodes nodes (16..8192; tables are
 * built with 512).  No reference counterpart: the pick table is this build's
 * interpolation of the per-step sigma set (scattering.py:440-527).  The node
 * count changes FP32 rounding of the mean free path, not the picks; a
 * per-configuration launch reproduces a batch item's records when both use
 * the same count.  Waits for the device before freeing the old table. */
int mcd_tables_set_pick_nodes(mcd_ctx* ctx, mcd_tables* t, int n_nodes);

/* Tolerance-driven table grids (probTable.findWseries / findEseries,
 * MC/tools/probTables.py:280-382): the float32 W grid geomspace(Wmin,
 * (Emax - Ef)/2, nW) and E grid over [Emin, Emax] the reference's search
 * picks at tol_W / tol_E, with an FP64 adaptive Gauss-Kronrod integral in
 * place of scipy.integrate.quad.  Host code only (no device, no context).
 * kind: MCD_TABLE_MOLLER (Wmin = Wc, n_el = n_val, Eb = 0) or MCD_TABLE_GRYZ
 * (Wmin = Eb, n_el = ns of the shell; Emin raised to Eb when below it).
 * nE / nW: capacities of Es / Ws on entry, grid sizes on return; Es = Ws =
 * NULL returns the sizes only. */
int mcd_grid_search(int kind, double Wmin, double n_el, double Eb, double Ef, double Emin, double Emax,
                    double tol_E, double tol_W, float* Es, int32_t* nE, float* Ws, int32_t* nW);
/* One table's search parameters for mcd_grid_search_batch (the arguments of
 * mcd_grid_search). */
typedef struct mcd_grid_spec {
  int32_t kind; /* MCD_TABLE_MOLLER or MCD_TABLE_GRYZ */
  int32_t pad;
  double Wmin, n_el, Eb, Ef, Emin, Emax;
} mcd_grid_spec;
/* mcd_grid_search for n tables in one device launch (findWseries /
 * findEseries, MC/tools/probTables.py:280-382, for every table of a batch):
 * one workgroup per table runs the search's steps in order, the same
 * candidate sizes tested in the same order against the same FP64
 * Gauss-Kronrod reference integral, sums formed in numpy's pairwise order;
 * the sizes equal mcd_grid_search's, and the grids follow from them
 * (geomspace for W, linspace for E, as mcd_grid_search returns them).
 * nE[i], nW[i]: table i's sizes.  status (n entries, may be NULL): MCD_OK or
 * MCD_ERR_TABLE per table (no grid meets the tolerance, empty W range, or a
 * loss grid reaching the incident energy); with status NULL a failing table
 * fails the call with a message naming it. */
int mcd_grid_search_batch(mcd_ctx* ctx, const mcd_grid_spec* specs, int n, double tol_E, double tol_W, int32_t* nE,
                          int32_t* nW, int32_t* status);
/* Tables of a configuration from its input-file parameters (probTables.genTables,
 * MC/tools/probTables.py:541-614): the grid search above for the Moller table
 * and the outermost shell's Gryzinski table (the one every Gryzinski event
 * reads, scattering.py:625), then the CDF rows filled on the device in FP64
 * (precision MCD_PREC_FP64) or FP32.  Continuous mode: tables without
 * energy-loss tables, as mcd_tables_create with NULL grids. */
int mcd_tables_build(mcd_ctx* ctx, const mcd_config* cfg, double tol_E, double tol_W, int precision,
                     mcd_tables** out);
/* mcd_tables_build for n configurations at once (genTables, probTables.py:541-614,
 * for every configuration of a batch): the grid searches of every table run
 * in one device launch (mcd_grid_search_batch; MCD_GRID_SEARCH=host in the
 * environment runs them on up to 16 host threads instead), then one device
 * launch fills the CDF rows of every table and one more their row records.  out[i] receives configuration i's tables, equal bit
 * for bit to mcd_tables_build's; on an error every out[i] is NULL and the
 * message names the configuration. */
int mcd_tables_build_batch(mcd_ctx* ctx, const mcd_config* cfgs, int n, double tol_E, double tol_W, int precision,
                           mcd_tables** out);
/* The grid of one table (MCD_TABLE_MOLLER / MCD_TABLE_GRYZ): sizes in nE, nW;
 * Es and Ws (capacity nE / nW floats), when not NULL, receive the grid. */
int mcd_tables_grid(mcd_ctx* ctx, const mcd_tables* t, int which, int32_t* nE, int32_t* nW, float* Es, float* Ws);
/* Table files (the reference's '{type}_{shell}.table' memmaps,
 * probTables.py:489-535, whose names omit the parameters they were built
 * with): one file holds both tables with the parameters they depend on
 * (Z, E0, Emin, Wc, n_val, E_F, the outer shell's E_b and n_s, tol_E, tol_W,
 * precision), their grids and CDFs.  mcd_tables_save takes tables from
 * mcd_tables_build or mcd_tables_load; mcd_tables_load refuses a file built
 * for other parameters (MCD_ERR_TABLE) and otherwise gives the saved tables
 * bit for bit. */
int mcd_tables_save(mcd_ctx* ctx, const mcd_tables* t, const char* path);
int mcd_tables_load(mcd_ctx* ctx, const mcd_config* cfg, double tol_E, double tol_W, int precision,
                    const char* path, mcd_tables** out);

/* Run electrons [first, first+n) of one configuration.  els: device array of
 * n records (index = electron - first).  scats may be NULL.  counters: device
 * array of MCD_NCOUNTERS uint64, accumulated into (not cleared).  Philox key =
 * seed, counter = (electron low 32 bits, config_id | electron bits 32..47 << 16,
 * step, 0), so first + n must not exceed 2^48.  Asynchronous on stream. */
int mcd_forward(mcd_ctx* ctx, const mcd_config* cfg, const mcd_tables* tables, uint64_t seed,
                uint32_t config_id, uint64_t first, uint64_t n, mcd_electron_rec* els,
                const mcd_scatter_buf* scats, uint64_t* counters, const mcd_launch* launch,
                void* stream);

/* mcd_forward that allocates nothing: no reference counterpart; for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, records and
 * counters as mcd_forward.  A DS FP32 launch runs the deferred guard unless a
 * guard flag asks for another form, and parks its electrons in the library's
 * static device scratch (MCD_NOALLOC_MAX_ELECTRONS park records, ~72 MiB held
 * by each GPU the library runs on until the process ends); every other launch
 * form uses no scratch.  n above MCD_NOALLOC_MAX_ELECTRONS is rejected with
 * MCD_ERR_INVALID (split the range: records are keyed by electron id).  All
 * calls on one GPU share the scratch, so calls on one GPU must not overlap in
 * time (one stream, or synchronise between them); the work counters are the
 * caller's counters block and are cleared by the call with a fill kernel. */
#define MCD_NOALLOC_MAX_ELECTRONS (1 << 20)
int mcd_forward_noalloc(mcd_ctx* ctx, const mcd_config* cfg, const mcd_tables* tables, uint64_t seed,
                        uint32_t config_id, uint64_t first, uint64_t n, mcd_electron_rec* els,
                        const mcd_scatter_buf* scats, uint64_t* counters, const mcd_launch* launch,
                        void* stream);

/* Batched launches (no reference counterpart: the reference runs one
 * configuration per process, doScatter.py:155-252).  Item i uses Philox
 * counter word config = i and electrons [0, n). */
typedef struct mcd_batch_item {
  const mcd_config* cfg;
  const mcd_tables* tables;  /* DS: from mcd_tables_create for cfg; continuous: may be NULL */
  uint64_t n;
  mcd_electron_rec* els;     /* device, n records */
} mcd_batch_item;

typedef struct mcd_batch mcd_batch;

/* Prepare a batch (uploads per-configuration constants and builds channel-pick
 * tables of pick_nodes nodes; 0 sizes each table to an 8 KB per-wavefront LDS
 * slice: 256 nodes for 3 shells, 170 for 4, 128 for 8).
 * The caller's tables and els buffers must outlive the batch. */
int mcd_batch_create(mcd_ctx* ctx, const mcd_batch_item* items, int n_items, int pick_nodes, mcd_batch** out);
/* Run every item in one persistent launch.  counters: device array of
 * n_items x MCD_NCOUNTERS uint64 (accumulated into).  use_graph != 0 records
 * the launch sequence into a hipGraph once (per seed, counters and stream
 * shape) and replays it.  Items may hold 0 electrons (no wavefront takes
 * them; their counters stay untouched).  Asynchronous on stream. */
int mcd_batch_run(mcd_ctx* ctx, mcd_batch* b, uint64_t seed, uint64_t* counters, const mcd_launch* launch,
                  int use_graph, void* stream);
int mcd_batch_destroy(mcd_batch* b);

/* The selected fields (a mask of MCD_SCAT_*) of n_rows scatter rows in device
 * memory, written to device memory as columns: one column per selected field
 * in record order, n_rows values of the field's type each, every column
 * starting at a multiple of 256 bytes of out.  *needed receives the bytes the
 * columns take; out = NULL asks for that size only.  This is the on-device
 * filter of the scats columns doScatter.py writes (scatter_output,
 * MC/doScatter.py:112-117; fileTools.py:193-242 keeps only the requested
 * names): only the requested fields cross to the host.  Asynchronous on
 * stream. */
int mcd_scatter_pack(mcd_ctx* ctx, const mcd_scatter_rec* rows, uint64_t n_rows, uint32_t fields, void* out,
                     size_t out_bytes, size_t* needed, void* stream);

/* Differentiation (no reference counterpart).  Derivatives of
 *   Y   = number of backscattered electrons (outcome 'bks', electron.py:68-70)
 *   H_b = number of backscattered electrons whose exit energy (the pre-step
 *         energy of the exit step) falls in bin b of n_bins equal bins over
 *         [0, E0], the last bin closed
 * with respect to the parameters below, estimated by replaying electrons
 * [first, first+n) from their Philox counters (no stored trajectories):
 * survival biasing at the surface with pathwise tangents, likelihood-ratio
 * scores for the channel picks, and the threshold terms at Emin and at bin
 * edges as central differences along the energy tangent. */
enum {
  MCD_GRAD_WC = 0,     /* d/dWc [1/eV], DS (Moller cross section, crossSections.py:183-208,
                          and Moller losses, probTables.py:109) */
  MCD_GRAD_TILT = 1,   /* d/ds_tilt [1/degree] (initial direction, multiScatter.py:97-100) */
  MCD_GRAD_J = 2,      /* d/dJ [1/eV], continuous classical and JL stopping powers
                          (stoppingPowers.py:112-173) */
  MCD_GRAD_LNRHO = 3,  /* d/d ln(density): n, E_F ~ n^(2/3), E_pl ~ n^(1/2) move together
                          (material.py:127-131); for a density scale s, d/ds = this / s */
  MCD_NGRAD = 4
};
#define MCD_MAX_BINS 256

typedef struct mcd_bwd_opts {
  int32_t n_bins;     /* histogram bins over [0, E0], 1..MCD_MAX_BINS; 0 = 64 */
  float rel_step;     /* parameter step of the threshold terms, relative to the parameter
                         (absolute for ln density); 0 = 1e-2 */
  float graze;        /* flights with d_z in (-graze, 0) exit unsmoothed (their exit
                         indicator is not differentiated); 0 = 0.005 */
  int32_t block;      /* threads per workgroup (multiple of 64, at most 256); 0 = 256 */
} mcd_bwd_opts;

/* Exactly one of out_jac / out_vjp is non-NULL (device memory, accumulated
 * into, not cleared).
 *   out_jac: (1 + n_bins) x (1 + MCD_NGRAD) doubles, row 0 = Y, row 1+b = H_b;
 *            column 0 = the replay's estimate of the quantity, column 1+g =
 *            its derivative with respect to parameter g.
 *   out_vjp: with grad_out (device, 1 + n_bins floats: dL/dY, dL/dH_b), the
 *            1 + MCD_NGRAD doubles L = dL/dY*Y + sum_b dL/dH_b*H_b and dL/dparam.
 * Sums are over the n electrons (divide by n for per-electron values).
 * Derivatives a mode does not depend on (J in DS, Wc in continuous mode) are
 * 0.  Needs a semi-infinite target and maxScatt off.  Asynchronous on stream. */
int mcd_backward(mcd_ctx* ctx, const mcd_config* cfg, const mcd_tables* tables, uint64_t seed,
                 uint32_t config_id, uint64_t first, uint64_t n, const float* grad_out, double* out_vjp,
                 double* out_jac, const mcd_bwd_opts* opts, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MCD_H */
