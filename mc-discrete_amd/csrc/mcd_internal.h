// mcd_internal.h — declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mcd_device.h"

namespace mcd {

// An electron parked by the deferred-guard transport at the start of the step
// whose channel pick needs FP64: the resume kernel restarts that step from
// this state and the same Philox counter.
struct ParkRec {
  float Ehi, Elo;
  float p[3], pc[3], d[3], y[3];
  uint32_t step, flags;
  uint64_t e;
};

struct RunArgs {
  uint64_t seed;
  uint64_t first, n;
  mcd_electron_rec* els;
  mcd_scatter_rec* rows;
  uint64_t row_cap, row_max_electron;
  unsigned long long* row_cursor;
  unsigned long long* counters;
  unsigned long long* queue;
  ParkRec* park;                     // deferred guard: parked electrons (capacity n)
  unsigned long long* park_count;    // electrons parked by the first kernel
  unsigned long long* resume_queue;  // work counter of the resume kernel
  uint32_t refill_mask;              // idle lanes are refilled every (refill_mask + 1) steps
  uint32_t guard_spins;              // served guard: polls a transport wavefront waits for a reply
};

// Default of RunArgs::guard_spins: s_sleep(2) polls of 128 clocks, ~0.2 s in
// all; past it the step takes the Rutherford channel and is flagged
// MCD_FLAG_GUARD_TIMEOUT (counted as an error).  The environment variable
// MCD_GUARD_SPIN_CAP (a positive integer) replaces it, so that a test can
// drive this error path from the host with every wait still bounded.
constexpr uint32_t kGuardSpinCap = 1u << 22;

// One configuration of a batched launch (device-resident).  The device array
// holds the items grouped by XCD (group g = items [group_start[g],
// group_start[g + 1]), largest estimated work first within a group).
struct BatchItemDev {
  DevCfg c;
  RunArgs R;                   // els, n, queue (per item); seed/counters given at launch
  float cost;                  // estimated work per electron (load balancing)
  float cum_lo, cum_hi;        // the item's share of its group's work: initial wavefront placement
  int slot;                    // the item's index in the caller's list: its counter block
  unsigned long long* active;  // wavefronts currently working on the item
};

// At most this many XCD groups (MI355X has 8 XCDs).
constexpr int kMaxXcdGroups = 16;

hipError_t launch_batched(const BatchItemDev* items, int n_items, const int* group_start, int n_groups,
                          int slice_floats, uint64_t seed, uint32_t refill_mask, unsigned long long* counters,
                          int phase, int grid, int block, hipStream_t s);
int batched_blocks_per_cu(int phase, int block, int slice_floats, bool period);
int batched_cfg_floats();


hipError_t launch_count_outcomes_batched(const BatchItemDev* items, int n_items, unsigned long long* counters,
                                         hipStream_t s);

// Derivative replay (mcd_backward.hip).
struct BwdArgs {
  uint64_t seed, first, n;
  const float* nodes;          // node table (global; staged into LDS)
  int n_nodes, stride, n_ch;
  float lnE0, inv_dln;
  int n_bins;
  float bin_w, inv_bin_w;
  float delta[3], inv_2delta[3];  // threshold terms: parameter step per tangent slot
  float graze;                 // survival biasing for flights with d_z < -graze
  const float* grad_out;       // VJP: 1 + n_bins upstream gradients
  double* out;                 // VJP: 1 + MCD_NGRAD; Jacobian: (1 + n_bins) x (1 + MCD_NGRAD)
  unsigned long long* queue;
  int col_of_slot[3];          // output column of tangent slots 0..2
  const float* m_score_wc;     // DS: Wc score of each Moller table column pick
  const float* m_score_e;      // DS: energy score of each Moller table column pick
  const float* g_score_e;      // DS: energy score of each Gryzinski table column pick
  float inv_wc, inv_nw1;       // DS: 1 / Wc, 1 / (nW_Moller - 1)
};

hipError_t launch_bwd_nodes(const DevCfg& c, int n_nodes, double lnE0, double dln, float* nodes, int stride,
                            hipStream_t s);
int bwd_stride(int mode, int n_ch);
int bwd_blocks_per_cu(const DevCfg& c, const BwdArgs& A, bool jac, int block);
hipError_t launch_backward(const DevCfg& c, const BwdArgs& A, bool jac, int grid, int block, hipStream_t s);

hipError_t launch_transport(const DevCfg& c, const RunArgs& R, int precision, int minw, int grid, int block,
                            hipStream_t s);
int max_blocks_per_cu(const DevCfg& c, int precision, int minw, int block);
// DS FP32 with the FP64 guard deferred: phase 1 parks guarded steps, phase 2 resumes them
hipError_t launch_transport_deferred(const DevCfg& c, const RunArgs& R, int phase, int grid, int block, hipStream_t s);
int deferred_blocks_per_cu(const DevCfg& c, int phase, int block);
// DS FP32 with the FP64 guard served by wavefront 0 of each workgroup
hipError_t launch_transport_svc(const DevCfg& c, const RunArgs& R, int minw, int grid, int block, hipStream_t s);
int svc_blocks_per_cu(const DevCfg& c, int minw, int block);
bool valid_svc_minw(int w);
bool valid_minw(int w);
hipError_t launch_count_outcomes(const mcd_electron_rec* els, uint64_t n, unsigned long long* counters, hipStream_t s);
// Fields of mcd_scatter_rec to pack: source offset and size in the record,
// destination offset of the field's column in the output.
constexpr int kScatFields = 18;
struct ScatPackPlan {
  int n;
  int src[kScatFields];
  int size[kScatFields];
  uint64_t dst[kScatFields];
};
hipError_t launch_pack_scatter(const mcd_scatter_rec* rows, uint64_t n, const ScatPackPlan& plan, void* out,
                               hipStream_t s);

// tables
hipError_t launch_build_cdf(const float* Es, const float* Ws, int nE, int nW, double Ef, int kind, double n_el,
                            double Eb, int precision, float* cdf, hipStream_t s);
hipError_t launch_ratio_score(const float* c0, const float* cp, const float* cm, int nE, int nW, double inv_2delta,
                              float* out, hipStream_t s);
hipError_t launch_row_score(const float* cdf, const float* Es, int nE, int nW, float* out, hipStream_t s);
hipError_t launch_build_rows(const float* cdf, const float* Es, int nE, int nW, uint8_t* rows, hipStream_t s);
// One table of a batched build: its device grids and outputs, and the first
// of its rows in the batch's row numbering (jobs in order, rows consecutive).
struct CdfJob {
  const float* Es;
  const float* Ws;
  float* cdf;
  uint8_t* rows;
  int nE, nW, kind;
  double Ef, n_el, Eb;
  long long row0;
};
// CDF rows, then row records, of every job's table.
hipError_t launch_build_tables_batched(const CdfJob* jobs, int n_jobs, long long total_rows, int precision,
                                       hipStream_t s);
hipError_t launch_pick_nodes(const DevCfg& c, int n_nodes, double lnE0, double dln, float* nodes, int stride,
                             hipStream_t s);
hipError_t launch_pick_check(const DevCfg& c, int n_checks, float* err, int* bad, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, int n, hipStream_t s);
hipError_t launch_unmask(const float* src, float* dst, size_t n, hipStream_t s);
hipError_t launch_mask(const float* src, float* dst, size_t n, hipStream_t s);

// Tolerance-driven grid search of one energy-loss table (mcd_grid.hip):
// float32 Es over [Emin, Emax] and Ws over [Wmin, (Emax - Ef)/2].
int search_grid(int kind, double Wmin, double n_el, double Eb, double Ef, double Emin, double Emax, double tol_E,
                double tol_W, std::vector<float>& Es, std::vector<float>& Ws, char* err, size_t err_len);
// The reference integral of the search: dsigma/dW over [Wmin, (E - Ef)/2] at E.
double grid_reference_integral(int kind, double Wmin, double n_el, double Eb, double Ef, double E);

// The same search on the device for many tables (mcd_grid.hip): one
// workgroup per table; results are the grid sizes (the grids follow from
// them, grid_materialize) and the reference integral.
struct GridSearchJob {
  int kind;
  double Wmin, n_el, Eb, Ef, Emin, Emax;
};
struct GridSearchResult {
  int32_t nE, nW, status, pad;
  double ref;
};
constexpr int kGridOrders = 5, kGridPercents = 10;
// candidate sizes: ordersOfMag(1e6) and percents(n) of each (probTables.py:54-64)
struct GridCandidates {
  int32_t orders[kGridOrders];
  int32_t pct[kGridOrders][kGridPercents];
};
enum { kGridEmptyW = 1, kGridNoW, kGridWRefine, kGridWReachesE, kGridNoE, kGridERefine };
size_t grid_search_scratch_bytes(int n_jobs);
hipError_t launch_grid_search(const GridSearchJob* jobs, int n_jobs, double tol_E, double tol_W, void* scratch,
                              GridSearchResult* res, hipStream_t s);
void grid_materialize(const GridSearchJob& j, int nE, int nW, std::vector<float>& Es, std::vector<float>& Ws);
const char* grid_status_text(int status);

}  // namespace mcd
