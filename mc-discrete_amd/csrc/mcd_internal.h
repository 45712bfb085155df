// mcd_internal.h — declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "mcd_device.h"

namespace mcd {

struct RunArgs {
  uint64_t seed;
  uint64_t first, n;
  mcd_electron_rec* els;
  mcd_scatter_rec* rows;
  uint64_t row_cap, row_max_electron;
  unsigned long long* row_cursor;
  unsigned long long* counters;
  unsigned long long* queue;
};

// One configuration of a batched launch (device-resident).
struct BatchItemDev {
  DevCfg c;
  RunArgs R;                   // els, n, queue (per item); seed/counters given at launch
  float cost;                  // estimated work per electron (load balancing)
  float cum_lo, cum_hi;        // work-fraction interval for the initial wave assignment
  unsigned long long* active;  // wavefronts currently working on the item
};

hipError_t launch_batched(const BatchItemDev* items, int n_items, int slice_floats, uint64_t seed,
                          unsigned long long* counters, int grid, int block, hipStream_t s);
int batched_blocks_per_cu(int block, int slice_floats);
int batched_cfg_floats();
hipError_t launch_count_outcomes_batched(const BatchItemDev* items, int n_items, unsigned long long* counters,
                                         hipStream_t s);

hipError_t launch_transport(const DevCfg& c, const RunArgs& R, int precision, int minw, int grid, int block,
                            hipStream_t s);
int max_blocks_per_cu(const DevCfg& c, int precision, int minw, int block);
bool valid_minw(int w);
hipError_t launch_count_outcomes(const mcd_electron_rec* els, uint64_t n, unsigned long long* counters, hipStream_t s);

// tables
hipError_t launch_build_cdf(const float* Es, const float* Ws, int nE, int nW, double Ef, int kind, double n_el,
                            double Eb, int precision, float* cdf, hipStream_t s);
hipError_t launch_build_guide(const float* cdf, int nE, int nW, uint16_t* guide, hipStream_t s);
hipError_t launch_pick_nodes(const DevCfg& c, int n_nodes, double lnE0, double dln, float* nodes, int stride,
                             hipStream_t s);
hipError_t launch_pick_check(const DevCfg& c, int n_checks, float* err, int* bad, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, int n, hipStream_t s);
hipError_t launch_unmask(const float* src, float* dst, size_t n, hipStream_t s);
hipError_t launch_mask(const float* src, float* dst, size_t n, hipStream_t s);

}  // namespace mcd
