// mcd_transport.hip — persistent electron-transport kernels for gfx950.
//
// One electron per lane.  Each wavefront refills its idle lanes from a device
// work counter (ballot + popcount prefix, one atomic per refill), so lanes
// whose electron was absorbed or backscattered pick up the next electron
// immediately instead of idling until the wavefront's longest trajectory ends.
// The step follows trajectory_DS (MC/tools/singleScatter.py:10-41) and
// trajectory_cont_cl (singleScatter.py:205-239); see mcd_device.h for the
// numerics.
#include <hip/hip_runtime.h>

#include "mcd_device.h"
#include "mcd_internal.h"

namespace mcd {

// LDS of the one-configuration kernels: the pick table (nf floats), then the
// two loss-table descriptors at a 16-byte aligned offset.
__host__ __device__ constexpr int tab_offset_floats(int nf) { return (nf + 3) & ~3; }
static size_t single_lds_bytes(const DevCfg& c) {
  return (size_t)tab_offset_floats(c.pick.n_nodes * c.pick.stride) * sizeof(float) + 2 * sizeof(DevTable);
}

// Stage the pick table and the two loss-table descriptors (Moller,
// Gryzinski) that a lane selects by its channel.
__device__ __forceinline__ DevTable* stage_single(const DevCfg& c, float* lds) {
  const int nf = c.pick.n_nodes * c.pick.stride;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) lds[i] = c.pick.nodes[i];
  DevTable* tabs = reinterpret_cast<DevTable*>(lds + tab_offset_floats(nf));
  if (threadIdx.x == 0) {
    tabs[0] = c.tm;
    tabs[1] = c.tg;
  }
  return tabs;
}

// The served kernel's LDS copy of the pick table in the wide record
// (pick_stride_wide): each node's record, then its differences to the next
// node, each the same float subtraction the compact record's interpolation
// does at every step, so the interpolated values are the same bit for bit.
template <int NCH>
__device__ __forceinline__ DevTable* stage_single_wide(const DevCfg& c, float* lds) {
  constexpr int S = pick_stride(NCH), SW = pick_stride_wide(NCH), NB = NCH - 1, D = NB + 3;
  const int n = c.pick.n_nodes;
  const float* __restrict__ g = c.pick.nodes;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const float* a = g + (size_t)k * S;
    float* o = lds + (size_t)k * SW;
    for (int i = 0; i < D; ++i) o[i] = a[i];
    for (int i = 0; i <= NB; ++i) o[D + i] = k + 1 < n ? a[S + i] - a[i] : 0.f;
  }
  DevTable* tabs = reinterpret_cast<DevTable*>(lds + tab_offset_floats(n * SW));
  if (threadIdx.x == 0) {
    tabs[0] = c.tm;
    tabs[1] = c.tg;
  }
  return tabs;
}

// The record is written once, when the trajectory ends, as nine dword
// stores marked non-temporal: records of one wavefront land in different
// cache lines at different times, and streaming them past L2 keeps them from
// evicting the loss tables and from being read back to merge partial lines.
__device__ __forceinline__ void write_rec(mcd_electron_rec* rec, float E, const float* d, const float* p, uint32_t nsc,
                                          int o1, int o2, uint32_t flags) {
  uint32_t* w = reinterpret_cast<uint32_t*>(rec);
  __builtin_nontemporal_store(__float_as_uint(E), w + 0);
  __builtin_nontemporal_store(__float_as_uint(d[0]), w + 1);
  __builtin_nontemporal_store(__float_as_uint(d[1]), w + 2);
  __builtin_nontemporal_store(__float_as_uint(d[2]), w + 3);
  __builtin_nontemporal_store(__float_as_uint(p[0]), w + 4);
  __builtin_nontemporal_store(__float_as_uint(p[1]), w + 5);
  __builtin_nontemporal_store(__float_as_uint(p[2]), w + 6);
  __builtin_nontemporal_store(nsc, w + 7);
  // outcome, outcome2, flags, pad (bytes 32..35, little endian)
  __builtin_nontemporal_store((uint32_t)(o1 & 0xFF) | ((uint32_t)(o2 & 0xFF) << 8) | ((flags & 0xFFu) << 16), w + 8);
}

// wave-level append of scatter rows; returns the slot or ~0 when full
__device__ __forceinline__ uint64_t append_slot(bool want, unsigned long long* cursor, uint64_t cap, int lane) {
  uint64_t m = __ballot(want);
  if (!m) return ~0ull;
  int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(cursor, (unsigned long long)__popcll(m));
  base = lane_u64(base, leader);
  uint64_t slot = base + __popcll(m & ((1ull << lane) - 1ull));
  return (want && slot < cap) ? slot : ~0ull;
}

__device__ __forceinline__ void wave_add_counters(unsigned long long* cnt, const uint64_t* v, int nv, int lane) {
  for (int i = 0; i < nv; ++i) {
    uint64_t x = v[i];
    for (int off = 32; off > 0; off >>= 1) {
      uint32_t lo = __shfl_xor((uint32_t)x, off), hi = __shfl_xor((uint32_t)(x >> 32), off);
      x += ((uint64_t)hi << 32) | lo;
    }
    if (lane == 0 && x) atomicAdd(&cnt[i], (unsigned long long)x);
  }
}

// Counts of records i0, i0 + step, ... < n: each thread reads only the
// records' outcome dword (byte 32: outcome, outcome2, flags), the wavefronts
// reduce by shuffles, the block sums its wavefronts in `part` (one row of 5
// per wavefront, blockDim.x <= 256) and adds once per counter to out[0..4]
// (electrons, bks, abs, trsm, long).
static_assert(sizeof(mcd_electron_rec) == 36 && offsetof(mcd_electron_rec, outcome) == 32 &&
                  offsetof(mcd_electron_rec, outcome2) == 33,
              "count_outcome_words reads the outcome bytes as dword 8 of a 9-dword record");
__device__ __forceinline__ void count_outcome_words(const mcd_electron_rec* __restrict__ els, uint64_t n, uint64_t i0,
                                                    uint64_t step, uint32_t (*part)[5], unsigned long long* out) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(els) + 8;
  uint32_t v[5] = {0u, 0u, 0u, 0u, 0u};
  for (uint64_t i = i0; i < n; i += step) {
    const uint32_t x = __builtin_nontemporal_load(w + 9 * i);
    const uint32_t o1 = x & 0xFFu, o2 = (x >> 8) & 0xFFu;
    v[0] += 1u;
    v[1] += o1 == MCD_OUT_BKS;
    v[2] += (uint32_t)(o1 == MCD_OUT_ABS) + (uint32_t)(o2 == MCD_OUT_ABS);
    v[3] += o1 == MCD_OUT_TRSM;
    v[4] += o1 == MCD_OUT_LONG;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    uint32_t x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (lane == 0) part[wv][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    unsigned long long t = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += part[q][threadIdx.x];
    if (t) atomicAdd(out + threadIdx.x, t);
  }
}

// ---------------------------------------------------------------------------
// Service guard (DS, FP32, one kernel).  Wavefront 0 of a workgroup is a
// service wavefront: a transport wavefront posts the guard-band picks of a step
// to its LDS mailbox and waits; the service wavefront settles them in FP64, one
// of its lanes per posted lane, and posts the channels back.  The FP64 code is
// on a path no transport wavefront executes, so the kernel keeps the register
// budget of an FP64-free step without a second kernel or a park buffer.
struct GuardSlot {
  float Ehi, Elo, u;
  uint32_t perm_lo, perm_hi;
  int meta;  // posted: swap position | exact << 8; replied: the channel
};
struct GuardBox {
  unsigned long long mask;  // posted lanes
  uint32_t state;           // kBoxPosted, then kBoxReplied
  uint32_t bit;             // this mailbox's bit in the workgroup's pending word
  GuardSlot lane[64];
};
constexpr uint32_t kBoxPosted = 1u, kBoxReplied = 2u;
// A transport wavefront waits at most R.guard_spins s_sleep(2) polls for a
// reply (kGuardSpinCap by default); past it the step takes the Rutherford
// channel and is flagged MCD_FLAG_GUARD_TIMEOUT (counted as an error).

__device__ __forceinline__ uint32_t box_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void box_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The channel order d_pick_ordered expects: a crossing cell's swap was
// applied to pe.perm by the interpolation; undo it so the FP64 pick sees node
// k's order and decides the pair exactly.
template <int NCH>
__device__ __forceinline__ void guard_perm(const PickEval<NCH>& pe, uint32_t& plo, uint32_t& phi) {
  plo = pe.perm_lo;
  phi = pe.perm_hi;
  if (pe.swap != (int)kMetaNoSwap && pe.pB > pe.pA) {
    uint64_t pm = ((uint64_t)phi << 32) | plo;
    const uint64_t la = (pm >> (4 * pe.swap)) & 15u, lb = (pm >> (4 * (pe.swap + 1))) & 15u;
    pm &= ~((uint64_t)0xFFu << (4 * pe.swap));
    pm |= (lb << (4 * pe.swap)) | (la << (4 * (pe.swap + 1)));
    plo = (uint32_t)pm;
    phi = (uint32_t)(pm >> 32);
  }
}

// The service wavefront's loop: answer posted mailboxes until every transport
// wavefront of the workgroup has finished (each adds one to *done on exit;
// a wavefront exits only after its last reply, so none is left waiting).
// A posting wavefront sets its mailbox's bit in *pending, so an idle poll is
// one LDS read between sleeps of 512 clocks (the idle service wavefront takes
// few issue slots from the transport wavefronts on its SIMD).  The loop has
// no poll limit of its own, so a launch may run as long as its work needs:
// it ends because every transport wavefront ends (the work counter is finite,
// every trajectory is capped at kMaxStepsOff steps and every wait for a reply
// is bounded by R.guard_spins), and each of them adds to *done on exit.

template <int NCH>
__device__ void guard_service(const DevCfg& c, GuardBox* boxes, uint32_t* pending, uint32_t* done, int n_boxes) {
  const int lane = threadIdx.x & 63;
  const uint32_t nb = (uint32_t)__builtin_amdgcn_readfirstlane(n_boxes);
  // a waiting transport wavefront is blocked until the reply: serve first
  __builtin_amdgcn_s_setprio(3);
  while (true) {
    if (__builtin_amdgcn_readfirstlane(box_load(pending)) == 0u) {
      if (__builtin_amdgcn_readfirstlane(box_load(done)) == nb) break;
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    // all 64 lanes are active here: lane 0 takes the pending set for the wavefront
    uint32_t q = 0;
    if (lane == 0) q = __hip_atomic_exchange(pending, 0u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    q = __builtin_amdgcn_readfirstlane(q);
    while (q) {
      GuardBox& b = boxes[__builtin_ctz(q)];
      q &= q - 1u;
      if (box_load(&b.state) != kBoxPosted) continue;
      if ((b.mask >> lane) & 1ull) {
        GuardSlot& g = b.lane[lane];
        const int meta = g.meta;
        g.meta = d_pick_ordered<NCH>((double)g.Ehi + (double)g.Elo, (double)g.u, c, g.perm_lo, g.perm_hi,
                                     meta & 0xFF, (meta >> 8) & 1);
      }
      box_store(&b.state, kBoxReplied);
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 transport (DS and continuous).
// The persistent loop of one wavefront over the electrons of one
// configuration: refill idle lanes from the configuration's work counter,
// step every active lane, write the record when a trajectory ends; returns
// when the counter is exhausted and every lane is idle.
// PHASE 0: the FP64 guard pick inline.  PHASE 1 (DS): no FP64 code; a step
// whose pick falls in the guard band parks the electron (state at the start
// of the step) and the lane takes the next electron.  PHASE 2: refill from the
// parked electrons instead of new ones, guard inline.  Phases 1 + 2 give the
// same records as phase 0: a parked step is redone from the same state and
// Philox counter.  PHASE 3 (DS): guard-band picks go to the workgroup's
// service wavefront through `box`; same records as phase 0.
// EXPL (continuous mode): the explicit Bethe law, evaluated in FP64.
template <int MODE, int NCH, int PHASE = 0, bool PERIOD = false, bool EXPL = false, bool WIDE = false>
__device__ __forceinline__ void transport_loop(const DevCfg& c, const float* __restrict__ lds,
                                               const DevTable* __restrict__ tabs, const RunArgs& R,
                                               uint32_t& n_events, uint32_t& n_fp64, GuardBox* box = nullptr,
                                               uint32_t* pending = nullptr) {
  const int lane = threadIdx.x & 63;
  uint32_t k0 = (uint32_t)R.seed, k1 = (uint32_t)(R.seed >> 32);
  const uint64_t n_work = PHASE == 2 ? (uint64_t)__atomic_load_n(R.park_count, __ATOMIC_RELAXED) : R.n;

  bool active = false, open = true;
  WaveQueue wq;
  uint64_t e = 0;
  uint32_t ctr1 = 0;  // second Philox counter word of the lane's electron
  uint32_t step = 0, flags = 0;
  F2 E = {0.f, 0.f};
  float p[3] = {0.f, 0.f, 0.f}, pc[3] = {0.f, 0.f, 0.f}, d[3] = {0.f, 0.f, 0.f}, y[3] = {0.f, 0.f, 0.f};

  uint32_t iter = 0;
  const uint32_t refill_mask = PERIOD ? R.refill_mask : 0u;
  while (true) {
    // the key words stay loop-carried SGPR values (set where all lanes are
    // converged): otherwise the compiler hoists all twenty round keys out of
    // the loop and spills them to VGPR lanes, paying a v_readlane per round
    asm volatile("" : "+s"(k0), "+s"(k1));
    // compaction period (PERIOD kernels): idle lanes take new electrons every
    // (refill_mask + 1) steps, and at once when every lane is idle
    if (open && (!PERIOD || (iter++ & refill_mask) == 0u || !__any(active))) {
      uint64_t idle = __ballot(!active);
      if (idle) {
        const uint64_t idx = refill_index(wq, PHASE == 2 ? R.resume_queue : R.queue, idle, lane, n_work, open);
        if (!active) {
          if (idx < n_work) {
            active = true;
            if constexpr (PHASE == 2) {
              const ParkRec& q = R.park[idx];
              e = q.e;
              if (e - R.first >= R.n) {
                // not an electron of this launch: its record cannot be placed,
                // so the launch is reported invalid instead of writing out of range
                atomicAdd(R.counters + MCD_CNT_GUARD_TIMEOUTS, 1ull);
                active = false;
              } else {
                ctr1 = c.config_id | (uint32_t)(e >> 32) << 16;
                step = q.step;
                flags = q.flags;
                E.hi = q.Ehi; E.lo = q.Elo;
#pragma unroll
                for (int j = 0; j < 3; ++j) { p[j] = q.p[j]; pc[j] = q.pc[j]; d[j] = q.d[j]; y[j] = q.y[j]; }
              }
            } else {
              e = R.first + idx;
              ctr1 = c.config_id | (uint32_t)(e >> 32) << 16;
              step = 0;
              flags = 0;
              E.hi = c.E0_hi; E.lo = c.E0_lo;
              p[0] = p[1] = p[2] = 0.f;
              pc[0] = pc[1] = pc[2] = 0.f;
              d[0] = c.dir0_f[0]; d[1] = c.dir0_f[1]; d[2] = c.dir0_f[2];
              y[0] = 0.f; y[1] = 1.f; y[2] = 0.f;
            }
          }
        }
      }
    }
    if (!__any(active)) break;
    if constexpr (PHASE == 3) {
      // Wavefronts holding high-energy electrons (the longest remaining
      // trajectories) issue before the others on their SIMD, so the electrons
      // started last are not the ones left running alone after the work
      // counter is spent (measured: 2^22-electron launch 16.1 -> 15.4 ms; the
      // same rule applied only past the end of the counter gave 16.0 ms).
      const float span = c.E0_hi - c.Emin_hi;
      if (__any(active && E.hi > c.Emin_hi + 0.5f * span)) __builtin_amdgcn_s_setprio(3);
      else if (__any(active && E.hi > c.Emin_hi + 0.2f * span)) __builtin_amdgcn_s_setprio(2);
      else __builtin_amdgcn_s_setprio(1);
    }
    if (!active) continue;

    // ---- one loop iteration of trajectory_DS / trajectory_cont_cl ----
    if ((int64_t)step >= c.max_scatt) {  // num_scatt > maxScatt -> 'long'
      const float sp[3] = {p[0] - pc[0], p[1] - pc[1], p[2] - pc[2]};  // compensated, as for other outcomes
      write_rec(&R.els[e - R.first], E.hi, d, sp, step, MCD_OUT_LONG, 0, flags);
      active = false;
      continue;
    }
    const U4 r = philox10((uint32_t)e, ctr1, step, 0u, k0, k1);
    const F2 Epre = E;
    float l, Wh, Wl = 0.f, c2, s2, u_az;
    int ch = MCD_CH_RUTH;
    uint32_t sflags = 0;
    if (MODE == MCD_MODE_DS) {
      const float u0 = u01f(r.x), u1 = u01f(r.y), u2 = u01f(r.z);
      u_az = u01f(r.w);
      float mfp_ln2;
      PickEval<NCH> pe;
      ch = f_pick_table<NCH, WIDE>(lds, c.pick, E.hi, u0, mfp_ln2, pe);
      if constexpr (PHASE == 1) {
        if (ch < 0) {  // guard band: park the electron for the resume kernel
          const uint64_t m = __ballot(true);
          const int leader = __ffsll((unsigned long long)m) - 1;
          unsigned long long base = 0;
          if (lane == leader) base = atomicAdd(R.park_count, (unsigned long long)__popcll(m));
          const uint64_t slot = lane_u64(base, leader) + __popcll(m & ((1ull << lane) - 1ull));
          ParkRec& q = R.park[slot];
          q.Ehi = E.hi; q.Elo = E.lo;
#pragma unroll
          for (int j = 0; j < 3; ++j) { q.p[j] = p[j]; q.pc[j] = pc[j]; q.d[j] = d[j]; q.y[j] = y[j]; }
          q.step = step;
          q.flags = flags;
          q.e = e;
          active = false;
          continue;
        }
      } else if constexpr (PHASE == 3) {
        if (__any(ch < 0)) {  // guard band: post to the service wavefront and wait
          if (ch < 0) {
            GuardSlot& g = box->lane[lane];
            uint32_t plo, phi;
            guard_perm<NCH>(pe, plo, phi);
            g.Ehi = E.hi; g.Elo = E.lo; g.u = u0;
            g.perm_lo = plo; g.perm_hi = phi;
            g.meta = (pe.swap & 0xFF) | (pe.exact ? 0x100 : 0);
          }
          const uint64_t m = __ballot(ch < 0);
          const int leader = __ffsll((unsigned long long)__ballot(true)) - 1;
          if (lane == leader) {
            box->mask = m;
            box_store(&box->state, kBoxPosted);
            __hip_atomic_fetch_or(pending, box->bit, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
          uint32_t spins = 0;
          while (box_load(&box->state) != kBoxReplied && ++spins < R.guard_spins) __builtin_amdgcn_s_sleep(2);
          if (ch < 0) {
            if (spins < R.guard_spins) {
              ch = box->lane[lane].meta;
              sflags |= MCD_FLAG_FP64_PICK;
              n_fp64++;
            } else {
              ch = MCD_CH_RUTH;
              sflags |= MCD_FLAG_GUARD_TIMEOUT;
              atomicAdd(R.counters + MCD_CNT_GUARD_TIMEOUTS, 1ull);
            }
          }
        }
      } else {
        if (ch < 0) {  // guard band: settle the pick in FP64
          uint32_t plo, phi;
          guard_perm<NCH>(pe, plo, phi);
          ch = d_pick_ordered<NCH>((double)E.hi + (double)E.lo, (double)u0, c, plo, phi, pe.swap, pe.exact);
          sflags |= MCD_FLAG_FP64_PICK;
          n_fp64++;
        }
      }
      l = -mfp_ln2 * log2_f(u1);
      // 1/E for both the screening parameter and the loss fraction: one
      // reciprocal for the wavefront instead of one in each channel branch
      const float rE = __builtin_amdgcn_rcpf(E.hi);
      if (ch == MCD_CH_RUTH) {  // Rutherford_halfPol, scattering.py:125-152
        Wh = 0.f;
        const float a = c.c_alpha_f * rE;
        const float om = 1.f - u2;
        const float inv = __builtin_amdgcn_rcpf(om + a);
        s2 = a * u2 * inv;
        c2 = om * (1.f + a) * inv;
      } else if (ch == MCD_CH_QUINN) {  // scattering.py:643-656
        Wh = c.Epl_hi; Wl = c.Epl_lo;
        c2 = 1.f; s2 = 0.f;
      } else {  // Moller table, or the outer-shell Gryzinski table for all shells (:602-640)
        const DevTable& t = tabs[ch == MCD_CH_MOLLER ? 0 : 1];  // LDS: no per-lane kernel-argument loads
        const F2 Ec = E;
        int row;
        const int col = table_lookup(t, E.hi, [Ec](float v) { return f_lt_f2(v, Ec); }, u2, row);
        Wh = ldg(t.Ws + col);
        if (!f_lt_f2(Wh, E)) sflags |= MCD_FLAG_LOSS_GE_E;
        // binaryCollModel (:156-179): W <= (E - E_F)/2
        const F2 D = f2_sub(E, c.Ef2_hi, c.Ef2_lo);
        const float w2 = 2.f * Wh;
        const bool ok = !(D.hi < w2 || (D.hi == w2 && D.lo < 0.f));
        if (ok) {
          const float x = Wh * rE;
          const float sq = __builtin_amdgcn_sqrtf(1.f - x);
          c2 = 0.5f * (1.f + sq);
          s2 = 0.5f * x * __builtin_amdgcn_rcpf(1.f + sq);
        } else {
          c2 = s2 = __builtin_nanf("");
          sflags |= MCD_FLAG_NAN_DIR;
        }
      }
    } else {
      const float u0 = u01f(r.x), u1 = u01f(r.y);
      u_az = u01f(r.z);
      // Rutherford-only mean free path (scattering.py:780-797)
      // Joy's screening: the polar angle (Rutherford_halfPol) in every model
      const float a = c.c_alpha_f * __builtin_amdgcn_rcpf(E.hi);
      float sig;
      if (c.elastic >= MCD_ELASTIC_NIGRAM) {  // ruther_N_sigma(_wDefl), crossSections.py:91-147
        const float b = 0.48f * (c.nig_zp_f * __builtin_amdgcn_rcpf(E.hi));
        sig = c.nig_num_f * __builtin_amdgcn_rcpf(4.f * b * (1.f + b) * (E.hi * E.hi));
      } else {
        const float Ek = E.hi * 1e-3f;
        const float rr = (Ek + 511.f) * __builtin_amdgcn_rcpf(Ek + 1024.f);
        sig = c.ruth_k_f * rr * rr * __builtin_amdgcn_rcpf(Ek * Ek * a * (1.f + a));
      }
      const float mfp = c.mfp_k_f * __builtin_amdgcn_rcpf(sig);
      l = -mfp * ln_f(u0);
      if (!(l < 1e4f)) sflags |= MCD_FLAG_LONG_PATH;
      // Bethe stopping power at the pre-step energy (stoppingPowers.py:112-215)
      if constexpr (EXPL) {
        // explicit law (:177-215), its shell sum folded to Z_eff (ln E - ln E*),
        // in FP64: near Emin ln E and ln E* are close, and their FP32
        // difference would lose most of its bits; the loss l*S is kept as a pair
        const double Ed = (double)E.hi + (double)E.lo;
        const double Wd = (double)l * (c.expl_k * (log(Ed) - c.lnEstar) / Ed);
        Wh = (float)Wd;
        Wl = (float)(Wd - (double)Wh);
      } else {
        float S;
        const float iE = __builtin_amdgcn_rcpf(E.hi);
        if (c.bethe == MCD_BETHE_CLASSICAL) S = c.bethe_k1_f * iE * ln_f(c.bethe_c1_f * E.hi);
        else S = c.bethe_k1_f * iE * ln_f(c.bethe_c1_f * (E.hi + c.bethe_kJ_f));
        Wh = l * S;
      }
      if (!f_lt_f2(Wh, E)) sflags |= MCD_FLAG_LOSS_GE_E;
      const float om = 1.f - u1;
      const float inv = __builtin_amdgcn_rcpf(om + a);
      s2 = a * u1 * inv;
      c2 = om * (1.f + a) * inv;
    }
    n_events++;

    // update_xyz (electron.py:64-84), Kahan-compensated position
    const float v0 = l * d[0] - pc[0], v1 = l * d[1] - pc[1], v2 = l * d[2] - pc[2];
    const float t0 = p[0] + v0, t1 = p[1] + v1, t2 = p[2] + v2;
    int o1 = MCD_OUT_NONE, o2 = MCD_OUT_NONE;
    const float pre0 = p[0], pre1 = p[1], pre2 = p[2];
    if (t2 <= 0.f) o1 = MCD_OUT_BKS;
    else if (t2 >= c.thickness_f) o1 = MCD_OUT_TRSM;
    else {
      pc[0] = (t0 - p[0]) - v0; pc[1] = (t1 - p[1]) - v1; pc[2] = (t2 - p[2]) - v2;
      p[0] = t0; p[1] = t1; p[2] = t2;
    }
    // update_energy (electron.py:44-59): absorbed when E - W <= Emin
    E = f2_sub(E, Wh, Wl);
    if (f2_le(E, c.Emin_hi, c.Emin_lo)) {
      if (o1 == MCD_OUT_NONE) o1 = MCD_OUT_ABS; else o2 = MCD_OUT_ABS;
    }
    flags |= sflags;
    const bool done = o1 != MCD_OUT_NONE;
    if (done) {
      // saveOutcomes: pre-loss energy, pre-rotation direction, last interior position
      const float sp[3] = {p[0] - pc[0], p[1] - pc[1], p[2] - pc[2]};
      write_rec(&R.els[e - R.first], Epre.hi, d, sp, step + 1, o1, o2, flags);
    }
    const bool rec_row = R.rows != nullptr && e < R.row_max_electron;
    if (!done || rec_row) f_rotate(c2, s2, u_az, d, y);  // update_direction
    if (R.rows != nullptr) {
      const uint64_t slot = append_slot(rec_row, R.row_cursor, R.row_cap, lane);
      if (slot != ~0ull) {
        mcd_scatter_rec* w = &R.rows[slot];
        w->electron = (uint32_t)e; w->step = step; w->config = (uint16_t)c.config_id;
        w->channel = (uint8_t)ch; w->outcome = (uint8_t)(o2 ? o2 : o1);
        w->E = Epre.hi; w->E_loss = Wh + Wl; w->pol = c2; w->az = kPiF * u_az; w->pathl = l;
        w->x = pre0; w->y = pre1; w->z = pre2;
        w->energy = E.hi;
        w->dx = d[0]; w->dy = d[1]; w->dz = d[2];
        w->step_flags = sflags;
      }
    }
    step++;
    if (done) active = false;
  }
}

// MINW: minimum resident wavefronts per SIMD the register allocation must
// allow (1 = compiler's choice, the default; others per launch).
template <int MODE, int NCH, int MINW, int PHASE = 0, bool EXPL = false>
__global__ __launch_bounds__(256, MINW) void k_transport_f32(DevCfg c, RunArgs R) {
  extern __shared__ float lds[];
  DevTable* tabs = nullptr;
  if (MODE == MCD_MODE_DS) {
    tabs = stage_single(c, lds);
    __syncthreads();
  }
  uint32_t n_events = 0, n_fp64 = 0;
  transport_loop<MODE, NCH, PHASE, true, EXPL>(c, lds, tabs, R, n_events, n_fp64);
  const int lane = threadIdx.x & 63;
  const uint64_t cnt[2] = {n_events, n_fp64};
  wave_add_counters(R.counters + MCD_CNT_EVENTS, cnt, 1, lane);
  wave_add_counters(R.counters + MCD_CNT_FP64_PICKS, cnt + 1, 1, lane);
}

// Service-guard kernel (DS, FP32).  LDS: the pick table (wide records), the two loss-table
// descriptors, one GuardBox per transport wavefront, the finished-wavefront
// count and the pending-mailbox bits.  Wavefront 0 serves the guard picks, the others run the transport.
// MINW: resident wavefronts per SIMD the register allocation must allow.
__host__ __device__ inline size_t svc_box_offset(const DevCfg& c) {
  return ((size_t)tab_offset_floats(c.pick.n_nodes * pick_stride_wide(c.pick.n_ch)) * sizeof(float) +
          2 * sizeof(DevTable) + 15) &
         ~(size_t)15;
}
static size_t svc_lds_bytes(const DevCfg& c, int block) {
  return svc_box_offset(c) + (size_t)(block / 64 - 1) * sizeof(GuardBox) + 16;
}

template <int NCH, int MINW>
__global__ __launch_bounds__(1024, MINW) void k_transport_svc(DevCfg c, RunArgs R) {
  extern __shared__ float lds[];
  DevTable* tabs = stage_single_wide<NCH>(c, lds);
  const int n_boxes = (int)(blockDim.x >> 6) - 1;
  GuardBox* boxes = reinterpret_cast<GuardBox*>(reinterpret_cast<char*>(lds) + svc_box_offset(c));
  uint32_t* done = reinterpret_cast<uint32_t*>(boxes + n_boxes);
  uint32_t* pending = done + 1;
  if (threadIdx.x == 0) {
    *done = 0u;
    *pending = 0u;
  }
  if ((int)threadIdx.x < n_boxes) {
    boxes[threadIdx.x].state = 0u;
    boxes[threadIdx.x].bit = 1u << threadIdx.x;
  }
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (w == 0) {
    guard_service<NCH>(c, boxes, pending, done, n_boxes);
    return;
  }
  uint32_t n_events = 0, n_fp64 = 0;
  transport_loop<MCD_MODE_DS, NCH, 3, true, false, true>(c, lds, tabs, R, n_events, n_fp64, boxes + (w - 1),
                                                         pending);
  const int lane = threadIdx.x & 63;
  // every lane is converged here; the service wavefront stops once all are counted
  if (lane == 0) __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  const uint64_t cnt[2] = {n_events, n_fp64};
  wave_add_counters(R.counters + MCD_CNT_EVENTS, cnt, 1, lane);
  wave_add_counters(R.counters + MCD_CNT_FP64_PICKS, cnt + 1, 1, lane);
}

// ---------------------------------------------------------------------------
// Batched launch: many configurations in one persistent grid.  A wavefront
// works on one configuration at a time (its constants and channel-pick table
// staged in the wavefront's own LDS slice), refills
// from that configuration's counter and, once the counter is exhausted and its
// lanes are idle, moves to the configuration with the most remaining work per
// active wavefront.  Initial placement spreads wavefronts over configurations
// in proportion to their estimated work.
// Work of an item in a phase: new electrons (phases 0, 1) or parked ones (2).
template <int PHASE>
__device__ __forceinline__ void item_work(const BatchItemDev& it, unsigned long long& done, unsigned long long& total) {
  if (PHASE == 2) {
    done = __atomic_load_n(it.R.resume_queue, __ATOMIC_RELAXED);
    total = __atomic_load_n(it.R.park_count, __ATOMIC_RELAXED);
  } else {
    done = __atomic_load_n(it.R.queue, __ATOMIC_RELAXED);
    total = it.R.n;
  }
}

// The item with the most estimated work left per wavefront already on it,
// among items [lo, hi); -1 when none has electrons left to claim.
template <int PHASE>
__device__ __forceinline__ int most_work_left(const BatchItemDev* __restrict__ items, int lo, int hi, int lane) {
  float best = 0.f;
  int bi = -1;
  for (int c = lo + lane; c < hi; c += 64) {
    unsigned long long q, n;
    item_work<PHASE>(items[c], q, n);
    if (q < n) {
      const unsigned long long act = __atomic_load_n(items[c].active, __ATOMIC_RELAXED);
      const float rem = (float)(n - q) * items[c].cost / (float)(act + 1);
      if (rem > best) { best = rem; bi = c; }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const float ob = __shfl_xor(best, off);
    const int oi = __shfl_xor(bi, off);
    if (ob > best || (ob == best && oi >= 0 && (bi < 0 || oi < bi))) { best = ob; bi = oi; }
  }
  return __builtin_amdgcn_readfirstlane(bi);
}

// The next item for a wavefront of XCD group [lo, hi): one of the group's
// items while any has electrons left to claim, so that the XCD's L2 holds the
// tables of its own group only; after that, help elsewhere.
template <int PHASE>
__device__ __forceinline__ int next_item(const BatchItemDev* __restrict__ items, int n_items, int lo, int hi,
                                         int lane) {
  const int ci = most_work_left<PHASE>(items, lo, hi, lane);
  return ci >= 0 ? ci : most_work_left<PHASE>(items, 0, n_items, lane);
}

// The XCD (accelerator complex die) the wavefront runs on: HW_REG_XCC_ID
// bits [3:0] (hardware register 20 of the gfx94x/gfx950 family).
__device__ __forceinline__ int xcc_id() {
  return (int)(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u);
}

// The wavefront's LDS slice holds a copy of the configuration constants
// followed by the channel-pick table, so the step loop reads both from LDS.
constexpr int kCfgFloats = (int)((sizeof(DevCfg) + 15) / 16 * 4);

int batched_cfg_floats() { return kCfgFloats; }

// PERIOD: idle lanes take new electrons every refill_mask + 1 steps (the
// launch's compaction period), and at once when all are idle; otherwise every
// step.  A separate instantiation, so the every-step kernels keep their
// register budget.  c: the item's constants (scalar loads); cl: the
// wavefront's LDS copy, whose two table descriptors the lanes select from.
template <int MODE, int NCH, int PHASE, bool PERIOD, bool EXPL = false>
__device__ __forceinline__ void run_item(const DevCfg& c, const DevCfg& cl, const RunArgs& itR, const float* pick,
                                         uint64_t seed, uint32_t refill_mask, unsigned long long* ic, uint32_t& ne,
                                         uint32_t& nf, GuardBox* box, uint32_t* pending) {
  RunArgs R = itR;
  R.seed = seed;
  R.refill_mask = refill_mask;
  // the item's counter block of this launch: phases 2 and 3 count errors in it
  if constexpr (PHASE >= 2) R.counters = ic;
  // the table descriptors from the LDS copy: the per-lane Moller / Gryzinski select reads LDS
  transport_loop<MODE, NCH, MODE == MCD_MODE_DS ? PHASE : 0, PERIOD, EXPL>(c, pick, &cl.tm, R, ne, nf, box, pending);
}

// The batched kernel's service wavefront: as guard_service, for mailboxes
// whose transport wavefronts may be on different configurations; the
// configuration of mailbox w is the LDS copy in wavefront w + 1's slice (that
// wavefront waits for the reply, so the copy stays in place).
__device__ void guard_service_batched(const float* lds, int slice_floats, GuardBox* boxes, uint32_t* pending,
                                      uint32_t* done, int n_boxes) {
  const int lane = threadIdx.x & 63;
  const uint32_t nb = (uint32_t)__builtin_amdgcn_readfirstlane(n_boxes);
  __builtin_amdgcn_s_setprio(3);
  while (true) {  // ends as guard_service does: every transport wavefront ends and adds to *done
    if (__builtin_amdgcn_readfirstlane(box_load(pending)) == 0u) {
      if (__builtin_amdgcn_readfirstlane(box_load(done)) == nb) break;
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    uint32_t q = 0;
    if (lane == 0) q = __hip_atomic_exchange(pending, 0u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    q = __builtin_amdgcn_readfirstlane(q);
    while (q) {
      const int w = __builtin_ctz(q);
      GuardBox& b = boxes[w];
      q &= q - 1u;
      if (box_load(&b.state) != kBoxPosted) continue;
      const DevCfg& c = *reinterpret_cast<const DevCfg*>(lds + (size_t)(w + 1) * slice_floats);
      const int nch = __builtin_amdgcn_readfirstlane(c.n_shells + 3);
      if ((b.mask >> lane) & 1ull) {
        GuardSlot& g = b.lane[lane];
        const int meta = g.meta;
        const double E = (double)g.Ehi + (double)g.Elo, u = (double)g.u;
        const int sw = meta & 0xFF;
        const bool ex = (meta >> 8) & 1;
        int ch;
        switch (nch) {
          case 4: ch = d_pick_ordered<4>(E, u, c, g.perm_lo, g.perm_hi, sw, ex); break;
          case 5: ch = d_pick_ordered<5>(E, u, c, g.perm_lo, g.perm_hi, sw, ex); break;
          case 6: ch = d_pick_ordered<6>(E, u, c, g.perm_lo, g.perm_hi, sw, ex); break;
          case 7: ch = d_pick_ordered<7>(E, u, c, g.perm_lo, g.perm_hi, sw, ex); break;
          case 8: ch = d_pick_ordered<8>(E, u, c, g.perm_lo, g.perm_hi, sw, ex); break;
          case 9: ch = d_pick_ordered<9>(E, u, c, g.perm_lo, g.perm_hi, sw, ex); break;
          case 10: ch = d_pick_ordered<10>(E, u, c, g.perm_lo, g.perm_hi, sw, ex); break;
          default: ch = d_pick_ordered<11>(E, u, c, g.perm_lo, g.perm_hi, sw, ex); break;
        }
        g.meta = ch;
      }
      box_store(&b.state, kBoxReplied);
    }
  }
}

// LDS of the batched kernel: one slice per wavefront, then (served guard)
// one mailbox per transport wavefront and the finished count and pending bits.
__host__ __device__ inline size_t batched_box_offset(int block, int slice_floats) {
  return ((size_t)(block / 64) * slice_floats * sizeof(float) + 15) & ~(size_t)15;
}
static size_t batched_lds_bytes(int phase, int block, int slice_floats) {
  if (phase != 3) return (size_t)(block / 64) * slice_floats * sizeof(float);
  return batched_box_offset(block, slice_floats) + (size_t)(block / 64 - 1) * sizeof(GuardBox) + 16;
}

// PHASE 0: guard inline.  PHASE 1: DS items park guarded steps (no FP64 code
// in the kernel).  PHASE 2: DS items resume their parked electrons.  PHASE 3:
// wavefront 0 of each workgroup serves the guard picks of the others (no FP64
// code on the transport path, one phase).
// The items are read through the constant address space: they do not change
// during the launch, so the item's configuration constants, the same for the
// whole wavefront, are scalar loads into SGPRs (reloaded where needed) rather
// than per-lane LDS reads held in VGPRs.
// Register budgets measured on MI355X for the 64-configuration batch: 5
// resident wavefronts per SIMD for phases 0, 1 and 3 (phase 0 spills at its
// refill, FP64-guard and item-switch code and still beats 4 and 6
// wavefronts), 4 for phase 2.
template <int PHASE, bool PERIOD>
__global__ __launch_bounds__(PHASE == 3 ? 1024 : 256, PHASE == 2 ? 4 : 5) void k_transport_batched(
    const __attribute__((address_space(4))) BatchItemDev* __restrict__ citems, int n_items,
    const int* __restrict__ group_start, int n_groups, int slice_floats, uint64_t seed, uint32_t refill_mask,
    unsigned long long* __restrict__ counters) {
  const BatchItemDev* items = (const BatchItemDev*)citems;
  extern __shared__ float lds[];
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  float* my = lds + (size_t)wib * slice_floats;
  float* pick = my + kCfgFloats;
  GuardBox* boxes = nullptr;
  uint32_t *done = nullptr, *pending = nullptr;
  if constexpr (PHASE == 3) {
    const int n_boxes = (int)(blockDim.x >> 6) - 1;
    boxes = reinterpret_cast<GuardBox*>(reinterpret_cast<char*>(lds) + batched_box_offset(blockDim.x, slice_floats));
    done = reinterpret_cast<uint32_t*>(boxes + n_boxes);
    pending = done + 1;
    if (threadIdx.x == 0) {
      *done = 0u;
      *pending = 0u;
    }
    if ((int)threadIdx.x < n_boxes) {
      boxes[threadIdx.x].state = 0u;
      boxes[threadIdx.x].bit = 1u << threadIdx.x;
    }
    __syncthreads();
    if (wib == 0) {
      guard_service_batched(lds, slice_floats, boxes, pending, done, n_boxes);
      return;
    }
  }
  GuardBox* box = PHASE == 3 ? boxes + (wib - 1) : nullptr;
  const int grp = xcc_id() % n_groups;
  const int glo = __builtin_amdgcn_readfirstlane(group_start[grp]);
  const int ghi = __builtin_amdgcn_readfirstlane(group_start[grp + 1]);
  int ci = -1;
  if (PHASE != 2) {
    // spread the group's wavefronts over its items in proportion to their
    // estimated work, so a wavefront mostly stays on one configuration
    // (blocks are placed on the XCDs round-robin: this XCD holds every
    // n_groups-th block)
    const int wpb = (int)(blockDim.x >> 6);
    const int local = (int)(blockIdx.x / n_groups) * wpb + wib;
    const int total = (int)((gridDim.x + n_groups - 1) / n_groups) * wpb;
    const float f = ((float)local + 0.5f) / (float)total;
    for (int c = glo; c < ghi; ++c)
      if (f >= items[c].cum_lo && f < items[c].cum_hi) ci = c;
    ci = __builtin_amdgcn_readfirstlane(ci);
  }
  if (ci < 0 || __atomic_load_n(items[ci].R.queue, __ATOMIC_RELAXED) >= items[ci].R.n)
    ci = next_item<PHASE>(items, n_items, glo, ghi, lane);
  while (ci >= 0) {
    const BatchItemDev& it = items[ci];
    if (lane == 0) atomicAdd(it.active, 1ull);
    // stage the configuration and its pick table in this wavefront's slice
    const float* src = reinterpret_cast<const float*>(&it.c);
    for (int i = lane; i < (int)(sizeof(DevCfg) / 4); i += 64) my[i] = src[i];
    const int mode = it.c.mode;
    const int nch = it.c.n_shells + 3;
    if (mode == MCD_MODE_DS) {
      const int nfl = it.c.pick.n_nodes * it.c.pick.stride;
      for (int i = lane; i < nfl; i += 64) pick[i] = it.c.pick.nodes[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const DevCfg& c = *reinterpret_cast<const DevCfg*>(my);
    const DevCfg& cg = (const DevCfg&)citems[ci].c;  // the same constants as scalar loads
    // set only where the loop uses it, so it is not live across the other phases' loops
    unsigned long long* icl = PHASE >= 2 ? counters + (size_t)it.slot * MCD_NCOUNTERS : nullptr;
    uint32_t ne = 0, nf = 0;
    if (mode == MCD_MODE_CONT) {
      if (PHASE != 2) {
        if (c.bethe == MCD_BETHE_EXPLICIT)
          run_item<MCD_MODE_CONT, 4, PHASE, PERIOD, true>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, nullptr, nullptr);
        else
          run_item<MCD_MODE_CONT, 4, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, nullptr, nullptr);
      }
    } else {
      switch (nch) {
        case 4: run_item<MCD_MODE_DS, 4, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, box, pending); break;
        case 5: run_item<MCD_MODE_DS, 5, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, box, pending); break;
        case 6: run_item<MCD_MODE_DS, 6, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, box, pending); break;
        case 7: run_item<MCD_MODE_DS, 7, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, box, pending); break;
        case 8: run_item<MCD_MODE_DS, 8, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, box, pending); break;
        case 9: run_item<MCD_MODE_DS, 9, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, box, pending); break;
        case 10: run_item<MCD_MODE_DS, 10, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, box, pending); break;
        default: run_item<MCD_MODE_DS, 11, PHASE, PERIOD>(cg, c, it.R, pick, seed, refill_mask, icl, ne, nf, box, pending); break;
      }
    }
    const uint64_t cnt[2] = {ne, nf};
    unsigned long long* ic = counters + (size_t)it.slot * MCD_NCOUNTERS;
    wave_add_counters(ic + MCD_CNT_EVENTS, cnt, 1, lane);
    wave_add_counters(ic + MCD_CNT_FP64_PICKS, cnt + 1, 1, lane);
    if (lane == 0) atomicSub(it.active, 1ull);
    ci = next_item<PHASE>(items, n_items, glo, ghi, lane);
    __builtin_amdgcn_wave_barrier();
  }
  if constexpr (PHASE == 3) {
    // every lane is converged here; the service wavefront stops once all are counted
    if (lane == 0) __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

template <int PHASE, bool PERIOD>
static const void* batched_kernel() {
  return reinterpret_cast<const void*>(&k_transport_batched<PHASE, PERIOD>);
}

// The kernel of a phase; the phases that take new electrons (0, 1, 3) have a
// compaction-period form, the resume phase (2) refills every step.
static const void* batched_kernel(int phase, bool period) {
  switch (phase) {
    case 1: return period ? batched_kernel<1, true>() : batched_kernel<1, false>();
    case 2: return batched_kernel<2, false>();
    case 3: return period ? batched_kernel<3, true>() : batched_kernel<3, false>();
    default: return period ? batched_kernel<0, true>() : batched_kernel<0, false>();
  }
}

hipError_t launch_batched(const BatchItemDev* items, int n_items, const int* group_start, int n_groups,
                          int slice_floats, uint64_t seed, uint32_t refill_mask, unsigned long long* counters,
                          int phase, int grid, int block, hipStream_t s) {
  if (phase == 3 && (block < 128 || block > 1024)) return hipErrorInvalidValue;
  if (n_groups < 1 || n_groups > kMaxXcdGroups) return hipErrorInvalidValue;
  const size_t lds = batched_lds_bytes(phase, block, slice_floats);
  const void* k = batched_kernel(phase, refill_mask != 0u);
  void* args[] = {(void*)&items, (void*)&n_items, (void*)&group_start, (void*)&n_groups, (void*)&slice_floats,
                  (void*)&seed, (void*)&refill_mask, (void*)&counters};
  return hipLaunchKernel(k, dim3(grid), dim3(block), args, lds, s);
}

int batched_blocks_per_cu(int phase, int block, int slice_floats, bool period) {
  int nb = 0;
  const size_t lds = batched_lds_bytes(phase, block, slice_floats);
  const void* k = batched_kernel(phase, period);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, block, lds) != hipSuccess || nb < 1) nb = 1;
  return nb;
}

__global__ __launch_bounds__(256) void k_count_outcomes_batched(const BatchItemDev* __restrict__ items,
                                                                unsigned long long* counters) {
  __shared__ uint32_t part[4][5];
  const BatchItemDev& it = items[blockIdx.y];
  count_outcome_words(it.R.els, it.R.n, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x,
                      (uint64_t)gridDim.x * blockDim.x, part,
                      counters + (size_t)it.slot * MCD_NCOUNTERS + MCD_CNT_ELECTRONS);
}

hipError_t launch_count_outcomes_batched(const BatchItemDev* items, int n_items, unsigned long long* counters,
                                         hipStream_t s) {
  hipLaunchKernelGGL(k_count_outcomes_batched, dim3(64, n_items), dim3(256), 0, s, items, counters);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 transport: the same loop with every quantity in double and the exact
// pick each step (the all-FP64 GPU run).  Arithmetic is not contracted into
// FMAs so the operation order matches the reference expressions.
struct DStep {
  double l, W, c2, hp;
  int ch;
  uint32_t sflags;
};

#pragma clang fp contract(off)
template <int MODE>
__device__ __forceinline__ DStep d_step_physics(const DevCfg& c, double E, const U4 r) {
  DStep o;
  o.sflags = 0;
  o.ch = MCD_CH_RUTH;
  if (MODE == MCD_MODE_DS) {
    const double u0 = u01d(r.x), u1 = u01d(r.y), u2 = u01d(r.z), u3 = u01d(r.w);
    double mfp;
    o.ch = d_pick(E, u0, c, &mfp);          // det_type, scattering.py:530-549
    o.l = -mfp * log(u1);                   // compute_pathl, :551-565
    if (o.ch == MCD_CH_RUTH) {              // :575-586
      o.W = 0.;
      const double a = c.c_alpha / E;
      o.c2 = 1. - (a * u2 / (1. + a - u2));
    } else if (o.ch == MCD_CH_QUINN) {      // :643-656
      o.W = c.Epl;
      o.c2 = 1.;
    } else {                                // :602-640
      const DevTable& t = (o.ch == MCD_CH_MOLLER) ? c.tm : c.tg;
      const double Ec = E;
      int row;
      const int col = table_lookup(t, (float)E, [Ec](float v) { return (double)v < Ec; }, (float)u2, row);
      o.W = (double)t.Ws[col];
      if (!(o.W < E)) o.sflags |= MCD_FLAG_LOSS_GE_E;
      if (o.W <= (E - c.Ef) * 0.5) o.c2 = 0.5 * (sqrt(1 - (o.W / E)) + 1);
      else { o.c2 = __builtin_nan(""); o.sflags |= MCD_FLAG_NAN_DIR; }
    }
    o.hp = kPi * u3;
  } else {
    const double u0 = u01d(r.x), u1 = u01d(r.y), u2 = u01d(r.z);
    const double sr = d_sigma_ruth(E, c);
    const double mfp = 1. / (c.atnd * sr) * 1e14;  // scattering.py:792
    o.l = -mfp * log(u0);                          // :802-813
    if (!(o.l < 1e4)) o.sflags |= MCD_FLAG_LONG_PATH;
    double S;                                      // stoppingPowers.py:112-215
    if (c.bethe == MCD_BETHE_CLASSICAL) S = 2. * kPiEFour * c.atnd * ((double)c.Z / E) * log(1.166 * (E) / c.J) * 1e-14;
    else if (c.bethe == MCD_BETHE_JL)
      S = 2. * kPiEFour * c.atnd * ((double)c.Z / E) * log(1.166 * (E + c.bethe_kk * c.J) / c.J) * 1e-14;
    else {
      const double pref = (2. * kPiEFour * c.atnd / E) * 1e-14;
      double sumi = c.n_val * log(E * 1. / c.E_val);
      for (int i = 0; i < c.n_shells; ++i) sumi = sumi + c.ns[i] * log(E * 1. / c.Eb[i]);
      S = pref * sumi;
    }
    o.W = o.l * S;                                 // compute_Eloss, :815-826
    if (!(o.W < E)) o.sflags |= MCD_FLAG_LOSS_GE_E;
    const double a = c.c_alpha / E;                // compute_sAngles, :828-840
    o.c2 = 1. - (a * u1 / (1. + a - u1));
    o.hp = kPi * u2;
  }
  return o;
}

__device__ __forceinline__ void d_advance(const double* p, double l, const double* d, double* np_) {
  for (int i = 0; i < 3; ++i) np_[i] = p[i] + l * d[i];
}
#pragma clang fp contract(on)

template <int MODE>
__global__ __launch_bounds__(256) void k_transport_f64(DevCfg c, RunArgs R) {
  const int lane = threadIdx.x & 63;
  const uint32_t k0 = (uint32_t)R.seed, k1 = (uint32_t)(R.seed >> 32);
  bool active = false, open = true;
  WaveQueue wq;
  uint64_t e = 0;
  uint32_t step = 0, flags = 0;
  double E = 0., p[3] = {0., 0., 0.}, d[3] = {0., 0., 0.}, y[3] = {0., 0., 0.};
  uint32_t n_events = 0;
  while (true) {
    if (open) {
      uint64_t idle = __ballot(!active);
      if (idle) {
        const uint64_t idx = refill_index(wq, R.queue, idle, lane, R.n, open);
        if (!active) {
          if (idx < R.n) {
            active = true;
            e = R.first + idx;
            step = 0; flags = 0;
            E = c.E0;
            p[0] = p[1] = p[2] = 0.;
            d[0] = c.dir0[0]; d[1] = c.dir0[1]; d[2] = c.dir0[2];
            y[0] = 0.; y[1] = 1.; y[2] = 0.;
          }
        }
      }
    }
    if (!__any(active)) break;
    if (!active) continue;
    if ((int64_t)step >= c.max_scatt) {
      const float df[3] = {(float)d[0], (float)d[1], (float)d[2]};
      const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
      write_rec(&R.els[e - R.first], (float)E, df, pf, step, MCD_OUT_LONG, 0, flags);
      active = false;
      continue;
    }
    const U4 r = philox10((uint32_t)e, c.config_id | (uint32_t)(e >> 32) << 16, step, 0u, k0, k1);
    const double Epre = E;
    const DStep o = d_step_physics<MODE>(c, E, r);
    n_events++;
    double np_[3];
    d_advance(p, o.l, d, np_);
    int o1 = MCD_OUT_NONE, o2 = MCD_OUT_NONE;
    const double pre0 = p[0], pre1 = p[1], pre2 = p[2];
    if (np_[2] <= 0.) o1 = MCD_OUT_BKS;
    else if (np_[2] >= c.thickness) o1 = MCD_OUT_TRSM;
    else { p[0] = np_[0]; p[1] = np_[1]; p[2] = np_[2]; }
    const double En = E - o.W;
    if (En <= c.Emin) { if (o1 == MCD_OUT_NONE) o1 = MCD_OUT_ABS; else o2 = MCD_OUT_ABS; }
    E = En;
    uint32_t sflags = o.sflags;
    if (o.c2 != o.c2) sflags |= MCD_FLAG_NAN_DIR;
    flags |= sflags;
    const bool done = o1 != MCD_OUT_NONE;
    if (done) {
      const float df[3] = {(float)d[0], (float)d[1], (float)d[2]};
      const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
      write_rec(&R.els[e - R.first], (float)Epre, df, pf, step + 1, o1, o2, flags);
    }
    const bool rec_row = R.rows != nullptr && e < R.row_max_electron;
    if (!done || rec_row) d_rotate(o.c2, o.hp, d, y);
    if (R.rows != nullptr) {
      const uint64_t slot = append_slot(rec_row, R.row_cursor, R.row_cap, lane);
      if (slot != ~0ull) {
        mcd_scatter_rec* w = &R.rows[slot];
        w->electron = (uint32_t)e; w->step = step; w->config = (uint16_t)c.config_id;
        w->channel = (uint8_t)o.ch; w->outcome = (uint8_t)(o2 ? o2 : o1);
        w->E = (float)Epre; w->E_loss = (float)o.W; w->pol = (float)o.c2; w->az = (float)o.hp;
        w->pathl = (float)o.l;
        w->x = (float)pre0; w->y = (float)pre1; w->z = (float)pre2;
        w->energy = (float)E;
        w->dx = (float)d[0]; w->dy = (float)d[1]; w->dz = (float)d[2];
        w->step_flags = sflags;
      }
    }
    step++;
    if (done) active = false;
  }
  const uint64_t cnt = n_events;
  wave_add_counters(R.counters + MCD_CNT_EVENTS, &cnt, 1, lane);
}

// ---------------------------------------------------------------------------
// Outcome counts from the electron records: electrons, bks, abs (both rows
// of an exit-then-absorb electron count), trsm, long.
__global__ __launch_bounds__(256) void k_count_outcomes(const mcd_electron_rec* __restrict__ els, uint64_t n,
                                                        unsigned long long* __restrict__ counters) {
  __shared__ uint32_t part[4][5];
  count_outcome_words(els, n, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, (uint64_t)gridDim.x * blockDim.x,
                      part, counters + MCD_CNT_ELECTRONS);
}

hipError_t launch_count_outcomes(const mcd_electron_rec* els, uint64_t n, unsigned long long* counters,
                                 hipStream_t s) {
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_count_outcomes, dim3((unsigned)blocks), dim3(256), 0, s, els, n, counters);
  return hipGetLastError();
}

// Selected fields of scatter rows as packed columns: field f of row i goes to
// out + plan.dst[f] + i * plan.size[f].  One thread per row, grid-stride; a
// row is one 64-byte line, so its fields are read from one cache line and
// each column is written with consecutive addresses across the wavefront.
__global__ __launch_bounds__(256) void k_pack_scatter(const uint8_t* __restrict__ rows, uint64_t n, ScatPackPlan plan,
                                                      uint8_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint8_t* r = rows + i * sizeof(mcd_scatter_rec);
    for (int f = 0; f < plan.n; ++f) {
      const int sz = plan.size[f];
      const uint8_t* src = r + plan.src[f];
      uint8_t* dst = out + plan.dst[f] + i * (uint64_t)sz;
      if (sz == 4) *reinterpret_cast<uint32_t*>(dst) = *reinterpret_cast<const uint32_t*>(src);
      else if (sz == 2) *reinterpret_cast<uint16_t*>(dst) = *reinterpret_cast<const uint16_t*>(src);
      else *dst = *src;
    }
  }
}

hipError_t launch_pack_scatter(const mcd_scatter_rec* rows, uint64_t n, const ScatPackPlan& plan, void* out,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_pack_scatter, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<const uint8_t*>(rows), n, plan, reinterpret_cast<uint8_t*>(out));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host-side launchers

template <int MODE, int NCH, int MINW, bool EXPL>
static hipError_t launch_f32(const DevCfg& c, const RunArgs& R, int grid, int block, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((k_transport_f32<MODE, NCH, MINW, 0, EXPL>), dim3(grid), dim3(block), lds, s, c, R);
  return hipGetLastError();
}

template <int MODE, int NCH, int MINW, bool EXPL>
static int occ_f32(int block, size_t lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_f32<MODE, NCH, MINW, 0, EXPL>, block, lds) !=
      hipSuccess)
    return 1;
  return nb < 1 ? 1 : nb;
}

// dispatch over (channel count, occupancy variant)
template <int MODE, int NCH, bool EXPL = false>
static hipError_t dispatch_w(int minw, const DevCfg& c, const RunArgs* R, int grid, int block, size_t lds,
                             hipStream_t s, int* occ) {
  switch (minw) {
#define MCD_W(W)                                                   \
  case W:                                                          \
    if (occ) { *occ = occ_f32<MODE, NCH, W, EXPL>(block, lds); return hipSuccess; } \
    return launch_f32<MODE, NCH, W, EXPL>(c, *R, grid, block, lds, s);
    MCD_W(1) MCD_W(4) MCD_W(5) MCD_W(6) MCD_W(8)
#undef MCD_W
    default: return hipErrorInvalidValue;
  }
}

static hipError_t dispatch_f32(int minw, const DevCfg& c, const RunArgs* R, int grid, int block, hipStream_t s,
                               int* occ) {
  if (c.mode == MCD_MODE_CONT) {
    if (c.bethe == MCD_BETHE_EXPLICIT) return dispatch_w<MCD_MODE_CONT, 4, true>(minw, c, R, grid, block, 0, s, occ);
    return dispatch_w<MCD_MODE_CONT, 4>(minw, c, R, grid, block, 0, s, occ);
  }
  const size_t lds = single_lds_bytes(c);
  switch (c.n_shells + 3) {
    case 4: return dispatch_w<MCD_MODE_DS, 4>(minw, c, R, grid, block, lds, s, occ);
    case 5: return dispatch_w<MCD_MODE_DS, 5>(minw, c, R, grid, block, lds, s, occ);
    case 6: return dispatch_w<MCD_MODE_DS, 6>(minw, c, R, grid, block, lds, s, occ);
    case 7: return dispatch_w<MCD_MODE_DS, 7>(minw, c, R, grid, block, lds, s, occ);
    case 8: return dispatch_w<MCD_MODE_DS, 8>(minw, c, R, grid, block, lds, s, occ);
    case 9: return dispatch_w<MCD_MODE_DS, 9>(minw, c, R, grid, block, lds, s, occ);
    case 10: return dispatch_w<MCD_MODE_DS, 10>(minw, c, R, grid, block, lds, s, occ);
    case 11: return dispatch_w<MCD_MODE_DS, 11>(minw, c, R, grid, block, lds, s, occ);
    default: return hipErrorInvalidValue;
  }
}

bool valid_minw(int w) { return w == 1 || w == 4 || w == 5 || w == 6 || w == 8; }

hipError_t launch_transport(const DevCfg& c, const RunArgs& R, int precision, int minw, int grid, int block,
                            hipStream_t s) {
  if (precision == MCD_PREC_FP64) {
    if (c.mode == MCD_MODE_DS)
      hipLaunchKernelGGL((k_transport_f64<MCD_MODE_DS>), dim3(grid), dim3(block), 0, s, c, R);
    else
      hipLaunchKernelGGL((k_transport_f64<MCD_MODE_CONT>), dim3(grid), dim3(block), 0, s, c, R);
    return hipGetLastError();
  }
  return dispatch_f32(minw, c, &R, grid, block, s, nullptr);
}

// resident blocks per CU for the persistent grid
int max_blocks_per_cu(const DevCfg& c, int precision, int minw, int block) {
  int nb = 1;
  if (precision == MCD_PREC_FP64) {
    hipError_t e = c.mode == MCD_MODE_DS
                       ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_f64<MCD_MODE_DS>, block, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_f64<MCD_MODE_CONT>, block, 0);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb;
  }
  if (dispatch_f32(minw, c, nullptr, 0, block, 0, &nb) != hipSuccess || nb < 1) nb = 1;
  return nb;
}

// ---------------------------------------------------------------------------
// Deferred guard (DS, FP32): phase 1 without FP64 code, phase 2 resumes.
template <int NCH, int PHASE>
static hipError_t deferred_nch(const DevCfg& c, const RunArgs* R, int grid, int block, size_t lds, hipStream_t s,
                               int* occ) {
  if (occ) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_f32<MCD_MODE_DS, NCH, 1, PHASE>, block, lds) !=
        hipSuccess)
      nb = 1;
    *occ = nb < 1 ? 1 : nb;
    return hipSuccess;
  }
  hipLaunchKernelGGL((k_transport_f32<MCD_MODE_DS, NCH, 1, PHASE>), dim3(grid), dim3(block), lds, s, c, *R);
  return hipGetLastError();
}

template <int PHASE>
static hipError_t deferred_dispatch(const DevCfg& c, const RunArgs* R, int grid, int block, hipStream_t s, int* occ) {
  const size_t lds = single_lds_bytes(c);
  switch (c.n_shells + 3) {
    case 4: return deferred_nch<4, PHASE>(c, R, grid, block, lds, s, occ);
    case 5: return deferred_nch<5, PHASE>(c, R, grid, block, lds, s, occ);
    case 6: return deferred_nch<6, PHASE>(c, R, grid, block, lds, s, occ);
    case 7: return deferred_nch<7, PHASE>(c, R, grid, block, lds, s, occ);
    case 8: return deferred_nch<8, PHASE>(c, R, grid, block, lds, s, occ);
    case 9: return deferred_nch<9, PHASE>(c, R, grid, block, lds, s, occ);
    case 10: return deferred_nch<10, PHASE>(c, R, grid, block, lds, s, occ);
    case 11: return deferred_nch<11, PHASE>(c, R, grid, block, lds, s, occ);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_transport_deferred(const DevCfg& c, const RunArgs& R, int phase, int grid, int block,
                                     hipStream_t s) {
  if (c.mode != MCD_MODE_DS) return hipErrorInvalidValue;
  return phase == 1 ? deferred_dispatch<1>(c, &R, grid, block, s, nullptr)
                    : deferred_dispatch<2>(c, &R, grid, block, s, nullptr);
}

// ---------------------------------------------------------------------------
// Service guard (DS, FP32): one kernel, wavefront 0 of each workgroup settles
// the guard-band picks of the others.
template <int NCH, int MINW>
static hipError_t svc_nch(const DevCfg& c, const RunArgs* R, int grid, int block, size_t lds, hipStream_t s,
                          int* occ) {
  if (occ) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_transport_svc<NCH, MINW>, block, lds) != hipSuccess)
      nb = 1;
    *occ = nb < 1 ? 1 : nb;
    return hipSuccess;
  }
  hipLaunchKernelGGL((k_transport_svc<NCH, MINW>), dim3(grid), dim3(block), lds, s, c, *R);
  return hipGetLastError();
}

template <int MINW>
static hipError_t svc_dispatch(const DevCfg& c, const RunArgs* R, int grid, int block, hipStream_t s, int* occ) {
  const size_t lds = svc_lds_bytes(c, block);
  switch (c.n_shells + 3) {
    case 4: return svc_nch<4, MINW>(c, R, grid, block, lds, s, occ);
    case 5: return svc_nch<5, MINW>(c, R, grid, block, lds, s, occ);
    case 6: return svc_nch<6, MINW>(c, R, grid, block, lds, s, occ);
    case 7: return svc_nch<7, MINW>(c, R, grid, block, lds, s, occ);
    case 8: return svc_nch<8, MINW>(c, R, grid, block, lds, s, occ);
    case 9: return svc_nch<9, MINW>(c, R, grid, block, lds, s, occ);
    case 10: return svc_nch<10, MINW>(c, R, grid, block, lds, s, occ);
    case 11: return svc_nch<11, MINW>(c, R, grid, block, lds, s, occ);
    default: return hipErrorInvalidValue;
  }
}

bool valid_svc_minw(int w) { return w == 6 || w == 7 || w == 8; }

static hipError_t svc_any(int minw, const DevCfg& c, const RunArgs* R, int grid, int block, hipStream_t s, int* occ) {
  switch (minw) {
    case 6: return svc_dispatch<6>(c, R, grid, block, s, occ);
    case 7: return svc_dispatch<7>(c, R, grid, block, s, occ);
    case 8: return svc_dispatch<8>(c, R, grid, block, s, occ);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_transport_svc(const DevCfg& c, const RunArgs& R, int minw, int grid, int block, hipStream_t s) {
  if (c.mode != MCD_MODE_DS || block < 128 || block > 1024 || block % 64) return hipErrorInvalidValue;
  return svc_any(minw, c, &R, grid, block, s, nullptr);
}

int svc_blocks_per_cu(const DevCfg& c, int minw, int block) {
  int nb = 1;
  hipError_t e = svc_any(minw, c, nullptr, 0, block, 0, &nb);
  return e == hipSuccess && nb > 0 ? nb : 1;
}

int deferred_blocks_per_cu(const DevCfg& c, int phase, int block) {
  int nb = 1;
  hipError_t e = phase == 1 ? deferred_dispatch<1>(c, nullptr, 0, block, 0, &nb)
                            : deferred_dispatch<2>(c, nullptr, 0, block, 0, &nb);
  return e == hipSuccess && nb > 0 ? nb : 1;
}

}  // namespace mcd
