// mcd_calib.hip — issue-rate calibration kernels for the roofline of the
// transport (libmcd_calib.so, loaded only by bench.py --valu-ceiling).
//
// Each kernel runs a known number of VALU instructions per wavefront in
// independent chains (enough independent work per wavefront that the issue
// rate, not the dependency latency, is measured at full occupancy).  Kinds:
//   0: v_fma_f32 only      (FP32 FMA, not packed: packed FP32 is disabled)
//   1: v_mad_u64_u32 only  (the 32 x 32 -> 64 multiply-add of Philox rounds)
//   2: v_bitop3_b32 only   (the three-input XOR of Philox rounds)
//   3: v_fma_f32 + v_exp_f32, 1:1   (an FMA beside a transcendental)
//   4: v_fma_f32 + v_mad_u64_u32, 1:1
//   5: v_fma_f32 + v_bitop3_b32, 1:1
//   6: v_fma_f32 + v_alignbit_b32 (a 32-bit integer rotate), 1:1
//   7: v_rcp_f32 only      (a transcendental alone)
// Every unrolled step issues 8 instructions of each listed kind; the result of
// every chain is stored, so no chain is removed.  Timing a kernel gives the
// wavefront-instructions per second the SIMDs issue for that mix; the
// SQ_INSTS_VALU and SQ_ACTIVE_INST_VALU2 (quad-cycles with two VALU
// instructions issued) counters of the same dispatch show how many issue
// slots the mix used.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kUnroll = 16;

struct OpFma {
  using T = float;
  __device__ static T init(int t, int j) { return t * 1e-7f + j; }
  __device__ static T step(T x, int u) { return __builtin_fmaf(x, 0.999999f, 1e-6f); }
  __device__ static uint32_t bits(T x) { return __float_as_uint(x); }
};
struct OpExp {
  using T = float;
  __device__ static T init(int t, int j) { return -1.f - t * 1e-7f - j; }
  __device__ static T step(T x, int u) { return __builtin_amdgcn_exp2f(x); }
  __device__ static uint32_t bits(T x) { return __float_as_uint(x); }
};
struct OpRcp {
  using T = float;
  __device__ static T init(int t, int j) { return 1.5f + t * 1e-7f + j; }
  __device__ static T step(T x, int u) { return __builtin_amdgcn_rcpf(x); }
  __device__ static uint32_t bits(T x) { return __float_as_uint(x); }
};
struct OpMad {
  using T = uint64_t;
  __device__ static T init(int t, int j) { return (uint64_t)t + 977u * j; }
  __device__ static T step(T x, int u) { return (uint64_t)0xD2511F53u * (uint32_t)x + x; }
  __device__ static uint32_t bits(T x) { return (uint32_t)x ^ (uint32_t)(x >> 32); }
};
struct OpXor3 {
  using T = uint32_t;
  __device__ static T init(int t, int j) { return t * 2654435761u + j; }
  __device__ static T step(T x, int u) { return __builtin_amdgcn_bitop3_b32(x, 0x9E3779B9u, 0xBB67AE85u + u, 0x96); }
  __device__ static uint32_t bits(T x) { return x; }
};
struct OpRot {  // funnel shift of a chain with its neighbour (a rotate of itself would fold)
  using T = uint32_t;
  __device__ static T init(int t, int j) { return t + 31u * j; }
  __device__ static T step2(T x, T y) { return __builtin_amdgcn_alignbit(x, y, 7u); }
  __device__ static uint32_t bits(T x) { return x; }
};
struct OpNone {
  using T = uint32_t;
  __device__ static T init(int, int) { return 0u; }
  __device__ static T step(T x, int) { return x; }
  __device__ static uint32_t bits(T x) { return x; }
};

template <class B> constexpr bool kCoupled = false;
template <> constexpr bool kCoupled<OpRot> = true;

// 8 chains of A and (unless B is OpNone) 8 chains of B per unrolled step.
template <class A, class B>
__global__ __launch_bounds__(256) void k_mix(uint32_t* out, int iters) {
  typename A::T a[8];
  typename B::T b[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    a[j] = A::init(threadIdx.x, j);
    b[j] = B::init(threadIdx.x, j);
  }
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        a[j] = A::step(a[j], u);
        if constexpr (kCoupled<B>) b[j] = B::step2(b[j], b[(j + 1) & 7]);
        else b[j] = B::step(b[j], u);
      }
    }
  }
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s ^= A::bits(a[j]) ^ B::bits(b[j]);
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// VALU instructions of the measured mix per wavefront for `iters`.
long long mcd_calib_insts_per_wave(int kind, int iters) {
  const long long one = (long long)iters * 8 * kUnroll;
  return (kind >= 3 && kind <= 6) ? 2 * one : one;
}

// Launch `blocks` workgroups of 256 threads of calibration kernel `kind` on
// `stream`; `out` must hold blocks * 256 * 4 bytes (device).  Returns a
// hipError_t value (0 = launched).
int mcd_calib_launch(int kind, int blocks, int iters, void* out, void* stream) {
  if (blocks < 1 || iters < 1 || !out) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  uint32_t* o = (uint32_t*)out;
  switch (kind) {
    case 0: hipLaunchKernelGGL((k_mix<OpFma, OpNone>), dim3(blocks), dim3(256), 0, s, o, iters); break;
    case 1: hipLaunchKernelGGL((k_mix<OpMad, OpNone>), dim3(blocks), dim3(256), 0, s, o, iters); break;
    case 2: hipLaunchKernelGGL((k_mix<OpXor3, OpNone>), dim3(blocks), dim3(256), 0, s, o, iters); break;
    case 3: hipLaunchKernelGGL((k_mix<OpFma, OpExp>), dim3(blocks), dim3(256), 0, s, o, iters); break;
    case 4: hipLaunchKernelGGL((k_mix<OpFma, OpMad>), dim3(blocks), dim3(256), 0, s, o, iters); break;
    case 5: hipLaunchKernelGGL((k_mix<OpFma, OpXor3>), dim3(blocks), dim3(256), 0, s, o, iters); break;
    case 6: hipLaunchKernelGGL((k_mix<OpFma, OpRot>), dim3(blocks), dim3(256), 0, s, o, iters); break;
    case 7: hipLaunchKernelGGL((k_mix<OpRcp, OpNone>), dim3(blocks), dim3(256), 0, s, o, iters); break;
    default: return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

}  // extern "C"
