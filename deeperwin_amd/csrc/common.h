// Shared definitions for the dwmi HIP sources (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DW_WAVE 64

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));

static inline int dw_pad(int x, int m) { return (x + m - 1) / m * m; }

// Wave-level reductions (wave64).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// FP32 tanh from the identity tanh x = 1 - 2 / (1 + e^{2x}) with the hardware base-2 exponential and
// reciprocal (5 instructions).  Saturates correctly (e^{2x} -> inf gives 1, -> 0 gives -1).  Absolute
// error ~1e-7 near 0 and a few ulp elsewhere: below the rounding of the FP32 pre-activation it is
// applied to.
__device__ __forceinline__ float tanh_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);  // 2 / ln 2
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + e);
}

// Index division: 32-bit unsigned arithmetic whenever the dividend fits (the common case), 64-bit otherwise.
__device__ __forceinline__ long div_idx(long t, int d) {
  return ((unsigned long)t >> 32) == 0 ? (long)((unsigned)t / (unsigned)d) : t / d;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Wavefront sums of G = 1, 2 or 4 values at once.  The first log2(G) exchange steps (xor 32, then 16) trade
// the values a lane is not responsible for, so every later step moves one value; lane l ends with the
// total of v[l / (64 / G)].  Each total is paired in the same order as wave_sum_d (xor 32, 16, ..., 1).
template <int G>
__device__ __forceinline__ double wave_sum_group(const double (&v)[G], int lane) {
  static_assert(G == 1 || G == 2 || G == 4, "groups of 1, 2 or 4");
  double s;
  if constexpr (G == 1) {
    return wave_sum_d(v[0]);
  } else if constexpr (G == 4) {
    const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0;
    const double a = (h5 ? v[2] : v[0]) + __shfl_xor(h5 ? v[0] : v[2], 32, 64);
    const double b = (h5 ? v[3] : v[1]) + __shfl_xor(h5 ? v[1] : v[3], 32, 64);
    s = (h4 ? b : a) + __shfl_xor(h4 ? a : b, 16, 64);
  } else {
    const bool h5 = (lane & 32) != 0;
    s = (h5 ? v[1] : v[0]) + __shfl_xor(h5 ? v[0] : v[1], 32, 64);
  }
#pragma unroll
  for (int o = (G == 4 ? 8 : 16); o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// Block sum for blockDim.x a multiple of 64 (<= 1024); `scratch` holds >= 16 floats.  All threads get
// the result.
__device__ __forceinline__ float block_sum(float v, float* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[w] = v;
  __syncthreads();
  float t = 0.f;
  for (int i = 0; i < nw; ++i) t += scratch[i];
  return t;
}
__device__ __forceinline__ double block_sum_d(double v, double* scratch) {
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += scratch[i];
  return t;
}

// Round-to-nearest-even float -> bf16 bits, and back.
__device__ __forceinline__ unsigned short f2bf_bits(float f) {
  unsigned int u = __float_as_uint(f);
  unsigned int r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(r >> 16);
}
__device__ __forceinline__ float bf_bits2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

// Per-walker flag bits (same values as DW_FLAG_* in include/dwmi.h).
#define DWF_NONFINITE 1u
#define DWF_FP64 2u
#define DWF_FORCED 4u
