// JAX's default random streams on the device (jax 0.4.23: threefry2x32, original split and bit generation),
// so that the sampler can replay the reference's draws instead of taking supplied ones.
//
//   threefry2x32: Threefry-2x32 with 20 rounds (Salmon et al., SC 2011; the Random123 reference).
//   split(k, num): counts 0 .. 2 num - 1 hashed as pairs (i, num + i); outputs concatenated -> [num, 2].
//   bits(k, m):    counts 0 .. m - 1 (one zero appended when m is odd) hashed as pairs (i, h + i), h = ceil(m/2);
//                  outputs concatenated, the padding word dropped.
//   uniform:       (bits >> 9 | 0x3f800000) as float - 1, scaled to [lo, hi), max with lo.
//   normal:        sqrt(2) erfinv(uniform on [nextafter(-1, 0), 1)), erfinv in single precision after
//                  M. Giles, "Approximating the erfinv function" (GPU Computing Gems, 2011).
// The sampler's per-walker stream (mcmc.py:175-180, 345-378): at each step the walker key k gives the
// subkey s = split(k)[1] and the next key split(k)[0]; the proposal is normal(s, (n, 3)) and the accept
// threshold uniform(s, ()) -- the same subkey for both, as in the reference.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace dwmi {

namespace {

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __forceinline__ void threefry2x32(uint32_t k0, uint32_t k1, uint32_t x0, uint32_t x1, uint32_t& y0,
                                             uint32_t& y1) {
  const uint32_t ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  constexpr int rot[2][4] = {{13, 15, 26, 6}, {17, 29, 16, 24}};
  x0 += ks[0];
  x1 += ks[1];
#pragma unroll
  for (int blk = 0; blk < 5; ++blk) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      x0 += x1;
      x1 = rotl32(x1, rot[blk & 1][q]);
      x1 ^= x0;
    }
    x0 += ks[(blk + 1) % 3];
    x1 += ks[(blk + 2) % 3] + (uint32_t)(blk + 1);
  }
  y0 = x0;
  y1 = x1;
}

__device__ __forceinline__ float unit_float(uint32_t bits) { return __uint_as_float((bits >> 9) | 0x3F800000u) - 1.0f; }

__device__ __forceinline__ float erfinv32(float x) {
  const float lt5[9] = {2.81022636e-08f, 3.43273939e-07f, -3.5233877e-06f, -4.39150654e-06f, 0.00021858087f,
                        -0.00125372503f, -0.00417768164f, 0.246640727f, 1.50140941f};
  const float ge5[9] = {-0.000200214257f, 0.000100950558f, 0.00134934322f, -0.00367342844f, 0.00573950773f,
                        -0.0076224613f, 0.00943887047f, 1.00167406f, 2.83297682f};
  float w = -log1pf(-x * x);
  const bool lt = w < 5.0f;
  w = lt ? w - 2.5f : sqrtf(w) - 3.0f;
  float p = lt ? lt5[0] : ge5[0];
#pragma unroll
  for (int i = 1; i < 9; ++i) p = (lt ? lt5[i] : ge5[i]) + p * w;
  return fabsf(x) == 1.0f ? x * 3.402823466e+38f : p * x;
}

// normal(key) value from one 32-bit word
__device__ __forceinline__ float normal_from_bits(uint32_t bits) {
  const float lo = -0.99999994f;  // nextafter(-1, 0) in float32
  const float u = fmaxf(lo, unit_float(bits) * 2.0f + lo);  // (hi - lo) rounds to 2 in float32
  return 1.41421354f * erfinv32(u);
}

__device__ __forceinline__ void split2(uint32_t k0, uint32_t k1, uint32_t& n0, uint32_t& n1, uint32_t& s0,
                                       uint32_t& s1) {
  // split(k) = [[T(0, 2).y0, T(1, 3).y0], [T(0, 2).y1, T(1, 3).y1]]
  uint32_t a0, a1, b0, b1;
  threefry2x32(k0, k1, 0u, 2u, a0, a1);
  threefry2x32(k0, k1, 1u, 3u, b0, b1);
  n0 = a0;
  n1 = b0;
  s0 = a1;
  s1 = b1;
}

__global__ void k_threefry_split(const uint32_t* __restrict__ key, int num, uint32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num) return;
  uint32_t y0, y1;
  threefry2x32(key[0], key[1], (uint32_t)i, (uint32_t)(num + i), y0, y1);
  out[i] = y0;
  out[num + i] = y1;
}

__global__ void k_threefry_normal(const uint32_t* __restrict__ key, long count, float* __restrict__ out) {
  const long half = (count + 1) / 2;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half) return;
  const uint32_t x1 = (half + i < count) ? (uint32_t)(half + i) : 0u;
  uint32_t y0, y1;
  threefry2x32(key[0], key[1], (uint32_t)i, x1, y0, y1);
  out[i] = normal_from_bits(y0);
  if (half + i < count) out[half + i] = normal_from_bits(y1);
}

// One thread per (walker b, counter pair i of the 3 n proposal words): walks the walker's key chain for
// n_steps steps and writes its pair of normals per step; thread i = 0 also writes the accept uniform and
// the final key.
__global__ void k_threefry_mcmc(const uint32_t* __restrict__ keys_in, uint32_t* __restrict__ keys_out, int B, int n,
                                int n_steps, float* __restrict__ normals, float* __restrict__ uniforms) {
  const int m = 3 * n, half = (m + 1) / 2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * half) return;
  const long b = t / half;
  const int i = (int)(t - b * half);
  uint32_t k0 = keys_in[2 * b], k1 = keys_in[2 * b + 1];
  const uint32_t x1 = (half + i < m) ? (uint32_t)(half + i) : 0u;
  for (int s = 0; s < n_steps; ++s) {
    uint32_t n0, n1, s0, s1;
    split2(k0, k1, n0, n1, s0, s1);
    uint32_t y0, y1;
    threefry2x32(s0, s1, (uint32_t)i, x1, y0, y1);
    float* out = normals + ((long)s * B + b) * m;
    out[i] = normal_from_bits(y0);
    if (half + i < m) out[half + i] = normal_from_bits(y1);
    if (i == 0) {
      uint32_t u0, u1;
      threefry2x32(s0, s1, 0u, 0u, u0, u1);  // one word: counts [0] padded to [0, 0]
      uniforms[(long)s * B + b] = fmaxf(0.0f, unit_float(u0));
    }
    k0 = n0;
    k1 = n1;
  }
  if (i == 0) {
    keys_out[2 * b] = k0;
    keys_out[2 * b + 1] = k1;
  }
}

}  // namespace

void launch_threefry_split(const uint32_t* key, int num, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_threefry_split, dim3((num + 255) / 256), dim3(256), 0, s, key, num, out);
}

void launch_threefry_normal(const uint32_t* key, long count, float* out, hipStream_t s) {
  const long half = (count + 1) / 2;
  hipLaunchKernelGGL(k_threefry_normal, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, s, key, count, out);
}

void launch_threefry_mcmc(const uint32_t* keys_in, uint32_t* keys_out, int B, int n, int n_steps, float* normals,
                          float* uniforms, hipStream_t s) {
  const long total = (long)B * ((3 * n + 1) / 2);
  hipLaunchKernelGGL(k_threefry_mcmc, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, keys_in, keys_out, B,
                     n, n_steps, normals, uniforms);
}

}  // namespace dwmi
