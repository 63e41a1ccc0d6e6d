// Position-dependent Metropolis-Hastings proposals of the reference sampler (mcmc.py:204-215, 218-236, 255-284).
// Per electron i the step scale is s_i = stepsize * clip(distance to the closest nucleus, r_min, r_max) and
//   local:    r'_i = r_i + s_i xi_i,
//             log q = sum_i 3 (log s_i - log s'_i) + |r'_i - r_i|^2 (1 / s_i^2 - 1 / s'_i^2) / 2;
//   langevin: r'_i = r_i + s_i xi_i + g_i s_i^2, with the drift g_i = -scale sum_I (r_i - R_I) Z_I / |r_i - R_I|
//             (towards the nuclei; it does not use the wavefunction),
//             log q = sum_i 3 (log s_i - log s'_i) + (|r'_i - r_i - g_i s_i^2|^2 / s_i^2
//                                                     - |r_i - r'_i - g'_i s'_i^2|^2 / s'_i^2) / 2,
// with s', g' evaluated at r'.  log q is the log ratio of the reverse to the forward proposal density and
// enters the accept test as exp(lp' - lp + log q).  FP32, in the reference's operation order.  One
// wavefront per walker, lane = electron; the per-walker sum over electrons is a fixed butterfly.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.h"
#include "dwmi.h"
#include "kernels.h"

namespace dwmi {

namespace {

// step scale s and drift g at the electron position x (drift only for the Langevin form)
template <bool DRIFT>
__device__ __forceinline__ void scale_and_drift(const float x[3], const float* __restrict__ R,
                                                const float* __restrict__ Z, int I, float stepsize, float r_min,
                                                float r_max, float lscale, float& s, float g[3]) {
  float dmin = INFINITY;
  g[0] = g[1] = g[2] = 0.f;
  for (int k = 0; k < I; ++k) {
    const float dx = x[0] - R[3 * k], dy = x[1] - R[3 * k + 1], dz = x[2] - R[3 * k + 2];
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    dmin = fminf(dmin, d);
    if (DRIFT) {
      const float z = Z[k];
      g[0] += dx * z / d;
      g[1] += dy * z / d;
      g[2] += dz * z / d;
    }
  }
  s = stepsize * fminf(fmaxf(dmin, r_min), r_max);
  if (DRIFT) {
    g[0] *= -lscale;
    g[1] *= -lscale;
    g[2] *= -lscale;
  }
}

template <bool LANGEVIN>
__global__ __launch_bounds__(64) void k_propose_scaled(const float* __restrict__ r, const float* __restrict__ normals,
                                                       const dw_chain_scalars* __restrict__ chain,
                                                       const float* __restrict__ R, const float* __restrict__ Z,
                                                       int I, float r_min, float r_max, float lscale,
                                                       float* __restrict__ r_prop, float* __restrict__ logq, int n) {
  const long b = blockIdx.x;
  const int i = threadIdx.x;
  const float step = chain->stepsize;
  float lq = 0.f;
  if (i < n) {
    const long o = (b * n + i) * 3;
    const float x[3] = {r[o], r[o + 1], r[o + 2]};
    float s, g[3];
    scale_and_drift<LANGEVIN>(x, R, Z, I, step, r_min, r_max, lscale, s, g);
    const float s2 = s * s;
    float xn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xn[c] = x[c] + normals[o + c] * s;
      if (LANGEVIN) xn[c] += g[c] * s2;
    }
    float sn, gn[3];
    scale_and_drift<LANGEVIN>(xn, R, Z, I, step, r_min, r_max, lscale, sn, gn);
    const float sn2 = sn * sn;
#pragma unroll
    for (int c = 0; c < 3; ++c) r_prop[o + c] = xn[c];
    lq = 3.f * (logf(s) - logf(sn));
    if (LANGEVIN) {
      float d_fwd = 0.f, d_rev = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float f = xn[c] - x[c] - g[c] * s2;
        const float q = x[c] - xn[c] - gn[c] * sn2;
        d_fwd += f * f;
        d_rev += q * q;
      }
      lq += 0.5f * (d_fwd / s2 - d_rev / sn2);
    } else {
      float d2 = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) d2 += (xn[c] - x[c]) * (xn[c] - x[c]);
      lq += 0.5f * d2 * (1.f / s2 - 1.f / sn2);
    }
  }
  lq = wave_sum(lq);
  if (i == 0) logq[b] = lq;
}

// One-electron moves (mcmc.py:183-192, 239-252): only electron k = step_nr mod n moves, by its 3 draws
// xi [B, 3]; local form: s = stepsize * clip(distance of electron k to the closest nucleus), r'_k = r_k +
// s xi and log q = 3 (log s - log s') + |r'_k - r_k|^2 (1/s^2 - 1/s'^2) / 2 of that electron alone (the
// reference does not sum over electrons here).  One thread per (walker, electron): every thread copies its
// electron, the thread of electron k moves it.
template <bool LOCAL>
__global__ __launch_bounds__(256) void k_propose_one_el(const float* __restrict__ r, const float* __restrict__ xi,
                                                        const dw_chain_scalars* __restrict__ chain,
                                                        const float* __restrict__ R, int I, float r_min, float r_max,
                                                        float* __restrict__ r_prop, float* __restrict__ logq, int B,
                                                        int n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * n) return;
  const long b = t / n;
  const int i = (int)(t - b * n);
  const int k = (int)(((long)chain->step_nr % n + n) % n);
  const float step = chain->stepsize;
  const long o = t * 3;
  float x[3] = {r[o], r[o + 1], r[o + 2]};
  if (i != k) {
    r_prop[o] = x[0];
    r_prop[o + 1] = x[1];
    r_prop[o + 2] = x[2];
    return;
  }
  const float* d = xi + b * 3;
  if (!LOCAL) {
    for (int c = 0; c < 3; ++c) r_prop[o + c] = x[c] + d[c] * step;
    return;
  }
  float s, g[3];
  scale_and_drift<false>(x, R, nullptr, I, step, r_min, r_max, 0.f, s, g);
  float xn[3];
  for (int c = 0; c < 3; ++c) xn[c] = x[c] + d[c] * s;
  float sn;
  scale_and_drift<false>(xn, R, nullptr, I, step, r_min, r_max, 0.f, sn, g);
  for (int c = 0; c < 3; ++c) r_prop[o + c] = xn[c];
  float d2 = 0.f;
  for (int c = 0; c < 3; ++c) d2 += (xn[c] - x[c]) * (xn[c] - x[c]);
  logq[b] = 3.f * (logf(s) - logf(sn)) + 0.5f * d2 * (1.f / (s * s) - 1.f / (sn * sn));
}

}  // namespace

void launch_propose_one_el(int kind, const float* r, const float* xi, const void* chain, const float* R, int I,
                           float r_min, float r_max, float* r_prop, float* logq, int B, int n, hipStream_t s) {
  const dw_chain_scalars* c = (const dw_chain_scalars*)chain;
  const unsigned grid = (unsigned)(((long)B * n + 255) / 256);
  if (kind == DW_PROPOSAL_LOCAL_ONE_EL)
    hipLaunchKernelGGL(k_propose_one_el<true>, dim3(grid), dim3(256), 0, s, r, xi, c, R, I, r_min, r_max, r_prop, logq,
                       B, n);
  else
    hipLaunchKernelGGL(k_propose_one_el<false>, dim3(grid), dim3(256), 0, s, r, xi, c, R, I, r_min, r_max, r_prop,
                       logq, B, n);
}

void launch_propose_scaled(int kind, const float* r, const float* normals, const void* chain, const float* R,
                           const float* Z, int I, float r_min, float r_max, float lscale, float* r_prop, float* logq,
                           int B, int n, hipStream_t s) {
  const dw_chain_scalars* c = (const dw_chain_scalars*)chain;
  if (kind == DW_PROPOSAL_LANGEVIN)
    hipLaunchKernelGGL(k_propose_scaled<true>, dim3(B), dim3(64), 0, s, r, normals, c, R, Z, I, r_min, r_max, lscale,
                       r_prop, logq, n);
  else
    hipLaunchKernelGGL(k_propose_scaled<false>, dim3(B), dim3(64), 0, s, r, normals, c, R, Z, I, r_min, r_max,
                       lscale, r_prop, logq, n);
}

}  // namespace dwmi
