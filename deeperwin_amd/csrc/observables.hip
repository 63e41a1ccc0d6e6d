// Energy observables of an evaluation step: E_mean = nanmean(E_loc), E_var = nanmean(|E_loc - E_mean|^2)
// (optimization/evaluation.py:143-151; one device, so pmean is the identity).  NaN walkers are dropped,
// +-inf are kept (jnp.nanmean semantics), so one infinite walker makes the mean infinite.
//
// One 1024-thread workgroup, FP64 throughout: each thread sums a strided slice, then a fixed-order tree in
// LDS combines the partials, so the result is deterministic (independent of timing) for a given n.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.h"

namespace dwmi {

namespace {

constexpr int kThreads = 1024;

__device__ void tree_sum2(double* a, double* c, int tid) {
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    __syncthreads();
    if (tid < w) {
      a[tid] += a[tid + w];
      c[tid] += c[tid + w];
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_energy_moments(const float* __restrict__ e, int n,
                                                             double* __restrict__ out) {
  __shared__ double s_sum[kThreads];
  __shared__ double s_cnt[kThreads];
  const int tid = threadIdx.x;
  double acc = 0.0, cnt = 0.0;
  for (int i = tid; i < n; i += kThreads) {
    const double x = (double)e[i];
    if (!isnan(x)) {
      acc += x;
      cnt += 1.0;
    }
  }
  s_sum[tid] = acc;
  s_cnt[tid] = cnt;
  tree_sum2(s_sum, s_cnt, tid);
  const double mean = s_sum[0] / s_cnt[0];  // 0 / 0 = NaN when every walker is NaN, like nanmean
  __syncthreads();
  acc = 0.0;
  cnt = 0.0;
  for (int i = tid; i < n; i += kThreads) {
    const double d = (double)e[i] - mean;
    const double q = d * d;
    if (!isnan(q)) {  // inf - inf is NaN and is dropped, as jnp.nanmean drops it
      acc += q;
      cnt += 1.0;
    }
  }
  s_sum[tid] = acc;
  s_cnt[tid] = cnt;
  tree_sum2(s_sum, s_cnt, tid);
  if (tid == 0) {
    out[0] = mean;
    out[1] = s_sum[0] / s_cnt[0];
  }
}

}  // namespace

void launch_energy_moments(const float* e, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_moments, dim3(1), dim3(kThreads), 0, s, e, n, out);
}

}  // namespace dwmi
