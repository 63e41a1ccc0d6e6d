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

// ---- clipped-energy loss (optimization/loss_function.py:12-110) ----
// One 1024-thread workgroup.  NaN walkers are dropped from every mean (nanmean / nanmedian).  The median
// sorts the finite-or-inf values of E in LDS (bitonic, at most kClipMax of them).
constexpr int kClipMax = 8192;

__device__ double block_sum_fixed(double v, double* sh, int tid) {
  sh[tid] = v;
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    __syncthreads();
    if (tid < w) sh[tid] += sh[tid + w];
  }
  __syncthreads();
  const double r = sh[0];
  __syncthreads();
  return r;
}

// nanmean over i of f(i), skipping the i for which f(i) is NaN
template <typename F>
__device__ double nanmean_of(int n, F f, double* s1, double* s2, int tid) {
  double a = 0.0, c = 0.0;
  for (int i = tid; i < n; i += kThreads) {
    const double v = f(i);
    if (!isnan(v)) {
      a += v;
      c += 1.0;
    }
  }
  const double A = block_sum_fixed(a, s1, tid);
  const double C = block_sum_fixed(c, s2, tid);
  return A / C;
}

// nanmedian of f(i): the non-NaN values sorted in LDS (bitonic network, NaN entries as +inf at the end and
// excluded from the count); the mean of the two middle values for an even count.
template <typename F>
__device__ double nanmedian_of(int n, F f, double* v, double* s1, int tid) {
  int P = 1;
  while (P < n) P <<= 1;
  double c = 0.0;
  for (int i = tid; i < P; i += kThreads) {
    const double x = i < n ? f(i) : NAN;
    v[i] = isnan(x) ? INFINITY : x;
    if (i < n && !isnan(x)) c += 1.0;
  }
  const int cnt = (int)block_sum_fixed(c, s1, tid);  // also orders the stores above before the sort
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += kThreads) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const double a = v[i], b = v[l];
          if ((a > b) == up) {
            v[i] = b;
            v[l] = a;
          }
        }
      }
      __syncthreads();
    }
  const double med = cnt == 0 ? NAN : ((cnt & 1) ? v[cnt / 2] : 0.5 * (v[cnt / 2 - 1] + v[cnt / 2]));
  __syncthreads();
  return med;
}

struct ClipCfg {
  int tanh_clip;   // 1: tanh, 0: hard
  int median;      // 1: nanmedian center, 0: nanmean
  int mae;         // 1: mean absolute deviation, 0: standard deviation
  double clip_by;
  int from_previous;
};

// _get_clipping_center_and_width (loss_function.py:19-31) of the values f(i)
template <typename F>
__device__ void center_width(int n, F f, const ClipCfg& cfg, double* v, double* s1, double* s2, int tid, double& c,
                             double& w) {
  c = cfg.median ? nanmedian_of(n, f, v, s1, tid) : nanmean_of(n, f, s1, s2, tid);
  const double cc = c;
  if (cfg.mae)
    w = nanmean_of(n, [&](int i) { return fabs(f(i) - cc); }, s1, s2, tid);
  else
    w = sqrt(nanmean_of(n, [&](int i) { const double d = f(i) - cc; return d * d; }, s1, s2, tid));
  w *= cfg.clip_by;
}

// stats: [loss = E_mean_clipped, E_mean, E_var, E_var_clipped, new center, new width, center used, width used]
__global__ __launch_bounds__(kThreads) void k_clip_energies(const float* __restrict__ e_in, int n, ClipCfg cfg,
                                                            const double* __restrict__ state_in,
                                                            float* __restrict__ weights, float* __restrict__ eclip_out,
                                                            double* __restrict__ stats) {
  __shared__ double s1[kThreads];
  __shared__ double s2[kThreads];
  extern __shared__ double sortbuf[];  // kClipMax doubles when the center is a median
  const int tid = threadIdx.x;
  auto E = [&](int i) { return (double)e_in[i]; };
  const double E_mean = nanmean_of(n, E, s1, s2, tid);
  const double E_var = nanmean_of(n, [&](int i) { const double d = E(i) - E_mean; return d * d; }, s1, s2, tid);
  double c, w;
  if (cfg.from_previous && state_in) {
    c = state_in[0];
    w = state_in[1];
  } else {
    center_width(n, E, cfg, sortbuf, s1, s2, tid, c, w);
  }
  // _get_clip_func (loss_function.py:43-60); NaN stays NaN
  auto EC = [&](int i) {
    const double x = E(i);
    if (isnan(x)) return x;
    return cfg.tanh_clip ? c + tanh((x - c) / w) * w : fmin(fmax(x, c - w), c + w);
  };
  const double Ecm = nanmean_of(n, EC, s1, s2, tid);
  const double Ecv = nanmean_of(n, [&](int i) { const double d = EC(i) - Ecm; return d * d; }, s1, s2, tid);
  double c2, w2;
  center_width(n, EC, cfg, sortbuf, s1, s2, tid, c2, w2);
  for (int i = tid; i < n; i += kThreads) {
    const double x = EC(i);
    weights[i] = (float)((x - Ecm) / (double)n);
    if (eclip_out) eclip_out[i] = (float)x;
  }
  if (tid == 0) {
    stats[0] = Ecm;
    stats[1] = E_mean;
    stats[2] = E_var;
    stats[3] = Ecv;
    stats[4] = c2;
    stats[5] = w2;
    stats[6] = c;
    stats[7] = w;
  }
}

}  // namespace

int clip_max_walkers() { return kClipMax; }

void launch_clip_energies(const float* e, int n, int tanh_clip, int median, int mae, double clip_by,
                          int from_previous, const double* state_in, float* weights, float* eclip, double* stats,
                          hipStream_t s) {
  const ClipCfg cfg{tanh_clip, median, mae, clip_by, from_previous};
  const size_t smem = median ? (size_t)kClipMax * sizeof(double) : 0;
  hipLaunchKernelGGL(k_clip_energies, dim3(1), dim3(kThreads), smem, s, e, n, cfg, state_in, weights, eclip, stats);
}

void launch_energy_moments(const float* e, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_moments, dim3(1), dim3(kThreads), 0, s, e, n, out);
}

}  // namespace dwmi
