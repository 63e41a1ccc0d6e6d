// Embedding of the dpe4 model (configuration.py:875-891, 422-441, 666-674; ferminet_embedding.py:97-188,
// 196-267): FermiNet iterations whose 1-el input is [h, mean_up h, mean_dn h, e_ee, e_ei] with
//   e_ee[i] = sum_j w_{s(i,j)}(p_l[i, j]) * hm_l[j],  hm_l = tanh(h W_map + b)      (SchNet-style, self pair in)
//   e_ei[i] = sum_I q_l[i, I] * tanh(h_ion W + b)[I],  h_ion = table[Z_I - Z_min]
// where p_l is the same/diff-spin pair stream (started from |r_j - r_i| alone) and q_l the el-ion stream
// (started from [|r_i - R_I|, r_i - R_I]), each taking one residual tanh layer between iterations, and
// w_same / w_diff one tanh layer applied to p_l.
//
// Every quantity carries the forward-mode Laplacian channels of the pass.  A pair quantity depends on
// d = r_j - r_i only and keeps 5 channels [value, d/dd_x, d/dd_y, d/dd_z, total Laplacian over all 3n
// coordinates] (d/dr_j = +grad, d/dr_i = -grad, so each |grad_c|^2 enters the Laplacian twice); an el-ion
// quantity depends on r_i only and keeps [value, d/dr_i (3), Laplacian in r_i].  The electron rows of X_l
// carry the pass's C = 3n+2 channels (value, 1 + 3e + c, Laplacian) padded to Cp rows.
//
// Kernels: k_dpe4_pairs (pair stream and the mapped w_l for every iteration, stored per pair),
// k_dpe4_elion (el-ion stream and e_ei of every iteration, written into X_l), k_dpe4_tanh (tanh rule on
// hm = X_l[:, h] W_map, whose GEMM runs on the library's GEMM kernels), k_dpe4_conv (e_ee into X_l).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.h"
#include "kernels.h"

namespace dwmi {

namespace {

constexpr int EMB = 32;  // emb_dim = width of the pair and el-ion streams

__device__ __forceinline__ float sqrt_t(float x) { return sqrtf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return sqrt(x); }
__device__ __forceinline__ float tanh_t(float x) { return tanh_fast(x); }
__device__ __forceinline__ double tanh_t(double x) { return tanh(x); }

// y = tanh(z) in NCH channels; lap_factor 2 for pair quantities (both electrons see the gradient), 1 for
// el-ion quantities.
template <int NCH, typename T>
__device__ __forceinline__ void tanh_channels(const T z[NCH], T y[NCH], T lap_factor) {
  const T t = tanh_t(z[0]);
  y[0] = t;
  if constexpr (NCH > 1) {
    const T d1 = T(1) - t * t, d2 = T(-2) * t * d1;
    y[1] = d1 * z[1];
    y[2] = d1 * z[2];
    y[3] = d1 * z[3];
    y[4] = d1 * z[4] + d2 * lap_factor * (z[1] * z[1] + z[2] * z[2] + z[3] * z[3]);
  }
}

// ------------------------------------------------------------------------------------------------------
// Pair streams.  One workgroup per (walker, 8 electrons i); thread (il, f) walks the partners j in order.
// Weight sets q = 0 .. L-1 are the maps w_same/w_diff of iteration q, q = L .. 2L-2 the stream layers
// h_same/h_diff of iteration q - L; each [2 (same, diff)][32 k][32 f] (rows k >= fan-in zero) with biases
// [2][32].  Output: wm[l][b][i][j][ch][f].
// ------------------------------------------------------------------------------------------------------
template <bool LAP, bool WLDS, typename T>
__global__ __launch_bounds__(256) void k_dpe4_pairs(const float* __restrict__ r, const T* __restrict__ Wp,
                                                    const T* __restrict__ bp, T* __restrict__ wm, Shape sh) {
  constexpr int NCH = LAP ? 5 : 1;
  extern __shared__ unsigned char smem_raw[];
  const int L = sh.L, NS = 2 * L - 1, n = sh.n, u = sh.u;
  // WLDS: the weight sets are staged in LDS (64 KB at L = 4 in FP32); otherwise read through the caches
  T* xs = (T*)smem_raw;                                   // [8][NCH][32]
  T* sW = xs + (size_t)8 * NCH * EMB;                     // [NS][2][32][32]
  T* sB = sW + (size_t)NS * 2 * EMB * EMB;                // [NS][2][32]
  if constexpr (WLDS) {
    for (int k = threadIdx.x; k < NS * 2 * EMB * EMB; k += 256) sW[k] = Wp[k];
    for (int k = threadIdx.x; k < NS * 2 * EMB; k += 256) sB[k] = bp[k];
  }
  const T* gW = WLDS ? sW : Wp;
  const T* gB = WLDS ? sB : bp;

  const int groups = (n + 7) >> 3;
  const long b = blockIdx.x / groups;
  const int il = threadIdx.x >> 5, f = threadIdx.x & 31;
  const int i = (int)(blockIdx.x % groups) * 8 + il;
  const bool active = i < n;
  const float* rb = r + b * n * 3;
  const T rix = active ? (T)rb[i * 3 + 0] : T(0), riy = active ? (T)rb[i * 3 + 1] : T(0),
          riz = active ? (T)rb[i * 3 + 2] : T(0);
  const T rs2 = (T)0.70710678118654752440;
  const long per_l = (long)sh.B * n * n * NCH * EMB;
  T* xrow = xs + (size_t)il * NCH * EMB;

  for (int j = 0; j < n; ++j) {
    const int s = ((i < u) == (j < u)) ? 0 : 1;
    T x[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) x[c] = T(0);
    if (active && j != i && f == 0) {
      const T dx = (T)rb[j * 3 + 0] - rix, dy = (T)rb[j * 3 + 1] - riy, dz = (T)rb[j * 3 + 2] - riz;
      const T rho = sqrt_t(dx * dx + dy * dy + dz * dz);
      x[0] = rho;
      if constexpr (LAP) {
        x[1] = dx / rho;
        x[2] = dy / rho;
        x[3] = dz / rho;
        x[4] = T(4) / rho;  // 2 / rho from each of the two electrons
      }
    }
    T* out = wm + ((((long)b * n + (active ? i : 0)) * n + j) * NCH) * EMB + f;
    for (int l = 0; l < L; ++l) {
      const int kin = (l == 0) ? 1 : EMB;
      __syncthreads();  // the weight tiles are loaded / the previous reads of xs are done
#pragma unroll
      for (int c = 0; c < NCH; ++c) xrow[c * EMB + f] = x[c];
      __syncthreads();
      const T* Wm = gW + ((size_t)l * 2 + s) * EMB * EMB;
      T zm[NCH], zh[NCH];
      zm[0] = gB[(l * 2 + s) * EMB + f];
#pragma unroll
      for (int c = 1; c < NCH; ++c) zm[c] = T(0);
      const bool stream = l < L - 1;
      const T* Wh = gW + ((size_t)(L + l) * 2 + s) * EMB * EMB;
      zh[0] = stream ? gB[((L + l) * 2 + s) * EMB + f] : T(0);
#pragma unroll
      for (int c = 1; c < NCH; ++c) zh[c] = T(0);
      for (int k = 0; k < kin; ++k) {
        const T wmk = Wm[k * EMB + f];
        const T whk = stream ? Wh[k * EMB + f] : T(0);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          const T xv = xrow[c * EMB + k];
          zm[c] += xv * wmk;
          zh[c] += xv * whk;
        }
      }
      T y[NCH];
      tanh_channels<NCH, T>(zm, y, T(2));
      if (active) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) out[(long)l * per_l + c * EMB] = y[c];
      }
      if (stream) {
        tanh_channels<NCH, T>(zh, y, T(2));
        if (kin == EMB) {
#pragma unroll
          for (int c = 0; c < NCH; ++c) y[c] = (x[c] + y[c]) * rs2;
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) x[c] = y[c];
      }
    }
  }
}

// ------------------------------------------------------------------------------------------------------
// El-ion stream and its convolution.  k_dpe4_him forms h_ion_map of every iteration, him[l][I][f] (zero
// beyond the iteration's width g_l; NaN for a Z outside the embedding table).  k_dpe4_elion: one workgroup
// per (walker, 8 electrons i); thread (il, f) walks the ions.  Stream layer l: Wei[l][32 k][32 f] (rows
// k >= fan-in zero).  e_ei of iteration l goes to the columns tail[l] + EMB + f (f < g_l) of X_l for every
// channel row.
// ------------------------------------------------------------------------------------------------------
template <typename T>
struct ElIonOut {
  T* X[MAXL_DPE4];
  int Kx[MAXL_DPE4];
  int tail[MAXL_DPE4];  // first column of [e_ee | e_ei] in X_l
};

template <typename T>
__global__ __launch_bounds__(256) void k_dpe4_him(const float* __restrict__ Z, const T* __restrict__ table,
                                                  const T* __restrict__ Whim, const T* __restrict__ bhim, int n_types,
                                                  int z_min, T* __restrict__ him, int L, int I) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L * I * EMB) return;
  const int l = q / (I * EMB), ion = (q / EMB) % I, f = q % EMB;
  const int g = (l == 0) ? 4 : EMB;
  T v = T(0);
  if (f < g) {
    const int idx = (int)lrintf(Z[ion]) - z_min;
    if (idx < 0 || idx >= n_types) {
      v = (T)NAN;  // a charge outside the embedding table: the walkers' results become NaN (flagged)
    } else {
      T z = bhim[l * EMB + f];
      for (int k = 0; k < EMB; ++k) z += table[idx * EMB + k] * Whim[((size_t)l * EMB + k) * EMB + f];
      v = tanh_t(z);
    }
  }
  him[q] = v;
}

template <bool LAP, typename T>
__global__ __launch_bounds__(256) void k_dpe4_elion(const float* __restrict__ r, const float* __restrict__ R,
                                                    const T* __restrict__ Wei, const T* __restrict__ bei,
                                                    const T* __restrict__ him, ElIonOut<T> out, Shape sh) {
  constexpr int NCH = LAP ? 5 : 1;
  __shared__ T xs[8 * NCH * EMB];
  const int L = sh.L, I = sh.I, n = sh.n, C = sh.C, Cp = sh.Cp;

  const int groups = (n + 7) >> 3;
  const long b = blockIdx.x / groups;
  const int il = threadIdx.x >> 5, f = threadIdx.x & 31;
  const int i = (int)(blockIdx.x % groups) * 8 + il;
  const bool active = i < n;
  const long bi = b * n + (active ? i : 0);
  const T rs2 = (T)0.70710678118654752440;
  T* xrow = xs + (size_t)il * NCH * EMB;
  T acc[MAXL_DPE4][NCH];
#pragma unroll
  for (int l = 0; l < MAXL_DPE4; ++l)
#pragma unroll
    for (int c = 0; c < NCH; ++c) acc[l][c] = T(0);
  const float* ri = r + bi * 3;

  for (int ion = 0; ion < I; ++ion) {
    T x[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) x[c] = T(0);
    if (active && f < 4) {
      const T d[3] = {(T)ri[0] - (T)R[ion * 3 + 0], (T)ri[1] - (T)R[ion * 3 + 1], (T)ri[2] - (T)R[ion * 3 + 2]};
      const T rho = sqrt_t(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (f == 0) {
        x[0] = rho;
        if constexpr (LAP) {
          x[1] = d[0] / rho;
          x[2] = d[1] / rho;
          x[3] = d[2] / rho;
          x[4] = T(2) / rho;
        }
      } else {
        x[0] = d[f - 1];
        if constexpr (LAP) x[f] = T(1);
      }
    }
#pragma unroll
    for (int l = 0; l < MAXL_DPE4; ++l) {
      if (l >= L) break;
      __syncthreads();  // the previous reads of xs are done
      const T hv = him[((size_t)l * I + ion) * EMB + f];
#pragma unroll
      for (int c = 0; c < NCH; ++c) acc[l][c] += x[c] * hv;
      if (l == L - 1) break;
      const int kin = (l == 0) ? 4 : EMB;
#pragma unroll
      for (int c = 0; c < NCH; ++c) xrow[c * EMB + f] = x[c];
      __syncthreads();
      const T* W = Wei + (size_t)l * EMB * EMB;
      T z[NCH];
      z[0] = bei[l * EMB + f];
#pragma unroll
      for (int c = 1; c < NCH; ++c) z[c] = T(0);
      for (int k = 0; k < kin; ++k) {
        const T w = W[k * EMB + f];
#pragma unroll
        for (int c = 0; c < NCH; ++c) z[c] += xrow[c * EMB + k] * w;
      }
      T y[NCH];
      tanh_channels<NCH, T>(z, y, T(1));
      if (kin == EMB) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) y[c] = (x[c] + y[c]) * rs2;
      }
#pragma unroll
      for (int c = 0; c < NCH; ++c) x[c] = y[c];
    }
  }
  if (!active) return;
#pragma unroll
  for (int l = 0; l < MAXL_DPE4; ++l) {
    if (l >= L) break;
    const int g = (l == 0) ? 4 : EMB;
    if (f >= g) continue;
    T* col = out.X[l] + bi * Cp * (long)out.Kx[l] + out.tail[l] + EMB + f;
    col[0] = acc[l][0];
    if constexpr (LAP) {
      for (int ch = 1; ch < C - 1; ++ch) {
        const int k = ch - 1, e = k / 3, c = k - 3 * e;
        col[(long)ch * out.Kx[l]] = (e == i) ? acc[l][1 + c] : T(0);
      }
      col[(long)(C - 1) * out.Kx[l]] = acc[l][4];
    }
    for (int ch = LAP ? C : 1; ch < Cp; ++ch) col[(long)ch * out.Kx[l]] = T(0);
  }
}

// ------------------------------------------------------------------------------------------------------
// hm = tanh(z + b) with the forward-Laplacian rule over an electron's channel rows of z = X_l[:, h] W_map.
// One thread per (electron row, f).
// ------------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_dpe4_tanh(T* __restrict__ hm, const T* __restrict__ bias, Shape sh) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)sh.B * sh.n * EMB) return;
  const long bi = t / EMB;
  const int f = (int)(t % EMB);
  T* p = hm + bi * sh.Cp * EMB + f;
  const T th = tanh_t(p[0] + bias[f]);
  p[0] = th;
  if (sh.C == 1) return;
  const T d1 = T(1) - th * th;
  T sq = T(0);
  for (int ch = 1; ch < sh.C - 1; ++ch) {
    const T z = p[(long)ch * EMB];
    sq += z * z;
    p[(long)ch * EMB] = d1 * z;
  }
  p[(long)(sh.C - 1) * EMB] = d1 * p[(long)(sh.C - 1) * EMB] - T(2) * th * d1 * sq;
  for (int ch = sh.C; ch < sh.Cp; ++ch) p[(long)ch * EMB] = T(0);
}

// ------------------------------------------------------------------------------------------------------
// e_ee[i][ch][f] = sum_j w[i, j][f] hm[j][ch][f] plus, in the Laplacian pass, the pair derivatives:
//   tangent row (e, c):  + (delta_{e j} - delta_{e i}) g_c[i, j] hm[j][0]
//   Laplacian row:       + L_w[i, j] hm[j][0] + 2 sum_c g_c[i, j] (hm[j][(j, c)] - hm[j][(i, c)])
// (g = d w / d d, L_w the pair's total Laplacian).  One workgroup per (walker, block of RB channel rows):
// the walker's hm rows of the block (n x RB x 32) and its value rows are staged in LDS once, then thread
// (il, f) forms electrons i = il, il + 8, ... reading each w[i, j] once per block.
// ------------------------------------------------------------------------------------------------------
template <bool LAP, int RB, typename T>
__global__ __launch_bounds__(256) void k_dpe4_conv(const T* __restrict__ wm, const T* __restrict__ hm,
                                                   T* __restrict__ X, int Kx, int col0, Shape sh) {
  constexpr int NCH = LAP ? 5 : 1;
  extern __shared__ unsigned char smem_raw[];
  T* hs = (T*)smem_raw;                         // [n][RB][32]
  const int n = sh.n, C = sh.C, Cp = sh.Cp;
  T* h0s = hs + (size_t)n * RB * EMB;           // [n][32]: value rows
  const int nRB = LAP ? (Cp + RB - 1) / RB : 1;
  const long b = blockIdx.x / nRB;
  const int ch0 = (int)(blockIdx.x % nRB) * RB;
  const T* hb = hm + (long)b * n * Cp * EMB;
  for (int q = threadIdx.x; q < n * RB * EMB; q += 256) {
    const int j = q / (RB * EMB), k = (q / EMB) % RB, f = q % EMB;
    hs[q] = (ch0 + k < C) ? hb[((long)j * Cp + ch0 + k) * EMB + f] : T(0);
  }
  for (int q = threadIdx.x; q < n * EMB; q += 256) h0s[q] = hb[(long)(q / EMB) * Cp * EMB + q % EMB];
  __syncthreads();
  const int il = threadIdx.x >> 5, f = threadIdx.x & 31;
  const bool has_lap = LAP && (C - 1) >= ch0 && (C - 1) < ch0 + RB;
  for (int i = il; i < n; i += 8) {
    T acc[RB];
#pragma unroll
    for (int k = 0; k < RB; ++k) acc[k] = T(0);
    const T* wrow = wm + ((long)(b * n + i) * n * NCH) * EMB + f;
    for (int j = 0; j < n; ++j) {
      const T* w = wrow + (long)j * NCH * EMB;
      const T w0 = w[0];
      const T* hj = hs + (size_t)j * RB * EMB + f;
#pragma unroll
      for (int k = 0; k < RB; ++k) acc[k] += w0 * hj[k * EMB];
      if constexpr (LAP) {
        if (j == i) continue;  // the self pair's w is constant
        const T h0 = h0s[j * EMB + f];
        const T g[3] = {w[EMB], w[2 * EMB], w[3 * EMB]};
#pragma unroll
        for (int k = 0; k < RB; ++k) {
          const int ch = ch0 + k;
          if (ch >= 1 && ch < C - 1) {
            const int kk = ch - 1, e = kk / 3, c = kk - 3 * e;
            if (e == j) acc[k] += g[c] * h0;
            else if (e == i) acc[k] -= g[c] * h0;
          }
        }
        if (has_lap) {
          const T* hg = hb + (long)j * Cp * EMB + f;
          T v = w[4 * EMB] * h0;
#pragma unroll
          for (int c = 0; c < 3; ++c)
            v += T(2) * g[c] * (hg[(long)(1 + 3 * j + c) * EMB] - hg[(long)(1 + 3 * i + c) * EMB]);
          acc[C - 1 - ch0] += v;
        }
      }
    }
    T* out = X + (long)(b * n + i) * Cp * Kx + col0 + f;
#pragma unroll
    for (int k = 0; k < RB; ++k)
      if (ch0 + k < Cp) out[(long)(ch0 + k) * Kx] = acc[k];
  }
}

// ------------------------------------------------------------------------------------------------------
// Value pass of the pair streams: one lane per pair.  The pairs of one spin relation (s = 0 same, 1 diff)
// are enumerated walker by walker (same: up-up then down-down; diff: up-down then down-up) and the blocks
// of each relation are contiguous, so a wavefront's weights are uniform (scalar loads).  Each lane keeps
// its pair's 32 features in registers and writes w_l[b][i][j][f] as 16-byte vectors.
// ------------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_dpe4_pairs_value(const float* __restrict__ r, const T* __restrict__ Wp,
                                                          const T* __restrict__ bp, T* __restrict__ wm, Shape sh,
                                                          int blocks_same) {
  const int n = sh.n, u = sh.u, d = sh.n - sh.u, L = sh.L;
  const int s = (int)blockIdx.x >= blocks_same ? 1 : 0;
  const long P = s == 0 ? (long)u * u + (long)d * d : 2L * u * d;  // pairs of this relation per walker
  const long t = (long)(blockIdx.x - (s ? blocks_same : 0)) * blockDim.x + threadIdx.x;
  if (t >= (long)sh.B * P) return;
  const long b = t / P;
  const int q = (int)(t - b * P);
  int i, j;
  if (s == 0) {
    if (q < u * u) {
      i = q / u;
      j = q - i * u;
    } else {
      const int q2 = q - u * u;
      i = u + q2 / d;
      j = u + q2 % d;
    }
  } else {
    if (q < u * d) {
      i = q / d;
      j = u + q % d;
    } else {
      const int q2 = q - u * d;
      i = u + q2 / u;
      j = q2 % u;
    }
  }
  const float* rb = r + b * n * 3;
  T x[EMB];
#pragma unroll
  for (int k = 0; k < EMB; ++k) x[k] = T(0);
  if (i != j) {
    const T dx = (T)rb[j * 3 + 0] - (T)rb[i * 3 + 0], dy = (T)rb[j * 3 + 1] - (T)rb[i * 3 + 1],
            dz = (T)rb[j * 3 + 2] - (T)rb[i * 3 + 2];
    x[0] = sqrt_t(dx * dx + dy * dy + dz * dz);
  }
  const long per_l = (long)sh.B * n * n * EMB;
  T* out = wm + ((long)(b * n + i) * n + j) * EMB;
  const T rs2 = (T)0.70710678118654752440;
  for (int l = 0; l < L; ++l) {
    const T* Wm = Wp + ((size_t)l * 2 + s) * EMB * EMB;
    const T* Bm = bp + ((size_t)l * 2 + s) * EMB;
    // the map w_l
#pragma unroll
    for (int f0 = 0; f0 < EMB; f0 += 4) {
      T y4[4];
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4) {
        const int fo = f0 + q4;
        T z = Bm[fo];
        if (l == 0) {
          z += Wm[fo] * x[0];
        } else {
#pragma unroll
          for (int k = 0; k < EMB; ++k) z += Wm[k * EMB + fo] * x[k];
        }
        y4[q4] = tanh_t(z);
      }
      if constexpr (sizeof(T) == 4) {
        *(float4*)(out + l * per_l + f0) = make_float4(y4[0], y4[1], y4[2], y4[3]);
      } else {
        *(double2*)(out + l * per_l + f0) = make_double2(y4[0], y4[1]);
        *(double2*)(out + l * per_l + f0 + 2) = make_double2(y4[2], y4[3]);
      }
    }
    if (l == L - 1) break;
    // the stream layer h_same / h_diff (residual from the second iteration on)
    const T* Wh = Wp + ((size_t)(L + l) * 2 + s) * EMB * EMB;
    const T* Bh = bp + ((size_t)(L + l) * 2 + s) * EMB;
    T y[EMB];
#pragma unroll
    for (int fo = 0; fo < EMB; ++fo) {
      T z = Bh[fo];
      if (l == 0) {
        z += Wh[fo] * x[0];
      } else {
#pragma unroll
        for (int k = 0; k < EMB; ++k) z += Wh[k * EMB + fo] * x[k];
      }
      y[fo] = tanh_t(z);
    }
#pragma unroll
    for (int k = 0; k < EMB; ++k) x[k] = (l == 0) ? y[k] : (x[k] + y[k]) * rs2;
  }
}

}  // namespace

template <typename T>
void launch_dpe4_pairs(const float* r, const T* Wp, const T* bp, T* wm, const Shape& sh, hipStream_t s) {
  if (sh.C == 1) {  // lane per pair, wave-uniform weights
    const long u = sh.u, d = sh.n - sh.u;
    const int blocks_same = (int)((sh.B * (u * u + d * d) + 255) / 256);
    const int blocks_diff = (int)((sh.B * 2 * u * d + 255) / 256);
    hipLaunchKernelGGL(k_dpe4_pairs_value<T>, dim3(blocks_same + blocks_diff), dim3(256), 0, s, r, Wp, bp, wm, sh,
                       blocks_same);
    return;
  }
  const unsigned grid = (unsigned)(sh.B * ((sh.n + 7) / 8));
  const int NS = 2 * sh.L - 1;
  constexpr bool WLDS = sizeof(T) == 4;
  const size_t smem = ((size_t)8 * 5 * EMB + (WLDS ? (size_t)NS * 2 * (EMB * EMB + EMB) : 0)) * sizeof(T);
  hipLaunchKernelGGL((k_dpe4_pairs<true, WLDS, T>), dim3(grid), dim3(256), smem, s, r, Wp, bp, wm, sh);
}

template <typename T>
void launch_dpe4_elion(const float* r, const float* R, const float* Z, const T* Wei, const T* bei, const T* Whim,
                       const T* bhim, const T* table, int n_types, int z_min, T* him, T* const* X, const int* Kx,
                       const int* tail, const Shape& sh, hipStream_t s) {
  const int total = sh.L * sh.I * EMB;
  hipLaunchKernelGGL(k_dpe4_him<T>, dim3((total + 255) / 256), dim3(256), 0, s, Z, table, Whim, bhim, n_types, z_min,
                     him, sh.L, sh.I);
  ElIonOut<T> o{};
  for (int l = 0; l < sh.L; ++l) {
    o.X[l] = X[l];
    o.Kx[l] = Kx[l];
    o.tail[l] = tail[l];
  }
  const unsigned grid = (unsigned)(sh.B * ((sh.n + 7) / 8));
  if (sh.C > 1)
    hipLaunchKernelGGL((k_dpe4_elion<true, T>), dim3(grid), dim3(256), 0, s, r, R, Wei, bei, him, o, sh);
  else
    hipLaunchKernelGGL((k_dpe4_elion<false, T>), dim3(grid), dim3(256), 0, s, r, R, Wei, bei, him, o, sh);
}

template <typename T>
void launch_dpe4_tanh(T* hm, const T* bias, const Shape& sh, hipStream_t s) {
  const long total = (long)sh.B * sh.n * EMB;
  hipLaunchKernelGGL(k_dpe4_tanh<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, hm, bias, sh);
}

template <typename T, int RB>
static void conv_launch(const T* wm, const T* hm, T* X, int Kx, int col0, const Shape& sh, hipStream_t s) {
  const int nRB = (sh.Cp + RB - 1) / RB;
  const size_t smem = ((size_t)sh.n * RB * EMB + (size_t)sh.n * EMB) * sizeof(T);
  hipLaunchKernelGGL((k_dpe4_conv<true, RB, T>), dim3((unsigned)(sh.B * nRB)), dim3(256), smem, s, wm, hm, X, Kx,
                     col0, sh);
}

template <typename T>
void launch_dpe4_conv(const T* wm, const T* hm, T* X, int Kx, int col0, const Shape& sh, hipStream_t s) {
  if (sh.C == 1) {
    const size_t smem = (size_t)2 * sh.n * EMB * sizeof(T);
    hipLaunchKernelGGL((k_dpe4_conv<false, 1, T>), dim3((unsigned)sh.B), dim3(256), smem, s, wm, hm, X, Kx, col0,
                       sh);
    return;
  }
  // channel rows per block: the staged hm block (n x RB x 32) stays within 48 KB of LDS
  const size_t row_bytes = (size_t)sh.n * EMB * sizeof(T);
  if (row_bytes * 16 <= 48 * 1024)
    conv_launch<T, 16>(wm, hm, X, Kx, col0, sh, s);
  else if (row_bytes * 8 <= 48 * 1024)
    conv_launch<T, 8>(wm, hm, X, Kx, col0, sh, s);
  else
    conv_launch<T, 4>(wm, hm, X, Kx, col0, sh, s);
}

#define DPE4_INST(T)                                                                                               \
  template void launch_dpe4_pairs<T>(const float*, const T*, const T*, T*, const Shape&, hipStream_t);             \
  template void launch_dpe4_elion<T>(const float*, const float*, const float*, const T*, const T*, const T*,       \
                                     const T*, const T*, int, int, T*, T* const*, const int*, const int*,          \
                                     const Shape&, hipStream_t);                                                   \
  template void launch_dpe4_tanh<T>(T*, const T*, const Shape&, hipStream_t);                                      \
  template void launch_dpe4_conv<T>(const T*, const T*, T*, int, int, const Shape&, hipStream_t);
DPE4_INST(float)
DPE4_INST(double)

}  // namespace dwmi
