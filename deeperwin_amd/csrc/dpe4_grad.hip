// Reverse mode through the dpe4 embedding (value pass), for the parameter gradient G = sum_b c_b d log|psi|^2 /
// d theta.  The 1-el layers, orbitals and determinants use the FermiNet backward kernels (grad.hip); this file
// pushes the gradients of the columns [e_ee | e_ei] of every layer input X_l and of hm_l back through the
// convolutions and the pair and el-ion streams (ferminet_embedding.py:97-188, 207-267):
//   e_ee[i] = sum_j w[i, j] * hm[j]:   dw[i, j] = de[i] * hm[j],   dhm[j] = sum_i de[i] * w[i, j]
//   hm = tanh(h W_map + b):           dz = dhm (1 - hm^2)
//   e_ei[i] = sum_I q[i, I] * him[I]:  dq[i, I] = de[i] * him[I],  dhim[I] = sum_(b, i) de[i] * q[i, I]
// and through the pair streams (w = tanh(W_s p + b) per iteration, p' = (p + tanh(W_h p + b)) / sqrt 2, the first
// stream layer without the residual) and the el-ion stream, recomputed per pair / per (electron, ion).  The
// per-pair and per-(electron, ion) layer inputs and pre-activation gradients are written out so that the weight
// gradients are GEMMs over those rows (wgrad in grad.hip), in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.h"
#include "kernels.h"

namespace dwmi {

namespace {

constexpr int EMB = 32;

__device__ __forceinline__ float sqrt_g(float x) { return sqrtf(x); }
__device__ __forceinline__ double sqrt_g(double x) { return sqrt(x); }
__device__ __forceinline__ float tanh_g(float x) { return tanh_fast(x); }
__device__ __forceinline__ double tanh_g(double x) { return tanh(x); }

// dhm[b n + j][f] = sum_i de[b n + i][f] w[b][i][j][f];  dw[b][i][j][f] = de[b n + i][f] hm[b n + j][f].
// One thread per (walker, partner j, feature).
template <typename T>
__global__ __launch_bounds__(256) void k_conv_bwd(const T* __restrict__ dX, int ldx, int col0, const T* __restrict__ wm,
                                                  const T* __restrict__ hm, T* __restrict__ dhm, T* __restrict__ dw,
                                                  int B, int n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * n * EMB) return;
  const int f = (int)(t % EMB);
  const long bj = t / EMB;
  const long b = bj / n;
  const int j = (int)(bj - b * n);
  const T hj = hm[bj * EMB + f];
  T acc = T(0);
  for (int i = 0; i < n; ++i) {
    const T de = dX[(b * n + i) * ldx + col0 + f];
    const long pij = ((b * n + i) * n + j) * EMB + f;
    acc += de * wm[pij];
    dw[pij] = de * hj;
  }
  dhm[bj * EMB + f] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_dtanh(T* __restrict__ d, const T* __restrict__ y, long count) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) d[t] *= T(1) - y[t] * y[t];
}

template <typename T>
__global__ __launch_bounds__(256) void k_add_cols(T* __restrict__ dst, int ldd, const T* __restrict__ src, int lds,
                                                  long rows, int ncols) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * ncols) return;
  const long r = t / ncols;
  const int c = (int)(t - r * ncols);
  dst[r * ldd + c] += src[r * lds + c];
}

// Pair streams backward: one lane per pair, pairs of one spin relation per block range (as the value-pass
// kernel), so the weights are wave-uniform.  For iteration l the lane reads dw_l of its pair (natural
// [b][i][j] order), and writes, at its row of the relation-grouped order, the layer input p_l and the
// pre-activation gradients of the map (DZm[l]) and of the stream layer (DZh[l]).
template <typename T>
struct PairBwdArgs {
  const T* dw[MAXL_DPE4];
  const T* wm[MAXL_DPE4];
  T* X2[MAXL_DPE4];
  T* DZm[MAXL_DPE4];
  T* DZh[MAXL_DPE4];
};

template <typename T>
__global__ __launch_bounds__(256) void k_pairs_bwd(const float* __restrict__ r, const T* __restrict__ Wp,
                                                   const T* __restrict__ bp, PairBwdArgs<T> a, Shape sh,
                                                   int blocks_same) {
  const int n = sh.n, u = sh.u, d = sh.n - sh.u, L = sh.L;
  const int s = (int)blockIdx.x >= blocks_same ? 1 : 0;
  const long P = s == 0 ? (long)u * u + (long)d * d : 2L * u * d;
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
  const long row = (s ? (long)sh.B * ((long)u * u + (long)d * d) : 0) + t;  // relation-grouped row
  const long nat = ((b * n + i) * n + j) * EMB;                            // natural pair offset
  const float* rb = r + b * n * 3;
  const T rs2 = (T)0.70710678118654752440;
  // forward: the stream inputs p_l of every iteration
  T p[MAXL_DPE4][EMB];
#pragma unroll
  for (int k = 0; k < EMB; ++k) p[0][k] = T(0);
  if (i != j) {
    const T dx = (T)rb[j * 3 + 0] - (T)rb[i * 3 + 0], dy = (T)rb[j * 3 + 1] - (T)rb[i * 3 + 1],
            dz = (T)rb[j * 3 + 2] - (T)rb[i * 3 + 2];
    p[0][0] = sqrt_g(dx * dx + dy * dy + dz * dz);
  }
#pragma unroll
  for (int l = 0; l + 1 < MAXL_DPE4; ++l) {
    if (l + 1 >= L) break;
    const T* Wh = Wp + ((size_t)(L + l) * 2 + s) * EMB * EMB;
    const T* Bh = bp + ((size_t)(L + l) * 2 + s) * EMB;
#pragma unroll
    for (int fo = 0; fo < EMB; ++fo) {
      T z = Bh[fo];
      if (l == 0) {
        z += Wh[fo] * p[0][0];
      } else {
#pragma unroll
        for (int k = 0; k < EMB; ++k) z += Wh[k * EMB + fo] * p[l][k];
      }
      const T y = tanh_g(z);
      p[l + 1][fo] = (l == 0) ? y : (p[l][fo] + y) * rs2;
    }
  }
  // backward, last iteration first; g = gradient of the stream state p_{l+1}
  T g[EMB];
#pragma unroll
  for (int k = 0; k < EMB; ++k) g[k] = T(0);
#pragma unroll
  for (int l = MAXL_DPE4 - 1; l >= 0; --l) {
    if (l >= L) continue;
    T gx[EMB];
#pragma unroll
    for (int k = 0; k < EMB; ++k) gx[k] = T(0);
    if (l < L - 1) {  // stream layer l: p_{l+1} = tanh(W_h p_l + b) [+ p_l, / sqrt 2 from the second on]
      const T* Wh = Wp + ((size_t)(L + l) * 2 + s) * EMB * EMB;
      const T* Bh = bp + ((size_t)(L + l) * 2 + s) * EMB;
      T* dzh = a.DZh[l] + row * EMB;
#pragma unroll
      for (int fo = 0; fo < EMB; ++fo) {
        T z = Bh[fo];
        if (l == 0) {
          z += Wh[fo] * p[0][0];
        } else {
#pragma unroll
          for (int k = 0; k < EMB; ++k) z += Wh[k * EMB + fo] * p[l][k];
        }
        const T y = tanh_g(z);
        const T dy = (l == 0) ? g[fo] : g[fo] * rs2;
        const T dzv = dy * (T(1) - y * y);
        dzh[fo] = dzv;
        if (l == 0) {
          gx[0] += Wh[fo] * dzv;
        } else {
#pragma unroll
          for (int k = 0; k < EMB; ++k) gx[k] += Wh[k * EMB + fo] * dzv;
        }
      }
      if (l > 0) {
#pragma unroll
        for (int k = 0; k < EMB; ++k) gx[k] += g[k] * rs2;
      }
    }
    // map of iteration l: w_l = tanh(W_s p_l + b)
    {
      const T* Wm = Wp + ((size_t)l * 2 + s) * EMB * EMB;
      const T* dwl = a.dw[l] + nat;
      const T* wml = a.wm[l] + nat;
      T* dzm = a.DZm[l] + row * EMB;
#pragma unroll
      for (int fo = 0; fo < EMB; ++fo) {
        const T w = wml[fo];
        const T dzv = dwl[fo] * (T(1) - w * w);
        dzm[fo] = dzv;
        if (l == 0) {
          gx[0] += Wm[fo] * dzv;
        } else {
#pragma unroll
          for (int k = 0; k < EMB; ++k) gx[k] += Wm[k * EMB + fo] * dzv;
        }
      }
    }
    T* x2 = a.X2[l] + row * EMB;
#pragma unroll
    for (int k = 0; k < EMB; ++k) {
      x2[k] = p[l][k];
      g[k] = gx[k];
    }
  }
}

// El-ion stream backward: one lane per (electron row, ion).  de_ei of iteration l is read from the columns
// col[l] .. col[l] + g_l of dX_l; writes the stream inputs Q[l], the pre-activation gradients DZQ[l] of the
// stream layers, and PH[l][row][ion][f] = de_ei_l[row][f] q_l[row, ion][f] (summed over rows into dhim).
template <typename T>
struct ElIonBwdArgs {
  const T* dX[MAXL_DPE4];
  int ldx[MAXL_DPE4];
  int col[MAXL_DPE4];
  T* Q[MAXL_DPE4];
  T* DZQ[MAXL_DPE4];
  T* PH[MAXL_DPE4];
};

template <typename T>
__global__ __launch_bounds__(256) void k_elion_bwd(const float* __restrict__ r, const float* __restrict__ R,
                                                   const T* __restrict__ Wei, const T* __restrict__ bei,
                                                   const T* __restrict__ him, ElIonBwdArgs<T> a, Shape sh) {
  const int I = sh.I, L = sh.L;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)sh.B * sh.n * I) return;
  const long row = t / I;
  const int ion = (int)(t - row * I);
  const T rs2 = (T)0.70710678118654752440;
  T q[MAXL_DPE4][EMB];
#pragma unroll
  for (int k = 0; k < EMB; ++k) q[0][k] = T(0);
  {
    const float* ri = r + row * 3;
    const T dx = (T)ri[0] - (T)R[ion * 3 + 0], dy = (T)ri[1] - (T)R[ion * 3 + 1], dz = (T)ri[2] - (T)R[ion * 3 + 2];
    q[0][0] = sqrt_g(dx * dx + dy * dy + dz * dz);
    q[0][1] = dx;
    q[0][2] = dy;
    q[0][3] = dz;
  }
#pragma unroll
  for (int l = 0; l + 1 < MAXL_DPE4; ++l) {
    if (l + 1 >= L) break;
    const T* W = Wei + (size_t)l * EMB * EMB;
    const T* Bb = bei + (size_t)l * EMB;
    const int kin = l == 0 ? 4 : EMB;
#pragma unroll
    for (int fo = 0; fo < EMB; ++fo) {
      T z = Bb[fo];
#pragma unroll
      for (int k = 0; k < EMB; ++k)
        if (k < kin) z += W[k * EMB + fo] * q[l][k];
      const T y = tanh_g(z);
      q[l + 1][fo] = (l == 0) ? y : (q[l][fo] + y) * rs2;
    }
  }
  T g[EMB];
#pragma unroll
  for (int k = 0; k < EMB; ++k) g[k] = T(0);
#pragma unroll
  for (int l = MAXL_DPE4 - 1; l >= 0; --l) {
    if (l >= L) continue;
    const int gw = l == 0 ? 4 : EMB;  // width of e_ei_l
    T gx[EMB];
#pragma unroll
    for (int k = 0; k < EMB; ++k) gx[k] = T(0);
    if (l < L - 1) {
      const T* W = Wei + (size_t)l * EMB * EMB;
      const T* Bb = bei + (size_t)l * EMB;
      const int kin = l == 0 ? 4 : EMB;
      T* dzq = a.DZQ[l] + t * EMB;
#pragma unroll
      for (int fo = 0; fo < EMB; ++fo) {
        T z = Bb[fo];
#pragma unroll
        for (int k = 0; k < EMB; ++k)
          if (k < kin) z += W[k * EMB + fo] * q[l][k];
        const T y = tanh_g(z);
        const T dy = (l == 0) ? g[fo] : g[fo] * rs2;
        const T dzv = dy * (T(1) - y * y);
        dzq[fo] = dzv;
#pragma unroll
        for (int k = 0; k < EMB; ++k)
          if (k < kin) gx[k] += W[k * EMB + fo] * dzv;
      }
      if (l > 0) {
#pragma unroll
        for (int k = 0; k < EMB; ++k) gx[k] += g[k] * rs2;
      }
    }
    // the convolution of iteration l
    const T* de = a.dX[l] + row * a.ldx[l] + a.col[l];
    const T* hv = him + ((size_t)l * I + ion) * EMB;
    T* ph = a.PH[l] + t * EMB;
#pragma unroll
    for (int f = 0; f < EMB; ++f) {
      const T d = f < gw ? de[f] : T(0);
      gx[f] += d * hv[f];
      ph[f] = d * q[l][f];
    }
    T* qo = a.Q[l] + t * EMB;
#pragma unroll
    for (int k = 0; k < EMB; ++k) {
      qo[k] = q[l][k];
      g[k] = gx[k];
    }
  }
}

// Gradients of the ion maps and the ion table from dhim[l][I][f] (float, summed over walkers and electrons):
// dz = dhim (1 - him^2); d W_him_l[k][f] = sum_I table[idx_I][k] dz[I][f]; d b_him_l[f] = sum_I dz[I][f];
// d table[t][k] = sum_(l, I: idx_I = t) sum_f dz_l[I][f] W_him_l[k][f].  One workgroup, fixed order.
struct HimGradOut {
  float* W[MAXL_DPE4];  // h_ion_map_l w leaves [32][g_l]
  float* B[MAXL_DPE4];  // h_ion_map_l b leaves [g_l]
  float* table;         // embeddings leaf [n_types][32]
};

template <typename T>
__global__ __launch_bounds__(256) void k_him_bwd(const float* __restrict__ dhim, const T* __restrict__ him,
                                                 const T* __restrict__ table, const T* __restrict__ Whim,
                                                 const float* __restrict__ Z, int z_min, int n_types, int L, int I,
                                                 HimGradOut out) {
  for (int q = threadIdx.x; q < L * EMB * EMB; q += blockDim.x) {
    const int l = q / (EMB * EMB), k = (q / EMB) % EMB, f = q % EMB;
    const int g = l == 0 ? 4 : EMB;
    if (f >= g) continue;
    double acc = 0.0;
    for (int ion = 0; ion < I; ++ion) {
      const int idx = (int)lrintf(Z[ion]) - z_min;
      const double hv = (double)him[((size_t)l * I + ion) * EMB + f];
      const double dz = (double)dhim[((size_t)l * I + ion) * EMB + f] * (1.0 - hv * hv);
      const double tk = (idx >= 0 && idx < n_types) ? (double)table[idx * EMB + k] : 0.0;
      acc += tk * dz;
    }
    out.W[l][k * g + f] = (float)acc;
  }
  for (int q = threadIdx.x; q < L * EMB; q += blockDim.x) {
    const int l = q / EMB, f = q % EMB;
    const int g = l == 0 ? 4 : EMB;
    if (f >= g) continue;
    double acc = 0.0;
    for (int ion = 0; ion < I; ++ion) {
      const double hv = (double)him[((size_t)l * I + ion) * EMB + f];
      acc += (double)dhim[((size_t)l * I + ion) * EMB + f] * (1.0 - hv * hv);
    }
    out.B[l][f] = (float)acc;
  }
  for (int q = threadIdx.x; q < n_types * EMB; q += blockDim.x) {
    const int tt = q / EMB, k = q % EMB;
    double acc = 0.0;
    for (int l = 0; l < L; ++l) {
      const int g = l == 0 ? 4 : EMB;
      for (int ion = 0; ion < I; ++ion) {
        if ((int)lrintf(Z[ion]) - z_min != tt) continue;
        for (int f = 0; f < g; ++f) {
          const double hv = (double)him[((size_t)l * I + ion) * EMB + f];
          const double dz = (double)dhim[((size_t)l * I + ion) * EMB + f] * (1.0 - hv * hv);
          acc += dz * (double)Whim[((size_t)l * EMB + k) * EMB + f];
        }
      }
    }
    out.table[q] = (float)acc;
  }
}

}  // namespace

template <typename T>
void launch_dpe4_conv_bwd(const T* dX, int ldx, int col0, const T* wm, const T* hm, T* dhm, T* dw, const Shape& sh,
                          hipStream_t s) {
  const long total = (long)sh.B * sh.n * EMB;
  hipLaunchKernelGGL(k_conv_bwd<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, dX, ldx, col0, wm, hm,
                     dhm, dw, sh.B, sh.n);
}

template <typename T>
void launch_dtanh(T* d, const T* y, long count, hipStream_t s) {
  hipLaunchKernelGGL(k_dtanh<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, d, y, count);
}

template <typename T>
void launch_add_cols(T* dst, int ldd, const T* src, int lds, long rows, int ncols, hipStream_t s) {
  const long total = rows * ncols;
  hipLaunchKernelGGL(k_add_cols<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, dst, ldd, src, lds, rows,
                     ncols);
}

template <typename T>
void launch_dpe4_pairs_bwd(const float* r, const T* Wp, const T* bp, T* const* dw, T* const* wm, T* const* X2,
                           T* const* DZm, T* const* DZh, const Shape& sh, hipStream_t s) {
  PairBwdArgs<T> a{};
  for (int l = 0; l < sh.L; ++l) {
    a.dw[l] = dw[l];
    a.wm[l] = wm[l];
    a.X2[l] = X2[l];
    a.DZm[l] = DZm[l];
    a.DZh[l] = l < sh.L - 1 ? DZh[l] : nullptr;
  }
  const long u = sh.u, d = sh.n - sh.u;
  const int blocks_same = (int)((sh.B * (u * u + d * d) + 255) / 256);
  const int blocks_diff = (int)((sh.B * 2 * u * d + 255) / 256);
  hipLaunchKernelGGL(k_pairs_bwd<T>, dim3(blocks_same + blocks_diff), dim3(256), 0, s, r, Wp, bp, a, sh, blocks_same);
}

template <typename T>
void launch_dpe4_elion_bwd(const float* r, const float* R, const T* Wei, const T* bei, const T* him,
                           T* const* dX, const int* ldx, const int* col, T* const* Q, T* const* DZQ, T* const* PH,
                           const Shape& sh, hipStream_t s) {
  ElIonBwdArgs<T> a{};
  for (int l = 0; l < sh.L; ++l) {
    a.dX[l] = dX[l];
    a.ldx[l] = ldx[l];
    a.col[l] = col[l];
    a.Q[l] = Q[l];
    a.DZQ[l] = l < sh.L - 1 ? DZQ[l] : nullptr;
    a.PH[l] = PH[l];
  }
  const long total = (long)sh.B * sh.n * sh.I;
  hipLaunchKernelGGL(k_elion_bwd<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, r, R, Wei, bei, him, a,
                     sh);
}

template <typename T>
void launch_dpe4_him_bwd(const float* dhim, const T* him, const T* table, const T* Whim, const float* Z, int z_min,
                         int n_types, int L, int I, float* const* gW, float* const* gB, float* gTable,
                         hipStream_t s) {
  HimGradOut o{};
  for (int l = 0; l < L; ++l) {
    o.W[l] = gW[l];
    o.B[l] = gB[l];
  }
  o.table = gTable;
  hipLaunchKernelGGL(k_him_bwd<T>, dim3(1), dim3(256), 0, s, dhim, him, table, Whim, Z, z_min, n_types, L, I, o);
}

#define DPE4_GRAD_INST(T)                                                                                          \
  template void launch_dpe4_conv_bwd<T>(const T*, int, int, const T*, const T*, T*, T*, const Shape&, hipStream_t); \
  template void launch_dtanh<T>(T*, const T*, long, hipStream_t);                                                  \
  template void launch_add_cols<T>(T*, int, const T*, int, long, int, hipStream_t);                               \
  template void launch_dpe4_pairs_bwd<T>(const float*, const T*, const T*, T* const*, T* const*, T* const*,        \
                                         T* const*, T* const*, const Shape&, hipStream_t);                         \
  template void launch_dpe4_elion_bwd<T>(const float*, const float*, const T*, const T*, const T*, T* const*,      \
                                         const int*, const int*, T* const*, T* const*, T* const*, const Shape&,    \
                                         hipStream_t);                                                             \
  template void launch_dpe4_him_bwd<T>(const float*, const T*, const T*, const T*, const float*, int, int, int, int, \
                                       float* const*, float* const*, float*, hipStream_t);
DPE4_GRAD_INST(float)
DPE4_GRAD_INST(double)

}  // namespace dwmi
