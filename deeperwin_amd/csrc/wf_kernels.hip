// Streaming, per-walker and per-determinant kernels of the FermiNet-style wavefunction on gfx950.
//
// Activations are stored one row per (walker b, electron i, channel ch) with C channels per electron:
//   C == 1        : value only (log|psi| / MCMC)
//   C == 3 n + 2  : ch 0 value, ch 1 + 3 e + c the tangent along coordinate c of electron e,
//                   ch 3 n + 1 the Laplacian (sum over all 3 n second derivatives)
// which is the forward-mode Laplacian of the local energy.  Rows of a 1-el layer input X_l hold
//   [ h_l (Kh) | mean over spin-up partners of the 2-el stream (Ke) | same over spin-down (Ke) | 0 pad ]
// The spin means of h (the per-walker part of the symmetric features) are applied through a separate
// per-(walker, channel) GEMM, so the 1-el GEMM contracts over Kh + 2 Ke instead of 3 Kh + 2 Ke.
#include "common.h"
#include "kernels.h"

namespace dwmi {

static constexpr int MAXL = 4;  // FermiNet iterations supported by the fused 2-el kernel

// The device chain record (dw_chain_scalars).
struct ChainScalars {
  float stepsize;
  float acc_rate;
  int32_t step_nr;
  int32_t pad;
};

// Scalar-type generic math (T = float for the FP32/BF16 forms, double for the FP64 form).
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dtanh(float x) { return tanh_fast(x); }
__device__ __forceinline__ double dtanh(double x) { return tanh(x); }
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }

// ------------------------------------------------------------------------------------------------
// Layer-0 one-electron features: per ion [|r_i - R_I|, r_i - R_I] (utils.py:262-273,
// input_features.py:41-68, 237-239), with their tangents (only along electron i) and Laplacians
// (Laplacian of |x| in 3D is 2/|x|; of a coordinate difference 0).
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_features0(const float* __restrict__ r, const float* __restrict__ R, T* __restrict__ X0, int Kx0,
                            int tail0, int tailw, Shape sh, const float* __restrict__ normals,
                            const ChainScalars* __restrict__ chain, float* __restrict__ r_out) {
  // one thread per (electron row bi, column): the distance to the column's ion is formed once and the
  // column is written for every channel row (consecutive threads write consecutive columns of a row).
  // With normals, the positions are the Gaussian proposal r + stepsize * normals (k_propose's expression),
  // stored to r_out by the thread of column 0 for the kernels after this one.
  const long total = (long)sh.B * sh.n * Kx0;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long bi = div_idx(t, Kx0);
  const int col = (int)(t - bi * Kx0);
  const int i = (int)(bi - div_idx(bi, sh.n) * sh.n);
  const int I4 = 4 * sh.I, C = sh.C, Cp = sh.Cp;
  T* out = X0 + bi * Cp * (long)Kx0 + col;
  float rp[3] = {0.f, 0.f, 0.f};
  if (normals && (col == 0 || col < I4)) {
    const float step = chain->stepsize;
#pragma unroll
    for (int c = 0; c < 3; ++c) rp[c] = r[bi * 3 + c] + normals[bi * 3 + c] * step;
    if (col == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) r_out[bi * 3 + c] = rp[c];
    }
  }
  if (col >= I4) {
    // the columns [tail0, tail0 + tailw) are written by the 2-el kernels for the real channels; padding
    // columns and rows are zero
    for (int ch = (col >= tail0 && col < tail0 + tailw) ? C : 0; ch < Cp; ++ch) out[(long)ch * Kx0] = T(0);
    return;
  }
  const int ion = col >> 2, f = col & 3;
  const float* ri = normals ? rp : r + bi * 3;
  const T d[3] = {(T)ri[0] - (T)R[ion * 3 + 0], (T)ri[1] - (T)R[ion * 3 + 1], (T)ri[2] - (T)R[ion * 3 + 2]};
  const T rho = dsqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  out[0] = (f == 0) ? rho : d[f - 1];
  if (C == 1) {
    for (int ch = 1; ch < Cp; ++ch) out[(long)ch * Kx0] = T(0);
    return;
  }
  // tangent rows: only electron i's own three coordinates are non-zero
  for (int ch = 1; ch < C - 1; ++ch) {
    const int k = ch - 1, e = k / 3, c = k - 3 * e;
    T v = T(0);
    if (e == i) v = (f == 0) ? d[c] / rho : ((f - 1 == c) ? T(1) : T(0));
    out[(long)ch * Kx0] = v;
  }
  out[(long)(C - 1) * Kx0] = (f == 0) ? T(2) / rho : T(0);  // Laplacian of |x| in 3D is 2/|x|
  for (int ch = C; ch < Cp; ++ch) out[(long)ch * Kx0] = T(0);
}

template <typename T>
void launch_features0(const float* r, const float* R, T* X0, int Kx0, int tail0, int tailw, const Shape& sh,
                      hipStream_t s, const float* normals, const void* chain, float* r_out) {
  const long total = (long)sh.B * sh.n * Kx0;
  hipLaunchKernelGGL(k_features0<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, r, R, X0, Kx0, tail0,
                     tailw, sh, normals, (const ChainScalars*)chain, r_out);
}
template void launch_features0<float>(const float*, const float*, float*, int, int, int, const Shape&, hipStream_t,
                                      const float*, const void*, float*);
template void launch_features0<double>(const float*, const float*, double*, int, int, int, const Shape&, hipStream_t,
                                       const float*, const void*, float*);

// ------------------------------------------------------------------------------------------------
// Fused two-electron stream.  One workgroup per (walker b, group of 8 electrons i); 256 threads =
// 8 electrons x 32 features.  Every thread walks the partners j in order: the pair features
// [|r_j - r_i|, r_j - r_i] (utils.py:287-305) go through the 2-el MLPs 'h_el_el_l' (tanh, residual
// (x+y)/sqrt(2) when widths agree; mlp.py:13-16, 45-68) carrying the value, the 3 tangents along
// d = r_j - r_i and the Laplacian.  A pair feature depends on r_j - r_i only, so d/dr_j = +T and
// d/dr_i = -T and the Laplacian over all 3 n directions sees each |T_c|^2 twice.  The spin means over
// partners (ferminet_embedding.py:79-81; the self pair is included) are accumulated in registers in a
// fixed order and written into the pair columns of every layer input X_l:
//   value / Laplacian rows : mean over j in the spin block
//   tangent row (e, c)     : T_c[i, e] / n_s(e) in block s(e)  (e != i);  -sum_j T_c[i, j] / n_s (e == i)
// ------------------------------------------------------------------------------------------------
template <typename T>
struct TwoElOut {
  T* X[MAXL];
  int Kx[MAXL];
  int Kh[MAXL];
};

template <bool LAP, int L, typename T>
__global__ __launch_bounds__(256) void k_twoel(const float* __restrict__ r, TwoElOut<T> out,
                                               const T* __restrict__ W2e, const T* __restrict__ b2e,
                                               Shape sh) {
  constexpr int NCH = LAP ? 5 : 1;
  constexpr int NL2 = (L > 1) ? L - 1 : 1;
  const int n = sh.n, u = sh.u, C = sh.C, Cp = sh.Cp;
  const int groups = (n + 7) >> 3;
  const long b = blockIdx.x / groups;
  const int il = threadIdx.x >> 5, f = threadIdx.x & 31;
  const int i = (int)(blockIdx.x % groups) * 8 + il;
  const bool active = i < n;
  const long bi = b * n + (active ? i : 0);
  __shared__ T xs[8][NCH][32];

  T w[NL2][32], bb[NL2];
#pragma unroll
  for (int l2 = 0; l2 < L - 1; ++l2) {
    bb[l2] = b2e[l2 * 32 + f];
#pragma unroll
    for (int k = 0; k < 32; ++k) w[l2][k] = W2e[(l2 * 32 + k) * 32 + f];
  }
  T acc0[L][NCH], acc1[L][NCH];
#pragma unroll
  for (int l = 0; l < L; ++l)
#pragma unroll
    for (int c = 0; c < NCH; ++c) acc0[l][c] = acc1[l][c] = T(0);

  const float* rb = r + b * n * 3;
  const T rix = active ? (T)rb[i * 3 + 0] : T(0), riy = active ? (T)rb[i * 3 + 1] : T(0),
          riz = active ? (T)rb[i * 3 + 2] : T(0);
  const T inv_up = T(1) / (T)u, inv_dn = T(1) / (T)(n - u);
  const T rs2 = (T)0.70710678118654752440;

  for (int j = 0; j < n; ++j) {
    const int sj = (j < u) ? 0 : 1;
    const T inv_s = sj == 0 ? inv_up : inv_dn;
    const bool offdiag = active && j != i;
    T x[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) x[c] = T(0);
    if (offdiag && f < 4) {
      const T dx = (T)rb[j * 3 + 0] - rix, dy = (T)rb[j * 3 + 1] - riy, dz = (T)rb[j * 3 + 2] - riz;
      const T rho = dsqrt(dx * dx + dy * dy + dz * dz);
      if (f == 0) {
        x[0] = rho;
        if constexpr (LAP) {
          x[1] = dx / rho;
          x[2] = dy / rho;
          x[3] = dz / rho;
          x[4] = T(4) / rho;
        }
      } else {
        x[0] = (f == 1) ? dx : (f == 2 ? dy : dz);
        if constexpr (LAP) {
          x[1] = (f == 1) ? T(1) : T(0);
          x[2] = (f == 2) ? T(1) : T(0);
          x[3] = (f == 3) ? T(1) : T(0);
        }
      }
    }
    if (sj == 0) {
#pragma unroll
      for (int c = 0; c < NCH; ++c) acc0[0][c] += x[c];
    } else {
#pragma unroll
      for (int c = 0; c < NCH; ++c) acc1[0][c] += x[c];
    }
    if constexpr (LAP) {
      if (offdiag && f < 4) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          T* row = out.X[0] + (bi * Cp + 1 + 3 * j + c) * (long)out.Kx[0] + out.Kh[0];
          row[sj * 4 + f] = x[1 + c] * inv_s;
          row[(1 - sj) * 4 + f] = T(0);
        }
      }
    }
#pragma unroll
    for (int l2 = 0; l2 < L - 1; ++l2) {
      const int kin = (l2 == 0) ? 4 : 32;
      __syncthreads();
#pragma unroll
      for (int c = 0; c < NCH; ++c) xs[il][c][f] = x[c];
      __syncthreads();
      T z[NCH];
      z[0] = bb[l2];
#pragma unroll
      for (int c = 1; c < NCH; ++c) z[c] = T(0);
#pragma unroll
      for (int k4 = 0; k4 < 32; k4 += 4) {
        if (k4 < kin) {
#pragma unroll
          for (int c = 0; c < NCH; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) z[c] += xs[il][c][k4 + q] * w[l2][k4 + q];
        }
      }
      const T th = dtanh(z[0]);
      T y[NCH];
      y[0] = th;
      if constexpr (LAP) {
        const T d1 = T(1) - th * th, d2 = T(-2) * th * d1;
        y[1] = d1 * z[1];
        y[2] = d1 * z[2];
        y[3] = d1 * z[3];
        y[4] = d1 * z[4] + d2 * T(2) * (z[1] * z[1] + z[2] * z[2] + z[3] * z[3]);
      }
      if (kin == 32) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) y[c] = (x[c] + y[c]) * rs2;
      }
#pragma unroll
      for (int c = 0; c < NCH; ++c) x[c] = y[c];
      if (sj == 0) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) acc0[l2 + 1][c] += y[c];
      } else {
#pragma unroll
        for (int c = 0; c < NCH; ++c) acc1[l2 + 1][c] += y[c];
      }
      if constexpr (LAP) {
        if (offdiag) {
          const int l = l2 + 1;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            T* row = out.X[l] + (bi * Cp + 1 + 3 * j + c) * (long)out.Kx[l] + out.Kh[l];
            row[sj * 32 + f] = y[1 + c] * inv_s;
            row[(1 - sj) * 32 + f] = T(0);
          }
        }
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int Ke = (l == 0) ? 4 : 32;
    if (f >= Ke) continue;
    T* X = out.X[l];
    const long Kx = out.Kx[l];
    const int Kh = out.Kh[l];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const T inv = (s == 0) ? inv_up : inv_dn;
      T a[NCH];
#pragma unroll
      for (int c = 0; c < NCH; ++c) a[c] = (s == 0) ? acc0[l][c] : acc1[l][c];
      X[(bi * Cp) * Kx + Kh + s * Ke + f] = a[0] * inv;
      if constexpr (LAP) {
        X[(bi * Cp + C - 1) * Kx + Kh + s * Ke + f] = a[4] * inv;
#pragma unroll
        for (int c = 0; c < 3; ++c) X[(bi * Cp + 1 + 3 * i + c) * Kx + Kh + s * Ke + f] = -a[1 + c] * inv;
      }
    }
  }
}

// ------------------------------------------------------------------------------------------------
// FP32 two-electron stream on MFMA.  One wavefront per (walker b, group of EPW electrons): the 32 rows
// of a v_mfma_f32_32x32x2_f32 tile are (electron e of the group, partner j) with P = 32 / EPW partners
// per electron (EPW = 4, 2, 1 for n <= 8, 16, > 16; n > 32 uses two row tiles), the 32 features are its
// columns; each layer's weights are the B operand, held in registers.  The value, the 3 tangents along
// r_j - r_i and the Laplacian are 5 channels sharing those weights (no bias on derivative channels).
// A layer's output (C layout: lane = feature, 16 rows per lane) is transposed through wave-private LDS
// into the A layout of the next layer.  Spin sums over partners accumulate per lane over its rows and
// meet across lanes in a fixed order.  Same quantities as k_twoel: the self pair is included in the
// means, padded rows are excluded.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int mfma_row(int q, int h) { return 8 * (q >> 2) + 4 * h + (q & 3); }

template <bool LAP, int L, int EPW>
__global__ __launch_bounds__(64) void k_twoel_mfma(const float* __restrict__ r, TwoElOut<float> out,
                                                   const float* __restrict__ W2e, const float* __restrict__ b2e,
                                                   Shape sh, int tpw) {
  constexpr int NCH = LAP ? 5 : 1;
  constexpr int NL2 = L - 1;  // 2-el layers
  constexpr int NW = NL2 > 1 ? NL2 - 1 : 1;
  constexpr int P = 32 / EPW;  // partner rows per electron in a tile
  __shared__ float xs[NCH][32][33];
  const int lane = threadIdx.x, h = lane >> 5, fl = lane & 31;
  const int n = sh.n, u = sh.u, C = sh.C, Cp = sh.Cp;
  const int groups = (n + EPW - 1) / EPW;
  const long ntiles = (long)sh.B * groups;
  const float inv_up = 1.f / (float)u, inv_dn = 1.f / (float)(n - u);
  const float rs2 = 0.70710678118654752440f;

  // B operands: lane holds W[k = 2 s + h][f = fl].  The value pass keeps the 32 -> 32 layers' weights in
  // LDS (read per MFMA step), which lowers its registers to 3 wavefronts per SIMD; the 5-channel Laplacian
  // pass keeps them in registers (one wavefront per SIMD either way).
  constexpr bool WLDS = !LAP;
  __shared__ float wsh[WLDS ? NW : 1][WLDS ? 32 : 1][WLDS ? 32 : 1];
  float w0[2], wB[WLDS ? 1 : NW][16], bias[NL2 > 0 ? NL2 : 1];
#pragma unroll
  for (int q = 0; q < 2; ++q) w0[q] = W2e[(2 * q + h) * 32 + fl];
#pragma unroll
  for (int l2 = 1; l2 < NL2; ++l2)
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const float w = W2e[(l2 * 32 + 2 * q + h) * 32 + fl];
      if constexpr (WLDS)
        wsh[l2 - 1][2 * q + h][fl] = w;
      else
        wB[l2 - 1][q] = w;
    }
  if constexpr (WLDS) __syncthreads();
#pragma unroll
  for (int l2 = 0; l2 < NL2; ++l2) bias[l2] = b2e[l2 * 32 + fl];

  // tpw consecutive (walker, electron group) tiles per wavefront share the weight loads above
  for (long tile = (long)blockIdx.x * tpw, tile_end = tile + tpw; tile < tile_end && tile < ntiles; ++tile) {
  const long b = tile / groups;
  const int i0 = (int)(tile - b * groups) * EPW;  // first electron of the group
  const float* rb = r + b * n * 3;
  // per-lane spin sums of the layer outputs, per electron of the group
  float su[NL2 > 0 ? NL2 : 1][EPW][NCH], sd[NL2 > 0 ? NL2 : 1][EPW][NCH];
  float ru[NCH][2], rd[NCH][2];  // raw pair features of this lane's row (k = h, 2 + h)
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    ru[c][0] = ru[c][1] = rd[c][0] = rd[c][1] = 0.f;
#pragma unroll
    for (int l2 = 0; l2 < NL2; ++l2)
#pragma unroll
      for (int e = 0; e < EPW; ++e) su[l2][e][c] = sd[l2][e][c] = 0.f;
  }
  // this lane's A-layout row
  const int eA = fl / P, iA = i0 + eA;

  for (int jt = 0; jt < n; jt += P) {
    // ---- layer-0 input in the A layout: row (eA, j = jt + fl % P), k-slots h and 2 + h
    const int j = jt + fl % P;
    const bool pair = (iA < n) && (j < n) && (j != iA);
    float a0[NCH][2];
    {
      float dx = 0.f, dy = 0.f, dz = 0.f, rho = 1.f;
      if (pair) {
        dx = rb[j * 3 + 0] - rb[iA * 3 + 0];
        dy = rb[j * 3 + 1] - rb[iA * 3 + 1];
        dz = rb[j * 3 + 2] - rb[iA * 3 + 2];
        rho = sqrtf(dx * dx + dy * dy + dz * dz);
      }
      const float dv[3] = {dx, dy, dz};
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int k = 2 * q + h;
        float v = 0.f;
        if (pair) v = (k == 0) ? rho : dv[k - 1];
        a0[0][q] = v;
        if constexpr (LAP) {
#pragma unroll
          for (int c = 0; c < 3; ++c) a0[1 + c][q] = pair ? ((k == 0) ? dv[c] / rho : ((k - 1 == c) ? 1.f : 0.f)) : 0.f;
          a0[4][q] = (pair && k == 0) ? 4.f / rho : 0.f;
        }
      }
    }
    const bool up_j = j < u;
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (j < n && up_j) ru[c][q] += a0[c][q];
        if (j < n && !up_j) rd[c][q] += a0[c][q];
      }
    if constexpr (LAP) {
      if (pair) {  // tangent rows of X_0
        const float inv_s = up_j ? inv_up : inv_dn;
        const int sj = up_j ? 0 : 1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          float* row = out.X[0] + (((long)b * n + iA) * Cp + 1 + 3 * j + c) * (long)out.Kx[0] + out.Kh[0];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            row[sj * 4 + 2 * q + h] = a0[1 + c][q] * inv_s;
            row[(1 - sj) * 4 + 2 * q + h] = 0.f;
          }
        }
      }
    }

    // ---- 2-el layers
#pragma unroll
    for (int l2 = 0; l2 < NL2; ++l2) {
      v16f acc[NCH];
#pragma unroll
      for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[c][q] = (c == 0) ? bias[l2] : 0.f;
      if (l2 == 0) {
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
          for (int c = 0; c < NCH; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[c][q], w0[q], acc[c], 0, 0, 0);
      } else {
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
          for (int c = 0; c < NCH; ++c)
            acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(xs[c][fl][2 * q + h],
                                                          WLDS ? wsh[l2 - 1][2 * q + h][fl] : wB[WLDS ? 0 : l2 - 1][q],
                                                          acc[c], 0, 0, 0);
      }
      // tanh with the forward Laplacian rule (in place), then the residual (x + y) / sqrt 2 on the
      // 32 -> 32 layers with x read back from the transposition buffer (it still holds the layer input)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const float t = tanh_fast(acc[0][q]);
        if constexpr (LAP) {
          const float d1 = 1.f - t * t, d2 = -2.f * t * d1;
          const float z1 = acc[1][q], z2 = acc[2][q], z3 = acc[3][q];
          acc[4][q] = d1 * acc[4][q] + d2 * 2.f * (z1 * z1 + z2 * z2 + z3 * z3);
          acc[1][q] = d1 * z1;
          acc[2][q] = d1 * z2;
          acc[3][q] = d1 * z3;
        }
        acc[0][q] = t;
        if (l2 > 0) {
#pragma unroll
          for (int c = 0; c < NCH; ++c) acc[c][q] = (xs[c][mfma_row(q, h)][fl] + acc[c][q]) * rs2;
        }
      }
      // spin sums and the tangent rows of X_{l2+1}
      const long Kxo = out.Kx[l2 + 1];
      const int Kho = out.Kh[l2 + 1];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int row = mfma_row(q, h);
        const int e = row / P;  // compile-time per q and h when P >= 4
        const int jr = jt + row % P;
        const int ie = i0 + e;
        if (jr < n && ie < n) {
          const bool up = jr < u;
#pragma unroll
          for (int c = 0; c < NCH; ++c) {
#pragma unroll
            for (int e2 = 0; e2 < EPW; ++e2) {
              if (e2 == e) {
                if (up)
                  su[l2][e2][c] += acc[c][q];
                else
                  sd[l2][e2][c] += acc[c][q];
              }
            }
          }
          if constexpr (LAP) {
            if (jr != ie) {
              const float inv_s = up ? inv_up : inv_dn;
              const int sj = up ? 0 : 1;
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                float* rw = out.X[l2 + 1] + (((long)b * n + ie) * Cp + 1 + 3 * jr + c) * Kxo + Kho;
                rw[sj * 32 + fl] = acc[1 + c][q] * inv_s;
                rw[(1 - sj) * 32 + fl] = 0.f;
              }
            }
          }
        }
      }
      if (l2 + 1 < NL2) {  // transpose into the A layout of the next layer (wave-private LDS)
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
          for (int c = 0; c < NCH; ++c) xs[c][mfma_row(q, h)][fl] = acc[c][q];
        __syncthreads();
      }
    }
    __syncthreads();  // xs is rewritten by the next row tile
  }

  // ---- means into the pair columns: half-wave h writes spin block h
#pragma unroll
  for (int l2 = 0; l2 < NL2; ++l2) {
    float* X = out.X[l2 + 1];
    const long Kx = out.Kx[l2 + 1];
    const int col = out.Kh[l2 + 1] + h * 32 + fl;
#pragma unroll
    for (int e = 0; e < EPW; ++e) {
      float m[NCH];
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const float tu = su[l2][e][c] + __shfl_xor(su[l2][e][c], 32, 64);
        const float td = sd[l2][e][c] + __shfl_xor(sd[l2][e][c], 32, 64);
        m[c] = h == 0 ? tu * inv_up : td * inv_dn;
      }
      const int ie = i0 + e;
      if (ie < n) {
        const long base = ((long)b * n + ie) * Cp;
        X[base * Kx + col] = m[0];
        if constexpr (LAP) {
          X[(base + C - 1) * Kx + col] = m[4];
#pragma unroll
          for (int c = 0; c < 3; ++c) X[(base + 1 + 3 * ie + c) * Kx + col] = -m[1 + c];
        }
      }
    }
  }
  // raw features: reduce over the P partner lanes of each electron (k = h, 2 + h)
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int o = P / 2; o > 0; o >>= 1) {
        ru[c][q] += __shfl_xor(ru[c][q], o, 32);
        rd[c][q] += __shfl_xor(rd[c][q], o, 32);
      }
  if (fl % P == 0 && iA < n) {
    float* X = out.X[0];
    const long Kx = out.Kx[0];
    const int Kh = out.Kh[0];
    const long base = ((long)b * n + iA) * Cp;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int k = 2 * q + h;
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) {
        const float inv = s2 == 0 ? inv_up : inv_dn;
        X[base * Kx + Kh + s2 * 4 + k] = (s2 == 0 ? ru[0][q] : rd[0][q]) * inv;
        if constexpr (LAP) {
          X[(base + C - 1) * Kx + Kh + s2 * 4 + k] = (s2 == 0 ? ru[4][q] : rd[4][q]) * inv;
#pragma unroll
          for (int c = 0; c < 3; ++c)
            X[(base + 1 + 3 * iA + c) * Kx + Kh + s2 * 4 + k] = -(s2 == 0 ? ru[1 + c][q] : rd[1 + c][q]) * inv;
        }
      }
    }
  }
  }  // tiles
}

template <bool LAP, int L>
static void twoel_mfma_dispatch(const float* r, const TwoElOut<float>& o, const float* W2e, const float* b2e,
                                const Shape& sh, hipStream_t s) {
  // value pass: several electrons per wavefront for small n; the 5-channel Laplacian pass keeps one
  // electron per wavefront (its per-electron sums would not fit in registers)
  // tiles per wavefront (weights loaded once per wavefront): 4 in the Laplacian pass (544 -> 512 us at
  // 2048 walkers x 14 electrons); no gain measured in the value pass
  const int tpw_v = 1;
  int tpw_l = 4;
  if constexpr (!LAP) {
    const int epw = sh.n <= 8 ? 4 : (sh.n <= 16 ? 2 : 1);
    const long tiles = (long)sh.B * ((sh.n + epw - 1) / epw);
    const unsigned grid = (unsigned)((tiles + tpw_v - 1) / tpw_v);
    if (epw == 4) {
      hipLaunchKernelGGL((k_twoel_mfma<false, L, 4>), dim3(grid), dim3(64), 0, s, r, o, W2e, b2e, sh, tpw_v);
      return;
    }
    if (epw == 2) {
      hipLaunchKernelGGL((k_twoel_mfma<false, L, 2>), dim3(grid), dim3(64), 0, s, r, o, W2e, b2e, sh, tpw_v);
      return;
    }
    tpw_l = tpw_v;
  }
  const long tiles = (long)sh.B * sh.n;
  hipLaunchKernelGGL((k_twoel_mfma<LAP, L, 1>), dim3((unsigned)((tiles + tpw_l - 1) / tpw_l)), dim3(64), 0, s, r, o,
                     W2e, b2e, sh, tpw_l);
}

template <bool LAP, typename T>
static void twoel_dispatch(const float* r, const TwoElOut<T>& o, const T* W2e, const T* b2e, const Shape& sh,
                           hipStream_t s) {
  // FP32: MFMA kernel (width2 == 32 is enforced at model creation); FP64 form: the VALU kernel
  if constexpr (sizeof(T) == 4) {
    switch (sh.L) {
      case 2: twoel_mfma_dispatch<LAP, 2>(r, o, W2e, b2e, sh, s); break;
      case 3: twoel_mfma_dispatch<LAP, 3>(r, o, W2e, b2e, sh, s); break;
      default: twoel_mfma_dispatch<LAP, 4>(r, o, W2e, b2e, sh, s); break;
    }
  } else {
    const unsigned grid = (unsigned)(sh.B * ((sh.n + 7) / 8));
    switch (sh.L) {
      case 2: hipLaunchKernelGGL((k_twoel<LAP, 2, T>), dim3(grid), dim3(256), 0, s, r, o, W2e, b2e, sh); break;
      case 3: hipLaunchKernelGGL((k_twoel<LAP, 3, T>), dim3(grid), dim3(256), 0, s, r, o, W2e, b2e, sh); break;
      default: hipLaunchKernelGGL((k_twoel<LAP, 4, T>), dim3(grid), dim3(256), 0, s, r, o, W2e, b2e, sh); break;
    }
  }
}

template <typename T>
void launch_twoel(const float* r, T* const* X, const int* Kx, const int* Kh, const T* W2e, const T* b2e,
                  const Shape& sh, hipStream_t s) {
  TwoElOut<T> o;
  for (int l = 0; l < MAXL; ++l) {
    o.X[l] = l < sh.L ? X[l] : nullptr;
    o.Kx[l] = l < sh.L ? Kx[l] : 0;
    o.Kh[l] = l < sh.L ? Kh[l] : 0;
  }
  if (sh.C > 1)
    twoel_dispatch<true, T>(r, o, W2e, b2e, sh, s);
  else
    twoel_dispatch<false, T>(r, o, W2e, b2e, sh, s);
}
template void launch_twoel<float>(const float*, float* const*, const int*, const int*, const float*, const float*,
                                  const Shape&, hipStream_t);
template void launch_twoel<double>(const float*, double* const*, const int*, const int*, const double*,
                                   const double*, const Shape&, hipStream_t);

// ------------------------------------------------------------------------------------------------
// Spin means of h over electrons (ferminet_embedding.py:61-62, 72): MU[b, ch] = [mean_{i<u} h, mean_{i>=u} h].
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_spin_mean(const T* __restrict__ X, int Kx, int Kh, T* __restrict__ MU, int Km, Shape sh) {
  // one thread per (walker, channel, 4 consecutive features); Kh, Kx and Km are multiples of 4
  const int Km4 = Km >> 2;
  const long total = (long)sh.B * sh.Cp * Km4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long row = div_idx(t, Km4);
  const int f = (int)(t - row * Km4) * 4;
  const long b = div_idx(row, sh.Cp);
  const int ch = (int)(row - b * sh.Cp);
  T v0 = T(0), v1 = T(0), v2 = T(0), v3 = T(0);
  if (f < 2 * Kh) {
    const int s = f >= Kh ? 1 : 0;
    const int fi = f - s * Kh;
    const int i0 = s ? sh.u : 0, i1 = s ? sh.n : sh.u;
    const T* xp = X + ((b * sh.n + i0) * sh.Cp + ch) * (long)Kx + fi;
    const long stride = (long)sh.Cp * Kx;
    if constexpr (sizeof(T) == 4) {
      // one 16-byte load per electron (rows and fi are multiples of 4 floats)
#pragma unroll 4
      for (int i = i0; i < i1; ++i, xp += stride) {
        const float4 x = *(const float4*)xp;
        v0 += x.x;
        v1 += x.y;
        v2 += x.z;
        v3 += x.w;
      }
    } else {
#pragma unroll 4
      for (int i = i0; i < i1; ++i, xp += stride) {
        v0 += xp[0];
        v1 += xp[1];
        v2 += xp[2];
        v3 += xp[3];
      }
    }
    const T inv = T(1) / (T)(i1 - i0);
    v0 *= inv;
    v1 *= inv;
    v2 *= inv;
    v3 *= inv;
  }
  T* mp = MU + row * Km + f;
  mp[0] = v0;
  mp[1] = v1;
  mp[2] = v2;
  mp[3] = v3;
}

template <typename T>
void launch_spin_mean(const T* X, int Kx, int Kh, T* MU, int Km, const Shape& sh, hipStream_t s) {
  const long total = (long)sh.B * sh.Cp * (Km / 4);
  hipLaunchKernelGGL(k_spin_mean<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, X, Kx, Kh, MU, Km, sh);
}
template void launch_spin_mean<float>(const float*, int, int, float*, int, const Shape&, hipStream_t);
template void launch_spin_mean<double>(const double*, int, int, double*, int, const Shape&, hipStream_t);

// ------------------------------------------------------------------------------------------------
// tanh with forward-mode Laplacian as a separate pass (used by the FP64 form, whose GEMM has no fused
// epilogue): y = tanh z; dy_k = t' dz_k; Lap y = t' Lap z + t'' sum_k dz_k^2.  Padding channels -> 0.
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_tanh_rule(const T* __restrict__ Z, T* __restrict__ out, int ldo, int permute, Shape sh) {
  const long total = (long)sh.B * sh.n * sh.W1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int f = (int)(t % sh.W1);
  const long bi = t / sh.W1;
  const int i = (int)(bi % sh.n);
  const long b = bi / sh.n;
  const int C = sh.C, Cp = sh.Cp;
  const T* z = Z + bi * Cp * (long)sh.W1 + f;
  long obase;
  if (permute) {
    const int dn = sh.n - sh.u;
    obase = (i < sh.u) ? (b * sh.u + i) * Cp : (long)sh.B * sh.u * Cp + (b * dn + (i - sh.u)) * Cp;
  } else {
    obase = bi * Cp;
  }
  const T th = dtanh(z[0]);
  const T d1 = T(1) - th * th, d2 = T(-2) * th * d1;
  out[obase * ldo + f] = th;
  T S = T(0);
  for (int ch = 1; ch < C - 1; ++ch) {
    const T zc = z[(long)ch * sh.W1];
    S += zc * zc;
    out[(obase + ch) * ldo + f] = d1 * zc;
  }
  out[(obase + C - 1) * ldo + f] = d1 * z[(long)(C - 1) * sh.W1] + d2 * S;
  for (int ch = C; ch < Cp; ++ch) out[(obase + ch) * ldo + f] = T(0);
}

template <typename T>
void launch_tanh_rule(const T* Z, T* out, int ldo, int permute, const Shape& sh, hipStream_t s) {
  const long total = (long)sh.B * sh.n * sh.W1;
  hipLaunchKernelGGL(k_tanh_rule<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, Z, out, ldo, permute, sh);
}
template void launch_tanh_rule<double>(const double*, double*, int, int, const Shape&, hipStream_t);

// ------------------------------------------------------------------------------------------------
// Envelope x backflow (envelope_orbitals.py:91-127): A[b, d, i, k] = env_s[i, c] * bf[i, c], c = d n + k,
// env_s[i, c] = sum_I w_s[I, c] exp(-kappa_s[I, c] |r_i - R_I|), kappa = softplus(alpha), with the product
// rule for tangents (env depends on electron i only) and the Laplacian
//   Lap A = Lap(env) bf + 2 grad_i(env) . grad_i(bf) + env Lap(bf),   Lap e^{-k rho} = (k^2 - 2k/rho) e^{-k rho}.
// O rows are spin-major (up block then down block, Cp rows per electron); A is [B, D, C, n, n].
// ------------------------------------------------------------------------------------------------
// Spin-major row of electron i's first channel in O (up block, then down block, Cp rows per electron).
__device__ __forceinline__ long orb_row(const Shape& sh, long b, int i) {
  const int dn = sh.n - sh.u;
  return (i < sh.u) ? (b * sh.u + i) * sh.Cp : (long)sh.B * sh.u * sh.Cp + (b * dn + (i - sh.u)) * sh.Cp;
}

// Envelope of electron i of walker b at orbital column col; with GRAD also its gradient in r_i and its
// Laplacian.  Shared by the envelope kernel and the determinant kernels that apply it while reading O.
template <typename T, bool GRAD>
__device__ __forceinline__ void envelope_at(const T* __restrict__ kappa, const T* __restrict__ wenv,
                                            const float* __restrict__ r, const float* __restrict__ R, const Shape& sh,
                                            long b, int i, int col, T& env, T& g0, T& g1, T& g2, T& lenv) {
  const int NO = sh.D * sh.n, s = (i < sh.u) ? 0 : 1;
  const float* ri = r + (b * sh.n + i) * 3;
  const T* kap = kappa + (long)s * sh.I * NO;
  const T* we = wenv + (long)s * sh.I * NO;
  env = T(0);
  g0 = T(0);
  g1 = T(0);
  g2 = T(0);
  lenv = T(0);
#pragma unroll 1
  for (int I = 0; I < sh.I; ++I) {
    const T dx = (T)ri[0] - (T)R[I * 3 + 0], dy = (T)ri[1] - (T)R[I * 3 + 1], dz = (T)ri[2] - (T)R[I * 3 + 2];
    const T rho = dsqrt(dx * dx + dy * dy + dz * dz);
    const T k = kap[I * NO + col];
    const T e = we[I * NO + col] * dexp(-k * rho);
    env += e;
    if constexpr (GRAD) {
      const T coef = -k * e / rho;
      g0 += coef * dx;
      g1 += coef * dy;
      g2 += coef * dz;
      lenv += e * (k * k - T(2) * k / rho);
    }
  }
}

template <typename T>
__global__ void k_envelope(const T* __restrict__ O, int ldo, const float* __restrict__ r, const float* __restrict__ R,
                           const T* __restrict__ kappa, const T* __restrict__ wenv, T* __restrict__ A, Shape sh) {
  // thread order (b, d, i, k): consecutive threads write consecutive elements of A for every channel
  const int n = sh.n, NO = sh.D * n, C = sh.C;
  const long total = (long)sh.B * n * NO;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int nn_i = n * n;
  const long bd_ = div_idx(t, nn_i);
  const int ik = (int)(t - bd_ * nn_i);
  const int i = ik / n, kk_ = ik - (ik / n) * n;
  const long b = div_idx(bd_, sh.D);
  const int d_ = (int)(bd_ - b * sh.D);
  const int col = d_ * n + kk_;
  const long orow = orb_row(sh, b, i);
  T env, g0, g1, g2, lenv;
  if (C > 1)
    envelope_at<T, true>(kappa, wenv, r, R, sh, b, i, col, env, g0, g1, g2, lenv);
  else
    envelope_at<T, false>(kappa, wenv, r, R, sh, b, i, col, env, g0, g1, g2, lenv);
  const int d = col / n, k = col % n;
  const long nn = (long)n * n;
  T* Ab = A + ((b * sh.D + d) * C) * nn + (long)i * n + k;
  const T bf0 = O[orow * ldo + col];
  Ab[0] = env * bf0;
  if (C > 1) {
    // tangent channels: loads of 8 channels are issued before their stores
    const T* Oc = O + (orow + 1) * ldo + col;
    T* Ac = Ab + nn;
    const int nt = C - 2;
    int ch = 0;
    for (; ch + 8 <= nt; ch += 8) {
      T v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = env * Oc[(long)(ch + q) * ldo];
#pragma unroll
      for (int q = 0; q < 8; ++q) Ac[(long)(ch + q) * nn] = v[q];
    }
    for (; ch < nt; ++ch) Ac[(long)ch * nn] = env * Oc[(long)ch * ldo];
    // electron i's own coordinates also carry the envelope gradient (one rounding: fma with g_c bf0)
    Ac[(long)(3 * i + 0) * nn] = fma(env, Oc[(long)(3 * i + 0) * ldo], g0 * bf0);
    Ac[(long)(3 * i + 1) * nn] = fma(env, Oc[(long)(3 * i + 1) * ldo], g1 * bf0);
    Ac[(long)(3 * i + 2) * nn] = fma(env, Oc[(long)(3 * i + 2) * ldo], g2 * bf0);
    const T t0 = O[(orow + 1 + 3 * i) * ldo + col];
    const T t1 = O[(orow + 2 + 3 * i) * ldo + col];
    const T t2 = O[(orow + 3 + 3 * i) * ldo + col];
    Ab[(long)(C - 1) * nn] = lenv * bf0 + env * O[(orow + C - 1) * ldo + col] + T(2) * (g0 * t0 + g1 * t1 + g2 * t2);
  }
}

template <typename T>
void launch_envelope(const T* O, int ldo, const float* r, const float* R, const T* kappa, const T* wenv, T* A,
                     const Shape& sh, hipStream_t s) {
  const long total = (long)sh.B * sh.n * sh.D * sh.n;
  hipLaunchKernelGGL(k_envelope<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, O, ldo, r, R, kappa, wenv,
                     A, sh);
}
template void launch_envelope<float>(const float*, int, const float*, const float*, const float*, const float*, float*,
                                     const Shape&, hipStream_t);
template void launch_envelope<double>(const double*, int, const float*, const float*, const double*, const double*,
                                      double*, const Shape&, hipStream_t);

// ------------------------------------------------------------------------------------------------
// Determinants, in FP64.  One wavefront per (walker, det) matrix; lane j holds column j in registers.
// Partial-pivot LU (first maximal |pivot| wins, as LAPACK getrf): the pivot search runs in the pivot
// column's lane, the multipliers are broadcast with readlane, every lane updates its own column.
// sign and log|det| as jnp.linalg.slogdet (wavefunction.py:69).  In Laplacian mode the inverse is
// formed column per lane (forward / back substitution against the broadcast factors) and stored in
// LDS; the tangent channels are then processed several per pass (lane = (channel, column)):
//   Y_k = M^-1 dM_k,  g_k = tr Y_k,  q_k = tr(Y_k^2),   Lap log|det| = tr(M^-1 Lap M) - sum_k q_k.
// The 1-norm condition number |M|_1 |M^-1|_1 (Laplacian mode) or the pivot ratio max|U_pp| / min|U_pp|
// (value mode) is written per matrix to dcond; results are FP64.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// In-register LU of the column-per-lane matrix.  Returns the number of row swaps; perm[] (uniform)
// records the row order; mydiag receives U_jj in lane j.
template <int NMAX>
__device__ __forceinline__ int lu_columns(double (&col)[NMAX], int (&perm)[NMAX], double& mydiag, bool& sing, int n,
                                          int lane) {
  int swaps = 0;
#pragma unroll
  for (int i = 0; i < NMAX; ++i) perm[i] = i;
#pragma unroll
  for (int p = 0; p < NMAX; ++p) {
    if (p < n) {
    double best = -1.0;
    int bidx = p;
#pragma unroll
    for (int i = p; i < NMAX; ++i) {
      if (i < n) {
        const double a = fabs(col[i]);
        if (a > best) {
          best = a;
          bidx = i;
        }
      }
    }
    const int piv = __builtin_amdgcn_readlane(bidx, p);
    const double bval = readlane_d(best, p);
    if (!(bval > 0.0)) sing = true;
    if (piv != p) {
      ++swaps;
#pragma unroll
      for (int i = p + 1; i < NMAX; ++i) {
        if (i == piv) {
          const double t = col[p];
          col[p] = col[i];
          col[i] = t;
          const int tp = perm[p];
          perm[p] = perm[i];
          perm[i] = tp;
        }
      }
    }
    const double pivot = readlane_d(col[p], p);
    if (lane == p) {
      mydiag = col[p];
      const double rp = 1.0 / pivot;  // multipliers by the reciprocal, as LAPACK's dgetf2 scales
#pragma unroll
      for (int i = p + 1; i < NMAX; ++i)
        if (i < n) col[i] = col[i] * rp;
    }
#pragma unroll
    for (int i = p + 1; i < NMAX; ++i) {
      if (i < n) {
        const double li = readlane_d(col[i], p);
        if (lane > p) col[i] -= li * col[p];
      }
    }
    }
  }
  return swaps;
}

template <int NMAX>
__device__ __forceinline__ void finish_slogdet(double mydiag, int swaps, bool sing, int n, int lane, double& logabs,
                                               double& sgn) {
  const bool act = lane < n;
  double lg = act ? log(fabs(mydiag)) : 0.0;
  lg = wave_sum_d(lg);
  const unsigned long long neg = __ballot(act && mydiag < 0.0);
  const int nneg = __popcll(neg) + swaps;
  sgn = (nneg & 1) ? -1.0 : 1.0;
  logabs = lg;
  if (sing) {
    sgn = 0.0;
    logabs = -INFINITY;
  }
}

// Value-pass combine of one walker's D determinants (log|det| l, sign sg, matrix conditioning dc).
__device__ __forceinline__ void combine_value_walker(const double* l, const double* sg, const float* dc, int D,
                                                     float cond_thr, long b, float* __restrict__ out_lp,
                                                     float* __restrict__ out_phase, uint32_t* __restrict__ out_flags,
                                                     float* __restrict__ out_kappa) {
  double m = -INFINITY;
  for (int d = 0; d < D; ++d) m = fmax(m, (double)l[d]);
  double S = 0.0, Sabs = 0.0, cmax = 0.0;
  uint32_t fl = 0;
  for (int d = 0; d < D; ++d) {
    const double e = sg[d] * exp(l[d] - m);
    S += e;
    Sabs += fabs(e);
    cmax = fmax(cmax, (double)dc[d]);
  }
  // conditioning of log|psi|: matrix condition times the cancellation of the determinant sum
  const double kappa = cmax * Sabs / fabs(S);
  if (out_kappa) out_kappa[b] = (float)kappa;
  if (cond_thr > 0.f && !(kappa <= (double)cond_thr)) fl |= DWF_FP64;
  const double eps = 1e-8;
  const double aS = fabs(S);
  const double fval = log(aS + eps) + m;
  const float lp = (float)(2.0 * fval);
  if (out_lp) out_lp[b] = lp;
  if (out_phase) out_phase[b] = (S < 0.0) ? 3.14159265358979323846f : 0.f;
  if (!isfinite(lp)) fl |= DWF_NONFINITE;
  if (out_flags) out_flags[b] = fl;
}

// Value mode: 64 / NMAX matrices per wavefront (lane group g holds matrix g), 4 wavefronts per workgroup.
// The matrix is formed with a lane per column (coalesced reads of O), staged in wave-private LDS, and
// factorised with a lane per row: a row interchange is then a relabelling (no data moves), the pivot is a
// butterfly arg-max over the group's lanes (largest |a|, lowest row on ties), and each remaining row takes
// the pivot row's entries by shuffle.  Multipliers a[i][p] * (1 / pivot) (the reciprocal, as LAPACK's
// dgetf2 scales) and updates a[i][c] -= l a[piv][c] are the same operations as a column-oriented
// elimination, so U, log|det| (summed over U's diagonal in a fixed lane order) and the conditioning are
// too; the sign counts the negative pivots and the parity of the pivot order.
template <int NMAX, typename TA>
__global__ __launch_bounds__(256) void k_det_value(OrbSrc<TA> src, double* __restrict__ ld, double* __restrict__ sd,
                                                   float* __restrict__ dcond, long nmat, int n, Shape sh,
                                                   CombineValue co) {
  constexpr int G = 64 / NMAX;
  const int lane = threadIdx.x & 63;
  const int g = lane / NMAX, t = lane % NMAX, base = g * NMAX;
  const long bd = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * G + g;
  const bool valid = bd < nmat;
  const long bdl = valid ? bd : nmat - 1;  // out-of-range groups factor a copy, never stored
  const long b = bdl / sh.D;
  const int oc = (int)(bdl - b * sh.D) * n + t;  // orbital column of matrix column t
  const bool act = t < n;                         // column t (forming), row t (factorising) exists
  // element (r, c) = env(r_r, column c) * O[electron r, column c]: the envelope is applied as O is read.
  // Lane t forms column t in a rolled loop (the element phase stays out of the unrolled factorisation);
  // element (r, c) of group g sits at slot base + (c ^ r) of LDS row r, so both this write (fixed r) and
  // the row read below (fixed c, lanes over r) touch distinct banks.
  __shared__ TA sA[4][NMAX][64];
  TA (*mine)[64] = sA[threadIdx.x >> 6];
#pragma unroll 1
  for (int r = 0; r < n; ++r) {
    TA v = TA(0);
    if (act) {
      TA env, g0, g1, g2, lv;
      envelope_at<TA, false>(src.kappa, src.wenv, src.r, src.R, sh, b, r, oc, env, g0, g1, g2, lv);
      v = env * src.O[orb_row(sh, b, r) * src.ldo + oc];
    }
    mine[r][base + (t ^ r)] = v;
  }
  __syncthreads();  // every wavefront runs the same iterations: no early exit above
  double a[NMAX];  // row t
#pragma unroll
  for (int c = 0; c < NMAX; ++c) a[c] = (act && c < n) ? (double)mine[t][base + (c ^ t)] : 0.0;
  bool used = !act;  // padding rows never pivot
  int step_of = 0;   // elimination step at which this row became the pivot row
  bool sing = false;
  double mydiag = 1.0;  // lane p: U[p][p]
#pragma unroll
  for (int p = 0; p < NMAX; ++p) {
    if (p < n) {
      const double ap = fabs(a[p]);
      double v = (used || !(ap >= 0.0)) ? -1.0 : ap;
      int idx = t;
#pragma unroll
      for (int o = NMAX / 2; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o, NMAX);
        const int i2 = __shfl_xor(idx, o, NMAX);
        if (v2 > v || (v2 == v && i2 < idx)) {
          v = v2;
          idx = i2;
        }
      }
      if (!(v > 0.0)) sing = true;
      const double pivot = __shfl(a[p], base + idx, 64);
      if (t == p) mydiag = pivot;
      if (t == idx) {
        used = true;
        step_of = p;
      }
      const double rp = 1.0 / pivot;
      const double l = a[p] * rp;
#pragma unroll
      for (int c = p + 1; c < NMAX; ++c) {
        if (c < n) {
          const double pc = __shfl(a[c], base + idx, 64);
          if (!used) a[c] -= l * pc;
        }
      }
    }
  }
  // parity of the pivot order: inversions of step_of over the rows
  int inv = 0;
#pragma unroll
  for (int c = 0; c < NMAX; ++c) {
    const int sc = __shfl(step_of, base + c, 64);
    if (act && c > t && c < n && sc < step_of) ++inv;
  }
  double lg = act ? log(fabs(mydiag)) : 0.0;
  int neg = (act && mydiag < 0.0) ? 1 : 0;
  double umax = act ? fabs(mydiag) : 0.0, umin = act ? fabs(mydiag) : 1e300;
#pragma unroll
  for (int o = NMAX / 2; o > 0; o >>= 1) {
    lg += __shfl_xor(lg, o, NMAX);
    neg += __shfl_xor(neg, o, NMAX);
    inv += __shfl_xor(inv, o, NMAX);
    umax = fmax(umax, __shfl_xor(umax, o, NMAX));
    umin = fmin(umin, __shfl_xor(umin, o, NMAX));
  }
  const double ldv = sing ? -INFINITY : lg, sdv = sing ? 0.0 : (((neg + inv) & 1) ? -1.0 : 1.0);
  const float dcv = (float)(umax / umin);
  if (t == 0 && valid) {
    ld[bd] = ldv;
    sd[bd] = sdv;
    dcond[bd] = dcv;
  }
  if (co.fuse) {
    // the workgroup's 256 / NMAX matrices are whole walkers (D divides that number): combine them here
    constexpr int MPW = 256 / NMAX;
    __shared__ double c_ld[MPW], c_sd[MPW];
    __shared__ float c_dc[MPW];
    const int mi = (int)(threadIdx.x >> 6) * G + g;
    if (t == 0) {
      c_ld[mi] = ldv;
      c_sd[mi] = sdv;
      c_dc[mi] = dcv;
    }
    __syncthreads();
    const int D = sh.D, wpw = MPW / D;
    if ((int)threadIdx.x < wpw) {
      const long b = (long)blockIdx.x * wpw + threadIdx.x;
      const int o = (int)threadIdx.x * D;
      if (b < sh.B)
        combine_value_walker(c_ld + o, c_sd + o, c_dc + o, D, co.cond_thr, b, co.out_lp, co.out_phase, co.out_flags,
                             co.out_kappa);
    }
  }
}

// Laplacian mode (n <= 32): WPB matrices per workgroup, one per wavefront; LU and inverse with a lane per
// column, M^-1 through LDS into MFMA operand registers, then the channel passes on FP64 MFMA.
// ENV: the matrices are read from the orbitals O and the envelope is applied on the way (no stored A);
// while the value channel is read, each wavefront tabulates in LDS the envelope, the own-electron
// gradient terms g_c bf0 (channel 3i + c of row i gains them) and the Laplacian channel, [row][column].
template <int NMAX, int WPB, typename TA, bool ENV>
__global__ __launch_bounds__(64 * WPB) void k_det_lap(const TA* __restrict__ A, OrbSrc<TA> src, Shape sh,
                                                      double* __restrict__ ld, double* __restrict__ sd,
                                                      double* __restrict__ g, double* __restrict__ lapd,
                                                      float* __restrict__ dcond, long nmat, int n, int C) {
  extern __shared__ double dsm[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  long bd = (long)blockIdx.x * WPB + wv;
  const bool valid_mat = bd < nmat;
  if (!valid_mat) bd = nmat - 1;  // keep the wave in the barriers; results are not stored
  static_assert(NMAX == 16 || NMAX == 32, "channel passes are written for 16 x 16 MFMA blocks");
  constexpr int NN = NMAX * NMAX;
  double* sMi = dsm + (size_t)wv * NN;
  TA* sE = (TA*)(dsm + (size_t)WPB * NN) + (size_t)wv * 5 * NN;  // ENV tables: [row][column]
  TA* sG = sE + NN;                                               // [c][row][column]
  TA* sL = sE + 4 * NN;
  const long nn = (long)n * n;
  const TA* Ab = ENV ? nullptr : A + bd * C * nn;
  const bool act = lane < n;
  const int nd = 3 * n;
  const long b = bd / sh.D;
  const int dcol = (int)(bd - b * sh.D) * n;  // first orbital column of this determinant

  double col[NMAX];
  int perm[NMAX];
  double csum = 0.0;
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    if constexpr (ENV) {
      double v = 0.0;
      TA e0 = TA(0), gb0 = TA(0), gb1 = TA(0), gb2 = TA(0), lp = TA(0);
      if (act && i < n) {
        const int oc = dcol + lane;
        TA env, g0, g1, g2, lv;
        envelope_at<TA, true>(src.kappa, src.wenv, src.r, src.R, sh, b, i, oc, env, g0, g1, g2, lv);
        const TA* Oi = src.O + orb_row(sh, b, i) * src.ldo + oc;
        const TA bf0 = Oi[0];
        const TA t0 = Oi[(long)(1 + 3 * i) * src.ldo];
        const TA t1 = Oi[(long)(2 + 3 * i) * src.ldo];
        const TA t2 = Oi[(long)(3 + 3 * i) * src.ldo];
        v = (double)(env * bf0);
        e0 = env;
        gb0 = g0 * bf0;
        gb1 = g1 * bf0;
        gb2 = g2 * bf0;
        lp = lv * bf0 + env * Oi[(long)(C - 1) * src.ldo] + TA(2) * (g0 * t0 + g1 * t1 + g2 * t2);
      }
      if (lane < NMAX) {  // zero outside n x n
        sE[i * NMAX + lane] = e0;
        sG[(0 * NMAX + i) * NMAX + lane] = gb0;
        sG[(1 * NMAX + i) * NMAX + lane] = gb1;
        sG[(2 * NMAX + i) * NMAX + lane] = gb2;
        sL[i * NMAX + lane] = lp;
      }
      col[i] = v;
    } else {
      col[i] = (act && i < n) ? (double)Ab[i * n + lane] : 0.0;
    }
    csum += fabs(col[i]);
  }
  const double normA = wave_max_d(csum);
  double mydiag = 1.0;
  bool sing = false;
  const int swaps = lu_columns<NMAX>(col, perm, mydiag, sing, n, lane);
  double logabs, sgn;
  finish_slogdet<NMAX>(mydiag, swaps, sing, n, lane, logabs, sgn);

  // inverse, column `lane`: solve L y = P e_lane, then U x = y
  double x[NMAX];
#pragma unroll
  for (int p = 0; p < NMAX; ++p) {
    if (p < n) {
      double y = (perm[p] == lane) ? 1.0 : 0.0;
#pragma unroll
      for (int q = 0; q < p; ++q) y -= readlane_d(col[p], q) * x[q];
      x[p] = y;
    } else {
      x[p] = 0.0;
    }
  }
#pragma unroll
  for (int p = NMAX - 1; p >= 0; --p) {
    if (p < n) {
      double y = x[p];
#pragma unroll
      for (int q = p + 1; q < NMAX; ++q)
        if (q < n) y -= readlane_d(col[p], q) * x[q];
      x[p] = y / readlane_d(col[p], p);
    }
  }
  double isum = 0.0;
#pragma unroll
  for (int a = 0; a < NMAX; ++a) isum += act ? fabs(x[a]) : 0.0;
  const double normMi = wave_max_d(isum);
  // M^-1 to LDS, row-major [a][e]
  if (lane < NMAX) {
#pragma unroll
    for (int a = 0; a < NMAX; ++a) sMi[a * NMAX + lane] = act ? x[a] : 0.0;
  }
  __syncthreads();

  // Channel passes on FP64 MFMA (v_mfma_f64_16x16x4f64) over NB x NB blocks of 16.  Lane l holds
  // M^-1[I*16 + l%16][K*16 + l/16 + 4s] (A operand of Y_k = M^-1 dM_k) and dM_k[K*16 + l/16 + 4s][J*16 + l%16]
  // (its B operand); the same registers are the B and A operands of (Y_k^T)_IJ = sum_K (dM_k)_KI^T (M^-1)_JK^T.
  // Both products come out in one layout, so tr Y_k^2 is the sum of their elementwise product, and tr Y_k
  // is the sum of the operand products.  Channels run in groups of G: the loads of the next group are
  // issued before the current group is multiplied, and the G traces tr Y_k leave the wavefront in one
  // grouped reduction (same pairing order as a single wave sum).
  double qacc = 0.0;
  {
    constexpr int NB = NMAX / 16, G = (NMAX == 16) ? 2 : 1;  // measured: 2 / 1 fastest at n = 14 / 26
    const int ri = lane & 15, kq = lane >> 4;
    double miA[NB][NB][4];
#pragma unroll
    for (int I = 0; I < NB; ++I)
#pragma unroll
      for (int K = 0; K < NB; ++K)
#pragma unroll
        for (int q = 0; q < 4; ++q) miA[I][K][q] = sMi[(I * 16 + ri) * NMAX + K * 16 + kq + 4 * q];
    // ENV: envelope and the O row base of each element this lane holds
    TA envq[NB][NB][4];
    const TA* ob[NB][NB][4];
#pragma unroll
    for (int K = 0; K < NB; ++K)
#pragma unroll
      for (int J = 0; J < NB; ++J)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rr = K * 16 + kq + 4 * q, cc = J * 16 + ri;
          envq[K][J][q] = ENV ? sE[rr * NMAX + cc] : TA(1);
          ob[K][J][q] = (ENV && rr < n && cc < n) ? src.O + orb_row(sh, b, rr) * src.ldo + dcol + cc : src.O;
        }
    TA cur[G][NB][NB][4], nxt[G][NB][NB][4];
    auto load = [&](TA (&d)[G][NB][NB][4], int k0) {
#pragma unroll
      for (int c = 0; c < G; ++c) {
        const int k = k0 + c;
        const TA* Tg = ENV ? nullptr : Ab + (long)(1 + (k < nd ? k : 0)) * nn;
#pragma unroll
        for (int K = 0; K < NB; ++K)
#pragma unroll
          for (int J = 0; J < NB; ++J)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int rr = K * 16 + kq + 4 * q, cc = J * 16 + ri;
              const bool in = k < nd && rr < n && cc < n;
              if constexpr (ENV)
                d[c][K][J][q] = in ? ob[K][J][q][(long)(1 + k) * src.ldo] : TA(0);
              else
                d[c][K][J][q] = in ? Tg[rr * n + cc] : TA(0);
            }
      }
    };
    load(cur, 0);
    for (int k0 = 0; k0 < nd; k0 += G) {
      load(nxt, k0 + G);
      // one channel at a time: only the loads of the next group overlap the products (register budget)
      __builtin_amdgcn_sched_barrier(0);
      double gp[G];
#pragma unroll
      for (int c = 0; c < G; ++c) {
        double dmB[NB][NB][4];
        const int kr = (k0 + c) / 3, kc3 = (k0 + c) - 3 * kr;  // the channel's electron and component
#pragma unroll
        for (int K = 0; K < NB; ++K)
#pragma unroll
          for (int J = 0; J < NB; ++J)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              if constexpr (ENV) {
                const int rr = K * 16 + kq + 4 * q, cc = J * 16 + ri;
                const TA own = (rr == kr) ? sG[(kc3 * NMAX + rr) * NMAX + cc] : TA(0);
                dmB[K][J][q] = (double)fma(envq[K][J][q], cur[c][K][J][q], own);
              } else {
                dmB[K][J][q] = (double)cur[c][K][J][q];
              }
            }
        double t = 0.0;
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
          for (int K = 0; K < NB; ++K)
#pragma unroll
            for (int q = 0; q < 4; ++q) t += miA[I][K][q] * dmB[K][I][q];
        gp[c] = t;
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
          for (int J = 0; J < NB; ++J) {
            v4d y = {0.0, 0.0, 0.0, 0.0}, yt = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int K = 0; K < NB; ++K)
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                y = __builtin_amdgcn_mfma_f64_16x16x4f64(miA[I][K][q], dmB[K][J][q], y, 0, 0, 0);
                yt = __builtin_amdgcn_mfma_f64_16x16x4f64(dmB[K][I][q], miA[J][K][q], yt, 0, 0, 0);
              }
            qacc += y[0] * yt[0] + y[1] * yt[1] + y[2] * yt[2] + y[3] * yt[3];
          }
        __builtin_amdgcn_sched_barrier(0);
      }
      const double s = wave_sum_group<G>(gp, lane);
      const int kc = k0 + lane / (64 / G);
      if (lane % (64 / G) == 0 && kc < nd && valid_mat) g[bd * nd + kc] = s;
#pragma unroll
      for (int c = 0; c < G; ++c)
#pragma unroll
        for (int K = 0; K < NB; ++K)
#pragma unroll
          for (int J = 0; J < NB; ++J)
#pragma unroll
            for (int q = 0; q < 4; ++q) cur[c][K][J][q] = nxt[c][K][J][q];
    }
    qacc = wave_sum_d(qacc);
  }
  // tr(M^-1 Lap M): lane c takes column c of Lap M and row c of M^-1
  const TA* Tl = ENV ? nullptr : Ab + (long)(C - 1) * nn;
  double tp = 0.0;
  if (act) {
#pragma unroll
    for (int e = 0; e < NMAX; ++e)
      if (e < n) tp += sMi[lane * NMAX + e] * (double)(ENV ? sL[e * NMAX + lane] : Tl[e * n + lane]);
  }
  tp = wave_sum_d(tp);
  if (lane == 0 && valid_mat) {
    ld[bd] = logabs;
    sd[bd] = sgn;
    lapd[bd] = tp - qacc;
    dcond[bd] = (float)(normA * normMi);
  }
}

// Laplacian mode, large matrices (n > 32): one 256-thread workgroup per matrix, FP64 in LDS.
// Element-parallel partial-pivot LU (wave 0 searches the pivot), inverse by per-column substitution, then
// the tangent channels on FP64 MFMA blocks (below).
template <typename TA>
__global__ __launch_bounds__(256) void k_det_lds(const TA* __restrict__ A, double* __restrict__ ld,
                                                 double* __restrict__ sd, double* __restrict__ g,
                                                 double* __restrict__ lapd, float* __restrict__ dcond, int n, int C) {
  extern __shared__ double sm[];
  const int P = n + 1;
  double* T = sm;           // LU factors, then channel matrices
  double* Mi = T + n * P;   // inverse
  double* lcol = Mi + n * P;
  double* red = lcol + 64;
  int* perm = (int*)(red + 32);
  __shared__ int piv_sh, sing_sh;
  const long bd = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const long nn = (long)n * n;
  const TA* Ab = A + bd * C * nn;
  double csum_local = 0.0;
  for (int e = tid; e < n * n; e += 256) T[(e / n) * P + (e % n)] = (double)Ab[e];
  if (tid < n) perm[tid] = tid;
  if (tid == 0) sing_sh = 0;
  __syncthreads();
  {  // 1-norm of M (max column sum)
    for (int c = tid; c < n; c += 256) {
      double cs = 0.0;
      for (int rr = 0; rr < n; ++rr) cs += fabs(T[rr * P + c]);
      csum_local = fmax(csum_local, cs);
    }
  }
  int swaps = 0;
  for (int p = 0; p < n; ++p) {
    if (wid == 0) {
      double best = -1.0;
      int bidx = p;
      const int rr = p + lane;
      if (rr < n) {
        best = fabs(T[rr * P + p]);
        bidx = rr;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bidx, o, 64);
        if (ob > best || (ob == best && oi < bidx)) {
          best = ob;
          bidx = oi;
        }
      }
      if (lane == 0) {
        piv_sh = bidx;
        if (!(best > 0.0)) sing_sh = 1;
      }
    }
    __syncthreads();
    const int piv = piv_sh;
    if (piv != p) {
      for (int j = tid; j < n; j += 256) {
        const double t = T[p * P + j];
        T[p * P + j] = T[piv * P + j];
        T[piv * P + j] = t;
      }
      if (tid == 0) {
        const int tp = perm[p];
        perm[p] = perm[piv];
        perm[piv] = tp;
      }
      ++swaps;
    }
    __syncthreads();
    const double pivot = T[p * P + p];
    for (int rr = p + 1 + tid; rr < n; rr += 256) {
      const double l = T[rr * P + p] / pivot;
      lcol[rr] = l;
      T[rr * P + p] = l;
    }
    __syncthreads();
    const int m = n - p - 1;
    for (int e = tid; e < m * m; e += 256) {
      const int rr = p + 1 + e / m, cc = p + 1 + e % m;
      T[rr * P + cc] -= lcol[rr] * T[p * P + cc];
    }
    __syncthreads();
  }
  double lsum = 0.0;
  int nneg = 0;
  double umax = 0.0, umin = 1e300;
  for (int q = tid; q < n; q += 256) {
    const double up = T[q * P + q];
    lsum += log(fabs(up));
    nneg += up < 0.0 ? 1 : 0;
    umax = fmax(umax, fabs(up));
    umin = fmin(umin, fabs(up));
  }
  lsum = block_sum_d(lsum, red);
  const double negs = block_sum_d((double)nneg, red + 8);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    umax = fmax(umax, __shfl_xor(umax, o, 64));
    umin = fmin(umin, __shfl_xor(umin, o, 64));
  }
  __syncthreads();
  if (lane == 0) {
    lcol[wid] = umax;
    lcol[8 + wid] = umin;
  }
  __syncthreads();
  double gmax = 0.0, gmin = 1e300;
  for (int w = 0; w < 4; ++w) {
    gmax = fmax(gmax, lcol[w]);
    gmin = fmin(gmin, lcol[8 + w]);
  }
  const bool sing = sing_sh != 0;
  const double logabs = sing ? -INFINITY : lsum;
  const double sgn = sing ? 0.0 : ((((int)negs + swaps) & 1) ? -1.0 : 1.0);
  __syncthreads();
  // inverse: column j solves L U x = P e_j
  for (int j = tid; j < n; j += 256) {
    for (int p = 0; p < n; ++p) {
      double y = (perm[p] == j) ? 1.0 : 0.0;
      for (int q = 0; q < p; ++q) y -= T[p * P + q] * Mi[q * P + j];
      Mi[p * P + j] = y;
    }
    for (int p = n - 1; p >= 0; --p) {
      double x = Mi[p * P + j];
      for (int q = p + 1; q < n; ++q) x -= T[p * P + q] * Mi[q * P + j];
      Mi[p * P + j] = x / T[p * P + p];
    }
  }
  __syncthreads();
  double isum = 0.0;
  for (int c = tid; c < n; c += 256) {
    double cs = 0.0;
    for (int rr = 0; rr < n; ++rr) cs += fabs(Mi[rr * P + c]);
    isum = fmax(isum, cs);
  }
  double nA = 0.0, nI = 0.0;
  {
    double a1 = csum_local, a2 = isum;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a1 = fmax(a1, __shfl_xor(a1, o, 64));
      a2 = fmax(a2, __shfl_xor(a2, o, 64));
    }
    __syncthreads();
    if (lane == 0) {
      lcol[wid] = a1;
      lcol[8 + wid] = a2;
    }
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
      nA = fmax(nA, lcol[w]);
      nI = fmax(nI, lcol[8 + w]);
    }
  }
  // Channel passes on FP64 MFMA over 16 x 16 blocks (nb = ceil(n / 16) <= 4): wave J computes the block
  // column J of Y_k = M^-1 dM_k and the matching blocks of Y_k^T = dM_k^T M^-T, so tr Y_k^2 is the sum of
  // their elementwise products; lane l holds operand elements [l%16][l/16 + 4s] (A) / [l/16 + 4s][l%16] (B).
  const int nd = 3 * n;
  const int nb = (n + 15) >> 4;
  const int J = wid, ri = lane & 15, kq = lane >> 4;
  const bool wact = J < nb;
  auto mi_at = [&](int rr, int cc) -> double { return (rr < n && cc < n) ? Mi[rr * P + cc] : 0.0; };
  auto dm_at = [&](int rr, int cc) -> double { return (rr < n && cc < n) ? T[rr * P + cc] : 0.0; };
  double miJ[4][4];  // M^-1[J*16 + ri][K*16 + kq + 4 q]
#pragma unroll
  for (int K = 0; K < 4; ++K)
#pragma unroll
    for (int q = 0; q < 4; ++q) miJ[K][q] = wact ? mi_at(J * 16 + ri, K * 16 + kq + 4 * q) : 0.0;
  double qacc = 0.0;
  for (int ch = 1; ch <= nd; ++ch) {
    __syncthreads();
    const TA* Tg = Ab + (long)ch * nn;
    for (int e = tid; e < n * n; e += 256) T[(e / n) * P + (e % n)] = (double)Tg[e];
    __syncthreads();
    double pd = 0.0;
    if (wact) {
      double dmJ[4][4];  // dM[K*16 + kq + 4 q][J*16 + ri]
#pragma unroll
      for (int K = 0; K < 4; ++K)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          dmJ[K][q] = (K < nb) ? dm_at(K * 16 + kq + 4 * q, J * 16 + ri) : 0.0;
          pd += miJ[K][q] * dmJ[K][q];
        }
      for (int I = 0; I < nb; ++I) {
        v4d y = {0.0, 0.0, 0.0, 0.0}, yt = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int K = 0; K < 4; ++K) {
          if (K < nb) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              y = __builtin_amdgcn_mfma_f64_16x16x4f64(mi_at(I * 16 + ri, K * 16 + kq + 4 * q), dmJ[K][q], y, 0, 0, 0);
              yt = __builtin_amdgcn_mfma_f64_16x16x4f64(dm_at(K * 16 + kq + 4 * q, I * 16 + ri), miJ[K][q], yt, 0, 0, 0);
            }
          }
        }
        qacc += y[0] * yt[0] + y[1] * yt[1] + y[2] * yt[2] + y[3] * yt[3];
      }
    }
    const double gs = block_sum_d(pd, red);
    if (tid == 0) g[bd * nd + (ch - 1)] = gs;
  }
  qacc = block_sum_d(qacc, red + 8);
  __syncthreads();
  const TA* Tl = Ab + (long)(C - 1) * nn;
  double tp = 0.0;
  for (int e = tid; e < n * n; e += 256) {
    const int a = e / n, q = e % n;
    tp += Mi[a * P + q] * (double)Tl[q * n + a];
  }
  const double tr = block_sum_d(tp, red);
  if (tid == 0) {
    ld[bd] = logabs;
    sd[bd] = sgn;
    lapd[bd] = tr - qacc;
    dcond[bd] = (float)(nA * nI);
  }
}

// Value mode, 32 < n <= 64: one wavefront per matrix, the matrix in LDS (FP64, pitch n + 1), lane j owns
// column j.  Partial-pivot LU: argmax of the pivot column by shuffles (lowest row on ties), row swap and
// rank-1 update column-parallel; a single-wavefront workgroup, so each barrier only orders the LDS
// traffic of one wavefront.
template <typename TA>
__global__ __launch_bounds__(64) void k_det_value_lds(const TA* __restrict__ A, double* __restrict__ ld,
                                                      double* __restrict__ sd, float* __restrict__ dcond, int n,
                                                      long stride) {
  extern __shared__ double tm[];
  const int P = n + 1, j = threadIdx.x;
  const long bd = blockIdx.x;
  const TA* Ab = A + bd * stride;
  for (int e = j; e < n * n; e += 64) tm[(e / n) * P + (e % n)] = (double)Ab[e];
  __syncthreads();
  int swaps = 0;
  bool sing = false;
  double lg = 0.0, umax = 0.0, umin = 1e300;
  int neg = 0;
  for (int p = 0; p < n; ++p) {
    double best = (j >= p && j < n) ? fabs(tm[j * P + p]) : -1.0;
    int bidx = (j >= p && j < n) ? j : n;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bidx, o, 64);
      if (ob > best || (ob == best && oi < bidx)) {
        best = ob;
        bidx = oi;
      }
    }
    if (!(best > 0.0)) sing = true;
    if (bidx != p && bidx < n) {
      ++swaps;
      if (j < n) {
        const double t = tm[p * P + j];
        tm[p * P + j] = tm[bidx * P + j];
        tm[bidx * P + j] = t;
      }
    }
    __syncthreads();
    const double pivot = tm[p * P + p];
    const double ap = fabs(pivot);
    lg += log(ap);
    neg += pivot < 0.0 ? 1 : 0;
    umax = fmax(umax, ap);
    umin = fmin(umin, ap);
    const double rp = 1.0 / pivot;  // multipliers by the reciprocal, as LAPACK's dgetf2 scales
    if (j > p && j < n) tm[j * P + p] *= rp;
    __syncthreads();
    if (j > p && j < n) {
      const double up = tm[p * P + j];
      int i = p + 1;
      for (; i + 8 <= n; i += 8) {  // 8 rows per batch: all loads before the stores
        double l8[8], a8[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          l8[q] = tm[(i + q) * P + p];
          a8[q] = tm[(i + q) * P + j];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) tm[(i + q) * P + j] = a8[q] - l8[q] * up;
      }
      for (; i < n; ++i) tm[i * P + j] -= tm[i * P + p] * up;
    }
    __syncthreads();
  }
  if (j == 0) {
    ld[bd] = sing ? -INFINITY : lg;
    sd[bd] = sing ? 0.0 : (((neg + swaps) & 1) ? -1.0 : 1.0);
    dcond[bd] = (float)(umax / umin);
  }
}

template <typename TA>
bool launch_det(const TA* A, const OrbSrc<TA>& src, double* ld, double* sd, double* g, double* lapd, float* dcond,
                const Shape& sh, hipStream_t s, const CombineValue* combine) {
  const int n = sh.n;
  const long nmat = (long)sh.B * sh.D;
  const long stride = (long)sh.C * n * n;
  if (n > 32) {
    const size_t shm = (size_t)(2 * n * (n + 1) + 64 + 32) * sizeof(double) + 64 * sizeof(int);
    if (sh.C == 1) {
      const size_t shv = (size_t)n * (n + 1) * sizeof(double);  // <= 31.5 KB for n <= 63
      hipLaunchKernelGGL((k_det_value_lds<TA>), dim3((unsigned)nmat), dim3(64), shv, s, A, ld, sd, dcond, n, stride);
    } else {
      if (shm > 65536) (void)hipFuncSetAttribute((const void*)k_det_lds<TA>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
      hipLaunchKernelGGL((k_det_lds<TA>), dim3((unsigned)nmat), dim3(256), shm, s, A, ld, sd, g, lapd, dcond, n,
                         sh.C);
    }
    return false;
  }
  if (sh.C == 1) {
    // 4 wavefronts per workgroup, 64 / NMAX matrices per wavefront; the envelope is applied as O is read.
    // When D divides the matrices per workgroup, the walkers' combine runs at the end of the same kernel.
    const int mpw = n <= 8 ? 32 : (n <= 16 ? 16 : 8);
    CombineValue co{};
    if (combine && mpw % sh.D == 0) {
      co = *combine;
      co.fuse = 1;
    }
    const dim3 grid((unsigned)((nmat + mpw - 1) / mpw));
    if (n <= 8)
      hipLaunchKernelGGL((k_det_value<8, TA>), grid, dim3(256), 0, s, src, ld, sd, dcond, nmat, n, sh, co);
    else if (n <= 16)
      hipLaunchKernelGGL((k_det_value<16, TA>), grid, dim3(256), 0, s, src, ld, sd, dcond, nmat, n, sh, co);
    else
      hipLaunchKernelGGL((k_det_value<32, TA>), grid, dim3(256), 0, s, src, ld, sd, dcond, nmat, n, sh, co);
    return co.fuse != 0;
  }
  if (n <= 16) {
    // M^-1 (FP64) plus the envelope tables (5 [16][16] arrays of TA) per wavefront
    constexpr int NM = 16, W = 4;
    const size_t shm = (size_t)W * (NM * NM) * (sizeof(double) + 5 * sizeof(TA));
    hipLaunchKernelGGL((k_det_lap<NM, W, TA, true>), dim3((unsigned)((nmat + W - 1) / W)), dim3(64 * W), shm, s, A,
                       src, sh, ld, sd, g, lapd, dcond, nmat, n, sh.C);
  } else {
    constexpr int NM = 32, W = 2;
    const size_t shm = (size_t)W * (NM * NM) * sizeof(double);
    hipLaunchKernelGGL((k_det_lap<NM, W, TA, false>), dim3((unsigned)((nmat + W - 1) / W)), dim3(64 * W), shm, s, A,
                       src, sh, ld, sd, g, lapd, dcond, nmat, n, sh.C);
  }
  return false;
}
template bool launch_det<float>(const float*, const OrbSrc<float>&, double*, double*, double*, double*, float*,
                                const Shape&, hipStream_t, const CombineValue*);
template bool launch_det<double>(const double*, const OrbSrc<double>&, double*, double*, double*, double*, float*,
                                 const Shape&, hipStream_t, const CombineValue*);

// ------------------------------------------------------------------------------------------------
// Per-walker combination (FP64).  Sum of determinants with max shift and LOG_EPSILON
// (wavefunction.py:63-83): f = log(|S| + 1e-8) + m, S = sum_d s_d exp(l_d - m), log_psi_sqr = 2 f.
// Derivatives follow the same expression literally (the max routes its derivative to the arg-max set,
// averaged over ties, like jnp.max):
//   df_k   = sgn(S) dS_k / (|S| + eps) + dm_k,          dS_k = sum_d e_d (g_dk - dm_k)
//   Lap f  = sgn(S) sum_k d2S_k / (|S| + eps) - sum_k dS_k^2 / (|S| + eps)^2 + Lap m
//   sum_k d2S_k = sum_d e_d [ sum_k (g_dk - dm_k)^2 + Lap l_d - Lap m ]
// E_kin = -1/2 (Lap f + |grad f|^2)  (hamiltonian.py:216, 266 with log psi^2 = 2 f); E_pot as
// hamiltonian.py:17-39.
// ------------------------------------------------------------------------------------------------
__global__ void k_combine_value(const double* __restrict__ ld, const double* __restrict__ sd, const double* __restrict__ g,
                          const double* __restrict__ lapd, const float* __restrict__ dcond, float cond_thr,
                          const float* __restrict__ r, const float* __restrict__ R, const float* __restrict__ Z,
                          float* __restrict__ out_lp, float* __restrict__ out_phase, float* __restrict__ out_eloc,
                          uint32_t* __restrict__ out_flags, float* __restrict__ out_kappa, Shape sh) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= sh.B) return;
  const int D = sh.D;
  combine_value_walker(ld + b * D, sd + b * D, dcond + b * D, D, cond_thr, b, out_lp, out_phase, out_flags, out_kappa);
}

// Laplacian pass: one wavefront per walker; the formulas above.  Lanes split the 3n gradient components
// and the electrons of the potential; partial sums meet in a fixed shuffle tree (deterministic).
__global__ __launch_bounds__(256) void k_combine_lap(const double* __restrict__ ld, const double* __restrict__ sd,
                                                     const double* __restrict__ g, const double* __restrict__ lapd,
                                                     const float* __restrict__ dcond, float cond_thr,
                                                     const float* __restrict__ r, const float* __restrict__ R,
                                                     const float* __restrict__ Z, float* __restrict__ out_lp,
                                                     float* __restrict__ out_phase, float* __restrict__ out_eloc,
                                                     uint32_t* __restrict__ out_flags, float* __restrict__ out_kappa,
                                                     Shape sh) {
  const int lane = threadIdx.x & 63;
  const long b = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= sh.B) return;  // uniform per wavefront
  const int D = sh.D, n = sh.n, nd = 3 * n;
  const double* l = ld + b * D;
  const double* sg = sd + b * D;
  double m = -INFINITY;
  for (int d = 0; d < D; ++d) m = fmax(m, l[d]);
  double S = 0.0, Sabs = 0.0, cmax = 0.0, lapm = 0.0;
  int cnt = 0;
  for (int d = 0; d < D; ++d) {
    const double e = sg[d] * exp(l[d] - m);
    S += e;
    Sabs += fabs(e);
    if (l[d] == m) {
      ++cnt;
      lapm += lapd[b * D + d];
    }
    cmax = fmax(cmax, (double)dcond[b * D + d]);
  }
  lapm /= cnt;
  const double kappa = cmax * Sabs / fabs(S);
  uint32_t fl = 0;
  if (cond_thr > 0.f && !(kappa <= (double)cond_thr)) fl |= DWF_FP64;
  const double eps = 1e-8;
  const double aS = fabs(S);
  const double fval = log(aS + eps) + m;
  const float lp = (float)(2.0 * fval);
  if (!isfinite(lp)) fl |= DWF_NONFINITE;
  const double sgnS = (S > 0.0) ? 1.0 : ((S < 0.0) ? -1.0 : 0.0);
  const double den = aS + eps;
  const double* gb = g + b * D * nd;
  double p_d2S = 0.0, p_dS2 = 0.0, p_df2 = 0.0;
  for (int k = lane; k < nd; k += 64) {
    double dm = 0.0;
    for (int d = 0; d < D; ++d)
      if (l[d] == m) dm += gb[d * nd + k];
    dm /= cnt;
    double dS = 0.0, d2 = 0.0;
    for (int d = 0; d < D; ++d) {
      const double e = sg[d] * exp(l[d] - m);
      const double x = gb[d * nd + k] - dm;
      dS += e * x;
      d2 += e * x * x;
    }
    p_d2S += d2;
    p_dS2 += dS * dS;
    const double df = sgnS * dS / den + dm;
    p_df2 += df * df;
  }
  // potential: electron pairs and electron-ion terms over lanes, ion-ion on every lane
  const float* rb = r + b * n * 3;
  double pe = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double xi = rb[i * 3], yi = rb[i * 3 + 1], zi = rb[i * 3 + 2];
    for (int j = i + 1; j < n; ++j) {
      const double dx = xi - rb[j * 3], dy = yi - rb[j * 3 + 1], dz = zi - rb[j * 3 + 2];
      pe += 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    }
    for (int I = 0; I < sh.I; ++I) {
      const double dx = xi - R[I * 3], dy = yi - R[I * 3 + 1], dz = zi - R[I * 3 + 2];
      pe -= (double)Z[I] / sqrt(dx * dx + dy * dy + dz * dz);
    }
  }
  double sum_d2S = wave_sum_d(p_d2S);
  const double sum_dS2 = wave_sum_d(p_dS2), sum_df2 = wave_sum_d(p_df2);
  double epot = wave_sum_d(pe);
  if (lane != 0) return;
  for (int d = 0; d < D; ++d) sum_d2S += sg[d] * exp(l[d] - m) * (lapd[b * D + d] - lapm);
  for (int I = 0; I < sh.I; ++I)
    for (int J = I + 1; J < sh.I; ++J) {
      const double dx = R[I * 3] - R[J * 3], dy = R[I * 3 + 1] - R[J * 3 + 1], dz = R[I * 3 + 2] - R[J * 3 + 2];
      epot += (double)Z[I] * Z[J] / sqrt(dx * dx + dy * dy + dz * dz);
    }
  const double lapf = sgnS * sum_d2S / den - sum_dS2 / (den * den) + lapm;
  const double ekin = -0.5 * (lapf + sum_df2);
  const float eloc = (float)(ekin + epot);
  if (!isfinite(eloc)) fl |= DWF_NONFINITE;
  if (out_kappa) out_kappa[b] = (float)kappa;
  if (out_lp) out_lp[b] = lp;
  if (out_phase) out_phase[b] = (S < 0.0) ? 3.14159265358979323846f : 0.f;
  if (out_eloc) out_eloc[b] = eloc;
  if (out_flags) out_flags[b] = fl;
}

void launch_combine(const double* ld, const double* sd, const double* g, const double* lapd, const float* dcond,
                    float cond_thr, const float* r, const float* R, const float* Z, float* out_lp, float* out_phase,
                    float* out_eloc, uint32_t* out_flags, float* out_kappa, const Shape& sh, hipStream_t s) {
  const unsigned grid = (unsigned)((sh.B + 127) / 128);
  if (sh.C > 1)
    hipLaunchKernelGGL(k_combine_lap, dim3((unsigned)((sh.B + 3) / 4)), dim3(256), 0, s, ld, sd, g, lapd, dcond,
                       cond_thr, r, R, Z, out_lp, out_phase, out_eloc, out_flags, out_kappa, sh);
  else
    hipLaunchKernelGGL(k_combine_value, dim3(grid), dim3(128), 0, s, ld, sd, g, lapd, dcond, cond_thr, r, R, Z,
                       out_lp, out_phase, out_eloc, out_flags, out_kappa, sh);
}

// ------------------------------------------------------------------------------------------------
// Metropolis-Hastings pieces (mcmc.py:175-180, 345-387).
// ------------------------------------------------------------------------------------------------

__global__ void k_fill_i32(int32_t* p, int32_t v, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = v;
}
void launch_fill_i32(int32_t* p, int32_t v, int count, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32, dim3((count + 255) / 256), dim3(256), 0, s, p, v, count);
}

// Compaction of flagged walkers for the FP64 recompute: out[k] = r[idx[k]]; and the inverse scatter.
__global__ void k_gather_walkers(const float* __restrict__ r, const int32_t* __restrict__ idx, int count, int n,
                                 float* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = 3L * n;
  if (t >= count * per) return;
  const long k = t / per, c = t % per;
  out[t] = r[(long)idx[k] * per + c];
}
void launch_gather_walkers(const float* r, const int32_t* idx, int count, int n, float* out, hipStream_t s) {
  const long total = (long)count * 3 * n;
  hipLaunchKernelGGL(k_gather_walkers, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, r, idx, count, n, out);
}
__global__ void k_scatter_values(const float* __restrict__ v, const int32_t* __restrict__ idx, int count,
                                 float* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) out[idx[t]] = v[t];
}
void launch_scatter_values(const float* v, const int32_t* idx, int count, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter_values, dim3((count + 255) / 256), dim3(256), 0, s, v, idx, count, out);
}

// r' = r + normal * stepsize
__global__ void k_propose(const float* __restrict__ r, const float* __restrict__ normals,
                          const ChainScalars* __restrict__ chain, float* __restrict__ r_prop, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) r_prop[t] = r[t] + normals[t] * chain->stepsize;
}
void launch_propose(const float* r, const float* normals, const void* chain, float* r_prop, int count, hipStream_t s) {
  hipLaunchKernelGGL(k_propose, dim3((count + 255) / 256), dim3(256), 0, s, r, normals, (const ChainScalars*)chain,
                     r_prop, count);
}

// accept = exp(lp' - lp + log q) > u  or  old age >= max_age;  age <- 0 / age + 1;  lp and r selected.
// log q is the proposal's log density ratio (absent, i.e. 0, for the symmetric Gaussian proposal).
__device__ __forceinline__ void chain_update(ChainScalars* chain, int accepted, int B, float* out_acc, int interval,
                                             float target, float lo, float hi);

__global__ __launch_bounds__(256) void k_accept(float* __restrict__ r, const float* __restrict__ r_prop,
                                                float* __restrict__ lp, const float* __restrict__ lp_prop,
                                                const float* __restrict__ logq,
                                                const float* __restrict__ uniforms, int32_t* __restrict__ age,
                                                int max_age, int32_t* __restrict__ acc_count,
                                                uint32_t* __restrict__ flags, int B, int n, ChainScalars* chain,
                                                int32_t* __restrict__ ticket, float* out_acc, int interval,
                                                float target, float lo, float hi) {
  // 64 walkers per workgroup: wave 0 decides (one lane per walker), all waves copy accepted positions
  __shared__ int acc_sh[64];
  const int tid = threadIdx.x;
  const long b0 = (long)blockIdx.x * 64;
  if (tid < 64) {
    const long b = b0 + tid;
    bool acc = false;
    if (b < B) {
      const float p_accept = expf(logq ? lp_prop[b] - lp[b] + logq[b] : lp_prop[b] - lp[b]);
      const bool by_ratio = p_accept > uniforms[b];
      const int old_age = age[b];
      const bool forced = old_age >= max_age;
      acc = by_ratio || forced;
      age[b] = acc ? 0 : old_age + 1;
      if (acc) lp[b] = lp_prop[b];
      if (flags) flags[b] = (forced && !by_ratio) ? DWF_FORCED : 0u;
    }
    acc_sh[tid] = acc ? 1 : 0;
    const unsigned long long bal = __ballot(acc);
    if (tid == 0 && bal) atomicAdd(acc_count, (int)__popcll(bal));
  }
  __syncthreads();
  const int per = 3 * n;
  const long nw = (B - b0) < 64 ? (B - b0) : 64;
  const long base = b0 * per;
  for (long t = tid; t < nw * per; t += 256) {
    const int w = (int)(t / per);
    if (acc_sh[w]) r[base + t] = r_prop[base + t];
  }
  // the last workgroup to finish (a ticket per step, zeroed before the chain) updates the chain from the
  // complete count; thread 0 made this workgroup's addition to the count above
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(ticket, 1) == (int)gridDim.x - 1) {
      __threadfence();
      chain_update(chain, atomicAdd(acc_count, 0), B, out_acc, interval, target, lo, hi);
    }
  }
}
void launch_accept(float* r, const float* r_prop, float* lp, const float* lp_prop, const float* logq,
                   const float* uniforms, int32_t* age, int max_age, int32_t* acc_count, uint32_t* flags, int B,
                   int n, void* chain, int32_t* ticket, float* out_acc, int interval, float target, float lo,
                   float hi, hipStream_t s) {
  hipLaunchKernelGGL(k_accept, dim3((unsigned)((B + 63) / 64)), dim3(256), 0, s, r, r_prop, lp, lp_prop, logq,
                     uniforms, age, max_age, acc_count, flags, B, n, (ChainScalars*)chain, ticket, out_acc,
                     interval, target, lo, hi);
}

// acc = mean(accept); step_nr += 1; acc_rate = 0.9 acc_rate + 0.1 acc; every `interval` steps the
// stepsize is scaled by 1/1.05 or 1.05 from the OLD acc_rate and clipped (mcmc.py:367-387).
__device__ __forceinline__ void chain_update(ChainScalars* chain, int accepted, int B, float* out_acc, int interval,
                                             float target, float lo, float hi) {
  const float acc = (float)accepted / (float)B;
  const float old_rate = chain->acc_rate;
  const float old_step = chain->stepsize;
  chain->step_nr += 1;
  chain->acc_rate = 0.9f * old_rate + 0.1f * acc;
  if (interval > 0 && (chain->step_nr % interval) == 0) {
    float s = (old_rate < target) ? old_step / 1.05f : old_step * 1.05f;
    s = fminf(fmaxf(s, lo), hi);
    chain->stepsize = s;
  }
  if (out_acc) *out_acc = acc;
}

}  // namespace dwmi
