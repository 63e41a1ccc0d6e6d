// Reverse-mode pass for the parameter gradient G = sum_b c_b d log|psi(r_b)|^2 / d theta: the weighted
// tangent that the reference's total-energy JVP contracts with the clipped energy deviations
// (optimization/loss_function.py:75-156).  It runs after a value pass whose activations (X_l, H, O, l_d,
// s_d) are kept, and walks the network backwards:
//   determinants (FP64 inverse) -> orbitals / envelope -> 1-el layers (tanh) -> spin means and pair means
//   -> 2-el stream (recomputed per pair).
// Weight gradients are reductions over walkers, electrons or pairs; every reduction here runs in a fixed
// order (chunk partials, then a chunk-ordered sum), so the result does not depend on timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "common.h"
#include "grad.h"

namespace dwmi {

namespace {

__device__ __forceinline__ float gexp(float x) { return __expf(x); }
__device__ __forceinline__ double gexp(double x) { return exp(x); }
__device__ __forceinline__ float gsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double gsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float gtanh(float x) { return tanh_fast(x); }
__device__ __forceinline__ double gtanh(double x) { return tanh(x); }

// Spin-major row of electron i (one row per electron): up block, then down block.
__device__ __forceinline__ long vrow(const Shape& sh, long b, int i) {
  const int dn = sh.n - sh.u;
  return (i < sh.u) ? b * sh.u + i : (long)sh.B * sh.u + b * dn + (i - sh.u);
}

// ---- seed -------------------------------------------------------------------------------------------
// log_psi_sqr = 2 (log(|S| + eps) + m), S = sum_d s_d e^{l_d - m}, m = max_d l_d.  With
// w_d = sgn(S) s_d e^{l_d - m} / (|S| + eps):  d/dl_d = 2 w_d  (+ the share of d/dm = 2 - 2 sum_d w_d that
// the max passes to its arguments, split evenly over ties as jnp.max's derivative does).
__global__ void k_grad_seed(const double* __restrict__ ld, const double* __restrict__ sd, const float* __restrict__ c,
                            double* __restrict__ gl, Shape sh) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= sh.B) return;
  const int D = sh.D;
  const double* l = ld + b * D;
  const double* sg = sd + b * D;
  double m = -INFINITY;
  for (int d = 0; d < D; ++d) m = fmax(m, l[d]);
  double S = 0.0;
  int ties = 0;
  for (int d = 0; d < D; ++d) {
    S += sg[d] * exp(l[d] - m);
    if (l[d] == m) ++ties;
  }
  const double den = fabs(S) + 1e-8;
  const double sgn = S < 0.0 ? -1.0 : 1.0;
  double wsum = 0.0;
  for (int d = 0; d < D; ++d) wsum += sgn * sg[d] * exp(l[d] - m) / den;
  const double dm = (2.0 - 2.0 * wsum) / (double)ties;
  const double cb = (double)c[b];
  for (int d = 0; d < D; ++d) {
    const double w = sgn * sg[d] * exp(l[d] - m) / den;
    gl[b * D + d] = cb * (2.0 * w + (l[d] == m ? dm : 0.0));
  }
}

// ---- determinants -----------------------------------------------------------------------------------
// One wavefront per (walker, det); the matrix and its inverse live in LDS as FP64 (Gauss-Jordan with
// partial pivoting on [A | I], lanes over the 2n columns).  n <= 63.
template <typename T>
__global__ __launch_bounds__(64) void k_det_bwd(OrbSrc<T> src, const double* __restrict__ gl, T* __restrict__ dbf,
                                                T* __restrict__ denv, int ldg, Shape sh) {
  extern __shared__ double lds[];
  const int n = sh.n, n2 = 2 * n, D = sh.D, NO = D * n;
  const long bd = blockIdx.x;
  const long b = bd / D;
  const int d = (int)(bd - b * D);
  const int lane = threadIdx.x;
  double* M = lds;                 // [n][2n]
  double* envs = lds + n * n2;     // [n][n] envelope of element (i, k)
  // form A (rows i = electrons, columns k = orbitals of det d) and the identity block
  for (int t = lane; t < n * n; t += 64) {
    const int i = t / n, k = t - (t / n) * n;
    const int col = d * n + k;
    const int s = (i < sh.u) ? 0 : 1;
    const float* ri = src.r + (b * n + i) * 3;
    T env = T(0);
    for (int I = 0; I < sh.I; ++I) {
      const T dx = (T)ri[0] - (T)src.R[I * 3 + 0], dy = (T)ri[1] - (T)src.R[I * 3 + 1],
              dz = (T)ri[2] - (T)src.R[I * 3 + 2];
      const T rho = gsqrt(dx * dx + dy * dy + dz * dz);
      env += src.wenv[((long)s * sh.I + I) * NO + col] * gexp(-src.kappa[((long)s * sh.I + I) * NO + col] * rho);
    }
    const T bf = src.O[vrow(sh, b, i) * src.ldo + col];
    M[i * n2 + k] = (double)(env * bf);
    M[i * n2 + n + k] = (i == k) ? 1.0 : 0.0;
    envs[i * n + k] = (double)env;
  }
  __syncthreads();
  __shared__ double fcol[64];
  for (int p = 0; p < n; ++p) {
    // pivot: the row r >= p with the largest |M[r][p]| (the first of equal ones), lanes over rows
    double v = -1.0;  // lanes outside [p, n) never win; a NaN entry counts as the largest
    if (lane >= p && lane < n) {
      const double a = fabs(M[lane * n2 + p]);
      v = isnan(a) ? 1.0e308 : a;
    }
    int idx = lane;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double v2 = __shfl_xor(v, o, 64);
      const int i2 = __shfl_xor(idx, o, 64);
      if (v2 > v || (v2 == v && i2 < idx)) {
        v = v2;
        idx = i2;
      }
    }
    const int pr = (idx >= p && idx < n) ? idx : p;
    if (pr != p)
      for (int j = lane; j < n2; j += 64) {
        const double t0 = M[p * n2 + j];
        M[p * n2 + j] = M[pr * n2 + j];
        M[pr * n2 + j] = t0;
      }
    __syncthreads();
    const double inv = 1.0 / M[p * n2 + p];
    if (lane < n) fcol[lane] = M[lane * n2 + p];  // factors of every row, read before any row changes
    __syncthreads();
    for (int j = lane; j < n2; j += 64) M[p * n2 + j] *= inv;
    __syncthreads();
    // eliminate column p from every other row: lanes over all (row, column) elements
    for (int t = lane; t < n * n2; t += 64) {
      const int rr = t / n2, j = t - (t / n2) * n2;
      if (rr != p) M[rr * n2 + j] -= fcol[rr] * M[p * n2 + j];
    }
    __syncthreads();
  }
  // the padding columns [NO, ldg) of this walker's rows are zero (they meet zero weight rows in dH)
  if (d == D - 1)
    for (int t = lane; t < n * (ldg - NO); t += 64) {
      const int i = t / (ldg - NO), cc = NO + t - (t / (ldg - NO)) * (ldg - NO);
      const long row = vrow(sh, b, i);
      dbf[row * ldg + cc] = T(0);
      denv[row * ldg + cc] = T(0);
    }
  // dL/dA[i][k] = g Ainv[k][i]; A = env * bf  ->  d bf = dA env, d env = dA bf
  const double g = gl[b * D + d];
  for (int t = lane; t < n * n; t += 64) {
    const int i = t / n, k = t - (t / n) * n;
    const double dA = g * M[k * n2 + n + i];
    const long row = vrow(sh, b, i);
    const int col = d * n + k;
    const double env = envs[i * n + k];
    const T bf = src.O[row * src.ldo + col];
    dbf[row * ldg + col] = (T)(dA * env);
    denv[row * ldg + col] = (T)(dA * (double)bf);
  }
}

// ---- envelope parameters ----------------------------------------------------------------------------
template <typename T>
__global__ void k_env_grad(const T* __restrict__ denv, int ldg, const float* __restrict__ r, const float* __restrict__ R,
                           const T* __restrict__ kappa, const T* __restrict__ wenv, T* __restrict__ part,
                           int wpc, Shape sh) {
  const int NO = sh.D * sh.n, I = sh.I;
  const int per_chunk = 2 * I * NO;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  if (t >= per_chunk) return;
  const int s = (int)(t / (I * NO));
  const int rem = (int)(t - (long)s * I * NO);
  const int Ii = rem / NO, col = rem - (rem / NO) * NO;
  const T k = kappa[((long)s * I + Ii) * NO + col];
  const T w = wenv[((long)s * I + Ii) * NO + col];
  const T dsig = T(1) - gexp(-k);  // d softplus / d alpha = sigmoid(alpha) = 1 - e^{-softplus(alpha)}
  const int i0 = s ? sh.u : 0, i1 = s ? sh.n : sh.u;
  const long b0 = (long)chunk * wpc;
  const long b1 = b0 + wpc < sh.B ? b0 + wpc : sh.B;
  T sw = T(0), sa = T(0);
  for (long b = b0; b < b1; ++b)
    for (int i = i0; i < i1; ++i) {
      const float* ri = r + (b * sh.n + i) * 3;
      const T dx = (T)ri[0] - (T)R[Ii * 3 + 0], dy = (T)ri[1] - (T)R[Ii * 3 + 1], dz = (T)ri[2] - (T)R[Ii * 3 + 2];
      const T rho = gsqrt(dx * dx + dy * dy + dz * dz);
      const T e = gexp(-k * rho);
      const T g = denv[vrow(sh, b, i) * ldg + col];
      sw += g * e;
      sa += g * w * e * (-rho);
    }
  T* p = part + (long)chunk * 2 * per_chunk;
  // layout [chunk][alpha | w][spin][I][NO]: the order of the leaves alpha_up, alpha_dn, weights_up, weights_dn
  p[((0L * 2 + s) * I + Ii) * NO + col] = sa * dsig;
  p[((1L * 2 + s) * I + Ii) * NO + col] = sw;
}

// ---- 1-el layers ------------------------------------------------------------------------------------
template <typename T>
__global__ void k_dz_last(const T* __restrict__ dH, const T* __restrict__ H, T* __restrict__ dz, Shape sh) {
  const int W = sh.W1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)sh.B * sh.n * W) return;
  const long be = t / W;
  const int o = (int)(t - be * W);
  const long b = be / sh.n;
  const int i = (int)(be - b * sh.n);
  const long row = vrow(sh, b, i);
  const T h = H[row * W + o];
  dz[t] = dH[row * W + o] * (T(1) - h * h);
}

template <typename T>
__global__ void k_dz(const T* __restrict__ dX, int ldx, const T* __restrict__ dMU, int ldmu, const T* __restrict__ X,
                     int ldX, int Kh, T* __restrict__ dz, Shape sh) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)sh.B * sh.n * Kh) return;
  const long be = t / Kh;
  const int o = (int)(t - be * Kh);
  const long b = be / sh.n;
  const int i = (int)(be - b * sh.n);
  const int s = i < sh.u ? 0 : 1;
  const T ns = (T)(s ? sh.n - sh.u : sh.u);
  const T h = X[be * ldX + o];
  const T dh = dX[be * ldx + o] + dMU[b * ldmu + s * Kh + o] / ns;
  dz[be * Kh + o] = dh * (T(1) - h * h);
}

template <typename T>
__global__ void k_walker_sum(const T* __restrict__ in, int ld, int ncols, T* __restrict__ out, Shape sh) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)sh.B * ncols) return;
  const long b = t / ncols;
  const int o = (int)(t - b * ncols);
  T acc = T(0);
  for (int i = 0; i < sh.n; ++i) acc += in[(b * sh.n + i) * ld + o];
  out[b * ncols + o] = acc;
}

}  // namespace

// Row chunks of a column sum: about 128 chunks of at least 256 rows.
int colsum_chunks(long rows) { return (int)std::min<long>(128, std::max<long>(1, (rows + 255) / 256)); }

namespace {

// Column sums of a chunk of rows: 32 columns x 8 row lanes per 256-thread block; the 8 row lanes meet in
// LDS in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k_colsum_part(const T* __restrict__ in, int ld, long rows, int ncols,
                                                     T* __restrict__ part) {
  __shared__ T red[8][32];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int o = blockIdx.x * 32 + tx;
  const long per = (rows + gridDim.y - 1) / gridDim.y;
  const long r0 = (long)blockIdx.y * per;
  const long r1 = r0 + per < rows ? r0 + per : rows;
  T acc = T(0);
  if (o < ncols)
    for (long r = r0 + ty; r < r1; r += 8) acc += in[r * ld + o];
  red[ty][tx] = acc;
  __syncthreads();
  if (ty == 0 && o < ncols) {
    T t = red[0][tx];
#pragma unroll
    for (int q = 1; q < 8; ++q) t += red[q][tx];
    part[(long)blockIdx.y * ncols + o] = t;
  }
}

template <typename T>
__global__ void k_sum_parts(const T* __restrict__ part, int chunks, long count, float* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  T acc = T(0);
  for (int c = 0; c < chunks; ++c) acc += part[(long)c * count + t];
  out[t] = (float)acc;
}

// ---- 2-el stream ------------------------------------------------------------------------------------
// One thread per pair (b, i, j).  h_0 = [|r_j - r_i|, r_j - r_i] (zero on the diagonal), layer 0:
// h_1 = tanh(h_0 W_0 + b_0); layers l >= 1: h_{l+1} = (h_l + tanh(h_l W_l + b_l)) / sqrt 2
// (model/mlp.py:13-68).  Gradient into h_l from the pair means of 1-el layer l: the mean over partners j of
// spin s has weight 1 / n_s, so h_l of pair (i, j) receives dX_l[b n + i, Kh_l + s(j) Ke_l + f] / n_s(j).
template <typename T>
__global__ __launch_bounds__(128) void k_twoel_bwd(const float* __restrict__ r, TwoElBwdArgs<T> a, Shape sh) {
  constexpr int W2 = 32, ML2 = 3;  // at most 3 2-el layers (4 iterations)
  __shared__ T Ws[ML2][W2][W2];
  __shared__ T bs[ML2][W2];
  const int L2 = sh.L - 1;  // 2-el layers
  for (int t = threadIdx.x; t < L2 * W2 * W2; t += blockDim.x) Ws[t / (W2 * W2)][(t / W2) % W2][t % W2] = a.W2e[t];
  for (int t = threadIdx.x; t < L2 * W2; t += blockDim.x) bs[t / W2][t % W2] = a.b2e[t];
  __syncthreads();
  const int n = sh.n;
  const long pairs = (long)sh.B * n * n;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pairs) return;
  const long be = p / n;
  const int j = (int)(p - be * n);
  const long b = be / n;
  const int i = (int)(be - b * n);
  const int up = j < sh.u ? 1 : 0;  // partner j is spin up
  const T inv_ns = T(1) / (T)(up ? sh.u : n - sh.u);
  const T rs = (T)0.70710678118654752440;
  T h[W2], y[ML2][W2];
#pragma unroll
  for (int k = 0; k < W2; ++k) h[k] = T(0);
  {
    const float* ri = r + (b * n + i) * 3;
    const float* rj = r + (b * n + j) * 3;
    const T dx = (T)rj[0] - (T)ri[0], dy = (T)rj[1] - (T)ri[1], dz = (T)rj[2] - (T)ri[2];
    h[0] = (i == j) ? T(0) : gsqrt(dx * dx + dy * dy + dz * dz);
    h[1] = dx;
    h[2] = dy;
    h[3] = dz;
  }
  // forward, writing each layer's input
#pragma unroll
  for (int l = 0; l < ML2; ++l) {
    if (l < L2) {
      const int kin = l == 0 ? 4 : W2;
      T* xo = a.X2[l] + p * kin;
#pragma unroll
      for (int k = 0; k < W2; ++k)
        if (k < kin) xo[k] = h[k];
#pragma unroll
      for (int f = 0; f < W2; ++f) {
        T z = bs[l][f];
#pragma unroll
        for (int k = 0; k < W2; ++k)
          if (k < kin) z += h[k] * Ws[l][k][f];
        y[l][f] = gtanh(z);
      }
#pragma unroll
      for (int f = 0; f < W2; ++f) h[f] = (l == 0) ? y[l][f] : (h[f] + y[l][f]) * rs;
    }
  }
  // backward: g is the gradient of h_{l+1}
  T g[W2];
  {
    const int l = sh.L - 1;
    const T* dpm = a.dX[l] + be * a.ldx[l] + a.Kh[l] + (up ? 0 : W2);
#pragma unroll
    for (int f = 0; f < W2; ++f) g[f] = dpm[f] * inv_ns;
  }
#pragma unroll
  for (int l = ML2 - 1; l >= 0; --l) {
    if (l < L2) {
      T dp[W2];
#pragma unroll
      for (int f = 0; f < W2; ++f) dp[f] = ((l == 0) ? g[f] : g[f] * rs) * (T(1) - y[l][f] * y[l][f]);
      T* dpo = a.DP[l] + p * W2;
#pragma unroll
      for (int f = 0; f < W2; ++f) dpo[f] = dp[f];
      if (l > 0) {
        // gradient into h_l: the residual branch, the layer, and the pair mean of 1-el layer l
        const T* dpm = a.dX[l] + be * a.ldx[l] + a.Kh[l] + (up ? 0 : W2);
        T gn[W2];
#pragma unroll
        for (int k = 0; k < W2; ++k) {
          T acc = g[k] * rs + dpm[k] * inv_ns;
#pragma unroll
          for (int f = 0; f < W2; ++f) acc += Ws[l][k][f] * dp[f];
          gn[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < W2; ++k) g[k] = gn[k];
      }
    }
  }
}

// ---- weight gradients -------------------------------------------------------------------------------
// C[m, n] = sum_k A[k, m] B[k, n].  The K rows are split into chunks (about 1024 workgroups in all); each
// chunk's tile goes to its own partial buffer, and k_sum_parts_ld adds the chunks in order.
// FP32, M or N > 32: 64 x 64 tiles per 256-thread workgroup, v_mfma_f32_32x32x2_f32 (4 waves of 32 x 32),
// K slices of 16 rows staged through LDS, the next slice's loads issued before the current slice's products.
// FP32, M and N <= 32 (2-el layers): one 32 x 32 block, no LDS staging: each wave streams every fourth
// 16-row slice of the chunk straight into MFMA operands (lane l takes row l / 32, column l % 32, which
// are coalesced 128-byte rows), and the four waves' accumulators meet in LDS in a fixed order.
// FP64: 64 x 64 tiles, 4 x 4 outputs per thread on FMA.
constexpr int kWgradBK = 16;

__global__ __launch_bounds__(256) void k_wgrad_f32(const float* __restrict__ A, int lda, const float* __restrict__ Bm,
                                                   int ldb, int M, int N, long K, long kchunk,
                                                   float* __restrict__ part, int tilesN) {
  constexpr int BM = 64, BN = 64, BK = kWgradBK;
  __shared__ float As[BK][BM + 4];
  __shared__ float Bs[BK][BN + 4];
  const int tile = blockIdx.x, chunk = blockIdx.y;
  const int m0 = (tile / tilesN) * BM, n0 = (tile % tilesN) * BN;
  const long k0 = (long)chunk * kchunk;
  const long k1 = k0 + kchunk < K ? k0 + kchunk : K;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  // one float4 of A and of B per thread and slice when the tile is full and aligned
  const bool vec = ((lda & 3) == 0) && ((ldb & 3) == 0) && ((((uintptr_t)A) & 15) == 0) &&
                   ((((uintptr_t)Bm) & 15) == 0) && m0 + BM <= M && n0 + BN <= N;
  const int vr = tid >> 4, vc = (tid & 15) * 4;
  v16f acc;
  for (int j = 0; j < 16; ++j) acc[j] = 0.f;
  float4 ra = make_float4(0.f, 0.f, 0.f, 0.f), rb = ra;
  auto gload = [&](long kb) {
    const long k = kb + vr;
    if (k < k1) {
      ra = *(const float4*)(A + k * lda + m0 + vc);
      rb = *(const float4*)(Bm + k * ldb + n0 + vc);
    } else {
      ra = make_float4(0.f, 0.f, 0.f, 0.f);
      rb = ra;
    }
  };
  if (vec) gload(k0);
  for (long kb = k0; kb < k1; kb += BK) {
    __syncthreads();
    if (vec) {
      *(float4*)&As[vr][vc] = ra;
      *(float4*)&Bs[vr][vc] = rb;
    } else {
      for (int t = tid; t < BK * BM; t += 256) {
        const int kk = t / BM, mm = t - (t / BM) * BM;
        const long k = kb + kk;
        As[kk][mm] = (k < k1 && m0 + mm < M) ? A[k * lda + m0 + mm] : 0.f;
        Bs[kk][mm] = (k < k1 && n0 + mm < N) ? Bm[k * ldb + n0 + mm] : 0.f;
      }
    }
    __syncthreads();
    if (vec && kb + BK < k1) gload(kb + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      const int kr = kk + (lane >> 5);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(As[kr][wm * 32 + (lane & 31)], Bs[kr][wn * 32 + (lane & 31)], acc,
                                                 0, 0, 0);
    }
  }
  float* P = part + (long)chunk * M * N;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int mm = m0 + wm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    const int nn = n0 + wn * 32 + (lane & 31);
    if (mm < M && nn < N) P[(long)mm * N + nn] = acc[j];
  }
}

__global__ __launch_bounds__(256) void k_wgrad_small_f32(const float* __restrict__ A, int lda,
                                                         const float* __restrict__ Bm, int ldb, int M, int N, long K,
                                                         long kchunk, float* __restrict__ part) {
  __shared__ float red[4][16][64];
  const int chunk = blockIdx.x;
  const long k0 = (long)chunk * kchunk;
  const long k1 = k0 + kchunk < K ? k0 + kchunk : K;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int col = lane & 31, half = lane >> 5;
  const bool am = col < M, bn = col < N;
  v16f acc;
  for (int j = 0; j < 16; ++j) acc[j] = 0.f;
  for (long kb = k0 + (long)wid * kWgradBK; kb < k1; kb += 4L * kWgradBK) {
    float a[kWgradBK / 2], b[kWgradBK / 2];
#pragma unroll
    for (int s = 0; s < kWgradBK / 2; ++s) {
      const long k = kb + 2 * s + half;
      a[s] = (am && k < k1) ? A[k * lda + col] : 0.f;
      b[s] = (bn && k < k1) ? Bm[k * ldb + col] : 0.f;
    }
#pragma unroll
    for (int s = 0; s < kWgradBK / 2; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], b[s], acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) red[wid][j][lane] = acc[j];
  __syncthreads();
  if (wid == 0) {
    float* P = part + (long)chunk * M * N;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float v = ((red[0][j][lane] + red[1][j][lane]) + red[2][j][lane]) + red[3][j][lane];
      const int mm = 8 * (j >> 2) + 4 * half + (j & 3);
      if (mm < M && col < N) P[(long)mm * N + col] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_wgrad_f64(const double* __restrict__ A, int lda, const double* __restrict__ Bm,
                                                   int ldb, int M, int N, long K, long kchunk,
                                                   double* __restrict__ part, int tilesN) {
  constexpr int BM = 64, BN = 64, BK = 16;
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  const int tile = blockIdx.x, chunk = blockIdx.y;
  const int m0 = (tile / tilesN) * BM, n0 = (tile % tilesN) * BN;
  const long k0 = (long)chunk * kchunk;
  const long k1 = k0 + kchunk < K ? k0 + kchunk : K;
  const int tid = threadIdx.x;
  const int tm = (tid >> 4) * 4, tn = (tid & 15) * 4;
  double acc[4][4] = {};
  for (long kb = k0; kb < k1; kb += BK) {
    __syncthreads();
    for (int t = tid; t < BK * BM; t += 256) {
      const int kk = t / BM, mm = t - (t / BM) * BM;
      const long k = kb + kk;
      As[kk][mm] = (k < k1 && m0 + mm < M) ? A[k * lda + m0 + mm] : 0.0;
      Bs[kk][mm] = (k < k1 && n0 + mm < N) ? Bm[k * ldb + n0 + mm] : 0.0;
    }
    __syncthreads();
    for (int kk = 0; kk < BK; ++kk)
      for (int x = 0; x < 4; ++x)
        for (int y = 0; y < 4; ++y) acc[x][y] = fma(As[kk][tm + x], Bs[kk][tn + y], acc[x][y]);
  }
  double* P = part + (long)chunk * M * N;
  for (int x = 0; x < 4; ++x)
    for (int y = 0; y < 4; ++y) {
      const int mm = m0 + tm + x, nn = n0 + tn + y;
      if (mm < M && nn < N) P[(long)mm * N + nn] = acc[x][y];
    }
}

template <typename T>
__global__ void k_sum_parts_ld(const T* __restrict__ part, int chunks, int M, int N, float* __restrict__ C, int ldc) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)M * N) return;
  T acc = T(0);
  for (int c = 0; c < chunks; ++c) acc += part[(long)c * M * N + t];
  const int m = (int)(t / N), nn = (int)(t - (long)m * N);
  C[(long)m * ldc + nn] = (float)acc;
}

}  // namespace

void launch_grad_seed(const double* ld, const double* sd, const float* c, double* gl, const Shape& sh, hipStream_t s) {
  hipLaunchKernelGGL(k_grad_seed, dim3((sh.B + 255) / 256), dim3(256), 0, s, ld, sd, c, gl, sh);
}

template <typename T>
void launch_det_bwd(const OrbSrc<T>& src, const double* gl, T* dbf, T* denv, int ld, const Shape& sh, hipStream_t s) {
  const size_t smem = (size_t)sh.n * 3 * sh.n * sizeof(double);
  hipLaunchKernelGGL(k_det_bwd<T>, dim3((unsigned)((long)sh.B * sh.D)), dim3(64), smem, s, src, gl, dbf, denv, ld, sh);
}

template <typename T>
int launch_env_grad(const T* denv, int ld, const float* r, const float* R, const T* kappa, const T* wenv, T* part,
                    int wpc, const Shape& sh, hipStream_t s) {
  const int chunks = (sh.B + wpc - 1) / wpc;
  const long per = 2L * sh.I * sh.D * sh.n;
  hipLaunchKernelGGL(k_env_grad<T>, dim3((unsigned)((per + 255) / 256), chunks), dim3(256), 0, s, denv, ld, r, R, kappa,
                     wenv, part, wpc, sh);
  return chunks;
}

template <typename T>
void launch_dz_last(const T* dH, const T* H, T* dz, const Shape& sh, hipStream_t s) {
  const long total = (long)sh.B * sh.n * sh.W1;
  hipLaunchKernelGGL(k_dz_last<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, dH, H, dz, sh);
}

template <typename T>
void launch_dz(const T* dX, int ldx, const T* dMU, int ldmu, const T* X, int ldX, int Kh, T* dz, const Shape& sh,
               hipStream_t s) {
  const long total = (long)sh.B * sh.n * Kh;
  hipLaunchKernelGGL(k_dz<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, dX, ldx, dMU, ldmu, X, ldX, Kh,
                     dz, sh);
}

template <typename T>
void launch_walker_sum(const T* in, int ld, int ncols, T* out, const Shape& sh, hipStream_t s) {
  const long total = (long)sh.B * ncols;
  hipLaunchKernelGGL(k_walker_sum<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, ld, ncols, out, sh);
}

template <typename T>
void launch_colsum(const T* in, int ld, long rows, int ncols, T* part, float* out, hipStream_t s) {
  const int chunks = colsum_chunks(rows);
  hipLaunchKernelGGL(k_colsum_part<T>, dim3((ncols + 31) / 32, chunks), dim3(256), 0, s, in, ld, rows, ncols, part);
  hipLaunchKernelGGL(k_sum_parts<T>, dim3((ncols + 255) / 256), dim3(256), 0, s, part, chunks, (long)ncols, out);
}

template <typename T>
void launch_sum_parts(const T* part, int chunks, long count, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_parts<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, part, chunks, count, out);
}

template <typename T>
void launch_twoel_bwd(const float* r, T* const* dX, const int* Kx, const int* Kh, const T* W2e, const T* b2e,
                      T* const* X2, T* const* DP, const Shape& sh, hipStream_t s) {
  TwoElBwdArgs<T> a{};
  for (int l = 0; l < sh.L; ++l) {
    a.dX[l] = dX[l];
    a.ldx[l] = Kx[l];
    a.Kh[l] = Kh[l];
  }
  for (int l = 0; l < sh.L - 1; ++l) {
    a.X2[l] = X2[l];
    a.DP[l] = DP[l];
  }
  a.W2e = W2e;
  a.b2e = b2e;
  const long pairs = (long)sh.B * sh.n * sh.n;
  hipLaunchKernelGGL(k_twoel_bwd<T>, dim3((unsigned)((pairs + 127) / 128)), dim3(128), 0, s, r, a, sh);
}

// Chunks of the K rows: about 1024 workgroups in all, at least 64 rows per chunk, at most 256 chunks.
static void wgrad_split(int M, int N, long K, int& chunks, long& kchunk) {
  const bool small = M <= 32 && N <= 32;
  const long tiles = small ? 1 : (long)((M + 63) / 64) * ((N + 63) / 64);
  long c = (1024 + tiles - 1) / tiles;
  c = std::min<long>(c, 256);
  c = std::min<long>(c, (K + 63) / 64);
  c = std::max<long>(c, 1);
  kchunk = (K + c - 1) / c;
  kchunk = (kchunk + kWgradBK - 1) / kWgradBK * kWgradBK;
  chunks = (int)((K + kchunk - 1) / kchunk);
}

size_t wgrad_part_elems(int M, int N, long K) {
  int chunks;
  long kchunk;
  wgrad_split(M, N, K, chunks, kchunk);
  return (size_t)chunks * (size_t)M * (size_t)N;
}

template <>
void wgrad<float>(const float* A, int lda, const float* Bm, int ldb, int M, int N, long K, float* part, float* C,
                  int ldc, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  int chunks;
  long kchunk;
  wgrad_split(M, N, K, chunks, kchunk);
  if (M <= 32 && N <= 32) {
    hipLaunchKernelGGL(k_wgrad_small_f32, dim3(chunks), dim3(256), 0, s, A, lda, Bm, ldb, M, N, K, kchunk, part);
  } else {
    const int tilesN = (N + 63) / 64, tiles = ((M + 63) / 64) * tilesN;
    hipLaunchKernelGGL(k_wgrad_f32, dim3(tiles, chunks), dim3(256), 0, s, A, lda, Bm, ldb, M, N, K, kchunk, part,
                       tilesN);
  }
  hipLaunchKernelGGL(k_sum_parts_ld<float>, dim3((unsigned)(((long)M * N + 255) / 256)), dim3(256), 0, s, part, chunks,
                     M, N, C, ldc);
}

template <>
void wgrad<double>(const double* A, int lda, const double* Bm, int ldb, int M, int N, long K, double* part, float* C,
                   int ldc, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  int chunks;
  long kchunk;
  wgrad_split(M, N, K, chunks, kchunk);
  const int tilesN = (N + 63) / 64, tiles = ((M + 63) / 64) * tilesN;
  hipLaunchKernelGGL(k_wgrad_f64, dim3(tiles, chunks), dim3(256), 0, s, A, lda, Bm, ldb, M, N, K, kchunk, part, tilesN);
  hipLaunchKernelGGL(k_sum_parts_ld<double>, dim3((unsigned)(((long)M * N + 255) / 256)), dim3(256), 0, s, part, chunks,
                     M, N, C, ldc);
}

#define DW_GRAD_INST(T)                                                                                          \
  template void launch_det_bwd<T>(const OrbSrc<T>&, const double*, T*, T*, int, const Shape&, hipStream_t);       \
  template int launch_env_grad<T>(const T*, int, const float*, const float*, const T*, const T*, T*, int,         \
                                  const Shape&, hipStream_t);                                                      \
  template void launch_dz_last<T>(const T*, const T*, T*, const Shape&, hipStream_t);                             \
  template void launch_dz<T>(const T*, int, const T*, int, const T*, int, int, T*, const Shape&, hipStream_t);    \
  template void launch_walker_sum<T>(const T*, int, int, T*, const Shape&, hipStream_t);                          \
  template void launch_colsum<T>(const T*, int, long, int, T*, float*, hipStream_t);                              \
  template void launch_sum_parts<T>(const T*, int, long, float*, hipStream_t);                                    \
  template void launch_twoel_bwd<T>(const float*, T* const*, const int*, const int*, const T*, const T*,          \
                                    T* const*, T* const*, const Shape&, hipStream_t);
DW_GRAD_INST(float)
DW_GRAD_INST(double)

}  // namespace dwmi
